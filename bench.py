"""Frames/s of MID-Fusion's per-frame dense path on one MI355X.

A step is one frame: preprocessing (inputs already in HBM), camera + moving
object tracking (ICP + photometric, 3 levels {10,5,4}), allocation +
integration of every instance, and raycasting.

Default workload (camera + K = 8 objects): 640x480 (fx=fy=525, cx=319.5,
cy=239.5), background volume 512^3 at 1 cm, 8 seeded moving boxes and
ellipsoids of 0.05-0.4 m, each in its own 64^3-256^3 volume at 2-5 mm, on
constant twists, seeded synthetic frames; all objects batched per launch.
The same run times the one-object-per-launch form of the same workload
(`vs_plain_gpu`), the single-object workload (one 0.2x0.15x0.1 m box in a
128^3 volume at 2 mm), the CPU oracle at the reference's compiler flags and
at -O3 on the host cores, and the pose and TSDF error of the GPU path against
the FP64 oracle.  Prints one JSON line.
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, 'mid-fusion_amd')


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


METRIC = 'frames/s (track+integrate+raycast, 640x480, cam+K objs) vs plain GPU; pose err'
HBM_PEAK_GBS = 8000.0
W, H = 640, 480
K = np.array([525.0, 525.0, 319.5, 239.5], np.float32)
BG_RES = 512            # background volume resolution (5.12 m cube)
HD = False              # --hd: the precision + scale workload


def make_scene(scene, n_objects):
    """One object: the textured 0.2x0.15x0.1 m box in a 128^3 volume at 2 mm.
    K > 1 objects: seeded boxes and ellipsoids of 0.05-0.4 m, each in a
    64^3-256^3 volume at 2-5 mm, on constant twists."""
    if HD:
        # 1280x720, 16 seeded boxes/ellipsoids, 2 % random invalid depth pixels
        movers = scene.batch_objects(n_objects, 5000)
        return scene.Scene(seed=5000, width=W, height=H, K=tuple(K), moving=movers, invalid_frac=0.02), movers
    if n_objects == 1:
        rng = np.random.default_rng(2)
        T0 = np.eye(4)
        T0[:3, 3] = [2.30 - 0.128, 2.45 - 0.128, 2.45]
        tw = np.r_[rng.uniform(-0.004, 0.004, 3), rng.uniform(-0.015, 0.015, 3)]
        movers = [scene.MovingBox([0.1, 0.075, 0.05], T0, tw, rng.uniform(0.3, 0.9, 3), 0.256, 128)]
        seed = 2
    else:
        movers = scene.batch_objects(n_objects, 4000 + n_objects)
        seed = 4
    return scene.Scene(seed=seed, width=W, height=H, K=tuple(K), moving=movers), movers


def workload_name(n_obj, plain):
    if HD:
        w = ('1280x720 RGB-D (fx=fy=1050, 2%% invalid depth), bg 1024^3 @5mm + %d moving boxes/ellipsoids '
             '0.05-0.4 m in 64^3-256^3 @2-5mm' % n_obj)
    elif n_obj == 1:
        w = '640x480 RGB-D, bg 512^3 @1cm + 1 moving box 0.2x0.15x0.1 m in 128^3 @2mm'
    else:
        w = '640x480 RGB-D, bg 512^3 @1cm + %d moving boxes/ellipsoids 0.05-0.4 m in 64^3-256^3 @2-5mm' % n_obj
    return w + ', ICP+RGB {10,5,4}, %s' % ('one object per launch' if plain else 'objects batched per launch')


_RENDER_SCENE = None


def _render_one(f):
    return _RENDER_SCENE.render(f)


def render_frames(sc, n_objects, n):
    """Frames 0..n-1 of the seeded scene (each frame depends only on its
    index), rendered by a pool of forked host processes when there are many;
    called before anything touches the GPU."""
    global _RENDER_SCENE
    scene = sys.modules['mf_scene']
    frames = []
    workers = min(16, os.cpu_count() or 1, n)
    t0 = time.time()
    if workers > 1 and n >= 16:
        import multiprocessing as mp
        _RENDER_SCENE = sc
        with mp.get_context('fork').Pool(workers) as pool:
            for f, (d, rgb, lab) in enumerate(pool.imap(_render_one, range(n), chunksize=1)):
                frames.append((d, rgb, scene.masks_for_labels(lab, n_objects)))
                if (f + 1) % 25 == 0:
                    print('rendered %d / %d frames (%.0f s)' % (f + 1, n, time.time() - t0), file=sys.stderr, flush=True)
        _RENDER_SCENE = None
    else:
        for f in range(n):
            d, rgb, lab = sc.render(f)
            frames.append((d, rgb, scene.masks_for_labels(lab, n_objects)))
    return frames


TB = 256        # threads per tracking workgroup
REPLAYS = 200   # back-to-back launches of the roofline kernel


def track_bytes(trace, rpt, objects, use_rgb=True):
    """Algorithmic bytes of the executed row launches of the camera
    (objects=False) or of the moving objects: per pixel of the level, ICP
    reads live v, n, cov (36 B) + model v, n (24 B) + mask (1 B); RGB reads
    rendered-reference depth, live/reference vertices, two intensities and two
    gradients (41 B); each workgroup writes a 128-B partial."""
    per_px = 61 + (41 if use_rgb else 0)
    tot = 0
    for r in trace:
        if (int(r[0]) > 0) != objects:
            continue
        lvl = int(r[1])
        n = (W >> lvl) * (H >> lvl)
        nblk = -(-(2 * n) // (TB * rpt[lvl]))
        tot += n * per_px + 128 * nblk
    return tot


def solve_bytes(trace, rpt):
    """Per executed iteration: the workgroup partials read (128 B each) and the
    GN state read and written (~1 KB)."""
    tot = 0
    for r in trace:
        lvl = int(r[1])
        n = (W >> lvl) * (H >> lvl)
        tot += 128 * -(-(2 * n) // (TB * rpt[lvl])) + 1024
    return tot


# kernel class -> kernel name in rocprofv3 output (traffic counters)
KERNEL_NAMES = {'track_rows': 'mf::k_track_px<3>',
                'track_rows_l0': 'mf::k_track_px<3> level 0',
                'track_rows_obj': 'mf::k_track<true, true, true, 3>',
                'track_solve': 'mf::k_solve', 'integrate': 'mf::k_integrate', 'raycast': 'mf::k_raycast',
                'alloc_band': 'mf::k_alloc'}
# committed rocprofv3 counter summary of each workload (FETCH_SIZE / WRITE_SIZE)
TRAFFIC_FILES = {8: 'hbm_traffic_640x480_8objects.json'}


def traffic_per_dispatch(kernel_class, n_obj):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE / WRITE_SIZE scaled by the
    copy-kernel calibration), or None."""
    name = TRAFFIC_FILES.get(n_obj)
    if HD or not name:
        return None
    try:
        with open(os.path.join(ROOT, 'profiles', name)) as f:
            t = json.load(f)
        return t['kernels'][KERNEL_NAMES[kernel_class]]['bytes_per_dispatch']
    except (OSError, KeyError, ValueError):
        return None


def apply_tuning(g, args):
    for lvl, v in enumerate(x for x in args.track_rpt.split(',') if x):
        g.set_tuning(g.TUNE_TRACK_ROWS_PER_THREAD_L0 + lvl, int(v))
    for lvl, v in enumerate(x for x in args.track_fused.split(',') if x):
        g.set_tuning(g.TUNE_TRACK_FUSED_L0 + lvl, int(v))
    for lvl, v in enumerate(x for x in args.obj_rpt.split(',') if x):
        g.set_tuning(g.TUNE_OBJ_ROWS_PER_THREAD_L0 + lvl, int(v))
    for lvl, v in enumerate(x for x in args.obj_ppt.split(',') if x):
        g.set_tuning(g.TUNE_OBJ_PIXELS_PER_THREAD_L0 + lvl, int(v))


def pingpong(f, n):
    """Data index of step f over n distinct frames played forward and then
    backward (0, 1, ..., n-1, n-2, ..., 1, 0, 1, ...): every step moves the
    camera and the objects by one frame of their smooth trajectories."""
    if n <= 1:
        return 0
    r = f % (2 * (n - 1))
    return r if r < n else 2 * (n - 1) - r


class Leg:
    """One GPU context over resident frames: warm-up, then K timed frames
    between two stream synchronisations."""

    def __init__(self, mf, hiprt, frames, movers, P0, plain, args):
        self.g = mf.Kfusion(W, H, P0, bg_size=5.12, bg_res=BG_RES, batch_objects=not plain)
        apply_tuning(self.g, args)
        for i, m in enumerate(movers):
            self.g.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
            self.g.set_moving(i + 1, True)
        # inputs resident in HBM before timing
        self.dev = [(hiprt.DeviceBuffer(d), hiprt.DeviceBuffer(rgb), [hiprt.DeviceBuffer(m) for m in masks])
                    for d, rgb, masks in frames]
        self.mask_ids = np.arange(len(self.dev[0][2]) if self.dev else 0, dtype=np.int32)
        self.mask_ptrs = [np.array([m.ptr.value for m in dm], np.uint64) for _, _, dm in self.dev]
        self.movers = movers

    def step(self, f):
        g = self.g
        fi = pingpong(f, len(self.dev))
        dd, dr, dm = self.dev[fi]
        # every instance's integration mask of the frame, one copy launch
        g.L.mf_objects_set_masks_device(g.ctx, len(dm), self.mask_ids, self.mask_ptrs[fi])
        g.preprocessing_device(dd.ptr, dr.ptr, W, H)
        g.tracking(K, f)
        g.integration(K, f)
        g.raycasting(K, f)

    def run(self, warmup, steps, autotune=False):
        g = self.g
        tuned = None
        for f in range(warmup):
            self.step(f)
            if autotune and f == 1:
                # both pyramids now hold real frames; the remaining warm-up
                # frames capture the graphs of the chosen shapes
                g.sync()
                chosen, us = g.autotune(K)
                tuned = {'rows_per_thread': [int(v) for v in chosen],
                         'us_per_iteration': {str(c): [round(float(x), 2) for x in us[:, i]]
                                              for i, c in enumerate(g.TUNE_CANDIDATES)}}
                print('autotune: %s' % json.dumps(tuned), file=sys.stderr, flush=True)
        g.sync()
        t = time.perf_counter()
        for f in range(warmup, warmup + steps):
            self.step(f)
        t_enq = time.perf_counter() - t
        g.sync()
        el = time.perf_counter() - t
        return {'fps': steps / el, 'ms': 1000.0 * el / steps, 'enqueue_ms': 1000.0 * t_enq / steps, 'tuned': tuned}

    def close(self):
        self.g.close()


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_leg(frames, movers, P0, n_frames, budget_s):
    """The CPU oracle (the float restatement of the reference path, OpenMP
    over every host thread) timed per stage with CLOCK_MONOTONIC, built on
    this host twice: with the reference's flags (-g -O1 -march=native
    -fopenmp) and with -O3.  Frame 0 only seeds the tracker and is not
    counted; each build runs up to n_frames frames or budget_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    out = {}
    with tempfile.TemporaryDirectory() as td:
        subprocess.check_call(['make', '-s', '-C', os.path.join(ROOT, 'oracle'), 'timing', 'OUT=' + td, '-j2'])
        for build in ('O1', 'O3'):
            o = oracle.Oracle(W, H, P0, lib_path=os.path.join(td, 'liborc32_%s.so' % build))
            o.add_object(5.12, BG_RES, np.eye(4))
            for i, m in enumerate(movers):
                o.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
                o.set_moving(i + 1, True)
            stages = {k: [] for k in ('preprocess', 'track', 'allocate', 'integrate', 'raycast', 'frame')}
            t_start = time.perf_counter()
            nf = 0
            for f in range(min(n_frames, len(frames))):
                d, rgb, masks = frames[f]
                for i, m in enumerate(masks):
                    o.set_mask(i, m)
                t0 = time.perf_counter()
                o.preprocess(d, rgb)
                t1 = time.perf_counter()
                o.track(K, f)
                t2 = time.perf_counter()
                o.stage_clock()
                o.integrate(K, f)
                ta, ti = o.stage_clock()
                t3 = time.perf_counter()
                o.raycast(K, f)
                t4 = time.perf_counter()
                if f > 0:
                    for k, v in (('preprocess', t1 - t0), ('track', t2 - t1), ('allocate', ta), ('integrate', ti),
                                 ('raycast', t4 - t3), ('frame', t4 - t0)):
                        stages[k].append(1000.0 * v)
                    nf += 1
                if time.perf_counter() - t_start > budget_s and nf >= 3:
                    break
            out[build] = {
                'frames_per_s': round(1000.0 / float(np.median(stages['frame'])), 4),
                'ms_median': {k: round(float(np.median(v)), 3) for k, v in stages.items()},
                'ms_p90': {k: round(float(np.percentile(v, 90)), 3) for k, v in stages.items()},
                'frames': nf, 'seconds': round(time.perf_counter() - t_start, 1)}
            threads = o.L.orc_num_threads()
            del o
    return {'value': out['O1']['frames_per_s'], 'unit': 'frames/s', 'cores': threads, 'kind': 'port',
            'sample': ('float oracle (OpenMP C++ restatement of the reference path) on frames 1..%d of the headline '
                       'workload per build, frame 0 as warm-up; value = 1000 / median frame ms of the build at the '
                       "reference's flags -g -O1 -march=native -fopenmp; -O3 build beside it"
                       % out['O1']['frames']),
            'cpu_model': cpu_model(), 'builds': {'-g -O1 -march=native -fopenmp': out['O1'],
                                                 '-O3 -march=native -fopenmp': out['O3']}}


def rot_err(A, B):
    R = A[:3, :3].astype(np.float64).T @ B[:3, :3].astype(np.float64)
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    return float(np.arctan2(s, (np.trace(R) - 1) / 2))


def accuracy_leg(mf, frames, movers, P0, n_frames, plain):
    """Pose and TSDF error of the GPU path against the FP64 oracle (checker
    of the CPU leg).  Each frame starts the GPU context, the FP32 oracle and
    the FP64 oracle from identical model maps and poses (the FP32 oracle's);
    camera and object poses after tracking are compared with the FP64
    oracle's, objects from the FP64 camera pose.  TSDF: after the last frame,
    the fraction of the background's voxels (4000 blocks evenly spaced over
    the FP64 oracle's keys) whose tsdf differs by more than 1e-4 or whose
    weight differs, GPU and FP32 oracle each against the FP64 oracle."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle
    g = mf.Kfusion(W, H, P0, bg_size=5.12, bg_res=BG_RES, batch_objects=not plain)
    o32 = oracle.Oracle(W, H, P0)
    o64 = oracle.Oracle(W, H, P0, fp64=True)
    for o in (o32, o64):
        o.add_object(5.12, BG_RES, np.eye(4))
    for i, m in enumerate(movers):
        T = m.pose(0).astype(np.float32)
        for x in (g, o32, o64):
            x.add_object(m.vol_size, m.vol_res, T)
            x.set_moving(i + 1, True)
    n = len(movers)
    cam, cam32, obj, obj32 = [], [], [], []
    for f in range(n_frames):
        d, rgb, masks = frames[f]
        for x in (g, o32, o64):
            for i, m in enumerate(masks):
                x.set_mask(i, m)
        g.preprocessing(d, rgb)
        o32.preprocess(d, rgb)
        o64.preprocess(d, rgb)
        if f > 0:
            vtx, nrm, lb = o32.maps()
            o64.set_maps(vtx, nrm, lb)
            g.set_maps(vtx, nrm, lb)
            for i in range(1, 1 + n):
                ov, on = o32.object_maps(i)
                o64.set_object_maps(i, ov, on)
                g.set_object_maps(i, ov, on)
            for i in [-1] + list(range(1, 1 + n)):
                T = o32.pose(i).astype(np.float32)
                o64.set_pose(i, T)
                g.setPose(i, T)
        g.track_camera(K, f)
        o32.track_camera(K, f)
        o64.track_camera(K, f)
        P64 = o64.pose()
        if f > 0:
            for P, acc in ((g.getPose(), cam), (o32.pose(), cam32)):
                acc.append((float(np.linalg.norm(P[:3, 3] - P64[:3, 3])), rot_err(P, P64)))
        g.setPose(-1, P64.astype(np.float32))
        o32.set_pose(-1, P64.astype(np.float32))
        g.track_objects(K, f)
        o32.track_objects(K, f)
        o64.track_objects(K, f)
        if f > 0:
            for i in range(1, 1 + n):
                Q64 = o64.pose(i)
                for Q, acc in ((g.getPose(i), obj), (o32.pose(i), obj32)):
                    acc.append((float(np.linalg.norm(Q[:3, 3] - Q64[:3, 3])), rot_err(Q, Q64)))
        g.setPose(-1, o32.pose().astype(np.float32))
        for i in range(1, 1 + n):
            g.setPose(i, o32.pose(i).astype(np.float32))
        for i in range(1 + n):
            g.set_trackresult(i, o32.trackresult(i)[0])
        g.integration(K, f)
        o32.integrate(K, f)
        o64.integrate(K, f)
        g.raycasting(K, f)
        o32.raycast(K, f)
        o64.raycast(K, f)
    L = oracle.lib(False)
    keys = o64.keys(0)
    keys = keys[np.linspace(0, len(keys) - 1, min(len(keys), 4000)).astype(int)]
    vg = g.voxels(0, keys)[0]
    xyz = np.zeros(3, np.uint32)
    bad_g = bad_o = 0
    for kk, key in enumerate(keys):
        L.orc_unpack(int(key) & ~0x1FF, xyz)
        b64 = o64.block(0, *map(int, xyz))
        b32 = o32.block(0, *map(int, xyz))
        for b, which in ((vg[kk], 0), (b32, 1)):
            if b is None:
                dd = 512
            else:
                dd = int(np.sum((np.abs(b[:, 0] - b64[:, 0]) > 1e-4) | (b[:, 1] != b64[:, 1])))
            if which == 0:
                bad_g += dd
            else:
                bad_o += dd
    g.close()
    nv = 512 * max(len(keys), 1)

    def summ(a):
        a = np.asarray(a) if a else np.zeros((1, 2))
        return {'max_m': float(a[:, 0].max()), 'max_rad': float(a[:, 1].max()),
                'median_m': float(np.median(a[:, 0])), 'median_rad': float(np.median(a[:, 1])),
                'within_1e-5': float(np.mean((a[:, 0] < 1e-5) & (a[:, 1] < 1e-5)))}
    return {'reference': 'FP64 oracle, %d tracked frames from identical inputs' % (n_frames - 1),
            'camera_gpu': summ(cam), 'camera_fp32_oracle': summ(cam32),
            'objects_gpu': summ(obj), 'objects_fp32_oracle': summ(obj32),
            'tsdf_voxels_differing_gpu': round(bad_g / nv, 6), 'tsdf_voxels_differing_fp32_oracle': round(bad_o / nv, 6)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=600)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--objects', type=int, default=8)
    ap.add_argument('--frames', type=int, default=16,
                    help='distinct frames rendered for the GPU legs; longer runs play them forward and backward '
                         '(16: the camera drifts to about 7 cm from its ground truth over 600 steps at 8 objects; '
                         'with 30 it can lose track near the turnaround at frames 25-29, and over 120 distinct '
                         'frames it loses track after about 50)')
    ap.add_argument('--plain', action='store_true', help='headline in the one-object-per-launch form')
    ap.add_argument('--no-plain', action='store_true', help='skip the one-object-per-launch leg')
    ap.add_argument('--no-single', action='store_true', help='skip the single-object workload leg')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU oracle leg and the FP64 accuracy check')
    ap.add_argument('--cpu-frames', type=int, default=0,
                    help='frames per CPU build (default 31: about 10-25 s of oracle work per build; 6 with --hd)')
    ap.add_argument('--accuracy-frames', type=int, default=4,
                    help='frames of the FP64 accuracy check (the first only seeds the tracker)')
    ap.add_argument('--render', action='store_true',
                    help='also time the display renderers, volume2mask, the background mesh and the '
                         'instance-mask generalisation (ms per call, results on the host)')
    ap.add_argument('--raw-stream', action='store_true',
                    help='also time the same frames read from a .raw sequence file through pinned staging '
                         '(disk + PCIe inclusive; never the headline value)')
    ap.add_argument('--track-fused', type=str, default='', metavar='F0,F1,F2',
                    help='1 per pyramid level: one fused launch per GN iteration (default: library default)')
    ap.add_argument('--track-rpt', type=str, default='', metavar='R0,R1,R2',
                    help='tracking rows per thread per pyramid level (default: library default)')
    ap.add_argument('--obj-rpt', type=str, default='', metavar='R0,R1,R2',
                    help='moving objects\' rows per thread per pyramid level (0: the camera\'s; default: library default)')
    ap.add_argument('--obj-ppt', type=str, default='', metavar='P0,P1,P2',
                    help='moving objects\' rows per pyramid level: pixels per thread of the shared pixel grid '
                         '(1, 2) or 0 for one row grid per object (default: library default)')
    ap.add_argument('--autotune', action='store_true',
                    help='choose the tracking rows per thread of every pyramid level with the library autotuner '
                         '(mf_autotune) after the second warm-up frame, before timing')
    ap.add_argument('--hd', action='store_true',
                    help='precision + scale workload instead: 1280x720 (fx=fy=1050), background 1024^3 at 5 mm, '
                         '16 moving objects, 2%% invalid depth pixels')
    ap.add_argument('--copy-ref', type=int, default=0, metavar='MB',
                    help='only launch a plain copy kernel of MB megabytes (traffic-counter calibration)')
    args = ap.parse_args()
    global W, H, K, BG_RES, HD
    if args.hd:
        HD = True
        W, H, BG_RES = 1280, 720, 1024
        K = np.array([1050.0, 1050.0, 639.5, 359.5], np.float32)
        if args.objects == 8:
            args.objects = 16
    if args.gpus != 1:
        print('bench.py measures one GPU: this build runs the dense path on a single MI355X '
              '(multi-GPU sharding by instance is a design note only); use --gpus 1', file=sys.stderr)
        sys.exit(2)

    mf = _load('midfusion', os.path.join(PKG, 'midfusion.py'))
    scene = _load('mf_scene', os.path.join(PKG, 'scene.py'))
    hiprt = _load('mf_hiprt', os.path.join(PKG, 'hiprt.py'))
    if args.copy_ref:
        g = mf.Kfusion(64, 48, np.eye(4, dtype=np.float32), bg_res=64)
        g._chk(g.L.mf_copy_reference(g.ctx, args.copy_ref << 20, 5))
        print(json.dumps({'copy_reference_bytes': args.copy_ref << 20, 'launches': 5}))
        return

    n_obj = args.objects
    sc, movers = make_scene(scene, n_obj)
    n_frames = args.warmup + args.steps
    n_distinct = max(2, min(n_frames + 3, args.frames))
    if args.cpu_frames <= 0:
        args.cpu_frames = 6 if HD else 31
    t0 = time.time()
    n_render = max(n_distinct, 0 if args.no_cpu else max(args.cpu_frames, args.accuracy_frames))
    frames = render_frames(sc, n_obj, n_render)
    print('rendered %d frames in %.1fs' % (len(frames), time.time() - t0), file=sys.stderr, flush=True)
    P0 = sc.camera_pose(0).astype(np.float32)

    # ---- headline: batched (or, with --plain, one object per launch)
    gframes = frames[:n_distinct]   # the GPU legs' frames, played forward and backward
    leg = Leg(mf, hiprt, gframes, movers, P0, args.plain, args)
    res = leg.run(args.warmup, args.steps, args.autotune)
    g = leg.g
    print('headline %.1f frames/s' % res['fps'], file=sys.stderr, flush=True)
    pose_gt = float(np.linalg.norm(g.getPose()[:3, 3] - sc.camera_pose(pingpong(n_frames - 1, n_distinct))[:3, 3]))
    if pose_gt >= 0.1:
        print('warning: the camera ended %.3f m from its ground truth: tracking lost' % pose_gt, file=sys.stderr, flush=True)

    # per-kernel timing + work counters on 3 more frames (events on the context stream)
    g.enable_timing(True)
    g.reset_stats()
    g.trace_clear()
    for f in range(n_frames, n_frames + 3):
        leg.step(f)
    g.sync()
    tm = g.timing()
    st = g.stats()
    tr = g.trace()
    tot_ms = sum(v[0] for v in tm.values())
    stage = {k: round(v[0] / 3, 4) for k, v in tm.items() if v[1]}
    N0 = W * H
    rpt = [g.get_tuning(g.TUNE_TRACK_ROWS_PER_THREAD_L0 + lvl) for lvl in range(4)]
    # the library's automatic object shape (knob 0): {8, 8, 4, 4} when two or
    # more objects share a launch, else the camera's
    batched = not args.plain and n_obj >= 2
    orpt = [g.get_tuning(g.TUNE_OBJ_ROWS_PER_THREAD_L0 + lvl) or ((8, 8, 4, 4)[lvl] if batched else rpt[lvl])
            for lvl in range(4)]
    alg = {
        'track_rows': track_bytes(tr, rpt, False),
        'track_rows_obj': track_bytes(tr, orpt, True),
        'track_solve': solve_bytes(tr, rpt),
        'integrate': 20 * N0 * (1 + n_obj) + 72 * st['voxel_updates'] + 16 * st['blocks_integrated'],
        'raycast': 12 * st['march_steps'] + 32 * st['trilinear_evals'] + 128 * st['hits'] + 28 * N0 * 3,
        'alloc_band': 8 * st['alloc_pixels'] + 8 * st['band_samples'],
    }
    kern = {k: v for k, v in tm.items() if v[1] and k in alg}
    # Dominant kernel: the camera's level-0 tracking rows (most HBM bytes and
    # the most launches of the frame).  Per-launch event pairs inflate short
    # kernels, so its duration is measured over a back-to-back burst of the
    # same launch, bracketed by two events on the context stream.
    g.enable_timing(False)
    rep_us, rep_bytes = g.replay_track_rows(K, 0, REPLAYS)
    achieved = rep_bytes / (rep_us * 1e-6) / 1e9
    roof = {'kernel': KERNEL_NAMES['track_rows'] + ' (level 0, %d workgroups)' % (-(-(2 * N0) // (TB * rpt[0]))),
            'bound': 'hbm', 'achieved': round(achieved, 2), 'peak': HBM_PEAK_GBS,
            'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 5),
            'traffic': traffic_per_dispatch('track_rows_l0', n_obj),
            'avg_launch_us': round(rep_us, 3), 'launches': REPLAYS, 'alg_bytes_per_launch': int(rep_bytes)}
    roof_all = {}
    for k, (ms_k, n_k) in kern.items():
        a = alg[k] / (ms_k / 1000.0) / 1e9
        roof_all[k] = {'achieved_gbs': round(a, 1), 'frac': round(a / HBM_PEAK_GBS, 4),
                       'avg_launch_us': round(1000.0 * ms_k / n_k, 2), 'launches': n_k}
    # allocated blocks of every volume against its pool capacity
    g.sync()
    capacity = []
    for i in range(1 + n_obj):
        nb = len(g.keys(i))
        res_i = BG_RES if i == 0 else movers[i - 1].vol_res
        cap = min((res_i // 8) ** 3, 262144)
        capacity.append({'id': i, 'blocks': nb, 'capacity': cap})
    leg.close()

    # ---- the one-object-per-launch form of the same workload
    plain = None
    if not args.plain and not args.no_plain and n_obj > 1:
        lp = Leg(mf, hiprt, gframes, movers, P0, True, args)
        plain = lp.run(args.warmup, args.steps)
        lp.close()
        print('plain %.1f frames/s' % plain['fps'], file=sys.stderr, flush=True)

    # ---- the single-object workload
    single = None
    if not args.no_single and not HD and n_obj != 1:
        sc1, mov1 = make_scene(scene, 1)
        fr1 = render_frames(sc1, 1, n_distinct)
        l1 = Leg(mf, hiprt, fr1, mov1, sc1.camera_pose(0).astype(np.float32), args.plain, args)
        single = l1.run(args.warmup, args.steps)
        l1.close()
        print('single object %.1f frames/s' % single['fps'], file=sys.stderr, flush=True)

    cpu = accuracy = None
    if not args.no_cpu:
        cpu = cpu_leg(frames, movers, P0, args.cpu_frames, 30.0 if HD else 25.0)
        if args.accuracy_frames > 1:
            accuracy = accuracy_leg(mf, frames, movers, P0, args.accuracy_frames, args.plain)

    render = None
    if args.render:
        g = mf.Kfusion(W, H, P0, bg_size=5.12, bg_res=BG_RES, batch_objects=not args.plain)
        for i, m in enumerate(movers):
            g.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
            g.set_moving(i + 1, True)
        for f in range(3):
            d, rgb, masks = frames[f]
            for i, m in enumerate(masks):
                g.set_mask(i, m)
            g.preprocessing(d, rgb)
            g.tracking(K, f)
            g.integration(K, f)
            g.raycasting(K, f)
        P = g.getPose()
        Pv = P.copy()
        Pv[:3, 3] += np.array([0.05, -0.03, 0.02], np.float32)
        calls = {'render_volume_maps': lambda: g.render_volume(K, P, largestep=0.075),
                 'render_volume_raycast': lambda: g.render_volume(K, Pv, largestep=0.075),
                 'volume2mask': lambda: g.volume2mask(K),
                 'render_track_rgbd': lambda: g.render_track(0, 2),
                 'render_depth': lambda: g.render_depth(),
                 'object_mesh_bg': lambda: g.mesh(0),
                 'instance_mask': lambda: g.set_instance_mask(1, (frames[0][2][1] > 0).astype(np.uint8) * 255)}
        render = {}
        for name, fn in calls.items():
            fn()
            t = time.perf_counter()
            for _ in range(10):
                fn()
            render[name] = round(1000.0 * (time.perf_counter() - t) / 10, 3)
        g.close()

    raw_leg = None
    if args.raw_stream:
        import struct
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, 'scene.raw')
            with open(path, 'wb') as fh:   # the SLAMBench frame layout (interface.h:254-256)
                for d, rgb, _ in gframes:
                    fh.write(struct.pack('<II', W, H) + np.ascontiguousarray(d, '<u2').tobytes())
                    fh.write(struct.pack('<II', W, H) + np.ascontiguousarray(rgb, np.uint8).tobytes())
            raw = mf.RawSequence(path)
            lr = Leg(mf, hiprt, gframes, movers, P0, args.plain, args)

            def step_raw(f):
                f = pingpong(f, len(lr.dev))
                for i, m in enumerate(lr.dev[f][2]):
                    lr.g.L.mf_object_set_mask_device(lr.g.ctx, i, m.ptr)
                lr.g.preprocessing_raw(raw, f)
                lr.g.tracking(K, f)
                lr.g.integration(K, f)
                lr.g.raycasting(K, f)
            for f in range(args.warmup):
                step_raw(f)
            lr.g.sync()
            t = time.perf_counter()
            for f in range(args.warmup, n_frames):
                step_raw(f)
            lr.g.sync()
            el2 = time.perf_counter() - t
            raw_leg = {'value': round(args.steps / el2, 3), 'unit': 'frames/s',
                       'path': '.raw file (page cache) -> pinned double-buffered staging -> PCIe -> preprocess'}
            raw.close()
            lr.close()

    out = {
        'metric': METRIC,
        'value': round(res['fps'], 3), 'unit': 'frames/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(res['ms'], 3), 'higher_is_better': True, 'scaling': 'weak',
        # no published number exists for this metric: the
        # comparison with the plain GPU path is vs_plain_gpu
        'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': workload_name(n_obj, args.plain), 'objects': n_obj, 'frames': args.steps,
                   'distinct_frames': n_distinct, 'frame_order': 'distinct frames played forward and backward'},
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    if plain:
        out['plain_gpu'] = {'value': round(plain['fps'], 3), 'unit': 'frames/s', 'ms_per_step': round(plain['ms'], 3),
                            'workload': workload_name(n_obj, True)}
        out['vs_plain_gpu'] = round(res['fps'] / plain['fps'], 4)
    if accuracy:
        out['pose_err_vs_fp64'] = accuracy
    if single:
        out['single_object'] = {'value': round(single['fps'], 3), 'unit': 'frames/s',
                                'ms_per_step': round(single['ms'], 3), 'workload': workload_name(1, args.plain)}
    out.update({
        'event_timed_by_kernel': roof_all,
        'stage_ms_per_frame': stage,
        'kernel_ms_per_frame_total': round(tot_ms / 3, 3),
        'host_enqueue_ms_per_frame': round(res['enqueue_ms'], 3),
        'work_per_frame': {k: v // 3 for k, v in st.items()},
        'camera_translation_error_vs_gt_m': round(pose_gt, 5),
        # a camera more than 10 cm off its ground truth has lost track: the
        # frames that follow are not the workload (rows end early)
        'camera_tracking_ok': bool(pose_gt < 0.1),
        'tracking_rows_per_thread': rpt[:3],
        'allocated_blocks': capacity,
    })
    if res['tuned']:
        out['autotune'] = res['tuned']
    if raw_leg:
        out['host_input_rate'] = raw_leg
    if render:
        out['render_ms_per_call'] = render
    print(json.dumps(out))


if __name__ == '__main__':
    main()
