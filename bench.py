"""Frames/s of MID-Fusion's per-frame dense path on one MI355X.

A step is one frame: preprocessing (inputs already in HBM), camera + moving
object tracking (ICP + photometric, 3 levels {10,5,4}), allocation +
integration of every instance, and raycasting.  Workload: 640x480
(fx=fy=525, cx=319.5, cy=239.5), background volume 512^3 at 1 cm, one
moving textured box 0.2x0.15x0.1 m with its own 128^3 volume at 2 mm,
seeded synthetic frames.  Prints one JSON line.
"""
import argparse
import importlib.util
import json
import os
import sys
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


def render_frames(sc, n_objects, n):
    scene = sys.modules['mf_scene']
    frames = []
    for f in range(n):
        d, rgb, lab = sc.render(f)
        frames.append((d, rgb, scene.masks_for_labels(lab, n_objects)))
    return frames


TB = 256   # threads per tracking workgroup
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
KERNEL_NAMES = {'track_rows': 'mf::k_track<true, true, false, 3>', 'track_rows_l0': 'mf::k_track<true, true, false, 3> level 0',
                'track_rows_obj': 'mf::k_track<true, true, true, 3>',
                'track_solve': 'mf::k_solve', 'integrate': 'mf::k_integrate', 'raycast': 'mf::k_raycast',
                'alloc_band': 'mf::k_alloc'}


def traffic_per_dispatch(kernel_class):
    """HBM bytes per dispatch of a kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, 'profiles', 'hbm_traffic_cfg2.json')
    try:
        with open(path) as f:
            t = json.load(f)
        return t['kernels'][KERNEL_NAMES[kernel_class]]['bytes_per_dispatch']
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=30)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--objects', type=int, default=1)
    ap.add_argument('--plain', action='store_true', help='one object per launch instead of batched launches')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU oracle baseline leg')
    ap.add_argument('--render', action='store_true',
                    help='also time the display renderers, volume2mask, the background mesh and the '
                         'instance-mask generalisation (ms per call, results on the host)')
    ap.add_argument('--raw-stream', action='store_true',
                    help='also time the same frames read from a .raw sequence file through pinned staging '
                         '(disk + PCIe inclusive; never the headline value)')
    ap.add_argument('--cpu-frames', type=int, default=0,
                    help='frames of the CPU baseline sample (default: 120, about 10 s of oracle work on 16 threads; '
                         '12 with --hd)')
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
                         '16 moving objects (default --objects), 2%% invalid depth pixels')
    ap.add_argument('--copy-ref', type=int, default=0, metavar='MB',
                    help='only launch a plain copy kernel of MB megabytes (traffic-counter calibration)')
    args = ap.parse_args()
    global W, H, K, BG_RES, HD
    if args.hd:
        HD = True
        W, H, BG_RES = 1280, 720, 1024
        K = np.array([1050.0, 1050.0, 639.5, 359.5], np.float32)
        if args.objects == 1:
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
    t0 = time.time()
    if args.cpu_frames <= 0:
        args.cpu_frames = 12 if HD else 120
    n_render = max(n_frames + 3, 0 if args.no_cpu else args.cpu_frames)
    frames = render_frames(sc, n_obj, n_render)
    print('rendered %d frames in %.1fs' % (len(frames), time.time() - t0), file=sys.stderr, flush=True)

    P0 = sc.camera_pose(0).astype(np.float32)
    g = mf.Kfusion(W, H, P0, bg_size=5.12, bg_res=BG_RES, batch_objects=not args.plain)
    for lvl, v in enumerate(x for x in args.track_rpt.split(',') if x):
        g.set_tuning(g.TUNE_TRACK_ROWS_PER_THREAD_L0 + lvl, int(v))
    for lvl, v in enumerate(x for x in args.track_fused.split(',') if x):
        g.set_tuning(g.TUNE_TRACK_FUSED_L0 + lvl, int(v))
    for lvl, v in enumerate(x for x in args.obj_rpt.split(',') if x):
        g.set_tuning(g.TUNE_OBJ_ROWS_PER_THREAD_L0 + lvl, int(v))
    for lvl, v in enumerate(x for x in args.obj_ppt.split(',') if x):
        g.set_tuning(g.TUNE_OBJ_PIXELS_PER_THREAD_L0 + lvl, int(v))
    for i, m in enumerate(movers):
        g.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
        g.set_moving(i + 1, True)
    # inputs resident in HBM before timing
    dev = []
    for d, rgb, masks in frames:
        dev.append((hiprt.DeviceBuffer(d), hiprt.DeviceBuffer(rgb), [hiprt.DeviceBuffer(m) for m in masks]))

    def step(f):
        dd, dr, dm = dev[f]
        for i, m in enumerate(dm):
            g.L.mf_object_set_mask_device(g.ctx, i, m.ptr)
        g.preprocessing_device(dd.ptr, dr.ptr, W, H)
        g.tracking(K, f)
        g.integration(K, f)
        g.raycasting(K, f)

    tuned = None
    for f in range(args.warmup):
        step(f)
        if args.autotune and f == 1:
            # both pyramids now hold real frames; the remaining warm-up frames
            # capture the graphs of the chosen shapes
            g.sync()
            chosen, us = g.autotune(K)
            tuned = {'rows_per_thread': [int(v) for v in chosen],
                     'us_per_iteration': {str(c): [round(float(x), 2) for x in us[:, i]]
                                          for i, c in enumerate(g.TUNE_CANDIDATES)}}
            print('autotune: %s' % json.dumps(tuned), file=sys.stderr, flush=True)
    g.sync()
    t = time.perf_counter()
    for f in range(args.warmup, n_frames):
        step(f)
    t_enq = time.perf_counter() - t
    g.sync()
    el = time.perf_counter() - t
    fps = args.steps / el
    ms = 1000.0 * el / args.steps
    pose_err = float(np.linalg.norm(g.getPose()[:3, 3] - sc.camera_pose(n_frames - 1)[:3, 3]))

    # per-kernel timing + work counters on 3 more frames (events on the context stream)
    g.enable_timing(True)
    g.reset_stats()
    g.trace_clear()
    for f in range(n_frames, n_frames + 3):
        step(f)
    g.sync()
    tm = g.timing()
    st = g.stats()
    tr = g.trace()
    tot_ms = sum(v[0] for v in tm.values())
    stage = {k: round(v[0] / 3, 4) for k, v in tm.items() if v[1]}
    # algorithmic bytes per kernel class over the 3 frames
    N0 = W * H
    rpt = [g.get_tuning(g.TUNE_TRACK_ROWS_PER_THREAD_L0 + lvl) for lvl in range(4)]
    orpt = [g.get_tuning(g.TUNE_OBJ_ROWS_PER_THREAD_L0 + lvl) or rpt[lvl] for lvl in range(4)]
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
            # the committed counter summary covers the 640x480 workload only
            'traffic': None if HD else traffic_per_dispatch('track_rows_l0'),
            'avg_launch_us': round(rep_us, 3), 'launches': REPLAYS, 'alg_bytes_per_launch': int(rep_bytes)}
    roof_all = {}
    for k, (ms_k, n_k) in kern.items():
        a = alg[k] / (ms_k / 1000.0) / 1e9
        roof_all[k] = {'achieved_gbs': round(a, 1), 'frac': round(a / HBM_PEAK_GBS, 4),
                       'avg_launch_us': round(1000.0 * ms_k / n_k, 2), 'launches': n_k}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle
        o = oracle.Oracle(W, H, P0)
        o.add_object(5.12, BG_RES, np.eye(4))
        for i, m in enumerate(movers):
            o.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
            o.set_moving(i + 1, True)
        nthreads = o.L.orc_num_threads()
        nc = max(2, args.cpu_frames)
        times = []
        for f in range(nc):
            d, rgb, masks = frames[f]
            t1 = time.perf_counter()
            for i, m in enumerate(masks):
                o.set_mask(i, m)
            o.preprocess(d, rgb)
            o.track(K, f)
            o.integrate(K, f)
            o.raycast(K, f)
            times.append(time.perf_counter() - t1)
        cpu_fps = 1.0 / float(np.mean(times[1:]))   # frame 0 only seeds the tracker
        cpu = {'value': round(cpu_fps, 4), 'unit': 'frames/s', 'cores': nthreads, 'kind': 'port',
               'sample': 'oracle (OpenMP C++ restatement, -O2) on frames 1..%d of the same workload (%.1f s of CPU '
                         'work), frame 0 as warm-up' % (nc - 1, float(np.sum(times[1:])))}

    render = None
    if args.render:
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

    raw_leg = None
    if args.raw_stream:
        import struct
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, 'scene.raw')
            with open(path, 'wb') as fh:   # the SLAMBench frame layout (interface.h:254-256)
                for d, rgb, _ in frames:
                    fh.write(struct.pack('<II', W, H) + np.ascontiguousarray(d, '<u2').tobytes())
                    fh.write(struct.pack('<II', W, H) + np.ascontiguousarray(rgb, np.uint8).tobytes())
            raw = mf.RawSequence(path)
            g2 = mf.Kfusion(W, H, P0, bg_size=5.12, bg_res=BG_RES, batch_objects=not args.plain)
            for i, m in enumerate(movers):
                g2.add_object(m.vol_size, m.vol_res, m.pose(0).astype(np.float32))
                g2.set_moving(i + 1, True)

            def step_raw(f):
                for i, m in enumerate(dev[f][2]):
                    g2.L.mf_object_set_mask_device(g2.ctx, i, m.ptr)
                g2.preprocessing_raw(raw, f)
                g2.tracking(K, f)
                g2.integration(K, f)
                g2.raycasting(K, f)
            for f in range(args.warmup):
                step_raw(f)
            g2.sync()
            t = time.perf_counter()
            for f in range(args.warmup, n_frames):
                step_raw(f)
            g2.sync()
            el2 = time.perf_counter() - t
            raw_leg = {'value': round(args.steps / el2, 3), 'unit': 'frames/s',
                       'path': '.raw file (page cache) -> pinned double-buffered staging -> PCIe -> preprocess'}
            raw.close()
            g2.close()

    out = {
        'metric': 'frames/s (track+integrate+raycast, %dx%d, cam+K objs)' % (W, H),
        'value': round(fps, 3), 'unit': 'frames/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(ms, 3), 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': ('1280x720 RGB-D (fx=fy=1050, 2%% invalid depth), bg 1024^3 @5mm + %d moving '
                                'boxes/ellipsoids 0.05-0.4 m in 64^3-256^3 @2-5mm' % n_obj if HD else
                                '640x480 RGB-D, bg 512^3 @1cm + 1 moving box 0.2x0.15x0.1 m in 128^3 @2mm'
                                if n_obj == 1 else
                                '640x480 RGB-D, bg 512^3 @1cm + %d moving boxes/ellipsoids 0.05-0.4 m in 64^3-256^3 '
                                '@2-5mm' % n_obj) + ', ICP+RGB {10,5,4}, %s' % (
                                    'one object per launch' if args.plain else 'objects batched per launch'),
                   'objects': n_obj, 'frames': args.steps},
        'roofline': roof,
        'event_timed_by_kernel': roof_all,
        'cpu_baseline': cpu,
        'stage_ms_per_frame': stage,
        'kernel_ms_per_frame_total': round(tot_ms / 3, 3),
        'host_enqueue_ms_per_frame': round(1000.0 * t_enq / args.steps, 3),
        'work_per_frame': {k: v // 3 for k, v in st.items()},
        'camera_translation_error_vs_gt_m': round(pose_err, 5),
    }
    out['tracking_rows_per_thread'] = rpt[:3]
    if tuned:
        out['autotune'] = tuned
    if raw_leg:
        out['host_input_rate'] = raw_leg
    if render:
        out['render_ms_per_call'] = render
    print(json.dumps(out))


if __name__ == '__main__':
    main()
