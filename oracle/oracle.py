"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker.  ``Oracle(fp64=True)`` loads the double build.
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

ROBUST = {'none': 0, 'huber': 1, 'tukey': 2, 'cauchy': 3}
ROBUST_K = {'none': 1.0, 'huber': 1.345, 'tukey': 4.6851, 'cauchy': 2.3849}


class OrcConfig(C.Structure):
    _fields_ = [('w', C.c_int), ('h', C.c_int), ('nlev', C.c_int), ('iters', C.c_int * 4),
                ('use_icp', C.c_int), ('use_rgb', C.c_int), ('robust', C.c_int), ('robust_k', C.c_double),
                ('mu', C.c_float), ('track_rate', C.c_int), ('integ_rate', C.c_int), ('init_frame', C.c_int),
                ('alloc_quirks', C.c_int)]


def build():
    import subprocess
    subprocess.check_call(['make', '-s', '-C', _HERE])


def _lib(fp64, path=None):
    if path is None:
        path = os.path.join(_HERE, 'liborc64.so' if fp64 else 'liborc32.so')
        if not os.path.exists(path):
            build()
    lib = C.CDLL(path)
    P = C.c_void_p
    f32p = np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS')
    f64p = np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS')
    u64p = np.ctypeslib.ndpointer(np.uint64, flags='C_CONTIGUOUS')
    i32p = np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS')
    u16p = np.ctypeslib.ndpointer(np.uint16, flags='C_CONTIGUOUS')
    u8p = np.ctypeslib.ndpointer(np.uint8, flags='C_CONTIGUOUS')
    sig = {
        'orc_create': (P, [C.POINTER(OrcConfig), f32p]),
        'orc_destroy': (None, [P]),
        'orc_object_create': (C.c_int, [P, C.c_float, C.c_int, f32p]),
        'orc_object_set_moving': (C.c_int, [P, C.c_int, C.c_int]),
        'orc_object_set_mask': (C.c_int, [P, C.c_int, f32p]),
        'orc_preprocess': (C.c_int, [P, u16p, u8p, C.c_int, C.c_int, C.c_int]),
        'orc_track': (C.c_int, [P, f32p, C.c_uint]),
        'orc_track_camera': (C.c_int, [P, f32p, C.c_uint]),
        'orc_track_objects': (C.c_int, [P, f32p, C.c_uint]),
        'orc_integrate': (C.c_int, [P, f32p, C.c_uint]),
        'orc_raycast': (C.c_int, [P, f32p, C.c_uint]),
        'orc_get_pose': (None, [P, C.c_int, f64p]),
        'orc_set_pose': (None, [P, C.c_int, f32p]),
        'orc_eval': (C.c_int, [P, C.c_int, C.c_int, f64p, f64p]),
        'orc_set_motion_auto': (None, [P, C.c_int]),
        'orc_motion_result': (C.c_int, [P, np.ctypeslib.ndpointer(np.int64, flags='C_CONTIGUOUS'),
                                        np.ctypeslib.ndpointer(np.int64, flags='C_CONTIGUOUS'), i32p, C.c_int]),
        'orc_motion_rows': (C.c_int, [P, i32p, f32p]),
        'orc_motion_state': (C.c_int, [P, i32p, i32p, i32p, C.c_int]),
        'orc_set_gt_poses': (None, [P, f32p, C.c_long]),
        'orc_object_mesh': (C.c_long, [P, C.c_int, P, C.c_long]),
        'orc_octree_mesh': (C.c_long, [P, P, C.c_long]),
        'orc_octree_set_full': (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_double, C.c_int]),
        'orc_render_track': (C.c_int, [P, C.c_int, C.c_int, u8p]),
        'orc_render_depth': (C.c_int, [P, u8p]),
        'orc_render_intensity': (C.c_int, [P, u8p]),
        'orc_volume2mask': (C.c_int, [P, f32p, i32p]),
        'orc_render_volume': (C.c_int, [P, f32p, f32p, C.c_int, C.c_int, C.c_float, C.c_int, u8p, C.c_int, u8p]),
        'orc_motion_residuals': (C.c_int, [P, f32p, f32p, f32p, f32p, f32p, i32p]),
        'orc_get_maps_bf': (None, [P, f32p, f32p, i32p]),
        'orc_object_vertex_bef': (None, [P, C.c_int, f32p]),
        'orc_virtual_cam': (None, [P, C.c_int, f64p]),
        'orc_get_maps': (None, [P, f32p, f32p, i32p]),
        'orc_get_object_maps': (None, [P, C.c_int, f32p, f32p]),
        'orc_set_maps': (None, [P, f32p, f32p, i32p]),
        'orc_set_object_maps': (None, [P, C.c_int, f32p, f32p]),
        'orc_object_keys': (C.c_long, [P, C.c_int, u64p, C.c_long]),
        'orc_object_node_count': (C.c_long, [P, C.c_int]),
        'orc_object_block': (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, f64p]),
        'orc_object_block_active': (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
        'orc_last_emitted': (C.c_long, [P, C.c_int]),
        'orc_last_keys': (C.c_long, [P, C.c_int, u64p, C.c_long]),
        'orc_object_trackresult': (None, [P, C.c_int, i32p, f32p]),
        'orc_trace': (C.c_long, [P, f64p, C.c_long]),
        'orc_get_rows': (None, [P, C.c_int, i32p, f32p, f32p]),
        'orc_trace_clear': (None, [P]),
        'orc_tracker_level': (C.c_long, [P, C.c_int, C.c_int, f32p]),
        'orc_get_frame': (None, [P, f32p, f32p, f32p]),
        'orc_bilateral': (C.c_int, [f32p, f32p, C.c_int, C.c_int]),
        'orc_half_sample': (C.c_int, [f32p, f32p, C.c_int, C.c_int]),
        'orc_bilinear': (C.c_float, [f32p, C.c_int, C.c_float, C.c_float]),
        'orc_encode': (C.c_uint64, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        'orc_morton': (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32]),
        'orc_unpack': (None, [C.c_uint64, np.ctypeslib.ndpointer(np.uint32, flags='C_CONTIGUOUS')]),
        'orc_unique_ms': (C.c_int, [u64p, C.c_int, C.c_int]),
        'orc_solve': (C.c_int, [f64p, f64p, f32p]),
        'orc_se3_exp': (None, [f64p, f64p]),
        'orc_inverse': (None, [f32p, f32p]),
        'orc_reduce_rows': (None, [i32p, f32p, f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double, f64p]),
        'orc_robust_weight': (C.c_float, [C.c_float, C.c_int, C.c_double]),
        'orc_robust_loss': (C.c_float, [C.c_float, C.c_int, C.c_double]),
        'orc_octree_create': (P, [C.c_int, C.c_float]),
        'orc_octree_destroy': (None, [P]),
        'orc_octree_alloc': (None, [P, u64p, C.c_int]),
        'orc_octree_keys': (C.c_long, [P, u64p, C.c_long]),
        'orc_octree_node_keys': (C.c_long, [P, u64p, C.c_long]),
        'orc_octree_set': (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]),
        'orc_octree_interp': (C.c_float, [P, C.c_float, C.c_float, C.c_float, C.c_int]),
        'orc_octree_grad': (None, [P, C.c_float, C.c_float, C.c_float, f32p]),
        'orc_octree_ray_tmin': (C.c_float, [P, f32p, f32p, C.c_float, C.c_float, C.c_float]),
        'orc_octree_raycast': (C.c_float, [P, f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                           f32p]),
        'orc_num_threads': (C.c_int, []),
        'orc_stage_clock': (None, [P, f64p]),
        'orc_real_bytes': (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIBS = {}


def lib(fp64=False, path=None):
    """The float (or double) oracle, or the float build at `path` (a timing
    build made by `make timing`)."""
    key = path or fp64
    if key not in _LIBS:
        _LIBS[key] = _lib(fp64, path)
    return _LIBS[key]


def f32(a):
    return np.ascontiguousarray(a, np.float32)


class Oracle:
    """Kfusion-style per-frame pipeline on the CPU oracle."""

    def __init__(self, w, h, init_pose, iters=(10, 5, 4), use_icp=True, use_rgb=True, robust='cauchy',
                 robust_k=None, mu=0.1, track_rate=1, integ_rate=1, init_frame=0, alloc_quirks=False, fp64=False,
                 lib_path=None):
        self.L = lib(fp64, lib_path)
        self.w, self.h = w, h
        cfg = OrcConfig()
        cfg.w, cfg.h, cfg.nlev = w, h, len(iters)
        for i, v in enumerate(iters):
            cfg.iters[i] = v
        cfg.use_icp, cfg.use_rgb = int(use_icp), int(use_rgb)
        cfg.robust = ROBUST[robust]
        cfg.robust_k = ROBUST_K[robust] if robust_k is None else robust_k
        cfg.mu, cfg.track_rate, cfg.integ_rate, cfg.init_frame = mu, track_rate, integ_rate, init_frame
        cfg.alloc_quirks = int(alloc_quirks)
        self.h_ = self.L.orc_create(C.byref(cfg), f32(np.asarray(init_pose).reshape(16)))
        self.n_obj = 0

    def __del__(self):
        if getattr(self, 'h_', None):
            self.L.orc_destroy(self.h_)
            self.h_ = None

    def add_object(self, size, res, pose):
        self.n_obj += 1
        return self.L.orc_object_create(self.h_, size, res, f32(np.asarray(pose).reshape(16)))

    def set_moving(self, oid, moving):
        self.L.orc_object_set_moving(self.h_, oid, int(moving))

    def set_mask(self, oid, mask):
        self.L.orc_object_set_mask(self.h_, oid, f32(mask).reshape(-1))

    def preprocess(self, depth, rgb, filt=False):
        d = np.ascontiguousarray(depth, np.uint16)
        r = np.ascontiguousarray(rgb, np.uint8)
        return self.L.orc_preprocess(self.h_, d.reshape(-1), r.reshape(-1), d.shape[1], d.shape[0], int(filt))

    def track(self, k, frame):
        return self.L.orc_track(self.h_, f32(k), frame)

    def track_camera(self, k, frame):
        return self.L.orc_track_camera(self.h_, f32(k), frame)

    def track_objects(self, k, frame):
        return self.L.orc_track_objects(self.h_, f32(k), frame)

    def integrate(self, k, frame):
        return self.L.orc_integrate(self.h_, f32(k), frame)

    def raycast(self, k, frame):
        return self.L.orc_raycast(self.h_, f32(k), frame)

    def stage_clock(self):
        """(allocation s, voxel update s) since the last call."""
        out = np.zeros(2)
        self.L.orc_stage_clock(self.h_, out)
        return float(out[0]), float(out[1])

    def pose(self, oid=-1):
        T = np.zeros(16)
        self.L.orc_get_pose(self.h_, oid, T)
        return T.reshape(4, 4)

    def eval(self, oid, level, T):
        """32 FP64 reduction sums of the residual of the camera (oid < 0, T =
        T_w_l) or object oid (T = T_c2_o2) at `level`; call between
        track_camera and track_objects."""
        out = np.zeros(32)
        r = self.L.orc_eval(self.h_, oid, level, np.ascontiguousarray(np.asarray(T, np.float64).reshape(16)), out)
        assert r == 0
        return out

    def set_motion_auto(self, on):
        self.L.orc_set_motion_auto(self.h_, int(on))

    def motion_result(self):
        cap = 64
        px, inl, mv = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(cap, np.int32)
        n = self.L.orc_motion_result(self.h_, px, inl, mv, cap)
        return px[:n], inl[:n], mv[:n]

    def render_track(self, oid, mode):
        out = np.zeros((self.h, self.w, 4), np.uint8)
        self.L.orc_render_track(self.h_, oid, mode, out.reshape(-1))
        return out

    def render_depth(self):
        out = np.zeros((self.h, self.w, 4), np.uint8)
        self.L.orc_render_depth(self.h_, out.reshape(-1))
        return out

    def render_intensity(self):
        out = np.zeros((self.h, self.w, 4), np.uint8)
        self.L.orc_render_intensity(self.h_, out.reshape(-1))
        return out

    def volume2mask(self, k):
        out = np.zeros((self.h, self.w), np.int32)
        self.L.orc_volume2mask(self.h_, f32(k), out.reshape(-1))
        return out

    def render_volume(self, k, view_pose, w=None, h=None, largestep=0.075, render_color=True, palette=None):
        w, h = w or self.w, h or self.h
        out = np.zeros((h, w, 4), np.uint8)
        pal = np.zeros((1, 4), np.uint8) if palette is None else np.ascontiguousarray(palette, np.uint8)
        r = self.L.orc_render_volume(self.h_, f32(k), f32(np.asarray(view_pose).reshape(16)), w, h, largestep,
                                     int(render_color), pal.reshape(-1), 0 if palette is None else len(pal),
                                     out.reshape(-1))
        return out, bool(r)

    def mesh(self, oid):
        n = self.L.orc_object_mesh(self.h_, oid, None, 0)
        out = np.zeros((max(n, 1), 3, 3), np.float32)
        if n:
            self.L.orc_object_mesh(self.h_, oid, out.ctypes.data, n)
        return out[:n]

    def set_gt_poses(self, poses):
        """Ground-truth camera poses (n x 4 x 4); an empty array turns the mode off."""
        p = f32(np.asarray(poses, np.float32).reshape(-1))
        self.L.orc_set_gt_poses(self.h_, p if p.size else np.zeros(16, np.float32), len(p) // 16)

    def motion_state(self):
        """(gate, rendered_n, in_view flags, moving set) of the last tracking call."""
        cap = 64
        g, iv, ms = np.zeros(2, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        n = self.L.orc_motion_state(self.h_, g, iv, ms, cap)
        return int(g[0]), int(g[1]), iv[:n], ms[:n]

    def motion_residuals(self, k, T, vtx, nrm, vbf):
        """compute_residuals codes (2*W*H) at pose T against the given maps."""
        out = np.zeros(2 * self.w * self.h, np.int32)
        self.L.orc_motion_residuals(self.h_, f32(k), f32(np.asarray(T).reshape(16)), f32(vtx).reshape(-1),
                                    f32(nrm).reshape(-1), f32(vbf).reshape(-1), out)
        return out

    def maps_bf(self):
        n = self.w * self.h
        v, nn, lab = np.zeros(3 * n, np.float32), np.zeros(3 * n, np.float32), np.zeros(n, np.int32)
        self.L.orc_get_maps_bf(self.h_, v, nn, lab)
        return v.reshape(self.h, self.w, 3), nn.reshape(self.h, self.w, 3), lab.reshape(self.h, self.w)

    def object_vertex_bef(self, oid):
        v = np.zeros(3 * self.w * self.h, np.float32)
        self.L.orc_object_vertex_bef(self.h_, oid, v)
        return v.reshape(self.h, self.w, 3)

    def virtual_cam(self, oid):
        T = np.zeros(16)
        self.L.orc_virtual_cam(self.h_, oid, T)
        return T.reshape(4, 4)

    def motion_rows(self):
        n = 2 * self.w * self.h
        r, e = np.zeros(n, np.int32), np.zeros(n, np.float32)
        self.L.orc_motion_rows(self.h_, r, e)
        return r, e

    def set_pose(self, oid, T):
        self.L.orc_set_pose(self.h_, oid, f32(np.asarray(T).reshape(16)))

    def maps(self):
        n = self.w * self.h
        v, nn, lab = np.zeros(3 * n, np.float32), np.zeros(3 * n, np.float32), np.zeros(n, np.int32)
        self.L.orc_get_maps(self.h_, v, nn, lab)
        return v.reshape(self.h, self.w, 3), nn.reshape(self.h, self.w, 3), lab.reshape(self.h, self.w)

    def set_maps(self, vtx, nrm, lab):
        self.L.orc_set_maps(self.h_, f32(vtx).reshape(-1), f32(nrm).reshape(-1),
                            np.ascontiguousarray(lab, np.int32).reshape(-1))

    def set_object_maps(self, oid, vtx, nrm):
        self.L.orc_set_object_maps(self.h_, oid, f32(vtx).reshape(-1), f32(nrm).reshape(-1))

    def object_maps(self, oid):
        n = self.w * self.h
        v, nn = np.zeros(3 * n, np.float32), np.zeros(3 * n, np.float32)
        self.L.orc_get_object_maps(self.h_, oid, v, nn)
        return v.reshape(self.h, self.w, 3), nn.reshape(self.h, self.w, 3)

    def keys(self, oid):
        n = self.L.orc_object_keys(self.h_, oid, np.zeros(1, np.uint64), 0)
        out = np.zeros(max(n, 1), np.uint64)
        self.L.orc_object_keys(self.h_, oid, out, n)
        return out[:n]

    def last_keys(self, oid):
        n = self.L.orc_last_keys(self.h_, oid, np.zeros(1, np.uint64), 0)
        out = np.zeros(max(n, 1), np.uint64)
        self.L.orc_last_keys(self.h_, oid, out, n)
        return out[:n]

    def block(self, oid, x, y, z):
        out = np.zeros(512 * 8)
        ok = self.L.orc_object_block(self.h_, oid, x, y, z, out)
        return out.reshape(512, 8) if ok else None

    def rows(self, level):
        n = 2 * self.w * self.h
        r, e, J = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(6 * n, np.float32)
        self.L.orc_get_rows(self.h_, level, r, e, J)
        return r, e, J.reshape(n, 6)

    def trace(self):
        n = self.L.orc_trace(self.h_, np.zeros(22), 0)
        out = np.zeros(22 * max(n, 1))
        self.L.orc_trace(self.h_, out, n)
        return out[:22 * n].reshape(n, 22)

    def trace_clear(self):
        self.L.orc_trace_clear(self.h_)

    def trackresult(self, oid):
        n = 2 * self.w * self.h
        r, e = np.zeros(n, np.int32), np.zeros(n, np.float32)
        self.L.orc_object_trackresult(self.h_, oid, r, e)
        return r, e

    def level(self, level, which):
        comp = 3 if which in (2, 3, 6, 8) else 1
        n = (self.w >> level) * (self.h >> level)
        out = np.zeros(n * comp, np.float32)
        self.L.orc_tracker_level(self.h_, level, which, out)
        return out.reshape((self.h >> level), (self.w >> level), comp) if comp == 3 else \
            out.reshape((self.h >> level), (self.w >> level))

    def frame_buffers(self):
        n = self.w * self.h
        d, g, r = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(3 * n, np.float32)
        self.L.orc_get_frame(self.h_, d, g, r)
        return d.reshape(self.h, self.w), g.reshape(self.h, self.w), r.reshape(self.h, self.w, 3)
