"""Python mirror of the reference `Kfusion` per-frame facade over libmidfusion.so.

Method names follow apps/kfusion/include/kernels.h:242-284 (preprocessing,
tracking, integration, raycasting, getPose).  Everything runs in the HIP
library; this module only marshals arrays.  It raises if the library is
missing: there is no CPU fallback.
"""
import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libmidfusion.so')

ROBUST = {'none': 0, 'huber': 1, 'tukey': 2, 'cauchy': 3}
ROBUST_K = {'none': 1.0, 'huber': 1.345, 'tukey': 4.6851, 'cauchy': 2.3849}


class MFConfig(C.Structure):
    _fields_ = [('width', C.c_int), ('height', C.c_int), ('n_levels', C.c_int), ('iters', C.c_int * 4),
                ('use_icp', C.c_int), ('use_rgb', C.c_int), ('robust', C.c_int), ('robust_k', C.c_double),
                ('mu', C.c_float), ('track_rate', C.c_int), ('integ_rate', C.c_int), ('init_frame', C.c_int),
                ('batch_objects', C.c_int)]


class MFError(RuntimeError):
    pass


_LIB = None

# name: (restype, argtypes) for every entry point of include/midfusion.h
P = C.c_void_p
f32p = np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS')
f64p = np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS')
u64p = np.ctypeslib.ndpointer(np.uint64, flags='C_CONTIGUOUS')
i32p = np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS')
i64p = np.ctypeslib.ndpointer(np.int64, flags='C_CONTIGUOUS')
u16p = np.ctypeslib.ndpointer(np.uint16, flags='C_CONTIGUOUS')
u8p = np.ctypeslib.ndpointer(np.uint8, flags='C_CONTIGUOUS')
SIGNATURES = {
    'mf_create': (P, [C.POINTER(MFConfig), f32p, C.c_float, C.c_int, C.c_long]),
    'mf_destroy': (None, [P]),
    'mf_last_error': (C.c_char_p, [P]),
    'mf_object_create': (C.c_int, [P, C.c_float, C.c_int, f32p, C.c_long]),
    'mf_object_set_moving': (C.c_int, [P, C.c_int, C.c_int]),
    'mf_set_motion_detection': (C.c_int, [P, C.c_int]),
    'mf_get_motion': (C.c_int, [P, i64p, i64p, i32p, C.c_int]),
    'mf_get_motion_state': (C.c_int, [P, i32p, i32p, i32p, C.c_int]),
    'mf_get_motion_codes': (C.c_int, [P, i32p]),
    'mf_motion_residuals': (C.c_int, [P, f32p, f32p, f32p, f32p, f32p, i32p]),
    'mf_set_trackresult': (C.c_int, [P, C.c_int, i32p]),
    'mf_raw_open': (P, [C.c_char_p]),
    'mf_raw_info': (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_long), f32p]),
    'mf_raw_read': (C.c_int, [P, C.c_long, P, P]),
    'mf_raw_close': (None, [P]),
    'mf_preprocess_raw': (C.c_int, [P, P, C.c_long, C.c_int]),
    'mf_parse_gt': (C.c_long, [C.c_char_p, P, C.c_long]),
    'mf_set_gt_poses': (C.c_int, [P, P, C.c_long]),
    'mf_save_poses': (C.c_int, [P, C.c_char_p, C.c_int]),
    'mf_format_tum': (C.c_int, [f32p, C.c_int, C.c_char_p, C.c_int]),
    'mf_render_track': (C.c_int, [P, C.c_int, C.c_int, u8p]),
    'mf_render_depth': (C.c_int, [P, u8p]),
    'mf_render_intensity': (C.c_int, [P, u8p]),
    'mf_volume2mask': (C.c_int, [P, f32p, i32p]),
    'mf_render_volume': (C.c_int, [P, f32p, f32p, C.c_int, C.c_int, C.c_float, C.c_int, P, C.c_int, u8p]),
    'mf_get_track_errors': (C.c_int, [P, C.c_int, f32p]),
    'mf_set_track_errors': (C.c_int, [P, C.c_int, f32p]),
    'mf_object_set_mask': (C.c_int, [P, C.c_int, f32p]),
    'mf_object_set_mask_device': (C.c_int, [P, C.c_int, P]),
    'mf_objects_set_masks_device': (C.c_int, [P, C.c_int, np.ctypeslib.ndpointer(np.int32, flags='C_CONTIGUOUS'),
                                              np.ctypeslib.ndpointer(np.uint64, flags='C_CONTIGUOUS')]),
    'mf_object_set_instance_mask': (C.c_int, [P, C.c_int, u8p, C.c_int, P]),
    'mf_object_set_instance_mask_device': (C.c_int, [P, C.c_int, P, C.c_int, P]),
    'mf_object_get_mask': (C.c_int, [P, C.c_int, f32p]),
    'mf_object_mesh': (C.c_long, [P, C.c_int, P, C.c_long]),
    'mf_dump_mesh': (C.c_int, [P, C.c_char_p]),
    'mf_preprocess': (C.c_int, [P, u16p, u8p, C.c_int, C.c_int, C.c_int]),
    'mf_preprocess_device': (C.c_int, [P, P, P, C.c_int, C.c_int, C.c_int]),
    'mf_track': (C.c_int, [P, f32p, C.c_uint]),
    'mf_track_camera': (C.c_int, [P, f32p, C.c_uint]),
    'mf_track_objects': (C.c_int, [P, f32p, C.c_uint]),
    'mf_track_object': (C.c_int, [P, C.c_int, f32p]),
    'mf_integrate': (C.c_int, [P, f32p, C.c_uint]),
    'mf_raycast': (C.c_int, [P, f32p, C.c_uint]),
    'mf_sync': (C.c_int, [P]),
    'mf_get_pose': (C.c_int, [P, C.c_int, f32p]),
    'mf_set_pose': (C.c_int, [P, C.c_int, f32p]),
    'mf_get_maps': (C.c_int, [P, f32p, f32p, i32p]),
    'mf_set_maps': (C.c_int, [P, f32p, f32p, i32p]),
    'mf_get_object_maps': (C.c_int, [P, C.c_int, f32p, f32p]),
    'mf_set_object_maps': (C.c_int, [P, C.c_int, f32p, f32p]),
    'mf_object_keys': (C.c_long, [P, C.c_int, u64p, C.c_long]),
    'mf_object_block': (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, f64p]),
    'mf_object_node_count': (C.c_long, [P, C.c_int]),
    'mf_get_trackresult': (C.c_int, [P, C.c_int, i32p]),
    'mf_get_trace': (C.c_long, [P, f64p, C.c_long]),
    'mf_solve6': (C.c_int, [P, f64p, C.c_long, f64p, i32p]),
    'mf_object_voxels': (C.c_long, [P, C.c_int, np.ctypeslib.ndpointer(np.uint64, flags='C_CONTIGUOUS'), C.c_long, f64p]),
    'mf_autotune': (C.c_int, [P, f32p, C.c_int, i32p, f64p]),
    'mf_object_voxels_device': (C.c_long, [P, C.c_int, P, C.c_long, P]),
    'mf_clear_trace': (None, [P]),
    'mf_get_level': (C.c_long, [P, C.c_int, C.c_int, f32p]),
    'mf_debug_rows': (C.c_int, [P, C.c_int]),
    'mf_get_rows': (C.c_int, [P, C.c_int, C.c_int, i32p, f32p, f32p]),
    'mf_get_frame': (C.c_int, [P, f32p, f32p, f32p]),
    'mf_enable_timing': (C.c_int, [P, C.c_int]),
    'mf_get_timing': (C.c_int, [P, f64p, C.c_int]),
    'mf_kernel_name': (C.c_char_p, [C.c_int]),
    'mf_get_stats': (C.c_int, [P, u64p, C.c_int]),
    'mf_reset_stats': (C.c_int, [P]),
    'mf_device_bytes': (C.c_long, [P]),
    'mf_copy_reference': (C.c_int, [P, C.c_size_t, C.c_int]),
    'mf_set_tuning': (C.c_int, [P, C.c_int, C.c_int]),
    'mf_residual_eval': (C.c_int, [P, C.c_int, f32p, C.c_int, f32p, np.ctypeslib.ndpointer(np.float64, flags='C_CONTIGUOUS')]),
    'mf_replay_track_rows': (C.c_int, [P, f32p, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    'mf_get_tuning': (C.c_int, [P, C.c_int, C.POINTER(C.c_int)]),
}


def load(path=LIB_PATH):
    """Load the HIP library; raises if it was not built (no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise MFError('libmidfusion.so is not built: run mid-fusion_amd/build.py (no CPU fallback exists)')
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class RawSequence:
    """A SLAMBench .raw sequence (RawDepthReader, interface.h:207-333)."""

    def __init__(self, path):
        self.L = load()
        self.h_ = self.L.mf_raw_open(path.encode())
        if not self.h_:
            raise MFError('Invalid Raw file: %s' % path)
        w, h, n = C.c_int(), C.c_int(), C.c_long()
        k = np.zeros(4, np.float32)
        self.L.mf_raw_info(self.h_, C.byref(w), C.byref(h), C.byref(n), k)
        self.width, self.height, self.frames, self.k = w.value, h.value, n.value, k

    def read(self, frame):
        """(depth uint16 [h, w], rgb uint8 [h, w, 3]); MFError past the end."""
        d = np.zeros((self.height, self.width), np.uint16)
        r = np.zeros((self.height, self.width, 3), np.uint8)
        rc = self.L.mf_raw_read(self.h_, frame, d.ctypes.data, r.ctypes.data)
        if rc != 0:
            raise MFError('end of the raw sequence (frame %d)' % frame)
        return d, r

    def close(self):
        if getattr(self, 'h_', None):
            self.L.mf_raw_close(self.h_)
            self.h_ = None

    def __del__(self):
        self.close()


def parse_gt(path):
    """TUM ground truth (parseGTFile, vtk-io.h:304-330): poses [n, 4, 4]."""
    L = load()
    n = L.mf_parse_gt(path.encode(), None, 0)
    if n < 0:
        raise MFError('cannot read ground truth %s' % path)
    out = np.zeros((n, 4, 4), np.float32)
    L.mf_parse_gt(path.encode(), out.ctypes.data, n)
    return out


def format_tum(T, frame):
    """One Kfusion::save_poses line for pose T."""
    buf = C.create_string_buffer(256)
    n = load().mf_format_tum(_f32(np.asarray(T).reshape(16)), frame, buf, 256)
    if n < 0:
        raise MFError('pose line')
    return buf.value.decode()


class Kfusion:
    """GPU per-frame dense path of one camera + object instances."""

    def __init__(self, width, height, init_pose, bg_size=5.12, bg_res=512, iters=(10, 5, 4), use_icp=True,
                 use_rgb=True, robust='cauchy', robust_k=None, mu=0.1, track_rate=1, integ_rate=1, init_frame=0,
                 batch_objects=True, bg_capacity=0):
        self.L = load()
        self.w, self.h = width, height
        self.n_levels = len(iters)
        cfg = MFConfig()
        cfg.width, cfg.height, cfg.n_levels = width, height, len(iters)
        for i, v in enumerate(iters):
            cfg.iters[i] = v
        cfg.use_icp, cfg.use_rgb = int(use_icp), int(use_rgb)
        cfg.robust = ROBUST[robust]
        cfg.robust_k = ROBUST_K[robust] if robust_k is None else robust_k
        cfg.mu, cfg.track_rate, cfg.integ_rate, cfg.init_frame = mu, track_rate, integ_rate, init_frame
        cfg.batch_objects = int(batch_objects)
        self.ctx = self.L.mf_create(C.byref(cfg), _f32(np.asarray(init_pose).reshape(16)), bg_size, bg_res,
                                    bg_capacity)
        if not self.ctx:
            raise MFError('mf_create failed (no HIP device or out of memory)')

    def close(self):
        if getattr(self, 'ctx', None):
            self.L.mf_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def _chk(self, r):
        if r < 0:
            raise MFError(self.L.mf_last_error(self.ctx).decode() + ' (code %d)' % r)
        return r

    # ---- objects (Object, object.h:48-51)
    def add_object(self, size, res, pose, capacity=0):
        return self._chk(self.L.mf_object_create(self.ctx, size, res, _f32(np.asarray(pose).reshape(16)), capacity))

    def set_moving(self, oid, moving):
        self._chk(self.L.mf_object_set_moving(self.ctx, oid, int(moving)))

    def set_motion_detection(self, on):
        """Motion detection on/off (off: the moving set is the caller's flags)."""
        self._chk(self.L.mf_set_motion_detection(self.ctx, int(on)))

    def motion_result(self):
        """Last motion check: per object (pixels in view, inliers, moving flag)."""
        cap = 64
        px, inl, mv = np.zeros(cap, np.int64), np.zeros(cap, np.int64), np.zeros(cap, np.int32)
        n = self.L.mf_get_motion(self.ctx, px, inl, mv, cap)
        if n < 0:
            self._chk(n)
        return px[:n], inl[:n], mv[:n]

    def motion_state(self):
        """(gate, rendered_n, in-view flags, moving set) of the last tracking call."""
        cap = 64
        g, iv, ms = np.zeros(2, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.int32)
        n = self.L.mf_get_motion_state(self.ctx, g, iv, ms, cap)
        if n < 0:
            self._chk(n)
        return int(g[0]), int(g[1]), iv[:n], ms[:n]

    def motion_codes(self):
        """Residual codes of the last motion check (2*W*H: ICP then RGB)."""
        r = np.zeros(2 * self.w * self.h, np.int32)
        self._chk(self.L.mf_get_motion_codes(self.ctx, r))
        return r

    def motion_residuals(self, k, T, vtx, nrm, vbf):
        """compute_residuals codes (2*W*H) at pose T against the given maps."""
        out = np.zeros(2 * self.w * self.h, np.int32)
        f = lambda a: np.ascontiguousarray(a, np.float32).reshape(-1)
        self._chk(self.L.mf_motion_residuals(self.ctx, f(k), f(T), f(vtx), f(nrm), f(vbf), out))
        return out

    def render_track(self, oid, mode):
        out = np.zeros((self.h, self.w, 4), np.uint8)
        self._chk(self.L.mf_render_track(self.ctx, oid, mode, out.reshape(-1)))
        return out

    def render_depth(self):
        out = np.zeros((self.h, self.w, 4), np.uint8)
        self._chk(self.L.mf_render_depth(self.ctx, out.reshape(-1)))
        return out

    def render_intensity(self):
        out = np.zeros((self.h, self.w, 4), np.uint8)
        self._chk(self.L.mf_render_intensity(self.ctx, out.reshape(-1)))
        return out

    def volume2mask(self, k):
        out = np.zeros((self.h, self.w), np.int32)
        self._chk(self.L.mf_volume2mask(self.ctx, _f32(k), out.reshape(-1)))
        return out

    def render_volume(self, k, view_pose, w=None, h=None, largestep=0.075, render_color=True, palette=None):
        """(image [h, w, 4] uint8, raycast?) of renderVolume at view_pose."""
        w, h = w or self.w, h or self.h
        out = np.zeros((h, w, 4), np.uint8)
        pal = np.zeros((0, 4), np.uint8) if palette is None else np.ascontiguousarray(palette, np.uint8)
        r = self._chk(self.L.mf_render_volume(self.ctx, _f32(k), _f32(np.asarray(view_pose).reshape(16)), w, h,
                                              largestep, int(render_color), pal.ctypes.data if len(pal) else None,
                                              len(pal), out.reshape(-1)))
        return out, bool(r)

    def track_errors(self, oid):
        out = np.zeros(self.w * self.h, np.float32)
        self._chk(self.L.mf_get_track_errors(self.ctx, oid, out))
        return out

    def set_track_errors(self, oid, err):
        self._chk(self.L.mf_set_track_errors(self.ctx, oid, _f32(err).reshape(-1)))

    def set_trackresult(self, oid, codes):
        self._chk(self.L.mf_set_trackresult(self.ctx, oid, np.ascontiguousarray(codes, np.int32).reshape(-1)))

    def set_instance_mask(self, oid, mask, rendered=False, model_mask=None):
        """generalize_label on the device from a 0/non-zero instance mask."""
        m = np.ascontiguousarray(mask, np.uint8).reshape(-1)
        g = None if model_mask is None else np.ascontiguousarray(model_mask, np.uint8).reshape(-1)
        self._chk(self.L.mf_object_set_instance_mask(self.ctx, oid, m, int(rendered),
                                                     None if g is None else g.ctypes.data))

    def mesh(self, oid):
        """Marching-cubes triangles of object oid: [n, 3, 3] float32 (object frame)."""
        n = self._chk(self.L.mf_object_mesh(self.ctx, oid, None, 0))
        out = np.zeros((max(n, 1), 3, 3), np.float32)
        if n:
            self._chk(self.L.mf_object_mesh(self.ctx, oid, out.ctypes.data, n))
        return out[:n]

    def dump_mesh(self, prefix):
        self._chk(self.L.mf_dump_mesh(self.ctx, prefix.encode()))

    def get_mask(self, oid):
        out = np.zeros(self.w * self.h, np.float32)
        self._chk(self.L.mf_object_get_mask(self.ctx, oid, out))
        return out.reshape(self.h, self.w)

    def set_mask(self, oid, mask):
        self._chk(self.L.mf_object_set_mask(self.ctx, oid, _f32(mask).reshape(-1)))

    def set_masks_device(self, ids, ptrs):
        """Integration masks of several objects from device pointers (ints;
        0 clears a mask), one copy launch."""
        ids = np.ascontiguousarray(ids, np.int32)
        ptrs = np.ascontiguousarray([int(p or 0) for p in ptrs], np.uint64)
        self._chk(self.L.mf_objects_set_masks_device(self.ctx, len(ids), ids, ptrs))

    # ---- per-frame stages (kernels.h:242-250)
    def preprocessing(self, depth, rgb, filt=False):
        d = np.ascontiguousarray(depth, np.uint16)
        r = np.ascontiguousarray(rgb, np.uint8)
        return self._chk(self.L.mf_preprocess(self.ctx, d.reshape(-1), r.reshape(-1), d.shape[1], d.shape[0],
                                              int(filt)))

    def preprocessing_raw(self, raw, frame, filt=False):
        """Frame `frame` of a RawSequence through pinned staging."""
        return self._chk(self.L.mf_preprocess_raw(self.ctx, raw.h_, frame, int(filt)))

    def set_gt_poses(self, poses):
        """Ground-truth camera poses [n, 4, 4] (n = 0: off)."""
        p = np.ascontiguousarray(np.asarray(poses, np.float32).reshape(-1, 16))
        return self._chk(self.L.mf_set_gt_poses(self.ctx, p.ctypes.data if len(p) else None, len(p)))

    def save_poses(self, prefix, frame):
        return self._chk(self.L.mf_save_poses(self.ctx, prefix.encode(), frame))

    def preprocessing_device(self, d_depth, d_rgb, in_w, in_h, filt=False):
        return self._chk(self.L.mf_preprocess_device(self.ctx, d_depth, d_rgb, in_w, in_h, int(filt)))

    def tracking(self, k, frame):
        return self._chk(self.L.mf_track(self.ctx, _f32(k), frame))

    def track_camera(self, k, frame):
        return self._chk(self.L.mf_track_camera(self.ctx, _f32(k), frame))

    def track_objects(self, k, frame):
        return self._chk(self.L.mf_track_objects(self.ctx, _f32(k), frame))

    def track_object(self, oid, k):
        """trackEachObject of one object (between track_camera and track_objects)."""
        return self._chk(self.L.mf_track_object(self.ctx, oid, _f32(k)))

    def integration(self, k, frame):
        return self._chk(self.L.mf_integrate(self.ctx, _f32(k), frame))

    def raycasting(self, k, frame):
        return self._chk(self.L.mf_raycast(self.ctx, _f32(k), frame))

    def sync(self):
        return self._chk(self.L.mf_sync(self.ctx))

    # ---- state
    def getPose(self, oid=-1):
        T = np.zeros(16, np.float32)
        self._chk(self.L.mf_get_pose(self.ctx, oid, T))
        return T.reshape(4, 4)

    def setPose(self, oid, T):
        self._chk(self.L.mf_set_pose(self.ctx, oid, _f32(np.asarray(T).reshape(16))))

    def maps(self):
        n = self.w * self.h
        v, nn, lab = np.zeros(3 * n, np.float32), np.zeros(3 * n, np.float32), np.zeros(n, np.int32)
        self._chk(self.L.mf_get_maps(self.ctx, v, nn, lab))
        return v.reshape(self.h, self.w, 3), nn.reshape(self.h, self.w, 3), lab.reshape(self.h, self.w)

    def set_maps(self, vtx, nrm, lab):
        self._chk(self.L.mf_set_maps(self.ctx, _f32(vtx).reshape(-1), _f32(nrm).reshape(-1),
                                     np.ascontiguousarray(lab, np.int32).reshape(-1)))

    def object_maps(self, oid):
        n = self.w * self.h
        v, nn = np.zeros(3 * n, np.float32), np.zeros(3 * n, np.float32)
        self._chk(self.L.mf_get_object_maps(self.ctx, oid, v, nn))
        return v.reshape(self.h, self.w, 3), nn.reshape(self.h, self.w, 3)

    def set_object_maps(self, oid, vtx, nrm):
        self._chk(self.L.mf_set_object_maps(self.ctx, oid, _f32(vtx).reshape(-1), _f32(nrm).reshape(-1)))

    def keys(self, oid):
        n = self._chk(self.L.mf_object_keys(self.ctx, oid, np.zeros(1, np.uint64), 0))
        out = np.zeros(max(n, 1), np.uint64)
        self.L.mf_object_keys(self.ctx, oid, out, n)
        return out[:n]

    def block(self, oid, x, y, z):
        out = np.zeros(512 * 8)
        ok = self._chk(self.L.mf_object_block(self.ctx, oid, x, y, z, out))
        return out.reshape(512, 8) if ok else None

    def voxels(self, oid, keys):
        """Voxels of the blocks with `keys` (mf_object_keys encoding):
        (array [n, 512, 8] of {tsdf, w, r, g, b, c, fg, fgw}, allocated count)."""
        keys = np.ascontiguousarray(keys, np.uint64)
        out = np.zeros(len(keys) * 512 * 8)
        n = self.L.mf_object_voxels(self.ctx, oid, keys, len(keys), out)
        if n < 0:
            self._chk(n)
        return out.reshape(len(keys), 512, 8), int(n)

    def voxels_device(self, oid, d_keys, n, d_out):
        """mf_object_voxels_device on device pointers; returns the allocated count."""
        return self._chk(self.L.mf_object_voxels_device(self.ctx, oid, d_keys, n, d_out))

    def node_count(self, oid):
        return self._chk(self.L.mf_object_node_count(self.ctx, oid))

    def trackresult(self, oid):
        out = np.zeros(2 * self.w * self.h, np.int32)
        self._chk(self.L.mf_get_trackresult(self.ctx, oid, out))
        return out

    def trace(self):
        n = self.L.mf_get_trace(self.ctx, np.zeros(25), 0)
        out = np.zeros(25 * max(n, 1))
        self.L.mf_get_trace(self.ctx, out, n)
        return out[:25 * n].reshape(n, 25)

    def trace_clear(self):
        self.L.mf_clear_trace(self.ctx)

    def level(self, level, which):
        comp = 3 if which in (2, 3, 6) else 1
        n = (self.w >> level) * (self.h >> level)
        out = np.zeros(n * comp, np.float32)
        self._chk(self.L.mf_get_level(self.ctx, level, which, out))
        shape = ((self.h >> level), (self.w >> level))
        return out.reshape(shape + (3,)) if comp == 3 else out.reshape(shape)

    def debug_rows(self, on=True):
        self._chk(self.L.mf_debug_rows(self.ctx, int(on)))

    def rows(self, state, level):
        n = 2 * self.w * self.h
        r, e, J = np.zeros(n, np.int32), np.zeros(n, np.float32), np.zeros(6 * n, np.float32)
        self._chk(self.L.mf_get_rows(self.ctx, state, level, r, e, J))
        return r, e, J.reshape(n, 6)

    def frame_buffers(self):
        n = self.w * self.h
        d, g, r = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(3 * n, np.float32)
        self._chk(self.L.mf_get_frame(self.ctx, d, g, r))
        return d.reshape(self.h, self.w), g.reshape(self.h, self.w), r.reshape(self.h, self.w, 3)

    def enable_timing(self, on=True):
        self.L.mf_enable_timing(self.ctx, int(on))

    def timing(self):
        out = np.zeros(64)
        n = self.L.mf_get_timing(self.ctx, out, 32)
        return {self.L.mf_kernel_name(i).decode(): (out[2 * i], int(out[2 * i + 1])) for i in range(n)}

    STAT_NAMES = ('voxel_updates', 'blocks_integrated', 'march_steps', 'trilinear_evals', 'hits', 'band_samples',
                  'alloc_pixels', 'ray_object_pairs', 'octree_nodes')

    def stats(self):
        out = np.zeros(len(self.STAT_NAMES), np.uint64)
        self.L.mf_get_stats(self.ctx, out, len(self.STAT_NAMES))
        return dict(zip(self.STAT_NAMES, (int(v) for v in out)))

    def reset_stats(self):
        self.L.mf_reset_stats(self.ctx)

    def device_bytes(self):
        return self.L.mf_device_bytes(self.ctx)

    def autotune(self, k, reps=30):
        """Launch-shape autotuner: times each pyramid level's rows-per-thread
        candidates TUNE_CANDIDATES on the last tracked frame and keeps the
        fastest.  Returns (chosen per level, us per iteration [level, cand])."""
        nl, nc = self.n_levels, len(self.TUNE_CANDIDATES)
        chosen, us = np.zeros(nl, np.int32), np.zeros(nl * nc)
        self._chk(self.L.mf_autotune(self.ctx, _f32(k), int(reps), chosen, us))
        return chosen, us.reshape(nl, nc)

    def solve6(self, s27):
        """Device 6x6 solve of systems s27 [n, 27] (JtWe, packed upper JtWJ):
        returns (x [n, 6], solver [n]; 1 Gauss-Jordan, 2 pseudo-inverse)."""
        s27 = np.ascontiguousarray(s27, np.float64).reshape(-1, 27)
        n = s27.shape[0]
        x, sv = np.zeros((n, 6)), np.zeros(n, np.int32)
        self._chk(self.L.mf_solve6(self.ctx, s27, n, x, sv))
        return x, sv

    def residual_eval(self, oid, k, level, T):
        """32 FP64 sums of the tracking residual of the camera (oid < 0, T =
        T_w_l) or object oid (T = T_c2_o2) at `level`: loss, J^T W e,
        J^T W J, counts.  Call between track_camera and track_objects."""
        out = np.zeros(32)
        self._chk(self.L.mf_residual_eval(self.ctx, oid, _f32(k), level, _f32(np.asarray(T).reshape(16)), out))
        return out

    def replay_track_rows(self, k, level, reps):
        us, b = C.c_double(), C.c_double()
        self._chk(self.L.mf_replay_track_rows(self.ctx, _f32(k), level, reps, C.byref(us), C.byref(b)))
        return us.value, b.value

    TUNE_TRACK_ROWS_PER_THREAD_L0 = 0   # + pyramid level
    TUNE_TRACK_FUSED_L0 = 8             # + pyramid level (1 fused, 0 two launches, -1 automatic)
    TUNE_OBJ_PIXELS_PER_THREAD_L0 = 16  # + pyramid level
    TUNE_OBJ_ROWS_PER_THREAD_L0 = 24    # + pyramid level (0: {8, 8, 4, 4} for 2+ objects per launch, else the camera's)
    TUNE_CANDIDATES = (1, 2, 3, 4, 8, 16)   # rows per thread tried by autotune()

    def set_tuning(self, knob, value):
        self._chk(self.L.mf_set_tuning(self.ctx, knob, int(value)))

    def get_tuning(self, knob):
        v = C.c_int()
        self._chk(self.L.mf_get_tuning(self.ctx, knob, C.byref(v)))
        return v.value
