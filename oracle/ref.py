"""ctypes binding of oracle/_ref/libref.so: the reference's own octree core
compiled in place (TEST INFRASTRUCTURE ONLY, used to pin the oracle)."""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(_HERE, '_ref', 'libref.so')
_LIB = None


def available():
    return os.path.exists(PATH)


def lib():
    global _LIB
    if _LIB is None:
        L = C.CDLL(PATH)
        P = C.c_void_p
        f32p = np.ctypeslib.ndpointer(np.float32, flags='C_CONTIGUOUS')
        u64p = np.ctypeslib.ndpointer(np.uint64, flags='C_CONTIGUOUS')
        u32p = np.ctypeslib.ndpointer(np.uint32, flags='C_CONTIGUOUS')
        sig = {
            'ref_encode': (C.c_uint64, [C.c_int] * 5),
            'ref_morton': (C.c_uint64, [C.c_uint32] * 3),
            'ref_unpack': (None, [C.c_uint64, u32p]),
            'ref_parent': (C.c_uint64, [C.c_uint64, C.c_int]),
            'ref_child_id': (C.c_int, [C.c_uint64, C.c_int, C.c_int]),
            'ref_unique_ms': (C.c_int, [u64p, C.c_int, C.c_int]),
            'ref_octree_create': (P, [C.c_int, C.c_float]),
            'ref_octree_alloc': (None, [P, u64p, C.c_int]),
            'ref_octree_hash': (C.c_uint64, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
            'ref_octree_keys': (C.c_long, [P, u64p, C.c_long]),
            'ref_octree_node_keys': (C.c_long, [P, u64p, C.c_long]),
            'ref_octree_set': (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]),
            'ref_octree_interp': (C.c_float, [P, C.c_float, C.c_float, C.c_float, C.c_int]),
            'ref_octree_grad': (None, [P, C.c_float, C.c_float, C.c_float, f32p]),
            'ref_octree_ray_tmin': (C.c_float, [P, f32p, f32p, C.c_float, C.c_float]),
            'ref_octree_destroy': (None, [P]),
            'ref_volume_create': (P, [C.c_int, C.c_float]),
            'ref_volume_alloc': (None, [P, u64p, C.c_int]),
            'ref_volume_set': (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]),
            'ref_volume_raycast': (C.c_float, [P, f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_float,
                                               C.c_float, f32p]),
            'ref_volume_destroy': (None, [P]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _LIB = L
    return _LIB


def build():
    import subprocess
    subprocess.check_call(['make', '-s', '-f', 'Makefile.ref'], cwd=_HERE)
