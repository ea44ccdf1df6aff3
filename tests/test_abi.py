"""CPU checks of the C-ABI boundary: the HIP library loads (no GPU needed)
and exports exactly the entry points include/midfusion.h declares, with the
Python mirror binding every one of them."""
import ctypes
import os
import re

from conftest import ROOT, midfusion

HEADER = os.path.join(ROOT, 'include', 'midfusion.h')


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(mf_[a-z0-9_]+)\s*\(', txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ('mf_create', 'mf_destroy', 'mf_preprocess', 'mf_track', 'mf_integrate', 'mf_raycast', 'mf_sync',
                 'mf_get_pose', 'mf_object_create', 'mf_object_keys'):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(midfusion.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    assert set(declared()) == set(midfusion.SIGNATURES)


def test_missing_library_fails_loudly(tmp_path):
    import pytest
    with pytest.raises(midfusion.MFError):
        midfusion.load.__wrapped__(str(tmp_path / 'nope.so')) if hasattr(midfusion.load, '__wrapped__') else \
            _load_fresh(str(tmp_path / 'nope.so'))


def _load_fresh(path):
    saved = midfusion._LIB
    midfusion._LIB = None
    try:
        return midfusion.load(path)
    finally:
        midfusion._LIB = saved
