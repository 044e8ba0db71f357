"""The C-ABI library loads and exports every symbol include/*.h declares."""
import ctypes
import os
import re

import pytest

import fluidtissue_amd as fa
from fluidtissue_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(ft_[a-z0-9_]+)\s*\(", src))
    return names


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    decl = _declared()
    assert len(decl) >= 15
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert decl <= set(_lib.exported_symbols()) | _lib.OPTIONAL


def test_euclidean_distance_shim_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    # void EuclideanDistance::init(const float*, const double*, int, int)
    assert hasattr(lib, "_ZN17EuclideanDistance4initEPKfPKdii")
    # void EuclideanDistance::resetPhi(const float*, int, double*)
    assert hasattr(lib, "_ZN17EuclideanDistance8resetPhiEPKfiPd")


def test_opts_default_matches_reference_constants():
    o = _lib.FtOpts()
    _lib.lib.ft_opts_default(ctypes.byref(o))
    assert (o.w_phi, o.w_ifn) == (0.75, 0.25)              # LevelSet.cpp:18
    assert o.mode == _lib.FT_DIST_POINTS and o.blend == _lib.FT_BLEND_INIT and o.band == 0.0


def test_create_without_device_fails_loudly(has_gpu):
    if has_gpu:
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    rc = _lib.lib.ft_create(0, ctypes.byref(h))
    assert rc == _lib.FT_ERR_NODEVICE and not h.value
    assert b"device" in _lib.lib.ft_last_error()
    with pytest.raises(_lib.FtError):
        fa.Reinit(0)


def test_null_handle_is_an_argument_error():
    rc = _lib.lib.ft_set_nodes(None, None, 0)
    assert rc == _lib.FT_ERR_ARG
