"""The C-ABI library loads and exports every entry point declared in include/ckmpm.h
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from ckmpm_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ckmpm.h")).read()
    return sorted(set(re.findall(r"\b(ckmpm_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _declared() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_create_without_device_fails_cleanly():
    """Without a gfx950 device ckmpm_create returns an error code, not a crash."""
    import torch

    if torch.cuda.is_available():
        return
    lib = native.load()
    d = native.DomainDesc((ctypes.c_int32 * 3)(16, 16, 16), (ctypes.c_int32 * 3)(0, 0, 0))
    s = native.SimDesc(1 / 64, 0, 0, 0.5, 1e-4, 48, 4096)
    h = ctypes.c_void_p()
    rc = lib.ckmpm_create(ctypes.byref(d), ctypes.byref(s), ctypes.byref(h))
    assert rc in (native.ERR_NO_DEVICE, native.ERR_HIP)
    assert not h


def test_invalid_descriptors_are_refused_before_any_device_call():
    """Domain and simulation descriptors are checked first: an empty or oversized
    domain, dx <= 0 or a zero block capacity give CKMPM_ERR_INVALID with or without a GPU."""
    lib = native.load()
    good_d, good_s = ((16, 16, 16), (0, 0, 0)), (1 / 64, 0, 0, 0.5, 1e-4, 48, 4096)
    cases = [(((0, 16, 16), (0, 0, 0)), good_s), (((16, 2048, 16), (0, 0, 0)), good_s),
             (((1025, 16, 16), (0, 0, 0)), good_s),
             (good_d, (0.0, 0, 0, 0.5, 1e-4, 48, 4096)), (good_d, (1 / 64, 0, 0, 0.5, 1e-4, 48, 0))]
    for (dim, off), sim in cases:
        d = native.DomainDesc((ctypes.c_int32 * 3)(*dim), (ctypes.c_int32 * 3)(*off))
        s = native.SimDesc(*sim)
        h = ctypes.c_void_p()
        assert lib.ckmpm_create(ctypes.byref(d), ctypes.byref(s), ctypes.byref(h)) == native.ERR_INVALID, (dim, sim)
        assert not h
    # the largest domain, 1024^3 blocks, passes the descriptor checks (a key stays 32-bit);
    # only where there is no device, which then refuses it (a device would allocate 8 GiB)
    import torch

    if torch.cuda.device_count() == 0:
        d = native.DomainDesc((ctypes.c_int32 * 3)(1024, 1024, 1024), (ctypes.c_int32 * 3)(0, 0, 0))
        h = ctypes.c_void_p()
        assert lib.ckmpm_create(ctypes.byref(d), ctypes.byref(native.SimDesc(*good_s)), ctypes.byref(h)) in (
            native.ERR_NO_DEVICE, native.ERR_HIP)
        assert not h
