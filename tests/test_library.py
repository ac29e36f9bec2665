"""The C-ABI library loads and exports every entry point the header declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "kagome_vmc.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(kvmc_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from kagome_vmc import _native
    if not os.path.exists(_native.LIB_PATH):
        import importlib.util
        spec = importlib.util.spec_from_file_location(
            "kvmc_build", os.path.join(ROOT, "kagomedsl.jl_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()
    return _native.load()


def test_header_declares_the_abi():
    syms = declared_symbols()
    for s in ["kvmc_create", "kvmc_sweep", "kvmc_measure", "kvmc_refresh", "kvmc_destroy"]:
        assert s in syms


def test_every_declared_symbol_is_exported(lib):
    from kagome_vmc import _native
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert sorted(_native.EXPORTS) == declared_symbols()


def test_version_and_error_string(lib):
    assert b"gfx950" in lib.kvmc_version()
    assert isinstance(lib.kvmc_last_error(), bytes)


def test_struct_layout_matches_header(lib):
    from kagome_vmc import _native
    # kvmc_desc: 4 int32, 3 pointers, int64, 6 int32, uint64, int32 + 4 bytes
    # of tail padding (x86-64 layout)
    assert ctypes.sizeof(_native.KvmcDesc) == 4 * 4 + 3 * 8 + 8 + 6 * 4 + 8 + 4 + 4
    assert ctypes.sizeof(_native.KvmcCounters) == 6 * 8


def test_invalid_descriptor_is_rejected_before_any_gpu_work(lib):
    """Argument validation happens on the host before HIP is touched."""
    from kagome_vmc import _native
    d = _native.KvmcDesc()
    d.ns, d.n_up, d.n_down, d.nb = 12, 5, 6, 0     # n_up + n_down != ns
    h = ctypes.c_void_p()
    rc = lib.kvmc_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _native.KVMC_ERR_INVALID
    assert b"n_up + n_down" in lib.kvmc_last_error()


def test_refresh_size_limit_is_checked_before_any_gpu_work(lib):
    """More than 768 rows of tildeU in the refresh (here complex orbitals at
    972 sites, 486 electrons per spin: a 972 x 972 real form) is rejected
    when the handle is created, with the reason."""
    import numpy as np
    from kagome_vmc import _native
    ns, n = 972, 486
    bonds = np.array([[1, 2]], dtype=np.int32)
    u = np.zeros(2 * ns * n)
    d = _native.KvmcDesc()
    d.ns, d.n_up, d.n_down, d.nb = ns, n, n, 1
    d.bonds = bonds.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    d.u_up = u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d.u_down = u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d.nwalkers = 1
    d.precision, d.kernel, d.refresh_policy = _native.PREC_FP64, _native.KERNEL_REFERENCE, _native.REFRESH_REFERENCE
    d.delay = 1
    d.orbitals = _native.ORB_COMPLEX
    h = ctypes.c_void_p()
    rc = lib.kvmc_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _native.KVMC_ERR_INVALID
    assert b"too large for the refresh kernel" in lib.kvmc_last_error()
