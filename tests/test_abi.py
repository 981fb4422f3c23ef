"""The C ABI library loads and exports every entry point of the header
(no GPU needed: symbols only, no compute calls)."""
import ctypes
import os

import pytest

from hpsint_amd import _ffi


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_ffi.LIB_PATH):
        pytest.fail("libhpsint_amd.so not built")
    L = ctypes.CDLL(_ffi.LIB_PATH)
    names = _ffi.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    bound = set(_ffi._SIGNATURES) | {"hps_last_error", "hps_newton_default_params"}
    assert set(names) <= bound, set(names) - bound


def test_invalid_arguments_fail_without_gpu():
    L = _ffi.lib()
    h = ctypes.c_void_p()
    p = _ffi.HpsParams()
    p.dim = 4
    assert L.hps_create(ctypes.byref(p), ctypes.byref(h)) == _ffi.HPS_ERR_UNSUPPORTED
    assert b"dim" in L.hps_last_error()
    p.dim = 3
    p.fe_degree = 1
    for d in range(3):
        p.n_cells[d] = 2
        p.cell_size[d] = 1.0
    p.n_components = 99
    assert L.hps_create(ctypes.byref(p), ctypes.byref(h)) == _ffi.HPS_ERR_INVALID_N_COMP
    with pytest.raises(_ffi.InvalidNumberOfComponents):
        _ffi.check(_ffi.HPS_ERR_INVALID_N_COMP)
    assert L.hps_vmult_f64(None, None, None) == _ffi.HPS_ERR_INVALID_ARGUMENT


def test_adapter_library_loads_in_process():
    """the C++ adapter compiled against the header links the library and
    exports its check entry points (loaded in this process, no compute)"""
    path = os.path.join(os.path.dirname(_ffi.LIB_PATH), "libhpsint_adapter_check.so")
    if not os.path.exists(path):
        pytest.fail("libhpsint_adapter_check.so not built")
    _ffi.lib()
    A = ctypes.CDLL(path)
    for name in ("hpsad_create", "hpsad_destroy", "hpsad_set_time", "hpsad_fill",
                 "hpsad_vmult", "hpsad_residual", "hpsad_precond", "hpsad_hook_calls"):
        assert hasattr(A, name), name


@pytest.mark.parametrize("name,value", [("HPSINT_AMD_ZCOL_MARCH", "0"),
                                        ("HPSINT_AMD_ZCOL_MARCH", "65"),
                                        ("HPSINT_AMD_ZCOL_MARCH_WG", "-1"),
                                        ("HPSINT_AMD_TILE_M", "0"),
                                        ("HPSINT_AMD_Q2_TILE", "-1")])
def test_tuning_switches_out_of_range_fail_without_gpu(monkeypatch, name, value):
    """hps_create checks the tuning switches of its environment before it
    touches the device: a value out of range is an invalid argument"""
    monkeypatch.setenv(name, value)
    L = _ffi.lib()
    h = ctypes.c_void_p()
    p = _ffi.HpsParams()
    p.dim, p.fe_degree, p.n_components = 3, 1, 6
    for d in range(3):
        p.n_cells[d] = 2
        p.cell_size[d] = 1.0
    assert L.hps_create(ctypes.byref(p), ctypes.byref(h)) == _ffi.HPS_ERR_INVALID_ARGUMENT
    assert name.encode() in L.hps_last_error()
