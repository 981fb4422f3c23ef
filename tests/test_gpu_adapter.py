"""The C++ adapter a maintainer adds to hpsint (hpsint_amd/adapter/
sintering_operator_mi355x.h), compiled against include/hpsint_amd.h by the
build, driven in-process through its C entry points
(libhpsint_adapter_check.so): the same vmult, residual (with its range
hooks) and block preconditioner as the Python binding, bitwise."""
import ctypes
import os

import numpy as np
import pytest
import torch

from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,  # noqa: E402
                        Preconditioner, SinteringOperatorGeneric, _ffi)

pytestmark = pytest.mark.gpu

ADAPTER = os.path.join(os.path.dirname(_ffi.LIB_PATH), "libhpsint_adapter_check.so")


def adapter():
    A = ctypes.CDLL(ADAPTER)
    A.hpsad_last_error.restype = ctypes.c_char_p
    A.hpsad_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                 ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_void_p)]
    for name in ("hpsad_fill",):
        getattr(A, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    for name in ("hpsad_vmult",):
        getattr(A, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    A.hpsad_residual.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    A.hpsad_precond.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    return A


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("cells,G", [((9, 5, 17), 4), ((13, 9), 2)])
def test_adapter_matches_the_binding(cells, G):
    _ffi.lib()  # the library the adapter links, loaded first
    A = adapter()
    dim = len(cells)
    mesh = CartesianMesh(cells, (0.0,) * dim, tuple(0.5 * n for n in cells))
    op = SinteringOperatorGeneric(mesh, G, FreeEnergyParams(), MobilityParams())
    h = ctypes.c_void_p()
    assert A.hpsad_create(ctypes.byref(op.params), ctypes.byref(h)) == 0, A.hpsad_last_error()
    try:
        assert A.hpsad_n_components(h) == G + 2
        dt = 1e-3
        rng = np.random.default_rng(5)
        u = torch.from_numpy(rng.uniform(0, 1, op.n_dofs)).cuda()
        old = torch.from_numpy(rng.uniform(0, 1, op.n_dofs)).cuda()
        src = torch.from_numpy(rng.uniform(-1, 1, op.n_dofs)).cuda()
        op.set_time_weights(1.0 / dt, (-1.0 / dt,))
        w = (ctypes.c_double * 1)(-1.0 / dt)
        olds = (ctypes.c_void_p * 1)(old.data_ptr())
        torch.cuda.synchronize()
        assert A.hpsad_set_time(h, 1.0 / dt, 1.0, w, 1, olds) == 0
        # residual, every wtd, with the range hooks called once each
        for wtd in (0, 1, 2):
            r_ref = op.initialize_dof_vector()
            op.evaluate_nonlinear_residual(r_ref, u, wtd, old=[old])
            r = torch.zeros_like(r_ref)
            torch.cuda.synchronize()
            assert A.hpsad_residual(h, wtd, ptr(r), ptr(u)) == 0, A.hpsad_last_error()
            assert torch.equal(r, r_ref)
        pre, post, end = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        A.hpsad_hook_calls(h, ctypes.byref(pre), ctypes.byref(post), ctypes.byref(end))
        assert (pre.value, post.value, end.value) == (3, 3, op.n_dofs)
        # tangent
        op.fill_quadrature_point_values(u)
        y_ref = op.initialize_dof_vector()
        op.vmult(y_ref, src)
        y = torch.zeros_like(y_ref)
        torch.cuda.synchronize()
        assert A.hpsad_fill(h, ptr(u)) == 0
        assert A.hpsad_vmult(h, ptr(y), ptr(src)) == 0
        assert torch.equal(y, y_ref)
        # the reference's default block preconditioner
        P = Preconditioner(op, "BlockPreconditioner2")
        P.do_update()
        z_ref = op.initialize_dof_vector()
        P.vmult(z_ref, src)
        z = torch.zeros_like(z_ref)
        torch.cuda.synchronize()
        assert A.hpsad_precond(h, _ffi.HPS_PRECON_BLOCK_CH_ILU, ptr(z), ptr(src)) == 0
        assert torch.equal(z, z_ref)
    finally:
        A.hpsad_destroy(h)


def test_adapter_maps_status_codes():
    _ffi.lib()
    A = adapter()
    p = _ffi.HpsParams()
    p.dim, p.fe_degree, p.n_components = 3, 1, 99
    for d in range(3):
        p.n_cells[d], p.cell_size[d] = 2, 1.0
    h = ctypes.c_void_p()
    assert A.hpsad_create(ctypes.byref(p), ctypes.byref(h)) == _ffi.HPS_ERR_INVALID_N_COMP
    assert b"n_components" in A.hpsad_last_error()
