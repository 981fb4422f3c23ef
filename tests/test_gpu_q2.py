"""Q2 (FE_Q(2), QGauss(3)) HIP path vs the CPU oracle through the C ABI.

Same bar as the Q1 parity tests: FP64 relative l2 error <= 1e-12 for the
vmult, residual, diagonal and node blocks; bitwise-identical repeats; FP32
vmult per-field relative l2 error <= 1e-5.  The 3D vmult runs the tile march
by default; a test runs it at several march lengths against the colour
passes.  Meshes mix odd and even cell
counts and include a single cell along a direction, so every colour pass,
empty colours and every first-colour rule of the scatter are exercised.
"""
import numpy as np
import pytest
import torch

import pyoracle as po

pytestmark = pytest.mark.gpu

from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,  # noqa: E402
                        Preconditioner, SinteringOperatorGeneric)

TOL64 = 1e-12


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev(x, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def make_pair(dim, cells, G, tensorial=False, h=None, w0=0.0, w_hist=()):
    h = h or (0.5, 0.75, 1.0)[:dim]
    fe = FreeEnergyParams()
    mob = MobilityParams(tensorial=tensorial)
    mesh = CartesianMesh(tuple(cells), (0.0,) * dim,
                         tuple(n * hh for n, hh in zip(cells, h)))
    op = SinteringOperatorGeneric(mesh, G, fe, mob, fe_degree=2)
    op.set_time_weights(w0, w_hist)
    p = po.make_params(dim, 2, cells, h, G + 2, tensorial=tensorial, A=fe.A, B=fe.B,
                       kappa_c=fe.kappa_c, kappa_p=fe.kappa_p, Mvol=mob.Mvol,
                       Mvap=mob.Mvap, Msurf=mob.Msurf, Mgb=mob.Mgb, L=mob.L, w0=w0,
                       w_hist=w_hist)
    assert op.n_nodes == po.n_nodes(p)
    return op, p


def fields(p, seed):
    rng = np.random.default_rng(seed)
    nn = po.n_nodes(p)
    u = rng.uniform(0.0, 1.0, size=(nn, p.n_comp))
    u[:, 1] = rng.uniform(-1.0, 1.0, size=nn)
    return u.ravel()


VMULT_CASES = [
    (3, (5, 4, 7), 2, False),
    (3, (1, 3, 2), 1, False),
    (3, (6, 5, 3), 4, False),
    (3, (4, 3, 5), 3, True),
    (3, (3, 2, 3), 10, True),
    (2, (9, 6), 2, False),
    (2, (1, 5), 4, False),
    (2, (7, 8), 3, True),
    # grain counts the benchmark configs do not use
    (3, (3, 4, 2), 6, False),
    (3, (2, 3, 3), 11, False),
    (3, (3, 2, 3), 7, True),
    (3, (2, 2, 3), 12, True),
    (2, (7, 8), 9, False),
    (2, (5, 6), 12, True),
]


@pytest.mark.parametrize("dim,cells,G,tens", VMULT_CASES)
def test_q2_vmult_f64_matches_oracle(dim, cells, G, tens):
    op, p = make_pair(dim, cells, G, tens, w0=100.0)
    u = fields(p, 1)
    src = np.random.default_rng(2).uniform(-1, 1, size=u.size)
    ref = po.vmult(p, po.fill_qp(p, u), src)
    op.fill_quadrature_point_values(dev(u))
    dst = op.initialize_dof_vector()
    op.vmult(dst, dev(src))
    got = host(dst)
    assert rel_l2(got, ref) <= TOL64, rel_l2(got, ref)
    dst2 = op.initialize_dof_vector()
    op.vmult(dst2, dev(src))
    assert np.array_equal(host(dst2), got)


@pytest.mark.parametrize("wtd", [0, 1, 2])
@pytest.mark.parametrize("dim,cells,G,tens", [
    (3, (5, 4, 7), 2, False), (3, (4, 3, 5), 3, True), (3, (3, 2, 3), 10, True),
    (2, (9, 6), 2, False), (2, (7, 8), 2, True)])
def test_q2_residual_matches_oracle(dim, cells, G, tens, wtd):
    dt0, dt1 = 1e-3, 2e-3
    w = [(2 * dt0 + dt1) / (dt0 * (dt0 + dt1)), -(dt0 + dt1) / (dt0 * dt1),
         dt0 / (dt1 * (dt0 + dt1))]
    op, p = make_pair(dim, cells, G, tens, w0=w[0], w_hist=(w[1], w[2]))
    u = fields(p, 3)
    old = [fields(p, 4), fields(p, 5)]
    ref = po.residual(p, wtd, u, old)
    dst = op.initialize_dof_vector()
    op.evaluate_nonlinear_residual(dst, dev(u), wtd, old=[dev(o) for o in old])
    assert rel_l2(host(dst), ref) <= TOL64, rel_l2(host(dst), ref)


@pytest.mark.parametrize("cells,G,tens", [
    ((9, 6, 13), 2, False), ((5, 5, 7), 10, True), ((8, 3, 4), 12, True)])
def test_q2_tile_march_matches_colour_passes(cells, G, tens, monkeypatch):
    """3D vmult by the tile march (HPSINT_AMD_Q2_TILE cell layers per march)
    against the colour passes (HPSINT_AMD_Q2_TILE=0) and the oracle, FP64 and
    FP32; meshes with full and partial tiles along every direction."""
    ops = {}
    for zc in (0, 1, 2, 6, 64):
        monkeypatch.setenv("HPSINT_AMD_Q2_TILE", str(zc))
        ops[zc], p = make_pair(3, cells, G, tens, w0=100.0)
    u = fields(p, 1)
    src = np.random.default_rng(2).uniform(-1, 1, size=u.size)
    ref = po.vmult(p, po.fill_qp(p, u), src)
    out64, out32 = {}, {}
    for zc, op in ops.items():
        op.fill_quadrature_point_values(dev(u))
        d = op.initialize_dof_vector()
        op.vmult(d, dev(src))
        out64[zc] = host(d)
        d32 = op.initialize_dof_vector(torch.float32)
        op.vmult(d32, dev(src, torch.float32))
        out32[zc] = host(d32)
        d2 = op.initialize_dof_vector()
        op.vmult(d2, dev(src))
        assert np.array_equal(host(d2), out64[zc])
    for zc in ops:
        assert rel_l2(out64[zc], ref) <= TOL64, (zc, rel_l2(out64[zc], ref))
        assert rel_l2(out64[zc], out64[0]) <= 1e-13, zc
        assert rel_l2(out32[zc], out32[0]) <= 1e-6, zc


@pytest.mark.parametrize("dim,cells,G,tens", [
    (3, (5, 4, 7), 2, False), (3, (3, 2, 3), 10, True), (2, (9, 6), 3, False)])
def test_q2_vmult_f32_error_per_field(dim, cells, G, tens):
    op, p = make_pair(dim, cells, G, tens, w0=100.0)
    u = fields(p, 1)
    src = np.random.default_rng(2).uniform(-1, 1, size=u.size)
    ref = po.vmult(p, po.fill_qp(p, u), src).reshape(-1, p.n_comp)
    op.fill_quadrature_point_values(dev(u))
    dst = op.initialize_dof_vector(torch.float32)
    op.vmult(dst, dev(src, torch.float32))
    got = host(dst).reshape(-1, p.n_comp)
    errs = [rel_l2(got[:, c], ref[:, c]) for c in range(p.n_comp)]
    assert max(errs) <= 1e-5, errs


@pytest.mark.parametrize("dim,cells,G,tens", [
    (3, (5, 4, 3), 2, False), (3, (3, 2, 3), 3, True), (2, (9, 6), 2, False)])
def test_q2_diagonal_and_node_blocks_match_oracle(dim, cells, G, tens):
    op, p = make_pair(dim, cells, G, tens, w0=1000.0)
    u = fields(p, 6)
    blocks_ref = po.node_blocks(p, po.fill_qp(p, u))
    op.fill_quadrature_point_values(dev(u))
    blocks = host(op.node_blocks()).reshape(blocks_ref.shape)
    assert rel_l2(blocks, blocks_ref) <= TOL64, rel_l2(blocks, blocks_ref)
    inv = op.initialize_dof_vector()
    op.compute_inverse_diagonal(inv)
    d = np.einsum("nii->ni", blocks_ref).ravel()
    inv_ref = np.where(np.abs(d) > 1e-10, 1.0 / d, 1.0)
    assert rel_l2(host(inv), inv_ref) <= TOL64
    P = Preconditioner(op, "InverseComponentBlockDiagonalMatrix")
    P.do_update()
    x = np.random.default_rng(9).standard_normal(u.size)
    y = op.initialize_dof_vector()
    P.vmult(y, dev(x))
    yr = np.linalg.solve(blocks_ref, x.reshape(-1, p.n_comp)[..., None])[..., 0].ravel()
    assert rel_l2(host(y), yr) <= 1e-11


@pytest.mark.parametrize("dim,cells,G,tens", [
    (3, (5, 4, 3), 2, False), (3, (3, 2, 3), 3, True), (2, (9, 6), 2, False)])
def test_q2_fp32_preconditioner_assembly(dim, cells, G, tens):
    """diagonal and node blocks assembled from the FP32 cache in FP32:
    relative l2 error against the FP64 oracle <= 1e-5"""
    import ctypes
    from hpsint_amd import _ffi
    op, p = make_pair(dim, cells, G, tens, w0=1000.0)
    u = fields(p, 6)
    blocks_ref = po.node_blocks(p, po.fill_qp(p, u))
    op.fill_quadrature_point_values(dev(u))
    P = Preconditioner(op, "InverseComponentBlockDiagonalMatrix")
    P.do_update(fp32=True)
    nc = p.n_comp
    blocks = torch.empty(op.n_nodes * nc * nc, dtype=torch.float64, device="cuda")
    _ffi.check(_ffi.lib().hps_precond_get_blocks(op.handle, ctypes.c_void_p(blocks.data_ptr())))
    b = host(blocks).reshape(blocks_ref.shape)
    assert rel_l2(b, blocks_ref) <= 1e-5, rel_l2(b, blocks_ref)
    D = Preconditioner(op, "InverseDiagonalMatrix")
    D.do_update(fp32=True)
    inv = op.initialize_dof_vector()
    d = np.einsum("nii->ni", blocks_ref).ravel()
    _ffi.check(_ffi.lib().hps_precond_get_diagonal(op.handle, ctypes.c_void_p(inv.data_ptr())))
    assert rel_l2(host(inv), np.where(np.abs(d) > 1e-10, 1.0 / d, 1.0)) <= 1e-5
