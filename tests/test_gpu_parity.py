"""HIP path vs the CPU oracle, through the C ABI, on the same seeded inputs.

Bar: FP64 relative l2 error <= 1e-12 (residual, vmult, diagonal, node
blocks); bitwise-identical repeats; FP32 vmult within the per-field bound
written in the test.  Meshes are ragged (not multiples of the tile) so that
partial tiles and every class of shared node are exercised.
"""
import ctypes

import numpy as np
import pytest
import torch

import pyoracle as po

pytestmark = pytest.mark.gpu

from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,  # noqa: E402
                        Preconditioner, SinteringOperatorGeneric)
from hpsint_amd import _ffi  # noqa: E402

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


def make_pair(dim, cells, G, tensorial=False, h=None, fe=None, mob=None, w0=0.0,
              w_hist=()):
    h = h or (0.5, 0.75, 1.0)[:dim]
    fe = fe or FreeEnergyParams()
    mob = mob or MobilityParams(tensorial=tensorial)
    mesh = CartesianMesh(tuple(cells), (0.0,) * dim,
                         tuple(n * hh for n, hh in zip(cells, h)))
    op = SinteringOperatorGeneric(mesh, G, fe, mob)
    op.set_time_weights(w0, w_hist)
    p = po.make_params(dim, 1, cells, h, G + 2, tensorial=tensorial, A=fe.A, B=fe.B,
                       kappa_c=fe.kappa_c, kappa_p=fe.kappa_p, Mvol=mob.Mvol,
                       Mvap=mob.Mvap, Msurf=mob.Msurf, Mgb=mob.Mgb, L=mob.L, w0=w0,
                       w_hist=w_hist)
    return op, p


def fields(p, seed):
    """random nodal state: c, eta in [0,1], mu in [-1,1]"""
    rng = np.random.default_rng(seed)
    nn = po.n_nodes(p)
    u = rng.uniform(0.0, 1.0, size=(nn, p.n_comp))
    u[:, 1] = rng.uniform(-1.0, 1.0, size=nn)
    return u.ravel()


VMULT_CASES = [
    (3, (9, 5, 17), 2, False),
    (3, (19, 13, 37), 4, False),
    (3, (8, 4, 16), 1, False),
    (3, (3, 7, 2), 6, False),
    (3, (10, 6, 19), 5, False),
    (3, (9, 5, 17), 8, False),
    (3, (7, 5, 18), 12, False),
    (3, (11, 6, 18), 3, True),
    (3, (5, 9, 20), 2, True),
    (2, (65, 17), 2, False),
    (2, (130, 33), 4, False),
    (2, (70, 19), 2, True),
    # grain counts the benchmark configs do not use (every G of
    # EXPAND_OPERATIONS, S/instantiation.h:97-152, is built)
    (3, (9, 5, 17), 7, False),
    (3, (10, 6, 19), 9, False),
    (3, (7, 5, 18), 11, False),
    (3, (9, 5, 11), 5, True),
    (3, (7, 5, 9), 8, True),
    (3, (5, 4, 6), 12, True),
    (2, (65, 17), 7, False),
    (2, (70, 19), 6, True),
    (2, (33, 9), 12, True),
]


@pytest.mark.parametrize("dim,cells,G,tens", VMULT_CASES)
def test_vmult_f64_matches_oracle(dim, cells, G, tens):
    op, p = make_pair(dim, cells, G, tens, w0=100.0)
    u = fields(p, 1)
    src = np.random.default_rng(2).uniform(-1, 1, size=u.size)
    cache = po.fill_qp(p, u)
    ref = po.vmult(p, cache, src)
    du = dev(u)
    op.fill_quadrature_point_values(du)
    dst = op.initialize_dof_vector()
    op.vmult(dst, dev(src))
    got = host(dst)
    assert rel_l2(got, ref) <= TOL64, rel_l2(got, ref)
    # deterministic: a second application is bitwise identical
    dst2 = op.initialize_dof_vector()
    op.vmult(dst2, dev(src))
    assert np.array_equal(host(dst2), got)


RES_CASES = [
    (3, (9, 5, 17), 2, False),
    (3, (19, 13, 37), 3, False),
    (3, (10, 6, 19), 5, False),
    (3, (9, 5, 17), 8, False),
    (3, (7, 5, 18), 12, False),
    (3, (11, 6, 18), 3, True),
    (2, (65, 17), 2, False),
    (2, (70, 19), 2, True),
    (3, (10, 6, 19), 9, False),
    (3, (7, 5, 18), 11, False),
    (3, (9, 5, 11), 5, True),
    (3, (7, 5, 9), 8, True),
    (2, (70, 19), 6, True),
]


@pytest.mark.parametrize("wtd", [0, 1, 2])
@pytest.mark.parametrize("dim,cells,G,tens", RES_CASES)
def test_residual_matches_oracle(dim, cells, G, tens, wtd):
    # BDF2 with unequal steps (time_schemes.h:87-105)
    dt0, dt1 = 1e-3, 2e-3
    w = [(2 * dt0 + dt1) / (dt0 * (dt0 + dt1)), -(dt0 + dt1) / (dt0 * dt1),
         dt0 / (dt1 * (dt0 + dt1))]
    op, p = make_pair(dim, cells, G, tens, w0=w[0], w_hist=(w[1], w[2]))
    u = fields(p, 3)
    old = [fields(p, 4), fields(p, 5)]
    ref = po.residual(p, wtd, u, old)
    dst = op.initialize_dof_vector()
    op.evaluate_nonlinear_residual(dst, dev(u), wtd, old=[dev(o) for o in old])
    got = host(dst)
    assert rel_l2(got, ref) <= TOL64, rel_l2(got, ref)


def test_residual_golden_2cell_on_gpu(golden_dir):
    """tests/residual_generic_{scalar,tensorial}_no_rbm.mpirun=1.output"""
    import json
    import os
    fx = json.load(open(os.path.join(golden_dir, "residual_generic_2cell.json")))
    s = fx["setup"]
    u = np.zeros((6, 4))
    for key, vals in s["nodal_values"].items():
        x, y = map(int, key.split(","))
        u[x + 3 * y] = vals
    perm = [pt[0] + 3 * pt[1] for pt in s["dof_points"]]
    for kind in ("scalar", "tensorial"):
        fe = FreeEnergyParams(A=s["A"], B=s["B"], kappa_c=s["kappa_c"], kappa_p=s["kappa_p"])
        mob = MobilityParams(Mvol=s["Mvol"], Mvap=s["Mvap"], Msurf=s["Msurf"], Mgb=s["Mgb"],
                             L=s["L"], tensorial=kind == "tensorial")
        mesh = CartesianMesh((2, 1), (0.0, 0.0), (2.0, 1.0))
        op = SinteringOperatorGeneric(mesh, 2, fe, mob)
        op.set_time_weights(1.0 / s["dt"], (-1.0 / s["dt"],))
        dst = op.initialize_dof_vector()
        du = dev(u.ravel())
        op.evaluate_nonlinear_residual(dst, du, 2, old=[du])
        r = host(dst).reshape(6, 4)
        for b in range(4):
            got = [float("%.3e" % r[perm[i], b]) for i in range(6)]
            np.testing.assert_allclose(got, fx[kind][b], rtol=0, atol=0)


@pytest.mark.parametrize("dim,cells,G,tens", [
    (3, (9, 5, 17), 2, False), (3, (5, 9, 20), 2, True), (2, (65, 17), 3, False),
    (3, (10, 6, 19), 4, False), (3, (7, 5, 18), 12, False)])
def test_vmult_f32_error_per_field(dim, cells, G, tens):
    """FP32 vmult (FP32 cache, FP32 arithmetic) against the FP64 oracle;
    bound: per-field relative l2 error <= 1e-5"""
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
    (3, (9, 5, 17), 2, False), (3, (5, 9, 20), 2, True), (2, (65, 17), 2, False),
    (3, (7, 6, 9), 4, False), (3, (5, 4, 6), 8, False)])
def test_diagonal_and_node_blocks_match_oracle(dim, cells, G, tens):
    op, p = make_pair(dim, cells, G, tens, w0=1000.0)
    u = fields(p, 6)
    cache = po.fill_qp(p, u)
    blocks_ref = po.node_blocks(p, cache)
    op.fill_quadrature_point_values(dev(u))
    blocks = host(op.node_blocks()).reshape(blocks_ref.shape)
    assert rel_l2(blocks, blocks_ref) <= TOL64
    inv = op.initialize_dof_vector()
    op.compute_inverse_diagonal(inv)
    d = np.einsum("nii->ni", blocks_ref).ravel()
    inv_ref = np.where(np.abs(d) > 1e-10, 1.0 / d, 1.0)
    assert rel_l2(host(inv), inv_ref) <= TOL64
    # block-Jacobi apply = solve with each node block
    P = Preconditioner(op, "InverseComponentBlockDiagonalMatrix")
    P.do_update()
    x = np.random.default_rng(9).standard_normal(u.size)
    y = op.initialize_dof_vector()
    P.vmult(y, dev(x))
    yr = np.linalg.solve(blocks_ref, x.reshape(-1, p.n_comp)[..., None])[..., 0].ravel()
    assert rel_l2(host(y), yr) <= 1e-11


def test_blas1_and_layout_conversions():
    mesh = CartesianMesh((7, 5, 3), (0, 0, 0), (7.0, 5.0, 3.0))
    op = SinteringOperatorGeneric(mesh, 3)
    h = op.handle
    L = _ffi.lib()
    n = op.n_dofs
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    dx, dy = dev(x), dev(y)
    r = ctypes.c_double()
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _ffi.check(L.hps_dot(h, n, P(dx), P(dy), ctypes.byref(r)))
    assert abs(r.value - x @ y) <= 1e-12 * abs(x) @ abs(y)
    _ffi.check(L.hps_norm2(h, n, P(dx), ctypes.byref(r)))
    assert abs(r.value - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
    bn = (ctypes.c_double * 5)()
    _ffi.check(L.hps_block_norms(h, P(dx), bn))
    np.testing.assert_allclose(list(bn), np.linalg.norm(x.reshape(-1, 5), axis=0), rtol=1e-13)
    _ffi.check(L.hps_axpy(h, n, 2.5, P(dx), P(dy)))
    np.testing.assert_allclose(host(dy), y + 2.5 * x, rtol=1e-15, atol=1e-15)
    _ffi.check(L.hps_sadd(h, n, -0.5, 3.0, P(dx), P(dy)))
    np.testing.assert_allclose(host(dy), -0.5 * (y + 2.5 * x) + 3.0 * x, rtol=1e-14, atol=1e-14)
    _ffi.check(L.hps_scale(h, n, 4.0, P(dx)))
    np.testing.assert_allclose(host(dx), 4.0 * x, rtol=0, atol=0)
    # multi-dot and fused CGS update
    k = 7
    V = rng.standard_normal((k, n))
    w = rng.standard_normal(n)
    dV, dw = dev(V.ravel()), dev(w)
    hh = (ctypes.c_double * k)()
    _ffi.check(L.hps_multi_dot(h, n, k, P(dV), P(dw), hh))
    np.testing.assert_allclose(list(hh), V @ w, rtol=1e-12, atol=1e-12)
    _ffi.check(L.hps_multi_axpy_norm(h, n, k, P(dV), hh, P(dw), ctypes.byref(r)))
    wr = w - V.T @ (V @ w)
    np.testing.assert_allclose(host(dw), wr, rtol=1e-11, atol=1e-11)
    assert abs(r.value - np.linalg.norm(wr)) <= 1e-11 * np.linalg.norm(wr)
    # layout conversion
    blocks = [torch.empty(op.n_nodes, dtype=torch.float64, device="cuda") for _ in range(5)]
    arr = (ctypes.c_void_p * 5)(*[b.data_ptr() for b in blocks])
    _ffi.check(L.hps_to_blocks(h, arr, P(dev(x))))
    for c in range(5):
        np.testing.assert_array_equal(host(blocks[c]), x.reshape(-1, 5)[:, c])
    back = op.initialize_dof_vector()
    _ffi.check(L.hps_from_blocks(h, P(back), arr))
    np.testing.assert_array_equal(host(back), x)


@pytest.mark.parametrize("ty,tz", [(8, 16), (4, 5), (8, 64), (4, 128)])
def test_vmult_and_residual_tile_geometries(monkeypatch, ty, tz):
    """same results for every workgroup tile geometry"""
    monkeypatch.setenv("HPSINT_AMD_TILE_Y", str(ty))
    monkeypatch.setenv("HPSINT_AMD_TILE_M", str(tz))
    op, p = make_pair(3, (19, 13, 37), 4, False, w0=100.0, w_hist=(-100.0,))
    u = fields(p, 1)
    src = np.random.default_rng(2).uniform(-1, 1, size=u.size)
    ref = po.vmult(p, po.fill_qp(p, u), src)
    op.fill_quadrature_point_values(dev(u))
    dst = op.initialize_dof_vector()
    op.vmult(dst, dev(src))
    assert rel_l2(host(dst), ref) <= TOL64
    old = fields(p, 7)
    ref = po.residual(p, 2, u, [old])
    op.evaluate_nonlinear_residual(dst, dev(u), 2, old=[dev(old)])
    assert rel_l2(host(dst), ref) <= TOL64


def test_fp32_storage_blas_accumulates_in_fp64():
    mesh = CartesianMesh((7, 5, 3), (0, 0, 0), (7.0, 5.0, 3.0))
    op = SinteringOperatorGeneric(mesh, 3)
    h, L, n = op.handle, _ffi.lib(), op.n_dofs
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    rng = np.random.default_rng(1)
    k = 5
    V = rng.standard_normal((k, n)).astype(np.float32)
    w = rng.standard_normal(n).astype(np.float32)
    V64, w64 = V.astype(np.float64), w.astype(np.float64)
    dV, dw = dev(V.ravel(), torch.float32), dev(w, torch.float32)
    r = ctypes.c_double()
    _ffi.check(L.hps_norm2_f32(h, n, P(dw), ctypes.byref(r)))
    assert abs(r.value - np.linalg.norm(w64)) <= 1e-13 * np.linalg.norm(w64)
    hh = (ctypes.c_double * k)()
    _ffi.check(L.hps_multi_dot_f32(h, n, k, P(dV), P(dw), hh))
    np.testing.assert_allclose(list(hh), V64 @ w64, rtol=1e-12, atol=1e-12)
    _ffi.check(L.hps_multi_axpy_norm_f32(h, n, k, P(dV), hh, P(dw), ctypes.byref(r)))
    wr = (w64 - V64.T @ (V64 @ w64)).astype(np.float32)
    np.testing.assert_allclose(host(dw), wr, rtol=1e-6, atol=1e-6)
    assert abs(r.value - np.linalg.norm(wr.astype(np.float64))) <= 1e-12 * np.linalg.norm(wr)
    y = dev(np.zeros(n))
    a = (ctypes.c_double * k)(*[0.5, -1.0, 2.0, 0.25, 3.0])
    _ffi.check(L.hps_axpy_multi_f32(h, n, k, a, P(dV), P(y)))
    np.testing.assert_allclose(host(y), np.array(list(a)) @ V64, rtol=1e-14, atol=1e-14)
    _ffi.check(L.hps_scale_f32(h, n, 0.5, P(dw)))
    np.testing.assert_allclose(host(dw), wr * np.float32(0.5), rtol=0, atol=0)


@pytest.mark.parametrize("k", list(range(1, 33)))
def test_multi_vector_kernels_every_k(k):
    """each basis size k has its own kernel: FP64 and FP32-storage multi-dot,
    fused update + norm and the multi-axpy against numpy (n odd, so the
    unrolled steps leave a remainder)"""
    mesh = CartesianMesh((7, 5, 3), (0, 0, 0), (7.0, 5.0, 3.0))
    op = SinteringOperatorGeneric(mesh, 3)
    h, L = op.handle, _ffi.lib()
    P = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    n = 100003
    rng = np.random.default_rng(k)
    V = rng.standard_normal((k, n))
    w = rng.standard_normal(n)
    r = ctypes.c_double()
    hh = (ctypes.c_double * k)()
    dV, dw = dev(V.ravel()), dev(w)
    _ffi.check(L.hps_multi_dot(h, n, k, P(dV), P(dw), hh))
    np.testing.assert_allclose(list(hh), V @ w, rtol=1e-12, atol=1e-11)
    _ffi.check(L.hps_multi_axpy_norm(h, n, k, P(dV), hh, P(dw), ctypes.byref(r)))
    wr = w - V.T @ np.array(list(hh))
    np.testing.assert_allclose(host(dw), wr, rtol=1e-12, atol=1e-12)
    assert abs(r.value - np.linalg.norm(wr)) <= 1e-12 * np.linalg.norm(wr)
    V32, w32 = V.astype(np.float32), w.astype(np.float32)
    dV, dw = dev(V32.ravel(), torch.float32), dev(w32, torch.float32)
    _ffi.check(L.hps_multi_dot_f32(h, n, k, P(dV), P(dw), hh))
    ref = V32.astype(np.float64) @ w32.astype(np.float64)
    np.testing.assert_allclose(list(hh), ref, rtol=1e-12, atol=1e-11)
    _ffi.check(L.hps_multi_axpy_norm_f32(h, n, k, P(dV), hh, P(dw), ctypes.byref(r)))
    wr = (w32.astype(np.float64) - V32.astype(np.float64).T @ np.array(list(hh))).astype(np.float32)
    np.testing.assert_allclose(host(dw), wr, rtol=1e-6, atol=1e-6)
    assert abs(r.value - np.linalg.norm(wr.astype(np.float64))) <= 1e-12 * np.linalg.norm(wr)
    a = (ctypes.c_double * k)(*rng.standard_normal(k))
    y = dev(np.ones(n))
    _ffi.check(L.hps_axpy_multi(h, n, k, a, P(dev(V.ravel())), P(y)))
    np.testing.assert_allclose(host(y), 1.0 + np.array(list(a)) @ V, rtol=1e-13, atol=1e-13)
    y = dev(np.ones(n))
    _ffi.check(L.hps_axpy_multi_f32(h, n, k, a, P(dV), P(y)))
    np.testing.assert_allclose(host(y), 1.0 + np.array(list(a)) @ V32.astype(np.float64),
                               rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dim,cells,G,tens", [
    (3, (9, 5, 17), 2, False), (3, (7, 6, 9), 4, False), (3, (5, 9, 20), 2, True),
    (2, (65, 17), 2, False)])
def test_fp32_preconditioner_assembly(dim, cells, G, tens):
    """diagonal and node blocks assembled from the FP32 cache in FP32:
    relative l2 error of the blocks against the FP64 oracle <= 1e-5, and the
    FP32 block-Jacobi apply solves with the oracle's blocks to 1e-5"""
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
    x = np.random.default_rng(9).standard_normal(u.size)
    y = op.initialize_dof_vector(torch.float32)
    P.vmult(y, dev(x, torch.float32))
    yr = np.linalg.solve(blocks_ref, x.reshape(-1, nc)[..., None])[..., 0].ravel()
    assert rel_l2(host(y), yr) <= 1e-5
    D = Preconditioner(op, "InverseDiagonalMatrix")
    D.do_update(fp32=True)
    inv = op.initialize_dof_vector()
    op_inv = np.where(np.abs(np.einsum("nii->ni", blocks_ref).ravel()) > 1e-10,
                      1.0 / np.einsum("nii->ni", blocks_ref).ravel(), 1.0)
    _ffi.check(_ffi.lib().hps_precond_get_diagonal(op.handle, ctypes.c_void_p(inv.data_ptr())))
    assert rel_l2(host(inv), op_inv) <= 1e-5


def boundary_and_random_nodes(p, seed, k=40):
    """the nodes of the first mesh face plus k random nodes, distinct"""
    nn = po.n_nodes(p)
    n0 = p.n_cells[0] + 1 if p.degree == 1 else 2 * p.n_cells[0] + 1
    face = np.arange(0, nn, n0)                      # x = 0 face
    rnd = np.random.default_rng(seed).choice(nn, size=min(k, nn), replace=False)
    return np.unique(np.concatenate([face, rnd]))


@pytest.mark.parametrize("degree,dim,cells,G,tens", [
    (1, 3, (9, 5, 17), 2, False), (1, 3, (11, 6, 18), 3, True), (1, 2, (65, 17), 2, False),
    (2, 3, (5, 4, 7), 2, False), (2, 3, (4, 3, 5), 3, True)])
def test_constrained_nodes_vmult_residual_preconditioner(degree, dim, cells, G, tens):
    """S/operator_base.h:294-353: zeroed constrained src entries, identity rows,
    src restored bitwise; residual rows of constrained nodes zero; diagonal 1
    and identity node blocks on them"""
    if degree == 1:
        op, p = make_pair(dim, cells, G, tens, w0=100.0, w_hist=(-100.0,))
    else:
        h = (0.5, 0.75, 1.0)[:dim]
        fe, mob = FreeEnergyParams(), MobilityParams(tensorial=tens)
        mesh = CartesianMesh(tuple(cells), (0.0,) * dim,
                             tuple(n * hh for n, hh in zip(cells, h)))
        op = SinteringOperatorGeneric(mesh, G, fe, mob, fe_degree=2)
        op.set_time_weights(100.0, (-100.0,))
        p = po.make_params(dim, 2, cells, h, G + 2, tensorial=tens, A=fe.A, B=fe.B,
                           kappa_c=fe.kappa_c, kappa_p=fe.kappa_p, Mvol=mob.Mvol,
                           Mvap=mob.Mvap, Msurf=mob.Msurf, Mgb=mob.Mgb, L=mob.L, w0=100.0,
                           w_hist=(-100.0,))
    nc = p.n_comp
    nodes = boundary_and_random_nodes(p, seed=5)
    u = fields(p, 1)
    src = np.random.default_rng(2).uniform(-1, 1, size=u.size)
    cache = po.fill_qp(p, u)
    op.fill_quadrature_point_values(dev(u))
    op.set_constrained_nodes(nodes)
    s = dev(src)
    dst = op.initialize_dof_vector()
    op.vmult(dst, s)
    ref = po.vmult_constrained(p, cache, src, nodes)
    got = host(dst)
    assert rel_l2(got, ref) <= TOL64, rel_l2(got, ref)
    assert np.array_equal(host(s), src)                      # src restored
    assert np.array_equal(got.reshape(-1, nc)[nodes], src.reshape(-1, nc)[nodes])
    # FP32
    s32 = dev(src, torch.float32)
    d32 = op.initialize_dof_vector(torch.float32)
    op.vmult(d32, s32)
    assert np.array_equal(host(s32), src.astype(np.float32).astype(np.float64))
    g32 = host(d32).reshape(-1, nc)
    r2 = ref.reshape(-1, nc)
    assert max(rel_l2(g32[:, c], r2[:, c]) for c in range(nc)) <= 1e-5
    # residual
    old = fields(p, 3)
    r = op.initialize_dof_vector()
    op.evaluate_nonlinear_residual(r, dev(u), 2, old=[dev(old)])
    rr = po.residual_constrained(p, 2, u, [old], nodes)
    assert rel_l2(host(r), rr) <= TOL64
    # preconditioners: identity on the constrained nodes
    inv = op.initialize_dof_vector()
    op.compute_inverse_diagonal(inv)
    assert np.all(host(inv).reshape(-1, nc)[nodes] == 1.0)
    B = host(op.node_blocks()).reshape(-1, nc, nc)
    assert np.array_equal(B[nodes], np.broadcast_to(np.eye(nc), (nodes.size, nc, nc)))
    # removing the constraints restores the plain operator
    op.set_constrained_nodes([])
    op.vmult(dst, s)
    assert rel_l2(host(dst), po.vmult(p, cache, src)) <= TOL64


def test_constrained_nodes_reject_bad_lists():
    op, p = make_pair(3, (3, 3, 3), 2)
    nn = po.n_nodes(p)
    for bad in ([0, 0], [-1], [nn]):
        with pytest.raises(RuntimeError):
            op.set_constrained_nodes(bad)


def test_every_grain_count_is_built():
    """hps_create accepts every grain count 1..12 for Q1 and Q2, 2D and 3D,
    scalar and tensorial mobility (EXPAND_OPERATIONS instantiates all of
    them, S/instantiation.h:97-152); 13 order parameters are refused with
    ExcInvalidNumberOfComponents"""
    for dim in (2, 3):
        mesh = CartesianMesh((3,) * dim, (0.0,) * dim, (3.0,) * dim)
        for deg in (1, 2):
            for tens in (False, True):
                for G in range(1, 13):
                    SinteringOperatorGeneric(mesh, G, FreeEnergyParams(),
                                             MobilityParams(tensorial=tens), fe_degree=deg)
                with pytest.raises(_ffi.InvalidNumberOfComponents):
                    SinteringOperatorGeneric(mesh, 13, FreeEnergyParams(),
                                             MobilityParams(tensorial=tens), fe_degree=deg)


@pytest.mark.parametrize("zcol", ["0", "1"])
@pytest.mark.parametrize("cells,G", [((19, 13, 37), 4), ((9, 5, 17), 2), ((3, 7, 2), 3),
                                     ((8, 12, 16), 1), ((17, 9, 33), 4), ((11, 9, 18), 6),
                                     ((19, 9, 20), 8), ((10, 17, 9), 12)])
def test_vmult_zcol_and_corner_lanes(monkeypatch, zcol, cells, G):
    """the 3D scalar tangent with z-column lanes (8 x 8 tiles; FP64 with up
    to 4 order parameters, FP32 with up to 12) and with one lane per cell
    corner give the oracle's vmult, FP64 at 1e-12 and FP32 per field at
    1e-5"""
    monkeypatch.setenv("HPSINT_AMD_ZCOL", zcol)
    op, p = make_pair(3, cells, G, False, w0=100.0)
    u = fields(p, 3)
    src = np.random.default_rng(4).uniform(-1, 1, size=u.size)
    ref = po.vmult(p, po.fill_qp(p, u), src)
    op.fill_quadrature_point_values(dev(u))
    dst = op.initialize_dof_vector()
    op.vmult(dst, dev(src))
    assert rel_l2(host(dst), ref) <= TOL64, rel_l2(host(dst), ref)
    d32 = op.initialize_dof_vector(torch.float32)
    op.vmult(d32, dev(src, torch.float32))
    g, r = host(d32).reshape(-1, G + 2), ref.reshape(-1, G + 2)
    errs = [rel_l2(g[:, c], r[:, c]) for c in range(G + 2)]
    assert max(errs) <= 1e-5, errs


@pytest.mark.parametrize("march", ["1", "2", "3"])
def test_zcol_march_depths(monkeypatch, march):
    """FP64 tangent and residual with z-column marches of one and of three
    cache tiles (16 layers each): 37 cells along z end inside a march and
    inside a cache tile; the oracle's vmult and residual at 1e-12 (no floor
    on the workgroups per CU, so the small grid marches that deep)"""
    monkeypatch.setenv("HPSINT_AMD_ZCOL_MARCH", march)
    monkeypatch.setenv("HPSINT_AMD_ZCOL_MARCH_WG", "0")
    cells, G = (19, 13, 37), 4
    op, p = make_pair(3, cells, G, False, w0=100.0, w_hist=(-100.0,))
    u = fields(p, 3)
    src = np.random.default_rng(4).uniform(-1, 1, size=u.size)
    ref = po.vmult(p, po.fill_qp(p, u), src)
    op.fill_quadrature_point_values(dev(u))
    dst = op.initialize_dof_vector()
    op.vmult(dst, dev(src))
    assert rel_l2(host(dst), ref) <= TOL64, rel_l2(host(dst), ref)
    olds = [fields(p, 6)]
    ref = po.residual(p, 2, u, olds)
    op.evaluate_nonlinear_residual(dst, dev(u), 2, old=[dev(o) for o in olds])
    assert rel_l2(host(dst), ref) <= TOL64, rel_l2(host(dst), ref)


@pytest.mark.parametrize("zcol", ["0", "1"])
@pytest.mark.parametrize("cells,G", [((19, 13, 37), 4), ((9, 5, 17), 2), ((3, 7, 2), 3),
                                     ((8, 12, 16), 1)])
def test_residual_zcol_and_corner_lanes(monkeypatch, zcol, cells, G):
    """the 3D scalar residual with z-column lanes and with one lane per cell
    corner give the oracle's residual at 1e-12 for every wtd, BDF2 history
    with unequal steps"""
    monkeypatch.setenv("HPSINT_AMD_ZCOL_RES", zcol)
    dt0, dt1 = 1e-2, 2e-2
    w1 = -(dt0 + dt1) / (dt0 * dt1)
    w2 = dt0 / (dt1 * (dt0 + dt1))
    w0 = -(w1 + w2)
    op, p = make_pair(3, cells, G, False, w0=w0, w_hist=(w1, w2))
    u = fields(p, 5)
    olds = [fields(p, 6), fields(p, 7)]
    for wtd in (0, 1, 2):
        ref = po.residual(p, wtd, u, olds)
        dst = op.initialize_dof_vector()
        op.evaluate_nonlinear_residual(dst, dev(u), wtd, old=[dev(o) for o in olds])
        assert rel_l2(host(dst), ref) <= TOL64, (wtd, rel_l2(host(dst), ref))
