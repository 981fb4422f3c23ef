"""The GPU's BlockPreconditioner2 (ILU(0) of the Cahn-Hilliard block in
parity-colour order, inverse diagonal on the order parameters) against its
restatement on the oracle's tangent (tests/precond_study.py): FP64 apply at
1e-12, FP32 apply at 1e-5, on random fields (scalar and tensorial
mobility) and on the 8-sphere packing's fields at the 128^3 cell size; and
a Newton step with it satisfies the oracle's residual."""
import numpy as np
import pytest
import torch

import config_fields as cf
import precond_study as ps
import pyoracle as po

pytestmark = pytest.mark.gpu

from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,  # noqa: E402
                        Preconditioner, SinteringOperatorGeneric, packing)


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def dev(x, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype)


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def check(mesh, G, u, dt=1e-3, seed=3, tens=False):
    dim = mesh.dim
    op = SinteringOperatorGeneric(mesh, G, FreeEnergyParams(), MobilityParams(tensorial=tens))
    op.set_time_weights(1.0 / dt, (-1.0 / dt,))
    p = po.make_params(dim, 1, mesh.n_cells, mesh.cell_size, G + 2, tensorial=tens,
                       w0=1.0 / dt, w_hist=(-1.0 / dt,))
    A = ps.sparse_tangent(p, po.fill_qp(p, u), keep_zeros=True)
    ref_apply = ps.parity_colour_ilu(A, G + 2, [n + 1 for n in mesh.n_cells])
    op.fill_quadrature_point_values(dev(u))
    P = Preconditioner(op, "BlockPreconditioner2")
    P.do_update()
    x = np.random.default_rng(seed).standard_normal(u.size)
    y = op.initialize_dof_vector()
    P.vmult(y, dev(x))
    ref = ref_apply(x)
    assert rel_l2(host(y), ref) <= 1e-12, rel_l2(host(y), ref)
    y32 = op.initialize_dof_vector(torch.float32)
    P.vmult(y32, dev(x, torch.float32))
    nc = G + 2
    g, r = host(y32).reshape(-1, nc), ref.reshape(-1, nc)
    errs = [rel_l2(g[:, c], r[:, c]) for c in range(nc)]
    assert max(errs) <= 1e-5, errs
    # deterministic
    y2 = op.initialize_dof_vector()
    P.vmult(y2, dev(x))
    assert np.array_equal(host(y2), host(y))


@pytest.mark.parametrize("cells,G", [((9, 5, 7), 2), ((6, 7, 5), 4), ((13, 9), 2), ((8, 11), 3),
                                     ((5, 6, 4), 8), ((4, 5, 3), 12)])
def test_ch_ilu_random_fields(cells, G):
    dim = len(cells)
    h = (0.5, 0.75, 1.0)[:dim]
    mesh = CartesianMesh(cells, (0.0,) * dim, tuple(n * hh for n, hh in zip(cells, h)))
    nn = int(np.prod([n + 1 for n in cells]))
    rng = np.random.default_rng(G)
    u = rng.uniform(0, 1, size=(nn, G + 2))
    u[:, 1] = rng.uniform(-1, 1, nn)
    check(mesh, G, u.ravel())


@pytest.mark.parametrize("cells,G", [((9, 5, 7), 2), ((6, 7, 5), 4), ((13, 9), 3),
                                     ((5, 6, 4), 7), ((4, 5, 3), 12)])
def test_ch_ilu_tensorial_mobility(cells, G):
    """the Cahn-Hilliard block of the tensorial tangent: its c row's flux
    coefficients found by evaluating the tangent at unit inputs"""
    dim = len(cells)
    h = (0.5, 0.75, 1.0)[:dim]
    mesh = CartesianMesh(cells, (0.0,) * dim, tuple(n * hh for n, hh in zip(cells, h)))
    nn = int(np.prod([n + 1 for n in cells]))
    rng = np.random.default_rng(10 + G)
    u = rng.uniform(0, 1, size=(nn, G + 2))
    u[:, 1] = rng.uniform(-1, 1, nn)
    check(mesh, G, u.ravel(), tens=True)


@pytest.mark.parametrize("cells,G,tens", [((9, 5, 7), 2, False), ((13, 9), 3, False),
                                          ((6, 7, 5), 4, True)])
def test_ch_ilu_with_constrained_nodes(cells, G, tens):
    """constrained nodes (x = 0 face plus random nodes): the ILU(0) of the
    constrained operator's Cahn-Hilliard block, whose constrained rows and
    columns are identity rows and columns; dst = src on those nodes"""
    import scipy.sparse as sps
    dim = len(cells)
    h = (0.5, 0.75, 1.0)[:dim]
    dt = 1e-3
    mesh = CartesianMesh(cells, (0.0,) * dim, tuple(n * hh for n, hh in zip(cells, h)))
    nn = int(np.prod([n + 1 for n in cells]))
    nc = G + 2
    rng = np.random.default_rng(20 + G)
    u = rng.uniform(0, 1, size=(nn, nc))
    u[:, 1] = rng.uniform(-1, 1, nn)
    u = u.ravel()
    nodes = np.unique(np.concatenate([np.arange(0, nn, cells[0] + 1),
                                      rng.choice(nn, size=15, replace=False)]))
    op = SinteringOperatorGeneric(mesh, G, FreeEnergyParams(), MobilityParams(tensorial=tens))
    op.set_time_weights(1.0 / dt, (-1.0 / dt,))
    p = po.make_params(dim, 1, mesh.n_cells, mesh.cell_size, nc, tensorial=tens,
                       w0=1.0 / dt, w_hist=(-1.0 / dt,))
    A = ps.sparse_tangent(p, po.fill_qp(p, u), keep_zeros=True)
    mask = np.zeros(nn * nc)
    mask[(nodes[:, None] * nc + np.arange(nc)).ravel()] = 1.0
    D, P = sps.diags(1.0 - mask), sps.diags(mask)
    Ac = (D @ A @ D + P).tocsr()
    ref_apply = ps.parity_colour_ilu(Ac, nc, [n + 1 for n in cells])
    op.fill_quadrature_point_values(dev(u))
    op.set_constrained_nodes(nodes)
    P2 = Preconditioner(op, "BlockPreconditioner2")
    P2.do_update()
    x = np.random.default_rng(7).standard_normal(u.size)
    y = op.initialize_dof_vector()
    P2.vmult(y, dev(x))
    got, ref = host(y), ref_apply(x)
    assert rel_l2(got, ref) <= 1e-12, rel_l2(got, ref)
    assert np.array_equal(got.reshape(-1, nc)[nodes], x.reshape(-1, nc)[nodes])


def test_ch_ilu_on_the_packing():
    big = packing.packing_cfg1(seed=0)
    h = (big.upper[0] - big.lower[0]) / 128
    n = 14
    pk = cf.sub_packing(big, cf.contact_point(big), n, h)
    u = cf.state(pk, (n, n, n), 4)
    mesh = CartesianMesh((n, n, n), pk.lower, pk.upper)
    check(mesh, 4, u)


@pytest.mark.parametrize("inner_fp32", [0, 1])
def test_newton_step_with_block_preconditioner(inner_fp32):
    from hpsint_amd.solvers import NewtonSolver
    pk = packing.two_particles_2d()
    cells = (160, 80)
    mesh = CartesianMesh(cells, pk.lower, pk.upper)
    op = SinteringOperatorGeneric(mesh, 2, FreeEnergyParams(), MobilityParams())
    u0 = packing.initial_fields(pk, cells)
    dt = 1e-3
    op.set_time_weights(1.0 / dt, (-1.0 / dt,))
    solver = NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2,
                          preconditioner="BlockPreconditioner2", inner_fp32=inner_fp32)
    old = dev(u0)
    x = old.clone()
    st = solver.solve(x, [old])
    assert st.converged
    p = po.make_params(2, 1, cells, mesh.cell_size, 4, w0=1.0 / dt, w_hist=(-1.0 / dt,))
    r0 = np.linalg.norm(po.residual(p, 2, u0, [u0]))
    r = np.linalg.norm(po.residual(p, 2, host(x), [u0]))
    assert r <= 2e-8 * r0, (r, r0)
