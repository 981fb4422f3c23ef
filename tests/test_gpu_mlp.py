"""The predictor's BF16-MFMA kernels against the numpy restatement.

- every GEMM variant (transposes, epilogues, split-K) on ragged sizes
  against float64 products of the same bf16 inputs (FP32 accumulation:
  relative error <= 1e-5; bf16 outputs within 1 bf16 ulp);
- forward, loss gradient and all weight/bias gradients of the full network
  against oracle/mlp_oracle.py with the same bf16 roundings (relative l2
  error <= 5e-3: the residual difference is FP32 accumulation order and the
  occasional 1-ulp flip of a bf16 activation);
- the stencil-feature kernel bit-exactly against a numpy restatement;
- training drives the loss down, and a trained predictor saves Newton
  iterations over the plain method at the same tolerance.
"""
import ctypes

import numpy as np
import pytest
import torch

import mlp_oracle as mo

pytestmark = pytest.mark.gpu

from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,  # noqa: E402
                        SinteringOperatorGeneric, packing)
from hpsint_amd._ffi import check, lib  # noqa: E402
from hpsint_amd.predictor import MLP, LearnedPredictor  # noqa: E402


def small_op():
    mesh = CartesianMesh((4, 4, 4), (0, 0, 0), (4.0, 4.0, 4.0))
    return SinteringOperatorGeneric(mesh, 2)


def tbf(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda().to(torch.bfloat16)


def P(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def as64(t):
    return t.float().cpu().numpy().astype(np.float64)


CASES = [  # ta, tb, epi, M, N, K, split
    (0, 1, 0, 100, 70, 50, 1), (0, 1, 1, 130, 128, 272, 1), (0, 1, 2, 77, 4, 128, 1),
    (0, 0, 0, 65, 33, 40, 1), (0, 0, 3, 100, 272, 128, 1),
    (1, 0, 0, 128, 272, 3000, 7), (1, 0, 0, 4, 128, 1000, 3), (1, 1, 0, 40, 50, 60, 2),
]


@pytest.mark.parametrize("ta,tb,epi,M,N,K,split", CASES)
def test_gemm_variants(ta, tb, epi, M, N, K, split):
    op = small_op()
    rng = np.random.default_rng(M + N + K)
    A = mo.bf16(rng.uniform(-1, 1, (K, M) if ta else (M, K)))
    B = mo.bf16(rng.uniform(-1, 1, (N, K) if tb else (K, N)))
    Am = A.T if ta else A
    Bm = B.T if tb else B
    ref = Am @ Bm
    bias = rng.uniform(-0.5, 0.5, N).astype(np.float32)
    H = mo.bf16(rng.uniform(-1, 1, (M, N)))
    dA, dB, dbias, dH = tbf(A), tbf(B), torch.from_numpy(bias).cuda(), tbf(H)
    lda = A.shape[1]
    ldb = B.shape[1]
    if epi == 0:
        C = torch.zeros(split * M * N, dtype=torch.float32, device="cuda")
    elif epi == 2:
        C = torch.zeros(M * N, dtype=torch.float32, device="cuda")
    else:
        C = torch.zeros(M * N, dtype=torch.bfloat16, device="cuda")
    check(lib().hps_mlp_gemm(small_op().handle if False else op.handle, ta, tb, epi, M, N, K,
                             P(dA), lda, P(dB), ldb, P(C), N, P(dbias), P(dH), split))
    torch.cuda.synchronize()
    if epi == 0:
        got = C.view(split, M, N).sum(dim=0).cpu().numpy().astype(np.float64)
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    elif epi == 2:
        got = C.view(M, N).cpu().numpy().astype(np.float64)
        np.testing.assert_allclose(got, ref + bias, rtol=1e-5, atol=1e-5 * np.abs(ref).max())
    else:
        got = as64(C.view(M, N))
        want = np.tanh(ref + bias) if epi == 1 else ref * (1 - H ** 2)
        ulp = np.maximum(np.abs(want), 1e-30) * 2.0 ** -7
        assert np.all(np.abs(got - mo.bf16(want)) <= ulp + 1e-6 * np.abs(ref).max())


def test_forward_backward_match_oracle():
    op = small_op()
    mlp = MLP(op, seed=3)
    # larger last layer so the gradients are not tiny
    mlp.W[-1].normal_(0, 0.1)
    mlp.Wb[-1].copy_(mlp.W[-1].to(torch.bfloat16))
    for b in mlp.b:
        b.uniform_(-0.1, 0.1)
    M = 3000
    rng = np.random.default_rng(4)
    X = mo.bf16(rng.uniform(-1, 1, (M, 272)))
    T = rng.uniform(-1, 1, (M, 4)).astype(np.float32)
    dX, dT = tbf(X), torch.from_numpy(T).cuda()
    Y, acts = mlp.forward(dX, M)
    loss, dY = mlp.loss_grad(Y, dT, M)
    gW, gb = mlp.backward(acts, dY, M)
    torch.cuda.synchronize()
    Ws = [as64(w) for w in mlp.Wb]
    bs = [b.cpu().numpy().astype(np.float64) for b in mlp.b]
    Yr, actsr = mo.forward(Ws, bs, X)
    lr_, dYr = mo.loss_and_grad(Yr, T.astype(np.float64))
    gWr, gbr = mo.backward(Ws, actsr, dYr)

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)

    assert rel(Y.cpu().numpy(), Yr) < 5e-3
    assert abs(loss - lr_) <= 5e-3 * lr_
    for l in range(len(Ws)):
        assert rel(gW[l].cpu().numpy(), gWr[l]) < 5e-3, l
        assert rel(gb[l].cpu().numpy(), gbr[l]) < 5e-3, l


def features_numpy(u_prev, u_cur, nn, nc, dt, dt_scale, inc_scale):
    G = nc - 2
    up = u_prev.reshape(nn[2], nn[1], nn[0], nc)
    uc = u_cur.reshape(nn[2], nn[1], nn[0], nc)
    n = nn[0] * nn[1] * nn[2]
    X = np.zeros((n, 272))
    ls_all = np.zeros(n)
    top = np.zeros((n, 2), dtype=np.int64)
    node = 0
    for z in range(nn[2]):
        for y in range(nn[1]):
            for x in range(nn[0]):
                e = uc[z, y, x, 2:]
                order = sorted(range(G), key=lambda i: (-e[i], i))
                i1, i2 = order[0], (order[1] if G > 1 else order[0])
                top[node] = (i1, i2)
                s = 0
                inc = np.zeros((27, 5))
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            zz = min(max(z + dz, 0), nn[2] - 1)
                            yy = min(max(y + dy, 0), nn[1] - 1)
                            xx = min(max(x + dx, 0), nn[0] - 1)
                            a, b = uc[zz, yy, xx], up[zz, yy, xx]
                            fa = [a[0], a[1], np.sum(a[2:] ** 2), a[2 + i1], a[2 + i2]]
                            fb = [b[0], b[1], np.sum(b[2:] ** 2), b[2 + i1], b[2 + i2]]
                            for f in range(5):
                                X[node, s * 10 + f] = fa[f]
                                inc[s, f] = fa[f] - fb[f]
                            s += 1
                ls = np.sqrt(np.sum(inc ** 2) / 135.0) + 1e-3 / inc_scale
                for q in range(27):
                    X[node, q * 10 + 5:q * 10 + 10] = inc[q] / ls
                X[node, 270] = dt * dt_scale
                X[node, 271] = np.log10(ls * inc_scale)
                ls_all[node] = ls
                node += 1
    return X, top, ls_all


def test_training_reduces_loss():
    op = small_op()
    mlp = MLP(op, seed=0)
    M = 8192
    rng = np.random.default_rng(6)
    X = mo.bf16(rng.uniform(-1, 1, (M, 272)))
    # a smooth target the network can represent
    Wt = rng.standard_normal((272, 4)) / 16
    T = np.tanh(X @ Wt).astype(np.float32)
    dX, dT = tbf(X), torch.from_numpy(T).cuda()
    first = mlp.train_step(dX, dT, M, 1e-3)
    for _ in range(300):
        last = mlp.train_step(dX, dT, M, 1e-3)
    assert last < 0.1 * first, (first, last)


def test_predictor_saves_newton_iterations():
    """2D two-particle problem: 16 plain steps for training, 6 held out"""
    from hpsint_amd.predictor_training import collect_trajectory, fit_predictor
    from hpsint_amd.solvers import NewtonSolver, TimeStepper
    pk = packing.two_particles_2d()
    cells = (160, 80)
    mesh = CartesianMesh(cells, pk.lower, pk.upper)
    op = SinteringOperatorGeneric(mesh, 2, FreeEnergyParams(), MobilityParams())
    u0 = torch.from_numpy(packing.initial_fields(pk, cells)).cuda()
    traj = collect_trajectory(op, u0, n_steps=22, dt=1e-3, nl_rel_tol=1e-8)
    pred = fit_predictor(op, traj[:17], dt=1e-3, iterations=400, rows_per_step=8000, seed=0)

    def run(predictor):
        solver = NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2)
        ts = TimeStepper(op, solver, order=1, dt=1e-3, predictor=predictor)
        ts.initialize(traj[16])
        ts.history = [traj[16].clone(), traj[15].clone()]
        return [ts.make_step().newton_iterations for _ in range(5)]

    plain = run(None)
    learned = run(pred)
    assert sum(learned) < sum(plain), (plain, learned)


@pytest.mark.parametrize("dim", [2, 3])
def test_features_kernel_matches_numpy(dim):
    """hps_mlp_features against features_numpy on a packing state: 2D on the
    two-particle fields, 3D on a sub-box of the 100-sphere packing (12 order
    parameters); bf16 features within one bf16 rounding, top-2 order
    parameters exact, local scales to 1e-6"""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import config_fields as cf
    if dim == 2:
        pk = packing.two_particles_2d()
        cells, G = (12, 7), 2
        pk = packing.Packing(pk.centers, pk.radii, (-8.0, -3.5), (4.0, 3.5), pk.colors)
        u_cur = packing.initial_fields(pk, cells)
    else:
        big = packing.packing_cfg4(seed=0)
        cells, G = (6, 5, 7), 12
        pk = cf.sub_packing(big, cf.contact_point(big), 6, 1.0)
        pk = packing.Packing(pk.centers, pk.radii, pk.lower,
                             (pk.upper[0], pk.upper[1] - 1.0, pk.upper[2] + 1.0), pk.colors)
        u_cur = cf.state(pk, cells, G)
    nc = G + 2
    mesh = CartesianMesh(cells, pk.lower, pk.upper)
    op = SinteringOperatorGeneric(mesh, G)
    rng = np.random.default_rng(11)
    u_prev = u_cur + 0.01 * rng.standard_normal(u_cur.size) * (np.abs(u_cur) > 1e-3)
    pred = LearnedPredictor(op, MLP(op), dt_scale=1e3)
    dp, dc = torch.from_numpy(u_prev).cuda(), torch.from_numpy(u_cur).cuda()
    X, top2 = pred.features(dp, dc, 1e-3)
    torch.cuda.synchronize()
    nn = [n + 1 for n in cells] + [1] * (3 - dim)
    Xr, topr, lsr = features_numpy(u_prev, u_cur, nn, nc, 1e-3, 1e3, pred.inc_scale)
    np.testing.assert_array_equal(top2.cpu().numpy(), topr)
    np.testing.assert_allclose(pred.lscale.cpu().numpy(), lsr, rtol=1e-6)
    got = as64(X)
    want = mo.bf16(Xr)
    ulp = np.maximum(np.abs(want), 1e-30) * 2.0 ** -7
    assert np.all(np.abs(got - want) <= ulp), np.abs(got - want).max()


def test_predictor_saves_newton_iterations_3d():
    """3D: a 40^3 sub-box of the 100-sphere packing (12 order parameters,
    cell size 1): 12 plain steps train the network, the next 4 steps are
    solved from both initial guesses at Newton rtol 1e-8 (relative to the
    plain method's initial residual in both runs)"""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import config_fields as cf
    from hpsint_amd.predictor_training import collect_trajectory, fit_predictor
    from hpsint_amd.solvers import NewtonSolver, TimeStepper
    big = packing.packing_cfg4(seed=0)
    n, G, dt = 40, 12, 1e-3
    pk = cf.sub_packing(big, cf.contact_point(big), n, 1.0)
    mesh = CartesianMesh((n, n, n), pk.lower, pk.upper)
    op = SinteringOperatorGeneric(mesh, G, FreeEnergyParams(), MobilityParams())
    u0 = torch.from_numpy(packing.initial_fields(pk, (n, n, n), n_order_params=G)).cuda()
    traj = collect_trajectory(op, u0, n_steps=16, dt=dt, nl_rel_tol=1e-8)
    pred = fit_predictor(op, traj[:13], dt=dt, iterations=1500,
                         rows_per_step=min(32768, op.n_nodes), seed=0,
                         column_weights=(3.0, 1.0, 1.0, 1.0))

    def run(predictor):
        solver = NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2)
        ts = TimeStepper(op, solver, order=1, dt=dt, predictor=predictor)
        ts.initialize(traj[12])
        ts.history = [traj[12].clone(), traj[11].clone()]
        return [ts.make_step().newton_iterations for _ in range(4)]

    plain = run(None)
    learned = run(pred)
    print("3D Newton iterations plain", plain, "learned", learned)
    assert sum(learned) < sum(plain), (plain, learned)
