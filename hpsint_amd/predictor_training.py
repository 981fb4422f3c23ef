"""Training data and fitting of the learned Newton predictor.

Trajectories come from the plain solver (Newton from u_n at the same
tolerance).  For every consecutive triple (u_{n-1}, u_n, u_{n+1}) the
library computes the stencil features of (u_{n-1}, u_n) and the targets
(u_{n+1} - (2 u_n - u_{n-1})) / out_scale on the 4 outputs; a subset of
nodes (the most active half by target magnitude plus a uniform sample) forms
the training set.  Fitting is Adam on mini-batches, every product on the
library's MFMA kernels.  torch only holds memory and selects rows.
"""
import math

import torch

from .predictor import MLP, N_IN, N_OUT, LearnedPredictor
from .solvers import NewtonSolver, TimeStepper


def collect_trajectory(op, u0, n_steps, dt, nl_rel_tol=1e-8, l_rel_tol=1e-2,
                       preconditioner="InverseComponentBlockDiagonalMatrix", records=None):
    solver = NewtonSolver(op, nl_rel_tol=nl_rel_tol, l_rel_tol=l_rel_tol,
                          preconditioner=preconditioner)
    ts = TimeStepper(op, solver, order=1, dt=dt)
    ts.initialize(u0)
    traj = [u0.clone()]
    for _ in range(n_steps):
        rec = ts.make_step()
        if records is not None:
            records.append(rec)
        traj.append(ts.u.clone())
    return traj


def build_dataset(pred: LearnedPredictor, traj, dt, rows_per_step, seed=0,
                  column_weights=(1.0, 1.0, 1.0, 1.0)):
    """rows of (features, targets) from every triple of traj (one trajectory,
    or a list of trajectories of the same mesh); sets pred.out_scale to the
    RMS of each target column divided by its weight and normalises by it (a
    column of weight w counts w^2 in the loss)"""
    op = pred.op
    n = op.n_nodes
    dev = pred.X.device
    g = torch.Generator(device="cpu").manual_seed(seed)
    pred.out_scale.fill_(1.0)
    Xs, Ts = [], []
    T = torch.empty(n, N_OUT, dtype=torch.float32, device=dev)
    trajs = traj if isinstance(traj[0], (list, tuple)) else [traj]
    triples = [(t[i - 1], t[i], t[i + 1]) for t in trajs for i in range(1, len(t) - 1)]
    for (u_prev, u_cur, u_next) in triples:
        X, _ = pred.features(u_prev, u_cur, dt)
        pred.targets(u_prev, u_cur, u_next, T)
        k = min(rows_per_step // 2, n)
        act = T.abs().sum(dim=1)
        top = torch.topk(act, k).indices
        rnd = torch.randint(0, n, (rows_per_step - k,), generator=g).to(dev)
        idx = torch.cat([top, rnd])
        Xs.append(X.index_select(0, idx))
        Ts.append(T.index_select(0, idx))
    X, T = torch.cat(Xs), torch.cat(Ts)
    scale = T.double().pow(2).mean(dim=0).sqrt().clamp_min(1e-12)
    scale = (scale / torch.tensor(column_weights, dtype=torch.float64, device=dev)).float()
    T /= scale
    pred.out_scale.copy_(scale)
    return X, T


def fit_predictor(op, traj, dt, iterations=1000, rows_per_step=32768, batch=16384,
                  lr=1e-3, seed=0, log=None, skip=2, column_weights=(1.0, 1.0, 1.0, 1.0)):
    """trains on the triples of traj[skip:] (the first steps of a run are the
    relaxation of the initial condition and are left out); traj may be a list
    of trajectories (several packings on one mesh).  column_weights
    weight the outputs (c, mu, eta_top1, eta_top2) in the loss: the c
    error enters the residual through the mass term w0 = 1/dt and dominates
    it, so weighting c up lowers the residual of the prediction."""
    if isinstance(traj[0], (list, tuple)):
        traj = [t[skip:] for t in traj]
    else:
        traj = traj[skip:]
    mlp = MLP(op, seed=seed)
    pred = LearnedPredictor(op, mlp, dt_scale=1.0 / dt)
    X, T = build_dataset(pred, traj, dt, rows_per_step, seed, column_weights)
    m = X.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed + 1)
    b = min(batch, m)
    Xb = torch.empty(b, N_IN, dtype=torch.bfloat16, device=X.device)
    Tb = torch.empty(b, N_OUT, dtype=torch.float32, device=X.device)
    losses = []
    for it in range(iterations):
        idx = torch.randint(0, m, (b,), generator=g).to(X.device)
        torch.index_select(X, 0, idx, out=Xb)
        torch.index_select(T, 0, idx, out=Tb)
        lr_it = lr * 0.5 * (1.0 + math.cos(math.pi * it / iterations))
        losses.append(mlp.train_step(Xb, Tb, b, lr_it))
        if log is not None and (it % 100 == 0 or it == iterations - 1):
            log(it, losses[-1])
    pred.losses = losses
    return pred
