"""GMRES iteration counts of the tangent under the preconditioners of the
hot path and under the reference's default block preconditioner, on the
CPU (test infrastructure: the oracle's tangent assembled as a sparse
matrix).

- "InverseDiagonalMatrix": point Jacobi (I/lac/preconditioners.h:160-215,
  S/operator_base.h:364-389);
- "InverseComponentBlockDiagonalMatrix": the node-block Jacobi the GPU path
  uses (n_c x n_c block per node, the Cahn-Hilliard 2x2 coupling included);
- "BlockPreconditioner2": the reference's default (S/preconditioners.h:
  983-1282 with BlockPreconditioner2Data, S/preconditioner_parameters.h:
  57-64): ILU(0) of the Cahn-Hilliard block (rows and columns of c and mu)
  and the inverse diagonal of the Allen-Cahn block (the eta rows);
- "BlockPreconditioner2 (exact CH)": the same with the Cahn-Hilliard block
  solved exactly, the limit of any better Cahn-Hilliard preconditioner.

GMRES(30) with left preconditioning and the ReductionControl rule on the
preconditioned residual, as in the Newton solves (rtol 1e-2).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import pyoracle as po


def sparse_tangent(p, cache, keep_zeros=False):
    """the tangent as a CSR matrix (node-interleaved numbering), assembled
    from oracle vmults of coloured probe vectors: nodes 3 apart along every
    direction have disjoint Q1 neighbourhoods.  keep_zeros: store the whole
    Q1 pattern (every neighbour pair), entries that vanish included"""
    assert p.degree == 1
    nc, dim = p.n_comp, p.dim
    nn = [p.n_cells[d] + 1 for d in range(dim)]
    n_nodes = int(np.prod(nn))
    idx = np.arange(n_nodes)
    coords = []
    r = idx.copy()
    for d in range(dim):
        coords.append(r % nn[d])
        r //= nn[d]
    rows, cols, vals = [], [], []
    for color in range(3 ** dim):
        cc = [(color // 3 ** d) % 3 for d in range(dim)]
        sel = np.ones(n_nodes, dtype=bool)
        for d in range(dim):
            sel &= coords[d] % 3 == cc[d]
        probe_nodes = idx[sel]
        for c in range(nc):
            src = np.zeros(n_nodes * nc)
            src[probe_nodes * nc + c] = 1.0
            y = po.vmult(p, cache, src).reshape(n_nodes, nc)
            # owner of every node: the probe node within distance 1
            owner = np.full(n_nodes, -1)
            for off in np.ndindex(*(3,) * dim):
                shift = [o - 1 for o in off]
                ok = np.ones(len(probe_nodes), dtype=bool)
                tgt = np.zeros(len(probe_nodes), dtype=np.int64)
                stride = 1
                for d in range(dim):
                    x = coords[d][probe_nodes] + shift[d]
                    ok &= (x >= 0) & (x < nn[d])
                    tgt += x * stride
                    stride *= nn[d]
                owner[tgt[ok]] = probe_nodes[ok]
            rr = np.nonzero(owner >= 0)[0]
            for a in range(nc):
                v = y[rr, a]
                nz = np.ones(v.shape, dtype=bool) if keep_zeros else v != 0.0
                rows.append(rr[nz] * nc + a)
                cols.append(owner[rr[nz]] * nc + c)
                vals.append(v[nz])
    n = n_nodes * nc
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    A.sort_indices()
    return A


def gmres_iterations(A, Minv, b, rtol=1e-2, restart=30, maxit=3000):
    """left-preconditioned GMRES(restart) from x = 0 (scipy's GMRES on
    M^-1 A x = M^-1 b); stops when the preconditioned residual falls below
    rtol times its initial value; returns the iteration count (maxit when
    it did not converge)"""
    n = A.shape[0]
    op = spla.LinearOperator((n, n), matvec=lambda v: Minv(A @ v), dtype=np.float64)
    rhs = Minv(b)
    count = [0]

    def cb(res):
        count[0] += 1

    x, info = spla.gmres(op, rhs, rtol=rtol, atol=0.0, restart=restart, maxiter=maxit // restart + 1,
                         callback=cb, callback_type="pr_norm")
    return count[0] if info == 0 else maxit


def preconditioners(A, nc):
    n = A.shape[0]
    nn = n // nc
    d = A.diagonal()
    inv_d = np.where(np.abs(d) > 1e-10, 1.0 / d, 1.0)
    out = {"InverseDiagonalMatrix": lambda v: inv_d * v}
    # node blocks
    Acoo = A.tocoo()
    same = (Acoo.row // nc) == (Acoo.col // nc)
    B = np.zeros((nn, nc, nc))
    B[Acoo.row[same] // nc, Acoo.row[same] % nc, Acoo.col[same] % nc] = Acoo.data[same]
    Binv = np.linalg.inv(B)
    out["InverseComponentBlockDiagonalMatrix"] = \
        lambda v: np.einsum("nij,nj->ni", Binv, v.reshape(nn, nc)).ravel()
    # Cahn-Hilliard block (c, mu) and Allen-Cahn diagonal
    ch = np.sort(np.concatenate([np.arange(nn) * nc, np.arange(nn) * nc + 1]))
    ac = np.setdiff1d(np.arange(n), ch)
    Ach = A[ch][:, ch].tocsr()
    ilu = po.ILU0(Ach)
    lu = spla.splu(Ach.tocsc())
    inv_ac = inv_d[ac]

    def bp2(solve_ch):
        def apply(v):
            out_ = np.empty_like(v)
            out_[ch] = solve_ch(v[ch])
            out_[ac] = inv_ac * v[ac]
            return out_
        return apply

    out["BlockPreconditioner2"] = bp2(ilu.solve)
    out["BlockPreconditioner2 (exact CH)"] = bp2(lu.solve)
    return out


def newton_state(p, u, old, iterations=2):
    """u after a few exact Newton iterations of the BDF step from `old`
    (sparse direct solves): a linearisation point in the regime of the
    time loop, with the chemical potential relaxed"""
    for _ in range(iterations):
        A = sparse_tangent(p, po.fill_qp(p, u))
        u = u + spla.spsolve(A.tocsc(), -po.residual(p, 2, u, [old]))
    return u


def study(p, u, old, seed=0):
    """iteration counts at the linearisation point u: the Newton system
    J du = -R(u) to rtol 1e-2 (the Newton solves' GMRES tolerance), and a
    random right-hand side to rtol 1e-6"""
    A = sparse_tangent(p, po.fill_qp(p, u))
    b = -po.residual(p, 2, u, [old])
    z = np.random.default_rng(seed).standard_normal(A.shape[0])
    out = {}
    for name, M in preconditioners(A, p.n_comp).items():
        out[name] = (gmres_iterations(A, M, b, rtol=1e-2),
                     gmres_iterations(A, M, z, rtol=1e-6))
    return out


def parity_colour_ilu(A, nc, nn_dim):
    """the GPU's BlockPreconditioner2 restated: ILU(0) of the Cahn-Hilliard
    block (rows and columns of c and mu, the whole Q1 pattern) with the
    nodes ordered by parity colour (x&1 | (y&1)<<1 | (z&1)<<2) and index,
    c before mu; inverse diagonal (|d| <= 1e-10 -> 1) on the order
    parameters.  Returns the apply function."""
    n = A.shape[0]
    nn = n // nc
    node = np.arange(nn)
    colour = np.zeros(nn, dtype=np.int64)
    r = node.copy()
    for d, m in enumerate(nn_dim):
        colour |= (r % m) % 2 << d
        r //= m
    order = np.argsort(colour * nn + node, kind="stable")
    ch = np.concatenate([[k * nc, k * nc + 1] for k in order])
    Ach = A.tocsr()[ch][:, ch].tocsr()
    ilu = po.ILU0(Ach)
    d = A.diagonal()
    inv_d = np.where(np.abs(d) > 1e-10, 1.0 / d, 1.0)
    ac = np.setdiff1d(np.arange(n), ch)

    def apply(v):
        out = np.empty_like(v)
        out[ch] = ilu.solve(v[ch])
        out[ac] = inv_d[ac] * v[ac]
        return out
    return apply
