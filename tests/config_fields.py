"""Inputs of the config-workload parity tests: the benchmark packings'
initial fields (hpsint_amd.packing) on the configs' own mesh resolution,
either on the whole domain or on a sub-box of it, and statistics of the
quadrature-point states they produce, used to show that the inputs enter
the threshold branches of the physics:

- phi clamp, c > 1 (MobilityScalar / MobilityTensorial,
  S/mobility.h:508-509, 769-770, 836);
- |grad c| < 1e-4, the norm mask of the tensorial mobility derivative
  (S/mobility.h:881), and |grad c| <= 1e-10, the zero branch of unit_vector
  (I/base/tensor.h:137-155);
- |grad eta_i - grad eta_j|^2 <= 1e-20, the zero branch of
  unit_vector_filter_2 (I/base/tensor.h:157-174).

(S/ = applications/sintering/include/pf-applications/sintering/,
I/ = include/pf-applications/.)
"""
import numpy as np

from hpsint_amd import packing
from hpsint_amd.packing import Packing

# 1D Gauss points on [0, 1]
GAUSS = {1: [0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)],
         2: [0.5 - 0.5 * np.sqrt(0.6), 0.5, 0.5 + 0.5 * np.sqrt(0.6)]}


def basis_1d(degree, t):
    """values and derivatives (on [0, 1]) of the Lagrange basis on the
    Gauss-Lobatto support points of FE_Q(degree) at t"""
    if degree == 1:
        return [1.0 - t, t], [-1.0, 1.0]
    return ([2.0 * (t - 0.5) * (t - 1.0), -4.0 * t * (t - 1.0), 2.0 * t * (t - 0.5)],
            [4.0 * t - 3.0, -8.0 * t + 4.0, 4.0 * t - 1.0])


def sub_packing(pk, center, n_cells, h):
    """the packing on a cube of n_cells^3 cells of size h around `center`"""
    lo = tuple(float(c - 0.5 * n_cells * h) for c in center)
    hi = tuple(float(c + 0.5 * n_cells * h) for c in center)
    return Packing(pk.centers, pk.radii, lo, hi, pk.colors)


def contact_point(pk, k=0):
    """the midpoint of the k-th closest pair of touching spheres of
    different order parameters"""
    pairs = []
    n = len(pk.radii)
    for i in range(n):
        for j in range(i + 1, n):
            if pk.colors[i] == pk.colors[j]:
                continue
            d = np.linalg.norm(pk.centers[i] - pk.centers[j])
            pairs.append((d - pk.radii[i] - pk.radii[j], i, j))
    pairs.sort()
    _, i, j = pairs[k]
    a, b = pk.centers[i], pk.centers[j]
    t = pk.radii[i] / (pk.radii[i] + pk.radii[j])
    return a + t * (b - a)


def with_mu(pk, u, n_nodes_dim, G):
    """the initial fields with a smooth nonzero chemical potential
    mu = 0.2 sin(x/5.3) cos(y/4.1) + 0.1 cos(z/6.7) (the initial fields have
    mu = 0, which would leave the grad-mu terms of the tangent untested)"""
    nc = G + 2
    u = u.reshape(-1, nc).copy()
    axes = [np.linspace(l, h, n) for l, h, n in zip(pk.lower, pk.upper, n_nodes_dim)]
    grid = np.meshgrid(*reversed(axes), indexing="ij")
    x, y = grid[-1].ravel(), grid[-2].ravel()
    z = grid[0].ravel() if len(axes) == 3 else 0.0 * x
    u[:, 1] = 0.2 * np.sin(x / 5.3) * np.cos(y / 4.1) + 0.1 * np.cos(z / 6.7)
    return u.ravel()


def state(pk, n_cells, G, degree=1):
    """interleaved nodal state of the packing on n_cells (c, mu, eta_i)"""
    nn = [degree * n + 1 for n in n_cells]
    u = packing.initial_fields(pk, [degree * n for n in n_cells], n_order_params=G)
    return with_mu(pk, u, nn, G)


def qp_states(u, n_cells, h, nc, degree=1):
    """yields, per quadrature point of the cell (tensor-product Gauss rule),
    the values (cells, nc) and gradients (cells, nc, dim) of the nodal state
    u (3D, x fastest)"""
    nx, ny, nz = n_cells
    U = np.asarray(u, dtype=np.float64).reshape(degree * nz + 1, degree * ny + 1,
                                                degree * nx + 1, nc)
    pts = GAUSS[degree]

    def contract(A, axis, w, n):
        # sum_a w[a] A[.. degree*i + a ..] for the n cells along axis
        out = 0.0
        for a, wa in enumerate(w):
            sl = [slice(None)] * 4
            sl[axis] = slice(a, a + degree * n, degree) if degree > 1 else slice(a, a + n)
            out = out + wa * A[tuple(sl)]
        return out

    for qz in pts:
        for qy in pts:
            for qx in pts:
                vals = []
                bx, dx = basis_1d(degree, qx)
                by, dy = basis_1d(degree, qy)
                bz, dz = basis_1d(degree, qz)
                v = contract(contract(contract(U, 2, bx, nx), 1, by, ny), 0, bz, nz)
                gx = contract(contract(contract(U, 2, dx, nx), 1, by, ny), 0, bz, nz) / h[0]
                gy = contract(contract(contract(U, 2, bx, nx), 1, dy, ny), 0, bz, nz) / h[1]
                gz = contract(contract(contract(U, 2, bx, nx), 1, by, ny), 0, dz, nz) / h[2]
                vals = v.reshape(-1, nc)
                grads = np.stack([gx.reshape(-1, nc), gy.reshape(-1, nc), gz.reshape(-1, nc)],
                                 axis=-1)
                yield vals, grads


def branch_stats(u, n_cells, h, nc, degree=1):
    """numbers of quadrature points (or point-pair combinations) that enter
    each threshold branch"""
    G = nc - 2
    st = {"qp": 0, "c_gt_1": 0, "grad_c_lt_1e-4": 0, "grad_c_le_1e-10": 0,
          "pair_grad_diff_zero": 0, "max_c": -np.inf}
    for vals, grads in qp_states(u, n_cells, h, nc, degree):
        c = vals[:, 0]
        gc = np.sqrt((grads[:, 0, :] ** 2).sum(axis=1))
        st["qp"] += c.size
        st["c_gt_1"] += int((c > 1.0).sum())
        st["max_c"] = max(st["max_c"], float(c.max()))
        st["grad_c_lt_1e-4"] += int((gc < 1e-4).sum())
        st["grad_c_le_1e-10"] += int((gc <= 1e-10).sum())
        for i in range(G):
            for j in range(i):
                e = grads[:, 2 + i, :] - grads[:, 2 + j, :]
                st["pair_grad_diff_zero"] += int(((e * e).sum(axis=1) <= 1e-20).sum())
    return st


def assert_branches_reached(st):
    assert st["c_gt_1"] > 0, st
    assert st["grad_c_lt_1e-4"] > 0, st
    assert st["pair_grad_diff_zero"] > 0, st
