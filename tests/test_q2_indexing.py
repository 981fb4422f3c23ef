"""CPU checks of the Q2 kernels' index arithmetic and constants
(hpsint_amd/csrc/fe_q2.hpp: Q2Lane, q2_gather, q2_scatter; ops.hpp:
q2_colour, q2_workgroups).

- the colour passes visit every cell exactly once, cells of one colour share
  no node, every global index is in range and every cache workgroup index is
  below the allocated count;
- every node is stored exactly once (by the first colour that reaches it,
  before any add) and added to by each of its other cells;
- every LDS read of a sum-factorisation pass stays inside its 64-entry row;
- the 1D coefficients written in fe_q2.hpp are the Q2 Lagrange basis at the
  3-point Gauss points, computed here independently by numpy (values,
  derivatives, Gauss weights) and tied together by Dq S = L'.
"""
import itertools
import os
import re

import numpy as np
import pytest

HDR = os.path.join(os.path.dirname(__file__), "..", "hpsint_amd", "csrc", "fe_q2.hpp")


def colour(nc, dim, col, cpw, wg):
    par = [(col >> d) & 1 if d < dim else 0 for d in range(3)]
    ncol = [(nc[d] - par[d] + 1) // 2 if d < dim else 1 for d in range(3)]
    n_col = ncol[0] * ncol[1] * ncol[2]
    return par, ncol, n_col, wg, wg + -(-n_col // cpw)


def cell_of(par, ncol, n_col, ci):
    if ci >= n_col:
        return None
    q = ci // ncol[0]
    ix = ci - q * ncol[0]
    iy, iz = q % ncol[1], q // ncol[1]
    return (2 * ix + par[0], 2 * iy + par[1], 2 * iz + par[2])


def node_of(dim, nn, c, p):
    gx = 2 * c[0] + p % 3
    gy = 2 * c[1] + (p // 3) % 3
    gz = 2 * c[2] + p // 9 if dim == 3 else 0
    return gx + nn[0] * (gy + nn[1] * gz)


def first(dim, nc, c, p):
    tt = [p % 3, (p // 3) % 3, p // 9 if dim == 3 else 1]
    return all((c[d] & 1) == 0 or tt[d] == 1 or (tt[d] == 0 and c[d] == 0) or
               (tt[d] == 2 and c[d] == nc[d] - 1) for d in range(dim))


@pytest.mark.parametrize("dim,cells", [
    (3, (5, 4, 7)), (3, (1, 3, 2)), (3, (1, 1, 1)), (3, (2, 2, 2)), (3, (6, 5, 3)),
    (2, (9, 6)), (2, (1, 5)), (2, (1, 1)), (2, (7, 8))])
def test_colour_passes_store_once_then_add(dim, cells):
    NP = 27 if dim == 3 else 9
    CPW = 64 // NP
    nc = [cells[0], cells[1], cells[2] if dim == 3 else 1]
    nn = [2 * nc[d] + 1 if d < dim else 1 for d in range(3)]
    n_nodes = nn[0] * nn[1] * nn[2]
    state = [0] * n_nodes          # 0 untouched, then number of contributions
    seen = set()
    wg = 0
    total_wg = 0
    for col in range(1 << dim):
        par, ncol, n_col, base, wg = colour(nc, dim, col, CPW, wg)
        total_wg = wg
        touched = set()
        for b in range(-(-n_col // CPW)):
            for s in range(CPW):
                c = cell_of(par, ncol, n_col, b * CPW + s)
                if c is None:
                    continue
                assert all(0 <= c[d] < nc[d] for d in range(3))
                assert c not in seen
                seen.add(c)
                assert base + b < wg
                for p in range(NP):
                    n = node_of(dim, nn, c, p)
                    assert 0 <= n < n_nodes
                    assert n not in touched, "two cells of one colour share a node"
                    touched.add(n)
                    if first(dim, nc, c, p):
                        assert state[n] == 0
                    else:
                        assert state[n] > 0
                    state[n] += 1
    assert len(seen) == nc[0] * nc[1] * nc[2]
    # every node reached, with as many contributions as cells hold it
    for n in range(n_nodes):
        gx, gy, gz = n % nn[0], (n // nn[0]) % nn[1], n // (nn[0] * nn[1])
        holders = 1
        for g, d in ((gx, 0), (gy, 1), (gz, 2)):
            if d < dim and g % 2 == 0 and 0 < g < 2 * nc[d]:
                holders *= 2
        assert state[n] == holders
    assert total_wg >= 1


@pytest.mark.parametrize("dim", [2, 3])
def test_lds_pass_reads_stay_in_row(dim):
    NP = 27 if dim == 3 else 9
    CPW = 64 // NP
    strides = (1, 3, 9)
    for l in range(64):
        s, p = l // NP, l % NP
        t = [p % 3, (p // 3) % 3, p // 9 if dim == 3 else 0]
        for d in range(dim):
            b0 = l - t[d] * strides[d] if s < CPW else 0
            for a in range(3):
                idx = b0 + a * strides[d]
                assert 0 <= idx < 64
                if s < CPW:
                    # a live lane reads its own cell's line
                    assert idx // NP == s
    # gather / scatter items map into the cell slots of a row
    for NC in (3, 4, 6, 12, 14):
        for f in range(CPW * NP * NC):
            sp, c = divmod(f, NC)
            s2, p2 = divmod(sp, NP)
            assert s2 < CPW and c * 64 + sp < NC * 64


def header_constant(name):
    txt = open(HDR).read()
    m = re.search(r"static constexpr T %s\s*=\s*T\(([-0-9.eE+]+)\)" % name, txt)
    assert m, name
    return float(m.group(1))


def test_q2_coefficients_are_the_lagrange_basis_at_gauss_points():
    z, w = np.polynomial.legendre.leggauss(3)
    x = np.sort(0.5 * (z + 1.0))
    w = 0.5 * w[np.argsort(z)]
    nodes = np.array([0.0, 0.5, 1.0])
    # Lagrange basis through the nodes and its derivative, by polyfit
    V = np.vander(nodes, 3)
    coef = np.linalg.solve(V, np.eye(3))          # columns: basis polynomials
    S = np.vander(x, 3) @ coef                     # S[q][n]
    dV = np.stack([2 * x, np.ones(3), np.zeros(3)], axis=1)
    E = dV @ coef                                  # L'[q][n]
    Vq = np.vander(x, 3)
    cq = np.linalg.solve(Vq, np.eye(3))
    Dq = dV @ cq                                   # collocation derivative
    np.testing.assert_allclose(Dq @ S, E, atol=1e-13)

    C = {k: header_constant(k) for k in
         ("s_a", "s_b", "s_c", "d_a", "d_b", "d_c", "e_a", "e_b", "e_c", "w_a", "w_b")}
    S_h = np.array([[C["s_a"], C["s_b"], C["s_c"]], [0, 1, 0], [C["s_c"], C["s_b"], C["s_a"]]])
    D_h = np.array([[C["d_a"], C["d_b"], C["d_c"]], [C["d_c"], 0, -C["d_c"]],
                    [-C["d_c"], -C["d_b"], -C["d_a"]]])
    E_h = np.array([[C["e_a"], C["e_b"], C["e_c"]], [-1, 0, 1],
                    [-C["e_c"], -C["e_b"], -C["e_a"]]])
    np.testing.assert_allclose(S_h, S, rtol=0, atol=2e-16)
    np.testing.assert_allclose(D_h, Dq, rtol=0, atol=4e-15)
    np.testing.assert_allclose(E_h, E, rtol=0, atol=2e-15)
    np.testing.assert_allclose([C["w_a"], C["w_b"], C["w_a"]], w, rtol=0, atol=1e-16)
