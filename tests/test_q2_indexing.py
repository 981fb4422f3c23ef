"""CPU checks of the Q2 kernels' index arithmetic and constants
(hpsint_amd/csrc/fe_q2.hpp: Q2Lane, q2_gather, q2_scatter, q2_vmult_tile;
ops.hpp: q2_colour, q2_workgroups, Q2Launch::tiles).

- the colour passes visit every cell exactly once, cells of one colour share
  no node, every global index is in range and every cache workgroup index is
  below the allocated count;
- every node is stored exactly once (by the first colour that reaches it,
  before any add) and added to by each of its other cells;
- the 3D tile march (q2_vmult_tile) adds every cell's contribution to each
  of its nodes exactly once, stores each node once before any add, and reads
  the cache entries the colour passes' fill wrote;
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


def tile_colours(nc, NY, zc):
    """the launches of q2_vmult_tile (ops.hpp Q2Launch::tiles): per colour
    the tile parities and the tiles of that colour along x, y, z"""
    ext = (4, NY, zc)
    nt = [-(-nc[d] // ext[d]) for d in range(3)]
    for col in range(8):
        tpar = [(col >> d) & 1 for d in range(3)]
        ntc = [(nt[d] - tpar[d] + 1) // 2 for d in range(3)]
        yield tpar, ntc


@pytest.mark.parametrize("cells,NY,zc", [
    ((5, 4, 7), 2, 3), ((9, 5, 6), 2, 6), ((8, 4, 12), 2, 6), ((3, 3, 2), 1, 1),
    ((12, 7, 5), 1, 2), ((1, 1, 1), 2, 6), ((4, 2, 9), 2, 4), ((7, 6, 4), 1, 5)])
def test_tile_march_sums_every_contribution_once(cells, NY, zc):
    """Emulates q2_vmult_tile lane by lane: every (cell, local node)
    contribution reaches its node exactly once, every node is stored once by
    the first colour and then only added to, tiles of one colour write
    disjoint nodes, the owner reads of other cells' rows hit a lane holding
    that cell and node, and each wave reads the cache entries the colour
    passes' fill wrote for its cells."""
    NP, CPW = 27, 2
    nc = list(cells)
    nn = [2 * n + 1 for n in nc]
    n_nodes = nn[0] * nn[1] * nn[2]
    # the fill's cache slot of every cell (colour passes, two cells per block)
    slot = {}
    wgb = []
    wg = 0
    for col in range(8):
        par, ncol, n_col, base, wg = colour(nc, 3, col, CPW, wg)
        wgb.append(base)
        for ci in range(n_col):
            slot[cell_of(par, ncol, n_col, ci)] = (base + ci // 2, ci % 2)
    NW, PXY = 2 * NY, 9 * (2 * NY + 1)
    state = [0] * n_nodes
    used = set()
    for tpar, ntc in tile_colours(nc, NY, zc):
        written = set()
        for b in range(ntc[0] * ntc[1] * ntc[2]):
            q = b // ntc[0]
            tx = 2 * (b - q * ntc[0]) + tpar[0]
            ty = 2 * (q % ntc[1]) + tpar[1]
            tz = 2 * (q // ntc[1]) + tpar[2]
            z0 = zc * tz
            nl = min(zc, nc[2] - z0)
            assert nl >= 1
            lanes = {}
            for w in range(NW):
                xp, cyl = w & 1, w >> 1
                for l in range(64):
                    sl = l // NP if l < 2 * NP else 0
                    cxl = xp + 2 * sl
                    cx, cy = 4 * tx + cxl, NY * ty + cyl
                    valid = l < 2 * NP and cx < nc[0] and cy < nc[1]
                    lanes[(w, l)] = (valid, cxl, cyl, cx, cy, l - NP * sl)
            carry = {}
            for k in range(nl):
                cz = z0 + k
                rows = {}
                for (w, l), (valid, cxl, cyl, cx, cy, p) in lanes.items():
                    if valid:
                        rows[(w, l)] = {((cx, cy, cz), p)}
                        # cache entry of this lane = the fill's
                        xp, py = w & 1, cy & 1
                        col = xp + 2 * py + 4 * (cz & 1)
                        ncx, ncy = (nc[0] - xp + 1) >> 1, (nc[1] - py + 1) >> 1
                        ci = ((cz >> 1) * ncy + (cy >> 1)) * ncx + (cx >> 1)
                        assert (wgb[col] + (ci >> 1), ci & 1) == slot[(cx, cy, cz)]
                new_carry = {}
                for (w, l), (valid, cxl, cyl, cx, cy, p) in lanes.items():
                    t0, t1, t2 = p % 3, (p // 3) % 3, p // 9
                    if not (valid and (t0 != 0 or cxl == 0) and (t1 != 0 or cyl == 0)):
                        continue
                    xr = t0 == 2 and cxl < 3 and cx + 1 < nc[0]
                    yu = t1 == 2 and cyl < NY - 1 and cy + 1 < nc[1]
                    wx, sx = w ^ 1, (cxl + 1) >> 1
                    v = set(rows[(w, l)])
                    reads = []
                    if xr:
                        reads.append(((wx, sx * NP + p - 2), (cx + 1, cy, cz), p - 2))
                    if yu:
                        reads.append(((w + 2, l - 6), (cx, cy + 1, cz), p - 6))
                    if xr and yu:
                        reads.append(((wx + 2, sx * NP + p - 8), (cx + 1, cy + 1, cz), p - 8))
                    for (rw, rl), cell, rp in reads:
                        assert 0 <= rl < 2 * NP and rw < NW
                        assert rows[(rw, rl)] == {(cell, rp)}
                        v |= rows[(rw, rl)]
                    pn = (2 * cxl + t0) + 9 * (2 * cyl + t1)
                    assert 0 <= pn < PXY
                    if t2 == 0 and k > 0:
                        v |= carry.pop(pn)
                    if t2 < 2 or k == nl - 1:
                        gx, gy, gz = 2 * cx + t0, 2 * cy + t1, 2 * cz + t2
                        shx = (gx == 8 * tx and tx > 0) or (gx == 8 * tx + 8 and gx < nn[0] - 1)
                        shy = (gy == 2 * NY * ty and ty > 0) or \
                              (gy == 2 * NY * (ty + 1) and gy < nn[1] - 1)
                        shz = (t2 == 0 and k == 0 and tz > 0) or (t2 == 2 and gz < nn[2] - 1)
                        fst = (not shx or tx % 2 == 0) and (not shy or ty % 2 == 0) and \
                              (not shz or tz % 2 == 0)
                        n = gx + nn[0] * (gy + nn[1] * gz)
                        assert 0 <= n < n_nodes
                        assert n not in written, "two writes of one node in one colour"
                        written.add(n)
                        if fst:
                            assert state[n] == 0
                        else:
                            assert state[n] > 0
                        state[n] += len(v)
                        assert not (used & v)
                        used |= v
                    else:
                        assert pn not in new_carry
                        new_carry[pn] = v
                assert not carry, "a carried plane node was not consumed"
                carry = new_carry
            assert not carry
    n_cells = nc[0] * nc[1] * nc[2]
    assert len(used) == n_cells * NP
    for n in range(n_nodes):
        gx, gy, gz = n % nn[0], (n // nn[0]) % nn[1], n // (nn[0] * nn[1])
        holders = 1
        for g, d in ((gx, 0), (gy, 1), (gz, 2)):
            if g % 2 == 0 and 0 < g < 2 * nc[d]:
                holders *= 2
        assert state[n] == holders


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
