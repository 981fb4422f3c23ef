"""CPU emulation of the index arithmetic of the Q1 kernels' tiling
(hpsint_amd/csrc/fe_q1.hpp: Tile, put, q1_fixup; hpsint_amd/csrc/api.hip:
hps_create) on small and ragged meshes: every node of dst is written exactly
once (by its only tile or by the fix-up), the fix-up reads only partials that
a tile wrote, every global/LDS index stays in bounds, and the fix-up sums
exactly the tiles whose closure holds the node."""
import itertools

import pytest


def tiling(dim, ncells, TX, TY, TZ):
    nc = [ncells[0], ncells[1] if dim == 3 else 1, ncells[dim - 1]]
    nn = [nc[0] + 1, nc[1] + 1 if dim == 3 else 1, nc[2] + 1]
    ty = TY if dim == 3 else 1
    nt = [-(-nc[0] // TX), -(-nc[1] // ty), -(-nc[2] // TZ)]
    PX, PY = TX + 1, (ty + 1 if dim == 3 else 1)
    return nc, nn, ty, nt, PX, PY


def emulate(dim, ncells, TX, TY, TZ):
    nc, nn, ty, nt, PX, PY = tiling(dim, ncells, TX, TY, TZ)
    PN = PX * PY
    n_nodes = nn[0] * nn[1] * nn[2]
    dst_count = [0] * n_nodes
    part_written = set()
    holders = {}  # node -> set of tiles with the node in their closure
    n_tiles = nt[0] * nt[1] * nt[2]
    for b in range(n_tiles):
        tix, r = b % nt[0], b // nt[0]
        tiy, tim = r % nt[1], r // nt[1]
        x0, y0, m0 = tix * TX, tiy * ty, tim * TZ
        tx_eff = min(TX, nc[0] - x0)
        ty_eff = min(ty, nc[1] - y0) if dim == 3 else 1
        tm_eff = min(TZ, nc[2] - m0)
        assert tx_eff >= 1 and ty_eff >= 1 and tm_eff >= 1
        # lane -> LDS slot index stays inside the plane
        for cx, cy, bx, by in itertools.product(range(TX), range(ty), (0, 1),
                                                (0, 1) if dim == 3 else (0,)):
            assert (cx + bx) + PX * (cy + by) < PN
        for k in range(tm_eff + 1):
            for pn in range(PN):
                nx, ny = pn % PX, pn // PX
                if not (nx <= tx_eff and (dim == 2 or ny <= ty_eff)):
                    continue
                gx, gy, gm = x0 + nx, y0 + ny, m0 + k
                assert gx < nn[0] and gy < nn[1] and gm < nn[2]
                node = (gm * nn[1] + gy) * nn[0] + gx
                holders.setdefault(node, set()).add(b)
                shared = ((nx == 0 and x0 > 0) or (nx == tx_eff and x0 + tx_eff < nc[0]) or
                          (k == 0 and m0 > 0) or (k == tm_eff and m0 + tm_eff < nc[2]))
                if dim == 3:
                    shared = shared or (ny == 0 and y0 > 0) or (ny == ty_eff and y0 + ty_eff < nc[1])
                if shared:
                    idx = (b * (TZ + 1) + k) * PN + pn
                    assert idx < n_tiles * (TZ + 1) * PN
                    part_written.add(idx)
                else:
                    dst_count[node] += 1

    tsz = [TX, ty, TZ]
    nb = [t - 1 for t in nt]
    n_m = nb[2] * nn[0] * nn[1]
    n_y = nb[1] * nn[0] * nn[2]
    n_x = nb[0] * nn[1] * nn[2]

    def is_bnd(g, t, n):
        return 0 < g < n and g % t == 0

    for idx in range(n_m + n_y + n_x):
        g = [0, 0, 0]
        if idx < n_m:
            g[0] = idx % nn[0]
            i = idx // nn[0]
            g[1] = i % nn[1]
            g[2] = (i // nn[1] + 1) * tsz[2]
        elif idx - n_m < n_y:
            i = idx - n_m
            g[0] = i % nn[0]
            i //= nn[0]
            g[2] = i % nn[2]
            g[1] = (i // nn[2] + 1) * tsz[1]
            if is_bnd(g[2], tsz[2], nc[2]):
                continue
        else:
            i = idx - n_m - n_y
            g[0] = (i % nb[0] + 1) * tsz[0]
            i //= nb[0]
            g[1] = i % nn[1]
            g[2] = i // nn[1]
            if is_bnd(g[2], tsz[2], nc[2]) or is_bnd(g[1], tsz[1], nc[1]):
                continue
        tl, lo = [], []
        for d in range(3):
            if is_bnd(g[d], tsz[d], nc[d]):
                tl.append([g[d] // tsz[d] - 1, g[d] // tsz[d]])
                lo.append([tsz[d], 0])
            else:
                t = min(g[d] // tsz[d], nt[d] - 1)
                tl.append([t])
                lo.append([g[d] - t * tsz[d]])
        node = (g[2] * nn[1] + g[1]) * nn[0] + g[0]
        tiles = set()
        for a in range(len(tl[2])):
            for bb in range(len(tl[1])):
                for e in range(len(tl[0])):
                    tile = (tl[2][a] * nt[1] + tl[1][bb]) * nt[0] + tl[0][e]
                    pn = lo[0][e] + PX * lo[1][bb]
                    key = (tile * (TZ + 1) + lo[2][a]) * PN + pn
                    assert key in part_written, (g, tile)
                    tiles.add(tile)
        assert tiles == holders[node], (g, tiles, holders[node])
        dst_count[node] += 1
    return dst_count


CASES = [
    (3, (1, 1, 1)), (3, (2, 2, 1)), (3, (8, 4, 16)), (3, (9, 5, 17)),
    (3, (19, 13, 37)), (3, (16, 8, 33)), (3, (3, 7, 2)),
    (2, (2, 1)), (2, (64, 16)), (2, (65, 17)), (2, (160, 80)), (2, (130, 3)),
]


@pytest.mark.parametrize("tz", [16, 5, 64])
@pytest.mark.parametrize("dim,ncells", CASES)
def test_every_node_written_exactly_once(dim, ncells, tz):
    if dim == 3:
        for ty in (4, 8):
            counts = emulate(3, ncells, 8, ty, tz)
            assert all(c == 1 for c in counts), [i for i, c in enumerate(counts) if c != 1][:10]
    else:
        counts = emulate(2, ncells, 64, 1, tz)
        assert all(c == 1 for c in counts), [i for i, c in enumerate(counts) if c != 1][:10]
