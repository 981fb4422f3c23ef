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


@pytest.mark.parametrize("march", [1, 2, 3])
@pytest.mark.parametrize("tz", [16, 5])
@pytest.mark.parametrize("ncells", [c for d, c in CASES if d == 3])
def test_zcol_cache_map(ncells, tz, march):
    """q1_vmult_zcol (hpsint_amd/csrc/fe_q1_zcol.hpp) reads the cache of the
    8 x 4 x tz tiles: for every lane of every 8 x 8 x (march tz) tile and
    every layer of its march, the cache layer and the lanes of its two Gauss
    points are those the 8 x 4 x tz tiling gives the same cell and corners;
    lanes of cells beyond the mesh read inside the cache."""
    nc, nn, ty, nt, PX, PY = tiling(3, ncells, 8, 4, tz)
    ntz1 = -(-nc[1] // 8)
    n_tiles = nt[0] * nt[1] * nt[2]
    mz = march * tz  # layers per march (kp.tz of the launch; kp.ctz = tz)
    ntm = -(-nc[2] // mz)
    for b in range(nt[0] * ntz1 * ntm):
        tix, r = b % nt[0], b // nt[0]
        tiy, tim = r % ntz1, r // ntz1
        x0, y0, m0 = tix * 8, tiy * 8, tim * mz
        tx_eff, ty_eff = min(8, nc[0] - x0), min(8, nc[1] - y0)
        tm_eff = min(mz, nc[2] - m0)
        for tid in range(256):
            xy, cell = tid & 3, tid >> 2
            cx, cy = cell & 7, cell >> 3
            tyc = 2 * tiy + (cy >> 2)
            if tyc >= nt[1]:
                tyc = 2 * tiy
            cxy = tix + nt[0] * tyc
            lane_lo = (((cy & 3) * 8 + cx) << 3) | xy
            assert 0 <= lane_lo and (lane_lo | 4) < 256
            for k in range(tm_eff):
                gz = m0 + k
                ct = gz // tz
                layer = (cxy + nt[0] * nt[1] * ct) * tz + (gz - ct * tz)
                assert 0 <= layer < n_tiles * tz
                if cx < tx_eff and cy < ty_eff:
                    gy = y0 + cy
                    ref_tile = tix + nt[0] * (gy // 4 + nt[1] * (gz // tz))
                    assert layer == ref_tile * tz + gz % tz
                    ref_lane = ((gy % 4) * 8 + cx) * 8
                    assert lane_lo == ref_lane + xy and (lane_lo | 4) == ref_lane + xy + 4


def zcol_slot_layout(ncells, tz):
    """Shared-node partials of q1_vmult_zcol (8 x 8 x tz tiles), slot-major
    per class: nodes on an interior z boundary (8 slots), else on a y
    boundary (4), else on an x boundary (2); slot bit d set when the writing
    tile is the upper one along d.  Returns the tiling, the class sizes and
    the record index function of the kernel's put()."""
    nc = list(ncells)
    nn = [n + 1 for n in nc]
    tsz = [8, 8, tz]
    nt = [-(-nc[d] // tsz[d]) for d in range(3)]
    nb = [t - 1 for t in nt]
    n_m, n_y, n_x = nb[2] * nn[1] * nn[0], nb[1] * nn[2] * nn[0], nb[0] * nn[2] * nn[1]

    def record(gx, gy, gz, sx, sy, sz, bxn, byn, bzk):
        if bzk:
            i = ((gz // tz - 1) * nn[1] + gy) * nn[0] + gx
            return (sx | sy << 1 | sz << 2) * n_m + i
        if byn:
            i = ((gy // 8 - 1) * nn[2] + gz) * nn[0] + gx
            return 8 * n_m + (sx | sy << 1) * n_y + i
        i = ((gx // 8 - 1) * nn[2] + gz) * nn[1] + gy
        return 8 * n_m + 4 * n_y + sx * n_x + i
    return nc, nn, tsz, nt, nb, (n_m, n_y, n_x), record


@pytest.mark.parametrize("tz", [32, 16, 5, 3])
@pytest.mark.parametrize("ncells", [c for d, c in CASES if d == 3] + [(17, 9, 33), (8, 16, 4)])
def test_zcol_slot_partials(ncells, tz):
    """put() of q1_vmult_zcol and q1_fixup_slots: every node written once,
    records in bounds and written once, the fix-up reads exactly the records
    the tiles holding the node wrote"""
    nc, nn, tsz, nt, nb, (n_m, n_y, n_x), record = zcol_slot_layout(ncells, tz)
    n_rec = 8 * n_m + 4 * n_y + 2 * n_x
    dst_count = {}
    written = {}
    for b in range(nt[0] * nt[1] * nt[2]):
        tix, r = b % nt[0], b // nt[0]
        tiy, tim = r % nt[1], r // nt[1]
        x0, y0, m0 = tix * 8, tiy * 8, tim * tz
        tx_eff, ty_eff, tm_eff = min(8, nc[0] - x0), min(8, nc[1] - y0), min(tz, nc[2] - m0)
        for k in range(tm_eff + 1):
            for ny in range(ty_eff + 1):
                for nx in range(tx_eff + 1):
                    gx, gy, gz = x0 + nx, y0 + ny, m0 + k
                    node = (gz * nn[1] + gy) * nn[0] + gx
                    bxn = (nx == 0 and x0 > 0) or (nx == tx_eff and x0 + tx_eff < nc[0])
                    byn = (ny == 0 and y0 > 0) or (ny == ty_eff and y0 + ty_eff < nc[1])
                    bzk = (k == 0 and m0 > 0) or (k == tm_eff and m0 + tm_eff < nc[2])
                    if not (bxn or byn or bzk):
                        dst_count[node] = dst_count.get(node, 0) + 1
                        continue
                    sx, sy, sz = int(bxn and nx == 0), int(byn and ny == 0), int(bzk and k == 0)
                    rec = record(gx, gy, gz, sx, sy, sz, bxn, byn, bzk)
                    assert 0 <= rec < n_rec
                    assert rec not in written, (gx, gy, gz, b)
                    written[rec] = b

    def is_bnd(g, t, n):
        return 0 < g < n and g % t == 0

    for idx in range(n_m + n_y + n_x):
        if idx < n_m:
            gx, gy, gz = idx % nn[0], (idx // nn[0]) % nn[1], (idx // (nn[0] * nn[1]) + 1) * tz
            sl = [sx | sy << 1 | sz << 2 for sz in (0, 1)
                  for sy in ((0, 1) if is_bnd(gy, 8, nc[1]) else (0,))
                  for sx in ((0, 1) if is_bnd(gx, 8, nc[0]) else (0,))]
            recs = [s * n_m + idx for s in sl]
        elif idx < n_m + n_y:
            i = idx - n_m
            gx, gz, gy = i % nn[0], (i // nn[0]) % nn[2], (i // (nn[0] * nn[2]) + 1) * 8
            if is_bnd(gz, tz, nc[2]):
                continue
            sl = [sx | sy << 1 for sy in (0, 1)
                  for sx in ((0, 1) if is_bnd(gx, 8, nc[0]) else (0,))]
            recs = [8 * n_m + s * n_y + i for s in sl]
        else:
            i = idx - n_m - n_y
            gy, gz, gx = i % nn[1], (i // nn[1]) % nn[2], (i // (nn[1] * nn[2]) + 1) * 8
            if is_bnd(gz, tz, nc[2]) or is_bnd(gy, 8, nc[1]):
                continue
            recs = [8 * n_m + 4 * n_y + s * n_x + i for s in (0, 1)]
        node = (gz * nn[1] + gy) * nn[0] + gx
        for rec in recs:
            assert rec in written, (gx, gy, gz, rec)
            del written[rec]
        dst_count[node] = dst_count.get(node, 0) + 1
    assert not written, list(written.items())[:5]
    assert len(dst_count) == nn[0] * nn[1] * nn[2]
    assert all(c == 1 for c in dst_count.values())
