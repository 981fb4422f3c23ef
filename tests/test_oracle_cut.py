"""The oracle's grain cut-off (CPU): per-cell masks by the reference's
l-infinity rule, and the cut vmult, residual and node blocks pinned by the
reduced-system identities of do_vmult_range_no_template and
do_evaluate_nonlinear_residual_nt (S/operator_sintering_generic.h:456-537,
1008-1081, S/ = applications/sintering/include/pf-applications/sintering/):

- every order parameter active in every cell: the full operator;
- order parameter g inactive in every cell: its rows are zero and the other
  rows are the operator of the system without g (one order parameter less);
- random per-cell masks: the C restatement agrees with the independent
  per-cell Python restatement (3D Q1).
"""
import numpy as np
import pytest

import pyoracle as po

CASES = [(3, 1, (3, 2, 2), False), (3, 1, (2, 3, 2), True), (2, 1, (4, 3), False),
         (2, 1, (3, 4), True), (3, 2, (2, 1, 2), False), (2, 2, (2, 3), True)]


def state(p, seed):
    rng = np.random.default_rng(seed)
    nn = po.n_nodes(p)
    u = rng.uniform(0, 1, size=(nn, p.n_comp))
    u[:, 1] = rng.uniform(-1, 1, nn)
    return u


def params(dim, deg, cells, nc, tens):
    h = (0.5, 0.75, 1.0)[:dim]
    return po.make_params(dim, deg, cells, h, nc, tensorial=tens, w0=100.0, w_hist=(-100.0,))


def test_cell_masks_linfty_rule():
    p = params(3, 1, (4, 3, 2), 6, False)
    u = state(p, 1)
    u[:, 2:] *= (np.random.default_rng(2).uniform(0, 1, size=(u.shape[0], 4)) > 0.5)
    for tol in (0.0, 0.3, 0.9):
        m = po.cell_masks(p, u.ravel(), tol)
        nx, ny, nz = 4, 3, 2
        U = u.reshape(nz + 1, ny + 1, nx + 1, 6)
        for cz in range(nz):
            for cy in range(ny):
                for cx in range(nx):
                    box = np.abs(U[cz:cz + 2, cy:cy + 2, cx:cx + 2, 2:]).reshape(-1, 4).max(0)
                    want = sum(1 << g for g in range(4) if box[g] > tol)
                    assert m[cx + nx * (cy + ny * cz)] == want


@pytest.mark.parametrize("dim,deg,cells,tens", CASES)
def test_cut_operators_all_active_are_the_full_operators(dim, deg, cells, tens):
    nc = 5
    p = params(dim, deg, cells, nc, tens)
    u = state(p, 3)
    old = state(p, 4)
    src = np.random.default_rng(5).uniform(-1, 1, u.size)
    cache = po.fill_qp(p, u.ravel())
    masks = np.full(po.n_cells(p), 0b111, dtype=np.uint32)

    def close(a, b):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-13 * np.abs(b).max())

    close(po.vmult_cut(p, cache, src, masks), po.vmult(p, cache, src))
    close(po.residual_cut(p, 2, u.ravel(), [old.ravel()], masks),
          po.residual(p, 2, u.ravel(), [old.ravel()]))
    close(po.node_blocks_cut(p, cache, masks), po.node_blocks(p, cache))


@pytest.mark.parametrize("dim,deg,cells,tens", CASES)
def test_cut_operators_reduced_system(dim, deg, cells, tens):
    nc, g = 5, 1
    p = params(dim, deg, cells, nc, tens)
    u = state(p, 6)
    old = state(p, 7)
    src = np.random.default_rng(8).uniform(-1, 1, u.shape)
    keep = [c for c in range(nc) if c != 2 + g]
    pr = params(dim, deg, cells, nc - 1, tens)
    masks = np.full(po.n_cells(p), 0b111 & ~(1 << g), dtype=np.uint32)
    cache = po.fill_qp(p, u.ravel())
    nn = u.shape[0]

    def check(cut, red):
        cut = cut.reshape(nn, -1)
        assert not cut[:, 2 + g].any()
        np.testing.assert_allclose(cut[:, keep], red.reshape(nn, -1), rtol=0,
                                   atol=1e-13 * np.abs(red).max())

    check(po.vmult_cut(p, cache, src.ravel(), masks),
          po.vmult(pr, po.fill_qp(pr, u[:, keep].ravel()), src[:, keep].ravel()))
    check(po.residual_cut(p, 2, u.ravel(), [old.ravel()], masks),
          po.residual(pr, 2, u[:, keep].ravel(), [old[:, keep].ravel()]))
    B = po.node_blocks_cut(p, cache, masks)
    Br = po.node_blocks(pr, po.fill_qp(pr, u[:, keep].ravel()))
    assert not B[:, 2 + g, :].any() and not B[:, :, 2 + g].any()
    np.testing.assert_allclose(B[:, keep][:, :, keep], Br, rtol=0, atol=1e-13 * np.abs(Br).max())


def test_cut_vmult_c_and_python_restatements_agree():
    """random per-cell masks, 3D Q1: orc_vmult_cut against the per-cell
    Python restatement vmult_cut_py"""
    nc = 6
    p = params(3, 1, (3, 2, 3), nc, False)
    u = state(p, 9)
    src = np.random.default_rng(10).uniform(-1, 1, u.size)
    masks = np.random.default_rng(11).integers(0, 16, size=po.n_cells(p)).astype(np.uint32)
    a = po.vmult_cut(p, po.fill_qp(p, u.ravel()), src, masks)
    b = po.vmult_cut_py(p, u.ravel(), src, masks)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())
