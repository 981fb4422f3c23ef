"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, where it is the checker.  The product
library never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MAX_COMP = 16


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("degree", ctypes.c_int),
        ("n_cells", ctypes.c_int * 3),
        ("h", ctypes.c_double * 3),
        ("n_comp", ctypes.c_int),
        ("tensorial", ctypes.c_int),
        ("A", ctypes.c_double),
        ("B", ctypes.c_double),
        ("kappa_c", ctypes.c_double),
        ("kappa_p", ctypes.c_double),
        ("Mvol", ctypes.c_double),
        ("Mvap", ctypes.c_double),
        ("Msurf", ctypes.c_double),
        ("Mgb", ctypes.c_double),
        ("L", ctypes.c_double),
        ("w0", ctypes.c_double),
        ("w_alg", ctypes.c_double),
        ("order", ctypes.c_int),
        ("w_hist", ctypes.c_double * 3),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(OrcParams)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_free_energy_eval.argtypes = [ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, dp, ctypes.c_int, dp]
        L.orc_d2f_detaidetaj.argtypes = [ctypes.c_double] * 3
        L.orc_d2f_detaidetaj.restype = ctypes.c_double
        L.orc_power_sum.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.orc_power_sum.restype = ctypes.c_double
        L.orc_unit_vector.argtypes = [dp, ctypes.c_int, ctypes.c_double, dp]
        L.orc_unit_vector_filter_2.argtypes = [dp, ctypes.c_int, ctypes.c_double]
        L.orc_unit_vector_filter_2.restype = ctypes.c_double
        L.orc_residual_qp.argtypes = [P, ctypes.c_int, dp, dp, dp, dp, dp]
        L.orc_tangent_qp.argtypes = [P, dp, dp, dp, dp, dp, dp]
        L.orc_n_nodes.argtypes = [P]
        L.orc_n_nodes.restype = ctypes.c_int64
        L.orc_n_cells.argtypes = [P]
        L.orc_n_cells.restype = ctypes.c_int64
        L.orc_n_q_points.argtypes = [P]
        L.orc_residual.argtypes = [P, ctypes.c_int, dp, ctypes.POINTER(dp), dp]
        L.orc_fill_qp.argtypes = [P, dp, dp]
        L.orc_vmult.argtypes = [P, dp, dp, dp]
        L.orc_node_blocks.argtypes = [P, dp, dp]
        L.orc_dense_tangent.argtypes = [P, dp, dp]
        up = ctypes.POINTER(ctypes.c_uint32)
        L.orc_cell_masks.argtypes = [P, dp, ctypes.c_double, up]
        L.orc_vmult_cut.argtypes = [P, dp, up, dp, dp]
        L.orc_residual_cut.argtypes = [P, ctypes.c_int, up, dp, ctypes.POINTER(dp), dp]
        L.orc_node_blocks_cut.argtypes = [P, dp, up, dp]
        ip = ctypes.POINTER(ctypes.c_int64)
        L.orc_ilu0.argtypes = [ctypes.c_int64, ip, ip, dp]
        L.orc_ilu0.restype = ctypes.c_int
        L.orc_ilu0_solve.argtypes = [ctypes.c_int64, ip, ip, dp, dp, dp]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(dim, degree, n_cells, h, n_comp, tensorial=False, A=16.0,
                B=1.0, kappa_c=1.0, kappa_p=0.5, Mvol=1e-2, Mvap=1e-10,
                Msurf=4.0, Mgb=0.4, L=1.0, w0=0.0, w_alg=1.0, w_hist=()):
    p = OrcParams()
    p.dim = dim
    p.degree = degree
    for d in range(3):
        p.n_cells[d] = int(n_cells[d]) if d < dim else 1
        p.h[d] = float(h[d]) if d < dim else 1.0
    p.n_comp = n_comp
    p.tensorial = int(bool(tensorial))
    p.A, p.B, p.kappa_c, p.kappa_p = A, B, kappa_c, kappa_p
    p.Mvol, p.Mvap, p.Msurf, p.Mgb, p.L = Mvol, Mvap, Msurf, Mgb, L
    p.w0 = w0
    p.w_alg = w_alg
    p.order = len(w_hist)
    for i, w in enumerate(w_hist):
        p.w_hist[i] = w
    return p


def n_nodes(p):
    return int(lib().orc_n_nodes(ctypes.byref(p)))


def n_cells(p):
    return int(lib().orc_n_cells(ctypes.byref(p)))


def n_q_points(p):
    return int(lib().orc_n_q_points(ctypes.byref(p)))


def free_energy_eval(A, B, c, etas):
    etas = np.ascontiguousarray(etas, dtype=np.float64)
    G = len(etas)
    out = np.zeros(3 + 3 * G)
    lib().orc_free_energy_eval(A, B, c, _dp(etas), G, _dp(out))
    return {
        "f": out[0], "df_dc": out[1], "d2f_dc2": out[2],
        "df_detai": out[3:3 + G].copy(),
        "d2f_dcdetai": out[3 + G:3 + 2 * G].copy(),
        "d2f_detai2": out[3 + 2 * G:3 + 3 * G].copy(),
    }


def d2f_detaidetaj(B, ei, ej):
    return lib().orc_d2f_detaidetaj(B, ei, ej)


def power_sum(v, p):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_power_sum(_dp(v), len(v), p)


def unit_vector(v, tol=1e-10):
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros_like(v)
    lib().orc_unit_vector(_dp(v), len(v), tol, _dp(out))
    return out


def unit_vector_filter_2(v, tol=1e-20):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return lib().orc_unit_vector_filter_2(_dp(v), len(v), tol)


def residual_qp(p, wtd, value, grad, hist=None):
    nc, dim = p.n_comp, p.dim
    value = np.ascontiguousarray(value, dtype=np.float64)
    grad = np.ascontiguousarray(grad, dtype=np.float64).reshape(nc * dim)
    hist = np.zeros(nc) if hist is None else np.ascontiguousarray(hist, dtype=np.float64)
    vr = np.zeros(nc)
    gr = np.zeros(nc * dim)
    lib().orc_residual_qp(ctypes.byref(p), wtd, _dp(value), _dp(grad), _dp(hist),
                          _dp(vr), _dp(gr))
    return vr, gr.reshape(nc, dim)


def tangent_qp(p, lin_value, lin_grad, value, grad):
    nc, dim = p.n_comp, p.dim
    lv = np.ascontiguousarray(lin_value, dtype=np.float64)
    lg = np.ascontiguousarray(lin_grad, dtype=np.float64).reshape(nc * dim)
    v = np.ascontiguousarray(value, dtype=np.float64)
    g = np.ascontiguousarray(grad, dtype=np.float64).reshape(nc * dim)
    vr = np.zeros(nc)
    gr = np.zeros(nc * dim)
    lib().orc_tangent_qp(ctypes.byref(p), _dp(lv), _dp(lg), _dp(v), _dp(g),
                         _dp(vr), _dp(gr))
    return vr, gr.reshape(nc, dim)


def residual(p, wtd, src, old=()):
    src = np.ascontiguousarray(src, dtype=np.float64).ravel()
    dst = np.empty_like(src)
    olds = [np.ascontiguousarray(o, dtype=np.float64).ravel() for o in old]
    arr = (ctypes.POINTER(ctypes.c_double) * max(1, len(olds)))()
    for i, o in enumerate(olds):
        arr[i] = _dp(o)
    lib().orc_residual(ctypes.byref(p), wtd, _dp(src), arr, _dp(dst))
    return dst


def fill_qp(p, lin):
    lin = np.ascontiguousarray(lin, dtype=np.float64).ravel()
    size = n_cells(p) * n_q_points(p) * p.n_comp * (1 + p.dim)
    cache = np.empty(size)
    lib().orc_fill_qp(ctypes.byref(p), _dp(lin), _dp(cache))
    return cache


def vmult(p, cache, src):
    src = np.ascontiguousarray(src, dtype=np.float64).ravel()
    dst = np.empty_like(src)
    lib().orc_vmult(ctypes.byref(p), _dp(cache), _dp(src), _dp(dst))
    return dst


def vmult_constrained(p, cache, src, nodes):
    """OperatorBase::vmult with constrained DoFs (S/operator_base.h:294-353):
    the constrained entries of src (all components of the listed nodes) are
    zero for the cell operator, and dst equals src on them afterwards."""
    nc = p.n_comp
    s = np.array(src, dtype=np.float64).reshape(-1, nc)
    saved = s[nodes].copy()
    s[nodes] = 0.0
    dst = vmult(p, cache, s.ravel()).reshape(-1, nc)
    dst[nodes] = saved
    return dst.ravel()


def residual_constrained(p, wtd, src, old, nodes):
    """residual with constrained DoFs: FEEvaluation::distribute_local_to_global
    (deal.II, called at S/operator_sintering_generic.h:1005) skips constrained
    entries, so their rows stay zero"""
    dst = residual(p, wtd, src, old).reshape(-1, p.n_comp)
    dst[nodes] = 0.0
    return dst.ravel()


def _up(a):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def cell_masks(p, u, tol):
    """SinteringOperatorData::set_component_mask (S/operator_sintering_data.h:
    150-298) with one cell per batch: bit g when max |eta_g| over the cell's
    nodes > tol (uint32 per cell, cells x fastest)"""
    u = np.ascontiguousarray(u, dtype=np.float64).ravel()
    m = np.empty(n_cells(p), dtype=np.uint32)
    lib().orc_cell_masks(ctypes.byref(p), _dp(u), float(tol), _up(m))
    return m


def vmult_cut(p, cache, src, masks):
    """tangent vmult with the grain cut-off (each cell is the system of c, mu
    and its active order parameters, do_vmult_range_no_template,
    S/operator_sintering_generic.h:456-537); cache from fill_qp of the full
    state"""
    src = np.ascontiguousarray(src, dtype=np.float64).ravel()
    masks = np.ascontiguousarray(masks, dtype=np.uint32)
    dst = np.empty_like(src)
    lib().orc_vmult_cut(ctypes.byref(p), _dp(cache), _up(masks), _dp(src), _dp(dst))
    return dst


def residual_cut(p, wtd, src, old, masks):
    """nonlinear residual with the grain cut-off
    (do_evaluate_nonlinear_residual_nt, S/operator_sintering_generic.h:
    1008-1081)"""
    src = np.ascontiguousarray(src, dtype=np.float64).ravel()
    masks = np.ascontiguousarray(masks, dtype=np.uint32)
    dst = np.empty_like(src)
    olds = [np.ascontiguousarray(o, dtype=np.float64).ravel() for o in old]
    arr = (ctypes.POINTER(ctypes.c_double) * max(1, len(olds)))()
    for i, o in enumerate(olds):
        arr[i] = _dp(o)
    lib().orc_residual_cut(ctypes.byref(p), wtd, _up(masks), _dp(src), arr, _dp(dst))
    return dst


def node_blocks_cut(p, cache, masks):
    """node blocks of the cut tangent"""
    masks = np.ascontiguousarray(masks, dtype=np.uint32)
    nn = n_nodes(p)
    out = np.empty(nn * p.n_comp * p.n_comp)
    lib().orc_node_blocks_cut(ctypes.byref(p), _dp(cache), _up(masks), _dp(out))
    return out.reshape(nn, p.n_comp, p.n_comp)


def vmult_cut_py(p, lin, src, cell_masks):
    """tangent vmult with a grain cut-off, do_vmult_range_no_template
    (S/operator_sintering_generic.h:456-537): each cell is evaluated with c,
    mu and only the order parameters whose bit is set in cell_masks[cell]
    (cells x fastest); the others get no contribution from the cell and do
    not enter its physics.  Q1, 3D, 2-point Gauss; the quadrature-point
    tangent is orc_tangent_qp on the reduced component set."""
    assert p.dim == 3 and p.degree == 1
    nc = p.n_comp
    ncx, ncy, ncz = p.n_cells[0], p.n_cells[1], p.n_cells[2]
    h = [p.h[0], p.h[1], p.h[2]]
    nnx, nny = ncx + 1, ncy + 1
    lin = np.asarray(lin, dtype=np.float64).reshape(-1, nc)
    src = np.asarray(src, dtype=np.float64).reshape(-1, nc)
    dst = np.zeros_like(src)
    xg = [(1.0 - 1.0 / np.sqrt(3.0)) / 2.0, (1.0 + 1.0 / np.sqrt(3.0)) / 2.0]
    corners = [(a, b, c) for c in (0, 1) for b in (0, 1) for a in (0, 1)]
    # shape values and gradients of the 8 corners at the 8 points
    S = np.zeros((8, 8))
    Gr = np.zeros((8, 8, 3))
    for q, (qa, qb, qc) in enumerate(corners):
        x = (xg[qa], xg[qb], xg[qc])
        for k, (a, b, c) in enumerate(corners):
            f = [x[d] if (a, b, c)[d] else 1.0 - x[d] for d in range(3)]
            df = [(1.0 if (a, b, c)[d] else -1.0) / h[d] for d in range(3)]
            S[q, k] = f[0] * f[1] * f[2]
            Gr[q, k] = [df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]]
    jxw = h[0] * h[1] * h[2] / 8.0
    for cz in range(ncz):
        for cy in range(ncy):
            for cx in range(ncx):
                cell = cx + ncx * (cy + ncy * cz)
                m = int(cell_masks[cell])
                L = [0, 1] + [2 + g for g in range(nc - 2) if (m >> g) & 1]
                p2 = type(p).from_buffer_copy(p)
                p2.n_comp = len(L)
                nodes = [(cx + a) + nnx * ((cy + b) + nny * (cz + c)) for (a, b, c) in corners]
                ul = lin[nodes][:, L]
                us = src[nodes][:, L]
                for q in range(8):
                    lv = S[q] @ ul
                    lg = (Gr[q].T @ ul).T
                    v = S[q] @ us
                    g = (Gr[q].T @ us).T
                    vr, gr = tangent_qp(p2, lv, lg, v, g)
                    contrib = jxw * (np.outer(S[q], vr) + Gr[q] @ gr.T)
                    for k, nd in enumerate(nodes):
                        dst[nd, L] += contrib[k]
    return dst.ravel()


def node_blocks(p, cache):
    nn = n_nodes(p)
    out = np.empty(nn * p.n_comp * p.n_comp)
    lib().orc_node_blocks(ctypes.byref(p), _dp(cache), _dp(out))
    return out.reshape(nn, p.n_comp, p.n_comp)


def dense_tangent(p, cache):
    n = n_nodes(p) * p.n_comp
    out = np.empty(n * n)
    lib().orc_dense_tangent(ctypes.byref(p), _dp(cache), _dp(out))
    return out.reshape(n, n)


def max_threads():
    return int(lib().orc_max_threads())


def set_threads(n):
    lib().orc_set_threads(int(n))


class ILU0:
    """ILU(0) of a scipy CSR matrix (the factors and the solve in C)"""

    def __init__(self, A):
        A = A.tocsr().copy()
        A.sort_indices()
        self.n = A.shape[0]
        self.rowptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        self.col = np.ascontiguousarray(A.indices, dtype=np.int64)
        self.val = np.ascontiguousarray(A.data, dtype=np.float64)
        ip = ctypes.POINTER(ctypes.c_int64)
        self._r = self.rowptr.ctypes.data_as(ip)
        self._c = self.col.ctypes.data_as(ip)
        rc = lib().orc_ilu0(self.n, self._r, self._c, _dp(self.val))
        if rc != 0:
            raise ValueError("ILU(0): zero or missing pivot")

    def solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        lib().orc_ilu0_solve(self.n, self._r, self._c, _dp(self.val), _dp(b), _dp(x))
        return x
