"""Host-side mirror of hpsint's SinteringOperatorGeneric over the C ABI.

Same operator concept as the reference (S/operator_sintering_generic.h:324-
1085 with S/ = applications/sintering/include/pf-applications/sintering/):
``evaluate_nonlinear_residual(dst, src, with_time_derivative)``,
``vmult(dst, src)``, ``initialize_dof_vector``, ``n_components``,
``n_grains``, ``equation_type``, ``compute_inverse_diagonal`` and the
linearisation setup of SinteringOperatorData.  Vectors are torch tensors on
the GPU in the component-interleaved layout (entry = node * n_components +
component); torch is used for device memory and the stream only, all
arithmetic runs in the HIP kernels of libhpsint_amd.so.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _ffi
from ._ffi import check, lib


@dataclass
class FreeEnergyParams:
    """Material.EnergyAbstract (S/parameters.h:113-118)"""
    A: float = 16.0
    B: float = 1.0
    kappa_c: float = 1.0
    kappa_p: float = 0.5


@dataclass
class MobilityParams:
    """Material.MobilityAbstract (S/parameters.h:120-134)"""
    Mvol: float = 1e-2
    Mvap: float = 1e-10
    Msurf: float = 4.0
    Mgb: float = 0.4
    L: float = 1.0
    tensorial: bool = False


@dataclass
class CartesianMesh:
    """Uniform Cartesian mesh ("Divisions" mode, S/driver.h:575-590)."""
    n_cells: tuple
    lower: tuple
    upper: tuple

    @property
    def dim(self):
        return len(self.n_cells)

    @property
    def cell_size(self):
        return tuple((u - l) / n for l, u, n in zip(self.lower, self.upper, self.n_cells))

    @property
    def n_nodes_per_dim(self):
        return tuple(n + 1 for n in self.n_cells)

    @property
    def n_nodes(self):
        p = 1
        for n in self.n_nodes_per_dim:
            p *= n
        return p

    def node_coordinates(self, device=None, dtype=torch.float64):
        """(n_nodes, dim) coordinates, x fastest"""
        axes = [torch.linspace(l, u, n + 1, dtype=dtype, device=device)
                for l, u, n in zip(self.lower, self.upper, self.n_cells)]
        grids = torch.meshgrid(*reversed(axes), indexing="ij")
        return torch.stack([g.reshape(-1) for g in reversed(grids)], dim=1)


class TimeIntegratorData:
    """BDF weights, I/time_integration/time_schemes.h:65-148 and
    time_integrator_data.h:25-227 (variable step sizes, effective order
    limited by the number of valid past steps)."""

    def __init__(self, order=1, dt=1e-3):
        self.order = order
        self.dt = [dt] + [0.0] * (order - 1)
        self.weights = [0.0] * (order + 1)
        self.update_weights()

    def effective_order(self):
        return sum(1 for d in self.dt if d > 0.0)

    def update_weights(self):
        dt = self.dt
        w = [0.0] * (self.order + 1)
        eo = self.effective_order()
        if eo >= 3 and self.order >= 3:
            w[1] = -(dt[0] + dt[1]) * (dt[0] + dt[1] + dt[2]) / (dt[0] * dt[1] * (dt[1] + dt[2]))
            w[2] = dt[0] * (dt[0] + dt[1] + dt[2]) / (dt[1] * dt[2] * (dt[0] + dt[1]))
            w[3] = -dt[0] * (dt[0] + dt[1]) / (dt[2] * (dt[1] + dt[2]) * (dt[0] + dt[1] + dt[2]))
            w[0] = -(w[1] + w[2] + w[3])
        elif eo >= 2 and self.order >= 2:
            w[0] = (2 * dt[0] + dt[1]) / (dt[0] * (dt[0] + dt[1]))
            w[1] = -(dt[0] + dt[1]) / (dt[0] * dt[1])
            w[2] = dt[0] / (dt[1] * (dt[0] + dt[1]))
        else:
            w[0] = 1.0 / dt[0]
            w[1] = -1.0 / dt[0]
        self.weights = w

    def update_dt(self, dt_new):
        """time_integrator_data.h:101-109"""
        for i in range(self.order - 2, -1, -1):
            self.dt[i + 1] = self.dt[i]
        self.dt[0] = dt_new
        self.update_weights()

    def replace_dt(self, dt_new):
        self.dt[0] = dt_new
        self.update_weights()

    def get_primary_weight(self):
        return self.weights[0]

    def get_algebraic_weight(self):
        return 1.0

    def history_weights(self):
        """weights[1..order] restricted to the effective order"""
        eo = min(self.order, self.effective_order())
        return self.weights[1:1 + eo]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class SinteringOperatorGeneric:
    """Residual and tangent of the coupled CH/AC system (c, mu, eta_i)."""

    def __init__(self, mesh: CartesianMesh, n_grains: int,
                 free_energy: FreeEnergyParams = None,
                 mobility: MobilityParams = None, device="cuda", fe_degree=1):
        """fe_degree 1: FE_Q(1) with QGauss(2); 2: FE_Q(2) with QGauss(3)"""
        free_energy = free_energy or FreeEnergyParams()
        mobility = mobility or MobilityParams()
        self.mesh = mesh
        self.fe_degree = fe_degree
        self.device = torch.device(device)
        p = _ffi.HpsParams()
        p.dim = mesh.dim
        p.fe_degree = fe_degree
        for d in range(3):
            p.n_cells[d] = mesh.n_cells[d] if d < mesh.dim else 1
            p.cell_size[d] = mesh.cell_size[d] if d < mesh.dim else 1.0
        p.n_components = n_grains + 2
        p.mobility = (_ffi.HPS_MOBILITY_TENSORIAL if mobility.tensorial
                      else _ffi.HPS_MOBILITY_SCALAR)
        p.A, p.B = free_energy.A, free_energy.B
        p.kappa_c, p.kappa_p = free_energy.kappa_c, free_energy.kappa_p
        p.Mvol, p.Mvap, p.Msurf, p.Mgb, p.L = (mobility.Mvol, mobility.Mvap, mobility.Msurf,
                                                mobility.Mgb, mobility.L)
        self.params = p
        h = ctypes.c_void_p()
        check(lib().hps_create(ctypes.byref(p), ctypes.byref(h)))
        self._h = h
        nn, nd = ctypes.c_int64(), ctypes.c_int64()
        check(lib().hps_sizes(h, ctypes.byref(nn), ctypes.byref(nd)))
        self._n_nodes, self._n_dofs = nn.value, nd.value
        self._bind_stream()
        self.time_data = None
        self._order = 0
        self._old = []
        self.cut_off_tolerance = 0.0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _ffi._LIB is not None:
            lib().hps_destroy(h)
            self._h = None

    # -- plumbing
    def _bind_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().hps_set_stream(self._h, ctypes.c_void_p(s)))

    @property
    def handle(self):
        return self._h

    def sync(self):
        check(lib().hps_sync(self._h))

    # -- the operator concept
    def n_components(self):
        return self.params.n_components

    def n_grains(self):
        return self.params.n_components - 2

    @staticmethod
    def n_grains_to_n_components(n_grains):
        return n_grains + 2

    def equation_type(self, component):
        """S/operator_sintering_generic.h:634-639 (mu is stationary)"""
        return "Stationary" if component == 1 else "TimeDependent"

    @property
    def n_nodes(self):
        return self._n_nodes

    @property
    def n_dofs(self):
        return self._n_dofs

    def initialize_dof_vector(self, dtype=torch.float64):
        return torch.zeros(self._n_dofs, dtype=dtype, device=self.device)

    def set_time_data(self, time_data: TimeIntegratorData):
        self.time_data = time_data
        w = time_data.history_weights()
        arr = (ctypes.c_double * 3)(*(list(w) + [0.0] * (3 - len(w))))
        check(lib().hps_set_time_weights(self._h, time_data.get_primary_weight(),
                                         time_data.get_algebraic_weight(), arr, len(w)))
        self._order = len(w)

    def set_time_weights(self, w0, w_hist=(), w_alg=1.0):
        arr = (ctypes.c_double * 3)(*(list(w_hist) + [0.0] * (3 - len(w_hist))))
        check(lib().hps_set_time_weights(self._h, w0, w_alg, arr, len(w_hist)))
        self._order = len(w_hist)

    @property
    def history_order(self):
        """number of old solutions the current time weights read"""
        return self._order

    def check_history(self, old):
        """raises unless `old` holds a vector for every history weight"""
        if len(old) < self._order:
            raise ValueError("%d old solutions given, the time weights need %d"
                             % (len(old), self._order))

    def set_old_solutions(self, old):
        self._old = list(old)

    def fill_quadrature_point_values(self, u):
        """SinteringOperatorData::fill_quadrature_point_values"""
        check(lib().hps_set_linearization(self._h, _ptr(u)))

    def evaluate_nonlinear_residual(self, dst, src, with_time_derivative=2, old=None):
        old = self._old if old is None else old
        if with_time_derivative == 2:
            self.check_history(old)
        arr = (ctypes.c_void_p * max(1, len(old)))(*[o.data_ptr() for o in old])
        check(lib().hps_residual(self._h, with_time_derivative, _ptr(dst), _ptr(src), arr))

    def set_constrained_nodes(self, nodes):
        """constrained nodes (OperatorBase::vmult's constrained_indices from
        MatrixFree::get_constrained_dofs, S/operator_base.h:300-337): vmult
        treats their src entries as zero and returns dst = src on them,
        the residual leaves their rows zero; an empty list removes them"""
        arr = np.ascontiguousarray(np.asarray(nodes, dtype=np.int64).ravel())
        check(lib().hps_set_constrained_nodes(
            self._h, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(arr.size)))

    def set_grain_cut_off(self, tol):
        """grain cut-off tolerance (GrainCutOffTolerance, S/parameters.h:314):
        tol > 0 enables the cut-off, tol <= 0 disables it; the component
        table comes from set_component_mask"""
        check(lib().hps_set_grain_cut_off(self._h, float(tol)))
        self.cut_off_tolerance = float(tol) if tol > 0 else 0.0

    def set_component_mask(self, u):
        """SinteringOperatorData::set_component_mask (S/operator_sintering_
        data.h:150-298): per cell, the order parameters with max |eta| over
        the cell's nodes above the tolerance stay in the cell's residual,
        tangent and preconditioner; the others are dropped.  The
        linearisation must be set again afterwards."""
        check(lib().hps_set_component_mask(self._h, _ptr(u)))

    def component_mask(self):
        """the component table: one int32 bit mask per cell (cells x fastest)"""
        n = 1
        for c in self.mesh.n_cells:
            n *= c
        out = torch.empty(n, dtype=torch.int32, device=self.device)
        check(lib().hps_get_component_mask(self._h, _ptr(out)))
        return out

    def vmult(self, dst, src):
        if src.dtype == torch.float32:
            check(lib().hps_vmult_f32(self._h, _ptr(dst), _ptr(src)))
        else:
            check(lib().hps_vmult_f64(self._h, _ptr(dst), _ptr(src)))

    def compute_inverse_diagonal(self, diagonal):
        check(lib().hps_precond_update(self._h, _ffi.HPS_PRECON_POINT_JACOBI))
        check(lib().hps_precond_get_diagonal(self._h, _ptr(diagonal)))

    def node_blocks(self):
        """n_nodes x n_c x n_c tangent blocks (before inversion)"""
        check(lib().hps_precond_update(self._h, _ffi.HPS_PRECON_NODE_BLOCK_JACOBI))
        nc = self.n_components()
        out = torch.empty(self._n_nodes * nc * nc, dtype=torch.float64, device=self.device)
        check(lib().hps_precond_get_blocks(self._h, _ptr(out)))
        return out.view(self._n_nodes, nc, nc)

    def norm(self, x):
        """l2 norm of a device vector (DynamicBlockVector::l2_norm)"""
        r = ctypes.c_double()
        check(lib().hps_norm2(self._h, x.numel(), _ptr(x), ctypes.byref(r)))
        return r.value

    def enable_timing(self, on=True, ring=1):
        """HIP-event timing of the main kernel and the fix-up of every vmult and
        residual, kept for the last `ring` calls without waiting"""
        check(lib().hps_enable_timing(self._h, int(ring) if on else 0))

    def kernel_ms_mean(self):
        """(main, fix-up, count): mean kernel durations in ms over the calls
        recorded since timing was enabled (at most `ring` of them)"""
        a, b, n = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        check(lib().hps_kernel_ms_mean(self._h, ctypes.byref(a), ctypes.byref(b),
                                       ctypes.byref(n)))
        return a.value, b.value, n.value

    def last_kernel_ms(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        check(lib().hps_last_kernel_ms(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Preconditioner:
    """PreconditionerBase (I/lac/preconditioners.h:77-112) over the
    operator's data: "InverseDiagonalMatrix", "InverseComponentBlockDiagonalMatrix"
    or "BlockPreconditioner2" (ILU(0) of the Cahn-Hilliard block in parity-
    colour order, inverse diagonal on the order parameters)."""

    KINDS = {"InverseDiagonalMatrix": _ffi.HPS_PRECON_POINT_JACOBI,
             "InverseComponentBlockDiagonalMatrix": _ffi.HPS_PRECON_NODE_BLOCK_JACOBI,
             "BlockPreconditioner2": _ffi.HPS_PRECON_BLOCK_CH_ILU}

    def __init__(self, op: SinteringOperatorGeneric, kind="InverseDiagonalMatrix"):
        self.op = op
        self.kind = self.KINDS[kind]

    def do_update(self, fp32=False):
        """fp32: assemble from the FP32 cache in FP32 arithmetic (the
        preconditioner of the FP32 inner solve)"""
        if fp32:
            check(lib().hps_precond_update_f32(self.op.handle, self.kind))
        else:
            check(lib().hps_precond_update(self.op.handle, self.kind))

    def vmult(self, dst, src):
        if src.dtype == torch.float32:
            check(lib().hps_precond_apply_f32(self.op.handle, _ptr(dst), _ptr(src)))
        else:
            check(lib().hps_precond_apply_f64(self.op.handle, _ptr(dst), _ptr(src)))
