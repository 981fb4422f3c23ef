"""Newton-Krylov and time stepping over the C ABI.

``NewtonSolver`` wraps hps_newton_* (the C++ restatement of
DampedNewtonSolver + SolverGMRESWrapper, I/lac/solvers_nonlinear.h:248-357,
I/lac/solvers_linear.h:58-117, with I/ = include/pf-applications/).
``TimeStepper`` is the implicit time loop of the driver for fixed steps
(S/driver.h:1932-2700 with S/ = applications/sintering/include/
pf-applications/sintering/): set time weights, predict the initial guess
(the predictor slot of S/driver.h:2002-2050), Newton solve, commit the old
solutions (TimeIntegration::SolutionHistory).  On non-convergence the step
is retried with dt/2 (process_failure, S/driver.h:2163-2200).
"""
import ctypes
import time
from dataclasses import dataclass

import torch

from . import _ffi
from ._ffi import check, lib
from .operator import SinteringOperatorGeneric, TimeIntegratorData


class NewtonSolver:
    PRECONDITIONERS = {"InverseDiagonalMatrix": _ffi.HPS_PRECON_POINT_JACOBI,
                       "InverseComponentBlockDiagonalMatrix": _ffi.HPS_PRECON_NODE_BLOCK_JACOBI,
                       "BlockPreconditioner2": _ffi.HPS_PRECON_BLOCK_CH_ILU}

    def __init__(self, op: SinteringOperatorGeneric, nl_rel_tol=1e-5, nl_abs_tol=1e-20,
                 nl_max_iter=10, l_rel_tol=1e-2, l_abs_tol=1e-10, l_max_iter=1000,
                 l_restart=30, preconditioner="InverseComponentBlockDiagonalMatrix",
                 use_damping=True, threshold_newton_iter=100, threshold_linear_iter=20,
                 inner_fp32=False):
        self.op = op
        p = _ffi.HpsNewtonParams()
        lib().hps_newton_default_params(ctypes.byref(p))
        p.nl_rel_tol, p.nl_abs_tol, p.nl_max_iter = nl_rel_tol, nl_abs_tol, nl_max_iter
        p.l_rel_tol, p.l_abs_tol, p.l_max_iter = l_rel_tol, l_abs_tol, l_max_iter
        p.l_restart = l_restart
        p.precond = self.PRECONDITIONERS[preconditioner]
        p.use_damping = int(use_damping)
        p.threshold_newton_iter = threshold_newton_iter
        p.threshold_linear_iter = threshold_linear_iter
        # inner_fp32: False/0 FP64; True/1 FP32 tangent, preconditioner and
        # Krylov basis; 2 FP32 tangent and preconditioner with an FP64 basis
        p.inner_fp32 = int(inner_fp32)
        self.params = p
        h = ctypes.c_void_p()
        check(lib().hps_newton_create(op.handle, ctypes.byref(p), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _ffi._LIB is not None:
            lib().hps_newton_destroy(h)
            self._h = None

    def clear(self):
        check(lib().hps_newton_clear(self._h))

    def solve(self, x, old, r0_ref=-1.0):
        """solves R(x) = 0 in place; returns the statistics.  r0_ref > 0 is
        the reference norm of the relative criterion (default: |R(x_0)|)"""
        self.op.check_history(old)
        arr = (ctypes.c_void_p * max(1, len(old)))(*[o.data_ptr() for o in old])
        st = _ffi.HpsNewtonStats()
        rc = lib().hps_newton_solve_ref(self._h, ctypes.c_void_p(x.data_ptr()), arr,
                                        float(r0_ref), ctypes.byref(st))
        self.last_stats = st
        check(rc)
        return st


@dataclass
class StepRecord:
    t: float
    dt: float
    newton_iterations: int
    linear_iterations: int
    residual_evaluations: int
    seconds: float
    predictor_seconds: float = 0.0


class TimeStepper:
    """Fixed-step implicit BDF time loop with a pluggable predictor."""

    def __init__(self, op: SinteringOperatorGeneric, solver: NewtonSolver,
                 order=1, dt=1e-3, predictor=None, max_failures=5):
        self.op = op
        self.solver = solver
        self.td = TimeIntegratorData(order=order, dt=dt)
        self.dt = dt
        self.order = order
        self.predictor = predictor
        self.max_failures = max_failures
        self.history = []     # old solutions, most recent first
        self.t = 0.0
        self.records = []
        self.keep_guess = False

    def initialize(self, u0):
        self.u = u0.clone()
        self.history = [u0.clone()]   # committed solutions, most recent first

    def make_step(self):
        dt = self.dt
        failures = 0
        while True:
            self.td.replace_dt(dt)
            self.op.set_time_data(self.td)
            eo = min(self.order, len(self.history))
            old = self.history[:eo]
            t0 = time.perf_counter()
            x = self.u.clone()
            tp = 0.0
            r0_ref = -1.0
            if self.predictor is not None:
                tp0 = time.perf_counter()
                # the plain method's initial residual |R(u_n)| stays the
                # reference of the relative Newton criterion
                r = self._scratch if getattr(self, "_scratch", None) is not None else x.clone()
                self._scratch = r
                self.op.evaluate_nonlinear_residual(r, x, 2, old=old)
                r0_ref = self.op.norm(r)
                self.predictor(x, self.history, dt)
                torch.cuda.synchronize()
                tp = time.perf_counter() - tp0
            self.last_guess = x.clone() if self.keep_guess else None
            if self.op.cut_off_tolerance > 0.0:
                # the component table of the step's initial guess
                # (S/driver.h:2212-2218)
                self.op.set_component_mask(x)
            try:
                st = self.solver.solve(x, old, r0_ref)
            except (_ffi.NewtonDidNotConverge, _ffi.LinearSolverNoConvergence):
                failures += 1
                if failures > self.max_failures:
                    raise
                dt *= 0.5
                self.solver.clear()
                continue
            torch.cuda.synchronize()
            sec = time.perf_counter() - t0
            break
        self.t += dt
        self.history = [x.clone()] + self.history[:max(self.order, 2) - 1]
        self.u = x
        self.td.update_dt(dt)
        rec = StepRecord(self.t, dt, st.newton_iterations, st.linear_iterations,
                         st.residual_evaluations, sec, tp)
        self.records.append(rec)
        return rec


class ExtrapolationPredictor:
    """The reference's predictors of the Newton initial guess
    (Parameters::TimeIntegrationData::predictor, S/driver.h:2002-2050):

    - "Euler": u + (-dt) R_0(u), one explicit Euler step with the residual
      without time terms (evaluate_nonlinear_residual<0>) and sadd;
    - "Midpoint": the explicit midpoint rule, two such residuals;
    - "Linear": u_n + (dt_n / dt_{n-1}) (u_n - u_{n-1}), the solution
      history's extrapolation (SolutionHistory::extrapolate).

    All vector work runs in the library (residual, hps_sadd, hps_axpy).
    Called by TimeStepper as predictor(x, states, dt) with x = u_n on entry
    and states = [u_n, u_{n-1}, ...]."""

    def __init__(self, op, kind="Euler"):
        if kind not in ("Euler", "Midpoint", "Linear"):
            raise ValueError("predictor must be Euler, Midpoint or Linear")
        self.op = op
        self.kind = kind
        self._r = None
        self._m = None
        self.dt_prev = None

    def _buf(self, name, like):
        b = getattr(self, name)
        if b is None or b.numel() != like.numel():
            b = torch.empty_like(like)
            setattr(self, name, b)
        return b

    def __call__(self, x, states, dt):
        h, n = self.op.handle, x.numel()
        p = ctypes.c_void_p
        if len(states) < 2:
            # the first step (t = 0) is not predicted
            self.dt_prev = dt
            return x
        if self.kind == "Linear":
            fac = dt / (self.dt_prev or dt)
            self.dt_prev = dt
            # x = u_n + fac (u_n - u_{n-1})
            check(lib().hps_sadd(h, n, 1.0 + fac, -fac, p(states[1].data_ptr()),
                                 p(x.data_ptr())))
            return x
        r = self._buf("_r", x)
        # evaluate_nonlinear_residual<0> has no time terms: the weights do not matter
        if self.kind == "Euler":
            self.op.evaluate_nonlinear_residual(r, x, 0)
            check(lib().hps_sadd(h, n, -dt, 1.0, p(x.data_ptr()), p(r.data_ptr())))
        else:
            m = self._buf("_m", x)
            self.op.evaluate_nonlinear_residual(m, x, 0)
            check(lib().hps_sadd(h, n, -dt / 2.0, 1.0, p(x.data_ptr()), p(m.data_ptr())))
            self.op.evaluate_nonlinear_residual(r, m, 0)
            check(lib().hps_sadd(h, n, -dt, 1.0, p(x.data_ptr()), p(r.data_ptr())))
        check(lib().hps_copy(h, n, p(x.data_ptr()), p(r.data_ptr())))
        return x
