"""Benchmark of the sintering-operator hot path on one MI355X.

Workload of the headline line: 3D 2x2x2 packing of 8 spheres, R = 10 (1 +
U(-0.1,0.1)), contact overlap 0.5, Q1 on a uniform 128^3 hex mesh (2.1M
cells), 4 order parameters -> 6 components (12.9M DoF), A=16, B=1,
kappa_c=1, kappa_p=0.5, Mvol=1e-2, Mvap=1e-10, Msurf=4, Mgb=0.4, L=1, BDF1
dt=1e-3.

A "step" is one FP64 tangent vmult (hps_vmult_f64) with the linearisation
cache filled at the initial fields and src = U(-1,1) (seed 1234), inputs
resident in HBM.  value = DoF * steps / time in GDoF/s.  Before the W
warm-up steps the vmult runs for --settle seconds: the GPU's clocks dip
during the first ~20 ms of sustained load and recover, and the K timed
steps measure the steady state.  Also reported: the dominant kernel's
roofline from HIP events on the operator's stream over K back-to-back
vmults, the FP64 residual and FP32 vmult, the CPU baseline (a SIMD/OpenMP
restatement of the same vmult on the host cores), the Q2 tensorial and
27-sphere mixed-precision workloads (with the grain cut-off on and off),
and seconds per timestep of plain FP64 Newton-GMRES against mixed precision
with the reference's Euler predictor and with the learned predictor, the
latter trained on the 100-sphere packings of seeds 0-3 and timed on the
packing of seed 4, which it has not seen.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the FP64 vmult main kernel on the 128^3, 4 order
# parameter workload: FETCH_SIZE and WRITE_SIZE (rocprofv3 --pmc), each
# scaled by the same counter's reading for a copy of known size; counts and
# factors in profiles/vmult_q1_128cubed_hbm_traffic.md (z-column kernel).
# Valid for that workload only.
VMULT_TRAFFIC_128 = 1.3612e9
# FP64 flop per launch of the residual main kernel (z-column lanes) on the
# same workload (SQ_INSTS_VALU_{ADD,MUL,FMA}_F64 x 64 lanes, FMA = 2;
# profiles/q1_kernels_128cubed_fp64_counters.md); FP64 vector peak
RESIDUAL_FLOP_128 = 7.858e9
FP64_PEAK_TFLOPS = 78.6


# the reference's default block preconditioner (ILU(0) on the Cahn-Hilliard
# block, inverse diagonal on the Allen-Cahn block)
BP2 = "BlockPreconditioner2"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--settle", type=float, default=0.5,
                    help="seconds of untimed vmults before the warm-up (clock settling)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="budget of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--timesteps", type=int, default=5,
                    help="timed steps per method of the timestep comparisons (0 = skip)")
    ap.add_argument("--train-iterations", type=int, default=6000)
    ap.add_argument("--q2-cells", type=int, default=96,
                    help="cells per direction of the Q2 tensorial workload (0 = skip)")
    ap.add_argument("--cut-tol", type=float, default=1e-5,
                    help="grain cut-off tolerance timed on the 27-sphere workload (<=0: not)")
    ap.add_argument("--q2-timesteps", type=int, default=2,
                    help="timed Newton steps per method on the Q2 workload (0: none)")
    ap.add_argument("--mp-cells", type=int, default=192,
                    help="cells per direction of the 27-sphere mixed-precision workload "
                         "(0 = skip)")
    ap.add_argument("--mp-timesteps", type=int, default=2,
                    help="timed Newton steps per method on the 27-sphere workload (0: none)")
    ap.add_argument("--study-cells", type=int, default=160,
                    help="cells per direction of the held-out predictor comparison "
                         "(100-sphere packings, 12 order parameters; 0 = skip)")
    ap.add_argument("--study-seeds", type=int, default=4,
                    help="packings (seeds 0..k-1) whose plain trajectories train; seed k "
                         "is the held-out packing")
    ap.add_argument("--study-train-steps", type=int, default=8)
    ap.add_argument("--predictor-study", action="store_true",
                    help="only run the held-out predictor comparison and exit")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run hps_copy of this many doubles 5 times (calibrates "
                         "the HBM byte counters against a copy of known size) and exit")
    return ap.parse_args()


def progress(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def timed(fn, k):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / k


def kernel_times(op, fn, k):
    """mean durations (s) of the main kernel and the fix-up of fn's operator
    calls, from HIP events on the operator's stream over k back-to-back
    calls (no waits between them)"""
    op.enable_timing(True, ring=k)
    for _ in range(k):
        fn()
    a, b, n = op.kernel_ms_mean()
    op.enable_timing(False)
    return a * 1e-3, b * 1e-3, n


def settle(fn, seconds):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(20):
            fn()
        torch.cuda.synchronize()
        n += 20
    return n


def copy_probe(n):
    """5 plain copies of n doubles (8 n bytes read + 8 n written each)"""
    import ctypes
    import torch
    from hpsint_amd import CartesianMesh, SinteringOperatorGeneric
    from hpsint_amd._ffi import check, lib
    op = SinteringOperatorGeneric(CartesianMesh((4, 4, 4), (0, 0, 0), (4.0, 4.0, 4.0)), 1)
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        check(lib().hps_copy(op.handle, n, ctypes.c_void_p(y.data_ptr()),
                             ctypes.c_void_p(x.data_ptr())))
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_doubles": n, "bytes_read": 8 * n, "bytes_written": 8 * n}))


def summary(rs):
    return {"s_per_step": sum(r.seconds for r in rs) / len(rs),
            "newton": [r.newton_iterations for r in rs],
            "linear": [r.linear_iterations for r in rs]}


def predictor_study(args):
    """Held-out timestep comparison on the 100-sphere packings (lognormal
    radii mean 8, sigma 0.2, 12 order parameters, Q1 study_cells^3 in a
    160^3 box, BDF1 dt = 1e-3): the plain FP64 trajectories of the packings
    with seeds 0..k-1 train one network; on the packing with seed k, after 3
    plain steps, `timesteps` steps are timed per method at Newton rtol 1e-8
    (relative to |R(u_n)| in every method):
      fp64_plain    FP64 Newton-GMRES, no predictor (the reference setup);
      mp_learned    FP32 inner Krylov (FP32 basis) + learned predictor;
      mp_only       FP32 inner Krylov, no predictor;
      fp64_learned  FP64 + learned predictor;
      mp_euler      FP32 inner Krylov + the reference's Euler predictor
                    (S/driver.h:2005-2011);
      mp64_learned  FP32 tangent and preconditioner with an FP64 Krylov basis
                    + learned predictor;
    and, with the reference's default block preconditioner
    (BlockPreconditioner2: ILU(0) of the Cahn-Hilliard block, inverse
    diagonal of the Allen-Cahn block) in place of the node-block Jacobi:
      fp64_bp2, mp_bp2, mp_learned_bp2."""
    import torch
    from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,
                            SinteringOperatorGeneric, packing)
    from hpsint_amd.predictor_training import collect_trajectory, fit_predictor
    from hpsint_amd.solvers import ExtrapolationPredictor, NewtonSolver, TimeStepper
    n = args.study_cells
    G = 12
    dt = 1e-3
    pk0 = packing.packing_cfg4(seed=0)
    mesh = CartesianMesh((n, n, n), pk0.lower, pk0.upper)
    op = SinteringOperatorGeneric(mesh, G, FreeEnergyParams(), MobilityParams())

    def fields(seed):
        pk = packing.packing_cfg4(seed=seed)
        assert pk.n_order_params <= G
        return torch.from_numpy(packing.initial_fields(pk, (n, n, n), n_order_params=G)).cuda()

    t0 = time.perf_counter()
    trajs = []
    for s in range(args.study_seeds):
        trajs.append(collect_trajectory(op, fields(s), n_steps=args.study_train_steps, dt=dt,
                                        nl_rel_tol=1e-8))
        progress("predictor: training packing %d, %d plain steps done"
                 % (s, args.study_train_steps))
    t_traj = time.perf_counter() - t0
    t0 = time.perf_counter()
    pred = fit_predictor(op, trajs, dt, iterations=args.train_iterations,
                         rows_per_step=min(131072, op.n_nodes), seed=0,
                         column_weights=(3.0, 1.0, 1.0, 1.0))
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    progress("predictor: network trained in %.1f s" % t_fit)
    del trajs
    torch.cuda.empty_cache()
    held = collect_trajectory(op, fields(args.study_seeds), n_steps=3, dt=dt, nl_rel_tol=1e-8)

    def run(fp32, predictor, prec="InverseComponentBlockDiagonalMatrix"):
        solver = NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2, preconditioner=prec,
                              inner_fp32=fp32)
        s = TimeStepper(op, solver, order=1, dt=dt, predictor=predictor)
        s.initialize(held[3])
        s.history = [held[3].clone(), held[2].clone()]
        s.t = 3 * dt
        rs = [s.make_step() for _ in range(args.timesteps)]
        progress("predictor: held-out steps done (inner_fp32 %s, predictor %s, %s)"
                 % (fp32, type(predictor).__name__ if predictor else None, prec))
        return summary(rs)

    euler = ExtrapolationPredictor(op, "Euler")
    out = {"workload": "100-sphere packings, Q1 %d^3, %d order parameters (%d DoF); one "
                       "network trained on the plain trajectories of packings 0-%d (%d steps "
                       "each), %d steps timed on packing %d (held out)"
                       % (n, G, op.n_dofs, args.study_seeds - 1, args.study_train_steps,
                          args.timesteps, args.study_seeds),
           "fp64_plain": run(0, None), "mp_learned": run(1, pred), "mp_only": run(1, None),
           "fp64_learned": run(0, pred), "mp_euler": run(1, euler),
           "mp64_learned": run(2, pred),
           "fp64_bp2": run(0, None, BP2), "mp_bp2": run(1, None, BP2),
           "mp_learned_bp2": run(1, pred, BP2),
           "trajectory_s": t_traj, "predictor_training_s": t_fit,
           "predictor_final_loss": pred.losses[-1]}
    base = out["fp64_plain"]["s_per_step"]
    for k in ("mp_learned", "mp_only", "fp64_learned", "mp_euler", "mp64_learned",
              "fp64_bp2", "mp_bp2", "mp_learned_bp2"):
        out["speedup_%s_vs_fp64" % k] = base / out[k]["s_per_step"]
    # both sides with the reference's default block preconditioner
    out["speedup_mp_learned_bp2_vs_fp64_bp2"] = (out["fp64_bp2"]["s_per_step"]
                                                 / out["mp_learned_bp2"]["s_per_step"])
    del op, held
    torch.cuda.empty_cache()
    return out


def mp_workload(args):
    """The 27-sphere workload: 3x3x3 packing (lognormal radii, mean 10,
    sigma 0.15), Q1 mp_cells^3, 8 order parameters (10 components).  FP64
    and FP32 tangent vmult (per-field relative l2 error of FP32 against
    FP64), FP64 residual; the same with the grain cut-off; Newton steps at
    rtol 1e-8: plain FP64 against FP32 inner Krylov, with the cut-off off and
    on (the component table from each step's initial guess)."""
    import torch
    from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,
                            SinteringOperatorGeneric, packing)
    from hpsint_amd.solvers import NewtonSolver, TimeStepper
    n = args.mp_cells
    pk = packing.packing_cfg2(seed=0)
    G = pk.n_order_params
    mesh = CartesianMesh((n, n, n), pk.lower, pk.upper)
    op = SinteringOperatorGeneric(mesh, G, FreeEnergyParams(), MobilityParams())
    dt = 1e-3
    op.set_time_weights(1.0 / dt, (-1.0 / dt,))
    u = torch.from_numpy(packing.initial_fields(pk, (n, n, n))).cuda()
    op.fill_quadrature_point_values(u)
    src = torch.from_numpy(np.random.default_rng(1234).uniform(-1, 1, op.n_dofs)).cuda()
    dst = op.initialize_dof_vector()
    k = max(5, args.steps // 5)
    settle(lambda: op.vmult(dst, src), 0.1)
    t64 = timed(lambda: op.vmult(dst, src), k)
    src32, dst32 = src.float(), op.initialize_dof_vector(torch.float32)
    for _ in range(3):
        op.vmult(dst32, src32)
    t32 = timed(lambda: op.vmult(dst32, src32), k)
    nc = G + 2
    d64 = dst.view(-1, nc)
    d32 = dst32.double().view(-1, nc)
    err = ((d32 - d64).norm(dim=0) / d64.norm(dim=0)).tolist()
    old = [u.clone()]
    r = op.initialize_dof_vector()
    op.evaluate_nonlinear_residual(r, u, 2, old=old)
    t_res = timed(lambda: op.evaluate_nonlinear_residual(r, u, 2, old=old), k)
    cells = n ** 3
    bytes64 = 8 * (2 * op.n_dofs + cells * 8 * (1 + G + 3))
    out = {"workload": "Q1 %d^3, 27 spheres, %d order parameters (%d components, %d DoF)"
                       % (n, G, nc, op.n_dofs),
           "vmult_f64_gdofs": op.n_dofs / t64 / 1e9, "vmult_f64_ms": t64 * 1e3,
           "vmult_f64_hbm_frac": bytes64 / t64 / 1e9 / HBM_PEAK_GBS,
           "vmult_f32_gdofs": op.n_dofs / t32 / 1e9,
           "vmult_f32_rel_l2_error_per_field": err,
           "residual_f64_gdofs": op.n_dofs / t_res / 1e9}
    progress("27 spheres: operators timed")
    if args.cut_tol > 0.0:
        # the reference's cut-off: per cell, the order parameters below the
        # tolerance at all the cell's nodes are dropped
        full = dst.clone()
        op.set_grain_cut_off(args.cut_tol)
        op.set_component_mask(u)
        op.fill_quadrature_point_values(u)
        masks = op.component_mask()
        active = torch.bitwise_and(masks[:, None] >> torch.arange(G, device=masks.device),
                                   1).sum(dim=1).float()
        for _ in range(3):
            op.vmult(dst, src)
        tc = timed(lambda: op.vmult(dst, src), k)
        tc32 = timed(lambda: op.vmult(dst32, src32), k)
        tcr = timed(lambda: op.evaluate_nonlinear_residual(r, u, 2, old=old), k)
        op.vmult(dst, src)
        out["cut_off"] = {
            "tol": args.cut_tol, "mean_active_order_parameters_per_cell": float(active.mean()),
            "vmult_f64_gdofs": op.n_dofs / tc / 1e9, "vmult_f64_ms": tc * 1e3,
            "vmult_f32_gdofs": op.n_dofs / tc32 / 1e9,
            "residual_f64_gdofs": op.n_dofs / tcr / 1e9,
            "speedup_vs_full_f64": t64 / tc,
            "rel_l2_change_vs_full_tangent": float((dst - full).norm() / full.norm())}
        op.set_grain_cut_off(0.0)
        del full, masks, active
        progress("27 spheres: cut-off operators timed")
    del src, dst, src32, dst32, r
    torch.cuda.empty_cache()
    if args.mp_timesteps > 0:
        first = TimeStepper(op, NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2), order=1,
                            dt=dt)
        first.initialize(u)
        first.make_step()
        u1 = first.u.clone()
        del first
        res = {"workload": "BDF1 dt=%g, Newton rtol 1e-8, GMRES rtol 1e-2, node-block Jacobi "
                           "(*_bp2: BlockPreconditioner2); one untimed step, then %d timed "
                           "steps per method"
                           % (dt, args.mp_timesteps)}
        nb = "InverseComponentBlockDiagonalMatrix"
        for name, fp32, tol, prec in (("fp64_plain", 0, 0.0, nb), ("mp_only", 1, 0.0, nb),
                                      ("mp64_only", 2, 0.0, nb),
                                      ("fp64_cut", 0, args.cut_tol, nb),
                                      ("mp_cut", 1, args.cut_tol, nb),
                                      ("fp64_bp2", 0, 0.0, BP2), ("mp_bp2", 1, 0.0, BP2)):
            if tol <= 0.0 and name.endswith("cut"):
                continue
            op.set_grain_cut_off(tol)
            s = TimeStepper(op, NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2,
                                             inner_fp32=fp32, preconditioner=prec),
                            order=1, dt=dt)
            s.initialize(u1)
            s.history = [u1.clone(), u.clone()]
            s.t = dt
            res[name] = summary([s.make_step() for _ in range(args.mp_timesteps)])
            del s
            torch.cuda.empty_cache()
            progress("27 spheres: %s steps done" % name)
        op.set_grain_cut_off(0.0)
        base = res["fp64_plain"]["s_per_step"]
        for name in ("mp_only", "mp64_only", "fp64_cut", "mp_cut", "fp64_bp2", "mp_bp2"):
            if name in res:
                res["speedup_%s_vs_fp64" % name] = base / res[name]["s_per_step"]
        res["speedup_mp_bp2_vs_fp64_bp2"] = (res["fp64_bp2"]["s_per_step"]
                                             / res["mp_bp2"]["s_per_step"])
        out["timesteps"] = res
    del op, u
    torch.cuda.empty_cache()
    return out


def q2_vmult(args):
    """FP64 and FP32 tangent vmult of the Q2 workload: random sequential
    packing of 64 spheres (lognormal radii, mean 8, sigma 0.2) in a 160^3
    box, Q2 on a uniform q2_cells^3 mesh (3^3 Gauss points), tensorial
    mobility, 10 order parameters (12 components)"""
    import torch
    from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,
                            SinteringOperatorGeneric, packing)
    n = args.q2_cells
    pk = packing.packing_cfg3(seed=0)
    G = 10
    if pk.n_order_params > G:
        raise RuntimeError("packing needs %d order parameters" % pk.n_order_params)
    mesh = CartesianMesh((n, n, n), pk.lower, pk.upper)
    op = SinteringOperatorGeneric(mesh, G, FreeEnergyParams(),
                                  MobilityParams(tensorial=True), fe_degree=2)
    op.set_time_weights(1e3, (-1e3,))
    # Q2 nodes of an n^3 mesh = Q1 nodes of a (2n)^3 mesh
    u = torch.from_numpy(packing.initial_fields(pk, (2 * n,) * 3, n_order_params=G)).cuda()
    op.fill_quadrature_point_values(u)
    src = torch.from_numpy(np.random.default_rng(1234).uniform(-1, 1, op.n_dofs)).cuda()
    dst = op.initialize_dof_vector()
    for _ in range(3):
        op.vmult(dst, src)
    k = max(5, args.steps // 5)
    t64 = timed(lambda: op.vmult(dst, src), k)
    src32, dst32 = src.float(), op.initialize_dof_vector(torch.float32)
    for _ in range(3):
        op.vmult(dst32, src32)
    t32 = timed(lambda: op.vmult(dst32, src32), k)
    nf = 1 + G + 2 * 3 + 1 + 6 + G * 3   # c, eta_i, grad c, grad mu, S, P, Q_i
    cells = n ** 3
    bytes64 = 8 * (2 * op.n_dofs + cells * 27 * nf)
    out = {"workload": "Q2 %d^3, 64 spheres, tensorial mobility, %d components, %d DoF"
                       % (n, G + 2, op.n_dofs),
           "n_dofs": op.n_dofs, "vmult_f64_gdofs": op.n_dofs / t64 / 1e9,
           "vmult_f64_ms": t64 * 1e3, "vmult_f32_gdofs": op.n_dofs / t32 / 1e9,
           "vmult_f64_hbm_frac": bytes64 / t64 / 1e9 / HBM_PEAK_GBS}
    del src, dst, src32, dst32
    progress("Q2: operators timed")
    if args.q2_timesteps > 0:
        out["timesteps"] = q2_timesteps(op, u, args)
    del op, u
    torch.cuda.empty_cache()
    return out


def q2_timesteps(op, u0, args):
    """s/timestep on the Q2 workload (BDF1, dt = 1e-3, Newton rtol 1e-8, GMRES
    rtol 1e-2, node-block Jacobi): plain FP64 and FP32 inner Krylov (no
    predictor), after one untimed step from the initial fields"""
    import torch
    from hpsint_amd.solvers import NewtonSolver, TimeStepper
    dt = 1e-3

    def stepper(fp32):
        solver = NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2,
                              preconditioner="InverseComponentBlockDiagonalMatrix",
                              inner_fp32=fp32)
        return TimeStepper(op, solver, order=1, dt=dt)

    first = stepper(0)
    first.initialize(u0)
    first.make_step()
    u1 = first.u.clone()
    del first
    res = {}
    for name, fp32 in (("fp64_plain", 0), ("mp_only", 1)):
        s = stepper(fp32)
        s.initialize(u1)
        s.history = [u1.clone(), u0.clone()]
        s.t = dt
        res[name] = summary([s.make_step() for _ in range(args.q2_timesteps)])
        del s
        torch.cuda.empty_cache()
    res["speedup_mp_vs_fp64"] = res["fp64_plain"]["s_per_step"] / res["mp_only"]["s_per_step"]
    progress("Q2: steps done")
    return res


def cfg1_timesteps(op, u0, dt, args):
    """The headline workload's timesteps (BDF1, Newton rtol 1e-8, GMRES rtol
    1e-2, node-block Jacobi), after 3 untimed plain steps: plain FP64, FP32
    inner Krylov, and FP32 inner Krylov with the reference's Euler
    predictor; then plain FP64 and FP32 inner Krylov with the reference's
    default block preconditioner (BlockPreconditioner2)"""
    import torch
    from hpsint_amd.solvers import ExtrapolationPredictor, NewtonSolver, TimeStepper
    s = TimeStepper(op, NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2), order=1, dt=dt)
    s.initialize(u0)
    traj = [u0.clone()]
    for _ in range(3):
        s.make_step()
        traj.append(s.u.clone())
    del s
    res = {"workload": "3 untimed plain steps, then %d timed steps per method" % args.timesteps}
    nb = "InverseComponentBlockDiagonalMatrix"
    for name, fp32, pred, prec in (("fp64_plain", 0, None, nb), ("mp_only", 1, None, nb),
                                   ("mp_euler", 1, ExtrapolationPredictor(op, "Euler"), nb),
                                   ("fp64_bp2", 0, None, BP2), ("mp_bp2", 1, None, BP2)):
        st = TimeStepper(op, NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2,
                                          inner_fp32=fp32, preconditioner=prec),
                         order=1, dt=dt, predictor=pred)
        st.initialize(traj[3])
        st.history = [traj[3].clone(), traj[2].clone()]
        st.t = 3 * dt
        res[name] = summary([st.make_step() for _ in range(args.timesteps)])
    base = res["fp64_plain"]["s_per_step"]
    res["speedup_mp_only_vs_fp64"] = base / res["mp_only"]["s_per_step"]
    res["speedup_mp_euler_vs_fp64"] = base / res["mp_euler"]["s_per_step"]
    res["speedup_fp64_bp2_vs_fp64"] = base / res["fp64_bp2"]["s_per_step"]
    res["speedup_mp_bp2_vs_fp64"] = base / res["mp_bp2"]["s_per_step"]
    res["speedup_mp_bp2_vs_fp64_bp2"] = (res["fp64_bp2"]["s_per_step"]
                                         / res["mp_bp2"]["s_per_step"])
    torch.cuda.empty_cache()
    progress("128^3: steps done")
    return res


def cpu_baseline(args, mesh, u0, src_h, nc, dt):
    """The CPU restatement of the same FP64 tangent vmult (QP cache,
    sum-factorised, cell batches on SIMD lanes, OpenMP over the host cores;
    oracle/cpu_fast.c, checked against the oracle), timed on the full
    128^3 workload for about --cpu-seconds"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pycpufast as cfb
    import pyoracle as po
    n = mesh.n_cells[0]
    p = po.make_params(3, 1, mesh.n_cells, mesh.cell_size, nc, A=16.0, B=1.0, kappa_c=1.0,
                       kappa_p=0.5, Mvol=1e-2, Mvap=1e-10, Msurf=4.0, Mgb=0.4, L=1.0,
                       w0=1.0 / dt, w_hist=(-1.0 / dt,))
    op = cfb.TangentVmult(p)
    op.fill(u0)
    dst = np.empty_like(src_h)
    op.vmult(src_h, dst)
    t0 = time.perf_counter()
    reps = 0
    while True:
        op.vmult(src_h, dst)
        reps += 1
        if time.perf_counter() - t0 > args.cpu_seconds or reps >= 1000:
            break
    t_cpu = (time.perf_counter() - t0) / reps
    return {"value": src_h.size / t_cpu / 1e9, "unit": "GDoF/s", "cores": cfb.threads(),
            "kind": "port",
            "sample": "%d FP64 tangent vmults of the full %d^3 workload by the CPU restatement "
                      "oracle/cpu_fast.c (QP cache, sum factorisation, %d cells per SIMD "
                      "batch, OpenMP over %d threads; not deal.II, which is not buildable "
                      "here) on %s" % (reps, n, cfb.simd_width(), cfb.threads(), _cpu_model())}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.predictor_study:
        print(json.dumps({"predictor_study": predictor_study(args)}))
        return

    import torch
    from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,
                            SinteringOperatorGeneric, packing)

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n = args.cells
    cells = (n, n, n)
    pk = packing.packing_cfg1(seed=0)
    G = pk.n_order_params
    mesh = CartesianMesh(cells, pk.lower, pk.upper)
    fe = FreeEnergyParams(A=16.0, B=1.0, kappa_c=1.0, kappa_p=0.5)
    mob = MobilityParams(Mvol=1e-2, Mvap=1e-10, Msurf=4.0, Mgb=0.4, L=1.0)
    op = SinteringOperatorGeneric(mesh, G, fe, mob)
    dt = 1e-3
    op.set_time_weights(1.0 / dt, (-1.0 / dt,))

    u0 = packing.initial_fields(pk, cells)
    nc = G + 2
    n_dofs = op.n_dofs
    n_nodes = op.n_nodes
    n_cells = n ** 3
    u = torch.from_numpy(u0).to(dev)
    src = torch.from_numpy(np.random.default_rng(1234).uniform(-1, 1, n_dofs)).to(dev)
    dst = op.initialize_dof_vector()
    op.fill_quadrature_point_values(u)

    # ---- headline: FP64 tangent vmult, steady state
    vm = lambda: op.vmult(dst, src)  # noqa: E731
    n_settle = settle(vm, args.settle)
    for _ in range(args.warmup):
        vm()
    t_vmult = timed(vm, args.steps)
    # dominant kernel: HIP events around each kernel of K back-to-back vmults
    t_main, t_fix, _ = kernel_times(op, vm, args.steps)
    progress("128^3: FP64 vmult timed")

    # ---- residual (BDF1 history) and FP32 vmult
    old = [u.clone()]
    r = op.initialize_dof_vector()
    rs = lambda: op.evaluate_nonlinear_residual(r, u, 2, old=old)  # noqa: E731
    settle(rs, 0.4 * args.settle)
    for _ in range(args.warmup):
        rs()
    t_res = timed(rs, args.steps)
    t_res_main, _, _ = kernel_times(op, rs, args.steps)
    src32 = src.float()
    dst32 = op.initialize_dof_vector(torch.float32)
    vm32 = lambda: op.vmult(dst32, src32)  # noqa: E731
    settle(vm32, 0.4 * args.settle)
    for _ in range(args.warmup):
        vm32()
    t_vm32 = timed(vm32, args.steps)
    t_main32, t_fix32, _ = kernel_times(op, vm32, args.steps)
    progress("128^3: residual and FP32 vmult timed")

    # ---- roofline of the vmult main kernel: algorithmic bytes per launch
    nq = 8
    nf = 1 + G + 3          # cached fields per q-point (c, eta_i, grad mu)
    bytes_vmult = 8 * (2 * nc * n_nodes + n_cells * nq * nf)
    bytes_vmult32 = 4 * (2 * nc * n_nodes + n_cells * nq * nf)
    bytes_vmult_ref = 8 * (2 * nc * n_nodes + n_cells * nq * (nc + 2 * 3))
    achieved = bytes_vmult / t_main / 1e9
    bytes_res = 8 * 3 * nc * n_nodes

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(args, mesh, u0, src.cpu().numpy(), nc, dt)
        progress("CPU baseline timed")

    steps = None
    if args.timesteps > 0:
        steps = {"headline_128cubed": cfg1_timesteps(op, u, dt, args)}
    del op, u, src, dst, src32, dst32, r, old
    torch.cuda.empty_cache()
    q2 = q2_vmult(args) if args.q2_cells > 0 else None
    mp = mp_workload(args) if args.mp_cells > 0 else None
    if args.timesteps > 0 and args.study_cells > 0:
        steps["held_out_100_spheres"] = predictor_study(args)

    out = {
        "metric": "tangent vmult GDoF/s (% HBM roofline); s/timestep FP64 vs MP+learned",
        "value": n_dofs / t_vmult / 1e9,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_vmult * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "3D 2x2x2 packing of 8 spheres, Q1 %d^3 hex mesh, %d order "
                        "parameters (%d components, %d DoF); FP64 tangent vmult, "
                        "QP-cache variant (cache: c, eta_i, grad mu per q-point)"
                        % (n, G, nc, n_dofs),
            "n_cells": n_cells, "n_dofs": n_dofs, "n_components": nc,
            "settle_vmults": n_settle,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": VMULT_TRAFFIC_128 if (n == 128 and G == 4) else None,
            "kernel": "%s FP64 tangent (main kernel; its mean duration from HIP events "
                      "over %d back-to-back vmults)"
                      % ("q1_apply" if os.environ.get("HPSINT_AMD_ZCOL") == "0" else "q1_vmult_zcol",
                         args.steps),
            "whole_vmult_frac": bytes_vmult / t_vmult / 1e9 / HBM_PEAK_GBS,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel on the "
                              "128^3, 4-order-parameter workload, scaled by a copy of known "
                              "size (profiles/vmult_q1_128cubed_hbm_traffic.md); a constant "
                              "for that workload",
        },
        "cpu_baseline": cpu,
        "kernel_ms": {"vmult_main": t_main * 1e3, "vmult_fixup": t_fix * 1e3,
                      "residual_main": t_res_main * 1e3, "vmult_f32_main": t_main32 * 1e3,
                      "vmult_f32_fixup": t_fix32 * 1e3},
        "bytes_per_vmult": bytes_vmult,
        "bytes_per_vmult_reference_cache": bytes_vmult_ref,
        "residual_gdofs": n_dofs / t_res / 1e9,
        "residual_hbm_frac": bytes_res / t_res_main / 1e9 / HBM_PEAK_GBS,
        "residual_fp64_frac": (RESIDUAL_FLOP_128 / t_res_main / 1e12 / FP64_PEAK_TFLOPS)
        if (n == 128 and G == 4) else None,
        "vmult_f32_gdofs": n_dofs / t_vm32 / 1e9,
        "vmult_f32_main_hbm_frac": bytes_vmult32 / t_main32 / 1e9 / HBM_PEAK_GBS,
        "timesteps": steps,
        "q2_tensorial": q2,
        "mp_27_spheres": mp,
    }
    print(json.dumps(out))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


if __name__ == "__main__":
    main()
