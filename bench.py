"""Benchmark of the sintering-operator hot path on one MI355X.

Workload (default): 3D 2x2x2 packing of 8 spheres, R = 10 (1 + U(-0.1,0.1)),
contact overlap 0.5, Q1 on a uniform 128^3 hex mesh (2.1M cells), 4 order
parameters -> 6 components (12.9M DoF), A=16, B=1, kappa_c=1, kappa_p=0.5,
Mvol=1e-2, Mvap=1e-10, Msurf=4, Mgb=0.4, L=1, BDF1 dt=1e-3.

A "step" is one FP64 tangent vmult (hps_vmult_f64) with the linearisation
cache filled at the initial fields and src = U(-1,1) (seed 1234), inputs
resident in HBM.  value = DoF * steps / time in GDoF/s.  Also reported: the
FP64 residual (BDF1 history) and the FP32 vmult, the dominant kernel's
roofline from HIP events on the operator's stream, and the CPU oracle timed
on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-seconds S]
"""
import argparse
import ctypes
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
# factors in profiles/vmult_q1_128cubed_hbm_traffic.md
VMULT_TRAFFIC_128 = 1.3839e9
# FP64 flop per launch of the residual main kernel on the same workload
# (SQ_INSTS_VALU_{ADD,MUL,FMA}_F64 x 64 lanes, FMA = 2;
# profiles/q1_kernels_128cubed_fp64_counters.md); FP64 vector peak
RESIDUAL_FLOP_128 = 8.25e9
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="budget of the CPU-oracle sample (0 = skip)")
    ap.add_argument("--cells", type=int, default=128)
    ap.add_argument("--timesteps", type=int, default=5,
                    help="held-out timesteps timed per method (0 = skip)")
    ap.add_argument("--train-steps", type=int, default=30,
                    help="plain-solver steps that train the predictor")
    ap.add_argument("--train-iterations", type=int, default=6000)
    ap.add_argument("--q2-cells", type=int, default=96,
                    help="cells per direction of the Q2 tensorial workload (0 = skip)")
    ap.add_argument("--cut-tol", type=float, default=1e-5,
                    help="grain cut-off tolerance timed on the 27-sphere workload (<0: not)")
    ap.add_argument("--q2-timesteps", type=int, default=2,
                    help="timed Newton steps per method on the Q2 workload (0: none)")
    ap.add_argument("--mp-cells", type=int, default=192,
                    help="cells per direction of the 27-sphere mixed-precision workload "
                         "(0 = skip)")
    ap.add_argument("--predictor-study", action="store_true",
                    help="only run the held-out predictor study on the 100-sphere "
                         "packings (12 order parameters) and exit")
    ap.add_argument("--study-cells", type=int, default=160)
    ap.add_argument("--study-seeds", type=int, default=4,
                    help="packings (seeds 0..k-1) whose plain trajectories train")
    ap.add_argument("--study-train-steps", type=int, default=12)
    ap.add_argument("--study-test-steps", type=int, default=5)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run hps_copy of this many doubles 5 times (calibrates "
                         "the HBM byte counters against a copy of known size) and exit")
    return ap.parse_args()


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


def predictor_study(args):
    """Learned predictor on held-out packings: plain FP64 trajectories of
    packings with seeds 0..k-1 (random sequential packings of 100 spheres,
    lognormal radii mean 8, sigma 0.2, 12 order parameters, Q1 on a
    study_cells^3 mesh, BDF1 dt = 1e-3) train one network; on the packing
    with seed k, after 3 plain steps, study_test_steps steps are timed per
    method at Newton rtol 1e-8."""
    import torch
    from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,
                            SinteringOperatorGeneric, packing)
    from hpsint_amd.predictor_training import collect_trajectory, fit_predictor
    from hpsint_amd.solvers import NewtonSolver, TimeStepper
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

    def progress(msg):
        print("predictor study: " + msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    trajs = []
    for s in range(args.study_seeds):
        trajs.append(collect_trajectory(op, fields(s), n_steps=args.study_train_steps, dt=dt,
                                        nl_rel_tol=1e-8))
        progress("training packing %d: %d plain steps done" % (s, args.study_train_steps))
    t_traj = time.perf_counter() - t0
    t0 = time.perf_counter()
    pred = fit_predictor(op, trajs, dt, iterations=args.train_iterations,
                         rows_per_step=min(131072, op.n_nodes), seed=0,
                         column_weights=(3.0, 1.0, 1.0, 1.0))
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    progress("network trained in %.1f s" % t_fit)
    del trajs
    torch.cuda.empty_cache()
    held = collect_trajectory(op, fields(args.study_seeds), n_steps=3, dt=dt, nl_rel_tol=1e-8)

    def run(fp32, predictor):
        solver = NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2,
                              preconditioner="InverseComponentBlockDiagonalMatrix",
                              inner_fp32=fp32)
        s = TimeStepper(op, solver, order=1, dt=dt, predictor=predictor)
        s.initialize(held[3])
        s.history = [held[3].clone(), held[2].clone()]
        s.t = 3 * dt
        rs = [s.make_step() for _ in range(args.study_test_steps)]
        progress("held-out steps done (fp32 inner %s, predictor %s)"
                 % (fp32, predictor is not None))
        return {"s_per_step": sum(r.seconds for r in rs) / len(rs),
                "newton": [r.newton_iterations for r in rs],
                "linear": [r.linear_iterations for r in rs]}

    out = {"workload": "100 spheres, Q1 %d^3, %d order parameters (%d DoF); trained on %d "
                       "packings x %d plain steps, tested on a held-out packing"
                       % (n, G, op.n_dofs, args.study_seeds, args.study_train_steps),
           "fp64_plain": run(False, None), "mp_learned": run(True, pred),
           "mp_only": run(True, None), "fp64_learned": run(False, pred),
           "trajectory_s": t_traj, "predictor_training_s": t_fit,
           "predictor_final_loss": pred.losses[-1]}
    out["speedup_mp_learned_vs_fp64"] = (out["fp64_plain"]["s_per_step"] /
                                         out["mp_learned"]["s_per_step"])
    return out


def mp_vmult(args, timed):
    """FP64 and FP32 tangent vmult of the mixed-precision workload: 3x3x3
    packing of 27 spheres (lognormal radii, mean 10, sigma 0.15), Q1 on a
    uniform mp_cells^3 mesh, 8 order parameters (10 components); per-field
    relative l2 error of the FP32 vmult (FP32 cache and arithmetic) against
    the FP64 vmult of the same source"""
    import torch
    from hpsint_amd import (CartesianMesh, FreeEnergyParams, MobilityParams,
                            SinteringOperatorGeneric, packing)
    n = args.mp_cells
    pk = packing.packing_cfg2(seed=0)
    G = pk.n_order_params
    mesh = CartesianMesh((n, n, n), pk.lower, pk.upper)
    op = SinteringOperatorGeneric(mesh, G, FreeEnergyParams(), MobilityParams())
    op.set_time_weights(1e3, (-1e3,))
    u = torch.from_numpy(packing.initial_fields(pk, (n, n, n))).cuda()
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
    if args.cut_tol >= 0.0:
        # the grain cut-off of the tangent: order parameters below the
        # tolerance on a whole tile are dropped from that tile's tangent
        full = dst.clone()
        op.set_grain_cut_off(args.cut_tol)
        op.fill_quadrature_point_values(u)
        for _ in range(3):
            op.vmult(dst, src)
        tc = timed(lambda: op.vmult(dst, src), k)
        tc32 = timed(lambda: op.vmult(dst32, src32), k)
        op.vmult(dst, src)
        out["cut_off"] = {
            "tol": args.cut_tol, "vmult_f64_gdofs": op.n_dofs / tc / 1e9,
            "vmult_f64_ms": tc * 1e3, "vmult_f32_gdofs": op.n_dofs / tc32 / 1e9,
            "speedup_vs_full_f64": t64 / tc,
            "rel_l2_change_vs_full_tangent": float((dst - full).norm() / full.norm())}
        op.set_grain_cut_off(-1.0)
        del full
    del op, u, src, dst, src32, dst32, r, old
    torch.cuda.empty_cache()
    return out


def q2_vmult(args, timed):
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

    first = stepper(False)
    first.initialize(u0)
    first.make_step()
    u1 = first.u.clone()
    del first
    res = {}
    for name, fp32 in (("fp64_plain", False), ("mp_only", True)):
        s = stepper(fp32)
        s.initialize(u1)
        s.history = [u1.clone(), u0.clone()]
        s.t = dt
        rs = [s.make_step() for _ in range(args.q2_timesteps)]
        res[name] = {"s_per_step": sum(r.seconds for r in rs) / len(rs),
                     "newton": [r.newton_iterations for r in rs],
                     "linear": [r.linear_iterations for r in rs]}
        del s
        torch.cuda.empty_cache()
    res["speedup_mp_vs_fp64"] = res["fp64_plain"]["s_per_step"] / res["mp_only"]["s_per_step"]
    return res


def timestep_comparison(op, u0, dt, args):
    """s/timestep of plain FP64 Newton-GMRES vs FP32 inner Krylov + learned
    predictor, Newton rtol 1e-8 (relative to |R(u_n)|, the plain method's
    initial residual), GMRES rtol 1e-2, node-block Jacobi"""
    import time
    import torch
    from hpsint_amd.predictor_training import fit_predictor
    from hpsint_amd.solvers import NewtonSolver, TimeStepper

    def stepper(fp32, predictor=None):
        solver = NewtonSolver(op, nl_rel_tol=1e-8, l_rel_tol=1e-2,
                              preconditioner="InverseComponentBlockDiagonalMatrix",
                              inner_fp32=fp32)
        return TimeStepper(op, solver, order=1, dt=dt, predictor=predictor)

    # plain FP64 trajectory: the training data, and the plain timings of the
    # steps that follow it
    ts = stepper(False)
    ts.initialize(u0)
    traj = [u0.clone()]
    recs = []
    for _ in range(args.train_steps + args.timesteps):
        recs.append(ts.make_step())
        traj.append(ts.u.clone())
    plain = recs[args.train_steps:]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pred = fit_predictor(op, traj[:args.train_steps + 1], dt,
                         iterations=args.train_iterations,
                         rows_per_step=min(131072, op.n_nodes), seed=0,
                         column_weights=(3.0, 1.0, 1.0, 1.0))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0

    def run(fp32, predictor):
        s = stepper(fp32, predictor)
        k = args.train_steps
        s.initialize(traj[k])
        s.history = [traj[k].clone(), traj[k - 1].clone()]
        s.t = k * dt
        out = [s.make_step() for _ in range(args.timesteps)]
        err = float(torch.linalg.vector_norm(s.u - traj[-1]) / torch.linalg.vector_norm(traj[-1]))
        return out, err

    mp, err_mp = run(True, pred)
    mp_only, _ = run(True, None)
    learned_only, _ = run(False, pred)

    def summary(rs):
        return {"s_per_step": sum(r.seconds for r in rs) / len(rs),
                "newton": [r.newton_iterations for r in rs],
                "linear": [r.linear_iterations for r in rs]}
    sp, sm = summary(plain), summary(mp)
    return {
        "workload": "%d plain FP64 steps (BDF1, dt=%g) train the predictor; the next %d "
                    "steps are timed per method" % (args.train_steps, dt, args.timesteps),
        "fp64_plain": sp,
        "mp_learned": sm,
        "mp_only": summary(mp_only),
        "fp64_learned": summary(learned_only),
        "speedup_mp_learned_vs_fp64": sp["s_per_step"] / sm["s_per_step"],
        "final_state_rel_diff_mp_learned_vs_fp64": err_mp,
        "predictor_training_s": t_train,
        "predictor_final_loss": pred.losses[-1],
    }


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

    def timed(fn, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k

    # ---- headline: FP64 tangent vmult
    for _ in range(args.warmup):
        op.vmult(dst, src)
    t_vmult = timed(lambda: op.vmult(dst, src), args.steps)

    # ---- dominant kernel duration (HIP events on the operator's stream)
    op.enable_timing(True)
    mains, fixes = [], []
    for _ in range(args.steps):
        op.vmult(dst, src)
        a, b = op.last_kernel_ms()
        mains.append(a)
        fixes.append(b)
    op.enable_timing(False)
    t_main = float(np.mean(mains)) * 1e-3
    t_fix = float(np.mean(fixes)) * 1e-3

    # ---- residual (BDF1 history) and FP32 vmult
    old = [u.clone()]
    r = op.initialize_dof_vector()
    for _ in range(args.warmup):
        op.evaluate_nonlinear_residual(r, u, 2, old=old)
    t_res = timed(lambda: op.evaluate_nonlinear_residual(r, u, 2, old=old), args.steps)
    op.enable_timing(True)
    res_main = []
    for _ in range(max(10, args.steps // 5)):
        op.evaluate_nonlinear_residual(r, u, 2, old=old)
        res_main.append(op.last_kernel_ms()[0])
    op.enable_timing(False)
    src32 = src.float()
    dst32 = op.initialize_dof_vector(torch.float32)
    for _ in range(args.warmup):
        op.vmult(dst32, src32)
    t_vm32 = timed(lambda: op.vmult(dst32, src32), args.steps)

    # ---- roofline of the vmult main kernel: algorithmic bytes per launch
    nq = 8
    nf = 1 + G + 3          # cached fields per q-point (c, eta_i, grad mu)
    s = 8
    bytes_vmult = s * (2 * nc * n_nodes + n_cells * nq * nf)
    bytes_vmult_ref = s * (2 * nc * n_nodes + n_cells * nq * (nc + 2 * 3))
    achieved = bytes_vmult / t_main / 1e9
    bytes_res = 8 * 3 * nc * n_nodes
    res_achieved = bytes_res / (float(np.mean(res_main)) * 1e-3) / 1e9

    # ---- CPU oracle on the host cores (bounded sample of the same workload)
    cpu = None
    if args.cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        p = po.make_params(3, 1, cells, mesh.cell_size, nc, A=16.0, B=1.0, kappa_c=1.0,
                           kappa_p=0.5, Mvol=1e-2, Mvap=1e-10, Msurf=4.0, Mgb=0.4, L=1.0,
                           w0=1.0 / dt, w_hist=(-1.0 / dt,))
        cache = po.fill_qp(p, u0)
        src_h = src.cpu().numpy()
        t0 = time.perf_counter()
        reps = 0
        while True:
            po.vmult(p, cache, src_h)
            reps += 1
            if time.perf_counter() - t0 > args.cpu_seconds or reps >= 100:
                break
        t_cpu = (time.perf_counter() - t0) / reps
        del cache
        cpu = {"value": n_dofs / t_cpu / 1e9, "unit": "GDoF/s",
               "cores": po.max_threads(), "kind": "port",
               "sample": "%d FP64 tangent vmults of the full 128^3 workload by the C "
                         "oracle (direct per-cell quadrature, OpenMP over colour classes; "
                         "CPU restatement, not deal.II) on %s" % (reps, _cpu_model())}

    q2 = q2_vmult(args, timed) if args.q2_cells > 0 else None
    mp = mp_vmult(args, timed) if args.mp_cells > 0 else None

    steps = None
    if args.timesteps > 0:
        steps = timestep_comparison(op, u, dt, args)

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
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": VMULT_TRAFFIC_128 if (n == 128 and G == 4) else None,
        },
        "cpu_baseline": cpu,
        "kernel_ms": {"vmult_main": t_main * 1e3, "vmult_fixup": t_fix * 1e3,
                      "residual_main": float(np.mean(res_main))},
        "bytes_per_vmult": bytes_vmult,
        "bytes_per_vmult_reference_cache": bytes_vmult_ref,
        "residual_gdofs": n_dofs / t_res / 1e9,
        "residual_hbm_frac": res_achieved / HBM_PEAK_GBS,
        "residual_fp64_frac": (RESIDUAL_FLOP_128 / (float(np.mean(res_main)) * 1e-3) / 1e12 /
                               FP64_PEAK_TFLOPS) if (n == 128 and G == 4) else None,
        "vmult_f32_gdofs": n_dofs / t_vm32 / 1e9,
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
