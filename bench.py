"""Benchmark: ms per Poisson solve to 1e-10 relative residual on the 257^3 stretched grid.

Grid of every Poisson workload: 257^3 nodes, periodic x/y, Neumann z, tanh z grid (beta 1),
right-hand side = 8 seeded Fourier modes (projected onto the operator's range by the solver),
V(2,2) red-black Gauss-Seidel cycles with vcDepth 7, stop on ||f - L u||_2 / ||f||_2 < 1e-10.

  rbc257      (default) one step = the pressure solve of one Rayleigh-Benard scalar_d3 time step on
              257^3 nodes (the same grid, tanh z beta 1, PPN): learned guess + mixed precision
              against plain FP64 multigrid warm-started as the reference does (rbc_bench)
  rbc257x129  the same on 257x257x129 nodes (L = 2, 2, 1)
  seq257      one step = one solve of a synthetic sequence whose right-hand side evolves in time
              (the 8 seeded Fourier modes with time-dependent amplitudes) (seq_bench)
  poisson257  one step = one solve of the single right-hand side from a zero guess: plain FP64,
              mixed-precision refinement, GCR-accelerated mixed precision
  tgv129      one step = one Taylor-Green hydro_d3 time step at 129^3
  mixed257x16 the 257^3 case for seeds 0..15: FP64, mixed refinement and GCR mixed solves, the
              mixed solutions' errors against the FP64 solve, seed 0 against the oracle
  unet129     the learned guess's network on 129^3: each layer's MFMA convolution, adjoint and weight
              gradient timed, TFLOP/s against the dense BF16 peak
  fields257   the field operators on caller arrays (include/saras_field.h) on 257^3: GB/s of each, the
              oracle's CPU time for the same call beside it

Inputs are resident in HBM before the timed region.  Prints one JSON line.  ``--gpus`` above 1 is
refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/Poisson solve to 1e-10 rel. residual, 257³ (learned+mixed vs FP64 MG)"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the roofline kernels, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (KiB per dispatch) scaled by what the same counters read for a device copy of known size
# (FETCH_SIZE x 2.005, WRITE_SIZE x 1.0); the summaries and commands are in profiles/
TRAFFIC_BYTES = {"k_rb_sweep<float> 257x257x257": 124.4e6, "k_rb_sweep<double> 257x257x257": 229.0e6}


def host_info():
    """The host the CPU baseline ran on: logical CPUs, the CPUs this process may run on, and the
    CPU model string."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = None
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "cpu_model": model}


def cpu_baseline(p, g, f, max_cycles, guess=None):
    """Oracle restatement of the reference algorithm (lexicographic Gauss-Seidel, one thread,
    -Ofast build like the reference's release flags) on the same workload and stop rule, from a
    zero initial guess or from `guess`."""
    import numpy as np
    from oracle import pyoracle as po
    prm = po.Params()
    prm.xInd, prm.yInd, prm.zInd = p.xInd, p.yInd, p.zInd
    prm.vcDepth, prm.vcCount = p.vcDepth, max_cycles
    prm.preSmooth, prm.postSmooth, prm.resType = p.preSmooth, p.postSmooth, 2
    prm.smoother = po.SMOOTH_LEX
    prm.xPer, prm.yPer, prm.zPer = int(p.xPer), int(p.yPer), int(p.zPer)
    prm.testPoisson, prm.semantics = 0, po.SEM_PROJECTED
    prm.mgTolerance, prm.relTolerance = p.mgTolerance, p.relTolerance
    prm.dXi, prm.dEt, prm.dZt = g.dXi, g.dEt, g.dZt
    met = {"xix2": g.xix2Staggr, "xixx": g.xixxStaggr, "ety2": g.ety2Staggr, "etyy": g.etyyStaggr,
           "ztz2": g.ztz2Staggr, "ztzz": g.ztzzStaggr}
    o = po.OracleMG(prm, met, fast=True)
    t0 = time.perf_counter()
    _, hist = o.solve(np.zeros_like(f) if guess is None else np.array(guess, dtype=np.float64), f)
    dt = time.perf_counter() - t0
    o.close()
    return dt * 1e3, len(hist), float(hist[-1]) if len(hist) else float("nan")


def mixed_seeds_bench(args):
    """The mixed-precision variant of the 257^3 stretched-grid case over 16 seeded right-hand sides:
    each seed solved to 1e-10 by plain FP64 multigrid, by FP32 V(2,2) cycles under FP64 iterative
    refinement and by the GCR-accelerated form, from a zero guess; the solutions' max-abs and L2
    differences against the FP64 solve (centred: the PPN problem fixes the solution up to a
    constant).  Seed 0 is also solved by the oracle's FP64 multigrid (red-black order, the device's
    algorithm) on the CPU, which pins the FP64 reference itself."""
    import numpy as np
    from saras_amd import mg, workloads
    from saras_amd._lib import DeviceArray
    p, g, _ = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)
    shape = g.shape()
    d_rhs, d_p = DeviceArray(shape), DeviceArray(shape)
    solvers = {}
    for name, prec, direct in (("fp64", mg._L.SR_PREC_FP64, 0), ("mixed", mg._L.SR_PREC_MIXED, 0),
                               ("gcr", mg._L.SR_PREC_MIXED_GCR, 1)):
        q = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)[0]
        q.precision, q.coarseDirect = prec, direct
        solvers[name] = mg.Multigrid(g, q)
    centred = lambda a: a[1:-1, 1:-1, 1:-1] - a[1:-1, 1:-1, 1:-1].mean()  # noqa: E731
    rows, sols0, f0 = [], {}, None
    for seed in range(16):
        f = workloads.fourier_rhs(g.xStaggr, g.yStaggr, g.zStaggr, seed)
        d_rhs.upload(f)
        rec, sols = {"seed": seed}, {}
        for name, s in solvers.items():
            d_p.upload(np.zeros(shape))
            t0 = time.perf_counter()
            st = s.mgSolve_device(d_p.ptr, d_rhs.ptr)
            rec[name + "_ms"] = round((time.perf_counter() - t0) * 1e3, 3)
            rec[name + "_cycles"] = int(st.cycles)
            rec[name + "_residual"] = float(st.residual)
            sols[name] = centred(d_p.download())
        b = sols["fp64"]
        for name in ("mixed", "gcr"):
            d = sols[name] - b
            rec[name + "_max_abs_err"] = float(np.max(np.abs(d)))
            rec[name + "_rel_l2_err"] = float(np.linalg.norm(d) / np.linalg.norm(b))
        rows.append(rec)
        if seed == 0:
            sols0, f0 = sols, f
        print(f"[mixed257x16] seed {seed}: cycles {rec['fp64_cycles']}/{rec['mixed_cycles']}/{rec['gcr_cycles']}, "
              f"max |mixed - fp64| {rec['mixed_max_abs_err']:.2e}, max |gcr - fp64| {rec['gcr_max_abs_err']:.2e}",
              file=sys.stderr, flush=True)
    for s in solvers.values():
        s.close()
    d_rhs.free()
    d_p.free()
    oracle = None
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        prm = po.Params()
        prm.xInd, prm.yInd, prm.zInd = p.xInd, p.yInd, p.zInd
        prm.vcDepth, prm.vcCount = p.vcDepth, 60
        prm.preSmooth, prm.postSmooth, prm.resType = p.preSmooth, p.postSmooth, 2
        prm.smoother = po.SMOOTH_RB
        prm.xPer, prm.yPer, prm.zPer = int(p.xPer), int(p.yPer), int(p.zPer)
        prm.testPoisson, prm.semantics = 0, po.SEM_PROJECTED
        prm.mgTolerance, prm.relTolerance = p.mgTolerance, p.relTolerance
        prm.dXi, prm.dEt, prm.dZt = g.dXi, g.dEt, g.dZt
        met = {"xix2": g.xix2Staggr, "xixx": g.xixxStaggr, "ety2": g.ety2Staggr, "etyy": g.etyyStaggr,
               "ztz2": g.ztz2Staggr, "ztzz": g.ztzzStaggr}
        o = po.OracleMG(prm, met, fast=True)
        t0 = time.perf_counter()
        po_sol, hist = o.solve(np.zeros(shape), f0)
        cpu_s = time.perf_counter() - t0
        o.close()
        ref = centred(np.asarray(po_sol))
        scale = float(np.max(np.abs(ref)))
        oracle = {"seed": 0, "cycles": len(hist), "final_rel_residual": float(hist[-1]),
                  **{f"{n}_max_abs_diff_vs_oracle": float(np.max(np.abs(sols0[n] - ref))) for n in sols0},
                  "solution_max_abs": scale}
        cpu = {"value": round(cpu_s * 1e3, 1), "unit": "ms/solve", "cores": 1, "kind": "port", "host": host_info(),
               "sample": "the seed-0 solve by the oracle's FP64 multigrid (red-black V(2,2), -Ofast, 1 thread)"}
    keys = ("fp64_ms", "mixed_ms", "gcr_ms", "fp64_cycles", "mixed_cycles", "gcr_cycles")
    summ = {k + "_mean": float(np.mean([r[k] for r in rows])) for k in keys}
    for n in ("mixed", "gcr"):
        summ[n + "_max_abs_err_max"] = max(r[n + "_max_abs_err"] for r in rows)
        summ[n + "_rel_l2_err_max"] = max(r[n + "_rel_l2_err"] for r in rows)
    out = {
        "metric": "ms/Poisson solve to 1e-10 rel. residual, 257^3, mixed precision over 16 seeded right-hand sides",
        "value": round(summ["mixed_ms_mean"], 4), "unit": "ms/solve", "n_gpus": 1, "steps": 16, "warmup": 0,
        "ms_per_step": round(summ["mixed_ms_mean"], 4), "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "stretched-grid Poisson 257^3 (PPN, tanh z beta=1), seeds 0..15 of the 8-mode "
                               "Fourier right-hand side, each solved from zero to 1e-10 rel. residual by FP64 MG, "
                               "FP32-in-FP64 iterative refinement and GCR-accelerated mixed precision",
                   "summary": summ, "per_seed": rows, "oracle_seed0": oracle},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def tgv_bench(args):
    """Taylor-Green vortex time steps: 129^3 nodes (128^3 periodic cells), L = 2 pi, Re 1600,
    dt 1e-3, Crank-Nicolson Jacobi viscous solve (tolerance 1e-6), multigrid projection with the
    reference's defaults (V(4,4), 5 cycles, stop on |change of max-residual| < 1e-6) at depth 6.
    One step = hydro_d3::timeAdvance; the kinetic energy (tseries definition) is read after each."""
    import numpy as np
    from saras_amd import ts
    desc = ts.ts_desc(ind=(7, 7, 7), depth=6)
    h = ts.Hydro(desc)
    h.initTGV()
    ke = [h.energy()]
    for _ in range(args.warmup):
        h.timeAdvance()
        ke.append(h.energy())
    step_ms, jac, cyc = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = h.timeAdvance()
        step_ms.append(st.step_ms)
        jac.append(list(st.jacobi_iters))
        cyc.append(st.mg_cycles)
        ke.append(h.energy())
    t1 = time.perf_counter()
    maxdiv = h.maxDivergence()
    # the step kernels one by one (HIP events on the handle's stream) against their algorithmic
    # bytes; timed after the stepped state was read, since they overwrite what they write
    kernels = {}
    for kid, name in ((0, "k_build_rhs x3"), (1, "k_jacobi_v (Vx)"), (2, "k_jacobi_res (Vx)"),
                      (6, "k_jacobi_vres (Vx: sweep + residual of its input)"), (3, "k_divergence"),
                      (4, "k_project x3"), (5, "k_energy")):
        kms = h.time_kernel(kid, reps=20)
        kb = h.kernel_bytes(kid)
        gbs = kb / (kms * 1e-3) / 1e9
        kernels[name] = {"ms": round(kms, 5), "bytes": kb, "GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    h.close()
    ms = (t1 - t0) * 1e3 / args.steps
    cpu = None
    ke_check = None
    ke_steps_checked = 0
    if not args.no_cpu:
        from oracle import pyoracle as po
        n_cpu = 2
        # the reference algorithm (lexicographic Gauss-Seidel), -Ofast, one thread
        o = po.OracleTS(po.ts_params(ind=(7, 7, 7), depth=6, smoother=po.SMOOTH_LEX), fast=True)
        o.init_tgv()
        c0 = time.perf_counter()
        for n in range(n_cpu):
            o.time_advance()
            if (n + 1) % 10 == 0:
                print(f"[tgv129] oracle (lexicographic) step {n + 1}", file=sys.stderr, flush=True)
        c1 = time.perf_counter()
        o.close()
        cpu = {"value": round((c1 - c0) * 1e3 / n_cpu, 1), "unit": "ms/step", "cores": 1, "kind": "port", "host": host_info(),
               "sample": f"{n_cpu} time steps of the same 129^3 Taylor-Green case by the oracle's restatement "
                         "of the reference step (lexicographic GS multigrid, -Ofast, 1 thread)"}
        if args.warmup == 0:
            # energy trace against the oracle with the device's red-black smoother
            o = po.OracleTS(po.ts_params(ind=(7, 7, 7), depth=6, smoother=po.SMOOTH_RB), fast=True)
            o.init_tgv()
            ref = [o.energy()]
            for n in range(min(args.cpu_steps, args.steps)):
                o.time_advance()
                ref.append(o.energy())
                if (n + 1) % 10 == 0:
                    print(f"[tgv129] oracle step {n + 1}", file=sys.stderr, flush=True)
            o.close()
            ke_check = max(abs(a - b) / abs(b) for a, b in zip(ke, ref))
            ke_steps_checked = len(ref) - 1
    out = {
        "metric": "ms/time step, Taylor-Green 129^3 (hydro_d3::timeAdvance)", "value": round(ms, 4),
        "unit": "ms/step", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "Taylor-Green vortex 129^3 nodes PPP, L=2pi, Re=1600, dt=1e-3, cnTolerance 1e-6, "
                               "MG V(4,4) x<=5, depth 6, mgTolerance 1e-6",
                   "median_step_ms": float(np.median(step_ms)), "jacobi_sweeps_per_step": jac[-1],
                   "mg_cycles_median": int(np.median(cyc)), "ke_first": ke[0], "ke_last": ke[-1],
                   "ke_trace_every_10": ke[::10], "max_divergence_end": maxdiv,
                   "ke_rel_diff_vs_oracle_first_steps": ke_check, "ke_steps_checked_vs_oracle": ke_steps_checked,
                   "step_kernels": kernels},
        "roofline": {"bound": "hbm", "kernel": "k_build_rhs x3 (the largest non-multigrid kernel of the step)",
                     "achieved": kernels["k_build_rhs x3"]["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kernels["k_build_rhs x3"]["frac"], "traffic": None,
                     "bytes_per_launch": kernels["k_build_rhs x3"]["bytes"],
                     "avg_launch_ms": kernels["k_build_rhs x3"]["ms"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def rbc_bench(args):
    """Rayleigh-Benard convection (scalar_d3, config C5 physics): pressure solves to 1e-10 relative
    residual inside the time step.  The learned guess (online least-squares predictor + network) is
    trained over the warm-up steps; from that state the same number of steps is timed with
    (a) learned guess + GCR-accelerated mixed precision, (b) the reference's warm start in plain FP64
    (the baseline), (c) warm start in mixed refinement, (d) warm start in GCR mixed precision, and
    with --pad-mode 1 (e) plain FP64 with the mirror residual pads that (a) and (d) use.
    A step's solve time covers the guess and the solve; the right-hand side is built on the device."""
    import numpy as np
    from saras_amd import _lib, guess, mg, ts
    nz_ind = 8 if args.workload == "rbc257" else 7
    kw = dict(ind=(8, 8, nz_ind), L=(1.0, 1.0, 1.0) if nz_ind == 8 else (2.0, 2.0, 1.0), domain="PPN",
              mesh="UUD", beta=(1.0, 1.0, 1.0), scalar=True, rbc_type=2, Ra=1e7, Pr=1.0, dt=args.dt, prob_type=5,
              Re=1.0, cn_tol=1e-6, semantics=_lib.SR_SEM_PROJECTED, rel_tol=1e-10, vc_count=100, pre=2, post=2,
              res_type=2, depth=nz_ind - 1)
    n = (257, 257, 257 if nz_ind == 8 else 129)
    log = lambda msg: print(f"[{args.workload}] {msg}", file=sys.stderr, flush=True)  # noqa: E731
    net = guess.GuessNet(tuple((k - 1) // 2 + 1 for k in n), fine=n, seed=0, lr=1e-3, history=args.history)
    # smoothing counts of the GCR preconditioner cycle: the reference's V(2,2) (V(3,3) measured no
    # faster on this workload: 7.86 against 7.82 ms per solve)
    inner = (2, 2)
    # residual pads of the mixed solves' restriction (0: the reference's zero pads, 1: mirror at the walls)
    pads = args.pad_mode
    run = ts.Hydro(ts.ts_desc(precision=_lib.SR_PREC_MIXED_GCR, coarse_direct=1, inner=inner, pad_mode=pads,
                              **kw))
    run.attach_guess(net, train=True)
    run.initRBC(seed=0, amp=1e-3)
    log(f"training over {args.train_steps} steps")
    t0 = time.perf_counter()
    train_loss = []
    for i in range(args.train_steps):
        st = run.timeAdvance()
        if st.train_loss >= 0:
            train_loss.append(st.train_loss)
        if (i + 1) % 100 == 0:
            log(f"trained {i + 1} steps")
    train_s = time.perf_counter() - t0
    snap = {w: run.download(w) for w in (ts.VX, ts.VY, ts.VZ, ts.P, ts.PP, ts.T)}
    # (a) learned + mixed, frozen network (used adaptively), from the trained state: W untimed
    # steps, then K timed ones
    log("learned + mixed steps")
    run.attach_guess(net, train=False)
    net.set_policy(2)
    for _ in range(args.warmup):
        run.timeAdvance()
    rec_l = [run.timeAdvance() for _ in range(args.steps)]
    learned = {"ms_per_solve": float(np.mean([r.solve_ms + r.guess_ms for r in rec_l])),
               "cycles_mean": float(np.mean([r.mg_cycles for r in rec_l])),
               "cycles_per_step": [int(r.mg_cycles) for r in rec_l],
               "guess_ms_mean": float(np.mean([r.guess_ms for r in rec_l])),
               "guess_predict_ms_mean": float(np.mean([r.guess_predict_ms for r in rec_l])),
               "choices": {name: int(sum(1 for r in rec_l if r.guess_choice == c))
                           for c, name in ((0, "warm"), (1, "extrapolation"), (2, "learned"), (3, "linear"))},
               "rel_guess_median": float(np.median([r.rel_guess for r in rec_l])),
               "rel_guess_per_step": [float(f"{r.rel_guess:.3e}") for r in rec_l],
               "step_ms_mean": float(np.mean([r.step_ms for r in rec_l])),
               "sweep0_ms_mean": float(np.mean([r.sweep0_ms for r in rec_l])),
               # Jacobi sweeps of solveVx/Vy/Vz/solveT per step: a change in a count shifts the
               # pressure right-hand side by the Jacobi tolerance, off the history's span
               "jacobi_iters_per_step": [list(r.jacobi_iters) for r in rec_l]}
    nu_end = run.nusselt()
    # one real solve of this workload for the CPU baseline: the last step's right-hand side from
    # the solution before it (the reference's warm start)
    cpu_rhs, cpu_guess = run.download(ts.MGRHS), run.download(ts.PP2)
    sol_l = run.download(ts.P)
    run.close()
    out = {}
    sols = {}
    legs = [("fp64_warm", _lib.SR_PREC_FP64, 0, (0, 0), 0),
            ("mixed_refinement_warm", _lib.SR_PREC_MIXED, 0, (0, 0), 0),
            ("gcr_warm", _lib.SR_PREC_MIXED_GCR, 1, inner, pads)]
    if pads:
        # the baseline stays on the reference's pads; this leg shows what the pad rule alone is worth
        legs.append(("fp64_warm_mirror_pads", _lib.SR_PREC_FP64, 0, (0, 0), pads))
    for name, prec, direct, inn, pm in legs:
        log(f"timed {name} steps")
        h = ts.Hydro(ts.ts_desc(precision=prec, coarse_direct=direct, inner=inn, pad_mode=pm, **kw))
        h.initRBC(seed=0, amp=1e-3)
        for w, a in snap.items():
            h.upload(w, a)
        for _ in range(args.warmup):
            h.timeAdvance()
        rec = [h.timeAdvance() for _ in range(args.steps)]
        out[name] = {"ms_per_solve": float(np.mean([r.solve_ms for r in rec])),
                     "cycles_mean": float(np.mean([r.mg_cycles for r in rec])),
                     "step_ms_mean": float(np.mean([r.step_ms for r in rec]))}
        sols[name] = h.download(ts.P)
        h.close()
    net.close()
    # algorithmic bytes of one FP64 V(2,2) cycle: 114 B per node on every level above the coarsest
    # (four sweeps of two half sweeps, 96 B; residual read on colour 0 with its restriction, 8 B, and
    # the coarse right-hand side written, 1 B; prolongation onto colour 1, 9 B) + 12 B per finest
    # node for the colour-0 norm
    depth = nz_ind - 1
    n_levels = [((n[0] - 1 >> l) + 1) * ((n[1] - 1 >> l) + 1) * ((n[2] - 1 >> l) + 1) for l in range(depth + 1)]
    cyc_bytes = 114.0 * sum(n_levels[:-1]) + 12.0 * n_levels[0]
    a_, b_ = sol_l[1:-1, 1:-1, 1:-1], sols["fp64_warm"][1:-1, 1:-1, 1:-1]
    dP = (a_ - a_.mean()) - (b_ - b_.mean())
    base = out["fp64_warm"]["ms_per_solve"]
    # roofline: the finest FP32 half sweep of the learned+mixed solves (HIP events in the solves)
    p = mg.Parser.from_strings(mesh="UUD", domain="PPN", xInd=8, yInd=8, zInd=nz_ind, vcDepth=nz_ind - 1,
                               semantics=_lib.SR_SEM_PROJECTED, precision=_lib.SR_PREC_MIXED_GCR, coarseDirect=1)
    if nz_ind == 7:
        p.Lx, p.Ly = 2.0, 2.0
    solver = mg.Multigrid(mg.Grid(p), p)
    bytes_sweep = solver.kernel_bytes(6, 0)
    iso_ms = solver.time_kernel(6, 0, 20)
    solver.close()
    sweep_ms = learned["sweep0_ms_mean"]
    achieved = bytes_sweep / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else None
    cpu = None
    if not args.no_cpu:
        log("CPU baseline solve")
        g = mg.Grid(p)
        cpu_ms, cpu_cycles, cpu_res = cpu_baseline(p, g, cpu_rhs, args.cpu_cycles, guess=cpu_guess)
        cpu = {"value": round(cpu_ms, 1), "unit": "ms/solve", "cores": 1, "kind": "port", "host": host_info(),
               "sample": f"one pressure solve of this workload (the last timed step's right-hand side, warm-started "
                         f"from the solution before it, as the reference's mgSolve is) by the oracle's restatement "
                         f"of the reference algorithm (lexicographic GS V(2,2), -Ofast, 1 thread): {cpu_cycles} "
                         f"cycles to rel. residual {cpu_res:.2e}"}
    res = {
        "metric": METRIC,
        "value": round(learned["ms_per_solve"], 4), "unit": "ms/solve", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(learned["ms_per_solve"], 4), "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64+f32+bf16", "data": "synthetic",
        "config": {"workload": f"Rayleigh-Benard {n[0]}x{n[1]}x{n[2]} nodes, PPN, tanh z beta=1, Ra=1e7, Pr=1, "
                               f"rbcType 2, dt={args.dt}, T perturbation 1e-3 (seed 0): the pressure solve of every "
                               f"time step to 1e-10 rel. residual; learned guess trained online over the first "
                               f"{args.train_steps} steps, then warm-up and timed steps from that state",
                   "nodes": list(n), "vcDepth": nz_ind - 1, "guess_history": args.history,
                   "learned_mixed": learned, "fp64_warm": out["fp64_warm"],
                   "mixed_refinement_warm": out["mixed_refinement_warm"], "gcr_warm": out["gcr_warm"],
                   **({"fp64_warm_mirror_pads": out["fp64_warm_mirror_pads"]} if pads else {}),
                   "mixed_solver": f"FP32 V({inner[0]},{inner[1]}) cycle preconditioning flexible GCR(1) under the FP64 residual, "
                                   "exact coarsest level, "
                                   + ("mirror residual pads at the walls" if pads else "the reference's residual pads"),
                   "train_steps": args.train_steps,
                   # the two ratios are co-equal: the first against the reference's algorithm (its zero
                   # residual pads), the second against plain FP64 given the same mirror pads as the mixed
                   # solves, which isolates what the guess and the mixed-precision solver are worth
                   "speedup_vs_fp64_warm": round(base / learned["ms_per_solve"], 3),
                   **({"speedup_vs_fp64_warm_mirror_pads":
                       round(out["fp64_warm_mirror_pads"]["ms_per_solve"] / learned["ms_per_solve"], 3)}
                      if pads else {}),
                   "fp64_cycle_bytes_algorithmic": cyc_bytes,
                   "fp64_warm_solve_GBs_algorithmic": round(cyc_bytes * out["fp64_warm"]["cycles_mean"] /
                                                            (base * 1e-3) / 1e9, 1),
                   "pressure_max_abs_diff_vs_fp64_after_timed_steps": float(np.max(np.abs(dP))),
                   "pressure_rel_l2_diff_vs_fp64_after_timed_steps": float(np.linalg.norm(dP) /
                                                                          np.linalg.norm(b_ - b_.mean())),
                   "train_seconds": round(train_s, 2), "train_loss_first": train_loss[:3],
                   "train_loss_last": train_loss[-3:], "nusselt_end": nu_end},
        "roofline": {"bound": "hbm", "kernel": "k_rb_sweep<float> finest level (FP32 V-cycle)",
                     "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": TRAFFIC_BYTES.get(f"k_rb_sweep<float> {n[0]}x{n[1]}x{n[2]}"),
                     "bytes_per_launch": bytes_sweep, "avg_launch_ms_in_solve": round(sweep_ms, 5),
                     "avg_launch_ms_isolated": round(iso_ms, 5)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


def seq_bench(args):
    """Workload seq257: a sequence of 257^3 stretched-grid Poisson solves whose right-hand side
    evolves smoothly in time (the seeded Fourier modes of the single-solve workload with amplitudes
    a_m cos(w_m t + psi_m), dt 0.05), as a time stepper produces them.  Every solve runs to the same
    stop rule (relative residual 1e-10 on the projected right-hand side).
      learned+mixed: learned initial guess (trained online by its own HIP kernels over the warm-up
                     solves, then frozen) + FP32 V(2,2) cycle preconditioning GCR(1) under the FP64
                     residual, exact coarsest level
      fp64 warm:     plain FP64 V(2,2) multigrid (the reference algorithm, red-black order) from the
                     previous solution, as hydro_d3 warm-starts mgSolve
    A step's time covers the guess and the solve; the right-hand sides are in HBM beforehand."""
    import numpy as np
    from saras_amd import sequence, workloads
    W, K = args.warmup, args.steps
    n = W + K
    log = lambda msg: print(f"[seq257] {msg}", file=sys.stderr, flush=True)  # noqa: E731
    seq = workloads.EvolvingPoisson(ind=8, dt=0.05)
    log(f"modes ready; uploading {n} right-hand sides")
    rhs = sequence.upload_rhs(seq, n)
    res = {}
    sols = {}
    # (leg, sequence variant, exact coarsest level, history of the linear predictor)
    legs = (("learned_mixed", "learned_gcr", 1, 8), ("fp64_warm", "fp64_warm", 0, 0), ("fp64_zero", "fp64_zero", 0, 0),
            ("mixed_refinement_warm", "mixed_warm", 0, 0), ("gcr_warm", "gcr_warm", 1, 0),
            ("gcr_extrap", "gcr_extrap", 1, 0), ("linear_predictor_gcr", "linear_gcr", 1, 8),
            ("learned_no_history_gcr", "learned_gcr", 1, 0))
    for name, var, direct, hist in legs:
        log(f"leg {name}")
        recs, net_v, last = sequence.run(seq, var, n, ntrain=W if var.startswith("learned") else 0, rhs=rhs,
                                         coarse_direct=direct, history=hist)
        s = sequence.summary(recs, W)
        s["cycles_per_step"] = [r["cycles"] for r in recs[W:]]
        if var.startswith("learned"):
            losses = [r["train_loss"] for r in recs if r.get("train_loss") is not None and r["train_loss"] >= 0]
            s["train_loss_first_last"] = [losses[0], losses[-1]] if losses else None
            s["kappa"] = net_v.kappa()[0]
        if net_v is not None:
            net_v.close()
        s["sweep0_ms_mean"] = float(np.mean([r["sweep0_ms"] for r in recs[W:]]))
        res[name] = s
        sols[name] = last
    for d in rhs:
        d.free()
    a, b = sols["learned_mixed"][1:-1, 1:-1, 1:-1], sols["fp64_warm"][1:-1, 1:-1, 1:-1]
    dm = (a - a.mean()) - (b - b.mean())
    lm, base = res["learned_mixed"], res["fp64_warm"]
    # roofline of the dominant kernel of the learned+mixed solve: the finest FP32 half sweep, its
    # average launch time from HIP events on the solver's stream over the timed solves
    from saras_amd import mg
    p = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)[0]
    p.precision, p.coarseDirect = mg._L.SR_PREC_MIXED_GCR, 1
    solver = mg.Multigrid(seq.g, p)
    sweep_ms = lm["sweep0_ms_mean"]
    bytes_sweep = solver.kernel_bytes(6, 0)
    iso_ms = solver.time_kernel(6, 0, 20)
    solver.close()
    achieved = bytes_sweep / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else None
    traffic = TRAFFIC_BYTES.get("k_rb_sweep<float> 257x257x257")
    cpu = None
    if not args.no_cpu:
        pc, gc, fc = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)
        cpu_ms, cpu_cycles, cpu_res = cpu_baseline(pc, gc, seq.rhs(W), args.cpu_cycles)
        cpu = {"value": round(cpu_ms, 1), "unit": "ms/solve", "cores": 1, "kind": "port", "host": host_info(),
               "sample": f"one solve (right-hand side of step {W}, zero guess) of the same 257^3 sequence by the "
                         f"oracle's restatement of the reference algorithm (lexicographic GS V(2,2), -Ofast, "
                         f"1 thread): {cpu_cycles} cycles to rel. residual {cpu_res:.2e}"}
    out = {
        "metric": METRIC, "value": round(lm["ms_per_solve"], 4), "unit": "ms/solve", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": round(lm["ms_per_solve"], 4), "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64+f32+bf16", "data": "synthetic",
        "config": {"workload": "sequence of 257^3 stretched-grid Poisson solves (PPN, tanh z beta=1, 8 seeded "
                               "Fourier modes with amplitudes a_m cos(w_m t + psi_m), dt=0.05), each to 1e-10 rel. "
                               "residual; learned guess trained online over the warm-up solves",
                   "nodes": [257, 257, 257], "vcDepth": seq.p.vcDepth,
                   "speedup_vs_fp64_warm": round(base["ms_per_solve"] / lm["ms_per_solve"], 3),
                   "speedup_vs_fp64_zero": round(res["fp64_zero"]["ms_per_solve"] / lm["ms_per_solve"], 3),
                   "learned_mixed_max_abs_diff_vs_fp64_last_step": float(np.max(np.abs(dm))),
                   "learned_mixed_rel_l2_diff_vs_fp64_last_step": float(np.linalg.norm(dm) /
                                                                       np.linalg.norm(b - b.mean())),
                   "legs": res},
        "roofline": {"bound": "hbm", "kernel": "k_rb_sweep<float> finest level (FP32 V-cycle)",
                     "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None, "traffic": traffic,
                     "bytes_per_launch": bytes_sweep, "avg_launch_ms_in_solve": round(sweep_ms, 5),
                     "avg_launch_ms_isolated": round(iso_ms, 5)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix peak (no sparsity)


def unet_bench(args):
    """The learned guess's network on its 129^3 grid (the network grid of the 257^3 solves): one
    forward and one backward pass through the host entry points to fill every tensor, then each of
    the 12 layers' convolution, adjoint convolution and weight gradient timed over --steps launches
    with HIP events on the network's stream.  FLOPs per launch = 2 * 27 * cin * cout * voxels.  The
    line's value is the time of one forward plus one backward pass of the convolutions; the roofline
    is priced on the dominant kernel against the dense BF16 MFMA peak."""
    import numpy as np
    from saras_amd import guess
    n = (129, 129, 129)
    net = guess.GuessNet(n, fine=(257, 257, 257), seed=0, history=0)
    rng = np.random.default_rng(0)
    net.forward(rng.standard_normal(n + (guess.IN_CH,)).astype(np.float32))
    net.backward(rng.standard_normal(n).astype(np.float32) * 1e-2)
    layers, fwd_ms, bwd_ms, fwd_fl, bwd_fl = [], 0.0, 0.0, 0.0, 0.0
    for l in range(12):
        ci, co, lv = guess.layer_info(l)
        rec = {"layer": guess.LAYER_NAMES[l], "cin": ci, "cout": co, "level": lv}
        for which, key in ((0, "conv"), (1, "adjoint"), (2, "wgrad")):
            if which == 1 and l == 0:
                continue   # the input layer needs no data gradient
            ms, fl = net.time_layer(l, which, reps=max(args.steps, 3))
            rec[key + "_ms"] = round(ms, 5)
            rec[key + "_TFLOPs"] = round(fl / (ms * 1e-3) / 1e12, 2)
            if which == 0:
                fwd_ms += ms
                fwd_fl += fl
            else:
                bwd_ms += ms
                bwd_fl += fl
        rec["GFLOP"] = round(fl / 1e9, 3)
        layers.append(rec)
    net.close()
    # dominant kernel: the largest total time over the three kinds
    dom = max(((r[k + "_ms"], r["layer"], k, r[k + "_TFLOPs"]) for r in layers for k in ("conv", "adjoint", "wgrad")
               if k + "_ms" in r))
    total_ms = fwd_ms + bwd_ms
    res = {
        "metric": "ms per forward + backward pass of the guess network's convolutions, 129^3 grid",
        "value": round(total_ms, 4), "unit": "ms/pass", "n_gpus": 1, "steps": args.steps, "warmup": 2,
        "ms_per_step": round(total_ms, 4), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
        "config": {"workload": "3D U-Net of the learned guess (16/32/64/64 channels, 12 3x3x3 convolutions) on "
                               "129^3 nodes: every layer's convolution, adjoint and weight gradient",
                   "forward_ms": round(fwd_ms, 4), "forward_TFLOPs": round(fwd_fl / (fwd_ms * 1e-3) / 1e12, 2),
                   "backward_ms": round(bwd_ms, 4), "backward_TFLOPs": round(bwd_fl / (bwd_ms * 1e-3) / 1e12, 2),
                   "layers": layers},
        "roofline": {"bound": "mfma", "kernel": f"{dom[2]} of layer {dom[1]}", "achieved": dom[3],
                     "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(dom[3] / MFMA_BF16_PEAK_TFLOPS, 4),
                     "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(res))


class HipTimer:
    """HIP events on a given stream, through the HIP runtime the library itself is linked against
    (libamdhip64.so.7 of the image), so that stream and events belong to one runtime."""

    def __init__(self, stream):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.s = C.c_void_p(stream)
        self.e0, self.e1 = C.c_void_p(), C.c_void_p()
        for e in (self.e0, self.e1):
            self._ok(self.hip.hipEventCreate(C.byref(e)))

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(f"HIP runtime call failed with status {rc}")

    def time(self, fn, reps):
        """Average ms per call of fn (which enqueues on the stream) over reps calls."""
        C = self.C
        self._ok(self.hip.hipEventRecord(self.e0, self.s))
        for _ in range(reps):
            fn()
        self._ok(self.hip.hipEventRecord(self.e1, self.s))
        self._ok(self.hip.hipEventSynchronize(self.e1))
        ms = C.c_float()
        self._ok(self.hip.hipEventElapsedTime(C.byref(ms), self.e0, self.e1))
        return ms.value / reps

    def close(self):
        for e in (self.e0, self.e1):
            if e.value:
                self.hip.hipEventDestroy(e)
                e.value = None


def fields_bench(args):
    """The field operators on caller device arrays (include/saras_field.h) on the 257^3 convection
    grid (PPN, tanh z beta 1): each operator of one time step timed with HIP events on the handle's
    stream over --steps launches after --warmup, against its algorithmic bytes; then the Jacobi
    viscous solve (a synchronous call: wall time and sweeps).  The CPU leg times the oracle's
    restatement of the same operators (lib/field, one thread, -Ofast) on the same inputs."""
    import numpy as np
    from saras_amd import field
    g = dict(ind=(8, 8, 8), L=(1.0, 1.0, 1.0), domain="PPN", mesh="UUD", beta=(1.0, 1.0, 1.0))
    fg = field.FieldGrid(field.fgrid_desc(**g))
    rng = np.random.default_rng(0)
    shp = {loc: fg.shape(loc) for loc in range(4)}
    host = {loc: rng.standard_normal(shp[loc]) for loc in range(4)}
    P = fg.array(field.NODE, host[0])
    V = [fg.array(c + 1, host[c + 1]) for c in range(3)]
    H = [fg.array(c + 1) for c in range(3)]
    G = [fg.array(c + 1) for c in range(3)]
    out = fg.array(field.NODE)
    fsz = lambda loc: float(np.prod(shp[loc]))                      # noqa: E731
    core = lambda loc: float(np.prod([n - 2 for n in shp[loc]]))    # noqa: E731
    w = 8.0
    ops = [
        ("sr_deriv1 (P, z)", lambda: fg.calcDerivative1(field.NODE, 2, P, out), 2 * w * fsz(0),
         "F read, out written over the padded array"),
        ("sr_deriv2 (P, z)", lambda: fg.calcDerivative2(field.NODE, 2, P, out), 2 * w * fsz(0),
         "F read, out written over the padded array"),
        ("sr_vf_diff", lambda: fg.vfComputeDiff(V, H), sum(3 * w * core(c + 1) for c in range(3)),
         "per component: V_c and H_c read, H_c written"),
        ("sr_vf_nlin (V advects V)", lambda: fg.vfComputeNLin(V, V, H), sum(5 * w * core(c + 1) for c in range(3)),
         "per component: Vx, Vy, Vz and H_c read, H_c written"),
        ("sr_divergence", lambda: fg.divergence(V, out), 4 * w * fsz(0),
         "three face arrays read, the node array written"),
        ("sr_gradient", lambda: fg.gradient(P, G), sum(2 * w * core(c + 1) for c in range(3)),
         "per component: P read, G_c written"),
    ]
    timer = HipTimer(fg.stream())
    reps = max(args.steps, 5)
    rows = {}
    for name, fn, nbytes, what in ops:
        for _ in range(max(args.warmup, 1)):
            fn()
        fg.sync()
        ms = timer.time(fn, reps)
        gbs = nbytes / (ms * 1e-3) / 1e9
        rows[name] = {"ms": round(ms, 5), "bytes": nbytes, "bytes_are": what, "GBs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
    timer.close()
    # the viscous solve of Vx with the convection run's coefficients (scalar_d3 form, nu = sqrt(Pr/Ra)
    # for Ra 1e7, Pr 1; dt 0.01): a call returns with the solve finished
    walls = field.bc([field.PERIODIC, field.PERIODIC, field.PERIODIC, field.PERIODIC, field.DIRICHLET,
                      field.DIRICHLET])
    R = fg.array(field.VX, host[1])
    Fj, Gj = fg.array(field.VX, host[1]), fg.array(field.VX)
    t0 = time.perf_counter()
    sweeps = fg.jacobi(field.VX, 1, 0.01, walls, R, Fj, Gj, coef=float(np.sqrt(1.0 / 1e7)), cn_tol=1e-6)
    jac_ms = (time.perf_counter() - t0) * 1e3
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        o = po.OracleTS(po.ts_params(**g), fast=True)
        Vh = [host[c + 1] for c in range(3)]
        cpu_ms = {}
        t0 = time.perf_counter()
        o.fd_deriv(0, 1, 2, host[0])
        cpu_ms["sr_deriv1 (P, z)"] = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        o.fd_deriv(0, 2, 2, host[0])
        cpu_ms["sr_deriv2 (P, z)"] = (time.perf_counter() - t0) * 1e3
        Hh = [np.zeros(shp[c + 1]) for c in range(3)]
        t0 = time.perf_counter()
        for c in range(3):
            o.fd_diff(c + 1, Vh[c], Hh[c])
        cpu_ms["sr_vf_diff"] = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        for c in range(3):
            o.fd_nlin(c + 1, Vh[c], Vh, Hh[c])
        cpu_ms["sr_vf_nlin (V advects V)"] = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        o.fd_divergence(Vh)
        cpu_ms["sr_divergence"] = (time.perf_counter() - t0) * 1e3
        Gh = [np.zeros(shp[c + 1]) for c in range(3)]
        t0 = time.perf_counter()
        o.fd_gradient(host[0], Gh)
        cpu_ms["sr_gradient"] = (time.perf_counter() - t0) * 1e3
        o.close()
        for k, v in cpu_ms.items():
            rows[k]["cpu_ms"] = round(v, 2)
        cpu = {"value": round(sum(cpu_ms.values()), 2), "unit": "ms/pass", "cores": 1, "kind": "port",
               "host": host_info(),
               "sample": "one call of each of the six operators on the same 257^3 inputs by the oracle's "
                         "restatement of lib/field (-Ofast, 1 thread)"}
    P.free(), out.free()
    for a in V + H + G + [R, Fj, Gj]:
        a.free()
    fg.close()
    total = sum(r["ms"] for r in rows.values())
    dom_name, dom = max(rows.items(), key=lambda kv: kv[1]["ms"])
    res = {
        "metric": "ms per pass over the field operators on 257^3 caller arrays (include/saras_field.h)",
        "value": round(total, 4), "unit": "ms/pass", "n_gpus": 1, "steps": reps, "warmup": max(args.warmup, 1),
        "ms_per_step": round(total, 4), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "field operators of one time step on 257^3 nodes (PPN, tanh z beta=1), seeded "
                               "random fields on every location, HIP events on the handle's stream",
                   "operators": rows,
                   "jacobi_visc_Vx": {"call_ms": round(jac_ms, 3), "sweeps": int(sweeps),
                                      "note": "synchronous: the residual is read on the host after every sweep"}},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": dom["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["frac"], "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before the timed ones (default 30 for the sequence workload, where the "
                         "learned guess trains over them; 2 otherwise)")
    ap.add_argument("--train-steps", type=int, default=500,
                    help="convection workloads: time steps over which the learned guess trains online before the "
                         "warm-up and timed steps of every leg (500: the first 500 steps of the convection run)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-cycles", type=int, default=60)
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="tgv129: time steps of the oracle's red-black step compared with the device's energy trace")
    ap.add_argument("--workload", default="rbc257",
                    choices=["rbc257", "rbc257x129", "seq257", "poisson257", "tgv129", "mixed257x16", "unet129",
                             "fields257"])
    ap.add_argument("--dt", type=float, default=0.01, help="time step of the convection workloads")
    ap.add_argument("--pad-mode", type=int, default=1, choices=[0, 1],
                    help="convection workloads: residual pads of the mixed solves' restriction (0 the reference's "
                         "zero pads, 1 mirror at the walls); the plain FP64 baseline keeps the reference's")
    ap.add_argument("--history", type=int, default=8,
                    help="converged pairs kept by the learned guess's least-squares predictor (0..8)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the solver on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.warmup is None:
        args.warmup = 30 if args.workload == "seq257" else 2
    if args.workload == "seq257":
        seq_bench(args)
        return
    if args.workload == "mixed257x16":
        mixed_seeds_bench(args)
        return
    if args.workload == "tgv129":
        tgv_bench(args)
        return
    if args.workload == "unet129":
        unet_bench(args)
        return
    if args.workload == "fields257":
        fields_bench(args)
        return
    if args.workload.startswith("rbc"):
        rbc_bench(args)
        return

    import numpy as np
    from saras_amd import mg, workloads
    from saras_amd._lib import DeviceArray

    p, g, f = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)
    # inputs resident in HBM before the timed region: the right-hand side and a zero initial guess
    d_rhs, d_p0, d_p = DeviceArray(f.shape), DeviceArray(f.shape), DeviceArray(f.shape)
    d_rhs.upload(f)
    d_p0.upload(np.zeros(f.shape))

    def run(solver):
        def step():
            # the solve overwrites its in/out argument: start every step from the zero guess
            mg._L.check(mg._L.load().sr_memcpy(d_p.ptr, d_p0.ptr, d_p.nbytes, 3))
            return solver.mgSolve_device(d_p.ptr, d_rhs.ptr)

        for _ in range(args.warmup):
            step()
        stats = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            stats.append(step())
        t1 = time.perf_counter()
        return (t1 - t0) * 1e3 / args.steps, stats

    # plain FP64 multigrid (the reference algorithm, red-black order)
    solver = mg.Multigrid(g, p)
    ms, stats = run(solver)
    sol64 = d_p.download()
    cycles = [s.cycles for s in stats]
    res = max(s.residual for s in stats)
    conv = all(s.converged for s in stats)
    sweep_ms = float(np.mean([s.sweep0_ms for s in stats]))
    # roofline of the dominant kernel: the finest-level red-black half sweep
    bytes_sweep = solver.kernel_bytes(0, 0)
    achieved = bytes_sweep / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else None
    iso_ms = solver.time_kernel(0, 0, 20)
    plan = {"coarse_cycle_first_level_fp64": solver.info(0), "coarse_cycle_lds_bytes_fp64": solver.info(2),
            "lds_limit_per_workgroup": solver.info(4)}
    solver.close()
    # whole-cycle algorithmic traffic (114 B per node per level: the residual on colour 0 only, 16 B,
    # and the prolongation onto colour 1 only, 9 B; + 12 B per finest node for the colour-0 norm)
    n_levels = [((1 << (8 - l)) + 1) ** 3 for l in range(p.vcDepth + 1)]
    cyc_bytes = 114.0 * sum(n_levels[:-1]) + 12.0 * n_levels[0]
    solve_gbs = cyc_bytes * np.mean(cycles) / (ms * 1e-3) / 1e9

    # mixed precision: FP32 V(2,2) cycles under FP64 iterative refinement, same stop rule
    pm = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)[0]
    pm.precision = mg._L.SR_PREC_MIXED
    solver_m = mg.Multigrid(g, pm)
    ms_m, stats_m = run(solver_m)
    solm = d_p.download()
    solver_m.close()
    a, b = solm[1:-1, 1:-1, 1:-1], sol64[1:-1, 1:-1, 1:-1]
    dm = (a - a.mean()) - (b - b.mean())
    mixed = {"ms_per_solve": round(ms_m, 4), "cycles": int(np.median([s.cycles for s in stats_m])),
             "final_rel_residual": max(s.residual for s in stats_m),
             "converged": all(s.converged for s in stats_m),
             "max_abs_diff_vs_fp64": float(np.max(np.abs(dm))),
             "rel_l2_diff_vs_fp64": float(np.linalg.norm(dm) / np.linalg.norm(b - b.mean())),
             "speedup_vs_fp64": round(ms / ms_m, 3)}
    # GCR-accelerated mixed precision with the exact coarsest level, same stop rule
    pg = workloads.stretched_poisson(ind=8, beta=1.0, seed=0)[0]
    pg.precision, pg.coarseDirect = mg._L.SR_PREC_MIXED_GCR, 1
    solver_g = mg.Multigrid(g, pg)
    ms_g, stats_g = run(solver_g)
    solg = d_p.download()
    sweep32_ms = float(np.mean([s.sweep0_ms for s in stats_g]))
    solver_g.close()
    a = solg[1:-1, 1:-1, 1:-1]
    dg = (a - a.mean()) - (b - b.mean())
    gcr = {"ms_per_solve": round(ms_g, 4), "cycles": int(np.median([s.cycles for s in stats_g])),
           "final_rel_residual": max(s.residual for s in stats_g),
           "converged": all(s.converged for s in stats_g),
           "max_abs_diff_vs_fp64": float(np.max(np.abs(dg))),
           "rel_l2_diff_vs_fp64": float(np.linalg.norm(dg) / np.linalg.norm(b - b.mean())),
           "speedup_vs_fp64": round(ms / ms_g, 3), "fp32_finest_half_sweep_ms_in_solve": round(sweep32_ms, 5)}

    cpu = None
    if not args.no_cpu:
        cpu_ms, cpu_cycles, cpu_res = cpu_baseline(p, g, f, args.cpu_cycles)
        cpu = {"value": round(cpu_ms, 1), "unit": "ms/solve", "cores": 1, "kind": "port", "host": host_info(),
               "sample": f"one full solve of the same 257^3 workload by the oracle's restatement of the "
                         f"reference algorithm (lexicographic GS V(2,2), -Ofast, 1 thread): {cpu_cycles} cycles "
                         f"to rel. residual {cpu_res:.2e}"}

    out = {
        "metric": METRIC, "value": round(ms, 4), "unit": "ms/solve", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "strong",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "stretched-grid Poisson 257^3, PPN, tanh z beta=1, 8 seeded Fourier modes, "
                               "V(2,2) RB-GS to 1e-10 rel. residual (plain FP64 multigrid)",
                   "nodes": [257, 257, 257], "vcDepth": p.vcDepth, "variant": "fp64-mg",
                   "cycles_per_solve": int(np.median(cycles)), "final_rel_residual": res, "converged": conv,
                   "solve_GBs_algorithmic": round(solve_gbs, 1), "mixed_precision": mixed,
                   "gcr_mixed_precision": gcr, "plan": plan},
        "roofline": {"bound": "hbm", "kernel": "k_rb_sweep<double> finest level",
                     "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": TRAFFIC_BYTES.get("k_rb_sweep<double> 257x257x257"),
                     "bytes_per_launch": bytes_sweep, "avg_launch_ms_in_solve": round(sweep_ms, 5),
                     "avg_launch_ms_isolated": round(iso_ms, 5)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
