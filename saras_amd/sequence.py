"""Sequences of pressure solves with different initial guesses, all in HBM.

A time-stepping run solves L p_n = f_n once per step with f_n changing smoothly.  The variants
below differ only in the initial guess handed to the multigrid solve and in its precision:

  fp64_zero     FP64 multigrid from p = 0
  fp64_warm     FP64 multigrid from the previous solution (what the reference's mgSolve gets:
                hydro_d3 passes the previous pressure correction in, hydro_d3.cc:343)
  mixed_warm    FP32 V-cycles under FP64 refinement, from the previous solution
  mixed_extrap  mixed precision from 2 p_{n-1} - p_{n-2}
  learned_mixed mixed precision from the learned guess (include/saras_guess.h), trained online on
                the solves of the first `ntrain` steps, then frozen
  learned_fp64  FP64 multigrid from the learned guess
  gcr_warm      FP32 V-cycle preconditioned GCR(1) under the FP64 residual (SR_PREC_MIXED_GCR),
                from the previous solution
  gcr_extrap    the same from 2 p_{n-1} - p_{n-2}
  learned_gcr   the same from the learned guess
  linear_gcr    the same from the linear predictor alone (the learned guess's base, no network)

Each solve runs to the same stop rule (relative residual 1e-10 on the projected right-hand side);
the time of a step is the guess (its prediction before the solve, its training and recording of the
solved pair after it) plus the solve, with the right-hand side already in HBM."""
import copy
import time

import numpy as np

from . import _lib as _L
from . import mg

VARIANTS = ("fp64_zero", "fp64_warm", "mixed_warm", "mixed_extrap", "learned_mixed", "learned_fp64", "gcr_warm",
            "gcr_extrap", "learned_gcr", "linear_gcr")


def precision_of(variant):
    if "gcr" in variant:
        return _L.SR_PREC_MIXED_GCR
    if "mixed" in variant:
        return _L.SR_PREC_MIXED
    return _L.SR_PREC_FP64


def net_grid_for(shape):
    """The guess network runs at half the solver's resolution: (N - 1) / 2 + 1 nodes."""
    return tuple((k - 2 - 1) // 2 + 1 for k in shape)


def upload_rhs(seq, nsteps):
    """The first `nsteps` right-hand sides of `seq`, each in its own HBM buffer."""
    out = []
    for n in range(nsteps):
        d = _L.DeviceArray(seq.shape)
        d.upload(seq.rhs(n))
        out.append(d)
    return out


def run(seq, variant, nsteps, ntrain=0, net=None, seed=0, lr=1e-3, on_step=None, rhs=None, coarse_direct=0,
        history=0):
    """Run `nsteps` solves of the sequence `seq` (workloads.EvolvingPoisson).  Returns a list of
    per-step records and the network handle used (if any).  `rhs`: optional list of device buffers
    holding the right-hand sides (upload_rhs); otherwise each is uploaded before its step.
    coarse_direct: solve the coarsest level exactly (sr_mg_desc.coarseDirect).
    history: pairs kept by the learned variants' linear predictor (sr_guess_desc.history); after
    the `ntrain` training solves the network is frozen and used adaptively (policy 2)."""
    assert variant in VARIANTS, variant
    p = copy.copy(seq.p)
    p.precision = precision_of(variant)
    p.coarseDirect = int(coarse_direct)
    solver = mg.Multigrid(seq.g, p)
    shape = seq.shape
    d_f = _L.DeviceArray(shape)
    d_zero = _L.DeviceArray(shape)
    d_zero.upload(np.zeros(shape))
    ring = [_L.DeviceArray(shape) for _ in range(3)]
    lib = _L.load()
    need_net = variant.endswith("extrap") or variant.startswith("learned") or variant.startswith("linear")
    if need_net and net is None:
        from .guess import GuessNet
        net = GuessNet(net_grid_for(shape), seed=seed, lr=lr, fine=tuple(k - 2 for k in shape), history=history)
        if variant.startswith("linear"):
            net.set_policy(3)
    recs = []
    for n in range(nsteps):
        if rhs is None:
            d_f.upload(seq.rhs(n))
            f_ptr = d_f.ptr
        else:
            f_ptr = rhs[n].ptr
        cur = ring[n % 3]
        p1 = ring[(n - 1) % 3] if n >= 1 else None
        p2 = ring[(n - 2) % 3] if n >= 2 else None
        rec = {"step": n}
        if variant.startswith("learned") and n == ntrain and ntrain > 0:
            net.set_policy(2)
        t0 = time.perf_counter()
        if variant == "fp64_zero" or p1 is None:
            _L.check(lib.sr_memcpy(cur.ptr, d_zero.ptr, cur.nbytes, 3))
        elif variant.endswith("warm"):
            _L.check(lib.sr_memcpy(cur.ptr, p1.ptr, cur.nbytes, 3))
        elif variant.endswith("extrap"):
            net.extrapolate(p1.ptr, p2.ptr if p2 else None, cur.ptr)
        else:
            st = net.predict(solver, p1.ptr, p2.ptr if p2 else None, f_ptr, cur.ptr)
            rec.update(choice=st.choice, rel_warm=st.rel_warm, rel_extrap=st.rel_extrap,
                       rel_learned=st.rel_learned, rel_linear=st.rel_linear, net_used=st.net_used,
                       predict_ms=st.predict_ms)
        t1 = time.perf_counter()
        stats = solver.mgSolve_device(cur.ptr, f_ptr)
        t2 = time.perf_counter()
        rec.update(ms=(t2 - t0) * 1e3, guess_ms=(t1 - t0) * 1e3, solve_ms=(t2 - t1) * 1e3, cycles=stats.cycles,
                   converged=bool(stats.converged), residual=stats.residual, sweep0_ms=stats.sweep0_ms,
                   residual0=stats.residual0)
        hist = solver.history()
        rec["first_cycle_residual"] = float(hist[0]) if len(hist) else None
        # the guess's bookkeeping after the solve (training, recording the pair) counts as guess time
        t3 = time.perf_counter()
        if variant.startswith("learned") and p1 is not None and n < ntrain:
            rec["train_loss"] = net.train(cur.ptr)
        if variant.startswith("learned") or variant.startswith("linear"):
            net.observe(solver, f_ptr, cur.ptr)
        post_ms = (time.perf_counter() - t3) * 1e3
        rec["ms"] += post_ms
        rec["guess_ms"] += post_ms
        recs.append(rec)
        if on_step:
            on_step(rec)
    last = [r.download() for r in (ring[(nsteps - 1) % 3],)]
    solver.close()
    for a in ring + [d_f, d_zero]:
        a.free()
    return recs, net, last[0]


def summary(recs, first):
    """Mean ms, cycles and solve time over steps >= first."""
    sel = [r for r in recs if r["step"] >= first]
    out = {"steps": len(sel), "ms_per_solve": float(np.mean([r["ms"] for r in sel])),
           "cycles_mean": float(np.mean([r["cycles"] for r in sel])),
           "all_converged": all(r["converged"] for r in sel),
           "max_final_residual": float(max(r["residual"] for r in sel)),
           "guess_ms_mean": float(np.mean([r["guess_ms"] for r in sel]))}
    if "choice" in sel[0]:
        out["choices"] = {name: int(sum(1 for r in sel if r.get("choice") == c))
                          for c, name in ((0, "warm"), (1, "extrapolation"), (2, "learned"), (3, "linear"))}
        out["net_used"] = int(sum(r.get("net_used", 0) for r in sel))
        for key in ("rel_extrap", "rel_learned", "rel_warm", "rel_linear"):
            vals = [r[key] for r in sel if r.get(key, -1.0) >= 0.0]
            out[key + "_median"] = float(np.median(vals)) if vals else None
    vals = [r["residual0"] for r in sel if r.get("residual0", -1.0) >= 0.0]
    out["guess_residual_median"] = float(np.median(vals)) if vals else None
    return out
