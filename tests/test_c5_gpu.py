"""Config C5: Rayleigh-Benard convection on 257x257x129 nodes (256x256x128 cells) with the learned
guess trained online over the first 500 steps: PPN, tanh z (beta 1), L = (2, 2, 1), Ra = 1e7,
Pr = 1, rbcType 2, dt = 0.01, a seeded 1e-3 temperature perturbation, every pressure solve to a
relative residual of 1e-10 (scalar_d3::timeAdvance, scalar_d3.cc:243-370; the guess replaces the
warm start of scalar_d3.cc:329).

Three runs with the GCR-accelerated mixed solver train the guess identically over 500 steps
(training is deterministic: they reach the same state bit for bit), then run the same steps with
the frozen network used adaptively (policy 2), with the least-squares predictor alone (policy 3),
and with the network tried on every step (policy 0 without training: its guess is kept only where
its residual is lower than the predictor's); a fourth run takes the reference's path (plain FP64
multigrid, the previous correction as the initial guess, the reference's residual pads) over all
the steps.  What is asserted: every solve of every run meets its tolerance; the runs' fields agree
(the guess and the solver change where each solve starts and how it iterates, not where it ends);
the gate hands the solver the lower of the two residuals; and what the network adds over the
predictor alone, in V-cycles per solve and in the guess's residual, is printed and bounded: the
adaptive policy costs at most a quarter cycle per solve on average."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TRAIN, STEPS = 500, 20


def _desc(**over):
    from saras_amd import _lib, ts
    kw = dict(ind=(8, 8, 7), L=(2.0, 2.0, 1.0), domain="PPN", mesh="UUD", beta=(1.0, 1.0, 1.0), scalar=True,
              rbc_type=2, Ra=1e7, Pr=1.0, dt=0.01, prob_type=5, Re=1.0, cn_tol=1e-6,
              semantics=_lib.SR_SEM_PROJECTED, rel_tol=1e-10, vc_count=100, pre=2, post=2, res_type=2, depth=6)
    kw.update(over)
    return ts.ts_desc(**kw)


def _learned_run(policy):
    from saras_amd import _lib, guess, ts
    h = ts.Hydro(_desc(precision=_lib.SR_PREC_MIXED_GCR, coarse_direct=1, inner=(2, 2), pad_mode=1))
    net = guess.GuessNet((129, 129, 65), fine=(257, 257, 129), seed=0, lr=1e-3, history=6)
    h.attach_guess(net, train=True)
    h.initRBC(seed=0, amp=1e-3)
    recs = [h.timeAdvance() for _ in range(TRAIN)]
    trained_P = h.download(ts.P)
    h.attach_guess(net, train=False)
    net.set_policy(policy)
    recs += [h.timeAdvance() for _ in range(STEPS)]
    fields = {w: h.download(w) for w in (ts.VX, ts.VY, ts.VZ, ts.T, ts.P)}
    h.close()
    net.close()
    return recs, trained_P, fields


def _centred(a):
    c = a[1:-1, 1:-1, 1:-1]
    return c - c.mean()


def test_c5_trained_guess_against_predictor_alone_and_plain_fp64():
    from saras_amd import _lib, ts
    adaptive, P_a, f_a = _learned_run(2)
    predictor, P_p, f_p = _learned_run(3)
    tried, P_t, f_t = _learned_run(0)
    plain = ts.Hydro(_desc(precision=_lib.SR_PREC_FP64))
    plain.initRBC(seed=0, amp=1e-3)
    recs_fp64 = [plain.timeAdvance() for _ in range(TRAIN + STEPS)]
    f_fp64 = {w: plain.download(w) for w in (ts.VX, ts.VY, ts.VZ, ts.T, ts.P)}
    plain.close()
    # every solve of every run meets its stop rule
    for name, recs in (("adaptive", adaptive), ("predictor", predictor), ("tried", tried), ("fp64", recs_fp64)):
        bad = [i for i, r in enumerate(recs) if not (r.mg_converged and r.mg_residual < 1e-10)]
        assert not bad, (name, bad[:5])
    # identical training: the learned runs reach the same state
    assert np.array_equal(P_a, P_p) and np.array_equal(P_a, P_t)
    # the fields agree with the reference's path after all the steps
    for w in (ts.VX, ts.VY, ts.VZ, ts.T):
        scale = max(float(np.max(np.abs(f_fp64[w]))), 1e-300)
        for f in (f_a, f_p, f_t):
            assert np.max(np.abs(f[w] - f_fp64[w])) <= 1e-6 * scale, w
    for f in (f_a, f_p, f_t):
        a, b = _centred(f[ts.P]), _centred(f_fp64[ts.P])
        assert np.max(np.abs(a - b)) <= 1e-6 * np.max(np.abs(b))
    # with the network tried on every step, both residuals are exact and the solver gets the lower
    steps = tried[TRAIN:]
    assert all(r.guess_rel_learned >= 0.0 and r.guess_rel_linear >= 0.0 for r in steps)
    for r in steps:
        low = min(r.guess_rel_linear, r.guess_rel_learned)
        assert abs(r.rel_guess - low) <= 1e-6 * low + 1e-300, (r.rel_guess, r.guess_rel_linear, r.guess_rel_learned)
    ratio = np.median([r.guess_rel_learned / r.guess_rel_linear for r in steps])
    # what the network adds over the predictor alone in the steps after training
    ca = np.mean([r.mg_cycles for r in adaptive[TRAIN:]])
    cp = np.mean([r.mg_cycles for r in predictor[TRAIN:]])
    ct = np.mean([r.mg_cycles for r in steps])
    cf = np.mean([r.mg_cycles for r in recs_fp64[TRAIN:]])
    chosen = sum(1 for r in steps if r.guess_choice == 2)
    ra = np.median([r.rel_guess for r in adaptive[TRAIN:]])
    rp = np.median([r.rel_guess for r in predictor[TRAIN:]])
    losses = [r.train_loss for r in adaptive[:TRAIN] if r.train_loss >= 0]
    print(f"C5 257x257x129 after {TRAIN} training steps, {STEPS} steps: cycles per solve adaptive network "
          f"{ca:.2f}, predictor alone {cp:.2f}, network tried every step {ct:.2f} (chosen {chosen}/{STEPS}), "
          f"plain FP64 warm start {cf:.2f}; median guess residual {ra:.2e} vs {rp:.2e}; network's guess residual "
          f"over the predictor's, median {ratio:.3g}; training loss first {np.round(losses[:3], 3)} "
          f"last {np.round(losses[-3:], 3)}; Jacobi sweeps per step {[list(r.jacobi_iters) for r in steps]}")
    assert ca <= cp + 0.25, (ca, cp)
    assert cp < cf
