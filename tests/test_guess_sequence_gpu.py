"""The learned guess inside sequences of solves (include/saras_guess.h with saras_mg.h/saras_ts.h).

What is checked: every solve still reaches its own tolerance and the converged fields agree with
the runs without the guess (the guess changes where the solve starts, not where it ends); the
predict entry point returns the best of its three candidates by relative residual; online
training lowers the network's loss; a time-stepping run with the guess attached follows the run
without it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _centred(a):
    c = a[1:-1, 1:-1, 1:-1]
    return c - c.mean()


def test_sequence_learned_guess_converges_and_trains():
    from saras_amd import sequence, workloads
    seq = workloads.EvolvingPoisson(ind=5, dt=0.05)
    recs_ref, _, sol_ref = sequence.run(seq, "fp64_zero", 24)
    recs, net, sol = sequence.run(seq, "learned_mixed", 24, ntrain=24, lr=1e-3)
    assert all(r["converged"] and r["residual"] < 1e-10 for r in recs)
    # both runs stop at a relative residual below 1e-10; the solutions then differ by at most the
    # residuals times the norm of the inverse operator (about 1e3 relative on this 33^3 grid)
    a, b = _centred(sol), _centred(sol_ref)
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
    for r in recs[1:]:
        # candidates that were evaluated (-1: not computed for this step)
        cands = {c: v for c, v in ((0, r["rel_warm"]), (1, r["rel_extrap"]), (2, r["rel_learned"])) if v >= 0.0}
        # policy 0: the guess handed to the solve has the smallest relative residual of those
        assert cands[r["choice"]] <= min(cands.values()) * (1 + 1e-12)
    losses = [r["train_loss"] for r in recs if r.get("train_loss") is not None]
    assert len(losses) == 23
    assert np.mean(losses[-5:]) < np.mean(losses[:5]), losses
    k, trained = net.kappa()
    assert trained == 23 and k > 0
    net.close()


@pytest.mark.parametrize("history,gcr", [(0, False), (4, False), (4, True)])
def test_time_steps_with_guess_follow_the_plain_run(history, gcr):
    from saras_amd import _lib, guess, ts
    kw = dict(ind=(5, 5, 4), L=(2.0, 2.0, 1.0), domain="PPN", mesh="UUD", scalar=True, rbc_type=2, Ra=1e5,
              Pr=1.0, dt=5e-3, prob_type=5, Re=1.0, semantics=_lib.SR_SEM_PROJECTED, rel_tol=1e-11, vc_count=60,
              pre=2, post=2, res_type=2)
    plain = ts.Hydro(ts.ts_desc(**kw))
    if gcr:   # Krylov-accelerated mixed precision with the exact coarsest level
        learned = ts.Hydro(ts.ts_desc(precision=_lib.SR_PREC_MIXED_GCR, coarse_direct=1, **kw))
    else:
        learned = ts.Hydro(ts.ts_desc(precision=_lib.SR_PREC_MIXED, **kw))
    net = guess.GuessNet((17, 17, 9), fine=(33, 33, 17), seed=3, history=history)
    learned.attach_guess(net, train=True)
    for h in (plain, learned):
        h.initRBC(seed=1, amp=0.05)
    choices = []
    for step in range(12):
        plain.timeAdvance()
        st = learned.timeAdvance()
        choices.append(st.guess_choice)
        if step >= 1:
            assert st.rel_guess >= 0.0 and st.train_loss >= 0.0
    assert choices[0] == -1 and all(c in (0, 1, 2, 3) for c in choices[1:])
    if history:
        # from the third step on the linear predictor (or the network on top of it) is the guess
        assert all(c in (2, 3) for c in choices[2:]), choices
    for which in (ts.VX, ts.VY, ts.VZ, ts.T):
        a, b = learned.download(which), plain.download(which)
        scale = max(float(np.max(np.abs(b))), 1e-300)
        assert np.max(np.abs(a - b)) <= 1e-8 * scale, which
    a, b = _centred(learned.download(ts.P)), _centred(plain.download(ts.P))
    assert np.max(np.abs(a - b)) <= 1e-8 * np.max(np.abs(b))
    for h in (plain, learned):
        h.close()
    net.close()
