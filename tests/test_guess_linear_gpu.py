"""The linear predictor of the learned guess (sr_guess_desc.history, sr_guess_observe).

What pins it: (1) its residual is known from the least-squares fit without a stencil pass, so the
fit's value must equal the solver's own computeError of the guess (sr_mg_residual), to the
cancellation floor of ||f||^2 - sum c_i^2 (about sqrt(eps) relative) and to the earlier solves'
own residuals; (2) on a
sequence whose right-hand sides span a two-dimensional space the predictor is exact after two
observed solves (the guess's residual is at the level of the solves' tolerance), and the solve
from it needs at most one cycle; (3) the bounded history (keep the newest half when full) keeps
(1) true after it wraps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(history, prec=2):
    from saras_amd import _lib, guess, mg, workloads
    seq = workloads.EvolvingPoisson(ind=5, dt=0.05)
    p = seq.p
    p.precision = prec
    p.coarseDirect = 1
    solver = mg.Multigrid(seq.g, p)
    shape = seq.shape
    net = guess.GuessNet(tuple((k - 3) // 2 + 1 for k in shape), fine=tuple(k - 2 for k in shape), history=history)
    net.set_policy(3)
    bufs = {k: _lib.DeviceArray(shape) for k in ("f", "out", "p1", "p2")}
    return seq, solver, net, bufs


def _step(solver, net, bufs, f, n, lib):
    bufs["f"].upload(f)
    st = None
    if n >= 1:
        st = net.predict(solver, bufs["p1"].ptr, bufs["p2"].ptr if n >= 2 else None, bufs["f"].ptr, bufs["out"].ptr)
    else:
        bufs["out"].upload(np.zeros(f.shape))
    rel_meas = solver.residual_device(bufs["out"].ptr, bufs["f"].ptr)
    stats = solver.mgSolve_device(bufs["out"].ptr, bufs["f"].ptr)
    assert stats.converged and stats.residual < 1e-10
    net.observe(solver, bufs["f"].ptr, bufs["out"].ptr)
    lib.sr_memcpy(bufs["p2"].ptr, bufs["p1"].ptr, bufs["p1"].nbytes, 3)
    lib.sr_memcpy(bufs["p1"].ptr, bufs["out"].ptr, bufs["p1"].nbytes, 3)
    return st, rel_meas, stats


def test_linear_predictor_exact_on_a_two_dimensional_sequence():
    from saras_amd import _lib
    lib = _lib.load()
    seq, solver, net, bufs = _setup(history=4)
    core = tuple(k - 2 for k in seq.shape)
    fa = seq.fields[0].reshape(core)
    fb = seq.fields[3].reshape(core)
    for n in range(6):
        f = np.zeros(seq.shape)
        f[1:-1, 1:-1, 1:-1] = np.cos(0.3 * n) * fa + np.sin(0.5 * n + 0.2) * fb
        st, rel_meas, stats = _step(solver, net, bufs, f, n, lib)
        if n >= 2:
            assert st.choice == 3
            assert rel_meas < 1e-8, (n, rel_meas)
            # the fit's value ||f||^2 - sum c_i^2 cancels to rounding here: its floor is ~sqrt(eps)
            assert abs(st.rel_linear - rel_meas) < 5e-8, (n, st.rel_linear, rel_meas)
            assert stats.cycles <= 1, (n, stats.cycles)
    for b in bufs.values():
        b.free()
    net.close()
    solver.close()


def test_linear_fit_residual_equals_the_solver_residual_through_wraps():
    from saras_amd import _lib
    lib = _lib.load()
    seq, solver, net, bufs = _setup(history=4)
    rels = []
    for n in range(12):
        st, rel_meas, stats = _step(solver, net, bufs, seq.rhs(n), n, lib)
        if n >= 2:
            assert st.choice == 3 and st.rel_linear >= 0.0
            # the fit's residual and computeError of the guess agree (both relative to ||P f||); the
            # fit assumes L x_i = q_i, true up to the earlier solves' residuals (1e-10) scaled by the
            # basis normalisation, a few per cent of a 1e-6 residual at most
            assert abs(st.rel_linear - rel_meas) <= 0.05 * rel_meas + 5e-8, (n, st.rel_linear, rel_meas)
            rels.append(rel_meas)
    # the span of earlier solutions predicts better than the plain warm start would
    assert max(rels) < 0.05, rels
    for b in bufs.values():
        b.free()
    net.close()
    solver.close()
