"""Pins of the oracle's Nusselt number, reference-semantics mgSolve and padMode-1 residual pads
(the checks live in pins.py so the mutation test can run them against broken builds), and the
behaviour of the reference's V-cycle on z-periodic anisotropic grids.  CPU only."""
import numpy as np
import pytest

import pins
from helpers import Problem
from oracle import pyoracle as po


@pytest.mark.parametrize("pin", [pins.nusselt_xmode, pins.nusselt_face_pair, pins.nusselt_tanh_metric,
                                 pins.reference_mean_over_padded, pins.reference_stop_rule,
                                 pins.reference_restype0_one_cycle, pins.pad_mode1_cosine_modes,
                                 pins.pad_mode1_weighted_transpose], ids=lambda f: f.__name__)
def test_pin(pin):
    pin()


def _growth(hist):
    """geometric mean growth factor of the residual over the last six cycles"""
    return float((hist[-1] / hist[-7]) ** (1.0 / 6.0))


def test_z_periodic_anisotropic_cycle_diverges_in_the_reference_algorithm():
    """33x33x17 nodes on the unit cube (hz = 2 hx), periodic z, V(2,2), depth 3: the reference's
    own lexicographic cycle grows the residual, as the red-black one does.  What drives it: the
    reference's zero z pads for restriction (poisson_d3.cc:91, 110) roughly double the growth --
    wrap pads (padMode 1) take it from about 2 to about 1.2 per cycle -- and the weak z coupling
    that four point sweeps leave unsmoothed does the rest: the same grid with equal spacings
    (Lz = 0.5) or with V(8,8) converges with the reference's pads."""
    def run(sm, pad_mode=0, L=(1.0, 1.0, 1.0), pre=2, domain="PPP"):
        pb = Problem(ind=(5, 5, 4), L=L, domain=domain, mesh="UUU", depth=3, semantics=po.SEM_PROJECTED,
                     rel_tol=1e-10, vc_count=12, pad_mode=pad_mode, pre=pre, post=pre)
        o = pb.oracle(smoother=sm)
        _, hist = o.solve(np.zeros(pb.shape), pb.random_rhs(0))
        o.close()
        return hist

    lex, rb = run(po.SMOOTH_LEX), run(po.SMOOTH_RB)
    assert _growth(lex) > 1.5 and _growth(rb) > 1.5 and lex[-1] > lex[0]
    rb_wrap = run(po.SMOOTH_RB, pad_mode=1)
    assert 1.0 < _growth(rb_wrap) < 0.7 * _growth(rb)
    assert _growth(run(po.SMOOTH_RB, L=(1.0, 1.0, 0.5))) < 0.5
    assert _growth(run(po.SMOOTH_RB, pre=8)) < 0.5
    # the same anisotropy with Neumann z converges with the reference's pads
    assert _growth(run(po.SMOOTH_RB, domain="PPN")) < 0.5
