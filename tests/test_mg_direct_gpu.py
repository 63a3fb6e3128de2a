"""The exact coarsest-level solve (sr_mg_desc.coarseDirect) on the device.

What pins it: the solution of the coarsest linear system itself.  For a right-hand side made as
f = L v from a random v (so f lies in the operator's range), the device's u must satisfy L u = f to
rounding; where L is non-singular (Dirichlet walls) u must equal v; where it is singular u may
differ from v only by a constant.  Whole solves with the direct coarsest level reach the same
tolerance and the same solution as the Gauss-Seidel coarsest level, in FP64 and in both
mixed-precision modes."""
import numpy as np
import pytest

from oracle import pyoracle as po
from helpers import Problem, r2_walls

pytestmark = pytest.mark.gpu

CASES = {
    "ppn_tanh": dict(ind=(5, 5, 4), domain="PPN", mesh="UUD"),
    "ppp": dict(ind=(4, 5, 3), domain="PPP", mesh="UUU", L=(6.283185307179586,) * 3),
    "nnn_stretched": dict(ind=(4, 4, 4), domain="NNN", mesh="DDD", beta=(1.2, 0.8, 1.5)),
    "dirichlet": dict(ind=(4, 4, 4), domain="NNN", mesh="UUU", test_poisson=True, walls=r2_walls((17, 17, 17))),
}


@pytest.mark.parametrize("case", list(CASES))
def test_direct_coarsest_solves_the_coarsest_system(case):
    from saras_amd import _lib as mg
    pb = Problem(**CASES[case])
    g = pb.gpu(coarse_direct=1)
    D = pb.depth
    g.vLevel = D
    g.set_zeroBC(1)
    shape = tuple(k + 2 for k in g.dims(D))
    rng = np.random.default_rng(3)
    v = np.zeros(shape)
    v[1:-1, 1:-1, 1:-1] = rng.standard_normal(tuple(k - 2 for k in shape))
    # f = L v with the coarse-level ghost images of v
    g.upload(mg.SR_PRESSURE, D, v)
    g.upload(mg.SR_RESIDUAL, D, np.zeros(shape))
    g.imposeBC()
    g.computeResidual()
    f = -g.download(mg.SR_TMP, D)
    f[0, :, :] = f[-1, :, :] = 0.0
    f[:, 0, :] = f[:, -1, :] = 0.0
    f[:, :, 0] = f[:, :, -1] = 0.0
    g.upload(mg.SR_RESIDUAL, D, f)
    g.upload(mg.SR_PRESSURE, D, np.zeros(shape))
    g.solve()
    assert g.info(5) == 0
    u = g.download(mg.SR_PRESSURE, D)
    g.computeResidual()
    r = g.download(mg.SR_TMP, D)[1:-1, 1:-1, 1:-1]
    fc = f[1:-1, 1:-1, 1:-1]
    assert np.max(np.abs(r)) <= 1e-11 * np.max(np.abs(fc)), (np.max(np.abs(r)), np.max(np.abs(fc)))
    uc, vc = u[1:-1, 1:-1, 1:-1], v[1:-1, 1:-1, 1:-1]
    if pb.test_poisson:
        assert np.max(np.abs(uc - vc)) <= 1e-10 * np.max(np.abs(vc))
    else:
        # singular: unique up to the constant; the bordered inverse returns the zero-sum solution
        assert abs(uc.sum()) <= 1e-10 * np.abs(uc).sum()
        d = (uc - uc.mean()) - (vc - vc.mean())
        assert np.max(np.abs(d)) <= 1e-9 * np.max(np.abs(vc))


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["fp64", "mixed", "gcr"])
# (the anisotropic fully periodic case is left out: its V-cycle diverges with either coarsest
# solve, in the oracle as on the device, so there is no converged solution to compare)
@pytest.mark.parametrize("case", ["ppn_tanh", "nnn_stretched", "dirichlet"])
def test_solves_with_direct_coarsest_match(case, precision):
    kw = dict(CASES[case])
    pb = Problem(semantics=po.SEM_PROJECTED, rel_tol=1e-10, vc_count=80, **kw)
    f = pb.random_rhs(5)
    if pb.test_poisson:
        f = np.zeros(pb.shape)
        f[1:-1, 1:-1, 1:-1] = 1.0
    pg, pd = np.zeros(pb.shape), np.zeros(pb.shape)
    sg = pb.gpu(precision=precision).mgSolve(pg, f)
    sd = pb.gpu(precision=precision, coarse_direct=1).mgSolve(pd, f)
    assert sd.converged and sd.residual < 1e-10, (sd.cycles, sd.residual)
    assert sg.converged
    assert sd.cycles <= sg.cycles + 1, (sd.cycles, sg.cycles)
    a, b = pd[1:-1, 1:-1, 1:-1], pg[1:-1, 1:-1, 1:-1]
    if not pb.test_poisson:
        a, b = a - a.mean(), b - b.mean()
    assert np.max(np.abs(a - b)) <= 1e-7 * np.max(np.abs(b))
    # the oracle confirms the FP64 residual of the direct-coarsest solution
    o = pb.oracle()
    o.level = 0
    o.array(po.P, 0)[...] = pd
    o.array(po.R, 0)[...] = o.project_rhs(f.copy()) if not pb.test_poisson else f
    assert o.compute_error(2) < 1.2e-10
