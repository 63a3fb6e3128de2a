"""A run from a parameter file through solvePDE's loop on the device: the time series it writes
against the run's own diagnostics step by step, and the Taylor-Green energy against the oracle's
time step (lib/io/tseries.cc, src/solvers/hydro/hydro_d3.cc:131-240)."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from saras_amd import run, ts
from test_run_params import _with

pytestmark = pytest.mark.gpu


def _rows(path):
    lines = path.read_text().splitlines()
    return lines[0], [[float(x) for x in l.split("\t")] for l in lines[1:]]


def test_taylor_green_time_series(tmp_path):
    # 17^3 nodes, depth 2, 5 steps of dt 1e-3, a line every 2 steps
    p = run.read_parameters(_with({"V-Cycle Depth": "2", "I/O Count": "2"}, X_Index="4", Y_Index="4",
                                  Z_Index="4", X_Number_of_Procs="1", Y_Number_of_Procs="1"))
    rows, steps = run.solve_pde(p, out_dir=str(tmp_path), log=lambda s: None)
    assert steps == 5
    header, got = _rows(tmp_path / "output" / "TimeSeries.dat")
    assert header == "#VARIABLES = Time, Energy, Divergence, dt"
    assert [r[0] for r in got] == [0.0, 0.002, 0.004]
    assert all(r[3] == 0.001 for r in got)
    # the same steps on a run driven directly: energy and divergence at the written times
    h = ts.Hydro(run.desc_from(p))
    h.initTGV()
    ref = [(h.energy(), h.maxDivergence())]
    for n in range(1, 6):
        h.timeAdvance()
        if n % 2 == 0:
            ref.append((h.energy(), h.maxDivergence()))
    h.close()
    for (t, e, d, _), (e_ref, d_ref) in zip(got, ref):
        assert abs(e - e_ref) <= 5e-9 and abs(d - d_ref) <= 5e-9
    # the energy column against the oracle's restated time step (red-black smoother, as on the device)
    o = po.OracleTS(po.ts_params(ind=(4, 4, 4), depth=2, Re=1600.0, dt=1e-3, pre=4, post=4, vc_count=5,
                                 res_type=1, mg_tol=1e-6, cn_tol=1e-6))
    o.init_tgv()
    e_o = [o.energy()]
    for n in range(1, 5):
        o.time_advance()
        if n % 2 == 0:
            e_o.append(o.energy())
    o.close()
    for r, e in zip(got, e_o):
        assert abs(r[1] - e) <= 5e-9 + 1e-12 * abs(e)


def test_zero_start_lid_driven_cavity(tmp_path):
    """icType 0 with the moving lid of probType 1: P = 1, V = 0 inside, lid BC imposed; the energy
    grows as the lid drives the flow."""
    p = run.read_parameters(_with(Problem_Type="1", Initial_Condition="0", Domain_Type='"NNN"',
                                  Reynolds_Number="100", X_Length="1.0", Y_Length="1.0", Z_Length="1.0",
                                  X_Index="4", Y_Index="4", Z_Index="4", X_Number_of_Procs="1",
                                  Y_Number_of_Procs="1", Final_Time="0.003"))
    rows, steps = run.solve_pde(p, out_dir=str(tmp_path), log=lambda s: None)
    assert steps == 3 and len(rows) == 4
    energies = [r[1] for r in rows]
    assert all(np.isfinite(energies)) and all(b > a for a, b in zip(energies, energies[1:]))


def test_rayleigh_benard_time_series(tmp_path):
    """A scalar run (probType 5, buoyancy): the columns are sqrt(E)/nu, the Nusselt number and the
    divergence of the run at the written times."""
    p = run.read_parameters(_with(Problem_Type="5", Force="3", RBC_Type="2", Domain_Type='"PPN"',
                                  Mesh_Type='"UUD"', Rayleigh_Number="1.0e5", Prandtl_Number="1.0",
                                  X_Length="1.0", Y_Length="1.0", Z_Length="1.0", X_Index="4", Y_Index="4",
                                  Z_Index="4", X_Number_of_Procs="1", Y_Number_of_Procs="1",
                                  Final_Time="0.003", **{"V-Cycle Depth": "2"}))
    rows, steps = run.solve_pde(p, out_dir=str(tmp_path), log=lambda s: None, amp=1e-3)
    header, got = _rows(tmp_path / "output" / "TimeSeries.dat")
    assert header == "#VARIABLES = Time, Energy, NusseltNo, Divergence, dt" and len(got) == 4
    h = ts.Hydro(run.desc_from(p))
    run.initialise(h, p, amp=1e-3)
    nu, kappa = h.coefficients()
    assert nu == pytest.approx(math.sqrt(1.0 / 1e5)) and kappa == pytest.approx(1.0 / math.sqrt(1e5))
    for n, r in enumerate(got):
        if n:
            h.timeAdvance()
        re, nus, div = math.sqrt(h.energy()) / nu, h.nusselt(), h.maxDivergence()
        assert abs(r[1] - re) <= 5e-9 * max(1.0, abs(re)) and abs(r[2] - nus) <= 5e-9 and abs(r[3] - div) <= 5e-9
    h.close()
    assert np.isfinite(np.array(got)).all()
