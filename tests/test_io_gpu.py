"""Probes and restart files on the device path: sr_probe_gather (include/saras_probes.h) against the
fields it reads, ProbesData.dat and restartFile.h5 written by the solvePDE loop
(hydro_d3.cc:131-240), and a run restarted from its restart file (hydro_d3.cc:70-81)."""
import numpy as np
import pytest

from saras_amd import _lib, probes, restart, run, ts
from test_run_params import _with

pytestmark = pytest.mark.gpu


def _small(**kw):
    """17^3 nodes, depth 2, one process; the probe block's second set [1:30:3, 7, 1:30:3] of the
    template becomes [1:12:3, 7, 1:12:3], inside this grid"""
    base = dict(X_Index="4", Y_Index="4", Z_Index="4", X_Number_of_Procs="1", Y_Number_of_Procs="1")
    base.update(kw)
    text = _with({"V-Cycle Depth": "2"}, **base).replace("[1:30:3, 7, 1:30:3]", "[1:12:3, 7, 1:12:3]")
    return run.read_parameters(text)


def test_probe_gather_reads_every_field_at_its_index():
    p = _small()
    h = ts.Hydro(run.desc_from(p))
    rng = np.random.default_rng(0)
    ws = [ts.VX, ts.VY, ts.VZ, ts.P]
    host = {w: rng.standard_normal(h.shape(w)) for w in ws}
    for w in ws:
        h.upload(w, host[w])
    fields = [(h.device_ptr(w), tuple(s - 2 for s in h.shape(w))) for w in ws]
    plist = [(0, 0, 0), (5, 7, 9), (15, 15, 15), (-1, 3, 16), (2, -1, -1)]
    vals = probes.gather(fields, plist)
    for n, (i, j, k) in enumerate(plist):
        for f, w in enumerate(ws):
            assert vals[n, f] == host[w][i + 1, j + 1, k + 1], (n, f)
    # z = 17 is a pad index of the node fields but lies beyond Vz's padded array
    with pytest.raises(_lib.SarasError, match="outside the padded range"):
        probes.gather(fields, [(1, 1, 17)])
    h.close()


def _lines(path):
    return path.read_text().splitlines()


def test_run_writes_probes_and_restart_and_continues_from_it(tmp_path):
    kw = dict(Record_Probes="true", Probe_Time_Interval="0.002", Restart_Write_Interval="0.003",
              Final_Time="0.006")
    # A: straight to t = 0.006
    a = tmp_path / "a"
    rows_a, steps = run.solve_pde(_small(**kw), out_dir=str(a), log=lambda s: None)
    assert steps == 6
    lines = _lines(a / "output" / "ProbesData.dat")
    assert lines[0] == "#VARIABLES = Time, X, Y, Z, Vx, Vy, Vz, P"
    plist = probes.parse_probes(" [5, 5, 5] [1:12:3, 7, 1:12:3]\n")
    assert len(plist) == 10 and len(lines) == 1 + 4 * 10            # t = 0, 0.002, 0.004, 0.006
    assert [l.split("\t")[0] for l in lines[1::10]] == ["0.0000", "0.0020", "0.0040", "0.0060"]
    # the last probe lines against the restart file written at the same time
    # 17 nodes per direction: 16 faces; padded shapes
    fields = {n: np.zeros(s) for n, s in (("Vx", (18, 19, 19)), ("Vy", (19, 18, 19)), ("Vz", (19, 19, 18)),
                                          ("P", (19, 19, 19)))}
    t_rs = restart.read(str(a / "output" / "restartFile.h5"), fields)
    assert abs(t_rs - 0.006) < 1e-12
    for line, pr in zip(lines[-10:], plist):
        vals = [float(x) for x in line.split("\t")[4:]]
        for v, name in zip(vals, ("Vx", "Vy", "Vz", "P")):
            assert abs(v - fields[name][pr[0] + 1, pr[1] + 1, pr[2] + 1]) <= 5e-9, (pr, name)
    # B: to t = 0.003 (restart file at 0.003), then restarted to t = 0.006
    b = tmp_path / "b"
    run.solve_pde(_small(**dict(kw, Final_Time="0.003")), out_dir=str(b), log=lambda s: None)
    rows_b, steps_b = run.solve_pde(_small(**dict(kw, Restart_Run="true")), out_dir=str(b), log=lambda s: None)
    assert steps_b == 3
    assert [round(r[0], 6) for r in rows_b] == [0.003, 0.004, 0.005, 0.006]
    # the restarted run starts from the saved velocity; its pressure correction starts from zero as
    # the reference's does, so the solves differ within their tolerance
    e_a, e_b = rows_a[-1][1], rows_b[-1][1]
    assert abs(e_a - e_b) <= 1e-6 * abs(e_a), (e_a, e_b)
    ts_lines = _lines(b / "output" / "TimeSeries.dat")
    assert sum(l.startswith("#VARIABLES") for l in ts_lines) == 2       # the header again on restart
    pb = _lines(b / "output" / "ProbesData.dat")
    assert sum(l.startswith("#VARIABLES") for l in pb) == 1             # no probe header on restart
    assert [l.split("\t")[0] for l in pb[1::10]] == ["0.0000", "0.0020", "0.0040", "0.0060"]


def test_restart_load_restores_the_saved_state(tmp_path):
    """save then load into a fresh handle: the velocity and temperature, their boundary values
    imposed after reading, are those saved; P's core is restored."""
    p = _small(Problem_Type="5", Force="3", RBC_Type="2", Domain_Type='"PPN"', Rayleigh_Number="1.0e5",
               Prandtl_Number="1.0", X_Length="1.0", Y_Length="1.0", Z_Length="1.0")
    h = ts.Hydro(run.desc_from(p))
    h.initRBC(seed=3, amp=1e-2)
    for _ in range(3):
        h.timeAdvance()
    path = str(tmp_path / "output" / "restartFile.h5")
    restart.save(h, path, 0.75)
    g = ts.Hydro(run.desc_from(p))
    assert restart.load(g, path) == 0.75
    for w in (ts.VX, ts.VY, ts.VZ, ts.T):
        assert np.array_equal(g.download(w), h.download(w)), w
    assert np.array_equal(g.download(ts.P)[1:-1, 1:-1, 1:-1], h.download(ts.P)[1:-1, 1:-1, 1:-1])
    h.close()
    g.close()
