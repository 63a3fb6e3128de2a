"""A SARAS run from the reference's parameter file, on the device path of saras_amd.ts.

* ``read_parameters``: input/parameters.yaml, the keys and sections lib/io/parser.cc:162-235 reads,
  into a ``Parameters`` record with parser.h's field names; ``check`` restates parser::checkData
  (parser.cc:249-340), raising ValueError where the reference prints and exits.
* ``desc_from``: the run descriptor (include/saras_ts.h) of a parameter set.  Options this build does
  not implement raise NotImplementedError instead of running something else (adaptive time step,
  channel flows and their forcings, rotation, the heating plate, single-sided stretching,
  fourth-order differences, Euler integration).
* ``TimeSeries``: output/TimeSeries.dat and the console line of tseries (lib/io/tseries.cc:62-246),
  the energy, divergence and Nusselt number reduced on the device.
* ``solve_pde``: the time loop of hydro_d3::solvePDE (src/solvers/hydro/hydro_d3.cc:131-240) and its
  scalar_d3 counterpart: the time series every ``ioCnt`` steps, probes every ``prInt``
  (output/ProbesData.dat, saras_amd.probes) and the restart file every ``rsInt``
  (output/restartFile.h5, saras_amd.restart) until ``tMax``; with ``restartFlag`` the run starts
  from the restart file.  Solution files (writeSolution / writeTarang) are not written.

The MPI decomposition keys are read and validated as the reference does, but one GPU holds the
whole domain.
"""
import dataclasses
import math
import os

import numpy as np

from . import ts

# (section, key, field, type) in the order of parser.cc:162-235
_KEYS = (
    ("Program", "Problem Type", "probType", int),
    ("Program", "Initial Condition", "icType", int),
    ("Program", "Domain Type", "domainType", str),
    ("Program", "RBC Type", "rbcType", int),
    ("Program", "Reynolds Number", "Re", float),
    ("Program", "Rossby Number", "Ro", float),
    ("Program", "Rayleigh Number", "Ra", float),
    ("Program", "Prandtl Number", "Pr", float),
    ("Program", "Taylor Number", "Ta", float),
    ("Program", "X Length", "Lx", float),
    ("Program", "Y Length", "Ly", float),
    ("Program", "Z Length", "Lz", float),
    ("Program", "Force", "forceType", int),
    ("Program", "Heating Plate", "nonHgBC", bool),
    ("Program", "Plate Radius", "patchRadius", float),
    ("Mesh", "Mesh Type", "meshType", str),
    ("Mesh", "X Beta", "betaX", float),
    ("Mesh", "Y Beta", "betaY", float),
    ("Mesh", "Z Beta", "betaZ", float),
    ("Mesh", "X Index", "xInd", int),
    ("Mesh", "Y Index", "yInd", int),
    ("Mesh", "Z Index", "zInd", int),
    ("Parallel", "Number of OMP threads", "nThreads", int),
    ("Parallel", "X Number of Procs", "npX", int),
    ("Parallel", "Y Number of Procs", "npY", int),
    ("Solver", "Differentiation Scheme", "dScheme", int),
    ("Solver", "Integration Scheme", "iScheme", int),
    ("Solver", "Solve Tolerance", "cnTolerance", float),
    ("Solver", "Restart Run", "restartFlag", bool),
    ("Solver", "Use CFL Condition", "useCFL", bool),
    ("Solver", "Courant Number", "courantNumber", float),
    ("Solver", "Time-Step", "tStp", float),
    ("Solver", "Final Time", "tMax", float),
    ("Solver", "I/O Count", "ioCnt", int),
    ("Solver", "Solution Format", "solnFormat", int),
    ("Solver", "Solution Write Interval", "fwInt", float),
    ("Solver", "Restart Write Interval", "rsInt", float),
    ("Solver", "Record Probes", "readProbes", bool),
    ("Solver", "Probe Time Interval", "prInt", float),
    ("Solver", "Probes", "probeCoords", str),
    ("Multigrid", "V-Cycle Depth", "vcDepth", int),
    ("Multigrid", "V-Cycle Count", "vcCount", int),
    ("Multigrid", "Solve Tolerance", "mgTolerance", float),
    ("Multigrid", "Smoothing Method", "gsSmooth", int),
    ("Multigrid", "Pre-Smoothing Count", "preSmooth", int),
    ("Multigrid", "Post-Smoothing Count", "postSmooth", int),
    ("Multigrid", "Residual Type", "resType", int),
    ("Multigrid", "Print Residual", "printResidual", bool),
)

Parameters = dataclasses.make_dataclass("Parameters", [(f, t) for _, _, f, t in _KEYS])
Parameters.__doc__ = "The parser fields of lib/io/parser.h filled from input/parameters.yaml."


def _convert(value, typ, where):
    """A scalar as yaml-cpp's as<T>() converts it: numbers from their text (so 1.0e5, which YAML 1.1
    loaders keep as a string, is a number), booleans only from the YAML boolean words."""
    if typ is bool:
        if not isinstance(value, bool):
            raise ValueError(f"{where}: expected true or false, got {value!r}")
        return value
    if typ is str:
        return str(value)
    if isinstance(value, str):
        try:
            value = int(value) if typ is int else float(value)
        except ValueError:
            raise ValueError(f"{where}: expected {'an integer' if typ is int else 'a number'}, got {value!r}")
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise ValueError(f"{where}: expected a number, got {value!r}")
    if typ is int and float(value) != int(value):
        raise ValueError(f"{where}: expected an integer, got {value!r}")
    return typ(value)


def read_parameters(source):
    """Parameters from a parameters.yaml path or its text (every key of parser.cc:162-235 required,
    as yaml-cpp's as<T>() requires it), then checked like parser::checkData."""
    import yaml
    text = source
    if "\n" not in source and os.path.exists(source):
        with open(source) as fh:
            text = fh.read()
    node = yaml.safe_load(text)
    if not isinstance(node, dict):
        raise ValueError("parameter file: expected the Program / Mesh / Parallel / Solver / Multigrid sections")
    vals = {}
    for sec, key, field, typ in _KEYS:
        if sec not in node or not isinstance(node[sec], dict) or key not in node[sec]:
            raise ValueError(f"parameter file: missing \"{sec}\" / \"{key}\"")
        vals[field] = _convert(node[sec][key], typ, f"{sec} / {key}")
    p = Parameters(**vals)
    check(p)
    return p


def check(p):
    """parser::checkData (parser.cc:249-340) for a 3D build in double precision; the reference's
    warnings (processor counts below 1 set to 1) are applied silently."""
    if p.yInd == 0:
        raise ValueError("Y Index parameter of YAML file is 0 for 3D simulation")
    p.npX = max(p.npX, 1)
    p.npY = max(p.npY, 1)
    if len(p.domainType) != 3:
        raise ValueError("Domain type string is not correct")
    if len(p.meshType) != 3:
        raise ValueError("Mesh type string is not correct")
    if p.tStp > p.tMax:
        raise ValueError("Time step is larger than the maximum duration assigned for simulation")
    coarsest = 2 ** (p.vcDepth + 1)
    for ind, nproc, axis in ((p.xInd, p.npX, "X"), (p.yInd, p.npY, "Y"), (p.zInd, 1, "Z")):
        if (2 ** ind) // nproc < coarsest:
            raise ValueError(f"The grid size and domain decomposition along {axis}-direction results in "
                             "sub-domains too coarse to reach the V-Cycle depth specified")
    if p.resType > 2:
        raise ValueError("The specified value for printing error at end of V-Cycles is not defined")
    return p


def is_scalar(p):
    """Problem types 5-7 solve the temperature too (scalar_d3)."""
    return p.probType >= 5


def desc_from(p, **overrides):
    """The sr_ts_desc of a parameter set (ts.ts_desc keywords), after the checks of what this build runs.
    Keyword overrides go to ts.ts_desc (e.g. semantics, precision, rel_tol for the additions)."""
    unsupported = []
    if p.useCFL:
        unsupported.append("Use CFL Condition (adaptive time step)")
    if p.dScheme != 1:
        unsupported.append("Differentiation Scheme other than 1")
    if p.iScheme != 1:
        unsupported.append("Integration Scheme other than 1 (Crank-Nicolson)")
    if p.nonHgBC and is_scalar(p):
        unsupported.append("Heating Plate")
    if "S" in p.meshType.upper():
        unsupported.append("single-sided tangent-hyperbolic stretching (S)")
    if p.probType in (3, 4, 7) or p.probType < 1 or p.probType > 7:
        unsupported.append(f"Problem Type {p.probType}")
    if is_scalar(p):
        if p.forceType != 3:
            unsupported.append(f"Force {p.forceType} for a scalar run (buoyancy, 3, is built)")
    elif p.forceType != 0:
        unsupported.append(f"Force {p.forceType} for a hydro run (none, 0, is built)")
    if p.icType not in (0, 1):
        unsupported.append(f"Initial Condition {p.icType}")
    if unsupported:
        raise NotImplementedError("not built for one GPU here: " + "; ".join(unsupported))
    mesh = "".join("U" if c.upper() == "U" else "D" for c in p.meshType)
    kw = dict(ind=(p.xInd, p.yInd, p.zInd), L=(p.Lx, p.Ly, p.Lz), domain=p.domainType.upper(), mesh=mesh,
              beta=(p.betaX, p.betaY, p.betaZ), Re=p.Re, dt=p.tStp, cn_tol=p.cnTolerance, depth=p.vcDepth,
              vc_count=p.vcCount, pre=p.preSmooth, post=p.postSmooth, res_type=p.resType, mg_tol=p.mgTolerance,
              prob_type=p.probType, scalar=is_scalar(p), rbc_type=p.rbcType, Ra=p.Ra, Pr=p.Pr,
              gs_smooth=p.gsSmooth)
    kw.update(overrides)
    return ts.ts_desc(**kw)


def format_line(values, time, dt=None):
    """One line of tseries output: time with std::fixed, setw(6), setprecision(4); the first value
    setw(16) and every value with setprecision(8) (tseries.cc:174-177, 238-243); dt appended for the
    file line."""
    out = f"{time:6.4f}\t{values[0]:16.8f}"
    for v in values[1:]:
        out += f"\t{v:.8f}"
    if dt is not None:
        out += f"\t{dt:.8f}"
    return out


class TimeSeries:
    """tseries: output/TimeSeries.dat (header, then one line per write) and the console line."""

    def __init__(self, hydro, p, path, log=print):
        self.h, self.p, self.log = hydro, p, log
        self.scalar = is_scalar(p)
        self.nu, self.kappa = hydro.coefficients()
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        self.fh = open(path, "a" if p.restartFlag else "w")
        if p.probType <= 4:
            self.log(f"{'Time':>6}\t{'Total energy':>16}\tDivergence")
            self.fh.write("#VARIABLES = Time, Energy, Divergence, dt\n")
        else:
            self.log(f"{'Time':>6}\t{'Re (Urms)':>16}\tNusselt No\tDivergence")
            self.fh.write("#VARIABLES = Time, Energy, NusseltNo, Divergence, dt\n")
        self.rows = []

    def write(self, time, dt):
        """writeTSData: maximum divergence (aborting above 1e5, as the reference does), volume-averaged
        energy, and for scalar runs sqrt(E)/nu and the Nusselt number."""
        div = self.h.maxDivergence()
        if div > 1.0e5:
            raise RuntimeError("Divergence exceeds permissible limits")
        energy = self.h.energy()
        if self.scalar:
            values = [math.sqrt(energy) / self.nu, self.h.nusselt(), div]
        else:
            values = [energy, div]
        self.log(format_line(values, time))
        self.fh.write(format_line(values, time, dt) + "\n")
        self.fh.flush()
        self.rows.append((time, *values, dt))
        return values

    def close(self):
        if self.fh:
            self.fh.close()
            self.fh = None


def initialise(h, p, seed=0, amp=0.0):
    """The initial state of hydro_d3's / scalar_d3's constructor (hydro_d3.cc:84-117): P = 1, V from
    the initial condition (0 zero, 1 Taylor-Green), velocity BCs imposed.  Scalar runs start from
    T = amp * U(-1, 1) (seeded) inside the plates; amp = 0 is the reference's T = 0."""
    if is_scalar(p):
        h.initRBC(seed=seed, amp=amp)
        return
    if p.icType == 1:
        h.initTGV()
        return
    for w in (ts.VX, ts.VY, ts.VZ, ts.PP):
        h.upload(w, np.zeros(h.shape(w)))
    h.upload(ts.P, np.ones(h.shape(ts.P)))
    for c in range(3):
        h.imposeVBC(c)


def solve_pde(p, out_dir=".", log=print, seed=0, amp=0.0, max_steps=None, **overrides):
    """The time loop of solvePDE (hydro_d3.cc:131-235): the initial state (or the restart file's),
    the time series, probes and, for fresh runs, nothing else at t = 0; then timeAdvance, t += dt, a
    time-series line every ioCnt steps, probes when |prTime - t| < dt / 2, the restart file when
    |rsTime - t| < dt / 2, until |tMax - t| < dt / 2.  The next probe and restart times start at
    t + interval - fmod(t, interval), as the reference sets them.  Returns the TimeSeries rows and the
    number of steps; max_steps stops early (for tests)."""
    from . import probes as probes_mod
    from . import restart
    h = ts.Hydro(desc_from(p, **overrides))
    series = prb = None
    try:
        out = os.path.join(out_dir, "output")
        rs_path = os.path.join(out, "restartFile.h5")
        if p.restartFlag:
            time = restart.load(h, rs_path)
        else:
            initialise(h, p, seed=seed, amp=amp)
            time = 0.0
        series = TimeSeries(h, p, os.path.join(out, "TimeSeries.dat"), log=log)
        if p.readProbes:
            plist = probes_mod.check_probes(probes_mod.parse_probes(p.probeCoords), (p.xInd, p.yInd, p.zInd))
            prb = probes_mod.Probes(h, plist, os.path.join(out, "ProbesData.dat"), restart=p.restartFlag)
        dt, steps = p.tStp, 0
        series.write(time, dt)
        if not p.restartFlag and prb is not None:
            prb.probe_data(time)
        pr_time = time + p.prInt - math.fmod(time, p.prInt)
        rs_time = time + p.rsInt - math.fmod(time, p.rsInt)
        while True:
            h.timeAdvance()
            steps += 1
            time += dt
            if steps % p.ioCnt == 0:
                series.write(time, dt)
            if prb is not None and abs(pr_time - time) < 0.5 * dt:
                prb.probe_data(time)
                pr_time += p.prInt
            if abs(rs_time - time) < 0.5 * dt:
                restart.save(h, rs_path, time)
                rs_time += p.rsInt
            if abs(p.tMax - time) < 0.5 * dt or (max_steps is not None and steps >= max_steps):
                break
        return series.rows, steps
    finally:
        if series is not None:
            series.close()
        if prb is not None:
            prb.close()
        h.close()


def main(argv=None):
    """python -m saras_amd.run [input/parameters.yaml] [--out DIR]: run the case on cuda:0."""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("parameters", nargs="?", default=os.path.join("input", "parameters.yaml"))
    ap.add_argument("--out", default=".")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--amp", type=float, default=0.0, help="temperature perturbation of scalar runs")
    a = ap.parse_args(argv)
    p = read_parameters(a.parameters)
    rows, steps = solve_pde(p, out_dir=a.out, seed=a.seed, amp=a.amp)
    return 0 if rows else 1


if __name__ == "__main__":
    raise SystemExit(main())
