"""Sweeps of runs, as the reference's multi/ programs describe them: each run
is a rootname with an input file rootname.in (options, ndspmhd_amd.infile)
and, for shock tubes, rootname.shk (left/right states).  The reference runs
them one after another (`ndspmhd run1 run2 ...`, src/ndspmhd.f90:335-360);
here every run of the sweep becomes one member of a batch stepped on the GPU
together (ndsp_batch_step), each writing its own .ev file and dumps.

    python -m ndspmhd_amd.multi run1 run2 ... [--precision 0|1] [--sequential]
    python -m ndspmhd_amd.multi --write-shock-sweep shk 7     (multi_shock)

Shock tubes are set up as src/setup_shock1D_mhd.f90 does in 1D (particle
placement, smoothed states, equal masses), with the first and last six
particles held fixed as the reference's boundary particles (ibound 1,
src/setup_shock1D_mhd.f90:47-48).
"""
from __future__ import annotations

import argparse
import math
import os
import sys
from dataclasses import dataclass, field

import numpy as np

from . import infile as inf
from .options import empty_state
from .setups import fixed_ends_1d

# the seven MHD shock tubes of multi/multi_shock.f90:57-220: psep, gamma,
# tmax, tout, iener override (None: the base file's), polyk, and the states
# rhoL rhoR prL prR vxL vxR vyL vyR vzL vzR Bx ByL ByR BzL BzR
_C = 1.0 / math.sqrt(4.0 * math.pi)
SHOCK_TESTS = [
    ("Brio/Wu", 0.0056, 2.0, 0.1, 0.05, None, None,
     (1.0, 0.125, 1.0, 0.1, 0., 0., 0., 0., 0., 0., 0.75, 1.0, -1.0, 0., 0.)),
    ("Balsara 1", 0.00375, 1.6666666666667, 0.15, 0.075, None, None,
     (1.0, 0.2, 1.0, 0.1, 0., 0., 0., 0., 0., 0., 1., 1.0, 0.0, 0., 0.)),
    ("7 discontinuities", 0.001485, 1.6666666666667, 0.2, 0.1, None, None,
     (1.08, 1.0, 0.95, 1.0, 1.2, 0., 0.01, 0., 0.5, 0., 2. * _C, 3.6 * _C, 4.0 * _C, 2. * _C, 2. * _C)),
    ("isothermal 6 discontinuities", 0.001485, 1.0, 0.2, 0.1, 0, 1.0,
     (1.08, 1.0, 0.95, 1.0, 1.2, 0., 0.01, 0., 0.5, 0., 2. * _C, 3.6 * _C, 4.0 * _C, 2. * _C, 2. * _C)),
    ("rarefaction", 0.002, 1.66666666666666667, 0.1, 0.05, "default", None,
     (1.0, 1.0, 1.0, 1.0, -1.0, 1., 0., 0., 0., 0., 0., 1., 1., 0., 0.)),
    ("two fast shocks", 0.0025, 1.66666666666667, 0.03, 0.015, None, None,
     (1.0, 1.0, 1.0, 1.0, 36.87, -36.87, -0.1546, 0., -0.03864, 0., 4. * _C, 4. * _C, 4. * _C, 1. * _C, 1. * _C)),
    ("2D shock problem", 0.0025, 1.66666666666667, 0.08, 0.04, None, None,
     (1.0, 1.0, 20.0, 1.0, 10.0, -10.0, 0., 0., 0., 0., 5. * _C, 5. * _C, 5. * _C, 0., 0.)),
]


def run_name(prefix: str, i: int) -> str:
    """file stem of run i (1-based) as multi_shock names it"""
    return prefix + (str(i) if i < 10 else "%02d" % i)


def write_shk(path: str, s):
    """the eight records of a .shk file (list-directed, as multi_shock
    writes them, multi/multi_shock.f90:222-231)"""
    rows = [(s[0], s[1]), (s[2], s[3]), (s[4], s[5]), (s[6], s[7]), (s[8], s[9]), (s[10],),
            (s[11], s[12]), (s[13], s[14])]
    with open(path, "w") as f:
        for r in rows:
            f.write(" " + "  ".join("%.17g" % v for v in r) + "\n")


def read_shk(path: str) -> tuple:
    """(rhoL rhoR prL prR vxL vxR vyL vyR vzL vzR Bx ByL ByR BzL BzR) read as
    setup_shock1D_mhd.f90:84-93 reads them: eight list-directed records"""
    counts = (2, 2, 2, 2, 2, 1, 2, 2)
    with open(path) as f:
        lines = [l for l in f.read().splitlines() if l.strip()]
    if len(lines) < 8:
        raise inf.InfileError("%s: expected 8 records, found %d" % (path, len(lines)))
    vals = []
    for line, c in zip(lines, counts):
        tok = line.replace(",", " ").split()
        if len(tok) < c:
            raise inf.InfileError("%s: short record %r" % (path, line))
        vals += [float(t.replace("d", "e").replace("D", "E")) for t in tok[:c]]
    return tuple(vals)


def write_shock_sweep(prefix: str, nruns: int, base: dict | None = None, directory: str = ".") -> list:
    """the .in and .shk files of multi_shock's sweep: run i takes test i's
    psep, gamma, tmax, tout and states over the base options (those of
    multirun.in); runs past the seventh repeat the seventh"""
    v = dict(inf.DEFAULTS if base is None else base)
    iener_default = v["iener"]
    names = []
    for i in range(1, nruns + 1):
        name, psep, gamma, tmax, tout, iener, polyk, states = SHOCK_TESTS[min(i, 7) - 1]
        v.update(psep=psep, gamma=gamma, tmax=tmax, tout=tout)
        if iener == "default":
            v["iener"] = iener_default
        elif iener is not None:
            v["iener"] = iener
        if polyk is not None:
            v["polyk"] = polyk
        stem = os.path.join(directory, run_name(prefix, i))
        write_shk(stem + ".shk", states)
        inf.write_infile(stem + ".in", v)
        names.append(stem)
    return names


def shock_setup_1d(values: dict, shk, dsmooth: float = 20.0):
    """(opts, state) of src/setup_shock1D_mhd.f90:45-175: x in [-0.5, 0.5],
    massp = rho_R psep, particles placed at x(i) = x(i-2) + 2 massp/rho(i-1)
    from the left wall, states joined by a logistic profile in
    delta = 2 (x - xc)/psep over |delta| <= dsmooth; the first and last six
    particles held fixed (ibound 1, nbpts 6, src/setup_shock1D_mhd.f90:47-48,
    src/set_fixedbound.f90:38-53)"""
    (rl, rr, pl, pr_, vxl, vxr, vyl, vyr, vzl, vzr, bx, byl, byr, bzl, bzr) = shk
    psep, gamma, imhd = values["psep"], values["gamma"], values["imhd"]
    xmin, xmax = -0.5, 0.5
    xc = 0.5 * (xmax + xmin)
    gam1 = gamma - 1.0
    if abs(gam1) > 1.0e-3:
        ul, ur = pl / (gam1 * rl), pr_ / (gam1 * rr)
    else:
        ul, ur = 3.0 * pl / (2.0 * rl), 3.0 * pr_ / (2.0 * rr)
    vl, vr = (vxl, vyl, vzl), (vxr, vyr, vzr)
    if imhd == 0:
        bx = byl = byr = bzl = bzr = 0.0
    massp = rr * psep
    x = [xmin + 0.5 * massp / rl, xmin + psep * rr / rl + 0.5 * massp / rl]
    dens, uu, vel, B = [rl, rl], [ul, ul], [vl, vl], [(bx, byl, bzl), (bx, byl, bzl)]
    while x[-1] < xmax:
        delta = 2.0 * (x[-1] - xc) / psep
        if delta > dsmooth:
            d, u, v, b = rr, ur, vr, (bx, byr, bzr)
        elif delta < -dsmooth:
            d, u, v, b = rl, ul, vl, (bx, byl, bzl)
        else:
            e = math.exp(delta)
            d = (rl + rr * e) / (1.0 + e)
            u = (pl + pr_ * e) / ((1.0 + e) * gam1 * d) if abs(gam1) > 1.0e-3 else \
                3.0 * (pl + pr_ * e) / ((1.0 + e) * 2.0 * d)
            v = vr if delta > 0.0 else vl
            b = (bx, (byl + byr * e) / (1.0 + e), (bzl + bzr * e) / (1.0 + e))
        x.append(x[-2] + 2.0 * massp / dens[-1])
        dens.append(d)
        uu.append(u)
        vel.append(v)
        B.append(b)
    n = len(x) - 1  # the last particle placed lies at or beyond xmax
    st = empty_state(n, 1)
    st["x"][:, 0] = x[:n]
    st["pmass"][:] = massp
    d = np.array(dens[:n])
    st["hh"][:] = values["hfact"] * massp / d
    st["en"][:] = uu[:n]
    st["vel"][:] = np.array(vel[:n])
    st["Bevol"][:] = np.array(B[:n])
    o = inf.to_options(values, 1, ibound=(1, 0, 0), xmin=(xmin, 0.0, 0.0), xmax=(xmax, 0.0, 0.0))
    st["alpha"][:, 0] = o.alphamin
    st["alpha"][:, 1] = o.alphaumin
    st["alpha"][:, 2] = o.alphaBmin
    return o, fixed_ends_1d(st, 6)


@dataclass
class Run:
    """one run of a sweep and its evolve-style output bookkeeping
    (src/evolve.f90:52-175)"""
    root: str
    values: dict
    sim: object = None
    ev: object = None
    nsteps: int = 0
    time: float = 0.0
    ifile: int = -1
    noutput: int = 1
    tprint: float = 0.0
    etotin: float = 0.0
    momtotin: float = 0.0
    detot: float = 0.0
    dmomtot: float = 0.0
    warnings: list = field(default_factory=list)

    @property
    def done(self) -> bool:
        return not (self.time < self.values["tmax"] and self.nsteps < self.values["nmax"])


def load_run(root: str, precision: int = 0, strict: bool = False) -> Run:
    """rootname.in and rootname.shk as a Simulation (derivs evaluated)"""
    from . import Simulation
    f = inf.read_infile(root + ".in", strict=strict)
    o, st = shock_setup_1d(f.values, read_shk(root + ".shk"))
    sim = Simulation(o, st, precision=precision)
    sim.derivs()
    r = Run(root, f.values, sim, warnings=f.warnings + (["missing options: " + ", ".join(f.missing)]
                                                          if f.missing else []))
    r.time = 0.0
    return r


def _output(r: Run):
    from . import dumpfile
    r.ifile += 1
    dumpfile.write_sim("%s_%05d.dat" % (r.root, r.ifile), r.sim, r.time)


def start(r: Run, outputs: bool = True):
    from . import evfile
    if outputs:
        r.ev = evfile.EvFile(r.root + ".ev", r.values["imhd"], r.values["idust"])
        row = r.ev.write_sim(r.sim, r.time, 0.0)
        r.etotin, r.momtotin = row["etot"], row["momtot"]
        _output(r)
    r.tprint = r.values["tout"]


def after_step(r: Run, info: dict, outputs: bool = True):
    """evolve's per-step bookkeeping: dump when due, .ev line, and the
    timestep shortened to reach the next output time exactly"""
    r.nsteps += 1
    r.time = info["time"]
    v = r.values
    dt = info["dt"]
    if outputs:
        if r.time >= r.tprint or r.time >= v["tmax"] or (v["nout"] > 0 and r.nsteps % v["nout"] == 0) \
                or r.nsteps >= v["nmax"]:
            _output(r)
            r.noutput += 1
            r.tprint = r.noutput * v["tout"]
        row = r.ev.write_sim(r.sim, r.time, dt)
        r.detot = max(r.detot, abs(row["etot"] - r.etotin))
        r.dmomtot = max(r.dmomtot, abs(row["momtot"] - r.momtotin))
    else:
        if r.time >= r.tprint:
            r.noutput += 1
            r.tprint = r.noutput * v["tout"]
    if dt >= r.tprint - r.time:
        r.sim.set_dt(r.tprint - r.time, r.time)


def finish(r: Run):
    if r.ev:
        r.ev.close()
    r.sim.close()


def run_sweep(roots, precision: int = 0, batched: bool = True, outputs: bool = True,
              max_steps: int | None = None, log=print) -> list:
    """every run to its tmax/nmax (or max_steps steps); batched: all
    unfinished runs step together in one batch"""
    from . import Batch
    runs = [load_run(r, precision) for r in roots]
    for r in runs:
        for w in r.warnings:
            log("%s: warning: %s" % (r.root, w))
        start(r, outputs)
    limit = max_steps if max_steps is not None else 1 << 62
    active = [r for r in runs if not r.done]
    batch = None
    steps = 0
    while active and steps < limit:
        if batched:
            if batch is None:
                batch = Batch([r.sim for r in active])
            batch.step()
            infos = [r.sim.dt_info for r in active]
        else:
            infos = [r.sim.step() for r in active]
        for r, info in zip(active, infos):
            after_step(r, info, outputs)
        steps += 1
        left = [r for r in active if not r.done]
        if len(left) != len(active):
            if batch is not None:
                batch.close()
                batch = None
            active = left
    if batch is not None:
        batch.close()
    for r in runs:
        if outputs and r.nsteps > 0:
            log("%s: %d steps to t = %.6g; max energy error %.2e, max momentum error %.2e" % (
                r.root, r.nsteps, r.time, r.detot / max(abs(r.etotin), 1e-300),
                r.dmomtot / (abs(r.momtotin) if abs(r.momtotin) > 1e-300 else 1.0)))
    return runs


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("runs", nargs="*", help="rootnames (rootname.in, rootname.shk)")
    ap.add_argument("--precision", type=int, default=0, help="0 mixed, 1 FP64")
    ap.add_argument("--sequential", action="store_true", help="step the runs one by one")
    ap.add_argument("--max-steps", type=int, default=None)
    ap.add_argument("--write-shock-sweep", nargs=2, metavar=("PREFIX", "NRUNS"),
                    help="write multi_shock's .in/.shk files (base options from multirun.in if present)")
    a = ap.parse_args(argv)
    if a.write_shock_sweep:
        prefix, n = a.write_shock_sweep[0], int(a.write_shock_sweep[1])
        base = inf.read_infile("multirun.in").values if os.path.exists("multirun.in") else None
        for s in write_shock_sweep(prefix, n, base):
            print("wrote", s + ".in", s + ".shk")
        return 0
    if not a.runs:
        ap.error("no runs given")
    runs = run_sweep(a.runs, a.precision, batched=not a.sequential, max_steps=a.max_steps)
    for r in runs:
        finish(r)
    return 0


if __name__ == "__main__":
    sys.exit(main())
