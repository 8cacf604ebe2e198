"""Reader/writer for the compiled reference harness (oracle/ref/harness.f90).

Input file (little-endian stream):
  int32  magic=1313100880, mode (1 derivs, 2 derivs + neighbour sets), npart, ndim
  int32  iopt[32]      (slots: ndspmhd_amd.options.IOPT)
  float64 dopt[32]     (slots: ndspmhd_amd.options.DOPT)
  float64 x(ndim,N) vel(3,N) pmass(N) hh(N) en(N) Bevol(3,N) alpha(3,N) psi(N) dustevol(N)
  int32   itype(N) ireal(N)  (partner of each fixed particle, 0-based; -1 none)
Mode 3 is the binding harness (oracle/ref/bindtest.f90: the reference with
its hot-path files replaced by the GPU binding, stepped as evolve does).
Output file:
  int32  npart, ntotal, itsdensity, 0
  float64 dtcourant, dtforce, dtdrag, dtav, vsig2max, hhmax
  then OUT_FIELDS below (float64, N*width each), numneigh int32(N);
  mode 2 appends: int32 npg, int32 pg(3,npg), int32 nps, int32 ps(3,nps)
  (i, j_real, image code), 1-based particle indices;
  mode 3 appends: int32 nsteps, float64 time, dt, then vel(3,N) en(N)
  Bevol(3,N) alpha(3,N) after the last step.
"""
from __future__ import annotations

import os
import resource
import subprocess
import tempfile

import numpy as np

MAGIC = 1313100880
OUT_FIELDS = [("x", None), ("hh", 1), ("rho", 1), ("gradh", 1), ("gradgradh", 1),
              ("drhodt", 1), ("dhdt", 1), ("dens", 1), ("pr", 1), ("spsound", 1),
              ("uu", 1), ("psi", 1), ("force", 3), ("dudt", 1), ("dendt", 1),
              ("dBevoldt", 3), ("daldt", 3), ("dpsidt", 1), ("divB", 1),
              ("curlB", 3), ("gradpsi", 3), ("dustfrac", 1), ("rhogas", 1),
              ("rhodust", 1), ("ddustevoldt", 1), ("deltav", 3), ("graddivv", 3)]
SCALARS = ["dtcourant", "dtforce", "dtdrag", "dtav", "vsig2max", "hhmax"]


def write_input(path, opts, st, mode=1):
    n = st["pmass"].shape[0]
    with open(path, "wb") as f:
        np.array([MAGIC, mode, n, opts.ndim], dtype="<i4").tofile(f)
        opts.iopt().astype("<i4").tofile(f)
        opts.dopt().astype("<f8").tofile(f)
        for k in ("x", "vel", "pmass", "hh", "en", "Bevol", "alpha", "psi", "dustevol"):
            np.ascontiguousarray(st[k], dtype="<f8").tofile(f)
        np.ascontiguousarray(st["itype"], dtype="<i4").tofile(f)
        np.ascontiguousarray(st.get("ireal", np.full(n, -1)), dtype="<i4").tofile(f)


def read_output(path, ndim, mode=1):
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    npart, ntotal, its, _ = take("<i4", 4)
    out = dict(npart=int(npart), ntotal=int(ntotal), itsdensity=int(its))
    sc = take("<f8", 6)
    out.update({k: float(v) for k, v in zip(SCALARS, sc)})
    for name, w in OUT_FIELDS:
        w = ndim if w is None else w
        a = take("<f8", npart * w)
        out[name] = a.reshape(npart, w) if w > 1 else a
    out["numneigh"] = take("<i4", npart)
    if mode == 3:
        out["nsteps"] = int(take("<i4", 1)[0])
        out["time"], out["dt"] = (float(v) for v in take("<f8", 2))
        for name, w in (("vel", 3), ("en", 1), ("Bevol", 3), ("alpha", 3)):
            a = take("<f8", npart * w)
            out[name] = a.reshape(npart, w) if w > 1 else a
    if mode == 2:
        npg = int(take("<i4", 1)[0])
        out["pairs_gather"] = take("<i4", 3 * npg).reshape(npg, 3)
        nps = int(take("<i4", 1)[0])
        out["pairs_sym"] = take("<i4", 3 * nps).reshape(nps, 3)
    assert off == len(raw), "trailing bytes in reference output"
    return out


def _unlimited_stack():
    # the reference keeps large automatic arrays on the stack
    # (src/allocateND.f90:62-100, src/iterate_density.f90:63-66)
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_reference(binary, opts, st, mode=1, timeout=600):
    """Run the compiled reference harness on (opts, st); returns its outputs."""
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        fout = os.path.join(d, "out.bin")
        write_input(fin, opts, st, mode)
        p = subprocess.run([os.path.abspath(binary), fin, fout], cwd=d, capture_output=True,
                           text=True, timeout=timeout, preexec_fn=_unlimited_stack)
        if p.returncode != 0:
            raise RuntimeError("reference harness failed (%d):\n%s\n%s"
                               % (p.returncode, p.stdout[-4000:], p.stderr[-4000:]))
        out = read_output(fout, opts.ndim, mode)
        with open(fout + ".ev") as f:
            out["ev"] = f.read()
        with open(fout + ".in") as f:
            out["infile"] = f.read()
        with open(fout + ".in2") as f:
            out["infile2"] = f.read()
        out["log"] = p.stdout
        return out


def run_binding(binary, opts, st, nsteps, precision="fp64", timeout=600):
    """Run the binding harness (oracle/ref/bindtest.f90) for nsteps steps;
    returns the final state and the .ev text written by the reference's
    evwrite."""
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "in.bin")
        fout = os.path.join(d, "out.bin")
        write_input(fin, opts, st, 3)
        env = dict(os.environ, NDSP_PRECISION=precision)
        p = subprocess.run([os.path.abspath(binary), fin, fout, str(nsteps)], cwd=d, capture_output=True,
                           text=True, timeout=timeout, env=env, preexec_fn=_unlimited_stack)
        if p.returncode != 0:
            raise RuntimeError("binding harness failed (%d):\n%s\n%s"
                               % (p.returncode, p.stdout[-4000:], p.stderr[-4000:]))
        out = read_output(fout, opts.ndim, 3)
        with open(fout + ".ev") as f:
            out["ev"] = f.read()
        out["log"] = p.stdout
        return out
