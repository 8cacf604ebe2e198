"""The Fortran host (ndspmhd_amd/fortran: iso_c_binding over the C-ABI, the
reference's argument-less derivs/step) runs the 1D Brio-Wu shock tube on the
GPU and agrees with the same steps driven from Python and with the CPU oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ndspmhd_amd", "fortran", "ndsp_run")


def _read_out(path, n, ndim):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw, dtype="<i4", count=4)
    t, dt = np.frombuffer(raw, dtype="<f8", count=2, offset=16)
    a = np.frombuffer(raw, dtype="<f8", offset=32)
    sizes = [("x", ndim), ("vel", 3), ("Bevol", 3), ("hh", 1), ("en", 1), ("alpha", 3),
             ("psi", 1), ("dustevol", 1), ("rho", 1)]
    out, off = {}, 0
    for k, w in sizes:
        v = a[off: off + n * w]
        off += n * w
        out[k] = v.reshape(n, w) if w > 1 else v
    assert off == a.size
    return int(hdr[2]), float(t), float(dt), out


def test_fortran_host_brio_wu():
    import ndspmhd_amd as nd
    import oracle
    from ndspmhd_amd.setups import brio_wu_1d
    from refio import write_input
    o, st = brio_wu_1d(400)
    nsteps = 10
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        write_input(fin, o, st)
        print("fortran host: brio-wu 400, %d steps" % nsteps, flush=True)
        p = subprocess.run([EXE, fin, fout, str(nsteps), "1e30", "1"], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        ns, t, dt, out = _read_out(fout, st["pmass"].shape[0], 1)
    assert ns == nsteps
    ev = [l.split() for l in p.stdout.splitlines() if l.startswith(" ev")]
    assert len(ev) == nsteps + 1
    # the same steps from Python through the same C-ABI (FP64 path)
    sim = nd.Simulation(o, st, precision=1)
    sim.derivs()
    for _ in range(nsteps):
        sim.step()
    for k in ("x", "vel", "Bevol", "hh", "en"):
        np.testing.assert_array_equal(out[k], sim.get(k).reshape(out[k].shape), err_msg=k)
    # and the CPU oracle
    orc = oracle.Oracle(o, st)
    orc.derivs()
    for _ in range(nsteps):
        orc.step()
    for k in ("x", "vel", "Bevol", "hh", "en"):
        b = orc.get(k).reshape(out[k].shape)
        err = np.max(np.abs(out[k] - b)) / np.max(np.abs(b))
        assert err < 1e-9, (k, err)
    assert t == pytest.approx(orc.scalar("time"), rel=1e-12)
