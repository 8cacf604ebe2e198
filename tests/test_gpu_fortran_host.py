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


def test_fortran_host_ev_and_dumps():
    """the Fortran host's .ev file and dumps (reference formats, written by
    Fortran I/O) equal what the Python writers make of the same steps, and a
    restart from the last dump takes up its state"""
    import ndspmhd_amd as nd
    from ndspmhd_amd import dumpfile as df
    from ndspmhd_amd import evfile as ev
    from ndspmhd_amd.setups import shock_tube_2d
    from refio import write_input
    o, st = shock_tube_2d(400, hfact=1.1)
    nsteps, nout = 6, 4
    with tempfile.TemporaryDirectory() as d:
        fin, fout, root = os.path.join(d, "in.bin"), os.path.join(d, "out.bin"), os.path.join(d, "run")
        write_input(fin, o, st)
        print("fortran host outputs: 2d shock tube 400, %d steps" % nsteps, flush=True)
        p = subprocess.run([EXE, fin, fout, str(nsteps), "1e30", "1", root, str(nout)],
                           capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        assert "Max energy error" in p.stdout
        dumps = sorted(f for f in os.listdir(d) if f.startswith("run_") and f.endswith(".dat"))
        assert dumps == ["run_00000.dat", "run_00001.dat", "run_00002.dat"], dumps
        evtext = open(root + ".ev").read()
        # the same run from Python
        sim = nd.Simulation(o, st, precision=1)
        sim.derivs()
        pyev = ev.EvFile(os.path.join(d, "py.ev"), o.imhd, o.idust)
        pyev.write_sim(sim, 0.0, o.dt)
        df.write_sim(os.path.join(d, "py_00000.dat"), sim, 0.0)
        t = 0.0
        for k in range(nsteps):
            info = sim.step()
            t = info["time"]
            if k + 1 == nout:
                df.write_sim(os.path.join(d, "py_00001.dat"), sim, t)
            pyev.write_sim(sim, t, info["dt"])
        df.write_sim(os.path.join(d, "py_00002.dat"), sim, t)
        pyev.close()
        names, rows_f = ev.parse(evtext)
        names_p, rows_p = ev.parse(open(os.path.join(d, "py.ev")).read())
        assert names == names_p and len(rows_f) == nsteps + 1
        np.testing.assert_allclose(np.array(rows_f), np.array(rows_p), rtol=1e-9, atol=0)
        assert evtext.splitlines()[0] == open(os.path.join(d, "py.ev")).read().splitlines()[0]
        for k in range(3):
            a = open(os.path.join(d, "run_%05d.dat" % k), "rb").read()
            b = open(os.path.join(d, "py_%05d.dat" % k), "rb").read()
            ra, rb = df.read_records(os.path.join(d, "run_%05d.dat" % k)), df.read_records(
                os.path.join(d, "py_%05d.dat" % k))
            assert [len(r) for r in ra] == [len(r) for r in rb]
            # header: time accumulates on the host in the Fortran program
            assert ra[0][8:] == rb[0][8:], k
            assert ra[1:] == rb[1:], ("dump", k, [i for i, (x, y) in enumerate(zip(ra, rb)) if x != y][:5])
        # restart from the last dump: the state is taken up exactly
        last = os.path.join(d, "run_00002.dat")
        fout2 = os.path.join(d, "out2.bin")
        p = subprocess.run([EXE, fin, fout2, "0", "1e30", "1", "", "0", last],
                           capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        ns, t2, dt2, out = _read_out(fout2, st["pmass"].shape[0], 2)
        dd = df.read_dump(last)
        assert ns == 0 and t2 == dd["t"]
        # fixed particles (the tube's x walls) take their partner's en, B and
        # alpha in that evaluation; every other particle keeps the dumped state
        fixed = st["itype"] == 1
        part = st["ireal"][fixed]
        for k in ("x", "vel", "en", "Bevol", "alpha"):
            a, b = out[k], dd["state"][k].reshape(out[k].shape)
            if k in ("x", "vel"):
                np.testing.assert_array_equal(a, b, err_msg=k)
            else:
                np.testing.assert_array_equal(a[~fixed], b[~fixed], err_msg=k)
                np.testing.assert_array_equal(a[fixed], a[part], err_msg=k)
        # derivs after the restart runs the h-rho iteration again from the
        # dumped h (which had converged to tolh); fixed particles take their
        # partner's
        h0 = dd["state"]["hh"]
        assert np.max(np.abs(out["hh"] - h0)[~fixed] / h0[~fixed]) <= 2 * o.tolh
        np.testing.assert_array_equal(out["hh"][fixed], out["hh"][part])
