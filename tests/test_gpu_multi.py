"""The sweep launcher (ndspmhd_amd.multi) on the GPU: multi_shock's runs
ingested from their .in/.shk files and stepped as one batch give the same
states, .ev files and dumps as the same runs stepped one after another."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sweep(d, batched, nsteps):
    from ndspmhd_amd import infile as inf
    from ndspmhd_amd import multi
    cwd = os.getcwd()
    os.chdir(d)
    try:
        # the base options of the sweep (multirun.in): MHD with psi cleaning
        base = dict(inf.DEFAULTS, imhd=11, idivbzero=2)
        names = multi.write_shock_sweep("shk", 7, base)
        runs = multi.run_sweep(names, precision=1, batched=batched, max_steps=nsteps,
                               log=lambda s: print(s, flush=True))
        out = {os.path.basename(r.root): {k: r.sim.get(k) for k in ("x", "vel", "Bevol", "hh", "en")}
               for r in runs}
        steps = {os.path.basename(r.root): (r.nsteps, r.time) for r in runs}
        for r in runs:
            multi.finish(r)
    finally:
        os.chdir(cwd)
    return out, steps


def test_sweep_batched_equals_sequential(tmp_path):
    from ndspmhd_amd import dumpfile, evfile
    a, b = tmp_path / "batched", tmp_path / "sequential"
    a.mkdir()
    b.mkdir()
    print("multi_shock sweep of 7, 12 steps, batched", flush=True)
    sa, na = _sweep(str(a), True, 12)
    print("multi_shock sweep of 7, 12 steps, one by one", flush=True)
    sb, nb = _sweep(str(b), False, 12)
    assert na == nb
    for run in sa:
        for k in sa[run]:
            np.testing.assert_array_equal(sa[run][k], sb[run][k], err_msg=(run, k))
        ta = (a / (run + ".ev")).read_text()
        assert ta == (b / (run + ".ev")).read_text(), run
        names, rows = evfile.parse(ta)
        assert tuple(names) == evfile.MHD and len(rows) == na[run][0] + 1
        dumps = sorted(p.name for p in a.iterdir() if p.name.startswith(run + "_"))
        assert dumps and dumps[0] == run + "_00000.dat"
        for dname in dumps:
            assert (a / dname).read_bytes() == (b / dname).read_bytes(), dname
        last = dumpfile.read_dump(str(a / dumps[-1]))
        assert last["npart"] == sa[run]["x"].shape[0]
