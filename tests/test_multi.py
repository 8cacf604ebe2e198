"""Sweep descriptions (ndspmhd_amd.multi): multi_shock's .in/.shk files
(multi/multi_shock.f90) and the 1D shock set-up (src/setup_shock1D_mhd.f90)."""
import os

import numpy as np
import pytest

from ndspmhd_amd import infile as inf
from ndspmhd_amd import multi


def test_shock_sweep_files(tmp_path):
    names = multi.write_shock_sweep("shk", 9, directory=str(tmp_path))
    assert [os.path.basename(n) for n in names] == ["shk%d" % i for i in range(1, 10)]
    for i, n in enumerate(names, 1):
        v = inf.read_infile(n + ".in", strict=True).values
        t = multi.SHOCK_TESTS[min(i, 7) - 1]
        assert (v["psep"], v["tmax"], v["tout"]) == pytest.approx(t[1:2] + t[3:5], rel=1e-9)
        assert v["gamma"] == pytest.approx(t[2], rel=1e-8)
        assert multi.read_shk(n + ".shk") == tuple(t[7])
    v4 = inf.read_infile(names[3] + ".in").values
    v5 = inf.read_infile(names[4] + ".in").values
    assert v4["iener"] == 0 and v4["polyk"] == 1.0 and v5["iener"] == 2
    big = multi.write_shock_sweep("x", 12, directory=str(tmp_path))
    assert os.path.basename(big[-1]) == "x12"


def test_shk_list_directed_forms(tmp_path):
    p = tmp_path / "a.shk"
    p.write_text(" 1.0, 0.125\n 1.D0 0.1\n0 0\n 0. 0.\n0 0\n 0.75\n 1.0 -1.0\n 0 0\n")
    s = multi.read_shk(str(p))
    assert s[:4] == (1.0, 0.125, 1.0, 0.1) and s[10:13] == (0.75, 1.0, -1.0)


@pytest.mark.parametrize("i", range(1, 8))
def test_shock_setup(i, tmp_path):
    (n,) = multi.write_shock_sweep("s", 1, directory=str(tmp_path)) if i == 1 else \
        multi.write_shock_sweep("s", i, directory=str(tmp_path))[-1:]
    v = inf.read_infile(n + ".in").values
    shk = multi.read_shk(n + ".shk")
    o, st = multi.shock_setup_1d(v, shk)
    x = st["x"][:, 0]
    assert np.all(np.diff(x) > 0) and x[0] > -0.5 and x[-1] < 0.5
    m = st["pmass"]
    assert np.all(m == m[0]) and m[0] == pytest.approx(shk[1] * v["psep"])
    # far from the discontinuity x(i) - x(i-2) = 2 m/rho on each side (the
    # two interleaved rows keep the offset they had through the transition)
    dl, dr = x[2:20] - x[:18], x[-18:] - x[-20:-2]
    assert dl == pytest.approx(2 * m[0] / shk[0], rel=1e-9)
    assert dr == pytest.approx(2 * m[0] / shk[1], rel=1e-9)
    # total mass matches the states over the two halves
    assert m.sum() == pytest.approx(0.5 * (shk[0] + shk[1]), rel=0.02)
    assert o.ibound[0] == 1 and o.ndim == 1 and o.gamma == v["gamma"]
    # the first and last six particles are fixed, copying the seventh from
    # their end, with its h
    n = x.shape[0]
    assert np.all(st["itype"][:6] == 1) and np.all(st["itype"][-6:] == 1) and st["itype"][6:-6].sum() == 0
    assert np.all(st["ireal"][:6] == 6) and np.all(st["ireal"][-6:] == n - 7)
    assert np.all(st["hh"][:6] == st["hh"][6]) and np.all(st["hh"][-6:] == st["hh"][n - 7])
    assert np.all(st["vel"][:5] == shk[4:10:2]) and np.all(st["vel"][-5:] == shk[5:10:2])
