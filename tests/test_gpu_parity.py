"""HIP path against the compiled reference's golden vectors and the CPU
oracle, through the C-ABI (libndsp_hip.so)."""
import numpy as np
import pytest

import golden_util as gu

pytestmark = pytest.mark.gpu

DENS = ["rho", "gradh", "drhodt", "dhdt", "pr", "spsound"]
RATES = ["force", "dudt", "dBevoldt", "daldt", "dpsidt", "divB", "curlB", "gradpsi",
         "ddustevoldt", "deltav"]
# tolerances: norm-relative L-infinity error against the FP64 reference
TOL_MIXED = {"rho": 2e-6, "gradh": 2e-6, "drhodt": 1e-5, "dhdt": 1e-5, "pr": 2e-6, "spsound": 2e-6,
             "hh": 1e-12, "force": 1e-5, "dudt": 1e-5, "dBevoldt": 1e-5, "daldt": 1e-5,
             "dpsidt": 1e-5, "divB": 1e-5, "curlB": 1e-5, "gradpsi": 1e-5, "ddustevoldt": 1e-5,
             "deltav": 1e-5, "gradgradh": 1e-5, "graddivv": 1e-5}
TOL_FP64 = 1e-11


def _rates(o):
    """the rates a case checks: graddivv (written to dumps) only for runs
    without MHD, where the GPU forms it"""
    return RATES + (["graddivv"] if o.imhd == 0 else [])


def _sim(name, prec):
    import ndspmhd_amd as nd
    o, st, ref = gu.load(name)
    print("case", name, "prec", prec, "N", st["pmass"].shape[0], flush=True)
    sim = nd.Simulation(o, st, precision=prec)
    sim.derivs()
    return o, st, ref, sim


@pytest.mark.parametrize("name", [n for n in gu.names() if n.endswith("fixedh")])
@pytest.mark.parametrize("prec", [0, 1])
def test_fixed_h_fields(name, prec):
    o, st, ref, sim = _sim(name, prec)
    for k in ["hh"] + DENS + ["gradgradh"] + _rates(o):
        tol = TOL_FP64 if prec == 1 else TOL_MIXED[k]
        err = gu.relerr(sim.get(k), ref[k])
        print(k, err, flush=True)
        assert err <= tol, (k, err)
    np.testing.assert_array_equal(sim.get("numneigh").astype(np.int32), ref["numneigh"])
    d = sim.dt_info
    rt = 1e-12 if prec == 1 else 2e-6
    assert d["dtcourant"] == pytest.approx(float(ref["dtcourant"]), rel=rt)
    assert d["dtforce"] == pytest.approx(float(ref["dtforce"]), rel=1e-5 if prec == 0 else 1e-11)
    assert d["vsig2max"] == pytest.approx(float(ref["vsig2max"]), rel=rt * 2)


@pytest.mark.parametrize("name", [n for n in gu.names() if n.endswith("fixedh")])
def test_neighbour_sets_bit_exact(name):
    o, st, ref, sim = _sim(name, 0)
    assert gu.pairset(sim.neighbours(False)) == gu.pairset(ref["pairs_gather"])
    assert gu.pairset(sim.neighbours(True)) == gu.pairset(ref["pairs_sym"])


@pytest.mark.parametrize("name", [n for n in gu.names() if not n.endswith("fixedh")])
def test_iterated_density_fp64(name):
    """FP64 pair sums follow the reference's h-rho iteration pass for pass,
    including which passes are symmetric (numneigh, psi)."""
    o, st, ref, sim = _sim(name, 1)
    assert sim.stats["itsdensity"] == int(ref["itsdensity"])
    np.testing.assert_array_equal(sim.get("numneigh").astype(np.int32), ref["numneigh"])
    for k in ["hh", "psi"] + DENS + _rates(o):
        err = gu.relerr(sim.get(k), ref[k])
        assert err <= 1e-10, (k, err)


@pytest.mark.parametrize("name", [n for n in gu.names() if not n.endswith("fixedh")])
def test_iterated_density_mixed(name):
    """FP32 pair sums: every particle converges to the same tolerance.  A
    convergence decision (|h_new - h|/h_in < tolh, src/iterate_density.f90:
    242-243) can flip only where FP32 sums move that ratio across tolh, so h
    agrees with the reference's to 2 tolh everywhere.  Every particle's
    fields then meet the mixed tolerances against the FP64 oracle evaluated
    at the h the GPU iteration stopped at (sums at a given h do not depend on
    the path to it; psi as the GPU's partial passes left it).  At tolh 1e-6
    the flips are rare enough that every particle also meets them against
    the golden vectors directly."""
    import oracle
    o, st, ref, sim = _sim(name, 0)
    h = sim.get("hh")
    dh = np.abs(h - ref["hh"]) / ref["hh"]
    print("max |dh|/h vs golden %.3e, particles off the golden path (>1e-9): %d of %d"
          % (dh.max(), int(np.sum(dh > 1e-9)), dh.size), flush=True)
    assert dh.max() <= 2 * o.tolh
    at = dict(st)
    at["hh"] = h.copy()
    at["psi"] = sim.get("psi")
    orc = oracle.Oracle(o.replace(maxdensits=0), at)
    orc.derivs()
    bad = []
    for k in DENS + ["gradgradh"] + _rates(o):
        err = gu.relerr(sim.get(k), orc.get(k))
        print(k, "vs oracle at the GPU's h", err, flush=True)
        if err > TOL_MIXED[k]:
            bad.append((k, err))
    assert not bad, bad
    if o.tolh <= 1e-6:
        for k in DENS + _rates(o):
            err = gu.relerr(sim.get(k), ref[k])
            print(k, "vs golden", err, flush=True)
            assert err <= TOL_MIXED[k], (k, err)
