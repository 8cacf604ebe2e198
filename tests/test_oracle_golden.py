"""The CPU oracle against golden vectors from the compiled reference
(3D periodic MHD and one-fluid dust boxes; tests/golden/make_golden.py)."""
import numpy as np
import pytest

import golden_util as gu

FIELDS = ["hh", "rho", "gradh", "gradgradh", "drhodt", "dhdt", "pr", "spsound", "psi",
          "force", "dudt", "dendt", "dBevoldt", "daldt", "dpsidt", "divB", "curlB",
          "gradpsi", "dustfrac", "rhogas", "rhodust", "ddustevoldt", "deltav"]
# the oracle restates the reference's arithmetic in the same order; what is
# left is libm pow (quintic table) and norm2 rounding
TOL = 1.0e-12


@pytest.mark.parametrize("name", gu.names())
def test_oracle_matches_reference(name, oracle_mod):
    o, st, ref = gu.load(name)
    orc = oracle_mod.Oracle(o, st)
    orc.derivs()
    assert int(orc.scalar("itsdensity")) == int(ref["itsdensity"])
    assert int(orc.scalar("ntotal")) == int(ref["ntotal"])
    # graddivv (the "curl v" sum of the switch block) is compared where the
    # dumps carry it, in runs without MHD
    for k in FIELDS + (["graddivv"] if o.imhd == 0 else []):
        assert gu.relerr(orc.get(k), ref[k]) <= TOL, k
    np.testing.assert_array_equal(orc.get("numneigh").astype(np.int32), ref["numneigh"])
    for k in ("dtcourant", "dtforce", "dtdrag", "vsig2max", "hhmax"):
        assert orc.scalar(k) == pytest.approx(float(ref[k]), rel=1e-13, abs=0), k


@pytest.mark.parametrize("name", gu.names())
def test_oracle_neighbour_sets_bit_exact(name, oracle_mod):
    o, st, ref = gu.load(name)
    orc = oracle_mod.Oracle(o, st)
    orc.derivs()
    assert gu.pairset(orc.neighbours(False)) == gu.pairset(ref["pairs_gather"])
    assert gu.pairset(orc.neighbours(True)) == gu.pairset(ref["pairs_sym"])


def test_planted_mistake_is_caught(oracle_mod):
    """A wrong sign in one force term must fail the golden comparison: plant it
    in the oracle's output (not in the GPU) and check the tolerance bites."""
    o, st, ref = gu.load("mhd3d_nr")
    orc = oracle_mod.Oracle(o, st)
    orc.derivs()
    f = orc.get("force")
    f[:, 1] = -f[:, 1]
    assert gu.relerr(f, ref["force"]) > 1e-3
