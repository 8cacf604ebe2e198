"""Fixed boundary particles (ibound 1) in the oracle, beyond what the golden
fixed-wall case pins (the reference's derivs): the leapfrog step moves a
fixed particle with its partner's starting velocity and force
(src/stepND_leapfrog_mhd.f90:110-186), so over a step it is displaced
exactly as its partner is, while its own velocity never changes."""
import numpy as np

import golden_util as gu


def test_fixed_particles_follow_their_partners(oracle_mod):
    from ndspmhd_amd.setups import brio_wu_1d
    o, st = brio_wu_1d(200, boundary="fixed")
    fixed = st["itype"] == 1
    part = st["ireal"][fixed]
    assert fixed.sum() == 12
    orc = oracle_mod.Oracle(o, st)
    orc.derivs()
    for _ in range(6):
        x0 = orc.get("x").copy()
        orc.step()
        dx = orc.get("x") - x0
        np.testing.assert_allclose(dx[fixed], dx[part], rtol=0, atol=4e-16)
        np.testing.assert_array_equal(orc.get("vel")[fixed], st["vel"][fixed])
        assert np.all(orc.get("force")[fixed] == 0.0)
    # an evaluation makes them replicas of their partners (the corrector then
    # restores their held starting values, so this holds right after derivs)
    orc.derivs()
    for k in ("rho", "hh", "en", "pr"):
        np.testing.assert_array_equal(orc.get(k)[fixed], orc.get(k)[part], err_msg=k)
    # the walls stay: the fixed particles are still the outermost ones
    x = orc.get("x").reshape(-1)
    inner = x[~fixed]
    assert x[fixed].min() < inner.min() and x[fixed].max() > inner.max()


def test_planted_gas_treatment_differs(oracle_mod):
    """the fixed-wall golden case tells fixed particles from gas: run as gas,
    the oracle misses the reference by O(1)"""
    o, st, ref = gu.load("mhd3d_fixedwall")
    st = dict(st)
    st["itype"] = np.zeros_like(st["itype"])
    st.pop("ireal")
    orc = oracle_mod.Oracle(o, st)
    orc.derivs()
    assert gu.relerr(orc.get("force"), ref["force"]) > 0.1


def test_fixed_near_walls_partners():
    """the 2D tube's fixed particles lie within 1.1 radkern hfact psep of an x
    wall, and each copies its closest gas particle, the lowest index among
    equally close ones (src/set_fixedbound.f90:87-106, by brute force here)"""
    from ndspmhd_amd.setups import shock_tube_2d
    o, st = shock_tube_2d(600, hfact=1.2)
    assert o.ibound[0] == 1
    x = st["x"]
    fixed = st["itype"] == 1
    assert 0 < fixed.sum() < x.shape[0] // 2
    fx, gx = x[fixed, 0], x[~fixed, 0]
    assert np.all((fx < gx.min()) | (fx > gx.max()))
    gas = np.flatnonzero(~fixed)
    for i in np.flatnonzero(fixed):
        best, bj = np.inf, -1
        for j in gas:
            rr = (x[i, 0] - x[j, 0]) ** 2 + (x[i, 1] - x[j, 1]) ** 2
            if rr < best:
                best, bj = rr, j
        assert st["ireal"][i] == bj
    assert np.all(st["ireal"][~fixed] == -1)


def test_fixed_wall_tube_2d_steps(oracle_mod):
    """2D tube with fixed x walls and periodic y in the oracle: fixed
    particles keep their velocity and carry zero rates, and after an
    evaluation their state is their partner's"""
    from ndspmhd_amd.setups import shock_tube_2d
    o, st = shock_tube_2d(300, hfact=1.2)
    fixed = st["itype"] == 1
    part = st["ireal"][fixed]
    orc = oracle_mod.Oracle(o, st)
    orc.derivs()
    for _ in range(3):
        orc.step()
    np.testing.assert_array_equal(orc.get("vel")[fixed], st["vel"][fixed])
    assert np.all(orc.get("force")[fixed] == 0.0)
    orc.derivs()
    for k in ("rho", "hh", "en", "pr", "alpha"):
        np.testing.assert_array_equal(orc.get(k)[fixed], orc.get(k)[part], err_msg=k)
