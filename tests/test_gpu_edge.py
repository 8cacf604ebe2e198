"""Edge cases of the HIP path through the C-ABI: a periodic box so small
that the search radius reaches half of it (no neighbour lists; particles see
their periodic images, as the reference's ghosts), a particle count that
leaves partial wavefronts, and the error the reference stops on (two
particles at one position, src/ratesND_mhd.f90:384-395) reported cleanly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "hh", "gradh", "drhodt", "pr", "force", "dBevoldt", "dudt", "divB")


def _compare(sim, orc, tol):
    for f in FIELDS:
        a, b = sim.get(f), orc.get(f)
        err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
        print(f, err, flush=True)
        assert err <= tol, (f, err)


@pytest.mark.parametrize("nside", [5, 7])
def test_small_periodic_box_fp64(nside):
    """5^3: radkern*h is about half the box (lists off, candidate walk over
    image cells); 7^3 = 343 particles, 5 full waves and a partial one"""
    import ndspmhd_amd as nd
    import oracle
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(nside, seed=5, vel_seed=6)
    print("small box", nside, flush=True)
    sim = nd.Simulation(o, st, precision=1)
    orc = oracle.Oracle(o, st)
    sim.derivs()
    orc.derivs()
    _compare(sim, orc, 1e-10)
    np.testing.assert_array_equal(sim.get("numneigh").astype(np.int32), orc.get("numneigh").astype(np.int32))


def test_coincident_particles_reported():
    """two particles at one position: the rates stop with the coincident-pair
    error on both sides, nothing faults"""
    import ndspmhd_amd as nd
    import oracle
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(8, seed=5, vel_seed=6)
    st = dict(st)
    st["x"] = st["x"].copy()
    st["x"][1] = st["x"][0]
    sim = nd.Simulation(o, st, precision=0)
    with pytest.raises(nd.NdspError) as e:
        sim.derivs()
    assert e.value.code == 8
    orc = oracle.Oracle(o, st)
    with pytest.raises(Exception):
        orc.derivs()


@pytest.mark.parametrize("hscale", [1.8, 2.6, 0.4])
def test_crowded_cells_fp64(hscale):
    """one particle with a large h makes every cell (side radkern*hmax, one
    per search radius) hold about 200 (1.8) or all 1728 (2.6) particles: the
    in-cell order is then built by the four-per-lane ranking or the
    single-lane fallback.  One with a small h (0.4) sits in the lists of
    about 60 particles that are not in its own, more transposed extras than
    the in-register sort takes.  Results as the oracle's"""
    import ndspmhd_amd as nd
    import oracle
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(12, seed=7, vel_seed=8)
    st = dict(st)
    st["hh"] = st["hh"].copy()
    st["hh"][5] *= hscale
    print("crowded cells, h scale", hscale, flush=True)
    sim = nd.Simulation(o, st, precision=1)
    sim.set_tile_shapes(256, 256, 1)
    orc = oracle.Oracle(o, st)
    sim.derivs()
    orc.derivs()
    _compare(sim, orc, 1e-10)
    np.testing.assert_array_equal(sim.get("numneigh").astype(np.int32), orc.get("numneigh").astype(np.int32))
