"""Neighbour lists against the per-pass cell scan on the GPU.

Every pair taken from a list still passes the same exact FP64 tests where it
is used, so both paths sum the same pairs; only the order differs for the
transposed part of a list, so the states agree to round-off (norm-relative
1e-12 in FP64; 1e-6 mixed, where psi, driven by the cancelling divB sums,
picks up order differences of ~2e-8 over four steps), and neighbour counts
and iteration counts are identical.  Also
covered: an evaluation whose h grows beyond the list's skin (the list is then
abandoned, with the same results)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "vel", "Bevol", "en", "hh", "alpha", "psi", "dustevol", "rho",
          "force", "dBevoldt", "dudt", "ddustevoldt")


def _pair(o, st, prec, nsteps):
    import ndspmhd_amd as nd
    out = []
    for cellscan in (0, 1):
        sim = nd.Simulation(o.replace(cellscan=cellscan), st, precision=prec)
        sim.set_timing(True)
        sim.derivs()
        names = set(sim.kernel_times())
        counts = (sim.get("numneigh").copy(), sim.get("nits").copy())
        for _ in range(nsteps):
            sim.step()
        out.append((sim, names, counts))
    return out


def _close(a, b, tol):
    for k in FIELDS:
        x, y = a.get(k), b.get(k)
        err = np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300)
        assert err <= tol, (k, err)


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 1e-6)])
def test_mhd_box_lists_match_scan(prec, tol):
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(12, seed=3, vel_seed=4, vrms=0.3)
    # uneven h, so that the transposed part of the lists is not empty
    rng = np.random.default_rng(8)
    st["hh"] *= np.exp(0.1 * rng.standard_normal(st["hh"].shape))
    print("mhd box 12^3 lists vs cell scan prec", prec, flush=True)
    (s0, n0, c0), (s1, n1, c1) = _pair(o, st, prec, 4)
    assert "nblist" in n0 and "nblist" not in n1
    assert np.array_equal(c0[0], c1[0]) and np.array_equal(c0[1], c1[1])
    _close(s0, s1, tol)


@pytest.mark.parametrize("ndim,nside", [(2, 24), (1, 200)])
def test_dust_box_lists_match_scan(ndim, nside):
    from ndspmhd_amd.setups import dusty_box
    o, st = dusty_box(nside, ndim=ndim, seed=9)
    print("dust box lists vs cell scan ndim", ndim, flush=True)
    (s0, n0, c0), (s1, n1, c1) = _pair(o, st, 0, 3)
    assert "nblist" in n0
    assert np.array_equal(c0[0], c1[0])
    _close(s0, s1, 1e-6)


def test_list_abandoned_when_h_outgrows_skin():
    """h seeded 30% below its converged value: the iteration grows h beyond
    the 4% skin, the force pass must fall back to the cell scan."""
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(10, seed=5, vel_seed=6)
    st["hh"] *= 0.7
    import ndspmhd_amd as nd
    res = []
    for cellscan in (0, 1):
        sim = nd.Simulation(o.replace(cellscan=cellscan), st, precision=1)
        sim.derivs()
        res.append(sim)
    assert np.array_equal(res[0].get("numneigh"), res[1].get("numneigh"))
    _close(res[0], res[1], 1e-12)


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 1e-6)])
def test_lists_fall_back_where_they_overflow(prec, tol):
    """a few particles with h 2.5 times their neighbours' at fixed h: their
    groups' candidates exceed the staging capacity and their rows the list
    capacity, so those particles fall back to the cell scan; the results
    equal the scan's"""
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(16, seed=7, vel_seed=8, vrms=0.3)
    rng = np.random.default_rng(9)
    big = rng.choice(st["hh"].size, 40, replace=False)
    st["hh"][big] *= 2.5
    o = o.replace(maxdensits=0)
    print("mhd box 16^3 with 40 particles of 2.5 h, prec", prec, flush=True)
    import ndspmhd_amd as nd
    probe = nd.Simulation(o, st, precision=prec)
    probe.derivs()
    L = probe.get("nblist_len")
    nn = probe.get("numneigh")
    print("list fallbacks", int((L < 0).sum()), "of", L.size, flush=True)
    assert (L < 0).any() and (L >= 0).mean() > 0.5
    # a list holds at least the symmetric neighbours of its particle
    assert np.all(L[L >= 0] >= nn[L >= 0])
    (s0, n0, c0), (s1, n1, c1) = _pair(o, st, prec, 2)
    assert "nblist" in n0
    assert np.array_equal(c0[0], c1[0])
    _close(s0, s1, tol)


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 1e-6)])
def test_lists_holding_both_images_of_a_neighbour(prec, tol):
    """2D tube (fixed x walls) whose periodic y extent, 10 lattice rows, lies
    between one and two list radii of its largest h: a list then holds the
    images of a neighbour on both sides, and images of its own particle, as
    the reference's ghost copies interact; the results equal the scan's"""
    from ndspmhd_amd.setups import shock_tube_2d
    o, st = shock_tube_2d(800, hfact=1.2, width_rows=10)
    Ly = o.xmax[1] - o.xmin[1]
    rlist = 2.0 * st["hh"].max() * 1.04  # radkern hmax (1 + skin), cubic spline
    print("2D tube, Ly %.4g, list radius %.4g, prec %d" % (Ly, rlist, prec), flush=True)
    assert rlist * 1.001 < Ly < 2.0 * rlist
    import ndspmhd_amd as nd
    probe = nd.Simulation(o, st, precision=prec)
    probe.derivs()
    L = probe.get("nblist_len")
    # lists are built (a few rows at the contact, where h jumps, overflow
    # their capacity and scan cells)
    print("list fallbacks", int((L < 0).sum()), "of", L.size, flush=True)
    assert (L >= 0).mean() > 0.9
    (s0, n0, c0), (s1, n1, c1) = _pair(o, st, prec, 4)
    assert "nblist" in n0
    assert np.array_equal(c0[0], c1[0]) and np.array_equal(c0[1], c1[1])
    _close(s0, s1, tol)


def test_list_lengths_at_the_fp32_screening_band():
    """Each particle's h puts its list radius within 2e-6 (relative) of its
    distance to one neighbour, inside the band where the builder's FP32
    screening defers to the FP64 test: every list length (gather part plus
    transposed extras) equals the count of the exact FP64 membership test
    replayed on the host, image by image"""
    import itertools
    import ndspmhd_amd as nd
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(8, seed=21, vel_seed=22, vrms=0.3)
    o = o.replace(maxdensits=0)
    x = st["x"]
    n = x.shape[0]
    lim2 = 2.0 * 2.0 * 1.04 * 1.04 * (1.0 + 1.0e-12)  # radkern^2 (1+skin)^2 (1+1e-12)
    shifts = np.array(list(itertools.product((-1.0, 0.0, 1.0), repeat=3)))  # box side 1
    d = (x[:, None, None, :] - x[None, :, None, :]) + shifts[None, None, :, :]
    r2 = d[..., 0] * d[..., 0]
    r2 = r2 + d[..., 1] * d[..., 1]
    r2 = r2 + d[..., 2] * d[..., 2]
    rng = np.random.default_rng(23)
    h0 = st["hh"][0]
    h = np.empty(n)
    for k in range(n):
        # the neighbour whose distance is nearest the natural list radius
        cand = r2[k][r2[k] > 0.0]
        rr = cand[np.argmin(np.abs(cand - lim2 * h0 * h0))]
        h[k] = np.sqrt(rr / lim2) * (1.0 + rng.uniform(-2.0e-6, 2.0e-6))
    st["hh"][:] = h
    hi = 1.0 / h
    hi21 = hi * hi
    ink = r2 * hi21[:, None, None] < lim2
    inj = r2 * hi21[None, :, None] < lim2
    expected = np.sum(ink | inj, axis=(1, 2))
    near = np.sum(np.abs(r2 * hi21[:, None, None] / lim2 - 1.0) < 1.6e-5)
    print("pairs within the screening band:", int(near), flush=True)
    assert near >= n
    sim = nd.Simulation(o, st, precision=1)
    sim.derivs()
    L = sim.get("nblist_len")
    assert (L >= 0).all()
    assert np.array_equal(L.astype(np.int64), expected), np.nonzero(L != expected)
