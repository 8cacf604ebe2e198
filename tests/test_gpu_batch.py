"""Batched sweep on the GPU: independent 1D/2D MHD shock tubes of uneven size
stepped together (ndsp_batch_step, two recorded graphs for all members)
against the same simulations stepped one by one, which must agree bit for
bit, and one member against the CPU oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "vel", "Bevol", "en", "hh", "alpha", "psi", "rho", "force",
          "dBevoldt", "dudt", "numneigh", "nits", "slot")


def _sweep(nsim, seed=11):
    from ndspmhd_amd.setups import shock_sweep
    return shock_sweep(nsim=nsim, seed=seed, nmin=200, nmax=3000)


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_matches_individual_steps(prec):
    import ndspmhd_amd as nd
    sims_cfg = _sweep(6)
    print("batch of", len(sims_cfg), "shock tubes, N =",
          [s[1]["x"].shape[0] for s in sims_cfg], "ndim =", [s[0].ndim for s in sims_cfg], flush=True)
    members = [nd.Simulation(o, st, precision=prec) for o, st in sims_cfg]
    alone = [nd.Simulation(o, st, precision=prec) for o, st in sims_cfg]
    batch = nd.Batch(members)
    for step in range(12):
        batch.step()
        for k, sim in enumerate(alone):
            sim.step()
            m = members[k]
            for f in FIELDS:
                a, b = m.get(f), sim.get(f)
                assert np.array_equal(a, b), ("step", step, "member", k, f, np.max(np.abs(a - b)))
            assert m.dt_info["dt"] == sim.dt_info["dt"]
            assert m.stats == sim.stats
    batch.close()


def test_batch_member_against_oracle():
    import ndspmhd_amd as nd
    import oracle
    from ndspmhd_amd.setups import brio_wu_1d, shock_tube_2d
    cfgs = [brio_wu_1d(300, hfact=1.3, gamma=5.0 / 3.0), shock_tube_2d(500, hfact=1.1)]
    members = [nd.Simulation(o, st, precision=1) for o, st in cfgs]
    orcs = [oracle.Oracle(o, st) for o, st in cfgs]
    for orc in orcs:
        orc.derivs()
    batch = nd.Batch(members)
    for step in range(4):
        for m, orc in zip(members, orcs):
            m.set_dt(orc.scalar("dt"), orc.scalar("time"))
            orc.step()
        batch.step()
    bad = []
    for m, orc in zip(members, orcs):
        for f in ("x", "vel", "Bevol", "en", "hh", "rho", "force"):
            a, b = m.get(f), orc.get(f)
            err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
            print(m.opts.ndim, f, err, flush=True)
            if err > (1e-9 if f != "x" else 1e-11):
                bad.append((m.opts.ndim, f, err))
    assert not bad, bad
    batch.close()


def test_autotune_keeps_state_and_parity():
    """the autotuner restores the evolved state bit for bit; the tuned
    simulation still matches the oracle"""
    import ndspmhd_amd as nd
    import oracle
    from ndspmhd_amd.setups import brio_wu_1d
    o, st = brio_wu_1d(600)
    sim = nd.Simulation(o, st, precision=1)
    orc = oracle.Oracle(o, st)
    sim.derivs()
    orc.derivs()
    before = {f: sim.get(f).copy() for f in FIELDS}
    choice = sim.autotune(reps=1)
    print("tile shapes", choice, flush=True)
    assert choice["force_workgroup"] in (64, 256) and choice["density_workgroup"] in (64, 256)
    assert choice["cells_per_radius"] in (1, 2)
    for f in FIELDS:
        assert np.array_equal(before[f], sim.get(f)), f
    for _ in range(5):
        sim.set_dt(orc.scalar("dt"), orc.scalar("time"))
        orc.step()
        sim.step()
    for f in ("x", "vel", "Bevol", "en", "hh", "rho", "force"):
        a, b = sim.get(f), orc.get(f)
        err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
        assert err <= (1e-9 if f != "x" else 1e-11), (f, err)


@pytest.mark.parametrize("shapes", [(64, 64, 1), (256, 256, 2)])
def test_set_tile_shapes_parity_with_graphs(shapes):
    """given tile shapes (ndsp_set_tile_shapes) change launch shapes only:
    graph-replayed steps after the change still match the oracle"""
    import ndspmhd_amd as nd
    import oracle
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(10, seed=3, vel_seed=4)
    sim = nd.Simulation(o, st, precision=1)
    orc = oracle.Oracle(o, st)
    sim.set_graph(True)
    sim.derivs()
    orc.derivs()
    for s in range(4):
        if s == 2:
            print("tile shapes", shapes, flush=True)
            sim.set_tile_shapes(*shapes)
        sim.set_dt(orc.scalar("dt"), orc.scalar("time"))
        orc.step()
        sim.step()
    for f in ("x", "vel", "Bevol", "en", "hh", "rho", "force"):
        a, b = sim.get(f), orc.get(f)
        err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
        assert err <= (1e-9 if f != "x" else 1e-11), (f, err)
    with pytest.raises(Exception):
        sim.set_tile_shapes(128, 64, 1)


def test_batch_autotune_keeps_state_and_parity():
    """the batch tuner restores every member bit for bit (layout included);
    the tuned batch still matches the oracle"""
    import ndspmhd_amd as nd
    import oracle
    from ndspmhd_amd.setups import brio_wu_1d, shock_tube_2d
    cfgs = [brio_wu_1d(300, hfact=1.3, gamma=5.0 / 3.0), shock_tube_2d(500, hfact=1.1)]
    members = [nd.Simulation(o, st, precision=1) for o, st in cfgs]
    orcs = [oracle.Oracle(o, st) for o, st in cfgs]
    for m, orc in zip(members, orcs):
        m.derivs()
        orc.derivs()
    batch = nd.Batch(members)
    before = [{f: m.get(f).copy() for f in FIELDS} for m in members]
    choice = batch.autotune(reps=1)
    print("batch tile shapes", choice, flush=True)
    assert choice["workgroup"] in (64, 256) and choice["cells_per_radius"] in (1, 2)
    for m, b in zip(members, before):
        for f in FIELDS:
            assert np.array_equal(b[f], m.get(f)), f
    for step in range(4):
        for m, orc in zip(members, orcs):
            m.set_dt(orc.scalar("dt"), orc.scalar("time"))
            orc.step()
        batch.step()
    for m, orc in zip(members, orcs):
        for f in ("x", "vel", "Bevol", "en", "hh", "rho", "force"):
            a, b = m.get(f), orc.get(f)
            err = np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
            assert err <= (1e-9 if f != "x" else 1e-11), (m.opts.ndim, f, err)
    batch.close()
