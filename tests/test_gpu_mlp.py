"""h-predictor MLP kernels (BF16 weights, FP32 accumulation) against the FP64
numpy reference oracle/mlp_ref.py at the same BF16-rounded weights."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sim():
    import ndspmhd_amd as nd
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(6)
    return nd.Simulation(o, st)


def test_forward_and_gradient_match_oracle():
    from oracle import mlp_ref as m
    sim = _sim()
    p = m.init_params(3)
    rng = np.random.default_rng(4)
    p[m.NH * m.NIN:m.NH * m.NIN + m.NH] = rng.normal(0, 0.1, m.NH)  # nonzero b1
    X = rng.standard_normal((1000, 6)).astype(np.float32)
    T = rng.standard_normal(1000).astype(np.float32)
    sim.mlp_set_params(p)
    pb = m.bf16_round(p)
    y_ref, _ = m.forward(pb, X.astype(np.float64))
    y = sim.mlp_forward(X)
    assert np.max(np.abs(y - y_ref)) / np.max(np.abs(y_ref)) < 1e-5
    L_ref, g_ref = m.loss_and_grad(pb, X.astype(np.float64), T.astype(np.float64))
    L, g = sim.mlp_grad(X, T)
    print("loss", L, L_ref, "grad err", np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref)), flush=True)
    assert L == pytest.approx(L_ref, rel=1e-5)
    assert np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref)) < 1e-4
    # a planted mistake (one wrong gradient entry) must be visible at this tolerance
    g_bad = g_ref.copy()
    g_bad[200] *= 1.01
    assert np.max(np.abs(g_bad - g_ref)) / np.max(np.abs(g_ref)) > 1e-4 or abs(g_ref[200]) < 1e-3


def test_adam_steps_match_oracle():
    from oracle import mlp_ref as m
    sim = _sim()
    p = m.init_params(5)
    rng = np.random.default_rng(6)
    X = rng.standard_normal((512, 6)).astype(np.float32)
    T = (0.5 * X[:, 0] - 0.2 * X[:, 3] ** 2).astype(np.float32)
    sim.mlp_set_params(p)
    lr, iters = 1e-2, 20
    loss = sim.mlp_train(X, T, iters, lr)
    q, k = sim.mlp_get_params()
    assert k == iters
    # reference: FP64 Adam on the FP32 master weights, forward with BF16 weights
    pm = p.astype(np.float64)
    mm = np.zeros_like(pm)
    vv = np.zeros_like(pm)
    for it in range(1, iters + 1):
        L, g = m.loss_and_grad(m.bf16_round(pm.astype(np.float32)), X.astype(np.float64), T.astype(np.float64))
        pm, mm, vv = m.adam(pm, mm, vv, g, it, lr)
    err = np.max(np.abs(q - pm))
    print("adam: max |w_gpu - w_ref|", err, "loss", loss, L, flush=True)
    assert err < 2e-3  # BF16 forward rounding differs in a handful of low bits
    assert loss < m.loss_and_grad(m.bf16_round(p), X.astype(np.float64), T.astype(np.float64))[0]


def test_online_seed_runs_and_converges():
    """With the MLP seeding the h-rho iteration, every particle still converges
    to tolh (the density sum is exact at the returned h)."""
    import ndspmhd_amd as nd
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(16, seed=2, vel_seed=3)
    sim = nd.Simulation(o, st)
    sim.mlp_config(enable=True, train=True, lr=2e-3, batch=2048, iters=4, seed_after=2)
    sim.derivs()
    for _ in range(6):
        sim.step()
    # at the converged h: h = hfact (m/rho)^(1/3) within tolh
    h, rho, m = sim.get("hh"), sim.get("rho"), sim.get("pmass")
    # hh after the step is the corrector's value; recompute density at fixed h
    sim.derivs()
    h, rho = sim.get("hh"), sim.get("rho")
    rel = np.abs(h - o.hfact * (m / rho) ** (1 / 3)) / h
    assert rel.max() < 2 * o.tolh


def test_trained_seed_saves_iterations():
    """What the learned seed gains: the same turbulent MHD box stepped with
    the reference seed h_in + dt dh/dt alone and with the online-trained
    network seeding the h-rho iteration.  After a short training period the
    extra Newton-Raphson passes (mean passes per particle - 1) must fall to
    at most a third of those of the plain seed (measured on a 48^3 box over
    40 steps: 0.060 -> 0.0047), while every particle still converges."""
    import ndspmhd_amd as nd
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(32)
    extra = {}
    for mode in ("plain", "mlp"):
        sim = nd.Simulation(o, st)
        if mode == "mlp":
            sim.mlp_config(enable=True, train=True, lr=2e-3, batch=8192, iters=4, seed_after=3)
        sim.derivs()
        nc = []
        for _ in range(30):
            sim.step()
            nc.append(sim.stats["ncalc_mean"])
        extra[mode] = float(np.mean(nc[12:]) - 1.0)
        print(mode, "ncalc per step", ["%.4f" % x for x in nc], flush=True)
        if mode == "mlp":
            sim.derivs()
            h, rho, m = sim.get("hh"), sim.get("rho"), sim.get("pmass")
            rel = np.abs(h - o.hfact * (m / rho) ** (1 / 3)) / h
            assert rel.max() < 2 * o.tolh
    print("extra passes per particle: plain %.5f, with the trained seed %.5f" % (extra["plain"], extra["mlp"]),
          flush=True)
    assert extra["plain"] > 0.01
    assert extra["mlp"] <= extra["plain"] / 3.0, extra
