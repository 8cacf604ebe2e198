"""ndsp_mlp_forward_noalloc (the MLP inference that allocates nothing; its
staging is a fixed device array) against ndsp_mlp_forward through the
C-ABI: bit-identical outputs on two calls in a row, the second smaller than
the first (the array keeps the first call's rows beyond the second's), and a
clear error for a batch above the array's 65536 rows and for a context
without weights."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def test_noalloc_forward_matches_forward():
    import ndspmhd_amd as nd
    from ndspmhd_amd.setups import periodic_mhd_box
    from oracle import mlp_ref as m
    L = nd.load_library()
    o, st = periodic_mhd_box(6)
    sim = nd.Simulation(o, st)
    # weights never set: rejected, nothing run
    X0 = np.zeros((4, 6), dtype=np.float32)
    y0 = np.zeros(4, dtype=np.float32)
    assert L.ndsp_mlp_forward_noalloc(sim.h, _fp(X0), 4, _fp(y0)) != 0
    assert b"weights" in L.ndsp_last_error()
    p = m.init_params(9)
    rng = np.random.default_rng(10)
    p[m.NH * m.NIN:m.NH * m.NIN + m.NH] = rng.normal(0, 0.1, m.NH)
    sim.mlp_set_params(p)
    for B in (50000, 1237):
        X = rng.standard_normal((B, 6)).astype(np.float32)
        y_ref = sim.mlp_forward(X)
        y = np.full(B, np.nan, dtype=np.float32)
        print("noalloc forward, B =", B, flush=True)
        rc = L.ndsp_mlp_forward_noalloc(sim.h, _fp(X), B, _fp(y))
        assert rc == 0, L.ndsp_last_error()
        np.testing.assert_array_equal(y, y_ref)
        # and against the FP64 checker at the same BF16 weights
        yc, _ = m.forward(m.bf16_round(p), X.astype(np.float64))
        assert np.max(np.abs(y - yc)) / np.max(np.abs(yc)) < 1e-5
    big = np.zeros((65537, 6), dtype=np.float32)
    yb = np.zeros(65537, dtype=np.float32)
    assert L.ndsp_mlp_forward_noalloc(sim.h, _fp(big), 65537, _fp(yb)) != 0
    assert b"65536" in L.ndsp_last_error()
    sim.close()
