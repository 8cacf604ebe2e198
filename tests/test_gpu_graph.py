"""Step graphs against the launch-by-launch path on the GPU.

ndsp_step with graphs on records its two halves once and replays them; the
kernels and their order are the same, so every replayed step must give
bit-identical states.  Each configuration replays several steps (with the
library's own eager work -- downloads, reorders of the state every ten
steps, host-side completion of the h-rho iteration -- between replays) and
every step is checked against the eager run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("x", "vel", "Bevol", "en", "hh", "alpha", "psi", "dustevol", "rho",
          "force", "dBevoldt", "dudt", "numneigh", "nits", "slot")


def _run_pair(o, st, prec, nsteps):
    import ndspmhd_amd as nd
    eager = nd.Simulation(o, st, precision=prec)
    graph = nd.Simulation(o, st, precision=prec)
    graph.set_graph(True)
    for s in range(nsteps):
        eager.step()
        graph.step()
        for k in FIELDS:
            a, b = eager.get(k), graph.get(k)
            assert np.array_equal(a, b), ("step", s, k, np.max(np.abs(a - b)))
        assert eager.stats == graph.stats, s
        assert eager.dt_info["dt"] == graph.dt_info["dt"], s
    return eager, graph


@pytest.mark.parametrize("prec", [0, 1])
def test_mhd_box_graph_replay(prec):
    """23 steps: the reorders at steps 0, 10 and 20 move the state between the
    two field sets, so the third replays the graphs recorded after the first"""
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(12, seed=3, vel_seed=4, vrms=0.5)
    print("mhd box 12^3 graph replay prec", prec, flush=True)
    _run_pair(o, st, prec, 23)


def test_dust_box_graph_replay():
    from ndspmhd_amd.setups import dusty_box
    o, st = dusty_box(24, ndim=2, seed=9)
    print("dust box 24^2 graph replay", flush=True)
    _run_pair(o, st, 0, 6)


def test_brio_wu_graph_replay():
    """reflecting walls: synchronous h-rho passes, completed on the host path
    after the recorded chunk when needed"""
    from ndspmhd_amd.setups import brio_wu_1d
    o, st = brio_wu_1d(400)
    print("brio-wu 1d graph replay", flush=True)
    _run_pair(o, st, 1, 15)
