#!/usr/bin/env python3
"""Throughput of the ndspmhd hot path on one MI355X.

Workload: 3D periodic unit box, nside^3 particles (default 128^3) on a cubic
lattice with uniform random offsets of up to 0.1 spacing, rho = 1, Gaussian
random velocity field P(k) ~ k^-4 at rms Mach 1, uniform B = (0.1, 0, 0),
gamma = 5/3, cubic spline, hfact = 1.2 (SPMHD with hyperbolic divB cleaning,
artificial viscosity/conductivity/resistivity with switches).  One step =
one leapfrog step (predictor, derivs = binning + h-rho iteration + fused
force sums, corrector, new dt), all on the device with the state resident in
HBM.  Synthetic data.

Metric: pair interactions/s = (density pairs summed over h-rho passes +
unordered force pairs) / wall time, plus steps/s; the same work in plain FP64
on the GPU and the compiled reference (serial FP64 CPU) are timed beside it.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic compulsory bytes per particle and launch, reference layout
# (8 B per real, 4 B per integer, each field read once and each output
# written once; neighbour re-reads, lists and sort scratch not counted)
BYTES_BIN = 28           # binning (cell lists): x in, permutation and cell ranges out
BYTES_DENSITY = 108      # per active particle and h-rho pass
BYTES_FORCE = 280        # 20 reals in, 15 reals out
# algorithmic FP32 flops per ordered pair evaluation (estimates from the
# operation counts of the pair sums: add, mul = 1, fma = 2)
FLOPS_DENSITY_PAIR = 40
FLOPS_FORCE_PAIR = 200


def log(*a):
    print(*a, file=sys.stderr, flush=True)


_SETUPS = {}


def setup(nside):
    from ndspmhd_amd.setups import periodic_mhd_box
    if nside not in _SETUPS:
        _SETUPS.clear()
        _SETUPS[nside] = periodic_mhd_box(nside)
    return _SETUPS[nside]


SHAPES = "auto"  # tile shapes: "auto" (autotune) or (force_wg, density_wg, cells_per_radius)


def run_gpu(nside, steps, warmup, prec, timing=True, graph=False):
    import ndspmhd_amd as nd
    o, st = setup(nside)
    sim = nd.Simulation(o, st, precision=prec)
    sim.derivs()
    # tile shapes chosen by timing derivs on this workload (state unchanged),
    # or given
    if SHAPES == "auto":
        shapes = sim.autotune(reps=2)
    else:
        sim.set_tile_shapes(*SHAPES)
        shapes = dict(zip(("force_workgroup", "density_workgroup", "cells_per_radius"), SHAPES))
    sim.set_graph(graph)
    for _ in range(warmup):
        sim.step()
    sim.set_timing(timing)
    cons0 = sim.conserved()
    pd = pf = 0
    nits = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
        pd += sim.stats["pairs_density"]
        pf += sim.stats["pairs_force"]
        nits.append(sim.stats["ncalc_mean"])
    t1 = time.perf_counter()
    kt = sim.kernel_times() if timing else {}
    cons1 = sim.conserved()
    numneigh = float(np.mean(sim.get("numneigh")))
    return dict(sim=sim, opts=o, wall=t1 - t0, pairs_density=pd, pairs_force=pf, shapes=shapes,
                numneigh_mean=numneigh,
                ncalc_mean=float(np.mean(nits)), kernel_ms=kt, n=nside ** 3,
                cons=(cons0, cons1, float(np.sum(st["pmass"]))))


def conservation(cons0, cons1, mass):
    """drift of the FP64 conserved sums over the timed steps: relative
    change of the total energy; change of the total momentum vector and of
    the angular momentum, in units of sqrt(2 M E_kin) (the momentum scale of
    the flow)"""
    pscale = np.sqrt(2.0 * mass * max(cons0["ekin"], 1e-300))
    dmom = np.sqrt(sum((cons1[k] - cons0[k]) ** 2 for k in ("momx", "momy", "momz")))
    dang = np.sqrt(sum((cons1[k] - cons0[k]) ** 2 for k in ("angx", "angy", "angz")))
    return {"etot_rel_drift": (cons1["etot"] - cons0["etot"]) / abs(cons0["etot"]),
            "momentum_drift": dmom / pscale, "angular_momentum_drift": dang / pscale,
            "etot_start": cons0["etot"], "etot_end": cons1["etot"]}


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary: the DRAM-side read and write request counters in 32-byte units
    (TCC_EA0_RDREQ_DRAM_32B, TCC_EA0_WRREQ_WRITE_DRAM_32B), which read the
    exact byte count for the plain copy kernel of the same run; None when no
    summary is committed"""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_mhd128_mixed.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] in ("dram_read_bytes", "dram_write_bytes"):
            vals[(f[0], f[1])] = float(f[2])
    try:
        return vals[("dram_read_bytes", kernel)] + vals[("dram_write_bytes", kernel)]
    except KeyError:
        return None


def run_mlp(nside, steps):
    """h-predictor leg: the same box stepped with the reference seed
    h_in + dt dh/dt (mixed precision), with the online-trained MLP seed
    (mixed), and with the reference seed in plain FP64; reports the mean
    Newton-Raphson passes per particle over the second half of the run and
    the deviation of h and rho of the MLP run from the plain FP64 run (both
    iterate every particle to tolh)."""
    import ndspmhd_amd as nd
    o, st = setup(nside)
    out = {"particles": nside ** 3, "steps": steps, "tolh": o.tolh}
    fields = {}
    for mode, prec in (("plain", 0), ("mlp", 0), ("plain_fp64", 1)):
        sim = nd.Simulation(o, st, precision=prec)
        if mode == "mlp":
            sim.mlp_config(enable=True, train=True, lr=2e-3, batch=16384, iters=4, seed_after=3)
        sim.derivs()
        nc, tw = [], []
        for _ in range(steps):
            t0 = time.perf_counter()
            sim.step()  # returns after the step's scalars are on the host
            tw.append(time.perf_counter() - t0)
            nc.append(sim.stats["ncalc_mean"])
        # the second half: the seed is applied from step seed_after on
        out[mode] = {"ncalc_mean_second_half": float(np.mean(nc[steps // 2:])),
                     "steps_per_s_second_half": 1.0 / float(np.mean(tw[steps // 2:])),
                     "steps_per_s": steps / float(np.sum(tw))}
        fields[mode] = (sim.get("hh").copy(), sim.get("rho").copy())
        # per-kernel time (ms per evaluation, one evaluation a step) over a few
        # more steps with events on: the seed's own cost beside what it saves
        # in the density passes
        sim.set_timing(True)
        for _ in range(4):
            sim.step()
        out[mode]["kernel_ms_per_evaluation"] = {
            k: round(float(v), 4) for k, v in sim.kernel_times().items()
            if k in ("mlp_seed", "mlp_target", "mlp_train", "density", "predict", "nblist", "force")}
        sim.close()
    for k, (a, b) in {"h": (fields["mlp"][0], fields["plain_fp64"][0]),
                      "rho": (fields["mlp"][1], fields["plain_fp64"][1])}.items():
        out["max_rel_dev_vs_fp64_" + k] = float(np.max(np.abs(a - b) / np.abs(b)))
    for k, (a, b) in {"h": (fields["plain"][0], fields["plain_fp64"][0]),
                      "rho": (fields["plain"][1], fields["plain_fp64"][1])}.items():
        out["plain_max_rel_dev_vs_fp64_" + k] = float(np.max(np.abs(a - b) / np.abs(b)))
    e_p, e_m = out["plain"]["ncalc_mean_second_half"] - 1.0, out["mlp"]["ncalc_mean_second_half"] - 1.0
    out["extra_passes_saved_fraction"] = (e_p - e_m) / e_p if e_p > 0 else 0.0
    log("[bench] h predictor on %d^3: %s" % (nside, out))
    return out


def run_sweep(nsim, steps):
    """Batched sweep: nsim independent 1D/2D MHD shock tubes of uneven N
    (log-uniform 200..20000), all stepped together with ndsp_batch_step, and
    the same members stepped one after another as the baseline."""
    import ndspmhd_amd as nd
    from ndspmhd_amd.setups import shock_sweep
    cfgs = shock_sweep(nsim=nsim)
    out = {"members": nsim, "particles": int(sum(st["x"].shape[0] for _, st in cfgs)),
           "steps": steps}
    for mode in ("batched_tuned", "batched", "one_by_one"):
        sims = [nd.Simulation(o, st, precision=0) for o, st in cfgs]
        batch = nd.Batch(sims) if mode != "one_by_one" else None
        if mode == "batched_tuned":
            t0 = time.perf_counter()
            out["tile_shapes"] = batch.autotune(reps=3)
            out["autotune_s"] = time.perf_counter() - t0
        for _ in range(2):  # warmup (records the graphs)
            if batch:
                batch.step()
            else:
                for s in sims:
                    s.step()
        pairs = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            if batch:
                batch.step()
            else:
                for s in sims:
                    s.step()
            pairs += sum(s.stats["pairs_density"] + s.stats["pairs_force"] for s in sims)
        wall = time.perf_counter() - t0
        out[mode] = {"sweep_steps_per_s": steps / wall, "pairs_per_s": pairs / wall}
        if batch:
            batch.close()
        for s in sims:
            s.close()
    out["speedup_batched"] = out["batched"]["sweep_steps_per_s"] / out["one_by_one"]["sweep_steps_per_s"]
    out["speedup_batched_tuned"] = out["batched_tuned"]["sweep_steps_per_s"] / out["one_by_one"]["sweep_steps_per_s"]
    log("[bench] sweep of %d: %s" % (nsim, out))
    return out


def field_errors(nside):
    """mixed vs FP64 GPU on one derivs of the same state (norm-relative L-inf),
    the box's uniform B perturbed by a random tangled component"""
    import ndspmhd_amd as nd
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(nside)
    # a tangled field on top of the uniform one, so that divB and the
    # Lorentz force are not zero
    rng = np.random.default_rng(17)
    st["Bevol"] = st["Bevol"] + 0.05 * rng.standard_normal(st["Bevol"].shape)
    o = o.replace(maxdensits=0)
    a = nd.Simulation(o, st, precision=0)
    b = nd.Simulation(o, st, precision=1)
    a.derivs()
    b.derivs()
    out = {}
    for k in ("rho", "force", "dBevoldt", "dudt", "divB"):
        x, y = a.get(k), b.get(k)
        out[k] = float(np.max(np.abs(x - y)) / max(np.max(np.abs(y)), 1e-300))
    a.close()
    b.close()
    return out


def cpu_baseline(sample_nside):
    """The compiled reference (serial FP64 Fortran, oracle/_ref) on a bounded
    sample of the same workload; pair counts of the sample come from the GPU
    run of the identical input (the count is a property of the input)."""
    import ndspmhd_amd as nd
    from ndspmhd_amd.setups import periodic_mhd_box
    o, st = periodic_mhd_box(sample_nside)
    g = nd.Simulation(o, st, precision=1)
    g.derivs()
    pairs = g.stats["pairs_density"] + g.stats["pairs_force"]
    g.close()
    ref = os.path.join(ROOT, "oracle", "_ref", "ndsp_ref3d")
    if os.path.exists(ref):
        sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
        from refio import run_reference
        out = run_reference(ref, o, st, mode=1, timeout=900)
        secs = [float(l.split()[1]) for l in out["log"].splitlines() if "NDSP_DERIVS_SECONDS" in l]
        t = secs[0]
        kind = "reference"
        what = "compiled reference derivs (amdflang -O2, serial)"
    else:
        import oracle
        orc = oracle.Oracle(o, st)
        t0 = time.perf_counter()
        orc.derivs()
        t = time.perf_counter() - t0
        kind = "port"
        what = "CPU oracle derivs (C restatement, serial)"
    return dict(value=pairs / t, unit="pair interactions/s", cores=1, kind=kind, cpu=cpu_model(),
                sample="%s on the %d^3 box (%d particles, %d pairs, %.2f s)" % (
                    what, sample_nside, sample_nside ** 3, pairs, t))


def cpu_model():
    """the host CPU's model name (first processor in /proc/cpuinfo)"""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nside", type=int, default=128)
    ap.add_argument("--cpu-nside", type=int, default=96)
    ap.add_argument("--big-nside", type=int, default=256,
                    help="second workload (mixed vs plain FP64 steps/s); 0 skips it")
    ap.add_argument("--big-steps", type=int, default=5)
    ap.add_argument("--mlp-nside", type=int, default=256,
                    help="box of the h-predictor leg (plain seed vs trained seed); 0 skips it")
    ap.add_argument("--mlp-steps", type=int, default=12)
    ap.add_argument("--sweep", type=int, default=64,
                    help="members of the batched shock-tube sweep leg; 0 skips it")
    ap.add_argument("--sweep-steps", type=int, default=20)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed region, one device copy of this many MiB "
                         "(reference kernel for scaling rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--shapes", default="auto",
                    help="tile shapes of the main runs: auto (autotune) or FORCE_WG,DENSITY_WG,CELLS_PER_RADIUS")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    global SHAPES
    if args.shapes != "auto":
        SHAPES = tuple(int(v) for v in args.shapes.split(","))
    if args.gpus != 1:
        print("bench.py: the library runs on one GPU only; sharding over several "
              "GPUs is a design, not implemented. Run with --gpus 1.", file=sys.stderr)
        return 2

    log("[bench] mixed precision, %d^3, warmup %d, steps %d" % (args.nside, args.warmup, args.steps))
    r = run_gpu(args.nside, args.steps, args.warmup, prec=0)
    n = r["n"]
    pairs = r["pairs_density"] + r["pairs_force"]
    value = pairs / r["wall"]
    ms_step = 1e3 * r["wall"] / args.steps
    kt = r["kernel_ms"]
    log("[bench] %.3f ms/step, %.3e pairs/s, kernels %s" % (ms_step, value, kt))
    pairs_force_step = r["pairs_force"] / args.steps
    pairs_density_step = r["pairs_density"] / args.steps
    # roofline of each pair kernel: algorithmic bytes and FP32 flops per
    # launch; t_roof = max(bytes / HBM peak, flops / FP32 peak)
    work = {"nblist": (BYTES_BIN * n, 0.0),
            "density": (BYTES_DENSITY * n * r["ncalc_mean"], FLOPS_DENSITY_PAIR * pairs_density_step),
            "force": (BYTES_FORCE * n, FLOPS_FORCE_PAIR * 2 * pairs_force_step)}
    per_kernel = {}
    for k, (b, f) in work.items():
        if k not in kt:
            continue
        t = kt[k] * 1e-3
        t_roof = max(b / (HBM_PEAK_GBS * 1e9), f / (FP32_PEAK_TFLOPS * 1e12))
        per_kernel[k] = {"ms": kt[k], "alg_bytes": b, "alg_fp32_flops": f,
                         "hbm_gbs": b / t / 1e9, "hbm_frac": b / t / 1e9 / HBM_PEAK_GBS,
                         "fp32_tflops": f / t / 1e12, "fp32_frac": f / t / 1e12 / FP32_PEAK_TFLOPS,
                         "roof_frac": t_roof / t,
                         "traffic": traffic_from_profile({"density": "density_evaluation"}.get(k, "k_" + k))}
    dom = max(per_kernel, key=lambda k: per_kernel[k]["ms"]) if per_kernel else "force"
    alg_bytes = per_kernel[dom]["alg_bytes"]
    dom_ms = per_kernel[dom]["ms"]
    achieved = per_kernel[dom]["hbm_gbs"]
    result = {
        "metric": "pair interactions/s (3D 128^3 SPMHD leapfrog step, density + force pairs)",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "steps_per_s": 1e3 / ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair math with f64 accumulation; f64 positions, neighbour tests, h-rho update, integrator, reductions",
        "data": "synthetic",
        "config": {"workload": "3D periodic MHD box %d^3 (%d particles), cubic spline, hfact 1.2, "
                               "gamma 5/3, v_rms Mach 1 P(k)~k^-4, B=(0.1,0,0)" % (args.nside, n),
                   "particles": n, "steps": args.steps},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "limited_by": "VALU issue and gather latency (SQ counters), not HBM bandwidth",
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": per_kernel[dom]["traffic"],
                     "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": dom_ms,
                     "per_kernel": per_kernel},
        "kernel_ms": kt,
        "tile_shapes": r["shapes"],
        "numneigh_mean": r["numneigh_mean"],
        "pairs_per_step": {"density": r["pairs_density"] / args.steps, "force": pairs_force_step},
        "ncalc_mean": r["ncalc_mean"],
        "conservation": conservation(*r["cons"]),
    }
    r["sim"].close()
    log("[bench] same workload, steps replayed as HIP graphs")
    rg = run_gpu(args.nside, args.steps, args.warmup, prec=0, timing=False, graph=True)
    # the headline is the production configuration (steps replayed as
    # graphs, as the Fortran host runs them); the run above, with an event
    # around every kernel group, gives the kernel times and the roofline
    result["events_run"] = {"pairs_per_s": value, "ms_per_step": ms_step}
    result["value"] = (rg["pairs_density"] + rg["pairs_force"]) / rg["wall"]
    result["ms_per_step"] = 1e3 * rg["wall"] / args.steps
    result["steps_per_s"] = args.steps / rg["wall"]
    result["graph_replay"] = {"steps_per_s": args.steps / rg["wall"],
                              "pairs_per_s": result["value"], "ms_per_step": result["ms_per_step"],
                              "conservation": conservation(*rg["cons"])}
    rg["sim"].close()
    if not args.no_fp64:
        log("[bench] plain FP64 GPU, same steps, graph replay")
        r64 = run_gpu(args.nside, args.steps, args.warmup, prec=1, timing=False, graph=True)
        v64 = (r64["pairs_density"] + r64["pairs_force"]) / r64["wall"]
        result["fp64_gpu"] = {"value": v64, "ms_per_step": 1e3 * r64["wall"] / args.steps,
                              "steps_per_s": args.steps / r64["wall"],
                              "speedup_mixed_over_fp64": (args.steps / rg["wall"]) / (args.steps / r64["wall"])}
        r64["sim"].close()
        # the same steps in FP64 with an event around every kernel group:
        # where the mixed path gains, kernel by kernel
        log("[bench] plain FP64 GPU, kernel times")
        r64t = run_gpu(args.nside, args.steps, args.warmup, prec=1, timing=True, graph=False)
        kt64 = r64t["kernel_ms"]
        result["fp64_gpu"]["kernel_ms"] = kt64
        result["fp64_gpu"]["fp64_over_mixed_per_kernel"] = {
            k: kt64[k] / kt[k] for k in kt if k in kt64 and kt[k] > 0}
        r64t["sim"].close()
        result["error_vs_fp64"] = field_errors(min(args.nside, 64))
    if args.big_nside > 0 and not args.no_fp64:
        nb = args.big_nside
        log("[bench] %d^3 mixed vs plain FP64, %d steps each" % (nb, args.big_steps))
        big = {}
        for prec, key in ((0, "mixed"), (1, "fp64")):
            rb = run_gpu(nb, args.big_steps, 1, prec=prec, timing=False, graph=True)
            big[key] = {"steps_per_s": args.big_steps / rb["wall"],
                        "pairs_per_s": (rb["pairs_density"] + rb["pairs_force"]) / rb["wall"],
                        "ncalc_mean": rb["ncalc_mean"]}
            rb["sim"].close()
            del rb
        big["speedup_mixed_over_fp64"] = big["mixed"]["steps_per_s"] / big["fp64"]["steps_per_s"]
        big["particles"] = nb ** 3
        result["box_%d" % nb] = big
    if args.mlp_nside > 0 and not args.no_fp64:
        result["mlp_seed"] = run_mlp(args.mlp_nside, args.mlp_steps)
    if args.sweep > 0:
        result["sweep"] = run_sweep(args.sweep, args.sweep_steps)
    if args.copy_probe_mib > 0:
        import ndspmhd_amd as nd
        o, st = setup(args.nside)
        probe = nd.Simulation(o, st, precision=0)
        probe.copy_probe(args.copy_probe_mib << 20)
        result["copy_probe_bytes"] = args.copy_probe_mib << 20
        probe.close()
    if not args.no_cpu:
        log("[bench] CPU baseline on %d^3" % args.cpu_nside)
        result["cpu_baseline"] = cpu_baseline(args.cpu_nside)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
