"""Throughput of the active-KA MD step on one MI355X.

Workload (the batched-ensemble configuration the headline metric is quoted
on): 64 replicas of the 80:20 Kob-Andersen mixture at density 1.2 and T=0.5,
N_r drawn log-uniformly in [1000, 5000] from a seed, active force grid
f0 in {0, 0.5, 1, 1.5, 2} x tau_p in {0.1, 1, 10} (tau_per 20/200/2000 steps),
pair forces in FP32 with FP64 positions/momenta/thermostat, all replicas packed
into one launch sequence per step and replayed as a hipGraph.  A "step" is one
MD step of every replica; value = sum_r N_r x steps / seconds.

Beside it: the plain single run (one replica, FP64 everywhere, eager launches)
on the same kind of system, the per-kernel times from HIP events on the
library's stream, the roofline of the dominant kernel and the oracle (the CPU
restatement) timed on a bounded sample on one host core.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas R]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# compulsory bytes per particle of the fused force kernel (k_force): own
# position 24 + type 4 + momentum read 24 + force write 24 + momentum write 24
FORCE_BYTES_PER_PARTICLE = 100.0
# whole step (k_first_half + k_force), FP64 positions/momenta, once each
STEP_BYTES_PER_PARTICLE = 100.0 + 172.0


def ensemble_params(ac4, R, seed=20261202):
    rng = np.random.default_rng(seed)
    ns = np.floor(np.exp(rng.uniform(math.log(1000), math.log(5000), R))).astype(int)
    f0s = [0.0, 0.5, 1.0, 1.5, 2.0]
    taus = [20, 200, 2000]
    ps = []
    for r in range(R):
        g = r % 15
        ps.append(ac4.ReplicaParams(int(ns[r]), 1.2, 0.5, f0s[g % 5], -1, taus[g // 5], seed + r))
    return ps


def load_states(ac4, md, ps):
    for r, p in enumerate(ps):
        s = ac4.init_lattice(p.n, p.T, p.seed)
        md.set_state(r, s["x"], s["y"], s["z"], s["xu"], s["yu"], s["zu"], s["px"], s["py"], s["pz"],
                     s["type"], s["act_tag"], s["kx"], s["ky"], s["kz"])


def cpu_baseline(ac4, n=2485, steps=1500):
    """Oracle (single-threaded C restatement of the reference step) on one core,
    from a state brought to temperature on the GPU (the oracle has no ramp)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    lib = OL.load()
    p = ac4.ReplicaParams(n, 1.2, 0.5, 1.0, -1, 0, 99)
    md = ac4.MD([p], precision=ac4.FP64)
    load_states(ac4, md, [p])
    md.relax(300, 1.5, 1.0)
    md.init_chain(0)
    g = md.get_state(0)
    chain = md.get_chain(0)
    md.close()
    s = ac4.init_lattice(n, 0.5, 99)
    s.update({k: g[k] for k in ("x", "y", "z", "xu", "yu", "zu", "px", "py", "pz")})
    t0 = time.perf_counter()
    _, _, out = OL.nhc_steps(s, lib.or_md_box_length(n, 1.2), 0.5, 1.0, steps, chain)
    dt = time.perf_counter() - t0
    assert np.isfinite(out[4]) and 0.2 < out[3] < 1.0, out
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle NHC step (FP64, half Verlet list, f0=1.0) on one N=%d replica, %d steps, %.1f s"
                      % (n, steps, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=400)
    ap.add_argument("--relax", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import activechi4_amd as ac4
    ac4.require_gpu()

    ps = ensemble_params(ac4, a.replicas)
    ntot = sum(p.n for p in ps)
    md = ac4.MD(ps, precision=ac4.FP32)
    load_states(ac4, md, ps)
    md.relax(a.relax, 1.5, 1.0)  # lattice start -> temperature (coldstart/breakxal)
    for r in range(a.replicas):
        md.init_chain(r, seed=r)
    print("bench: %d replicas, %d particles; relaxed %d steps, warmup %d steps" % (a.replicas, ntot, a.relax,
                                                                                a.warmup), file=sys.stderr,
          flush=True)
    md.step(a.warmup, graph=True)
    md.sync()
    t0 = time.perf_counter()
    md.step(a.steps, graph=True)
    md.sync()
    t_batched = time.perf_counter() - t0
    value = ntot * a.steps / t_batched
    print("bench: batched %.3f s for %d steps" % (t_batched, a.steps), file=sys.stderr, flush=True)
    kt = md.kernel_times(200)
    rebuilds = sum(md.rebuilds(r) for r in range(a.replicas))
    md.close()

    # plain single run: one replica, FP64, eager launches
    p0 = ac4.ReplicaParams(ps[0].n, 1.2, 0.5, 1.0, -1, 200, 7)
    plain = ac4.MD([p0], precision=ac4.FP64)
    load_states(ac4, plain, [p0])
    plain.relax(a.relax, 1.5, 1.0)
    plain.init_chain(0)
    plain.step(100)
    plain.sync()
    t0 = time.perf_counter()
    plain.step(a.plain_steps)
    plain.sync()
    t_plain = time.perf_counter() - t0
    plain_value = p0.n * a.plain_steps / t_plain
    plain.close()

    t_force = kt["force"] * 1e-3
    achieved = FORCE_BYTES_PER_PARTICLE * ntot / t_force / 1e9
    line = {
        "metric": "particle-steps/s, batched replica server vs plain single run; p50/p99 latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_batched / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded lattice starts)",
        "config": {"workload": "batched ensemble: %d replicas KA 80:20 rho=1.2 T=0.5, N logU[1000,5000] "
                               "(sum %d), f0 x tau_p grid, FP32 pair forces + FP64 state, hipGraph"
                               % (a.replicas, ntot),
                   "replicas": a.replicas, "particles": ntot, "forces": "fp32", "positions": "fp64"},
        "plain_single_run": {"value": plain_value, "n": p0.n, "dtype": "fp64", "steps": a.plain_steps,
                             "ms_per_step": 1e3 * t_plain / a.plain_steps},
        "speedup_vs_plain": value / plain_value,
        "kernel_ms_per_step": kt,
        "rebuilds_total": rebuilds,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_force<fp32,kick>", "bytes_per_particle": FORCE_BYTES_PER_PARTICLE},
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ac4)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
