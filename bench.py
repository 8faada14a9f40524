"""Throughput of the active-KA MD step on one MI355X.

Workload (the batched-ensemble configuration the headline metric is quoted
on): 64 replicas of the 80:20 Kob-Andersen mixture at density 1.2 and T=0.5,
N_r drawn log-uniformly in [1000, 5000] from a seed, active force grid
f0 in {0, 0.5, 1, 1.5, 2} x tau_p in {0.1, 1, 10} (tau_per 20/200/2000 steps),
pair forces in FP32 with FP64 positions/momenta/thermostat, all replicas packed
into one launch sequence per step and replayed as a hipGraph, kernel shape
chosen by the autotuner (ac4_autotune), 5e4 steps with the overlap Q(t)
(a = 0.3) of every replica accumulated on the device every 100 steps.  A
"step" is one MD step of every replica; value = sum_r N_r x steps / seconds of
the timed region.

Beside it:
- chi4(t) of every (f0, tau_p) point from an untimed fixed-length run of the
  same ensemble (5e4 steps), once with FP32 pair forces and once in FP64: the
  chi4 peak height and time of each point with jackknife errors over its
  replicas, and their FP32-vs-FP64 deviation;
- the plain single run: the same 64 replicas run one after another, FP64
  everywhere, eager launches (one launch sequence per step);
- a serving stream: jobs with N logU[500, 20000], T U[0.45, 0.8], f0 U[0, 2],
  tau_p logU[0.1, 10], steps U[1e4, 1e5] arriving as a seeded Poisson process,
  frames every 100 steps counted into each job's Q(t) on the device, slots
  planned from an HBM budget, p50/p99 time to completion, against the same
  jobs run one after another with the plain kernel (each one measured);
- per-kernel times (HIP events on the library's stream), the roofline of the
  force kernel, and the oracle (the CPU restatement) on a bounded sample on
  the host's cores.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--replicas R] [--serve-jobs J]
"""
import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# compulsory HBM bytes per particle of the fused force kernel (k_force, FP32
# pair math, half kick fused): own fixed-point position with the species bit
# 12 + list length 4 + active index 4 + momentum read 24 + force write 24 +
# momentum write 24 = 92 B.  Neighbour positions are gathered from cache; the
# list entries (4 B each) are the kernel's other stream and show in the
# measured traffic.
FORCE_BYTES_PER_PARTICLE = 92.0
# compulsory HBM bytes of one whole mixed-precision step per particle (force
# kernel: packed FP32 position + type 16, force 12; integration: read v, F, x,
# x_u 84, write v, x, x_u and the FP32 position copy 88): 200 B.
STEP_BYTES_PER_PARTICLE = 200.0
PMC_FILE = os.path.join(ROOT, "profiles", "k_force_batched64_pmc.txt")


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


def start(ac4, ps, precision, relax):
    """Lattice starts brought to temperature (coldstart/breakxal ramp), then
    the Nose-Hoover chains drawn from each replica's seed."""
    md = ac4.MD(ps, precision=precision)
    load_states(ac4, md, ps)
    md.relax(relax, 1.5, 1.0)
    for r, p in enumerate(ps):
        md.init_chain(r, seed=p.seed)
    return md


def default_shape(ac4):
    """Kernel shape chosen automatically from the handle size, reference skin 0.35."""
    ac4.set_tuning(0, 3)
    ac4.set_skin(0.35)


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


CHI4_STEPS = 50000                      # the batched ensemble's run length
CHI4_OV = (100, 1000, 26, 250, 0.3)     # sample every, origin every, origins kept, lags, a


def chi4_leg(ac4, ps, precision, relax, steps=CHI4_STEPS):
    """chi4(t) of every (f0, tau_p) point of the ensemble from a fixed-length
    run (untimed): overlap Q(t) of every replica accumulated on the device
    every 100 steps from origins every 1000 steps, lags up to 2.5e4 steps;
    chi4 = Nbar (<Q^2> - <Q>^2) pooled over the point's replicas and origins
    (FP64 on the device), its peak and jackknife errors over the replicas."""
    se, oe, no, nl, a = CHI4_OV
    md = start(ac4, ps, precision, relax)
    md.overlap_enable(se, oe, no, nl, a)
    t0 = time.perf_counter()
    md.overlap_sample()  # step 0: the first origin
    for _ in range(steps // se):
        md.step(se, graph=True)
        md.overlap_sample()
    md.sync()
    wall = time.perf_counter() - t0
    groups = {}
    for r, p in enumerate(ps):
        groups.setdefault((p.f0, p.tau_per), []).append(r)
    res = {key: ac4.chi4_peak(md, rs, se * 0.005) for key, rs in sorted(groups.items())}
    md.close()
    return res, groups, wall


def chi4_compare(ac4, ps, relax):
    """The batched FP32-force ensemble against the same replicas in FP64 (the
    plain form: a replica's results do not depend on the batch): chi4 peak
    height and time per (f0, tau_p) point, with jackknife errors."""
    r32, groups, w32 = chi4_leg(ac4, ps, ac4.FP32, relax)
    r64, _, w64 = chi4_leg(ac4, ps, ac4.FP64, relax)
    rows = []
    for key in sorted(groups):
        a, b = r32[key], r64[key]
        row = {"f0": key[0], "tau_p": key[1] * 0.005, "replicas": len(groups[key])}
        for tag, v in (("fp32", a), ("fp64", b)):
            row[tag] = {"peak": v["peak"], "peak_err": v.get("peak_err"), "peak_time": v["peak_time"],
                        "peak_time_err": v.get("peak_time_err"), "Q_at_peak": v["Q_at_peak"]}
        row["peak_rel_dev"] = (a["peak"] - b["peak"]) / b["peak"]
        if a.get("peak_err") is not None and b.get("peak_err") is not None:
            row["peak_rel_sigma"] = math.hypot(a["peak_err"], b["peak_err"]) / b["peak"]
        row["peak_time_dev"] = a["peak_time"] - b["peak_time"]
        row["resolved"] = b["peak_lag"] < CHI4_OV[3]  # the peak lies inside the lag window
        rows.append(row)
    dev = [abs(r["peak_rel_dev"]) for r in rows]
    zs = [abs(r["peak_rel_dev"]) / r["peak_rel_sigma"] for r in rows if r.get("peak_rel_sigma")]
    return {"a": CHI4_OV[4], "steps": CHI4_STEPS, "sample_every": CHI4_OV[0], "origin_every": CHI4_OV[1],
            "max_lag_time": CHI4_OV[0] * CHI4_OV[3] * 0.005, "wall_s": {"fp32": w32, "fp64": w64},
            "by_param": rows, "max_abs_peak_rel_dev": max(dev), "median_abs_peak_rel_dev": float(np.median(dev)),
            "points_within_2pct": int(sum(d <= 0.02 for d in dev)),
            "points_within_2sigma": int(sum(z <= 2.0 for z in zs)), "points": len(rows)}


def plain_sequential(ac4, ps, relax, steps):
    """The same replicas one after another: FP64, eager launches."""
    default_shape(ac4)
    tot_ps, tot_t = 0.0, 0.0
    for p in ps:
        md = start(ac4, [p], ac4.FP64, relax)
        md.step(50)
        md.sync()
        t0 = time.perf_counter()
        md.step(steps)
        md.sync()
        tot_t += time.perf_counter() - t0
        tot_ps += p.n * steps
        md.close()
    return tot_ps / tot_t, tot_t / steps * 1e3


SERVE_BUDGET = 512e6          # HBM bytes the server may use for slots
SERVE_OVERLAP = (1000, 10, 100, 0.3)  # per-job Q(t): origins every 1000 steps, 10 kept, lags to 1e4 steps


def serving_jobs(njobs, rate_per_s, seed=4242):
    """The serving stream's requests: N logU[500, 20000], T U[0.45, 0.8],
    f0 U[0, 2], tau_p logU[0.1, 10], steps U[1e4, 1e5], Poisson arrivals."""
    rng = np.random.default_rng(seed)
    jobs = []
    t = 0.0
    for k in range(njobs):
        t += rng.exponential(1.0 / rate_per_s)
        n = int(math.exp(rng.uniform(math.log(500), math.log(20000))))
        jobs.append(dict(n=n, T=float(rng.uniform(0.45, 0.8)), f0=float(rng.uniform(0.0, 2.0)),
                         tau_p=float(math.exp(rng.uniform(math.log(0.1), math.log(10.0)))),
                         steps=int(rng.integers(10000, 100001)), seed=1000 + k, t=t))
    return jobs


def plain_job(ac4, j, relax=300):
    """One job with the plain kernel (FP64, eager launches): seeded lattice,
    the relaxation ramp, then its production steps; seconds of GPU work."""
    default_shape(ac4)
    tau = max(1, int(round(j["tau_p"] / 0.005)))
    md = ac4.MD([ac4.ReplicaParams(j["n"], 1.2, j["T"], j["f0"], -1, tau, j["seed"])], precision=ac4.FP64)
    load_states(ac4, md, [ac4.ReplicaParams(j["n"], 1.2, j["T"], j["f0"], -1, tau, j["seed"])])
    md.sync()
    t0 = time.perf_counter()
    md.relax(relax, 1.5, 1.0)
    md.init_chain(0, seed=j["seed"])
    left = ((j["steps"] + 99) // 100) * 100
    while left > 0:
        k = min(left, 5000)
        md.step(k)
        left -= k
    md.sync()
    dt = time.perf_counter() - t0
    md.close()
    return dt


def serving(ac4, njobs, rate_per_s, seed=4242):
    """Seeded Poisson stream of MD jobs through the continuous-batching
    server (slots planned from an HBM budget, every job's Q(t) kept on the
    device), against the same jobs run one after another with the plain
    kernel, measured job by job."""
    jobs = serving_jobs(njobs, rate_per_s, seed)
    default_shape(ac4)
    srv = ac4.Server(precision=ac4.FP32, chunk=100, relax=300, budget=SERVE_BUDGET, n_range=(500, 20000),
                     overlap=SERVE_OVERLAP)
    caps = srv.slots()
    groups = srv.groups()
    ids = [srv.submit(j["n"], j["T"], j["f0"], j["tau_p"], j["steps"], j["seed"], j["t"]) for j in jobs]
    trace = []  # (time, particle-steps done, jobs waiting) after every tick
    t0 = time.perf_counter()
    while srv.pending():
        now = time.perf_counter() - t0
        if srv.tick(now) == 0:
            time.sleep(0.001)  # nothing admitted yet: wait for the next arrival
        waiting = sum(1 for j, i in zip(jobs, ids) if j["t"] <= now and srv.job(i)["state"] == 0)
        trace.append((time.perf_counter() - t0, srv.particle_steps(), waiting))
    wall = time.perf_counter() - t0
    lat = []
    for j, i in zip(jobs, ids):
        st = srv.job(i)
        assert st["state"] == 2, (i, st)
        lat.append(st["t_done"] - j["t"])
    q1 = [srv.job_overlap(i) for i in ids[:1]]
    srv.close()
    work = sum(j["n"] * j["steps"] for j in jobs)
    serve_rate = work / (wall - jobs[0]["t"])
    # saturated window: from the first tick with a job waiting for a slot to
    # the last such tick (the server full and requests queued)
    sat = [k for k, (_, _, w) in enumerate(trace) if w > 0]
    sat_rate = None
    if len(sat) >= 2 and trace[sat[-1]][0] > trace[sat[0]][0]:
        sat_rate = (trace[sat[-1]][1] - trace[sat[0]][1]) / (trace[sat[-1]][0] - trace[sat[0]][0])
    log("serving: %d jobs in %.1f s on %d slots; plain sequential next" % (njobs, wall, len(caps)))
    t_plain = [plain_job(ac4, j) for j in jobs]
    t_seq = sum(t_plain)
    lat = np.array(lat)
    # the same arrivals served first-come first-served, one job at a time,
    # with the plain kernel (each job's measured time)
    free, seq_lat = 0.0, []
    for j, tj in zip(jobs, t_plain):
        free = max(free, j["t"]) + tj
        seq_lat.append(free - j["t"])
    qq, _, qc = q1[0]
    return {"jobs": njobs, "arrival_rate_per_s": rate_per_s, "hbm_budget_bytes": SERVE_BUDGET,
            "slots": len(caps), "launch_groups": groups,
            "slot_classes": {str(c): int((caps == c).sum()) for c in np.unique(caps)},
            "slot_capacity_particles": int(caps.sum()), "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "value": serve_rate, "unit": "particle-steps/s",
            "value_note": "all jobs' particle-steps / (last completion - first arrival)",
            "saturated_value": sat_rate,
            "saturated_window_s": (trace[sat[-1]][0] - trace[sat[0]][0]) if len(sat) >= 2 else 0.0,
            "plain_sequential_value": work / t_seq, "plain_sequential_makespan_s": t_seq,
            "plain_sequential_measured": True,
            "plain_sequential_p50_latency_s": float(np.percentile(seq_lat, 50)),
            "plain_sequential_p99_latency_s": float(np.percentile(seq_lat, 99)),
            "speedup_vs_plain_sequential": serve_rate * t_seq / work, "wall_s": wall,
            "production_particle_steps": work,
            "job0_Q_at_lag_1e4_steps": float(qq[-1] / qc[-1]) if qc[-1] else None}


def _oracle_worker(args):
    """One host core: the oracle's NHC step on one replica (spawned process)."""
    s, L, chain, steps = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    t0 = time.perf_counter()
    _, _, out = OL.nhc_steps(s, L, 0.5, 1.0, steps, chain)
    return time.perf_counter() - t0, out


def cpu_baseline(ac4, n=2485, steps=2000, max_workers=16):
    """The oracle (single-threaded C restatement of the reference step) on
    every host core the box gives this GPU (at most 16): one replica per
    core, in spawned processes, from a state brought to temperature on the
    GPU (the oracle has no ramp).  Replicas are independent, so the host's
    rate is the sum over cores."""
    import multiprocessing as mp
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as OL
    L = OL.load().or_md_box_length(n, 1.2)
    p = ac4.ReplicaParams(n, 1.2, 0.5, 1.0, -1, 0, 99)
    default_shape(ac4)
    md = start(ac4, [p], ac4.FP64, 300)
    g = md.get_state(0)
    chain = md.get_chain(0)
    md.close()
    s = ac4.init_lattice(n, 0.5, 99)
    s.update({k: g[k] for k in ("x", "y", "z", "xu", "yu", "zu", "px", "py", "pz")})
    nproc = os.cpu_count() or 1
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    w = max(1, min(nproc, max_workers))
    with mp.get_context("spawn").Pool(w) as pool:
        pool.map(_oracle_worker, [(s, L, chain, 5)] * w)  # interpreter and library start-up, untimed
        t0 = time.perf_counter()
        res = pool.map(_oracle_worker, [(s, L, chain, steps)] * w)
        wall = time.perf_counter() - t0
    for _, out in res:
        assert np.isfinite(out[4]) and 0.2 < out[3] < 1.0, out
    return {"value": w * n * steps / wall, "unit": "particle-steps/s", "cores": w, "host_cpus": nproc,
            "kind": "port", "per_core_value": float(np.mean([n * steps / t for t, _ in res])),
            "sample": "oracle NHC step (FP64, half Verlet list, f0=1.0): one N=%d replica per core on %d cores, "
                      "%d steps each, %.1f s wall" % (n, w, steps, wall)}


def cfg5(ac4, n=1_000_000, steps=20000, nframes=500, relax=500, ndir=64, seed=5150, eager=False):
    """Large single box with S4: KA N=10^6, rho=1.2, T=0.5, c=0.1, f0=1.0,
    tau_p=1.0, FP32 forces, 2e4 steps with about 500 log-spaced frames (the
    reference schedule with the smallest origin count that gives that many),
    every frame kept in HBM.  Q(t) and chi4(t) over the reference's
    origin/lag walk (a=0.3) give the peak lag tau*; S4(q, tau*) then follows
    on |q| shells from 0.1 to 7.25 in steps of 2 pi/L with ndir seeded random
    directions per shell, weighted by the overlap of each origin at tau*.
    eager: launch every MD step's kernels one by one (for kernel-trace
    profiles, which are not taken over graph replay)."""
    default_shape(ac4)
    p = ac4.ReplicaParams(n, 1.2, 0.5, 1.0, -1, 200, seed)
    t_all = time.perf_counter()
    md = start(ac4, [p], ac4.FP32, relax)
    for no in range(5, 200):
        fidx = np.asarray(ac4.schedule(steps, no)[0], dtype=np.int64)
        if len(fidx) >= nframes:
            break
    L_ana = (n / 1.2) ** (1.0 / 3.0)  # Analysis/main.c:15
    fr = ac4.Frames(n, L_ana, len(fidx))
    an = ac4.Analysis(n, L_ana, fidx, obs=ac4.OBS_OVERLAP)
    log("cfg5: N=%d, %d frames (%d origins), frame store %.1f GB" % (n, len(fidx), no, 24e-9 * n * len(fidx)))
    md.sync()
    t0 = time.perf_counter()
    t = 0
    for f, fi in enumerate(fidx):
        if fi > t:
            md.step(int(fi - t), graph=not eager)
            t = int(fi)
        fr.push_md(md, 0, int(fi))
        an.push_md(f, md, 0)
    if t < steps:
        md.step(steps - t, graph=not eager)
    md.sync()
    t_md = time.perf_counter() - t0
    log("cfg5: %d steps with frames in %.1f s" % (steps, t_md))
    an.finalize()
    ov = an.overlap()  # per slot: sum_o Q, sum_o Q^2, origin counts (Analysis/overlap.c:89-118)
    cnt = ov["counters"][: ov["noi"]].astype(float)
    ok = np.arange(len(cnt)) >= 1
    ok &= cnt > 0
    qm = np.where(ok, ov["c"][: len(cnt)] / np.maximum(cnt, 1), np.nan)
    x4 = np.where(ok, ov["x4"][: len(cnt)] / np.maximum(cnt, 1) - qm * qm, np.nan)
    chi4 = n * x4
    k = int(np.nanargmax(chi4))
    tau_star = float(ov["delta_t"][k])
    lag = int(round(tau_star / 0.005))
    gap = an.gap
    pairs = [(fr.find(o), fr.find(o + lag)) for o in range(0, steps + 1, gap)]
    pairs = [q for q in pairs if q[0] >= 0 and q[1] >= 0]
    dq = 2.0 * math.pi / L_ana
    nshell = int((7.25 - 0.1) / dq + 1e-9) + 1
    t1 = time.perf_counter()
    s4r, s4d = fr.s4_shells([q[0] for q in pairs], [q[1] for q in pairs], 0.1, dq, nshell, ndir, seed)
    t_s4 = time.perf_counter() - t1
    qk = 0.1 + dq * np.arange(nshell)
    pick = [0, 1, 2, 4, 8, 16, 32, 64, nshell - 1]
    fr.close()
    an.close()
    md.close()
    return {"metric": "cfg5 large box", "n": n, "steps": steps, "frames": int(len(fidx)), "origins": no,
            "frame_store_GB": 24e-9 * n * len(fidx), "md_particle_steps_per_s": n * steps / t_md,
            "launch": "eager" if eager else "graph",
            "md_seconds": t_md, "chi4_peak": float(chi4[k]), "tau_star": tau_star,
            "Q_at_tau_star": float(qm[k]), "chi4_t": [[float(ov["delta_t"][i]), float(qm[i]), float(chi4[i])]
                                                      for i in range(len(cnt)) if ok[i]],
            "pairs_at_tau_star": len(pairs), "shells": nshell,
            "directions": ndir, "s4_seconds": t_s4,
            "S4_q": [float(qk[i]) for i in pick], "S4_ref_form": [float(s4r[i]) for i in pick],
            "S4_dir_avg": [float(s4d[i]) for i in pick], "total_seconds": time.perf_counter() - t_all}


def rows(ac4, n=50000, n_origins=20, gap=100, seed=777):
    """The analysis rows beside the MD step, each on the GPU (libac4) and on
    the oracle (one host core) over the same input:
      - trajectory text I/O of one N=10^6 frame (ac4_format_frame /
        ac4_parse_frame, and the device parser ac4_textframes, against glibc
        sprintf / strtod per line);
      - block analysis (ac4_frames_block, Analysis_Block: MSD, overlap and van
        Hove per block) for the 17 block edges L/3 ... L/19; the oracle runs
        the edge L/3 and the GPU is timed on that edge too;
      - g_uu(r) of the active particles (ac4_frames_guu, Analysis/guu.c);
      - cage-relative overlap (ac4_frames_overlap_cage,
        Analysis/overlap_cage_rel.c).
    Input: a seeded random walk of N particles (trajgen), unwrapped box
    fractions quantised as the %.5lf text, on the two-level frame schedule with
    n_origins origins gap steps apart, every frame in HBM."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import oracle_lib as OL
    import trajgen
    lib = OL.load()
    out = {"metric": "analysis rows: GPU seconds vs oracle seconds (one host core), same input"}
    # trajectory text I/O
    rng = np.random.default_rng(seed)
    m = 1_000_000
    x, y, z = (trajgen.quantize(rng.random(m) * 3.0 - 1.0) for _ in range(3))
    t = (rng.random(m) < 0.2).astype(np.int32)
    import ctypes
    buf = ctypes.create_string_buffer(80 * m + 1)
    t0 = time.perf_counter()
    k_ref = lib.or_frame_printf(m, OL.dp(x), OL.dp(y), OL.dp(z), OL.ip(t), buf)
    t_printf = time.perf_counter() - t0
    t0 = time.perf_counter()
    txt = ac4.format_frame(x, y, z, t)
    t_fmt = time.perf_counter() - t0
    assert txt == buf.raw[:k_ref]
    xs, ys, zs = np.empty(m), np.empty(m), np.empty(m)
    ts = np.empty(m, dtype=np.int32)
    t0 = time.perf_counter()
    assert lib.or_frame_strtod(txt, m, OL.dp(xs), OL.dp(ys), OL.dp(zs), OL.ip(ts)) == len(txt)
    t_strtod = time.perf_counter() - t0
    t0 = time.perf_counter()
    xp, yp, zp, tp, used = ac4.parse_frame(txt, m)
    t_parse = time.perf_counter() - t0
    assert used == len(txt) and np.array_equal(xp, xs) and np.array_equal(tp, ts) and np.array_equal(zp, zs)
    # the same text parsed on the GPU (ac4_textframes: the chunk copied from
    # pinned memory, line ends found, one line per thread); timed from the
    # text in the pinned chunk to the parsed frame on the device
    tf = ac4.TextFrames(m)
    ctypes.memmove(tf.buf, txt, len(txt))
    nfr, usedp = ctypes.c_int(), ctypes.c_longlong()
    ac4.lib.ac4_textframes_parse(tf.h, len(txt), 1, ctypes.byref(nfr), ctypes.byref(usedp))  # first launches
    t0 = time.perf_counter()
    rc = ac4.lib.ac4_textframes_parse(tf.h, len(txt), 1, ctypes.byref(nfr), ctypes.byref(usedp))
    t_gparse = time.perf_counter() - t0
    assert rc == 0 and nfr.value == 1 and usedp.value == len(txt)
    xg, yg, zg, tg = tf.frame(0)
    tf.close()
    assert np.array_equal(xg, xs) and np.array_equal(yg, ys) and np.array_equal(zg, zs) and np.array_equal(tg, ts)
    # the same frame as the unwrapped positions of an MD handle, formatted on
    # the device (ac4_md_format_frame: what bin/run writes per frame)
    md = ac4.MD([ac4.ReplicaParams(m, 1.2, 0.6, 1.5, -1, 0, 1)])
    w = [rng.random(m) for _ in range(3)]
    zero = np.zeros(m)
    md.set_state(0, w[0], w[1], w[2], x, y, z, zero, zero, zero, t, np.full(m, -1, dtype=np.int32))
    assert md.format_frame(0) == txt
    md.sync()
    t0 = time.perf_counter()
    dev = md.format_frame(0)
    t_dev = time.perf_counter() - t0
    md.close()
    assert dev == txt
    out["trajectory_io"] = {"frame_bytes": len(txt), "format_s": t_fmt, "sprintf_s": t_printf,
                            "device_format_s": t_dev, "parse_s": t_parse, "strtod_s": t_strtod,
                            "device_parse_s": t_gparse,
                            "format_GBps": len(txt) / t_fmt / 1e9, "device_format_GBps": len(txt) / t_dev / 1e9,
                            "parse_GBps": len(txt) / t_parse / 1e9, "device_parse_GBps": len(txt) / t_gparse / 1e9}
    log("rows: frame text %.1f MB: format %.3f s on the host, %.4f s on the device (sprintf %.3f s), "
        "parse %.3f s on the host, %.4f s on the device (strtod %.3f s)"
        % (len(txt) / 1e6, t_fmt, t_dev, t_printf, t_parse, t_gparse, t_strtod))
    # trajectory
    frames = trajgen.case_frames(n_origins, gap)
    X, Y, Z, _ = trajgen.make_frames(n, frames, seed, 0.06, 0.5)
    L = (n / 1.2) ** (1.0 / 3.0)
    act = np.zeros(n, dtype=np.int32)
    act[np.arange(0, n, 10)] = 1
    fr = ac4.Frames(n, L, len(frames))
    for k_, st in enumerate(frames):
        fr.push(X[k_], Y[k_], Z[k_], st)
    pc, _, _, _ = OL.walk(np.array(frames, dtype=np.int32), gap, 1)
    work = {"N": n, "frames": len(frames), "origins": n_origins, "gap_steps": gap, "pairs": len(pc),
            "active": int(act.sum())}
    out["workload"] = work
    log("rows: N=%d, %d frames, %d (origin, lag) pairs" % (n, len(frames), len(pc)))
    with tempfile.TemporaryDirectory() as tmp:
        cuts = OL.block_scales(L)
        fr.block(gap, cuts[0], os.path.join(tmp, "warm"))  # first call: module load
        t0 = time.perf_counter()
        fr.block(gap, cuts[0], os.path.join(tmp, "g1"))
        t_g1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        for c in cuts:
            fr.block(gap, c, os.path.join(tmp, "g"))
        t_gall = time.perf_counter() - t0
        t0 = time.perf_counter()
        OL.block_all(n, L, X, Y, Z, frames, gap, os.path.join(tmp, "o"), cuts=cuts[:1])
        t_o1 = time.perf_counter() - t0
        for d in ("MSD", "Relax"):
            for f_ in os.listdir(os.path.join(tmp, "o", "BlockAnalysis", d)):
                a_ = open(os.path.join(tmp, "o", "BlockAnalysis", d, f_), "rb").read()
                assert a_ == open(os.path.join(tmp, "g1", "BlockAnalysis", d, f_), "rb").read(), f_
        out["block"] = {"gpu_one_edge_s": t_g1, "oracle_one_edge_s": t_o1, "speedup": t_o1 / t_g1,
                        "gpu_17_edges_s": t_gall, "files_identical": True}
        log("rows: block analysis, edge L/3: GPU %.3f s, oracle %.2f s; 17 edges on the GPU %.2f s"
            % (t_g1, t_o1, t_gall))
        tau = 0.005 * gap * 2
        t0 = time.perf_counter()
        hg, fn, no = fr.guu(act, gap, tau, path=os.path.join(tmp, "guu_g.dat"))
        t_gg = time.perf_counter() - t0
        t0 = time.perf_counter()
        ho, fno, noo = OL.guu(n, L, X, Y, Z, act, frames, gap, tau, path=os.path.join(tmp, "guu_o.dat"))
        t_go = time.perf_counter() - t0
        assert (fn, no) == (fno, noo) and np.array_equal(hg[0], ho[0])
        out["guu"] = {"gpu_s": t_gg, "oracle_s": t_go, "speedup": t_go / t_gg, "lag_steps": fn, "origins": no,
                      "active_pairs_per_origin": int(act.sum() * (act.sum() - 1) // 2),
                      "max_rel_dev_weighted": float(np.max(np.abs(hg[1:] - ho[1:]) / (np.abs(ho[1:]) + 1e-300)))}
        log("rows: g_uu, lag %d steps, %d origins: GPU %.3f s, oracle %.2f s" % (fn, no, t_gg, t_go))
        t0 = time.perf_counter()
        cg = fr.overlap_cage(gap, path=os.path.join(tmp, "cage_g.dat"))
        t_cg = time.perf_counter() - t0
        t0 = time.perf_counter()
        co = OL.overlap_cage(n, L, X, Y, Z, frames, gap, path=os.path.join(tmp, "cage_o.dat"))
        t_co = time.perf_counter() - t0
        assert np.array_equal(cg, co)
        assert open(os.path.join(tmp, "cage_g.dat"), "rb").read() == open(os.path.join(tmp, "cage_o.dat"),
                                                                           "rb").read()
        out["cage_overlap"] = {"gpu_s": t_cg, "oracle_s": t_co, "speedup": t_co / t_cg, "files_identical": True}
        log("rows: cage overlap: GPU %.3f s, oracle %.2f s" % (t_cg, t_co))
    fr.close()
    out["cores"] = 1
    out["kind"] = "port"
    out["reference_programs"] = reference_programs(ac4, n, frames, X, Y, Z, act)
    return out


def reference_programs(ac4, n, frames, X, Y, Z, act):
    """The reference's own Analysis_Block program and its Analysis overlap
    and MSD (built from the reference sources into oracle/_ref/ when those
    were present at build time) against the drop-in executables on the GPU,
    on the same job directory: time to completion, text parsing included,
    and the output files compared byte for byte."""
    import shutil
    import subprocess
    import tempfile
    ref = os.path.join(ROOT, "oracle", "_ref")
    blk_ref, ana_ref = os.path.join(ref, "analyse_block_ref"), os.path.join(ref, "analyse_ref")
    blk_ours, ana_ours = (os.path.join(ROOT, "activechi4_amd", "bin", e) for e in ("analyse_block", "analyse"))
    if not (os.path.exists(blk_ref) and os.path.exists(ana_ref)) or n != 50000:
        return None  # the reference programs fix N = 50000 at compile time
    res = {"kind": "reference", "cores": 1, "sample": "N=%d, %d frames of %%.5lf text" % (n, len(frames))}
    with tempfile.TemporaryDirectory() as tmp:
        run = os.path.join(tmp, "run")
        job = os.path.join(tmp, "0.800", "data_000")
        os.makedirs(run)
        os.makedirs(job)
        with open(os.path.join(run, "para_file"), "w") as f:
            f.write("DENSITY\tT\tP\tDIAMETER\tnb\ttau_alpha\n1.2 .80000    0.1\t0.3\t1\t.0\n"
                    "start_ensemble\tend_ensemble\n0\t0\n")
        with open(os.path.join(job, "time_index_file.dat"), "w") as f:
            f.write("total timesteps equals %d\n" % len(frames) + "".join("%d\n" % v for v in frames))
        with open(os.path.join(job, "ActiveFileIndex.dat"), "w") as f:
            f.write("".join("%d\n" % i for i in np.nonzero(act)[0]))
        t0 = time.perf_counter()
        ty = np.zeros(n, dtype=np.int32)
        ty[: n // 5] = 1
        with open(os.path.join(job, "data_coor.out"), "wb") as f:
            for k in range(len(frames)):
                f.write(ac4.format_frame(X[k], Y[k], Z[k], ty))
        res["write_text_s"] = time.perf_counter() - t0
        res["text_MB"] = os.path.getsize(os.path.join(job, "data_coor.out")) / 1e6
        log("rows: reference programs on %.0f MB of frame text" % res["text_MB"])

        def timed(cmd, **kw):
            t = time.perf_counter()
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=900, **kw)
            return time.perf_counter() - t

        def tree(root):
            return {os.path.relpath(os.path.join(d, f_), root): open(os.path.join(d, f_), "rb").read()
                    for d, _, fs in os.walk(root) for f_ in fs}
        ba = os.path.join(job, "BlockAnalysis")
        t_ref = timed([blk_ref, "0", "0.8", "0"], cwd=run)
        want = tree(ba)
        shutil.rmtree(ba)
        t_gpu = timed([blk_ours, "0", "0.8", "0", str(n)], cwd=run)
        got = tree(ba)
        res["analysis_block"] = {"reference_s": t_ref, "gpu_s": t_gpu, "speedup": t_ref / t_gpu,
                                 "files": len(want), "files_identical": got == want}
        log("rows: Analysis_Block: reference %.1f s, drop-in on the GPU %.2f s, %d files identical: %s"
            % (t_ref, t_gpu, len(want), got == want))
        t_ref = timed([ana_ref, job, "mo"])
        want = {f_: open(os.path.join(job, f_), "rb").read() for f_ in ("msd.dat", "relax_a_0.30.dat")}
        t_gpu = timed([ana_ours, "0", "0.8", "1.0", str(n)], env=dict(os.environ, AC4_DIR=job))
        a = np.loadtxt(os.path.join(job, "msd.dat"))
        b = np.loadtxt(__import__("io").BytesIO(want["msd.dat"]))
        dev = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) if a.shape == b.shape else None
        res["analysis_msd_overlap"] = {"reference_s": t_ref, "gpu_s": t_gpu, "speedup": t_ref / t_gpu,
                                       "gpu_also_computes": "S4qt_a_0.30.dat",
                                       "relax_identical": open(os.path.join(job, "relax_a_0.30.dat"), "rb").read()
                                       == want["relax_a_0.30.dat"], "msd_max_rel_dev": dev}
        log("rows: Analysis MSD + overlap: reference %.1f s, drop-in on the GPU (also S4) %.2f s"
            % (t_ref, t_gpu))
    return res


def s4_row(ac4, n=50000, steps=10000, seed=4321, eager=False):
    """S4(q, t) of Analysis/S4Qt.c (the 162 GetQ vectors, 573 with their
    permutations, weighted by the overlap a=0.3) in situ on the GPU, at the
    reference Analysis' default N=50000 and its frame schedule for 1e4 steps
    (200 origins): the MD run pushes every frame of time_index_file.dat into
    the streaming analysis as it is reached (MSD and overlap beside S4).
    Reports the (origin, lag) pairs, the sincos evaluations the device makes
    (576 columns, the 573 permuted vectors padded to whole tiles, per
    particle per origin) and those a per-pair evaluation would make (573 per
    overlapping particle per pair); the S4 kernel times come from a
    kernel-trace profile of this leg."""
    default_shape(ac4)
    p = ac4.ReplicaParams(n, 1.2, 0.6, 1.5, -1, 200, seed)
    md = start(ac4, [p], ac4.FP32, 300)
    n_origins = 200
    fidx = [int(v) for v in ac4.schedule(steps, n_origins)[0]]
    L_ana = (n / 1.2) ** (1.0 / 3.0)  # Analysis/main.c:15
    an = ac4.Analysis(n, L_ana, fidx, obs=ac4.OBS_MSD | ac4.OBS_OVERLAP | ac4.OBS_S4)
    md.sync()
    t0 = time.perf_counter()
    t = 0
    for f, fi in enumerate(fidx):
        if fi > t:
            md.step(fi - t, graph=not eager)
            t = fi
        an.push_md(f, md, 0)
    an.finalize()
    wall = time.perf_counter() - t0
    _, masks = an.pairs(ac4.OBS_S4)
    md.close()
    n_ov = int(np.sum(masks))
    return {"row": "S4(q,t) in situ, reference schedule", "n": n, "frames": len(fidx), "origins": n_origins,
            "pairs": int(len(masks)), "overlapping_particle_pairs": n_ov, "sincos": 576 * n * n_origins,
            "sincos_per_pair_form": 573 * n_ov, "wall_s_md_and_analysis": wall,
            "launch": "eager" if eager else "graph"}


def pmc_probe(ac4):
    """For hardware-counter runs (rocprofv3 --pmc): the batched workload with
    eager launches (every dispatch is counted), then three device-to-device
    copies of 256 MiB that calibrate FETCH_SIZE/WRITE_SIZE against a known
    byte count."""
    import ctypes
    ps = ensemble_params(ac4, 64)
    md = start(ac4, ps, ac4.FP32, 100)
    md.step(40)
    md.sync()
    md.close()
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(256 << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    assert hip.hipMemset(a, 1, n) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("pmc probe done")


def pmc_traffic():
    """HBM bytes per k_force launch from the committed rocprofv3 PMC summary."""
    try:
        txt = open(PMC_FILE).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    return float(m.group(1)) if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--replicas", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=300)
    ap.add_argument("--relax", type=int, default=300)
    ap.add_argument("--serve-jobs", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=16.0, help="job arrivals per second")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--groups", type=int, default=0,
                    help="launch groups of the batched ensemble (0: the autotuner's choice, 1 without tuning)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-chi4", action="store_true", help="skip the untimed fixed-length chi4 leg")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="short eager run + 256 MiB copies, for rocprofv3 --pmc (prints no bench line)")
    ap.add_argument("--cfg5", action="store_true",
                    help="the N=10^6 box with frames in HBM and S4 on |q| shells (its own JSON line)")
    ap.add_argument("--s4", action="store_true",
                    help="S4(q,t) in situ at N=50000 on the reference schedule (its own JSON line; with --eager "
                         "for kernel-trace profiles)")
    ap.add_argument("--rows", action="store_true",
                    help="the analysis rows (frame text I/O, block analysis, g_uu, cage overlap), GPU vs oracle")
    ap.add_argument("--only-batched", action="store_true",
                    help="batched ensemble only (no plain, serving or CPU legs), e.g. under a profiler")
    ap.add_argument("--tpp", type=int, default=0,
                    help="force lanes per particle for --no-tune and --pmc-probe runs (0: automatic), e.g. to "
                         "profile the shape the autotuner chose")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="Verlet skin for --no-tune and --pmc-probe runs (0: the reference's 0.35), e.g. to "
                         "profile the shape the autotuner chose")
    ap.add_argument("--eager", action="store_true",
                    help="timed steps as eager launches instead of graph replays (for kernel-trace profiles)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import activechi4_amd as ac4
    ac4.require_gpu()
    if a.tpp:
        ac4.set_tuning(a.tpp, 3)
    if a.skin:
        ac4.set_skin(a.skin)
    if a.pmc_probe:
        pmc_probe(ac4)
        return
    if a.cfg5:
        print(json.dumps(cfg5(ac4, eager=a.eager)), flush=True)
        return
    if a.rows:
        print(json.dumps(rows(ac4)), flush=True)
        return
    if a.s4:
        print(json.dumps(s4_row(ac4, eager=a.eager)), flush=True)
        return

    ps = ensemble_params(ac4, a.replicas)
    ntot = sum(p.n for p in ps)
    tune = None
    groups = a.groups
    if not a.no_tune:
        best, table = ac4.autotune(ps, ac4.FP32, relax=a.relax, steps=200)
        groups = best[3] if a.groups == 0 else a.groups
        tune = {"tpp": best[0], "nlist_mode": best[1], "skin": best[2], "launch_groups": best[3],
                "candidates": [[int(t), int(m), float(sk), int(g), float(v)] for t, m, sk, g, v in table]}
        log("autotune: tpp=%d nlist_mode=%d skin=%.2f launch groups=%d" % best)
    groups = max(1, groups)
    # replicas per launch: the autotuner's launch groups, one handle and
    # stream each, stepped together (ac4_md_step_graph_multi)
    gof = ac4.group_plan(ps, groups)
    gps = [[p for p, g in zip(ps, gof) if g == k] for k in range(groups)]
    mds = [start(ac4, gp, ac4.FP32, a.relax) for gp in gps]
    sample_every = CHI4_OV[0]
    for md in mds:
        md.overlap_enable(*CHI4_OV)
    log("%d replicas, %d particles in %d launch group(s) %s; relaxed %d steps, warmup %d steps"
        % (a.replicas, ntot, groups, [sum(p.n for p in gp) for gp in gps], a.relax, a.warmup))
    graph = not a.eager

    def advance(k):
        if graph:
            ac4.step_multi(mds, k)
        else:
            for md in mds:
                md.step(k)

    def sync():
        for md in mds:
            md.sync()

    advance(a.warmup)  # chain steps before the first origin
    sync()
    # step counters restart at the origin of the accumulators: the first
    # sample is taken at the first multiple of sample_every
    done = a.warmup
    if done % sample_every:
        advance(sample_every - done % sample_every)
    sync()
    t0 = time.perf_counter()
    left = a.steps
    for md in mds:
        md.overlap_sample()
    while left > 0:
        k = min(sample_every, left)
        advance(k)
        left -= k
        if k == sample_every:
            for md in mds:
                md.overlap_sample()
    sync()
    t_batched = time.perf_counter() - t0
    value = ntot * a.steps / t_batched
    log("batched %.3f s for %d steps" % (t_batched, a.steps))
    rebuilds = sum(md.rebuilds(r) for md, gp in zip(mds, gps) for r in range(len(gp)))
    for md in mds:
        md.close()
    # per-kernel HIP-event times (eager, on the library's stream) of one
    # handle holding the whole ensemble: one launch of each kernel over all
    # particles, the kernels the launch groups run on their shares
    md = start(ac4, ps, ac4.FP32, a.relax)
    md.step(200, graph=graph)
    kt = md.kernel_times(200)
    md.close()

    if a.only_batched:
        print(json.dumps({"metric": "particle-steps/s (batched only)", "value": value, "unit": "particle-steps/s",
                          "steps": a.steps, "launch": "graph" if graph else "eager",
                          "kernel_ms_per_step": kt}), flush=True)
        return
    chi4 = None
    if not a.no_chi4:
        chi4 = chi4_compare(ac4, ps, a.relax)
        log("chi4: max |peak dev| FP32 vs FP64 %.3f, %d/%d points within 2 sigma"
            % (chi4["max_abs_peak_rel_dev"], chi4["points_within_2sigma"], chi4["points"]))
    plain_value, plain_ms = plain_sequential(ac4, ps, a.relax, a.plain_steps)
    log("plain sequential %.3e particle-steps/s" % plain_value)
    serve = None
    if a.serve_jobs > 0:
        serve = serving(ac4, a.serve_jobs, a.serve_rate)
        log("serving p50 %.2f s p99 %.2f s" % (serve["p50_latency_s"], serve["p99_latency_s"]))

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
        "data": "synthetic (seeded lattice starts, relaxed on the GPU)",
        "config": {"workload": "batched ensemble: %d replicas KA 80:20 rho=1.2 T=0.5, N logU[1000,5000] "
                               "(sum %d), f0 x tau_p grid, FP32 pair forces + FP64 state, hipGraph, "
                               "in-situ Q(t) a=0.3 every 100 steps" % (a.replicas, ntot),
                   "replicas": a.replicas, "particles": ntot, "forces": "fp32", "positions": "fp64",
                   "launch_groups": groups, "particles_per_group": [sum(p.n for p in gp) for gp in gps]},
        "plain_single_run": {"value": plain_value, "replicas_run_sequentially": a.replicas, "dtype": "fp64",
                             "launch": "eager", "steps_each": a.plain_steps,
                             "ms_per_step_all_replicas": plain_ms},
        "speedup_vs_plain": value / plain_value,
        "autotune": tune,
        "chi4": chi4,
        "serving": serve,
        "kernel_ms_per_step": kt,
        "rebuilds_total": rebuilds,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": pmc_traffic(),
                     "kernel": "k_force<fp32,kick>", "bytes_per_particle": FORCE_BYTES_PER_PARTICLE,
                     "particles_per_launch": ntot,
                     "note": "one k_force launch over the whole ensemble at the timed run's kernel shape "
                             "(HIP events, eager; the per-replica sums and chain update, k_force_end, are "
                             "timed apart); the timed run splits the same particles into %d concurrent "
                             "launch groups" % groups,
                     "whole_step": {"bytes_per_particle_step": STEP_BYTES_PER_PARTICLE,
                                    "achieved": value * STEP_BYTES_PER_PARTICLE / 1e9,
                                    "frac": value * STEP_BYTES_PER_PARTICLE / 1e9 / HBM_PEAK_GBS}},
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(ac4)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
