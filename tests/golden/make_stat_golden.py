"""CPU ensemble of the reference's CPU-runnable configuration, run with the
oracle (oracle/liboracle.so, the C restatement of the reference NHC step)
on this machine's cores, and its MSD, Q(t) and chi4 curves with error bars
written to tests/golden/stat_cfg1.json (stat_ensemble.py has the protocol).
The GPU test (tests/test_gpu_stat.py) runs the same protocol with the
library and checks agreement within the error bars.

Also the NVE energy drift of the oracle: one replica at f0 = 0 brought to
temperature the same way, then 1e4 velocity-Verlet steps, E = KE + PE per
particle at the start and the end.

Seeds: replica r uses lattice seed 7000 + r and shuffle stream seed 9000 + r.

usage: python tests/golden/make_stat_golden.py [replicas] [workers]
"""
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import stat_ensemble as SE  # noqa: E402


def _kinetic2(s):
    return float((s["px"] ** 2 + s["py"] ** 2 + s["pz"] ** 2).sum())


def oracle_replica(r, f0=SE.F0, prod=SE.PROD, nve_steps=0):
    """One replica through the protocol with the oracle.  Returns its frames
    (or, with nve_steps, the NVE energies per particle at start and end)."""
    import activechi4_amd as ac4
    import oracle_lib as OL
    lib = OL.load()
    L = lib.or_md_box_length(SE.N, SE.RHO)
    s = ac4.init_lattice(SE.N, SE.T, 7000 + r)
    for k in ("fx", "fy", "fz"):
        s[k] = np.zeros(SE.N)
    dof = 3.0 * SE.N - 3.0
    rng = np.random.default_rng(9000 + r)
    nact = len(s["kx"])
    # velocity-rescale ramp: one velocity-Verlet step, then momenta scaled to
    # the ramped temperature (the library's ac4_md_relax does the same)
    zero_chain = np.zeros(9)
    for k in range(SE.RAMP):
        s, _, _ = OL.nhc_steps(s, L, SE.T, f0, 1, zero_chain, fresh=1 if k == 0 else 0, nve=1)
        f = 1.5 + (1.0 - 1.5) * k / (SE.RAMP - 1)
        sc = np.sqrt(SE.T * f * dof / _kinetic2(s))
        for a in ("px", "py", "pz"):
            s[a] = s[a] * sc
    chain = ac4.chain_init(SE.N, SE.T, 7000 + r)

    def shuffle():
        for a in ("kx", "ky", "kz"):
            u = rng.random(nact)
            lib.or_shuffle(nact, OL.ip(s[a]), OL.dp(u))

    def advance(nsteps, step0):
        nonlocal s, chain
        done = 0
        while done < nsteps:
            st = step0 + done
            if f0 > 0 and st % SE.TAU_STEPS == 0:
                shuffle()  # before the step, MDcode/main.c:247-249
            k = min(nsteps - done, SE.TAU_STEPS - st % SE.TAU_STEPS)
            s, chain, out = OL.nhc_steps(s, L, SE.T, f0, k, chain, fresh=0)
            assert np.isfinite(out[4]), "oracle run blew up"
            done += k

    advance(SE.EQUIL, 0)
    if nve_steps:
        _, _, o0 = OL.nhc_steps(s, L, SE.T, 0.0, 1, np.zeros(9), fresh=1, nve=1)
        s1, _, o1 = OL.nhc_steps(s, L, SE.T, 0.0, nve_steps, np.zeros(9), fresh=1, nve=1)
        return o0[4] / SE.N, o1[4] / SE.N
    frames = [np.stack([s["xu"], s["yu"], s["zu"]])]
    for k in range(prod // SE.SAMPLE):
        advance(SE.SAMPLE, SE.EQUIL + k * SE.SAMPLE)
        frames.append(np.stack([s["xu"], s["yu"], s["zu"]]))
    return np.array(frames)


def _curves(r):
    fr = oracle_replica(r)
    return SE.replica_curves(fr, SE.box_length())


def _nve(_):
    return oracle_replica(0, f0=0.0, nve_steps=10000)


def main():
    R = int(sys.argv[1]) if len(sys.argv) > 1 else 32
    W = int(sys.argv[2]) if len(sys.argv) > 2 else min(8, os.cpu_count() or 1)
    t0 = time.time()
    with mp.get_context("spawn").Pool(W) as pool:
        per = pool.map(_curves, range(R))
        e0, e1 = pool.map(_nve, [0])[0]
    out = SE.summarize(per)
    out["config"] = {"N": SE.N, "rho": SE.RHO, "T": SE.T, "f0": SE.F0, "tau_p_steps": SE.TAU_STEPS,
                     "ramp": SE.RAMP, "equil": SE.EQUIL, "prod": SE.PROD, "sample": SE.SAMPLE,
                     "origin_every_frames": SE.ORIGIN_EVERY, "nlag": SE.NLAG, "a": SE.A,
                     "producer": "oracle NHC step (oracle/liboracle.so), tests/golden/make_stat_golden.py"}
    out["nve"] = {"f0": 0.0, "steps": 10000, "E_start_per_particle": e0, "E_end_per_particle": e1,
                  "drift_per_particle": e1 - e0}
    with open(os.path.join(HERE, "stat_cfg1.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("replicas %d, %.0f s; chi4 peak %.3f +- %.3f at %d steps; NVE drift %.2e per particle"
          % (R, time.time() - t0, out["chi4_peak"], out["chi4_peak_err"], out["chi4_peak_lag_steps"], e1 - e0))


if __name__ == "__main__":
    main()
