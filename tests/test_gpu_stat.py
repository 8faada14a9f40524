"""Statistical parity with CPU runs: the library runs the ensemble protocol of
tests/golden/stat_ensemble.py (KA N=1000, T=0.6, f0=1.5, tau_p=1.0; 1e4
production steps, frames every 100 steps) for 64 replicas packed in one
handle, in FP64 and with FP32 pair forces, and its MSD, Q(t) and chi4(t)
agree with the oracle's 32-replica CPU ensemble (tests/golden/stat_cfg1.json,
written by make_stat_golden.py) within the combined error bars.  Chaos rules
out matching trajectories; the ensemble averages must agree.  Also the NVE
energy drift over 1e4 steps (f0 = 0), reported beside the oracle's."""
import json
import os

import numpy as np
import pytest

import stat_ensemble as SE

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "stat_cfg1.json")
R_GPU = 64
K_SIGMA = 4.0  # per lag; 50 correlated lags per curve


def gpu_ensemble(ac4, prec, R=R_GPU, seed0=17000):
    ps = [ac4.ReplicaParams(SE.N, SE.RHO, SE.T, SE.F0, -1, SE.TAU_STEPS, seed0 + r) for r in range(R)]
    md = ac4.MD(ps, precision=prec)
    for r in range(R):
        s = ac4.init_lattice(SE.N, SE.T, seed0 + r)
        md.set_state(r, s["x"], s["y"], s["z"], s["xu"], s["yu"], s["zu"], s["px"], s["py"], s["pz"], s["type"],
                     s["act_tag"], s["kx"], s["ky"], s["kz"])
    md.relax(SE.RAMP, 1.5, 1.0)
    for r in range(R):
        md.init_chain(r, seed=seed0 + r)
    md.step(SE.EQUIL, graph=True)
    frames = [md.export_xu()]
    for _ in range(SE.PROD // SE.SAMPLE):
        md.step(SE.SAMPLE, graph=True)
        frames.append(md.export_xu())
    md.close()
    L = SE.box_length()
    per = [SE.replica_curves(np.array([f[r] for f in frames]), L) for r in range(R)]
    return SE.summarize(per)


def _agree(g, c, key):
    a, b = np.array(g[key]), np.array(c[key])
    ea, eb = np.array(g[key + "_err"]), np.array(c[key + "_err"])
    z = np.abs(a - b) / np.sqrt(ea * ea + eb * eb)
    return z


@pytest.mark.parametrize("prec", [0, 1])
def test_ensemble_statistics_match_cpu_runs(gpu, prec):
    cpu = json.load(open(GOLD))
    assert cpu["lag_steps"] == [l * SE.SAMPLE for l in range(1, SE.NLAG + 1)]
    g = gpu_ensemble(gpu, prec)
    for key in ("msd", "Q", "chi4"):
        z = _agree(g, cpu, key)
        print("%s %s: max z %.2f, mean z %.2f" % (("FP64", "FP32")[prec], key, z.max(), z.mean()))
        assert z.max() < K_SIGMA, (key, z.max(), int(np.argmax(z)))
    zp = abs(g["chi4_peak"] - cpu["chi4_peak"]) / np.hypot(g["chi4_peak_err"], cpu["chi4_peak_err"])
    print("chi4 peak GPU %.3f +- %.3f at %d, CPU %.3f +- %.3f at %d steps"
          % (g["chi4_peak"], g["chi4_peak_err"], g["chi4_peak_lag_steps"], cpu["chi4_peak"], cpu["chi4_peak_err"],
             cpu["chi4_peak_lag_steps"]))
    assert zp < K_SIGMA
    tol_t = K_SIGMA * np.hypot(g["chi4_peak_lag_err_steps"], cpu["chi4_peak_lag_err_steps"]) + SE.SAMPLE
    assert abs(g["chi4_peak_lag_steps"] - cpu["chi4_peak_lag_steps"]) <= tol_t
    # the curves themselves are physical: Q decays from 1, MSD grows
    assert g["Q"][0] > g["Q"][-1] and g["msd"][-1] > g["msd"][0]


def test_nve_drift_beside_cpu(gpu):
    """NVE (f0 = 0) energy drift per particle over 1e4 steps from an
    equilibrated start, FP64 and FP32 forces, beside the oracle's."""
    cpu = json.load(open(GOLD))["nve"]
    out = {}
    for prec in (0, 1):
        p = gpu.ReplicaParams(SE.N, SE.RHO, SE.T, 0.0, -1, 0, 7000)
        md = gpu.MD([p], precision=prec)
        s = gpu.init_lattice(SE.N, SE.T, 7000)
        md.set_state(0, s["x"], s["y"], s["z"], s["xu"], s["yu"], s["zu"], s["px"], s["py"], s["pz"], s["type"],
                     s["act_tag"], s["kx"], s["ky"], s["kz"])
        md.relax(SE.RAMP, 1.5, 1.0)
        md.init_chain(0, seed=7000)
        md.step(SE.EQUIL, graph=True)
        md.set_nve(0)
        md.step(1, graph=True)
        md.sync()
        e0 = md.stats(0)["Heff"] / SE.N
        md.step(9999, graph=True)
        md.sync()
        e1 = md.stats(0)["Heff"] / SE.N
        md.close()
        out[prec] = e1 - e0
    print("NVE drift per particle over 1e4 steps: GPU FP64 %.2e, FP32 %.2e; oracle %.2e"
          % (out[0], out[1], cpu["drift_per_particle"]))
    for v in (out[0], out[1], cpu["drift_per_particle"]):
        assert abs(v) < 1e-3
