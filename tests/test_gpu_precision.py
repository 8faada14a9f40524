"""Mixed precision against FP64 on the single-large-run configuration
(KA N=10,000, T=0.5, c=0.1, f0=1.0, tau_p=1.0, dt=0.005): NVE trajectory
parity with the oracle, NVE energy drift over 1e4 steps, per-particle FP32
force error on 100 configurations of a run, and the overlap Q(t) of an
ensemble in FP32 against the same ensemble in FP64."""
import numpy as np
import pytest

import oracle_lib as OL
from test_gpu_md import _chain0, make_md
from test_oracle_md import _rel_force_err, lattice_state

pytestmark = pytest.mark.gpu


def test_nve_trajectory_matches_oracle(gpu):
    n, T, f0 = 1000, 0.6, 1.5
    s, L = lattice_state(n, T=T, jitter=0.02, seed=23)
    s, ch, _ = OL.nhc_steps(s, L, T, f0, 200, _chain0(n, T))  # thermalise on the oracle
    s = {k: v for k, v in s.items() if k not in ("fx", "fy", "fz")}
    md = make_md(gpu, [(s, L)], f0=f0, T=T)
    md.set_chain(0, ch)
    md.set_nve(0)
    md.step(60)
    g = md.get_state(0)
    o, och, oout = OL.nhc_steps(s, L, T, f0, 60, ch, nve=1)
    for k in ("xu", "yu", "zu"):
        assert np.abs(g[k] - o[k]).max() < 1e-11
    for k in ("px", "py", "pz"):
        assert np.abs(g[k] - o[k]).max() < 1e-9
    assert np.array_equal(md.get_chain(0), ch)  # the chain is not touched in NVE
    st = md.stats(0)
    assert st["Heff"] == pytest.approx(oout[4], rel=1e-10)  # KE + PE
    assert st["pe"] == pytest.approx(oout[0], rel=1e-10)


def _equilibrated(gpu, prec, n=10000, T=0.5, f0=0.0, seed=5, nhc_steps=1000):
    p = gpu.ReplicaParams(n, 1.2, T, f0, -1, 200, seed)
    md = gpu.MD([p], precision=prec)
    s = gpu.init_lattice(n, T, seed)
    md.set_state(0, s["x"], s["y"], s["z"], s["xu"], s["yu"], s["zu"], s["px"], s["py"], s["pz"], s["type"],
                 s["act_tag"], s["kx"], s["ky"], s["kz"])
    md.relax(300, 1.5, 1.0)
    md.init_chain(0, seed=seed)
    md.step(nhc_steps, graph=True)
    md.sync()
    return md, s


@pytest.mark.parametrize("prec", [0, 1])
def test_nve_energy_drift_10k(gpu, prec):
    """NVE from an equilibrated N=10,000 state (f0 = 0): energy per particle
    over 1e4 steps, sampled every 500 steps."""
    md, _ = _equilibrated(gpu, prec)
    md.set_nve(0)
    md.step(1, graph=True)
    md.sync()
    e0 = md.stats(0)["Heff"]
    es = []
    for _ in range(20):
        md.step(500, graph=True)
        md.sync()
        es.append(md.stats(0)["Heff"])
    n = 10000
    dev = (np.array(es) - e0) / n
    print("NVE %s: drift over 1e4 steps %.3e per particle, max |dE| %.3e per particle, T %.4f"
          % ("FP32 forces" if prec else "FP64", dev[-1], np.abs(dev).max(), md.stats(0)["T"]))
    assert np.abs(dev).max() < 1e-3


def test_fp32_force_error_on_100_configurations(gpu):
    """Configurations every 100 steps of an FP64 N=10,000 active run (f0=1.0,
    tau_p=1.0): FP32 forces against the FP64 forces of the same state, and
    the FP64 forces of five of them against the oracle's."""
    n, T, f0 = 10000, 0.5, 1.0
    md64, s0 = _equilibrated(gpu, gpu.FP64, n=n, T=T, f0=f0, seed=17, nhc_steps=300)
    md32 = gpu.MD([gpu.ReplicaParams(n, 1.2, T, f0, -1, 200, 17)], precision=gpu.FP32)
    L = (n / 1.2) ** 0.33333333
    errs, oracle_errs = [], []
    for k in range(100):
        md64.step(100, graph=True)
        g = md64.get_state(0)
        f64 = md64.forces(0)
        md32.set_state(0, g["x"], g["y"], g["z"], g["xu"], g["yu"], g["zu"], g["px"], g["py"], g["pz"],
                       s0["type"], s0["act_tag"], g["kx"], g["ky"], g["kz"])
        md32.calculate_forces()
        errs.append(_rel_force_err(md32.forces(0), f64))
        if k % 20 == 0:
            st = {k: np.ascontiguousarray(g[k]) for k in ("x", "y", "z", "xu", "yu", "zu")}
            st.update({k: np.ascontiguousarray(g[k], dtype=np.int32) for k in ("kx", "ky", "kz")})
            st.update({k: np.ascontiguousarray(s0[k], dtype=np.int32) for k in ("type", "act_tag")})
            cnt, lst, _ = OL.half_list(st["x"], st["y"], st["z"], L)
            fx, fy, fz, _ = OL.forces_list(st, L, f0, cnt, lst)
            oracle_errs.append(_rel_force_err(f64, (fx, fy, fz)))
    errs = np.array(errs)
    print("FP32 force error over 100 configurations: max %.3e, mean %.3e; FP64 vs oracle max %.3e"
          % (errs.max(), errs.mean(), max(oracle_errs)))
    assert max(oracle_errs) < 1e-12
    assert errs.max() <= 1e-5


def test_fp32_overlap_statistics_match_fp64(gpu):
    """The same seeded ensemble (8 replicas, N=2000, active f0=1.0) run 2e4
    steps in FP32 and in FP64: the replica-mean overlap Q(t) (a=0.3, all
    origins) agrees within the ensemble's error bars at every lag."""
    specs = [gpu.ReplicaParams(2000, 1.2, 0.5, 1.0, -1, 200, 300 + r) for r in range(8)]
    res = {}
    for prec in (gpu.FP64, gpu.FP32):
        md = gpu.MD(specs, precision=prec)
        for r, p in enumerate(specs):
            s = gpu.init_lattice(p.n, p.T, p.seed)
            md.set_state(r, s["x"], s["y"], s["z"], s["xu"], s["yu"], s["zu"], s["px"], s["py"], s["pz"],
                         s["type"], s["act_tag"], s["kx"], s["ky"], s["kz"])
        md.relax(300, 1.5, 1.0)
        for r, p in enumerate(specs):
            md.init_chain(r, seed=p.seed)
        md.step(2000, graph=True)  # leave the ramp behind
        md.sync()
        md.overlap_enable(100, 1000, 11, 100, 0.3)
        md.overlap_sample()
        for _ in range(200):
            md.step(100, graph=True)
            md.overlap_sample()
        md.sync()
        res[prec] = np.array([md.overlap(r)[0][1:] / md.overlap(r)[2][1:] for r in range(len(specs))])
        md.close()
    q64, q32 = res[gpu.FP64], res[gpu.FP32]
    se = np.sqrt(q64.var(axis=0, ddof=1) / len(specs) + q32.var(axis=0, ddof=1) / len(specs))
    d = q32.mean(axis=0) - q64.mean(axis=0)
    print("Q(t) FP32 - FP64: max |dQ| %.4f, max |dQ|/SE %.2f" % (np.abs(d).max(), (np.abs(d) / se).max()))
    assert np.all(np.abs(d) <= 4.0 * se + 1e-3)


def test_forces_at_one_million_particles(gpu):
    """The single large box (N=10^6, L=94.1): the GPU's FP64 forces against the
    oracle's O(N) brute force on 256 sampled particles, FP32 forces against
    FP64 for every particle, and Newton-3 (the pair forces sum to zero)."""
    n, T, f0 = 1_000_000, 0.5, 1.0
    s, L = lattice_state(n, T=T, jitter=0.03, seed=31)
    md64 = make_md(gpu, [(s, L)], f0=f0, T=T, prec=gpu.FP64)
    md64.calculate_forces()
    F = md64.forces(0)
    md64.close()
    idx = np.random.default_rng(4).choice(n, 256, replace=False)
    fo = OL.forces_sample(s, L, f0, idx)
    err64 = _rel_force_err((F[0][idx], F[1][idx], F[2][idx]), (fo[:, 0], fo[:, 1], fo[:, 2]))
    act = s["act_tag"] >= 0
    a = s["act_tag"][act]
    fact = [f0 * np.asarray(s[k], dtype=float)[a].sum() for k in ("kx", "ky", "kz")]
    fscale = np.sqrt((F[0] ** 2 + F[1] ** 2 + F[2] ** 2).mean())
    net = max(abs(F[c].sum() - fact[c]) for c in range(3)) / (fscale * n)
    md32 = make_md(gpu, [(s, L)], f0=f0, T=T, prec=gpu.FP32)
    md32.calculate_forces()
    err32 = _rel_force_err(md32.forces(0), F)
    md32.close()
    print("N=1e6: FP64 vs oracle %.2e (256 particles), FP32 vs FP64 %.2e, net pair force %.1e per particle"
          % (err64, err32, net))
    assert err64 < 1e-12
    assert err32 <= 1e-5
    assert net < 1e-10
