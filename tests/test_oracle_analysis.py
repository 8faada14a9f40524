"""Pins of the Analysis oracle: byte-identical outputs against the reference
Analysis program on the golden trajectories, plus closed forms."""
import filecmp
import json
import os

import numpy as np
import pytest

import oracle_lib as OL
import trajgen

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = ["msd.dat", "relax_a_0.30.dat", "S4qt_a_0.30.dat", "Fsqt_a_0.30.dat"]


def golden_case(name):
    with open(os.path.join(GOLD, name, "case.json")) as f:
        meta = json.load(f)
    frames = trajgen.case_frames(meta["n_origins"], meta["gap"])
    assert frames == meta["frames"]
    X, Y, Z, T = trajgen.make_frames(meta["n"], frames, meta["seed"], meta["step_sigma"], meta["wild"])
    return meta, frames, X, Y, Z, T


@pytest.mark.parametrize("name", [c[0] for c in trajgen.CASES])
def test_oracle_matches_reference_bytes(oracle, tmp_path, name):
    meta, frames, X, Y, Z, T = golden_case(name)
    job = str(tmp_path / "job")
    trajgen.write_job(job, frames, X, Y, Z, T)
    assert oracle.or_analyse_dir(job.encode(), meta["n"], 1, 1, 1, 1) == 0
    for f in FILES:
        assert filecmp.cmp(os.path.join(job, f), os.path.join(GOLD, name, f), shallow=False), f


@pytest.mark.parametrize("name,gap", [("walk_g20", 20), ("wild_g24", 24)])
def test_origin_gap_golden(oracle, ac4, name, gap):
    meta, frames = json.load(open(os.path.join(GOLD, name, "case.json"))), None
    f = np.array(meta["frames"], dtype=np.int32)
    assert oracle.or_origin_gap(OL.ip(f), len(f)) == gap
    assert ac4.origin_gap(f) == gap


def small_traj(n=300, frames=None, seed=0):
    frames = frames or trajgen.case_frames(3, 30)
    return frames, n


def test_msd_ballistic_closed_form(oracle):
    n = 400
    frames = trajgen.case_frames(3, 30)
    L = OL.load().or_ana_box_length(n)
    rng = np.random.default_rng(2)
    v = rng.normal(0, 1e-3, (3, n))
    x0 = rng.random((3, n))
    t = np.array(frames, dtype=float)
    X, Y, Z = (x0[k][None, :] + v[k][None, :] * t[:, None] for k in range(3))
    T = np.zeros((len(frames), n), dtype=np.int32)
    r = OL.msd(n, L, X, Y, Z, T, frames, 30)
    v2 = (L * L * (v ** 2).sum(0)).mean()
    v4 = (L ** 4 * (v ** 4).sum(0)).mean()
    r4 = ((L * L * (v ** 2).sum(0)) ** 2).mean()
    for k in range(1, len(r["counters"])):
        if r["counters"][k]:
            lag = r["delta_t"][k] / 0.005
            assert r["m0"][k] / r["counters"][k] == pytest.approx(v2 * lag ** 2, rel=1e-9)
            assert r["m1"][k] / r["counters"][k] == pytest.approx(v4 * lag ** 4, rel=1e-9)
            assert r["m2"][k] / r["counters"][k] == pytest.approx(r4 * lag ** 4, rel=1e-9)


def test_overlap_static_and_full_shift(oracle):
    n = 300
    frames = trajgen.case_frames(3, 30)
    L = OL.load().or_ana_box_length(n)
    rng = np.random.default_rng(3)
    x0 = rng.random((3, n)) * 3 - 1  # unwrapped, outside [0,1) too
    X, Y, Z = (np.repeat(x0[k][None, :], len(frames), 0) for k in range(3))
    T = np.zeros((len(frames), n), dtype=np.int32)
    r = OL.overlap(n, L, X, Y, Z, T, frames, 30)
    m = r["counters"] > 0
    assert np.all(r["pairs"] == n)
    assert np.allclose(r["c"][m] / r["counters"][m], 1.0) and np.allclose(r["x4"][m] / r["counters"][m] - 1.0, 0)
    # a displacement of exactly one box length is the minimum image of zero
    # (overlap.c:67-78 corrects once), two box lengths are not
    X2 = X + (np.arange(len(frames)) % 2)[:, None]
    r2 = OL.overlap(n, L, X2, Y, Z, T, frames, 30)
    assert np.all(r2["pairs"] == n)
    X3 = X + 2.0 * (np.arange(len(frames)) % 2)[:, None]
    r3 = OL.overlap(n, L, X3, Y, Z, T, frames, 30)
    odd = np.array([(a - b) % 2 for a, b in zip(*OL.walk(frames, 30, 1)[:2])])
    assert np.array_equal(r3["pairs"], np.where(odd == 1, 0, n))


def test_fs_uniform_shift_closed_form(oracle):
    n = 200
    frames = trajgen.case_frames(3, 30)
    L = OL.load().or_ana_box_length(n)
    rng = np.random.default_rng(4)
    x0 = rng.random((3, n))
    d = np.array([0.013, -0.021, 0.008])  # box units per frame step
    t = np.array(frames, dtype=float)
    X, Y, Z = (x0[k][None, :] + d[k] * t[:, None] for k in range(3))
    T = np.zeros((len(frames), n), dtype=np.int32)
    r = OL.phase("fs", n, L, X, Y, Z, T, frames, 30)
    vx, vy, vz, q = (np.zeros(170, dtype=np.int32) for _ in range(4))
    OL.load().or_getq(OL.ip(vx), OL.ip(vy), OL.ip(vz), OL.ip(q))
    import itertools
    k = 5
    lag = r["delta_t"][k] / 0.005
    disp = d * lag
    disp = np.where(disp >= 0.5, disp - 1, np.where(disp < -0.5, disp + 1, disp))
    for qi in range(162):
        perms = set(itertools.permutations((vx[qi], vy[qi], vz[qi])))
        want = np.mean([np.cos(2 * np.pi * np.dot(p, disp)) for p in perms])
        assert r["cs"][k, qi] / r["counters"][k] / n == pytest.approx(want, abs=1e-10)


def test_s4_lattice_and_zero_overlap(oracle):
    tsite = 6
    n = tsite ** 3
    frames = trajgen.case_frames(3, 30)
    L = OL.load().or_ana_box_length(n)
    g = np.arange(n)
    x0 = np.stack([g % tsite, (g // tsite) % tsite, g // tsite ** 2]) / tsite
    X, Y, Z = (np.repeat(x0[k][None, :], len(frames), 0) for k in range(3))
    T = np.zeros((len(frames), n), dtype=np.int32)
    r = OL.phase("s4", n, L, X, Y, Z, T, frames, 30)
    vx, vy, vz, q = (np.zeros(170, dtype=np.int32) for _ in range(4))
    OL.load().or_getq(OL.ip(vx), OL.ip(vy), OL.ip(vz), OL.ip(q))
    k = 3
    for qi in range(162):
        on = all(v % tsite == 0 for v in (vx[qi], vy[qi], vz[qi]))
        assert r["cs"][k, qi] / r["counters"][k] == pytest.approx(n if on else 0.0, abs=1e-9)
    assert np.all(r["mask"] == n)
    # every particle of the non-origin frames moved by 0.5 sigma: those pairs
    # have no overlapping particle and contribute nothing
    is_org = np.array([f % 30 == 0 for f in frames])
    X2 = X + (~is_org)[:, None] * (0.5 / L)
    r2 = OL.phase("s4", n, L, X2, Y, Z, T, frames, 30)
    oth = OL.walk(frames, 30, 1)[1]
    assert np.array_equal(r2["mask"], np.where(is_org[oth], n, 0))
    assert np.all(r2["pairs"][~is_org[oth]] == 0)


def test_getq_table(oracle):
    vx, vy, vz, q = (np.zeros(170, dtype=np.int32) for _ in range(4))
    assert oracle.or_getq(OL.ip(vx), OL.ip(vy), OL.ip(vz), OL.ip(q)) == 162
    v = np.stack([vx[:162], vy[:162], vz[:162]], 1)
    assert np.array_equal(q[:162], (v ** 2).sum(1))
    assert np.all(np.diff(q[:162]) >= 0)
    dg = [6 if len({a, b, c}) == 3 else (1 if a == b == c else 3) for a, b, c in v]
    assert (dg.count(1), dg.count(3), dg.count(6)) == (15, 108, 39)
    assert sum(dg) == 573
    # the set is fixed by S4Qt.c:299-323; the sort only orders it
    want = {(1, 0, 0)} | {(0, k, i) for k in (1, 2) for i in range(k, 22)} | \
        {(i, j, j) for i in range(1, 16) for j in range(1, i + 1)}
    assert set(map(tuple, v.tolist())) == want


def test_chi4_ensemble_closed_forms(oracle):
    """or_chi4_ensemble from per-lag sums against the variance of the
    per-origin samples the sums were made of (numpy), the reference's one-run
    X4 (overlap.c:116-118, no factor N) for a single replica, and zero for
    identical samples."""
    rng = np.random.default_rng(5)
    nrep, nor, nlag = 6, 20, 4
    q = rng.random((nrep, nor, nlag))
    nr = np.array([1000, 2000, 1500, 1200, 3000, 2500], dtype=np.int32)
    cnt = np.full((nrep, nlag), nor, dtype=np.int32)
    chi, qm, c = OL.chi4_ensemble(q.sum(1), (q * q).sum(1), cnt, nr)
    for l in range(nlag):
        assert chi[l] == pytest.approx(nr.mean() * q[:, :, l].var(), rel=1e-12)
        assert qm[l] == pytest.approx(q[:, :, l].mean(), rel=1e-14)
        assert c[l] == nrep * nor
    # one replica: N * X4 of Analysis/overlap.c:116-118
    x4 = (q[0] * q[0]).sum(0) / nor - q[0].sum(0) * q[0].sum(0) / (nor * nor)
    chi1, _, _ = OL.chi4_ensemble(q[:1].sum(1), (q[:1] * q[:1]).sum(1), cnt[:1], nr[:1])
    assert np.array_equal(chi1, nr[0] * x4)
    # uneven origin counts: lags with fewer origins pool only what exists
    cnt2 = cnt.copy()
    cnt2[:, 3] = 0
    chi2, qm2, c2 = OL.chi4_ensemble(q.sum(1), (q * q).sum(1), cnt2, nr)
    assert chi2[3] == 0.0 and qm2[3] == 0.0 and c2[3] == 0
    same = np.repeat(q[:1], nrep, 0)
    chi3, _, _ = OL.chi4_ensemble(same.sum(1), (same * same).sum(1), cnt, nr)
    assert np.all(np.abs(chi3 - nr.mean() * q[0].var(0)) <= 1e-12 * nr.mean())
    flat = np.full((nrep, nor, nlag), 0.625)  # exactly representable: no fluctuation at all
    chi4z, _, _ = OL.chi4_ensemble(flat.sum(1), (flat * flat).sum(1), cnt, nr)
    assert np.all(chi4z == 0.0)

