"""Pins of the pair-observable oracle (oracle/oracle_pairs.c): g_uu(r)
byte-identical against the reference Analysis program's guu (golden
guuTauAlphaAct.dat, tests/golden/make_golden.py); the cage-relative overlap,
which has no runnable reference, against an O(N^2) numpy restatement of its
neighbour sets and counts, and a rigid translation of the box (Q = 1)."""
import filecmp
import os

import numpy as np
import pytest

import oracle_lib as OL
import trajgen
from test_oracle_analysis import GOLD, golden_case


def act_tags(n):
    a = np.zeros(n, dtype=np.int32)
    a[np.arange(0, n, 10)] = 1  # trajgen.write_job's ActiveFileIndex.dat
    return a


@pytest.mark.parametrize("name", [c[0] for c in trajgen.CASES])
def test_oracle_guu_matches_reference_bytes(oracle, tmp_path, name):
    meta, frames, X, Y, Z, T = golden_case(name)
    n = meta["n"]
    L = oracle.or_ana_box_length(n)
    out = str(tmp_path / "guu.dat")
    h, fn, no = OL.guu(n, L, X, Y, Z, act_tags(n), frames, meta["gap"], meta["guu_tau"], path=out)
    assert fn > 0 and no >= 2
    assert filecmp.cmp(out, os.path.join(GOLD, name, "guuTauAlphaAct.dat"), shallow=False)


def test_oracle_guu_rigid_translation(oracle):
    """Every active particle moves by the same u: hist1 = hist2 = hist*|u|^2 L^2
    in every bin, and the bin counts are those of the origin's pair distances."""
    n = 2000
    frames = trajgen.case_frames(3, 30)
    rng = np.random.default_rng(8)
    base = rng.random((3, n))
    u = np.array([2e-3, 1e-3, -1.5e-3])
    t = np.array(frames, dtype=float)
    X, Y, Z = (base[k][None, :] + u[k] * t[:, None] for k in range(3))
    L = (n / 1.2) ** (1.0 / 3.0)
    act = act_tags(n)
    h, fn, no = OL.guu(n, L, X, Y, Z, act, frames, 30, 0.06)
    assert fn == 12 and no == 4
    d2 = (fn * u * L) @ (fn * u * L)
    assert np.allclose(h[1], h[0] * d2, rtol=1e-9, atol=1e-12)
    assert np.allclose(h[2], h[0] * d2, rtol=1e-9, atol=1e-12)
    w = base[:, act == 1]
    dd = w[:, None, :] - w[:, :, None]
    dd = np.where(dd >= 0.5, dd - 1.0, np.where(dd < -0.5, dd + 1.0, dd))
    r = np.sqrt(L * L * (dd * dd).sum(0))[np.triu_indices(w.shape[1], 1)]
    idx = (r / 0.005).astype(int)
    want = np.bincount(idx[idx < h.shape[1]], minlength=h.shape[1]) * 2.0 * no  # same origin geometry each time
    assert np.array_equal(h[0], want)


def brute_cage(x0, y0, z0, L):
    """Neighbour sets within 1.3 (len_check arithmetic), O(N^2)."""
    def mi(d):
        d = np.where(d < -0.5, d + 1.0, d)
        return np.where(d > 0.5, d - 1.0, d)
    dx, dy, dz = (mi(a[:, None] - a[None, :]) for a in (x0, y0, z0))
    near = (dx * dx + dy * dy + dz * dz) * L * L < 1.3 * 1.3
    np.fill_diagonal(near, False)
    return near


def test_oracle_cage_lists_are_the_neighbour_sets(oracle):
    n = 1500
    rng = np.random.default_rng(5)
    x0, y0, z0 = rng.random((3, n))
    L = (n / 1.2) ** (1.0 / 3.0)
    nn, nb = OL.cage_list(x0, y0, z0, L)
    near = brute_cage(x0, y0, z0, L)
    assert np.array_equal(nn, near.sum(1))
    for i in range(n):
        lst = nb[i, :nn[i]]
        assert len(set(lst)) == nn[i] and set(lst) == set(np.nonzero(near[i])[0])


def test_oracle_cage_counts_match_brute_force(oracle):
    n = 1500
    frames = trajgen.case_frames(3, 30)
    X, Y, Z, _ = trajgen.make_frames(n, frames, 12, 0.08, 1.0)
    L = (n / 1.2) ** (1.0 / 3.0)
    counts = OL.overlap_cage(n, L, X, Y, Z, frames, 30)
    pc, po, ps, _ = OL.walk(np.array(frames, dtype=np.int32), 30, 1)
    assert len(counts) == len(pc)

    def wrap(v):
        return v - np.floor(v)
    want = []
    for p in range(len(pc)):
        c, o = pc[p], po[p]
        near = brute_cage(wrap(X[c]), wrap(Y[c]), wrap(Z[c]), L)
        D = [L * np.where(d >= 0.5, d - 1.0, np.where(d < -0.5, d + 1.0, d)) for d in (X[o] - X[c], Y[o] - Y[c],
                                                                                       Z[o] - Z[c])]
        m = near.sum(1)
        cm = [(near * d[None, :]).sum(1) / m for d in D]
        r2 = sum(d * d for d in D) + sum(c_ * c_ for c_ in cm) - 2.0 * sum(d * c_ for d, c_ in zip(D, cm))
        want.append(int((r2 < 0.09).sum()))
    assert np.array_equal(counts, np.array(want))
    assert 0 < counts.min() < n  # both outcomes occur


def test_oracle_cage_rigid_translation(oracle, tmp_path):
    n = 1500
    frames = trajgen.case_frames(3, 30)
    rng = np.random.default_rng(6)
    base = rng.random((3, n))
    u = np.array([4e-3, -3e-3, 2e-3])  # up to 0.6 sigma per step lag: plain overlap would drop to 0
    t = np.array(frames, dtype=float)
    X, Y, Z = (base[k][None, :] + u[k] * t[:, None] for k in range(3))
    L = (n / 1.2) ** (1.0 / 3.0)
    out = str(tmp_path / "cage.dat")
    counts = OL.overlap_cage(n, L, X, Y, Z, frames, 30, path=out)
    assert np.all(counts == n)
    rows = np.loadtxt(out)
    assert np.all(rows[1:, 1] == 1.0) and np.all(rows[1:, 2] == 0.0)
