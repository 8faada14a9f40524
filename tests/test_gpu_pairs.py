"""GPU pair observables (libac4.so) against the reference and the oracle:
g_uu(r) (ac4_frames_guu) on the golden trajectories against the reference
Analysis program's guuTauAlphaAct.dat (r and pair-count columns exact, the two
weighted columns to rounding: their FP64 sums are accumulated by atomics) and
the oracle's histograms; the cage-relative overlap (ac4_frames_overlap_cage)
byte-identical with the oracle, whose neighbour order it reproduces."""
import os

import numpy as np
import pytest

import oracle_lib as OL
import trajgen
from test_oracle_analysis import GOLD, golden_case
from test_oracle_pairs import act_tags

pytestmark = pytest.mark.gpu


def store(gpu, n, L, frames, X, Y, Z):
    fr = gpu.Frames(n, L, len(frames))
    for k, t in enumerate(frames):
        fr.push(X[k], Y[k], Z[k], t)
    return fr


@pytest.mark.parametrize("name", [c[0] for c in trajgen.CASES])
def test_guu_matches_reference(gpu, tmp_path, name):
    meta, frames, X, Y, Z, _ = golden_case(name)
    n, gap, tau = meta["n"], meta["gap"], meta["guu_tau"]
    L = (n / 1.2) ** (1.0 / 3.0)
    act = act_tags(n)
    fr = store(gpu, n, L, frames, X, Y, Z)
    out = str(tmp_path / "guu.dat")
    h, fn, no = fr.guu(act, gap, tau, path=out)
    fr.close()
    ho, fno, noo = OL.guu(n, L, X, Y, Z, act, frames, gap, tau)
    assert (fn, no) == (fno, noo)
    assert np.array_equal(h[0], ho[0])  # pair counts: exact bin indices
    assert np.allclose(h[1:], ho[1:], rtol=1e-10, atol=1e-9)
    got = [line.split("\t") for line in open(out).read().splitlines()]
    want = [line.split("\t") for line in open(os.path.join(GOLD, name, "guuTauAlphaAct.dat")).read().splitlines()]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g[:2] == w[:2]
        assert np.allclose([float(v) for v in g[2:]], [float(v) for v in w[2:]], rtol=1e-8, atol=0)


@pytest.mark.parametrize("n,seed,wild", [(3000, 21, 1.0), (1500, 22, 0.0)])
def test_cage_overlap_matches_oracle(gpu, tmp_path, n, seed, wild):
    frames = trajgen.case_frames(4, 40)
    X, Y, Z, _ = trajgen.make_frames(n, frames, seed, 0.08, wild)
    L = (n / 1.2) ** (1.0 / 3.0)
    want = OL.overlap_cage(n, L, X, Y, Z, frames, 40, path=str(tmp_path / "or.dat"))
    fr = store(gpu, n, L, frames, X, Y, Z)
    got = fr.overlap_cage(40, path=str(tmp_path / "gpu.dat"))
    fr.close()
    assert np.array_equal(got, want)
    assert open(tmp_path / "gpu.dat", "rb").read() == open(tmp_path / "or.dat", "rb").read()


def test_cage_overlap_rigid_translation(gpu):
    n = 2000
    frames = trajgen.case_frames(3, 30)
    rng = np.random.default_rng(6)
    base = rng.random((3, n))
    u = np.array([4e-3, -3e-3, 2e-3])
    t = np.array(frames, dtype=float)
    X, Y, Z = (base[k][None, :] + u[k] * t[:, None] for k in range(3))
    L = (n / 1.2) ** (1.0 / 3.0)
    fr = store(gpu, n, L, frames, X, Y, Z)
    assert np.all(fr.overlap_cage(30) == n)
    fr.close()


def test_cage_overlap_needs_three_cells(gpu):
    n = 500  # L = 7.46: two cells of 2.85 per edge
    frames = trajgen.case_frames(2, 20)
    X, Y, Z, _ = trajgen.make_frames(n, frames, 1, 0.05, 0.0)
    fr = store(gpu, n, (n / 1.2) ** (1.0 / 3.0), frames, X, Y, Z)
    with pytest.raises(gpu.AC4Error, match="at least 3"):
        fr.overlap_cage(20)
    fr.close()


def test_analyse_dropin_guu_and_cage(gpu, tmp_path):
    """bin/analyse with AC4_GUU=1 and AC4_CAGE=1 on a golden job directory."""
    import subprocess
    name = "walk_g20"
    meta, frames, X, Y, Z, T = golden_case(name)
    n = meta["n"]
    job = tmp_path / "job"
    trajgen.write_job(str(job), frames, X, Y, Z, T)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "activechi4_amd", "bin",
                       "analyse")
    env = dict(os.environ, AC4_DIR=str(job), AC4_GUU="1", AC4_CAGE="1")
    out = subprocess.run([exe, "0", "0.8", repr(meta["guu_tau"]), str(n)], env=env, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    got = [line.split("\t") for line in open(job / "guuTauAlphaAct.dat").read().splitlines()]
    want = [line.split("\t") for line in open(os.path.join(GOLD, name, "guuTauAlphaAct.dat")).read().splitlines()]
    assert len(got) == len(want) and all(g[:2] == w[:2] for g, w in zip(got, want))
    L = (n / 1.2) ** (1.0 / 3.0)
    OL.overlap_cage(n, L, X, Y, Z, frames, meta["gap"], path=str(tmp_path / "cage.dat"))
    assert open(job / "relaxMPICR1_a_0.30.dat", "rb").read() == open(tmp_path / "cage.dat", "rb").read()
