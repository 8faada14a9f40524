"""GPU parity of the analysis (libac4.so) against the oracle and the golden
reference outputs."""
import filecmp
import json
import os

import numpy as np
import pytest

import oracle_lib as OL
import trajgen
from test_oracle_analysis import golden_case, GOLD

pytestmark = pytest.mark.gpu


def run_product(ac4, n, L, frames, X, Y, Z, T, obs=15, flag=0):
    an = ac4.Analysis(n, L, frames, obs=obs, type_flag=flag)
    for f in range(len(frames)):
        an.push(f, X[f], Y[f], Z[f], T[f])
    an.finalize()
    return an


@pytest.mark.parametrize("wild,flag", [(0.0, 0), (1.5, 0), (0.7, 1)])
def test_pairs_match_oracle(gpu, wild, flag):
    n = 700
    frames = trajgen.case_frames(3, 30)
    X, Y, Z, T = trajgen.make_frames(n, frames, 123, 0.08, wild)
    L = OL.load().or_ana_box_length(n)
    an = run_product(gpu, n, L, frames, X, Y, Z, T, flag=flag)
    assert an.gap == 30
    om = OL.msd(n, L, X, Y, Z, T, frames, 30, flag=flag)
    assert np.allclose(an.pairs(gpu.OBS_MSD), om["pairs"], rtol=1e-12, atol=0)
    oo = OL.overlap(n, L, X, Y, Z, T, frames, 30, flag=flag)
    assert np.array_equal(an.pairs(gpu.OBS_OVERLAP), oo["pairs"])
    os4 = OL.phase("s4", n, L, X, Y, Z, T, frames, 30)
    v, m = an.pairs(gpu.OBS_S4)
    assert np.array_equal(m, os4["mask"])
    assert np.abs(v - os4["pairs"]).max() < 1e-9 * n
    ofs = OL.phase("fs", n, L, X, Y, Z, T, frames, 30, flag=flag)
    assert np.abs(an.pairs(gpu.OBS_FS) - ofs["pairs"]).max() < 1e-9 * n
    # slot accumulation in the reference's origin order
    ov = an.overlap()
    k = len(oo["c"])
    assert np.array_equal(ov["c"][:k][ov["counters"][:k] > 0], oo["c"][oo["counters"] > 0])


def test_s4_origins_with_more_lags_than_one_pass(gpu):
    """S4 is evaluated per origin over all of its lags, in passes of at most
    64 lags.  With the walk factor 1.05 on 1000 consecutive frames and
    origins 400 frames apart the origins have 100, 100 and 76 lags: passes of
    64, 36 and 12 lags (4, 3 and 1 row tiles of 16).  N=300 is not a multiple
    of the 32-particle mask word.  Per-pair sums and overlap counts against
    the oracle."""
    n, gap, factor = 300, 400, 1.05
    frames = list(range(1000))
    X, Y, Z, T = trajgen.make_frames(n, frames, 99, 0.05, 1.0)
    L = OL.load().or_ana_box_length(n)
    an = gpu.Analysis(n, L, frames, gap=gap, factor=factor, obs=gpu.OBS_S4)
    for f in range(len(frames)):
        an.push(f, X[f], Y[f], Z[f], T[f])
    an.finalize()
    os4 = OL.phase("s4", n, L, X, Y, Z, T, frames, gap, factor=factor)
    v, m = an.pairs(gpu.OBS_S4)
    assert len(m) == len(os4["mask"]) > 3 * 64
    assert np.array_equal(m, os4["mask"])
    assert np.abs(v - os4["pairs"]).max() < 1e-9 * n


def _numbers(path):
    return np.array([float(t) for line in open(path) for t in line.split()])


@pytest.mark.parametrize("name", [c[0] for c in trajgen.CASES])
def test_golden_outputs(gpu, tmp_path, name):
    meta, frames, X, Y, Z, T = golden_case(name)
    n = meta["n"]
    L = (n / 1.2) ** (1.0 / 3.0)
    an = run_product(gpu, n, L, frames, X, Y, Z, T)
    d = str(tmp_path)
    an.write(d)
    # integer-count observable: byte-identical with the reference program
    assert filecmp.cmp(os.path.join(d, "relax_a_0.30.dat"), os.path.join(GOLD, name, "relax_a_0.30.dat"),
                       shallow=False)
    for f, tol in (("msd.dat", 1e-12), ("S4qt_a_0.30.dat", 1e-8), ("Fsqt_a_0.30.dat", 1e-8)):
        a, b = _numbers(os.path.join(d, f)), _numbers(os.path.join(GOLD, name, f))
        assert a.shape == b.shape, f
        assert np.all(np.abs(a - b) <= tol * np.maximum(1.0, np.abs(b))), f


def test_insitu_push_equals_host_push(gpu):
    from test_gpu_md import make_relaxed_md, relaxed_state
    n = 1000
    s, L = relaxed_state(gpu, n, 61)
    md = make_relaxed_md(gpu, [(n, 61, 4)], tau=20, prec=0)
    frames = trajgen.case_frames(2, 20)
    La = OL.load().or_ana_box_length(n)
    a1 = gpu.Analysis(n, La, frames)
    a2 = gpu.Analysis(n, La, frames)
    t = 0
    for f, fi in enumerate(frames):
        md.step(fi - t)
        t = fi
        g = md.get_state(0)
        a1.push_md(f, md, 0)
        a2.push(f, g["xu"], g["yu"], g["zu"], s["type"])
    md.sync()
    a1.finalize()
    a2.finalize()
    for o in (gpu.OBS_MSD, gpu.OBS_OVERLAP):
        assert np.array_equal(a1.pairs(o), a2.pairs(o))
    assert np.array_equal(a1.pairs(gpu.OBS_S4)[0], a2.pairs(gpu.OBS_S4)[0])


def test_cfg1_full_schedule_matches_oracle(gpu):
    """The reference's CPU-runnable configuration at full length: KA N=1000,
    T=0.6, f0=1.5, tau_p=1.0 (200 steps), 1e4 steps on the GPU with frames on
    the reference schedule (4449 frames, 200 origins).  The MSD and overlap of
    those frames on the GPU against the oracle: all 4891 (origin, lag)
    overlap counts and the per-slot sums exact, MSD within 1e-12."""
    from test_gpu_md import make_relaxed_md
    n, steps = 1000, 10000
    fidx = [int(v) for v in gpu.schedule(steps, 200)[0]]
    assert len(fidx) == 4449
    md = make_relaxed_md(gpu, [(n, 7, 5)], tau=200, prec=0)
    X, Y, Z = (np.empty((len(fidx), n)) for _ in range(3))
    t = 0
    for f, fi in enumerate(fidx):
        if fi > t:
            md.step(fi - t, graph=True)
            t = fi
        g = md.get_state(0)
        X[f], Y[f], Z[f] = g["xu"], g["yu"], g["zu"]
    md.close()
    T = np.zeros((len(fidx), n), dtype=np.int32)
    L = OL.load().or_ana_box_length(n)
    an = run_product(gpu, n, L, fidx, X, Y, Z, T, obs=gpu.OBS_MSD | gpu.OBS_OVERLAP)
    assert an.gap == 50
    oo = OL.overlap(n, L, X, Y, Z, T, fidx, 50)
    assert len(oo["pairs"]) == 4891
    assert np.array_equal(an.pairs(gpu.OBS_OVERLAP), oo["pairs"])
    ov = an.overlap()
    mg, mo = ov["counters"] > 0, oo["counters"] > 0
    assert np.array_equal(np.nonzero(mg)[0], np.nonzero(mo)[0])
    assert np.array_equal(ov["counters"][mg], oo["counters"][mo])
    assert np.array_equal(ov["c"][mg], oo["c"][mo]) and np.array_equal(ov["x4"][mg], oo["x4"][mo])
    om = OL.msd(n, L, X, Y, Z, T, fidx, 50)
    assert np.allclose(an.pairs(gpu.OBS_MSD), om["pairs"], rtol=1e-12, atol=0)
    q = oo["c"][mo] / oo["counters"][mo]
    assert q[0] > 0.9 and q[-1] < q[0]  # the overlap decays over the run
