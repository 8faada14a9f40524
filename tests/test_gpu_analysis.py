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
