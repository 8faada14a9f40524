"""Pins of the block-analysis oracle (oracle/oracle_block.c, the restatement of
Analysis_Block): byte-identical output files against the reference program on
the golden trajectories (SHA-256 of every file in
tests/golden/<case>/block/manifest.json, made by make_block_golden.py), and
closed forms for a rigid translation of the whole box."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as OL
import trajgen

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BLOCK_NAMES = [c[0] for c in trajgen.BLOCK_CASES]


def block_case(name):
    """(n, L, frames, gap, X, Y, Z, manifest) of a golden block case."""
    _, n, n_or, gap, seed, sig, wild = [c for c in trajgen.BLOCK_CASES if c[0] == name][0]
    frames = trajgen.case_frames(n_or, gap)
    X, Y, Z, _ = trajgen.make_frames(n, frames, seed, sig, wild)
    with open(os.path.join(GOLD, name, "block", "manifest.json")) as f:
        man = json.load(f)
    assert man["frames"] == len(frames)
    L = (n / 1.2) ** (1.0 / 3.00)  # Analysis_Block/main.c:15, glibc pow
    return n, L, frames, gap, X, Y, Z, man


def check_block_files(out_dir, name, man):
    """Every file under out_dir (a BlockAnalysis directory) has the digest the
    reference's file has; for a mismatch of a kept file, show the first
    differing line."""
    got = {}
    for dirpath, _, files in os.walk(out_dir):
        for fn in files:
            p = os.path.join(dirpath, fn)
            got[os.path.relpath(p, out_dir)] = p
    assert sorted(got) == sorted(man["files"])
    for rel, d in sorted(man["files"].items()):
        data = open(got[rel], "rb").read()
        if hashlib.sha256(data).hexdigest() != d["sha256"]:
            ref = os.path.join(GOLD, name, "block", rel)
            if os.path.exists(ref):
                for k, (a, b) in enumerate(zip(data.decode().splitlines(), open(ref).read().splitlines())):
                    assert a == b, "%s line %d" % (rel, k + 1)
            assert data.count(b"\n") == d["lines"], rel
            pytest.fail("%s differs from the reference (%d lines)" % (rel, d["lines"]))


@pytest.mark.parametrize("name", BLOCK_NAMES)
def test_oracle_block_matches_reference_bytes(oracle, tmp_path, name):
    n, L, frames, gap, X, Y, Z, man = block_case(name)
    assert oracle.or_origin_gap(OL.ip(np.array(frames, dtype=np.int32)), len(frames)) == gap
    OL.block_all(n, L, X, Y, Z, frames, gap, str(tmp_path))
    check_block_files(str(tmp_path / "BlockAnalysis"), name, man)


def read_rows(path):
    return np.array([[float(v) for v in line.split("\t")] for line in open(path).read().splitlines()])


def rigid_shift_case(n=3000, seed=4):
    """Every particle moves by the same vector v*t (box units), crossing the
    periodic boundary; positions unwrapped as the MD writes them."""
    frames = trajgen.case_frames(3, 30)
    rng = np.random.default_rng(seed)
    base = rng.random((3, n))
    v = np.array([1.3e-3, -0.7e-3, 0.4e-3])
    t = np.array(frames, dtype=float)
    X, Y, Z = (base[k][None, :] + v[k] * t[:, None] for k in range(3))
    return frames, X, Y, Z, v


@pytest.mark.parametrize("vh_time", [1500.0, 0.06])
def test_oracle_block_rigid_shift_closed_form(oracle, tmp_path, vh_time):
    n = 3000
    frames, X, Y, Z, v = rigid_shift_case(n)
    L = (n / 1.2) ** (1.0 / 3.0)
    cuts = [L / 3.0, L / 4.0]
    OL.block_all(n, L, X, Y, Z, frames, 30, str(tmp_path), vh_time=vh_time, cuts=cuts)
    ba = tmp_path / "BlockAnalysis"
    for cut in cuts:
        lb = "%.3f" % (L / int(L / cut))
        m = read_rows(ba / "MSD" / ("MSDLb%s.dat" % lb))[1:]
        lag = np.rint(m[:, 0] / 0.005)
        d = lag[:, None] * v[None, :]                       # displacement, box units (all < 1/2)
        r2 = L * L * (d * d).sum(1)
        assert np.allclose(m[:, 1], r2, rtol=1e-9, atol=1e-14)                    # block mean r^2
        assert np.allclose(m[:, 2], r2, rtol=1e-6, atol=1e-12)                    # block COM |d|^2
        assert np.allclose(m[:, 3], L ** 4 * (d ** 4).sum(1), rtol=1e-5, atol=1e-12)
        q = (r2 < 0.09).astype(float)
        for fn in ("RelaxLb%s.dat" % lb, "RelaxTotLb%s.dat" % lb):
            r = read_rows(ba / "Relax" / fn)
            assert np.array_equal(r[:, 0], m[:, 0])
            assert np.array_equal(r[:, 1], q)                # every particle of every block alike
            assert np.abs(r[:, 2]).max() < 1e-12             # no fluctuation
        vh = np.array([float(s) for s in open(ba / ("vanHove%s.dat" % lb)).read().split()])
        k = int(L / cut) ** 3
        nor = 4  # the do-while visits every frame k*30 present: steps 0, 30, 60, 90
        if vh_time == 1500.0:  # no lag within 1000 time units: each origin with itself
            assert vh.size == nor * 3 * k and np.all(vh == 0.0)
        else:  # lag 12 steps (0.06) from every origin (step 102 = 90 + 12 is a frame)
            assert vh.size == nor * 3 * k
            assert np.allclose(vh.reshape(-1, 3), (12 * v * L)[None, :], rtol=1e-9, atol=1e-12)
