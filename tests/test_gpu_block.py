"""GPU block analysis (ac4_frames_block in libac4.so) against the reference:
every output file byte-identical with the reference Analysis_Block program's
on the golden trajectories, and with the oracle on a random walk at N=3000
with a van Hove lag inside the run; the drop-in executable on the reference's
directory layout."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as OL
import trajgen
from test_oracle_block import BLOCK_NAMES, block_case, check_block_files

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def frames_store(gpu, n, L, frames, X, Y, Z):
    fr = gpu.Frames(n, L, len(frames))
    for k, t in enumerate(frames):
        fr.push(X[k], Y[k], Z[k], t)
    return fr


@pytest.mark.parametrize("name", BLOCK_NAMES)
def test_block_files_match_reference(gpu, tmp_path, name):
    n, L, frames, gap, X, Y, Z, man = block_case(name)
    assert gpu.origin_gap(frames) == gap
    fr = frames_store(gpu, n, L, frames, X, Y, Z)
    for cut in OL.block_scales(L):
        fr.block(gap, cut, str(tmp_path))
    fr.close()
    check_block_files(str(tmp_path / "BlockAnalysis"), name, man)


@pytest.mark.parametrize("vh_time", [1500.0, 0.045])
def test_block_files_match_oracle_random_walk(gpu, tmp_path, vh_time):
    n = 3000
    frames = trajgen.case_frames(4, 40)
    X, Y, Z, _ = trajgen.make_frames(n, frames, 91, 0.08, 1.0)
    L = (n / 1.2) ** (1.0 / 3.0)
    cuts = [L / s for s in (3.0, 5.0, 8.0, 12.0)]  # 8^3 and 12^3 blocks leave some empty (NaN rows)
    OL.block_all(n, L, X, Y, Z, frames, 40, str(tmp_path / "or"), vh_time=vh_time, cuts=cuts)
    fr = frames_store(gpu, n, L, frames, X, Y, Z)
    for cut in cuts:
        fr.block(40, cut, str(tmp_path / "gpu"), vh_time=vh_time)
    fr.close()
    want = {}
    for base in ("or", "gpu"):
        root = tmp_path / base / "BlockAnalysis"
        want[base] = {os.path.relpath(os.path.join(d, f), root): open(os.path.join(d, f), "rb").read()
                      for d, _, fs in os.walk(root) for f in fs}
    assert sorted(want["or"]) == sorted(want["gpu"]) and len(want["or"]) == 4 * len(cuts)
    for rel in want["or"]:
        assert want["gpu"][rel] == want["or"][rel], rel
    assert any(b"-nan" in v for v in want["or"].values())  # the empty-block path ran


def test_analyse_block_dropin(gpu, tmp_path):
    name = "wild_g24"
    n, L, frames, gap, X, Y, Z, man = block_case(name)
    run = tmp_path / "run"
    job = tmp_path / "0.800" / "data_000"
    run.mkdir()
    (run / "para_file").write_text("DENSITY\tT\tP\tDIAMETER\tnb\ttau_alpha\n1.2 .80000    0.1\t0.3\t1\t.0\n"
                                   "start_ensemble\tend_ensemble\n0\t0\n")
    T = np.zeros((len(frames), n), dtype=np.int32)
    trajgen.write_job(str(job), frames, X, Y, Z, T)
    exe = os.path.join(ROOT, "activechi4_amd", "bin", "analyse_block")
    out = subprocess.run([exe, "0", "0.8", "0"], cwd=str(run), capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    check_block_files(str(job / "BlockAnalysis"), name, man)
