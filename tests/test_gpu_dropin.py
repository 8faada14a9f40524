"""The drop-in executables end to end on the GPU: bin/run writes the
reference's job directory, bin/analyse reads it back; the oracle (the CPU
restatement of the reference Analysis) analyses the same files."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as OL

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "activechi4_amd", "bin")


def test_run_then_analyse(gpu, tmp_path):
    n = 1000
    with open(tmp_path / "para_file", "w") as f:
        # eqDuration prodDuration frame_step / DENSITY T P / f0 (MDcode/main.c:417-425)
        f.write("eqDuration\tprodDuration\tframe_step\n2.0\t50.0\t500\nDENSITY\tT\tP\n1.20\t0.60\t1.0\nf0\n1.5\n")
    env = dict(os.environ, AC4_SEED="12345")
    r = subprocess.run([os.path.join(BIN, "run"), "1", "0", "0", str(n)], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    job = tmp_path / "0.600" / "data_001"
    idx = open(job / "time_index_file.dat").read().split("\n")
    nframes = int(idx[0].split()[-1])
    assert nframes == 4449 - 0 or nframes == 4448  # App. B: 4449 frames written for 1e4 steps
    lines = sum(1 for _ in open(job / "data_coor.out"))
    assert lines in (nframes * n, (nframes + 1) * n)
    first = open(job / "data_coor.out").readline().split("\t")
    assert len(first) == 4 and len(first[0].split(".")[1]) == 5  # %.5lf\t%.5lf\t%.5lf\t%d
    stuff = [l.split("\t") for l in open(job / "stuff.dat")]
    assert len(stuff) > 10 and all(len(s) == 6 for s in stuff)
    T_inst = np.array([float(s[2]) for s in stuff[-5:]])
    assert np.all((T_inst > 0.4) & (T_inst < 0.8))
    assert sum(1 for _ in open(job / "ActiveFileIndex.dat")) == 100
    # product analysis
    env = dict(os.environ, AC4_DIR=str(job))
    r = subprocess.run([os.path.join(BIN, "analyse"), "1", "0.6", "1.0", str(n)], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    # the oracle on the same files, into a copy of the directory
    ref = tmp_path / "ref"
    ref.mkdir()
    for f in ("time_index_file.dat", "data_coor.out", "ActiveFileIndex.dat"):
        os.symlink(job / f, ref / f)
    assert OL.load().or_analyse_dir(str(ref).encode(), n, 1, 1, 0, 0) == 0
    assert filecmp.cmp(job / "relax_a_0.30.dat", ref / "relax_a_0.30.dat", shallow=False)
    a = np.loadtxt(job / "msd.dat")
    b = np.loadtxt(ref / "msd.dat")
    assert a.shape == b.shape and np.allclose(a, b, rtol=1e-12, atol=1e-14)
    assert os.path.getsize(job / "S4qt_a_0.30.dat") > 1000


def _run(cwd, args, **env):
    e = dict(os.environ, AC4_SEED="777", **env)
    r = subprocess.run([os.path.join(BIN, "run")] + args, cwd=cwd, env=e, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return r


def test_run_resumes_exactly_and_sidecar_matches_text(gpu, tmp_path):
    """A production run stopped after a restart file and resumed with
    AC4_RESUME=1 writes byte for byte the files of the uninterrupted run; the
    binary sidecar holds exactly the values the text parses to, so analyse
    on it writes the same output files."""
    n = 1000
    para = "eqDuration\tprodDuration\tframe_step\n1.0\t10.0\t500\nDENSITY\tT\tP\n1.20\t0.60\t1.0\nf0\n1.5\n"
    for d in ("a", "b"):
        (tmp_path / d).mkdir()
        (tmp_path / d / "para_file").write_text(para)
    _run(tmp_path / "a", ["1", "0", "0", str(n)], AC4_BIN="1", AC4_RESTART_EVERY="500")
    r = _run(tmp_path / "b", ["1", "0", "0", str(n)], AC4_BIN="1", AC4_RESTART_EVERY="500", AC4_STOP_AT="1000")
    assert "Stopped after step 1000" in r.stdout
    r = _run(tmp_path / "b", ["1", "0", "1", str(n)], AC4_BIN="1", AC4_RESTART_EVERY="500", AC4_RESUME="1")
    assert "resumed exactly at step 1001" in r.stdout
    ja, jb = tmp_path / "a" / "0.600" / "data_001", tmp_path / "b" / "0.600" / "data_001"
    for f in ("data_coor.out", "data_coor.bin", "stuff.dat", "time_index_file.dat", "ActiveDirectionRun.dat"):
        assert filecmp.cmp(ja / f, jb / f, shallow=False), f
    # analyse from the text and from the sidecar
    outs = {}
    for mode in ("0", "1"):
        env = dict(os.environ, AC4_DIR=str(ja), AC4_BIN=mode)
        r = subprocess.run([os.path.join(BIN, "analyse"), "1", "0.6", "1.0", str(n)], cwd=tmp_path, env=env,
                           capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        outs[mode] = {f: (ja / f).read_bytes() for f in ("msd.dat", "relax_a_0.30.dat", "S4qt_a_0.30.dat")}
    assert outs["0"] == outs["1"]
