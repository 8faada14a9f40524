"""Regenerate the golden block-analysis vectors from the reference program.

Runs in the build container only (it needs the reference sources): builds the
reference Analysis_Block program into oracle/_ref/ (oracle/ref/Makefile),
writes each seeded trajectory of trajgen.BLOCK_CASES into a scratch tree laid
out as that program expects (run/para_file, 0.800/data_000/), runs it there
(block edges L/3 ... L/19, van Hove time 1500), and keeps, under
tests/golden/<case>/block/:
  - MSD/MSDLb*.dat, Relax/RelaxLb*.dat, Relax/RelaxTotLb*.dat (copied whole);
  - manifest.json: the SHA-256 and line count of every output file, the
    vanHove*.dat files included (those are kept only as digests: up to
    41,154 lines each).
The trajectories are not committed: tests regenerate them from trajgen.

usage: python tests/golden/make_block_golden.py
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import trajgen  # noqa: E402

PARA = ("DENSITY\tT\tP\tDIAMETER\tnb\ttau_alpha\n1.2 .80000    0.1\t0.3\t1\t.0\n"
        "start_ensemble\tend_ensemble\n0\t0\n")


def digest(path):
    with open(path, "rb") as f:
        data = f.read()
    return {"sha256": hashlib.sha256(data).hexdigest(), "lines": data.count(b"\n")}


def main():
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    exe = os.path.join(ROOT, "oracle", "_ref", "analyse_block_ref")
    for name, n, n_or, gap, seed, sig, wild in trajgen.BLOCK_CASES:
        frames = trajgen.case_frames(n_or, gap)
        X, Y, Z, T = trajgen.make_frames(n, frames, seed, sig, wild)
        with tempfile.TemporaryDirectory() as tmp:
            run = os.path.join(tmp, "run")
            job = os.path.join(tmp, "0.800", "data_000")
            os.makedirs(run)
            with open(os.path.join(run, "para_file"), "w") as f:
                f.write(PARA)
            trajgen.write_job(job, frames, X, Y, Z, T)
            subprocess.check_call([exe, "0", "0.8", "0"], cwd=run, stdout=subprocess.DEVNULL)
            src = os.path.join(job, "BlockAnalysis")
            dst = os.path.join(HERE, name, "block")
            shutil.rmtree(dst, ignore_errors=True)
            manifest = {"case": [name, n, n_or, gap, seed, sig, wild], "frames": len(frames),
                        "producer": "reference Analysis_Block (oracle/ref/Makefile)", "files": {}}
            for dirpath, _, files in os.walk(src):
                for fn in sorted(files):
                    rel = os.path.relpath(os.path.join(dirpath, fn), src)
                    manifest["files"][rel] = digest(os.path.join(dirpath, fn))
                    if not fn.startswith("vanHove"):
                        os.makedirs(os.path.dirname(os.path.join(dst, rel)), exist_ok=True)
                        shutil.copy(os.path.join(dirpath, fn), os.path.join(dst, rel))
            os.makedirs(dst, exist_ok=True)
            with open(os.path.join(dst, "manifest.json"), "w") as f:
                json.dump(manifest, f, indent=1, sort_keys=True)
        print("golden block", name, "frames", len(frames), "files", len(manifest["files"]))


if __name__ == "__main__":
    main()
