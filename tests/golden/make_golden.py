"""Regenerate the golden Analysis vectors from the reference program.

Runs in the build container only (it needs the reference sources): builds the
reference Analysis into oracle/_ref/ (oracle/ref/Makefile), writes each seeded
trajectory of trajgen.CASES into a scratch job directory, runs the reference
there with all four observables, and copies the output files into
tests/golden/<case>/.  The trajectories themselves are not committed: tests
regenerate them from trajgen with the same seeds.

usage: python tests/golden/make_golden.py
"""
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

OUTPUTS = ["msd.dat", "relax_a_0.30.dat", "S4qt_a_0.30.dat", "Fsqt_a_0.30.dat"]
# g_uu(r) at the lag nearest tau (Analysis/guu.c, called through the harness)
GUU_TAU = {"walk_g20": 0.05, "wild_g24": 0.07}


def main():
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])
    exe = os.path.join(ROOT, "oracle", "_ref", "analyse_ref")
    for name, n, n_or, gap, seed, sig, wild in trajgen.CASES:
        frames = trajgen.case_frames(n_or, gap)
        X, Y, Z, T = trajgen.make_frames(n, frames, seed, sig, wild)
        with tempfile.TemporaryDirectory() as tmp:
            job = os.path.join(tmp, "job")
            trajgen.write_job(job, frames, X, Y, Z, T)
            subprocess.check_call([exe, job, "mosf"], stdout=subprocess.DEVNULL)
            subprocess.check_call([exe, job, "g", repr(GUU_TAU[name])], stdout=subprocess.DEVNULL)
            dst = os.path.join(HERE, name)
            os.makedirs(dst, exist_ok=True)
            for o in OUTPUTS + ["guuTauAlphaAct.dat"]:
                shutil.copy(os.path.join(job, o), os.path.join(dst, o))
        with open(os.path.join(dst, "case.json"), "w") as f:
            json.dump({"n": n, "n_origins": n_or, "gap": gap, "seed": seed,
                       "step_sigma": sig, "wild": wild, "frames": frames, "guu_tau": GUU_TAU[name],
                       "producer": "reference Analysis (oracle/ref/Makefile)"}, f, indent=1)
        print("golden", name, "frames", len(frames))


if __name__ == "__main__":
    main()
