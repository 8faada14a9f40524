"""The drop-in executables' frame reader (activechi4_amd/host/frame_reader.h,
untrusted text) under AddressSanitizer and UndefinedBehaviorSanitizer on the
CPU: well-formed job directories read back exactly, and malformed ones
(truncated frames, garbage lines, short or lying index files, lines longer
than the read buffer, a bad binary sidecar) are refused with an error, with
no sanitizer report."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "frame_reader_harness.c")


@pytest.fixture(scope="module")
def harness(tmp_path_factory, ac4):
    exe = str(tmp_path_factory.mktemp("fr") / "frame_reader_harness")
    pkg = os.path.dirname(ac4.LIB_PATH)
    subprocess.check_call(["gcc", "-std=gnu99", "-O1", "-g", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
                           "-I" + os.path.join(ROOT, "include"), SRC, "-L" + pkg, "-l:" + os.path.basename(ac4.LIB_PATH),
                           "-Wl,-rpath," + pkg, "-lm", "-o", exe])
    return exe


def run(exe, d, n, header=1, binary=0):
    # the host reader (the GPU text parser has its own tests)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1", AC4_HOST_PARSE="1")
    env.pop("LD_PRELOAD", None)
    p = subprocess.run([exe, str(d), str(n), str(header), str(binary)], capture_output=True, text=True, env=env,
                       timeout=120)
    assert "Sanitizer" not in p.stderr and "runtime error" not in p.stderr, p.stderr[-3000:]
    return p


def write_job(d, frames, n, idx_count=None, extra=b""):
    d.mkdir(parents=True, exist_ok=True)
    nf = len(frames)
    with open(d / "time_index_file.dat", "w") as f:
        f.write("total timesteps equals %d\n" % (nf if idx_count is None else idx_count))
        for k in range(nf):
            f.write("%d\n" % (10 * k))
    with open(d / "data_coor.out", "wb") as f:
        for fr in frames:
            for i in range(n):
                f.write(b"%.5f\t%.5f\t%.5f\t%d\n" % (fr[0][i], fr[1][i], fr[2][i], i % 2))
        f.write(extra)


def frames_of(n, nf, seed=3):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-3, 4, (3, n)) for _ in range(nf)]


def test_well_formed_text_reads_back(harness, tmp_path):
    n, frs = 7, frames_of(7, 4)
    write_job(tmp_path, frs, n)
    p = run(harness, tmp_path, n)
    assert p.returncode == 0, p.stderr
    nf, s, t = p.stdout.split()
    want = sum(float(b"%.5f" % v) * w for fr in frs for w, a in zip((1.0, 2.0, 3.0), fr) for v in a)
    assert int(nf) == 4 and abs(float(s) - want) < 1e-9 * max(1.0, abs(want)) and int(t) == 4 * (n // 2)


@pytest.mark.parametrize("case", ["truncated_frame", "garbage_line", "index_lies", "index_header", "empty_data",
                                  "no_final_newline", "long_line", "missing_type"])
def test_malformed_text_is_refused(harness, tmp_path, case):
    n, frs = 6, frames_of(6, 3, seed=9)
    write_job(tmp_path, frs, n)
    data = tmp_path / "data_coor.out"
    txt = data.read_bytes()
    if case == "truncated_frame":
        data.write_bytes(txt[: len(txt) - 30])
    elif case == "garbage_line":
        lines = txt.split(b"\n")
        lines[8] = b"0.12345\tnot-a-number\t0.5\t1"
        data.write_bytes(b"\n".join(lines))
    elif case == "index_lies":
        write_job(tmp_path, frs, n, idx_count=50)
    elif case == "index_header":
        (tmp_path / "time_index_file.dat").write_text("garbage\n1\n2\n")
    elif case == "empty_data":
        data.write_bytes(b"")
    elif case == "no_final_newline":
        data.write_bytes(txt.rstrip(b"\n"))
    elif case == "long_line":
        lines = txt.split(b"\n")
        lines[3] = b"0." + b"1" * (3 << 20) + b"\t0.5\t0.5\t0"
        data.write_bytes(b"\n".join(lines))
    elif case == "missing_type":
        lines = txt.split(b"\n")
        lines[2] = b"0.50000\t0.50000\t0.50000"
        data.write_bytes(b"\n".join(lines))
    p = run(harness, tmp_path, n)
    assert p.returncode in (2, 3, 4), (case, p.returncode, p.stdout, p.stderr[-500:])


def test_binary_sidecar_bad_header_and_truncation(harness, tmp_path):
    n, nf = 5, 2
    write_job(tmp_path, frames_of(n, nf), n)
    side = tmp_path / "data_coor.bin"
    side.write_bytes(b"AC4FRM01" + struct.pack("<ii", n + 1, 0))
    assert run(harness, tmp_path, n, binary=1).returncode == 3
    body = b"AC4FRM01" + struct.pack("<ii", n, 0)
    frame = struct.pack("<Q", 0) + np.zeros(3 * n).tobytes() + np.zeros(n, dtype=np.int32).tobytes()
    side.write_bytes(body + frame + frame[:-7])
    assert run(harness, tmp_path, n, binary=1).returncode == 4
    side.write_bytes(body + frame + frame)
    assert run(harness, tmp_path, n, binary=1).returncode == 0
