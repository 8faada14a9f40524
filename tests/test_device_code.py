"""Properties of the gfx950 device code itself, read from the assembly of
this build's sources (no GPU needed).

An agent-scope release or acquire fence (``__threadfence``, or an atomic with
agent-scope release/acquire ordering) compiles on gfx950 to a write-back and
an invalidation of the XCD's L2 (``buffer_wbl2`` / ``buffer_inv``).  Issued by
every workgroup of the force kernel it cost the neighbour gathers of all the
other workgroups their L2 lines (k_force 42.7 against 55.8 us on the
64-replica ensemble without it), so the per-replica reductions run as kernels
of their own; this test keeps such instructions out of every kernel."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "activechi4_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["ac4_md.hip", "ac4_analysis.hip", "ac4_frames.hip", "ac4_block.hip", "ac4_pairs.hip", "ac4_text.hip"]


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not available")
def test_no_l2_writeback_or_invalidate_in_device_code(tmp_path):
    flags = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
             "-Xclang", "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-S",
             "--cuda-device-only"]
    procs = []
    for src in SOURCES:
        extra = ["-mllvm", "-amdgpu-mfma-vgpr-form"] if src == "ac4_analysis.hip" else []
        out = tmp_path / (src + ".s")
        procs.append((src, out, subprocess.Popen([HIPCC] + flags + extra + [os.path.join(CSRC, src), "-o", str(out)],
                                                 stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)))
    for src, out, p in procs:
        _, err = p.communicate(timeout=600)
        assert p.returncode == 0, (src, err.decode(errors="replace")[-2000:])
        text = out.read_text()
        assert "s_endpgm" in text, src  # kernels were emitted
        for ins in ("buffer_wbl2", "buffer_inv"):
            assert ins not in text, (src, ins)
