"""The allocation-free S4-on-shells entry point (ac4_frames_s4_shells_noalloc,
through the C ABI) against the allocating one (ac4_frames_s4_shells) on the
same frame store: bit for bit, on two calls in a row, the second smaller
(fewer pairs, shells and directions); the largest shape the kernels take
(64 directions x 108 shells, the |q|-shell workload of the N=10^6 box) fits
the scratch array, and a larger one is refused."""
import numpy as np
import pytest

from test_gpu_frames import _frames

pytestmark = pytest.mark.gpu


def test_noalloc_s4_shells_equals_allocating(gpu):
    n, L, nf = 3000, (3000 / 1.2) ** (1 / 3), 6
    fr_list = _frames(n, L, nf, 17)
    fr = gpu.Frames(n, L, nf)
    for k, f in enumerate(fr_list):
        fr.push(f[:, 0], f[:, 1], f[:, 2], 10 * k)
    calls = [  # (origins, lag frames, q0, nshell, ndir, seed): the second is smaller
        ([0, 1, 2, 0], [3, 4, 5, 5], 0.2, 108, 64, 5150),
        ([1, 2], [2, 4], 0.9, 20, 7, 77),
    ]
    for orig, lagf, q0, nshell, ndir, seed in calls:
        dq = 2 * np.pi / L
        want = fr.s4_shells(orig, lagf, q0, dq, nshell, ndir, seed, phases=True)
        got = fr.s4_shells(orig, lagf, q0, dq, nshell, ndir, seed, phases=True, noalloc=True)
        for a, b in zip(got, want):
            assert a.shape == b.shape
            assert np.array_equal(a, b)
        assert np.all(np.isfinite(got[0])) and got[0].max() > 0
    with pytest.raises(gpu.AC4Error, match="must be <= 256"):
        fr.s4_shells([0, 1], [2, 3], 0.5, 0.1, 200, 41, 1, noalloc=True)  # 41 * ceil(200/27) = 328 > 256
    fr.close()
