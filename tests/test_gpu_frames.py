"""Device frame store and S4 on |q| shells (libac4.so through the C ABI)
against a direct numpy restatement: overlap counts exact, shell phase sums
(computed on the GPU by a recurrence from shell to shell) against cos/sin of
every shell, and both S4 forms from those sums."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(n, L, nf, seed):
    rng = np.random.default_rng(seed)
    s0 = rng.random((n, 3))
    frames = [s0.copy()]
    for k in range(1, nf):
        step = rng.normal(0.0, 0.25 / L, (n, 3))
        jump = rng.integers(-2, 3, (n, 3)) * (rng.random((n, 3)) < 0.05)  # whole-box jumps of x_u
        frames.append(frames[-1] + step + jump)
    return frames


def _overlap_w(f0, f1, L, a=0.3):
    d = f1 - f0
    d = np.where(d >= 0.5, d - 1.0, np.where(d < -0.5, d + 1.0, d))
    return (L * L * ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) < a * a)


def test_frame_store_overlap_and_s4_shells(gpu):
    n, L, nf = 1500, (1500 / 1.2) ** (1 / 3), 4
    fr_list = _frames(n, L, nf, 3)
    fr = gpu.Frames(n, L, 8)
    for k, f in enumerate(fr_list):
        fr.push(f[:, 0], f[:, 1], f[:, 2], 10 * k)
    assert fr.count() == nf and fr.find(20) == 2 and fr.find(5) == -1
    for i0, i1 in ((0, 1), (0, 3), (2, 3)):
        assert fr.overlap(i0, i1) == int(_overlap_w(fr_list[i0], fr_list[i1], L).sum())

    pairs = [(0, 2), (1, 3), (0, 3)]
    q0, dq, nshell, ndir, seed = 0.7, 2 * np.pi / L, 30, 5, 11
    s4r, s4d, ph = fr.s4_shells([p[0] for p in pairs], [p[1] for p in pairs], q0, dq, nshell, ndir, seed,
                                phases=True)
    dirs = gpu.shell_directions(ndir, seed)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-14)
    q = q0 + dq * np.arange(nshell)
    want = np.zeros((len(pairs), ndir, nshell, 2))
    for p, (i0, i1) in enumerate(pairs):
        w = _overlap_w(fr_list[i0], fr_list[i1], L)
        s = fr_list[i0] - np.floor(fr_list[i0])
        proj = s[w] @ dirs.T  # (m, ndir)
        for d in range(ndir):
            arg = np.outer(proj[:, d], q * L)  # (m, nshell)
            want[p, d, :, 0] = np.cos(arg).sum(axis=0)
            want[p, d, :, 1] = np.sin(arg).sum(axis=0)
    assert np.abs(ph - want).max() < 1e-9 * n
    c, s = want[..., 0], want[..., 1]
    cbar, sbar = c.mean(axis=1), s.mean(axis=1)  # (pairs, nshell)
    ref = ndir * (cbar.var(axis=0) + sbar.var(axis=0)) / n
    dirv = (c.var(axis=0) + s.var(axis=0)).mean(axis=0) / n
    assert np.allclose(s4r, ref, rtol=1e-9, atol=1e-12)
    assert np.allclose(s4d, dirv, rtol=1e-9, atol=1e-12)
    # one pair: no fluctuation over origins
    s4r1, s4d1 = fr.s4_shells([0], [2], q0, dq, nshell, ndir, seed)
    assert np.all(s4r1 == 0.0) and np.all(s4d1 == 0.0)
    fr.close()


def test_shell_directions_cover_the_sphere(gpu):
    d = gpu.shell_directions(4096, 7)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-14)
    assert np.abs(d.mean(axis=0)).max() < 0.05  # isotropic: mean direction ~ 0
    assert np.abs((d * d).mean(axis=0) - 1 / 3).max() < 0.03
    assert np.array_equal(d[:10], gpu.shell_directions(10, 7))  # counter-based: prefix-stable
