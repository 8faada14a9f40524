"""data_coor.out text parsed on the GPU (ac4_textframes) against the host
reader (ac4_parse_frame, itself equal to strtod): every value bit for bit,
across chunk boundaries, for the reference's "%.5lf\\t%.5lf\\t%.5lf\\t%d" lines
and for lines outside that form (which take the host path)."""
import numpy as np
import pytest

import trajgen

pytestmark = pytest.mark.gpu


def _text(X, Y, Z, T):
    out = []
    for f in range(X.shape[0]):
        out.append("".join("%.5f\t%.5f\t%.5f\t%d\n" % (X[f, i], Y[f, i], Z[f, i], T[f, i])
                           for i in range(X.shape[1])))
    return "".join(out).encode()


def _gpu_frames(ac4, text, n, nframes, chunk):
    tf = ac4.TextFrames(n, chunk)
    frames, pos, calls = [], 0, 0
    while len(frames) < nframes:
        k, used = tf.parse(text[pos:pos + tf.cap])
        assert k > 0 and used > 0
        frames += [tf.frame(i) for i in range(k)]
        pos += used
        calls += 1
    tf.close()
    return frames[:nframes], pos, calls


def _host_frames(ac4, text, n, nframes):
    frames, pos = [], 0
    for _ in range(nframes):
        x, y, z, t, used = ac4.parse_frame(text[pos:], n)
        frames.append((x, y, z, t))
        pos += used
    return frames, pos


def _same(a, b):
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
        if u.dtype == np.float64:
            assert np.array_equal(np.signbit(u), np.signbit(v))  # -0.00000 stays -0.0


def test_text_frames_equal_host_reader(gpu):
    """12 frames of N=3000 unwrapped coordinates several boxes from [0,1) in
    the smallest chunk (128 bytes per line of a frame: about four frames a
    call, frames cut by the chunk end carried to the next call)."""
    n, nf = 3000, 12
    X, Y, Z, T = trajgen.make_frames(n, list(range(0, 10 * nf, 10)), 31, 0.3, 2.5)
    text = _text(X, Y, Z, T)
    g, gpos, calls = _gpu_frames(gpu, text, n, nf, 1)
    h, hpos = _host_frames(gpu, text, n, nf)
    assert gpos == hpos == len(text) and calls >= 3
    for a, b in zip(g, h):
        _same(a, b)
    assert np.array_equal(g[5][0], X[5]) and np.array_equal(g[11][3], T[11])


def test_text_frames_outside_the_fast_form(gpu):
    """Exponents, 17-digit mantissas, signs, bare fractions, CR LF endings,
    leading blanks, negative zero, and a last line without its newline: the
    frame goes through the host reader and the values are the same."""
    n = 8
    lines = ["1.5e-3\t-2.00000\t0.12500\t1", "0.12345678901234567\t3.1\t-0.00000\t0",
             "+2.5\t.5\t7.\t-3", "  4.25\t 5.5\t6.75\t 2", "1.00000\t2.00000\t3.00000\t0\r",
             "-0.00000\t-0.00000\t0.00000\t1", "123456789.12345\t-98765.4321\t0.0\t7", "9.99999\t8.88888\t7.77777\t1"]
    text = ("\n".join(lines) + "\n" + "\n".join(lines)).encode()  # second frame ends without a newline
    tf = gpu.TextFrames(n)
    k, used = tf.parse(text + b"\n")
    assert k == 2 and used == len(text) + 1
    h0 = gpu.parse_frame(text, n)
    h1 = gpu.parse_frame(text[h0[4]:], n)
    _same(tf.frame(0), h0[:4])
    _same(tf.frame(1), h1[:4])
    tf.close()


def test_text_frames_fast_form_only_and_partial_frame(gpu):
    """Two whole frames and part of a third in one call: two frames, the
    bytes of those two used; max_frames limits a call."""
    n = 1000
    X, Y, Z, T = trajgen.make_frames(n, [0, 5, 10], 5, 0.1, 0.0)
    text = _text(X, Y, Z, T)
    cut = len(text) - 100
    tf = gpu.TextFrames(n)
    k, used = tf.parse(text[:cut])
    h, hpos = _host_frames(gpu, text, n, 2)
    assert k == 2 and used == hpos
    for i in range(2):
        _same(tf.frame(i), h[i])
    k1, used1 = tf.parse(text, max_frames=1)
    assert k1 == 1 and used1 == _host_frames(gpu, text, n, 1)[1]
    tf.close()


def test_text_frames_refuse_bad_text(gpu):
    n = 4
    text = b"0.1\t0.2\t0.3\t0\n0.4\tabc\t0.6\t1\n0.7\t0.8\t0.9\t0\n1.0\t1.1\t1.2\t1\n"
    with pytest.raises(gpu.AC4Error):
        gpu.parse_frame(text, n)
    tf = gpu.TextFrames(n)
    with pytest.raises(gpu.AC4Error):
        tf.parse(text)
    tf.close()
