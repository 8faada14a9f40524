"""C-ABI library checks that need no GPU: exported symbols, loud failure
without a device, host-side logic (schedule, origin gap, walk) against the
oracle, Philox known-answer vectors."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as OL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(ac4_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol(ac4):
    out = subprocess.check_output(["nm", "-D", "--defined-only", ac4.LIB_PATH]).decode()
    exported = set(re.findall(r" T (ac4_\w+)", out))
    missing = declared() - exported
    assert not missing, missing
    assert len(declared()) >= 35


def test_no_silent_cpu_path(ac4):
    if ac4.lib.ac4_device_count() > 0:
        pytest.skip("a GPU is visible")
    p = ac4.ReplicaParams(1000, 1.2, 0.6, 1.5, -1, 200, 1)
    with pytest.raises(ac4.AC4Error, match="no HIP device"):
        ac4.MD([p])
    with pytest.raises(ac4.AC4Error, match="no HIP device"):
        ac4.Analysis(100, 4.4, [0, 1, 2, 3, 4, 5, 20, 21, 22, 23, 24, 25])


def test_server_group_count_is_checked(ac4):
    """Launch-group counts outside 0..64 are refused before anything touches
    the device (a tick steps at most 64 groups)."""
    for g in (-1, 65):
        with pytest.raises(ac4.AC4Error, match="groups in 0..64"):
            ac4.Server([2048, 2048, 4096], groups=g)


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ac4, ctr, key, want):
    # Random123 philox4x32-10 known-answer vectors (Salmon et al. 2011)
    assert ac4.philox(ctr, key).tolist() == want


@pytest.mark.parametrize("run", [400, 1000, 10000, 20000, 50000, 100000, 123457])
def test_product_schedule_matches_oracle(ac4, oracle, run):
    a, araw = ac4.schedule(run)
    b, braw = OL.schedule(run)
    assert np.array_equal(a, b)
    assert int(araw[len(a)]) == int(braw[len(b)])


def test_product_walk_matches_oracle(ac4, oracle):
    import trajgen
    for fr, gap in ((trajgen.case_frames(3, 30), 30), (ac4.schedule(10000)[0].astype(np.int64).tolist(), 50)):
        f = np.array(fr, dtype=np.int32)
        assert ac4.origin_gap(f) == gap == oracle.or_origin_gap(OL.ip(f), len(f))


@pytest.mark.parametrize("n,T,want", __import__("test_oracle_md").CHAIN_MASS_KAT)
def test_chain_init_masses_known_answers(ac4, n, T, want):
    # ac4_chain_init: the masses of InitialiseNoseHoover (time_evolve.c:477-482),
    # eta = 0, p_eta finite and seed-determined
    c = ac4.chain_init(n, T, 12345)
    assert c[6:9] == pytest.approx(want, rel=1e-14, abs=0)
    assert np.all(c[0:3] == 0.0) and np.all(np.isfinite(c[3:6]))
    assert np.array_equal(c, ac4.chain_init(n, T, 12345))
    assert not np.array_equal(c[3:6], ac4.chain_init(n, T, 12346)[3:6])


def test_init_lattice_invariants(ac4):
    n, T = 1000, 0.5
    s = ac4.init_lattice(n, T, 42)
    assert (s["type"] == 1).sum() == n - int(0.8 * n)
    p2 = (s["px"] ** 2 + s["py"] ** 2 + s["pz"] ** 2).sum()
    assert p2 / (3 * n - 3) == pytest.approx(T, rel=1e-12)
    assert abs(s["px"].sum()) < 1e-9 and abs(s["py"].sum()) < 1e-9
    assert sorted(s["act_tag"][s["act_tag"] >= 0].tolist()) == list(range(100))
    assert s["kx"].sum() == 0 and set(s["kx"].tolist()) == {1, -1}
    pos = np.stack([s["x"], s["y"], s["z"]], 1)
    assert len({tuple(p) for p in pos.tolist()}) == n  # distinct lattice sites
    assert np.all((pos >= 0) & (pos < 1))
    s2 = ac4.init_lattice(n, T, 42)
    assert np.array_equal(s["x"], s2["x"]) and np.array_equal(s["px"], s2["px"])


def _glibc_line(libc, x, y, z, t):
    buf = ctypes.create_string_buffer(256)
    libc.snprintf(buf, 256, b"%.5lf\t%.5lf\t%.5lf\t%d\n", ctypes.c_double(x), ctypes.c_double(y), ctypes.c_double(z),
                  ctypes.c_int(t))
    return buf.value


def test_frame_format_is_glibc_printf(ac4):
    """ac4_format_frame against glibc's own %.5lf on values a trajectory holds
    and on the hard cases: exact ties (k/2^j), values that round to -0.00000,
    subnormals, large unwrapped coordinates."""
    libc = ctypes.CDLL(None)
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        rng.uniform(-3, 3, 4000), rng.uniform(-5000, 5000, 1000), rng.normal(0, 1e-6, 500),
        np.arange(-4096, 4096) / 2.0 ** 12, np.arange(0, 2000) / 2.0 ** 17 - 0.0076,
        [0.0, -0.0, 5e-324, -5e-324, 1e-300, 0.000005, -0.000005, 0.015625, 0.999995, 1 - 1e-17, 12345.678905,
         2.5e-6, -2.5e-6, 7.5e-6, 0.000015, 1e11 + 0.5],
    ])
    n = len(vals) // 3 * 3
    v = vals[:n].reshape(3, -1)
    t = (rng.random(v.shape[1]) < 0.2).astype(np.int32)
    text = ac4.format_frame(v[0], v[1], v[2], t)
    want = b"".join(_glibc_line(libc, a, b, c, int(d)) for a, b, c, d in zip(v[0], v[1], v[2], t))
    assert text == want


def test_frame_parse_is_strtod(ac4):
    libc = ctypes.CDLL(None)
    libc.strtod.restype = ctypes.c_double
    libc.strtod.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
    rng = np.random.default_rng(6)
    m = 3000
    v = rng.uniform(-40, 40, (3, m))
    t = (rng.random(m) < 0.2).astype(np.int32)
    text = ac4.format_frame(v[0], v[1], v[2], t)
    # also other spellings the reader must accept: exponents, long mantissas
    text += b"1e-3\t-2.5E1\t0.12345678901234567\t1\n"
    x, y, z, tt, k = ac4.parse_frame(text, m + 1)
    assert k == len(text)
    toks = text.split()
    want = np.array([libc.strtod(tok, None) for tok in toks]).reshape(-1, 4)
    assert np.array_equal(x, want[:, 0]) and np.array_equal(y, want[:, 1]) and np.array_equal(z, want[:, 2])
    assert np.array_equal(np.signbit(x), np.signbit(want[:, 0]))
    assert np.array_equal(tt, want[:, 3].astype(np.int32))


def test_frame_format_of_huge_values(ac4):
    """Values of 1e12 or more (a run that blew up) go through glibc's %.5lf;
    a frame whose text would not fit 80 bytes a line is refused with
    AC4_ECAP instead of overrunning the buffer."""
    libc = ctypes.CDLL(None)
    x = np.array([1e15, -2.5e13, np.inf, -np.inf, np.nan, 0.5])
    y = np.array([0.25, 1e12, 3.0, -1e-7, 7.0, 123456789012.0])
    z = np.zeros(6)
    t = np.array([0, 1, 0, 1, 0, 1], dtype=np.int32)
    text = ac4.format_frame(x, y, z, t)
    want = b"".join(_glibc_line(libc, a, b, c, int(d)) for a, b, c, d in zip(x, y, z, t))
    assert text == want
    x[1], y[1] = 1e300, -1e300
    with pytest.raises(ac4.AC4Error, match="1e12 or more"):
        ac4.format_frame(x, y, z, t)


def test_frame_parse_stops_at_the_text_end(ac4):
    """ac4_parse_frame reads only the bytes it is given: a last line without
    its newline is parsed up to the end, and a line that is not four numbers
    is an error, not a read into the next line."""
    good = b"0.50000\t-1.25000\t3.00000\t1\n0.1\t0.2\t0.3\t0"
    x, y, z, t, k = ac4.parse_frame(good, 2)
    assert k == len(good) and list(x) == [0.5, 0.1] and list(t) == [1, 0]
    for bad in (b"0.5\t0.5\n1.0\t2.0\t3.0\t1\n", b"0.5\t0.5\t0.5\tx\n", b"\n"):
        with pytest.raises(ac4.AC4Error, match="not"):
            ac4.parse_frame(bad, 1)
    with pytest.raises(ac4.AC4Error, match="ends at line 1"):
        ac4.parse_frame(b"1\t2\t3\t0\n", 2)


def test_sincos_turns_against_extended_precision(ac4):
    """The sin/cos of the S4 and Fs kernels (host build of the same source):
    within 2.5e-16 of sin/cos(2 pi f) formed in x87 extended precision, for
    phases in [-1/2, 1/2] and with up to 100 whole turns added; exact at the
    quarter turns."""
    rng = np.random.default_rng(11)
    f = np.concatenate([rng.uniform(-0.5, 0.5, 200000), [0.0, 0.125, -0.125, 0.375, -0.375, 1e-300]])
    f = (f[:, None] + np.array([0.0, 1.0, -3.0, 17.0, -100.0])[None, :]).ravel()
    s, c = np.empty_like(f), np.empty_like(f)
    dp = ctypes.POINTER(ctypes.c_double)
    assert ac4.lib.ac4_sincos_turns_host(f.ctypes.data_as(dp), len(f), s.ctypes.data_as(dp),
                                         c.ctypes.data_as(dp)) == 0
    pi = np.longdouble("3.14159265358979323846264338327950288")
    x = 2 * pi * f.astype(np.longdouble)
    assert float(np.abs(s - np.sin(x)).max()) < 2.5e-16
    assert float(np.abs(c - np.cos(x)).max()) < 2.5e-16
    q = np.array([0.0, 0.25, 0.5, 0.75, -0.25, -0.5, 3.0])
    sq, cq = np.empty_like(q), np.empty_like(q)
    ac4.lib.ac4_sincos_turns_host(q.ctypes.data_as(dp), len(q), sq.ctypes.data_as(dp), cq.ctypes.data_as(dp))
    assert np.array_equal(sq, [0.0, 1.0, 0.0, -1.0, -1.0, 0.0, 0.0])
    assert np.array_equal(cq, [1.0, 0.0, -1.0, 0.0, 0.0, -1.0, 1.0])
