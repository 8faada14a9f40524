"""ctypes binding of the oracle (oracle/liboracle.so): the CPU restatement of
the reference used as the checker.  Only tests, smoke() and bench.py's
cpu_baseline load it."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.environ.get("AC4_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
U64 = ctypes.POINTER(ctypes.c_uint64)
_lib = None


class Cells(ctypes.Structure):
    _fields_ = [("ncx", ctypes.c_int), ("ncy", ctypes.c_int), ("ncz", ctypes.c_int), ("ncell", ctypes.c_int),
                ("inv_cx", ctypes.c_double), ("inv_cy", ctypes.c_double), ("inv_cz", ctypes.c_double)]


SIGS = {
    "or_md_box_length": (ctypes.c_double, [ctypes.c_int, ctypes.c_double]),
    "or_ana_box_length": (ctypes.c_double, [ctypes.c_int]),
    "or_cells_setup": (None, [ctypes.c_double, ctypes.POINTER(Cells)]),
    "or_cell_of": (ctypes.c_int, [ctypes.POINTER(Cells), ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "or_maps": (None, [ctypes.POINTER(Cells), I]),
    "or_half_list": (ctypes.c_int, [ctypes.POINTER(Cells), ctypes.c_int, D, D, D, ctypes.c_double, ctypes.c_int,
                                    I, I, I]),
    "or_full_counts_brute": (None, [ctypes.c_int, D, D, D, ctypes.c_double, I]),
    "or_forces_list": (None, [ctypes.c_int, D, D, D, D, D, D, I, ctypes.c_double, I, I, ctypes.c_int, I, I, I, I,
                              ctypes.c_double, D, D, D, D]),
    "or_forces_brute": (None, [ctypes.c_int, D, D, D, I, ctypes.c_double, I, I, I, I, ctypes.c_double, D, D, D, D]),
    "or_forces_sample": (None, [ctypes.c_int, D, D, D, I, ctypes.c_double, I, I, I, I, ctypes.c_double,
                                ctypes.c_int, I, D]),
    "or_nhc_masses": (None, [ctypes.c_int, ctypes.c_double, D]),
    "or_nhc_steps": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int] + [D] * 12 + [I] * 5 + [D] * 4 + [D]),
    "or_shuffle": (None, [ctypes.c_int, I, D]),
    "or_schedule": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, U64,
                                   ctypes.c_int]),
    "or_origin_gap": (ctypes.c_int, [I, ctypes.c_int]),
    "or_getq": (ctypes.c_int, [I, I, I, I]),
    "or_walk": (ctypes.c_int, [I, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, I, I, I,
                               I]),
    "or_msd": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, I, ctypes.c_int, D, D, D, I,
                              ctypes.c_double, ctypes.c_int, ctypes.c_int, D, D, D, I, D, D, I]),
    "or_overlap": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, I, ctypes.c_int, D, D, D, I,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, D, D, I, D, I, I]),
    "or_s4": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, I, ctypes.c_int, D, D, D,
                             ctypes.c_double, ctypes.c_double, ctypes.c_int, D, D, D, D, I, D, D, D, I]),
    "or_fsqt": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, I, ctypes.c_int, D, D, D, I,
                               ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, D, D, D, D, I, D, D, D]),
    "or_analyse_dir": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]),
    "or_chi4_ensemble": (None, [ctypes.c_int, ctypes.c_int, D, D, I, I, D]),
    "or_block_scale": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, I, ctypes.c_int, D, D, D,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_char_p]),
    "or_block_analyse_dir": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_double]),
    "or_guu": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, I, ctypes.c_int, D, D, D, I,
                              ctypes.c_double, ctypes.c_double, ctypes.c_char_p, D, I, I, I]),
    "or_frame_printf": (ctypes.c_longlong, [ctypes.c_int, D, D, D, I, ctypes.c_char_p]),
    "or_frame_strtod": (ctypes.c_longlong, [ctypes.c_char_p, ctypes.c_int, D, D, D, I]),
    "or_cage_list": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, D, D, D, ctypes.c_int, I, I]),
    "or_overlap_cage": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, I, ctypes.c_int, D, D, D,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_char_p, I,
                                       ctypes.c_int, I]),
}


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(PATH)
        for k, (r, a) in SIGS.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def dp(a):
    return None if a is None else a.ctypes.data_as(D)


def ip(a):
    return None if a is None else a.ctypes.data_as(I)


def cells(L):
    c = Cells()
    load().or_cells_setup(L, ctypes.byref(c))
    return c


def half_list(x, y, z, L, maxneib=500):
    lib = load()
    n = len(x)
    c = cells(L)
    cnt = np.zeros(n, dtype=np.int32)
    lst = np.zeros(n * maxneib, dtype=np.int32)
    cell = np.zeros(n, dtype=np.int32)
    rc = lib.or_half_list(ctypes.byref(c), n, dp(x), dp(y), dp(z), L, maxneib, ip(cnt), ip(lst), ip(cell))
    assert rc == 0
    return cnt, lst.reshape(n, maxneib), cell


def full_counts(x, y, z, L):
    n = len(x)
    cnt = np.zeros(n, dtype=np.int32)
    load().or_full_counts_brute(n, dp(x), dp(y), dp(z), L, ip(cnt))
    return cnt


def forces_list(s, L, f0, cnt, lst, maxneib=500):
    n = len(s["x"])
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    o = np.zeros(3)
    load().or_forces_list(n, dp(s["x"]), dp(s["y"]), dp(s["z"]), dp(s["xu"]), dp(s["yu"]), dp(s["zu"]),
                          ip(s["type"]), L, ip(cnt), ip(lst), maxneib, ip(s["act_tag"]), ip(s["kx"]),
                          ip(s["ky"]), ip(s["kz"]), f0, dp(fx), dp(fy), dp(fz), dp(o))
    return fx, fy, fz, o


def forces_brute(s, L, f0):
    n = len(s["x"])
    fx, fy, fz = np.zeros(n), np.zeros(n), np.zeros(n)
    o = np.zeros(2)
    load().or_forces_brute(n, dp(s["x"]), dp(s["y"]), dp(s["z"]), ip(s["type"]), L, ip(s["act_tag"]),
                           ip(s["kx"]), ip(s["ky"]), ip(s["kz"]), f0, dp(fx), dp(fy), dp(fz), dp(o))
    return fx, fy, fz, o


def nhc_steps(s, L, T, f0, nsteps, chain, fresh=1, safe=1, nve=0):
    """Runs the oracle NHC in place on copies; returns (state, chain, out8)."""
    lib = load()
    st = {k: np.array(v, copy=True) for k, v in s.items()}
    n = len(st["x"])
    for k in ("fx", "fy", "fz"):
        st.setdefault(k, np.zeros(n))
    xo, yo, zo = st["xu"].copy(), st["yu"].copy(), st["zu"].copy()
    ch = np.array(chain, dtype=np.float64, copy=True)
    out = np.zeros(8)
    rc = lib.or_nhc_steps(n, L, T, f0, nsteps, fresh, safe, nve,
                          *[dp(st[k]) for k in ("x", "y", "z", "xu", "yu", "zu", "px", "py", "pz", "fx", "fy", "fz")],
                          ip(st["type"]), ip(st["act_tag"]), ip(st["kx"]), ip(st["ky"]), ip(st["kz"]),
                          dp(ch), dp(xo), dp(yo), dp(zo), dp(out))
    assert rc == 0
    return st, ch, out


def schedule(run_length, n_origins=200, factor=1.4, tail=0):
    cap = n_origins * 64 + 64
    out = np.zeros(cap, dtype=np.uint64)
    n = load().or_schedule(run_length, n_origins, factor, tail, out.ctypes.data_as(U64), cap)
    return out[:n].copy(), out


def walk(fidx, gap, rule, factor=1.4):
    f = np.ascontiguousarray(fidx, dtype=np.int32)
    cap = len(f) * 64 + 64
    pc, po, ps = (np.zeros(cap, dtype=np.int32) for _ in range(3))
    mi = ctypes.c_int()
    n = load().or_walk(ip(f), len(f), gap, factor, rule, cap, ip(pc), ip(po), ip(ps), ctypes.byref(mi))
    return pc[:n], po[:n], ps[:n], mi.value


def traj_args(X, Y, Z, T, fidx):
    return (np.ascontiguousarray(X, dtype=np.float64), np.ascontiguousarray(Y, dtype=np.float64),
            np.ascontiguousarray(Z, dtype=np.float64), np.ascontiguousarray(T, dtype=np.int32),
            np.ascontiguousarray(fidx, dtype=np.int32))


def msd(N, L, X, Y, Z, T, fidx, gap, flag=0, factor=1.4):
    X, Y, Z, T, f = traj_args(X, Y, Z, T, fidx)
    ns = len(f) + 64
    m0, m1, m2, dt = (np.zeros(ns) for _ in range(4))
    cn = np.zeros(ns, dtype=np.int32)
    pv = np.zeros((len(f) * 64 + 64) * 3)
    noi = ctypes.c_int()
    npairs = load().or_msd(N, L, len(f), ip(f), gap, dp(X), dp(Y), dp(Z), ip(T), factor, flag, ns, dp(m0), dp(m1),
                           dp(m2), ip(cn), dp(dt), dp(pv), ctypes.byref(noi))
    return dict(m0=m0, m1=m1, m2=m2, counters=cn, delta_t=dt, pairs=pv[:3 * npairs].reshape(-1, 3), noi=noi.value)


def overlap(N, L, X, Y, Z, T, fidx, gap, a=0.3, flag=0, factor=1.4):
    X, Y, Z, T, f = traj_args(X, Y, Z, T, fidx)
    ns = len(f) + 64
    c, x4, dt = (np.zeros(ns) for _ in range(3))
    cn = np.zeros(ns, dtype=np.int32)
    pc = np.zeros(len(f) * 64 + 64, dtype=np.int32)
    noi = ctypes.c_int()
    npairs = load().or_overlap(N, L, len(f), ip(f), gap, dp(X), dp(Y), dp(Z), ip(T), factor, a, flag, ns, dp(c),
                               dp(x4), ip(cn), dp(dt), ip(pc), ctypes.byref(noi))
    return dict(c=c, x4=x4, counters=cn, delta_t=dt, pairs=pc[:npairs], noi=noi.value)


def phase(kind, N, L, X, Y, Z, T, fidx, gap, a=0.3, flag=0, factor=1.4):
    X, Y, Z, T, f = traj_args(X, Y, Z, T, fidx)
    ns = len(f)
    cs, ss, cs2, ss2 = (np.zeros(ns * 162) for _ in range(4))
    cn = np.zeros(ns, dtype=np.int32)
    dt = np.zeros(ns)
    qv = np.zeros(162)
    cap = len(f) * 64 + 64
    pcs = np.zeros(cap * 324)
    pm = np.zeros(cap, dtype=np.int32)
    if kind == "s4":
        npairs = load().or_s4(N, L, len(f), ip(f), gap, dp(X), dp(Y), dp(Z), factor, a, ns, dp(cs), dp(ss), dp(cs2),
                              dp(ss2), ip(cn), dp(dt), dp(qv), dp(pcs), ip(pm))
    else:
        npairs = load().or_fsqt(N, L, len(f), ip(f), gap, dp(X), dp(Y), dp(Z), ip(T), factor, a, flag, ns, dp(cs),
                                dp(ss), dp(cs2), dp(ss2), ip(cn), dp(dt), dp(qv), dp(pcs))
    return dict(cs=cs.reshape(ns, 162), ss=ss.reshape(ns, 162), cs2=cs2.reshape(ns, 162), ss2=ss2.reshape(ns, 162),
                counters=cn, delta_t=dt, qval=qv, pairs=pcs[:npairs * 324].reshape(-1, 324), mask=pm[:npairs])


def block_scales(L):
    """Block edges of Analysis_Block/main.c:45-63: L/scale, scale = 3..19."""
    out, scale = [], 3.0
    while scale < 20.0:
        out.append(L / scale)
        scale = scale + 1.0
    return out


def block_all(N, L, X, Y, Z, fidx, gap, sub_dir, vh_time=1500.0, factor=1.4, a=0.3, cuts=None):
    """or_block_scale for every block edge; files under sub_dir/BlockAnalysis."""
    for d in ("BlockAnalysis", "BlockAnalysis/MSD", "BlockAnalysis/Relax"):
        os.makedirs(os.path.join(sub_dir, d), exist_ok=True)
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.float64) for v in (X, Y, Z))
    f = np.ascontiguousarray(fidx, dtype=np.int32)
    for cut in (cuts if cuts is not None else block_scales(L)):
        rc = load().or_block_scale(N, L, len(f), ip(f), gap, dp(X), dp(Y), dp(Z), factor, a * a, cut, vh_time,
                                   os.fsencode(sub_dir))
        assert rc == 0


def guu(N, L, X, Y, Z, act, fidx, gap, tau, path=None, factor=1.4):
    """or_guu: returns (hist[3][maxbin], frame_num, n_origins)."""
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.float64) for v in (X, Y, Z))
    f = np.ascontiguousarray(fidx, dtype=np.int32)
    act = np.ascontiguousarray(act, dtype=np.int32)
    mb = int(0.5 * L / 0.005)
    h = np.zeros(3 * mb)
    out = [ctypes.c_int() for _ in range(3)]
    rc = load().or_guu(N, L, len(f), ip(f), gap, dp(X), dp(Y), dp(Z), ip(act), factor, tau,
                       os.fsencode(path) if path else None, dp(h), *[ctypes.byref(o) for o in out])
    assert rc == 0 and out[0].value == mb
    return h.reshape(3, mb), out[1].value, out[2].value


CAGE_CUT = 2.5 + 0.35  # MDcode/global.h:30 (Lcuttoff + skin)


def cage_list(x0, y0, z0, L, cap=400):
    n = len(x0)
    nn = np.zeros(n, dtype=np.int32)
    nb = np.full(n * cap, -1, dtype=np.int32)
    x0, y0, z0 = (np.ascontiguousarray(v, dtype=np.float64) for v in (x0, y0, z0))
    assert load().or_cage_list(n, L, CAGE_CUT, dp(x0), dp(y0), dp(z0), cap, ip(nn), ip(nb)) == 0
    return nn, nb.reshape(n, cap)


def overlap_cage(N, L, X, Y, Z, fidx, gap, path=None, a=0.3, factor=1.4):
    """or_overlap_cage: returns the per-pair counts in walk order."""
    X, Y, Z = (np.ascontiguousarray(v, dtype=np.float64) for v in (X, Y, Z))
    f = np.ascontiguousarray(fidx, dtype=np.int32)
    cap = len(f) * 64 + 64
    pc = np.zeros(cap, dtype=np.int32)
    npairs = ctypes.c_int()
    rc = load().or_overlap_cage(N, L, len(f), ip(f), gap, dp(X), dp(Y), dp(Z), factor, a, CAGE_CUT,
                                os.fsencode(path) if path else None, ip(pc), cap, ctypes.byref(npairs))
    assert rc == 0
    return pc[:npairs.value]


def forces_sample(s, L, f0, idx):
    """or_forces_sample: forces on particles idx by an O(N) pass each, (m, 3)."""
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.zeros((len(idx), 3))
    g = {k: np.ascontiguousarray(s[k], dtype=np.float64) for k in ("x", "y", "z")}
    g.update({k: np.ascontiguousarray(s[k], dtype=np.int32) for k in ("type", "act_tag", "kx", "ky", "kz")})
    load().or_forces_sample(len(g["x"]), dp(g["x"]), dp(g["y"]), dp(g["z"]), ip(g["type"]), L, ip(g["act_tag"]),
                            ip(g["kx"]), ip(g["ky"]), ip(g["kz"]), f0, len(idx), ip(idx), dp(out))
    return out


def chi4_ensemble(sq, sq2, cnt, n_r):
    """or_chi4_ensemble over per-replica per-lag sums (arrays [nrep, nlag]);
    returns (chi4, Q, origins) per lag."""
    sq, sq2 = (np.ascontiguousarray(a, dtype=np.float64) for a in (sq, sq2))
    cnt = np.ascontiguousarray(cnt, dtype=np.int32)
    n_r = np.ascontiguousarray(n_r, dtype=np.int32)
    nrep, nlag = sq.shape
    out = np.zeros(3 * nlag)
    load().or_chi4_ensemble(nrep, nlag, dp(sq), dp(sq2), ip(cnt), ip(n_r), dp(out))
    out = out.reshape(nlag, 3)
    return out[:, 0].copy(), out[:, 1].copy(), out[:, 2].astype(np.int64)
