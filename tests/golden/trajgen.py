"""Seeded synthetic trajectories in the reference's on-disk layout.

A job directory holds ``time_index_file.dat`` ("total timesteps equals %d" then
one step index per line, MDcode/main.c:111-116), ``ActiveFileIndex.dat`` (one
index per line, MDcode/active.c:27) and ``data_coor.out`` (N lines
``%.5lf\t%.5lf\t%.5lf\t%d`` of unwrapped box-scaled x_u,y_u,z_u,type per frame,
MDcode/print.c:10).  The coordinates are random walks chosen so that the
overlap, S4 and Fs observables see every branch: particles that stay put,
particles that move a few sigma, and particles whose unwrapped coordinate
crosses one or more box lengths.
"""
import numpy as np


def schedule_two_level(n_origins, gap, max_interval):
    """Frame step indices: origins k*gap and offsets 1,2,3,4,5,7,9,12,16,...
    (x1.4 truncated, +1 when the truncation stalls, MDcode/main.c:84-94)."""
    idx = set()
    for k in range(n_origins):
        idx.add(k * gap)
        off = 1
        while off < max_interval:
            idx.add(k * gap + off)
            off = off + 1 if int(off * 1.4) == off else int(off * 1.4)
    return sorted(idx)


def make_frames(n, frames, seed, step_sigma=0.06, wild=0.0):
    """Return (X, Y, Z, type) arrays [nframes, n] of unwrapped box-scaled
    coordinates, already rounded to the 5 decimals the text file holds."""
    rng = np.random.default_rng(seed)
    L = (n / 1.2) ** (1.0 / 3.0)
    nf = len(frames)
    typ = np.zeros(n, dtype=np.int32)
    typ[rng.permutation(n)[: n // 5]] = 1
    base = rng.random((3, n))
    # per-particle mobility: most particles are caged, some are fast
    mob = np.where(rng.random(n) < 0.15, 6.0, 1.0) * step_sigma / L
    out = np.empty((3, nf, n))
    prev_t = frames[0]
    pos = base.copy()
    if wild:
        # unwrapped coordinates several box lengths away from [0,1)
        pos += np.round(rng.normal(0.0, wild, size=(3, n)))
    for f, t in enumerate(frames):
        dt = t - prev_t
        if dt > 0:
            pos = pos + rng.normal(0.0, 1.0, size=(3, n)) * mob * np.sqrt(dt)
        prev_t = t
        out[:, f, :] = pos
    out = quantize(out)
    return out[0], out[1], out[2], np.broadcast_to(typ, (nf, n)).copy()


def quantize(a):
    """The doubles a reader gets back from the ``%.5lf`` text of ``a``.

    ``%.5f`` rounds the exact value to k * 1e-5 and the reader returns the
    double nearest k * 1e-5, which is k / 1e5 (both exact, one IEEE division).
    rint(a * 1e5) finds k except within a rounding error of a tie; those few
    values go through the text itself."""
    a = np.asarray(a, dtype=np.float64)
    s = a * 1e5
    k = np.rint(s)
    out = k / 1e5
    amb = np.abs(np.abs(s - k) - 0.5) < 1e-6
    if amb.any():
        out[amb] = np.char.mod("%.5f", a[amb]).astype(np.float64)
    return out


def write_job(sub_dir, frames, X, Y, Z, T, active=None):
    import os
    os.makedirs(sub_dir, exist_ok=True)
    with open(os.path.join(sub_dir, "time_index_file.dat"), "w") as f:
        f.write("total timesteps equals %d\n" % len(frames))
        for v in frames:
            f.write("%d\n" % v)
    n = X.shape[1]
    if active is None:
        active = np.arange(0, n, 10)
    with open(os.path.join(sub_dir, "ActiveFileIndex.dat"), "w") as f:
        for v in active:
            f.write("%d\n" % v)
    with open(os.path.join(sub_dir, "data_coor.out"), "w") as f:
        for fr in range(X.shape[0]):
            lines = ["%.5f\t%.5f\t%.5f\t%d\n" % (X[fr, i], Y[fr, i], Z[fr, i], T[fr, i])
                     for i in range(n)]
            f.write("".join(lines))


# Golden cases: (name, n, n_origins, gap, seed, step_sigma, wild)
CASES = [
    ("walk_g20", 50000, 2, 20, 20261202, 0.06, 0.0),
    ("wild_g24", 50000, 3, 24, 7, 0.05, 1.5),
]


# Block analysis (Analysis_Block, N fixed at 50000 there): the two cases above
# plus lags long enough (up to 1900 time units) for the van Hove lag search.
BLOCK_CASES = [
    ("walk_g20", 50000, 2, 20, 20261202, 0.06, 0.0),
    ("wild_g24", 50000, 3, 24, 7, 0.05, 1.5),
    ("long_g400k", 50000, 2, 400000, 11, 0.002, 0.0),
]


def case_frames(n_origins, gap):
    return schedule_two_level(n_origins, gap, n_origins * gap // 2)
