"""Ensemble statistics of the reference's CPU-runnable configuration, shared
by the CPU-side fixture generator (make_stat_golden.py, the oracle's NHC
step) and the GPU test (the library): KA N=1000 at density 1.2, T=0.6,
active fraction 0.1 with f0=1.5 and directions reshuffled every 200 steps
(tau_p = 1.0), dt = 0.005.

Each replica starts from a seeded lattice, is brought to temperature by a
300-step velocity-rescale ramp (1.5 T -> T; the coldstart/breakxal idea of
MDcode/initialise.c:166-201), runs 2000 Nose-Hoover steps to equilibrate and
then PROD steps with x_u saved every SAMPLE steps.  From those frames, for
origins every ORIGIN_EVERY frames and lags l = 1..NLAG frames:
  MSD(l) = mean over origins of (1/N) sum_i L^2 |x_u(t0+l) - x_u(t0)|^2
           (Analysis/msd_pos.c:64-76),
  Q(l)   = mean over origins of (1/N) #{i : L*L*((dx*dx+dy*dy)+dz*dz) < a*a}
           with one minimum-image correction (Analysis/overlap.c:61-84),
  chi4(l) = N (<Q^2> - <Q>^2) pooled over replicas and origins,
with standard errors over the replicas (independent runs) and, for chi4,
jackknife errors over the replicas.
"""
import numpy as np

N, RHO, T, F0, TAU_STEPS = 1000, 1.2, 0.6, 1.5, 200
RAMP, EQUIL, PROD, SAMPLE = 300, 2000, 10000, 100
ORIGIN_EVERY, NLAG, A = 10, 50, 0.3
DT = 0.005


def box_length():
    return (N / RHO) ** 0.33333333  # MDcode/main.c:465


def replica_curves(frames, L):
    """frames: array [nframes, 3, N] of box-scaled x_u.  Per-lag sums over
    origins of MSD, Q, Q^2 and the origin counts (lags 0..NLAG)."""
    nf = frames.shape[0]
    msd, q, q2 = (np.zeros(NLAG + 1) for _ in range(3))
    cnt = np.zeros(NLAG + 1, dtype=np.int64)
    for o in range(0, nf, ORIGIN_EVERY):
        for lag in range(1, NLAG + 1):
            if o + lag >= nf:
                break
            d = frames[o + lag] - frames[o]
            msd[lag] += (L * L * (d * d).sum(axis=0)).mean()
            dm = np.where(d >= 0.5, d - 1.0, np.where(d < -0.5, d + 1.0, d))
            r2 = L * L * ((dm[0] * dm[0] + dm[1] * dm[1]) + dm[2] * dm[2])
            qv = float((r2 < A * A).sum()) / N
            q[lag] += qv
            q2[lag] += qv * qv
            cnt[lag] += 1
    return msd, q, q2, cnt


def _chi4(q, q2, c):
    s, s2, cc = q.sum(0), q2.sum(0), c.sum(0)
    out = np.zeros(q.shape[1])
    ok = cc > 0
    out[ok] = N * (s2[ok] / cc[ok] - (s[ok] / cc[ok]) ** 2)
    return out


def summarize(per_replica):
    """per_replica: list of replica_curves outputs.  Means and errors per lag."""
    msd = np.array([m for m, _, _, _ in per_replica])
    q = np.array([v for _, v, _, _ in per_replica])
    q2 = np.array([v for _, _, v, _ in per_replica])
    c = np.array([v for _, _, _, v in per_replica]).astype(float)
    R = len(per_replica)
    with np.errstate(invalid="ignore", divide="ignore"):
        m_r, q_r = msd / c, q / c
    lags = np.arange(1, NLAG + 1)
    chi = _chi4(q, q2, c)
    jk = np.array([_chi4(np.delete(q, r, 0), np.delete(q2, r, 0), np.delete(c, r, 0)) for r in range(R)])
    chi_err = np.sqrt((R - 1) / R * ((jk - jk.mean(0)) ** 2).sum(0))
    k = int(np.argmax(chi[1:])) + 1
    kj = np.argmax(jk[:, 1:], axis=1) + 1
    hj = jk[np.arange(R), kj]
    return {
        "lag_steps": (lags * SAMPLE).tolist(),
        "msd": m_r[:, 1:].mean(0).tolist(), "msd_err": (m_r[:, 1:].std(0, ddof=1) / np.sqrt(R)).tolist(),
        "Q": q_r[:, 1:].mean(0).tolist(), "Q_err": (q_r[:, 1:].std(0, ddof=1) / np.sqrt(R)).tolist(),
        "chi4": chi[1:].tolist(), "chi4_err": chi_err[1:].tolist(),
        "chi4_peak": float(chi[k]), "chi4_peak_err": float(np.sqrt((R - 1) / R * ((hj - hj.mean()) ** 2).sum())),
        "chi4_peak_lag_steps": int(k * SAMPLE),
        "chi4_peak_lag_err_steps": float(SAMPLE * np.sqrt((R - 1) / R * ((kj - kj.mean()) ** 2).sum())),
        "replicas": R,
    }
