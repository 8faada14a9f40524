"""The batched-ensemble configuration (64 replicas, N logU[1000, 5000], T=0.5,
f0 in {0, 0.5, 1, 1.5, 2} x tau_p in {0.1, 1, 10}) with its ensemble chi4:
the same replicas run once with FP32 pair forces and once in FP64 (the plain
form; a replica's results do not depend on the batch), 5e4 steps each, and
the chi4(t) peak of every (f0, tau_p) point compared: height within 2% or
within 3 combined jackknife standard errors, time within 2 lags or 3
standard errors.  The device's chi4 reduction (ac4_md_overlap_chi4) equals
the oracle's on the same accumulator sums bit for bit."""
import numpy as np
import pytest

import bench
import oracle_lib as OL

pytestmark = pytest.mark.gpu


def test_device_chi4_equals_oracle_on_the_same_sums(gpu):
    ps = bench.ensemble_params(gpu, 12)
    md = bench.start(gpu, ps, gpu.FP32, 300)
    md.overlap_enable(100, 500, 8, 30, 0.3)
    md.overlap_sample()
    for _ in range(40):
        md.step(100, graph=True)
        md.overlap_sample()
    groups = {}
    for r, p in enumerate(ps):
        groups.setdefault(p.f0, []).append(r)
    for f0, rs in sorted(groups.items()):
        chi, q, c = md.chi4(rs)
        acc = [md.overlap(r) for r in rs]
        och, oq, oc = OL.chi4_ensemble([a[0] for a in acc], [a[1] for a in acc], [a[2] for a in acc],
                                       [ps[r].n for r in rs])
        assert np.array_equal(c, oc) and np.array_equal(q, oq) and np.array_equal(chi, och), f0
        assert c[1] > 0 and q[1] > q[-1] and chi[1:].max() > 0
    md.close()


def test_chi4_peaks_fp32_against_fp64(gpu):
    ps = bench.ensemble_params(gpu, 64)
    bench.default_shape(gpu)
    res = bench.chi4_compare(gpu, ps, 300)
    lag_t = res["sample_every"] * 0.005
    for row in res["by_param"]:
        a, b = row["fp32"], row["fp64"]
        print("f0 %.1f tau_p %5.1f: chi4 peak FP32 %7.2f +- %5.2f at %5.1f, FP64 %7.2f +- %5.2f at %5.1f (dev %+.3f)"
              % (row["f0"], row["tau_p"], a["peak"], a["peak_err"], a["peak_time"], b["peak"], b["peak_err"],
                 b["peak_time"], row["peak_rel_dev"]))
        assert row["resolved"], row  # the peak lies inside the lag window
        assert abs(row["peak_rel_dev"]) <= max(0.02, 3.0 * row["peak_rel_sigma"]), row
        tol_t = max(2 * lag_t, 3.0 * np.hypot(a["peak_time_err"], b["peak_time_err"]))
        assert abs(row["peak_time_dev"]) <= tol_t, row
    print("points within 2%%: %d of %d; median |dev| %.3f"
          % (res["points_within_2pct"], res["points"], res["median_abs_peak_rel_dev"]))
