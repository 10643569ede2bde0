"""The learned threshold surrogate end to end, judged by GPU bisection.

A 4-member ensemble is trained on bisection thresholds of 600,000 seeded
fibers and evaluated on 100,000 held-out fibers (a disjoint range of the
seeded stream).  Bars: median threshold error <= 2 % after fallback; every
fallback value is exactly the bisection value; the surrogate pipeline
(features, ensemble, fallback) beats bisection of the same fibers by >= 10x
in wall time (the fallback bisection has a fixed latency of one full
bisection, so the ratio grows with the batch; bench.py measures 1,000,000
held-out fibers against a 2,000,000-fiber training set); inputs outside
the training range always fall back.
"""

import time

import numpy as np
import pytest
import torch

from dendra_amd.workloads import threshold_fibers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    from dendra_amd.surrogate import ThresholdSurrogate
    from dendra_amd.thresholder import GPUThresholder

    tr = threshold_fibers(600_000, seed=9)
    ytr, _ = GPUThresholder(tr.spec, tr.profile).calculate_thresholds()
    print(f"training labels: {np.isfinite(ytr).mean() * 100:.2f}% finite", flush=True)
    sur = ThresholdSurrogate(n_members=4, seed=0)
    sur.fit(tr.profile, tr.diam.astype(np.float32), ytr, epochs=40)
    return sur


def test_surrogate_error_fallback_and_speed(trained):
    from dendra_amd.thresholder import GPUThresholder, bisection_judge

    te = threshold_fibers(100_000, seed=9, start=10_000_000)
    judge = bisection_judge(te.spec, te.profile)
    d = te.diam.astype(np.float32)
    trained.thresholds(te.profile[:512], d[:512], judge)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    thr, fb = trained.thresholds(te.profile, d, judge)
    torch.cuda.synchronize()
    t_sur = time.perf_counter() - t
    t = time.perf_counter()
    ref, _ = GPUThresholder(te.spec, te.profile).calculate_thresholds()
    torch.cuda.synchronize()
    t_bis = time.perf_counter() - t
    err = np.abs(thr - ref) / ref
    print(f"median error {np.median(err) * 100:.3f}%, p90 {np.quantile(err, 0.9) * 100:.3f}%, "
          f"fallback {fb.mean() * 100:.2f}%, surrogate {t_sur * 1e3:.1f} ms vs bisection "
          f"{t_bis * 1e3:.1f} ms ({t_bis / t_sur:.0f}x)", flush=True)
    assert np.median(err) <= 0.02
    assert np.array_equal(thr[fb], ref[fb])
    assert fb.mean() < 0.05
    assert t_bis / t_sur >= 10.0


def test_out_of_range_inputs_fall_back(trained):
    from dendra_amd.thresholder import bisection_judge

    te = threshold_fibers(300, seed=9, start=20_000_000, dmin=17.0, dmax=20.0)
    judge = bisection_judge(te.spec, te.profile)
    thr, fb = trained.thresholds(te.profile, te.diam.astype(np.float32), judge)
    assert fb.all()
    assert np.array_equal(thr, judge(np.arange(300)))
