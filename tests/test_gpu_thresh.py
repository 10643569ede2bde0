"""GPU bisection thresholds against the oracle's bisection (float64).

Both run the same bisection; the trial outcomes agree, so the brackets
(ub, lb), the status codes and the trial counts are equal exactly.
"""

import numpy as np
import pytest

from dendra_amd.model import HH, rect_wave
from dendra_amd.workloads import threshold_fibers

pytestmark = pytest.mark.gpu


def _oracle(orc, fs, **kw):
    wave = rect_wave(1000, 0.005, 0.5, 0.1, amp=-1.0, dtype=np.float64)
    return orc.threshold(fs.spec, HH(celsius=6.3).array(),
                         fs.profile.astype(np.float64).reshape(-1), wave, 0.005, **kw)


def test_gpu_thresholds_equal_oracle(orc):
    from dendra_amd.thresholder import GPUThresholder

    fs = threshold_fibers(150, seed=9)
    gt = GPUThresholder(fs.spec, fs.profile, rtol=0.005)
    ub, lb = gt.calculate_thresholds()
    oub, olb, ost, otr = _oracle(orc, fs, rtol=0.005, threads=8)
    assert np.array_equal(gt.status, ost)
    assert np.array_equal(gt.trials, otr)
    assert np.array_equal(ub, oub) and np.array_equal(lb, olb)


def test_gpu_threshold_failures_match_oracle(orc):
    from dendra_amd.thresholder import GPUThresholder

    fs = threshold_fibers(40, seed=31, zmin=0.3, zmax=1.5)
    kw = dict(rtol=0.001)
    gt = GPUThresholder(fs.spec, fs.profile, max_tries_bound_fix=5, max_tries_thresh=6, **kw)
    ub, lb = gt.calculate_thresholds()
    oub, olb, ost, otr = _oracle(orc, fs, max_fix=5, max_bisect=6, threads=8, **kw)
    assert np.array_equal(gt.status, ost) and set(ost.tolist()) >= {1, 2}
    ok = ost == 0
    assert np.array_equal(ub[ok], oub[ok]) and np.isnan(ub[~ok]).all()
    assert np.array_equal(gt.trials, otr)
