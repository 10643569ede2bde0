"""Batched activation thresholds by bisection on the GPU (float64).

Counterpart of the reference's ``Thresholder.calculate_thresholds``
(dendra/models/instruments/thresholder.py:873-979) for unbranched HH cables
under an extracellular field ``Ve = lead * amp * w(t)`` with the ``Active``
criterion (models/callbacks.py:1946-1990).  The reference re-runs the whole
batch once per bisection step and synchronises with the host each time
(thresholder.py:954); here each wavefront runs the complete bound fixing and
bisection of its fibers inside one kernel launch (``dn_threshold_f64``), and
a trial ends as soon as its outcome is decided.

Defaults follow the reference's documented protocol (docs/basics/
01d_thresholding.ipynb): mono_rect(amp=-1, pw=0.1, delay=0.5) cathodic
pulse, Active(threshold=0, node_check=[5, -5]), ub = 0.1 mA, tstop 5 ms,
dt 0.005 ms; rtol 0.005 by default.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .model import (HH, CableSpec, cable_workspace, hh_steady_state, plan_cables,
                    rect_wave, slot_map)

_RCAP = 4


class GPUThresholder:
    """Thresholds (mA) of every cable of ``spec`` under per-cable lead fields.

    lead: (n_cells, K) or (total,) Ve per mA at each compartment.
    """

    def __init__(self, spec: CableSpec, lead, *, dt: float = 0.005, tstop: float = 5.0,
                 delay: float = 0.5, pw: float = 0.1, amp: float = -1.0,
                 node_check=(5, -5), threshold: float = 0.0, ub=0.1, rtol: float = 0.005,
                 fix_bound_up: float = 2.0, max_tries_bound_fix: int = 10,
                 max_tries_thresh: int = 25, hh: HH | None = None, device="cuda"):
        self.lib = _lib.load()
        self.device = torch.device(device)
        _lib.require_gpu(torch.empty(0, device=self.device))
        if not fix_bound_up > 1.0:
            raise ValueError("fix_bound_up should be > 1.")
        if not rtol > 0:
            raise ValueError("rtol must be positive")
        self.spec, self.hh = spec, hh or HH(celsius=6.3)
        self.dt, self.rtol, self.vthr = float(dt), float(rtol), float(threshold)
        self.fix_up, self.max_fix, self.max_bisect = (float(fix_bound_up),
                                                      int(max_tries_bound_fix),
                                                      int(max_tries_thresh))
        self.n_steps = int(round(tstop / dt))
        sizes = spec.sizes
        lead = np.asarray(lead, dtype=np.float64).reshape(-1)
        if lead.size != spec.total:
            raise ValueError("lead needs one value per compartment")
        self.plan = plan = plan_cables(sizes, _RCAP, 64)
        nt = len(plan.task_p)
        self.n_tasks, self.slots = nt, nt * 64 * _RCAP
        slot = slot_map(plan, sizes, spec.off, _RCAP)
        driven = CableSpec(spec.off, spec.diam, spec.dx, spec.rhoa, spec.cm, lead=lead,
                           ve_amp=np.ones(spec.n_cells), v_init=spec.v_init)
        ws = cable_workspace(driven)
        dev = self.device

        def s_arr(x, dtype=np.float64):
            out = np.zeros(self.slots, dtype=dtype)
            out[slot] = x
            return torch.as_tensor(out, device=dev)

        self.gl, self.grp, self.sc, self.dr = (s_arr(ws.gL), s_arr(ws.gRp), s_arr(ws.scale),
                                               s_arr(ws.drive))
        cell = np.repeat(np.arange(spec.n_cells), sizes)
        j = np.arange(spec.total) - spec.off[cell]
        K = sizes[cell]
        chk = np.zeros(spec.total, dtype=np.int32)
        for c in node_check:
            chk |= (j == (c % K)).astype(np.int32)
        self.chk = s_arr(chk, np.int32)
        self.wave = torch.as_tensor(rect_wave(self.n_steps, dt, delay, pw, amp=amp,
                                              dtype=np.float64), device=dev)
        i32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.int32, device=dev)
        self.task_nr, self.task_p = i32(plan.task_nr), i32(plan.task_p)
        self.task_cell0, self.task_ncell = i32(plan.task_cell0), i32(plan.task_ncell)
        ub = np.broadcast_to(np.asarray(ub, dtype=np.float64), (spec.n_cells,))
        self.ub_initial = ub.copy()
        m0, h0, n0 = hh_steady_state(spec.v_init)
        self.state0 = np.array([spec.v_init, m0, h0, n0], dtype=np.float64)
        self.hh_arr = self.hh.array().astype(np.float64)

    def calculate_thresholds(self, stream=None):
        """Return (upper, lower) bounds per cable (spec order, numpy float64);
        NaN where bound fixing or the tolerance failed.  ``status`` and
        ``trials`` (simulations per cable) are kept as attributes."""
        dev = self.device
        order = self.plan.order
        ub = torch.as_tensor(self.ub_initial[order].copy(), device=dev)
        lb = torch.zeros_like(ub)
        status = torch.zeros(len(order), dtype=torch.int32, device=dev)
        trials = torch.zeros_like(status)
        strm = (stream or torch.cuda.current_stream(dev)).cuda_stream
        st = self.lib.dn_threshold_f64(
            self.n_tasks, _RCAP, _lib.ptr(self.task_nr), _lib.ptr(self.task_p),
            _lib.ptr(self.task_cell0), _lib.ptr(self.task_ncell), _lib.ptr(self.gl),
            _lib.ptr(self.grp), _lib.ptr(self.sc), _lib.ptr(self.dr), _lib.ptr(self.chk),
            _lib.ptr(self.wave), self.n_steps, self.dt, self.hh_arr.ctypes.data,
            self.state0.ctypes.data, self.vthr, self.rtol, self.fix_up, self.max_fix,
            self.max_bisect, _lib.ptr(ub), _lib.ptr(lb), _lib.ptr(status), _lib.ptr(trials),
            strm)
        _lib.check(st, "threshold bisection")
        n = len(order)
        out = [np.empty(n), np.empty(n), np.empty(n, np.int32), np.empty(n, np.int32)]
        for o, t in zip(out, (ub, lb, status, trials)):
            o[order] = t.cpu().numpy()
        self.status, self.trials = out[2], out[3]
        bad = self.status != 0
        out[0][bad] = np.nan
        out[1][bad] = np.nan
        return out[0], out[1]


def bisection_judge(spec: CableSpec, lead, **kw):
    """judge(indices) -> thresholds (mA) of those cables by GPU bisection
    (the upper bounds; NaN where bisection failed)."""
    lead = np.asarray(lead, dtype=np.float64).reshape(spec.n_cells, -1)

    def judge(idx):
        idx = np.asarray(idx, dtype=np.int64)
        ub, _ = GPUThresholder(spec.subset(idx), lead[idx], **kw).calculate_thresholds()
        return ub

    return judge
