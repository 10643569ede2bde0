"""Trace fitting of HH cables: loss and hand-written adjoint on the GPU.

The reference gets gradients of a loss over a long cable run by autograd
through every time step of ``Population.run`` (dendra/models/core.py:4555-4614):
``ThomasSolve.backward`` re-solves the transposed band system each step
(dendra/models/integrators/triton/t_kernel_thomas.py:147-189) and autograd
differentiates the gating arithmetic.  It can checkpoint the run
(``Population.longrun_checkpointed``, core.py:7802; ``dn.compile(...,
checkpoint_every=K)``, func/_convenience.py:959,1005-1007).

``CableFit`` keeps that contract behind a ``torch.autograd.Function`` and
replaces its arithmetic with three HIP launches per evaluation:

* forward: one persistent launch over all steps; it accumulates the loss
  and writes a state checkpoint every ``ckpt_every`` steps;
* backward, for each checkpoint segment from the last to the first: a
  recomputation launch from the checkpoint that writes the segment's
  trajectory, then the reverse-time adjoint launch (transposed partition
  solve + gating vector-Jacobian products) which carries the adjoint state
  across segments and accumulates dL/dsNa, dL/dsK and dL/d(activating knots);
* a deterministic CSR product maps the knot gradients back to the Ve table.

Loss (over the post-step voltages V_{n+1}, n = 0..n_steps-1):
    L = 1/(B S N) sum_{n,b,s} (V_{n+1}[b, site_s] - target[n, b, s])^2
      + margin_weight/(B N) sum_{n,b} sign_b sigmoid((V_{n+1}[b, margin] - theta)/tau)
The second term is the smooth above-threshold gate of the reference's spike
detector (mod/SPIKEDETECT.py:133-143, sigmoid((v - thr)/tau)).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .model import (HH, CableSpec, cable_workspace, hh_steady_state, plan_cables,
                    plan_cells_per_task, slot_map)

_NP = {torch.float32: np.float32, torch.float64: np.float64}


def knot_weights(knot_times, dt, n_steps):
    """Interpolation segment k_n and weight alpha_n of t_n = n dt."""
    tk = np.asarray(knot_times, dtype=np.float64)
    if tk.ndim != 1 or len(tk) < 2 or np.any(np.diff(tk) <= 0):
        raise ValueError("knot_times must be at least two strictly increasing times")
    t = np.arange(n_steps, dtype=np.float64) * dt
    if t[0] < tk[0] or t[-1] > tk[-1]:
        raise ValueError("knot_times must cover every start-of-step time of the run")
    k = np.clip(np.searchsorted(tk, t, side="right") - 1, 0, len(tk) - 2)
    return k.astype(np.int32), (t - tk[k]) / (tk[k + 1] - tk[k])


class CableFit:
    """Loss of a batch of HH cables driven by a Ve knot table, with its adjoint.

    spec: geometry (diam, dx, rhoa, cm per compartment; no lead/iamp).
    sites: (S,) compartment indices recorded in every cell.
    targets: (n_steps, n_cells, S) target voltages after each step.
    ve (at call time): (K, Q) shared by all cells (``shared_ve``, equal K)
    or (total, Q) per compartment in spec order; knot_times (Q,) in ms.
    (waves, rmax): (1, 4) or (1, 8) packs 64/P cells of P lanes into a
    wavefront with up to rmax rows per lane; (2, 2) or (4, 1) spreads each
    cell over a workgroup of that many wavefronts, so that a few hundred
    cells still fill every SIMD.
    """

    def __init__(self, spec: CableSpec, dt: float, n_steps: int, sites, targets, *,
                 knot_times=None, n_knots: int = 64, margin_site: int | None = None,
                 margin_sign=None, margin_weight: float = 0.0, theta: float = 0.0,
                 tau: float = 1.0, ckpt_every: int = 500, hh: HH | None = None,
                 shared_ve: bool = True, dtype=torch.float32, device="cuda",
                 rmax: int = 4, waves: int = 1):
        lib = _lib.load()
        if spec.lead is not None or spec.iamp is not None:
            raise ValueError("CableFit drives cables through the Ve table only")
        if (waves, rmax) not in ((1, 4), (1, 8), (2, 2), (4, 1)):
            raise ValueError("(waves, rmax) must be (1, 4), (1, 8), (2, 2) or (4, 1)")
        if n_steps < 1 or ckpt_every < 1:
            raise ValueError("n_steps and ckpt_every must be positive")
        if not tau > 0:
            raise ValueError("tau must be positive")
        self.lib = lib
        self.spec, self.dt, self.n_steps = spec, float(dt), int(n_steps)
        self.hh = hh or HH()
        self.dtype, self.device = dtype, torch.device(device)
        _lib.require_gpu(torch.empty(0, device=self.device))
        self.rmax, self.waves, self.ckpt_every = rmax, int(waves), int(ckpt_every)
        self.theta, self.tau = float(theta), float(tau)
        sizes = spec.sizes
        B = spec.n_cells
        self.shared_ve = bool(shared_ve)
        if self.shared_ve and len(set(sizes.tolist())) != 1:
            raise ValueError("a shared Ve table needs equal compartment counts")
        width = 64 * self.waves
        # one wavefront per task holds 64 / P cells; with several wavefronts
        # per task each cell spans all of them
        plan = (plan_cables(sizes, rmax, 64) if self.waves == 1
                else plan_cells_per_task(sizes, rmax, width))
        self.plan = plan
        nt = len(plan.task_p)
        R = rmax
        self.n_tasks, self.slots = nt, nt * width * R

        # slot of every compartment (spec order)
        comp_slot = slot_map(plan, sizes, spec.off, R, width)
        self.comp_slot = comp_slot

        ws = cable_workspace(spec)
        npd = _NP[dtype]

        def slot_arr(x, fill=0.0, dt_=npd):
            out = np.full(self.slots, fill, dtype=dt_)
            out[comp_slot] = x
            return out

        dev = self.device
        tens = lambda a, d=dtype: torch.as_tensor(np.ascontiguousarray(a), dtype=d,
                                                  device=dev)
        self.gl = tens(slot_arr(ws.gL))
        self.grp = tens(slot_arr(ws.gRp))
        self.sc = tens(slot_arr(ws.scale))
        self.task_nr = tens(plan.task_nr, torch.int32)
        self.task_p = tens(plan.task_p, torch.int32)
        self.comp_slot_t = tens(comp_slot, torch.int64)

        # sites, targets, margin
        sites = np.asarray(sites, dtype=np.int64).reshape(-1)
        S = len(sites)
        if S == 0 or sites.min() < 0 or sites.max() >= sizes.min():
            raise ValueError("sites must index compartments present in every cell")
        targets = np.asarray(targets)
        if targets.shape != (self.n_steps, B, S):
            raise ValueError(f"targets must have shape {(self.n_steps, B, S)}")
        site_col = np.full(self.slots, -1, dtype=np.int32)
        cells = np.arange(B)
        for s in range(S):
            site_col[comp_slot[spec.off[:-1] + sites[s]]] = cells * S + s
        self.site = tens(site_col, torch.int32)
        self.target = tens(targets.reshape(self.n_steps, B * S))
        self.tstride = B * S
        self.mse_scale = 1.0 / (B * S * self.n_steps)
        mw = np.zeros(self.slots, dtype=np.float64)
        if margin_site is not None and margin_weight != 0.0:
            sign = np.ones(B) if margin_sign is None else np.asarray(margin_sign, np.float64)
            if sign.shape != (B,):
                raise ValueError("margin_sign must have one entry per cell")
            mw[comp_slot[spec.off[:-1] + sites[int(margin_site)]]] = (
                margin_weight * sign / (B * self.n_steps))
        self.mw = tens(mw)

        # knot interpolation
        if knot_times is None:
            knot_times = np.linspace(0.0, self.n_steps * self.dt, n_knots)
        self.knot_times = np.asarray(knot_times, dtype=np.float64)
        self.n_knots = len(self.knot_times)
        kk, ka = knot_weights(self.knot_times, self.dt, self.n_steps)
        self.knot_k = tens(kk, torch.int32)
        self.knot_a = tens(ka)

        # activating-function operator as CSR in both directions
        K0 = int(sizes[0])
        rows_of = (np.tile(np.arange(K0), B) if self.shared_ve
                   else np.arange(spec.total))
        self.ve_rows = K0 if self.shared_ve else spec.total
        j = np.concatenate([np.arange(k) for k in sizes])
        Kc = np.repeat(sizes, sizes)
        has_r, has_l = j < Kc - 1, j > 0
        slot_e = np.concatenate([comp_slot[has_r], comp_slot, comp_slot[has_l]])
        row_e = np.concatenate([rows_of[has_r] + 1, rows_of, rows_of[has_l] - 1])
        w_e = np.concatenate([ws.gL[has_r], -(ws.gL * has_r) - (ws.gRp * has_l),
                              ws.gRp[has_l]])
        o = np.lexsort((row_e, slot_e))
        self.fwd_ptr = tens(np.concatenate([[0], np.cumsum(np.bincount(
            slot_e[o], minlength=self.slots))]), torch.int64)
        self.fwd_idx = tens(row_e[o], torch.int32)
        self.fwd_w = tens(w_e[o])
        o = np.lexsort((slot_e, row_e))
        self.bwd_ptr = tens(np.concatenate([[0], np.cumsum(np.bincount(
            row_e[o], minlength=self.ve_rows))]), torch.int64)
        self.bwd_idx = tens(slot_e[o], torch.int32)
        self.bwd_w = tens(w_e[o])

        # initial state (resting), checkpoints, loss partials
        v0 = spec.v_init
        m0, h0, n0 = hh_steady_state(v0)
        x0 = np.stack([np.full(self.slots, v0), slot_arr(m0), slot_arr(h0), slot_arr(n0)])
        self.x0 = tens(x0)
        self.n_seg = -(-self.n_steps // self.ckpt_every)
        self.ckpt = torch.empty((self.n_steps // self.ckpt_every + 1, 4, self.slots),
                                dtype=dtype, device=dev)
        self.loss_part = torch.zeros((nt * self.waves, 2), dtype=dtype, device=dev)
        self.hh_arr = self.hh.array().astype(npd)
        self._traj = None

    # -- parameters <-> slots ------------------------------------------------
    def _params_to_slots(self, sna, sk, ve):
        dev, dt = self.device, self.dtype
        out = []
        for x in (sna, sk):
            x = torch.as_tensor(x, device=dev).detach().to(dt).reshape(-1)
            if x.numel() != self.spec.total:
                raise ValueError("conductance scales need one value per compartment")
            s = torch.ones(self.slots, dtype=dt, device=dev)
            s[self.comp_slot_t] = x
            out.append(s)
        ve = torch.as_tensor(ve, device=dev).detach().to(dt).contiguous()
        if ve.shape != (self.ve_rows, self.n_knots):
            raise ValueError(f"ve must have shape {(self.ve_rows, self.n_knots)}")
        act = torch.empty((self.n_knots, self.slots), dtype=dt, device=dev)
        self._csr(self.slots, self.fwd_ptr, self.fwd_idx, self.fwd_w, ve, self.n_knots, 1,
                  act, 1, self.slots)
        return out[0], out[1], act

    def _csr(self, n_out, ptr, idx, w, inp, in_row, in_k, out, out_row, out_k):
        f64 = self.dtype == torch.float64
        fn = self.lib.dn_fit_csr_knots_f64 if f64 else self.lib.dn_fit_csr_knots_f32
        st = fn(n_out, self.n_knots, _lib.ptr(ptr), _lib.ptr(idx), _lib.ptr(w),
                _lib.ptr(inp), in_row, in_k, _lib.ptr(out), out_row, out_k,
                torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(st, "knot CSR product")

    def _common(self, sna_s, sk_s, act, step0, n):
        return [self.n_tasks, self.rmax, self.waves, _lib.ptr(self.task_nr),
                _lib.ptr(self.task_p),
                _lib.ptr(self.gl), _lib.ptr(self.grp), _lib.ptr(self.sc), _lib.ptr(sna_s),
                _lib.ptr(sk_s), _lib.ptr(act), self.n_knots, _lib.ptr(self.knot_k),
                _lib.ptr(self.knot_a), _lib.ptr(self.site), _lib.ptr(self.mw),
                _lib.ptr(self.target), self.tstride, self.mse_scale, self.theta, self.tau,
                step0, n, self.dt, self.hh_arr.ctypes.data]

    def _fwd(self, sna_s, sk_s, act, x0, step0, n, *, xout=None, ckpt=None, traj=None,
             loss=None, stream=None):
        f64 = self.dtype == torch.float64
        fn = self.lib.dn_fit_forward_f64 if f64 else self.lib.dn_fit_forward_f32
        strm = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        st = fn(*self._common(sna_s, sk_s, act, step0, n), _lib.ptr(x0), _lib.ptr(xout),
                _lib.ptr(ckpt), self.ckpt_every, _lib.ptr(traj), _lib.ptr(loss), strm)
        _lib.check(st, "fit forward")

    def _adj(self, sna_s, sk_s, act, traj, adj, g_sna, g_sk, g_act, step0, n,
             stream=None):
        f64 = self.dtype == torch.float64
        fn = self.lib.dn_fit_adjoint_f64 if f64 else self.lib.dn_fit_adjoint_f32
        strm = (stream or torch.cuda.current_stream(self.device)).cuda_stream
        st = fn(*self._common(sna_s, sk_s, act, step0, n), _lib.ptr(traj), _lib.ptr(adj),
                _lib.ptr(g_sna), _lib.ptr(g_sk), _lib.ptr(g_act), strm)
        _lib.check(st, "fit adjoint")

    # -- public --------------------------------------------------------------
    def forward(self, sna, sk, ve):
        """Loss (0-d tensor) of the run; leaves checkpoints for ``backward``."""
        sna_s, sk_s, act = self._params_to_slots(sna, sk, ve)
        self._fwd(sna_s, sk_s, act, self.x0, 0, self.n_steps, ckpt=self.ckpt,
                  loss=self.loss_part)
        self._saved = (sna_s, sk_s, act)
        return self.loss_part.sum()

    def backward(self):
        """Gradients (dL/dsNa, dL/dsK, dL/dve) of the last ``forward``.

        Segments run from the last to the first.  The recomputation of
        segment c-1 (from its checkpoint) runs on a second stream while the
        adjoint of segment c runs, with two trajectory buffers, so that the
        two latency-bound launches share the GPU."""
        sna_s, sk_s, act = self._saved
        dev, dt = self.device, self.dtype
        if self._traj is None:
            self._traj = [torch.empty((self.ckpt_every + 1, 4, self.slots), dtype=dt,
                                      device=dev) for _ in range(2)]
            self._streams = (torch.cuda.Stream(dev), torch.cuda.Stream(dev))
        caller = torch.cuda.current_stream(dev)
        s_adj, s_rec = self._streams
        adj = torch.zeros((4, self.slots), dtype=dt, device=dev)
        g_sna = torch.zeros(self.slots, dtype=dt, device=dev)
        g_sk = torch.zeros(self.slots, dtype=dt, device=dev)
        g_act = torch.zeros((self.n_knots, self.slots), dtype=dt, device=dev)
        start = torch.cuda.Event()
        start.record(caller)
        s_adj.wait_event(start)
        s_rec.wait_event(start)
        seg = lambda c: (c * self.ckpt_every,
                         min(self.ckpt_every, self.n_steps - c * self.ckpt_every))
        rec_done = {}
        adj_done = {}
        last = self.n_seg - 1

        def recompute(c):
            if c + 2 in adj_done:
                s_rec.wait_event(adj_done[c + 2])   # its buffer is free again
            s0, n = seg(c)
            self._fwd(sna_s, sk_s, act, self.ckpt[c], s0, n, traj=self._traj[c % 2],
                      stream=s_rec)
            ev = torch.cuda.Event()
            ev.record(s_rec)
            rec_done[c] = ev

        recompute(last)
        for c in range(last, -1, -1):
            if c >= 1:
                recompute(c - 1)
            s_adj.wait_event(rec_done[c])
            s0, n = seg(c)
            self._adj(sna_s, sk_s, act, self._traj[c % 2], adj, g_sna, g_sk, g_act, s0, n,
                      stream=s_adj)
            ev = torch.cuda.Event()
            ev.record(s_adj)
            adj_done[c] = ev
        caller.wait_event(adj_done[0])
        caller.wait_event(rec_done[0])
        for t in (adj, g_sna, g_sk, g_act, sna_s, sk_s, act, *self._traj):
            t.record_stream(s_adj)
            t.record_stream(s_rec)
        g_ve = torch.empty((self.ve_rows, self.n_knots), dtype=dt, device=dev)
        self._csr(self.ve_rows, self.bwd_ptr, self.bwd_idx, self.bwd_w, g_act, 1,
                  self.slots, g_ve, self.n_knots, 1)
        self.last_adjoint = adj
        return g_sna[self.comp_slot_t], g_sk[self.comp_slot_t], g_ve

    def loss(self, sna, sk, ve):
        """Differentiable loss: gradients flow to whichever inputs require them."""
        return _FitLoss.apply(self, sna, sk, ve)

    def final_state(self, sna, sk, ve):
        """(v, m, h, n) after the run, each (total,) in spec order."""
        sna_s, sk_s, act = self._params_to_slots(sna, sk, ve)
        xout = torch.empty((4, self.slots), dtype=self.dtype, device=self.device)
        self._fwd(sna_s, sk_s, act, self.x0, 0, self.n_steps, xout=xout)
        return xout[:, self.comp_slot_t]


class _FitLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fit, sna, sk, ve):
        ctx.fit = fit
        ctx.shapes = (sna.shape, sk.shape)
        return fit.forward(sna, sk, ve)

    @staticmethod
    def backward(ctx, g):
        g_sna, g_sk, g_ve = ctx.fit.backward()
        s1, s2 = ctx.shapes
        return None, (g * g_sna).reshape(s1), (g * g_sk).reshape(s2), g * g_ve
