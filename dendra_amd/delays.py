"""NetCon delivery with fractional (two-tap) delays on the GPU.

Drop-in for the reference's dense interpolated-delay transition
(dendra/models/networks/_event_delivery.py:200-310) in its two-tap form, with
the ordered native enqueue (networks/_two_tap_enqueue_cpu.cpp:26-152) as a
HIP kernel:

* ``linear_delay_taps`` / ``delay_interpolation``: positions
  ``lam = max(delay / dt, 1)`` split into ``floor(lam)`` and the upper-tap
  fraction ``lam - floor(lam)`` (differentiable in the delays);
* ``dense_interpolated_delay_transition(queue, slot, global_step, weights,
  connection_gates, delays, dt, *, post_idx, depth, width, taps=2, ...)``:
  clears today's ring row, then adds ``w g (1 - f)`` and ``w g f`` at rows
  ``slot + floor(lam)`` and ``slot + floor(lam) + 1`` (modulo depth) of every
  connection's target column; returns the reference's ``DenseTransition``.
  The result is bitwise the reference's ordered enqueue (all lower taps, then
  all upper taps, each cell in ascending connection order).  Gradients flow
  to the queue (zero on the cleared row), the weights, the gates and, through
  the fraction, the delays.  Inputs are never modified.

The three-tap (Gaussian) form is not provided on the GPU.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np
import torch

from . import _lib


class DenseTransition(NamedTuple):
    queue: torch.Tensor
    slot: torch.Tensor
    global_step: torch.Tensor
    event_queue: torch.Tensor | None


def linear_delay_taps(positions):
    """Lower index and upper-tap fraction of positions in steps
    (_event_delivery.py:200-203)."""
    lower = positions.floor()
    return lower.to(torch.long), positions - lower


def delay_interpolation(delays, dt, *, dtype, taps=2, sigma=0.0):
    """Offsets and fractions of the two-tap delay split
    (_event_delivery.py:206-216, two-tap branch)."""
    if taps != 2:
        raise NotImplementedError("the GPU delay transition supports two taps")
    dt = torch.as_tensor(dt, dtype=dtype, device=delays.device)
    lam = (delays.to(dtype) / dt.reshape(1)).clamp_min(1.0)
    return linear_delay_taps(lam)


_plans: dict = {}


def two_tap_plan(post_idx: torch.Tensor, width: int):
    """Target-major CSR of the connections (ascending connection index within
    each target) and int32 targets; cached per post_idx tensor version."""
    key = (post_idx.data_ptr(), post_idx._version, post_idx.numel(), int(width),
           post_idx.device)
    plan = _plans.get(key)
    if plan is None:
        post = post_idx.detach().to("cpu", torch.int64).numpy().reshape(-1)
        if post.size and (post.min() < 0 or post.max() >= width):
            raise ValueError("post_idx entries must lie in [0, width)")
        order = np.argsort(post, kind="stable").astype(np.int32)
        ptr = np.zeros(width + 1, dtype=np.int32)
        np.cumsum(np.bincount(post, minlength=width), out=ptr[1:])
        dev = post_idx.device
        i32 = lambda a: torch.as_tensor(a if a.size else np.zeros(1, np.int32), device=dev)
        plan = (i32(ptr), i32(order), i32(post.astype(np.int32)))
        if len(_plans) > 64:
            _plans.clear()
        _plans[key] = plan
    return plan


def _fn(name, like):
    return getattr(_lib.load(), f"{name}_{'f64' if like.dtype == torch.float64 else 'f32'}")


class _TwoTapEnqueue(torch.autograd.Function):
    @staticmethod
    def forward(ctx, queue, slot, weights, gates, fraction, off, in_ptr, in_edge, post):
        depth, width = map(int, queue.shape)
        out = queue.clone(memory_format=torch.contiguous_format)
        w, g, f = (v.contiguous() for v in (weights, gates, fraction))
        p = _lib.ptr
        st = _fn("dn_two_tap_enqueue", out)(p(out), depth, width, p(slot), p(in_ptr),
                                            p(in_edge), p(w), p(g), p(f), p(off),
                                            _lib.stream_handle(out.device))
        _lib.check(st, "two-tap enqueue")
        ctx.save_for_backward(slot, w, g, f, off, post)
        ctx.shape = (depth, width)
        return out

    @staticmethod
    def backward(ctx, grad):
        slot, w, g, f, off, post = ctx.saved_tensors
        depth, width = ctx.shape
        grad = grad.contiguous()
        dq = grad.clone()
        dq.index_fill_(0, slot.remainder(depth), 0)
        n = w.numel()
        dw, dg, df = (torch.empty_like(w) for _ in range(3))
        p = _lib.ptr
        st = _fn("dn_two_tap_vjp", grad)(p(grad), depth, width, p(slot), n, p(post), p(w),
                                         p(g), p(f), p(off), p(dw), p(dg), p(df),
                                         _lib.stream_handle(grad.device))
        _lib.check(st, "two-tap enqueue cotangent")
        return dq, None, dw, dg, df, None, None, None, None


def dense_interpolated_delay_transition(queue, slot, global_step, weights, connection_gates,
                                        delays, dt, *, post_idx, depth, width, taps=2,
                                        sigma=0.0, delay_taps=None, _clone_once=False):
    """One dense-ring step with fractional delays (see the module doc).  Read
    today's row (``queue[slot]``) before calling: this transition clears it."""
    if taps != 2:
        raise NotImplementedError("the GPU delay transition supports two taps")
    if queue.ndim != 2 or tuple(queue.shape) != (int(depth), int(width)):
        raise ValueError(f"queue must have shape ({depth}, {width}), got {tuple(queue.shape)}")
    if queue.device.type != "cuda":
        raise ValueError("the GPU delay transition requires GPU tensors")
    if queue.dtype not in (torch.float32, torch.float64):
        raise TypeError("the GPU delay transition supports float32 and float64")
    n = post_idx.numel()
    for name, v in (("weights", weights), ("connection_gates", connection_gates)):
        if v.numel() != n or v.dtype != queue.dtype or v.device != queue.device:
            raise ValueError(f"{name} must hold one {queue.dtype} value per connection "
                             f"on {queue.device}")
    if delay_taps is None:
        delay_taps = delay_interpolation(delays, dt, dtype=queue.dtype, taps=2, sigma=sigma)
    offsets, fraction = delay_taps
    off = offsets.reshape(-1).remainder(int(depth)).to(torch.int32).contiguous()
    fraction = fraction.reshape(-1).to(queue.dtype)
    in_ptr, in_edge, post = two_tap_plan(post_idx.reshape(-1), int(width))
    slot_t = slot.reshape(1).to(torch.int64)
    out = _TwoTapEnqueue.apply(queue, slot_t, weights.reshape(-1),
                               connection_gates.reshape(-1), fraction, off, in_ptr, in_edge,
                               post)
    return DenseTransition(out, (slot + 1).remainder(depth), global_step + 1, None)
