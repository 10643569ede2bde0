"""NetCon delivery with fractional (interpolated) delays on the GPU.

Drop-in for the reference's dense interpolated-delay transition
(dendra/models/networks/_event_delivery.py:200-310), with its ordered native
two-tap enqueue (networks/_two_tap_enqueue_cpu.cpp:26-152) and its three-tap
scatter as HIP kernels:

* ``linear_delay_taps`` / ``delay_interpolation``: positions
  ``lam = max(delay / dt, 1)``; two taps split into ``floor(lam)`` and the
  upper-tap fraction ``lam - floor(lam)``; three taps give offsets
  ``floor(lam) + (-1, 0, 1)`` and softmax weights of
  ``-((lam - offset) / (sigma + 1e-6))^2 / 2`` (differentiable in the delays);
* ``dense_interpolated_delay_transition(queue, slot, global_step, weights,
  connection_gates, delays, dt, *, post_idx, depth, width, taps, sigma, ...)``
  clears today's ring row, then adds ``w g`` times each tap weight at the
  tap rows (modulo depth) of every connection's target column, and returns
  the reference's ``DenseTransition``.  Each ring cell receives its
  contributions in the reference's order (two taps: all lower taps, then all
  upper taps; three taps: connection-major), so the result is bitwise the
  reference's.  Gradients flow to the queue (zero on the cleared row), the
  weights, the gates and, through the tap weights, the delays.  Inputs are
  never modified.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np
import torch
from torch.utils.weak import WeakIdKeyDictionary

from . import _lib


class DenseTransition(NamedTuple):
    queue: torch.Tensor
    slot: torch.Tensor
    global_step: torch.Tensor
    event_queue: torch.Tensor | None


class DenseDue(NamedTuple):
    payload: torch.Tensor
    event_counts: torch.Tensor | None


def linear_delay_taps(positions):
    """Lower index and upper-tap fraction of positions in steps
    (_event_delivery.py:200-203)."""
    lower = positions.floor()
    return lower.to(torch.long), positions - lower


def delay_interpolation(delays, dt, *, dtype, taps, sigma):
    """Tap offsets and coefficients (_event_delivery.py:206-223): two taps
    return (floor, upper fraction); three taps return the offsets
    floor + (-1, 0, 1) and their normalised Gaussian weights."""
    dt = torch.as_tensor(dt, dtype=dtype, device=delays.device)
    lam = (delays.to(dtype) / dt.reshape(1)).clamp_min(1.0)
    if taps == 2:
        return linear_delay_taps(lam)
    if taps != 3:
        raise ValueError("taps must be 2 or 3")
    lower = lam.floor().to(torch.long)
    offsets = torch.stack([lower - 1, lower, lower + 1], dim=-1)
    centers = offsets.to(dtype)
    fractions = torch.softmax(
        -0.5 * ((lam.unsqueeze(-1) - centers) / (sigma + 1e-6)) ** 2, dim=-1)
    return offsets, fractions


# Plans are cached per post_idx tensor object (weakly, so a freed tensor whose
# storage is reused can never hand its plan to another) and checked against
# the tensor's version, size and the ring width.
_plans = WeakIdKeyDictionary()


def delay_plan(post_idx: torch.Tensor, width: int):
    """Target-major CSR of the connections (ascending connection index within
    each target) and int32 targets."""
    stamp = (post_idx._version, post_idx.numel(), int(width), post_idx.device)
    hit = _plans.get(post_idx)
    if hit is not None and hit[0] == stamp:
        return hit[1]
    post = post_idx.detach().to("cpu", torch.int64).numpy().reshape(-1)
    if post.size and (post.min() < 0 or post.max() >= width):
        raise ValueError("post_idx entries must lie in [0, width)")
    order = np.argsort(post, kind="stable").astype(np.int32)
    ptr = np.zeros(width + 1, dtype=np.int32)
    np.cumsum(np.bincount(post, minlength=width), out=ptr[1:])
    dev = post_idx.device
    i32 = lambda a: torch.as_tensor(a if a.size else np.zeros(1, np.int32), device=dev)
    plan = (i32(ptr), i32(order), i32(post.astype(np.int32)))
    _plans[post_idx] = (stamp, plan)
    return plan


# Index tensors derived from a caller's tensor (tap rows reduced modulo the
# ring depth, as int32) are cached the same way: per tensor object, checked
# against its version and size and the derivation's own key, so a step whose
# delays did not change reuses them instead of recomputing.
_derived = WeakIdKeyDictionary()


def derived_index(t: torch.Tensor, key, make):
    """make(t) cached per (t, t._version, t.numel(), key)."""
    stamp = (t._version, t.numel(), t.device, key)
    hit = _derived.get(t)
    if hit is not None and hit[0] == stamp:
        return hit[1]
    out = make(t)
    _derived[t] = (stamp, out)
    return out


def csr_positions(in_edge: torch.Tensor) -> torch.Tensor:
    """pos[in_edge[p]] = p (int32): each connection's position in the
    target-major CSR, cached with the plan; the record-packing operators
    write each connection's record there so that their target-major passes
    read the records in order."""
    def make(t):
        pos = torch.empty_like(t)
        pos[t.long()] = torch.arange(t.numel(), dtype=torch.int32, device=t.device)
        return pos
    return derived_index(in_edge, "csr positions", make)


def _fn(name, like):
    return getattr(_lib.load(), f"{name}_{'f64' if like.dtype == torch.float64 else 'f32'}")


class _InterpEnqueue(torch.autograd.Function):
    @staticmethod
    def forward(ctx, queue, slot, weights, gates, fraction, base, in_ptr, in_edge, post, taps):
        depth, width = map(int, queue.shape)
        qin = queue.contiguous()
        out = torch.empty_like(qin)
        w, g = weights.contiguous(), gates.contiguous()
        f = None if fraction is None else fraction.contiguous()  # None: one tap
        p, lib, n = _lib.ptr, _lib.load(), w.numel()
        nbytes = lib.dn_interp_delay_work_bytes(n, taps, int(out.dtype == torch.float64))
        work = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=out.device)
        pos = csr_positions(in_edge) if n else in_edge
        st = _fn("dn_interp_delay_enqueue_ws", out)(p(qin), p(out), depth, width, p(slot), taps,
                                                    p(in_ptr), p(in_edge), p(pos), p(w), p(g),
                                                    p(f), p(base), n, p(work), nbytes,
                                                    _lib.stream_handle(out.device))
        _lib.check(st, "interpolated-delay enqueue")
        ctx.save_for_backward(slot, w, g, f, base, post)
        ctx.shape, ctx.taps = (depth, width), taps
        return out

    @staticmethod
    def backward(ctx, grad):
        slot, w, g, f, base, post = ctx.saved_tensors
        depth, width = ctx.shape
        grad = grad.contiguous()
        dq = grad.clone()
        dq.index_fill_(0, slot.remainder(depth), 0)
        n = w.numel()
        dw, dg = torch.empty_like(w), torch.empty_like(w)
        df = None if f is None else torch.empty_like(f)
        p = _lib.ptr
        st = _fn("dn_interp_delay_vjp", grad)(p(grad), depth, width, p(slot), ctx.taps, n,
                                              p(post), p(w), p(g), p(f), p(base), p(dw),
                                              p(dg), p(df), _lib.stream_handle(grad.device))
        _lib.check(st, "interpolated-delay enqueue cotangent")
        return dq, None, dw, dg, df, None, None, None, None, None


def dense_interpolated_delay_transition(queue, slot, global_step, weights, connection_gates,
                                        delays, dt, *, post_idx, depth, width, taps, sigma,
                                        delay_taps=None, _clone_once=False):
    """One dense-ring step with fractional delays (see the module doc).  Read
    today's row (``queue[slot]``) before calling: this transition clears it."""
    if taps not in (2, 3):
        raise ValueError("taps must be 2 or 3")
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
        delay_taps = delay_interpolation(delays, dt, dtype=queue.dtype, taps=taps, sigma=sigma)
    offsets, fraction = delay_taps
    base = derived_index(
        offsets, ("first-tap rows", int(depth), taps),
        lambda o: (o.reshape(-1) if taps == 2 else o.reshape(-1, 3)[:, 0])
        .remainder(int(depth)).to(torch.int32).contiguous())
    fraction = fraction.reshape(-1).to(queue.dtype)
    in_ptr, in_edge, post = delay_plan(post_idx, int(width))
    slot_t = slot.reshape(1).to(torch.int64)
    out = _InterpEnqueue.apply(queue, slot_t, weights.reshape(-1),
                               connection_gates.reshape(-1), fraction, base, in_ptr, in_edge,
                               post, int(taps))
    return DenseTransition(out, (slot + 1).remainder(depth), global_step + 1, None)


# -- fixed integer delays: the default dense backend ------------------------
#
# Drop-ins for dense_due, gather_source_gates, connection_event_counts,
# _fixed_delay_indices, _clear_dense_slot, _enqueue_dense_values and
# dense_fixed_delay_transition (_event_delivery.py:30-199), which
# NetCon.advance_non_diff_dense_uniform/_mixed and the dense training path
# call (netcon.py:5183-5205).  The scatters run as HIP kernels whose every
# ring cell receives its values in connection order, bitwise the sequential
# scatter_add; rows and index arithmetic are plain tensor indexing.


def dense_due(queue, slot, event_queue=None):
    """Today's target payload (and per-connection event counts): independent
    copies of row ``slot`` (_event_delivery.py:30-43)."""
    payload = queue.index_select(0, slot.reshape(1)).squeeze(0)
    events = None if event_queue is None else \
        event_queue.index_select(0, slot.reshape(1)).squeeze(0)
    return DenseDue(payload, events)


def gather_source_gates(source_gates, pre_idx):
    """Source gates in connection order, values and gradients unchanged."""
    return source_gates.flatten().index_select(0, pre_idx)


def connection_event_counts(intrinsic_gates, scheduled_counts=None):
    """int32 event counts: positive intrinsic gates plus scheduled events."""
    counts = intrinsic_gates.gt(0).to(torch.int32)
    return counts if scheduled_counts is None else counts + scheduled_counts


def _fixed_delay_indices(slot, post_idx, delay_steps, *, depth, width, flat_offsets=None,
                         uniform_delay=None):
    """Flat ring index of each connection's destination cell: row
    (slot + delay) mod depth, column post (_event_delivery.py:46-59)."""
    if uniform_delay is not None:
        return ((slot + uniform_delay) % depth) * width + post_idx
    if flat_offsets is not None:
        return (slot * width + flat_offsets) % (depth * width)
    return ((slot + delay_steps) % depth) * width + post_idx


def _clear_dense_slot(queue, slot, *, inplace=False):
    """Zero row ``slot``: in place on caller-owned storage, else a new ring."""
    if inplace:
        return queue.index_fill_(0, slot.reshape(-1), 0)
    keep = torch.arange(queue.shape[0], device=queue.device).ne(slot.reshape(1))
    return queue * keep.to(queue.dtype).reshape(-1, *([1] * (queue.dim() - 1)))


class _OrderedScatter(torch.autograd.Function):
    @staticmethod
    def forward(ctx, queue, destination, values, inplace):
        out = queue if inplace else queue.contiguous().clone()
        n = destination.numel()
        lib = _lib.load()
        f64 = out.dtype == torch.float64
        nbytes = lib.dn_ordered_scatter_work_bytes(n, out.numel(), int(f64))
        work = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=out.device)
        fn = lib.dn_ordered_scatter_add_f64 if f64 else lib.dn_ordered_scatter_add_f32
        st = fn(_lib.ptr(out), out.numel(), _lib.ptr(destination), _lib.ptr(values), n,
                _lib.ptr(work), work.numel(), _lib.stream_handle(out.device))
        _lib.check(st, "ordered scatter-add into the delay ring")
        ctx.save_for_backward(destination)
        if inplace:
            ctx.mark_dirty(out)
        return out

    @staticmethod
    def backward(ctx, grad):
        (destination,) = ctx.saved_tensors
        return grad, None, grad.reshape(-1).index_select(0, destination), None


def _enqueue_dense_values(queue, destination, values, *, inplace=False):
    """queue.flat[destination[i]] += values[i] in index order, bitwise the
    sequential scatter_add (_event_delivery.py:61-76).  ``inplace`` writes
    into the caller's contiguous ring; otherwise a new ring is returned."""
    if queue.device.type != "cuda":
        raise ValueError("the GPU ring scatter requires GPU tensors")
    if queue.dtype not in (torch.float32, torch.float64):
        raise TypeError("the GPU ring scatter supports float32 and float64 rings")
    if inplace and not queue.is_contiguous():
        raise ValueError("an in-place ring must be contiguous")
    dest = destination.reshape(-1).to(torch.int64).contiguous()
    vals = values.reshape(-1).to(queue.dtype).contiguous()
    if vals.numel() != dest.numel():
        raise ValueError("destination and values must have one entry each")
    return _OrderedScatter.apply(queue, dest, vals, bool(inplace))


def _event_ring_step(event_queue, slot, delay_steps, event_counts):
    if event_queue.dtype != torch.int32 or event_counts.dtype != torch.int32:
        raise TypeError("event_queue and event_counts must be int32")
    n = event_counts.numel()
    if event_queue.ndim != 2 or event_queue.shape[1] != n:
        raise ValueError("event_queue must be [depth, connections]")
    qin = event_queue.contiguous()
    out = torch.empty_like(qin)
    delay = delay_steps.reshape(-1).to(torch.int64).contiguous()
    st = _lib.load().dn_dense_event_step(
        _lib.ptr(qin), _lib.ptr(out), qin.shape[0], n, _lib.ptr(slot), _lib.ptr(delay),
        _lib.ptr(event_counts.reshape(-1).contiguous()), _lib.stream_handle(qin.device))
    _lib.check(st, "event ring step")
    return out


def dense_fixed_delay_transition(queue, slot, global_step, weights, connection_gates, *,
                                 post_idx, delay_steps, depth, width, event_queue=None,
                                 event_counts=None, _clone_once=False):
    """Clear today's row, enqueue ``weights * connection_gates`` at row
    (slot + delay) mod depth of each connection's target, advance the
    counters (_event_delivery.py:125-199).  Pure: inputs are not modified and
    a new ring is returned (``_clone_once`` gives the same result).  Each ring
    cell receives its connections in connection order, so the ring is
    bitwise the reference's.  Gradients flow to the queue (zero on the
    cleared row), the weights and the gates.  With event tracking, the int32
    ``event_queue`` [depth, connections] gets each connection's count in its
    own column."""
    if (event_queue is None) != (event_counts is None):
        raise ValueError("event tracking needs both event_queue and event_counts")
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
    delay = torch.as_tensor(delay_steps, device=queue.device).reshape(-1)
    if delay.numel() == 1 and n != 1:
        delay = delay.expand(n)
    if delay.numel() != n:
        raise ValueError("delay_steps must hold one delay per connection (or one for all)")
    in_ptr, in_edge, post = delay_plan(post_idx, int(width))
    base = derived_index(delay, ("fixed rows", int(depth), n),
                         lambda d: d.to(torch.int64).expand(n).remainder(int(depth))
                         .to(torch.int32).contiguous())
    slot_t = slot.reshape(1).to(torch.int64)
    out = _InterpEnqueue.apply(queue, slot_t, weights.reshape(-1),
                               connection_gates.reshape(-1), None, base, in_ptr, in_edge,
                               post, 1)
    events = None
    if event_queue is not None:
        events = _event_ring_step(event_queue, slot_t, delay, event_counts)
    return DenseTransition(out, (slot + 1).remainder(depth), global_step + 1, events)
