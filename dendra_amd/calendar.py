"""Bounded pending calendar of the sparse NetCon backend, on the GPU.

Drop-in for the reference's pure pending calendar
(dendra/models/networks/_pending_calendar.py:16-161), the storage of the
``sparse_calendar`` delay backend (networks/netcon.py:3309-3450, 5223-5300):
pending deliveries as compact (slot, index, value) entries in enqueue order,
so storage follows the pending contributions instead of delay depth x width.

* ``pending_calendar_empty(capacity, *, device, dtype)``;
* ``pending_calendar_append(state, slots, indices, values, active=None)``:
  the active candidates, in order, after the valid prefix; all or nothing --
  if they do not fit, or the calendar already overflowed, nothing is written
  and the sticky overflow flag is set;
* ``pending_calendar_pop(state, slot, *, width, base=None, _scratch=None)``:
  the valid entries of ``slot`` are added per index, in position order, to a
  zero payload (or to ``base``, or into the donated ``_scratch``), and the
  rest are compacted to the front of fresh arrays with a zero tail;
* ``pending_calendar_from_legacy`` / ``pending_calendar_to_legacy``: the
  slot -> [(indices, values)] dictionary form.

The transitions are HIP stream compactions (dendra_amd/csrc/dn_calendar.hip);
the ordered per-index sums make a pop's payload bitwise the reference's
sequential CPU scatter_add.  Inputs are never modified (``_scratch`` aside).
Gradients flow to the values (and ``base``).  One difference of the donated
``_scratch`` form: the reference adds +0.0 at index 0 for every non-due entry,
which can only turn a -0.0 payload[0] into +0.0; this one does not.
"""

from __future__ import annotations

from typing import NamedTuple

import torch

from . import _lib

_KIND = {torch.float32: (0, "f32"), torch.float64: (1, "f64"), torch.int32: (2, "i32")}


class CalendarState(NamedTuple):
    slots: torch.Tensor
    indices: torch.Tensor
    values: torch.Tensor
    size: torch.Tensor
    overflow: torch.Tensor


def pending_calendar_empty(capacity, *, device, dtype):
    """An empty calendar of fixed capacity."""
    if not isinstance(capacity, int) or isinstance(capacity, bool) or capacity < 0:
        raise ValueError("Pending calendar capacity must be a nonnegative integer.")
    return CalendarState(
        torch.zeros(capacity, device=device, dtype=torch.long),
        torch.zeros(capacity, device=device, dtype=torch.long),
        torch.zeros(capacity, device=device, dtype=dtype),
        torch.zeros(1, device=device, dtype=torch.long),
        torch.zeros(1, device=device, dtype=torch.bool),
    )


def _kind(dtype):
    if dtype not in _KIND:
        raise TypeError("the GPU pending calendar supports float32, float64 and int32 values")
    return _KIND[dtype]


def _work(n, cap, dtype, device):
    kind, _ = _kind(dtype)
    nbytes = int(_lib.load().dn_calendar_work_bytes(int(n), int(cap), kind))
    if nbytes < 0:
        raise ValueError("pending calendar too large for the GPU kernels")
    return torch.empty(nbytes, dtype=torch.uint8, device=device), nbytes


def _check_state(state):
    if state.values.device.type != "cuda":
        raise ValueError("the GPU pending calendar requires GPU tensors")
    cap = state.values.numel()
    for name in ("slots", "indices", "values"):
        t = getattr(state, name)
        if t.shape != (cap,) or not t.is_contiguous() or t.device != state.values.device:
            raise ValueError(f"calendar {name} must be a contiguous ({cap},) tensor")
    if state.slots.dtype != torch.long or state.indices.dtype != torch.long:
        raise TypeError("calendar slots and indices must be int64")
    if state.size.numel() != 1 or state.overflow.numel() != 1:
        raise ValueError("calendar size and overflow must hold one element")
    return cap


class _Append(torch.autograd.Function):
    @staticmethod
    def forward(ctx, old_values, cand_values, state, c_slots, c_indices, active):
        cap, n = old_values.numel(), cand_values.numel()
        slots, indices = state.slots.clone(), state.indices.clone()
        values = old_values.clone()
        size = torch.empty_like(state.size)
        ovf = torch.empty(1, dtype=torch.bool, device=values.device)
        dest = torch.empty(max(n, 1), dtype=torch.long, device=values.device)
        work, nbytes = _work(n, cap, values.dtype, values.device)
        p = _lib.ptr
        fn = getattr(_lib.load(), f"dn_calendar_append_{_kind(values.dtype)[1]}")
        st = fn(p(slots), p(indices), p(values), cap, p(state.size),
                p(state.overflow.to(torch.bool).contiguous()), p(size), p(ovf), n, p(c_slots),
                p(c_indices), p(cand_values), p(active), p(dest), p(work), nbytes,
                _lib.stream_handle(values.device))
        _lib.check(st, "pending calendar append")
        ctx.save_for_backward(dest[:n])
        ctx.mark_non_differentiable(slots, indices, size, ovf)
        return slots, indices, values, size, ovf

    @staticmethod
    def backward(ctx, _gs, _gi, grad_values, _gz, _go):
        (dest,) = ctx.saved_tensors
        if grad_values is None:
            return None, None, None, None, None, None
        written = dest >= 0
        g_old = grad_values.clone()
        g_old.index_fill_(0, dest[written], 0)
        g_cand = torch.where(written, grad_values.index_select(0, dest.clamp_min(0)),
                             torch.zeros((), dtype=grad_values.dtype,
                                         device=grad_values.device)) \
            if dest.numel() else grad_values.new_zeros(0)
        return g_old, g_cand, None, None, None, None


def pending_calendar_append(state, slots, indices, values, active=None):
    """Append the active candidates in order, without changing any input."""
    cap = _check_state(state)
    dev = state.values.device
    n = slots.numel()
    if indices.numel() != n or values.numel() != n:
        raise ValueError("Pending calendar candidates must align.")
    act = torch.ones(n, dtype=torch.bool, device=dev) if active is None else \
        active.reshape(n).to(device=dev, dtype=torch.bool)
    c_slots = slots.reshape(n).to(device=dev, dtype=torch.long).contiguous()
    c_indices = indices.reshape(n).to(device=dev, dtype=torch.long).contiguous()
    c_values = values.reshape(n).to(device=dev, dtype=state.values.dtype).contiguous()
    if n == 0:
        c_slots = c_indices = torch.zeros(1, dtype=torch.long, device=dev)
    out = _Append.apply(state.values, c_values, state, c_slots, c_indices,
                        act.contiguous().view(torch.uint8) if n else
                        torch.zeros(1, dtype=torch.uint8, device=dev))
    return CalendarState(*out)


class _Pop(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, base, state, slot, width, scratch):
        cap = values.numel()
        dev = values.device
        n_slots = torch.zeros(cap, dtype=torch.long, device=dev)
        n_indices = torch.zeros(cap, dtype=torch.long, device=dev)
        n_values = torch.zeros(cap, dtype=values.dtype, device=dev)
        size = torch.empty_like(state.size)
        if scratch is not None:
            payload = scratch
        elif base is not None:
            payload = base.clone(memory_format=torch.contiguous_format)
        else:
            payload = torch.zeros(width, dtype=values.dtype, device=dev)
        keys = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        dest = torch.empty(max(cap, 1), dtype=torch.long, device=dev)
        work, nbytes = _work(0, cap, values.dtype, dev)
        p = _lib.ptr
        fn = getattr(_lib.load(), f"dn_calendar_pop_{_kind(values.dtype)[1]}")
        st = fn(p(state.slots), p(state.indices), p(values), cap, p(state.size), p(slot),
                width, p(n_slots), p(n_indices), p(n_values), p(size),
                p(payload) if width else None, p(keys), p(dest), p(work), nbytes,
                _lib.stream_handle(dev))
        _lib.check(st, "pending calendar pop")
        ctx.save_for_backward(keys[:cap], dest[:cap], state.indices)
        ctx.width, ctx.has_base = width, base is not None
        ctx.mark_non_differentiable(n_slots, n_indices, size)
        return n_slots, n_indices, n_values, size, payload

    @staticmethod
    def backward(ctx, _gs, _gi, grad_values, _gz, grad_payload):
        keys, dest, indices = ctx.saved_tensors
        g = None
        if grad_values is not None or grad_payload is not None:
            zero = torch.zeros((), dtype=keys.dtype if grad_values is None else
                               grad_values.dtype, device=keys.device)
            due = keys < ctx.width
            kept = grad_values.index_select(0, dest.clamp_min(0)) \
                if grad_values is not None and dest.numel() else None
            g = torch.where(dest >= 0, kept, zero) if kept is not None else None
            if grad_payload is not None and ctx.width:
                from_payload = grad_payload.index_select(0, indices.clamp(0, ctx.width - 1))
                g = torch.where(due, from_payload, g if g is not None else
                                torch.zeros_like(from_payload))
        g_base = grad_payload if ctx.has_base else None
        return g, g_base, None, None, None, None


def pending_calendar_pop(state, slot, *, width, base=None, _scratch=None):
    """Reduce the due entries in enqueue order and compact the rest."""
    cap = _check_state(state)
    dev = state.values.device
    width = int(width)
    slot_t = slot.reshape(-1)[:1] if torch.is_tensor(slot) else torch.tensor([int(slot)])
    slot_t = slot_t.to(device=dev, dtype=torch.long).contiguous()
    if base is not None and (base.shape != (width,) or base.dtype != state.values.dtype):
        raise ValueError(f"base must be a ({width},) tensor of the calendar's dtype")
    if _scratch is not None:
        if _scratch.shape != (width,) or _scratch.dtype != state.values.dtype or \
                not _scratch.is_contiguous():
            raise ValueError(f"_scratch must be a contiguous ({width},) tensor of the "
                             "calendar's dtype")
        with torch.no_grad():
            out = _Pop.apply(state.values, None, state, slot_t, width, _scratch)
    else:
        out = _Pop.apply(state.values, base, state, slot_t, width, None)
    n_slots, n_indices, n_values, size, payload = out
    return CalendarState(n_slots, n_indices, n_values, size, state.overflow), payload


def pending_calendar_from_legacy(chunks, *, capacity=None, device, dtype):
    """Import slot -> [(indices, values)] chunks (values keep their graph)."""
    slots, indices, values = [], [], []
    for slot, bucket in chunks.items():
        for idx, value in bucket:
            idx = idx.to(device=device, dtype=torch.long).reshape(idx.numel())
            value = value.to(device=device, dtype=dtype).reshape(value.numel())
            if idx.numel() != value.numel():
                raise ValueError("Pending calendar indices and values must align.")
            slots.append(torch.full_like(idx, int(slot)))
            indices.append(idx)
            values.append(value)
    count = sum(v.numel() for v in values)
    capacity = count if capacity is None else capacity
    state = pending_calendar_empty(capacity, device=device, dtype=dtype)
    if count > capacity:
        raise ValueError(f"Pending calendar capacity {capacity} cannot hold {count} entries.")
    if count == 0:
        return state
    return pending_calendar_append(state, torch.cat(slots), torch.cat(indices),
                                   torch.cat(values))


def pending_calendar_to_legacy(state):
    """Export buckets in first-occurrence order, keeping value gradients."""
    if bool(state.overflow.any()):
        raise RuntimeError("Pending calendar overflow; increase pending_capacity.")
    size = int(state.size.item())
    buckets = {}
    for position, slot in enumerate(state.slots[:size].detach().cpu().tolist()):
        buckets.setdefault(int(slot), []).append(position)
    return {
        slot: [(state.indices.index_select(0, torch.tensor(pos, device=state.indices.device)),
                state.values.index_select(0, torch.tensor(pos, device=state.values.device)))]
        for slot, pos in buckets.items()
    }
