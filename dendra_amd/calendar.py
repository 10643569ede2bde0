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


def _work(n, cap, dtype, device, width=None):
    """Workspace of an append (n candidates) or, with `width`, of a pop that
    adds the due values through per-index hit lists."""
    kind, _ = _kind(dtype)
    lib = _lib.load()
    nbytes = int(lib.dn_calendar_work_bytes(int(n), int(cap), kind) if width is None else
                 lib.dn_calendar_pop_work_bytes(int(cap), int(width), kind))
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
        # the kernel zeroes the entries past the kept ones
        n_slots = torch.empty(cap, dtype=torch.long, device=dev)
        n_indices = torch.empty(cap, dtype=torch.long, device=dev)
        n_values = torch.empty(cap, dtype=values.dtype, device=dev)
        size = torch.empty_like(state.size)
        if scratch is not None:
            payload = scratch
        elif base is not None:
            payload = base.clone(memory_format=torch.contiguous_format)
        else:
            payload = torch.zeros(width, dtype=values.dtype, device=dev)
        keys = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        dest = torch.empty(max(cap, 1), dtype=torch.long, device=dev)
        work, nbytes = _work(0, cap, values.dtype, dev, width=width)
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
    """Build a calendar from the legacy ``{slot: [(indices, values), ...]}``
    buckets (networks/_pending_calendar.py:120-140 semantics): entries keep
    the bucket order and, within a bucket, the chunk order; floating values
    keep their autograd graph.  ``capacity`` defaults to the entry count."""
    pieces = [(int(slot), idx, val) for slot, bucket in chunks.items() for idx, val in bucket]
    idx_cols = [p[1].reshape(-1).to(device=device, dtype=torch.long) for p in pieces]
    val_cols = [p[2].reshape(-1).to(device=device, dtype=dtype) for p in pieces]
    if any(i.numel() != v.numel() for i, v in zip(idx_cols, val_cols)):
        raise ValueError("legacy calendar chunk: index and value counts differ")
    lengths = [i.numel() for i in idx_cols]
    total = sum(lengths)
    cap = total if capacity is None else capacity
    empty = pending_calendar_empty(cap, device=device, dtype=dtype)
    if total > cap:
        raise ValueError(f"capacity {cap} cannot hold the {total} legacy calendar entries")
    if total == 0:
        return empty
    slot_col = torch.repeat_interleave(
        torch.tensor([p[0] for p in pieces], dtype=torch.long, device=device),
        torch.tensor(lengths, dtype=torch.long, device=device))
    return pending_calendar_append(empty, slot_col, torch.cat(idx_cols), torch.cat(val_cols))


def pending_calendar_to_legacy(state):
    """Export the live entries as legacy buckets ``{slot: [(indices, values)]}``
    (networks/_pending_calendar.py:143-161 semantics): one chunk per slot,
    entries in position order, buckets ordered by their slot's first position;
    values keep their gradients.  An overflowed calendar cannot be exported."""
    if bool(state.overflow.any()):
        raise RuntimeError("pending calendar overflowed: raise its capacity before export")
    live = state.slots[: int(state.size.item())].detach()
    # A stable sort groups the positions by slot and keeps them in position
    # order inside each group; the groups are then ranked by first position.
    by_slot = torch.argsort(live, stable=True)
    keys, counts = torch.unique_consecutive(live[by_slot], return_counts=True)
    groups = torch.split(by_slot, counts.tolist())
    heads = [int(g[0]) for g in groups]
    out = {}
    for k in sorted(range(len(groups)), key=heads.__getitem__):
        pos = groups[k]
        out[int(keys[k])] = [(state.indices.index_select(0, pos),
                              state.values.index_select(0, pos))]
    return out
