"""Source-history transport of NetCon on the GPU.

Drop-in for the reference's pure source-history kernels
(dendra/models/networks/_source_history.py:14-186), the transport of its
source-history training backend (networks/netcon.py:3660-3725): instead of a
delay ring per target, a ring ``history[depth, n_col]`` keeps one gate per
unique source and step, and today's payload is rebuilt from it with the
current weights and delays.

* ``source_history_project(history, slot, weights, *, post_idx, depth, width,
  source_pos | word_idx + bit_mask, delay_steps | delays + dt | delay_taps,
  taps, sigma, _owned_output, _scratch)``: connection ``e`` reads its source
  at rows ``(slot - offset) mod depth`` (fixed delays: one tap; fractional
  delays: two linear or three Gaussian taps) and adds ``w_e * gate`` to its
  target.  A HIP kernel walks each target's connections in ascending order
  (target-major CSR), rounding each product and sum as the reference's torch
  ops do, so the payload is bitwise the reference's sequential index_add.
  Gradients reach the weights, the tap weights (hence the delays) and a
  floating history.
* ``source_history_record(history, slot, gates, *, source_word_idx,
  source_bit_mask, _inplace, _clone_once)``: overwrite row ``slot`` with the
  gates (floating; gradients to the unwritten history and to the gates), or
  with packed words ``sum mask_s [gate_s != 0]`` (signed int64 layout).
* ``source_history_gates``, ``read_float_history``, ``read_packed_history``:
  the observation and read helpers, elementwise torch on the caller's device.

Inputs are never modified except where ``_inplace`` / ``_scratch`` donate
storage, as in the reference.
"""

from __future__ import annotations

import numpy as np
import torch
from torch.utils.weak import WeakIdKeyDictionary

from . import _lib
from .delays import csr_positions, delay_interpolation, delay_plan, derived_index

_FLOAT = (torch.float32, torch.float64)
_checked = WeakIdKeyDictionary()


def _int32_index(t, bound, name):
    """int32 copy of an index tensor, checked once per tensor object (and
    version) to lie in [0, bound)."""
    stamp = (t._version, t.numel(), int(bound), t.device)
    hit = _checked.get(t)
    if hit is not None and hit[0] == stamp:
        return hit[1]
    v = t.detach().reshape(-1)
    if v.numel():
        lo, hi = int(v.min()), int(v.max())
        if lo < 0 or hi >= bound:
            raise ValueError(f"{name} entries must lie in [0, {bound})")
    out = v.to(torch.int32).contiguous()
    _checked[t] = (stamp, out)
    return out


def _suffix(dtype):
    return "f64" if dtype == torch.float64 else "f32"


def _slot_tensor(slot, device):
    if torch.is_tensor(slot):
        return slot.reshape(-1)[:1].to(device=device, dtype=torch.int64).contiguous()
    return torch.tensor([int(slot)], dtype=torch.int64, device=device)


# -- observation and reads (elementwise; _source_history.py:14-80) ----------

def _update_active(has_spiked, values, threshold):
    ge = values >= threshold
    return ge, torch.logical_and(ge, ~has_spiked)


def _threshold_transition(values, previous, thresholds, *, diff_spiking, tau):
    """Rising-edge detector with a straight-through sigmoid surrogate
    (networks/spiking.py:7-73)."""
    if not diff_spiking:
        nxt, gates = _update_active(previous, values, thresholds)
        return nxt.detach(), gates
    s = torch.sigmoid((values - thresholds) / tau)
    ge = values >= thresholds
    spiked = ge & ~previous
    rise = s * (1.0 - previous.to(s.dtype))
    return ge.detach(), spiked.to(s.dtype) + (rise - rise.detach())


def _netstim_gates(sel, *, binary, dtype, **_):
    # generated events pass through; a hard gate is their truth value
    return sel.to(torch.bool) if binary else sel.to(dtype)


def _pre_var_gates(sel, *, binary, dtype, diff_spiking, **_):
    if binary and sel.dtype == torch.bool:
        return sel
    sel = sel.to(dtype)
    if diff_spiking:
        return sel
    positive = sel > 0
    return positive if binary else positive.to(dtype)


# mode -> (gate function, does it run the rising-edge detector)
_GATE_MODES = {"netstim": _netstim_gates, "pre_var": _pre_var_gates, "threshold": None}


def source_history_gates(source_values, has_spiked, *, source_idx, mode, threshold=None,
                         diff_spiking=False, tau=0.1, dtype=None, binary=False):
    """Observe one gate per unique source (semantics of
    _source_history.py:14-57): returns (detector latch, gates).  Modes:
    "netstim" (the source is an event stream), "pre_var" (gate = value > 0,
    or the raw value under surrogate spiking), "threshold" (rising edge of
    value >= threshold with the latch ``has_spiked``).  ``binary`` asks for
    hard boolean gates and excludes surrogate spiking."""
    if mode not in _GATE_MODES:
        raise ValueError(f"unknown source-history mode {mode!r} "
                         "(expected netstim, pre_var or threshold)")
    if binary and diff_spiking:
        raise ValueError("binary source-history gates exclude surrogate spiking")
    sel = torch.index_select(source_values.flatten(), 0, source_idx)
    work = source_values.dtype if dtype is None else dtype
    fn = _GATE_MODES[mode]
    if fn is not None:
        return has_spiked, fn(sel, binary=binary, dtype=work, diff_spiking=diff_spiking)
    return _threshold_transition(sel.to(work), has_spiked, threshold,
                                 diff_spiking=diff_spiking, tau=tau)


def _tap_columns(rows, cols):
    """Broadcast a per-connection column index over the tap rows."""
    return cols if rows.dim() == 1 else cols[:, None].expand(rows.shape)


def read_float_history(history, rows, source_pos):
    """Floating gates history[row, source] in connection order, for one tap
    (rows of shape (n,)) or several (rows of shape (n, taps))."""
    return history[rows, _tap_columns(rows, source_pos)]


def read_packed_history(history, rows, word_idx, bit_mask, *, dtype):
    """Hard gates from packed signed-int64 history words: bit ``bit_mask``
    of word ``word_idx`` at each tap row, as 0/1 of ``dtype``."""
    words = history[rows, _tap_columns(rows, word_idx)]
    return (words & _tap_columns(rows, bit_mask)).ne(0).to(dtype)


# -- projection --------------------------------------------------------------

# Per weights tensor: the projection's workspace and the tensors its packed
# records were built from.  The records hold tap offsets, not rows, so they
# stay valid while the slot advances and the history changes; the steps of a
# simulation with fixed weights and delays then only read them.  A change of
# any source (version, identity), of the sizes or of the stream rebuilds
# them.  The entry keeps its sources alive, so identity checks are sound.
_records = WeakIdKeyDictionary()


def _record_workspace(owner, sources, key, nbytes, device):
    stream = torch.cuda.current_stream(device).cuda_stream
    stamp = (owner._version, key, int(nbytes), stream, tuple(t._version for t in sources))
    hit = _records.get(owner)
    if hit is not None and hit[0] == stamp and len(hit[1]) == len(sources) and \
            all(a is b for a, b in zip(hit[1], sources)):
        return hit[2], True
    work = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
    _records[owner] = (stamp, tuple(sources), work)
    return work, False


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, history, weights, frac, slot, plan, col, mask, off, taps, width, cache):
        in_ptr, in_edge, post = plan
        packed = history.dtype == torch.int64
        depth, n_col = map(int, history.shape)
        out = torch.empty(width, dtype=weights.dtype, device=weights.device)
        p, lib, n = _lib.ptr, _lib.load(), weights.numel()
        f64 = int(weights.dtype == torch.float64)
        nbytes = lib.dn_history_work_bytes(n, taps, int(packed), f64, depth, n_col)
        pos = csr_positions(in_edge) if n else in_edge
        owner, sources, key = cache
        work, reuse = _record_workspace(owner, sources + (in_edge, pos), key, nbytes,
                                        weights.device)
        fn = getattr(lib, f"dn_history_project_ws_{_suffix(weights.dtype)}")
        st = fn(p(history), int(packed), depth, n_col, p(slot), taps, p(in_ptr), p(in_edge),
                p(pos), p(col), p(mask) if packed else None, p(off), p(weights),
                p(frac) if taps > 1 else None, p(out), width, n, p(work), nbytes, int(reuse),
                _lib.stream_handle(weights.device))
        _lib.check(st, "source-history projection")
        ctx.save_for_backward(history, weights, frac, slot, post, col, mask, off)
        ctx.taps, ctx.packed = taps, packed
        return out

    @staticmethod
    def backward(ctx, grad):
        history, weights, frac, slot, post, col, mask, off = ctx.saved_tensors
        taps, packed = ctx.taps, ctx.packed
        grad = grad.contiguous()
        depth, n_col = map(int, history.shape)
        n = weights.numel()
        dw = torch.empty_like(weights)
        dfrac = torch.empty_like(frac) if taps > 1 else None
        need_h = not packed and ctx.needs_input_grad[0]
        dhist = torch.zeros_like(history) if need_h else None
        p = _lib.ptr
        fn = getattr(_lib.load(), f"dn_history_project_vjp_{_suffix(weights.dtype)}")
        st = fn(p(grad), p(history), int(packed), depth, n_col, p(slot), taps, n, p(post),
                p(col), p(mask) if packed else None, p(off), p(weights),
                p(frac) if taps > 1 else None, p(dw), p(dfrac) if taps > 1 else None,
                p(dhist) if need_h else None, _lib.stream_handle(weights.device))
        _lib.check(st, "source-history projection cotangent")
        return dhist, dw, dfrac, None, None, None, None, None, None, None, None


def source_history_project(history, slot, weights, *, post_idx, depth, width, source_pos=None,
                           word_idx=None, bit_mask=None, delay_steps=None, delays=None,
                           dt=None, taps=2, sigma=0.35, delay_taps=None, _owned_output=False,
                           _scratch=None):
    """Today's payload from the source history (see the module doc)."""
    del _owned_output  # every result is a fresh tensor or the donated scratch
    if weights.dtype not in _FLOAT:
        raise TypeError("the GPU source-history projection supports float32 and float64")
    if weights.device.type != "cuda":
        raise ValueError("the GPU source-history projection requires GPU tensors")
    width, depth = int(width), int(depth)
    if weights.numel() == 0:
        return weights.new_zeros((width,)) if _scratch is None else _scratch.zero_()
    packed = source_pos is None
    if packed and history.dtype != torch.int64:
        raise TypeError("a packed source history must hold int64 words")
    if not packed and history.dtype != weights.dtype:
        raise TypeError("a floating source history must have the weights' dtype")
    if history.ndim != 2 or int(history.shape[0]) != depth:
        raise ValueError(f"history must have shape ({depth}, n); got {tuple(history.shape)}")
    if delay_taps is None and delays is not None:
        delay_taps = delay_interpolation(delays, dt, dtype=weights.dtype, taps=taps,
                                         sigma=sigma)
    n = weights.numel()
    index = word_idx if packed else source_pos
    if post_idx.numel() != n or index is None or index.numel() != n or \
            (packed and (bit_mask is None or bit_mask.numel() != n)):
        raise ValueError("post_idx and the source metadata must hold one entry per weight")
    if delay_taps is not None:
        if taps not in (2, 3):
            raise ValueError("taps must be 2 or 3")
        offsets, interp = delay_taps
        if taps == 2:
            tap_offsets = lambda o: torch.stack((o.reshape(n), o.reshape(n) + 1), dim=-1)
            frac = interp.reshape(n).to(weights.dtype)
        else:
            tap_offsets = lambda o: o.reshape(n, 3)
            frac = interp.reshape(n, 3).to(weights.dtype)
        n_taps = taps
    else:
        if delay_steps is None:
            raise ValueError("source_history_project needs delay_steps, delays or delay_taps")
        offsets = delay_steps
        tap_offsets = lambda o: o.reshape(n, 1)
        frac = weights.new_empty(0)
        n_taps = 1
    off = derived_index(offsets, ("history tap rows", depth, n_taps, n),
                        lambda o: tap_offsets(o).remainder(depth).to(torch.int32)
                        .contiguous())
    n_col = int(history.shape[1])
    if packed:
        col = _int32_index(word_idx, n_col, "word_idx")
        mask = bit_mask.reshape(n).to(torch.int64).contiguous()
    else:
        col = _int32_index(source_pos, n_col, "source_pos")
        mask = col  # unused placeholder tensor
    plan = delay_plan(post_idx, width)
    # the records' sources, by the caller's tensors where the operands are
    # fresh views or conversions of them (the weights are the cache's key)
    sources = tuple(t for t in (off, col, bit_mask if packed else None,
                                delay_taps[1] if delay_taps is not None else None)
                    if t is not None)
    out = _Project.apply(history.contiguous(), weights.reshape(n).contiguous(),
                         frac.contiguous(), _slot_tensor(slot, weights.device), plan, col, mask,
                         off, n_taps, width, (weights, sources, (n_taps, packed, n)))
    if _scratch is None:
        return out
    return _scratch.copy_(out)


# -- recording ---------------------------------------------------------------

class _RecordFloat(torch.autograd.Function):
    @staticmethod
    def forward(ctx, history, gates, slot, inplace):
        out = history if inplace else history.clone(memory_format=torch.contiguous_format)
        if inplace:
            ctx.mark_dirty(history)
        depth, n_col = map(int, out.shape)
        g = gates.reshape(n_col).to(out.dtype).contiguous()
        p = _lib.ptr
        fn = getattr(_lib.load(), f"dn_history_record_{_suffix(out.dtype)}")
        _lib.check(fn(p(out), depth, n_col, p(slot), p(g), _lib.stream_handle(out.device)),
                   "source-history record")
        ctx.save_for_backward(slot)
        ctx.depth, ctx.gshape = depth, gates.shape
        return out

    @staticmethod
    def backward(ctx, grad):
        (slot,) = ctx.saved_tensors
        row = slot.remainder(ctx.depth)
        gh = grad.clone()
        gh.index_fill_(0, row, 0)
        gg = grad.index_select(0, row).reshape(ctx.gshape)
        return gh, gg, None, None


def source_history_record(history, slot, gates, *, source_word_idx=None, source_bit_mask=None,
                          _inplace=False, _clone_once=False):
    """Overwrite row ``slot`` of the history (see the module doc)."""
    if history.device.type != "cuda":
        raise ValueError("the GPU source-history record requires GPU tensors")
    slot_t = _slot_tensor(slot, history.device)
    depth = int(history.shape[0])
    if source_word_idx is None:
        if history.dtype not in _FLOAT:
            raise TypeError("a floating source history must be float32 or float64")
        if gates.numel() != int(history.shape[1]):
            raise ValueError("gates must hold one value per history column")
        if _clone_once:
            history = history.clone(memory_format=torch.contiguous_format)
        inplace = (_inplace or _clone_once) and history.is_contiguous()
        return _RecordFloat.apply(history, gates, slot_t, inplace)
    if history.dtype != torch.int64:
        raise TypeError("a packed source history must hold int64 words")
    n_words = int(history.shape[1])
    n_src = gates.numel()
    if source_word_idx.numel() != n_src or source_bit_mask.numel() != n_src:
        raise ValueError("source_word_idx and source_bit_mask must hold one entry per gate")
    word = _int32_index(source_word_idx, max(n_words, 1) if n_src == 0 else n_words,
                        "source_word_idx")
    mask = source_bit_mask.reshape(n_src).to(torch.int64).contiguous()
    active = (gates.reshape(n_src) != 0).to(torch.uint8).contiguous()
    own = _inplace or _clone_once
    out = history if (own and history.is_contiguous() and not _clone_once) else \
        history.clone(memory_format=torch.contiguous_format)
    p = _lib.ptr
    st = _lib.load().dn_history_record_packed(p(out), depth, n_words, p(slot_t), n_src,
                                              p(active), p(word), p(mask),
                                              _lib.stream_handle(out.device))
    _lib.check(st, "packed source-history record")
    if _inplace and out is not history:
        history.copy_(out)
        return history
    return out
