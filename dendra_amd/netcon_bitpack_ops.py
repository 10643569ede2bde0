"""Drop-in for the reference's ``netcon_bitpack_ops`` extension module.

``NetCon`` (dendra/models/networks/netcon.py:3758-3827) imports
``netcon_bitpack_ops`` and calls ``is_available()``, then
``pack_source_spikes`` / ``build_delivery`` / ``build_delivery_uniform``
(netcon_bitpack_ops.py:559-700), and ``last_error()`` when the module is not
available.  This module keeps those names, argument lists, dtypes and
exceptions, and runs the gfx950 kernels of ``libdnhip.so``
(``dn_bitpack_pack``, ``dn_bitpack_deliver_{f32,f64}``) on the current
stream without synchronising:

* ``pack_source_spikes(source_spikes bool[n_src], packed_history
  int64[D, ceil(n_src/63)], current_time_step int64[1])``: writes row
  ``current_time_step`` (bits 0..62 of word w = sources 63w..63w+62);
* ``build_delivery(packed_history, current_time_step, delay_steps int64[n],
  conn_word_idx int64[n], conn_bit_mask int64[n], post_idx int64[n], weight
  float[n], delivery_out float[n_syn])``: ``delivery_out[post] += weight``
  for every connection whose source bit is set in row
  ``(current_time_step - delay) mod D``; the caller zeroes ``delivery_out``;
* ``build_delivery_uniform(..., delay_step int, ...)``: one delay for all
  connections (below 1 taken as 1, at or beyond D taken modulo D).

Structural checks read no tensor values (no device synchronisation), as
the reference's netcon_bitpack_contracts.py does; value errors (a delay,
word or target out of range) are skipped by the kernels, not reported.
Hits that share a target in one call are added in an unspecified order, as
with the reference's atomics.
"""

from __future__ import annotations

from typing import Optional

import torch

from . import _lib

BITS_PER_WORD = 63
_FLOATS = (torch.float32, torch.float64)

_LAST_ERROR: Optional[BaseException] = None


def is_available() -> bool:
    """True when a GPU is present and libdnhip.so loads (the first call
    loads it); after a failure ``last_error()`` holds the reason."""
    global _LAST_ERROR
    if not torch.cuda.is_available():
        _LAST_ERROR = RuntimeError("no ROCm GPU is available")
        return False
    try:
        _lib.load()
    except Exception as exc:  # the library is missing or does not load
        _LAST_ERROR = exc
        return False
    return True


def last_error() -> Optional[BaseException]:
    return _LAST_ERROR


def build_info() -> dict:
    """Where the kernels come from (the in-tree gfx950 library)."""
    info = {"library": str(_lib.LIB_PATH), "loaded": _lib.loaded(), "arch": "gfx950"}
    if _LAST_ERROR is not None:
        info["last_error"] = repr(_LAST_ERROR)
    return info


def _expect_tensors(**named):
    for name, t in named.items():
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")


def _same_device(**named):
    devs = {t.device for t in named.values()}
    if len(devs) > 1:
        raise ValueError("all bitpack tensors must share one device, got "
                         + ", ".join(f"{k}={t.device}" for k, t in named.items()))


def _gpu_contiguous(**named):
    for name, t in named.items():
        if t.device.type != "cuda":
            raise RuntimeError(f"{name} must be a GPU tensor")
        if not t.is_contiguous():
            raise RuntimeError(f"{name} must be contiguous")


def _check_step(packed_history, current_time_step):
    if packed_history.dtype != torch.int64:
        raise TypeError("packed_history must be torch.int64")
    if packed_history.ndim != 2:
        raise ValueError("packed_history must be 2-D [depth, words]")
    if current_time_step.dtype != torch.int64 or current_time_step.numel() != 1:
        raise TypeError("current_time_step must be a one-element torch.long tensor")


def pack_source_spikes(source_spikes: torch.Tensor, packed_history: torch.Tensor,
                       current_time_step: torch.Tensor) -> None:
    """Write this step's source spikes into history row ``current_time_step``."""
    _expect_tensors(source_spikes=source_spikes, packed_history=packed_history,
                    current_time_step=current_time_step)
    if source_spikes.dtype != torch.bool:
        raise TypeError("source_spikes must be torch.bool")
    if source_spikes.ndim != 1:
        raise ValueError("source_spikes must be 1-D")
    _check_step(packed_history, current_time_step)
    n_src = source_spikes.numel()
    depth, n_words = map(int, packed_history.shape)
    want = -(-n_src // BITS_PER_WORD)
    if n_words != want:
        raise ValueError(f"packed_history holds {n_words} words per row; {n_src} sources "
                         f"need {want}")
    if n_words and depth < 1:
        raise ValueError("packed_history needs at least one row")
    _same_device(source_spikes=source_spikes, packed_history=packed_history,
                 current_time_step=current_time_step)
    _gpu_contiguous(source_spikes=source_spikes, packed_history=packed_history,
                    current_time_step=current_time_step)
    if n_words == 0:
        return
    st = _lib.load().dn_bitpack_pack(
        _lib.ptr(source_spikes), n_src, _lib.ptr(packed_history), depth, n_words,
        _lib.ptr(current_time_step), _lib.stream_handle(source_spikes.device))
    _lib.check(st, "bitpack source-spike packing")


def _deliver(packed_history, current_time_step, delay_steps, uniform_delay, conn_word_idx,
             conn_bit_mask, post_idx, weight, delivery_out):
    named = dict(packed_history=packed_history, current_time_step=current_time_step,
                 conn_word_idx=conn_word_idx, conn_bit_mask=conn_bit_mask, post_idx=post_idx,
                 weight=weight, delivery_out=delivery_out)
    if delay_steps is not None:
        named["delay_steps"] = delay_steps
    _expect_tensors(**named)
    _check_step(packed_history, current_time_step)
    ints = {k: named[k] for k in ("delay_steps", "conn_word_idx", "conn_bit_mask",
                                  "post_idx") if k in named}
    for name, t in ints.items():
        if t.dtype != torch.int64:
            raise TypeError(f"{name} must be torch.long")
    if weight.dtype != delivery_out.dtype:
        raise TypeError("weight and delivery_out must share a dtype")
    if weight.dtype not in _FLOATS:
        raise TypeError("weight and delivery_out must be float32 or float64")
    n_conn = (delay_steps if delay_steps is not None else conn_word_idx).numel()
    for name, t in list(ints.items()) + [("weight", weight)]:
        if t.numel() != n_conn:
            raise ValueError(f"{name} holds {t.numel()} values for {n_conn} connections")
    depth, n_words = map(int, packed_history.shape)
    if n_conn and (depth < 1 or n_words < 1):
        raise ValueError("packed_history needs rows and words when connections exist")
    _same_device(**named)
    _gpu_contiguous(**named)
    if n_conn == 0:
        return
    fn = _lib.load().dn_bitpack_deliver_f64 if weight.dtype == torch.float64 else \
        _lib.load().dn_bitpack_deliver_f32
    st = fn(_lib.ptr(packed_history), depth, n_words, _lib.ptr(current_time_step),
            _lib.ptr(delay_steps), int(uniform_delay), _lib.ptr(conn_word_idx),
            _lib.ptr(conn_bit_mask), _lib.ptr(post_idx), _lib.ptr(weight),
            _lib.ptr(delivery_out), delivery_out.numel(), n_conn,
            _lib.stream_handle(weight.device))
    _lib.check(st, "bitpack delivery")


def build_delivery(packed_history: torch.Tensor, current_time_step: torch.Tensor,
                   delay_steps: torch.Tensor, conn_word_idx: torch.Tensor,
                   conn_bit_mask: torch.Tensor, post_idx: torch.Tensor, weight: torch.Tensor,
                   delivery_out: torch.Tensor) -> None:
    """Add today's deliveries (per-connection delays) into ``delivery_out``."""
    _deliver(packed_history, current_time_step, delay_steps, 0, conn_word_idx, conn_bit_mask,
             post_idx, weight, delivery_out)


def build_delivery_uniform(packed_history: torch.Tensor, current_time_step: torch.Tensor,
                           delay_step: int, conn_word_idx: torch.Tensor,
                           conn_bit_mask: torch.Tensor, post_idx: torch.Tensor,
                           weight: torch.Tensor, delivery_out: torch.Tensor) -> None:
    """Add today's deliveries (one delay for every connection)."""
    _deliver(packed_history, current_time_step, None, int(delay_step), conn_word_idx,
             conn_bit_mask, post_idx, weight, delivery_out)
