"""Oracle (test infrastructure only): the reference's source-history
projection and recording.

Restates dendra/models/networks/_source_history.py:
  * source_history_project (:83-153): connection e reads its source gate at
    rows (slot - offset_j) mod depth -- floating history[row, source_pos_e]
    or packed (history[row, word_e] & mask_e) != 0 as 0/1 -- and contributes
        one tap   v = w g0
        two taps  v = w (g0 (1 - a) + g1 a)        (offsets k, k + 1)
        three     v = w ((g0 f0 + g1 f1) + g2 f2)  (torch's 3-element sum)
    then out = zeros(width); out[post_e] += v_e in connection order (CPU
    index_add).  Every product and sum is rounded in the weights' dtype
    (numpy scalars), so float32 results are the exact float32 sequence.
  * source_history_record (:156-186): row slot := gates (floating) or the
    packed words sum_s mask_s [gate_s != 0] (int64).
Pure-Python loops: small cases.
"""

import numpy as np


def _gate(history, row, col, mask, e, dt, packed):
    if packed:
        return dt(1) if (int(history[row, col[e]]) & int(mask[e])) != 0 else dt(0)
    return dt(history[row, col[e]])


def project(history, slot, weights, post, width, *, col, offsets, frac=None, mask=None,
            packed=False):
    """offsets: (n, taps) int tap offsets (taps 1, 2 or 3); frac: None (one
    tap), (n,) upper-tap weights (two taps) or (n, 3) weights (three)."""
    w = np.asarray(weights)
    dt = w.dtype.type
    depth = history.shape[0]
    offsets = np.asarray(offsets).reshape(len(w), -1)
    taps = offsets.shape[1]
    out = np.zeros(width, dtype=w.dtype)
    for e in range(len(w)):
        g = [_gate(history, (slot - int(offsets[e, j])) % depth, col, mask, e, dt, packed)
             for j in range(taps)]
        if taps == 1:
            s = g[0]
        elif taps == 2:
            a = dt(frac[e])
            s = dt(dt(g[0] * dt(dt(1) - a)) + dt(g[1] * a))
        else:
            p = [dt(g[j] * dt(frac[e, j])) for j in range(3)]
            s = dt(dt(p[0] + p[1]) + p[2])
        out[post[e]] = dt(out[post[e]] + dt(dt(w[e]) * s))
    return out


def record_float(history, slot, gates):
    h = np.array(history, copy=True)
    h[slot % h.shape[0], :] = np.asarray(gates, dtype=h.dtype)
    return h


def record_packed(history, slot, gates, word, mask):
    h = np.array(history, dtype=np.int64, copy=True)
    row = np.zeros(h.shape[1], dtype=np.int64)
    for s, g in enumerate(np.asarray(gates)):
        if g != 0:
            row[word[s]] = np.int64(int(row[word[s]]) + int(mask[s]))
    h[slot % h.shape[0], :] = row
    return h
