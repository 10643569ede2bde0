"""Oracle (test infrastructure only): the reference's ordered two-tap enqueue
and its three-tap scatter.

Restates ``enqueue_two_tap`` of dendra/models/networks/_two_tap_enqueue_cpu.cpp
:61-125 (with add_tap, :26-58) and the two-tap split of
networks/_event_delivery.py:200-216:
  lam = max(delay / dt, 1); lower = floor(lam); f = lam - lower
  out = copy of queue; out[slot] = 0
  for e in order:  out[(slot + lower_e) % depth, post_e] += (w_e g_e)(1 - f_e)
  for e in order:  out[(slot + lower_e + 1) % depth, post_e] += (w_e g_e) f_e
with every product and sum rounded in the queue's dtype (numpy scalars), so a
float32 result is the exact float32 sequence.  Pure-Python loops: small cases.
"""

import numpy as np


def two_tap_split(delays, dt, dtype):
    lam = np.maximum(np.asarray(delays, dtype=dtype) / dtype(dt), dtype(1))
    lower = np.floor(lam)
    return lower.astype(np.int64), (lam - lower).astype(dtype)


def three_tap_split(delays, dt, sigma, dtype):
    """Offsets floor(lam) + (-1, 0, 1) and softmax weights of
    -((lam - offset) / (sigma + 1e-6))^2 / 2 (_event_delivery.py:217-223)."""
    lam = np.maximum(np.asarray(delays, dtype=np.float64) / dt, 1.0)
    lower = np.floor(lam).astype(np.int64)
    offsets = lower[:, None] + np.array([-1, 0, 1])
    z = -0.5 * ((lam[:, None] - offsets) / (sigma + 1e-6)) ** 2
    z = np.exp(z - z.max(axis=1, keepdims=True))
    return offsets, (z / z.sum(axis=1, keepdims=True)).astype(dtype)


def enqueue3(queue, slot, weights, gates, fractions, offsets, post):
    """Three-tap scatter in the flattened (connection, tap) order of
    _event_delivery.py:296-308."""
    q = np.array(queue, copy=True)
    dt = q.dtype.type
    depth, _ = q.shape
    q[slot, :] = dt(0)
    for e in range(len(weights)):
        wg = dt(weights[e]) * dt(gates[e])
        for j in range(3):
            row = (slot + int(offsets[e, j])) % depth
            q[row, post[e]] = dt(q[row, post[e]] + wg * dt(fractions[e, j]))
    return q


def enqueue(queue, slot, weights, gates, fraction, lower, post):
    q = np.array(queue, copy=True)
    dt = q.dtype.type
    depth, _ = q.shape
    q[slot, :] = dt(0)
    for tap in (0, 1):
        for e in range(len(weights)):
            row = (slot + int(lower[e]) + tap) % depth
            wg = dt(weights[e]) * dt(gates[e])
            frac = dt(1) - dt(fraction[e]) if tap == 0 else dt(fraction[e])
            q[row, post[e]] = dt(q[row, post[e]] + wg * frac)
    return q
