"""Oracle (test infrastructure only): the reference's ordered two-tap enqueue.

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
