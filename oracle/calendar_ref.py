"""Oracle (test infrastructure only): the reference's pending calendar.

Restates dendra/models/networks/_pending_calendar.py:37-113 with plain
Python loops over numpy arrays:
  append: if overflow or size + count(active) > capacity -> overflow = True
          and nothing is written; else the active candidates, in order, are
          written at size, size + 1, ... and size += count;
  pop:    payload = copy of base (or zeros); for p < size in order, if
          slots[p] == slot: payload[indices[p]] += values[p] (rounded in the
          values' dtype, the sequential scatter_add); the other valid
          entries are compacted in order into zero-filled arrays.
A state is a dict {slots, indices, values, size, overflow}.
"""

import numpy as np


def empty(capacity, dtype):
    return {"slots": np.zeros(capacity, np.int64), "indices": np.zeros(capacity, np.int64),
            "values": np.zeros(capacity, dtype), "size": 0, "overflow": False}


def append(state, slots, indices, values, active=None):
    s = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in state.items()}
    n = len(slots)
    active = np.ones(n, bool) if active is None else np.asarray(active, bool)
    count = int(active.sum())
    cap = len(s["values"])
    if s["overflow"] or s["size"] + count > cap:
        s["overflow"] = True
        return s
    pos = s["size"]
    for i in range(n):
        if active[i]:
            s["slots"][pos] = slots[i]
            s["indices"][pos] = indices[i]
            s["values"][pos] = values[i]
            pos += 1
    s["size"] = pos
    return s


def pop(state, slot, width, base=None):
    vals = state["values"]
    dt = vals.dtype.type
    payload = np.zeros(width, vals.dtype) if base is None else np.array(base, vals.dtype)
    cap = len(vals)
    out = empty(cap, vals.dtype)
    k = 0
    for p in range(state["size"]):
        if state["slots"][p] == slot:
            i = int(state["indices"][p])
            payload[i] = dt(payload[i] + vals[p])
        else:
            out["slots"][k] = state["slots"][p]
            out["indices"][k] = state["indices"][p]
            out["values"][k] = vals[p]
            k += 1
    out["size"] = k
    out["overflow"] = state["overflow"]
    return out, payload
