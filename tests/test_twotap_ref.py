"""Pins of the two-tap delay oracle (oracle/twotap_ref.py): a hand-computed
known answer, agreement with torch's sequential CPU index_add form of the
reference's transition, and the delay split's clamp and fraction."""

import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import twotap_ref  # noqa: E402


def torch_transition(queue, slot, weights, gates, fraction, lower, post):
    """The reference's dense two-tap transition written with index_add
    (_event_delivery.py:282-300)."""
    depth, width = queue.shape
    out = queue.clone()
    out[slot] = 0
    wg = weights * gates
    row0 = (slot + lower) % depth
    row1 = (row0 + 1) % depth
    flat = out.reshape(-1)
    flat = flat.index_add(0, row0 * width + post, wg * (1 - fraction))
    flat = flat.index_add(0, row1 * width + post, wg * fraction)
    return flat.reshape(depth, width)


def test_known_answer():
    q = np.zeros((4, 2), dtype=np.float64)
    q[1, 0] = 5.0  # slot row: cleared
    q[3, 1] = 1.0
    out = twotap_ref.enqueue(q, 1, [2.0, 1.0, 3.0], [0.5, 1.0, 1.0], [0.25, 0.0, 0.5],
                             [1, 2, 3], [0, 1, 0])
    want = np.zeros((4, 2))
    want[3, 1] = 1.0
    want[2, 0] += 1.0 * 0.75   # e0 lower: row 2
    want[3, 0] += 1.0 * 0.25   # e0 upper: row 3
    want[3, 1] += 1.0          # e1 lower: row 3 (fraction 0)
    want[0, 0] += 3.0 * 0.5    # e2 lower: row (1+3)%4 = 0
    want[1, 0] += 3.0 * 0.5    # e2 upper: the cleared slot row (wraps)
    np.testing.assert_array_equal(out, want)


def test_oracle_equals_sequential_index_add_bitwise():
    rng = np.random.default_rng(4)
    for dtype in (np.float32, np.float64):
        depth, width, n = 7, 5, 60
        q = rng.standard_normal((depth, width)).astype(dtype)
        w = rng.standard_normal(n).astype(dtype)
        g = rng.uniform(0, 1, n).astype(dtype)
        delays = rng.uniform(0.0, 9.0, n)
        lower, frac = twotap_ref.two_tap_split(delays, 1.0, dtype)
        post = rng.integers(0, width, n)
        slot = 3
        want = twotap_ref.enqueue(q, slot, w, g, frac, lower, post)
        got = torch_transition(*(torch.as_tensor(v) for v in (q,)), slot,
                               torch.as_tensor(w), torch.as_tensor(g), torch.as_tensor(frac),
                               torch.as_tensor(lower), torch.as_tensor(post)).numpy()
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_delay_split_clamps_to_one_step():
    lower, frac = twotap_ref.two_tap_split([0.1, 1.0, 2.5, 7.25], 0.5, np.float64)
    np.testing.assert_array_equal(lower, [1, 2, 5, 14])
    np.testing.assert_allclose(frac, [0.0, 0.0, 0.0, 0.5])


def test_three_tap_split_and_scatter_order():
    offsets, fr = twotap_ref.three_tap_split([0.1, 1.3], 1.0, 0.5, np.float64)
    np.testing.assert_array_equal(offsets, [[0, 1, 2], [0, 1, 2]])
    np.testing.assert_allclose(fr.sum(1), 1.0)
    assert fr[0, 1] > fr[0, 0] and fr[1, 1] > fr[1, 2] > fr[1, 0]
    rng = np.random.default_rng(3)
    q = rng.standard_normal((6, 4)).astype(np.float32)
    w, g = rng.standard_normal(30).astype(np.float32), rng.uniform(0, 1, 30).astype(np.float32)
    offs, fr = twotap_ref.three_tap_split(rng.uniform(0, 8, 30), 1.0, 0.6, np.float32)
    post = rng.integers(0, 4, 30)
    want = twotap_ref.enqueue3(q, 2, w, g, fr, offs, post)
    out = torch.as_tensor(q).clone()
    out[2] = 0
    dest = ((2 + torch.as_tensor(offs)) % 6) * 4 + torch.as_tensor(post)[:, None]
    vals = (torch.as_tensor(w) * torch.as_tensor(g))[:, None] * torch.as_tensor(fr)
    got = out.reshape(-1).scatter_add(0, dest.reshape(-1), vals.reshape(-1)).reshape(6, 4)
    assert np.array_equal(got.numpy().view(np.uint8), want.view(np.uint8))
