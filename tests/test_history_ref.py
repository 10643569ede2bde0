"""Pins of the source-history oracle (oracle/history_ref.py): known answers
from the reference's own fixture (tests/networks/test_source_history_kernels.py
:233-252, :175-200), bitwise agreement with the reference's torch formula
evaluated on the CPU (index_add in connection order), and the order of torch's
three-element sum that the three-tap projection follows."""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import history_ref  # noqa: E402

# the reference test module's _fixture: depth 5, three sources, five connections
SOURCE_POS = np.array([0, 1, 2, 0, 2])
POST = np.array([0, 0, 1, 1, 0])
VALUES = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 1], [1, 1, 0], [0, 1, 0]], dtype=np.float64)
WEIGHTS = np.array([0.7, -0.2, 1.3, 0.5, 2.1])
FIXED = np.array([1, 2, 3, 4, 1])


def packing(size):
    pos = np.arange(size)
    return pos // 63, np.left_shift(np.int64(1), (pos % 63).astype(np.int64))


def torch_project(history, slot, w, post, width, col, offsets, frac=None, mask=None,
                  packed=False):
    """The reference's projection formula (_source_history.py:118-153) on
    CPU tensors."""
    h, w = torch.as_tensor(history), torch.as_tensor(w)
    offsets = torch.as_tensor(offsets).reshape(len(w), -1)
    taps = offsets.shape[1]
    col = torch.as_tensor(col)

    def read(rows):
        if packed:
            words = h[rows, col]
            return (torch.bitwise_and(words, torch.as_tensor(mask)) != 0).to(w.dtype)
        return h[rows, col]

    depth = h.shape[0]
    if taps == 1:
        values = w * read((slot - offsets[:, 0]).remainder(depth))
    elif taps == 2:
        alpha = torch.as_tensor(frac)
        row0 = (slot - offsets[:, 0]).remainder(depth)
        row1 = (row0 - 1).remainder(depth)
        values = w * (read(row0) * (1.0 - alpha) + read(row1) * alpha)
    else:
        rows = (slot - offsets).remainder(depth)
        g = torch.stack([read(rows[:, j]) for j in range(3)], -1)
        values = w * (g * torch.as_tensor(frac)).sum(dim=-1)
    return torch.zeros(width, dtype=w.dtype).index_add(0, torch.as_tensor(post), values)


def test_fixed_delay_known_answer_of_the_reference_fixture():
    # rows (4 - fixed) % 5 = 3, 2, 1, 0, 3 -> gates 1, 1, 0, 0, 0
    out = history_ref.project(VALUES, 4, WEIGHTS, POST, 2, col=SOURCE_POS,
                              offsets=FIXED[:, None])
    assert out[0] == (0.0 + 0.7) + -0.2 and out[1] == 0.0
    word, mask = packing(3)
    packed = np.zeros((5, 1), dtype=np.int64)
    for s in range(5):
        packed = history_ref.record_packed(packed, s, VALUES[s], word, mask)
    assert packed[:, 0].tolist() == [4, 1, 6, 3, 2]
    out_p = history_ref.project(packed, 4, WEIGHTS, POST, 2, col=word[SOURCE_POS],
                                offsets=FIXED[:, None], mask=mask[SOURCE_POS], packed=True)
    assert np.array_equal(out_p, out)


def test_two_tap_known_answer():
    # connection 0 at offsets 1, 2 from slot 4: rows 3 and 2 of source 0 (1, 0)
    out = history_ref.project(VALUES, 4, WEIGHTS[:1], POST[:1], 2, col=SOURCE_POS[:1],
                              offsets=[[1, 2]], frac=np.array([0.25]))
    assert out[0] == 0.7 * (1 * 0.75 + 0 * 0.25)


def test_packed_record_crosses_word_boundaries():
    """tests/networks/test_source_history_kernels.py:185-209."""
    size, depth = 130, 4
    word, mask = packing(size)
    history = np.full((depth, 3), 17, dtype=np.int64)
    gates = np.zeros(size)
    active = [0, 61, 62, 63, 64, 125, 126, 129]
    gates[active] = 1
    out = history_ref.record_packed(history, 2, gates, word, mask)
    want = history.copy()
    want[2] = [sum(1 << (v % 63) for v in active if v // 63 == w) for w in range(3)]
    assert np.array_equal(out, want)
    assert (history == 17).all()


def test_float_record_known_answer():
    h = np.arange(15, dtype=np.float64).reshape(5, 3)
    out = history_ref.record_float(h, 7, [-0.2, 0.0, 0.8])
    want = h.copy()
    want[2] = [-0.2, 0.0, 0.8]
    assert np.array_equal(out, want)


def test_three_element_sum_order_is_torchs():
    rng = np.random.default_rng(0)
    for dtype in (torch.float32, torch.float64):
        x = torch.as_tensor(rng.standard_normal((4000, 3)) * 10.0 ** rng.integers(-3, 5, (4000, 3)),
                            dtype=dtype)
        left = (x[:, 0] + x[:, 1]) + x[:, 2]
        right = x[:, 0] + (x[:, 1] + x[:, 2])
        assert bool((left != right).any())  # the orders do differ on these inputs
        assert torch.equal(x.sum(dim=-1), left)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("taps", [1, 2, 3])
@pytest.mark.parametrize("packed", [False, True])
def test_oracle_equals_reference_formula_bitwise(dtype, taps, packed):
    rng = np.random.default_rng(10 * taps + packed)
    depth, n_src, n, width = 9, 70, 400, 13
    w = rng.standard_normal(n).astype(dtype)
    src = rng.integers(0, n_src, n)
    post = rng.integers(0, width, n)
    slot = 6
    lower = rng.integers(1, depth - 1, n)
    if taps == 1:
        offsets, frac = lower[:, None], None
    elif taps == 2:
        offsets, frac = np.stack([lower, lower + 1], -1), rng.uniform(0, 1, n).astype(dtype)
    else:
        offsets = np.stack([lower - 1, lower, lower + 1], -1)
        f = rng.uniform(0.1, 1, (n, 3))
        frac = (f / f.sum(1, keepdims=True)).astype(dtype)
    if packed:
        word, mask = packing(n_src)
        history = np.zeros((depth, int(word.max()) + 1), dtype=np.int64)
        for s in range(depth):
            history = history_ref.record_packed(history, s, rng.uniform(0, 1, n_src) < 0.4,
                                                word, mask)
        col, cmask = word[src], mask[src]
    else:
        history = rng.uniform(-1, 1, (depth, n_src)).astype(dtype)
        col, cmask = src, None
    got = history_ref.project(history, slot, w, post, width, col=col, offsets=offsets,
                              frac=frac, mask=cmask, packed=packed)
    want = torch_project(history, slot, w, post, width, col, offsets, frac, cmask,
                         packed).numpy()
    assert got.dtype == want.dtype
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
