"""Pins of the pending-calendar oracle (oracle/calendar_ref.py) against the
expectations of the reference's own tests
(tests/networks/test_pending_calendar.py:29-129) and against the legacy
chunk reduction the reference checks it with (index_add per chunk, in order)."""

import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import calendar_ref  # noqa: E402


def test_append_overflow_is_atomic_and_sticky():
    """test_append_overflow_is_atomic_sticky_and_does_not_mutate_inputs."""
    initial = calendar_ref.empty(3, np.float64)
    first = calendar_ref.append(initial, [2, 1], [0, 1], [4.0, 7.0])
    assert initial["size"] == 0 and not initial["values"].any()
    assert first["size"] == 2 and first["values"].tolist() == [4.0, 7.0, 0.0]
    failed = calendar_ref.append(first, [1, 2], [0, 0], [8.0, 9.0])
    assert failed["overflow"] and failed["size"] == 2
    assert failed["values"].tolist() == first["values"].tolist()
    later = calendar_ref.append(failed, [1], [0], [1.0])
    assert later["overflow"] and later["size"] == 2
    assert later["values"].tolist() == first["values"].tolist()


def test_masked_candidates_never_enter():
    """test_masked_nan_and_invalid_indices_never_enter_storage_or_payload."""
    state = calendar_ref.append(calendar_ref.empty(4, np.float64), [1, -999, 2, 999],
                                [0, -999, 1, 999], [3.0, np.nan, 5.0, np.inf],
                                [True, False, True, False])
    assert state["size"] == 2
    nxt, payload = calendar_ref.pop(state, 1, 2)
    assert payload.tolist() == [3.0, 0.0]
    assert nxt["size"] == 1
    assert nxt["slots"].tolist() == [2, 0, 0, 0]
    assert nxt["indices"].tolist() == [1, 0, 0, 0]
    assert nxt["values"].tolist() == [5.0, 0.0, 0.0, 0.0]


def test_ordered_pop_matches_chunk_reduction_with_cancellation():
    """test_ordered_pop_matches_legacy_chunk_reduction_with_cancellation
    (float32 and float64)."""
    for dtype, tdtype in ((np.float32, torch.float32), (np.float64, torch.float64)):
        gen = torch.Generator().manual_seed(31)
        chunks = []
        for _ in range(42):
            idx = torch.randint(0, 4, (7,), generator=gen)
            chunks.append((idx, (torch.randn(7, generator=gen) * 7).to(tdtype)))
        chunks.append((torch.tensor([0, 0, 0]),
                       torch.tensor([1000.0, 0.25, -1000.0], dtype=tdtype)))
        idx = torch.cat([c[0] for c in chunks]).numpy()
        val = torch.cat([c[1] for c in chunks]).numpy()
        state = calendar_ref.append(calendar_ref.empty(len(idx), dtype), np.full(len(idx), 3),
                                    idx, val)
        base = np.array([2.0, -3.0, 0.25, 8.0], dtype=dtype)
        expected = torch.as_tensor(base.copy())
        for i, v in chunks:
            expected.index_add_(0, i, v)
        nxt, payload = calendar_ref.pop(state, 3, 4, base=base)
        assert np.array_equal(payload.view(np.uint8), expected.numpy().view(np.uint8))
        assert nxt["size"] == 0 and not nxt["values"].any()
        assert base.tolist() == [2.0, -3.0, 0.25, 8.0]


def test_transition_known_answers():
    """test_calendar_transitions_compile_vmap_and_preserve_value_gradients:
    slots (2, 1, 2), indices (0, 1, 0); popping slot 2 of values (3, 4, 5)
    gives payload (8, 0), slot 1 of (6, 7, 8) gives (0, 7)."""
    for vals, slot, want in (([3.0, 4.0, 5.0], 2, [8.0, 0.0]), ([6.0, 7.0, 8.0], 1, [0.0, 7.0])):
        st = calendar_ref.append(calendar_ref.empty(5, np.float64), [2, 1, 2], [0, 1, 0], vals)
        _, payload = calendar_ref.pop(st, slot, 2)
        assert payload.tolist() == want


def test_zero_capacity():
    st = calendar_ref.append(calendar_ref.empty(0, np.float64), [], [], [])
    nxt, payload = calendar_ref.pop(st, 0, 0)
    assert nxt["size"] == 0 and payload.shape == (0,)
