"""GPU pending calendar (dendra_amd.calendar) against the oracle, bitwise,
and the reference's own cases (tests/networks/test_pending_calendar.py)."""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import calendar_ref  # noqa: E402

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    return torch.as_tensor(np.asarray(a), device="cuda", dtype=dtype)


def same_state(gpu, ref):
    size = int(gpu.size.item())
    assert size == ref["size"] and bool(gpu.overflow.item()) == ref["overflow"]
    assert np.array_equal(gpu.slots.cpu().numpy(), ref["slots"])
    assert np.array_equal(gpu.indices.cpu().numpy(), ref["indices"])
    v = gpu.values.cpu().numpy()
    assert np.array_equal(v.view(np.uint8), ref["values"].view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
def test_random_append_pop_sequence_is_bitwise_the_oracle(dtype):
    from dendra_amd.calendar import (pending_calendar_append, pending_calendar_empty,
                                     pending_calendar_pop)

    rng = np.random.default_rng({np.float32: 1, np.float64: 2, np.int32: 3}[dtype])
    cap, width, depth = 6000, 257, 40
    tdt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32}[dtype]
    g = pending_calendar_empty(cap, device="cuda", dtype=tdt)
    r = calendar_ref.empty(cap, dtype)
    for step in range(60):
        n = int(rng.integers(0, 400))
        slots = rng.integers(0, depth, n)
        idx = rng.integers(0, width, n)
        if dtype == np.int32:
            vals = rng.integers(-5, 6, n).astype(dtype)
        else:
            vals = (rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)).astype(dtype)
        active = rng.uniform(0, 1, n) < 0.8
        g = pending_calendar_append(g, dev(slots), dev(idx), dev(vals), dev(active))
        r = calendar_ref.append(r, slots, idx, vals, active)
        same_state(g, r)
        slot = step % depth
        base = rng.standard_normal(width).astype(dtype) if (step % 3 == 0 and dtype != np.int32) \
            else None
        g, payload = pending_calendar_pop(g, dev([slot]), width=width,
                                          base=None if base is None else dev(base))
        r, want = calendar_ref.pop(r, slot, width, base=base)
        same_state(g, r)
        got = payload.cpu().numpy()
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), step
    # overflow: a too-large append leaves the calendar intact and sticks
    size = int(g.size.item())
    n = cap - size + 1
    big = pending_calendar_append(g, dev(np.zeros(n, np.int64)), dev(np.zeros(n, np.int64)),
                                  dev(np.ones(n, dtype)))
    assert bool(big.overflow.item()) and int(big.size.item()) == size
    assert torch.equal(big.values, g.values)
    later = pending_calendar_append(big, dev([0]), dev([0]), dev(np.ones(1, dtype)))
    assert bool(later.overflow.item()) and int(later.size.item()) == size


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("width", [3, 1000])
def test_pop_hit_lists_and_sort_agree_with_the_oracle(dtype, width):
    """The pop's two ordered reductions through the C ABI: per-index hit
    lists (the workspace of dn_calendar_pop_work_bytes; indices with more
    than 16 due entries take the key scan) and the stable radix sort (the
    smaller workspace of dn_calendar_work_bytes) give the oracle's payload
    and calendar bit for bit; width 3 puts hundreds of due entries on each
    index, width 1000 about one."""
    from dendra_amd import _lib

    rng = np.random.default_rng(width + {np.float32: 0, np.float64: 1, np.int32: 2}[dtype])
    cap, size, slot = 5000, 4200, 7
    slots = np.zeros(cap, np.int64)
    slots[:size] = rng.integers(0, 12, size)
    idx = np.zeros(cap, np.int64)
    idx[:size] = rng.integers(0, width, size)
    if dtype == np.int32:
        vals = rng.integers(-5, 6, cap).astype(dtype)
    else:
        vals = (rng.standard_normal(cap) * 10.0 ** rng.integers(-3, 4, cap)).astype(dtype)
    base = (rng.integers(-3, 4, width) if dtype == np.int32 else
            rng.standard_normal(width)).astype(dtype)
    ref = {"slots": slots.copy(), "indices": idx.copy(), "values": vals.copy(), "size": size,
           "overflow": False}
    r, want = calendar_ref.pop(ref, slot, width, base=base)
    kind, sfx = {np.float32: (0, "f32"), np.float64: (1, "f64"), np.int32: (2, "i32")}[dtype]
    lib, p = _lib.load(), _lib.ptr
    fn = getattr(lib, f"dn_calendar_pop_{sfx}")
    for nbytes in (lib.dn_calendar_pop_work_bytes(cap, width, kind),
                   lib.dn_calendar_work_bytes(0, cap, kind)):
        ts, ti, tv = dev(slots), dev(idx), dev(vals)
        tsize, tslot = dev([size], torch.int64), dev([slot], torch.int64)
        # outputs start dirty: the kernel zeroes the entries past the kept ones
        ns, ni = torch.full((cap,), 7, dtype=torch.long, device="cuda"), torch.full(
            (cap,), 7, dtype=torch.long, device="cuda")
        nv = torch.full((cap,), 7, dtype=tv.dtype, device="cuda")
        nsize = torch.empty(1, dtype=torch.long, device="cuda")
        payload = dev(base)
        work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        st = fn(p(ts), p(ti), p(tv), cap, p(tsize), p(tslot), width, p(ns), p(ni), p(nv),
                p(nsize), p(payload), None, None, p(work), nbytes, _lib.stream_handle(tv.device))
        assert st == 0
        torch.cuda.synchronize()
        assert np.array_equal(payload.cpu().numpy().view(np.uint8), want.view(np.uint8))
        assert int(nsize.item()) == r["size"]
        assert np.array_equal(nv.cpu().numpy().view(np.uint8), r["values"].view(np.uint8))
        assert np.array_equal(ni.cpu().numpy(), r["indices"])
        assert np.array_equal(ns.cpu().numpy(), r["slots"])


def test_reference_overflow_mask_and_ordering_cases():
    from dendra_amd.calendar import (pending_calendar_append, pending_calendar_empty,
                                     pending_calendar_pop, pending_calendar_from_legacy)

    initial = pending_calendar_empty(3, device="cuda", dtype=torch.float64)
    original = tuple(v.clone() for v in initial)
    first = pending_calendar_append(initial, dev([2, 1]), dev([0, 1]), dev([4.0, 7.0]))
    for a, b in zip(initial, original):
        assert torch.equal(a, b)
    failed = pending_calendar_append(first, dev([1, 2]), dev([0, 0]), dev([8.0, 9.0]))
    assert failed.overflow.item()
    for a, b in zip(failed[:-1], first[:-1]):
        assert torch.equal(a, b)
    state = pending_calendar_append(
        pending_calendar_empty(4, device="cuda", dtype=torch.float64), dev([1, -999, 2, 999]),
        dev([0, -999, 1, 999]), dev([3.0, float("nan"), 5.0, float("inf")]),
        dev([True, False, True, False]))
    assert state.size.item() == 2
    nxt, payload = pending_calendar_pop(state, dev([1]), width=2)
    assert payload.tolist() == [3.0, 0.0] and nxt.size.item() == 1
    assert nxt.slots.tolist() == [2, 0, 0, 0] and nxt.indices.tolist() == [1, 0, 0, 0]
    assert nxt.values.tolist() == [5.0, 0.0, 0.0, 0.0]
    # ordered pop with cancellation, against the legacy chunk reduction
    for tdt in (torch.float32, torch.float64):
        gen = torch.Generator().manual_seed(31)
        chunks = []
        for _ in range(42):
            chunks.append((torch.randint(0, 4, (7,), generator=gen),
                           (torch.randn(7, generator=gen) * 7).to(tdt)))
        chunks.append((torch.tensor([0, 0, 0]), torch.tensor([1000.0, 0.25, -1000.0],
                                                             dtype=tdt)))
        st = pending_calendar_from_legacy({3: chunks}, device="cuda", dtype=tdt)
        base = torch.tensor([2.0, -3.0, 0.25, 8.0], dtype=tdt)
        expected = base.clone()
        for i, v in chunks:
            expected.index_add_(0, i, v)
        b = base.cuda()
        nxt, payload = pending_calendar_pop(st, 3, width=4, base=b)
        assert torch.equal(payload.cpu(), expected)
        assert torch.equal(b.cpu(), base) and nxt.size.item() == 0
        assert not bool(nxt.values.any())


def test_value_gradients_and_legacy_round_trip():
    from dendra_amd.calendar import (pending_calendar_append, pending_calendar_empty,
                                     pending_calendar_from_legacy, pending_calendar_pop,
                                     pending_calendar_to_legacy)

    slots, indices = dev([2, 1, 2]), dev([0, 1, 0])
    values = torch.tensor([[3.0, 4.0, 5.0], [6.0, 7.0, 8.0]], dtype=torch.float64,
                          device="cuda", requires_grad=True)
    outs = []
    for lane, slot in enumerate((2, 1)):
        st = pending_calendar_append(
            pending_calendar_empty(5, device="cuda", dtype=torch.float64), slots, indices,
            values[lane])
        outs.append(pending_calendar_pop(st, dev([slot]), width=2)[1])
    out = torch.stack(outs)
    assert out.tolist() == [[8.0, 0.0], [0.0, 7.0]]
    out.sum().backward()
    assert values.grad.tolist() == [[1.0, 0.0, 1.0], [0.0, 1.0, 0.0]]
    # survivors carry their gradient through the compaction too
    v = torch.tensor([1.0, 3.0, 2.0, 4.0], dtype=torch.float64, device="cuda",
                     requires_grad=True)
    st = pending_calendar_append(pending_calendar_empty(8, device="cuda",
                                                        dtype=torch.float64),
                                 dev([1, 3, 2, 1]), dev([0, 0, 0, 0]), v)
    nxt, pay = pending_calendar_pop(st, 1, width=1)
    (pay.sum() + 10 * nxt.values.sum()).backward()
    assert v.grad.tolist() == [1.0, 10.0, 10.0, 1.0]
    legacy = pending_calendar_to_legacy(st)
    assert list(legacy) == [1, 3, 2]
    assert legacy[1][0][1].tolist() == [1.0, 4.0]
    restored = pending_calendar_from_legacy(legacy, capacity=8, device="cuda",
                                            dtype=torch.float64)
    _, payload = pending_calendar_pop(restored, 1, width=1)
    assert payload.tolist() == [5.0]
    with pytest.raises(ValueError, match="cannot hold"):
        pending_calendar_from_legacy(legacy, capacity=3, device="cuda", dtype=torch.float64)


def test_zero_capacity_scratch_and_large_int32_conservation():
    from dendra_amd.calendar import (pending_calendar_append, pending_calendar_empty,
                                     pending_calendar_pop)

    st = pending_calendar_empty(0, device="cuda", dtype=torch.float64)
    st = pending_calendar_append(st, dev(np.zeros(0, np.int64)), dev(np.zeros(0, np.int64)),
                                 torch.empty(0, dtype=torch.float64, device="cuda"))
    nxt, payload = pending_calendar_pop(st, 0, width=0)
    assert nxt.size.item() == 0 and payload.shape == (0,)
    # donated scratch: in place, added to what it holds
    st = pending_calendar_append(pending_calendar_empty(4, device="cuda", dtype=torch.float64),
                                 dev([2, 2]), dev([0, 1]), dev([1.0, 2.0]))
    scratch = dev([0.25, -3.0], torch.float64)
    ptr = scratch.data_ptr()
    _, out = pending_calendar_pop(st, 2, width=2, _scratch=scratch)
    assert out.data_ptr() == ptr and out.tolist() == [1.25, -1.0]
    # a million entries over 200 slots: each slot's payload sums exactly
    rng = np.random.default_rng(9)
    n, width, depth = 1_000_000, 5000, 200
    slots = rng.integers(0, depth, n)
    idx = rng.integers(0, width, n)
    vals = rng.integers(-3, 4, n).astype(np.int32)
    st = pending_calendar_append(pending_calendar_empty(n, device="cuda", dtype=torch.int32),
                                 dev(slots), dev(idx), dev(vals))
    left = n
    for s in (0, 57, 199):
        st, pay = pending_calendar_pop(st, s, width=width)
        m = slots == s
        want = np.bincount(idx[m], weights=vals[m], minlength=width).astype(np.int64)
        assert np.array_equal(pay.cpu().numpy().astype(np.int64), want)
        left -= int(m.sum())
        assert int(st.size.item()) == left
        slots, idx, vals = slots[~m], idx[~m], vals[~m]
        assert np.array_equal(st.slots[:left].cpu().numpy(), slots)
