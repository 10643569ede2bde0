"""GPU dense delay ring with fixed integer delays (dendra_amd.delays):
dense_fixed_delay_transition, dense_due, _enqueue_dense_values and the event
ring, against the oracle's sequential float32 ring replay (orc_ring_replay_gates,
pinned by the reference's tapes and known answers) and, in float64, against
the reference's own scatter_add evaluated on CPU tensors (sequential)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def dev(a, t=None):
    return torch.as_tensor(np.asarray(a), device="cuda", dtype=t)


def _conn(rs, n, width, depth):
    post = rs.integers(0, width, n)
    delay = rs.integers(1, depth, n)
    w = rs.lognormal(np.log(1e-3), 0.5, n)
    return post, delay, w


def _cpu_transition(q, slot, w, g, post, delay, depth, width):
    """The reference's transition on CPU tensors (clear, then the flat
    scatter_add of w * g in connection order)."""
    dest = ((slot + delay) % depth) * width + post
    q = q.clone()
    q[slot] = 0
    return q.reshape(-1).scatter_add(0, dest, w * g).reshape(depth, width)


def test_fixed_delay_transition_replays_the_oracle_ring_float32(orc):
    """Twelve consecutive steps: today's payload (dense_due), then the
    transition; payloads and the final ring are bitwise the oracle's."""
    from dendra_amd.delays import dense_due, dense_fixed_delay_transition

    rs = np.random.default_rng(5)
    n, width, depth, steps = 30000, 400, 41, 12
    post, delay, w = _conn(rs, n, width, depth)
    w = w.astype(np.float32)
    gates = ((rs.random((steps, n)) < 0.2) * rs.uniform(0.5, 1.5, (steps, n))).astype(np.float32)
    ring0 = rs.standard_normal((depth, width)).astype(np.float32)
    want_pay, want_ring, want_cur = orc.ring_replay_gates(width, post, delay, w, depth, gates,
                                                           ring=ring0.copy(), cur=37)
    q, slot, step = dev(ring0), dev([37], torch.int64), dev([0], torch.int64)
    post_t, delay_t, w_t = dev(post, torch.int64), dev(delay, torch.int64), dev(w)
    for s in range(steps):
        due = dense_due(q, slot)
        assert np.array_equal(due.payload.cpu().numpy(), want_pay[s])
        tr = dense_fixed_delay_transition(q, slot, step, w_t, dev(gates[s]), post_idx=post_t,
                                          delay_steps=delay_t, depth=depth, width=width)
        q, slot, step = tr.queue, tr.slot, tr.global_step
        assert tr.event_queue is None
    assert int(slot) == want_cur and int(step) == steps
    assert np.array_equal(q.cpu().numpy().view(np.uint32), want_ring.view(np.uint32))


def test_fixed_delay_transition_float64_is_the_sequential_scatter():
    from dendra_amd.delays import dense_fixed_delay_transition

    rs = np.random.default_rng(6)
    n, width, depth = 20000, 300, 23
    post, delay, w = _conn(rs, n, width, depth)
    g = (rs.random(n) < 0.3) * rs.uniform(0.5, 1.5, n)
    q = rs.standard_normal((depth, width))
    for slot in (0, 22):
        tr = dense_fixed_delay_transition(dev(q), dev([slot], torch.int64), dev([3]), dev(w),
                                          dev(g), post_idx=dev(post, torch.int64),
                                          delay_steps=dev(delay, torch.int64), depth=depth,
                                          width=width)
        want = _cpu_transition(torch.tensor(q), slot, torch.tensor(w), torch.tensor(g),
                               torch.tensor(post), torch.tensor(delay), depth, width)
        assert np.array_equal(tr.queue.cpu().numpy().view(np.uint64),
                              want.numpy().view(np.uint64))
        assert int(tr.slot) == (slot + 1) % depth and int(tr.global_step) == 4


def test_fixed_delay_transition_gradients():
    from dendra_amd.delays import dense_fixed_delay_transition

    rs = np.random.default_rng(8)
    n, width, depth = 60, 7, 5
    post, delay, w = _conn(rs, n, width, depth)
    g = rs.uniform(0.1, 1.0, n)
    args = [dev(rs.standard_normal((depth, width))).requires_grad_(),
            dev(w).requires_grad_(), dev(g).requires_grad_()]
    kw = dict(post_idx=dev(post, torch.int64), delay_steps=dev(delay, torch.int64),
              depth=depth, width=width)
    slot, step = dev([2], torch.int64), dev([0], torch.int64)
    fn = lambda q, w_, g_: dense_fixed_delay_transition(q, slot, step, w_, g_, **kw).queue
    assert torch.autograd.gradcheck(fn, args)
    # the cleared row carries no gradient back to the input ring
    out = fn(*args)
    (gq,) = torch.autograd.grad(out.sum(), args[0])
    assert torch.equal(gq[2], torch.zeros_like(gq[2])) and bool(gq[0].eq(1).all())


def test_event_ring_and_uniform_delay():
    from dendra_amd.delays import connection_event_counts, dense_due, \
        dense_fixed_delay_transition

    rs = np.random.default_rng(9)
    n, width, depth = 500, 30, 9
    post, _, w = _conn(rs, n, width, depth)
    g = dev(((rs.random(n) < 0.4) * 1.0).astype(np.float32))
    ev0 = rs.integers(0, 5, (depth, n)).astype(np.int32)
    counts = connection_event_counts(g, dev(rs.integers(0, 3, n).astype(np.int32)))
    q = torch.zeros((depth, width), device="cuda")
    tr = dense_fixed_delay_transition(q, dev([4], torch.int64), dev([0]), dev(w, torch.float32),
                                      g, post_idx=dev(post, torch.int64),
                                      delay_steps=dev([3], torch.int64), depth=depth,
                                      width=width, event_queue=dev(ev0), event_counts=counts)
    want = ev0.copy()
    want[4] = 0
    want[7] += counts.cpu().numpy()
    np.testing.assert_array_equal(tr.event_queue.cpu().numpy(), want)
    due = dense_due(tr.queue, dev([7], torch.int64), tr.event_queue)
    np.testing.assert_array_equal(due.event_counts.cpu().numpy(), want[7])
    cpu = _cpu_transition(torch.zeros(depth, width), 4, torch.tensor(w, dtype=torch.float32),
                          g.cpu(), torch.tensor(post), torch.full((n,), 3), depth, width)
    assert torch.equal(tr.queue.cpu(), cpu)
    with pytest.raises(ValueError, match="both"):
        dense_fixed_delay_transition(q, dev([4], torch.int64), dev([0]), dev(w, torch.float32),
                                     g, post_idx=dev(post, torch.int64),
                                     delay_steps=dev([3], torch.int64), depth=depth,
                                     width=width, event_queue=dev(ev0))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_enqueue_dense_values_is_the_sequential_scatter(dtype):
    """Arbitrary flat destinations with heavy collisions (and some outside
    the ring, dropped), out of place and in place, with gradients."""
    from dendra_amd.delays import _enqueue_dense_values

    rs = np.random.default_rng(10)
    depth, width, n = 11, 50, 40000
    q = torch.tensor(rs.standard_normal((depth, width)), dtype=dtype)
    dest = torch.tensor(rs.integers(0, depth * width, n))
    vals = torch.tensor(rs.standard_normal(n), dtype=dtype)
    want = q.reshape(-1).scatter_add(0, dest, vals).reshape(depth, width)
    got = _enqueue_dense_values(q.cuda(), dest.cuda(), vals.cuda())
    assert torch.equal(got.cpu(), want)
    ring = q.cuda()
    out = _enqueue_dense_values(ring, dest.cuda(), vals.cuda(), inplace=True)
    assert out.data_ptr() == ring.data_ptr() and torch.equal(ring.cpu(), want)
    qg = q.cuda().requires_grad_()
    vg = vals.cuda().requires_grad_()
    res = _enqueue_dense_values(qg, dest.cuda(), vg)
    wts = torch.randn_like(res)
    dq, dv = torch.autograd.grad((res * wts).sum(), (qg, vg))
    assert torch.equal(dq, wts) and torch.equal(dv, wts.reshape(-1)[dest.cuda()])


def test_fixed_delay_transition_at_network_size(orc):
    """The network's ring: 4.2M connections onto 21,000 targets, depth 201,
    four steps, bitwise the oracle's float32 replay."""
    from dendra_amd.delays import dense_fixed_delay_transition

    rs = np.random.default_rng(12)
    n, width, depth, steps = 4_200_000, 21_000, 201, 4
    post = np.sort(rs.integers(0, width, n))
    delay = rs.integers(40, 201, n)
    w = rs.lognormal(np.log(1e-3), 0.5, n).astype(np.float32)
    gates = (rs.random((steps, n)) < 0.01).astype(np.float32)
    _, want_ring, _ = orc.ring_replay_gates(width, post, delay, w, depth, gates, cur=200)
    q, slot, step = torch.zeros((depth, width), device="cuda"), dev([200], torch.int64), \
        dev([0], torch.int64)
    post_t, delay_t, w_t = dev(post, torch.int64), dev(delay, torch.int64), dev(w)
    for s in range(steps):
        tr = dense_fixed_delay_transition(q, slot, step, w_t, dev(gates[s]), post_idx=post_t,
                                          delay_steps=delay_t, depth=depth, width=width)
        q, slot, step = tr.queue, tr.slot, tr.global_step
    assert np.array_equal(q.cpu().numpy().view(np.uint32), want_ring.view(np.uint32))
