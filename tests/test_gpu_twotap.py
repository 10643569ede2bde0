"""GPU two-tap delay transition (dendra_amd.delays) against the ordered
oracle, bitwise in float32 and float64, and its gradients against autograd of
the reference's index_add form."""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from test_twotap_ref import torch_transition

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import twotap_ref  # noqa: E402

pytestmark = pytest.mark.gpu


def case(rng, dtype, depth=201, width=300, n=5000):
    q = rng.standard_normal((depth, width)).astype(dtype)
    w = rng.lognormal(np.log(1e-3), 0.5, n).astype(dtype)
    g = (rng.uniform(0, 1, n) < 0.3).astype(dtype) * rng.uniform(0.5, 1.5, n).astype(dtype)
    delays = rng.uniform(0.5, 5.0, n)            # ms
    post = rng.integers(0, width, n)
    return q, w, g, delays, post


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_two_tap_transition_is_bitwise_the_ordered_enqueue(dtype):
    from dendra_amd.delays import dense_interpolated_delay_transition

    rng = np.random.default_rng(7)
    q, w, g, delays, post = case(rng, dtype)
    dt = 0.025
    lower, frac = twotap_ref.two_tap_split(delays, dt, dtype)
    dev = lambda a, t=None: torch.as_tensor(a, device="cuda", dtype=t)
    tq = dev(q)
    slot = dev(np.array([195]), torch.int64)
    step = dev(np.array([0]), torch.int64)
    want = q
    s = 195
    for _ in range(8):  # consecutive steps, the slot wrapping past depth
        tr = dense_interpolated_delay_transition(
            tq, slot, step, dev(w), dev(g), dev(delays.astype(dtype)), dt,
            post_idx=dev(post, torch.int64), depth=q.shape[0], width=q.shape[1], taps=2,
            sigma=0.0)
        want = twotap_ref.enqueue(want, s, w, g, frac, lower, post)
        got = tr.queue.cpu().numpy()
        assert got.dtype == want.dtype
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
        s = (s + 1) % q.shape[0]
        assert int(tr.slot) == s and tr.event_queue is None
        tq, slot, step = tr.queue, tr.slot, tr.global_step
    assert int(step) == 8


def test_two_tap_gradients_match_autograd():
    from dendra_amd.delays import dense_interpolated_delay_transition, delay_interpolation

    rng = np.random.default_rng(8)
    q, w, g, delays, post = case(rng, np.float64, depth=31, width=20, n=200)
    delays = rng.uniform(0.3, 0.7, 200)  # dt 0.025: 12..28 steps, all inside the ring
    dt = 0.025
    cot = torch.as_tensor(rng.standard_normal(q.shape))
    slot = 29

    def run(dev, fn):
        args = [torch.as_tensor(v, device=dev).requires_grad_() for v in (q, w, g, delays)]
        out = fn(*args)
        return torch.autograd.grad((out * cot.to(dev)).sum(), args)

    def ref(tq, tw, tg, td):
        lam = (td / dt).clamp_min(1.0)
        lower = lam.floor()
        return torch_transition(tq, slot, tw, tg, lam - lower, lower.long(),
                                torch.as_tensor(post))

    def gpu(tq, tw, tg, td):
        return dense_interpolated_delay_transition(
            tq, torch.tensor([slot], device="cuda"), torch.tensor([0], device="cuda"),
            tw, tg, td, dt, post_idx=torch.as_tensor(post, device="cuda"), depth=31,
            width=20, taps=2, sigma=0.0).queue

    exp = run("cpu", ref)
    got = run("cuda", gpu)
    for a, b in zip(got, exp):
        torch.testing.assert_close(a.cpu(), b, rtol=1e-12, atol=1e-14)
    offsets, fractions = delay_interpolation(torch.as_tensor(delays), torch.tensor(dt),
                                             dtype=torch.float64, taps=2, sigma=0.0)
    assert offsets.dtype == torch.int64 and (offsets >= 1).all()


def test_two_tap_no_connections_and_validation():
    from dendra_amd.delays import dense_interpolated_delay_transition

    q = torch.randn(5, 3, device="cuda")
    e = torch.empty(0, device="cuda")
    tr = dense_interpolated_delay_transition(
        q, torch.tensor([2], device="cuda"), torch.tensor([0], device="cuda"), e, e, e, 0.1,
        post_idx=torch.empty(0, dtype=torch.int64, device="cuda"), depth=5, width=3, taps=2,
        sigma=0.0)
    want = q.clone()
    want[2] = 0
    assert torch.equal(tr.queue, want)
    with pytest.raises(ValueError, match="queue must have shape"):
        dense_interpolated_delay_transition(
            q, torch.tensor([0], device="cuda"), torch.tensor([0], device="cuda"), e, e, e,
            0.1, post_idx=torch.empty(0, dtype=torch.int64, device="cuda"), depth=4, width=3,
            taps=2, sigma=0.0)
    with pytest.raises(ValueError, match=r"post_idx entries must lie in \[0, width\)"):
        one = torch.ones(1, device="cuda")
        dense_interpolated_delay_transition(
            q, torch.tensor([0], device="cuda"), torch.tensor([0], device="cuda"), one, one,
            one, 0.1, post_idx=torch.tensor([3], device="cuda"), depth=5, width=3, taps=2,
            sigma=0.0)
    with pytest.raises(ValueError, match="taps must be 2 or 3"):
        dense_interpolated_delay_transition(
            q, torch.tensor([0], device="cuda"), torch.tensor([0], device="cuda"), e, e, e,
            0.1, post_idx=torch.empty(0, dtype=torch.int64, device="cuda"), depth=5, width=3,
            taps=4, sigma=0.0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_three_tap_transition_is_bitwise_the_flattened_scatter(dtype):
    from dendra_amd.delays import delay_interpolation, dense_interpolated_delay_transition

    rng = np.random.default_rng(9)
    q, w, g, delays, post = case(rng, dtype)
    dt, sigma = 0.025, 0.7
    offsets, fractions = twotap_ref.three_tap_split(delays, dt, sigma, dtype)
    dev = lambda a, t=None: torch.as_tensor(a, device="cuda", dtype=t)
    tr = dense_interpolated_delay_transition(
        dev(q), dev(np.array([200]), torch.int64), dev(np.array([0]), torch.int64), dev(w),
        dev(g), dev(delays.astype(dtype)), dt, post_idx=dev(post, torch.int64),
        depth=q.shape[0], width=q.shape[1], taps=3, sigma=sigma,
        delay_taps=(dev(offsets, torch.int64), dev(fractions)))
    want = twotap_ref.enqueue3(q, 200, w, g, fractions, offsets, post)
    assert np.array_equal(tr.queue.cpu().numpy().view(np.uint8), want.view(np.uint8))
    off_t, frac_t = delay_interpolation(dev(delays), torch.tensor(dt, dtype=torch.float64),
                                        dtype=torch.float64, taps=3, sigma=sigma)
    assert torch.equal(off_t.cpu(), torch.as_tensor(offsets))
    np.testing.assert_allclose(frac_t.cpu().numpy(), fractions.astype(np.float64),
                               rtol=1e-6 if dtype == np.float32 else 1e-12)


def test_three_tap_gradients_match_autograd():
    from dendra_amd.delays import dense_interpolated_delay_transition

    rng = np.random.default_rng(10)
    q, w, g, _, post = case(rng, np.float64, depth=31, width=20, n=200)
    delays = rng.uniform(0.3, 0.7, 200)
    dt, sigma, slot = 0.025, 0.8, 29
    cot = torch.as_tensor(rng.standard_normal(q.shape))

    def run(dev, fn):
        args = [torch.as_tensor(v, device=dev).requires_grad_() for v in (q, w, g, delays)]
        return torch.autograd.grad((fn(*args) * cot.to(dev)).sum(), args)

    def ref(tq, tw, tg, td):
        lam = (td / dt).clamp_min(1.0)
        lower = lam.floor().long()
        offs = torch.stack([lower - 1, lower, lower + 1], -1)
        fr = torch.softmax(-0.5 * ((lam[:, None] - offs.double()) / (sigma + 1e-6)) ** 2, -1)
        out = tq.clone()
        out[slot] = 0
        rows = (slot + offs) % 31
        dest = rows * 20 + torch.as_tensor(post)[:, None]
        vals = (tw * tg)[:, None] * fr
        return out.reshape(-1).index_add(0, dest.reshape(-1), vals.reshape(-1)).reshape(31, 20)

    def gpu(tq, tw, tg, td):
        return dense_interpolated_delay_transition(
            tq, torch.tensor([slot], device="cuda"), torch.tensor([0], device="cuda"), tw, tg,
            td, dt, post_idx=torch.as_tensor(post, device="cuda"), depth=31, width=20, taps=3,
            sigma=sigma).queue

    for a, b in zip(run("cuda", gpu), run("cpu", ref)):
        torch.testing.assert_close(a.cpu(), b, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("taps", [2, 3])
@pytest.mark.parametrize("shape", [(201, 45, 40000, "uniform"), (201, 70, 12000, "skewed"),
                                   (3000, 10, 600, "uniform"), (2, 40, 3000, "uniform")])
def test_long_fan_in_and_deep_rings_are_bitwise_the_oracle(dtype, taps, shape):
    """Targets with hundreds of connections (the tile's connections span many
    staged chunks, so two taps take one pass per tap), one target holding
    half of all connections, widths that are not a multiple of the tile, and
    a ring too deep for the LDS tile (the global-memory form)."""
    from dendra_amd.delays import dense_interpolated_delay_transition

    depth, width, n, kind = shape
    rng = np.random.default_rng(depth + width + taps)
    q, w, g, delays, post = case(rng, dtype, depth=depth, width=width, n=n)
    if kind == "skewed":
        post[: n // 2] = 3
        rng.shuffle(post)
    dt, slot = 0.025, depth - 2
    dev = lambda a, t=None: torch.as_tensor(a, device="cuda", dtype=t)
    kw = dict(post_idx=dev(post, torch.int64), depth=depth, width=width, taps=taps)
    if taps == 2:
        lower, frac = twotap_ref.two_tap_split(delays, dt, dtype)
        tr = dense_interpolated_delay_transition(
            dev(q), dev([slot], torch.int64), dev([0], torch.int64), dev(w), dev(g),
            dev(delays.astype(dtype)), dt, sigma=0.0, **kw)
        want = twotap_ref.enqueue(q, slot, w, g, frac, lower, post)
    else:
        offsets, fractions = twotap_ref.three_tap_split(delays, dt, 0.6, dtype)
        tr = dense_interpolated_delay_transition(
            dev(q), dev([slot], torch.int64), dev([0], torch.int64), dev(w), dev(g),
            dev(delays.astype(dtype)), dt, sigma=0.6,
            delay_taps=(dev(offsets, torch.int64), dev(fractions)), **kw)
        want = twotap_ref.enqueue3(q, slot, w, g, fractions, offsets, post)
    assert np.array_equal(tr.queue.cpu().numpy().view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("taps", [1, 2, 3])
def test_workspace_and_direct_entries_agree_with_the_oracle(dtype, taps):
    """dn_interp_delay_enqueue_ws_* (records packed at their CSR positions,
    the positions passed in or computed in the workspace) and
    dn_interp_delay_enqueue_* (four gathers per connection) give the oracle's
    ring bit for bit; one target holds half of the 12,000 connections, so its
    tile spans several staged chunks."""
    from dendra_amd import _lib
    from dendra_amd.delays import csr_positions, delay_plan

    rng = np.random.default_rng(40 + taps)
    depth, width, n, slot = 201, 70, 12000, 199
    q, w, g, _, post = case(rng, dtype, depth=depth, width=width, n=n)
    post[: n // 2] = 3
    rng.shuffle(post)
    if taps == 3:
        lower = rng.integers(1, 190, n)
        offsets = np.stack([lower - 1, lower, lower + 1], -1)
        f = rng.dirichlet([2.0, 3.0, 2.0], n).astype(dtype)
        base = (offsets[:, 0] % depth).astype(np.int32)
        want = twotap_ref.enqueue3(q, slot, w, g, f, offsets, post)
    else:
        lower = rng.integers(0, 200, n)
        base = (lower % depth).astype(np.int32)
        f = rng.uniform(0, 1, n).astype(dtype)
        if taps == 2:
            want = twotap_ref.enqueue(q, slot, w, g, f, lower, post)
        else:
            want = np.array(q, copy=True)
            want[slot] = 0
            for e in np.argsort(post, kind="stable"):
                r = (slot + lower[e]) % depth
                want[r, post[e]] = dtype(want[r, post[e]] + dtype(w[e]) * dtype(g[e]))
    dev = lambda a, t=None: torch.as_tensor(a, device="cuda", dtype=t)
    in_ptr, in_edge, _ = delay_plan(dev(post, torch.int64), width)
    tw, tg, tf, tb, ts, src = dev(w), dev(g), dev(f), dev(base), dev([slot], torch.int64), dev(q)
    lib, p, sfx = _lib.load(), _lib.ptr, "f32" if dtype == np.float32 else "f64"
    stream = _lib.stream_handle(src.device)
    fp = None if taps == 1 else p(tf)
    direct = torch.empty_like(src)
    assert getattr(lib, f"dn_interp_delay_enqueue_{sfx}")(
        p(src), p(direct), depth, width, p(ts), taps, p(in_ptr), p(in_edge), p(tw), p(tg), fp,
        p(tb), stream) == 0
    nbytes = lib.dn_interp_delay_work_bytes(n, taps, int(dtype == np.float64))
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    packed, own = torch.empty_like(src), torch.empty_like(src)
    fn = getattr(lib, f"dn_interp_delay_enqueue_ws_{sfx}")
    pos = csr_positions(in_edge)
    assert np.array_equal(pos.cpu().numpy()[in_edge.cpu().numpy()], np.arange(n))
    assert fn(p(src), p(packed), depth, width, p(ts), taps, p(in_ptr), p(in_edge), p(pos), p(tw),
              p(tg), fp, p(tb), n, p(work), nbytes, stream) == 0
    assert fn(p(src), p(own), depth, width, p(ts), taps, p(in_ptr), p(in_edge), None, p(tw),
              p(tg), fp, p(tb), n, p(work), nbytes, stream) == 0
    assert fn(p(src), p(packed), depth, width, p(ts), taps, p(in_ptr), p(in_edge), p(pos), p(tw),
              p(tg), fp, p(tb), n, p(work), nbytes - 1, stream) == -1  # too small: DN_EINVAL
    torch.cuda.synchronize()
    for got in (direct, packed, own):
        assert np.array_equal(got.cpu().numpy().view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("qin_null", [True, False])
def test_in_place_enqueue_through_the_c_abi(qin_null):
    """qin NULL or equal to q: the ring is transformed in place (one tap,
    fixed delays), equal to the out-of-place result."""
    from dendra_amd import _lib
    from dendra_amd.delays import delay_plan

    rng = np.random.default_rng(12)
    depth, width, n = 201, 300, 30000
    q, w, g, _, post = case(rng, np.float32, depth=depth, width=width, n=n)
    base = rng.integers(0, depth, n).astype(np.int32)
    dev = lambda a, t=None: torch.as_tensor(a, device="cuda", dtype=t)
    post_t = dev(post, torch.int64)
    in_ptr, in_edge, _ = delay_plan(post_t, width)
    tw, tg, tb, slot = dev(w), dev(g), dev(base), dev([17], torch.int64)
    lib, p = _lib.load(), _lib.ptr
    ref = torch.empty((depth, width), device="cuda")
    src = dev(q)
    st = lib.dn_interp_delay_enqueue_f32(p(src), p(ref), depth, width, p(slot), 1, p(in_ptr),
                                         p(in_edge), p(tw), p(tg), None, p(tb),
                                         _lib.stream_handle(src.device))
    assert st == 0
    inplace = dev(q)
    st = lib.dn_interp_delay_enqueue_f32(None if qin_null else p(inplace), p(inplace), depth,
                                         width, p(slot), 1, p(in_ptr), p(in_edge), p(tw),
                                         p(tg), None, p(tb), _lib.stream_handle(src.device))
    assert st == 0
    torch.cuda.synchronize()
    assert torch.equal(inplace, ref)
    assert torch.equal(src, dev(q))
    want = np.array(q, copy=True)
    want[17] = 0
    for e in np.argsort(post, kind="stable"):
        r = (17 + base[e]) % depth
        want[r, post[e]] = np.float32(want[r, post[e]] + np.float32(w[e]) * np.float32(g[e]))
    assert np.array_equal(ref.cpu().numpy().view(np.uint8), want.view(np.uint8))


def test_delay_plan_follows_each_post_idx_tensor():
    """A freed post_idx whose storage the allocator hands to a new tensor of
    the same size must not lend its plan to it."""
    from dendra_amd.delays import dense_interpolated_delay_transition

    rng = np.random.default_rng(11)
    q, w, g, delays, _ = case(rng, np.float32, depth=31, width=40, n=300)
    delays = rng.uniform(0.3, 0.7, 300).astype(np.float32)
    lower, frac = twotap_ref.two_tap_split(delays, 0.025, np.float32)
    dev = lambda a, t=None: torch.as_tensor(a, device="cuda", dtype=t)
    for seed in range(3):
        post = np.random.default_rng(100 + seed).integers(0, 40, 300)
        p_t = dev(post, torch.int64)
        tr = dense_interpolated_delay_transition(
            dev(q), dev(np.array([3]), torch.int64), dev(np.array([0]), torch.int64), dev(w),
            dev(g), dev(delays), 0.025, post_idx=p_t, depth=31, width=40, taps=2, sigma=0.0)
        want = twotap_ref.enqueue(q, 3, w, g, frac, lower, post)
        assert np.array_equal(tr.queue.cpu().numpy().view(np.uint8), want.view(np.uint8))
        del p_t, tr


def test_cached_tap_rows_follow_in_place_edits_of_the_offsets():
    """The ring's first-tap rows are derived once per offsets tensor; an
    in-place edit of that tensor (a new version) must be seen by the next
    step."""
    from dendra_amd.delays import dense_interpolated_delay_transition

    rng = np.random.default_rng(13)
    q, w, g, delays, post = case(rng, np.float32, depth=31, width=40, n=300)
    dev = lambda a, t=None: torch.as_tensor(a, device="cuda", dtype=t)
    lower = rng.integers(1, 20, 300)
    frac = rng.uniform(0, 1, 300).astype(np.float32)
    off_t, frac_t = dev(lower, torch.int64), dev(frac)
    for edit in range(3):
        tr = dense_interpolated_delay_transition(
            dev(q), dev([5], torch.int64), dev([0], torch.int64), dev(w), dev(g),
            dev(delays.astype(np.float32)), 0.025, post_idx=dev(post, torch.int64), depth=31,
            width=40, taps=2, sigma=0.0, delay_taps=(off_t, frac_t))
        want = twotap_ref.enqueue(q, 5, w, g, frac, off_t.cpu().numpy(), post)
        assert np.array_equal(tr.queue.cpu().numpy().view(np.uint8), want.view(np.uint8))
        off_t.add_(3 + edit)

