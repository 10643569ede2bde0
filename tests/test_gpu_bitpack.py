"""GPU parity of the netcon_bitpack_ops drop-in (packed source history and
dense delivery) against the oracle: history_ref (the reference's packed
recording and projection, pinned by its fixture) and the dense-ring replay."""

import numpy as np
import pytest
import torch

import history_ref

pytestmark = pytest.mark.gpu

BITS = 63


def _layout(n_src):
    word = np.arange(n_src) // BITS
    mask = np.left_shift(np.int64(1), (np.arange(n_src) % BITS).astype(np.int64))
    return word.astype(np.int64), mask.astype(np.int64)


def dev(a, dtype=None):
    return torch.as_tensor(np.asarray(a), device="cuda", dtype=dtype)


@pytest.mark.parametrize("n_src", [1, 62, 63, 64, 130, 1000, 21000])
def test_pack_source_spikes_matches_packed_recording(n_src):
    from dendra_amd import netcon_bitpack_ops as ops

    assert ops.is_available()
    rs = np.random.default_rng(n_src)
    words = -(-n_src // BITS)
    depth = 5
    hist0 = rs.integers(-2**62, 2**62, size=(depth, words), dtype=np.int64)
    hist = dev(hist0)
    word, mask = _layout(n_src)
    for cur, p in ((3, 0.3), (0, 1.0), (4, 0.0)):
        spikes = rs.random(n_src) < p
        ops.pack_source_spikes(dev(spikes), hist, dev([cur], torch.int64))
        hist0 = history_ref.record_packed(hist0, cur, spikes.astype(np.int8), word, mask)
        np.testing.assert_array_equal(hist.cpu().numpy(), hist0)
    assert (hist.cpu().numpy()[0] >= 0).all(), "bit 63 stays clear"


def _connections(rs, n_src, n_tgt, n_conn, depth):
    pre = rs.integers(0, n_src, n_conn)
    post = rs.integers(0, n_tgt, n_conn)
    delay = rs.integers(1, depth, n_conn)
    return pre, post, delay


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_build_delivery_matches_projection(dtype):
    """Dyadic weights make every sum exact in any order, so the atomic
    delivery is compared bitwise; random weights within float rounding."""
    from dendra_amd import netcon_bitpack_ops as ops

    rs = np.random.default_rng(7)
    n_src, n_tgt, n_conn, depth = 700, 90, 20000, 13
    words = -(-n_src // BITS)
    word, mask = _layout(n_src)
    pre, post, delay = _connections(rs, n_src, n_tgt, n_conn, depth)
    hist = np.zeros((depth, words), np.int64)
    for r in range(depth):
        hist = history_ref.record_packed(hist, r, (rs.random(n_src) < 0.2).astype(np.int8),
                                         word, mask)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    for kind in ("dyadic", "random"):
        w = (rs.integers(1, 512, n_conn) / 256.0 if kind == "dyadic"
             else rs.lognormal(0, 0.5, n_conn)).astype(npdt)
        for cur in (0, 5, depth - 1):
            out = torch.zeros(n_tgt, dtype=dtype, device="cuda")
            ops.build_delivery(dev(hist), dev([cur], torch.int64), dev(delay, torch.int64),
                               dev(word[pre]), dev(mask[pre]), dev(post, torch.int64),
                               dev(w), out)
            want = history_ref.project(hist, cur, w, post, n_tgt, col=word[pre],
                                       offsets=delay, mask=mask[pre], packed=True)
            got = out.cpu().numpy()
            assert want.any()
            if kind == "dyadic":
                np.testing.assert_array_equal(got, want)
            else:
                np.testing.assert_allclose(got, want, rtol=64 * np.finfo(npdt).eps)


@pytest.mark.parametrize("delay_step,used", [(3, 3), (0, 1), (-2, 1), (13, 0), (30, 4)])
def test_build_delivery_uniform_delay_rule(delay_step, used):
    """One delay for all connections: below 1 -> 1, at or beyond the depth ->
    modulo the depth (so 13 on a 13-row ring reads today's row)."""
    from dendra_amd import netcon_bitpack_ops as ops

    rs = np.random.default_rng(11)
    n_src, n_tgt, n_conn, depth = 200, 40, 3000, 13
    word, mask = _layout(n_src)
    pre, post, _ = _connections(rs, n_src, n_tgt, n_conn, depth)
    hist = rs.integers(0, 2**62, size=(depth, -(-n_src // BITS)), dtype=np.int64)
    w = (rs.integers(1, 64, n_conn) / 32.0).astype(np.float32)
    cur = 6
    out = torch.zeros(n_tgt, device="cuda")
    ops.build_delivery_uniform(dev(hist), dev([cur], torch.int64), delay_step, dev(word[pre]),
                               dev(mask[pre]), dev(post, torch.int64), dev(w), out)
    want = history_ref.project(hist, cur, w, post, n_tgt, col=word[pre],
                               offsets=np.full(n_conn, used), mask=mask[pre], packed=True)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_pack_then_deliver_replays_the_dense_ring(orc):
    """Step by step -- deliver today's row from history, then record this
    step's spikes, then advance -- the payloads equal the dense delay ring's
    replay of the same spike tape (reference order of netcon.py:3758-3883)."""
    from dendra_amd import netcon_bitpack_ops as ops

    rs = np.random.default_rng(3)
    n_src, n_tgt, n_conn, depth, n_steps = 500, 60, 8000, 41, 120
    word, mask = _layout(n_src)
    pre, post, delay = _connections(rs, n_src, n_tgt, n_conn, depth)
    w = (rs.integers(1, 256, n_conn) / 128.0).astype(np.float32)
    tape = (rs.random((n_steps, n_src)) < 0.05).astype(np.uint8)
    want, _, _ = orc.ring_replay(n_src, n_tgt, pre.astype(np.int32), post.astype(np.int32),
                                 delay.astype(np.int32), w, depth, tape)
    hist = torch.zeros((depth, -(-n_src // BITS)), dtype=torch.int64, device="cuda")
    cur = torch.zeros(1, dtype=torch.int64, device="cuda")
    args = [dev(delay, torch.int64), dev(word[pre]), dev(mask[pre]), dev(post, torch.int64),
            dev(w)]
    tape_d = dev(tape.astype(bool))
    got = torch.zeros((n_steps, n_tgt), device="cuda")
    for s in range(n_steps):
        ops.build_delivery(hist, cur, *args, got[s])
        ops.pack_source_spikes(tape_d[s], hist, cur)
        cur.add_(1).remainder_(depth)
    assert want.any()
    np.testing.assert_array_equal(got.cpu().numpy(), want)
