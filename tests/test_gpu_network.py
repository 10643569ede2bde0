"""GPU event delivery and network stepping against the oracle.

Delivery bar: bit-exact.  For identical spike tapes the payload delivered to
every target at every step equals the oracle's ordered float32 ring
(np.array_equal).  Network bar: the GPU's event list replayed through the
oracle ring reproduces the GPU's delivered payloads bit for bit, and its
spike times agree with the binary64 oracle network within one dt.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fan_conn():
    """The reference's _fan_network edges renumbered target-major; the
    per-target accumulation order is unchanged."""
    from dendra_amd.network import Connectivity

    pre = [0, 0, 1, 1, 2, 2]
    post = [0, 1, 0, 1, 0, 1]
    delay = [1, 2, 3, 1, 2, 3]
    w = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    order = np.argsort(post, kind="stable")
    arr = lambda x, dt: np.asarray(x, dt)[order]
    return Connectivity(arr(pre, np.int32), arr(post, np.int32), arr(delay, np.int32),
                        arr(w, np.float32), n_src=3, n_tgt=2), (pre, post, delay, w)


def test_tape_delivery_reference_fan_tape(orc):
    from dendra_amd.network import deliver_tape

    tape = [(1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 0, 0), (0, 1, 0), (1, 0, 0), (0, 0, 0),
            (0, 0, 0), (0, 0, 0)]
    conn, (pre, post, delay, w) = _fan_conn()
    got = deliver_tape(conn, tape)
    want, _, _ = orc.ring_replay(3, 2, pre, post, delay, w, conn.depth, tape)
    assert np.array_equal(got, want)


def test_tape_delivery_63bit_word_boundary(orc):
    """Sources 0, 62, 63, 64 of test_backend_equivalence.py:164 (one-to-one)."""
    from dendra_amd.network import Connectivity, deliver_tape

    n = 65
    conn = Connectivity(np.arange(n, dtype=np.int32), np.arange(n, dtype=np.int32),
                        np.ones(n, np.int32), np.ones(n, np.float32), n_src=n, n_tgt=n)
    tape = np.zeros((2, n), np.uint8)
    tape[0, [0, 62, 63, 64]] = 1
    got = deliver_tape(conn, tape)
    assert np.flatnonzero(got[1]).tolist() == [0, 62, 63, 64]
    want, _, _ = orc.ring_replay(n, n, conn.pre, conn.post, conn.delay, conn.w, 2, tape)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seed", [0, 1])
def test_tape_delivery_random_collisions_bitexact(orc, seed):
    from dendra_amd.network import Connectivity, deliver_tape

    rs = np.random.default_rng(seed)
    n_src, n_tgt, fan = 700, 300, 60
    post = np.repeat(np.arange(n_tgt), fan).astype(np.int32)
    pre = rs.integers(0, n_src, n_tgt * fan).astype(np.int32)
    delay = rs.integers(1, 12, n_tgt * fan).astype(np.int32)
    w = (rs.lognormal(0, 2, n_tgt * fan) * rs.choice([-1, 1], n_tgt * fan)).astype(np.float32)
    conn = Connectivity(pre, post, delay, w, n_src=n_src, n_tgt=n_tgt)
    tape = (rs.random((300, n_src)) < 0.08).astype(np.uint8)
    got = deliver_tape(conn, tape)
    want, _, _ = orc.ring_replay(n_src, n_tgt, pre, post, delay, w, conn.depth, tape)
    assert np.array_equal(got, want)
    assert np.count_nonzero(want) > 1000


def _small_net():
    from dendra_amd.workloads import sparse_network

    return sparse_network(n_point=400, n_bs=24, fan_in=40, seed=5, w_bg=0.02)


def test_network_delivery_bitexact_and_spike_times(orc):
    from dendra_amd.model import HH
    from dendra_amd.network import Network

    spec = _small_net()
    n_steps, dt = 1600, 0.025
    net = Network(spec, dt, HH(celsius=6.3))
    events, delivered = net.run(n_steps, record_events=True, record_delivered=True)
    assert len(events) > 200, "the network must be active"
    n_src = spec.conn.n_src
    tape = np.zeros((n_steps, n_src), np.uint8)
    tape[events[:, 0], events[:, 1]] = 1
    want, _, _ = orc.ring_replay(n_src, spec.cells.n_cells, spec.conn.pre,
                                 spec.conn.post, spec.conn.delay, spec.conn.w,
                                 spec.conn.depth, tape)
    assert np.array_equal(delivered, want), "delivered payloads differ from the oracle"
    # spike times against the binary64 network, within one dt
    _, ref = orc.network_run(spec, HH(celsius=6.3).array(), dt, n_steps)
    ref_ev = np.argwhere(ref)
    got_set = {(int(s), int(c)) for s, c in events}
    ref_set = {(int(s), int(c)) for s, c in ref_ev}
    def covered(a, b):
        return all(any((s + d, c) in b for d in (-1, 0, 1)) for s, c in a)
    horizon = 800  # before chaotic divergence of the recurrent network
    early_got = {e for e in got_set if e[0] < horizon}
    early_ref = {e for e in ref_set if e[0] < horizon}
    assert early_ref and covered(early_got, ref_set) and covered(early_ref, got_set)


def test_network_graph_replay_equals_direct_launches():
    from dendra_amd.model import HH
    from dendra_amd.network import Network

    spec = _small_net()
    a = Network(spec, 0.025, HH(celsius=6.3))
    ea, _ = a.run(400, record_events=True)
    b = Network(spec, 0.025, HH(celsius=6.3))
    eb, _ = b.run(400, record_events=True, graph_steps=50)
    assert np.array_equal(ea, eb)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
