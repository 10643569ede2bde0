"""Pins of the CPU oracle against known answers, closed forms and invariants.

Every oracle function is checked against something other than its own
formula: dense linear algebra, the reference's known-answer tests and
tapes, exact solutions of the ODEs it integrates, or brute force.
"""

import math

import numpy as np
import pytest


def _bands(seed, B, K):
    rs = np.random.default_rng(seed)
    a = 0.35 * rs.standard_normal((B, K - 1))
    c = 0.35 * rs.standard_normal((B, K - 1))
    b = 1 + rs.random((B, K)) + np.pad(np.abs(a), ((0, 0), (1, 0))) + np.pad(np.abs(c), ((0, 0), (0, 1)))
    d = rs.standard_normal((B, K))
    return a, b, c, d


@pytest.mark.parametrize("K", [1, 2, 3, 7, 16, 33, 101, 257])
def test_thomas_and_pcr_match_dense(orc, K):
    a, b, c, d = _bands(K, 3, K)
    x = orc.thomas(a, b, c, d)
    y = orc.pcr(a, b, c, d)
    for r in range(3):
        A = np.diag(b[r]) + np.diag(a[r], -1) + np.diag(c[r], 1)
        np.testing.assert_allclose(x[r], np.linalg.solve(A, d[r]), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(y, x, rtol=1e-11, atol=1e-12)


def test_thomas_known_answer(orc):
    """tests/test_voltage_solvers.py:98-128 of the reference."""
    x = orc.thomas([[1.0, 2.0]], [[4.0, 5.0, 6.0]], [[3.0, 4.0]], [[1.0, 2.0, 3.0]])
    A = np.array([[4, 3, 0], [1, 5, 4], [0, 2, 6]], dtype=float)
    np.testing.assert_allclose(x[0], np.linalg.solve(A, [1, 2, 3]), rtol=1e-14)


def test_dhs_known_answers_and_dense(orc):
    from dendra_amd.topology import build_dhs_layers, build_morphology

    # 2-node KAT, tests/test_voltage_solvers.py:22-49
    x = orc.dhs([[1.0, 2.0]], [[0.0, 0.5]], [[5.0, 3.0]], [-1, 0], [1, 0], [0, 1, 2])
    np.testing.assert_allclose(x[0], np.linalg.solve([[1.5, -0.5], [-0.5, 2.5]], [5, 3]))
    # 7-node tree of tests/test_cuda_solver_parity.py:32-55
    parent = [-1, 0, 0, 1, 1, 2, 5]
    _, _, depth = build_morphology(parent)
    for threads in (1, 2, 32):
        order, lp = build_dhs_layers(depth, threads)
        d = np.linspace(1.3, 2.6, 35).reshape(5, 7)
        a = np.linspace(0.015, 0.095, 35).reshape(5, 7)
        a[:, 0] = 0
        rhs = np.linspace(-1.7, 2.1, 35).reshape(5, 7)
        x = orc.dhs(d, a, rhs, parent, order, lp)
        for r in range(5):
            M = np.diag(d[r] + a[r])
            for ch, p in enumerate(parent):
                if p >= 0:
                    M[p, p] += a[r, ch]; M[ch, p] -= a[r, ch]; M[p, ch] -= a[r, ch]
            np.testing.assert_allclose(x[r], np.linalg.solve(M, rhs[r]), rtol=1e-13)


def test_vtrap_and_phi1_against_direct_forms(orc):
    for x in (-30.0, -7.0, -2.0, -1e-3, 1e-9, 0.3, 4.9, 5.1, 25.0):
        want = x / math.expm1(x / 10.0) if x != 0 else 10.0
        assert orc.vtrap(x, 10.0) == pytest.approx(want, rel=1e-14)
    assert orc.vtrap(0.0, 10.0) == 10.0          # removable singularity: y - x/2
    for z in (-0.5, -0.2, -1e-4, 1e-6, 0.25, 0.5):
        assert orc.phi1(z) == pytest.approx(math.expm1(z) / z, rel=1e-15)
    assert orc.phi1(0.0) == 1.0


@pytest.mark.parametrize("xa,xb,dt", [(0.2, 1.0, 0.1), (3.0, 40.0, 0.025), (0.0, 2.0, 0.5),
                                      (0.05, 0.3, 0.01), (7.0, 9.0, 1.0)])
def test_cnexp_is_the_exact_relaxation(orc, xa, xb, dt):
    x0 = 0.37
    eq = xa / xb
    want = eq + (x0 - eq) * math.exp(-xb * dt)
    assert orc.cnexp(x0, xa, xb, dt) == pytest.approx(want, rel=1e-13, abs=1e-15)


def _hh_rates_numpy(v):
    """HH (1952) rates in the modern sign convention, independent of the oracle."""
    am = 0.1 * (v + 40) / (1 - np.exp(-(v + 40) / 10))
    bm = 4 * np.exp(-(v + 65) / 18)
    ah = 0.07 * np.exp(-(v + 65) / 20)
    bh = 1 / (1 + np.exp(-(v + 35) / 10))
    an = 0.01 * (v + 55) / (1 - np.exp(-(v + 55) / 10))
    bn = 0.125 * np.exp(-(v + 65) / 80)
    return am, bm, ah, bh, an, bn


def test_hh_steady_state_and_one_step(orc):
    from dendra_amd.model import HH

    hh = HH(celsius=16.3)  # q10 = 3
    for v in (-80.0, -65.0, -52.0, -20.0, 10.0):
        am, bm, ah, bh, an, bn = _hh_rates_numpy(v)
        np.testing.assert_allclose(orc.hh_steady(v), (am / (am + bm), ah / (ah + bh),
                                                      an / (an + bn)), rtol=1e-12)
        m0, h0, n0 = 0.3, 0.5, 0.4
        dt = 0.025
        m, h, n, G, J = orc.hh_advance(v, dt, hh.array(), m0, h0, n0)
        q = 3.0
        for x0, x1, a_, b_ in ((m0, m, am, bm), (h0, h, ah, bh), (n0, n, an, bn)):
            inf = a_ / (a_ + b_)
            assert x1 == pytest.approx(inf + (x0 - inf) * math.exp(-q * (a_ + b_) * dt),
                                       rel=1e-12)
        gna, gk = 0.12 * m**3 * h, 0.036 * n**4
        assert G == pytest.approx(0.0003 + gna + gk, rel=1e-14)
        assert J == pytest.approx(0.0003 * -54.3 + gna * 50 + gk * -77, rel=1e-13)


def test_exp2syn_factor_normalises_the_peak(orc):
    for tau1, tau2 in ((0.5, 5.0), (0.1, 10.0), (1.0, 3.0)):
        f = orc.exp2syn_factor(tau1, tau2)
        # peak of exp(-t/tau2) - exp(-t/tau1) by golden-section search
        lo, hi = 0.0, 5 * tau2
        g = (math.sqrt(5) - 1) / 2
        fn = lambda t: math.exp(-t / tau2) - math.exp(-t / tau1)
        for _ in range(200):
            x1, x2 = hi - g * (hi - lo), lo + g * (hi - lo)
            if fn(x1) < fn(x2):
                lo = x1
            else:
                hi = x2
        assert f * fn(0.5 * (lo + hi)) == pytest.approx(1.0, rel=1e-12)


def _passive_spec(K, batch=2):
    from dendra_amd.model import CableSpec

    off = np.arange(batch + 1, dtype=np.int64) * K
    total = batch * K
    rs = np.random.default_rng(4)
    return CableSpec(off, 1.5 + rs.random(total), np.full(total, 12.0),
                     np.full(total, 90.0), np.ones(total), sna=np.zeros(total),
                     sk=np.zeros(total), v_init=-65.0)


def test_passive_cable_equals_matrix_power(orc):
    """With gNa = gK = 0 each step is v <- (I - dt A)^-1 (v + dt b): compare the
    oracle run with the dense linear recurrence built from geometry."""
    from dendra_amd.model import HH, cable_workspace

    K, n_steps, dt = 9, 40, 0.05
    spec = _passive_spec(K)
    spec.iamp = np.zeros(spec.total)
    spec.iamp[0] = 2e-7
    wave = np.ones(n_steps)
    hh = HH()
    st, _ = orc.cable_run(spec, hh.array(), dt, n_steps, i_wave=wave)
    ws = cable_workspace(spec)
    dts = dt * 1e-3
    for c in range(2):
        s = slice(c * K, (c + 1) * K)
        Lap = np.diag(-(ws.gL[s] + ws.gRp[s])) + np.diag(ws.gL[s][:-1], 1) + np.diag(ws.gRp[s][1:], -1)
        A = Lap - np.diag(hh.gl * ws.scale[s])
        b = hh.gl * hh.el * ws.scale[s] + ws.drive[s]
        v = np.full(K, -65.0)
        M = np.eye(K) - dts * A
        for _ in range(n_steps):
            v = np.linalg.solve(M, v + dts * b)
        np.testing.assert_allclose(st["v"][s], v, rtol=0, atol=1e-10)


def test_resting_hh_cable_is_stationary(orc):
    """A uniform HH cable at its resting potential stays there (the increment
    form cancels the capacitive term exactly)."""
    from dendra_amd.model import HH, CableSpec

    hh = HH()
    # find the resting potential of the point model
    v = -65.0
    for _ in range(200):
        st, _ = orc.cable_run(CableSpec(np.array([0, 1]), np.array([10.0]),
                                        np.array([10.0]), np.array([100.0]),
                                        np.array([1.0]), v_init=v), hh.array(), 1.0, 200)
        if abs(st["v"][0] - v) < 1e-12:
            break
        v = st["v"][0]
    K = 31
    spec = CableSpec(np.array([0, K]), np.full(K, 3.0), np.full(K, 10.0),
                     np.full(K, 100.0), np.ones(K), v_init=v)
    st, _ = orc.cable_run(spec, hh.array(), 0.025, 400)
    assert np.abs(st["v"] - v).max() < 1e-9


def test_unbranched_tree_equals_cable(orc):
    """The tree path (current form, DHS) on a path graph gives the cable
    path's (row-normalised form, Thomas) trajectory."""
    from dendra_amd.model import HH, CableSpec, TreeSpec
    from dendra_amd.workloads import extracellular_wave

    K = 60
    rs = np.random.default_rng(2)
    d = 2 + rs.random(K)
    lead = 300 / (4 * np.pi * np.sqrt((np.arange(K) * 10.0 - 200) ** 2 + 100.0**2) * 1e-4)
    cab = CableSpec(np.array([0, K]), d, np.full(K, 10.0), np.full(K, 100.0), np.ones(K),
                    lead=lead, ve_amp=np.array([-2.0]), v_init=-65.0)
    tre = TreeSpec(np.array([0, K]), np.arange(-1, K - 1), d, np.full(K, 10.0),
                   np.full(K, 100.0), np.ones(K), lead=lead, ve_amp=np.array([-2.0]),
                   v_init=-65.0)
    wave = extracellular_wave(600, 0.01, np.float64)
    hh = HH()
    _, rc = orc.cable_run(cab, hh.array(), 0.01, 600, ve_wave=wave, record_every=20)
    _, rt = orc.tree_run(tre, hh.array(), 0.01, 600, ve_wave=wave, record_every=20)
    assert rc.max() > 0, "the pulse must excite the fibre"
    np.testing.assert_allclose(rt, rc, rtol=0, atol=1e-8)


# ---- event delivery: the reference's own tapes ----------------------------------

PRE = [0, 0, 1, 1, 2, 2]            # source-major edges of _fan_network
POST = [0, 1, 0, 1, 0, 1]
DELAY = [1, 2, 3, 1, 2, 3]
WEIGHTS = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]


def _reference_trace(tape):
    """tests/networks/test_backend_equivalence.py:73-90 (brute-force dict)."""
    pending, state, trace = {}, np.zeros(2), []
    for step, spikes in enumerate(tape):
        state = state + pending.pop(step, np.zeros(2))
        trace.append(state.copy())
        for e, (s, t, d) in enumerate(zip(PRE, POST, DELAY)):
            if spikes[s]:
                due = pending.setdefault(step + d, np.zeros(2))
                due[t] += WEIGHTS[e]
    return np.array(trace)


FAN_TAPE = [(1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 0, 0), (0, 1, 0), (1, 0, 0), (0, 0, 0),
            (0, 0, 0), (0, 0, 0)]


def test_ring_replay_matches_reference_fan_tape(orc):
    delivered, _, cur = orc.ring_replay(3, 2, PRE, POST, DELAY, WEIGHTS, 4, FAN_TAPE)
    np.testing.assert_allclose(np.cumsum(delivered, axis=0), _reference_trace(FAN_TAPE),
                               atol=1e-6)
    assert cur == len(FAN_TAPE) % 4


def test_ring_replay_scheduled_collisions_known_answer(orc):
    """Scheduled events with collisions and a negative weight end at exactly
    [11.25, 7.0] (test_backend_equivalence.py:148-161); dt = 0.1 ms."""
    sched = [(0, 0.0, 2.0), (3, 0.1, -0.5), (5, 0.2, 1.5), (0, 0.4, 0.25), (2, 0.0, 3.0)]
    gates = np.zeros((9, 6), np.float32)
    for e, t_ms, g in sched:
        gates[int(round(t_ms / 0.1)), e] += g
    delivered, _, _ = orc.ring_replay_gates(2, POST, DELAY, WEIGHTS, 4, gates)
    assert np.cumsum(delivered, axis=0)[-1].tolist() == [11.25, 7.0]


def test_ring_replay_is_sequential_float32(orc):
    """Collisions accumulate in connection order with float32 rounding."""
    n = 7
    w = np.array([1e8, 1.0, -1e8, 1.0, 3.0, 1e-3, 2.0], np.float32)
    delivered, _, _ = orc.ring_replay(1, 1, [0] * n, [0] * n, [1] * n, w, 2, [[1], [0]])
    acc = np.float32(0)
    for x in w:
        acc = np.float32(acc + x)
    assert delivered[1, 0] == acc


def test_delay_discretisation_round_half_even(orc):
    from dendra_amd.network import delay_steps

    ms = np.array([0.0, 0.01, 0.0125, 0.025, 0.0375, 0.05, 0.0625, 1.0, 5.0])
    want = np.maximum(np.rint(ms / 0.025), 1).astype(np.int64)
    np.testing.assert_array_equal(orc.delay_steps(ms, 0.025), want)
    np.testing.assert_array_equal(delay_steps(ms, 0.025), want)
    # exact binary halves: 1.5 -> 2, 2.5 -> 2, 0.5 -> 0 -> clamped to 1
    halves = np.array([0.375, 0.625, 0.125])
    assert orc.delay_steps(halves, 0.25).tolist() == [2, 2, 1]
    assert delay_steps(halves, 0.25).tolist() == [2, 2, 1]


def test_counter_rng_matches_oracle(orc):
    from dendra_amd import rng

    for seed, stream, i in ((1, 77, 0), (11, 77, 123456), (3, 5, 2**40)):
        assert rng.uniform(seed, stream, [i])[0] == orc.uniform(seed, stream, i)


def test_network_oracle_single_cell_fires_from_background(orc):
    """One point cell with a strong Poisson source: every background event
    arrives one step later, and a spike follows."""
    from dendra_amd.model import HH, CableSpec
    from dendra_amd.network import Connectivity, NetworkSpec

    cells = CableSpec(np.array([0, 1]), np.array([20.0]), np.array([20.0]),
                      np.array([100.0]), np.array([1.0]), v_init=-65.0)
    conn = Connectivity(np.array([1], np.int32), np.array([0], np.int32),
                        np.array([1], np.int32), np.array([0.05], np.float32), 2, 1)
    spec = NetworkSpec(cells, 1, np.array([0]), conn, bg_rate_hz=100.0, seed=3)
    st, spikes = orc.network_run(spec, HH().array(), 0.025, 4000)
    n_bg = int(spikes[:, 1].sum())
    n_cell = int(spikes[:, 0].sum())
    assert n_bg > 3 and n_cell >= 1
    assert spikes[:, 0].argmax() > spikes[:, 1].argmax()


def _one_event_net(k_tgt, soma_local, delay, w):
    """Two passive cells: cell 0 (one compartment, started at +20 mV above the
    0 mV threshold) spikes exactly once, at step 0, and projects one event of
    weight ``w`` (uS) with ``delay`` steps onto the exp2syn of cell 1
    (``k_tgt`` compartments, synapse on compartment ``soma_local``)."""
    from dendra_amd.model import CableSpec
    from dendra_amd.network import Connectivity, NetworkSpec

    off = np.array([0, 1, 1 + k_tgt], dtype=np.int64)
    total = 1 + k_tgt
    diam = np.concatenate([[20.0], np.linspace(2.0, 3.0, k_tgt)])
    dx = np.concatenate([[20.0], np.full(k_tgt, 15.0)])
    cells = CableSpec(off, diam, dx, np.full(total, 110.0), np.full(total, 1.3),
                      sna=np.zeros(total), sk=np.zeros(total), v_init=-54.3)
    conn = Connectivity(np.array([0], np.int32), np.array([1], np.int32),
                        np.array([delay], np.int32), np.array([w], np.float64), 2, 2)
    spec = NetworkSpec(cells, 1 if k_tgt > 1 else 2, np.array([0, 1 + soma_local]), conn,
                       tau1=0.5, tau2=5.0, e_syn=-10.0, theta=0.0, bg_rate_hz=0.0, seed=5)
    v0 = np.full(total, -54.3)
    v0[0] = 20.0
    return spec, v0


@pytest.mark.parametrize("n_steps", [1, 4, 5, 6, 7, 20, 161])
def test_network_oracle_exp2syn_single_event_closed_form(orc, n_steps):
    """One event of weight w sent at step 0 with a 5-step delay is delivered
    at the start of step 5 and then decays by the exact exponential in the
    same step (synapses advance before the populations, net.py:335-377), so
    after N steps B - A = w * factor * (exp(-k dt/tau2) - exp(-k dt/tau1)) with
    k = N - 5 (zero for N <= 5); the factor makes the peak of that difference
    equal to w."""
    from dendra_amd.model import HH

    d, w, dt = 5, 0.004, 0.025
    spec, v0 = _one_event_net(1, 0, d, w)
    st, spikes = orc.network_run(spec, HH().array(), dt, n_steps, v0=v0)
    assert spikes[:, 0].tolist() == [1] + [0] * (n_steps - 1)
    assert not spikes[:, 1].any()
    k = n_steps - d
    tau1, tau2 = 0.5, 5.0
    # factor from the peak time of the two-exponential difference
    tp = tau1 * tau2 / (tau2 - tau1) * math.log(tau2 / tau1)
    factor = 1.0 / (math.exp(-tp / tau2) - math.exp(-tp / tau1))
    a = w * factor * math.exp(-k * dt / tau1) if k > 0 else 0.0
    b = w * factor * math.exp(-k * dt / tau2) if k > 0 else 0.0
    assert st["A"][1] == pytest.approx(a, rel=1e-12, abs=1e-300)
    assert st["B"][1] == pytest.approx(b, rel=1e-12, abs=1e-300)
    assert st["A"][0] == 0.0 and st["B"][0] == 0.0


def _geometry_dense_step(diam, dx, rhoa, cm, gl, el, e_syn, soma, dt):
    """Backward-Euler matrix pieces of a passive cable from geometry alone:
    conductances in S, capacitances in F, potentials in mV, time in s."""
    area = np.pi * diam * 1e-4 * dx * 1e-4                 # cm^2
    r = 0.5 * diam * 1e-4
    R = rhoa * dx * 1e-4 / (np.pi * r * r)                 # ohm, each half-cable pair
    g_ax = 2.0 / (R[:-1] + R[1:])
    C = cm * 1e-6 * area
    K = len(diam)
    Lax = np.zeros((K, K))
    for j, g in enumerate(g_ax):
        Lax[j, j] -= g
        Lax[j + 1, j + 1] -= g
        Lax[j, j + 1] += g
        Lax[j + 1, j] += g
    dts = dt * 1e-3

    def step(v, g_syn_uS):
        gs = g_syn_uS * 1e-6                                # S
        L = Lax.copy()
        L[np.diag_indices(K)] -= gl * area
        L[soma, soma] -= gs
        b = gl * el * area
        b[soma] += gs * e_syn
        L /= C[:, None]
        return np.linalg.solve(np.eye(K) - dts * L, v + dts * b / C)

    return step


@pytest.mark.parametrize("k_tgt,soma", [(1, 0), (6, 2)])
def test_network_oracle_exp2syn_drives_passive_cell_like_dense_recurrence(orc, k_tgt, soma):
    """The synaptic current i = (B - A)(v - e) nA enters the membrane as a
    conductance (B - A) uS on the synapse's compartment (mod/EXP2SYN.py:147-152,
    converted with the compartment area): a passive target cell, driven only by
    one event, follows the dense backward-Euler recurrence built from its
    geometry with g_syn = B - A after each step's decay."""
    from dendra_amd.model import HH

    d, w, dt, n = 3, 0.02, 0.025, 240
    spec, v0 = _one_event_net(k_tgt, soma, d, w)
    hh = HH()
    st, _ = orc.network_run(spec, hh.array(), dt, n, v0=v0)
    c = spec.cells
    sl = slice(1, 1 + k_tgt)
    step = _geometry_dense_step(c.diam[sl], c.dx[sl], c.rhoa[sl], c.cm[sl], hh.gl, hh.el,
                                spec.e_syn, soma, dt)
    tau1, tau2 = 0.5, 5.0
    tp = tau1 * tau2 / (tau2 - tau1) * math.log(tau2 / tau1)
    factor = 1.0 / (math.exp(-tp / tau2) - math.exp(-tp / tau1))
    v = v0[sl].copy()
    for s in range(n):
        k = s - d + 1
        g = w * factor * (math.exp(-k * dt / tau2) - math.exp(-k * dt / tau1)) if k > 0 else 0.0
        v = step(v, g)
    assert np.abs(v - v0[sl]).max() > 1e-3           # the event moved the cell
    np.testing.assert_allclose(st["v"][sl], v, rtol=0, atol=1e-11)
