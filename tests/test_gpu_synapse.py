"""exp2syn as a per-step receiver (dendra_amd.synapse) and the reference's
Network.step order composed from the drop-ins -- dense ring delivery
(dense_due / dense_fixed_delay_transition), net_receive, rising-edge
detection, the per-step cable integrator with the synapses in its membrane
frame -- against the binary64 network oracle."""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _factor(tau1, tau2):
    tp = tau1 * tau2 / (tau2 - tau1) * math.log(tau2 / tau1)
    return 1.0 / (math.exp(-tp / tau2) - math.exp(-tp / tau1))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-6)])
def test_single_event_gives_the_two_exponential(dtype, tol):
    """One weight w received, then k cnexp steps: B - A = w factor
    (exp(-k dt/tau2) - exp(-k dt/tau1)) (peak w), per synapse parameters."""
    from dendra_amd.synapse import Exp2Syn

    tau1 = np.array([0.5, 0.2, 1.0])
    tau2 = np.array([5.0, 3.0, 1.5])
    syn = Exp2Syn(np.arange(3), tau1, tau2, 0.0, dtype=dtype)
    w = torch.tensor([0.01, 0.5, 2.0], dtype=dtype, device="cuda")
    syn.net_receive(w)
    dt = 0.025
    for k in range(1, 401):
        syn.advance(dt)
        if k in (1, 7, 40, 400):
            got = (syn.B - syn.A).double().cpu().numpy()
            want = np.array([w_ * _factor(a, b) * (math.exp(-k * dt / b) - math.exp(-k * dt / a))
                             for w_, a, b in zip(w.tolist(), tau1, tau2)])
            np.testing.assert_allclose(got, want, rtol=tol * 10, atol=tol * np.abs(want).max())
    vals = {"A": syn.A, "B": syn.B}
    out = syn.net_receive_values(w, dt, vals)["states"]
    f = torch.tensor([_factor(a, b) for a, b in zip(tau1, tau2)], dtype=dtype, device="cuda")
    assert torch.allclose(out["A"], syn.A + w * f, rtol=tol) and out["A"] is not syn.A
    i, g = syn.i_with_conductance(torch.full((3,), -65.0, dtype=dtype, device="cuda"))
    assert torch.allclose(i, g * -65.0)


def test_network_step_order_from_the_drop_ins_matches_the_oracle(orc):
    from dendra_amd import rng
    from dendra_amd.delays import dense_due, dense_fixed_delay_transition
    from dendra_amd.model import HH
    from dendra_amd.stepping import CableStepper
    from dendra_amd.synapse import Exp2Syn
    from dendra_amd.workloads import sparse_network

    spec = sparse_network(n_point=300, n_bs=12, fan_in=30, seed=5, w_bg=0.02)
    n_steps, dt = 700, 0.025
    hh = HH(celsius=6.3)
    want_state, ref = orc.network_run(spec, hh.array(), dt, n_steps)
    f64 = torch.float64
    cells, conn = spec.cells, spec.conn
    n_cells, n_src = cells.n_cells, conn.n_src
    dev = lambda a, t=None: torch.as_tensor(np.asarray(a), device="cuda", dtype=t)
    stepper = CableStepper(cells, hh, dtype=f64)
    syn = Exp2Syn(spec.soma, spec.tau1, spec.tau2, spec.e_syn, dtype=f64)
    stepper.add_point_process(syn)
    depth = conn.depth
    ring = torch.zeros((depth, n_cells), dtype=f64, device="cuda")
    slot, gstep = dev([0], torch.int64), dev([0], torch.int64)
    pre, post = dev(conn.pre, torch.int64), dev(conn.post, torch.int64)
    delay, w = dev(conn.delay, torch.int64), dev(conn.w.astype(np.float64))
    soma = dev(spec.soma, torch.int64)
    v = torch.full((cells.total,), float(cells.v_init), dtype=f64, device="cuda")
    latch = torch.zeros(n_cells, dtype=torch.bool, device="cuda")
    n_bg = n_src - n_cells
    p_bg = spec.bg_rate_hz * dt * 1e-3
    spikes = torch.zeros((n_steps, n_src), dtype=torch.bool, device="cuda")
    for s in range(n_steps):
        syn.net_receive(dense_due(ring, slot).payload)          # deliver today's row
        above = v.index_select(0, soma) >= spec.theta           # detect on v_old
        spikes[s, :n_cells] = above & ~latch
        latch = above
        bg = rng.uniform(spec.seed, spec.bg_stream, s * n_bg + np.arange(n_bg)) < p_bg
        spikes[s, n_cells:] = dev(bg)
        gates = spikes[s].index_select(0, pre).to(f64)
        tr = dense_fixed_delay_transition(ring, slot, gstep, w, gates, post_idx=post,
                                          delay_steps=delay, depth=depth, width=n_cells)
        ring, slot, gstep = tr.queue, tr.slot, tr.global_step
        v, _ = stepper._step(v, dt, hh.celsius)                  # populations
    got = spikes.cpu().numpy()
    assert ref[:, :n_cells].sum() > 30, "the network must fire"
    assert np.array_equal(got, ref.astype(bool))
    err = np.abs(v.cpu().numpy() - want_state["v"]).max()
    gA = np.abs(syn.A.cpu().numpy() - want_state["A"]).max()
    print(f"network step order (float64): spikes equal, max |dV| {err:.3g} mV, "
          f"max |dA| {gA:.3g} uS")
    assert err < 1e-7 and gA < 1e-12
