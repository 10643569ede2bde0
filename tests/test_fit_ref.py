"""Pins of the binary64 autograd restatement (oracle/fit_ref.py) on the CPU.

* its voltages equal the C oracle's cable_run for a rank-one Ve table
  (lead field x waveform), a different formulation of the same step;
* its knot interpolation equals numpy.interp;
* its loss equals the loss recomputed in numpy from the oracle's voltages;
* its autograd gradient equals central differences of itself.
"""

import numpy as np
import pytest
import torch

import fit_ref
from dendra_amd.model import HH, CableSpec, point_source_lead
from dendra_amd.workloads import fit_cables


def _small(n_cells=3, K=31, n_steps=300, n_knots=9, seed=3, gain=20.0):
    # short cables need a stronger field than the 201-compartment workload
    w = fit_cables(n_cells=n_cells, n_comp=K, n_steps=n_steps, dt=0.01, n_knots=n_knots,
                   seed=seed, ve_gain=gain)
    return w


def _ref_inputs(w):
    geom = fit_ref.geometry(w.spec.diam, w.spec.dx, w.spec.rhoa, w.spec.cm, w.spec.off)
    B, K = geom["gL"].shape
    sna = torch.as_tensor(w.sna, dtype=torch.float64)
    sk = torch.as_tensor(w.sk, dtype=torch.float64)
    ve = torch.as_tensor(np.broadcast_to(w.ve, (B,) + w.ve.shape).copy(), dtype=torch.float64)
    kk, ka = fit_ref.knot_table(w.knot_times, w.dt, w.n_steps)
    mw = np.zeros((B, K))
    mw[:, w.sites[w.margin_site]] = w.margin_weight * w.margin_sign / (B * w.n_steps)
    mse = 1.0 / (B * len(w.sites) * w.n_steps)
    return geom, sna, sk, ve, kk, ka, torch.as_tensor(mw), mse


def test_knot_interpolation_equals_numpy_interp():
    tk = np.array([0.0, 0.3, 1.1, 1.2, 2.5])
    k, a = fit_ref.knot_table(tk, 0.01, 250)
    y = np.random.default_rng(0).normal(size=len(tk))
    t = np.arange(250) * 0.01
    got = (1 - a) * y[k] + a * y[k + 1]
    assert np.allclose(got, np.interp(t, tk, y), rtol=0, atol=1e-14)


def test_voltages_equal_c_oracle_rank_one_table(orc):
    w = _small()
    geom, sna, sk, ve, kk, ka, mw, mse = _ref_inputs(w)
    hh = HH(celsius=6.3).array()
    _, rec = fit_ref.fit_loss(geom, sna, sk, ve, kk, ka, hh, w.dt, w.n_steps, w.sites,
                              w.targets.astype(np.float64), mse, mw, w.theta, w.tau,
                              record=True)
    # the same run through the C oracle: Ve = lead_j * I(t), I on the knots
    K = w.ve.shape[0]
    x = np.arange(K) * 10.0
    lead = point_source_lead(x, 0.0, 0.0, x[K // 2], 0.0, 300.0)
    cur = w.ve[K // 2] / lead[K // 2]
    assert np.allclose(np.outer(lead, cur), w.ve, rtol=1e-12, atol=1e-12)
    t = np.arange(w.n_steps) * w.dt
    wave = np.interp(t, w.knot_times, cur)
    spec = CableSpec(off=w.spec.off, diam=w.spec.diam, dx=w.spec.dx, rhoa=w.spec.rhoa,
                     cm=w.spec.cm, sna=w.sna.reshape(-1), sk=w.sk.reshape(-1),
                     lead=np.tile(lead, len(w.spec.off) - 1),
                     ve_amp=np.ones(len(w.spec.off) - 1), v_init=-65.0)
    _, orec = orc.cable_run(spec, hh, w.dt, w.n_steps, ve_wave=wave, record_every=1)
    got = rec.numpy().reshape(w.n_steps, -1)
    assert orec.max() > 0.0, "the pulses must fire at least one cell"
    assert np.abs(got - orec).max() < 1e-9


def test_loss_equals_numpy_loss_of_recorded_voltages():
    w = _small()
    geom, sna, sk, ve, kk, ka, mw, mse = _ref_inputs(w)
    hh = HH(celsius=6.3).array()
    loss, rec = fit_ref.fit_loss(geom, sna, sk, ve, kk, ka, hh, w.dt, w.n_steps, w.sites,
                                 w.targets.astype(np.float64), mse, mw, w.theta, w.tau,
                                 record=True)
    V = rec.numpy()
    d = V[:, :, w.sites] - w.targets.astype(np.float64)
    vm = V[:, :, w.sites[w.margin_site]]
    sig = 1.0 / (1.0 + np.exp(-(vm - w.theta) / w.tau))
    B = V.shape[1]
    want = (d ** 2).sum() * mse + (w.margin_weight / (B * w.n_steps)) * (
        sig * w.margin_sign[None, :]).sum()
    assert abs(loss.item() - want) <= 1e-12 * abs(want)


def test_autograd_equals_central_differences():
    w = _small(n_cells=2, K=11, n_steps=120, n_knots=5, seed=4, gain=400.0)
    geom, sna, sk, ve, kk, ka, mw, mse = _ref_inputs(w)
    hh = HH(celsius=6.3).array()
    ve = ve[0]  # shared table, broadcast inside
    tg = w.targets.astype(np.float64)

    def L(a, b, c):
        return fit_ref.fit_loss(geom, a, b, c, kk, ka, hh, w.dt, w.n_steps, w.sites, tg,
                                mse, mw, w.theta, w.tau)

    p = [sna.clone().requires_grad_(), sk.clone().requires_grad_(),
         ve.clone().requires_grad_()]
    loss = L(*p)
    g = torch.autograd.grad(loss, p)
    rs = np.random.default_rng(1)
    h = 1e-4
    for _ in range(3):
        u = [torch.as_tensor(rs.normal(size=x.shape)) for x in p]
        nrm = sum((x ** 2).sum() for x in u).sqrt()
        u = [x / nrm for x in u]
        with torch.no_grad():
            lp = L(*[x + h * d for x, d in zip(p, u)])
            lm = L(*[x - h * d for x, d in zip(p, u)])
        fd = (lp - lm).item() / (2 * h)
        an = sum((gi * d).sum() for gi, d in zip(g, u)).item()
        assert abs(fd - an) <= 1e-6 * abs(fd) + 1e-12, (fd, an)
