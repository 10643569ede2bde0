"""Binary64 PyTorch restatement of the cable trace-fitting loss, for autograd.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by the ``cpu_baseline`` leg
of bench.py, never by the dendra_amd package.

It restates, in eager torch arithmetic, the path through which the reference
obtains config-3 gradients by autograd (``Population.run`` in train mode and
``loss.backward()``, dendra/models/core.py:4555-4614):
  gating        mechanisms/_cnexp.py:96-132 (Taylor phi1 for |z| <= 0.5),
                mechanisms/ops.py:13-128 (binary64 exprelr series |q| <= 0.5),
                mod/HH.py:24-56 and :109-116
  voltage step  integrators/implicit.py:569-664, increment form, with the
                extracellular term as the axial operator applied to Ve
                (:618-638)
  solve         the functional Thomas sweep of integrators/tridiag/pcr.py:125-183
                (autograd runs through the sweep)
  Ve            piecewise-linear knot table in time per compartment
                (stim/waveform/implementations.py:748-796,
                utils/interpolation.py:1140-1240), evaluated on the
                start-of-step grid t_n = n dt
Geometry (area, capacitance, half-cylinder edge conductance) is recomputed
here from diam/dx/Ra/cm (integrators/cable.py:10-29, implicit.py:484-522).

Autograd of ``fit_loss`` is the stand-in for the reference's autograd
gradient.  Pins (tests/test_fit_ref.py): its voltages equal the C oracle's
``cable_run`` for rank-one tables (lead field x waveform), its interpolation
equals numpy.interp, its loss equals the loss recomputed in numpy from the
oracle's recorded voltages, and its gradient equals central differences of
itself.
"""

from __future__ import annotations

import math

import numpy as np
import torch

PI = math.pi


def geometry(diam, dx, rhoa, cm, off):
    """Padded (B, Kmax) gL, gRp, scale and mask from per-compartment arrays."""
    off = np.asarray(off, dtype=np.int64)
    B = len(off) - 1
    sizes = np.diff(off)
    Kmax = int(sizes.max())
    gL = np.zeros((B, Kmax))
    gRp = np.zeros((B, Kmax))
    scale = np.zeros((B, Kmax))
    mask = np.zeros((B, Kmax), dtype=bool)
    for b in range(B):
        o, K = int(off[b]), int(sizes[b])
        d = np.asarray(diam[o:o + K], np.float64)
        L = np.asarray(dx[o:o + K], np.float64)
        ra = np.asarray(rhoa[o:o + K], np.float64)
        c = np.asarray(cm[o:o + K], np.float64)
        area = d * 1e-4 * PI * L * 1e-4
        cap = 1e-6 * c * area
        r = 1e-4 * d / 2.0
        R = ra * (1e-4 * L) / (PI * r * r)
        if K > 1:
            g = 2.0 / (R[:-1] + R[1:])
            gL[b, :K - 1] = g / cap[:-1]
            gRp[b, 1:K] = g / cap[1:]
        scale[b, :K] = area / cap
        mask[b, :K] = True
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    return {"gL": t(gL), "gRp": t(gRp), "scale": t(scale), "mask": torch.as_tensor(mask),
            "sizes": sizes}


def knot_table(knot_times, dt, n_steps):
    """Segment index and weight of every start-of-step time t_n = n dt."""
    tk = np.asarray(knot_times, dtype=np.float64)
    t = np.arange(n_steps, dtype=np.float64) * dt
    k = np.clip(np.searchsorted(tk, t, side="right") - 1, 0, len(tk) - 2)
    a = (t - tk[k]) / (tk[k + 1] - tk[k])
    return k, a


def _vtrap10(x):
    q = x / 10.0
    q2 = q * q
    e = q2 * (43867.0 / 5109094217170944000.0) - 3617.0 / 10670622842880000.0
    for c in (1.0 / 74724249600.0, -691.0 / 1307674368000.0, 1.0 / 47900160.0,
              -1.0 / 1209600.0, 1.0 / 30240.0, -1.0 / 720.0, 1.0 / 12.0):
        e = e * q2 + c
    series = (10.0 - x / 2.0) + x * (q * e)
    small = q.abs() <= 0.5
    qs = torch.where(small, torch.ones_like(q), q)
    xs = torch.where(small, torch.full_like(x, 10.0), x)
    return torch.where(small, series, xs / (torch.exp(qs) - 1.0))


def _phi1(z):
    v = z * (1.0 / math.factorial(17)) + 1.0 / math.factorial(16)
    for k in range(15, 0, -1):
        v = v * z + 1.0 / math.factorial(k)
    return v


def _cnexp(x, xa, xb, dt):
    z = -xb * dt
    inc = x + dt * (xa - xb * x) * _phi1(z)
    eq = xa / xb
    rel = eq + (x - eq) * torch.exp(z)
    return torch.where(z.abs() <= 0.5, inc, rel)


def _axial(g_left, g_right, x):
    """Row j: gL_j (x_{j+1} - x_j) - gRp_j (x_j - x_{j-1}) on the padded grid."""
    delta = x[:, 1:] - x[:, :-1]
    left = g_left[:, :-1] * delta
    right = g_right[:, 1:] * delta
    return torch.cat((left[:, :1], left[:, 1:] - right[:, :-1], -right[:, -1:]), dim=1)


def _thomas(a, b, c, d):
    """Sweep over columns, vectorised over rows; a/c are the (B, K-1) bands."""
    K = b.shape[1]
    if K == 1:
        return d / b
    cp = [c[:, 0] / b[:, 0]]
    dp = [d[:, 0] / b[:, 0]]
    for i in range(1, K):
        den = b[:, i] - a[:, i - 1] * cp[-1]
        if i < K - 1:
            cp.append(c[:, i] / den)
        dp.append((d[:, i] - a[:, i - 1] * dp[-1]) / den)
    x = [dp[-1]]
    for i in range(K - 2, -1, -1):
        x.append(dp[i] - cp[i] * x[-1])
    return torch.stack(x[::-1], dim=1)


def hh_steady(v):
    am = 0.1 * _vtrap10(torch.as_tensor(-(v + 40.0), dtype=torch.float64))
    bm = 4.0 * math.exp(-(v + 65.0) / 18.0)
    ah = 0.07 * math.exp(-(v + 65.0) / 20.0)
    bh = 1.0 / (math.exp(-(v + 35.0) / 10.0) + 1.0)
    an = 0.01 * _vtrap10(torch.as_tensor(-(v + 55.0), dtype=torch.float64))
    bn = 0.125 * math.exp(-(v + 65.0) / 80.0)
    return float(am / (am + bm)), ah / (ah + bh), float(an / (an + bn))


def fit_loss(geom, sna, sk, ve, knot_k, knot_a, hh, dt, n_steps, sites, targets,
             mse_scale, mw, theta, tau, v_init=-65.0, record=False):
    """Loss of an n_steps run; tensors (B, Kmax) float64, ve (B, Kmax, Q).

    sites: (S,) compartment indices; targets: (n_steps, B, S); mw: (B, Kmax)
    margin weights.  hh = (gnabar, gkbar, gl, ena, ek, el, q10).
    Returns the loss (and the (n_steps, B, Kmax) voltage record if asked).
    """
    gnabar, gkbar, gl, ena, ek, el, q10 = (float(x) for x in hh)
    gL, gRp, scale, mask = geom["gL"], geom["gRp"], geom["scale"], geom["mask"]
    B, K = gL.shape
    if ve.dim() == 2:
        ve = ve.expand(B, -1, -1)
    dts = dt * 1e-3
    m0, h0, n0 = hh_steady(v_init)
    v = torch.full((B, K), v_init, dtype=torch.float64)
    m = torch.full((B, K), m0, dtype=torch.float64)
    h = torch.full((B, K), h0, dtype=torch.float64)
    n = torch.full((B, K), n0, dtype=torch.float64)
    sites = torch.as_tensor(np.asarray(sites), dtype=torch.long)
    targets = torch.as_tensor(targets, dtype=torch.float64)
    lower = -dts * gRp[:, 1:]
    upper = -dts * gL[:, :-1]
    loss = torch.zeros((), dtype=torch.float64)
    rec = []
    for s in range(n_steps):
        am = 0.1 * _vtrap10(-(v + 40.0))
        bm = 4.0 * torch.exp(-(v + 65.0) / 18.0)
        ah = 0.07 * torch.exp(-(v + 65.0) / 20.0)
        bh = 1.0 / (torch.exp(-(v + 35.0) / 10.0) + 1.0)
        an = 0.01 * _vtrap10(-(v + 55.0))
        bn = 0.125 * torch.exp(-(v + 65.0) / 80.0)
        m = _cnexp(m, q10 * am, q10 * (am + bm), dt)
        h = _cnexp(h, q10 * ah, q10 * (ah + bh), dt)
        n = _cnexp(n, q10 * an, q10 * (an + bn), dt)
        gna = gnabar * sna * m ** 3 * h
        gk = gkbar * sk * n ** 4
        G = gl + gna + gk
        J = gl * el + gna * ena + gk * ek
        k, a = int(knot_k[s]), float(knot_a[s])
        vet = (1.0 - a) * ve[..., k] + a * ve[..., k + 1]
        f = (J - G * v) * scale + _axial(gL, gRp, v) + _axial(gL, gRp, vet)
        f = torch.where(mask, f, torch.zeros_like(f))
        main = 1.0 + dts * (gL + gRp + G * scale)
        v = v + _thomas(lower, main, upper, dts * f)
        vs = v[:, sites]
        loss = loss + mse_scale * ((vs - targets[s]) ** 2).sum()
        loss = loss + (mw * torch.sigmoid((v - theta) / tau)).sum()
        if record:
            rec.append(v.detach().clone())
    if record:
        return loss, torch.stack(rec)
    return loss
