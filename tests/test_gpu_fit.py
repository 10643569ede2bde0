"""GPU loss and hand-written adjoint of the cable trace fit.

Bars:
* float64 loss and gradients equal the binary64 autograd restatement
  (oracle/fit_ref.py, itself pinned to the C oracle) to 1e-9 relative, over
  several checkpoint segments, for shared and per-compartment Ve tables,
  equal and ragged cells, 4 and 8 rows per lane;
* at the full 512 x 201 x 10,000-step size, the float64 adjoint agrees with
  central finite differences of the float64 forward along 32 seeded unit
  directions: the differences at h = 1e-3 and 5e-4 show the pure O(h^2)
  truncation signature around the adjoint value, and at h = 1e-5 they agree
  within 1e-4 relative;
* the float32 path (the benchmarked one) is compared with float64: on a
  small run gradients within 2e-3 relative; at full size the loss within
  1e-3 and the per-cell gradient errors within 1e-3 (median) and 1e-2
  (90th percentile).
"""

import numpy as np
import pytest
import torch

import fit_ref
from dendra_amd.model import HH, CableSpec
from dendra_amd.workloads import fit_cables

pytestmark = pytest.mark.gpu


def _fit(w, dtype, ckpt_every=500, rmax=4, shared=True, spec=None, waves=1):
    from dendra_amd.fit import CableFit

    return CableFit(spec or w.spec, w.dt, w.n_steps, w.sites, w.targets,
                    knot_times=w.knot_times, margin_site=w.margin_site,
                    margin_sign=w.margin_sign, margin_weight=w.margin_weight,
                    theta=w.theta, tau=w.tau, ckpt_every=ckpt_every, hh=HH(celsius=6.3),
                    shared_ve=shared, dtype=dtype, rmax=rmax, waves=waves)


def _ref(w, spec, sna, sk, ve_rows, shared):
    """Loss and gradients of the restatement; sna/sk/ve_rows in spec order."""
    geom = fit_ref.geometry(spec.diam, spec.dx, spec.rhoa, spec.cm, spec.off)
    B, K = geom["gL"].shape
    sizes = np.diff(spec.off)
    pad = lambda flat, fill: torch.as_tensor(np.stack([
        np.concatenate([flat[spec.off[b]:spec.off[b + 1]], np.full(K - sizes[b], fill)])
        for b in range(B)]))
    a = pad(np.asarray(sna, np.float64).reshape(-1), 1.0).requires_grad_()
    b_ = pad(np.asarray(sk, np.float64).reshape(-1), 1.0).requires_grad_()
    if shared:
        ve = torch.as_tensor(np.asarray(ve_rows, np.float64)).requires_grad_()
        ve_in = ve
    else:
        ve = torch.as_tensor(np.asarray(ve_rows, np.float64)).requires_grad_()
        rows = [torch.cat([ve[spec.off[c]:spec.off[c + 1]],
                           torch.zeros(K - sizes[c], ve.shape[1], dtype=torch.float64)])
                for c in range(B)]
        ve_in = torch.stack(rows)
    kk, ka = fit_ref.knot_table(w.knot_times, w.dt, w.n_steps)
    mw = np.zeros((B, K))
    mw[:, w.sites[w.margin_site]] = w.margin_weight * w.margin_sign / (B * w.n_steps)
    mse = 1.0 / (B * len(w.sites) * w.n_steps)
    loss = fit_ref.fit_loss(geom, a, b_, ve_in, kk, ka, HH(celsius=6.3).array(), w.dt,
                            w.n_steps, w.sites, w.targets.astype(np.float64), mse,
                            torch.as_tensor(mw), w.theta, w.tau)
    ga, gb, gv = torch.autograd.grad(loss, [a, b_, ve])
    unpad = lambda g: np.concatenate([g[c, :sizes[c]].numpy() for c in range(B)])
    return loss.item(), unpad(ga), unpad(gb), gv.numpy()


def _gpu(fit, sna, sk, ve, dtype):
    dev = "cuda"
    p = [torch.tensor(np.asarray(x).reshape(-1), dtype=dtype, device=dev,
                      requires_grad=True) for x in (sna, sk)]
    v = torch.tensor(np.asarray(ve), dtype=dtype, device=dev, requires_grad=True)
    loss = fit.loss(p[0], p[1], v)
    loss.backward()
    return (loss.item(), p[0].grad.double().cpu().numpy(), p[1].grad.double().cpu().numpy(),
            v.grad.double().cpu().numpy())


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b))


@pytest.mark.parametrize("rmax,ckpt,waves", [(4, 70, 1), (8, 300, 1), (4, 1, 1), (2, 70, 2),
                                             (2, 1, 2), (1, 70, 4)])
def test_fp64_adjoint_equals_autograd_restatement(rmax, ckpt, waves):
    w = fit_cables(n_cells=5, n_comp=31, n_steps=300, dt=0.01, n_knots=9, seed=3,
                   ve_gain=20.0)
    print(f"shared table, rmax={rmax}, ckpt_every={ckpt}, waves={waves}", flush=True)
    fit = _fit(w, torch.float64, ckpt_every=ckpt, rmax=rmax, waves=waves)
    got = _gpu(fit, w.sna, w.sk, w.ve, torch.float64)
    want = _ref(w, w.spec, w.sna, w.sk, w.ve, True)
    assert abs(got[0] - want[0]) <= 1e-10 * abs(want[0])
    for g, r in zip(got[1:], want[1:]):
        assert _rel(g, r) <= 1e-9, _rel(g, r)


@pytest.mark.parametrize("waves", [1, 2])
def test_fp64_adjoint_ragged_cells_per_compartment_table(waves):
    """Ragged cells; with two wavefronts per cell the sizes straddle the
    one-row (<= 128) and two-row (129..256) forms of the cross-wavefront solve
    and the interface rows of the 201-compartment cells."""
    K = 45 if waves == 1 else 201
    base = fit_cables(n_cells=6, n_comp=K, n_steps=240, dt=0.01, n_knots=7, seed=8,
                      ve_gain=20.0 if waves == 1 else 3.0)
    sizes = (np.array([45, 12, 33, 20, 7, 45]) if waves == 1
             else np.array([201, 128, 129, 64, 65, 200]))
    off = np.zeros(7, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    total = int(off[-1])
    pick = np.concatenate([np.arange(c * K, c * K + k) for c, k in enumerate(sizes)])
    spec = CableSpec(off=off, diam=base.spec.diam[pick], dx=base.spec.dx[pick],
                     rhoa=base.spec.rhoa[pick], cm=base.spec.cm[pick], v_init=-65.0)
    sna = base.sna.reshape(-1)[pick]
    sk = base.sk.reshape(-1)[pick]
    rs = np.random.default_rng(2)
    ve = np.concatenate([base.ve[:k] for k in sizes]) * (1 + 0.3 * rs.random((total, 1)))
    w = base
    w.sites = np.array([0, 3, 6])
    w.margin_site = 2
    w.targets = base.targets[:, :, :3].copy()
    print(f"ragged cells, per-compartment table, waves={waves}", flush=True)
    fit = _fit(w, torch.float64, ckpt_every=50, shared=False, spec=spec,
               rmax=4 if waves == 1 else 2, waves=waves)
    got = _gpu(fit, sna, sk, ve, torch.float64)
    want = _ref(w, spec, sna, sk, ve, False)
    assert abs(got[0] - want[0]) <= 1e-10 * abs(want[0])
    for g, r in zip(got[1:], want[1:]):
        assert _rel(g, r) <= 1e-9, _rel(g, r)


def test_checkpoint_interval_does_not_change_the_gradient():
    w = fit_cables(n_cells=4, n_comp=31, n_steps=300, dt=0.01, n_knots=9, seed=6,
                   ve_gain=20.0)
    a = _gpu(_fit(w, torch.float64, ckpt_every=300), w.sna, w.sk, w.ve, torch.float64)
    b = _gpu(_fit(w, torch.float64, ckpt_every=37), w.sna, w.sk, w.ve, torch.float64)
    # identical trajectories; only the knot-gradient partial sums are grouped
    # differently across segments
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert _rel(x, y) <= 1e-12


@pytest.fixture(scope="module")
def full_workload():
    return fit_cables()


def test_full_size_fp64_adjoint_vs_central_differences(full_workload):
    """Central differences D(h) along 32 seeded unit directions.

    On this spiking loss D(1e-3) itself carries an O(h^2) truncation error of
    up to ~8e-2 relative (the error falls 4x per halving of h), so the test
    checks (1) that signature -- |D(h) - g.u| / |D(h/2) - g.u| close to 4,
    which an error in g.u would break -- and (2) agreement within 1e-4 at
    h = 1e-5, where truncation is below 1e-5 and float64 round-off of the
    10,000-step loss (~1e-8 absolute) is the floor: the relative error is
    taken against max(|D|, rms of D over the directions)."""
    w = full_workload
    fit = _fit(w, torch.float64)
    dev = "cuda"
    p = [torch.tensor(np.asarray(x).reshape(-1), dtype=torch.float64, device=dev)
         for x in (w.sna, w.sk)] + [torch.tensor(w.ve, dtype=torch.float64, device=dev)]
    loss = fit.forward(*p)
    g = fit.backward()
    rs = np.random.default_rng(2024)
    rows = []
    for i in range(32):
        u = [torch.tensor(rs.normal(size=x.shape), dtype=torch.float64, device=dev)
             for x in p]
        nrm = torch.sqrt(sum((x * x).sum() for x in u))
        u = [x / nrm for x in u]
        an = sum((gi * d).sum() for gi, d in zip(g, u)).item()
        fd = {}
        for h in (1e-3, 5e-4, 1e-5):
            lp = fit.forward(*[x + h * d for x, d in zip(p, u)]).item()
            lm = fit.forward(*[x - h * d for x, d in zip(p, u)]).item()
            fd[h] = (lp - lm) / (2 * h)
        rows.append((an, fd))
        print(f"direction {i}: adjoint={an:.9e} " + " ".join(
            f"D({h:.0e})={v:.9e}" for h, v in fd.items()), flush=True)
    rms = float(np.sqrt(np.mean([r[1][1e-5] ** 2 for r in rows])))
    worst_fine, ratios = 0.0, []
    for an, fd in rows:
        worst_fine = max(worst_fine, abs(fd[1e-5] - an) / max(abs(fd[1e-5]), rms))
        e1, e2 = abs(fd[1e-3] - an), abs(fd[5e-4] - an)
        if e1 > 1e-3 * max(abs(an), rms):
            ratios.append(e1 / e2)
    print(f"loss={loss.item():.9e}; worst relative error at h=1e-5: {worst_fine:.2e}; "
          f"error ratios D(1e-3)/D(5e-4): {np.round(ratios, 2).tolist()}", flush=True)
    assert worst_fine <= 1e-4
    assert len(ratios) >= 8 and all(3.0 <= r <= 5.5 for r in ratios)


def test_full_size_fp32_adjoint_vs_central_differences(full_workload):
    """The benchmarked float32 adjoint (four wavefronts per cell) against
    central differences of the float64 forward at h = 1e-5 (which agree with
    the float64 adjoint within 1e-4, test above), one direction per cell:
    32 seeded cells, each a random unit direction in that cell's gNa/gK
    scales.  The float32 error of a cell's gradient is that of its float32
    trajectory over 10,000 steps; a cell near a spike threshold is chaotic in
    float32 (whether a spike appears depends on the last bits), so the bars
    are on the distribution, as in the per-cell comparison below: median
    <= 1e-3 and at least 29 of the 32 within 1e-2."""
    w = full_workload
    fit64 = _fit(w, torch.float64)
    fit32 = _fit(w, torch.float32, rmax=1, waves=4)
    dev = "cuda"
    mk = lambda dt: [torch.tensor(np.asarray(x).reshape(-1), dtype=dt, device=dev)
                     for x in (w.sna, w.sk)] + [torch.tensor(w.ve, dtype=dt, device=dev)]
    p64, p32 = mk(torch.float64), mk(torch.float32)
    fit32.forward(*p32)
    g32 = [x.double() for x in fit32.backward()]
    B = w.spec.n_cells
    K = p64[0].numel() // B
    rs = np.random.default_rng(77)
    cells = rs.choice(B, 32, replace=False)
    h = 1e-5
    err = []
    for c in cells:
        u = [torch.zeros_like(x) for x in p64]
        d = rs.normal(size=2 * K)
        d /= np.linalg.norm(d)
        u[0][c * K:(c + 1) * K] = torch.as_tensor(d[:K], device=dev)
        u[1][c * K:(c + 1) * K] = torch.as_tensor(d[K:], device=dev)
        an = sum((gi * di).sum() for gi, di in zip(g32, u)).item()
        lp = fit64.forward(*[x + h * di for x, di in zip(p64, u)]).item()
        lm = fit64.forward(*[x - h * di for x, di in zip(p64, u)]).item()
        D = (lp - lm) / (2 * h)
        err.append(abs(an - D) / abs(D))
        print(f"cell {c}: float32 adjoint={an:.7e} D(1e-5)={D:.9e} rel {err[-1]:.2e}",
              flush=True)
    err = np.asarray(err)
    print(f"float32 adjoint vs float64 central differences over 32 cells: median "
          f"{np.median(err):.2e}, 90th percentile {np.quantile(err, 0.9):.2e}, "
          f"worst {err.max():.2e}", flush=True)
    assert np.median(err) <= 1e-3
    assert (err <= 1e-2).sum() >= 29


@pytest.mark.parametrize("rmax,waves", [(4, 1), (2, 2), (1, 4)])
def test_fp64_adjoint_201_compartments_equals_restatement(rmax, waves):
    """The layouts of the 201-compartment workload: 64 lanes x 4 rows, two
    wavefronts x 2 rows, four wavefronts x 1 row."""
    w = fit_cables(n_cells=2, n_comp=201, n_steps=300, dt=0.01, n_knots=9, seed=3,
                   ve_gain=3.0)
    print(f"201 compartments, rmax={rmax}, waves={waves}", flush=True)
    got = _gpu(_fit(w, torch.float64, ckpt_every=70, rmax=rmax, waves=waves), w.sna, w.sk,
               w.ve, torch.float64)
    want = _ref(w, w.spec, w.sna, w.sk, w.ve, True)
    assert abs(got[0] - want[0]) <= 1e-10 * abs(want[0])
    for g, r in zip(got[1:], want[1:]):
        assert _rel(g, r) <= 1e-9, _rel(g, r)


def test_fp32_adjoint_small_against_fp64():
    w = fit_cables(n_cells=5, n_comp=31, n_steps=300, dt=0.01, n_knots=9, seed=3,
                   ve_gain=20.0)
    g64 = _gpu(_fit(w, torch.float64, ckpt_every=70), w.sna, w.sk, w.ve, torch.float64)
    g32 = _gpu(_fit(w, torch.float32, ckpt_every=70), w.sna, w.sk, w.ve, torch.float32)
    assert abs(g32[0] - g64[0]) <= 1e-4 * abs(g64[0])
    for a, b in zip(g32[1:], g64[1:]):
        assert _rel(a, b) <= 2e-3, _rel(a, b)


@pytest.mark.parametrize("rmax,waves", [(4, 1), (1, 4)])
def test_full_size_fp32_against_fp64(full_workload, rmax, waves):
    """Per-cell comparison: over 10,000 steps a cell sitting on a spike
    threshold has a gradient orders of magnitude above the others (its spike
    appears or vanishes with the last bits of the trajectory), so the bar is
    on the distribution of per-cell relative errors, not on the global norm."""
    w = full_workload
    g64 = _gpu(_fit(w, torch.float64), w.sna, w.sk, w.ve, torch.float64)
    g32 = _gpu(_fit(w, torch.float32, rmax=rmax, waves=waves), w.sna, w.sk, w.ve,
               torch.float32)
    rl = abs(g32[0] - g64[0]) / abs(g64[0])
    B = w.spec.n_cells
    out = []
    for a, b in zip(g32[1:3], g64[1:3]):
        a, b = a.reshape(B, -1), b.reshape(B, -1)
        err = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
        out.append(np.quantile(err, [0.5, 0.9]))
    print(f"fp32 loss rel {rl:.2e}; per-cell rel error quantiles (50%, 90%) "
          f"sNa {out[0]}, sK {out[1]}; Ve rel L2 {_rel(g32[3], g64[3]):.3e}", flush=True)
    assert rl <= 1e-3
    for q50, q90 in out:
        assert q50 <= 1e-3 and q90 <= 1e-2
