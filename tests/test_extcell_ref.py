"""Pins of the ExtCell (double-layer extracellular) workspace and oracle:
Kirchhoff's invariant of the prepared block matrix, a dense 3K x 3K restatement
of one step assembled from the reference's equations, and the limit of
shorted extracellular layers, which must reproduce the plain cable oracle
(a different formulation) including its extracellular drive."""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import oracle  # noqa: E402

from dendra_amd.extcell import (ExtCellSpec, extcell_workspace, layered_edge_conductance,
                                prepare_workspace)
from dendra_amd.model import HH, CableSpec, hh_steady_state


def spec(rng, B, K, *, shorted=False, stim=True):
    diam = rng.uniform(2.0, 8.0, (B, K))
    dx = rng.uniform(20.0, 60.0, (B, K))
    rhoa = np.full((B, K), 100.0)
    cm = rng.uniform(0.8, 1.2, (B, K))
    if shorted:
        xraxial = np.full((B, K, 2), 1e9)
        xc = np.zeros((B, K, 2))
        xg = np.full((B, K, 2), 1e9)
    else:
        xraxial = rng.uniform(1e3, 5e4, (B, K, 2))
        xc = rng.uniform(0.0, 0.5, (B, K, 2))
        xg = rng.uniform(1e-2, 1.0, (B, K, 2))
    lead = rng.uniform(50.0, 400.0, (B, K)) if stim else None
    amp = rng.uniform(-0.5, -0.2, B) if stim else None
    iamp = np.zeros((B, K)) if stim else None
    if stim:
        iamp[:, K // 2] = rng.uniform(5e-7, 1e-6, B)
    return ExtCellSpec(diam, dx, rhoa, cm, xraxial, xc, xg, lead=lead, ve_amp=amp, iamp=iamp)


def init_state(s):
    B, K = s.shape
    m0, h0, n0 = hh_steady_state(s.v_init)
    vc = np.zeros((B, K, 3))
    vc[..., 0] = s.v_init
    return {"vc": vc, "m": np.full((B, K), m0), "h": np.full((B, K), h0),
            "n": np.full((B, K), n0)}


def np_ws(ws):
    return {k: v.numpy() for k, v in ws.items()}


def test_block_matrix_obeys_kirchhoff_and_symmetry():
    """A uniform potential drives no internal current: A 1 = (0, 0, xc_dt1 +
    radial_g1) per compartment; main blocks symmetric, lower == upper."""
    rng = np.random.default_rng(1)
    s = spec(rng, 3, 7)
    ws = extcell_workspace(s, 0.005)
    main, off = ws["maind"].numpy(), ws["lower"].numpy()
    assert np.allclose(main, np.swapaxes(main, -1, -2)) and np.array_equal(off, ws["upper"].numpy())
    row = main.sum(-1)
    row[:, 1:] += off
    row[:, :-1] += off
    want = np.zeros_like(row)
    want[..., 2] = ws["xc_dt"].numpy()[..., 1] + ws["xg"].numpy()[..., 1]
    assert np.abs(row - want).max() <= 1e-12 * np.abs(main).max()
    # K = 1: sealed ends on both sides, empty couplings
    one = prepare_workspace(0.01, cm=torch.ones(2, 1), area=torch.full((2, 1), 1e-5),
                            intracellular_edge_conductance=torch.zeros(2, 0),
                            extracellular_edge_conductance=torch.zeros(2, 0, 2),
                            xc=torch.ones(2, 1, 2), xg=torch.ones(2, 1, 2))
    assert one["lower"].shape == (2, 0, 3) and one["maind"].shape == (2, 1, 3, 3)
    cm_dt = 1e-6 * 1e-5 / 1e-5
    assert one["maind"][0, 0, 0, 0].item() == pytest.approx(cm_dt)
    assert one["maind"][0, 0, 1, 2].item() == pytest.approx(-(cm_dt + 1e-5))


def test_layered_edge_conductance_halves():
    r = torch.tensor([[[2.0, 4.0], [6.0, 8.0]]])
    L = torch.tensor([[10.0, 30.0]])
    g = layered_edge_conductance(r, L)
    seg = r * L.unsqueeze(-1) * 1e-4 * 1e6
    assert torch.allclose(g, 2.0 / (seg[:, :1] + seg[:, 1:]))
    assert g.shape == (1, 1, 2)


def dense_step(ws, st, hh_arr, dt, bath, intra):
    """One _step as a dense 3K x 3K solve per cell (implicit.py:1021-1136)."""
    B, K = st["vc"].shape[:2]
    out = {k: v.copy() for k, v in st.items()}
    for b in range(B):
        A = np.zeros((3 * K, 3 * K))
        R = np.zeros(3 * K)
        V = st["vc"][b]
        for k in range(K):
            v = V[k, 0] - V[k, 1]
            m, h, n, G, J = oracle.hh_advance(v, dt, hh_arr, st["m"][b, k], st["h"][b, k],
                                              st["n"][b, k])
            out["m"][b, k], out["h"][b, k], out["n"][b, k] = m, h, n
            g = G * ws["area"][b, k]
            d = (J - G * v) * ws["area"][b, k] + intra[b, k]
            blk = ws["maind"][b, k] + np.array([[g, -g, 0], [-g, g, 0], [0, 0, 0]])
            A[3 * k:3 * k + 3, 3 * k:3 * k + 3] = blk
            shell = ws["xg"][b, k, 0] * (V[k, 1] - V[k, 2])
            bathf = ws["xg"][b, k, 1] * (bath[b, k] - V[k, 2])
            R[3 * k:3 * k + 3] = [d, -d - shell, shell + bathf]
        for k in range(K - 1):
            lo, up = ws["lower"][b, k], ws["upper"][b, k]
            A[3 * (k + 1):3 * (k + 2), 3 * k:3 * k + 3] = np.diag(lo)
            A[3 * k:3 * k + 3, 3 * (k + 1):3 * (k + 2)] = np.diag(up)
            flux = -lo * (V[k + 1] - V[k])
            R[3 * k:3 * k + 3] += flux
            R[3 * (k + 1):3 * (k + 2)] -= flux
        out["vc"][b] = V + np.linalg.solve(A, R).reshape(K, 3)
    return out


@pytest.mark.parametrize("K", [1, 2, 9])
def test_one_step_equals_dense_restatement(K):
    rng = np.random.default_rng(10 + K)
    s = spec(rng, 3, K)
    ws = np_ws(extcell_workspace(s, 0.01))
    st = init_state(s)
    st["vc"] += rng.uniform(-5, 5, st["vc"].shape)  # off rest, shells charged
    hh_arr = HH(celsius=6.3).array()
    wave = np.array([0.7])
    iw = np.array([1.0])
    got, _ = oracle.extcell_run(ws, hh_arr, 0.01, 1, st, lead=s.lead, amp=s.ve_amp, wave=wave,
                                iamp=s.iamp, i_wave=iw)
    bath = s.lead * s.ve_amp[:, None] * wave[0]
    want = dense_step(ws, st, hh_arr, 0.01, bath, s.iamp * iw[0])
    for k in ("vc", "m", "h", "n"):
        assert np.abs(got[k] - want[k]).max() <= 1e-10 * max(1.0, np.abs(want[k]).max()), k


def test_shorted_layers_reduce_to_the_cable_oracle():
    """xg -> 1e9 S/cm^2 ties both shells to the bath: vi - ve0 follows the
    plain cable with the bath potential as its extracellular drive."""
    rng = np.random.default_rng(3)
    B, K, dt, steps = 4, 15, 0.005, 600
    s = spec(rng, B, K, shorted=True)
    ws = np_ws(extcell_workspace(s, dt))
    hh_arr = HH(celsius=6.3).array()
    wave = np.zeros(steps)
    wave[40:240] = 1.0
    iw = np.zeros(steps)
    iw[300:400] = 1.0
    st, rec = oracle.extcell_run(ws, hh_arr, dt, steps, init_state(s), lead=s.lead,
                                 amp=s.ve_amp, wave=wave, iamp=s.iamp, i_wave=iw,
                                 record_every=1)
    off = np.arange(B + 1) * K
    cab = CableSpec(off, s.diam.ravel(), s.dx.ravel(), s.rhoa.ravel(), s.cm.ravel(),
                    lead=s.lead.ravel(), ve_amp=s.ve_amp, iamp=s.iamp.ravel())
    _, crec = oracle.cable_run(cab, hh_arr, dt, steps, ve_wave=wave, i_wave=iw,
                               record_every=1)
    assert np.abs(rec - crec).max() <= 1e-6
    assert crec.max() > 0.0  # the drive fires the cables: the comparison covers spikes


# -- branched trees ------------------------------------------------------------

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import extcell_tree_ref  # noqa: E402


def tree_spec(rng, B, parent, *, shorted=False, stim=True):
    from dendra_amd.extcell import ExtCellTreeSpec

    K = len(parent)
    c = spec(rng, B, K, shorted=shorted, stim=stim)
    return ExtCellTreeSpec(np.asarray(parent), c.diam, c.dx, c.rhoa, c.cm, c.xraxial, c.xc,
                           c.xg, lead=c.lead, ve_amp=c.ve_amp, iamp=c.iamp), c


def random_parents(K, rng):
    return [-1] + [int(rng.integers(max(0, i - 3), i)) for i in range(1, K)]


def test_path_tree_equals_the_cable_oracle():
    """An unbranched ExtCell tree (dense block solve, axial fluxes along parent
    edges) equals the ExtCell cable (block Thomas, banded couplings)."""
    from dendra_amd.extcell import extcell_tree_workspace

    rng = np.random.default_rng(21)
    B, K, dt, steps = 2, 12, 0.005, 300
    ts, cs = tree_spec(rng, B, list(range(-1, K - 1)))
    wave = np.zeros(steps)
    wave[40:140] = 1.0
    iw = np.zeros(steps)
    iw[150:200] = 1.0
    hh_arr = HH(celsius=6.3).array()
    _, trec = extcell_tree_ref.run(np_ws(extcell_tree_workspace(ts, dt)), ts.parent, hh_arr,
                                   dt, steps, init_state(cs), lead=ts.lead, amp=ts.ve_amp,
                                   wave=wave, iamp=ts.iamp, i_wave=iw, record_every=10)
    _, crec = oracle.extcell_run(np_ws(extcell_workspace(cs, dt)), hh_arr, dt, steps,
                                 init_state(cs), lead=cs.lead, amp=cs.ve_amp, wave=wave,
                                 iamp=cs.iamp, i_wave=iw, record_every=10)
    assert np.abs(trec.reshape(trec.shape[0], -1) - crec).max() <= 1e-9
    assert crec.max() > 0.0


def test_shorted_tree_reduces_to_the_tree_oracle():
    """Shorted layers on a branched tree: vi - ve0 follows the plain tree
    oracle with the bath potential as its extracellular drive."""
    from dendra_amd.extcell import extcell_tree_workspace
    from dendra_amd.model import TreeSpec

    rng = np.random.default_rng(22)
    B, K, dt, steps = 2, 14, 0.005, 400
    par = random_parents(K, rng)
    ts, cs = tree_spec(rng, B, par, shorted=True)
    wave = np.zeros(steps)
    wave[40:200] = 1.0
    iw = np.zeros(steps)
    iw[220:300] = 1.0
    hh_arr = HH(celsius=6.3).array()
    _, trec = extcell_tree_ref.run(np_ws(extcell_tree_workspace(ts, dt)), par, hh_arr, dt,
                                   steps, init_state(cs), lead=ts.lead, amp=ts.ve_amp,
                                   wave=wave, iamp=ts.iamp, i_wave=iw, record_every=1)
    tree = TreeSpec(np.arange(B + 1) * K, np.tile(par, B), ts.diam.ravel(), ts.length.ravel(),
                    ts.ra.ravel(), ts.cm.ravel(), lead=ts.lead.ravel(), ve_amp=ts.ve_amp,
                    iamp=ts.iamp.ravel())
    _, orec = oracle.tree_run(tree, hh_arr, dt, steps, ve_wave=wave, i_wave=iw, record_every=1)
    assert np.abs(trec.reshape(steps, -1) - orec).max() <= 1e-6
    assert orec.max() > 0.0
