"""GPU ExtCell cables (dendra_amd.extcell) against the binary64 oracle."""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from test_extcell_ref import init_state, np_ws, spec

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import oracle  # noqa: E402

pytestmark = pytest.mark.gpu


def drive(steps):
    wave = np.zeros(steps)
    wave[40:240] = 1.0
    iw = np.zeros(steps)
    iw[300:400] = 1.0
    return wave, iw


@pytest.mark.parametrize("B,K", [(37, 23), (1, 1), (70, 2), (5, 101)])
def test_float64_matches_oracle(B, K):
    from dendra_amd.extcell import ExtCellCables
    from dendra_amd.model import HH

    rng = np.random.default_rng(B * 1000 + K)
    s = spec(rng, B, K)
    dt, steps = 0.005, 500
    wave, iw = drive(steps)
    hh = HH(celsius=6.3)
    cells = ExtCellCables(s, dt, hh, torch.float64, "cuda")
    rec = cells.run(steps, wave=wave, i_wave=iw, record_every=5).cpu().numpy()
    ws = np_ws(cells.workspace)
    st, orec = oracle.extcell_run(ws, hh.array(), dt, steps, init_state(s), lead=s.lead,
                                  amp=s.ve_amp, wave=wave, iamp=s.iamp, i_wave=iw,
                                  record_every=5)
    got = cells.state()
    assert np.abs(rec.reshape(rec.shape[0], -1) - orec).max() <= 1e-8
    assert np.abs(got["vc"] - st["vc"]).max() <= 1e-8
    for g in ("m", "h", "n"):
        assert np.abs(got[g] - st[g]).max() <= 1e-10
    if K >= 23:
        assert orec.max() > 0.0  # spikes are covered


def test_float32_spike_times_and_subthreshold_error():
    from dendra_amd.extcell import ExtCellCables
    from dendra_amd.model import HH

    rng = np.random.default_rng(7)
    B, K, dt, steps = 64, 31, 0.005, 600
    s = spec(rng, B, K)
    wave, iw = drive(steps)
    hh = HH(celsius=6.3)
    cells = ExtCellCables(s, dt, hh, torch.float32, "cuda")
    rec = cells.run(steps, wave=wave, i_wave=iw, record_every=1).double().cpu().numpy()
    _, orec = oracle.extcell_run(np_ws(cells.workspace), hh.array(), dt, steps,
                                 init_state(s), lead=s.lead, amp=s.ve_amp, wave=wave,
                                 iamp=s.iamp, i_wave=iw, record_every=1)
    rec = rec.reshape(steps, -1)
    # before anything fires the float32 run stays within 1e-3 mV
    quiet = np.flatnonzero((orec > -40).any(axis=1))
    first = quiet[0] if len(quiet) else steps
    assert np.abs(rec[:first] - orec[:first]).max() <= 1e-3
    # upward 0 mV crossings agree within one step
    up = lambda r: [np.flatnonzero((r[:-1, j] < 0) & (r[1:, j] >= 0)) for j in range(r.shape[1])]
    for a, b in zip(up(rec), up(orec)):
        assert len(a) == len(b) and (len(a) == 0 or np.abs(a - b).max() <= 1)
    assert sum(len(a) for a in up(orec)) > 0


def test_reset_state_round_trip_and_steps_continue():
    from dendra_amd.extcell import ExtCellCables

    rng = np.random.default_rng(2)
    s = spec(rng, 9, 5, stim=False)
    cells = ExtCellCables(s, 0.01, dtype=torch.float64, device="cuda")
    st0 = cells.state()
    st0["vc"][..., 1] = rng.uniform(-1, 1, (9, 5))
    cells.set_state(st0["vc"], st0["m"], st0["h"], st0["n"])
    assert np.array_equal(cells.state()["vc"], st0["vc"])
    cells.run(10)
    a = cells.state()
    cells.set_state(st0["vc"], st0["m"], st0["h"], st0["n"])
    cells.t_step = 0
    cells.run(4)
    cells.run(6)
    assert np.array_equal(cells.state()["vc"], a["vc"])


@pytest.mark.parametrize("kind,K", [("random", 30), ("binary", 31), ("path", 9), ("single", 1)])
def test_tree_float64_matches_dense_oracle(kind, K):
    from dendra_amd.extcell import ExtCellTrees
    from dendra_amd.model import HH
    from test_extcell_ref import extcell_tree_ref, random_parents, tree_spec

    rng = np.random.default_rng(50 + K)
    par = {"random": lambda: random_parents(K, rng),
           "binary": lambda: [-1] + [(i - 1) // 2 for i in range(1, K)],
           "path": lambda: list(range(-1, K - 1)), "single": lambda: [-1]}[kind]()
    B, dt, steps = 3, 0.005, 150
    ts, cs = tree_spec(rng, B, par)
    wave, iw = drive(steps)
    hh = HH(celsius=6.3)
    trees = ExtCellTrees(ts, dt, hh, torch.float64, "cuda")
    rec = trees.run(steps, wave=wave, i_wave=iw, record_every=5).cpu().numpy()
    st, orec = extcell_tree_ref.run(np_ws(trees.workspace), par, hh.array(), dt, steps,
                                    init_state(cs), lead=ts.lead, amp=ts.ve_amp, wave=wave,
                                    iamp=ts.iamp, i_wave=iw, record_every=5)
    got = trees.state()
    assert np.abs(rec - orec).max() <= 1e-8
    assert np.abs(got["vc"] - st["vc"]).max() <= 1e-8
    for g in ("m", "h", "n"):
        assert np.abs(got[g] - st[g]).max() <= 1e-10


def test_path_tree_equals_cable_on_the_gpu():
    from dendra_amd.extcell import ExtCellCables, ExtCellTrees
    from dendra_amd.model import HH
    from test_extcell_ref import tree_spec

    rng = np.random.default_rng(60)
    B, K, dt, steps = 40, 45, 0.005, 500
    ts, cs = tree_spec(rng, B, list(range(-1, K - 1)))
    wave, iw = drive(steps)
    hh = HH(celsius=6.3)
    t = ExtCellTrees(ts, dt, hh, torch.float64, "cuda").run(steps, wave=wave, i_wave=iw,
                                                            record_every=10)
    c = ExtCellCables(cs, dt, hh, torch.float64, "cuda").run(steps, wave=wave, i_wave=iw,
                                                             record_every=10)
    assert torch.max(torch.abs(t - c)).item() <= 1e-8
    assert c.max().item() > 0.0
