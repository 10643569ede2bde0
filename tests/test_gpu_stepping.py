"""The per-step integrator interface (dendra_amd.stepping) against the
binary64 oracle stepped one step at a time with the same arbitrary
(non-separable) extracellular potential and injected current per step."""

import dataclasses

import numpy as np
import pytest
import torch

import bars

pytestmark = pytest.mark.gpu


def _drives(rs, n_steps, total):
    """Per-step fields that are not one spatial profile times one waveform:
    a random smooth potential whose shape changes every step, and current
    pulses at random compartments."""
    base = np.cumsum(rs.standard_normal((n_steps, total)), axis=1) * 0.4
    ve = base - base.mean(axis=1, keepdims=True)
    intra = np.zeros((n_steps, total))
    hits = rs.random((n_steps, total)) < 0.004
    intra[hits] = rs.uniform(2e-6, 1e-5, hits.sum())
    return ve, intra


def _oracle_steps(orc, run, spec, hh, dt, ve, intra):
    st = None
    out = []
    ones = np.ones(1)
    n = spec.n_cells if hasattr(spec, "n_cells") else spec.n_trees
    for s in range(len(ve)):
        sp = dataclasses.replace(spec, lead=ve[s], ve_amp=np.ones(n), iamp=intra[s])
        st, _ = run(sp, hh.array(), dt, 1, ve_wave=ones, i_wave=ones, state=st)
        out.append(st["v"].copy())
    return np.array(out)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_cable_stepper_matches_oracle_with_arbitrary_drives(orc, dtype):
    from dendra_amd.model import HH, CableSpec
    from dendra_amd.stepping import CableStepper

    rs = np.random.default_rng(4)
    sizes = np.array([1, 5, 64, 65, 130, 301])
    off = np.concatenate([[0], np.cumsum(sizes)])
    total = int(off[-1])
    spec = CableSpec(off, rs.uniform(1.0, 6.0, total), np.full(total, 10.0),
                     np.full(total, 100.0), np.ones(total), v_init=-65.0)
    hh = HH(celsius=16.3)
    dt, n_steps = 0.01, 300
    ve, intra = _drives(rs, n_steps, total)
    want = _oracle_steps(orc, orc.cable_run, spec, hh, dt, ve, intra)
    stp = CableStepper(spec, hh, dtype=dtype)
    v = torch.full((total,), -65.0, dtype=dtype, device="cuda")
    got = []
    for s in range(n_steps):
        v_in = v.clone()
        v_new, imem = stp._step(v, dt, 16.3, torch.as_tensor(ve[s], device="cuda"),
                                torch.as_tensor(intra[s], device="cuda"))
        assert imem is None and torch.equal(v, v_in)  # the input is not modified
        v = v_new
        got.append(v.double().cpu().numpy())
    got = np.array(got)
    err = np.abs(got - want).max()
    print(f"cable stepper {dtype}: max |dV| {err:.3g} mV, range {np.ptp(want):.1f} mV")
    assert np.ptp(want) > 20.0
    # observed on MI355X: 7.4e-13 mV (float64), 1.9e-4 mV (float32)
    assert err <= (bars.F64 if dtype == torch.float64 else bars.F32_NORTH_STAR), err


def test_cable_stepper_contract():
    """v is not modified, v_new is new and keeps v's (B, K) shape, the gates
    are the stepper's state; a solver override is refused."""
    from dendra_amd.model import HH, CableSpec
    from dendra_amd.stepping import CableStepper

    B, K = 4, 33
    off = np.arange(B + 1) * K
    spec = CableSpec(off, np.full(B * K, 2.0), np.full(B * K, 10.0), np.full(B * K, 100.0),
                     np.ones(B * K), v_init=-65.0)
    stp = CableStepper(spec, HH(), dtype=torch.float64)
    v = torch.full((B, K), -60.0, dtype=torch.float64, device="cuda")
    v0 = v.clone()
    m0 = stp.gates()["m"].copy()
    v1, _ = stp._step(v, 0.025, 6.3)
    assert v1.shape == (B, K) and v1.data_ptr() != v.data_ptr() and torch.equal(v, v0)
    assert not np.array_equal(stp.gates()["m"], m0)
    with pytest.raises(ValueError, match="solver"):
        stp._step(v, 0.025, 6.3, solver=lambda *a: a)
    with pytest.raises(ValueError, match="values"):
        stp._step(v[:, :5], 0.025, 6.3)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_tree_stepper_matches_oracle_with_arbitrary_drives(orc, dtype):
    from dendra_amd.model import HH
    from dendra_amd.stepping import TreeStepper
    from dendra_amd.workloads import random_trees

    rs = np.random.default_rng(5)
    spec = random_trees(3, seed=21, tmin=60, tmax=400)
    spec = dataclasses.replace(spec, lead=None, ve_amp=None)
    hh = HH(celsius=6.3)
    dt, n_steps = 0.01, 300
    ve, intra = _drives(rs, n_steps, spec.total)
    want = _oracle_steps(orc, orc.tree_run, spec, hh, dt, ve, intra)
    stp = TreeStepper(spec, hh, dtype=dtype)
    v = torch.full((spec.total,), -65.0, dtype=dtype, device="cuda")
    got = []
    for s in range(n_steps):
        v, _ = stp._step(v, dt, 6.3, torch.as_tensor(ve[s], device="cuda"),
                         torch.as_tensor(intra[s], device="cuda"))
        got.append(v.double().cpu().numpy())
    err = np.abs(np.array(got) - want).max()
    print(f"tree stepper {dtype}: max |dV| {err:.3g} mV, range {np.ptp(want):.1f} mV")
    assert np.ptp(want) > 20.0
    # observed on MI355X: 3.5e-14 mV (float64), 4.7e-5 mV (float32)
    assert err <= (bars.F64 if dtype == torch.float64 else bars.F32_NORTH_STAR), err
