"""The launch-parameter autotuner picks a valid plan, and the tuned
populations give the oracle's voltages (the plan changes the partition of
the solves, not the physics)."""

import numpy as np
import pytest
import torch

from dendra_amd.model import HH

pytestmark = pytest.mark.gpu


def test_tuned_cables_match_oracle(orc):
    from dendra_amd.autotune import CABLE_WAVES, tune_cables
    from dendra_amd.workloads import extracellular_wave, uneven_cables

    spec = uneven_cables(96, seed=4, kmin=21, kmax=400)
    dt, n = 0.005, 400
    wave = extracellular_wave(n, dt)
    t, pop = tune_cables(spec, dt, HH(celsius=6.3), torch.float64, wave=wave, steps=40)
    assert t.waves_per_block in CABLE_WAVES and t.max_lanes * t.rmax >= 400
    assert t.candidates >= 3
    pop.reset()
    pop.run(n, wave=wave, waves_per_block=t.waves_per_block)
    got = pop.state()["v"]
    want, _ = orc.cable_run(spec, HH(celsius=6.3).array(), dt, n, ve_wave=wave)
    assert np.abs(got - want["v"]).max() < 1e-6


def test_tuned_trees_match_oracle(orc):
    from dendra_amd.autotune import TREE_BLOCKS, tune_trees
    from dendra_amd.population import TreePopulation
    from dendra_amd.workloads import extracellular_wave, random_trees

    spec = random_trees(24, seed=5, tmin=100, tmax=900)
    dt, n = 0.005, 300
    wave = extracellular_wave(n, dt)
    pop = TreePopulation(spec, dt, HH(celsius=6.3), torch.float64)
    t = tune_trees(pop, wave=wave, steps=20)
    assert t.block in TREE_BLOCKS and pop.block == t.block
    pop.reset()
    pop.run(n, wave=wave)
    want, _ = orc.tree_run(spec, HH(celsius=6.3).array(), dt, n, ve_wave=wave)
    assert np.abs(pop.state()["v"] - want["v"]).max() < 1e-6
