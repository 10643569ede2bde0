"""GPU checks of the persistent kernels' launch contracts and of the
node-slot tree kernel (layout 0, trees of more than 512 sections)."""

import numpy as np
import pytest
import torch

import bars

pytestmark = pytest.mark.gpu


def _caterpillar(n_spine: int, n_trees: int = 2):
    """Trees whose spine nodes each carry one leaf: every node is its own
    section, so a tree has about 2 * n_spine sections (layout 0 above 512)."""
    from dendra_amd.model import TreeSpec, point_source_lead

    parents, diams, leads = [], [], []
    for t in range(n_trees):
        par = [-1]
        pos = [(0.0, 0.0)]
        d = [3.0]
        prev = 0
        for i in range(1, n_spine):
            sid = len(par)
            par.append(prev)
            pos.append((10.0 * i, 0.0))
            d.append(3.0)
            par.append(sid)
            pos.append((10.0 * i, 10.0 + 2.0 * t))
            d.append(1.2)
            prev = sid
        pos = np.asarray(pos)
        parents.append(np.asarray(par, dtype=np.int64))
        diams.append(np.asarray(d))
        leads.append(point_source_lead(pos[:, 0], pos[:, 1], 0.0, 0.3 * n_spine * 10.0,
                                       0.0, 400.0 + 300.0 * t))
    sizes = [len(p) for p in parents]
    off = np.zeros(n_trees + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    total = int(off[-1])
    return TreeSpec(off, np.concatenate(parents), np.concatenate(diams),
                    np.full(total, 10.0), np.full(total, 100.0), np.ones(total),
                    lead=np.concatenate(leads), ve_amp=np.array([-1.5, 1.0])[:n_trees],
                    v_init=-65.0)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_node_slot_tree_kernel_matches_oracle(orc, dtype):
    from dendra_amd.model import HH
    from dendra_amd.population import TreePopulation
    from dendra_amd.workloads import extracellular_wave

    spec = _caterpillar(420)
    hh = HH(celsius=6.3)
    n_steps, dt = 2000, 0.005
    wave = extracellular_wave(n_steps, dt)
    pop = TreePopulation(spec, dt, hh, dtype=dtype)
    assert (pop.n_sec > 512).all()
    assert [b.layout for b in pop.buckets] == [0]
    _, want = orc.tree_run(spec, hh.array(), dt, n_steps, ve_wave=wave, record_every=25)
    rec = pop.unpermute(pop.run(n_steps, wave=wave, record_every=25, steps_per_launch=700))
    err = np.abs(rec - want).max()
    print(f"node-slot trees {dtype}: max |dV| {err:.3g} mV, range {np.ptp(want):.1f} mV")
    assert np.ptp(want) > 10.0
    assert err <= (bars.F64 if dtype == torch.float64 else bars.F32_NORTH_STAR), err


def test_cable_run_reports_too_many_rows_in_status():
    """A launch bounded at 8 rows per lane over tasks of 13 rows: the kernel
    leaves those cables untouched, sets bit 1 of the status word, and the
    population's check raises."""
    from dendra_amd import _lib
    from dendra_amd.model import HH, CableSpec
    from dendra_amd.population import CablePopulation

    K = 832
    spec = CableSpec(np.array([0, K]), np.full(K, 4.0), np.full(K, 10.0),
                     np.full(K, 100.0), np.ones(K), v_init=-65.0)
    pop = CablePopulation(spec, 0.005, HH(), dtype=torch.float32)
    assert int(pop.task_nr.max()) == 13
    v0 = pop.v.clone()
    pop.v += 1.0  # off rest: an advanced cable would move
    v1 = pop.v.clone()
    lib = _lib.load()
    st = lib.dn_hh_cable_run_f32(
        len(pop.task_nr), _lib.ptr(pop.task_cell0), _lib.ptr(pop.task_ncell),
        _lib.ptr(pop.task_nr), _lib.ptr(pop.task_p), _lib.ptr(pop.off), _lib.ptr(pop.v),
        _lib.ptr(pop.m), _lib.ptr(pop.h), _lib.ptr(pop.n), _lib.ptr(pop.gL),
        _lib.ptr(pop.gRp), _lib.ptr(pop.scale), None, None, None, None, 0, 10, 0.005,
        pop.hh_arr.ctypes.data, None, 1, pop.total, 4, 8, _lib.ptr(pop.status),
        torch.cuda.current_stream().cuda_stream)
    assert st == 0
    torch.cuda.synchronize()
    assert int(pop.status.item()) == 1
    assert torch.equal(pop.v, v1) and not torch.equal(pop.v, v0)
    with pytest.raises(RuntimeError, match="rows per lane"):
        pop.check()
    assert int(pop.status.item()) == 0
    pop.run(10)
    pop.check()
    assert not torch.equal(pop.v, v1)


def test_tree_run_reports_sections_longer_than_the_kernel_takes(monkeypatch):
    """A host plan that splits chains at 16 nodes against a kernel that takes
    dn_hh_tree_sec_max() (13): the section kernel leaves the group
    unadvanced, sets bit 2 of the status word, and the population's check
    raises; with the plan the kernel takes, the same tree advances."""
    from dendra_amd import _lib, population, topology
    from dendra_amd.model import HH, TreeSpec

    cap = _lib.load().dn_hh_tree_sec_max()
    assert cap == topology.SEC_MAX and cap < 16
    K = 40  # one unbranched chain: sections of 16, 16, 8 under the wider plan
    spec = TreeSpec(np.array([0, K]), np.arange(-1, K - 1), np.full(K, 2.0),
                    np.full(K, 10.0), np.full(K, 100.0), np.ones(K), v_init=-60.0)
    real = population.section_plan
    monkeypatch.setattr(population, "section_plan",
                        lambda parent, sec_max=None: real(parent, sec_max=16))
    pop = population.TreePopulation(spec, 0.005, HH(), dtype=torch.float32)
    assert int(pop.sec["len"].max()) == 16
    v0 = pop.v.clone()
    pop.run(20)
    torch.cuda.synchronize()
    assert torch.equal(pop.v, v0)
    with pytest.raises(RuntimeError, match="section longer"):
        pop.check()
    monkeypatch.setattr(population, "section_plan", real)
    ok = population.TreePopulation(spec, 0.005, HH(), dtype=torch.float32)
    assert int(ok.sec["len"].max()) <= cap
    ok.run(20)
    ok.check()
    assert not torch.equal(ok.v, v0)
