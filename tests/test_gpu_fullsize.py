"""Checks at the sizes the benchmark runs.

* the benchmark's launch plan (cables: 13 rows per lane, up to 64 lanes, 4
  wavefronts per workgroup; trees: 512-thread blocks) on the population's
  largest cables and on trees of 3500-4000 nodes, 4000 steps, against the
  binary64 oracle;
* the full 8192-cable + 1024-tree run: finite, the same compartments fire as
  in the oracle on every 64th cell, and without drive every compartment
  stays with the single-compartment oracle trajectory;
* the full network (21,000 cells, 4.2M connections): delivered payloads
  bitwise equal to the oracle's ring replay of the GPU's own spike tape.
"""

import numpy as np
import pytest
import torch

import bars

pytestmark = pytest.mark.gpu

DT, STEPS = 0.005, 4000


@pytest.fixture(scope="module")
def population():
    from dendra_amd.workloads import uneven_population

    return uneven_population()


def _pops(cables, trees, hh):
    from dendra_amd.population import CablePopulation, TreePopulation

    cp = CablePopulation(cables, DT, hh, dtype=torch.float32, rmax=13, max_lanes=64)
    tp = TreePopulation(trees, DT, hh, dtype=torch.float32, block=512)
    return cp, tp


def _fired(rec, theta=0.0):
    return (rec >= theta).any(axis=0)


def test_bench_plan_on_the_largest_cells_matches_oracle(orc, population):
    from dendra_amd.model import HH

    hh = HH(celsius=6.3)
    wave = population.wave
    c, t = population.cables, population.trees
    big_c = np.argsort(-c.sizes, kind="stable")[:12]
    sizes_t = t.sizes
    big_t = np.flatnonzero((sizes_t >= 3500) & (sizes_t <= 4000))[:4]
    assert len(big_t) == 4 and c.sizes[big_c].max() >= 790
    cs, ts = c.subset(big_c), t.subset(big_t)
    cp, tp = _pops(cs, ts, hh)
    rc = cp.unpermute(cp.run(STEPS, wave=wave, record_every=25, waves_per_block=4))
    rt = tp.unpermute(tp.run(STEPS, wave=wave, record_every=25))
    _, wc = orc.cable_run(cs, hh.array(), DT, STEPS, ve_wave=wave, record_every=25)
    _, wt = orc.tree_run(ts, hh.array(), DT, STEPS, ve_wave=wave, record_every=25)
    ec, et = np.abs(rc - wc).max(), np.abs(rt - wt).max()
    # the same cells through the reference's float32 step restated in binary32
    _, fc = orc.cable_run_f32(cs, hh.array(), DT, STEPS, ve_wave=wave, record_every=25)
    _, ft = orc.tree_run_f32(ts, hh.array(), DT, STEPS, ve_wave=wave, record_every=25)
    rc32, rt32 = np.abs(fc - wc).max(), np.abs(ft - wt).max()
    print(f"bench plan: cables max |dV| {ec:.3g} mV (binary32 restatement {rc32:.3g}), "
          f"trees {et:.3g} mV (restatement {rt32:.3g}); fired {int(_fired(wc).sum())} / "
          f"{int(_fired(wt).sum())} compartments")
    assert np.array_equal(_fired(rc), _fired(wc)) and np.array_equal(_fired(rt), _fired(wt))
    # observed on MI355X: cables 1.5e-2 mV (restatement 6.9e-2), trees 2.4e-3 mV
    # (restatement 2.4e-3): the GPU is as close as correctly rounded float32
    assert ec <= max(bars.F32_NORTH_STAR, bars.F32_VS_BINARY32 * rc32)
    assert et <= max(bars.F32_NORTH_STAR, bars.F32_VS_BINARY32 * rt32)


def test_full_population_properties(orc, population):
    from dendra_amd.model import HH

    hh = HH(celsius=6.3)
    c, t = population.cables, population.trees
    cp, tp = _pops(c, t, hh)
    ev = 40
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    rc = cp.run(STEPS, wave=population.wave, record_every=ev)
    rt = tp.run(STEPS, wave=population.wave, record_every=ev, stream=side)
    torch.cuda.current_stream().wait_stream(side)
    fired_c = torch.zeros(c.total, dtype=torch.bool, device="cuda")
    fired_c[torch.as_tensor(cp.perm, device="cuda")] = (rc >= 0).any(0)
    fired_t = torch.zeros(t.total, dtype=torch.bool, device="cuda")
    fired_t[torch.as_tensor(tp.perm, device="cuda")] = (rt >= 0).any(0)
    assert bool(torch.isfinite(rc).all()) and bool(torch.isfinite(rt).all())
    sc, st_ = cp.state(), tp.state()
    del rc, rt
    fired_c, fired_t = fired_c.cpu().numpy(), fired_t.cpu().numpy()
    # every 64th cable and tree through the binary64 oracle
    pick_c, pick_t = np.arange(0, c.n_cells, 64), np.arange(0, t.n_trees, 64)
    sub_c, sub_t = c.subset(pick_c), t.subset(pick_t)
    idx_c = np.concatenate([np.arange(c.off[i], c.off[i + 1]) for i in pick_c])
    idx_t = np.concatenate([np.arange(t.off[i], t.off[i + 1]) for i in pick_t])
    m0, h0, n0 = orc.hh_steady(-65.0)
    errs = []
    for run, sub, got, fired, idx in ((orc.cable_run, sub_c, sc, fired_c, idx_c),
                                      (orc.tree_run, sub_t, st_, fired_t, idx_t)):
        want, rec = run(sub, hh.array(), DT, STEPS, ve_wave=population.wave, record_every=ev)
        # what half a float32 ulp of initial voltage does to the binary64 run
        T = sub.total
        nudged = {"v": np.full(T, -65.0 + 2.0**-18), "m": np.full(T, m0),
                  "h": np.full(T, h0), "n": np.full(T, n0)}
        moved, _ = run(sub, hh.array(), DT, STEPS, ve_wave=population.wave, state=nudged)
        err = np.abs(got["v"][idx] - want["v"]).max()
        cond = np.abs(moved["v"] - want["v"]).max()
        errs.append((err, cond))
        mism = int((fired[idx] != _fired(rec)).sum())
        assert mism == 0, f"{mism} compartments fire differently from the oracle"
        assert err <= max(bars.F32, bars.F32_VS_BINARY32 * cond), (err, cond)
    n_fired = int(fired_c.sum() + fired_t.sum())
    print(f"full run: {n_fired} of {c.total + t.total} compartments fired; every 64th cell: "
          f"max |dV| cables {errs[0][0]:.3g} mV (half-ulp response {errs[0][1]:.3g}), "
          f"trees {errs[1][0]:.3g} mV ({errs[1][1]:.3g})")
    assert n_fired > 10000


def test_full_population_without_drive_stays_with_a_point_cell(orc, population):
    """With no drive every compartment starts alike and no axial current
    flows, so the whole population follows one compartment's trajectory."""
    from dendra_amd.model import HH, CableSpec

    hh = HH(celsius=6.3)
    c, t = population.cables, population.trees
    cp, tp = _pops(c, t, hh)
    cp.run(STEPS)
    tp.run(STEPS)
    point = CableSpec(np.array([0, 1]), np.array([8.0]), np.array([10.0]),
                      np.array([100.0]), np.array([1.0]), v_init=-65.0)
    ref, _ = orc.cable_run(point, hh.array(), DT, STEPS)
    ref32, _ = orc.cable_run_f32(point, hh.array(), DT, STEPS)
    # The slow approach to rest moves V by less than half a float32 ulp per
    # step near the end, so any float32 state stalls short of the binary64
    # value (here by 1.6e-3 mV); the binary32 restatement stalls alike.
    for pop in (cp, tp):
        v = pop.v.double()
        lo, hi = float(v.min()), float(v.max())
        print(f"undriven {type(pop).__name__}: V in [{lo:.7f}, {hi:.7f}] mV, point cell "
              f"{ref['v'][0]:.7f} mV (binary32 restatement {ref32['v'][0]:.7f})")
        assert hi - lo <= 1e-4 and abs(lo - ref32["v"][0]) <= 1e-4
        assert abs(lo - ref["v"][0]) <= bars.F32


def test_full_network_delivery_is_bitexact(orc):
    from dendra_amd.model import HH
    from dendra_amd.network import Network
    from dendra_amd.workloads import sparse_network

    spec = sparse_network()
    n_steps = 2000
    net = Network(spec, 0.025, HH(celsius=6.3))
    events, delivered = net.run(n_steps, record_events=True, record_delivered=True,
                                graph_steps=100)
    conn = spec.conn
    tape = np.zeros((n_steps, conn.n_src), np.uint8)
    tape[events[:, 0], events[:, 1]] = 1
    assert spec.cells.n_cells == 21000 and len(conn.pre) == 21000 * 201
    want, _, _ = orc.ring_replay(conn.n_src, spec.cells.n_cells, conn.pre, conn.post,
                                 conn.delay, conn.w, conn.depth, tape)
    print(f"full network: {len(events)} spikes, {int(np.count_nonzero(want))} nonzero "
          f"payloads over {n_steps} steps")
    assert len(events) > 1000
    assert np.array_equal(delivered, want)
