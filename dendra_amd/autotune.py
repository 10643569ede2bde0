"""Launch-parameter autotuner of the persistent population kernels.

The cable kernel's plan has three knobs:

* ``rmax``, rows per lane, which is the Thomas/PCR crossover: a cable of K
  compartments gets the fewest lanes P in {4, ..., 64} with ceil(K/P) <= rmax,
  so a larger rmax means longer per-lane Thomas sweeps and fewer PCR stages
  across lanes (and fewer wavefronts);
* ``max_lanes``, the widest PCR (cables longer than max_lanes * rmax are not
  allowed);
* ``waves_per_block``, wavefronts per workgroup, which with 64/P cables per
  wavefront sets the cells per workgroup.

The tree kernel's knob is the thread count of the workgroups its trees are
packed into (``block``).  Steps per launch do not change the throughput of a
persistent kernel (its state stays on chip); runs use one launch unless the
caller records.

``tune_cables`` / ``tune_trees`` time a short run of every candidate on the
GPU (HIP events on the launch stream) and return the fastest.  Results are
kept per population signature in memory and, with ``cache_path``, in a JSON
file.
"""

from __future__ import annotations

import hashlib
import json
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np
import torch

from .model import HH, CableSpec

CABLE_RMAX = (6, 8, 10, 13)
CABLE_LANES = (32, 64)
CABLE_WAVES = (1, 2, 4)
TREE_BLOCKS = (256, 512)

_memo: dict = {}


@dataclass
class CableTuning:
    rmax: int
    max_lanes: int
    waves_per_block: int
    ms_per_step: float
    candidates: int


@dataclass
class TreeTuning:
    block: int
    ms_per_step: float
    candidates: int


def _signature(kind, sizes, dtype, extra=()) -> str:
    h = hashlib.sha1()
    h.update(kind.encode())
    h.update(np.asarray(sizes, dtype=np.int64).tobytes())
    h.update(str(dtype).encode())
    h.update(repr(extra).encode())
    h.update(torch.cuda.get_device_name(0).encode())
    return h.hexdigest()


def _load(path):
    if path is None:
        return {}
    try:
        return json.loads(Path(path).read_text())
    except (OSError, ValueError):
        return {}


def _store(path, key, value):
    if path is None:
        return
    data = _load(path)
    data[key] = value
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    Path(path).write_text(json.dumps(data, indent=1))


def _time(fn, steps) -> float:
    """Average ms per step of fn(n) run for `steps` steps (after one warm-up)."""
    fn(max(1, steps // 4))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    fn(steps)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def tune_cables(spec: CableSpec, dt: float, hh: HH | None = None, dtype=torch.float32,
                device="cuda", wave=None, steps: int = 100, rmax=CABLE_RMAX,
                max_lanes=CABLE_LANES, waves=CABLE_WAVES, cache_path=None):
    """Fastest (rmax, max_lanes, waves_per_block) for this cable population.

    Returns (CableTuning, CablePopulation built with the chosen plan)."""
    from .population import CablePopulation

    key = _signature("cable", spec.sizes, dtype, (tuple(rmax), tuple(max_lanes),
                                                  tuple(waves), float(dt)))
    hit = _memo.get(key) or _load(cache_path).get(key)
    if hit:
        t = CableTuning(**hit)
        return t, CablePopulation(spec, dt, hh, dtype, device, rmax=t.rmax,
                                  max_lanes=t.max_lanes)
    kmax = int(spec.sizes.max())
    best, best_pop, n = None, None, 0
    for r in rmax:
        for lanes in max_lanes:
            if lanes * r < kmax:
                continue
            pop = CablePopulation(spec, dt, hh, dtype, device, rmax=r, max_lanes=lanes)
            won = False
            for w in waves:
                def run(k, pop=pop, w=w):
                    pop.reset()
                    pop.run(k, wave=wave, waves_per_block=w)
                ms = _time(run, steps)
                n += 1
                if best is None or ms < best.ms_per_step:
                    best, won = CableTuning(r, lanes, w, ms, 0), True
            if won:
                best_pop = pop
    if best is None:
        raise ValueError("no candidate plan covers the longest cable")
    best.candidates = n
    best_pop.reset()
    _memo[key] = asdict(best)
    _store(cache_path, key, asdict(best))
    return best, best_pop


def tune_trees(pop, wave=None, steps: int = 50, blocks=TREE_BLOCKS,
               cache_path=None) -> TreeTuning:
    """Fastest workgroup size for the packed trees of a TreePopulation; the
    population is left packed for it."""
    spec = pop.spec
    key = _signature("tree", np.concatenate([spec.sizes, spec.parent]), pop.dtype,
                     (tuple(blocks), float(pop.dt)))
    hit = _memo.get(key) or _load(cache_path).get(key)
    if hit:
        t = TreeTuning(**hit)
        pop.set_block(t.block)
        return t
    best, n = None, 0
    for b in blocks:
        pop.set_block(b)

        def run(k):
            pop.reset()
            pop.run(k, wave=wave)
        ms = _time(run, steps)
        n += 1
        if best is None or ms < best.ms_per_step:
            best = TreeTuning(b, ms, 0)
    best.candidates = n
    pop.set_block(best.block)
    pop.reset()
    _memo[key] = asdict(best)
    _store(cache_path, key, asdict(best))
    return best
