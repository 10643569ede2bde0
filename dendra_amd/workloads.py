"""Synthetic workloads of the benchmark configurations (seeded, counter RNG).

* ``single_cable``: one unbranched squid-HH cable, 101 compartments of
  10 um (the reference's Unmyelinated(L=1000 um, dx=10) discretisation,
  models/core.py:10926-10945), d 2 um, Ra 100 ohm cm, cm 1 uF/cm^2,
  6.3 degC, v0 -65 mV, with a batch of intracellular 1 ms pulses at
  compartment 0, amplitudes uniform in [0.05, 1] nA.
* ``uneven_population``: cables with compartment counts uniform in
  [kmin, kmax], 10 um compartments, lognormal diameters (median 8 um,
  sigma 0.4), each driven by a point electrode in a 300 ohm cm medium at a
  perpendicular distance uniform in 0.1-3 mm above a uniform position
  along the cable, one monophasic 0.1 ms pulse of amplitude uniform in
  +-2 mA at 0.5 ms; plus seeded random binary trees grown by a branching
  process (sections of 5-20 compartments of 10 um, a uniformly chosen leaf
  section splits into two children of 0.8x its diameter until the size
  target, uniform in [tmin, tmax], is reached).
* ``fit_cables``: cables of equal compartment count for trace fitting
  with per-compartment gNa/gK scales (lognormal, sigma 0.2) and a shared
  Ve table of (compartments x knots): a point electrode 0.3 mm above the
  middle compartment delivering cathodic triangular pulses (one knot wide
  on each side) of amplitude uniform in [0.02, 0.12] mA at every fourth
  knot; targets are seeded spike-shaped traces at 8 recording sites.
* ``threshold_fibers``: HH fibers of 101 compartments of 10 um, diameters
  uniform in [dmin, dmax] um, each under a point electrode in a 300 ohm cm
  medium at a position uniform in x over the middle half of the fiber and
  in perpendicular distance over [zmin, zmax] mm; the Ve profile is the
  float32 lead field for 1 mA.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import rng
from .model import CableSpec, TreeSpec, point_source_lead, rect_wave

# RNG stream ids
S_K, S_DIAM, S_DIST, S_POS, S_AMP, S_TK, S_TREE, S_TDIAM = range(1, 9)
S_SNA, S_SK, S_PULSE, S_TGT, S_SIGN = range(40, 45)
S_FD, S_FX, S_FZ = range(50, 53)


def single_cable(batch: int = 16, seed: int = 1, n_comp: int = 101,
                 dx: float = 10.0) -> CableSpec:
    off = np.arange(batch + 1, dtype=np.int64) * n_comp
    total = batch * n_comp
    amps_na = 0.05 + 0.95 * rng.uniform_n(seed, S_AMP, batch)
    iamp = np.zeros(total)
    iamp[off[:-1]] = amps_na * 1e-6  # nA -> mA
    return CableSpec(
        off=off, diam=np.full(total, 2.0), dx=np.full(total, dx),
        rhoa=np.full(total, 100.0), cm=np.full(total, 1.0), iamp=iamp, v_init=-65.0,
    )


def single_cable_wave(n_steps: int = 800, dt: float = 0.025, dtype=np.float64):
    """1 ms intracellular pulse from t = 0 on the start-of-step grid."""
    return rect_wave(n_steps, dt, delay=0.0, pw=1.0, dtype=dtype)


def extracellular_wave(n_steps: int, dt: float, dtype=np.float32, delay: float = 0.5,
                       pw: float = 0.1):
    return rect_wave(n_steps, dt, delay=delay, pw=pw, dtype=dtype)


def uneven_cables(n: int, seed: int = 2, kmin: int = 21, kmax: int = 801,
                  dx: float = 10.0, rhoa: float = 100.0) -> CableSpec:
    K = rng.integers_n(seed, S_K, n, kmin, kmax)
    diam_c = np.exp(np.log(8.0) + 0.4 * rng.normal_n(seed, S_DIAM, n))
    dist_um = 1e3 * (0.1 + 2.9 * rng.uniform_n(seed, S_DIST, n))
    pos_frac = rng.uniform_n(seed, S_POS, n)
    amp = -2.0 + 4.0 * rng.uniform_n(seed, S_AMP, n)
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum(K)
    total = int(off[-1])
    cell = np.repeat(np.arange(n), K)
    j = np.arange(total) - off[cell]
    x = j * dx
    ex = pos_frac[cell] * (K[cell] - 1) * dx
    lead = point_source_lead(x, 0.0, 0.0, ex, 0.0, dist_um[cell])
    return CableSpec(
        off=off, diam=diam_c[cell], dx=np.full(total, dx), rhoa=np.full(total, rhoa),
        cm=np.ones(total), lead=lead, ve_amp=amp, v_init=-65.0,
    )


def uneven_parity_sample() -> CableSpec:
    """41 extracellularly driven cables for parity runs: 24 cables of the
    uneven population (seed 5) plus edge-case sizes -- one compartment, the
    lane and tile boundaries of the cable plan, and the largest plan (801,
    832 compartments).  Geometry and lead fields repeat the first 24 cables'."""
    base = uneven_cables(64, seed=5)
    extra_k = [1, 2, 3, 4, 21, 52, 53, 64, 65, 104, 105, 208, 209, 416, 417, 801, 832]
    sizes = np.concatenate([base.sizes[:24], extra_k])
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    total = int(off[-1])
    cell = np.repeat(np.arange(len(sizes)), sizes)
    src = base.off[cell % 24] + (np.arange(total) - off[cell]) % base.sizes[cell % 24]
    return CableSpec(off, base.diam[src], base.dx[src], base.rhoa[src], base.cm[src],
                     lead=base.lead[src], ve_amp=base.ve_amp[np.arange(len(sizes)) % 24],
                     v_init=-65.0)


def random_binary_tree(target: int, key: int, seed: int):
    """One tree: parent (parent < child), per-node diameter, x/y/z (um)."""
    u = lambda n, stream: rng.uniform_n(seed, stream, n, start=key * 1_000_003)
    draws = u(4 * target + 16, S_TREE)
    di = 0

    def nxt():
        nonlocal di
        val = draws[di % len(draws)]
        di += 1
        return val

    d0 = float(np.exp(np.log(8.0) + 0.4 * rng.normal_n(seed + key, S_TDIAM, 1)[0]))
    parent, diam, pos = [], [], []
    # section: (first node, last node, diameter, end position, direction)
    leaves = []

    def grow(par_node, d, start, direction):
        n = 5 + int(nxt() * 16)
        n = min(n, target - len(parent))
        p = par_node
        cur = np.array(start, dtype=np.float64)
        for _ in range(n):
            cur = cur + 10.0 * direction
            parent.append(p)
            diam.append(d)
            pos.append(cur.copy())
            p = len(parent) - 1
        return p, cur

    def rand_dir():
        z = 2.0 * nxt() - 1.0
        phi = 2.0 * np.pi * nxt()
        s = np.sqrt(max(0.0, 1.0 - z * z))
        return np.array([s * np.cos(phi), s * np.sin(phi), z])

    last, end = grow(-1, d0, (0.0, 0.0, 0.0), rand_dir())
    leaves.append((last, d0, end))
    while len(parent) < target:
        k = int(nxt() * len(leaves))
        node, d, end = leaves.pop(k)
        for _ in range(2):
            if len(parent) >= target:
                break
            last, e2 = grow(node, 0.8 * d, end, rand_dir())
            leaves.append((last, 0.8 * d, e2))
    return (np.asarray(parent, dtype=np.int64), np.asarray(diam),
            np.asarray(pos, dtype=np.float64))


def random_trees(n: int, seed: int = 3, tmin: int = 200, tmax: int = 4000,
                 ra: float = 100.0, dx: float = 10.0) -> TreeSpec:
    sizes = rng.integers_n(seed, S_TK, n, tmin, tmax)
    dist_um = 1e3 * (0.1 + 2.9 * rng.uniform_n(seed, S_DIST, n))
    amp = -2.0 + 4.0 * rng.uniform_n(seed, S_AMP, n)
    parents, diams, leads = [], [], []
    for t in range(n):
        par, d, pos = random_binary_tree(int(sizes[t]), t, seed)
        parents.append(par)
        diams.append(d)
        leads.append(point_source_lead(pos[:, 0], pos[:, 1], pos[:, 2],
                                       0.0, 0.0, dist_um[t]))
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    total = int(off[-1])
    return TreeSpec(
        off=off, parent=np.concatenate(parents), diam=np.concatenate(diams),
        length=np.full(total, dx), ra=np.full(total, ra), cm=np.ones(total),
        lead=np.concatenate(leads), ve_amp=amp, v_init=-65.0,
    )


@dataclass
class UnevenPopulation:
    cables: CableSpec
    trees: TreeSpec
    dt: float
    n_steps: int
    wave: np.ndarray

    @property
    def compartments(self) -> int:
        return self.cables.total + self.trees.total

    @property
    def compartment_steps(self) -> int:
        return self.compartments * self.n_steps


def uneven_population(n_cables: int = 8192, n_trees: int = 1024, n_steps: int = 4000,
                      dt: float = 0.005, seed: int = 7) -> UnevenPopulation:
    cables = uneven_cables(n_cables, seed=seed)
    trees = random_trees(n_trees, seed=seed + 1)
    return UnevenPopulation(cables, trees, dt, n_steps,
                            extracellular_wave(n_steps, dt, np.float32))


S_PRE, S_W, S_DELAY = 21, 22, 23


def sparse_network(n_point: int = 20000, n_bs: int = 1000, fan_in: int = 200,
                   dt: float = 0.025, seed: int = 11, w_bg: float = 0.01,
                   delay_ms=(1.0, 5.0)):
    """Recurrent HH network of the network-delivery benchmark.

    Point cells: one HH compartment, 20 um x 20 um.  Ball-and-stick cells: an
    HH soma (20 um x 20 um) and a passive 100-compartment dendrite of 1 um x
    500 um (5 um compartments; gNa = gK = 0).  Ra 100 ohm cm, cm 1 uF/cm^2.
    Each cell gets `fan_in` sources uniform over all cells, exp2syn weights
    lognormal (median 1 nS, sigma 0.5) and delays that are integer multiples of
    dt uniform in [1, 5] ms, plus one Poisson background source of weight
    w_bg uS and a one-step delay.  Connections are numbered target-major.
    """
    from .model import CableSpec
    from .network import Connectivity, NetworkSpec, delay_steps

    n_cells = n_point + n_bs
    sizes = np.concatenate([np.ones(n_point, np.int64), np.full(n_bs, 101, np.int64)])
    off = np.zeros(n_cells + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    total = int(off[-1])
    diam = np.full(total, 20.0)
    dx = np.full(total, 20.0)
    sna = np.ones(total)
    sk = np.ones(total)
    for c in range(n_point, n_cells):
        o = off[c]
        diam[o + 1:o + 101] = 1.0
        dx[o + 1:o + 101] = 5.0
        sna[o + 1:o + 101] = 0.0
        sk[o + 1:o + 101] = 0.0
    cells = CableSpec(off, diam, dx, np.full(total, 100.0), np.ones(total), sna=sna,
                      sk=sk, v_init=-65.0)
    n_rec = n_cells * fan_in
    pre_rec = rng.integers_n(seed, S_PRE, n_rec, 0, n_cells - 1).reshape(n_cells, fan_in)
    w_rec = np.exp(np.log(1e-3) + 0.5 * rng.normal_n(seed, S_W, n_rec)).reshape(n_cells,
                                                                               fan_in)
    lo, hi = (int(round(x / dt)) for x in delay_ms)
    d_rec = delay_steps(rng.integers_n(seed, S_DELAY, n_rec, lo, hi) * dt, dt)
    d_rec = d_rec.reshape(n_cells, fan_in)
    pre = np.concatenate([pre_rec, n_cells + np.arange(n_cells)[:, None]], axis=1)
    w = np.concatenate([w_rec, np.full((n_cells, 1), w_bg)], axis=1)
    dly = np.concatenate([d_rec, np.ones((n_cells, 1), np.int64)], axis=1)
    post = np.repeat(np.arange(n_cells), fan_in + 1)
    conn = Connectivity(pre.ravel().astype(np.int32), post.astype(np.int32),
                        dly.ravel().astype(np.int32), w.ravel().astype(np.float32),
                        n_src=2 * n_cells, n_tgt=n_cells)
    return NetworkSpec(cells=cells, n_point=n_point, soma=off[:-1].copy(), conn=conn,
                       seed=seed)


@dataclass
class FitWorkload:
    spec: CableSpec
    dt: float
    n_steps: int
    sna: np.ndarray          # (n_cells, K)
    sk: np.ndarray           # (n_cells, K)
    ve: np.ndarray           # (K, n_knots), mV
    knot_times: np.ndarray   # (n_knots,), ms
    sites: np.ndarray        # (8,)
    targets: np.ndarray      # (n_steps, n_cells, 8), float32
    margin_site: int
    margin_sign: np.ndarray  # (n_cells,) +-1
    margin_weight: float
    theta: float
    tau: float


def fit_cables(n_cells: int = 512, n_comp: int = 201, n_steps: int = 10000,
               dt: float = 0.01, n_knots: int = 64, seed: int = 5,
               n_sites: int = 8, ve_gain: float = 1.0) -> FitWorkload:
    dx = 10.0
    off = np.arange(n_cells + 1, dtype=np.int64) * n_comp
    total = n_cells * n_comp
    diam = np.exp(np.log(4.0) + 0.2 * rng.normal_n(seed, S_DIAM, n_cells))
    spec = CableSpec(off=off, diam=np.repeat(diam, n_comp), dx=np.full(total, dx),
                     rhoa=np.full(total, 100.0), cm=np.ones(total), v_init=-65.0)
    sna = np.exp(0.2 * rng.normal_n(seed, S_SNA, total)).reshape(n_cells, n_comp)
    sk = np.exp(0.2 * rng.normal_n(seed, S_SK, total)).reshape(n_cells, n_comp)
    x = np.arange(n_comp) * dx
    lead = point_source_lead(x, 0.0, 0.0, x[n_comp // 2], 0.0, 300.0)
    knot_times = np.linspace(0.0, n_steps * dt, n_knots)
    cur = np.zeros(n_knots)
    pulse = np.arange(2, n_knots - 1, 4)
    cur[pulse] = -(0.02 + 0.10 * rng.uniform_n(seed, S_PULSE, len(pulse)))
    ve = ve_gain * lead[:, None] * cur[None, :]
    sites = np.round(np.linspace(0, n_comp - 1, n_sites)).astype(np.int64)
    # targets: resting trace with spike-shaped bumps 1 ms after a random
    # subset of the pulses, per cell and site
    t = (np.arange(n_steps) + 1) * dt
    hit = rng.uniform_n(seed, S_TGT, n_cells * len(pulse)).reshape(n_cells, len(pulse)) < 0.5
    targets = np.full((n_steps, n_cells, n_sites), -65.0, dtype=np.float32)
    for i, kp in enumerate(pulse):
        bump = (105.0 * np.exp(-((t - knot_times[kp] - 1.0) / 0.6) ** 2)).astype(np.float32)
        targets += bump[:, None, None] * hit[None, :, i, None]
    sign = np.where(rng.uniform_n(seed, S_SIGN, n_cells) < 0.5, -1.0, 1.0)
    return FitWorkload(spec, dt, n_steps, sna, sk, ve, knot_times, sites, targets,
                       n_sites - 1, sign, 10.0, 0.0, 1.0)


@dataclass
class FiberSet:
    spec: CableSpec
    profile: np.ndarray   # (n, K) float32, mV per mA
    diam: np.ndarray      # (n,) um


def threshold_fibers(n: int, seed: int = 9, n_comp: int = 101, dx: float = 10.0,
                     dmin: float = 2.0, dmax: float = 16.0, zmin: float = 0.1,
                     zmax: float = 1.5, start: int = 0) -> FiberSet:
    """Fibers start..start+n-1 of the seeded stream (disjoint ranges give
    disjoint training / held-out sets)."""
    u = lambda stream: rng.uniform_n(seed, stream, n, start=start)
    diam = dmin + (dmax - dmin) * u(S_FD)
    L = (n_comp - 1) * dx
    ex = L * (0.25 + 0.5 * u(S_FX))
    ez = 1e3 * (zmin + (zmax - zmin) * u(S_FZ))
    x = np.arange(n_comp) * dx
    prof = point_source_lead(x[None, :], 0.0, 0.0, ex[:, None], 0.0, ez[:, None])
    prof = prof.astype(np.float32)
    off = np.arange(n + 1, dtype=np.int64) * n_comp
    total = n * n_comp
    spec = CableSpec(off=off, diam=np.repeat(diam, n_comp), dx=np.full(total, dx),
                     rhoa=np.full(total, 100.0), cm=np.ones(total), v_init=-65.0)
    return FiberSet(spec, prof, diam)
