"""Host-side model description and workspace preparation (float64 numpy).

Geometry and workspace follow the reference:
  * cylinder area and half-cylinder edge conductance
    (dendra/models/integrators/cable.py:10-29),
  * the unbranched implicit-Euler workspace gL = g/C_j, gR = g/C_{j+1},
    scale = area/C (integrators/implicit.py:484-522),
  * child-indexed tree conductance 1 / (R_half_child + R_half_parent) and
    cm_dt = C / dt_s (integrators/tree.py:477-612,1010-1044),
  * HH steady state alpha/(alpha+beta) at v_init without q10
    (mod/HH.py:42-56) and q10 = 3^((celsius-6.3)/10) (mod/HH.py:21-22),
  * isotropic point-source lead field rho_e / (4 pi r) in mV/mA
    (fields/analytic.py:59-62) and half-open rectangular pulses
    (stim/waveform/implementations.py:512-558) on the start-of-step grid
    t_i = t0 + i dt (models/core.py:1280-1305).
The extracellular and intracellular drives are linear in their waveform, so
they are folded into one per-compartment "drive" coefficient (the axial
operator applied to the lead field, plus the intracellular amplitude), which
the kernels scale by the waveform value each step.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PI = np.pi


@dataclass
class HH:
    """Squid HH channel parameters (mod/HH.py:59-116 defaults)."""

    gnabar: float = 0.12
    gkbar: float = 0.036
    gl: float = 0.0003
    ena: float = 50.0
    ek: float = -77.0
    el: float = -54.3
    celsius: float = 6.3

    @property
    def q10(self) -> float:
        return 3.0 ** ((self.celsius - 6.3) / 10.0)

    def array(self) -> np.ndarray:
        return np.array(
            [self.gnabar, self.gkbar, self.gl, self.ena, self.ek, self.el, self.q10],
            dtype=np.float64,
        )


def _vtrap(x, y):
    q = x / y
    small = np.abs(q) <= 0.5
    qs = np.where(small, q, 0.0)
    q2 = qs * qs
    e = q2 * (43867 / 5109094217170944000) - 3617 / 10670622842880000
    for coef in (1 / 74724249600, -691 / 1307674368000, 1 / 47900160, -1 / 1209600,
                 1 / 30240, -1 / 720, 1 / 12):
        e = e * q2 + coef
    series = (y - x / 2) + x * (qs * e)
    qsafe = np.where(small, 1.0, q)
    return np.where(small, series, x / (np.exp(qsafe) - 1.0))


def hh_steady_state(v):
    """Gate steady state at voltage v (float64)."""
    v = np.asarray(v, dtype=np.float64)
    am = 0.1 * _vtrap(-(v + 40.0), 10.0)
    bm = 4.0 * np.exp(-(v + 65.0) / 18.0)
    ah = 0.07 * np.exp(-(v + 65.0) / 20.0)
    bh = 1.0 / (np.exp(-(v + 35.0) / 10.0) + 1.0)
    an = 0.01 * _vtrap(-(v + 55.0), 10.0)
    bn = 0.125 * np.exp(-(v + 65.0) / 80.0)
    return am / (am + bm), ah / (ah + bh), an / (an + bn)


def point_source_lead(x_um, y_um, z_um, ex, ey, ez, rhoe=300.0):
    """Lead field (mV per mA) of an isotropic point source."""
    r = np.sqrt((x_um - ex) ** 2 + (y_um - ey) ** 2 + (z_um - ez) ** 2) * 1e-4
    return rhoe / (4.0 * PI * r)


def rect_wave(n_steps: int, dt: float, delay: float, pw: float, amp: float = 1.0,
              t0: float = 0.0, dtype=np.float32) -> np.ndarray:
    """Rectangular pulse sampled on the start-of-step grid in the model dtype:
    amp where delay <= t < delay + pw."""
    t = (t0 + np.arange(n_steps, dtype=np.float64) * dt).astype(dtype)
    d = dtype(delay)
    e = dtype(d + dtype(pw))
    gate = (t >= d) & (t < e)
    return np.where(gate, amp, 0.0).astype(np.float64)


@dataclass
class CableSpec:
    """Ragged batch of unbranched cables; compartment j of cell c at off[c]+j."""

    off: np.ndarray
    diam: np.ndarray
    dx: np.ndarray
    rhoa: np.ndarray
    cm: np.ndarray
    sna: np.ndarray | None = None
    sk: np.ndarray | None = None
    lead: np.ndarray | None = None     # mV/mA per compartment
    ve_amp: np.ndarray | None = None   # mA per cell
    iamp: np.ndarray | None = None     # mA per compartment
    v_init: float = -65.0

    @property
    def n_cells(self) -> int:
        return len(self.off) - 1

    @property
    def total(self) -> int:
        return int(self.off[-1])

    @property
    def sizes(self) -> np.ndarray:
        return np.diff(self.off)

    def subset(self, cells) -> "CableSpec":
        cells = np.asarray(cells, dtype=np.int64)
        idx = np.concatenate([np.arange(self.off[c], self.off[c + 1]) for c in cells])
        off = np.zeros(len(cells) + 1, dtype=np.int64)
        off[1:] = np.cumsum(self.sizes[cells])
        pick = lambda a: None if a is None else a[idx]
        return CableSpec(off, self.diam[idx], self.dx[idx], self.rhoa[idx],
                         self.cm[idx], pick(self.sna), pick(self.sk), pick(self.lead),
                         None if self.ve_amp is None else self.ve_amp[cells],
                         pick(self.iamp), self.v_init)


@dataclass
class CableWorkspace:
    gL: np.ndarray
    gRp: np.ndarray
    scale: np.ndarray
    cm_inv: np.ndarray
    drive: np.ndarray | None


def cable_workspace(spec: CableSpec) -> CableWorkspace:
    area = spec.diam * 1e-4 * PI * spec.dx * 1e-4
    cap = 1e-6 * spec.cm * area
    cm_inv = 1.0 / cap
    r = 1e-4 * spec.diam / 2.0
    R = spec.rhoa * (1e-4 * spec.dx) / (PI * r * r)
    total = spec.total
    gL = np.zeros(total)
    gRp = np.zeros(total)
    for c in range(spec.n_cells):
        o, e = int(spec.off[c]), int(spec.off[c + 1])
        if e - o < 2:
            continue
        g = 2.0 / (R[o:e - 1] + R[o + 1:e])
        gL[o:e - 1] = g / cap[o:e - 1]
        gRp[o + 1:e] = g / cap[o + 1:e]
    drive = None
    if spec.lead is not None:
        amp = np.repeat(spec.ve_amp, spec.sizes) if spec.ve_amp is not None else 1.0
        ve = spec.lead * amp
        right = np.zeros(total)
        left = np.zeros(total)
        right[:-1] = ve[1:] - ve[:-1]
        left[1:] = ve[1:] - ve[:-1]
        drive = gL * right - gRp * left  # zero conductance masks cell boundaries
    if spec.iamp is not None:
        drive = (0.0 if drive is None else drive) + spec.iamp * cm_inv
    return CableWorkspace(gL, gRp, area * cm_inv, cm_inv, drive)


@dataclass
class TreeSpec:
    """Ragged batch of trees; parent holds local indices with parent < child."""

    off: np.ndarray
    parent: np.ndarray
    diam: np.ndarray
    length: np.ndarray
    ra: np.ndarray
    cm: np.ndarray
    sna: np.ndarray | None = None
    sk: np.ndarray | None = None
    lead: np.ndarray | None = None
    ve_amp: np.ndarray | None = None
    iamp: np.ndarray | None = None
    v_init: float = -65.0

    @property
    def n_trees(self) -> int:
        return len(self.off) - 1

    @property
    def total(self) -> int:
        return int(self.off[-1])

    @property
    def sizes(self) -> np.ndarray:
        return np.diff(self.off)

    def subset(self, trees) -> "TreeSpec":
        trees = np.asarray(trees, dtype=np.int64)
        idx = np.concatenate([np.arange(self.off[t], self.off[t + 1]) for t in trees])
        off = np.zeros(len(trees) + 1, dtype=np.int64)
        off[1:] = np.cumsum(self.sizes[trees])
        pick = lambda a: None if a is None else a[idx]
        return TreeSpec(off, self.parent[idx], self.diam[idx], self.length[idx],
                        self.ra[idx], self.cm[idx], pick(self.sna), pick(self.sk),
                        pick(self.lead),
                        None if self.ve_amp is None else self.ve_amp[trees],
                        pick(self.iamp), self.v_init)


@dataclass
class TreeWorkspace:
    area: np.ndarray
    cm_dt: np.ndarray
    g: np.ndarray        # child-indexed, 0 at roots (original numbering)
    drive: np.ndarray | None


def tree_workspace(spec: TreeSpec, dt: float) -> TreeWorkspace:
    dt_s = dt * 1e-3
    area = PI * spec.diam * spec.length * 1e-8
    cm_dt = 1e-6 * spec.cm * area / dt_s
    total = spec.total
    gpar = np.full(total, -1, dtype=np.int64)
    for t in range(spec.n_trees):
        o, e = int(spec.off[t]), int(spec.off[t + 1])
        p = spec.parent[o:e]
        gpar[o:e] = np.where(p >= 0, p + o, -1)
    has = gpar >= 0
    pp = np.where(has, gpar, 0)
    r = 0.5 * spec.diam * 1e-4
    rhalf = spec.ra * (spec.length * 1e-4 / 2.0) / (PI * r * r)
    g = np.where(has, 1.0 / (rhalf + rhalf[pp]), 0.0)
    drive = None
    if spec.lead is not None:
        amp = np.repeat(spec.ve_amp, spec.sizes) if spec.ve_amp is not None else 1.0
        ve = spec.lead * amp
        flow = np.where(has, g * (ve - ve[pp]), 0.0)
        drive = -flow
        np.add.at(drive, pp[has], flow[has])
    if spec.iamp is not None:
        drive = (0.0 if drive is None else drive) + spec.iamp
    return TreeWorkspace(area, cm_dt, g, drive)


@dataclass
class CablePlan:
    """Wavefront tasks of the persistent cable kernel.

    Each cell gets P lanes (the smallest of 4..64 with ceil(K/P) <= rmax) and
    nr = ceil(K/P) rows per lane; cells are sorted by (P, nr) and packed 64/P
    to a wavefront; tasks are ordered by estimated cost, largest first.
    """

    order: np.ndarray       # plan position -> user cell index
    off: np.ndarray         # offsets in plan order
    task_cell0: np.ndarray
    task_ncell: np.ndarray
    task_nr: np.ndarray
    task_p: np.ndarray
    cell_p: np.ndarray = field(default=None)


LANE_CHOICES = (4, 8, 16, 32, 64)


def plan_cables(sizes: np.ndarray, rmax: int = 13, max_lanes: int = 64) -> CablePlan:
    sizes = np.asarray(sizes, dtype=np.int64)
    if (sizes < 1).any():
        raise ValueError("every cable needs at least one compartment")
    lanes = np.full(len(sizes), -1, dtype=np.int64)
    for p_ in reversed([p for p in LANE_CHOICES if p <= max_lanes]):
        lanes = np.where(-(-sizes // p_) <= rmax, p_, lanes)
    if (lanes < 0).any():
        k = int(sizes[lanes < 0].max())
        raise ValueError(f"cable of {k} compartments exceeds {max_lanes * rmax}")
    nr = -(-sizes // lanes)
    # sort by lanes desc, then nr desc (stable in user order)
    order = np.lexsort((np.arange(len(sizes)), -nr, -lanes))
    task_cell0, task_ncell, task_nr, task_p = [], [], [], []
    pos = 0
    while pos < len(order):
        p_ = int(lanes[order[pos]])
        per = 64 // p_
        grp = [pos]
        while len(grp) < per and pos + len(grp) < len(order) and \
                lanes[order[pos + len(grp)]] == p_:
            grp.append(pos + len(grp))
        task_cell0.append(pos)
        task_ncell.append(len(grp))
        task_nr.append(int(nr[order[pos]]))
        task_p.append(p_)
        pos += len(grp)
    task_cell0 = np.asarray(task_cell0)
    task_ncell = np.asarray(task_ncell)
    task_nr = np.asarray(task_nr)
    task_p = np.asarray(task_p)
    cost = task_nr * 1.0 + np.log2(task_p) * 0.6
    t_order = np.argsort(-cost, kind="stable")
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes[order])
    return CablePlan(order, off, task_cell0[t_order], task_ncell[t_order],
                     task_nr[t_order], task_p[t_order], lanes[order])


def plan_cells_per_task(sizes: np.ndarray, rcap: int, width: int) -> CablePlan:
    """One cell per task of `width` threads (several wavefronts of a
    workgroup), nr = ceil(K / width) <= rcap rows per thread; tasks in cell
    order."""
    sizes = np.asarray(sizes, dtype=np.int64)
    if (sizes < 1).any():
        raise ValueError("every cable needs at least one compartment")
    nr = -(-sizes // width)
    if (nr > rcap).any():
        raise ValueError(f"cable of {int(sizes.max())} compartments exceeds {width * rcap}")
    n = len(sizes)
    order = np.arange(n)
    off = np.zeros(n + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    return CablePlan(order, off, np.arange(n), np.ones(n, dtype=np.int64), nr,
                     np.full(n, width, dtype=np.int64), np.full(n, width, dtype=np.int64))


def slot_map(plan: CablePlan, sizes: np.ndarray, off: np.ndarray, rcap: int,
             width: int = 64) -> np.ndarray:
    """Slot of every compartment (spec order) in the row-slot layout of the
    fitting and threshold kernels: task t owns width*rcap slots (width = the
    task's threads: a wavefront, or several when one cell spans them) and row
    r of lane `sub*P + pl` is slot t*width*rcap + r*width + sub*P + pl, where
    compartment j of the task's sub-th cell sits on lane pl = j // nr, row
    r = j % nr."""
    sizes = np.asarray(sizes, dtype=np.int64)
    nt = len(plan.task_p)
    ncell = np.asarray(plan.task_ncell, dtype=np.int64)
    k = np.arange(int(ncell.sum())) - np.repeat(np.cumsum(ncell) - ncell, ncell)
    pos = np.repeat(np.asarray(plan.task_cell0, dtype=np.int64), ncell) + k
    cell = np.asarray(plan.order, dtype=np.int64)[pos]
    # per cell (spec order): its task and its index inside the task
    task_c = np.empty(len(cell), dtype=np.int64)
    sub_c = np.empty(len(cell), dtype=np.int64)
    task_c[cell] = np.repeat(np.arange(nt), ncell)
    sub_c[cell] = k
    P_c = np.asarray(plan.task_p, dtype=np.int64)[task_c]
    nr_c = np.asarray(plan.task_nr, dtype=np.int64)[task_c]
    c_of = np.repeat(np.arange(len(sizes)), sizes)
    j = np.arange(int(sizes.sum())) - np.asarray(off, dtype=np.int64)[c_of]
    nr, P = nr_c[c_of], P_c[c_of]
    return task_c[c_of] * width * rcap + (j % nr) * width + sub_c[c_of] * P + j // nr
