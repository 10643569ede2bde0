"""GPU populations: the host drivers of the fused cable and tree kernels.

``CablePopulation.run`` and ``TreePopulation.run`` are the counterparts of
the reference's ``Population.run`` hot loop (dendra/models/core.py:4555-4614)
over the ``_bwd_euler_ub`` / ``_dhs`` integrators: state lives on the GPU in
plan order (structure-of-arrays), and a run is either

* ``path="persistent"``: the multi-step kernels, each wavefront / workgroup
  advancing its cells for ``steps_per_launch`` steps with state on chip, or
* ``path="plain"``: one gating+assembly launch and one solve launch per
  time step, state in HBM (the structure of the reference's GPU path).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .model import (
    HH,
    CableSpec,
    TreeSpec,
    cable_workspace,
    hh_steady_state,
    plan_cables,
    tree_workspace,
)
from .topology import section_plan

_DT = {torch.float32: np.float32, torch.float64: np.float64}


# upper row bounds of the cable launch classes (kernel instances for 8, 10
# and 13 rows per lane exist; 10 and 13 hold 4 and 3 wavefronts per SIMD)
ROW_CLASSES = (10, 13)


_STATUS_TEXT = {
    1: "a cable task has more rows per lane than its kernel instance holds",
    2: "a tree group exceeds its block's threads or node capacity, or holds a "
       "section longer than the kernel takes",
    4: "a tree exceeds its launch's node, section or level capacity",
}


def validate_cable_launches(task_nr, task_p, launches, rmax: int) -> None:
    """Host check of a cable launch list before anything is enqueued: each
    launch (t0, t1, rows) covers tasks t0..t1 whose row counts are <= rows
    <= rmax, with 4..64 lanes per cable.  Raises ValueError otherwise (the
    kernel would leave such a task unadvanced)."""
    task_nr, task_p = np.asarray(task_nr), np.asarray(task_p)
    if not set(np.unique(task_p).tolist()) <= {4, 8, 16, 32, 64}:
        raise ValueError("cable tasks must use 4, 8, 16, 32 or 64 lanes per cable")
    for t0, t1, rows in launches:
        if not 1 <= rows <= rmax:
            raise ValueError(f"launch row bound {rows} outside 1..{rmax}")
        if len(task_nr[t0:t1]) and int(task_nr[t0:t1].max()) > rows:
            raise ValueError(f"cable task with {int(task_nr[t0:t1].max())} rows per lane "
                             f"in a launch bounded by {rows}")


def validate_tree_groups(groups, n_sec, sizes, threads: int, node_cap: int) -> None:
    """Host check of one layout-1 tree launch: every group's section count
    fits the block's threads and its node count the LDS node capacity."""
    for g in groups:
        if int(np.sum(n_sec[g])) > threads:
            raise ValueError(f"tree group of {int(np.sum(n_sec[g]))} sections exceeds "
                             f"{threads} threads")
        if int(np.sum(sizes[g])) > node_cap:
            raise ValueError(f"tree group of {int(np.sum(sizes[g]))} nodes exceeds the "
                             f"node capacity {node_cap}")


class _Status:
    """The launch-contract status word the persistent kernels report into."""

    def _init_status(self):
        self.status = torch.zeros(1, dtype=torch.int32, device=self.device)

    def check(self) -> None:
        """Raise RuntimeError if a kernel of this population reported a
        launch-contract violation (work left unadvanced); waits for the
        work already enqueued on the current stream."""
        flags = int(self.status.item())
        if flags:
            self.status.zero_()
            why = "; ".join(t for b, t in _STATUS_TEXT.items() if flags & b)
            raise RuntimeError(f"persistent run left work unadvanced ({why})")


def _dev(a, dtype, device):
    if a is None:
        return None
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


class CablePopulation(_Status):
    """Ragged batch of unbranched HH cables (backward Euler, increment form)."""

    def __init__(self, spec: CableSpec, dt: float, hh: HH | None = None,
                 dtype=torch.float32, device="cuda", rmax: int | None = None,
                 max_lanes: int = 64):
        lib = _lib.load()
        self.spec, self.dt, self.hh = spec, float(dt), hh or HH()
        self.dtype, self.device = dtype, torch.device(device)
        _lib.require_gpu(torch.empty(0, device=self.device))
        self.rmax = rmax or lib.dn_hh_cable_rmax()
        self.plan = plan_cables(spec.sizes, self.rmax, max_lanes)
        sizes = spec.sizes
        # flat permutation user -> plan order
        self.perm = np.concatenate(
            [np.arange(spec.off[c], spec.off[c + 1]) for c in self.plan.order]
        ) if spec.n_cells else np.zeros(0, dtype=np.int64)
        ws = cable_workspace(spec)
        P = self.perm
        dev, dt_ = self.device, dtype
        self.gL = _dev(ws.gL[P], dt_, dev)
        self.gRp = _dev(ws.gRp[P], dt_, dev)
        self.scale = _dev(ws.scale[P], dt_, dev)
        self.drive = _dev(None if ws.drive is None else ws.drive[P], dt_, dev)
        self.sna = _dev(None if spec.sna is None else spec.sna[P], dt_, dev)
        self.sk = _dev(None if spec.sk is None else spec.sk[P], dt_, dev)
        self.off = torch.as_tensor(self.plan.off, dtype=torch.int64, device=dev)
        i32 = lambda a: torch.as_tensor(a, dtype=torch.int32, device=dev)
        # Tasks are grouped into row classes, each one launch of a kernel
        # instance sized for its rows (fewer rows: fewer VGPRs and less LDS
        # per wavefront, so more wavefronts per CU); cost order is kept
        # within a class.  The launches run concurrently on two streams.
        nr = self.plan.task_nr
        cls = np.searchsorted(np.asarray(ROW_CLASSES), nr)
        cls = cls.max() - cls if len(nr) else cls  # the widest class first
        t_order = np.argsort(cls, kind="stable")
        self._launches = []
        for c in np.unique(cls):
            sel = np.flatnonzero(cls[t_order] == c)
            self._launches.append((int(sel[0]), int(sel[-1]) + 1, int(nr[t_order][sel].max())))
        validate_cable_launches(nr[t_order], self.plan.task_p[t_order], self._launches,
                                self.rmax)
        self.task_cell0 = i32(self.plan.task_cell0[t_order])
        self.task_ncell = i32(self.plan.task_ncell[t_order])
        self.task_nr = i32(nr[t_order])
        self.task_p = i32(self.plan.task_p[t_order])
        self._side = None
        self.max_nr_plain = int(-(-int(sizes.max()) // 64)) if len(sizes) else 1
        self.hh_arr = self.hh.array().astype(_DT[dtype])
        self._init_status()
        self.reset()

    @property
    def total(self) -> int:
        return self.spec.total

    def reset(self, v_init: float | None = None):
        v0 = self.spec.v_init if v_init is None else v_init
        m0, h0, n0 = hh_steady_state(v0)
        n = self.total
        mk = lambda x: torch.full((n,), float(x), dtype=self.dtype, device=self.device)
        self.v, self.m, self.h, self.n = mk(v0), mk(m0), mk(h0), mk(n0)
        self.t_step = 0

    def set_state(self, v, m, h, n):
        P = torch.as_tensor(self.perm, device=self.device)
        for name, val in (("v", v), ("m", m), ("h", h), ("n", n)):
            t = torch.as_tensor(np.asarray(val), dtype=self.dtype, device=self.device)
            getattr(self, name).copy_(t[P])

    def state(self) -> dict:
        """State in user (spec) order as float64 numpy arrays."""
        self.check()
        out = {}
        for name in ("v", "m", "h", "n"):
            a = np.empty(self.total)
            a[self.perm] = getattr(self, name).double().cpu().numpy()
            out[name] = a
        return out

    def unpermute(self, rec: torch.Tensor) -> np.ndarray:
        self.check()
        r = rec.double().cpu().numpy()
        out = np.empty_like(r)
        out[:, self.perm] = r
        return out

    def run(self, n_steps: int, wave=None, record_every: int = 0,
            path: str = "persistent", waves_per_block: int = 4,
            steps_per_launch: int | None = None, stream=None):
        """Advance n_steps.  ``wave`` is the per-step drive waveform (length
        >= t_step + n_steps) or None.  Returns the V record [n_rec, total] in
        plan order when record_every > 0, else None.  ``stream`` (a
        torch.cuda.Stream, None = the current stream) becomes the current
        stream for the run: the run's tensors are allocated and its kernels
        ordered on it."""
        if stream is not None:
            with torch.cuda.stream(stream):
                return self.run(n_steps, wave, record_every, path, waves_per_block,
                                steps_per_launch)
        lib = _lib.load()
        f64 = self.dtype == torch.float64
        wave_t = None
        if wave is not None:
            wave_t = torch.as_tensor(np.asarray(wave), dtype=self.dtype, device=self.device)
            if wave_t.numel() < self.t_step + n_steps:
                raise ValueError("waveform shorter than the run")
        rec = None
        if record_every > 0:
            n_rec = (self.t_step + n_steps) // record_every - self.t_step // record_every
            rec_base = self.t_step // record_every
            rec = torch.empty((n_rec, self.total), dtype=self.dtype, device=self.device)
        caller = torch.cuda.current_stream(self.device)
        strm = caller.cuda_stream
        if path == "persistent":
            fn = lib.dn_hh_cable_run_f64 if f64 else lib.dn_hh_cable_run_f32
            if len(self._launches) > 1 and self._side is None:
                self._side = torch.cuda.Stream(self.device)
            chunk = steps_per_launch or n_steps
            done = 0
            while done < n_steps:
                k = min(chunk, n_steps - done)
                rec_ptr = None
                if rec is not None:
                    # the kernel indexes rows by absolute step; shift the base
                    rec_ptr = rec.data_ptr() - rec_base * self.total * rec.element_size()
                if len(self._launches) > 1:
                    self._side.wait_stream(caller)
                for i, (t0, t1, rows) in enumerate(self._launches):
                    s_ = caller if i == 0 else self._side
                    at = lambda t: t.data_ptr() + t0 * t.element_size()
                    st = fn(
                        t1 - t0, at(self.task_cell0), at(self.task_ncell),
                        at(self.task_nr), at(self.task_p), _lib.ptr(self.off),
                        _lib.ptr(self.v), _lib.ptr(self.m), _lib.ptr(self.h),
                        _lib.ptr(self.n), _lib.ptr(self.gL), _lib.ptr(self.gRp),
                        _lib.ptr(self.scale), _lib.ptr(self.drive), _lib.ptr(self.sna),
                        _lib.ptr(self.sk), _lib.ptr(wave_t), self.t_step, k, self.dt,
                        self.hh_arr.ctypes.data, rec_ptr, max(record_every, 1), self.total,
                        waves_per_block, rows, _lib.ptr(self.status), s_.cuda_stream,
                    )
                    _lib.check(st, "persistent cable run")
                if len(self._launches) > 1:
                    caller.wait_stream(self._side)
                self.t_step += k
                done += k
        elif path == "plain":
            fn = lib.dn_hh_cable_step_f64 if f64 else lib.dn_hh_cable_step_f32
            mainv = torch.empty_like(self.v)
            rhs = torch.empty_like(self.v)
            w_host = None if wave is None else np.asarray(wave, dtype=np.float64)
            for s in range(n_steps):
                w = 0.0 if w_host is None else float(w_host[self.t_step])
                st = fn(
                    self.spec.n_cells, _lib.ptr(self.off), self.total, self.max_nr_plain,
                    _lib.ptr(self.v), _lib.ptr(self.m), _lib.ptr(self.h), _lib.ptr(self.n),
                    _lib.ptr(self.gL), _lib.ptr(self.gRp), _lib.ptr(self.scale),
                    _lib.ptr(self.drive), _lib.ptr(self.sna), _lib.ptr(self.sk),
                    w, self.dt, self.hh_arr.ctypes.data, None, None, None, None, None,
                    _lib.ptr(mainv), _lib.ptr(rhs), strm,
                )
                _lib.check(st, "plain cable step")
                self.t_step += 1
                if rec is not None and self.t_step % record_every == 0:
                    rec[self.t_step // record_every - 1 - rec_base].copy_(self.v)
        else:
            raise ValueError(f"unknown path {path!r}")
        return rec


@dataclass
class _TreeBucket:
    order: torch.Tensor
    n: int
    kmax: int
    threads: int
    layout: int
    grp_off: torch.Tensor | None
    n_groups: int


class TreePopulation(_Status):
    """Ragged batch of HH trees (backward Euler in current units, DHS)."""

    def __init__(self, spec: TreeSpec, dt: float, hh: HH | None = None,
                 dtype=torch.float32, device="cuda", bucket_edges=(1024, 2048, 4096),
                 block: int = 512):
        lib = _lib.load()
        self.spec, self.dt, self.hh = spec, float(dt), hh or HH()
        self.dtype, self.device = dtype, torch.device(device)
        dev = self.device
        slots = lib.dn_hh_tree_slots()
        sec_max = lib.dn_hh_tree_sec_max()  # the kernels' section capacity
        ws = tree_workspace(spec, dt)
        perm, parent, ch_ptr, ch_idx = [], [], [0], []
        sec = {k: [] for k in ("head", "len", "child", "nchild", "ptail", "parent")}
        sec_off, lvl_off, lvl_ptr, node_tree = [0], [0], [], []
        n_sec, n_lvl = [], []
        for t in range(spec.n_trees):
            o, e = int(spec.off[t]), int(spec.off[t + 1])
            sp = section_plan(spec.parent[o:e], sec_max=sec_max)
            perm.append(sp.perm + o)
            parent.append(sp.parent)
            ch_ptr.extend((sp.ch_ptr[1:] + len(ch_idx)).tolist())
            ch_idx.extend(sp.ch_idx.tolist())
            for k in sec:
                sec[k].append(getattr(sp, "sec_" + k))
            sec_off.append(sec_off[-1] + len(sp.sec_head))
            lvl_ptr.append(sp.lvl_ptr)
            lvl_off.append(lvl_off[-1] + len(sp.lvl_ptr))
            node_tree.append(np.full(e - o, t))
            n_sec.append(len(sp.sec_head))
            n_lvl.append(sp.n_levels)
        self.perm = np.concatenate(perm)
        P = self.perm
        i32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.int32, device=dev)
        self.node_off = torch.as_tensor(spec.off, dtype=torch.int64, device=dev)
        self.parent = i32(np.concatenate(parent))
        self.ch_ptr = i32(np.asarray(ch_ptr))
        self.ch_idx = i32(np.asarray(ch_idx) if ch_idx else np.zeros(1))
        self.sec_off = i32(sec_off)
        self.sec = {k: i32(np.concatenate(v)) for k, v in sec.items()}
        self.lvl_off = i32(lvl_off)
        self.lvl_ptr = i32(np.concatenate(lvl_ptr))
        self.node_tree = i32(np.concatenate(node_tree))
        self.n_sec = np.asarray(n_sec)
        self.n_levels = np.asarray(n_lvl)
        self.smax = int(self.n_sec.max())
        self.lmax = int(self.n_levels.max())
        dt_ = dtype
        self.area = _dev(ws.area[P], dt_, dev)
        self.cm_dt = _dev(ws.cm_dt[P], dt_, dev)
        self.g = _dev(ws.g[P], dt_, dev)
        self.drive = _dev(None if ws.drive is None else ws.drive[P], dt_, dev)
        self.sna = _dev(None if spec.sna is None else spec.sna[P], dt_, dev)
        self.sk = _dev(None if spec.sk is None else spec.sk[P], dt_, dev)
        self.hh_arr = self.hh.array().astype(_DT[dtype])
        sizes = spec.sizes
        self.kmax = int(sizes.max())
        self.bucket_edges = tuple(bucket_edges)
        self.slots = slots
        self._init_status()
        self.set_block(block)
        self.reset()

    @property
    def total(self) -> int:
        return self.spec.total

    def _plan_args(self):
        return (self.spec.n_trees, _lib.ptr(self.node_off), _lib.ptr(self.parent),
                _lib.ptr(self.ch_ptr), _lib.ptr(self.ch_idx), _lib.ptr(self.sec_off),
                _lib.ptr(self.sec["head"]), _lib.ptr(self.sec["len"]),
                _lib.ptr(self.sec["child"]), _lib.ptr(self.sec["nchild"]),
                _lib.ptr(self.sec["ptail"]), _lib.ptr(self.sec["parent"]),
                _lib.ptr(self.lvl_off),
                _lib.ptr(self.lvl_ptr), self.smax, self.lmax)

    def set_block(self, block: int):
        """(Re)pack the trees into launch buckets for workgroups of `block`
        threads.  Trees of at most 512 sections run one thread per section
        (layout 1), packed first-fit-decreasing into blocks of `block`
        threads (trees above `block` sections get 512-thread blocks of their
        own bucket); larger trees run eight node slots per thread (layout 0),
        bucketed by size."""
        if block not in (64, 128, 256, 512):
            raise ValueError("block must be 64, 128, 256 or 512")
        dev = self.device
        i32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.int32, device=dev)
        sizes = self.spec.sizes
        self.block = block
        self.buckets = []
        self._streams = None
        by_sec = self.n_sec <= 512
        for lo_s, hi_s, threads in ((0, block, block), (block, 512, 512)):
            sel = np.flatnonzero(by_sec & (self.n_sec > lo_s) & (self.n_sec <= hi_s))
            if len(sel) == 0:
                continue
            sel = sel[np.argsort(-self.n_sec[sel], kind="stable")]
            groups, loads, nodes = [], [], []
            for t in sel:
                for gi in range(len(groups)):
                    if loads[gi] + self.n_sec[t] <= threads:
                        groups[gi].append(t); loads[gi] += self.n_sec[t]
                        nodes[gi] += int(sizes[t])
                        break
                else:
                    groups.append([t]); loads.append(int(self.n_sec[t]))
                    nodes.append(int(sizes[t]))
            validate_tree_groups(groups, self.n_sec, sizes, threads, int(max(nodes)))
            order = np.concatenate([np.asarray(g_) for g_ in groups])
            goff = np.zeros(len(groups) + 1, dtype=np.int64)
            goff[1:] = np.cumsum([len(g_) for g_ in groups])
            self.buckets.append(_TreeBucket(i32(order), len(order), int(max(nodes)),
                                            threads, 1, i32(goff), len(groups)))
        lo = 0
        for hi in self.bucket_edges:
            sel = np.flatnonzero(~by_sec & (sizes > lo) & (sizes <= hi))
            lo = hi
            if len(sel):
                sel = sel[np.argsort(-sizes[sel], kind="stable")]
                kmax = int(sizes[sel].max())
                threads = int(min(512, max(64, -(-kmax // (self.slots * 64)) * 64)))
                self.buckets.append(_TreeBucket(i32(sel), len(sel), kmax, threads, 0,
                                                None, 0))
        if (~by_sec & (sizes > self.bucket_edges[-1])).any():
            raise ValueError(f"tree larger than {self.bucket_edges[-1]} compartments")

    def reset(self, v_init: float | None = None):
        v0 = self.spec.v_init if v_init is None else v_init
        m0, h0, n0 = hh_steady_state(v0)
        n = self.total
        mk = lambda x: torch.full((n,), float(x), dtype=self.dtype, device=self.device)
        self.v, self.m, self.h, self.n = mk(v0), mk(m0), mk(h0), mk(n0)
        self.t_step = 0

    def _bucket_streams(self):
        """One default-priority stream per launch bucket, so buckets overlap."""
        if getattr(self, "_streams", None) is None:
            self._streams = [torch.cuda.Stream(self.device) for _ in self.buckets]
        return self._streams

    def state(self) -> dict:
        self.check()
        out = {}
        for name in ("v", "m", "h", "n"):
            a = np.empty(self.total)
            a[self.perm] = getattr(self, name).double().cpu().numpy()
            out[name] = a
        return out

    def unpermute(self, rec: torch.Tensor) -> np.ndarray:
        self.check()
        r = rec.double().cpu().numpy()
        out = np.empty_like(r)
        out[:, self.perm] = r
        return out

    def run(self, n_steps: int, wave=None, record_every: int = 0,
            path: str = "persistent", steps_per_launch: int | None = None, stream=None):
        """Advance n_steps (as CablePopulation.run; ``stream`` likewise becomes
        the current stream for the run)."""
        if stream is not None:
            with torch.cuda.stream(stream):
                return self.run(n_steps, wave, record_every, path, steps_per_launch)
        lib = _lib.load()
        f64 = self.dtype == torch.float64
        wave_t = None
        if wave is not None:
            wave_t = torch.as_tensor(np.asarray(wave), dtype=self.dtype, device=self.device)
            if wave_t.numel() < self.t_step + n_steps:
                raise ValueError("waveform shorter than the run")
        rec = None
        rec_base = 0
        if record_every > 0:
            n_rec = (self.t_step + n_steps) // record_every - self.t_step // record_every
            rec_base = self.t_step // record_every
            rec = torch.empty((n_rec, self.total), dtype=self.dtype, device=self.device)
        caller = torch.cuda.current_stream(self.device)
        strm = caller.cuda_stream
        if path == "persistent":
            fn = lib.dn_hh_tree_run_f64 if f64 else lib.dn_hh_tree_run_f32
            chunk = steps_per_launch or n_steps
            done = 0
            while done < n_steps:
                k = min(chunk, n_steps - done)
                rec_ptr = None
                if rec is not None:
                    rec_ptr = rec.data_ptr() - rec_base * self.total * rec.element_size()
                # a single bucket runs on the caller's stream (no extra queue)
                streams = [caller] if len(self.buckets) == 1 else self._bucket_streams()
                start = torch.cuda.Event()
                start.record(caller)
                for b, bs in zip(self.buckets, streams):
                    if bs is not caller:
                        bs.wait_event(start)
                    st = fn(
                        *self._plan_args(), _lib.ptr(b.order), b.n, _lib.ptr(b.grp_off),
                        b.n_groups, b.layout, b.kmax, b.threads,
                        _lib.ptr(self.v), _lib.ptr(self.m), _lib.ptr(self.h),
                        _lib.ptr(self.n), _lib.ptr(self.area), _lib.ptr(self.cm_dt),
                        _lib.ptr(self.g), _lib.ptr(self.drive), _lib.ptr(self.sna),
                        _lib.ptr(self.sk), _lib.ptr(wave_t), self.t_step, k, self.dt,
                        self.hh_arr.ctypes.data, rec_ptr, max(record_every, 1),
                        self.total, _lib.ptr(self.status), bs.cuda_stream,
                    )
                    _lib.check(st, "persistent tree run")
                    if bs is not caller:
                        done_ev = torch.cuda.Event()
                        done_ev.record(bs)
                        caller.wait_event(done_ev)
                self.t_step += k
                done += k
        elif path == "plain":
            fn = lib.dn_hh_tree_step_f64 if f64 else lib.dn_hh_tree_step_f32
            mainv = torch.empty_like(self.v)
            rhs = torch.empty_like(self.v)
            w_host = None if wave is None else np.asarray(wave, dtype=np.float64)
            for s in range(n_steps):
                w = 0.0 if w_host is None else float(w_host[self.t_step])
                st = fn(
                    *self._plan_args(), _lib.ptr(self.node_tree), self.kmax, self.total,
                    _lib.ptr(self.v), _lib.ptr(self.m), _lib.ptr(self.h), _lib.ptr(self.n),
                    _lib.ptr(self.area), _lib.ptr(self.cm_dt), _lib.ptr(self.g),
                    _lib.ptr(self.drive), _lib.ptr(self.sna), _lib.ptr(self.sk), w,
                    self.dt, self.hh_arr.ctypes.data, None, None, None, None,
                    _lib.ptr(mainv), _lib.ptr(rhs), strm,
                )
                _lib.check(st, "plain tree step")
                self.t_step += 1
                if rec is not None and self.t_step % record_every == 0:
                    rec[self.t_step // record_every - 1 - rec_base].copy_(self.v)
        else:
            raise ValueError(f"unknown path {path!r}")
        return rec
