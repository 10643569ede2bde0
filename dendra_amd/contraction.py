"""Tree solve by contraction, behind the reference's tree-contraction facade.

Drop-in for ``dendra/models/integrators/triton/tree_contraction.py``:
``TreeContractionPlan(parent, max_segment_interior_nodes=16)`` (the plan the
reference takes from dendra-solvers, tree.py:841-867), ``prepare(plan,
device)`` (tree_contraction.py:651-656) and ``solve(d, g, rhs, plan)``
(:748-770) with the same shapes, validation messages, fresh output and
autograd contract: the matrix is symmetric, so the adjoint is the same solve,
and ``grad_d = -y x``, ``grad_g = -(y_i - y_p)(x_i - x_p)`` (``-y x`` at the
root, whose g is a conductance to ground), ``grad_rhs = y``
(tree_contraction.py:596-609).

The plan is this package's own (the reference's comes from the absent
dendra-solvers wheel):

* sections: chains of at most ``max_segment_interior_nodes + 1`` nodes in
  which every node but the last has one child (``None``: one node per
  section, i.e. fine-grained contraction), numbered breadth-first;
* a rake/compress schedule over the section tree: each round eliminates
  every current leaf and alternate one-child sections whose child is not a
  leaf (an independent set), and lists, for every survivor touched, the
  eliminated neighbours it pulls from, in ascending order.

The device kernel (``csrc/dn_contract.hip``) runs one workgroup per system.
"""

from __future__ import annotations

from dataclasses import dataclass
from numbers import Integral

import numpy as np
import torch

from . import _lib
from .topology import build_morphology, section_plan


@dataclass
class ContractionSchedule:
    """Rake/compress rounds over a tree given by ``parent`` (root 0, and
    ``parent[i] < i``).  Round r eliminates ``sec[ptr[r]:ptr[r+1]]`` with
    their current parent ``par`` and, for a compress, their only child
    ``child`` (-1 for a rake); survivors ``recv[rptr[r]:rptr[r+1]]`` pull
    from ``src[sptr[i]:sptr[i+1]]`` (j >= 0: eliminated child j; -1-j:
    compressed parent j)."""

    ptr: np.ndarray
    sec: np.ndarray
    par: np.ndarray
    child: np.ndarray
    rptr: np.ndarray
    recv: np.ndarray
    sptr: np.ndarray
    src: np.ndarray

    @property
    def n_rounds(self) -> int:
        return len(self.ptr) - 1


def contraction_schedule(parent) -> ContractionSchedule:
    parent = np.asarray(parent, dtype=np.int64)
    n = len(parent)
    if n == 0 or parent[0] != -1 or (n > 1 and (parent[1:] < 0).any()):
        raise ValueError("schedule needs the root at 0 and one parent per other node")
    if n > 1 and (parent[1:] >= np.arange(1, n)).any():
        raise ValueError("schedule needs parent[i] < i")
    idx = np.arange(n)
    par = parent.copy()
    alive = np.ones(n, dtype=bool)
    ptr, sec, epar, echild = [0], [], [], []
    rptr, recv, sptr, src = [0], [], [0], []
    while alive[1:].any():
        live = alive & (idx != 0)
        nch = np.bincount(par[live], minlength=n)
        one = np.full(n, -1, dtype=np.int64)
        one[par[live]] = idx[live]          # the child, where there is exactly one
        leaf = live & (nch == 0)
        cand = live & (nch == 1)
        cand[cand] = ~leaf[one[cand]]       # a compressed node's child survives
        picked = np.zeros(n, dtype=bool)
        for s in np.flatnonzero(cand):      # parents come first: alternate down chains
            picked[s] = not picked[par[s]]
        gone = np.flatnonzero(leaf | picked)
        kids = np.where(picked[gone], one[gone], -1)
        ptr.append(ptr[-1] + len(gone))
        sec.extend(gone.tolist())
        epar.extend(par[gone].tolist())
        echild.extend(kids.tolist())
        pulls = [(int(par[j]), int(j)) for j in gone]
        pulls += [(int(c), -1 - int(j)) for j, c in zip(gone, kids) if c >= 0]
        pulls.sort(key=lambda t: (t[0], t[1] if t[1] >= 0 else n + (-1 - t[1])))
        last = None
        for q, s_ in pulls:
            if q != last:
                if last is not None:
                    sptr.append(len(src))
                recv.append(q)
                last = q
            src.append(s_)
        if last is not None:
            sptr.append(len(src))
        rptr.append(len(recv))
        for j, c in zip(gone, kids):
            if c >= 0:
                par[c] = par[j]
        alive[gone] = False
    i32 = lambda a: np.asarray(a, dtype=np.int32)
    return ContractionSchedule(i32(ptr), i32(sec), i32(epar), i32(echild), i32(rptr),
                               i32(recv), i32(sptr), i32(src))


def _validate_segment(max_segment_interior_nodes):
    """Same checks as the reference's factory (integrators/__init__.py:93-103)."""
    if max_segment_interior_nodes is None:
        return None
    if isinstance(max_segment_interior_nodes, bool) or not isinstance(
        max_segment_interior_nodes, Integral
    ):
        raise TypeError("max_segment_interior_nodes must be a positive integer or None")
    if max_segment_interior_nodes <= 0:
        raise ValueError("max_segment_interior_nodes must be positive")
    return int(max_segment_interior_nodes)


class TreeContractionPlan:
    """Immutable contraction plan of one tree topology (``parent[i]`` is the
    parent of node i in the caller's node order, -1 at the root)."""

    def __init__(self, parent, max_segment_interior_nodes=16):
        self.max_segment_interior_nodes = _validate_segment(max_segment_interior_nodes)
        par = np.asarray(parent.detach().cpu() if isinstance(parent, torch.Tensor)
                         else parent, dtype=np.int64).reshape(-1)
        build_morphology(par)  # raises on an invalid tree
        self.size = int(len(par))
        sec_max = 1 if self.max_segment_interior_nodes is None else \
            self.max_segment_interior_nodes + 1
        sp = section_plan(par, sec_max=sec_max)
        self.n_sections = int(len(sp.sec_head))
        self.schedule = contraction_schedule(sp.sec_parent)
        self.parent = par
        self._host = {
            "perm": sp.perm, "sec_head": sp.sec_head, "sec_len": sp.sec_len,
            "sec_child": sp.sec_child, "sec_nchild": sp.sec_nchild,
            "sec_parent": sp.sec_parent,
        }
        self._devices: dict = {}
        self._caps = None

    @property
    def n_rounds(self) -> int:
        return self.schedule.n_rounds

    def prepare(self, device):
        device = torch.device(device)
        if device.type != "cuda":
            raise ValueError("CUDA tree contraction requires a CUDA device")
        if device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.section_cap(4)
        return self.tables_on(device)

    def section_cap(self, elem_bytes: int) -> int:
        """Most sections one system may have at this element size (the
        kernel's LDS bound); read from the library once per plan."""
        if self._caps is None:
            lib = _lib.load()
            self._caps = {e: int(lib.dn_tree_contract_max_sections(e)) for e in (4, 8)}
        return self._caps[elem_bytes]

    def tables_on(self, device):
        """The plan's int32 tables (and the int64 parent) on any device."""
        device = torch.device(device)
        if device not in self._devices:
            i32 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int32), device=device)
            sc = self.schedule
            t = {k: i32(v) for k, v in self._host.items()}
            for k in ("ptr", "sec", "par", "child", "rptr", "recv", "sptr", "src"):
                v = getattr(sc, k)
                t[k] = i32(v if len(v) else np.zeros(1))
            par = torch.as_tensor(self.parent, dtype=torch.int64, device=device)
            t["parent_i64"] = par
            t["is_root"] = par < 0
            t["parent_safe"] = torch.where(t["is_root"], torch.zeros_like(par), par)
            self._devices[device] = t
        return self._devices[device]


def prepare(plan: TreeContractionPlan, device):
    """Upload the plan's tables to ``device`` (tree_contraction.py:651-656)."""
    return plan.prepare(device)


TABLE_NAMES = ("perm", "sec_head", "sec_len", "sec_child", "sec_nchild", "sec_parent",
               "ptr", "sec", "par", "child", "rptr", "recv", "sptr", "src")


def plan_tables(plan: TreeContractionPlan, device) -> tuple:
    """The plan's device tables in the entry point's argument order."""
    t = plan.prepare(device)
    return tuple(t[k] for k in TABLE_NAMES)


def contract_tables(d, g, rhs, tables, n_sec: int, n_rounds: int):
    """One launch of the contraction kernel on explicit plan tables (the
    opaque operator of dendra_amd.functional calls this)."""
    lib = _lib.load()
    (perm, sec_head, sec_len, sec_child, sec_nchild, sec_parent,
     ptr, sec, par, child, rptr, recv, sptr, src) = (v.contiguous() for v in tables)
    d, g, rhs = (v.contiguous() for v in (d, g, rhs))
    B, K = map(int, d.shape)
    x = torch.empty_like(rhs, memory_format=torch.contiguous_format)
    n_work = int(lib.dn_tree_contract_work_elems(B, K))
    work = torch.empty(max(n_work, 1), dtype=rhs.dtype, device=rhs.device)
    f64 = rhs.dtype == torch.float64
    fn = lib.dn_tree_contract_f64 if f64 else lib.dn_tree_contract_f32
    p = _lib.ptr
    st = fn(p(d), p(g), p(rhs), p(x), B, K, p(perm), int(n_sec), p(sec_head), p(sec_len),
            p(sec_child), p(sec_nchild), p(sec_parent), int(n_rounds), p(ptr), p(sec),
            p(par), p(child), p(rptr), p(recv), p(sptr), p(src), p(work), n_work,
            _lib.stream_handle(rhs.device))
    _lib.check(st, "tree contraction solve")
    return x


def _contract_raw(d, g, rhs, plan: TreeContractionPlan):
    return contract_tables(d, g, rhs, plan_tables(plan, d.device), plan.n_sections,
                           plan.n_rounds)


def _gradients(y, x, plan, device, need):
    t = plan.prepare(device)
    gd = -(y * x) if need[0] else None
    gg = None
    if need[1]:
        root, ps = t["is_root"], t["parent_safe"]
        x_drop = torch.where(root, x, x - x.index_select(-1, ps))
        y_drop = torch.where(root, y, y - y.index_select(-1, ps))
        gg = -(x_drop * y_drop)
    return gd, gg, (y if need[2] else None)


class _TreeContraction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, g, rhs, plan):
        x = _contract_raw(d, g, rhs, plan)
        ctx.plan = plan
        ctx.save_for_backward(d, g, x)
        return x

    @staticmethod
    def backward(ctx, cotangent):
        d, g, x = ctx.saved_tensors
        y = _contract_raw(d, g, cotangent.contiguous(), ctx.plan)
        gd, gg, gr = _gradients(y, x, ctx.plan, x.device, ctx.needs_input_grad[:3])
        return gd, gg, gr, None


def _check_inputs(d, g, rhs, plan: TreeContractionPlan):
    """The reference's argument checks (tree_contraction.py:750-766), plus the
    LDS bound on one system's section count."""
    if d.ndim != 2 or d.shape[1] != plan.size:
        raise ValueError("tree contraction inputs must have shape (B, K) matching the plan")
    if d.dtype not in (torch.float32, torch.float64):
        raise ValueError("CUDA tree contraction requires float32 or float64")
    for value in (d, g, rhs):
        if (value.shape != d.shape or value.device != d.device
                or value.device.type != "cuda" or value.dtype != d.dtype
                or value.layout != torch.strided):
            raise ValueError("tree contraction inputs must be same-shaped strided CUDA "
                             "tensors with matching dtypes")
    cap = plan.section_cap(d.element_size())
    if plan.n_sections > cap:
        raise ValueError(f"tree contraction plan has {plan.n_sections} sections; at most "
                         f"{cap} fit one workgroup in {d.dtype}")


def solve(d, g, rhs, plan: TreeContractionPlan):
    """Solve a scalar Hines right-hand side in float32 or float64 on the GPU."""
    _check_inputs(d, g, rhs, plan)
    if d.shape[0] < 1:
        return torch.empty_like(rhs, memory_format=torch.contiguous_format)
    return _TreeContraction.apply(d, g, rhs, plan)


tree_contraction_solve = solve
