"""Host-side tree preprocessing (integer work, done once per morphology).

* ``build_morphology`` / ``build_dhs_layers`` restate the reference's
  preprocessing (dendra/models/integrators/tree.py:414-474,615-673): parent
  array validation, BFS depth, and depth layers deepest -> root chunked to at
  most ``threads`` nodes.  They feed the drop-in ``dhs_solve``.
* ``section_plan`` builds the section-chain layout the fused tree kernels
  use: maximal unbranched chains renumbered contiguously, sections ordered by
  their depth in the section tree.
"""

from __future__ import annotations

from collections import deque
from dataclasses import dataclass

import numpy as np


def build_morphology(parent_idx):
    """Return (parent int32[K], children lists, depth int32[K]).

    Raises like the reference on an empty input, several roots, out-of-range
    or self parents, cycles and unreachable nodes.
    """
    parent_idx = [int(p) for p in parent_idx]
    K = len(parent_idx)
    if K == 0:
        raise ValueError("parent_idx must contain at least one node")
    roots = [i for i, p in enumerate(parent_idx) if p == -1]
    if len(roots) != 1:
        raise ValueError(f"parent_idx must contain exactly one root; found {len(roots)}")
    children = [[] for _ in range(K)]
    for i, p in enumerate(parent_idx):
        if p == -1:
            continue
        if p < 0 or p >= K:
            raise ValueError(f"parent index at node {i} must be in [0, {K - 1}]; got {p}")
        if p == i:
            raise ValueError(f"node {i} cannot be its own parent")
        children[p].append(i)
    depth = np.zeros(K, dtype=np.int32)
    seen = np.zeros(K, dtype=bool)
    q = deque([roots[0]])
    while q:
        u = q.popleft()
        if seen[u]:
            raise ValueError("parent_idx contains a cycle")
        seen[u] = True
        for c in children[u]:
            depth[c] = depth[u] + 1
            q.append(c)
    if not seen.all():
        raise ValueError("parent_idx contains a cycle or nodes unreachable from the root")
    return np.asarray(parent_idx, dtype=np.int32), children, depth


def build_dhs_layers(depth, k_threads: int):
    """Elimination order deepest -> root, layers chunked to <= k_threads nodes."""
    if k_threads <= 0:
        raise ValueError("k_threads must be positive")
    depth = np.asarray(depth)
    if depth.ndim != 1 or depth.size == 0:
        raise ValueError("depth must be a non-empty one-dimensional array")
    if (depth < 0).any():
        raise ValueError("depth values must be nonnegative")
    order, layer_ptr = [], [0]
    max_d = int(depth.max())
    bins = [[] for _ in range(max_d + 1)]
    for i, d in enumerate(depth.tolist()):
        bins[d].append(i)
    for d in range(max_d, -1, -1):
        bucket = bins[d]
        for s in range(0, len(bucket), k_threads):
            order.extend(bucket[s : s + k_threads])
            layer_ptr.append(len(order))
    return np.asarray(order, dtype=np.int32), np.asarray(layer_ptr, dtype=np.int32)


SEC_MAX = 13  # longest section the fused tree kernels hold (dn_hh_tree_sec_max)


@dataclass
class SectionPlan:
    """Section-chain numbering of one tree.

    perm[new] = old node index; inv[old] = new.  A section is a chain of at
    most ``SEC_MAX`` nodes in which every node but the last has exactly one
    child; sections are contiguous head..tail in the new numbering, listed
    level by level (breadth-first over the section tree), and the child
    sections of a section are contiguous.
    """

    perm: np.ndarray
    inv: np.ndarray
    parent: np.ndarray      # new numbering, -1 at the root
    ch_ptr: np.ndarray      # CSR over new numbering
    ch_idx: np.ndarray
    sec_head: np.ndarray
    sec_len: np.ndarray
    sec_child: np.ndarray   # first child section
    sec_nchild: np.ndarray
    sec_ptail: np.ndarray   # parent node of the head (new numbering), -1 at root
    sec_parent: np.ndarray  # parent section, -1 at the root
    lvl_ptr: np.ndarray     # section offsets per level

    @property
    def n_levels(self) -> int:
        return len(self.lvl_ptr) - 1

    @property
    def sec_tail(self) -> np.ndarray:
        return self.sec_head + self.sec_len - 1


def section_plan(parent_idx, sec_max: int = SEC_MAX) -> SectionPlan:
    parent, children, _ = build_morphology(parent_idx)
    K = len(parent)
    root = int(np.flatnonzero(parent == -1)[0])
    nchild = np.array([len(c) for c in children])
    sections = []   # (level, [old nodes], parent section)
    q = deque([(root, 0, -1)])
    while q:
        head, lvl, psec = q.popleft()
        nodes = [head]
        cur = head
        while nchild[cur] == 1 and len(nodes) < sec_max:
            cur = children[cur][0]
            nodes.append(cur)
        me = len(sections)
        sections.append((lvl, nodes, psec))
        if nchild[cur] == 1:  # split chain: the continuation is the only child
            q.append((children[cur][0], lvl + 1, me))
        else:
            for c in children[cur]:
                q.append((c, lvl + 1, me))
    perm = np.concatenate([np.asarray(nodes, dtype=np.int64) for _, nodes, _ in sections])
    inv = np.empty(K, dtype=np.int64)
    inv[perm] = np.arange(K)
    new_parent = np.where(parent[perm] >= 0, inv[np.maximum(parent[perm], 0)], -1)
    n_sec = len(sections)
    sec_head = np.zeros(n_sec, dtype=np.int64)
    sec_len = np.zeros(n_sec, dtype=np.int64)
    sec_child = np.zeros(n_sec, dtype=np.int64)
    sec_nchild = np.zeros(n_sec, dtype=np.int64)
    sec_ptail = np.full(n_sec, -1, dtype=np.int64)
    sec_parent = np.full(n_sec, -1, dtype=np.int64)
    levels = np.zeros(n_sec, dtype=np.int64)
    pos = 0
    for s_i, (lvl, nodes, psec) in enumerate(sections):
        sec_head[s_i] = pos
        sec_len[s_i] = len(nodes)
        levels[s_i] = lvl
        pos += len(nodes)
    for s_i, (_, _, psec) in enumerate(sections):
        if psec >= 0:
            if sec_nchild[psec] == 0:
                sec_child[psec] = s_i
            sec_nchild[psec] += 1
            sec_ptail[s_i] = sec_head[psec] + sec_len[psec] - 1
            sec_parent[s_i] = psec
    for s_i in range(n_sec):  # childless sections point past themselves
        if sec_nchild[s_i] == 0:
            sec_child[s_i] = 0
    assert (np.diff(levels) >= 0).all()
    n_lvl = int(levels.max()) + 1
    lvl_ptr = np.searchsorted(levels, np.arange(n_lvl + 1), side="left")
    order = np.argsort(new_parent, kind="stable")
    order = order[new_parent[order] >= 0]
    counts = np.bincount(new_parent[order], minlength=K)
    ch_ptr = np.zeros(K + 1, dtype=np.int64)
    ch_ptr[1:] = np.cumsum(counts)
    return SectionPlan(
        perm=perm, inv=inv, parent=new_parent.astype(np.int64), ch_ptr=ch_ptr,
        ch_idx=order.astype(np.int64), sec_head=sec_head, sec_len=sec_len,
        sec_child=sec_child, sec_nchild=sec_nchild, sec_ptail=sec_ptail,
        sec_parent=sec_parent,
        lvl_ptr=lvl_ptr.astype(np.int64),
    )


def multi_dhs_plan(parents, rows, threads: int = 16):
    """Packed plan of the multi-morphology tree solve for morphologies in
    solver order (parent < child): one entry of ``parents`` per morphology,
    ``rows[g]`` consecutive plane rows for morphology g (in order), each plan
    row covering up to 32 // threads of them (the reference's neurons per
    warp, _dhs_multi._get_tiled_plan, integrators/tree.py:2017-2084).
    Returns the tables as int64 numpy arrays with K_stride and L_max."""
    P, O, LP, p_off, o_off, l_off, nl, row0, rcount = ([] for _ in range(9))
    per_warp = max(1, 32 // threads)
    row = 0
    for par, n in zip(parents, rows):
        par = np.asarray(par, dtype=np.int64)
        _, _, depth = build_morphology(par)
        order, lptr = build_dhs_layers(depth, threads)
        base_p, base_l = len(P), len(LP)
        P.extend(par.tolist())
        O.extend(order.tolist())
        LP.extend(lptr.tolist())
        for k in range(0, int(n), per_warp):
            p_off.append(base_p)
            o_off.append(base_p)
            l_off.append(base_l)
            nl.append(len(lptr) - 1)
            row0.append(row + k)
            rcount.append(min(per_warp, int(n) - k))
        row += int(n)
    a = lambda v: np.asarray(v, dtype=np.int64)
    return dict(P_cat=a(P), ORDER_cat=a(O), LAYER_PTR_cat=a(LP), WARP_P_OFF=a(p_off),
                WARP_ORDER_OFF=a(o_off), WARP_LPTR_OFF=a(l_off), WARP_L=a(nl),
                WARP_ROW_BASE=a(row0), WARP_ROW_COUNT=a(rcount),
                K_stride=max(len(p) for p in parents), L_max=int(max(nl)),
                grid_x=len(row0), rows_total=row)
