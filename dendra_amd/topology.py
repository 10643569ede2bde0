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


@dataclass
class SectionPlan:
    """Section-chain numbering of one tree.

    perm[new] = old node index; inverse inv[old] = new.  Sections are
    contiguous head..tail in the new numbering and listed level by level.
    """

    perm: np.ndarray
    inv: np.ndarray
    parent: np.ndarray  # new numbering, -1 at the root
    ch_ptr: np.ndarray  # CSR over new numbering
    ch_idx: np.ndarray
    sec_head: np.ndarray
    sec_tail: np.ndarray
    lvl_ptr: np.ndarray  # section offsets per level

    @property
    def n_levels(self) -> int:
        return len(self.lvl_ptr) - 1


def section_plan(parent_idx) -> SectionPlan:
    parent, children, _ = build_morphology(parent_idx)
    K = len(parent)
    root = int(np.flatnonzero(parent == -1)[0])
    nchild = np.array([len(c) for c in children])
    # heads: root, or a node whose parent does not have exactly one child
    sections = []  # (level, head_old, [nodes...])
    q = deque([(root, 0)])
    while q:
        head, lvl = q.popleft()
        nodes = [head]
        cur = head
        while nchild[cur] == 1:
            cur = children[cur][0]
            nodes.append(cur)
        sections.append((lvl, nodes))
        for c in children[cur]:
            q.append((c, lvl + 1))
    # BFS already visits sections level by level
    perm = np.concatenate([np.asarray(nodes, dtype=np.int64) for _, nodes in sections])
    inv = np.empty(K, dtype=np.int64)
    inv[perm] = np.arange(K)
    new_parent = np.where(parent[perm] >= 0, inv[np.maximum(parent[perm], 0)], -1)
    sec_head, sec_tail, levels = [], [], []
    pos = 0
    for lvl, nodes in sections:
        sec_head.append(pos)
        sec_tail.append(pos + len(nodes) - 1)
        levels.append(lvl)
        pos += len(nodes)
    n_lvl = max(levels) + 1
    lvl_ptr = np.searchsorted(np.asarray(levels), np.arange(n_lvl + 1), side="left")
    ch_lists = [[] for _ in range(K)]
    for new_i in range(K):
        p = new_parent[new_i]
        if p >= 0:
            ch_lists[p].append(new_i)
    ch_ptr = np.zeros(K + 1, dtype=np.int64)
    ch_ptr[1:] = np.cumsum([len(c) for c in ch_lists])
    ch_idx = np.asarray([c for lst in ch_lists for c in sorted(lst)], dtype=np.int64)
    return SectionPlan(
        perm=perm,
        inv=inv,
        parent=new_parent.astype(np.int64),
        ch_ptr=ch_ptr,
        ch_idx=ch_idx,
        sec_head=np.asarray(sec_head, dtype=np.int64),
        sec_tail=np.asarray(sec_tail, dtype=np.int64),
        lvl_ptr=lvl_ptr.astype(np.int64),
    )
