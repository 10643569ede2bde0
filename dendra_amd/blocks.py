"""3x3 block solvers of the double-layer extracellular (ExtCell) model.

Drop-ins for the reference's block solvers, with the same shapes, argument
checks, non-mutating inputs, fresh outputs and autograd contracts:

* ``thomas_solve_cuda_bt(lower, main, upper, rhs)`` / ``thomas_bt3_solve``
  (dendra/models/integrators/triton/bt_kernel.py:192-298): main (B,K,3,3),
  diagonal coupling blocks lower/upper (B,K-1,3), rhs (B,K,3).  Backward: the
  transposed system (upper and lower exchanged, main blocks transposed) and
  ``grad_main = -(g x^T)``, ``grad_lower = -(g[1:] x[:-1])``,
  ``grad_upper = -(g[:-1] x[1:])`` (bt_kernel.py:266-292).
* ``block_spd_solve(lower, main, upper, rhs)`` and the low-level
  ``solve_bt_spd_fwd`` / ``solve_bt_spd_with_chol`` (triton/bt_spd_kernel.py:
  575-700): block Cholesky; the backward reuses the factors and returns the
  symmetric part of ``-(g x^T)`` for main (integrators/_cuda_functional.py:
  306-333).  A block still indefinite after two diagonal shifts raises.
* ``dhs_bt_solve(main, edge, rhs, parent_idx, order, layer_ptr, threads)``
  (triton/dhs_kernel_bt.py): tree with 3x3 node blocks and per-component edge
  conductances (``A_ii = main_i + diag(e_i) + sum_c diag(e_c)``,
  ``A_ip = -diag(e_i)``); backward: the same solve with transposed main
  blocks, ``grad_main = -(g x^T)``, ``grad_edge = -(g_i - g_p)(x_i - x_p)``
  per component (``-g x`` at the root) (_cuda_functional.py:365-374).
"""

from __future__ import annotations

import numpy as np
import torch
from torch.utils.weak import WeakIdKeyDictionary

from . import _lib
from .solvers import validate_threads

_FLOAT = (torch.float32, torch.float64)
_INDEX = (torch.int32, torch.int64)
KIND_THOMAS, KIND_SPD, KIND_TREE = 0, 1, 2


def _float_group(named, what="dendra_amd block solvers"):
    first = named[0][1]
    for name, t in named:
        if t.device != first.device:
            raise ValueError(f"{name} must be on device {first.device}; got {t.device}")
        if t.dtype != first.dtype:
            raise TypeError(f"{name} must have dtype {first.dtype}; got {t.dtype}")
    if first.device.type != "cuda":
        raise ValueError(f"{what} require GPU tensors")
    if first.dtype not in _FLOAT:
        raise TypeError(f"{what} support only float32 and float64")


def validate_block_tridiagonal(lower, main, upper, rhs):
    """Contract of the reference's validate_block_tridiagonal
    (triton/_contracts.py:162-196)."""
    named = (("lower", lower), ("main", main), ("upper", upper), ("rhs", rhs))
    for name, t in named:
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
    if rhs.ndim != 3 or rhs.shape[-1] != 3:
        raise ValueError(f"rhs must have shape (B, K, 3), got {tuple(rhs.shape)}")
    batch, size, _ = map(int, rhs.shape)
    if batch < 1 or size < 1:
        raise ValueError("B and K must both be >= 1")
    band = (batch, size - 1, 3)
    if tuple(lower.shape) != band:
        raise ValueError(f"lower must have shape {band}, got {tuple(lower.shape)}")
    if tuple(upper.shape) != band:
        raise ValueError(f"upper must have shape {band}, got {tuple(upper.shape)}")
    if tuple(main.shape) != (batch, size, 3, 3):
        raise ValueError(f"main must have shape {(batch, size, 3, 3)}, "
                         f"got {tuple(main.shape)}")
    _float_group(named)
    return batch, size


def validate_tree_block(main, edge, rhs, parent_idx, order, layer_ptr, threads):
    """Contract of the reference's validate_tree_block
    (triton/_contracts.py:232-268)."""
    data = (("main", main), ("edge", edge), ("rhs", rhs))
    topo = (("parent_idx", parent_idx), ("order", order), ("layer_ptr", layer_ptr))
    for name, t in data + topo:
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
    if rhs.ndim != 3 or rhs.shape[-1] != 3:
        raise ValueError(f"rhs must have shape (B, K, 3), got {tuple(rhs.shape)}")
    batch, size, _ = map(int, rhs.shape)
    if batch < 1 or size < 1:
        raise ValueError("B and K must both be >= 1")
    if tuple(edge.shape) != (batch, size, 3):
        raise ValueError(f"edge must have shape {(batch, size, 3)}, got {tuple(edge.shape)}")
    if tuple(main.shape) != (batch, size, 3, 3):
        raise ValueError(f"main must have shape {(batch, size, 3, 3)}, "
                         f"got {tuple(main.shape)}")
    _float_group(data)
    for name, t in topo:
        if t.ndim != 1:
            raise ValueError(f"{name} must be 1-D")
        if t.dtype not in _INDEX:
            raise TypeError(f"{name} must have dtype int32 or int64")
        if t.device != rhs.device:
            raise ValueError(f"{name} must be on device {rhs.device}; got {t.device}")
    if parent_idx.numel() != size or order.numel() != size:
        raise ValueError("parent_idx and order must each contain K entries")
    if layer_ptr.numel() < 2:
        raise ValueError("layer_ptr must contain at least two entries")
    return batch, size, validate_threads(threads)


def _work(kind, B, K, like):
    n = int(_lib.load().dn_block_work_elems(kind, B, K))
    return torch.empty(max(n, 1), dtype=like.dtype, device=like.device), n


def _fn(name, like):
    return getattr(_lib.load(), f"{name}_{'f64' if like.dtype == torch.float64 else 'f32'}")


def _bt3_raw(lower, main, upper, rhs):
    lower, main, upper, rhs = (t.contiguous() for t in (lower, main, upper, rhs))
    B, K = map(int, rhs.shape[:2])
    x = torch.empty_like(rhs, memory_format=torch.contiguous_format)
    work, n = _work(KIND_THOMAS, B, K, rhs)
    p = _lib.ptr
    st = _fn("dn_bt3_thomas", rhs)(p(lower), p(main), p(upper), p(rhs), p(x), B, K,
                                   p(work), n, _lib.stream_handle(rhs.device))
    _lib.check(st, "block Thomas solve")
    return x


def solve_bt_spd_fwd(lower, main, upper, rhs):
    """Factor and solve; returns (x, chol).  Raises when a block stays
    indefinite after the diagonal shifts (bt_spd_kernel.py:571-581)."""
    lower, main, upper, rhs = (t.contiguous() for t in (lower, main, upper, rhs))
    B, K = map(int, rhs.shape[:2])
    x = torch.empty_like(rhs, memory_format=torch.contiguous_format)
    chol = torch.empty(B, K, 9, dtype=rhs.dtype, device=rhs.device)
    info = torch.empty(B, dtype=torch.int32, device=rhs.device)
    work, n = _work(KIND_SPD, B, K, rhs)
    p = _lib.ptr
    st = _fn("dn_bt3_spd", rhs)(p(lower), p(main), p(upper), p(rhs), p(x), p(chol), p(info),
                                B, K, p(work), n, 1, _lib.stream_handle(rhs.device))
    _lib.check(st, "block SPD solve")
    if int(info.max()) > 0:
        raise RuntimeError("solve_bt_spd_cuda: Cholesky failed after all diagonal bumps")
    return x, chol


def solve_bt_spd_with_chol(lower, upper, chol, rhs):
    """Solve with the factors of a previous solve_bt_spd_fwd."""
    lower, upper, chol, rhs = (t.contiguous() for t in (lower, upper, chol, rhs))
    B, K = map(int, rhs.shape[:2])
    if tuple(chol.shape) != (B, K, 9):
        raise ValueError(f"chol must have shape {(B, K, 9)}, got {tuple(chol.shape)}")
    x = torch.empty_like(rhs, memory_format=torch.contiguous_format)
    work, n = _work(KIND_SPD, B, K, rhs)
    p = _lib.ptr
    st = _fn("dn_bt3_spd", rhs)(p(lower), None, p(upper), p(rhs), p(x), p(chol), None, B, K,
                                p(work), n, 0, _lib.stream_handle(rhs.device))
    _lib.check(st, "block SPD solve with factors")
    return x


# Children CSR per parent tensor object (weakly held, so reused storage never
# inherits another tensor's plan), checked against the tensor's version.
_children = WeakIdKeyDictionary()


def _children_csr(parent):
    stamp = (parent._version, parent.numel(), parent.device)
    hit = _children.get(parent)
    if hit is not None and hit[0] == stamp:
        return hit[1]
    par = parent.detach().to("cpu", torch.int64).numpy().reshape(-1)
    K = par.size
    par = np.where((par >= 0) & (par < K), par, -1)
    kids = np.flatnonzero(par >= 0)
    order = kids[np.argsort(par[kids], kind="stable")].astype(np.int32)
    ptr = np.zeros(K + 1, dtype=np.int32)
    np.cumsum(np.bincount(par[kids], minlength=K), out=ptr[1:])
    dev = parent.device
    csr = (torch.as_tensor(ptr, device=dev),
           torch.as_tensor(order if order.size else np.zeros(1, np.int32), device=dev))
    _children[parent] = (stamp, csr)
    return csr


_levels = WeakIdKeyDictionary()


def _depth_levels(parent):
    """Elimination order deepest -> root in whole depth levels (int64 order,
    layer_ptr on the parent's device), cached per parent tensor; None when
    the parent array is not a forest."""
    stamp = (parent._version, parent.numel(), parent.device)
    hit = _levels.get(parent)
    if hit is not None and hit[0] == stamp:
        return hit[1]
    par = parent.detach().to("cpu", torch.int64).numpy().reshape(-1)
    K = par.size
    par = np.where((par >= 0) & (par < K), par, -1)
    # depth by pointer jumping: after round r, anc is the 2^r-th ancestor and
    # depth counts the edges passed; a pointer still live after
    # ceil(log2 K) + 1 rounds means a cycle
    depth = (par >= 0).astype(np.int64)
    anc = par.copy()
    for _ in range(max(1, int(K).bit_length()) + 1):
        live = anc >= 0
        if not live.any():
            break
        a = anc[live]
        depth[live] += depth[a]
        anc[live] = anc[a]
    if (anc >= 0).any():
        depth[:] = -1
    d = int(depth.max()) + 1 if K else 0
    res = None
    if (depth >= 0).all():
        order = np.argsort(-depth, kind="stable")  # deepest first, ascending index within a level
        layer_ptr = np.concatenate([[0], np.cumsum(np.bincount(depth, minlength=d)[::-1])])
        dev = parent.device
        res = (torch.as_tensor(order, dtype=torch.int64, device=dev),
               torch.as_tensor(layer_ptr, dtype=torch.int64, device=dev))
    _levels[parent] = (stamp, res)
    return res


# Contraction plans per parent tensor object (version-checked, weakly held):
# the block tree solve then runs as chain condensation plus rake/compress
# rounds (its dependent chain is the longest section plus twice the rounds,
# not the tree's depth); None when the parents do not form one tree.
_bc_plans = WeakIdKeyDictionary()


def _contraction_plan(parent):
    from .contraction import TreeContractionPlan

    stamp = (parent._version, parent.numel(), parent.device)
    hit = _bc_plans.get(parent)
    if hit is not None and hit[0] == stamp:
        return hit[1]
    try:
        plan = TreeContractionPlan(parent.detach().to("cpu", torch.int64), 16)
    except ValueError:
        plan = None
    _bc_plans[parent] = (stamp, plan)
    return plan


def _contract_cap(elem_bytes: int, K: int) -> int:
    return int(_lib.load().dn_bt3_contract_max_sections(int(elem_bytes), int(K)))


def contraction_tables(plan, device):
    """The contraction plan's int32 tables in the block entry points' order."""
    from .contraction import plan_tables

    return plan_tables(plan, device)


def _tree_contract_raw(main, edge, rhs, plan):
    main, edge, rhs = (t.contiguous() for t in (main, edge, rhs))
    (perm, sec_head, sec_len, sec_child, sec_nchild, sec_parent,
     ptr, sec, par, child, rptr, recv, sptr, src) = contraction_tables(plan, rhs.device)
    B, K = map(int, rhs.shape[:2])
    x = torch.empty_like(rhs, memory_format=torch.contiguous_format)
    n = int(_lib.load().dn_bt3_contract_work_elems(B, K))
    work = torch.empty(max(n, 1), dtype=rhs.dtype, device=rhs.device)
    p = _lib.ptr
    st = _fn("dn_bt3_tree_contract", rhs)(
        p(main), p(edge), p(rhs), None, p(x), B, K, p(perm), plan.n_sections, p(sec_head),
        p(sec_len), p(sec_child), p(sec_nchild), p(sec_parent), plan.n_rounds, p(ptr), p(sec),
        p(par), p(child), p(rptr), p(recv), p(sptr), p(src), p(work), n,
        _lib.stream_handle(rhs.device))
    _lib.check(st, "block tree contraction solve")
    return x


def _tree_raw(main, edge, rhs, parent, order, layer_ptr):
    plan = _contraction_plan(parent)
    if plan is not None and plan.n_sections <= _contract_cap(rhs.element_size(), plan.size):
        return _tree_contract_raw(main, edge, rhs, plan)
    return _tree_layers_raw(main, edge, rhs, parent, order, layer_ptr)


def _tree_layers_raw(main, edge, rhs, parent, order, layer_ptr):
    main, edge, rhs = (t.contiguous() for t in (main, edge, rhs))
    child_ptr, child_idx = _children_csr(parent)
    # whole depth levels let the kernel pass the children's terms through
    # LDS; the solve's results do not depend on the layering
    lv = _depth_levels(parent)
    if lv is not None:
        order, layer_ptr = lv
    parent, order, layer_ptr = (t.to(torch.int64).contiguous()
                                for t in (parent, order, layer_ptr))
    B, K = map(int, rhs.shape[:2])
    x = torch.empty_like(rhs, memory_format=torch.contiguous_format)
    work, n = _work(KIND_TREE, B, K, rhs)
    p = _lib.ptr
    st = _fn("dn_bt3_tree", rhs)(p(main), p(edge), p(rhs), p(parent), p(order), p(layer_ptr),
                                 layer_ptr.numel() - 1, p(child_ptr), p(child_idx), p(x), B,
                                 K, p(work), n, _lib.stream_handle(rhs.device))
    _lib.check(st, "block tree solve")
    return x


def _sym(m):
    return 0.5 * (m + m.transpose(-1, -2))


class _BT3Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lower, main, upper, rhs):
        x = _bt3_raw(lower, main, upper, rhs)
        ctx.save_for_backward(lower, main, upper, x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        lower, main, upper, x = ctx.saved_tensors
        g = _bt3_raw(upper, main.transpose(-1, -2), lower, grad_x)
        return (-(g[:, 1:] * x[:, :-1]), -(g.unsqueeze(-1) * x.unsqueeze(-2)),
                -(g[:, :-1] * x[:, 1:]), g)


class _SPDSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lower, main, upper, rhs):
        x, chol = solve_bt_spd_fwd(lower, main, upper, rhs)
        ctx.save_for_backward(lower, upper, x, chol)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        lower, upper, x, chol = ctx.saved_tensors
        g = solve_bt_spd_with_chol(lower, upper, chol, grad_x)
        return (-(g[:, 1:] * x[:, :-1]), -_sym(g.unsqueeze(-1) * x.unsqueeze(-2)),
                -(g[:, :-1] * x[:, 1:]), g)


class _TreeBT3Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, main, edge, rhs, parent, order, layer_ptr):
        x = _tree_raw(main, edge, rhs, parent, order, layer_ptr)
        ctx.save_for_backward(main, edge, x, parent, order, layer_ptr)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        main, edge, x, parent, order, layer_ptr = ctx.saved_tensors
        g = _tree_raw(main.transpose(-1, -2), edge, grad_x, parent, order, layer_ptr)
        par = parent.to(torch.long)
        root = (par < 0).reshape(1, -1, 1)
        safe = par.clamp_min(0)
        ge = torch.where(root, -(g * x),
                         -((g - g.index_select(1, safe)) * (x - x.index_select(1, safe))))
        return -(g.unsqueeze(-1) * x.unsqueeze(-2)), ge, g, None, None, None


def thomas_solve_cuda_bt(lower, main, upper, rhs):
    """Block Thomas solve of a 3x3-block tridiagonal system."""
    validate_block_tridiagonal(lower, main, upper, rhs)
    return _BT3Solve.apply(lower, main, upper, rhs)


thomas_bt3_solve = thomas_solve_cuda_bt


def block_spd_solve(lower, main, upper, rhs):
    """Block Cholesky solve of a symmetric positive definite 3x3-block
    tridiagonal system (main blocks are symmetrised)."""
    validate_block_tridiagonal(lower, main, upper, rhs)
    return _SPDSolve.apply(lower, main, upper, rhs)


def dhs_bt_solve(main, edge, rhs, parent_idx, order, layer_ptr, threads: int = 16):
    """Block tree (Hines) solve with 3x3 node blocks over the caller's DHS
    order (children before parents, as build_dhs_layers produces)."""
    validate_tree_block(main, edge, rhs, parent_idx, order, layer_ptr, threads)
    return _TreeBT3Solve.apply(main, edge, rhs, parent_idx, order, layer_ptr)
