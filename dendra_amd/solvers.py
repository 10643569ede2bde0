"""Batched tridiagonal and tree solves on gfx950 behind the reference's facades.

Drop-in replacements for the reference's
``thomas_solve_cuda_t`` / ``pcr_solve_cuda_t``
(dendra/models/integrators/triton/t_kernel_thomas.py:218-232,
pcr_kernel_thomas.py:224-226) and ``dhs_solve_cuda``
(triton/dhs_kernel.py:634-651): same argument order, shapes, validation
errors, non-mutating inputs, fresh output tensor, and the same autograd
contract (the transposed solve plus outer-product band gradients,
t_kernel_thomas.py:147-189; the symmetric tree adjoint with
``grad_a = -(y_i - y_p)(x_i - x_p)``, dhs_kernel_multi.py:138-206).
"""

from __future__ import annotations

import torch
from torch.utils.weak import WeakIdKeyDictionary

from . import _lib

_FLOAT = (torch.float32, torch.float64)
_INDEX = (torch.int32, torch.int64)
METHOD_AUTO, METHOD_THOMAS, METHOD_PCR = 0, 1, 2


def _float_group(named):
    first_name, first = named[0]
    if first.device.type != "cuda":
        raise ValueError("dendra_amd solver inputs must be GPU tensors")
    if first.dtype not in _FLOAT:
        raise TypeError("dendra_amd solvers support only float32 and float64")
    for name, t in named[1:]:
        if t.device != first.device:
            raise ValueError(
                f"{name} must be on the same device as {first_name}; "
                f"got {t.device} and {first.device}"
            )
        if t.dtype != first.dtype:
            raise ValueError(
                f"{name} must have the same dtype as {first_name}; "
                f"got {t.dtype} and {first.dtype}"
            )


def validate_tridiagonal(a, b, c, d):
    """Shape/dtype contract of the reference's validate_tridiagonal
    (triton/_contracts.py:137-160)."""
    named = (("a", a), ("b", b), ("c", c), ("d", d))
    for name, t in named:
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
    if b.ndim != 2:
        raise ValueError(f"b must have shape (B, K), got {tuple(b.shape)}")
    batch, size = map(int, b.shape)
    if batch < 1 or size < 1:
        raise ValueError("B and K must both be >= 1")
    band = (batch, size - 1)
    if tuple(a.shape) != band:
        raise ValueError(f"a must have shape {band}, got {tuple(a.shape)}")
    if tuple(c.shape) != band:
        raise ValueError(f"c must have shape {band}, got {tuple(c.shape)}")
    if tuple(d.shape) != (batch, size):
        raise ValueError(f"d must have shape {(batch, size)}, got {tuple(d.shape)}")
    _float_group(named)
    return batch, size


def _tridiag_raw(a, b, c, d, method):
    lib = _lib.load()
    a, b, c, d = (t.contiguous() for t in (a, b, c, d))
    B, K = b.shape
    x = torch.empty_like(d)
    f64 = b.dtype == torch.float64
    n_scr = lib.dn_tridiag_scratch_elems(B, K, method, 8 if f64 else 4)
    scratch = torch.empty(max(int(n_scr), 1), dtype=b.dtype, device=b.device)
    fn = lib.dn_tridiag_solve_f64 if f64 else lib.dn_tridiag_solve_f32
    st = fn(
        _lib.ptr(a), _lib.ptr(b), _lib.ptr(c), _lib.ptr(d), _lib.ptr(x), B, K, method,
        _lib.ptr(scratch), int(n_scr), _lib.stream_handle(b.device),
    )
    _lib.check(st, "tridiagonal solve")
    return x


class _TridiagSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b, c, d, method):
        x = _tridiag_raw(a, b, c, d, method)
        ctx.save_for_backward(a, b, c, x)
        ctx.method = method
        return x

    @staticmethod
    def backward(ctx, grad_x):
        a, b, c, x = ctx.saved_tensors
        need = ctx.needs_input_grad
        if not any(need[:4]):
            return None, None, None, None, None
        # A^T has sub-diagonal c and super-diagonal a.
        y = _tridiag_raw(c, b, a, grad_x, ctx.method)
        ga = -y[:, 1:] * x[:, :-1] if need[0] else None
        gb = -y * x if need[1] else None
        gc = -y[:, :-1] * x[:, 1:] if need[2] else None
        gd = y if need[3] else None
        return ga, gb, gc, gd, None


def thomas_solve(a, b, c, d):
    """Batched tridiagonal solve; lane-per-system Thomas or wavefront hybrid
    chosen by size (``solve_tridiagonal`` exposes the choice)."""
    validate_tridiagonal(a, b, c, d)
    return _TridiagSolve.apply(a, b, c, d, METHOD_AUTO)


def pcr_solve(a, b, c, d):
    """Batched tridiagonal solve on the wavefront PCR hybrid."""
    validate_tridiagonal(a, b, c, d)
    return _TridiagSolve.apply(a, b, c, d, METHOD_PCR)


def solve_tridiagonal(a, b, c, d, method: str = "auto"):
    validate_tridiagonal(a, b, c, d)
    code = {"auto": METHOD_AUTO, "thomas": METHOD_THOMAS, "pcr": METHOD_PCR}[method]
    return _TridiagSolve.apply(a, b, c, d, code)


# aliases under the reference's names
thomas_solve_cuda_t = thomas_solve
pcr_solve_cuda_t = pcr_solve


# ---------------------------------------------------------------------------
# tree (DHS) solve
# ---------------------------------------------------------------------------


def validate_threads(threads: int) -> int:
    """The reference requires ``threads`` to divide its 32-lane warp
    (triton/_contracts.py:126-135).  A CDNA4 wavefront has 64 lanes, so any
    divisor of 64 is accepted; scheduling always uses one full wavefront."""
    if isinstance(threads, bool) or not isinstance(threads, int):
        raise TypeError("threads must be a positive integer that divides 64")
    if threads <= 0 or threads > 64 or 64 % threads != 0:
        raise ValueError("threads must divide 64 (wavefront size)")
    return threads


def validate_tree(d_mem, a_geom, b, parent_idx, order, layer_ptr, threads):
    """Contract of the reference's validate_tree (triton/_contracts.py:199-230)."""
    data = (("d_mem", d_mem), ("a_geom", a_geom), ("b", b))
    topo = (("parent_idx", parent_idx), ("order", order), ("layer_ptr", layer_ptr))
    for name, t in data + topo:
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
    if d_mem.ndim != 2:
        raise ValueError(f"d_mem must have shape (B, K), got {tuple(d_mem.shape)}")
    batch, size = map(int, d_mem.shape)
    if batch < 1 or size < 1:
        raise ValueError("B and K must both be >= 1")
    if tuple(a_geom.shape) != (batch, size) or tuple(b.shape) != (batch, size):
        raise ValueError("d_mem, a_geom, and b must have identical (B, K) shapes")
    _float_group(data)
    for name, t in topo:
        if t.ndim != 1:
            raise ValueError(f"{name} must be 1-D")
        if t.dtype not in _INDEX:
            raise TypeError(f"{name} must have dtype int32 or int64")
        if t.device != d_mem.device:
            raise ValueError(f"{name} must be on device {d_mem.device}; got {t.device}")
    if parent_idx.numel() != size or order.numel() != size:
        raise ValueError("parent_idx and order must each contain K entries")
    if layer_ptr.numel() < 2:
        raise ValueError("layer_ptr must contain at least two entries")
    return batch, size, validate_threads(threads)


def _dhs_raw(d_mem, a_geom, b, parent, order, layer_ptr):
    lib = _lib.load()
    d_mem, a_geom, b = (t.contiguous() for t in (d_mem, a_geom, b))
    parent, order, layer_ptr = (
        t.to(torch.int64).contiguous() for t in (parent, order, layer_ptr)
    )
    B, K = d_mem.shape
    L = layer_ptr.numel() - 1
    x = torch.empty_like(b)
    f64 = b.dtype == torch.float64
    n_scr = lib.dn_tree_dhs_scratch_elems(B, K, 8 if f64 else 4)
    scratch = torch.empty(max(int(n_scr), 1), dtype=b.dtype, device=b.device)
    fn = lib.dn_tree_dhs_f64 if f64 else lib.dn_tree_dhs_f32
    st = fn(
        _lib.ptr(d_mem), _lib.ptr(a_geom), _lib.ptr(b), _lib.ptr(parent),
        _lib.ptr(order), _lib.ptr(layer_ptr), _lib.ptr(x), B, K, L,
        _lib.ptr(scratch), int(n_scr), _lib.stream_handle(b.device),
    )
    _lib.check(st, "DHS tree solve")
    return x


class _DHSSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d_mem, a_geom, b, parent, order, layer_ptr):
        x = _dhs_raw(d_mem, a_geom, b, parent, order, layer_ptr)
        ctx.save_for_backward(d_mem, a_geom, parent, order, layer_ptr, x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        d_mem, a_geom, parent, order, layer_ptr, x = ctx.saved_tensors
        need = ctx.needs_input_grad
        # The tree matrix is symmetric: the adjoint is the same solve.
        y = _dhs_raw(d_mem, a_geom, grad_x.contiguous(), parent, order, layer_ptr)
        gd = -y * x if need[0] else None
        ga = None
        if need[1]:
            par = parent.to(torch.int64)
            has = par >= 0
            safe = torch.where(has, par, torch.zeros_like(par))
            yp = torch.where(has, y.index_select(1, safe), torch.zeros_like(y))
            xp = torch.where(has, x.index_select(1, safe), torch.zeros_like(x))
            ga = -(y - yp) * (x - xp)
        gb = y if need[2] else None
        return gd, ga, gb, None, None, None


# Contraction plans of the tree topologies the drop-in has seen, cached per
# parent_idx tensor object (weakly) and checked against its version and size.
_contract_plans = WeakIdKeyDictionary()


def _contraction_plan(parent_idx):
    """The topology's contraction plan, or None for a forest (several roots)
    or a tree beyond the contraction kernel's section capacity.  Building it
    reads parent_idx to the host once per tensor."""
    stamp = (parent_idx._version, parent_idx.numel(), parent_idx.device)
    hit = _contract_plans.get(parent_idx)
    if hit is not None and hit[0] == stamp:
        return hit[1]
    from .contraction import TreeContractionPlan

    par = parent_idx.detach().to("cpu", torch.int64).numpy().reshape(-1)
    try:
        plan = TreeContractionPlan(par, max_segment_interior_nodes=16)
    except ValueError:
        plan = None
    _contract_plans[parent_idx] = (stamp, plan)
    return plan


def dhs_solve(d_mem, a_geom, b, parent_idx, order, layer_ptr, threads: int = 16):
    """Tree (Hines) solve; see the module doc.  The solution does not depend
    on the caller's layering, so a single tree is solved by contraction over
    the topology that parent_idx gives (O(log) dependent rounds instead of one
    per layer; same gradients, including the root's a_geom as a conductance
    to ground); a forest, or a tree beyond the contraction kernel's capacity,
    is eliminated over the caller's layers."""
    validate_tree(d_mem, a_geom, b, parent_idx, order, layer_ptr, threads)
    plan = _contraction_plan(parent_idx)
    if plan is not None and plan.n_sections <= plan.section_cap(d_mem.element_size()):
        from .contraction import _TreeContraction

        return _TreeContraction.apply(d_mem, a_geom, b, plan)
    return _DHSSolve.apply(d_mem, a_geom, b, parent_idx, order, layer_ptr)


def dhs_solve_layers(d_mem, a_geom, b, parent_idx, order, layer_ptr, threads: int = 16):
    """The same solve always eliminated over the caller's layers
    (dn_tree_dhs_*, one wavefront per system)."""
    validate_tree(d_mem, a_geom, b, parent_idx, order, layer_ptr, threads)
    return _DHSSolve.apply(d_mem, a_geom, b, parent_idx, order, layer_ptr)


dhs_solve_cuda = dhs_solve


# -- packed multi-morphology tree solve (_dhs_multi) ------------------------

_PLAN_NAMES = ("P_cat", "ORDER_cat", "LAYER_PTR_cat", "WARP_P_OFF", "WARP_ORDER_OFF",
               "WARP_LPTR_OFF", "WARP_L", "WARP_ROW_BASE", "WARP_ROW_COUNT")


def validate_tree_multi(d_mem, a_geom, b, plan, K_stride, L_max, threads, grid_x):
    """Checks of the reference's packed launch contract (triton/_contracts.py:
    271-341): (B, K_stride) planes of one float dtype, integer plan tables on
    the planes' device, one entry per plan row in each WARP_* table, grid_x
    equal to that count.  Reads no tensor values."""
    data = (("d_mem", d_mem), ("a_geom", a_geom), ("b", b))
    for name, t in data + tuple(zip(_PLAN_NAMES, plan)):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
    if d_mem.ndim != 2:
        raise ValueError(f"d_mem must have shape (B, K_stride), got {tuple(d_mem.shape)}")
    if a_geom.shape != d_mem.shape or b.shape != d_mem.shape:
        raise ValueError("d_mem, a_geom and b must share one (B, K_stride) shape")
    batch, width = map(int, d_mem.shape)
    if batch < 1 or width < 1:
        raise ValueError("the planes need at least one row and one column")
    for name, v in (("K_stride", K_stride), ("L_max", L_max)):
        if isinstance(v, bool) or not isinstance(v, int):
            raise TypeError(f"{name} must be an integer")
    if K_stride != width:
        raise ValueError(f"K_stride {K_stride} differs from the plane width {width}")
    if L_max < 1:
        raise ValueError("L_max must be at least 1")
    _float_group(data)
    for name, t in zip(_PLAN_NAMES, plan):
        if t.dtype not in _INDEX:
            raise TypeError(f"{name} must have dtype int32 or int64")
        if t.device != d_mem.device:
            raise ValueError(f"{name} must be on device {d_mem.device}; got {t.device}")
    P_cat, order_cat, lptr_cat = plan[:3]
    if P_cat.numel() != order_cat.numel() or P_cat.numel() < 1:
        raise ValueError("P_cat and ORDER_cat must be non-empty and of equal length")
    if lptr_cat.numel() < 2:
        raise ValueError("LAYER_PTR_cat needs at least two entries")
    n_warps = int(plan[7].numel())
    if n_warps < 1:
        raise ValueError("the warp plan has no entries")
    for name, t in zip(_PLAN_NAMES[3:], plan[3:]):
        if t.numel() != n_warps:
            raise ValueError(f"{name} must hold one entry per plan row ({n_warps})")
    if grid_x is not None:
        if isinstance(grid_x, bool) or not isinstance(grid_x, int):
            raise TypeError("grid_x must be an integer or None")
        if grid_x != n_warps:
            raise ValueError(f"grid_x {grid_x} differs from the plan's {n_warps} entries")
    return batch, n_warps, validate_threads(threads)


def _multi_args(plan, B_total, K_stride, L_max):
    P_cat, order_cat, lptr_cat = (t.to(torch.int64).contiguous() for t in plan[:3])
    warps = [t.to(torch.int64).contiguous() for t in plan[3:]]
    keep = (P_cat, order_cat, lptr_cat, *warps)
    args = (_lib.ptr(P_cat), _lib.ptr(order_cat), P_cat.numel(), _lib.ptr(lptr_cat),
            lptr_cat.numel(), *(_lib.ptr(w) for w in warps), warps[0].numel(), B_total,
            K_stride, L_max)
    return keep, args


def _multi_solve_raw(d_mem, a_geom, b, plan, K_stride, L_max, max_rows):
    d_mem, a_geom, b = (t.contiguous() for t in (d_mem, a_geom, b))
    B = d_mem.shape[0]
    keep, args = _multi_args(plan, B, K_stride, L_max)
    x = torch.zeros_like(b)
    work = torch.empty_like(d_mem)
    lib = _lib.load()
    fn = lib.dn_tree_dhs_multi_f64 if b.dtype == torch.float64 else lib.dn_tree_dhs_multi_f32
    st = fn(_lib.ptr(d_mem), _lib.ptr(a_geom), _lib.ptr(b), *args, int(max_rows),
            _lib.ptr(work), _lib.ptr(x), _lib.stream_handle(b.device))
    _lib.check(st, "packed multi-tree solve")
    del keep
    return x


class _DHSSolveMulti(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d_mem, a_geom, b, plan, K_stride, L_max, max_rows):
        x = _multi_solve_raw(d_mem, a_geom, b, plan, K_stride, L_max, max_rows)
        ctx.save_for_backward(d_mem, a_geom, x, *plan)
        ctx.meta = (K_stride, L_max, max_rows)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        d_mem, a_geom, x, *plan = ctx.saved_tensors
        K_stride, L_max, max_rows = ctx.meta
        need = ctx.needs_input_grad
        # symmetric matrix: the adjoint is the same solve; padding stays zero
        y = _multi_solve_raw(d_mem, a_geom, grad_x.contiguous(), plan, K_stride, L_max,
                             max_rows)
        gd = -y * x if need[0] else None
        ga = None
        if need[1]:
            keep, args = _multi_args(plan, d_mem.shape[0], K_stride, L_max)
            ga = torch.zeros_like(a_geom)
            lib = _lib.load()
            fn = lib.dn_tree_dhs_multi_grad_a_f64 if x.dtype == torch.float64 else \
                lib.dn_tree_dhs_multi_grad_a_f32
            st = fn(_lib.ptr(x), _lib.ptr(y), *args, _lib.ptr(ga), _lib.stream_handle(x.device))
            _lib.check(st, "packed multi-tree gradient")
            del keep
        gb = y if need[2] else None
        return gd, ga, gb, None, None, None, None


def dhs_solve_multi(d_mem, a_geom, b, P_cat, ORDER_cat, LAYER_PTR_cat, WARP_P_OFF,
                    WARP_ORDER_OFF, WARP_LPTR_OFF, WARP_L, WARP_ROW_BASE, WARP_ROW_COUNT,
                    K_stride: int, L_max: int, threads: int = 16, grid_x: int | None = None):
    """Packed multi-morphology tree solve over row-padded planes, with the
    reference's arguments (dhs_kernel_multi.py:831-888).  Entries the plan
    does not cover -- padding columns and rows -- come back as zero
    (_cuda_functional.py:207-211).  Gradients: b (the adjoint solve), d_mem
    (-y x) and a_geom (-(y_i - y_p)(x_i - x_p), -y x at roots)."""
    plan = (P_cat, ORDER_cat, LAYER_PTR_cat, WARP_P_OFF, WARP_ORDER_OFF, WARP_LPTR_OFF,
            WARP_L, WARP_ROW_BASE, WARP_ROW_COUNT)
    _, _, threads = validate_tree_multi(d_mem, a_geom, b, plan, K_stride, L_max, threads,
                                        grid_x)
    max_rows = max(1, 32 // threads)
    return _DHSSolveMulti.apply(d_mem, a_geom, b, plan, int(K_stride), int(L_max), max_rows)


dhs_solve_multi_cuda = dhs_solve_multi
