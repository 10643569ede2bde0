"""Transform-compatible facades over the HIP solvers (torch.library operators).

Restates, for the scalar solvers, the interface of the reference's
``dendra/models/integrators/_cuda_functional.py``: every solve is an opaque,
non-mutating dispatcher operator

* ``dendra_amd::tridiagonal(a, b, c, rhs, pcr)`` (Thomas or PCR),
* ``dendra_amd::tree(main, edge, rhs, parent, order, layers, threads)`` (DHS),
* ``dendra_amd::tree_contraction(d, g, rhs, parent, <14 plan tables>, n_sec,
  n_rounds)``

with a fake (shape) rule, so ``make_fx`` and ``torch.compile`` capture the call
instead of tracing into it.  Reverse AD, forward AD and vmap are the rules of
split-context autograd Functions, registered on the operators as well, so a
captured graph keeps them (_cuda_functional.py:547-568):

* backward: the transposed solve (the symmetric tree's adjoint is the same
  solve) and the band / edge outer-product gradients (:234-262, 365-374,
  390-412);
* JVP: ``A dx = d rhs - dA x`` with the same operator (:264-282, 376-389,
  414-420);
* vmap: the mapped dimension joins the solver's batch; the topology must be
  shared (:284-294, 422-431).

The public entry points ``thomas_solve``, ``pcr_solve``, ``dhs_solve`` and
``tree_contraction_solve`` validate like the imperative facades and return
a fresh tensor.
"""

from __future__ import annotations

import torch

from .contraction import TABLE_NAMES, TreeContractionPlan, _check_inputs, contract_tables
from .solvers import (METHOD_PCR, METHOD_THOMAS, _dhs_raw, _tridiag_raw, validate_tree,
                      validate_tridiagonal)


# --- launchers: the only calls into the HIP library --------------------------
def _launch_tridiagonal(a, b, c, rhs, pcr):
    return _tridiag_raw(a, b, c, rhs, METHOD_PCR if pcr else METHOD_THOMAS)


def _launch_tree(main, edge, rhs, parent, order, layers):
    return _dhs_raw(main, edge, rhs, parent, order, layers)


def _launch_contraction(d, g, rhs, tables, n_sec, n_rounds):
    return contract_tables(d, g, rhs, tables, n_sec, n_rounds)


# --- opaque operators ---------------------------------------------------------
@torch.library.custom_op("dendra_amd::tridiagonal", mutates_args=())
def _raw_tridiagonal(a: torch.Tensor, b: torch.Tensor, c: torch.Tensor,
                     rhs: torch.Tensor, pcr: bool) -> torch.Tensor:
    return _launch_tridiagonal(a, b, c, rhs, pcr)


@_raw_tridiagonal.register_fake
def _fake_tridiagonal(a, b, c, rhs, pcr):
    return rhs.new_empty(rhs.shape)


@torch.library.custom_op("dendra_amd::tree", mutates_args=())
def _raw_tree(main: torch.Tensor, edge: torch.Tensor, rhs: torch.Tensor,
              parent: torch.Tensor, order: torch.Tensor, layers: torch.Tensor,
              threads: int) -> torch.Tensor:
    return _launch_tree(main, edge, rhs, parent.contiguous(), order.contiguous(),
                        layers.contiguous())


@_raw_tree.register_fake
def _fake_tree(main, edge, rhs, parent, order, layers, threads):
    return rhs.new_empty(rhs.shape)


@torch.library.custom_op("dendra_amd::tree_contraction", mutates_args=())
def _raw_contraction(d: torch.Tensor, g: torch.Tensor, rhs: torch.Tensor,
                     parent: torch.Tensor, perm: torch.Tensor, sec_head: torch.Tensor,
                     sec_len: torch.Tensor, sec_child: torch.Tensor,
                     sec_nchild: torch.Tensor, sec_parent: torch.Tensor,
                     elim_ptr: torch.Tensor, elim_sec: torch.Tensor,
                     elim_par: torch.Tensor, elim_child: torch.Tensor,
                     recv_ptr: torch.Tensor, recv_sec: torch.Tensor,
                     src_ptr: torch.Tensor, src: torch.Tensor, n_sec: int,
                     n_rounds: int) -> torch.Tensor:
    tables = [perm, sec_head, sec_len, sec_child, sec_nchild, sec_parent, elim_ptr,
              elim_sec, elim_par, elim_child, recv_ptr, recv_sec, src_ptr, src]
    return _launch_contraction(d, g, rhs, tables, n_sec, n_rounds)


@_raw_contraction.register_fake
def _fake_contraction(d, g, rhs, parent, *rest):
    return rhs.new_empty(rhs.shape)


N_TABLES = len(TABLE_NAMES)


# --- transform helpers ----------------------------------------------------------
def _save(ctx, *values):
    ctx.save_for_backward(*values)
    ctx.save_for_forward(*values)


def _require_shared(in_dims):
    for d in in_dims:
        if isinstance(d, (list, tuple)):
            _require_shared(d)
        elif d is not None:
            raise ValueError("solver topology must be shared across vmap")


def _flatten(info, in_dims, *tensors):
    """Fold the mapped dimension V into the solver batch B: (V, B, ...) ->
    (V*B, ...); unmapped operands are broadcast over V."""
    V = info.batch_size
    moved, B = [], None
    for t, d in zip(tensors, in_dims):
        v = t.unsqueeze(0).expand(V, *t.shape) if d is None else t.movedim(d, 0)
        if B is None:
            B = v.shape[1]
        elif v.shape[1] != B:
            raise ValueError("all solver operands must have the same batch size")
        moved.append(v)
    return tuple(v.reshape(V * B, *v.shape[2:]) for v in moved), B


def _restore(x, V, B):
    return x.reshape(V, B, *x.shape[1:])


def _empty_solution(flat):
    # a differentiable zero-sized result, without a zero-grid launch
    out = flat[-1].clone()
    for v in flat[:-1]:
        out = out + v.sum() * 0.0
    return out


def _tree_edge_gradient(x, y, parent):
    safe = parent.clamp_min(0).to(torch.long)
    root = (parent < 0).reshape(1, -1)
    return torch.where(root, -(y * x),
                       -((y - y.index_select(-1, safe)) * (x - x.index_select(-1, safe))))


def _tree_tangent_rhs(main_t, edge_t, rhs_t, x, parent):
    """d rhs - dA x for the tree matrix (edge to the parent; ground at the root)."""
    tangent = torch.zeros_like(x) if rhs_t is None else rhs_t
    if main_t is not None:
        tangent = tangent - main_t * x
    if edge_t is None:
        return tangent
    safe = parent.clamp_min(0).to(torch.long)
    root = (parent < 0).reshape(1, -1)
    action = torch.where(root, torch.zeros_like(x), edge_t * (x - x.index_select(-1, safe)))
    tangent = tangent - action
    tangent = tangent.scatter_add(-1, safe.reshape(1, -1).expand_as(x), action)
    return tangent - torch.where(root, edge_t * x, torch.zeros_like(x))


# --- autograd Functions ---------------------------------------------------------
class _TridiagonalSolve(torch.autograd.Function):
    @staticmethod
    def forward(a, b, c, rhs, pcr):
        return _raw_tridiagonal(a, b, c, rhs, pcr)

    @staticmethod
    def setup_context(ctx, inputs, output):
        a, b, c, _rhs, ctx.pcr = inputs
        _save(ctx, a, b, c, output)

    @staticmethod
    def backward(ctx, grad_x):
        a, b, c, x = ctx.saved_tensors
        y = _apply_tridiagonal(c, b, a, grad_x, ctx.pcr)  # transposed bands
        return (-(y[..., 1:] * x[..., :-1]), -(y * x), -(y[..., :-1] * x[..., 1:]), y,
                None)

    @staticmethod
    def jvp(ctx, a_t, b_t, c_t, rhs_t, _pcr_t):
        a, b, c, x = ctx.saved_tensors
        tangent = torch.zeros_like(x) if rhs_t is None else rhs_t
        pad = torch.zeros_like(x[..., :1])
        if b_t is not None:
            tangent = tangent - b_t * x
        if a_t is not None:
            tangent = tangent - torch.cat((pad, a_t * x[..., :-1]), dim=-1)
        if c_t is not None:
            tangent = tangent - torch.cat((c_t * x[..., 1:], pad), dim=-1)
        return _apply_tridiagonal(a, b, c, tangent, ctx.pcr)

    @staticmethod
    def vmap(info, in_dims, a, b, c, rhs, pcr):
        flat, B = _flatten(info, in_dims[:4], a, b, c, rhs)
        out = _empty_solution(flat) if info.batch_size == 0 else \
            _apply_tridiagonal(*flat, pcr)
        return _restore(out, info.batch_size, B), 0


class _TreeSolve(torch.autograd.Function):
    @staticmethod
    def forward(main, edge, rhs, parent, order, layers, threads):
        return _raw_tree(main, edge, rhs, parent, order, layers, threads)

    @staticmethod
    def setup_context(ctx, inputs, output):
        main, edge, _rhs, parent, order, layers, ctx.threads = inputs
        _save(ctx, main, edge, output, parent, order, layers)

    @staticmethod
    def backward(ctx, grad_x):
        main, edge, x, parent, order, layers = ctx.saved_tensors
        y = _apply_tree(main, edge, grad_x, parent, order, layers, ctx.threads)
        return -(y * x), _tree_edge_gradient(x, y, parent), y, None, None, None, None

    @staticmethod
    def jvp(ctx, main_t, edge_t, rhs_t, *_topology_t):
        main, edge, x, parent, order, layers = ctx.saved_tensors
        tangent = _tree_tangent_rhs(main_t, edge_t, rhs_t, x, parent)
        return _apply_tree(main, edge, tangent, parent, order, layers, ctx.threads)

    @staticmethod
    def vmap(info, in_dims, main, edge, rhs, parent, order, layers, threads):
        _require_shared(in_dims[3:])
        flat, B = _flatten(info, in_dims[:3], main, edge, rhs)
        out = _empty_solution(flat) if info.batch_size == 0 else \
            _apply_tree(*flat, parent, order, layers, threads)
        return _restore(out, info.batch_size, B), 0


class _ContractionSolve(torch.autograd.Function):
    # inputs: d, g, rhs, parent, the N_TABLES plan tables, n_sec, n_rounds
    @staticmethod
    def forward(d, g, rhs, parent, *rest):
        return _raw_contraction(d, g, rhs, parent, *rest)

    @staticmethod
    def setup_context(ctx, inputs, output):
        d, g, _rhs, parent, *rest = inputs
        ctx.n_sec, ctx.n_rounds = rest[N_TABLES:]
        _save(ctx, d, g, output, parent, *rest[:N_TABLES])

    @staticmethod
    def backward(ctx, grad_x):
        d, g, x, parent, *tables = ctx.saved_tensors
        y = _apply_contraction(d, g, grad_x, parent, tables, ctx.n_sec, ctx.n_rounds)
        return (-(y * x), _tree_edge_gradient(x, y, parent), y) + (None,) * (N_TABLES + 3)

    @staticmethod
    def jvp(ctx, d_t, g_t, rhs_t, *_topology_t):
        d, g, x, parent, *tables = ctx.saved_tensors
        tangent = _tree_tangent_rhs(d_t, g_t, rhs_t, x, parent)
        return _apply_contraction(d, g, tangent, parent, tables, ctx.n_sec, ctx.n_rounds)

    @staticmethod
    def vmap(info, in_dims, d, g, rhs, parent, *rest):
        _require_shared(in_dims[3:])
        flat, B = _flatten(info, in_dims[:3], d, g, rhs)
        out = _empty_solution(flat) if info.batch_size == 0 else \
            _apply_contraction(*flat, parent, rest[:N_TABLES], *rest[N_TABLES:])
        return _restore(out, info.batch_size, B), 0


# make_fx erases Python autograd Functions during capture; the operators carry
# the same rules so that a captured graph still differentiates and maps.
_raw_tridiagonal.register_autograd(_TridiagonalSolve.backward,
                                   setup_context=_TridiagonalSolve.setup_context)
_raw_tridiagonal.register_vmap(_TridiagonalSolve.vmap)
_raw_tree.register_autograd(_TreeSolve.backward, setup_context=_TreeSolve.setup_context)
_raw_tree.register_vmap(_TreeSolve.vmap)
_raw_contraction.register_autograd(_ContractionSolve.backward,
                                   setup_context=_ContractionSolve.setup_context)
_raw_contraction.register_vmap(_ContractionSolve.vmap)


@torch.compiler.allow_in_graph
def _apply_tridiagonal(a, b, c, rhs, pcr):
    return _TridiagonalSolve.apply(a, b, c, rhs, pcr)


@torch.compiler.allow_in_graph
def _apply_tree(main, edge, rhs, parent, order, layers, threads):
    return _TreeSolve.apply(main, edge, rhs, parent, order, layers, threads)


@torch.compiler.allow_in_graph
def _apply_contraction(d, g, rhs, parent, tables, n_sec, n_rounds):
    return _ContractionSolve.apply(d, g, rhs, parent, *tables, n_sec, n_rounds)


# --- public entry points ------------------------------------------------------------
def thomas_solve(a, b, c, rhs):
    """Transform-compatible Thomas solve (_cuda_functional.py:594-597)."""
    validate_tridiagonal(a, b, c, rhs)
    return _apply_tridiagonal(a, b, c, rhs, False)


def pcr_solve(a, b, c, rhs):
    """Transform-compatible PCR solve (_cuda_functional.py:599-602)."""
    validate_tridiagonal(a, b, c, rhs)
    return _apply_tridiagonal(a, b, c, rhs, True)


def dhs_solve(main, edge, rhs, parent, order, layers, *, threads):
    """Transform-compatible tree solve (_cuda_functional.py:615-618)."""
    validate_tree(main, edge, rhs, parent, order, layers, threads)
    return _apply_tree(main, edge, rhs, parent, order, layers, int(threads))


def tree_contraction_solve(d, g, rhs, plan: TreeContractionPlan):
    """Transform-compatible tree-contraction solve; same checks as
    ``dendra_amd.contraction.solve``."""
    if torch.compiler.is_compiling() and d.device not in plan._devices:
        raise RuntimeError("prepare the CUDA tree contraction plan before compilation")
    _check_inputs(d, g, rhs, plan)
    t = plan.prepare(d.device)
    return _apply_contraction(d, g, rhs, t["parent_i64"], [t[k] for k in TABLE_NAMES],
                              plan.n_sections, plan.n_rounds)
