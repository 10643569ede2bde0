"""CPU contracts of the transform-compatible solver operators
(dendra_amd.functional): gradients, JVP, vmap, capture by make_fx and
torch.compile, with dense float64 solves injected in place of the HIP
launchers.  Mirrors tests/func/test_cuda_functional_reference.py of the
reference; the GPU kernels behind the same operators are checked in
tests/test_gpu_functional.py."""

from __future__ import annotations

import pytest
import torch
from torch.fx.experimental.proxy_tensor import make_fx

import dendra_amd.functional as F
from dendra_amd.contraction import TABLE_NAMES, TreeContractionPlan

FAMILIES = ("thomas", "pcr", "dhs", "contraction")
PARENTS = (2, 0, -1, 2, 3)  # the reference's five-node test tree (root 2)


def tridiagonal_matrix(a, b, c):
    return torch.diag_embed(b) + torch.diag_embed(a, -1) + torch.diag_embed(c, 1)


def tree_matrix(main, edge, parents):
    n = len(parents)
    eye = torch.eye(n, dtype=main.dtype)
    m = torch.diag_embed(main + edge)
    for i, p in enumerate(parents):
        if p >= 0:
            e = (eye[p].unsqueeze(-1) * eye[p]) - (eye[i].unsqueeze(-1) * eye[p]) - \
                (eye[p].unsqueeze(-1) * eye[i])
            m = m + edge[..., i, None, None] * e
    return m


def dense(matrix, rhs):
    return torch.linalg.solve(matrix, rhs.unsqueeze(-1)).squeeze(-1)


def values(shape, start):
    n = int(torch.tensor(shape).prod())
    return torch.linspace(start, start + 0.5, n, dtype=torch.float64).reshape(shape)


class Case:
    def __init__(self, family):
        self.family = family
        B, K = 2, len(PARENTS)
        if family in ("thomas", "pcr"):
            a, c = values((B, K - 1), -0.3), values((B, K - 1), -0.2)
            b = values((B, K), 2.0)
            self.values = (a, b, c, values((B, K), -0.4))
            pcr = family == "pcr"
            self.solve = lambda a, b, c, r: F._apply_tridiagonal(a, b, c, r, pcr)
            self.reference = lambda a, b, c, r: dense(tridiagonal_matrix(a, b, c), r)
            self.metadata = ()
        else:
            main, edge = values((B, K), 1.0), values((B, K), 0.2)
            edge = edge.clone()
            edge[:, 2] = 0.0  # the conventional zero root coefficient
            self.values = (main, edge, values((B, K), -0.4))
            parent = torch.tensor(PARENTS, dtype=torch.int64)
            self.reference = lambda m, e, r: dense(tree_matrix(m, e, PARENTS), r)
            if family == "dhs":
                from dendra_amd.topology import build_dhs_layers, build_morphology

                _, _, depth = build_morphology(PARENTS)
                order, layers = build_dhs_layers(depth, 16)
                order = torch.as_tensor(order, dtype=torch.int64)
                layers = torch.as_tensor(layers, dtype=torch.int64)
                self.metadata = (parent, order, layers)
                self.solve = lambda m, e, r: F._apply_tree(m, e, r, parent, order, layers, 16)
            else:
                plan = TreeContractionPlan(PARENTS, max_segment_interior_nodes=2)
                t = plan.tables_on("cpu")
                tables = [t[k] for k in TABLE_NAMES]
                self.metadata = (t["parent_i64"], *tables)
                self.solve = lambda m, e, r: F._apply_contraction(
                    m, e, r, t["parent_i64"], tables, plan.n_sections, plan.n_rounds)


@pytest.fixture
def launches(monkeypatch):
    """Replace the HIP launchers by dense float64 solves; count the calls."""
    count = {"n": 0}

    def tri(a, b, c, rhs, pcr):
        count["n"] += 1
        return dense(tridiagonal_matrix(a, b, c), rhs)

    def tree(main, edge, rhs, parent, order, layers):
        count["n"] += 1
        return dense(tree_matrix(main, edge, parent.tolist()), rhs)

    def contraction(d, g, rhs, tables, n_sec, n_rounds):
        count["n"] += 1
        return dense(tree_matrix(d, g, PARENTS), rhs)

    monkeypatch.setattr(F, "_launch_tridiagonal", tri)
    monkeypatch.setattr(F, "_launch_tree", tree)
    monkeypatch.setattr(F, "_launch_contraction", contraction)
    return count


def close(x, y):
    torch.testing.assert_close(x, y, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("family", FAMILIES)
def test_primal_gradients_and_input_ownership(family, launches):
    case = Case(family)
    ins = tuple(v.clone().requires_grad_() for v in case.values)
    ref_ins = tuple(v.clone().requires_grad_() for v in case.values)
    snaps = [(v.clone(), v._version) for v in (*ins, *case.metadata)]
    x, want = case.solve(*ins), case.reference(*ref_ins)
    close(x, want)
    w = values(x.shape, 0.2)
    got = torch.autograd.grad((x.square() * w).sum(), ins)
    exp = torch.autograd.grad((want.square() * w).sum(), ref_ins)
    for g_, e_ in zip(got, exp):
        close(g_, e_)
    for v, (snap, ver) in zip((*ins, *case.metadata), snaps):
        assert v._version == ver
        torch.testing.assert_close(v, snap, rtol=0, atol=0)
    if family in ("dhs", "contraction"):
        assert torch.count_nonzero(got[1][:, 2])  # the root edge's gradient


@pytest.mark.parametrize("family", FAMILIES)
def test_joint_jvp_matches_dense_algebra(family, launches):
    case = Case(family)
    tangents = tuple(values(v.shape, 0.02) for v in case.values)
    got = torch.func.jvp(case.solve, case.values, tangents)
    exp = torch.func.jvp(case.reference, case.values, tangents)
    for g_, e_ in zip(got, exp):
        close(g_, e_)


@pytest.mark.parametrize("family", FAMILIES)
def test_rhs_vmap_nested_vmap_and_mapped_coefficients(family, launches):
    case = Case(family)
    coef, rhs = case.values[:-1], case.values[-1]
    src = torch.stack((rhs, rhs * 0.8, rhs * 1.2))
    solve = lambda r: case.solve(*coef, r)
    exp = torch.stack([case.reference(*coef, r) for r in src])
    close(torch.vmap(solve)(src), exp)
    nested = torch.stack((src, src * 0.9))
    close(torch.vmap(torch.vmap(solve))(nested), torch.stack((exp, exp * 0.9)))
    first = torch.stack((coef[0], coef[0] * 1.1))
    solve1 = lambda c0: case.solve(c0, *case.values[1:])
    close(torch.vmap(solve1)(first),
          torch.stack([case.reference(c0, *case.values[1:]) for c0 in first]))


@pytest.mark.parametrize("family", FAMILIES)
def test_empty_mapped_coefficients_are_differentiable(family, launches):
    case = Case(family)
    mapped = case.values[0].new_empty((0, *case.values[0].shape)).requires_grad_()
    out = torch.vmap(lambda c0: case.solve(c0, *case.values[1:]))(mapped)
    assert out.shape == (0, *case.values[-1].shape)
    grad = torch.autograd.grad(out.sum(), mapped)[0]
    assert grad.shape == mapped.shape
    assert launches["n"] == 0


@pytest.mark.parametrize("family", ["dhs", "contraction"])
def test_topology_mapped_vmap_fails_before_dispatch(family, launches):
    case = Case(family)
    parent = case.metadata[0]
    stacked = torch.stack((parent, parent))
    if family == "dhs":
        _, order, layers = case.metadata
        fn = lambda p: F._apply_tree(*case.values, p, order, layers, 16)
    else:
        plan = TreeContractionPlan(PARENTS, max_segment_interior_nodes=2)
        t = plan.tables_on("cpu")
        fn = lambda p: F._apply_contraction(*case.values, p, [t[k] for k in TABLE_NAMES],
                                            plan.n_sections, plan.n_rounds)
    with pytest.raises(ValueError, match="topology must be shared"):
        torch.vmap(fn)(stacked)
    assert launches["n"] == 0


@pytest.mark.parametrize("family", FAMILIES)
def test_no_grad_capture_keeps_the_opaque_operator_and_its_reverse_ad(family, launches):
    case = Case(family)
    with torch.no_grad():
        graph = make_fx(case.solve)(*case.values)
    ops = [n for n in graph.graph.nodes
           if n.op == "call_function" and str(n.target).startswith("dendra_amd.")]
    assert ops, "the solver operator must survive make_fx"
    assert not any("linalg_solve" in str(n.target) for n in graph.graph.nodes
                   if n.op == "call_function")
    ins = tuple(v.clone().requires_grad_() for v in case.values)
    ref_ins = tuple(v.clone().requires_grad_() for v in case.values)
    x, want = graph(*ins), case.reference(*ref_ins)
    close(x, want)
    for g_, e_ in zip(torch.autograd.grad(x.square().sum(), ins),
                      torch.autograd.grad(want.square().sum(), ref_ins)):
        close(g_, e_)


@pytest.mark.parametrize("family", FAMILIES)
def test_fullgraph_aot_forward_and_backward(family, launches):
    case = Case(family)
    compiled = torch.compile(case.solve, backend="aot_eager", fullgraph=True)
    ins = tuple(v.clone().requires_grad_() for v in case.values)
    ref_ins = tuple(v.clone().requires_grad_() for v in case.values)
    x, want = compiled(*ins), case.reference(*ref_ins)
    close(x, want)
    for g_, e_ in zip(torch.autograd.grad(x.square().sum(), ins),
                      torch.autograd.grad(want.square().sum(), ref_ins)):
        close(g_, e_)
