"""The transform-compatible solver operators on the GPU kernels: values,
gradients, JVP, vmap, make_fx capture and fullgraph compilation against dense
float64 solves, and torch.library.opcheck of each operator."""

import pytest
import torch
from torch.fx.experimental.proxy_tensor import make_fx

import dendra_amd.functional as F
from dendra_amd.contraction import TABLE_NAMES, TreeContractionPlan
from test_functional import PARENTS, dense, tree_matrix, tridiagonal_matrix, values

pytestmark = pytest.mark.gpu

FAMILIES = ("thomas", "pcr", "dhs", "contraction")


class GpuCase:
    def __init__(self, family):
        B, K = 3, len(PARENTS)
        dev = lambda t: t.cuda()
        if family in ("thomas", "pcr"):
            vals = (values((B, K - 1), -0.3), values((B, K), 2.0), values((B, K - 1), -0.2),
                    values((B, K), -0.4))
            pcr = family == "pcr"
            self.solve = lambda a, b, c, r: F.pcr_solve(a, b, c, r) if pcr else \
                F.thomas_solve(a, b, c, r)
            self.reference = lambda a, b, c, r: dense(tridiagonal_matrix(a, b, c), r)
        else:
            main, edge = values((B, K), 1.0), values((B, K), 0.2).clone()
            edge[:, 2] = 0.0
            vals = (main, edge, values((B, K), -0.4))
            self.reference = lambda m, e, r: dense(tree_matrix(m, e, PARENTS), r)
            if family == "dhs":
                from dendra_amd.topology import build_dhs_layers, build_morphology

                _, _, depth = build_morphology(PARENTS)
                order, layers = build_dhs_layers(depth, 16)
                topo = [torch.as_tensor(v, dtype=torch.int64, device="cuda")
                        for v in (PARENTS, order, layers)]
                self.solve = lambda m, e, r: F.dhs_solve(m, e, r, *topo, threads=16)
                self.op = (torch.ops.dendra_amd.tree.default, (*topo, 16))
            else:
                plan = TreeContractionPlan(PARENTS, max_segment_interior_nodes=2)
                self.solve = lambda m, e, r: F.tree_contraction_solve(m, e, r, plan)
                t = plan.prepare("cuda")
                self.op = (torch.ops.dendra_amd.tree_contraction.default,
                           (t["parent_i64"], *[t[k] for k in TABLE_NAMES],
                            plan.n_sections, plan.n_rounds))
        if family in ("thomas", "pcr"):
            self.op = (torch.ops.dendra_amd.tridiagonal.default, (family == "pcr",))
        self.cpu = vals
        self.values = tuple(dev(v) for v in vals)


def close(x, y):
    torch.testing.assert_close(x.cpu(), y, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("family", FAMILIES)
def test_gpu_operator_values_gradients_jvp_and_vmap(family):
    case = GpuCase(family)
    ins = tuple(v.clone().requires_grad_() for v in case.values)
    ref = tuple(v.clone().requires_grad_() for v in case.cpu)
    x, want = case.solve(*ins), case.reference(*ref)
    close(x, want)
    w = values(x.shape, 0.2)
    for g_, e_ in zip(torch.autograd.grad((x.square() * w.cuda()).sum(), ins),
                      torch.autograd.grad((want.square() * w).sum(), ref)):
        close(g_, e_)
    tangents = tuple(values(v.shape, 0.02) for v in case.cpu)
    got = torch.func.jvp(case.solve, case.values, tuple(t.cuda() for t in tangents))
    exp = torch.func.jvp(case.reference, case.cpu, tangents)
    for g_, e_ in zip(got, exp):
        close(g_, e_)
    coef, rhs = case.values[:-1], case.values[-1]
    src = torch.stack((rhs, rhs * 0.8, rhs * 1.2))
    exp = torch.stack([case.reference(*case.cpu[:-1], r.cpu()) for r in src])
    close(torch.vmap(lambda r: case.solve(*coef, r))(src), exp)


@pytest.mark.parametrize("family", FAMILIES)
def test_gpu_operator_survives_make_fx_and_fullgraph_compile(family):
    case = GpuCase(family)
    with torch.no_grad():
        graph = make_fx(case.solve)(*case.values)
    assert any(n.op == "call_function" and str(n.target).startswith("dendra_amd.")
               for n in graph.graph.nodes)
    compiled = torch.compile(case.solve, backend="aot_eager", fullgraph=True)
    for fn in (graph, compiled):
        ins = tuple(v.clone().requires_grad_() for v in case.values)
        ref = tuple(v.clone().requires_grad_() for v in case.cpu)
        x, want = fn(*ins), case.reference(*ref)
        close(x, want)
        for g_, e_ in zip(torch.autograd.grad(x.square().sum(), ins),
                          torch.autograd.grad(want.square().sum(), ref)):
            close(g_, e_)


@pytest.mark.parametrize("family", FAMILIES)
def test_gpu_opcheck(family):
    case = GpuCase(family)
    op, extra = case.op
    args = (*[v.clone().requires_grad_() for v in case.values], *extra)
    torch.library.opcheck(op, args)
