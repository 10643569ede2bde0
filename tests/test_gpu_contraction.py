"""GPU parity of the tree-contraction solve (dendra_amd.contraction.solve)
against dense float64 solves, the oracle's DHS and this package's DHS drop-in,
with the reference's tolerances where its tests state them
(tests/func/test_tree_contraction_integration.py:131-160,
tests/func/test_cuda_solver_precision.py:46-105)."""

import numpy as np
import pytest
import torch

from test_contraction_plan import dense_matrix, random_tree, shuffled

pytestmark = pytest.mark.gpu

# the reference integration test's tree: edges (7,1) (1,2) (2,3) (3,4) (4,0)
# (4,5) (7,6), root 7
PARENT8 = [4, 7, 1, 2, 3, 4, 7, -1]


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a), dtype=dtype, device="cuda")


def systems(par, B, rng, root_g=0.0):
    n = len(par)
    d = rng.uniform(0.01, 2.0, (B, n))
    g = rng.uniform(0.1, 5.0, (B, n))
    g[:, np.asarray(par) < 0] = root_g
    rhs = rng.standard_normal((B, n))
    want = np.stack([np.linalg.solve(dense_matrix(d[b], g[b], par), rhs[b])
                     for b in range(B)])
    return d, g, rhs, want


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("kind", ["random", "path", "star", "caterpillar", "binary"])
@pytest.mark.parametrize("seg", [16, 2, None])
def test_contraction_matches_dense(dtype, kind, seg):
    from dendra_amd.contraction import TreeContractionPlan, solve

    rng = np.random.default_rng(5)
    for n in (1, 3, 33, 700):
        par = shuffled(random_tree(n, rng, kind), rng)
        plan = TreeContractionPlan(par, max_segment_interior_nodes=seg)
        d, g, rhs, want = systems(par, 3, rng, root_g=0.4)
        x = solve(dev(d, dtype), dev(g, dtype), dev(rhs, dtype), plan)
        assert x.dtype == dtype and x.shape == (3, n)
        got = x.double().cpu().numpy()
        if dtype == torch.float64:
            assert np.abs(got - want).max() <= 1e-11 * max(1.0, np.abs(want).max())
        else:
            # normwise backward error of the float32 solution
            for b in range(3):
                A = dense_matrix(d[b], g[b], par)
                res = np.abs(A @ got[b] - rhs[b]).max()
                scale = np.abs(A).sum(1).max() * np.abs(got[b]).max() + np.abs(rhs[b]).max()
                assert res <= 64 * np.finfo(np.float32).eps * scale, (n, res / scale)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("cap", [2, None])
def test_contraction_matches_dhs_drop_in_and_oracle(orc, dtype, cap):
    """The reference's contraction-vs-DHS parity case: its 8-node tree, caps
    2 and None, its tolerances for the state and the gradients."""
    from dendra_amd.contraction import TreeContractionPlan, solve
    from dendra_amd.solvers import dhs_solve
    from dendra_amd.topology import build_dhs_layers, build_morphology

    rng = np.random.default_rng(8)
    _, _, depth = build_morphology(PARENT8)
    order, layer_ptr = build_dhs_layers(depth, 16)
    d, g, rhs, want = systems(PARENT8, 2, rng)
    oracle = orc.dhs(d, g, rhs, PARENT8, order, layer_ptr)
    assert np.abs(oracle - want).max() <= 1e-12
    plan = TreeContractionPlan(PARENT8, max_segment_interior_nodes=cap)
    w = dev(rng.standard_normal((2, 8)), dtype)

    def run(f):
        args = [dev(v, dtype).requires_grad_() for v in (d, g, rhs)]
        x = f(*args)
        return x, torch.autograd.grad((x * w).sum(), args)

    xc, gc = run(lambda a, b, c: solve(a, b, c, plan))
    xd, gd = run(lambda a, b, c: dhs_solve(a, b, c, dev(PARENT8, torch.int64),
                                           dev(order, torch.int64),
                                           dev(layer_ptr, torch.int64)))
    if dtype == torch.float64:
        (v_rtol, v_atol), (g_rtol, g_atol) = (2e-11, 1e-11), (2e-9, 5e-12)
    else:
        (v_rtol, v_atol), (g_rtol, g_atol) = (1e-6, 1e-5), (4e-4, 5e-6)
    torch.testing.assert_close(xc, xd, rtol=v_rtol, atol=v_atol)
    torch.testing.assert_close(xc.double().cpu(), torch.as_tensor(oracle),
                               rtol=max(v_rtol, 1e-6 if dtype == torch.float32 else 0),
                               atol=v_atol)
    for a, b in zip(gc, gd):
        assert torch.count_nonzero(a).item() > 0
        torch.testing.assert_close(a, b, rtol=g_rtol, atol=g_atol)


@pytest.mark.parametrize("signed", [False, True], ids=["spd", "indefinite"])
def test_contraction_float32_primal_and_gradients_match_dense(signed):
    """tests/func/test_cuda_solver_precision.py:67-105: a 3-node chain, SPD
    and indefinite (negative internal diagonal), float32 against dense
    float64 values and gradients; the root's d and g have equal gradients."""
    from dendra_amd.contraction import TreeContractionPlan, solve

    parents = [-1, 0, 1]
    if signed:
        main = [[2.0, -0.5, 1.0], [2.4, -0.6, 1.3]]
        edge = [[0.2, 0.2, 0.1], [0.3, 0.15, 0.2]]
    else:
        main = [[1.0, 2.0, 1.5], [1.2, 1.8, 2.4]]
        edge = [[0.2, 0.3, 0.4], [0.15, 0.4, 0.2]]
    rhs = [[0.7, -1.2, 0.3], [-0.2, 0.9, 1.1]]
    weights = torch.tensor([[0.3, -0.7, 0.9], [-0.6, 0.8, 0.2]], dtype=torch.float64)
    ref = [torch.tensor(v, dtype=torch.float64, requires_grad=True)
           for v in (main, edge, rhs)]
    mats = []
    for b in range(2):
        M = torch.diag(ref[0][b] + ref[1][b])
        for i, p in enumerate(parents):
            if p >= 0:
                e = torch.zeros(3, 3, dtype=torch.float64)
                e[p, p], e[i, p], e[p, i] = 1.0, -1.0, -1.0
                M = M + ref[1][b, i] * e
        mats.append(M)
    expected = torch.stack([torch.linalg.solve(mats[b], ref[2][b]) for b in range(2)])
    exp_grads = torch.autograd.grad((expected * weights).sum(), ref)

    plan = TreeContractionPlan(parents, max_segment_interior_nodes=16)
    args = [torch.tensor(v, dtype=torch.float32, device="cuda", requires_grad=True)
            for v in (main, edge, rhs)]
    actual = solve(*args, plan)
    grads = torch.autograd.grad((actual * weights.cuda().float()).sum(), args)
    torch.testing.assert_close(actual.cpu().double(), expected.detach(), rtol=1e-5,
                               atol=1e-6)
    for a, e in zip(grads, exp_grads):
        torch.testing.assert_close(a.cpu().double(), e, rtol=2e-5, atol=2e-6)
    torch.testing.assert_close(grads[0][:, 0], grads[1][:, 0], rtol=0, atol=0)


def test_contraction_float32_stiff_edge_keeps_uniform_solution():
    """tests/func/test_cuda_solver_precision.py:46-64: a 2^26 edge."""
    from dendra_amd.contraction import TreeContractionPlan, solve

    plan = TreeContractionPlan([-1, 0])
    main = torch.ones(1, 2, dtype=torch.float32, device="cuda")
    edge = torch.tensor([[0.0, float(2 ** 26)]], dtype=torch.float32, device="cuda")
    x = solve(main, edge, main.clone(), plan)
    assert bool(torch.isfinite(x).all())
    torch.testing.assert_close(x.cpu().double(), torch.ones(1, 2, dtype=torch.float64),
                               rtol=2e-6, atol=2e-7)


def test_contraction_deep_path_and_large_random_tree():
    """A 4000-node path (250 sections, logarithmic rounds) and a 3000-node
    random tree, float64 against dense solves, several systems at once;
    repeated calls are bitwise equal."""
    from dendra_amd.contraction import TreeContractionPlan, solve

    rng = np.random.default_rng(21)
    for par in (np.arange(-1, 3999), random_tree(3000, rng, "random")):
        plan = TreeContractionPlan(par)
        assert plan.n_rounds <= 2 * int(np.ceil(np.log2(plan.n_sections))) + 1
        d, g, rhs, want = systems(par, 4, rng)
        args = (dev(d), dev(g), dev(rhs))
        x = solve(*args, plan)
        got = x.cpu().numpy()
        assert np.abs(got - want).max() <= 1e-10 * max(1.0, np.abs(want).max())
        assert torch.equal(x, solve(*args, plan))


def test_contraction_validation_errors():
    from dendra_amd.contraction import TreeContractionPlan, solve

    plan = TreeContractionPlan([-1, 0, 0])
    ok = torch.ones(2, 3, device="cuda")
    with pytest.raises(ValueError, match="shape \\(B, K\\) matching the plan"):
        solve(torch.ones(2, 4, device="cuda"), ok, ok, plan)
    with pytest.raises(ValueError, match="float32 or float64"):
        solve(ok.half(), ok.half(), ok.half(), plan)
    with pytest.raises(ValueError, match="same-shaped strided CUDA"):
        solve(ok, ok.double(), ok, plan)
    with pytest.raises(ValueError, match="same-shaped strided CUDA"):
        solve(ok, ok.cpu(), ok, plan)
    big = TreeContractionPlan(np.concatenate([[-1], np.zeros(2999, dtype=np.int64)]),
                              max_segment_interior_nodes=None)
    z = torch.ones(1, 3000, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="sections; at most"):
        solve(z, z, z, big)
    x = solve(z.float(), z.float(), z.float(), big)  # 3000 sections fit in float32
    assert bool(torch.isfinite(x).all())
