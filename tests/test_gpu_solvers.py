"""GPU parity of the drop-in tridiagonal and tree solves (through the C ABI).

Tolerances follow the reference's dense-oracle checks
(tests/test_cuda_adversarial_solvers.py:113-135: float32 rtol 4e-4,
atol 6e-5, backward error 3e-5; float64 5e-11 / 5e-12 / 2e-12) and its
CPU/GPU parity (tests/test_cuda_solver_parity.py:68-105).
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = {torch.float32: (4e-4, 6e-5, 3e-5), torch.float64: (5e-11, 5e-12, 2e-12)}
# K = 4100 stages in LDS, K = 12000 (beyond a CU's 160 KiB) runs in scratch
CASES = [(101, 1, 1), (103, 3, 2), (107, 2, 7), (109, 4, 33), (113, 2, 65),
         (5, 3, 64), (6, 2, 129), (7, 1, 832), (8, 2, 4100), (9, 8, 3), (10, 1, 12000)]


def _bands(seed, B, K):
    rs = np.random.default_rng(seed)
    lower = 0.35 * rs.standard_normal((B, K - 1))
    upper = 0.35 * rs.standard_normal((B, K - 1))
    load = np.zeros((B, K))
    load[:, 1:] += np.abs(lower)
    load[:, :-1] += np.abs(upper)
    diag = load + 0.25 + rs.random((B, K))
    rhs = rs.standard_normal((B, K))
    return lower, diag, upper, rhs


def _backward_error(a, b, c, d, x):
    B, K = b.shape
    worst = 0.0
    for r in range(B):
        A = np.diag(b[r]) + np.diag(a[r], -1) + np.diag(c[r], 1)
        res = A @ x[r] - d[r]
        scale = np.abs(A) @ np.abs(x[r]) + np.abs(d[r])
        worst = max(worst, float((np.abs(res) / np.maximum(scale, 1e-30)).max()))
    return worst


@pytest.mark.parametrize("method", ["thomas", "pcr", "auto"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_tridiagonal_matches_oracle(orc, method, dtype):
    from dendra_amd.solvers import solve_tridiagonal

    rtol, atol, berr = TOL[dtype]
    for seed, B, K in CASES:
        a, b, c, d = _bands(seed, B, K)
        want = orc.thomas(a, b, c, d)
        t = [torch.tensor(x, dtype=dtype, device="cuda") for x in (a, b, c, d)]
        copies = [x.clone() for x in t]
        got = solve_tridiagonal(*t, method=method).double().cpu().numpy()
        for x, y in zip(t, copies):
            assert torch.equal(x, y), "inputs must not be modified"
        np.testing.assert_allclose(got, want, rtol=rtol, atol=atol,
                                   err_msg=f"K={K} B={B}")
        assert _backward_error(a, b, c, d, got) <= berr


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_lane_thomas_lds_many_short_systems(orc, dtype):
    """The LDS-staged lane-per-system sweep (K <= 32 in float32, 16 in
    float64): several workgroups of 256 systems and a partial last one, every
    row count up to the limit (odd and even row pitches), and the automatic
    choice for >= 4096 systems of <= 32 rows."""
    from dendra_amd.solvers import solve_tridiagonal

    rtol, atol, berr = TOL[dtype]
    kmax = 32 if dtype == torch.float32 else 16
    cases = [(785, K) for K in range(1, kmax + 1)] + [(4100, kmax)]
    for seed, (B, K) in enumerate(cases):
        a, b, c, d = _bands(100 + seed, B, K)
        want = orc.thomas(a, b, c, d)
        t = [torch.tensor(x, dtype=dtype, device="cuda") for x in (a, b, c, d)]
        got = solve_tridiagonal(*t, method="thomas" if B < 4096 else "auto")
        got = got.double().cpu().numpy()
        np.testing.assert_allclose(got, want, rtol=rtol, atol=atol, err_msg=f"K={K} B={B}")
        assert _backward_error(a[:64], b[:64], c[:64], d[:64], got[:64]) <= berr


def test_thomas_known_answer():
    """3x3 system of tests/test_voltage_solvers.py:98-128."""
    from dendra_amd.solvers import thomas_solve

    b = torch.tensor([[4.0, 5.0, 6.0]], device="cuda")
    a = torch.tensor([[1.0, 2.0]], device="cuda")
    c = torch.tensor([[3.0, 4.0]], device="cuda")
    d = torch.tensor([[1.0, 2.0, 3.0]], device="cuda")
    A = np.array([[4, 3, 0], [1, 5, 4], [0, 2, 6]], dtype=np.float64)
    want = np.linalg.solve(A, [1, 2, 3])
    np.testing.assert_allclose(thomas_solve(a, b, c, d).cpu().numpy()[0], want, atol=1e-6)


@pytest.mark.parametrize("solver", ["thomas_solve", "pcr_solve"])
def test_tridiagonal_gradients_match_finite_differences(solver):
    import dendra_amd.solvers as S

    fn = getattr(S, solver)
    g = torch.Generator().manual_seed(0)
    B, K = 2, 9
    b = ((torch.rand(B, K, generator=g, dtype=torch.float64) + 1.0) * 2).cuda()
    a = torch.rand(B, K - 1, generator=g, dtype=torch.float64).cuda()
    c = torch.rand(B, K - 1, generator=g, dtype=torch.float64).cuda()
    d = torch.randn(B, K, generator=g, dtype=torch.float64).cuda()
    for t in (a, b, c, d):
        t.requires_grad_(True)
    assert torch.autograd.gradcheck(fn, (a, b, c, d), eps=1e-6, atol=1e-5)


def _tree_inputs(batch, parent, dtype):
    size = len(parent)
    d_mem = np.linspace(1.3, 2.6, batch * size).reshape(batch, size)
    a_geom = np.linspace(0.015, 0.095, batch * size).reshape(batch, size)
    a_geom[:, 0] = 0.0
    rhs = np.linspace(-1.7, 2.1, batch * size).reshape(batch, size)
    return d_mem, a_geom, rhs


def _dense_tree(d_mem, a_geom, rhs, parent):
    out = np.empty_like(rhs)
    for r in range(d_mem.shape[0]):
        A = np.diag(d_mem[r] + a_geom[r])
        for ch, p in enumerate(parent):
            if p < 0:
                continue
            A[p, p] += a_geom[r, ch]
            A[ch, p] -= a_geom[r, ch]
            A[p, ch] -= a_geom[r, ch]
        out[r] = np.linalg.solve(A, rhs[r])
    return out


PARENT7 = [-1, 0, 0, 1, 1, 2, 5]  # tests/test_cuda_solver_parity.py:32


@pytest.mark.parametrize("layers", [False, True], ids=["drop-in", "layer-kernel"])
@pytest.mark.parametrize("threads", [1, 8, 32])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_dhs_matches_dense_and_oracle(orc, threads, dtype, layers):
    from dendra_amd import solvers
    from dendra_amd.topology import build_dhs_layers, build_morphology

    dhs_solve = solvers.dhs_solve_layers if layers else solvers.dhs_solve

    _, _, depth = build_morphology(PARENT7)
    order, layer_ptr = build_dhs_layers(depth, threads)
    d_mem, a_geom, rhs = _tree_inputs(5, PARENT7, dtype)
    dense = _dense_tree(d_mem, a_geom, rhs, PARENT7)
    ora = orc.dhs(d_mem, a_geom, rhs, PARENT7, order, layer_ptr)
    dev = lambda x, dt: torch.tensor(np.asarray(x), dtype=dt, device="cuda")
    got = dhs_solve(dev(d_mem, dtype), dev(a_geom, dtype), dev(rhs, dtype),
                    dev(PARENT7, torch.int64), dev(order, torch.int64),
                    dev(layer_ptr, torch.int64), threads=threads)
    rtol = 5e-5 if dtype == torch.float32 else 3e-12
    np.testing.assert_allclose(got.double().cpu().numpy(), dense, rtol=rtol, atol=rtol)
    np.testing.assert_allclose(ora, dense, rtol=1e-13, atol=1e-13)


def test_dhs_two_node_known_answer():
    """tests/test_voltage_solvers.py:22-49."""
    from dendra_amd.solvers import dhs_solve

    dev = lambda x, dt=torch.float32: torch.tensor(x, dtype=dt, device="cuda")
    x = dhs_solve(dev([[1.0, 2.0]]), dev([[0.0, 0.5]]), dev([[5.0, 3.0]]),
                  dev([-1, 0], torch.int64), dev([1, 0], torch.int64),
                  dev([0, 1, 2], torch.int64))
    np.testing.assert_allclose(x.cpu().numpy()[0], np.linalg.solve(
        [[1.5, -0.5], [-0.5, 2.5]], [5.0, 3.0]), atol=1e-6)


@pytest.mark.parametrize("layers", [False, True], ids=["drop-in", "layer-kernel"])
def test_dhs_large_random_tree_and_gradients(orc, layers):
    from dendra_amd import solvers
    from dendra_amd.topology import build_dhs_layers, build_morphology
    from dendra_amd.workloads import random_binary_tree

    dhs_solve = solvers.dhs_solve_layers if layers else solvers.dhs_solve

    parent, _, _ = random_binary_tree(3000, 0, 11)
    _, _, depth = build_morphology(parent)
    order, layer_ptr = build_dhs_layers(depth, 16)
    rs = np.random.default_rng(3)
    B, K = 3, len(parent)
    a_geom = 0.05 + rs.random((B, K))
    a_geom[:, 0] = 0
    d_mem = 0.1 + rs.random((B, K))
    rhs = rs.standard_normal((B, K))
    want = orc.dhs(d_mem, a_geom, rhs, parent, order, layer_ptr)
    dev = lambda x, dt=torch.float64: torch.tensor(np.asarray(x), dtype=dt, device="cuda")
    got = dhs_solve(dev(d_mem), dev(a_geom), dev(rhs), dev(parent, torch.int64),
                    dev(order, torch.int64), dev(layer_ptr, torch.int64))
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-11, atol=1e-11)
    # bitwise reproducible run to run (no atomics)
    again = dhs_solve(dev(d_mem), dev(a_geom), dev(rhs), dev(parent, torch.int64),
                      dev(order, torch.int64), dev(layer_ptr, torch.int64))
    assert torch.equal(got, again)
    # gradients against the dense oracle on the 7-node tree
    _, _, depth7 = build_morphology(PARENT7)
    o7, l7 = build_dhs_layers(depth7, 8)
    d7, a7, r7 = (dev(x).requires_grad_() for x in _tree_inputs(3, PARENT7, None))
    x = dhs_solve(d7, a7, r7, dev(PARENT7, torch.int64), dev(o7, torch.int64),
                  dev(l7, torch.int64), threads=8)
    w = torch.linspace(0.4, 1.6, x.numel(), device="cuda", dtype=torch.float64).reshape_as(x)
    got_g = torch.autograd.grad((x.square() * w).sum(), (d7, a7, r7))
    dc, ac, rc = (t.detach().cpu().requires_grad_() for t in (d7, a7, r7))
    A = torch.diag_embed(dc + ac)
    for ch, p in enumerate(PARENT7):
        if p < 0:
            continue
        e = ac[:, ch]
        A = A.clone()
        A[:, p, p] = A[:, p, p] + e
        A[:, ch, p] = A[:, ch, p] - e
        A[:, p, ch] = A[:, p, ch] - e
    xd = torch.linalg.solve(A, rc.unsqueeze(-1)).squeeze(-1)
    want_g = torch.autograd.grad((xd.square() * w.cpu()).sum(), (dc, ac, rc))
    for gg, wg in zip(got_g, want_g):
        np.testing.assert_allclose(gg.cpu().numpy(), wg.numpy(), rtol=3e-10, atol=3e-10)


def _chain_matrix(main, edge, parents):
    """Dense A = diag(main + edge + sum of children's edges) - offdiag(edge)."""
    B, K = main.shape
    A = torch.diag_embed(main + edge)
    for ch, p in enumerate(parents):
        if p < 0:
            continue
        e = edge[:, ch]
        A = A.clone()
        A[:, p, p] = A[:, p, p] + e
        A[:, ch, p] = A[:, ch, p] - e
        A[:, p, ch] = A[:, p, ch] - e
    return A


def _chain_topology(parents):
    dev = lambda x: torch.tensor(x, dtype=torch.int64, device="cuda")
    K = len(parents)
    return dev(list(parents)), dev(list(range(K - 1, -1, -1))), dev(list(range(K + 1)))


@pytest.mark.parametrize("layers", [False, True], ids=["drop-in", "layer-kernel"])
def test_dhs_float32_stiff_edge_keeps_the_uniform_solution(layers):
    """tests/func/test_cuda_solver_precision.py:46-65: a 2-node chain with an
    axial conductance of 2^26 keeps the passive uniform solution (all ones)
    finite and accurate in float32."""
    from dendra_amd import solvers

    solve = solvers.dhs_solve_layers if layers else solvers.dhs_solve
    main = torch.ones(1, 2, dtype=torch.float32, device="cuda")
    edge = torch.tensor([[0.0, float(2**26)]], dtype=torch.float32, device="cuda")
    x = solve(main, edge, main.clone(), *_chain_topology((-1, 0)), threads=32)
    assert x.dtype == torch.float32 and bool(torch.isfinite(x).all())
    torch.testing.assert_close(x.cpu().double(), torch.ones(1, 2, dtype=torch.float64),
                               rtol=2e-6, atol=2e-7)


@pytest.mark.parametrize("layers", [False, True], ids=["drop-in", "layer-kernel"])
@pytest.mark.parametrize("signed", [False, True], ids=["spd", "indefinite"])
def test_dhs_float32_schur_primal_gradients_and_root_placeholder(layers, signed):
    """tests/func/test_cuda_solver_precision.py:68-105: a 3-node chain, SPD or
    nonsingular indefinite (negative internal diagonal), against dense float64
    solves and autograd; the root's edge coefficient acts on its own diagonal,
    so its gradient equals the root diagonal's exactly."""
    from dendra_amd import solvers

    solve = solvers.dhs_solve_layers if layers else solvers.dhs_solve
    if signed:
        main = [[2.0, -0.5, 1.0], [2.4, -0.6, 1.3]]
        edge = [[0.2, 0.2, 0.1], [0.3, 0.15, 0.2]]
    else:
        main = [[1.0, 2.0, 1.5], [1.2, 1.8, 2.4]]
        edge = [[0.2, 0.3, 0.4], [0.15, 0.4, 0.2]]
    rhs = [[0.7, -1.2, 0.3], [-0.2, 0.9, 1.1]]
    values = [torch.tensor(v, dtype=torch.float32) for v in (main, edge, rhs)]
    parents = (-1, 0, 1)
    ref = [v.double().requires_grad_() for v in values]
    want = torch.linalg.solve(_chain_matrix(ref[0], ref[1], parents),
                              ref[2].unsqueeze(-1)).squeeze(-1)
    w = torch.tensor([[0.3, -0.7, 0.9], [-0.6, 0.8, 0.2]], dtype=torch.float64)
    want_g = torch.autograd.grad((want * w).sum(), ref)
    args = [v.cuda().requires_grad_() for v in values]
    got = solve(*args, *_chain_topology(parents), threads=32)
    got_g = torch.autograd.grad((got * w.cuda().float()).sum(), args)
    torch.testing.assert_close(got.detach().cpu().double(), want.detach(), rtol=1e-5,
                               atol=1e-6)
    for g_, wg in zip(got_g, want_g):
        torch.testing.assert_close(g_.cpu().double(), wg, rtol=2e-5, atol=2e-6)
    assert torch.equal(got_g[0][:, 0], got_g[1][:, 0])
