"""GPU parity of the 3x3 block solvers (dendra_amd.blocks) against dense
float64 solves of the same block matrices, with gradients against dense
autograd (the block tree matrix is the reference's own definition,
tests/func/test_cuda_functional_reference.py:57-76)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def block_tridiagonal_matrix(lower, main, upper):
    B, K = main.shape[:2]
    A = torch.zeros(B, 3 * K, 3 * K, dtype=main.dtype)
    for k in range(K):
        A[:, 3 * k:3 * k + 3, 3 * k:3 * k + 3] = main[:, k]
        if k > 0:
            A[:, 3 * k:3 * k + 3, 3 * k - 3:3 * k] = torch.diag_embed(lower[:, k - 1])
        if k + 1 < K:
            A[:, 3 * k:3 * k + 3, 3 * k + 3:3 * k + 6] = torch.diag_embed(upper[:, k])
    return A


def block_tree_matrix(main, edge, parents):
    B, K = main.shape[:2]
    A = torch.zeros(B, 3 * K, 3 * K, dtype=main.dtype)
    for i in range(K):
        A[:, 3 * i:3 * i + 3, 3 * i:3 * i + 3] = main[:, i] + torch.diag_embed(edge[:, i])
        p = parents[i]
        if p >= 0:
            e = torch.diag_embed(edge[:, i])
            A[:, 3 * p:3 * p + 3, 3 * p:3 * p + 3] = A[:, 3 * p:3 * p + 3, 3 * p:3 * p + 3] + e
            A[:, 3 * i:3 * i + 3, 3 * p:3 * p + 3] = -e
            A[:, 3 * p:3 * p + 3, 3 * i:3 * i + 3] = -e
    return A


def dense_solve(A, rhs):
    return torch.linalg.solve(A, rhs.reshape(rhs.shape[0], -1, 1)).reshape(rhs.shape)


def gen_tridiagonal(B, K, rng, spd=False):
    lower = rng.uniform(-1, 1, (B, max(K - 1, 0), 3))
    upper = lower.copy() if spd else rng.uniform(-1, 1, (B, max(K - 1, 0), 3))
    m = rng.uniform(-0.5, 0.5, (B, K, 3, 3))
    if spd:
        m = m @ np.swapaxes(m, -1, -2)
    main = m + np.eye(3) * 4.0
    rhs = rng.standard_normal((B, K, 3))
    return [torch.as_tensor(v) for v in (lower, main, upper, rhs)]


def random_parents(K, rng):
    return [-1] + [int(rng.integers(max(0, i - 4), i)) for i in range(1, K)]


@pytest.mark.parametrize("K", [1, 2, 7, 300])
def test_block_thomas_matches_dense_with_gradients(K):
    from dendra_amd.blocks import thomas_solve_cuda_bt

    rng = np.random.default_rng(K)
    vals = gen_tridiagonal(5, K, rng)
    want = dense_solve(block_tridiagonal_matrix(*vals[:3]), vals[3])
    args = [v.cuda() for v in vals]
    snap = [a.clone() for a in args]
    x = thomas_solve_cuda_bt(*args)
    assert (x.cpu() - want).abs().max() <= 1e-11 * max(1.0, want.abs().max())
    assert all(torch.equal(a, b) for a, b in zip(args, snap))
    xf = thomas_solve_cuda_bt(*[a.float() for a in args])
    A = block_tridiagonal_matrix(*vals[:3])
    res = (A @ xf.double().cpu().reshape(5, -1, 1)).reshape(vals[3].shape) - vals[3]
    assert res.abs().max() <= 64 * np.finfo(np.float32).eps * (
        A.abs().sum(-1).max() * xf.abs().max().cpu() + vals[3].abs().max())
    if K > 1:
        w = torch.as_tensor(rng.standard_normal((5, K, 3)))
        ref = [v.clone().requires_grad_() for v in vals]
        exp = torch.autograd.grad(
            (dense_solve(block_tridiagonal_matrix(*ref[:3]), ref[3]) * w).sum(), ref)
        gin = [a.clone().requires_grad_() for a in args]
        got = torch.autograd.grad((thomas_solve_cuda_bt(*gin) * w.cuda()).sum(), gin)
        for g_, e_ in zip(got, exp):
            torch.testing.assert_close(g_.cpu(), e_, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("K", [1, 5, 200])
def test_block_spd_matches_dense_with_gradients(K):
    from dendra_amd.blocks import block_spd_solve, solve_bt_spd_fwd, solve_bt_spd_with_chol

    rng = np.random.default_rng(100 + K)
    vals = gen_tridiagonal(4, K, rng, spd=True)
    A = block_tridiagonal_matrix(*vals[:3])
    assert torch.linalg.eigvalsh(A).min() > 0
    want = dense_solve(A, vals[3])
    args = [v.cuda() for v in vals]
    x, chol = solve_bt_spd_fwd(*args)
    assert (x.cpu() - want).abs().max() <= 1e-11 * max(1.0, want.abs().max())
    other = torch.as_tensor(rng.standard_normal((4, K, 3)))
    x2 = solve_bt_spd_with_chol(args[0], args[2], chol, other.cuda())
    assert (x2.cpu() - dense_solve(A, other)).abs().max() <= 1e-11
    w = torch.as_tensor(rng.standard_normal((4, K, 3)))
    ref = [v.clone().requires_grad_() for v in vals]
    exp = torch.autograd.grad(
        (dense_solve(block_tridiagonal_matrix(*ref[:3]), ref[3]) * w).sum(), ref,
        allow_unused=True)  # K = 1: the empty bands do not enter the matrix
    gin = [a.clone().requires_grad_() for a in args]
    got = torch.autograd.grad((block_spd_solve(*gin) * w.cuda()).sum(), gin)
    sym = lambda m: 0.5 * (m + m.transpose(-1, -2))
    torch.testing.assert_close(got[1].cpu(), sym(exp[1]), rtol=1e-9, atol=1e-11)
    for i in (0, 2, 3):
        if exp[i] is None:
            assert got[i].numel() == 0
        else:
            torch.testing.assert_close(got[i].cpu(), exp[i], rtol=1e-9, atol=1e-11)


def test_block_spd_shifts_an_indefinite_block_and_rejects_nan():
    """One indefinite block: the solve is that of the block shifted by its
    Gershgorin deficit plus max(eps_abs, eps_rel * sum |a_rr|)."""
    from dendra_amd.blocks import solve_bt_spd_fwd

    m = torch.tensor([[[[1.0, 2.0, 0.0], [2.0, 1.0, 0.5], [0.0, 0.5, 3.0]]]],
                     dtype=torch.float64)
    rhs = torch.tensor([[[1.0, -1.0, 0.5]]], dtype=torch.float64)
    empty = torch.zeros(1, 0, 3, dtype=torch.float64)
    x, _ = solve_bt_spd_fwd(empty.cuda(), m.cuda(), empty.cuda(), rhs.cuda())
    a = m[0, 0].numpy()
    gmin = min(a[r, r] - sum(abs(a[r, c]) for c in range(3) if c != r) for r in range(3))
    margin = max(1e-30, 1e-12 * np.abs(np.diag(a)).sum())
    tau = margin - gmin
    want = np.linalg.solve(a + tau * np.eye(3), rhs[0, 0].numpy())
    np.testing.assert_allclose(x[0, 0].cpu().numpy(), want, rtol=1e-12)
    bad = m.clone()
    bad[0, 0, 1, 1] = float("nan")
    with pytest.raises(RuntimeError, match="Cholesky failed"):
        solve_bt_spd_fwd(empty.cuda(), bad.cuda(), empty.cuda(), rhs.cuda())


@pytest.mark.parametrize("K", [1, 2, 9, 40, 400])
@pytest.mark.parametrize("kind", ["near", "binary", "path"])
def test_block_tree_matches_dense_with_gradients(K, kind):
    """Layered trees (binary; layers average two or more nodes) take the
    wavefront-per-system layer kernel, paths and single nodes the
    lane-per-system kernel; both against dense float64 solves."""
    from dendra_amd.blocks import dhs_bt_solve
    from dendra_amd.topology import build_dhs_layers, build_morphology

    rng = np.random.default_rng(200 + K)
    if kind == "near":
        parents = random_parents(K, rng)
    elif kind == "binary":
        parents = [-1] + [(i - 1) // 2 for i in range(1, K)]
    else:
        parents = list(range(-1, K - 1))
    _, _, depth = build_morphology(parents)
    order, layers = build_dhs_layers(depth, 16)
    main = torch.as_tensor(rng.uniform(-0.3, 0.3, (3, K, 3, 3)) + np.eye(3) * 2.0)
    edge = torch.as_tensor(rng.uniform(0.1, 2.0, (3, K, 3)))
    rhs = torch.as_tensor(rng.standard_normal((3, K, 3)))
    want = dense_solve(block_tree_matrix(main, edge, parents), rhs)
    topo = [torch.as_tensor(np.asarray(v), dtype=torch.int64, device="cuda")
            for v in (parents, order, layers)]
    args = [v.cuda() for v in (main, edge, rhs)]
    x = dhs_bt_solve(*args, *topo)
    assert (x.cpu() - want).abs().max() <= 1e-11 * max(1.0, want.abs().max())
    xf = dhs_bt_solve(*[a.float() for a in args], *topo)
    assert (xf.double().cpu() - want).abs().max() <= 1e-4 * max(1.0, want.abs().max())
    if K <= 40:
        w = torch.as_tensor(rng.standard_normal((3, K, 3)))
        ref = [v.clone().requires_grad_() for v in (main, edge, rhs)]
        exp = torch.autograd.grad(
            (dense_solve(block_tree_matrix(ref[0], ref[1], parents), ref[2]) * w).sum(), ref)
        gin = [a.clone().requires_grad_() for a in args]
        got = torch.autograd.grad((dhs_bt_solve(*gin, *topo) * w.cuda()).sum(), gin)
        for g_, e_ in zip(got, exp):
            torch.testing.assert_close(g_.cpu(), e_, rtol=1e-9, atol=1e-11)


def test_block_validation_messages():
    from dendra_amd.blocks import dhs_bt_solve, thomas_solve_cuda_bt

    z = lambda *s: torch.zeros(*s, device="cuda")
    with pytest.raises(ValueError, match=r"rhs must have shape \(B, K, 3\)"):
        thomas_solve_cuda_bt(z(2, 3, 3), z(2, 4, 3, 3), z(2, 3, 3), z(2, 4))
    with pytest.raises(ValueError, match="lower must have shape"):
        thomas_solve_cuda_bt(z(2, 4, 3), z(2, 4, 3, 3), z(2, 3, 3), z(2, 4, 3))
    with pytest.raises(ValueError, match="main must have shape"):
        thomas_solve_cuda_bt(z(2, 3, 3), z(2, 4, 9), z(2, 3, 3), z(2, 4, 3))
    with pytest.raises(TypeError, match="float32 and float64"):
        thomas_solve_cuda_bt(*(t.half() for t in (z(2, 3, 3), z(2, 4, 3, 3), z(2, 3, 3),
                                                  z(2, 4, 3))))
    idx = torch.zeros(4, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError, match="edge must have shape"):
        dhs_bt_solve(z(2, 4, 3, 3), z(2, 4), z(2, 4, 3), idx, idx, idx[:2])
    with pytest.raises(ValueError, match="parent_idx and order must each contain K"):
        dhs_bt_solve(z(2, 4, 3, 3), z(2, 4, 3), z(2, 4, 3), idx[:3], idx, idx[:2])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_block_tree_ring_and_workspace_forms_agree_bitwise(dtype):
    """Whole depth levels of <= 64 nodes take the LDS-ring form; the same
    tree with its levels split into chunks of three fails the kernel's check
    and takes the workspace form (still the layer kernel: at least two nodes
    per layer on average).  Both follow the same elimination; the compiler
    may contract their products into fused multiply-adds differently, so the
    bar between them is a few units of rounding, and each matches a dense
    float64 solve."""
    from dendra_amd import _lib
    from dendra_amd.blocks import _children_csr
    from dendra_amd.topology import build_dhs_layers, build_morphology
    from dendra_amd.workloads import random_binary_tree

    parents = [int(p) for p in random_binary_tree(300, 0, 4)[0]]
    K, B = len(parents), 5
    _, _, depth = build_morphology(parents)
    rng = np.random.default_rng(9)
    main = torch.as_tensor(rng.uniform(-0.3, 0.3, (B, K, 3, 3)) + np.eye(3) * 2.0)
    edge = torch.as_tensor(rng.uniform(0.1, 2.0, (B, K, 3)))
    rhs = torch.as_tensor(rng.standard_normal((B, K, 3)))
    want = dense_solve(block_tree_matrix(main, edge, parents), rhs)
    par = torch.as_tensor(parents, dtype=torch.int64, device="cuda")
    cp, ci = _children_csr(par)
    args = [v.to("cuda", dtype).contiguous() for v in (main, edge, rhs)]
    lib = _lib.load()
    fn = lib.dn_bt3_tree_f64 if dtype == torch.float64 else lib.dn_bt3_tree_f32
    n = int(lib.dn_block_work_elems(2, B, K))
    outs = []
    for chunk in (K, 3):
        order, layers = build_dhs_layers(depth, chunk)
        assert (np.diff(layers) <= 64).all() and 2 * (len(layers) - 1) <= K
        o = torch.as_tensor(order, dtype=torch.int64, device="cuda")
        lp = torch.as_tensor(layers, dtype=torch.int64, device="cuda")
        x = torch.empty_like(args[2])
        work = torch.empty(n, dtype=dtype, device="cuda")
        p = _lib.ptr
        st = fn(p(args[0]), p(args[1]), p(args[2]), p(par), p(o), p(lp), lp.numel() - 1,
                p(cp), p(ci), p(x), B, K, p(work), n, _lib.stream_handle(args[2].device))
        _lib.check(st, "block tree solve")
        outs.append(x)
    eps = torch.finfo(dtype).eps
    scale = outs[1].abs().max().item()
    assert (outs[0] - outs[1]).abs().max().item() <= 64 * eps * max(1.0, scale)
    tol = 1e-11 if dtype == torch.float64 else 1e-4
    assert (outs[0].double().cpu() - want).abs().max() <= tol * max(1.0, want.abs().max())
