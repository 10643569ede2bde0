"""Packed multi-morphology tree solve (dendra_amd.solvers.dhs_solve_multi,
the dhs_solve_multi_cuda drop-in) against dense float64 solves of every
covered row, the oracle's DHS, zero padding, and gradients."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _morphs():
    from dendra_amd.workloads import random_binary_tree

    trees = [random_binary_tree(n, k, 3)[0] for k, n in enumerate((300, 57, 1000))]
    path = np.arange(-1, 199)                       # a 200-node path
    star = np.concatenate([[-1], np.zeros(40, np.int64)])
    return trees + [path, star, np.array([-1])]


def _case(dtype, threads, rows=(5, 1, 7, 3, 2, 4), extra_rows=3, seed=0):
    from dendra_amd.topology import multi_dhs_plan

    morphs = _morphs()
    plan = multi_dhs_plan(morphs, rows, threads)
    B, W = plan["rows_total"] + extra_rows, plan["K_stride"]
    rs = np.random.default_rng(seed)
    d = np.zeros((B, W)); a = np.zeros((B, W)); b = np.zeros((B, W))
    row = 0
    owner = []
    for par, n in zip(morphs, rows):
        K = len(par)
        for _ in range(n):
            d[row, :K] = rs.uniform(0.05, 1.0, K)
            a[row, :K] = np.where(par >= 0, rs.uniform(0.5, 20.0, K), 0.0)
            b[row, :K] = rs.standard_normal(K)
            owner.append((row, par))
            row += 1
    # garbage in the padding must not leak into the result
    for r in range(B):
        K = len(owner[r][1]) if r < len(owner) else 0
        d[r, K:] = 7.0; a[r, K:] = 3.0; b[r, K:] = 11.0
    t = lambda v: torch.as_tensor(v, dtype=dtype, device="cuda")
    tabs = {k: torch.as_tensor(plan[k], device="cuda") for k in
            ("P_cat", "ORDER_cat", "LAYER_PTR_cat", "WARP_P_OFF", "WARP_ORDER_OFF",
             "WARP_LPTR_OFF", "WARP_L", "WARP_ROW_BASE", "WARP_ROW_COUNT")}
    return plan, tabs, (d, a, b), (t(d), t(a), t(b)), owner


def _dense(par, d, a, b):
    K = len(par)
    A = np.diag(d + a)
    for i, p in enumerate(par):
        if p >= 0:
            A[p, p] += a[i]
            A[i, p] -= a[i]
            A[p, i] -= a[i]
    return np.linalg.solve(A, b)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("threads", [1, 8, 32])
def test_multi_solve_matches_dense_and_zero_pads(orc, dtype, threads):
    from dendra_amd.solvers import dhs_solve_multi_cuda

    plan, tabs, (d, a, b), (td, ta, tb), owner = _case(dtype, threads)
    x = dhs_solve_multi_cuda(td, ta, tb, *tabs.values(), K_stride=plan["K_stride"],
                             L_max=plan["L_max"], threads=threads, grid_x=plan["grid_x"])
    got = x.double().cpu().numpy()
    for r, par in owner:
        K = len(par)
        want = _dense(par, d[r, :K], a[r, :K], b[r, :K])
        scale = np.abs(want).max()
        tol = 1e-11 if dtype == torch.float64 else 2e-5
        assert np.abs(got[r, :K] - want).max() <= tol * scale, (r, K)
        assert not got[r, K:].any()
    assert not got[len(owner):].any()
    if dtype == torch.float64:
        # the oracle's DHS on one row of the 1000-node tree, same layers
        from dendra_amd.topology import build_dhs_layers, build_morphology

        r, par = next((r, p) for r, p in owner if len(p) == 1000)
        _, _, depth = build_morphology(par)
        order, lptr = build_dhs_layers(depth, threads)
        want = orc.dhs(d[r:r + 1, :1000], a[r:r + 1, :1000], b[r:r + 1, :1000], par, order,
                       lptr)
        assert np.abs(got[r, :1000] - want[0]).max() < 1e-12 * np.abs(want).max()


def test_multi_solve_gradients():
    from dendra_amd.solvers import dhs_solve_multi

    rows = (2, 1, 1, 2, 1, 1)
    plan, tabs, _, (td, ta, tb), owner = _case(torch.float64, 16, rows=rows, extra_rows=1)
    args = [t.clone().requires_grad_() for t in (td, ta, tb)]
    kw = dict(K_stride=plan["K_stride"], L_max=plan["L_max"], threads=16)
    x = dhs_solve_multi(*args, *tabs.values(), **kw)
    w = torch.randn_like(x)
    gd, ga, gb = torch.autograd.grad((x * w).sum(), args)
    # reference: dense autograd of every covered row, padding gradients zero
    for r, par in owner:
        K = len(par)
        dd, aa, bb = (t.detach()[r, :K].clone().requires_grad_() for t in args)
        A = torch.diag(dd + aa)
        pi = torch.as_tensor(par)
        has = pi >= 0
        idx = torch.arange(K)[has]
        A = A.index_put((pi[has], pi[has]), aa[has], accumulate=True)
        A = A.index_put((idx, pi[has]), -aa[has], accumulate=True)
        A = A.index_put((pi[has], idx), -aa[has], accumulate=True)
        xr = torch.linalg.solve(A, bb)
        ref = torch.autograd.grad((xr * w[r, :K]).sum(), (dd, aa, bb))
        for g_, want in zip((gd, ga, gb), ref):
            assert torch.allclose(g_[r, :K], want, rtol=1e-9, atol=1e-11)
            assert not g_[r, K:].any()
    for g_ in (gd, ga, gb):
        assert not g_[len(owner):].any()


def test_multi_solve_contract_errors():
    from dendra_amd.solvers import dhs_solve_multi

    plan, tabs, _, (td, ta, tb), _ = _case(torch.float32, 16)
    kw = dict(K_stride=plan["K_stride"], L_max=plan["L_max"], threads=16)
    with pytest.raises(ValueError, match="K_stride"):
        dhs_solve_multi(td, ta, tb, *tabs.values(), K_stride=plan["K_stride"] + 1,
                        L_max=plan["L_max"], threads=16)
    with pytest.raises(ValueError, match="grid_x"):
        dhs_solve_multi(td, ta, tb, *tabs.values(), grid_x=plan["grid_x"] + 1, **kw)
    bad = dict(tabs)
    bad["WARP_L"] = bad["WARP_L"][:-1]
    with pytest.raises(ValueError, match="one entry per plan row"):
        dhs_solve_multi(td, ta, tb, *bad.values(), **kw)
    with pytest.raises(TypeError, match="threads"):
        dhs_solve_multi(td, ta, tb, *tabs.values(), K_stride=plan["K_stride"],
                        L_max=plan["L_max"], threads=3.0)
