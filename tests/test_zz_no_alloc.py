"""The allocation-free tridiagonal entry point (dn_tridiag_solve_noalloc_f32)
against dn_tridiag_solve_f32 with caller scratch, through the C ABI: the
same kernels must give bitwise the same solutions, on two calls in a row with
the second smaller, for both scratch-using methods; oversize solves are
refused with DN_ERANGE before any launch."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DN_ERANGE = -3


def _system(rs, B, K):
    a = rs.uniform(-0.4, 0.4, (B, K - 1))
    c = rs.uniform(-0.4, 0.4, (B, K - 1))
    b = 1.0 + rs.random((B, K)) + np.pad(np.abs(a), ((0, 0), (1, 0))) + \
        np.pad(np.abs(c), ((0, 0), (0, 1)))
    d = rs.standard_normal((B, K))
    return [torch.as_tensor(v, dtype=torch.float32, device="cuda") for v in (a, b, c, d)]


def _with_scratch(lib, p, a, b, c, d, B, K, method):
    x = torch.empty((B, K), dtype=torch.float32, device="cuda")
    n = lib.dn_tridiag_scratch_elems(B, K, method, 4)
    scratch = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
    st = lib.dn_tridiag_solve_f32(p(a), p(b), p(c), p(d), p(x), B, K, method, p(scratch), n,
                                  torch.cuda.current_stream().cuda_stream)
    assert st == 0
    return x


def _no_alloc(lib, p, a, b, c, d, B, K, method):
    x = torch.empty((B, K), dtype=torch.float32, device="cuda")
    st = lib.dn_tridiag_solve_noalloc_f32(p(a), p(b), p(c), p(d), p(x), B, K, method,
                                          torch.cuda.current_stream().cuda_stream)
    assert st == 0
    return x


@pytest.mark.parametrize("method,shapes", [(1, [(16384, 64), (1000, 40)]),
                                           (2, [(40, 12000), (9, 10300)])])
def test_no_alloc_solve_equals_scratch_solve(method, shapes):
    from dendra_amd import _lib

    lib, p = _lib.load(), _lib.ptr
    cap = lib.dn_tridiag_noalloc_max_scratch_elems()
    rs = np.random.default_rng(method)
    for B, K in shapes:
        assert 0 < lib.dn_tridiag_scratch_elems(B, K, method, 4) <= cap
        a, b, c, d = _system(rs, B, K)
        want = _with_scratch(lib, p, a, b, c, d, B, K, method)
        got = _no_alloc(lib, p, a, b, c, d, B, K, method)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        # the solve is right, not only repeatable: residual of the system
        r = b * got
        r[:, 1:] += a * got[:, :-1]
        r[:, :-1] += c * got[:, 1:]
        assert float((r - d).abs().max()) < 1e-4


def test_no_alloc_refuses_oversize_and_runs_without_scratch():
    from dendra_amd import _lib

    lib, p = _lib.load(), _lib.ptr
    cap = lib.dn_tridiag_noalloc_max_scratch_elems()
    B, K = cap // 64 + 1, 64
    assert lib.dn_tridiag_scratch_elems(B, K, 1, 4) > cap
    st = lib.dn_tridiag_solve_noalloc_f32(None, None, None, None, None, B, K, 1,
                                          torch.cuda.current_stream().cuda_stream)
    assert st == DN_ERANGE
    rs = np.random.default_rng(3)
    a, b, c, d = _system(rs, 300, 201)   # LDS-staged: no workspace at all
    assert lib.dn_tridiag_scratch_elems(300, 201, 2, 4) == 0
    got = _no_alloc(lib, p, a, b, c, d, 300, 201, 2)
    want = _with_scratch(lib, p, a, b, c, d, 300, 201, 2)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    # lane-per-system sweeps of <= 32 rows stage in LDS: no workspace either
    a, b, c, d = _system(rs, 70000, 32)
    assert lib.dn_tridiag_scratch_elems(70000, 32, 1, 4) == 0
    got = _no_alloc(lib, p, a, b, c, d, 70000, 32, 1)
    want = _with_scratch(lib, p, a, b, c, d, 70000, 32, 1)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
