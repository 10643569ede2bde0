"""Host side of the tree-contraction solve: plan validation, the rake/compress
schedule's invariants, and a float64 replay of the device algorithm against
a dense solve (CPU only)."""

import math

import numpy as np
import pytest

from dendra_amd.contraction import TreeContractionPlan, contraction_schedule


def random_tree(n, rng, kind="random"):
    if kind == "path":
        return np.arange(-1, n - 1)
    if kind == "star":
        return np.concatenate([[-1], np.zeros(n - 1, dtype=np.int64)])
    if kind == "caterpillar":  # a spine with one leaf on every spine node
        par = [-1]
        spine = 0
        while len(par) < n:
            par.append(spine)          # leaf
            if len(par) < n:
                par.append(spine)      # next spine node
                spine = len(par) - 1
        return np.asarray(par[:n])
    if kind == "binary":
        return np.concatenate([[-1], (np.arange(1, n) - 1) // 2])
    # random recursive tree with long chains: attach to the last node with
    # probability 0.8, else to a random earlier node
    par = [-1]
    for i in range(1, n):
        par.append(i - 1 if rng.random() < 0.8 else int(rng.integers(0, i)))
    return np.asarray(par)


def shuffled(par, rng):
    """The same tree with nodes relabelled at random (root anywhere)."""
    n = len(par)
    lab = rng.permutation(n)
    out = np.empty(n, dtype=np.int64)
    out[lab] = np.where(par >= 0, lab[np.maximum(par, 0)], -1)
    return out


def dense_matrix(d, g, parent):
    n = len(d)
    A = np.diag(np.asarray(d, dtype=np.float64) + g)
    for i, p in enumerate(parent):
        if p >= 0:
            A[p, p] += g[i]
            A[i, p] -= g[i]
            A[p, i] -= g[i]
    return A


def replay(plan, d, g, rhs):
    """Float64 replay of the kernel's phases (condensation, rounds, recovery)
    on the plan's tables."""
    h = plan._host
    perm, head, ln = h["perm"], h["sec_head"], h["sec_len"]
    ch0, nch, spar = h["sec_child"], h["sec_nchild"], h["sec_parent"]
    n = plan.n_sections
    SM, SF, SE, PM, PF = (np.zeros(n) for _ in range(5))
    CM, CF, CE, XS = (np.zeros(n) for _ in range(4))
    wa, wb, wc = (np.zeros(plan.size) for _ in range(3))
    for s in range(n):
        o = perm[head[s]]
        e, Mc, Fc, dM, dF = g[o], d[o], rhs[o], 0.0, 0.0
        for r in range(ln[s] - 1):
            o1 = perm[head[s] + r + 1]
            inv = 1.0 / (Mc + e + g[o1])
            be, ce = e * inv, g[o1] * inv
            dM += be * Mc
            dF += be * Fc
            wa[head[s] + r], wb[head[s] + r], wc[head[s] + r] = Fc * inv, be, ce
            Mc, Fc, e = d[o1] + ce * Mc, rhs[o1] + ce * Fc, e * ce
        SM[s], SF[s], SE[s], PM[s], PF[s] = Mc, Fc, e, dM, dF
    for s in range(n):
        for c in range(ch0[s], ch0[s] + nch[s]):
            SM[s] += PM[c]
            SF[s] += PF[c]
    sc = plan.schedule
    for r in range(sc.n_rounds):
        for i in range(sc.ptr[r], sc.ptr[r + 1]):
            j, c = sc.sec[i], sc.child[i]
            if c < 0:
                inv = 1.0 / (SM[j] + SE[j])
                fac = SE[j] * inv
                PM[j], PF[j] = fac * SM[j], fac * SF[j]
                SF[j], SM[j], SE[j] = SF[j] * inv, fac, 0.0
            else:
                inv = 1.0 / (SM[j] + SE[j] + SE[c])
                bj, cj = SE[j] * inv, SE[c] * inv
                PM[j], PF[j] = bj * SM[j], bj * SF[j]
                CM[j], CF[j], CE[j] = cj * SM[j], cj * SF[j], SE[j] * cj
                SF[j], SM[j], SE[j] = SF[j] * inv, bj, cj
        for i in range(sc.rptr[r], sc.rptr[r + 1]):
            q = sc.recv[i]
            for k in range(sc.sptr[i], sc.sptr[i + 1]):
                sj = sc.src[k]
                if sj >= 0:
                    SM[q] += PM[sj]
                    SF[q] += PF[sj]
                else:
                    j = -1 - sj
                    SM[q] += CM[j]
                    SF[q] += CF[j]
                    SE[q] = CE[j]
    XS[0] = SF[0] / (SM[0] + SE[0])
    for r in range(sc.n_rounds - 1, -1, -1):
        for i in range(sc.ptr[r], sc.ptr[r + 1]):
            j, p, c = sc.sec[i], sc.par[i], sc.child[i]
            XS[j] = SF[j] + SM[j] * XS[p] + (SE[j] * XS[c] if c >= 0 else 0.0)
    x = np.zeros(plan.size)
    for s in range(n):
        xp = XS[spar[s]] if spar[s] >= 0 else 0.0
        xv = XS[s]
        x[perm[head[s] + ln[s] - 1]] = xv
        for r in range(ln[s] - 2, -1, -1):
            xv = wa[head[s] + r] + wb[head[s] + r] * xp + wc[head[s] + r] * xv
            x[perm[head[s] + r]] = xv
    return x


def check_schedule(par):
    """Every non-root node is eliminated once; each round's eliminated set is
    independent; its neighbours survive the round and pull from it."""
    par = np.asarray(par)
    n = len(par)
    sc = contraction_schedule(par)
    cur = par.copy()
    alive = np.ones(n, dtype=bool)
    seen = np.zeros(n, dtype=int)
    for r in range(sc.n_rounds):
        gone = sc.sec[sc.ptr[r]:sc.ptr[r + 1]]
        seen[gone] += 1
        gset = set(gone.tolist())
        assert 0 not in gset
        for i in range(sc.ptr[r], sc.ptr[r + 1]):
            j, p, c = sc.sec[i], sc.par[i], sc.child[i]
            assert alive[j] and p == cur[j] and p not in gset
            kids = np.flatnonzero(alive & (cur == j))
            if c < 0:
                assert len(kids) == 0
            else:
                assert list(kids) == [c] and c not in gset
                assert (alive & (cur == c)).any(), "a compressed node's child is no leaf"
        pulls = {}
        for i in range(sc.rptr[r], sc.rptr[r + 1]):
            pulls[int(sc.recv[i])] = sorted(int(v) for v in sc.src[sc.sptr[i]:sc.sptr[i + 1]])
        want = {}
        for i in range(sc.ptr[r], sc.ptr[r + 1]):
            j, p, c = int(sc.sec[i]), int(sc.par[i]), int(sc.child[i])
            want.setdefault(p, []).append(j)
            if c >= 0:
                want.setdefault(c, []).append(-1 - j)
        assert pulls == {k: sorted(v) for k, v in want.items()}
        for i in range(sc.ptr[r], sc.ptr[r + 1]):
            if sc.child[i] >= 0:
                cur[sc.child[i]] = sc.par[i]
        alive[gone] = False
    assert seen[0] == 0 and (seen[1:] == 1).all()
    return sc


@pytest.mark.parametrize("kind", ["random", "path", "star", "caterpillar", "binary"])
def test_schedule_invariants(kind):
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 7, 64, 301):
        check_schedule(random_tree(n, rng, kind))


def test_schedule_rounds_are_logarithmic_on_paths_and_binary_trees():
    for n in (2, 16, 1000, 4096):
        sc = check_schedule(np.arange(-1, n - 1))
        assert sc.n_rounds <= 2 * math.ceil(math.log2(n)) + 1, (n, sc.n_rounds)
    sc = check_schedule(random_tree(4095, None, "binary"))
    assert sc.n_rounds <= 12


@pytest.mark.parametrize("kind", ["random", "path", "star", "caterpillar", "binary"])
@pytest.mark.parametrize("seg", [16, 1, 3, None])
def test_replay_matches_dense_solve(kind, seg):
    rng = np.random.default_rng(11)
    for n in (1, 2, 5, 40, 257):
        par = shuffled(random_tree(n, rng, kind), rng)
        plan = TreeContractionPlan(par, max_segment_interior_nodes=seg)
        d = rng.uniform(0.01, 2.0, n)
        g = rng.uniform(0.1, 5.0, n)
        g[par < 0] = rng.choice([0.0, 0.7])
        rhs = rng.standard_normal(n)
        want = np.linalg.solve(dense_matrix(d, g, par), rhs)
        got = replay(plan, d, g, rhs)
        assert np.abs(got - want).max() <= 1e-10 * max(1.0, np.abs(want).max())


def test_plan_sections_bound_chain_length():
    par = np.arange(-1, 99)
    for seg, n_sec in ((16, 6), (1, 50), (None, 100)):
        plan = TreeContractionPlan(par, max_segment_interior_nodes=seg)
        assert plan.n_sections == n_sec
        assert plan._host["sec_len"].max() <= (1 if seg is None else seg + 1)


def test_plan_validation_matches_reference_messages():
    with pytest.raises(TypeError, match="positive integer or None"):
        TreeContractionPlan([-1, 0], max_segment_interior_nodes=True)
    with pytest.raises(TypeError, match="positive integer or None"):
        TreeContractionPlan([-1, 0], max_segment_interior_nodes=2.0)
    with pytest.raises(ValueError, match="must be positive"):
        TreeContractionPlan([-1, 0], max_segment_interior_nodes=0)
    with pytest.raises(ValueError, match="exactly one root"):
        TreeContractionPlan([-1, -1])
    with pytest.raises(ValueError, match="cycle"):
        TreeContractionPlan([-1, 2, 1])
    with pytest.raises(ValueError, match="own parent"):
        TreeContractionPlan([-1, 1])
