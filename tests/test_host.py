"""Host-side logic: topology preprocessing, plans, workspaces, RNG, C ABI.

CPU only.  The library checks load libdnhip.so and compare its exported
symbols with include/dendra_hip.h; no kernel is called.
"""

import ctypes
import re
from pathlib import Path

import networkx as nx
import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_header_symbols_are_exported_and_bound():
    from dendra_amd import _lib

    hdr = (ROOT / "include" / "dendra_hip.h").read_text()
    declared = set(re.findall(r"\b(dn_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared but not exported"
    assert declared == set(_lib.exported_symbols()), "ctypes binding out of sync"


def test_library_load_fails_loudly_when_missing(monkeypatch, tmp_path):
    from dendra_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load()


def test_build_dhs_layers_matches_reference_semantics():
    from dendra_amd.topology import build_dhs_layers, build_morphology

    parent = [-1, 0, 0, 1, 1, 2, 5]
    _, children, depth = build_morphology(parent)
    assert depth.tolist() == [0, 1, 1, 2, 2, 2, 3]
    order, layer_ptr = build_dhs_layers(depth, 2)
    # deepest first, chunks of at most 2 nodes, ascending node order per depth
    assert order.tolist() == [6, 3, 4, 5, 1, 2, 0]
    assert layer_ptr.tolist() == [0, 1, 3, 4, 6, 7]
    pos = {n: i for i, n in enumerate(order.tolist())}
    for i, p in enumerate(parent):
        if p >= 0:
            assert pos[i] < pos[p], "children eliminated before parents"


@pytest.mark.parametrize("bad,msg", [
    ([], "at least one"), ([-1, -1], "exactly one root"), ([-1, 5], "must be in"),
    ([-1, 1], "own parent"), ([2, 2, -1, -1], "exactly one root"),
])
def test_build_morphology_rejects_invalid(bad, msg):
    from dendra_amd.topology import build_morphology

    with pytest.raises(ValueError, match=msg):
        build_morphology(bad)


def test_build_morphology_rejects_cycles():
    from dendra_amd.topology import build_morphology

    with pytest.raises(ValueError, match="cycle|unreachable"):
        build_morphology([-1, 2, 1])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_section_plan_invariants(seed):
    from dendra_amd.topology import SEC_MAX, section_plan
    from dendra_amd.workloads import random_binary_tree

    parent, _, _ = random_binary_tree(1500, seed, 17)
    sp = section_plan(parent)
    K = len(parent)
    assert sorted(sp.perm.tolist()) == list(range(K))
    # topology preserved under renumbering
    for new_i, old_i in enumerate(sp.perm):
        p_old = parent[old_i]
        p_new = sp.parent[new_i]
        assert (p_old < 0 and p_new < 0) or sp.perm[p_new] == p_old
    G = nx.DiGraph([(int(sp.parent[i]), i) for i in range(K) if sp.parent[i] >= 0])
    assert nx.is_arborescence(G)
    # sections: contiguous chains, <= SEC_MAX, internal nodes have one child
    nch = np.bincount(sp.parent[sp.parent >= 0], minlength=K)
    covered = np.zeros(K, dtype=int)
    for s, (h, ln) in enumerate(zip(sp.sec_head, sp.sec_len)):
        assert 1 <= ln <= SEC_MAX
        covered[h:h + ln] += 1
        for k in range(h + 1, h + ln):
            assert sp.parent[k] == k - 1
        for k in range(h, h + ln - 1):
            assert nch[k] == 1
        t = h + ln - 1
        kids = np.flatnonzero(sp.parent == t)
        heads = sp.sec_head[sp.sec_child[s]:sp.sec_child[s] + sp.sec_nchild[s]]
        assert sorted(kids.tolist()) == sorted(heads.tolist())
        assert sp.sec_ptail[s] == sp.parent[h]
    assert (covered == 1).all()
    # level order: a section's parent section lies on the previous level
    lvl = np.repeat(np.arange(sp.n_levels), np.diff(sp.lvl_ptr))
    for s in range(len(sp.sec_head)):
        for c in range(sp.sec_child[s], sp.sec_child[s] + sp.sec_nchild[s]):
            assert lvl[c] == lvl[s] + 1
    # CSR children
    for i in range(K):
        kids = sp.ch_idx[sp.ch_ptr[i]:sp.ch_ptr[i + 1]]
        assert (sp.parent[kids] == i).all() and len(kids) == nch[i]


def test_long_chain_is_split():
    from dendra_amd.topology import SEC_MAX, section_plan

    K = 5 * SEC_MAX + 3
    sp = section_plan([-1] + list(range(K - 1)))
    assert sp.n_levels == 6 and sp.sec_len.tolist() == [SEC_MAX] * 5 + [3]


def test_cable_plan_covers_every_cell_once():
    from dendra_amd.model import plan_cables

    sizes = np.array([1, 2, 3, 21, 52, 53, 64, 65, 104, 208, 209, 416, 417, 801, 832, 7])
    pl = plan_cables(sizes, rmax=13)
    seen = np.zeros(len(sizes), dtype=int)
    for c0, nc, nr, p in zip(pl.task_cell0, pl.task_ncell, pl.task_nr, pl.task_p):
        assert 1 <= nc <= 64 // p and 1 <= nr <= 13
        for k in range(nc):
            cell = pl.order[c0 + k]
            seen[cell] += 1
            assert sizes[cell] <= p * nr and pl.cell_p[c0 + k] == p
    assert (seen == 1).all()
    assert (np.diff(pl.off) == sizes[pl.order]).all()
    with pytest.raises(ValueError):
        plan_cables(np.array([833]), rmax=13)


def test_cable_workspace_matches_dense_operator():
    """gL/gRp/scale/drive against the cable equation built independently."""
    from dendra_amd.model import CableSpec, cable_workspace

    K = 7
    rs = np.random.default_rng(0)
    d = 1 + rs.random(K)
    dx = 5 + 10 * rs.random(K)
    ra = 70 + 60 * rs.random(K)
    cm = 0.9 + 0.2 * rs.random(K)
    lead = rs.standard_normal(K) * 100
    spec = CableSpec(np.array([0, K]), d, dx, ra, cm, lead=lead, ve_amp=np.array([0.7]))
    ws = cable_workspace(spec)
    area = np.pi * (d * 1e-4) * (dx * 1e-4)
    C = 1e-6 * cm * area
    r = d * 1e-4 / 2
    half = ra * (dx * 1e-4 / 2) / (np.pi * r * r)
    g = 1 / (half[:-1] + half[1:])  # series half-compartment resistances
    A = np.zeros((K, K))
    for j in range(K - 1):
        A[j, j] -= g[j]; A[j, j + 1] += g[j]
        A[j + 1, j + 1] -= g[j]; A[j + 1, j] += g[j]
    A = A / C[:, None]
    np.testing.assert_allclose(ws.gL[:-1], np.diag(A, 1), rtol=1e-13)
    np.testing.assert_allclose(ws.gRp[1:], np.diag(A, -1), rtol=1e-13)
    assert ws.gL[-1] == 0 and ws.gRp[0] == 0
    np.testing.assert_allclose(ws.scale, area / C, rtol=1e-13)
    np.testing.assert_allclose(ws.drive, A @ (0.7 * lead), rtol=1e-10, atol=1e-9)


def test_tree_workspace_drive_is_edge_current_divergence():
    from dendra_amd.model import TreeSpec, tree_workspace

    parent = np.array([-1, 0, 0, 1, 1, 2, 5])
    K = len(parent)
    rs = np.random.default_rng(1)
    spec = TreeSpec(np.array([0, K]), parent, 1 + rs.random(K), 10 + rs.random(K),
                    np.full(K, 100.0), np.ones(K), lead=rs.standard_normal(K),
                    ve_amp=np.array([2.0]))
    ws = tree_workspace(spec, 0.01)
    L = np.zeros((K, K))
    for c, p in enumerate(parent):
        if p >= 0:
            g = ws.g[c]
            L[c, c] -= g; L[c, p] += g; L[p, p] -= g; L[p, c] += g
    np.testing.assert_allclose(ws.drive, L @ (2.0 * spec.lead), rtol=1e-12, atol=1e-15)
    assert ws.g[0] == 0
    np.testing.assert_allclose(ws.cm_dt, 1e-6 * np.pi * spec.diam * spec.length * 1e-8 / 1e-5)


def test_counter_rng_is_deterministic_and_uniform():
    from dendra_amd import rng

    a = rng.uniform_n(5, 1, 100000)
    b = rng.uniform(5, 1, np.arange(100000))
    assert np.array_equal(a, b)
    assert not np.array_equal(a, rng.uniform_n(5, 2, 100000))
    assert 0.0 <= a.min() and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.005 and abs(a.var() - 1 / 12) < 0.002
    z = rng.normal_n(3, 4, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    k = rng.integers_n(1, 2, 50000, 21, 801)
    assert k.min() == 21 and k.max() == 801


def test_rect_wave_is_half_open():
    from dendra_amd.model import rect_wave

    w = rect_wave(200, 0.005, delay=0.5, pw=0.1, dtype=np.float32)
    on = np.flatnonzero(w)
    assert on[0] == 100 and len(on) == 20


def test_hh_steady_state_matches_oracle(orc):
    from dendra_amd.model import hh_steady_state

    for v in (-80.0, -65.0, -55.0, -40.0, -40.000001, 0.0, 30.0):
        np.testing.assert_allclose(hh_steady_state(v), orc.hh_steady(v), rtol=1e-14)


def test_slot_map_matches_definition():
    """Row-slot layout of the fitting/threshold kernels on ragged plans."""
    from dendra_amd.model import plan_cables, slot_map

    rs = np.random.default_rng(5)
    sizes = rs.integers(1, 250, 200)
    off = np.zeros(201, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    for R in (4, 8):
        plan = plan_cables(sizes, R, 64)
        got = slot_map(plan, sizes, off, R)
        assert len(np.unique(got)) == len(got) and got.max() < len(plan.task_p) * 64 * R
        for t in range(len(plan.task_p)):
            P, nr = int(plan.task_p[t]), int(plan.task_nr[t])
            for sub in range(int(plan.task_ncell[t])):
                cell = int(plan.order[plan.task_cell0[t] + sub])
                j = np.arange(sizes[cell])
                want = t * 64 * R + (j % nr) * 64 + sub * P + j // nr
                assert np.array_equal(got[off[cell] + j], want)


def test_cell_per_task_slot_map():
    """Cells spanning several wavefronts: one cell per task of `width`
    threads, compartment j on thread j // nr, row j % nr, with
    nr = ceil(K / width); rows beyond a cell stay free."""
    from dendra_amd.model import plan_cells_per_task, slot_map

    rs = np.random.default_rng(6)
    sizes = np.concatenate([[1, 128, 129, 256, 255], rs.integers(1, 257, 40)])
    off = np.zeros(len(sizes) + 1, dtype=np.int64)
    off[1:] = np.cumsum(sizes)
    for R, width in ((2, 128), (1, 256)):
        plan = plan_cells_per_task(sizes, R, width)
        assert np.array_equal(plan.task_nr, -(-sizes // width))
        assert (plan.task_p == width).all() and (plan.task_ncell == 1).all()
        got = slot_map(plan, sizes, off, R, width)
        assert len(np.unique(got)) == len(got) and got.max() < len(sizes) * width * R
        for c in range(len(sizes)):
            nr = -(-int(sizes[c]) // width)
            j = np.arange(sizes[c])
            assert np.array_equal(got[off[c] + j], c * width * R + (j % nr) * width + j // nr)
    with pytest.raises(ValueError):
        plan_cells_per_task(np.array([257]), 1, 256)


def test_entry_points_reject_invalid_arguments_before_any_launch():
    """Argument validation happens on the host before any HIP call, so the
    error contract (DN_EINVAL = -1, DN_ERANGE = -3) is checkable without a GPU."""
    from dendra_amd import _lib

    lib = _lib.load()
    hh = np.array([0.12, 0.036, 3e-4, 50.0, -77.0, -54.3, 1.0])
    hp = hh.ctypes.data
    fake = 1 << 20  # never dereferenced: validation fails first
    # fit forward: (rcap, waves) in {(4, 1), (8, 1), (2, 2), (1, 4)}; null task arrays
    args = [1, 5, 1, fake, fake] + [fake] * 6 + [4, fake, fake, fake, fake, fake, 8, 1.0,
                                                 0.0, 1.0, 0, 10, 0.01, hp, fake, None,
                                                 None, 1, None, None, None]
    assert lib.dn_fit_forward_f32(*args) == -1
    for rcap, waves in ((2, 1), (4, 2), (1, 2), (2, 4), (8, 4)):
        args[1], args[2] = rcap, waves
        assert lib.dn_fit_forward_f32(*args) == -1
    args[1], args[2] = 4, 1
    args[3] = None
    assert lib.dn_fit_forward_f32(*args) == -1
    # tau must be positive
    args[3] = fake
    args[20] = 0.0
    assert lib.dn_fit_forward_f64(*args) == -1
    # adjoint needs a trajectory
    adj = [1, 4, 1, fake, fake] + [fake] * 6 + [4, fake, fake, fake, fake, fake, 8, 1.0,
                                                0.0, 1.0, 0, 10, 0.01, hp, None, fake, fake,
                                                fake, fake, None]
    assert lib.dn_fit_adjoint_f32(*adj) == -1
    adj[1], adj[2] = 1, 4
    assert lib.dn_fit_adjoint_f32(*adj) == -1
    # knot CSR: null pointer table
    assert lib.dn_fit_csr_knots_f32(4, 2, None, fake, fake, fake, 1, 1, fake, 1, 1, None) == -1
    # threshold: rcap 4 only, rtol > 0, fix_up > 1
    st0 = np.zeros(4).ctypes.data
    thr = [1, 4] + [fake] * 10 + [100, 0.005, hp, st0, 0.0, 0.005, 2.0, 10, 25] + \
        [fake] * 4 + [None]
    thr[1] = 8
    assert lib.dn_threshold_f64(*thr) == -1
    thr[1] = 4
    thr[17] = 0.0
    assert lib.dn_threshold_f64(*thr) == -1
    thr[17] = 0.005
    thr[18] = 1.0
    assert lib.dn_threshold_f64(*thr) == -1
    # MLP products: M, N multiples of 128, K of 64 * splits
    g = [0, 1, 100, 128, 128, 1, fake, 128, 0, fake, 128, 0, fake, 128, None, 0, 0, fake,
         128, 0, None, 0, 0, None, 0, 0, None]
    assert lib.dn_mlp_gemm(*g) == -1
    g[2], g[4] = 128, 96
    assert lib.dn_mlp_gemm(*g) == -1
    g[4], g[5] = 128, 3
    assert lib.dn_mlp_gemm(2, *g[1:5], 3, *g[6:]) == -1
    # epilogue 0 needs a bias
    g[5] = 1
    g[12] = None
    assert lib.dn_mlp_gemm(*g) == -1
    assert lib.dn_mlp_out_bwd(1, 100, 512, fake, 100, fake, 512, fake, 1, fake, fake, 1,
                              None) == -1
    assert lib.dn_mlp_gather(1, 64, 129, fake, fake, fake, fake, fake, fake, None) == -1
    assert lib.dn_mlp_adam(10, 3, 3, fake, fake, fake, fake, 1e-3, 0.9, 0.999, 1e-8, None,
                           None, None, None, None) == -1
    # packed copies need rows % 32 == 0 and cols % 16 == 0 (W^T: the reverse)
    assert lib.dn_mlp_adam(48 * 16, 48, 16, fake, fake, fake, fake, 1e-3, 0.9, 0.999, 1e-8,
                           None, None, fake, None, None) == -1
    assert lib.dn_mlp_adam(64 * 48, 64, 48, fake, fake, fake, fake, 1e-3, 0.9, 0.999, 1e-8,
                           None, None, None, fake, None) == -1
    # fused forward / backward: tile 64 or 128; training needs tile 64 and
    # every training output; the backward takes tile 64 only
    fw = [1, 256, 96, fake, 0, fake, fake, fake, 1, 1, fake, fake, fake, fake, fake, fake, 256]
    assert lib.dn_mlp_forward(*fw, *[None] * 9, None) == -1
    fw[2] = 128
    assert lib.dn_mlp_forward(*fw, *[fake] * 9, None) == -1
    fw[2] = 64
    assert lib.dn_mlp_forward(*fw, *[fake] * 8, None, None) == -1
    assert lib.dn_mlp_backward(1, 256, 128, fake, 256, fake, fake, fake, fake, fake, fake, 1,
                               fake, fake, fake, fake, None) == -1
    assert lib.dn_mlp_backward(1, 96, 64, fake, 96, fake, fake, fake, fake, fake, fake, 1,
                               fake, fake, fake, fake, None) == -1
    assert lib.dn_mlp_partials_reduce(1, 0, fake, fake, fake, fake, fake, fake, None) == -1
    n = np.array([10], np.int64)
    rc = np.array([2, 5], np.int32)
    a = np.array([fake], np.uint64)
    z = np.array([0], np.uint64)
    assert lib.dn_mlp_adam_multi(9, n.ctypes.data, rc.ctypes.data, rc.ctypes.data,
                                 *[a.ctypes.data] * 4, *[z.ctypes.data] * 4, 1e-3, 0.9, 0.999,
                                 1e-8, None) == -1
    # rows * cols must divide n
    n7 = np.array([7], np.int64)
    assert lib.dn_mlp_adam_multi(1, n7.ctypes.data, rc.ctypes.data, rc[1:].ctypes.data,
                                 *[a.ctypes.data] * 4, *[z.ctypes.data] * 4, 1e-3, 0.9, 0.999,
                                 1e-8, None) == -1
    assert lib.dn_mlp_ensemble(10, 9, fake, 10, fake, 0.0, 1.0, 0.03, fake, fake,
                               None) == -1


def test_depth_levels_are_whole_levels_deepest_first():
    """The block tree solve's host layering: every node once, whole depth
    levels from the deepest, ascending node index within a level, each
    parent in the next layer; a cycle gives None.  Depths are checked
    against a breadth-first search from the roots (networkx)."""
    import torch

    from dendra_amd.blocks import _depth_levels

    rng = np.random.default_rng(4)
    cases = [np.arange(-1, 299), np.array([-1]), np.array([-1, 0, 0, -1, 3, 4, 4])]
    for K in (50, 400):
        p = np.full(K, -1)
        for i in range(1, K):
            p[i] = rng.integers(max(0, i - 8), i)
        perm = rng.permutation(K)
        inv = np.argsort(perm)
        cases.append(np.where(p[perm] >= 0, inv[np.maximum(p[perm], 0)], -1))
    for par in cases:
        K = len(par)
        order, lp = (t.numpy() for t in _depth_levels(torch.as_tensor(par, dtype=torch.int64)))
        assert sorted(order.tolist()) == list(range(K)) and lp[0] == 0 and lp[-1] == K
        g = nx.DiGraph()
        g.add_nodes_from(range(K))
        g.add_edges_from((int(q), i) for i, q in enumerate(par) if q >= 0)
        depth = np.zeros(K, int)
        for r in np.flatnonzero(par < 0):
            for node, dist in nx.single_source_shortest_path_length(g, int(r)).items():
                depth[node] = dist
        n_lvl = len(lp) - 1
        for l in range(n_lvl):
            seg = order[lp[l]:lp[l + 1]]
            assert (depth[seg] == n_lvl - 1 - l).all() and (np.diff(seg) > 0).all()
            for i in seg:
                if par[i] >= 0:
                    assert lp[l + 1] <= np.flatnonzero(order == par[i])[0] < lp[l + 2]
    assert _depth_levels(torch.tensor([1, 2, 0])) is None
    assert _depth_levels(torch.tensor([-1, 2, 1])) is None


def test_cable_launch_list_is_checked_before_launch():
    """A launch whose tasks need more rows per lane than its bound (or than
    the largest kernel instance) is rejected on the host, before any work
    is enqueued."""
    from dendra_amd.model import plan_cables
    from dendra_amd.population import validate_cable_launches

    plan = plan_cables(np.array([21, 200, 801, 832, 5]), 13, 64)
    nr, tp = plan.task_nr, plan.task_p
    validate_cable_launches(nr, tp, [(0, len(nr), int(nr.max()))], 13)
    with pytest.raises(ValueError, match="rows per lane"):
        validate_cable_launches(nr, tp, [(0, len(nr), int(nr.max()) - 1)], 13)
    with pytest.raises(ValueError, match="outside"):
        validate_cable_launches(nr, tp, [(0, len(nr), 14)], 13)
    with pytest.raises(ValueError, match="lanes"):
        validate_cable_launches(nr, np.full_like(tp, 2), [(0, len(nr), 13)], 13)


def test_tree_groups_are_checked_before_launch():
    from dendra_amd.population import validate_tree_groups

    n_sec = np.array([300, 200, 13])
    sizes = np.array([2000, 1500, 100])
    validate_tree_groups([[0, 1], [2]], n_sec, sizes, 512, 3500)
    with pytest.raises(ValueError, match="threads"):
        validate_tree_groups([[0, 1, 2]], n_sec, sizes, 512, 4000)
    with pytest.raises(ValueError, match="node capacity"):
        validate_tree_groups([[0, 1]], n_sec, sizes, 512, 3000)


def test_multi_dhs_plan_covers_each_row_once():
    from dendra_amd.topology import multi_dhs_plan

    morphs = [np.array([-1, 0, 0, 1, 1, 2]), np.arange(-1, 9), np.array([-1])]
    for threads in (1, 4, 32):
        pl = multi_dhs_plan(morphs, (70, 3, 1), threads)
        per = max(1, 32 // threads)
        assert pl["rows_total"] == 74 and pl["K_stride"] == 10 and pl["grid_x"] == len(
            pl["WARP_ROW_BASE"])
        covered = np.concatenate([np.arange(r, r + c) for r, c in
                                  zip(pl["WARP_ROW_BASE"], pl["WARP_ROW_COUNT"])])
        assert np.array_equal(np.sort(covered), np.arange(74))
        assert pl["WARP_ROW_COUNT"].max() <= per
        for w in range(pl["grid_x"]):
            lp = pl["LAYER_PTR_cat"][pl["WARP_LPTR_OFF"][w]:][: pl["WARP_L"][w] + 1]
            K = lp[-1]
            order = pl["ORDER_cat"][pl["WARP_ORDER_OFF"][w]:][:K]
            par = pl["P_cat"][pl["WARP_P_OFF"][w]:][:K]
            assert sorted(order.tolist()) == list(range(K))
            pos = np.empty(K, int)
            pos[order] = np.arange(K)
            assert all(pos[i] < pos[p] for i, p in enumerate(par) if p >= 0)  # children first
            assert (np.diff(lp) <= threads).all()
