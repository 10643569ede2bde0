"""The block-tree solve and the ExtCell tree run at the batch sizes where
their kernels group several systems per workgroup (B >= 256; B >= 1024 in the
workspace form, with wavefronts past the batch when B is not a multiple of
the group), on a tree with a depth level of more than 64 nodes (the workspace
form, membrane term inside the solve) and one whose levels fit the LDS ring:
every sampled system equals the same system solved alone, and the dense
float64 solve or the binary64 oracle."""

import dataclasses

import numpy as np
import pytest
import torch

from test_extcell_ref import extcell_tree_ref, init_state, np_ws, tree_spec
from test_gpu_blocks import block_tree_matrix, dense_solve

pytestmark = pytest.mark.gpu

COMPLETE = [-1] + [(i - 1) // 2 for i in range(1, 255)]   # level 7 holds 128 nodes


def _layered(parents, threads=64):
    from dendra_amd.topology import build_dhs_layers, build_morphology

    _, _, depth = build_morphology(parents)
    order, layers = build_dhs_layers(depth, threads)
    return [torch.as_tensor(np.asarray(v), dtype=torch.int64, device="cuda")
            for v in (parents, order, layers)], depth


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("B", [257, 2049])
@pytest.mark.parametrize("tree", ["deep-level", "ring"])
def test_block_tree_batch_equals_single_solves(dtype, B, tree):
    from dendra_amd.blocks import dhs_bt_solve
    from dendra_amd.workloads import random_binary_tree

    parents = COMPLETE if tree == "deep-level" else [int(p) for p in
                                                     random_binary_tree(300, 0, 4)[0]]
    topo, depth = _layered(parents)
    assert (np.bincount(depth).max() > 64) == (tree == "deep-level")
    K = len(parents)
    rng = np.random.default_rng(B + K)
    main = torch.as_tensor(rng.uniform(-0.3, 0.3, (B, K, 3, 3)) + np.eye(3) * 2.0)
    edge = torch.as_tensor(rng.uniform(0.1, 2.0, (B, K, 3)))
    rhs = torch.as_tensor(rng.standard_normal((B, K, 3)))
    args = [v.to("cuda", dtype) for v in (main, edge, rhs)]
    x = dhs_bt_solve(*args, *topo)
    sample = sorted({0, 1, 3, 4, 255, 256, B // 2, B - 2, B - 1})
    for b in sample:
        alone = dhs_bt_solve(*[a[b:b + 1].contiguous() for a in args], *topo)
        assert torch.allclose(x[b:b + 1], alone, rtol=0,
                              atol=64 * torch.finfo(dtype).eps * float(alone.abs().max())), b
    want = dense_solve(block_tree_matrix(main[sample], edge[sample], parents), rhs[sample])
    tol = 1e-11 if dtype == torch.float64 else 1e-4
    assert (x[sample].double().cpu() - want).abs().max() <= tol * max(1.0, want.abs().max())


def _one(spec, b):
    pick = lambda a: None if a is None else a[b:b + 1]
    return dataclasses.replace(spec, diam=pick(spec.diam), length=pick(spec.length),
                               ra=pick(spec.ra), cm=pick(spec.cm), xraxial=pick(spec.xraxial),
                               xc=pick(spec.xc), xg=pick(spec.xg), sna=pick(spec.sna),
                               sk=pick(spec.sk), lead=pick(spec.lead),
                               ve_amp=pick(spec.ve_amp), iamp=pick(spec.iamp))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("B", [257, 2049])
def test_extcell_tree_batch_equals_single_runs_and_oracle(dtype, B):
    from dendra_amd.extcell import ExtCellTrees
    from dendra_amd.model import HH

    rng = np.random.default_rng(B)
    ts, cs = tree_spec(rng, B, COMPLETE)
    steps, dt = 160, 0.005
    wave = np.zeros(steps)
    wave[20:80] = 1.0
    iw = np.zeros(steps)
    iw[90:130] = 1.0
    hh = HH(celsius=6.3)
    trees = ExtCellTrees(ts, dt, hh, dtype, "cuda")
    rec = trees.run(steps, wave=wave, i_wave=iw, record_every=8).double().cpu().numpy()
    scale = np.abs(rec).max()
    tol = 1e-10 if dtype == torch.float64 else 2e-5
    for b in sorted({0, 1, 4, 255, 256, B // 2, B - 1}):
        one = ExtCellTrees(_one(ts, b), dt, hh, dtype, "cuda")
        r1 = one.run(steps, wave=wave, i_wave=iw, record_every=8).double().cpu().numpy()
        np.testing.assert_allclose(rec.reshape(rec.shape[0], B, -1)[:, b], r1.reshape(
            r1.shape[0], -1), rtol=0, atol=tol * scale, err_msg=f"system {b}")
    if dtype == torch.float64:
        b = B - 1
        one = ExtCellTrees(_one(ts, b), dt, hh, dtype, "cuda")
        st, orec = extcell_tree_ref.run(np_ws(one.workspace), COMPLETE, hh.array(), dt, steps,
                                        init_state(_one(ts, b)), lead=ts.lead[b:b + 1],
                                        amp=ts.ve_amp[b:b + 1], wave=wave,
                                        iamp=ts.iamp[b:b + 1], i_wave=iw, record_every=8)
        got = rec.reshape(rec.shape[0], B, -1)[:, b]
        assert np.abs(got - orec.reshape(orec.shape[0], -1)).max() <= 1e-8


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_extcell_tree_contraction_equals_the_layer_solve(dtype):
    """The block tree solve by contraction (sections, rake/compress rounds)
    and over whole depth levels give the same ExtCell tree run on a
    1000-node random binary tree (102 sections of up to 17 nodes, 168 depth
    levels), to rounding."""
    from dendra_amd.extcell import ExtCellTrees
    from dendra_amd.model import HH
    from dendra_amd.workloads import random_binary_tree

    parents = [int(p) for p in random_binary_tree(1000, 1, 7)[0]]
    rng = np.random.default_rng(77)
    ts, _ = tree_spec(rng, 6, parents)
    steps, dt = 120, 0.005
    wave = np.zeros(steps)
    wave[10:60] = 1.0
    hh = HH(celsius=6.3)
    recs = []
    for solver in ("contraction", "layers"):
        trees = ExtCellTrees(ts, dt, hh, dtype, "cuda", solver=solver)
        assert trees.solver == solver
        recs.append(trees.run(steps, wave=wave, record_every=10).double().cpu().numpy())
    scale = np.abs(recs[1]).max()
    tol = 1e-10 if dtype == torch.float64 else 2e-5
    assert np.abs(recs[0] - recs[1]).max() <= tol * scale

