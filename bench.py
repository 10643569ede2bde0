"""Benchmark: batched forward HH cable/tree population on one MI355X.

Workload (one "step" = one full forward simulation of the population):
8192 unbranched HH cables with compartment counts uniform in [21, 801] plus
1024 random binary trees of 200-4000 compartments, float32 state, 4000 time
steps of 0.005 ms with an extracellular point-source pulse (see
dendra_amd/workloads.py).  Inputs are resident in HBM before the timed region.

Printed JSON line: metric compartment-steps/s of the persistent path (cable
and tree kernels on two streams), with
  * roofline: the binding bound, VALU issue: the two concurrent persistent
    kernels' VALU instructions (per compartment-step from rocprofv3
    SQ_INSTS_VALU) over the run time, against 6.144e11 wave64 instructions/s;
  * hbm_roofline: the cable kernel's per-step algorithmic bytes (48 B per
    compartment-step: read+write V,m,h,n and read four coefficient streams,
    float32) over its launch time measured with HIP events on its stream --
    a secondary figure, as the kernel keeps its state on chip;
  * plain: the per-step GPU path (2 launches per cable step, 2 per tree step)
    timed on the same workload;
  * cpu_baseline: the binary64 C oracle (OpenMP) on a bounded sample of the
    same workload.
"""

from __future__ import annotations

import argparse
import json
from dataclasses import asdict
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CS = 48.0       # float32 cable compartment-step
BYTES_PER_CS_TREE = 52.0  # float32 tree compartment-step (+ int32 parent index)

# VALU issue bound of the persistent kernels (state stays on chip, so they
# are not HBM-bound).  Peak: 256 CUs x 4 SIMDs, each issuing one wave64
# VALU instruction per 4 cycles (16 lanes per cycle) at 2.4 GHz = 6.144e11
# wave-instructions/s (the packed-FP32 forms, which double the published
# 157.3 TFLOP/s figure's 78.6, are disabled in this build).  Instructions per
# compartment-step: SQ_INSTS_VALU of each kernel divided by its
# compartment-steps, profiles/hh_population_uneven_valu.md.
VALU_PEAK_PER_S = 256 * 4 * 2.4e9 / 4
VALU_PER_CS = {"cable": 2.099, "tree": 2.304}
# The trace-fitting kernels (four wavefronts per cell, one row per lane):
# SQ_INSTS_VALU per compartment-step of the forward (also the recomputation)
# and of the adjoint, profiles/fit_adjoint_512x201_kernels.md.
FIT_VALU_PER_CS = {"forward": 6.74, "adjoint": 10.22}


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cables", type=int, default=8192)
    ap.add_argument("--trees", type=int, default=1024)
    ap.add_argument("--time-steps", type=int, default=4000)
    ap.add_argument("--plain", type=int, default=1, help="plain-path runs to time (0 = skip)")
    ap.add_argument("--cpu-sample", type=float, default=0.125,
                    help="fraction of cells in the CPU oracle sample (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=800)
    ap.add_argument("--waves-per-block", type=int, default=4)
    ap.add_argument("--tune", type=int, default=1,
                    help="autotune the launch parameters before the timed region (0 = defaults)")
    ap.add_argument("--only", choices=["all", "cables", "trees", "none"], default="all",
                    help="parts of the uneven population to run (none: only the other lines)")
    ap.add_argument("--network-steps", type=int, default=40000,
                    help="time steps of the network workload line (0 = skip)")
    ap.add_argument("--surrogate-train", type=int, default=2_000_000,
                    help="fibers in the surrogate training set (0 = skip the surrogate)")
    ap.add_argument("--surrogate-test", type=int, default=1_000_000)
    ap.add_argument("--surrogate-epochs", type=int, default=40)
    ap.add_argument("--fit", type=int, default=1,
                    help="also time the adjoint workload (0 = skip)")
    ap.add_argument("--solvers", type=int, default=1,
                    help="also time the drop-in solver and delay operators (0 = skip)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="MiB copied once by a plain device copy before the timed region "
                         "(reference kernel for scaling FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def main():
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(cells are independent; a multi-GPU run would partition them over ranks)",
              file=sys.stderr)
        return 2
    import torch

    from dendra_amd import _lib
    from dendra_amd.model import HH

    _lib.load()
    dev = torch.device("cuda:0")
    hh = HH(celsius=6.3)
    if args.only == "none":
        out = {"metric": "compartment-steps/s (batched forward, uneven HH population)",
               "value": None, "unit": "compartment-steps/s", "n_gpus": 1,
               "note": "uneven population skipped (--only none)"}
    else:
        out = population_bench(args, dev, hh)
    if args.network_steps > 0:
        out["network"] = network_bench(args.network_steps, hh, dev, args.cpu_sample > 0)
    if args.fit:
        out["adjoint"] = fit_bench(dev, args.cpu_sample > 0)
    if args.solvers:
        out["operators"] = operator_bench(dev)
    if args.surrogate_train > 0:
        out["surrogate"] = surrogate_bench(args.surrogate_train, args.surrogate_test,
                                           args.surrogate_epochs, args.cpu_sample > 0)
    print(json.dumps(out))
    return 0


def population_bench(args, dev, hh):
    """Primary line: the uneven population (cables and trees on two streams)."""
    import torch

    from dendra_amd.population import CablePopulation, TreePopulation
    from dendra_amd.workloads import uneven_population

    t0 = time.time()
    work = uneven_population(args.cables, args.trees, args.time_steps)
    print(f"# workload built in {time.time() - t0:.1f}s: {work.cables.total} cable + "
          f"{work.trees.total} tree compartments, {work.n_steps} steps", file=sys.stderr)
    use_c = args.only in ("all", "cables")
    use_t = args.only in ("all", "trees")
    wave = work.wave
    tuned = {}
    t_tune = time.time()
    if args.tune:
        from dendra_amd.autotune import tune_cables, tune_trees

        cab = tre = None
        if use_c:
            ct, cab = tune_cables(work.cables, work.dt, hh, torch.float32, dev, wave=wave)
            args.waves_per_block = ct.waves_per_block
            tuned["cables"] = asdict(ct)
        if use_t:
            tre = TreePopulation(work.trees, work.dt, hh, torch.float32, dev)
            tuned["trees"] = asdict(tune_trees(tre, wave=wave))
        tuned["tune_s"] = time.time() - t_tune
    else:
        cab = CablePopulation(work.cables, work.dt, hh, torch.float32, dev) if use_c else None
        tre = TreePopulation(work.trees, work.dt, hh, torch.float32, dev) if use_t else None
    main_stream = torch.cuda.current_stream(dev)
    side = torch.cuda.Stream(dev)
    ev = {k: [] for k in ("c0", "c1", "t0", "t1")}

    def one_run(record=False):
        if cab is not None:
            cab.reset()
        if tre is not None:
            tre.reset()
        side.wait_stream(main_stream)
        if tre is not None:
            with torch.cuda.stream(side):
                if record:
                    e = torch.cuda.Event(enable_timing=True); e.record(side); ev["t0"].append(e)
                tre.run(work.n_steps, wave=wave, stream=side)
                if record:
                    e = torch.cuda.Event(enable_timing=True); e.record(side); ev["t1"].append(e)
        if cab is not None:
            if record:
                e = torch.cuda.Event(enable_timing=True); e.record(main_stream); ev["c0"].append(e)
            cab.run(work.n_steps, wave=wave, waves_per_block=args.waves_per_block)
            if record:
                e = torch.cuda.Event(enable_timing=True); e.record(main_stream); ev["c1"].append(e)
        main_stream.wait_stream(side)

    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        one_run()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        one_run(record=True)
        print(f"# run {i + 1}/{args.steps} enqueued", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t_start) / max(args.steps, 1)
    comps = (work.cables.total if use_c else 0) + (work.trees.total if use_t else 0)
    cs_per_run = comps * work.n_steps
    value = cs_per_run / wall
    c_ms = [a.elapsed_time(b) for a, b in zip(ev["c0"], ev["c1"])]
    t_ms = [a.elapsed_time(b) for a, b in zip(ev["t0"], ev["t1"])]
    out = {
        "metric": "compartment-steps/s (batched forward, uneven HH population)",
        "value": value,
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded counter RNG)",
        "config": {
            "workload": "uneven HH population: 8192 cables K~U[21,801] + 1024 binary "
                        "trees K~U[200,4000], 4000 steps dt 0.005 ms, point-source Ve",
            "cables": args.cables if use_c else 0, "trees": args.trees if use_t else 0,
            "compartments": comps, "time_steps": work.n_steps,
            "compartment_steps_per_run": cs_per_run, "path": "persistent",
            "tuned": tuned,
        },
    }
    job_bytes = (BYTES_PER_CS * (work.cables.total if use_c else 0)
                 + BYTES_PER_CS_TREE * (work.trees.total if use_t else 0)) * work.n_steps
    out["job_roofline"] = {"bound": "hbm", "achieved": job_bytes / wall / 1e9,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": job_bytes / wall / 1e9 / HBM_PEAK_GBS,
                           "algorithmic_bytes_per_run": job_bytes,
                           "note": "algorithmic bytes of the per-step path (state and "
                                   "coefficients read and written every step); the persistent "
                                   "kernels keep the state on chip, so frac can exceed 1"}
    # VALU bound of the job: both persistent kernels share the GPU for the
    # whole run, so their instructions are counted together over the run.
    valu = (VALU_PER_CS["cable"] * (work.cables.total if use_c else 0)
            + VALU_PER_CS["tree"] * (work.trees.total if use_t else 0)) * work.n_steps
    out["roofline"] = {
        "bound": "valu", "achieved": valu / wall, "peak": VALU_PEAK_PER_S,
        "unit": "wave64 VALU instructions/s", "frac": valu / wall / VALU_PEAK_PER_S,
        "traffic": _job_traffic(use_c, use_t),
        "kernel": "hh_cable_run_kernel<float> + hh_tree_sec_kernel<float> (concurrent)",
        "valu_instructions_per_run": valu,
        "valu_per_compartment_step": VALU_PER_CS,
        "note": "instructions per compartment-step from rocprofv3 SQ_INSTS_VALU "
                "(profiles/hh_population_uneven_valu.md); time = the run",
    }
    if c_ms:
        c_avg = float(np.mean(c_ms)) * 1e-3
        c_bytes = BYTES_PER_CS * work.cables.total * work.n_steps
        ach = c_bytes / c_avg / 1e9
        out["hbm_roofline"] = {
            "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": _measured_traffic("hh_cable_run_kernel"),
            "kernel": "hh_cable_run_kernel<float>",
            "kernel_ms": c_avg * 1e3,
            "algorithmic_bytes_per_launch": c_bytes,
            "note": "per-step algorithmic bytes (48 B per compartment-step) over the cable "
                    "kernel's launch time; a secondary figure, the kernel keeps its state "
                    "on chip and moves `traffic` bytes",
        }
        cv = VALU_PER_CS["cable"] * work.cables.total * work.n_steps / c_avg
        out["cable_valu"] = {"achieved": cv, "frac_of_gpu": cv / VALU_PEAK_PER_S,
                             "kernel_ms": c_avg * 1e3}
        out["cable_cs_per_s"] = work.cables.total * work.n_steps / c_avg
    if t_ms:
        t_avg = float(np.mean(t_ms)) * 1e-3
        out["tree_kernel_ms"] = t_avg * 1e3
        out["tree_cs_per_s"] = work.trees.total * work.n_steps / t_avg
        t_bytes = BYTES_PER_CS_TREE * work.trees.total * work.n_steps
        out["tree_roofline"] = {"achieved": t_bytes / t_avg / 1e9, "unit": "GB/s",
                                "traffic": _measured_traffic("hh_tree_sec_kernel"),
                                "frac": t_bytes / t_avg / 1e9 / HBM_PEAK_GBS,
                                "kernel": "hh_tree_sec_kernel<float>",
                                "note": "per-step algorithmic bytes, secondary"}
        tv = VALU_PER_CS["tree"] * work.trees.total * work.n_steps / t_avg
        out["tree_valu"] = {"achieved": tv, "frac_of_gpu": tv / VALU_PEAK_PER_S,
                            "kernel_ms": t_avg * 1e3}
        if c_ms and t_avg > c_avg:
            out["roofline_note"] = "tree kernel is the longer of the two concurrent kernels"
    if args.plain > 0:
        torch.cuda.synchronize()
        n_pl = work.n_steps
        if cab is not None:
            cab.reset(); cab.run(20, wave=wave, path="plain"); cab.reset()
        if tre is not None:
            tre.reset(); tre.run(20, wave=wave, path="plain"); tre.reset()
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for _ in range(args.plain):
            if cab is not None:
                cab.reset(); cab.run(n_pl, wave=wave, path="plain")
            if tre is not None:
                tre.reset(); tre.run(n_pl, wave=wave, path="plain")
        torch.cuda.synchronize()
        plain_wall = (time.perf_counter() - tp) / args.plain
        out["plain"] = {"value": cs_per_run / plain_wall, "ms_per_run": plain_wall * 1e3,
                        "speedup_persistent_over_plain": plain_wall / wall}
    if args.cpu_sample > 0:
        out["cpu_baseline"] = cpu_baseline(work, hh, args.cpu_sample, args.cpu_steps)
    return out


def network_bench(n_steps, hh, dev, with_cpu):
    """Secondary line: the event-driven network (20,000 point cells + 1,000
    ball-and-stick cells, 200 sources per cell, exp2syn, Poisson background),
    dt 0.025 ms, steps replayed as HIP graphs of 100 steps."""
    import torch

    from dendra_amd.network import Network
    from dendra_amd.workloads import sparse_network

    spec = sparse_network()
    net = Network(spec, 0.025, hh, dev)
    net.run(200, graph_steps=100)
    torch.cuda.synchronize()
    net.reset()
    t = time.perf_counter()
    events, _ = net.run(n_steps, record_events=True, graph_steps=100,
                        ev_cap=n_steps * 4096)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    out_deg = np.bincount(spec.conn.pre, minlength=spec.conn.n_src)
    deliveries = int(out_deg[events[:, 1]].sum())
    res = {
        "workload": "20000 point + 1000 ball-and-stick HH cells, 4.2M connections, "
                    "exp2syn, 5 Hz Poisson background, dt 0.025 ms",
        "time_steps": n_steps, "wall_s": wall, "steps_per_s": n_steps / wall,
        "spikes": int(len(events)), "deliveries": deliveries,
        "deliveries_per_s": deliveries / wall,
        "compartment_steps_per_s": spec.cells.total * n_steps / wall,
    }
    # per-step algorithmic bytes: every compartment's state and coefficients
    # (48 B), today's ring row read and cleared (8 B per cell), 16 B per
    # delivered event (target, weight, ring write)
    net_bytes = (48.0 * spec.cells.total + 8.0 * spec.cells.n_cells) * n_steps + \
        16.0 * deliveries
    res["roofline"] = {"bound": "hbm", "achieved": net_bytes / wall / 1e9,
                       "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": net_bytes / wall / 1e9 / HBM_PEAK_GBS,
                       "us_per_step": wall / n_steps * 1e6,
                       "note": "three dependent launches per step (HIP graphs of 100 "
                               "steps): launch latency, not bandwidth, sets the step time"}
    if with_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        cpu_steps = 100
        t = time.perf_counter()
        _, sp = oracle.network_run(spec, hh.array(), 0.025, cpu_steps)
        dt_cpu = time.perf_counter() - t
        res["cpu_baseline"] = {"steps_per_s": cpu_steps / dt_cpu, "cores": 1,
                               "kind": "port",
                               "sample": f"first {cpu_steps} steps of the same network, "
                                         f"float64 C oracle, 1 thread, {dt_cpu:.1f}s"}
    return res


def fit_bench(dev, with_cpu, reps=3):
    """Secondary line: loss and hand-written adjoint of 512 HH cables x 201
    compartments, 10,000 steps of 0.01 ms in float32, gradients w.r.t. the
    per-compartment gNa/gK scales and the 201 x 64 Ve knot table, state
    checkpoints every 500 steps.  Each cell runs on a workgroup of four
    wavefronts, one compartment row per lane (2048 wavefronts).  Roofline:
    VALU issue (the kernels keep their state on chip), from the VALU
    instructions per compartment-step of the forward, the recomputation and
    the adjoint; secondary: algorithmic bytes, 56 per compartment-step
    forward, 144 per compartment-step forward + adjoint."""
    import torch

    from dendra_amd.fit import CableFit
    from dendra_amd.model import HH
    from dendra_amd.workloads import fit_cables

    w = fit_cables()
    fit = CableFit(w.spec, w.dt, w.n_steps, w.sites, w.targets, knot_times=w.knot_times,
                   margin_site=w.margin_site, margin_sign=w.margin_sign,
                   margin_weight=w.margin_weight, theta=w.theta, tau=w.tau,
                   hh=HH(celsius=6.3), dtype=torch.float32, device=dev, rmax=1, waves=4)
    p = [torch.tensor(np.asarray(x).reshape(-1), dtype=torch.float32, device=dev)
         for x in (w.sna, w.sk)] + [torch.tensor(w.ve, dtype=torch.float32, device=dev)]
    fit.forward(*p)
    fit.backward()
    torch.cuda.synchronize()
    tf, tb = [], []
    for _ in range(reps):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record()
        loss = fit.forward(*p)
        e[1].record()
        fit.backward()
        e[2].record()
        torch.cuda.synchronize()
        tf.append(e[0].elapsed_time(e[1]))
        tb.append(e[1].elapsed_time(e[2]))
    tf, tb = float(np.median(tf)), float(np.median(tb))
    cs = w.spec.total * w.n_steps
    both = tf + tb
    res = {
        "workload": "512 HH cables x 201 compartments, 10000 steps of 0.01 ms, per-"
                    "compartment gNa/gK scales, 201x64 Ve knot table, MSE at 8 sites + "
                    "sigmoid margin, checkpoints every 500 steps",
        "dtype": "f32", "loss": loss.item(),
        "layout": "one cell per workgroup of 4 wavefronts, one row per lane (2048 wavefronts)",
        "forward": {"ms": tf, "compartment_steps_per_s": cs / (tf * 1e-3)},
        "forward_adjoint": {"ms": both, "compartment_steps_per_s": cs / (both * 1e-3)},
    }
    # forward + recomputation + adjoint VALU instructions over the evaluation
    valu = (2 * FIT_VALU_PER_CS["forward"] + FIT_VALU_PER_CS["adjoint"]) * cs
    res["roofline"] = {
        "bound": "valu", "achieved": valu / (both * 1e-3), "peak": VALU_PEAK_PER_S,
        "unit": "wave64 VALU instructions/s", "frac": valu / (both * 1e-3) / VALU_PEAK_PER_S,
        "valu_per_compartment_step": FIT_VALU_PER_CS,
        "forward_frac": FIT_VALU_PER_CS["forward"] * cs / (tf * 1e-3) / VALU_PEAK_PER_S,
        "note": "instructions per compartment-step from rocprofv3 SQ_INSTS_VALU "
                "(profiles/fit_adjoint_512x201_kernels.md); forward + adjoint = forward, "
                "recomputation and adjoint sweeps"}
    res["hbm_roofline"] = {
        "bound": "hbm", "bytes_per_cs": 144, "achieved": 144 * cs / (both * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": 144 * cs / (both * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "note": "per-step algorithmic bytes; a secondary figure, the state stays on chip"}
    if with_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import fit_ref

        sw = fit_cables(n_cells=64, n_steps=100, n_knots=8)
        geom = fit_ref.geometry(sw.spec.diam, sw.spec.dx, sw.spec.rhoa, sw.spec.cm,
                                sw.spec.off)
        kk, ka = fit_ref.knot_table(sw.knot_times, sw.dt, sw.n_steps)
        B, K = geom["gL"].shape
        mw = torch.zeros((B, K), dtype=torch.float64)
        mw[:, sw.sites[-1]] = torch.as_tensor(sw.margin_weight * sw.margin_sign
                                              / (B * sw.n_steps))
        a = torch.as_tensor(sw.sna).clone().requires_grad_()
        b = torch.as_tensor(sw.sk).clone().requires_grad_()
        v = torch.as_tensor(sw.ve).clone().requires_grad_()
        t = time.perf_counter()
        lc = fit_ref.fit_loss(geom, a, b, v, kk, ka, HH(celsius=6.3).array(), sw.dt,
                              sw.n_steps, sw.sites, sw.targets.astype(np.float64),
                              1.0 / (B * 8 * sw.n_steps), mw, sw.theta, sw.tau)
        torch.autograd.grad(lc, [a, b, v])
        dt_cpu = time.perf_counter() - t
        scs = B * K * sw.n_steps
        res["cpu_baseline"] = {
            "value": scs / dt_cpu, "unit": "compartment-steps/s (forward + adjoint)",
            "cores": torch.get_num_threads(), "kind": "port",
            "sample": f"64 cells x 201 compartments x 100 steps, float64 eager PyTorch "
                      f"autograd through a Thomas sweep, {dt_cpu:.1f}s"}
    return res


def _mlp_flops_per_sample(train):
    """GEMM flops per sample per member: forward 2*(128*512 + 2*512*512 + 512);
    training adds the two data products and the three weight products."""
    fwd = 2 * (128 * 512 + 2 * 512 * 512 + 512)
    if not train:
        return fwd
    return fwd + 2 * (2 * 512 * 512 + 512) + 2 * (128 * 512 + 2 * 512 * 512 + 512)


def surrogate_bench(n_train, n_test, epochs, with_cpu):
    """Surrogate line: GPU bisection labels for n_train seeded fibers (101
    compartments, d ~ U[2, 16] um, point electrode), a 4-member BF16 MLP
    ensemble trained on them, then thresholds of n_test held-out fibers by the
    ensemble with bisection fallback (spread > 3 % or out of range).  Error is
    measured against bisection on the first 50,000 held-out fibers."""
    import torch

    from dendra_amd.surrogate import ThresholdSurrogate
    from dendra_amd.thresholder import GPUThresholder, bisection_judge
    from dendra_amd.workloads import threshold_fibers

    tr = threshold_fibers(n_train, seed=9)
    gt = GPUThresholder(tr.spec, tr.profile)
    torch.cuda.synchronize()
    t = time.perf_counter()
    ytr, _ = gt.calculate_thresholds()
    torch.cuda.synchronize()
    t_label = time.perf_counter() - t
    trials = float(gt.trials.mean())
    del gt
    sur = ThresholdSurrogate(n_members=4, seed=0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    steps = sur.fit(tr.profile, tr.diam.astype(np.float32), ytr, epochs=epochs, batch=8192)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t
    del tr
    te = threshold_fibers(n_test, seed=9, start=10_000_000)
    judge = bisection_judge(te.spec, te.profile)
    d = te.diam.astype(np.float32)
    sur.thresholds(te.profile[:8192], d[:8192], judge)
    torch.cuda.synchronize()
    t = time.perf_counter()
    thr, fb = sur.thresholds(te.profile, d, judge)
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t
    # the network alone (features, 4 members, ensemble) on the same inputs
    torch.cuda.synchronize()
    t = time.perf_counter()
    sur.predict(te.profile, d)
    torch.cuda.synchronize()
    t_net = time.perf_counter() - t
    # kernel-level rates: the fused forward alone on a 65,536-row chunk (one
    # launch), and training steps on one minibatch, timed with HIP events on
    # the stream they run on
    Xc, _, _ = sur.featurize(te.profile[:65536], d[:65536])
    strm = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    sur._forward_chunk(Xc, 65536)
    ev[0].record(strm)
    for _ in range(10):
        sur._forward_chunk(Xc, 65536)
    ev[1].record(strm)
    g = torch.Generator(device=Xc.device).manual_seed(5)
    idx_b = torch.stack([torch.randperm(sur.X.shape[0], generator=g, device=Xc.device)[:8192]
                         for _ in range(4)]).to(torch.int32).contiguous()
    sur.train_step(idx_b, 8192, 0.0)
    ev[2].record(strm)
    for _ in range(20):
        sur.train_step(idx_b, 8192, 0.0)
    ev[3].record(strm)
    torch.cuda.synchronize()
    t_fwd_k = ev[0].elapsed_time(ev[1]) / 10 * 1e-3
    t_step_k = ev[2].elapsed_time(ev[3]) / 20 * 1e-3
    ns = min(50_000, n_test)
    idx = np.arange(ns)
    ref = judge(idx)
    err = np.abs(thr[:ns] - ref) / ref
    bis_rate = n_train / t_label
    sur_rate = n_test / t_inf
    train_flops = steps * 8192 * 4 * _mlp_flops_per_sample(True)
    inf_flops = n_test * 4 * _mlp_flops_per_sample(False)
    res = {
        "workload": f"{n_train} training fibers (bisection labels, rtol 0.005, FP64), "
                    f"{n_test} held-out fibers; 101 compartments, d ~ U[2,16] um, point "
                    "electrode 0.1-1.5 mm, 0.1 ms cathodic pulse",
        "thresholds_per_s": sur_rate, "unit": "thresholds/s",
        "bisection_thresholds_per_s": bis_rate, "speedup_vs_bisection": sur_rate / bis_rate,
        "bisection_trials_per_threshold": trials,
        "median_error": float(np.median(err)), "p90_error": float(np.quantile(err, 0.9)),
        "error_sample": ns, "fallback_fraction": float(fb.mean()),
        "train_s": t_train, "train_steps": steps, "label_s": t_label,
        "network_thresholds_per_s": n_test / t_net,
        "roofline": {"bound": "mfma", "unit": "TFLOP/s", "peak": 2500.0,
                     "achieved_train": train_flops / t_train / 1e12,
                     "frac_train": train_flops / t_train / 1e12 / 2500.0,
                     "achieved_inference": inf_flops / t_net / 1e12,
                     "frac_inference": inf_flops / t_net / 1e12 / 2500.0,
                     "kernel_inference": {
                         "kernel": "mlp_fwd_kernel (fused forward, 64-row tiles), 65536 rows "
                                   "x 4 members per launch",
                         "us": t_fwd_k * 1e6,
                         "achieved": 65536 * 4 * _mlp_flops_per_sample(False) / t_fwd_k / 1e12,
                         "frac": 65536 * 4 * _mlp_flops_per_sample(False) / t_fwd_k / 1e12
                                 / 2500.0},
                     "kernel_train_step": {
                         "launches": "gather, fused forward, fused backward, partial sums, "
                                     "3 weight-gradient products + reductions, one Adam",
                         "us": t_step_k * 1e6,
                         "achieved": 8192 * 4 * _mlp_flops_per_sample(True) / t_step_k / 1e12,
                         "frac": 8192 * 4 * _mlp_flops_per_sample(True) / t_step_k / 1e12
                                 / 2500.0},
                     "note": "achieved_train / achieved_inference are whole-loop rates (host "
                             "loop, features, ensemble included); kernel_* are the GPU-timed "
                             "launches"},
    }
    if with_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle

        from dendra_amd.model import HH, rect_wave

        cs = threshold_fibers(256, seed=9, start=10_000_000)
        wave = rect_wave(1000, 0.005, 0.5, 0.1, amp=-1.0, dtype=np.float64)
        t = time.perf_counter()
        oracle.threshold(cs.spec, HH(celsius=6.3).array(),
                         cs.profile.astype(np.float64).reshape(-1), wave, 0.005,
                         threads=oracle.cpu_threads())
        dt_cpu = time.perf_counter() - t
        res["cpu_baseline"] = {"value": 256 / dt_cpu, "unit": "thresholds/s",
                               "cores": oracle.cpu_threads(), "kind": "port",
                               "sample": f"256 held-out fibers, float64 C oracle bisection, "
                                         f"OpenMP, {dt_cpu:.1f}s"}
    return res


def _time_op(fn, reps=20, warmup=3):
    """Average milliseconds of fn() on torch's current stream (HIP events)."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / reps


def _rate(ms, nbytes, **extra):
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"ms": round(ms, 4), "bytes": int(nbytes), "GB_s": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 3), **extra}


def operator_bench(dev):
    """Secondary lines: the drop-in operators beside the cable path, float32,
    inputs resident; algorithmic bytes per call over the average call time.

    * tree solves of 2048 systems on a 2000-node random binary tree and on a
      2000-node path: the DHS drop-in (one lane group per system, layer by
      layer) and the tree contraction (O(log depth) rounds), 16 B per node
      (read d, g, rhs; write x);
    * 3x3-block solves, 4096 systems x 201 blocks: block Thomas and block
      Cholesky, 84 B per block row (lower 3, main 9, upper 3, rhs 3, x 3);
      block tree over a 2000-node binary tree x 512 systems, 72 B per node;
    * the dense interpolated-delay transition at the network's size (4.2M
      connections, 21,000 targets, ring depth 201): 16 B per connection
      (w, g, tap base, CSR edge) + 4 B per stored fraction (1 or 3) + 8 B
      per tap (ring cell read and write) + 4 B per target (cleared row) + the
      copy of the ring that the pure transition returns (8 B per cell).
    """
    import torch

    from dendra_amd import blocks, contraction, delays, solvers
    from dendra_amd.topology import build_dhs_layers, build_morphology
    from dendra_amd.workloads import random_binary_tree

    gen = torch.Generator(device=dev).manual_seed(5)
    u = lambda shape, lo, hi: torch.rand(shape, device=dev, generator=gen) * (hi - lo) + lo
    res = {}

    # batched tridiagonal drop-ins (thomas_solve_cuda_t / pcr_solve_cuda_t):
    # 20 B per row (read a, b, c, d; write x)
    # many very short systems: the LDS-staged lane-per-system sweep (auto)
    B, K = 262144, 32
    a, c = u((B, K - 1), -0.4, -0.1), u((B, K - 1), -0.4, -0.1)
    b, d = u((B, K), 1.0, 2.0), u((B, K), -1.0, 1.0)
    with torch.no_grad():
        line = {"systems": B, "rows": K}
        for meth in ("thomas", "pcr", "auto"):
            ms = _time_op(lambda: solvers.solve_tridiagonal(a, b, c, d, meth))
            line[meth] = _rate(ms, 20.0 * B * K)
    res["tridiagonal_very_short"] = line
    for name, (B, K) in (("short", (65536, 201)), ("long", (2048, 4001))):
        a, c = u((B, K - 1), -0.4, -0.1), u((B, K - 1), -0.4, -0.1)
        b, d = u((B, K), 1.0, 2.0), u((B, K), -1.0, 1.0)
        line = {"systems": B, "rows": K}
        with torch.no_grad():
            for meth in ("thomas", "pcr", "auto"):
                ms = _time_op(lambda: solvers.solve_tridiagonal(a, b, c, d, meth))
                line[meth] = _rate(ms, 20.0 * B * K)
        res[f"tridiagonal_{name}"] = line

    B, K = 2048, 2000
    trees = {"binary_tree": np.asarray(random_binary_tree(K, 0, 11)[0]),
             "path": np.arange(-1, K - 1)}
    for name, par in trees.items():
        K = len(par)
        _, _, depth = build_morphology(par)
        order, layers = build_dhs_layers(depth, 16)
        topo = [torch.as_tensor(v, dtype=torch.int64, device=dev) for v in (par, order, layers)]
        plan = contraction.TreeContractionPlan(par, max_segment_interior_nodes=16)
        contraction.prepare(plan, dev)
        d, g, rhs = u((B, K), 0.01, 2.0), u((B, K), 0.1, 5.0), u((B, K), -1.0, 1.0)
        with torch.no_grad():
            solvers.dhs_solve(d, g, rhs, *topo, threads=16)  # builds the drop-in's plan
            ms_dhs = _time_op(lambda: solvers.dhs_solve(d, g, rhs, *topo, threads=16))
            ms_lay = _time_op(lambda: solvers.dhs_solve_layers(d, g, rhs, *topo, threads=16))
            ms_con = _time_op(lambda: contraction.solve(d, g, rhs, plan))
            diff = (solvers.dhs_solve_layers(d, g, rhs, *topo) -
                    solvers.dhs_solve(d, g, rhs, *topo))
        nbytes = 16.0 * B * K
        res[f"tree_solve_{name}"] = {
            "systems": B, "nodes": K, "dhs_layers": len(layers) - 1,
            "dhs_drop_in": _rate(ms_dhs, nbytes), "dhs_layer_kernel": _rate(ms_lay, nbytes),
            "contraction": _rate(ms_con, nbytes),
            "drop_in_speedup_over_layers": round(ms_lay / ms_dhs, 2),
            "max_abs_diff_drop_in_vs_layers": float(diff.abs().max()),
        }

    B, K = 4096, 201
    low, up = u((B, K - 1, 3), -1.0, 1.0), u((B, K - 1, 3), -1.0, 1.0)
    m = u((B, K, 3, 3), -0.5, 0.5)
    eye = torch.eye(3, device=dev)
    main = m + 4.0 * eye
    spd_main = m @ m.transpose(-1, -2) + 4.0 * eye
    rhs = u((B, K, 3), -1.0, 1.0)
    with torch.no_grad():
        ms_bt = _time_op(lambda: blocks.thomas_solve_cuda_bt(low, main, up, rhs))
        ms_spd = _time_op(lambda: blocks.block_spd_solve(low, spd_main, low, rhs))
    res["block_tridiagonal"] = {"systems": B, "blocks": K,
                                "thomas": _rate(ms_bt, 84.0 * B * K),
                                "cholesky": _rate(ms_spd, 84.0 * B * K)}
    B = 512
    par = np.asarray(random_binary_tree(2000, 1, 11)[0])
    K = len(par)
    _, _, depth = build_morphology(par)
    order, layers = build_dhs_layers(depth, 16)
    topo = [torch.as_tensor(v, dtype=torch.int64, device=dev) for v in (par, order, layers)]
    tmain = u((B, K, 3, 3), -0.3, 0.3) + 2.0 * eye
    tedge, trhs = u((B, K, 3), 0.1, 2.0), u((B, K, 3), -1.0, 1.0)
    with torch.no_grad():
        ms_bt_tree = _time_op(lambda: blocks.dhs_bt_solve(tmain, tedge, trhs, *topo))
    res["block_tree"] = {"systems": B, "nodes": K, **_rate(ms_bt_tree, 72.0 * B * K)}

    n, width, depth, dt = 4_200_000, 21_000, 201, 0.025
    post = torch.randint(0, width, (n,), device=dev, generator=gen)
    w = u((n,), 1e-4, 2e-3)
    gates = (u((n,), 0.0, 1.0) < 0.3).float()
    dly = u((n,), 1.0, 5.0)
    q = u((depth, width), 0.0, 1e-3)
    slot = torch.zeros(1, dtype=torch.int64, device=dev)
    step = torch.zeros(1, dtype=torch.int64, device=dev)
    for taps, sigma in ((2, 0.0), (3, 0.35)):
        dtaps = delays.delay_interpolation(dly, dt, dtype=torch.float32, taps=taps,
                                           sigma=sigma)
        call = lambda: delays.dense_interpolated_delay_transition(
            q, slot, step, w, gates, dly, dt, post_idx=post, depth=depth, width=width,
            taps=taps, sigma=sigma, delay_taps=dtaps)
        with torch.no_grad():
            ms = _time_op(call)
        # w, g, base, edge 16 B + fractions (1 or 3 floats) + 8 B per tap, the
        # cleared row, and the copy of the ring the pure transition returns
        nbytes = n * (16.0 + 4.0 * (1 if taps == 2 else 3) + 8.0 * taps) \
            + 4.0 * width + 8.0 * depth * width
        res[f"delay_transition_{taps}tap"] = {
            "connections": n, "targets": width, "depth": depth, **_rate(ms, nbytes)}

    from dendra_amd import history

    n_src = 21_000
    src = torch.randint(0, n_src, (n,), device=dev, generator=gen)
    hist = u((depth, n_src), 0.0, 1.0)
    for taps, sigma in ((2, 0.0), (3, 0.35)):
        dtaps = delays.delay_interpolation(dly, dt, dtype=torch.float32, taps=taps,
                                           sigma=sigma)
        call = lambda: history.source_history_project(
            hist, slot, w, post_idx=post, depth=depth, width=width, source_pos=src,
            delay_taps=dtaps, taps=taps)
        with torch.no_grad():
            ms = _time_op(call)
        # per connection: CSR edge, source column, w 4 B each, tap offset and
        # gate read 4 B each per tap, tap weights (1 or 3 floats); per target
        # the payload
        nbytes = n * (12.0 + 8.0 * taps + 4.0 * (1 if taps == 2 else 3)) + 4.0 * width
        res[f"history_projection_{taps}tap"] = {
            "connections": n, "sources": n_src, "targets": width, "depth": depth,
            **_rate(ms, nbytes)}

    # sparse calendar in steady state: each step pops today's slot and appends
    # the step's deliveries (750 per step, delays 40-200 steps)
    from dendra_amd import calendar

    cap, per_step, n_fill, reps = 262_144, 750, 400, 50
    total_steps = n_fill + reps + 3
    delay_s = torch.randint(40, 201, (total_steps, per_step), device=dev, generator=gen)
    cand_idx = torch.randint(0, width, (total_steps, per_step), device=dev, generator=gen)
    cand_val = u((total_steps, per_step), 1e-4, 2e-3)
    slots_of = [torch.full((1,), k % depth, dtype=torch.int64, device=dev)
                for k in range(total_steps)]
    fut = (torch.arange(total_steps, device=dev)[:, None] + delay_s) % depth
    st = calendar.pending_calendar_empty(cap, device=dev, dtype=torch.float32)
    cur = [0]

    def cal_step():
        nonlocal st
        k = cur[0]
        st, _ = calendar.pending_calendar_pop(st, slots_of[k], width=width)
        st = calendar.pending_calendar_append(st, fut[k], cand_idx[k], cand_val[k])
        cur[0] += 1

    with torch.no_grad():
        for _ in range(n_fill):  # fill to steady state
            cal_step()
        ms = _time_op(cal_step, reps=reps)
    size = int(st.size.item())
    # a pop reads the valid prefix (slot, index, value: 20 B) and writes the
    # capacity of fresh arrays; an append copies the capacity once more
    nbytes = 20.0 * size + 2 * 20.0 * cap
    res["sparse_calendar_step"] = {"capacity": cap, "pending": size,
                                   "appended_per_step": per_step, "targets": width,
                                   **_rate(ms, nbytes)}
    res["extcell_cables"] = extcell_bench(dev)
    res["extcell_trees"] = extcell_tree_bench(dev)
    return res


def extcell_tree_bench(dev, B=1024, n_steps=1000, dt=0.005):
    """ExtCell trees: B copies of one 1000-node random binary tree (per-copy
    diameters and stimulus), float32, per step an assembly launch, the block
    tree solve and the update.  Algorithmic bytes 112 B per node-step: state
    vc and gates read and written (48 B), the prepared radial block (9),
    conductances to the parent (3), area, radial conductances (2) and the
    lead field read (64 B)."""
    import torch

    from dendra_amd.extcell import ExtCellTrees, ExtCellTreeSpec
    from dendra_amd.model import HH
    from dendra_amd.workloads import random_binary_tree

    rng = np.random.default_rng(23)
    par, diam0, pos = random_binary_tree(1000, 0, 13)
    K = len(par)
    lead = 300.0 / (4 * np.pi * 1e-4 * np.hypot(
        np.linalg.norm(pos - pos.mean(0), axis=1)[None, :], rng.uniform(200.0, 1500.0, (B, 1))))
    spec = ExtCellTreeSpec(parent=par, diam=np.tile(diam0, (B, 1)) * rng.uniform(0.8, 1.2, (B, 1)),
                           length=np.full((B, K), 10.0), ra=np.full((B, K), 100.0),
                           cm=np.ones((B, K)), xraxial=np.full((B, K, 2), 1e5),
                           xc=np.zeros((B, K, 2)), xg=np.full((B, K, 2), 10.0), lead=lead,
                           ve_amp=rng.uniform(-1.0, -0.05, B) * 1e-3)
    wave = np.zeros(n_steps)
    wave[200:220] = 1.0
    trees = ExtCellTrees(spec, dt, HH(celsius=6.3), torch.float32, dev)
    trees.run(20, wave=wave)
    torch.cuda.synchronize()
    trees.reset()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    trees.run(n_steps, wave=wave)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    cs = B * K * n_steps
    return {"trees": B, "nodes": K, "time_steps": n_steps,
            "node_steps_per_s": cs / (ms * 1e-3), **_rate(ms, 112.0 * cs)}


def _extcell_spec(B, K, seed):
    from dendra_amd.extcell import ExtCellSpec

    rng = np.random.default_rng(seed)
    lead = 300.0 / (4 * np.pi * 1e-4 * np.hypot(
        (np.arange(K) - K / 2) * 50.0, rng.uniform(200.0, 1500.0, (B, 1))))
    return ExtCellSpec(diam=rng.uniform(4.0, 12.0, (B, K)).round(1), dx=np.full((B, K), 50.0),
                       rhoa=np.full((B, K), 100.0), cm=np.ones((B, K)),
                       xraxial=np.full((B, K, 2), 1e5), xc=np.zeros((B, K, 2)),
                       xg=np.full((B, K, 2), 10.0), lead=lead,
                       ve_amp=rng.uniform(-1.0, -0.05, B) * 1e-3)


def extcell_bench(dev, B=32768, K=101, n_steps=2000, dt=0.005, cpu_cells=512,
                  cpu_steps=2000):
    """HH cables with two extracellular layers (the ExtCell integrator step):
    B x K compartments, float32, a 0.1 ms cathodic point-source pulse.
    Algorithmic bytes 124 B per compartment-step: state vc and gates read and
    written (48 B) and the prepared workspace read (maind 9, couplings 6, area,
    radial conductances 2, lead field: 76 B); the block elimination's own
    factors (12 values written and read back) are not counted."""
    import torch

    from dendra_amd.extcell import ExtCellCables
    from dendra_amd.model import HH

    hh = HH(celsius=6.3)
    spec = _extcell_spec(B, K, 21)
    wave = np.zeros(n_steps)
    wave[200:220] = 1.0
    cells = ExtCellCables(spec, dt, hh, torch.float32, dev)
    cells.run(50, wave=wave)
    torch.cuda.synchronize()
    cells.reset()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    cells.run(n_steps, wave=wave)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    cs = B * K * n_steps
    res = {"cells": B, "compartments": K, "time_steps": n_steps,
           "compartment_steps_per_s": cs / (ms * 1e-3),
           **_rate(ms, 124.0 * cs)}
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    from dendra_amd.extcell import extcell_workspace
    from dendra_amd.model import hh_steady_state

    small = _extcell_spec(cpu_cells, K, 21)
    ws = {k: v.numpy() for k, v in extcell_workspace(small, dt).items()}

    m0, h0, n0 = hh_steady_state(-65.0)
    st = {"vc": np.zeros((cpu_cells, K, 3)), "m": np.full((cpu_cells, K), m0),
          "h": np.full((cpu_cells, K), h0), "n": np.full((cpu_cells, K), n0)}
    st["vc"][..., 0] = -65.0
    threads = oracle.cpu_threads()
    t = time.perf_counter()
    oracle.extcell_run(ws, hh.array(), dt, cpu_steps, st, lead=small.lead, amp=small.ve_amp,
                       wave=wave[:cpu_steps], threads=threads)
    cpu_s = time.perf_counter() - t
    res["cpu_baseline"] = {"value": cpu_cells * K * cpu_steps / cpu_s,
                           "unit": "compartment-steps/s", "cores": threads, "kind": "port",
                           "sample": f"{cpu_cells} cells x {K} compartments x {cpu_steps} "
                                     f"steps, float64 C oracle, {cpu_s:.1f}s"}
    return res


def _measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy), or None."""
    path = ROOT / "profiles" / "hh_population_uneven_traffic.json"
    try:
        return float(json.loads(path.read_text())["traffic_bytes_per_launch"][kernel])
    except (OSError, KeyError, ValueError):
        return None


def _job_traffic(use_c, use_t):
    """HBM bytes of one run's persistent launches (the kernels that run)."""
    parts = ([_measured_traffic("hh_cable_run_kernel")] if use_c else []) + \
            ([_measured_traffic("hh_tree_sec_kernel")] if use_t else [])
    return None if any(p is None for p in parts) else float(sum(parts))


def cpu_baseline(work, hh, frac, n_steps):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    threads = oracle.cpu_threads()
    stride = max(1, int(round(1.0 / frac)))
    cs = work.cables.subset(np.arange(0, work.cables.n_cells, stride))
    ts = work.trees.subset(np.arange(0, work.trees.n_trees, stride))
    wave = work.wave[:n_steps].astype(np.float64)
    t = time.perf_counter()
    oracle.cable_run(cs, hh.array(), work.dt, n_steps, ve_wave=wave, threads=threads)
    oracle.tree_run(ts, hh.array(), work.dt, n_steps, ve_wave=wave, threads=threads)
    dt = time.perf_counter() - t
    n_cs = (cs.total + ts.total) * n_steps
    # one thread on an eighth of that sample (every 8th cell of it)
    cs1 = cs.subset(np.arange(0, cs.n_cells, 8))
    ts1 = ts.subset(np.arange(0, ts.n_trees, 8))
    t = time.perf_counter()
    oracle.cable_run(cs1, hh.array(), work.dt, n_steps, ve_wave=wave, threads=1)
    oracle.tree_run(ts1, hh.array(), work.dt, n_steps, ve_wave=wave, threads=1)
    dt1 = time.perf_counter() - t
    return {
        "value": n_cs / dt, "unit": "compartment-steps/s", "cores": threads,
        "kind": "port",
        "sample": f"every {stride}th cable and tree ({cs.n_cells} cables, {ts.n_trees} "
                  f"trees, {cs.total + ts.total} compartments) x {n_steps} steps, float64 "
                  f"C oracle, OpenMP {threads} threads, {dt:.1f}s",
        "value_1_thread": (cs1.total + ts1.total) * n_steps / dt1,
        "sample_1_thread": f"{cs1.n_cells} cables and {ts1.n_trees} trees of that sample, "
                           f"{n_steps} steps, 1 thread, {dt1:.1f}s",
        "cpu_model": _cpu_model(),
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
