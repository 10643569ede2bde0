"""Event-driven networks on gfx950: the host driver of the network step.

Counterpart of the reference's ``Network.run`` loop with dense NetCon
delivery (dendra/models/networks/net.py:4485-4545,335-377;
netcon.py:5183-5205) and exp2syn receivers (mod/EXP2SYN.py).  Cells are
unbranched cables (point cells have one compartment) with one exp2syn at the
soma; sources are the cells' somata (threshold crossing) followed by one
Poisson background source per cell.  Connections are numbered target-major,
which is the order the ring accumulates them in.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .model import HH, CableSpec, cable_workspace, hh_steady_state


def delay_steps(delay_ms, dt: float) -> np.ndarray:
    """round(delay / dt) half-to-even, clamped to >= 1 (netcon.py:3930,2776-2778)."""
    steps = np.rint(np.asarray(delay_ms, dtype=np.float64) / dt).astype(np.int64)
    return np.maximum(steps, 1)


def exp2syn_factor(tau1: float, tau2: float) -> tuple[float, float]:
    """Peak normalisation and effective tau1 of exp2syn (mod/EXP2SYN.py:112-145)."""
    ratio = min(max(tau1 / tau2, 1e-9), 0.9999)
    logden = np.log1p(-ratio) + ratio / (1.0 - ratio) * np.log(ratio)
    return float(np.exp(-logden)), float(ratio * tau2)


@dataclass
class Connectivity:
    """Target-major connection list plus the by-source CSR."""

    pre: np.ndarray     # int32 [n_conn]
    post: np.ndarray    # int32 [n_conn], non-decreasing
    delay: np.ndarray   # int32 steps [n_conn]
    w: np.ndarray       # float weights (uS)
    n_src: int
    n_tgt: int

    def __post_init__(self):
        if len(self.post) and (np.diff(self.post) < 0).any():
            raise ValueError("connections must be numbered target-major")
        if (self.delay < 1).any():
            raise ValueError("delays must be >= 1 step")
        if len(self.pre) and (self.pre.min() < 0 or self.pre.max() >= self.n_src):
            raise ValueError("pre index out of range")
        if len(self.post) and (self.post.min() < 0 or self.post.max() >= self.n_tgt):
            raise ValueError("post index out of range")

    @property
    def depth(self) -> int:
        return int(self.delay.max()) + 1

    def csr(self):
        in_ptr = np.zeros(self.n_tgt + 1, dtype=np.int64)
        in_ptr[1:] = np.cumsum(np.bincount(self.post, minlength=self.n_tgt))
        order = np.argsort(self.pre, kind="stable")
        out_ptr = np.zeros(self.n_src + 1, dtype=np.int64)
        out_ptr[1:] = np.cumsum(np.bincount(self.pre, minlength=self.n_src))
        # by source: target and connection index of every connection
        return in_ptr, out_ptr, self.post[order], order


class _Delivery:
    """Device-side connectivity and per-step scratch shared by both modes."""

    def __init__(self, conn: Connectivity, device):
        i32 = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.int32, device=device)
        in_ptr, out_ptr, out_tgt, out_e = conn.csr()
        if conn.pre.size >= 2**31:
            raise ValueError("too many connections for int32 indices")
        self.conn = conn
        self.in_ptr, self.out_ptr, self.out_tgt = i32(in_ptr), i32(out_ptr), i32(out_tgt)
        self.out_e = i32(out_e)
        self.pre, self.delay = i32(conn.pre), i32(conn.delay)
        self.w = torch.as_tensor(conn.w, dtype=torch.float32, device=device)
        self.depth = conn.depth
        self.ring = torch.zeros((self.depth, conn.n_tgt), dtype=torch.float32, device=device)
        self.spk_bits = torch.zeros((conn.n_src + 31) // 32, dtype=torch.int32, device=device)
        self.spk_list = torch.zeros(conn.n_src, dtype=torch.int32, device=device)
        # per target: hits filed this step (zero between steps) and their
        # connection indices (16 per target; more fall back to a scan)
        self.hit_cnt = torch.zeros(conn.n_tgt, dtype=torch.int32, device=device)
        self.hit_e = torch.zeros(conn.n_tgt * 16, dtype=torch.int32, device=device)
        self.tgt_list = torch.zeros(conn.n_tgt, dtype=torch.int32, device=device)
        self.ctr = torch.zeros(8, dtype=torch.int32, device=device)


def deliver_tape(conn: Connectivity, tape, device="cuda") -> np.ndarray:
    """Deliver a source spike tape [n_steps][n_src] through the delay ring on
    the GPU; return the payload delivered to every target at every step."""
    lib = _lib.load()
    tape = np.ascontiguousarray(np.asarray(tape, dtype=np.uint8))
    n_steps = tape.shape[0]
    if tape.shape[1] != conn.n_src:
        raise ValueError("tape width must equal the number of sources")
    d = _Delivery(conn, torch.device(device))
    tape_t = torch.as_tensor(tape, device=device)
    out = torch.empty((n_steps, conn.n_tgt), dtype=torch.float32, device=device)
    st = lib.dn_net_deliver_tape_f32(
        conn.n_tgt, conn.n_src, _lib.ptr(d.out_ptr), _lib.ptr(d.out_tgt), _lib.ptr(d.out_e),
        _lib.ptr(d.in_ptr), _lib.ptr(d.pre), _lib.ptr(d.delay), _lib.ptr(d.w), d.depth,
        _lib.ptr(d.ring), _lib.ptr(d.spk_bits), _lib.ptr(d.spk_list), _lib.ptr(d.hit_cnt),
        _lib.ptr(d.hit_e), _lib.ptr(d.tgt_list), _lib.ptr(d.ctr), _lib.ptr(tape_t),
        _lib.ptr(out), n_steps,
        torch.cuda.current_stream(device).cuda_stream)
    _lib.check(st, "tape delivery")
    return out.cpu().numpy()


@dataclass
class NetworkSpec:
    cells: CableSpec        # point cells first (K = 1), then multi-compartment cells
    n_point: int
    soma: np.ndarray        # flat compartment index of each cell's synapse
    conn: Connectivity      # sources: cells then Poisson background sources
    tau1: float = 0.5
    tau2: float = 5.0
    e_syn: float = 0.0
    theta: float = 0.0
    bg_rate_hz: float = 5.0
    seed: int = 1
    bg_stream: int = 77


class Network:
    """GPU network: float32 cells, exp2syn synapses, dense delay ring."""

    def __init__(self, spec: NetworkSpec, dt: float, hh: HH | None = None, device="cuda"):
        lib = _lib.load()
        self.spec, self.dt, self.hh = spec, float(dt), hh or HH()
        dev = self.device = torch.device(device)
        cells = spec.cells
        if cells.sizes.max() > 256:
            raise ValueError("multi-compartment network cells are limited to 256 compartments")
        if (cells.sizes[: spec.n_point] != 1).any():
            raise ValueError("the first n_point cells must have one compartment")
        ws = cable_workspace(cells)
        f32 = lambda a: None if a is None else torch.as_tensor(
            np.ascontiguousarray(a), dtype=torch.float32, device=dev)
        self.off = torch.as_tensor(cells.off, dtype=torch.int64, device=dev)
        self.soma = torch.as_tensor(spec.soma, dtype=torch.int32, device=dev)
        self.gL, self.gRp, self.scale = f32(ws.gL), f32(ws.gRp), f32(ws.scale)
        self.area = f32(np.pi * cells.diam * 1e-4 * cells.dx * 1e-4)
        self.sna, self.sk = f32(cells.sna), f32(cells.sk)
        self.d = _Delivery(spec.conn, dev)
        factor, tau1_eff = exp2syn_factor(spec.tau1, spec.tau2)
        self.syn = np.array([factor, tau1_eff, spec.tau2, spec.e_syn], dtype=np.float32)
        self.hh_arr = self.hh.array().astype(np.float32)
        self.bg_p = spec.bg_rate_hz * self.dt * 1e-3
        self.stream = torch.cuda.Stream(dev)
        self.reset()

    def reset(self):
        n = self.spec.cells.total
        c = self.spec.cells.n_cells
        v0 = self.spec.cells.v_init
        m0, h0, n0 = hh_steady_state(v0)
        full = lambda x, k: torch.full((k,), float(x), dtype=torch.float32, device=self.device)
        self.v, self.m, self.h, self.n = full(v0, n), full(m0, n), full(h0, n), full(n0, n)
        self.A, self.B = full(0.0, c), full(0.0, c)
        self.latch = torch.zeros(c, dtype=torch.uint8, device=self.device)
        self.d.ring.zero_()
        self.d.ctr.zero_()
        self.d.spk_bits.zero_()
        self.d.hit_cnt.zero_()
        self.t_step = 0

    def run(self, n_steps: int, record_events: bool = False, record_delivered: bool = False,
            graph_steps: int = 0, ev_cap: int | None = None):
        """Advance n_steps; returns (events [n,2] (step, source) sorted, delivered
        [n_steps, n_cells]) for what was requested."""
        lib = _lib.load()
        d = self.d
        ev_log = ev_count = delivered = None
        cap = 0
        if record_events:
            cap = ev_cap or max(1024, n_steps * 64)
            ev_log = torch.zeros((cap, 2), dtype=torch.int32, device=self.device)
            ev_count = torch.zeros(1, dtype=torch.int32, device=self.device)
        if record_delivered:
            if self.t_step:
                raise ValueError("record_delivered needs a run from reset()")
            delivered = torch.zeros((n_steps, self.spec.cells.n_cells), dtype=torch.float32,
                                    device=self.device)
        self.stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.stream):
            st = lib.dn_net_run_f32(
                self.spec.cells.n_cells, self.spec.n_point, _lib.ptr(self.off),
                _lib.ptr(self.soma), _lib.ptr(self.v), _lib.ptr(self.m), _lib.ptr(self.h),
                _lib.ptr(self.n), _lib.ptr(self.gL), _lib.ptr(self.gRp),
                _lib.ptr(self.scale), _lib.ptr(self.sna), _lib.ptr(self.sk),
                _lib.ptr(self.area), self.dt, self.hh_arr.ctypes.data, _lib.ptr(self.A),
                _lib.ptr(self.B), self.syn.ctypes.data, self.spec.conn.n_src,
                _lib.ptr(self.latch), self.spec.theta, self.spec.seed, self.spec.bg_stream,
                self.bg_p, _lib.ptr(d.out_ptr), _lib.ptr(d.out_tgt), _lib.ptr(d.out_e),
                _lib.ptr(d.in_ptr), _lib.ptr(d.pre), _lib.ptr(d.delay), _lib.ptr(d.w), d.depth,
                _lib.ptr(d.ring), _lib.ptr(d.spk_bits), _lib.ptr(d.spk_list),
                _lib.ptr(d.hit_cnt), _lib.ptr(d.hit_e), _lib.ptr(d.tgt_list), _lib.ptr(d.ctr),
                _lib.ptr(ev_log), _lib.ptr(ev_count),
                cap, _lib.ptr(delivered), n_steps if record_delivered else 0, n_steps,
                graph_steps, self.stream.cuda_stream)
        torch.cuda.current_stream(self.device).wait_stream(self.stream)
        _lib.check(st, "network run")
        self.t_step += n_steps
        events = None
        if record_events:
            n_ev = int(ev_count.item())
            if n_ev > cap:
                raise RuntimeError(f"event log overflow ({n_ev} > {cap})")
            ev = ev_log[:n_ev].cpu().numpy()
            events = ev[np.lexsort((ev[:, 1], ev[:, 0]))]
        return events, (None if delivered is None else delivered.cpu().numpy())

    def state(self) -> dict:
        return {k: getattr(self, k).double().cpu().numpy() for k in ("v", "m", "h", "n", "A", "B")}
