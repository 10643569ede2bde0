"""ctypes wrapper of the CPU oracle ``liboracle.so``.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, never by the dendra_amd package.  The
oracle restates the reference's algorithms in binary64 C (see
dn_oracle.c / dn_oracle_net.c for the per-function citations).

Parity pins (what checks each function against something other than itself):
  thomas / pcr      dense numpy solves, the reference's 3x3 known answer,
                    backward-error bounds, PCR == Thomas on odd/prime sizes
  dhs               dense tree matrix solve, the reference's 2-node and
                    7-node known answers
  vtrap / phi1      direct x/expm1 and expm1(z)/z away from the singularity,
                    mpmath-free closed limits at 0
  cnexp             exact exponential relaxation in closed form
  hh steady/advance HH (1952) rates recomputed independently in numpy;
                    stationarity of the resting state
  exp2syn factor    1 / peak of the normalised two-exponential in closed form
  cable_run         dense backward-Euler step built from the matrix; passive
                    cable equals (I - dt A)^-n v0; uniform equilibrium is
                    stationary; no-drive HH rest stays at rest
  tree_run          an unbranched tree equals the cable run; dense step
  ring_replay       the reference's own brute-force tape (_reference_trace,
                    tests/networks/test_backend_equivalence.py:73-161)
                    including its exact [11.25, 7.0] collision result
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
vp = C.c_void_p


def build() -> Path:
    res = subprocess.run(["make", "-C", str(HERE)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        i64, f64, i32 = C.c_int64, C.c_double, C.c_int
        lib.orc_thomas.argtypes = [dp, dp, dp, dp, dp, i64, i64]
        lib.orc_pcr.argtypes = [dp, dp, dp, dp, dp, i64, i64]
        lib.orc_dhs.argtypes = [dp, dp, dp, lp, lp, lp, dp, i64, i64, i64]
        lib.orc_vtrap.argtypes = [f64, f64]
        lib.orc_vtrap.restype = f64
        lib.orc_phi1.argtypes = [f64]
        lib.orc_phi1.restype = f64
        lib.orc_cnexp.argtypes = [f64, f64, f64, f64]
        lib.orc_cnexp.restype = f64
        lib.orc_exp2syn_factor.argtypes = [f64, f64]
        lib.orc_exp2syn_factor.restype = f64
        lib.orc_hh_steady.argtypes = [f64, C.POINTER(f64), C.POINTER(f64), C.POINTER(f64)]
        lib.orc_hh_advance.argtypes = [f64, f64, dp, f64, f64] + [C.POINTER(f64)] * 5
        run_args = [i64, lp, dp, dp, dp, dp, vp, vp, dp, vp, vp, vp, vp, vp, f64, i64,
                    dp, dp, dp, dp, vp, i64, i32]
        lib.orc_cable_run.argtypes = run_args
        lib.orc_cable_run.restype = i32
        prec_args = [i64, lp, dp, dp, dp, dp, vp, vp, dp, vp, vp, vp, f64, i64,
                     dp, dp, dp, dp, vp, i64, i32]
        for fn in (lib.orc_cable_run_f32, lib.orc_cable_run_f32gates):
            fn.argtypes = prec_args
            fn.restype = i32
        lib.orc_tree_run_f32.argtypes = [i64, lp, lp] + prec_args[2:]
        lib.orc_tree_run_f32.restype = i32
        lib.orc_tree_run.argtypes = [i64, lp, lp] + run_args[2:]
        lib.orc_tree_run.restype = i32
        lib.orc_extcell_run.argtypes = [i64, i64, dp, dp, dp, dp, dp, vp, vp, dp, vp, vp, vp,
                                        vp, vp, f64, i64, dp, dp, dp, dp, vp, i64, i32]
        lib.orc_extcell_run.restype = i32
        lib.orc_threshold.argtypes = [i64, lp, dp, dp, dp, dp, dp, dp, dp, i64, f64, f64, lp,
                                      i32, f64, f64, f64, i32, i32, dp, dp, ip, ip, i32]
        lib.orc_threshold.restype = i32
        lib.orc_delay_steps.argtypes = [dp, f64, i64, lp]
        lib.orc_ring_replay_f32.argtypes = [i64, i64, i64, ip, ip, ip, fp, i64, i64, bp,
                                            fp, fp, C.POINTER(i64)]
        lib.orc_ring_replay_f32.restype = i32
        lib.orc_detect_f32.argtypes = [i64, fp, C.c_float, bp, bp]
        _lib = lib
    return _lib


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def _opt(a, dtype=np.float64):
    if a is None:
        return None, None
    arr = _c(a, dtype)
    return arr, arr.ctypes.data_as(vp)


def thomas(a, b, c, d):
    b = _c(b)
    B, K = b.shape
    x = np.empty_like(b)
    load().orc_thomas(_c(a).reshape(B, K - 1), b, _c(c).reshape(B, K - 1), _c(d), x, B, K)
    return x


def pcr(a, b, c, d):
    b = _c(b)
    B, K = b.shape
    x = np.empty_like(b)
    load().orc_pcr(_c(a).reshape(B, K - 1), b, _c(c).reshape(B, K - 1), _c(d), x, B, K)
    return x


def dhs(d_mem, a_geom, b, parent, order, layer_ptr):
    d_mem = _c(d_mem)
    B, K = d_mem.shape
    x = np.empty_like(d_mem)
    lp_ = _c(layer_ptr, np.int64)
    load().orc_dhs(d_mem, _c(a_geom), _c(b), _c(parent, np.int64), _c(order, np.int64),
                   lp_, x, B, K, len(lp_) - 1)
    return x


def vtrap(x, y):
    return load().orc_vtrap(float(x), float(y))


def phi1(z):
    return load().orc_phi1(float(z))


def cnexp(x, xa, xb, dt):
    return load().orc_cnexp(float(x), float(xa), float(xb), float(dt))


def exp2syn_factor(tau1, tau2):
    return load().orc_exp2syn_factor(float(tau1), float(tau2))


def hh_steady(v):
    m, h, n = C.c_double(), C.c_double(), C.c_double()
    load().orc_hh_steady(float(v), C.byref(m), C.byref(h), C.byref(n))
    return m.value, h.value, n.value


def hh_advance(v, dt, hh_arr, m, h, n, sna=1.0, sk=1.0):
    mm, hh_, nn, G, J = (C.c_double(m), C.c_double(h), C.c_double(n), C.c_double(),
                         C.c_double())
    load().orc_hh_advance(float(v), float(dt), _c(hh_arr), float(sna), float(sk),
                          C.byref(mm), C.byref(hh_), C.byref(nn), C.byref(G), C.byref(J))
    return mm.value, hh_.value, nn.value, G.value, J.value


def _run(fn, n, off, extra_first, spec_arrays, spec, hh_arr, dt, n_steps, ve_wave,
         i_wave, record_every, threads, state):
    total = int(off[-1])
    sna, sna_p = _opt(spec.sna)
    sk, sk_p = _opt(spec.sk)
    lead, lead_p = _opt(spec.lead)
    vea, vea_p = _opt(spec.ve_amp)
    vew, vew_p = _opt(ve_wave)
    iamp, iamp_p = _opt(spec.iamp)
    iw, iw_p = _opt(i_wave)
    if state is None:
        m0, h0, n0 = hh_steady(spec.v_init)
        V = np.full(total, float(spec.v_init))
        M, H, N = np.full(total, m0), np.full(total, h0), np.full(total, n0)
    else:
        V, M, H, N = (np.array(state[k], dtype=np.float64) for k in ("v", "m", "h", "n"))
    rec = rec_p = None
    if record_every > 0:
        rec = np.zeros((n_steps // record_every, total))
        rec_p = rec.ctypes.data_as(vp)
    st = fn(n, _c(off, np.int64), *extra_first, *spec_arrays, sna_p, sk_p, _c(hh_arr),
            lead_p, vea_p, vew_p, iamp_p, iw_p, float(dt), int(n_steps), V, M, H, N,
            rec_p, int(record_every), int(threads))
    if st != 0:
        raise RuntimeError(f"oracle run failed ({st})")
    return {"v": V, "m": M, "h": H, "n": N}, rec


def cable_run(spec, hh_arr, dt, n_steps, ve_wave=None, i_wave=None, record_every=0,
              threads=0, state=None):
    arrays = (_c(spec.diam), _c(spec.dx), _c(spec.rhoa), _c(spec.cm))
    return _run(load().orc_cable_run, spec.n_cells, spec.off, (), arrays, spec, hh_arr,
                dt, n_steps, ve_wave, i_wave, record_every, threads, state)


def cable_run_f32(spec, hh_arr, dt, n_steps, ve_wave=None, record_every=0, threads=0,
                  state=None, binary64_voltage=False):
    """The reference's float32 cable step restated in binary32
    (orc_cable_run_f32), or with binary32 gating and a binary64 voltage,
    assembly and elimination (orc_cable_run_f32gates).  Same arguments and
    results as cable_run (no intracellular current)."""
    lib = load()
    fn = lib.orc_cable_run_f32gates if binary64_voltage else lib.orc_cable_run_f32
    total = spec.total
    sna, sna_p = _opt(spec.sna)
    sk, sk_p = _opt(spec.sk)
    lead, lead_p = _opt(spec.lead)
    vea, vea_p = _opt(spec.ve_amp)
    vew, vew_p = _opt(ve_wave)
    if state is None:
        m0, h0, n0 = hh_steady(spec.v_init)
        V = np.full(total, float(spec.v_init))
        M, H, N = np.full(total, m0), np.full(total, h0), np.full(total, n0)
    else:
        V, M, H, N = (np.array(state[k], dtype=np.float64) for k in ("v", "m", "h", "n"))
    rec = rec_p = None
    if record_every > 0:
        rec = np.zeros((n_steps // record_every, total))
        rec_p = rec.ctypes.data_as(vp)
    st = fn(spec.n_cells, _c(spec.off, np.int64), _c(spec.diam), _c(spec.dx), _c(spec.rhoa),
            _c(spec.cm), sna_p, sk_p, _c(hh_arr), lead_p, vea_p, vew_p, float(dt),
            int(n_steps), V, M, H, N, rec_p, int(record_every), int(threads))
    if st != 0:
        raise RuntimeError(f"oracle run failed ({st})")
    return {"v": V, "m": M, "h": H, "n": N}, rec


def tree_run_f32(spec, hh_arr, dt, n_steps, ve_wave=None, record_every=0, threads=0,
                 state=None):
    """The reference's float32 tree step restated in binary32
    (orc_tree_run_f32); arguments and results as tree_run (no intracellular
    current)."""
    lib = load()
    total = spec.total
    sna, sna_p = _opt(spec.sna)
    sk, sk_p = _opt(spec.sk)
    lead, lead_p = _opt(spec.lead)
    vea, vea_p = _opt(spec.ve_amp)
    vew, vew_p = _opt(ve_wave)
    if state is None:
        m0, h0, n0 = hh_steady(spec.v_init)
        V = np.full(total, float(spec.v_init))
        M, H, N = np.full(total, m0), np.full(total, h0), np.full(total, n0)
    else:
        V, M, H, N = (np.array(state[k], dtype=np.float64) for k in ("v", "m", "h", "n"))
    rec = rec_p = None
    if record_every > 0:
        rec = np.zeros((n_steps // record_every, total))
        rec_p = rec.ctypes.data_as(vp)
    st = lib.orc_tree_run_f32(spec.n_trees, _c(spec.off, np.int64), _c(spec.parent, np.int64),
                              _c(spec.diam), _c(spec.length), _c(spec.ra), _c(spec.cm), sna_p,
                              sk_p, _c(hh_arr), lead_p, vea_p, vew_p, float(dt), int(n_steps),
                              V, M, H, N, rec_p, int(record_every), int(threads))
    if st != 0:
        raise RuntimeError(f"oracle run failed ({st})")
    return {"v": V, "m": M, "h": H, "n": N}, rec


def extcell_run(ws, hh_arr, dt, n_steps, state, *, sna=None, sk=None, lead=None, amp=None,
                wave=None, iamp=None, i_wave=None, record_every=0, threads=0):
    """Double-layer extracellular HH cables (orc_extcell_run).  ws: the
    prepared workspace as numpy arrays -- maind (B,K,3,3), lower/upper
    (B,K-1,3), area (B,K), xg (B,K,2); state: vc (B,K,3), m, h, n (B,K),
    advanced in copies.  Returns (state, rec of v = vi - ve0)."""
    maind = _c(ws["maind"])
    B, K = maind.shape[:2]
    vc = _c(state["vc"]).copy()
    M, H, N = (_c(state[k]).copy() for k in ("m", "h", "n"))
    opt = [_opt(a)[1] for a in (sna, sk)]
    keep = [_opt(a) for a in (lead, amp, wave, iamp, i_wave)]
    rec = rec_p = None
    if record_every > 0:
        rec = np.zeros((n_steps // record_every, B * K))
        rec_p = rec.ctypes.data_as(vp)
    st = load().orc_extcell_run(B, K, maind, _c(ws["lower"]), _c(ws["upper"]), _c(ws["area"]),
                                _c(ws["xg"]), *opt, _c(hh_arr), *(k[1] for k in keep),
                                float(dt), int(n_steps), vc, M, H, N, rec_p,
                                int(record_every), int(threads))
    if st != 0:
        raise RuntimeError(f"oracle run failed ({st})")
    return {"vc": vc, "m": M, "h": H, "n": N}, rec


def tree_run(spec, hh_arr, dt, n_steps, ve_wave=None, i_wave=None, record_every=0,
             threads=0, state=None):
    arrays = (_c(spec.diam), _c(spec.length), _c(spec.ra), _c(spec.cm))
    return _run(load().orc_tree_run, spec.n_trees, spec.off,
                (_c(spec.parent, np.int64),), arrays, spec, hh_arr, dt, n_steps,
                ve_wave, i_wave, record_every, threads, state)


def threshold(spec, hh_arr, lead, wave, dt, node_check=(5, -5), vthr=0.0, ub0=0.1,
              rtol=0.005, fix_up=2.0, max_fix=10, max_bisect=25, threads=0):
    """Bisection thresholds (mA) of every cable of `spec` for Ve = lead*amp*wave.

    Returns (ub, lb, status, trials); node_check entries are resolved per
    fiber like the reference's negative node indices."""
    sizes = np.diff(spec.off)
    if len(set(sizes.tolist())) != 1:
        raise ValueError("oracle thresholds need equal compartment counts")
    K = int(sizes[0])
    chk = np.array([k % K for k in node_check], dtype=np.int64)
    n = spec.n_cells
    ub = np.full(n, float(ub0)) if np.ndim(ub0) == 0 else np.array(ub0, dtype=np.float64)
    lb = np.zeros(n)
    status = np.zeros(n, np.int32)
    trials = np.zeros(n, np.int32)
    wave = _c(wave)
    st = load().orc_threshold(n, _c(spec.off, np.int64), _c(spec.diam), _c(spec.dx),
                              _c(spec.rhoa), _c(spec.cm), _c(hh_arr), _c(lead), wave,
                              len(wave), float(dt), float(spec.v_init), chk, len(chk),
                              float(vthr), float(rtol), float(fix_up), int(max_fix),
                              int(max_bisect), ub, lb, status, trials, int(threads))
    if st != 0:
        raise RuntimeError("oracle threshold failed")
    return ub, lb, status, trials


def delay_steps(delay_ms, dt):
    d = _c(delay_ms)
    out = np.empty(len(d), dtype=np.int64)
    load().orc_delay_steps(d, float(dt), len(d), out)
    return out


def ring_replay(n_src, n_syn, pre, post, delay, w, depth, spikes, ring=None, cur=0):
    spikes = _c(spikes, np.uint8)
    n_steps = spikes.shape[0]
    delivered = np.zeros((n_steps, n_syn), dtype=np.float32)
    ring = np.zeros((depth, n_syn), np.float32) if ring is None else _c(ring, np.float32)
    c = C.c_int64(cur)
    st = load().orc_ring_replay_f32(n_src, n_syn, len(pre), _c(pre, np.int32),
                                    _c(post, np.int32), _c(delay, np.int32),
                                    _c(w, np.float32), depth, n_steps, spikes, delivered,
                                    ring, C.byref(c))
    if st != 0:
        raise ValueError("invalid ring replay arguments")
    return delivered, ring, c.value


def cpu_threads() -> int:
    return len(os.sched_getaffinity(0))


def ring_replay_gates(n_syn, post, delay, w, depth, gates, ring=None, cur=0):
    lib = load()
    lib.orc_ring_replay_gates_f32.argtypes = [C.c_int64, C.c_int64, ip, ip, fp, C.c_int64,
                                              C.c_int64, fp, fp, fp, C.POINTER(C.c_int64)]
    lib.orc_ring_replay_gates_f32.restype = C.c_int
    gates = _c(gates, np.float32)
    n_steps = gates.shape[0]
    delivered = np.zeros((n_steps, n_syn), dtype=np.float32)
    ring = np.zeros((depth, n_syn), np.float32) if ring is None else _c(ring, np.float32)
    c = C.c_int64(cur)
    st = lib.orc_ring_replay_gates_f32(n_syn, len(post), _c(post, np.int32),
                                       _c(delay, np.int32), _c(w, np.float32), depth,
                                       n_steps, gates, delivered, ring, C.byref(c))
    if st != 0:
        raise ValueError("invalid ring replay arguments")
    return delivered, ring, c.value


def uniform(seed, stream, i):
    lib = load()
    lib.orc_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    lib.orc_uniform.restype = C.c_double
    return lib.orc_uniform(seed, stream, i)


def network_run(spec, hh_arr, dt, n_steps, record_spikes=True, v0=None):
    """Binary64 network simulation; returns (state dict, spikes [n_steps][n_src]).

    v0 (optional, one value per compartment) replaces the initial voltage; the
    gates still start at the steady state of spec.cells.v_init."""
    lib = load()
    i64, f64 = C.c_int64, C.c_double
    lib.orc_network_run.argtypes = [i64, lp, lp, dp, dp, dp, dp, vp, vp, dp, dp, f64, i64,
                                    C.c_uint64, C.c_uint32, f64, i64, ip, ip, ip, dp, i64,
                                    f64, i64, dp, dp, dp, dp, dp, dp, bp, vp]
    lib.orc_network_run.restype = C.c_int
    cells, conn = spec.cells, spec.conn
    total, n_cells = cells.total, cells.n_cells
    m0, h0, n0 = hh_steady(cells.v_init)
    V = np.full(total, float(cells.v_init))
    if v0 is not None:
        V[:] = np.asarray(v0, dtype=np.float64)
    M, H, N = np.full(total, m0), np.full(total, h0), np.full(total, n0)
    A, B = np.zeros(n_cells), np.zeros(n_cells)
    latch = np.zeros(n_cells, dtype=np.uint8)
    sna, sna_p = _opt(cells.sna)
    sk, sk_p = _opt(cells.sk)
    ratio = min(max(spec.tau1 / spec.tau2, 1e-9), 0.9999)
    factor = exp2syn_factor(spec.tau1, spec.tau2)
    syn = np.array([factor, ratio * spec.tau2, spec.tau2, spec.e_syn])
    spikes = np.zeros((n_steps, conn.n_src), dtype=np.uint8) if record_spikes else None
    st = lib.orc_network_run(
        n_cells, _c(cells.off, np.int64), _c(spec.soma, np.int64), _c(cells.diam),
        _c(cells.dx), _c(cells.rhoa), _c(cells.cm), sna_p, sk_p, _c(hh_arr), syn,
        float(spec.theta), conn.n_src, spec.seed, spec.bg_stream,
        spec.bg_rate_hz * dt * 1e-3, len(conn.pre), _c(conn.pre, np.int32),
        _c(conn.post, np.int32), _c(conn.delay, np.int32), _c(conn.w), conn.depth,
        float(dt), int(n_steps), V, M, H, N, A, B, latch,
        None if spikes is None else spikes.ctypes.data_as(vp))
    if st != 0:
        raise RuntimeError(f"oracle network run failed ({st})")
    return {"v": V, "m": M, "h": H, "n": N, "A": A, "B": B}, spikes
