"""ctypes binding of the gfx950 C-ABI library ``libdnhip.so``.

The library is the only compute path of this package: if it is missing or
cannot be loaded, every entry point raises instead of falling back to a
CPU or PyTorch implementation.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libdnhip.so"

_lock = threading.Lock()
_lib = None

_ERRORS = {
    -1: "invalid argument",
    -2: "kernel launch failed",
    -3: "size out of range for this kernel",
    -4: "unsupported configuration",
}

p = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
f32 = C.c_float
f64 = C.c_double

_CABLE_RUN = [i32, p, p, p, p, p, p, p, p, p, p, p, p, p, p, p, p, i64, i64]
_TREE_PLAN = [i32, p, p, p, p, p, p, p, p, p, p, p, p, p, i32, i32]

_SIGNATURES = {
    "dn_tridiag_scratch_elems": (i64, [i64, i64, i32, i32]),
    "dn_tridiag_solve_f32": (i32, [p, p, p, p, p, i64, i64, i32, p, i64, p]),
    "dn_tridiag_solve_f64": (i32, [p, p, p, p, p, i64, i64, i32, p, i64, p]),
    "dn_tree_dhs_scratch_elems": (i64, [i64, i64, i32]),
    "dn_tree_dhs_f32": (i32, [p, p, p, p, p, p, p, i64, i64, i64, p, i64, p]),
    "dn_tree_dhs_f64": (i32, [p, p, p, p, p, p, p, i64, i64, i64, p, i64, p]),
    "dn_hh_cable_rmax": (i32, []),
    "dn_hh_cable_run_f32": (i32, _CABLE_RUN + [f32, p, p, i64, i64, i32, i32, p, p]),
    "dn_hh_cable_run_f64": (i32, _CABLE_RUN + [f64, p, p, i64, i64, i32, i32, p, p]),
    "dn_hh_cable_step_f32": (
        i32,
        [i64, p, i64, i32, p, p, p, p, p, p, p, p, p, p, f32, f32, p, p, p, p, p, p, p, p, p],
    ),
    "dn_hh_cable_step_f64": (
        i32,
        [i64, p, i64, i32, p, p, p, p, p, p, p, p, p, p, f64, f64, p, p, p, p, p, p, p, p, p],
    ),
    "dn_hh_tree_slots": (i32, []),
    "dn_hh_tree_sec_max": (i32, []),
    "dn_hh_tree_run_f32": (
        i32,
        _TREE_PLAN
        + [p, i32, p, i32, i32, i32, i32, p, p, p, p, p, p, p, p, p, p, p, i64, i64, f32]
        + [p, p]
        + [i64, i64, p, p],
    ),
    "dn_hh_tree_run_f64": (
        i32,
        _TREE_PLAN
        + [p, i32, p, i32, i32, i32, i32, p, p, p, p, p, p, p, p, p, p, p, i64, i64, f64]
        + [p, p]
        + [i64, i64, p, p],
    ),
    "dn_hh_tree_step_f32": (
        i32,
        _TREE_PLAN + [p, i32, i64, p, p, p, p, p, p, p, p, p, p, f32, f32, p, p, p, p, p, p, p, p],
    ),
    "dn_hh_tree_step_f64": (
        i32,
        _TREE_PLAN + [p, i32, i64, p, p, p, p, p, p, p, p, p, p, f64, f64, p, p, p, p, p, p, p, p],
    ),
}


u32 = C.c_uint32
u64 = C.c_uint64
_SIGNATURES["dn_net_run_f32"] = (
    i32,
    [i32, i32] + [p] * 12 + [f32, p, p, p, p, i32, p, f32, u64, u32, f64]
    + [p] * 7 + [i32] + [p] * 7 + [p, p, i64, p, i64, i64, i32, p],
)
_SIGNATURES["dn_net_deliver_tape_f32"] = (
    i32, [i32, i32] + [p] * 7 + [i32] + [p] * 9 + [i64, p],
)

# n_tasks, rcap, waves, task_nr, task_p, gl, grp, scale, sna, sk, act, n_knots,
# knot_k, knot_a, site, mw, target, tstride, mse_scale, theta, tau, step0, n_steps,
# dt, hh
def _fit_common(f):
    return [i32, i32, i32, p, p, p, p, p, p, p, p, i32, p, p, p, p, p, i64, f, f, f, i64,
            i64, f, p]


_SIGNATURES.update({
    "dn_mlp_gemm": (i32, [i32, i32, i32, i32, i32, i32, p, i64, i64, p, i64, i64, p, i64,
                          p, i64, i64, p, i64, i64, p, i64, i64, p, i64, i64, p]),
    "dn_mlp_out": (i32, [i32, i32, i32, p, i64, p, i64, p, p, i64, p, p, i64, p]),
    "dn_mlp_out_bwd": (i32, [i32, i32, i32, p, i64, p, i64, p, i64, p, p, i64, p]),
    "dn_mlp_row_dot": (i32, [i32, i32, i32, i32, p, i64, p, i64, p, i64, p]),
    "dn_mlp_splitk_reduce": (i32, [i32, i64, i32, p, p, p]),
    "dn_mlp_adam": (i32, [i64, i32, i32, p, p, p, p, f32, f32, f32, f32, p, p, p, p, p]),
    "dn_mlp_forward": (i32, [i32, i32, i32, p, i64, p, p, p, i64, i64, p, p, p, p, p, p, i64]
                       + [p] * 10),
    "dn_mlp_backward": (i32, [i32, i32, i32, p, i64, p, p, p, p, p, p, i64, p, p, p, p, p]),
    "dn_mlp_adam_multi": (i32, [i32] + [p] * 11 + [f32, f32, f32, f32, p]),
    "dn_mlp_partials_reduce": (i32, [i32, i32, p, p, p, p, p, p, p]),
    "dn_mlp_featurize": (i32, [i64, i32, p, p, f32, f32, p, i32, p, p, p, f32, p, p]),
    "dn_mlp_gather": (i32, [i32, i32, i32, p, p, p, p, p, p, p]),
    "dn_mlp_ensemble": (i32, [i64, i32, p, i64, p, f32, f32, f32, p, p, p]),
})

_SIGNATURES["dn_threshold_f64"] = (
    i32, [i32, i32] + [p] * 10 + [i64, f64, p, p, f64, f64, f64, i32, i32] + [p] * 5)

for _sfx, _f in (("f32", f32), ("f64", f64)):
    _SIGNATURES[f"dn_fit_forward_{_sfx}"] = (i32, _fit_common(_f) + [p, p, p, i64, p, p, p])
    _SIGNATURES[f"dn_fit_adjoint_{_sfx}"] = (i32, _fit_common(_f) + [p, p, p, p, p, p])
    _SIGNATURES[f"dn_fit_csr_knots_{_sfx}"] = (
        i32, [i64, i32, p, p, p, p, i64, i64, p, i64, i64, p])
    _SIGNATURES[f"dn_tree_contract_{_sfx}"] = (
        i32, [p, p, p, p, i64, i64, p, i32] + [p] * 5 + [i32] + [p] * 8 + [p, i64, p])

_SIGNATURES["dn_tree_contract_work_elems"] = (i64, [i64, i64])
_SIGNATURES["dn_block_work_elems"] = (i64, [i32, i64, i64])
for _sfx in ("f32", "f64"):
    _SIGNATURES[f"dn_bt3_thomas_{_sfx}"] = (i32, [p] * 5 + [i64, i64, p, i64, p])
    _SIGNATURES[f"dn_bt3_spd_{_sfx}"] = (i32, [p] * 7 + [i64, i64, p, i64, i32, p])
    _SIGNATURES[f"dn_bt3_tree_{_sfx}"] = (i32, [p] * 6 + [i64, p, p, p, i64, i64, p, i64, p])
_SIGNATURES["dn_tree_contract_max_sections"] = (i32, [i32])
_SIGNATURES["dn_bt3_contract_max_sections"] = (i32, [i32, i64])
_SIGNATURES["dn_bt3_contract_work_elems"] = (i64, [i64, i64])
_SIGNATURES["dn_extcell_tree_contract_work_elems"] = (i64, [i64, i64])
for _sfx in ("f32", "f64"):
    _SIGNATURES[f"dn_bt3_tree_contract_{_sfx}"] = (
        i32, [p] * 5 + [i64, i64, p, i32] + [p] * 5 + [i32] + [p] * 8 + [p, i64, p])
_SIGNATURES["dn_interp_delay_work_bytes"] = (i64, [i64, i32, i32])
_SIGNATURES["dn_history_work_bytes"] = (i64, [i64, i32, i32, i32, i64, i64])
for _sfx in ("f32", "f64"):
    _SIGNATURES[f"dn_interp_delay_enqueue_{_sfx}"] = (i32, [p, p, i64, i64, p, i32] + [p] * 6 + [p])
    _SIGNATURES[f"dn_interp_delay_enqueue_ws_{_sfx}"] = (
        i32, [p, p, i64, i64, p, i32] + [p] * 7 + [i64, p, i64, p])
    _SIGNATURES[f"dn_interp_delay_vjp_{_sfx}"] = (
        i32, [p, i64, i64, p, i32, i64] + [p] * 8 + [p])
    _SIGNATURES[f"dn_history_project_{_sfx}"] = (
        i32, [p, i32, i64, i64, p, i32] + [p] * 8 + [i64, p])
    _SIGNATURES[f"dn_history_project_ws_{_sfx}"] = (
        i32, [p, i32, i64, i64, p, i32] + [p] * 9 + [i64, i64, p, i64, i32, p])
    _SIGNATURES[f"dn_history_project_vjp_{_sfx}"] = (
        i32, [p, p, i32, i64, i64, p, i32, i64] + [p] * 9 + [p])
    _SIGNATURES[f"dn_history_record_{_sfx}"] = (i32, [p, i64, i64, p, p, p])
_SIGNATURES["dn_history_record_packed"] = (i32, [p, i64, i64, p, i64, p, p, p, p])
_SIGNATURES["dn_calendar_work_bytes"] = (i64, [i64, i64, i32])
_SIGNATURES["dn_calendar_pop_work_bytes"] = (i64, [i64, i64, i32])
_SIGNATURES["dn_extcell_work_elems"] = (i64, [i64, i64])
_SIGNATURES["dn_extcell_tree_work_elems"] = (i64, [i64, i64])
for _sfx, _ft in (("f32", f32), ("f64", f64)):
    _SIGNATURES[f"dn_extcell_tree_run_{_sfx}"] = (
        i32, [i64, i64] + [p] * 8 + [i64] + [p] * 10 + [_ft, i64] + [p] * 5 + [i64, p, i64, p])
    _SIGNATURES[f"dn_extcell_tree_run_contract_{_sfx}"] = (
        i32, [i64, i64] + [p] * 8 + [i32] + [p] * 5 + [i32] + [p] * 8 + [p] * 8
        + [_ft, i64] + [p] * 5 + [i64, p, i64, p])
for _sfx, _ft in (("f32", f32), ("f64", f64)):
    _SIGNATURES[f"dn_extcell_run_{_sfx}"] = (
        i32, [i64, i64] + [p] * 13 + [_ft, i64] + [p] * 5 + [i64, p, i64, p])
for _sfx in ("f32", "f64", "i32"):
    _SIGNATURES[f"dn_calendar_append_{_sfx}"] = (
        i32, [p, p, p, i64, p, p, p, p, i64, p, p, p, p, p, p, i64, p])
    _SIGNATURES[f"dn_calendar_pop_{_sfx}"] = (
        i32, [p, p, p, i64, p, p, i64, p, p, p, p, p, p, p, p, i64, p])


_SIGNATURES["dn_bitpack_pack"] = (i32, [p, i64, p, i64, i64, p, p])
for _sfx in ("f32", "f64"):
    _SIGNATURES[f"dn_bitpack_deliver_{_sfx}"] = (
        i32, [p, i64, i64, p, p, i64, p, p, p, p, p, i64, i64, p])


_SIGNATURES["dn_dense_event_step"] = (i32, [p, p, i64, i64, p, p, p, p])
_SIGNATURES["dn_ordered_scatter_work_bytes"] = (i64, [i64, i64, i32])
for _sfx in ("f32", "f64"):
    _SIGNATURES[f"dn_ordered_scatter_add_{_sfx}"] = (i32, [p, i64, p, p, i64, p, i64, p])


_SIGNATURES["dn_tridiag_noalloc_max_scratch_elems"] = (i64, [])
_SIGNATURES["dn_tridiag_solve_noalloc_f32"] = (i32, [p, p, p, p, p, i64, i64, i32, p])


for _sfx, _ft in (("f32", f32), ("f64", f64)):
    _SIGNATURES[f"dn_exp2syn_receive_{_sfx}"] = (i32, [p, p, p, p, _ft, i64, p])
    _SIGNATURES[f"dn_exp2syn_advance_{_sfx}"] = (i32, [p, p, p, _ft, p, _ft, _ft, i64, p])
    _SIGNATURES[f"dn_exp2syn_membrane_{_sfx}"] = (
        i32, [p, p, p, _ft, p, p, p, p, i64, p, p, p])


_MULTI_PLAN = [p, p, i64, p, i64, p, p, p, p, p, p, i64, i64, i64, i64]
for _sfx in ("f32", "f64"):
    _SIGNATURES[f"dn_tree_dhs_multi_{_sfx}"] = (i32, [p, p, p] + _MULTI_PLAN + [i32, p, p, p])
    _SIGNATURES[f"dn_tree_dhs_multi_grad_a_{_sfx}"] = (i32, [p, p] + _MULTI_PLAN + [p, p])


def _bind(lib) -> None:
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args


def exported_symbols() -> list[str]:
    """Names this binding knows how to call (every declared C-ABI symbol)."""
    return sorted(_SIGNATURES)


def load():
    """Load ``libdnhip.so`` once; raise ``RuntimeError`` if it is unavailable."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH.name} is not built; run "
                    "`python -c 'import __graft_entry__ as g; g.build()'` first"
                )
            lib = C.CDLL(str(LIB_PATH))
            _bind(lib)
            _lib = lib
    return _lib


def loaded() -> bool:
    return _lib is not None


def check(status: int, what: str) -> None:
    if status != 0:
        msg = _ERRORS.get(int(status), f"error {status}")
        raise RuntimeError(f"{what}: {msg}")


def ptr(t) -> int | None:
    """Raw device (or host) address of a tensor, None for None."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def device_of(device) -> torch.device:
    """torch.device with the index of the current GPU filled in."""
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def require_gpu(*tensors) -> None:
    for t in tensors:
        if t is not None and t.device.type != "cuda":
            raise ValueError("dendra_amd kernels require tensors on a ROCm GPU device")
