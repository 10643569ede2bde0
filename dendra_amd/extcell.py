"""HH cables with two extracellular layers (ExtCell) on the GPU.

Drop-in for the reference's block backward-Euler integrator ``_bwd_euler_bt``
(dendra/models/integrators/implicit.py:669-1136) on HH cables of equal
length, the double-layer extracellular model of the stimulation workflows:

* ``layered_edge_conductance(xraxial, dx)`` (integrators/cable.py:32-45) and
  ``prepare_workspace(dt, *, cm, area, intracellular_edge_conductance,
  extracellular_edge_conductance, xc, xg)`` (implicit.py:800-872): the
  prepared block workspace (lower, upper, maind, area, cm_dt, xc_dt, c_rad,
  xg) with the reference's keys, shapes and units;
* ``ExtCellCables(spec, dt, hh, dtype, device)``: state vc (B, K, 3) =
  (vi, ve0, ve1) and the HH gates; ``run(n_steps, wave=, i_wave=)`` advances
  every cell by the reference's ``_step`` (gates at v = vi - ve0, membrane
  block and residual, axial fluxes, block tridiagonal solve of the
  increment) in one HIP launch: one lane per cell for all steps.  The bath
  potential is ``lead * ve_amp * wave[s]`` (mV), the intracellular current
  ``iamp * i_wave[s]`` (mA).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .model import HH, PI, hh_steady_state

_DT = {torch.float32: np.float32, torch.float64: np.float64}


def layered_edge_conductance(axial_resistance_mohm_per_cm, length_um):
    """Centre-to-centre conductance (S) of each extracellular layer: half of
    each adjoining compartment's longitudinal resistance (cable.py:32-45)."""
    length_cm = length_um * 1.0e-4
    seg = axial_resistance_mohm_per_cm * length_cm.unsqueeze(-1) * 1.0e6
    return 2.0 / (seg[..., :-1, :] + seg[..., 1:, :])


def prepare_workspace(dt, *, cm, area, intracellular_edge_conductance,
                      extracellular_edge_conductance, xc, xg):
    """The two-layer block workspace of _bwd_euler_bt._prepare_workspace
    (implicit.py:800-872); cm, xc in uF/cm^2, xg in S/cm^2, area in cm^2,
    edge conductances in S."""
    dt_s = dt.reshape(1) * 1.0e-3 if torch.is_tensor(dt) else dt * 1.0e-3
    cm_dt = cm * 1.0e-6 * area / dt_s
    xc_dt = xc * 1.0e-6 * area.unsqueeze(-1) / dt_s
    radial_g = xg * area.unsqueeze(-1)
    zero_node = torch.zeros_like(cm_dt[..., :1])
    gi = intracellular_edge_conductance
    gi_left = torch.cat((zero_node, gi), dim=-1)
    gi_right = torch.cat((gi, zero_node), dim=-1)
    zero_shell = torch.zeros_like(xc_dt[..., :1, :])
    gx = extracellular_edge_conductance
    gx_left = torch.cat((zero_shell, gx), dim=-2)
    gx_right = torch.cat((gx, zero_shell), dim=-2)
    shell_0 = xc_dt[..., 0] + radial_g[..., 0]
    d0 = cm_dt + gi_left + gi_right
    d1 = cm_dt + shell_0 + gx_left[..., 0] + gx_right[..., 0]
    d2 = shell_0 + xc_dt[..., 1] + radial_g[..., 1] + gx_left[..., 1] + gx_right[..., 1]
    zero = torch.zeros_like(cm_dt)
    main = torch.stack((torch.stack((d0, -cm_dt, zero), dim=-1),
                        torch.stack((-cm_dt, d1, -shell_0), dim=-1),
                        torch.stack((zero, -shell_0, d2), dim=-1)), dim=-2)
    off = -torch.cat((gi.unsqueeze(-1), gx), dim=-1)
    return {"lower": off.clone(), "upper": off.clone(), "maind": main, "area": area.clone(),
            "cm_dt": cm_dt, "xc_dt": xc_dt,
            "c_rad": torch.cat((cm_dt.unsqueeze(-1), xc_dt), dim=-1), "xg": radial_g}


@dataclass
class ExtCellSpec:
    """B cables of K compartments: diam, dx (um), rhoa (ohm cm), cm
    (uF/cm^2) as (B, K); xraxial (MOhm/cm), xc (uF/cm^2), xg (S/cm^2) as
    (B, K, 2); optional gNa/gK scales (B, K), lead field (B, K, mV/mA) with
    per-cell amplitude ve_amp (B, mA), intracellular amplitude iamp (B, K, mA)."""

    diam: np.ndarray
    dx: np.ndarray
    rhoa: np.ndarray
    cm: np.ndarray
    xraxial: np.ndarray
    xc: np.ndarray
    xg: np.ndarray
    sna: np.ndarray | None = None
    sk: np.ndarray | None = None
    lead: np.ndarray | None = None
    ve_amp: np.ndarray | None = None
    iamp: np.ndarray | None = None
    v_init: float = -65.0

    @property
    def shape(self):
        return tuple(np.shape(self.diam))


def extcell_workspace(spec: ExtCellSpec, dt: float) -> dict:
    """The prepared workspace (float64, CPU) from the cable geometry: lateral
    cylinder area and the half-cylinder intracellular edge conductance
    2 / (R_k + R_{k+1}) (integrators/cable.py:10-29)."""
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    diam, dx, rhoa = t(spec.diam), t(spec.dx), t(spec.rhoa)
    area = diam * 1e-4 * PI * dx * 1e-4
    r = 1e-4 * diam / 2.0
    R = rhoa * (1e-4 * dx) / (PI * r * r)
    gi = 2.0 / (R[:, :-1] + R[:, 1:])
    gx = layered_edge_conductance(t(spec.xraxial), dx)
    return prepare_workspace(dt, cm=t(spec.cm), area=area, intracellular_edge_conductance=gi,
                             extracellular_edge_conductance=gx, xc=t(spec.xc), xg=t(spec.xg))


def _kernel_layout(a, device, dtype):
    """(B, K, *c) -> contiguous (K, prod(c), B) on the device."""
    a = torch.as_tensor(a)
    B, K = a.shape[:2]
    return a.reshape(B, K, -1).permute(1, 2, 0).contiguous().to(device=device, dtype=dtype)


class ExtCellCables:
    """Equal-length HH cables with two extracellular layers (see the module
    doc).  State lives on the device in [compartment][component][cell]
    layout; ``vc``, ``v`` and ``state()`` return the reference's (B, K, ...)
    views."""

    def __init__(self, spec: ExtCellSpec, dt: float, hh: HH | None = None,
                 dtype=torch.float32, device="cuda", workspace: dict | None = None):
        if dtype not in _DT:
            raise TypeError("ExtCellCables supports float32 and float64")
        self.spec, self.dt, self.dtype = spec, float(dt), dtype
        self.device = torch.device(device)
        self.hh = hh or HH()
        B, K = spec.shape
        if B < 1 or K < 1:
            raise ValueError("ExtCellCables needs at least one cell of one compartment")
        self.B, self.K = B, K
        ws = workspace if workspace is not None else extcell_workspace(spec, dt)
        self.workspace = ws
        lay = lambda a: _kernel_layout(a, self.device, dtype)
        self._maind = lay(ws["maind"])
        edge = lambda a: lay(a) if K > 1 else torch.zeros(1, 3, B, dtype=dtype,
                                                          device=self.device)
        self._lower, self._upper = edge(ws["lower"]), edge(ws["upper"])
        self._area, self._xg = lay(ws["area"]), lay(ws["xg"])
        opt = lambda a: None if a is None else lay(a)
        self._sna, self._sk = opt(spec.sna), opt(spec.sk)
        self._lead, self._iamp = opt(spec.lead), opt(spec.iamp)
        self._amp = None if spec.ve_amp is None else torch.as_tensor(
            np.asarray(spec.ve_amp), dtype=dtype, device=self.device).contiguous()
        self._hh = torch.as_tensor(self.hh.array(), dtype=dtype, device=self.device)
        n = int(_lib.load().dn_extcell_work_elems(B, K))
        self._work = torch.empty(n, dtype=dtype, device=self.device)
        self.reset()

    def reset(self, v_init: float | None = None):
        v0 = self.spec.v_init if v_init is None else v_init
        m0, h0, n0 = hh_steady_state(v0)
        K, B = self.K, self.B
        mk = lambda x: torch.full((K, B), float(x), dtype=self.dtype, device=self.device)
        self._vc = torch.zeros(K, 3, B, dtype=self.dtype, device=self.device)
        self._vc[:, 0] = v0
        self._m, self._h, self._n = mk(m0), mk(h0), mk(n0)
        self.t_step = 0

    def set_state(self, vc, m, h, n):
        """State in the reference's shapes: vc (B, K, 3), gates (B, K)."""
        self._vc.copy_(_kernel_layout(vc, self.device, self.dtype))
        for dst, src in ((self._m, m), (self._h, h), (self._n, n)):
            dst.copy_(_kernel_layout(src, self.device, self.dtype)[:, 0])

    @property
    def vc(self) -> torch.Tensor:
        return self._vc.permute(2, 0, 1)

    @property
    def v(self) -> torch.Tensor:
        return (self._vc[:, 0] - self._vc[:, 1]).t()

    def state(self) -> dict:
        """vc (B, K, 3) and the gates (B, K) as float64 numpy arrays."""
        g = lambda t: t.t().double().cpu().numpy()
        return {"vc": self.vc.double().cpu().numpy(), "m": g(self._m), "h": g(self._h),
                "n": g(self._n)}

    def run(self, n_steps: int, wave=None, i_wave=None, record_every: int = 0,
            stream=None):
        """Advance n_steps; returns v = vi - ve0 every record_every steps as
        (n_rec, B, K) when record_every > 0."""
        lib = _lib.load()
        t = lambda w: None if w is None else torch.as_tensor(
            np.asarray(w)[self.t_step:self.t_step + n_steps], dtype=self.dtype,
            device=self.device).contiguous()
        wv, iw = t(wave), t(i_wave)
        for w, name in ((wv, "wave"), (iw, "i_wave")):
            if w is not None and w.numel() < n_steps:
                raise ValueError(f"{name} is shorter than the run")
        rec = None
        if record_every > 0:
            rec = torch.empty(n_steps // record_every, self.K, self.B, dtype=self.dtype,
                              device=self.device)
        p = _lib.ptr
        fn = lib.dn_extcell_run_f64 if self.dtype == torch.float64 else lib.dn_extcell_run_f32
        st = fn(self.B, self.K, p(self._maind), p(self._lower), p(self._upper), p(self._area),
                p(self._xg), p(self._sna), p(self._sk), p(self._hh), p(self._lead),
                p(self._amp), p(wv), p(self._iamp), p(iw), self.dt, int(n_steps),
                p(self._vc), p(self._m), p(self._h), p(self._n), p(rec), int(record_every),
                p(self._work), self._work.numel(),
                _lib.stream_handle(self.device) if stream is None else stream.cuda_stream)
        _lib.check(st, "ExtCell run")
        self.t_step += n_steps
        return None if rec is None else rec.permute(0, 2, 1)
