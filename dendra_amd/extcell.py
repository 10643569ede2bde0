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
  increment) in one HIP launch for all steps (two lanes per cell, eliminating
  from both ends).  The bath
  potential is ``lead * ve_amp * wave[s]`` (mV), the intracellular current
  ``iamp * i_wave[s]`` (mA).
* branched trees (``_dhs_bt``, integrators/tree_bt.py:91-654):
  ``tree_layered_edge_conductance``, ``prepare_tree_workspace`` (the block-tree
  workspace with the reference's keys) and ``ExtCellTrees(spec, dt, hh)``,
  whose ``run`` takes per step an assembly launch, the block tree solve over
  the DHS layers and an update.
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


# -- branched trees (tree_bt._dhs_bt) -------------------------------------------

def tree_layered_edge_conductance(xraxial, dx, edge_child_orig, edge_parent_orig):
    """Extracellular edge conductance (S) of each tree edge: the two
    half-segment resistances in series; invalid or non-positive resistance
    gives a sealed edge (tree_bt.py:91-121)."""
    cd, pd = dx.index_select(-1, edge_child_orig), dx.index_select(-1, edge_parent_orig)
    cx = xraxial.index_select(-2, edge_child_orig)
    px = xraxial.index_select(-2, edge_parent_orig)
    r = 0.5 * (cx * cd.unsqueeze(-1) + px * pd.unsqueeze(-1)) * 1.0e2
    valid = torch.isfinite(r) & (r > 0)
    return torch.where(valid, r.reciprocal(), torch.zeros_like(r))


def prepare_tree_workspace(dt, *, cm, area, intracellular_edge_conductance,
                           extracellular_edge_conductance, xc, xg, edge_child_orig,
                           solver_order):
    """The block-tree workspace of _dhs_bt._prepare_workspace
    (tree_bt.py:288-372): radial main blocks and per-node conductances to the
    parent (intracellular, two shells) gathered into solver order; zero-area
    branch points carry no radial terms."""
    dt_s = dt.reshape(1) * 1.0e-3 if torch.is_tensor(dt) else dt * 1.0e-3
    cm_dt = 1.0e-6 * cm * area / dt_s
    xc_dt = 1.0e-6 * xc * area.unsqueeze(-1) / dt_s
    radial_g = xg * area.unsqueeze(-1)
    zero_area = area == 0
    cm_dt = torch.where(zero_area, torch.zeros_like(cm_dt), cm_dt)
    xc_dt = torch.where(zero_area.unsqueeze(-1), torch.zeros_like(xc_dt), xc_dt)
    radial_g = torch.where(zero_area.unsqueeze(-1), torch.zeros_like(radial_g), radial_g)
    shell_0 = xc_dt[..., 0] + radial_g[..., 0]
    zero = torch.zeros_like(cm_dt)
    main = torch.stack((torch.stack((cm_dt, -cm_dt, zero), dim=-1),
                        torch.stack((-cm_dt, cm_dt + shell_0, -shell_0), dim=-1),
                        torch.stack((zero, -shell_0, shell_0 + xc_dt[..., 1] + radial_g[..., 1]),
                                    dim=-1)), dim=-2)
    gi, gx = intracellular_edge_conductance, extracellular_edge_conductance
    idx = edge_child_orig.reshape((1,) * (gi.ndim - 1) + (-1,)).expand_as(gi)
    gi_node = torch.zeros_like(area).scatter(-1, idx, gi)
    sidx = edge_child_orig.reshape((1,) * (gx.ndim - 2) + (-1, 1)).expand_as(gx)
    gx_node = torch.zeros_like(xc).scatter(-2, sidx, gx)
    g_to_parent = torch.cat((gi_node.unsqueeze(-1), gx_node), dim=-1).index_select(-2,
                                                                                   solver_order)
    return {"area": area.clone(), "cm_dt": cm_dt, "xc_dt": xc_dt,
            "c_rad": torch.cat((cm_dt.unsqueeze(-1), xc_dt), dim=-1), "xg": radial_g,
            "main_blocks": main.index_select(-3, solver_order), "g_to_parent": g_to_parent}


@dataclass
class ExtCellTreeSpec:
    """B trees of one topology (parent (K,), -1 at the root): per node diam,
    length (um), ra (ohm cm), cm (uF/cm^2) as (B, K); xraxial (MOhm/cm), xc,
    xg as (B, K, 2); optional gNa/gK scales, lead field (B, K) with ve_amp
    (B), intracellular amplitude iamp (B, K)."""

    parent: np.ndarray
    diam: np.ndarray
    length: np.ndarray
    ra: np.ndarray
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


def extcell_tree_workspace(spec: ExtCellTreeSpec, dt: float) -> dict:
    """The prepared block-tree workspace (float64, CPU, node order): lateral
    cylinder areas, intracellular conductance 1 / (R_half_child + R_half_parent)
    with R_half = ra (L/2) / (pi r^2) (integrators/tree.py:560-585), and the
    layered extracellular conductance of each edge."""
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64))
    par = np.asarray(spec.parent, dtype=np.int64)
    diam, length, ra = t(spec.diam), t(spec.length), t(spec.ra)
    area = diam * 1e-4 * PI * length * 1e-4
    r = 1e-4 * diam / 2.0
    rhalf = ra * (1e-4 * length / 2.0) / (PI * r * r)
    child = torch.as_tensor(np.flatnonzero(par >= 0))
    parent = torch.as_tensor(par[par >= 0])
    gi = 1.0 / (rhalf.index_select(-1, child) + rhalf.index_select(-1, parent))
    gx = tree_layered_edge_conductance(t(spec.xraxial), length, child, parent)
    K = len(par)
    return prepare_tree_workspace(dt, cm=t(spec.cm), area=area, intracellular_edge_conductance=gi,
                                  extracellular_edge_conductance=gx, xc=t(spec.xc),
                                  xg=t(spec.xg), edge_child_orig=child,
                                  solver_order=torch.arange(K))


class ExtCellTrees:
    """B HH trees of one topology with two extracellular layers: the
    reference's block-tree step (tree_bt.py:504-654) for all steps in one
    call: per step an assembly kernel (one thread per node: gates, membrane
    block, shell/bath fluxes, axial fluxes pulled from the children), the
    block tree solve over the DHS layers, and the update.  State in node
    order, (B, K, 3)."""

    def __init__(self, spec: ExtCellTreeSpec, dt: float, hh: HH | None = None,
                 dtype=torch.float32, device="cuda", threads: int = 16,
                 workspace: dict | None = None, solver: str = "auto"):
        from .blocks import _children_csr, _contract_cap, contraction_tables
        from .contraction import TreeContractionPlan
        from .topology import build_dhs_layers, build_morphology

        if dtype not in _DT:
            raise TypeError("ExtCellTrees supports float32 and float64")
        self.spec, self.dt, self.dtype = spec, float(dt), dtype
        self.device = torch.device(device)
        self.hh = hh or HH()
        B, K = spec.shape
        par = [int(p) for p in np.asarray(spec.parent)]
        if len(par) != K:
            raise ValueError("parent must hold one entry per node")
        _, _, depth = build_morphology(par)
        # whole depth levels: the block tree kernel then hands the children's
        # eliminated terms over through LDS (its results do not depend on the
        # layering; `threads` only bounds the reference's chunking)
        order, layers = build_dhs_layers(depth, max(K, threads))
        self.B, self.K = B, K
        ws = workspace if workspace is not None else extcell_tree_workspace(spec, dt)
        self.workspace = ws
        dev = lambda a, dt_=dtype: torch.as_tensor(np.asarray(a) if not torch.is_tensor(a) else a,
                                                   dtype=dt_, device=self.device).contiguous()
        self._main = dev(ws["main_blocks"]).reshape(B, K, 9)
        self._edge = dev(ws["g_to_parent"])
        self._area, self._xg = dev(ws["area"]), dev(ws["xg"])
        self._parent = dev(par, torch.int64)
        self._parent32 = dev(par, torch.int32)
        self._order = dev(order, torch.int64)
        self._layers = dev(layers, torch.int64)
        self._child_ptr, self._child_idx = _children_csr(self._parent)
        # the block tree solve: by contraction (sections, then rake/compress
        # rounds) when the plan fits one workgroup's LDS, else over the
        # depth levels
        if solver not in ("auto", "layers", "contraction"):
            raise ValueError("solver must be 'auto', 'layers' or 'contraction'")
        self._plan = None
        if solver != "layers":
            plan = TreeContractionPlan(np.asarray(par, dtype=np.int64), 16)
            elem = torch.tensor([], dtype=dtype).element_size()
            fits = plan.n_sections <= _contract_cap(elem, K)
            if fits:
                self._plan = plan
                self._tables = contraction_tables(plan, self.device)
            elif solver == "contraction":
                raise ValueError("the contraction plan does not fit one workgroup")
        self.solver = "contraction" if self._plan is not None else "layers"
        opt = lambda a: None if a is None else dev(a)
        self._sna, self._sk, self._lead, self._iamp = (opt(spec.sna), opt(spec.sk),
                                                       opt(spec.lead), opt(spec.iamp))
        self._amp = opt(spec.ve_amp)
        self._hh = torch.as_tensor(self.hh.array(), dtype=dtype, device=self.device)
        lib = _lib.load()
        n = int(lib.dn_extcell_tree_work_elems(B, K) if self._plan is None else
                lib.dn_extcell_tree_contract_work_elems(B, K))
        self._work = torch.empty(n, dtype=dtype, device=self.device)
        self.reset()

    def reset(self, v_init: float | None = None):
        v0 = self.spec.v_init if v_init is None else v_init
        m0, h0, n0 = hh_steady_state(v0)
        mk = lambda x: torch.full((self.B, self.K), float(x), dtype=self.dtype,
                                  device=self.device)
        self.vc = torch.zeros(self.B, self.K, 3, dtype=self.dtype, device=self.device)
        self.vc[..., 0] = v0
        self.m, self.h, self.n = mk(m0), mk(h0), mk(n0)
        self.t_step = 0

    @property
    def v(self) -> torch.Tensor:
        return self.vc[..., 0] - self.vc[..., 1]

    def state(self) -> dict:
        g = lambda t: t.double().cpu().numpy()
        return {"vc": g(self.vc), "m": g(self.m), "h": g(self.h), "n": g(self.n)}

    def run(self, n_steps: int, wave=None, i_wave=None, record_every: int = 0, stream=None):
        """Advance n_steps; returns vi - ve0 every record_every steps as
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
            rec = torch.empty(n_steps // record_every, self.B, self.K, dtype=self.dtype,
                              device=self.device)
        p = _lib.ptr
        sfx = "f64" if self.dtype == torch.float64 else "f32"
        handle = _lib.stream_handle(self.device) if stream is None else stream.cuda_stream
        tail = (p(self._sna), p(self._sk), p(self._hh), p(self._lead), p(self._amp), p(wv),
                p(self._iamp), p(iw), self.dt, int(n_steps), p(self.vc), p(self.m),
                p(self.h), p(self.n), p(rec), int(record_every), p(self._work),
                self._work.numel(), handle)
        if self._plan is not None:
            (perm, sec_head, sec_len, sec_child, sec_nchild, sec_parent,
             ptr, sec, par, child, rptr, recv, sptr, src) = self._tables
            st = getattr(lib, f"dn_extcell_tree_run_contract_{sfx}")(
                self.B, self.K, p(self._main), p(self._edge), p(self._area), p(self._xg),
                p(self._parent32), p(self._child_ptr), p(self._child_idx), p(perm),
                self._plan.n_sections, p(sec_head), p(sec_len), p(sec_child), p(sec_nchild),
                p(sec_parent), self._plan.n_rounds, p(ptr), p(sec), p(par), p(child), p(rptr),
                p(recv), p(sptr), p(src), *tail)
        else:
            st = getattr(lib, f"dn_extcell_tree_run_{sfx}")(
                self.B, self.K, p(self._main), p(self._edge), p(self._area), p(self._xg),
                p(self._parent32), p(self._parent), p(self._order), p(self._layers),
                self._layers.numel() - 1, p(self._child_ptr), p(self._child_idx), *tail)
        _lib.check(st, "ExtCell tree run")
        self.t_step += n_steps
        return rec
