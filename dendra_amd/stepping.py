"""Per-step integrators: the stepping interface of the reference's
``_bwd_euler_ub`` and ``_dhs`` integrators on gfx950.

The reference advances a population one step at a time through
``Integrator._step(v, dt, temp, ve=None, intra=None, *, solver=None,
call_local_currents=False) -> (v_new, i_membrane)``
(dendra/models/integrators/core.py:769-834 for unbranched cables with
``_bwd_euler_ub._voltage_update``, implicit.py:569-664; tree.py:1202-1357
for ``_dhs``), called by ``Integrator.step`` from ``Population.run``
(core.py:4555-4614) and by the network driver (net.py:229-239).  ``ve``
(extracellular potential, mV) and ``intra`` (injected current, mA) are
arbitrary tensors per step, so drives that are not one field times one
waveform work.

``CableStepper`` and ``TreeStepper`` keep that contract: ``v`` is not
modified, ``v_new`` is a new tensor of the same shape, the HH gates are the
integrator's own state (the reference keeps them in its mechanism handler)
and advance with every call.  Each call is one gating + assembly launch and
one solve launch (``dn_hh_cable_step_*`` / ``dn_hh_tree_step_*``); ``temp``
sets the rate factor q10 = 3^((temp - 6.3)/10).  Voltages are in user order
(cell by cell, compartments in spec order); a uniform batch may pass
``(B, K)`` tensors.  The membrane-current output of the reference's
``imem`` option is not produced (``i_membrane`` is None).  Point processes
(``dendra_amd.synapse.Exp2Syn``) attached with ``add_point_process`` are
advanced inside each step and enter its membrane frame.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .model import HH, CableSpec, TreeSpec, cable_workspace, hh_steady_state, tree_workspace
from .population import CablePopulation, TreePopulation

_NP = {torch.float32: np.float32, torch.float64: np.float64}


def _hh_array(hh: HH, temp, dtype):
    arr = hh.array()
    if temp is not None:
        arr[6] = 3.0 ** ((float(temp) - 6.3) / 10.0)
    return np.ascontiguousarray(arr.astype(_NP[dtype]))


class _Stepper:
    def _setup_order(self, perm: np.ndarray):
        self._identity = bool(np.array_equal(perm, np.arange(len(perm))))
        dev = self.device
        self._perm = torch.as_tensor(perm, dtype=torch.int64, device=dev)
        inv = np.empty_like(perm)
        inv[perm] = np.arange(len(perm))
        self._inv_np = inv
        self._inv = torch.as_tensor(inv, dtype=torch.int64, device=dev)
        self._pp = None

    def add_point_process(self, syn):
        """Attach an ``Exp2Syn`` whose synapses sit on this population's
        compartments: each step advances it (cnexp) and adds its
        conductances to the membrane frame, as the reference's mechanism
        handler does for point processes."""
        if self._pp is not None:
            raise ValueError("one point-process object per stepper")
        if syn.dtype != self.dtype or syn.device != self.device:
            raise ValueError("the synapses must share the stepper's dtype and device")
        syn._attach(self._inv_np, self._area_plan, self.total)
        self._gx = torch.zeros(self.total, dtype=self.dtype, device=self.device)
        self._gex = torch.zeros_like(self._gx)
        self._pp = syn

    def _point_terms(self, dt):
        if self._pp is None:
            return None, None
        self._pp.advance(dt)
        self._pp.membrane_terms(self._gx, self._gex)
        return _lib.ptr(self._gx), _lib.ptr(self._gex)

    def _to_plan(self, x, name):
        if x is None:
            return None
        if not isinstance(x, torch.Tensor):
            raise TypeError(f"{name} must be a tensor")
        if x.numel() != self.total or x.device != self.device:
            raise ValueError(f"{name} must hold {self.total} values on {self.device}")
        x = x.reshape(-1).to(self.dtype)
        return x.contiguous() if self._identity else x.index_select(0, self._perm)

    def _from_plan(self, x, like):
        x = x if self._identity else x.index_select(0, self._inv)
        return x.reshape(like.shape)

    def initialize(self, v0=None):
        """Gates at the steady state of ``v0`` (scalar or one value per
        compartment, user order; default the spec's v_init), as the
        reference's mechanism initialisation (mod/HH.py:42-56)."""
        if v0 is None:
            v0 = self.spec.v_init
        if np.ndim(v0) == 0:
            m, h, n = hh_steady_state(float(v0))
            for name, val in (("m", m), ("h", h), ("n", n)):
                getattr(self, name).fill_(float(val))
            return
        v0 = torch.as_tensor(v0, dtype=torch.float64).reshape(-1).cpu().numpy()
        m, h, n = hh_steady_state(v0)
        for name, val in (("m", m), ("h", h), ("n", n)):
            t = torch.as_tensor(np.asarray(val), dtype=self.dtype, device=self.device)
            getattr(self, name).copy_(t if self._identity else t[self._perm])

    def gates(self) -> dict:
        """Gate state in user order."""
        return {k: self._from_plan(getattr(self, k), getattr(self, k)).double().cpu().numpy()
                for k in ("m", "h", "n")}

    def _check_call(self, v, solver, call_local_currents):
        if solver is not None:
            raise ValueError("the fused step solves with its own kernels; solver must be None")
        if call_local_currents:
            raise ValueError("functional execution goes through dendra_amd.functional; "
                             "call_local_currents must be False")
        if v.device != self.device:
            raise ValueError(f"v must be on {self.device}")

    def step(self, v, dt, ve=None, intra=None, temp=None):
        """Integrator.step without a model object: returns the new voltage."""
        return self._step(v, dt, self.hh.celsius if temp is None else temp, ve, intra)[0]


class CableStepper(_Stepper):
    """Per-step ``_bwd_euler_ub`` for a ragged batch of unbranched HH cables
    (backward Euler, increment form)."""

    def __init__(self, spec: CableSpec, hh: HH | None = None, dtype=torch.float32,
                 device="cuda"):
        self.spec, self.hh = spec, hh or HH()
        self.dtype, self.device = dtype, _lib.device_of(device)
        _lib.require_gpu(torch.empty(0, device=self.device))
        self._pop = CablePopulation(spec, 0.025, self.hh, dtype=dtype, device=device)
        self.total = spec.total
        self._setup_order(self._pop.perm)
        ws = cable_workspace(spec)
        self._icap = torch.as_tensor(ws.cm_inv[self._pop.perm], dtype=dtype,
                                     device=self.device)
        area = np.pi * spec.diam * 1e-4 * spec.dx * 1e-4
        self._area_plan = torch.as_tensor(area[self._pop.perm], dtype=dtype,
                                          device=self.device)
        self.m, self.h, self.n = self._pop.m, self._pop.h, self._pop.n
        self._mainv = torch.empty(self.total, dtype=dtype, device=self.device)
        self._rhs = torch.empty_like(self._mainv)
        self.initialize()

    def _step(self, v, dt, temp, ve=None, intra=None, *, solver=None,
              call_local_currents=False):
        """One backward-Euler step; returns (v_new, None)."""
        self._check_call(v, solver, call_local_currents)
        p = self._pop
        vn = self._to_plan(v, "v")
        if vn.data_ptr() == v.data_ptr():
            vn = vn.clone()
        ve_p, in_p = self._to_plan(ve, "ve"), self._to_plan(intra, "intra")
        gx, gex = self._point_terms(dt)
        lib = _lib.load()
        fn = lib.dn_hh_cable_step_f64 if self.dtype == torch.float64 else \
            lib.dn_hh_cable_step_f32
        hh = _hh_array(self.hh, temp, self.dtype)
        st = fn(self.spec.n_cells, _lib.ptr(p.off), self.total, p.max_nr_plain, _lib.ptr(vn),
                _lib.ptr(self.m), _lib.ptr(self.h), _lib.ptr(self.n), _lib.ptr(p.gL),
                _lib.ptr(p.gRp), _lib.ptr(p.scale), None, _lib.ptr(p.sna), _lib.ptr(p.sk),
                0.0, float(dt), hh.ctypes.data, _lib.ptr(ve_p), _lib.ptr(in_p),
                _lib.ptr(self._icap), gx, gex, _lib.ptr(self._mainv), _lib.ptr(self._rhs),
                _lib.stream_handle(self.device))
        _lib.check(st, "cable step")
        return self._from_plan(vn, v), None


class TreeStepper(_Stepper):
    """Per-step ``_dhs`` for a ragged batch of HH trees (backward Euler in
    current units, section-chain elimination)."""

    def __init__(self, spec: TreeSpec, hh: HH | None = None, dtype=torch.float32,
                 device="cuda"):
        self.spec, self.hh = spec, hh or HH()
        self.dtype, self.device = dtype, _lib.device_of(device)
        _lib.require_gpu(torch.empty(0, device=self.device))
        self._pop = TreePopulation(spec, 0.025, self.hh, dtype=dtype, device=device)
        self.total = spec.total
        self._setup_order(self._pop.perm)
        self.m, self.h, self.n = self._pop.m, self._pop.h, self._pop.n
        self._area_plan = self._pop.area
        self._mainv = torch.empty(self.total, dtype=dtype, device=self.device)
        self._rhs = torch.empty_like(self._mainv)
        self._dt = None
        self.initialize()

    def _cm_dt(self, dt):
        if self._dt != float(dt):
            ws = tree_workspace(self.spec, float(dt))
            self._cmdt = torch.as_tensor(ws.cm_dt[self._pop.perm], dtype=self.dtype,
                                         device=self.device)
            self._dt = float(dt)
        return self._cmdt

    def _step(self, v, dt, temp, ve=None, intra=None, *, solver=None,
              call_local_currents=False):
        """One backward-Euler step; returns (v_new, None)."""
        self._check_call(v, solver, call_local_currents)
        p = self._pop
        vn = self._to_plan(v, "v")
        if vn.data_ptr() == v.data_ptr():
            vn = vn.clone()
        ve_p, in_p = self._to_plan(ve, "ve"), self._to_plan(intra, "intra")
        gx, gex = self._point_terms(dt)
        lib = _lib.load()
        fn = lib.dn_hh_tree_step_f64 if self.dtype == torch.float64 else \
            lib.dn_hh_tree_step_f32
        hh = _hh_array(self.hh, temp, self.dtype)
        st = fn(*p._plan_args(), _lib.ptr(p.node_tree), p.kmax, self.total, _lib.ptr(vn),
                _lib.ptr(self.m), _lib.ptr(self.h), _lib.ptr(self.n), _lib.ptr(p.area),
                _lib.ptr(self._cm_dt(dt)), _lib.ptr(p.g), None, _lib.ptr(p.sna),
                _lib.ptr(p.sk), 0.0, float(dt), hh.ctypes.data, _lib.ptr(ve_p),
                _lib.ptr(in_p), gx, gex, _lib.ptr(self._mainv), _lib.ptr(self._rhs),
                _lib.stream_handle(self.device))
        _lib.check(st, "tree step")
        return self._from_plan(vn, v), None
