"""exp2syn point processes on the GPU, stepped the way the reference steps
them (dendra/models/mod/EXP2SYN.py; Network.step order, net.py:335-377).

``Exp2Syn(comp, tau1=0.1, tau2=10.0, e=0.0)`` holds one synapse per entry of
``comp`` (the compartment it sits on, in the population's user order):

* ``derive_buffers()``: the peak normalisation ``factor`` and
  ``tau1_effective`` with the reference's ratio clamp [1e-9, 0.9999]
  (:112-145), evaluated in binary64;
* ``net_receive(weights)`` (Synapse.net_receive, mechanisms/_mechanism.py:
  4518-4524) and the pure ``net_receive_values(weights, dt, values)``
  (:154-156): A += w factor, B += w factor;
* ``advance(dt)``: the cnexp step of A and B (the integrator calls it before
  the membrane frame, as the reference's mechanism handler does);
* ``i(v)`` / ``i_with_conductance(v)`` (:147-152): (B - A)(v - e) nA and the
  conductance B - A (uS), v being the voltage at each synapse;
* attached to a ``CableStepper`` / ``TreeStepper`` with ``add_point_process``,
  the synapses' conductances enter each step's membrane frame as affine
  terms (G += (B - A) 1e-6 / area, J += G e per compartment).

receive / advance / membrane terms are HIP kernels (``dn_exp2syn_*``).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib


def exp2syn_buffers(tau1, tau2):
    """factor and tau1_effective (EXP2SYN.py:112-145) in binary64."""
    tau1 = np.asarray(tau1, dtype=np.float64)
    tau2 = np.asarray(tau2, dtype=np.float64)
    if not (np.all(np.isfinite(tau1)) and np.all(np.isfinite(tau2)) and np.all(tau1 > 0)
            and np.all(tau2 > 0)):
        raise ValueError("exp2syn tau1 and tau2 must be positive and finite")
    ratio = np.clip(tau1 / tau2, 1e-9, 0.9999)
    log_den = np.log1p(-ratio) + ratio / (1.0 - ratio) * np.log(ratio)
    return np.exp(-log_den), ratio * tau2


def _fn(name, dtype):
    return getattr(_lib.load(), f"{name}_{'f64' if dtype == torch.float64 else 'f32'}")


class Exp2Syn:
    def __init__(self, comp, tau1=0.1, tau2=10.0, e=0.0, *, dtype=torch.float32,
                 device="cuda"):
        self.comp = np.asarray(comp, dtype=np.int64).reshape(-1)
        self.n = len(self.comp)
        self.dtype, self.device = dtype, _lib.device_of(device)
        _lib.require_gpu(torch.empty(0, device=self.device))
        self.tau1, self.tau2, self.e = tau1, tau2, e
        self.derive_buffers()
        self.A = torch.zeros(self.n, dtype=dtype, device=self.device)
        self.B = torch.zeros_like(self.A)

    def _vec_or_scalar(self, x):
        x = np.asarray(x, dtype=np.float64)
        if x.ndim == 0:
            return None, float(x)
        if x.shape != (self.n,):
            raise ValueError(f"per-synapse parameters need {self.n} values")
        return torch.as_tensor(x, dtype=self.dtype, device=self.device), 0.0

    def derive_buffers(self):
        factor, tau1_eff = exp2syn_buffers(self.tau1, self.tau2)
        self.factor, self._factor_all = self._vec_or_scalar(factor)
        self.tau1_effective, self._tau1_all = self._vec_or_scalar(tau1_eff)
        self._tau2, self._tau2_all = self._vec_or_scalar(self.tau2)
        self._e, self._e_all = self._vec_or_scalar(self.e)
        return {"factor": factor, "tau1_effective": tau1_eff}

    def _payload(self, weights):
        w = torch.as_tensor(weights, device=self.device).reshape(-1).to(self.dtype)
        if w.numel() != self.n:
            raise ValueError(f"the payload must hold one weight per synapse ({self.n})")
        return w.contiguous()

    def _receive(self, A, B, w):
        st = _fn("dn_exp2syn_receive", self.dtype)(
            _lib.ptr(A), _lib.ptr(B), _lib.ptr(w), _lib.ptr(self.factor), self._factor_all,
            self.n, _lib.stream_handle(self.device))
        _lib.check(st, "exp2syn receive")

    def net_receive(self, weights, netcon=None):
        """Add this step's delivered weights to A and B (in place)."""
        del netcon
        self._receive(self.A, self.B, self._payload(weights))

    def net_receive_values(self, weights, dt, values):
        """Pure form: new A and B from ``values["A"]``, ``values["B"]``."""
        del dt
        A = values["A"].to(self.dtype).contiguous().clone()
        B = values["B"].to(self.dtype).contiguous().clone()
        self._receive(A, B, self._payload(weights))
        return {"states": {"A": A, "B": B}}

    def advance(self, dt):
        st = _fn("dn_exp2syn_advance", self.dtype)(
            _lib.ptr(self.A), _lib.ptr(self.B), _lib.ptr(self.tau1_effective), self._tau1_all,
            _lib.ptr(self._tau2), self._tau2_all, float(dt), self.n,
            _lib.stream_handle(self.device))
        _lib.check(st, "exp2syn advance")

    def i_with_conductance(self, v):
        g = self.B - self.A
        e = self._e if self._e is not None else self._e_all
        return g * (v - e), g

    def i(self, v):
        return self.i_with_conductance(v)[0]

    # -- membrane frame of a stepper ------------------------------------------

    def _attach(self, plan_index: np.ndarray, area_plan: torch.Tensor, total: int):
        """Compartment CSR in plan order for the stepper that owns the cells."""
        if self.n and (self.comp.min() < 0 or self.comp.max() >= total):
            raise ValueError("synapse compartment index out of range")
        where = plan_index[self.comp]
        order = np.argsort(where, kind="stable")
        comps, counts = np.unique(where[order], return_counts=True)
        ptr = np.zeros(len(comps) + 1, dtype=np.int32)
        np.cumsum(counts, out=ptr[1:])
        i32 = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int32), device=self.device)
        self._csr = (i32(ptr), i32(comps), i32(order), len(comps))
        self._area = area_plan

    def membrane_terms(self, g_out, ge_out):
        ptr, comps, order, n_comp = self._csr
        st = _fn("dn_exp2syn_membrane", self.dtype)(
            _lib.ptr(self.A), _lib.ptr(self.B), _lib.ptr(self._e), self._e_all,
            _lib.ptr(self._area), _lib.ptr(ptr), _lib.ptr(comps), _lib.ptr(order), n_comp,
            _lib.ptr(g_out), _lib.ptr(ge_out), _lib.stream_handle(self.device))
        _lib.check(st, "exp2syn membrane terms")
