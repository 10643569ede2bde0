// dn_thresh.hip -- activation thresholds of HH cables by on-device bisection.
//
// Restates, per fiber, Thresholder.calculate_thresholds
// (dendra/models/instruments/thresholder.py:873-979): _fix_bounds without
// conduction block (:835-871: ub *= fix_bound_up while inactive, at most
// max_tries_bound_fix times), then arithmetic bisection (calc_stimamp,
// :981-993) while the relative window (ub - lb) / ub (:476-478) is >= rtol,
// at most max_tries_thresh times; fibers that fail either stage are
// reported with status != 0 (the reference's "ignored", NaN).  A trial is
// active under the Active callback (models/callbacks.py:1946-1990 over
// models/_threshold_callbacks.py:50-53): some checked compartment has
// v >= threshold after some step.
//
// Each trial restarts the cable from rest and runs the unbranched
// backward-Euler HH step of dn_cable.hip (implicit.py:569-664) with the
// extracellular potential Ve_j(t) = lead_j * amp * w(t), folded into the
// activating-function coefficient dr_j = gL_j (lead_{j+1} - lead_j) -
// gRp_j (lead_j - lead_{j-1}).
//
// A wavefront holds 64/P fibers and carries all their trials to the end in
// one launch, so the batch needs no host round trip per bisection step (the
// reference synchronises once per step, thresholder.py:954).  A trial stops
// as soon as every unfinished fiber of the wavefront has crossed (the
// outcome is then fixed); inactive trials run all n_steps.
#include "dn_segsolve.hpp"

namespace dn {

template <typename T> struct ThrArgs {
  int n_tasks;
  const int32_t *task_nr, *task_p, *task_cell0, *task_ncell;
  const T *gl, *grp, *sc, *dr;
  const int32_t* chk;
  const T* wave;
  int64_t n_steps;
  T dt;
  HHParams<T> hh;
  T v0, m0, h0, n0, vthr, rtol, fix_up;
  int max_fix, max_bisect;
  T *ub, *lb;
  int32_t *status, *trials;
};

template <typename T, int P, int R>
__device__ void thr_wave(const ThrArgs<T>& A, int task) {
  const int lane = __lane_id();
  const int sub = lane / P;
  const int nr = A.task_nr[task];
  const bool has = sub < A.task_ncell[task];
  const int fiber = A.task_cell0[task] + sub;
  const int64_t base = (int64_t)task * 64 * R + lane;
  T gl[R], grp[R], sc[R], dr[R];
  bool chk[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = base + r * 64;
    const bool in = r < nr;
    gl[r] = in ? A.gl[i] : T(0);
    grp[r] = in ? A.grp[i] : T(0);
    sc[r] = in ? A.sc[i] : T(0);
    dr[r] = in ? A.dr[i] : T(0);
    chk[r] = in && A.chk[i] != 0;
  }
  T ub = has ? A.ub[fiber] : T(1);
  T lb = has ? A.lb[fiber] : T(0);
  // phase 0: fixing the upper bound, 1: bisection, 2: finished
  int phase = has ? 0 : 2;
  int tries = 0, st = 0, ntr = 0;
  const T dt = A.dt, dts = dt * T(1e-3);
  const uint64_t seg = (P == 64) ? ~0ull : ((1ull << P) - 1ull) << (sub * P);
  while (__ballot(phase < 2) != 0ull) {
    const T amp = phase == 0 ? ub : (ub + lb) / T(2);
    T v[R], m[R], h[R], n[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = A.v0; m[r] = A.m0; h[r] = A.h0; n[r] = A.n0;
    }
    bool act = false;
    for (int64_t s = 0; s < A.n_steps; ++s) {
      const T w = amp * A.wave[s];
      T vlast = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r)
        if (r < nr) vlast = v[r];
      const T vprev = shfl_up_seg<P>(vlast, 1);
      const T vnext = shfl_down_seg<P>(v[0], 1);
      T a[R], b[R], c[R], d[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        a[r] = T(0); b[r] = T(1); c[r] = T(0); d[r] = T(0);
        if (r < nr) {
          T G, J;
          hh_advance<T>(v[r], dt, A.hh, T(1), T(1), m[r], h[r], n[r], G, J);
          const T vl = (r == 0) ? vprev : v[r > 0 ? r - 1 : 0];
          const T vr = (r + 1 < nr) ? v[r + 1 < R ? r + 1 : r] : vnext;
          const T f = (J - G * v[r]) * sc[r] + gl[r] * (vr - v[r]) - grp[r] * (v[r] - vl) +
                      dr[r] * w;
          a[r] = -dts * grp[r];
          b[r] = T(1) + dts * (gl[r] + grp[r] + G * sc[r]);
          c[r] = -dts * gl[r];
          d[r] = dts * f;
        }
      }
      seg_solve<T, P, R>(nr, a, b, c, d);
      bool hit = false;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < nr) {
          v[r] += d[r];
          hit = hit || (chk[r] && v[r] >= A.vthr);
        }
      act = act || ((__ballot(hit) & seg) != 0ull);
      if (__ballot(phase < 2 && !act) == 0ull) break;
    }
    if (phase < 2) ++ntr;
    if (phase == 0) {
      if (act) {
        phase = 1;
        tries = 0;
        if ((ub - lb) / ub < A.rtol) phase = 2;
      } else if (tries >= A.max_fix) {
        phase = 2;
        st = 1;
      } else {
        ub *= A.fix_up;
        ++tries;
      }
    } else if (phase == 1) {
      if (act) ub = amp;
      else lb = amp;
      ++tries;
      if ((ub - lb) / ub < A.rtol) {
        phase = 2;
      } else if (tries >= A.max_bisect) {
        phase = 2;
        st = 2;
      }
    }
  }
  if (has && (lane % P) == 0) {
    A.ub[fiber] = ub;
    A.lb[fiber] = lb;
    A.status[fiber] = st;
    A.trials[fiber] = ntr;
  }
}

template <typename T, int R>
__global__ __launch_bounds__(64) void thr_kernel(ThrArgs<T> A) {
  const int task = blockIdx.x;
  if (task >= A.n_tasks) return;
  switch (A.task_p[task]) {
    case 4: thr_wave<T, 4, R>(A, task); break;
    case 8: thr_wave<T, 8, R>(A, task); break;
    case 16: thr_wave<T, 16, R>(A, task); break;
    case 32: thr_wave<T, 32, R>(A, task); break;
    default: thr_wave<T, 64, R>(A, task); break;
  }
}

}  // namespace dn

// Thresholds of a planned fiber batch in slot layout (rcap 4: slot
// task*256 + r*64 + lane).  ub/lb (plan order) hold the initial bounds on
// entry and the final bounds on exit; status 0 = converged, 1 = upper bound
// never activated, 2 = tolerance not met; trials = simulations run.
// hh (host) = {gnabar, gkbar, gl, ena, ek, el, q10}; state0 (host) =
// {v, m, h, n} at rest.
DN_EXPORT int dn_threshold_f64(int n_tasks, int rcap, const int32_t* task_nr,
                               const int32_t* task_p, const int32_t* task_cell0,
                               const int32_t* task_ncell, const double* gl,
                               const double* grp, const double* scale, const double* dr,
                               const int32_t* chk, const double* wave, int64_t n_steps,
                               double dt, const double* hh, const double* state0,
                               double vthr, double rtol, double fix_up, int max_fix,
                               int max_bisect, double* ub, double* lb, int32_t* status,
                               int32_t* trials, hipStream_t stream) {
  if (n_tasks < 0 || rcap != 4 || !task_nr || !task_p || !task_cell0 || !task_ncell ||
      !gl || !grp || !scale || !dr || !chk || !wave || !hh || !state0 || !ub || !lb ||
      !status || !trials || n_steps < 1 || !(rtol > 0.0) || !(fix_up > 1.0) ||
      max_fix < 0 || max_bisect < 0)
    return DN_EINVAL;
  if (n_tasks == 0) return DN_OK;
  dn::ThrArgs<double> A;
  A.n_tasks = n_tasks;
  A.task_nr = task_nr; A.task_p = task_p; A.task_cell0 = task_cell0;
  A.task_ncell = task_ncell;
  A.gl = gl; A.grp = grp; A.sc = scale; A.dr = dr; A.chk = chk; A.wave = wave;
  A.n_steps = n_steps; A.dt = dt;
  A.hh = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};
  A.v0 = state0[0]; A.m0 = state0[1]; A.h0 = state0[2]; A.n0 = state0[3];
  A.vthr = vthr; A.rtol = rtol; A.fix_up = fix_up;
  A.max_fix = max_fix; A.max_bisect = max_bisect;
  A.ub = ub; A.lb = lb; A.status = status; A.trials = trials;
  hipLaunchKernelGGL((dn::thr_kernel<double, 4>), dim3(n_tasks), dim3(64), 0, stream, A);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}
