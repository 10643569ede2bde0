// dn_fit.hip -- trace-fitting loss of a fused HH cable run and its
// hand-written adjoint (reverse-time sweep with checkpoint recomputation).
//
// Forward transition: the unbranched backward-Euler step of the reference's
// _bwd_euler_ub (dendra/models/integrators/implicit.py:569-664) in the order
// of Integrator._step (dendra/models/integrators/core.py:769-834), exactly as
// the persistent kernel of dn_cable.hip, with two differences:
//  * the extracellular potential is a piecewise-linear table in time per
//    compartment (the reference's `arbitrary` waveform,
//    stim/waveform/implementations.py:748-796, interpolated by
//    utils/interpolation.py:1140-1240).  Its activating function (the axial
//    operator applied to Ve, implicit.py:618-638) is precomputed per knot,
//    E_k = A ve_k, so the drive of step n is (1-alpha_n) E_k + alpha_n E_{k+1};
//  * the kernel accumulates the loss over the post-step voltages V_{n+1}
//      L = mse_scale * sum_{n, site} (V_{n+1} - target_n)^2
//        + sum_{n, margin rows} mw * sigmoid((V_{n+1} - theta) / tau)
//    and writes state checkpoints (and optionally the full trajectory).
//
// Adjoint of one step x_n -> x_{n+1} (lam = dL/dV, mu = dL/dgate):
//    lam~  = lam_{n+1} + dl_n/dV_{n+1}
//    y     = M^-T lam~        (transposed band solve, as ThomasSolve.backward,
//                              dendra/models/integrators/triton/t_kernel_thomas.py:147-189)
//    w     = dt_s scale y ;   dL/dG = -w V_{n+1},   dL/dJ = w
//    mu~_g = mu_g + dG/dg' dL/dG + dJ/dg' dL/dJ          (g' = gates of x_{n+1})
//    lam_n = y + sum_g mu~_g dF_g/dv ,   mu_g,n = mu~_g dF_g/dg
//    dL/dsNa += gnabar m'^3 h' w (ENa - V'),  dL/dsK += gkbar n'^4 w (EK - V')
//    dL/dE_k += (1 - alpha) dt_s y,  dL/dE_{k+1} += alpha dt_s y
// F_g is the cnexp update (mechanisms/_cnexp.py:96-132) at v_n; the voltage
// part of the Jacobian of the increment-form step is exactly M^-1, so lam_n
// starts from y.  Every derivative is of the arithmetic the forward runs
// (the Taylor polynomial of phi1, the Bernoulli polynomial of vtrap).
//
// Layout ("slots"): a task (one workgroup of NW wavefronts, W = 64 NW
// threads) owns W * R slots; slot task*W*R + r*W + tid holds row r of
// thread tid.  With NW = 1 a wavefront holds 64 / P cells of P lanes; with
// NW = 2 or 4 (R = 2 or 1) one cell spans the W threads, so that a
// population of few cells still fills every SIMD.  Every per-row array --
// coefficients, state [4][slots], trajectory [(steps+1)][4][slots],
// activating knots [Q][slots] and their gradients -- uses it, so each row
// access of a wavefront is one coalesced 256-byte (512-byte) transaction.
#include "dn_segsolve.hpp"

namespace dn {

template <typename T> struct FitArgs {
  int n_tasks;
  const int32_t* task_nr;
  const int32_t* task_p;
  int64_t slots;
  const T *gl, *grp, *sc, *sna, *sk;
  const T* act;
  int n_knots;
  const int32_t* knot_k;
  const T* knot_a;
  const int32_t* site;
  const T* mw;
  const T* target;
  int64_t tstride;
  T mse_scale, theta, inv_tau;
  int64_t step0, n_steps;
  T dt;
  HHParams<T> hh;
  // forward
  const T* x0;
  T* xout;
  T* ckpt;
  int64_t ckpt_every;
  T* traj;
  T* loss;
  // adjoint
  T* adj;
  T *g_sna, *g_sk, *g_act;
};

// ---------------------------------------------------------------------------
// derivative helpers (exact derivatives of the forward's polynomials)
// ---------------------------------------------------------------------------

// d/dx vtrap10(x): series branch is the derivative of the Bernoulli
// polynomial used by vtrap10, direct branch (e^q - 1 - q e^q) / (e^q - 1)^2.
__device__ __forceinline__ float dvtrap10(float x) {
  const float q = x * 0.1f, q2 = q * q;
  float s = fmaf(q2, -6.0f * 691.0f / 1307674368000.0f, 5.0f / 47900160.0f);
  s = fmaf(s, q2, -4.0f / 1209600.0f);
  s = fmaf(s, q2, 3.0f / 30240.0f);
  s = fmaf(s, q2, -2.0f / 720.0f);
  s = fmaf(s, q2, 1.0f / 12.0f);
  const float series = fmaf(2.0f * q, s, -0.5f);
  const float em = fexp(q) - 1.0f;
  const float ie = frcp(em);
  const float direct = (em - q * (em + 1.0f)) * ie * ie;
  return fabsf(q) <= 1.0f ? series : direct;
}
__device__ __forceinline__ double dvtrap10(double x) {
  const double q = x / 10.0, q2 = q * q;
  double s = q2 * (9.0 * 43867.0 / 5109094217170944000.0) -
             8.0 * 3617.0 / 10670622842880000.0;
  s = s * q2 + 7.0 / 74724249600.0;
  s = s * q2 - 6.0 * 691.0 / 1307674368000.0;
  s = s * q2 + 5.0 / 47900160.0;
  s = s * q2 - 4.0 / 1209600.0;
  s = s * q2 + 3.0 / 30240.0;
  s = s * q2 - 2.0 / 720.0;
  s = s * q2 + 1.0 / 12.0;
  const double series = 2.0 * q * s - 0.5;
  const double em = exp(q) - 1.0;
  const double direct = (em - q * (em + 1.0)) / (em * em);
  return fabs(q) <= 0.5 ? series : direct;
}

// phi1'(z) of the Taylor polynomial phi1 (degree 9 / 16): sum k z^(k-1)/(k+1)!.
// (The float32 forward cuts phi1 after z^4 when a wavefront's |z| <= 1/16; the
// two polynomials differ there by < 1.3e-9, below float32 resolution.)
__device__ __forceinline__ float dphi1(float z) {
  float v = fmaf(z, 1.0f / 403200.0f, 1.0f / 45360.0f);
  v = fmaf(v, z, 1.0f / 5760.0f);
  v = fmaf(v, z, 1.0f / 840.0f);
  v = fmaf(v, z, 1.0f / 144.0f);
  v = fmaf(v, z, 1.0f / 30.0f);
  v = fmaf(v, z, 1.0f / 8.0f);
  v = fmaf(v, z, 1.0f / 3.0f);
  v = fmaf(v, z, 0.5f);
  return v;
}
__device__ __forceinline__ double dphi1(double z) {
  double v = z * (1.0 / 22230464256000.0) + 1.0 / 1394852659200.0;
  v = v * z + 1.0 / 93405312000.0;
  v = v * z + 1.0 / 6706022400.0;
  v = v * z + 1.0 / 518918400.0;
  v = v * z + 1.0 / 43545600.0;
  v = v * z + 1.0 / 3991680.0;
  v = v * z + 1.0 / 403200.0;
  v = v * z + 1.0 / 45360.0;
  v = v * z + 1.0 / 5760.0;
  v = v * z + 1.0 / 840.0;
  v = v * z + 1.0 / 144.0;
  v = v * z + 1.0 / 30.0;
  v = v * z + 1.0 / 8.0;
  v = v * z + 1.0 / 3.0;
  v = v * z + 0.5;
  return v;
}

// Partial derivatives of F = cnexp(x, xa, xb, dt) (both branches of the
// forward): dF/dx and dF/dxa * dxa + dF/dxb * dxb.
template <typename T>
__device__ __forceinline__ void cnexp_jvp(T x, T xa, T xb, T dxa, T dxb, T dt, T& fx,
                                          T& fv) {
  const T z = -xb * dt;
  const T p = phi1(z), dp = dphi1(z);
  const T rate = xa - xb * x;
  const T t_x = T(1) + z * p;
  const T t_v = dt * p * dxa - dt * (x * p + dt * rate * dp) * dxb;
  const T e = fexp(z);
  const T ib = frcp(xb);
  const T eq = xa * ib;
  const T e_v = (T(1) - e) * ib * dxa - (eq * ib * (T(1) - e) + dt * (x - eq) * e) * dxb;
  const bool tay = fabs_(z) <= T(0.5);
  fx = tay ? t_x : e;
  fv = tay ? t_v : e_v;
}

// Adjoint of the HH gate advance at (v, m, h, n): given the adjoints of the
// new gates (am, ah, an), return their contribution to dL/dv and the
// adjoints of the old gates (in place).
template <typename T>
__device__ __forceinline__ T hh_gates_vjp(T v, T dt, const HHParams<T>& p, T m, T h, T n,
                                          T& am_, T& ah_, T& an_) {
  const T q = p.q10;
  const T xm = -(v + T(40)), xn = -(v + T(55));
  const T am = T(0.1) * vtrap10(xm), dam = T(-0.1) * dvtrap10(xm);
  const T bm = T(4) * fexp((v + T(65)) * T(-1.0 / 18.0)), dbm = bm * T(-1.0 / 18.0);
  const T ah = T(0.07) * fexp((v + T(65)) * T(-1.0 / 20.0)), dah = ah * T(-1.0 / 20.0);
  const T bh = frcp(fexp((v + T(35)) * T(-0.1)) + T(1));
  const T dbh = T(0.1) * bh * (T(1) - bh);
  const T an = T(0.01) * vtrap10(xn), dan = T(-0.01) * dvtrap10(xn);
  const T bn = T(0.125) * fexp((v + T(65)) * T(-1.0 / 80.0)), dbn = bn * T(-1.0 / 80.0);
  T fx, fv, gv = T(0);
  cnexp_jvp(m, q * am, q * (am + bm), q * dam, q * (dam + dbm), dt, fx, fv);
  gv += am_ * fv;
  am_ *= fx;
  cnexp_jvp(h, q * ah, q * (ah + bh), q * dah, q * (dah + dbh), dt, fx, fv);
  gv += ah_ * fv;
  ah_ *= fx;
  cnexp_jvp(n, q * an, q * (an + bn), q * dan, q * (dan + dbn), dt, fx, fv);
  gv += an_ * fv;
  an_ *= fx;
  return gv;
}

__device__ __forceinline__ float sigm(float x) { return frcp(1.0f + fexp(-x)); }
__device__ __forceinline__ double sigm(double x) { return 1.0 / (1.0 + exp(-x)); }

template <typename T> __device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}

__device__ __forceinline__ int clamp_knot(int k, int n_knots) {
  return k < 0 ? 0 : (k > n_knots - 2 ? n_knots - 2 : k);
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
template <typename T, int P, int R, int NW>
__device__ __forceinline__ void fit_forward_wave(const FitArgs<T>& A, int task, T* xs) {
  constexpr int W = 64 * NW;  // threads of the task
  const int lane = __lane_id();
  const int nr = A.task_nr[task];
  const int64_t S = A.slots;
  const int64_t base = (int64_t)task * W * R + threadIdx.x;
  T v[R], m[R], h[R], n[R], gl[R], grp[R], sc[R], sna[R], sk[R], e0[R], e1[R], mw[R];
  int site[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = base + r * W;
    const bool in = r < nr;
    v[r] = in ? A.x0[i] : T(0);
    m[r] = in ? A.x0[S + i] : T(0);
    h[r] = in ? A.x0[2 * S + i] : T(0);
    n[r] = in ? A.x0[3 * S + i] : T(0);
    gl[r] = in ? A.gl[i] : T(0);
    grp[r] = in ? A.grp[i] : T(0);
    sc[r] = in ? A.sc[i] : T(0);
    sna[r] = in ? A.sna[i] : T(1);
    sk[r] = in ? A.sk[i] : T(1);
    mw[r] = in ? A.mw[i] : T(0);
    site[r] = in ? A.site[i] : -1;
    e0[r] = e1[r] = T(0);
  }
  const T dt = A.dt, dts = dt * T(1e-3);
  int kc = -1;
  T lmse = T(0), lmar = T(0);
  // per-step inputs are fetched one step ahead, so that their global-memory
  // latency overlaps the step before the one that uses them
  T tg[R];
#pragma unroll
  for (int r = 0; r < R; ++r)
    tg[r] = (r < nr && site[r] >= 0) ? A.target[A.step0 * A.tstride + site[r]] : T(0);
  int kn = A.n_steps > 0 ? clamp_knot(A.knot_k[A.step0], A.n_knots) : 0;
  T aln = A.n_steps > 0 ? A.knot_a[A.step0] : T(0);
  for (int64_t s = 0; s < A.n_steps; ++s) {
    const int64_t ns = A.step0 + s;
    const int k = kn;
    const T al = aln;
    const bool more = s + 1 < A.n_steps;
    T tgn[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      tgn[r] = (more && r < nr && site[r] >= 0) ? A.target[(ns + 1) * A.tstride + site[r]]
                                                 : T(0);
    if (more) {
      kn = clamp_knot(A.knot_k[ns + 1], A.n_knots);
      aln = A.knot_a[ns + 1];
    }
    if (A.traj) {
      T* tr = A.traj + s * 4 * S;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < nr) {
          const int64_t i = base + r * W;
          tr[i] = v[r]; tr[S + i] = m[r]; tr[2 * S + i] = h[r]; tr[3 * S + i] = n[r];
        }
    }
    if (A.ckpt && ns % A.ckpt_every == 0) {
      T* ck = A.ckpt + (ns / A.ckpt_every) * 4 * S;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < nr) {
          const int64_t i = base + r * W;
          ck[i] = v[r]; ck[S + i] = m[r]; ck[2 * S + i] = h[r]; ck[3 * S + i] = n[r];
        }
    }
    if (k != kc) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < nr) {
          e0[r] = A.act[(int64_t)k * S + base + r * W];
          e1[r] = A.act[(int64_t)(k + 1) * S + base + r * W];
        }
      kc = k;
    }
    T vlast = v[0];
#pragma unroll
    for (int r = 1; r < R; ++r)
      if (r < nr) vlast = v[r];
    T vprev, vnext;
    if constexpr (NW == 1) {
      vprev = shfl_up_seg<P>(vlast, 1);
      vnext = shfl_down_seg<P>(v[0], 1);
    } else {
      blk_neighbours<T, NW>(vlast, v[0], xs + 24 * NW, vprev, vnext);
    }
    T a[R], b[R], c[R], d[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      a[r] = T(0); b[r] = T(1); c[r] = T(0); d[r] = T(0);
      if (r < nr) {
        T G, J;
        hh_advance<T>(v[r], dt, A.hh, sna[r], sk[r], m[r], h[r], n[r], G, J);
        const T vl = (r == 0) ? vprev : v[r > 0 ? r - 1 : 0];
        const T vr = (r + 1 < nr) ? v[r + 1 < R ? r + 1 : r] : vnext;
        const T e = (T(1) - al) * e0[r] + al * e1[r];
        const T f = (J - G * v[r]) * sc[r] + gl[r] * (vr - v[r]) - grp[r] * (v[r] - vl) + e;
        a[r] = -dts * grp[r];
        b[r] = T(1) + dts * (gl[r] + grp[r] + G * sc[r]);
        c[r] = -dts * gl[r];
        d[r] = dts * f;
      }
    }
    if constexpr (NW == 1) seg_solve<T, P, R>(nr, a, b, c, d);
    else blk_solve<T, R, NW>(nr, a, b, c, d, xs + (s & 1) * 12 * NW);
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r < nr) {
        v[r] += d[r];
        if (site[r] >= 0) {
          const T df = v[r] - tg[r];
          lmse += df * df;
        }
        if (mw[r] != T(0)) lmar += mw[r] * sigm((v[r] - A.theta) * A.inv_tau);
      }
#pragma unroll
    for (int r = 0; r < R; ++r) tg[r] = tgn[r];
  }
  const int64_t nend = A.step0 + A.n_steps;
  if (A.traj) {
    T* tr = A.traj + A.n_steps * 4 * S;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r < nr) {
        const int64_t i = base + r * W;
        tr[i] = v[r]; tr[S + i] = m[r]; tr[2 * S + i] = h[r]; tr[3 * S + i] = n[r];
      }
  }
  if (A.ckpt && nend % A.ckpt_every == 0) {
    T* ck = A.ckpt + (nend / A.ckpt_every) * 4 * S;
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r < nr) {
        const int64_t i = base + r * W;
        ck[i] = v[r]; ck[S + i] = m[r]; ck[2 * S + i] = h[r]; ck[3 * S + i] = n[r];
      }
  }
  if (A.xout) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r < nr) {
        const int64_t i = base + r * W;
        A.xout[i] = v[r]; A.xout[S + i] = m[r]; A.xout[2 * S + i] = h[r];
        A.xout[3 * S + i] = n[r];
      }
  }
  if (A.loss) {
    lmse = wave_sum(lmse);
    lmar = wave_sum(lmar);
    if (lane == 0) {
      const int64_t o = 2 * ((int64_t)task * NW + (threadIdx.x >> 6));
      A.loss[o] = lmse * A.mse_scale;
      A.loss[o + 1] = lmar;
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint over one segment [step0, step0 + n_steps), trajectory from traj
// ---------------------------------------------------------------------------
template <typename T, int P, int R, int NW>
__device__ __forceinline__ void fit_adjoint_wave(const FitArgs<T>& A, int task, T* xs) {
  constexpr int W = 64 * NW;  // threads of the task
  const int lane = __lane_id();
  // position in the cell's lane segment (P lanes in a wavefront, or all W
  // threads of the workgroup)
  const int pl = NW == 1 ? (lane & (P - 1)) : (int)threadIdx.x;
  constexpr int PL = NW == 1 ? P : W;
  const int nr = A.task_nr[task];
  const int64_t S = A.slots;
  const int64_t base = (int64_t)task * W * R + threadIdx.x;
  const HHParams<T> hp = A.hh;
  T gl[R], grp[R], sc[R], sna[R], sk[R], mw[R], glm[R], grp1[R];
  int site[R];
  T lam[R], mum[R], muh[R], mun[R];
  T vp[R], mp[R], hpp[R], np_[R];
  T gsna[R], gsk[R], acc0[R], acc1[R];
  const T* tre = A.traj + A.n_steps * 4 * S;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = base + r * W;
    const bool in = r < nr;
    gl[r] = in ? A.gl[i] : T(0);
    grp[r] = in ? A.grp[i] : T(0);
    sc[r] = in ? A.sc[i] : T(0);
    sna[r] = in ? A.sna[i] : T(1);
    sk[r] = in ? A.sk[i] : T(1);
    mw[r] = in ? A.mw[i] : T(0);
    site[r] = in ? A.site[i] : -1;
    lam[r] = in ? A.adj[i] : T(0);
    mum[r] = in ? A.adj[S + i] : T(0);
    muh[r] = in ? A.adj[2 * S + i] : T(0);
    mun[r] = in ? A.adj[3 * S + i] : T(0);
    vp[r] = in ? tre[i] : T(0);
    mp[r] = in ? tre[S + i] : T(0);
    hpp[r] = in ? tre[2 * S + i] : T(0);
    np_[r] = in ? tre[3 * S + i] : T(0);
    gsna[r] = gsk[r] = acc0[r] = acc1[r] = T(0);
  }
  // transposed bands: row j couples to j-1 through gL_{j-1}, to j+1 through gRp_{j+1}
  {
    T gl_last = gl[0];
#pragma unroll
    for (int r = 1; r < R; ++r)
      if (r < nr) gl_last = gl[r];
    T glp, grn;
    if constexpr (NW == 1) {
      glp = shfl_up_seg<P>(gl_last, 1);
      grn = shfl_down_seg<P>(grp[0], 1);
    } else {
      blk_neighbours<T, NW>(gl_last, grp[0], xs + 24 * NW, glp, grn);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      glm[r] = r == 0 ? (pl == 0 ? T(0) : glp) : gl[r > 0 ? r - 1 : 0];
      grp1[r] = (r + 1 < nr) ? grp[r + 1 < R ? r + 1 : r] : (pl == PL - 1 ? T(0) : grn);
    }
  }
  const T dt = A.dt, dts = dt * T(1e-3);
  int kc = clamp_knot(A.knot_k[A.step0 + A.n_steps - 1], A.n_knots);
  // the inputs of step s -- knots, targets and the trajectory row (pre-step
  // state) -- are fetched during step s + 1, so that their global-memory
  // latency overlaps the step before the one that uses them
  int kn = kc;
  T aln = A.knot_a[A.step0 + A.n_steps - 1];
  T tg[R], q0[R], q1[R], q2[R], q3[R];
  {
    const int64_t ns = A.step0 + A.n_steps - 1;
    const T* tr = A.traj + (A.n_steps - 1) * 4 * S;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = base + r * W;
      const bool in = r < nr;
      tg[r] = (in && site[r] >= 0) ? A.target[ns * A.tstride + site[r]] : T(0);
      q0[r] = in ? tr[i] : T(0);
      q1[r] = in ? tr[S + i] : T(0);
      q2[r] = in ? tr[2 * S + i] : T(0);
      q3[r] = in ? tr[3 * S + i] : T(0);
    }
  }
  for (int64_t s = A.n_steps - 1; s >= 0; --s) {
    const int64_t ns = A.step0 + s;
    const int k = kn;
    const T al = aln;
    T tgc[R], v0r[R], m0r[R], h0r[R], n0r[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      tgc[r] = tg[r];
      v0r[r] = q0[r];
      m0r[r] = q1[r];
      h0r[r] = q2[r];
      n0r[r] = q3[r];
    }
    if (s > 0) {
      kn = clamp_knot(A.knot_k[ns - 1], A.n_knots);
      aln = A.knot_a[ns - 1];
      const T* tr = A.traj + (s - 1) * 4 * S;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int64_t i = base + r * W;
        const bool in = r < nr;
        tg[r] = (in && site[r] >= 0) ? A.target[(ns - 1) * A.tstride + site[r]] : T(0);
        q0[r] = in ? tr[i] : T(0);
        q1[r] = in ? tr[S + i] : T(0);
        q2[r] = in ? tr[2 * S + i] : T(0);
        q3[r] = in ? tr[3 * S + i] : T(0);
      }
    }
    if (k != kc) {
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < nr) {
          const int64_t i = base + r * W;
          A.g_act[(int64_t)(kc + 1) * S + i] += acc1[r];
          if (k == kc - 1) {
            acc1[r] = acc0[r];
          } else {
            A.g_act[(int64_t)kc * S + i] += acc0[r];
            acc1[r] = T(0);
          }
          acc0[r] = T(0);
        }
      kc = k;
    }
    T a[R], b[R], c[R], d[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      a[r] = T(0); b[r] = T(1); c[r] = T(0); d[r] = T(0);
      if (r < nr) {
        T lt = lam[r];
        if (site[r] >= 0) lt += T(2) * A.mse_scale * (vp[r] - tgc[r]);
        if (mw[r] != T(0)) {
          const T sg = sigm((vp[r] - A.theta) * A.inv_tau);
          lt += mw[r] * sg * (T(1) - sg) * A.inv_tau;
        }
        const T m3h = mp[r] * mp[r] * mp[r] * hpp[r];
        const T n2 = np_[r] * np_[r];
        const T G = hp.gl + hp.gnabar * sna[r] * m3h + hp.gkbar * sk[r] * (n2 * n2);
        a[r] = -dts * glm[r];
        b[r] = T(1) + dts * (gl[r] + grp[r] + G * sc[r]);
        c[r] = -dts * grp1[r];
        d[r] = lt;
      }
    }
    if constexpr (NW == 1) seg_solve<T, P, R>(nr, a, b, c, d);
    else blk_solve<T, R, NW>(nr, a, b, c, d, xs + (s & 1) * 12 * NW);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r < nr) {
        const T v0 = v0r[r], m0 = m0r[r], h0 = h0r[r], n0 = n0r[r];
        const T y = d[r];
        const T w = dts * sc[r] * y;
        const T tna = w * (hp.ena - vp[r]);
        const T tk = w * (hp.ek - vp[r]);
        const T m2 = mp[r] * mp[r];
        const T n2 = np_[r] * np_[r];
        const T n3 = n2 * np_[r];
        gsna[r] += hp.gnabar * (m2 * mp[r] * hpp[r]) * tna;
        gsk[r] += hp.gkbar * (n2 * n2) * tk;
        T am = mum[r] + T(3) * hp.gnabar * sna[r] * m2 * hpp[r] * tna;
        T ah = muh[r] + hp.gnabar * sna[r] * (m2 * mp[r]) * tna;
        T an = mun[r] + T(4) * hp.gkbar * sk[r] * n3 * tk;
        const T gv = hh_gates_vjp<T>(v0, dt, hp, m0, h0, n0, am, ah, an);
        lam[r] = y + gv;
        mum[r] = am;
        muh[r] = ah;
        mun[r] = an;
        const T gf = dts * y;
        acc0[r] += (T(1) - al) * gf;
        acc1[r] += al * gf;
        vp[r] = v0;
        mp[r] = m0;
        hpp[r] = h0;
        np_[r] = n0;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (r < nr) {
      const int64_t i = base + r * W;
      A.adj[i] = lam[r];
      A.adj[S + i] = mum[r];
      A.adj[2 * S + i] = muh[r];
      A.adj[3 * S + i] = mun[r];
      A.g_sna[i] += gsna[r];
      A.g_sk[i] += gsk[r];
      A.g_act[(int64_t)kc * S + i] += acc0[r];
      A.g_act[(int64_t)(kc + 1) * S + i] += acc1[r];
    }
}

// One task per workgroup: NW = 1, a wavefront of cells of task_p lanes each;
// NW > 1, one cell over the 64 NW threads.
template <typename T, int R, bool ADJ, int NW>
__global__ __launch_bounds__(64 * NW) void fit_kernel(FitArgs<T> A) {
  // cross-wavefront exchange areas (NW > 1): two alternating solve areas of
  // 12 NW values and one neighbour area of 2 NW
  __shared__ T xs[NW > 1 ? 26 * NW : 1];
  const int task = blockIdx.x;
  if (task >= A.n_tasks) return;
  if constexpr (NW > 1) {
    if constexpr (ADJ) fit_adjoint_wave<T, 64 * NW, R, NW>(A, task, xs);
    else fit_forward_wave<T, 64 * NW, R, NW>(A, task, xs);
    return;
  }
  switch (A.task_p[task]) {
#define DN_FIT_CASE(PP)                                             \
  case PP:                                                          \
    if constexpr (ADJ) fit_adjoint_wave<T, PP, R, 1>(A, task, xs);  \
    else fit_forward_wave<T, PP, R, 1>(A, task, xs);                \
    break;
    DN_FIT_CASE(4)
    DN_FIT_CASE(8)
    DN_FIT_CASE(16)
    DN_FIT_CASE(32)
    default:
    DN_FIT_CASE(64)
#undef DN_FIT_CASE
  }
}

// out[i*out_row + k*out_k] = sum_{e in [ptr[i], ptr[i+1])} w[e] in[idx[e]*in_row + k*in_k]
template <typename T>
__global__ __launch_bounds__(256) void csr_knots_kernel(int64_t n_out, int n_k,
                                                        const int64_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ idx,
                                                        const T* __restrict__ w,
                                                        const T* __restrict__ in,
                                                        int64_t in_row, int64_t in_k,
                                                        T* __restrict__ out,
                                                        int64_t out_row, int64_t out_k) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_out * n_k) return;
  const int64_t i = t % n_out;
  const int64_t k = t / n_out;
  T acc = T(0);
  for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) acc += w[e] * in[idx[e] * in_row + k * in_k];
  out[i * out_row + k * out_k] = acc;
}

template <typename T>
static int fit_launch(const FitArgs<T>& A, int rcap, int waves, bool adj,
                      hipStream_t stream) {
  if (A.n_tasks < 0 || !A.task_nr || !A.task_p || !A.gl || !A.grp || !A.sc || !A.sna ||
      !A.sk || !A.act || !A.knot_k || !A.knot_a || !A.site || !A.mw || !A.target)
    return DN_EINVAL;
  if (!((waves == 1 && (rcap == 4 || rcap == 8)) || (waves == 2 && rcap == 2) ||
        (waves == 4 && rcap == 1)))
    return DN_EINVAL;
  if (A.slots != (int64_t)A.n_tasks * 64 * waves * rcap || A.n_knots < 2 || A.tstride < 0 ||
      A.step0 < 0 || A.n_steps < 0)
    return DN_EINVAL;
  if (adj) {
    if (!A.traj || !A.adj || !A.g_sna || !A.g_sk || !A.g_act) return DN_EINVAL;
  } else {
    if (!A.x0 || (A.ckpt && A.ckpt_every < 1)) return DN_EINVAL;
  }
  if (A.n_tasks == 0 || A.n_steps == 0) return DN_OK;
  const dim3 grid(A.n_tasks), block(64 * waves);
#define DN_FIT_GO(RR, NW)                                                                \
  {                                                                                     \
    if (adj) hipLaunchKernelGGL((fit_kernel<T, RR, true, NW>), grid, block, 0, stream, A); \
    else hipLaunchKernelGGL((fit_kernel<T, RR, false, NW>), grid, block, 0, stream, A);    \
  }
  if (waves == 4) DN_FIT_GO(1, 4)
  else if (waves == 2) DN_FIT_GO(2, 2)
  else if (rcap == 4) DN_FIT_GO(4, 1)
  else DN_FIT_GO(8, 1)
#undef DN_FIT_GO
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
static int csr_knots_launch(int64_t n_out, int n_k, const int64_t* ptr, const int32_t* idx,
                            const T* w, const T* in, int64_t in_row, int64_t in_k, T* out,
                            int64_t out_row, int64_t out_k, hipStream_t stream) {
  if (n_out < 0 || n_k < 0 || !ptr || !out || (n_out > 0 && n_k > 0 && (!in)))
    return DN_EINVAL;
  const int64_t total = n_out * (int64_t)n_k;
  if (total == 0) return DN_OK;
  const int64_t blocks = (total + 255) / 256;
  if (blocks > 0x7fffffff) return DN_ERANGE;
  hipLaunchKernelGGL(csr_knots_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, stream,
                     n_out, n_k, ptr, idx, w, in, in_row, in_k, out, out_row, out_k);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace dn

#define DN_FIT_COMMON(T)                                                              \
  int n_tasks, int rcap, int waves, const int32_t *task_nr, const int32_t *task_p,   \
      const T *gl, const T *grp, const T *scale, const T *sna, const T *sk,          \
      const T *act, int n_knots, const int32_t *knot_k, const T *knot_a,             \
      const int32_t *site, const T *mw, const T *target, int64_t tstride,            \
      T mse_scale, T theta, T tau, int64_t step0, int64_t n_steps, T dt, const T *hh

#define DN_FIT_FILL(T)                                                                \
  dn::FitArgs<T> A = {};                                                              \
  if (!hh || !(tau > T(0))) return DN_EINVAL;                                         \
  A.n_tasks = n_tasks; A.task_nr = task_nr; A.task_p = task_p;                        \
  A.slots = (int64_t)n_tasks * 64 * waves * rcap;                                     \
  A.gl = gl; A.grp = grp; A.sc = scale; A.sna = sna; A.sk = sk;                       \
  A.act = act; A.n_knots = n_knots; A.knot_k = knot_k; A.knot_a = knot_a;             \
  A.site = site; A.mw = mw; A.target = target; A.tstride = tstride;                   \
  A.mse_scale = mse_scale; A.theta = theta; A.inv_tau = T(1) / tau;                   \
  A.step0 = step0; A.n_steps = n_steps; A.dt = dt;                                    \
  A.hh = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};

#define DN_FIT_FWD(T, NAME)                                                           \
  DN_EXPORT int NAME(DN_FIT_COMMON(T), const T *x0, T *xout, T *ckpt,                \
                     int64_t ckpt_every, T *traj, T *loss, hipStream_t stream) {      \
    DN_FIT_FILL(T)                                                                    \
    A.x0 = x0; A.xout = xout; A.ckpt = ckpt; A.ckpt_every = ckpt_every;               \
    A.traj = traj; A.loss = loss;                                                     \
    return dn::fit_launch<T>(A, rcap, waves, false, stream);                                 \
  }

#define DN_FIT_ADJ(T, NAME)                                                           \
  DN_EXPORT int NAME(DN_FIT_COMMON(T), const T *traj, T *adj, T *g_sna, T *g_sk,     \
                     T *g_act, hipStream_t stream) {                                  \
    DN_FIT_FILL(T)                                                                    \
    A.traj = const_cast<T *>(traj); A.adj = adj;                                      \
    A.g_sna = g_sna; A.g_sk = g_sk; A.g_act = g_act;                                  \
    return dn::fit_launch<T>(A, rcap, waves, true, stream);                                  \
  }

DN_FIT_FWD(float, dn_fit_forward_f32)
DN_FIT_FWD(double, dn_fit_forward_f64)
DN_FIT_ADJ(float, dn_fit_adjoint_f32)
DN_FIT_ADJ(double, dn_fit_adjoint_f64)

DN_EXPORT int dn_fit_csr_knots_f32(int64_t n_out, int n_k, const int64_t* ptr,
                                   const int32_t* idx, const float* w, const float* in,
                                   int64_t in_row, int64_t in_k, float* out,
                                   int64_t out_row, int64_t out_k, hipStream_t stream) {
  return dn::csr_knots_launch<float>(n_out, n_k, ptr, idx, w, in, in_row, in_k, out,
                                     out_row, out_k, stream);
}
DN_EXPORT int dn_fit_csr_knots_f64(int64_t n_out, int n_k, const int64_t* ptr,
                                   const int32_t* idx, const double* w, const double* in,
                                   int64_t in_row, int64_t in_k, double* out,
                                   int64_t out_row, int64_t out_k, hipStream_t stream) {
  return dn::csr_knots_launch<double>(n_out, n_k, ptr, idx, w, in, in_row, in_k, out,
                                      out_row, out_k, stream);
}
