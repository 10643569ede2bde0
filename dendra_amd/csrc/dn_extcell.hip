// dn_extcell.hip -- HH cables with two extracellular layers (ExtCell), fused
// backward-Euler steps.
//
// Replaces _bwd_euler_bt._step (dendra/models/integrators/implicit.py:
// 1021-1136) for HH cables: per step and compartment k (unknowns
// vc = (vi, ve0, ve1)), with v = vi - ve0,
//   gates by cnexp at v -> G, J (mod/HH.py:109-116; the cable kernels' code)
//   g = G area, d = (J - G v) area + intra
//   A_k = maind_k + [[g, -g, 0], [-g, g, 0], [0, 0, 0]]
//   R_k = (d, -d - xg0 (ve0 - ve1), xg0 (ve0 - ve1) + xg1 (e_bath - ve1))
//         + f_k - f_{k-1},  f_k = -lower_k * (vc_{k+1} - vc_k)
//   (lower at (k, k-1), upper at (k-1, k): diagonal 3x3 couplings)
//   block Thomas for the increment, vc += increment.
// maind, lower, upper, area, xg are the prepared workspace of
// _prepare_workspace (implicit.py:800-872).
//
// One lane owns one cell for all steps (no synchronisation between lanes),
// or, for two or more compartments, two adjacent lanes share it and
// eliminate from both ends (extcell_pair_kernel below).
// Every array is laid out [compartment][component][cell], so the accesses of
// a wavefront are coalesced; the forward sweep keeps the previous block's
// inverse in registers, prefetches the next compartment's inputs, and stores
// (A'^-1, R') for the backward sweep in a [K][12][B] workspace.
#include "dn_common.hpp"

namespace dn {

template <typename T>
__device__ __forceinline__ void xc_inv3(const T* a, T* r) {
  const T c0 = a[4] * a[8] - a[5] * a[7];
  const T c1 = a[5] * a[6] - a[3] * a[8];
  const T c2 = a[3] * a[7] - a[4] * a[6];
  const T id = T(1) / (a[0] * c0 + a[1] * c1 + a[2] * c2);
  r[0] = c0 * id;
  r[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  r[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  r[3] = c1 * id;
  r[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  r[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  r[6] = c2 * id;
  r[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  r[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

template <typename T>
struct XcIn {
  T mai[9], lo[3], up[3], area, xg0, xg1, m, h, n, vn[3], lead, iamp, sna, sk;
};

template <typename T>
struct XcArgs {
  const T *maind, *lower, *upper, *area, *xg, *sna, *sk, *lead, *amp, *wave, *iamp, *iwave;
  const T* hh;
  T *vc, *m, *h, *n, *work, *rec;
  int64_t B, K, n_steps, rec_every;
  T dt;
};

template <typename T>
__device__ __forceinline__ void xc_load(const XcArgs<T>& a, int64_t k, int64_t b, XcIn<T>& in) {
  const int64_t B = a.B, K = a.K;
#pragma unroll
  for (int c = 0; c < 9; ++c) in.mai[c] = a.maind[(k * 9 + c) * B + b];
  const bool edge = k + 1 < K;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    in.lo[c] = edge ? a.lower[(k * 3 + c) * B + b] : T(0);
    in.up[c] = edge ? a.upper[(k * 3 + c) * B + b] : T(0);
    in.vn[c] = edge ? a.vc[((k + 1) * 3 + c) * B + b] : T(0);
  }
  in.area = a.area[k * B + b];
  in.xg0 = a.xg[(k * 2) * B + b];
  in.xg1 = a.xg[(k * 2 + 1) * B + b];
  in.m = a.m[k * B + b];
  in.h = a.h[k * B + b];
  in.n = a.n[k * B + b];
  in.lead = a.lead ? a.lead[k * B + b] : T(0);
  in.iamp = a.iamp ? a.iamp[k * B + b] : T(0);
  in.sna = a.sna ? a.sna[k * B + b] : T(1);
  in.sk = a.sk ? a.sk[k * B + b] : T(1);
}

template <typename T>
__global__ __launch_bounds__(256) void extcell_run_kernel(XcArgs<T> a) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.B) return;
  const int64_t B = a.B, K = a.K;
  const HHParams<T> hp{a.hh[0], a.hh[1], a.hh[2], a.hh[3], a.hh[4], a.hh[5], a.hh[6]};
  const T amp = (a.amp && a.wave && a.lead) ? a.amp[b] : T(0);
  T* W = a.work;
  for (int64_t s = 0; s < a.n_steps; ++s) {
    const T bath_s = (a.amp && a.wave && a.lead) ? a.wave[s] * amp : T(0);
    const T intra_s = (a.iamp && a.iwave) ? a.iwave[s] : T(0);
    T vk[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) vk[c] = a.vc[c * B + b];
    T Ainv[9], Rp[3], fprev[3] = {T(0), T(0), T(0)}, lprev[3], uprev[3];
    XcIn<T> cur, nxt;
    xc_load(a, 0, b, cur);
    for (int64_t k = 0; k < K; ++k) {
      if (k + 1 < K) xc_load(a, k + 1, b, nxt);
      // membrane: gates at v, then G and J
      const T v = vk[0] - vk[1];
      T G, J;
      hh_advance<T>(v, a.dt, hp, cur.sna, cur.sk, cur.m, cur.h, cur.n, G, J);
      a.m[k * B + b] = cur.m;
      a.h[k * B + b] = cur.h;
      a.n[k * B + b] = cur.n;
      const T g = G * cur.area;
      const T d = (J - G * v) * cur.area + cur.iamp * intra_s;
      T A[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) A[c] = cur.mai[c];
      A[0] += g;
      A[1] -= g;
      A[3] -= g;
      A[4] += g;
      const T shell = cur.xg0 * (vk[1] - vk[2]);
      const T bathf = cur.xg1 * (cur.lead * bath_s - vk[2]);
      T R[3] = {d, -d - shell, shell + bathf};
      T f[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        f[c] = -cur.lo[c] * (cur.vn[c] - vk[c]);
        R[c] += f[c] - fprev[c];
      }
      if (k > 0) {
        // A -= diag(l) Ainv_prev diag(u);  R -= diag(l) Ainv_prev Rp
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
          for (int c = 0; c < 3; ++c) A[r * 3 + c] -= lprev[r] * Ainv[r * 3 + c] * uprev[c];
          R[r] -= lprev[r] * (Ainv[r * 3] * Rp[0] + Ainv[r * 3 + 1] * Rp[1] +
                              Ainv[r * 3 + 2] * Rp[2]);
        }
      }
      xc_inv3(A, Ainv);
#pragma unroll
      for (int c = 0; c < 9; ++c) W[(k * 12 + c) * B + b] = Ainv[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        W[(k * 12 + 9 + c) * B + b] = R[c];
        Rp[c] = R[c];
        fprev[c] = f[c];
        lprev[c] = cur.lo[c];
        uprev[c] = cur.up[c];
        vk[c] = cur.vn[c];
      }
      cur = nxt;
    }
    // back-substitution: x_k = Ainv_k (R'_k - diag(u_k) x_{k+1}); vc += x
    T xn[3] = {T(0), T(0), T(0)};
    for (int64_t k = K - 1; k >= 0; --k) {
      T Ai[9], r[3], x[3];
#pragma unroll
      for (int c = 0; c < 9; ++c) Ai[c] = W[(k * 12 + c) * B + b];
      const bool edge = k + 1 < K;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        r[c] = W[(k * 12 + 9 + c) * B + b] - (edge ? a.upper[(k * 3 + c) * B + b] * xn[c] : T(0));
#pragma unroll
      for (int c = 0; c < 3; ++c) x[c] = Ai[c * 3] * r[0] + Ai[c * 3 + 1] * r[1] + Ai[c * 3 + 2] * r[2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        a.vc[(k * 3 + c) * B + b] += x[c];
        xn[c] = x[c];
      }
    }
    if (a.rec && a.rec_every > 0 && (s + 1) % a.rec_every == 0) {
      const int64_t rr = (s + 1) / a.rec_every - 1;
      for (int64_t k = 0; k < K; ++k)
        a.rec[(rr * K + k) * B + b] = a.vc[(k * 3) * B + b] - a.vc[(k * 3 + 1) * B + b];
    }
  }
}

// Two lanes per cell (adjacent lanes 2b, 2b + 1): elimination from both
// ends.  The "down" lane eliminates compartments 0 .. s-1 top-down as above,
// the "up" lane compartments K-1 .. s bottom-up,
//   A''_k = A_k - diag(u_k) A''^-1_{k+1} diag(l_k),
//   R''_k = R_k - diag(u_k) A''^-1_{k+1} R''_{k+1}.
// At the interface, with y_d = A'^-1_{s-1} R'_{s-1}, C_d = A'^-1_{s-1} diag(u),
// y_u = A''^-1_s R''_s, C_u = A''^-1_s diag(l) (u, l the couplings of edge
// (s-1, s)), both lanes solve (I - C_d C_u) x_{s-1} = y_d - C_d y_u and
// x_s = y_u - C_u x_{s-1} from the exchanged values, then back-substitute
// outward.  Halves the sequential chain and doubles the lanes.
template <typename T>
__device__ __forceinline__ void xc_assemble(const XcArgs<T>& a, int64_t k, int64_t b,
                                            const HHParams<T>& hp, T bath_s, T intra_s, T* A,
                                            T* R) {
  const int64_t B = a.B, K = a.K;
  T vk[3], vp[3], vn[3], lp[3], ln[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    vk[c] = a.vc[(k * 3 + c) * B + b];
    vp[c] = k > 0 ? a.vc[((k - 1) * 3 + c) * B + b] : vk[c];
    vn[c] = k + 1 < K ? a.vc[((k + 1) * 3 + c) * B + b] : vk[c];
    lp[c] = k > 0 ? a.lower[((k - 1) * 3 + c) * B + b] : T(0);
    ln[c] = k + 1 < K ? a.lower[(k * 3 + c) * B + b] : T(0);
  }
  const int64_t i = k * B + b;
  const T v = vk[0] - vk[1];
  T m = a.m[i], h = a.h[i], n = a.n[i], G, J;
  hh_advance<T>(v, a.dt, hp, a.sna ? a.sna[i] : T(1), a.sk ? a.sk[i] : T(1), m, h, n, G, J);
  a.m[i] = m;
  a.h[i] = h;
  a.n[i] = n;
  const T area = a.area[i];
  const T g = G * area;
  const T d = (J - G * v) * area + (a.iamp ? a.iamp[i] : T(0)) * intra_s;
#pragma unroll
  for (int c = 0; c < 9; ++c) A[c] = a.maind[(k * 9 + c) * B + b];
  A[0] += g;
  A[1] -= g;
  A[3] -= g;
  A[4] += g;
  const T xg0 = a.xg[(k * 2) * B + b], xg1 = a.xg[(k * 2 + 1) * B + b];
  const T lead = a.lead ? a.lead[i] : T(0);
  const T shell = xg0 * (vk[1] - vk[2]);
  const T bathf = xg1 * (lead * bath_s - vk[2]);
  R[0] = d;
  R[1] = -d - shell;
  R[2] = shell + bathf;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T f = -ln[c] * (vn[c] - vk[c]);
    const T fprev = -lp[c] * (vk[c] - vp[c]);
    R[c] += f - fprev;
  }
}

// Inputs of one compartment's step in the two-lane form: voltages of the
// compartment and its neighbours, the couplings, gates and coefficients.
// Loaded one compartment ahead (every index clamped into the cable, so all
// loads are unconditional and in flight together) while the current one is
// assembled and eliminated.
template <typename T>
struct XpIn {
  T vk[3], vp[3], vn[3], lp[3], ln[3], mai[9], cr[3], cn[3];
  T m, h, n, sna, sk, area, iamp, xg0, xg1, lead;
};

template <typename T>
__device__ __forceinline__ void xp_load(const XcArgs<T>& a, int64_t k, int64_t b, bool up,
                                        XpIn<T>& in) {
  const int64_t B = a.B, K = a.K;  // K >= 2
  const int64_t kp = k > 0 ? k - 1 : 0, kn = k + 1 < K ? k + 1 : K - 1;
  const int64_t en = k + 1 < K ? k : K - 2;
  // coupling applied from the left at step k (t > 0): down diag(l_{k-1}),
  // up diag(u_k); the one the next block multiplies from the right: down
  // diag(u_k), up diag(l_{k-1}) (absent at the cable's ends)
  const T* left = up ? a.upper : a.lower;
  const T* right = up ? a.lower : a.upper;
  const int64_t el = up ? (k < K - 1 ? k : K - 2) : kp;
  const int64_t er = up ? k - 1 : k;
  const bool has_r = er >= 0 && er + 1 < K;
  const int64_t erc = er < 0 ? 0 : (er > K - 2 ? K - 2 : er);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    in.vk[c] = a.vc[(k * 3 + c) * B + b];
    in.vp[c] = a.vc[(kp * 3 + c) * B + b];
    in.vn[c] = a.vc[(kn * 3 + c) * B + b];
    in.lp[c] = keep_if(a.lower[(kp * 3 + c) * B + b], k > 0);
    in.ln[c] = keep_if(a.lower[(en * 3 + c) * B + b], k + 1 < K);
    in.cr[c] = left[(el * 3 + c) * B + b];
    in.cn[c] = keep_if(right[(erc * 3 + c) * B + b], has_r);
  }
  const int64_t i = k * B + b;
#pragma unroll
  for (int c = 0; c < 9; ++c) in.mai[c] = a.maind[(k * 9 + c) * B + b];
  in.m = a.m[i];
  in.h = a.h[i];
  in.n = a.n[i];
  in.area = a.area[i];
  in.xg0 = a.xg[(k * 2) * B + b];
  in.xg1 = a.xg[(k * 2 + 1) * B + b];
  in.sna = a.sna ? a.sna[i] : T(1);
  in.sk = a.sk ? a.sk[i] : T(1);
  in.iamp = a.iamp ? a.iamp[i] : T(0);
  in.lead = a.lead ? a.lead[i] : T(0);
}

// The block row of compartment k (membrane step, then the assembled 3x3
// block and right-hand side), as xc_assemble but from prefetched inputs.
template <typename T>
__device__ __forceinline__ void xp_assemble(const XcArgs<T>& a, int64_t k, int64_t b,
                                            const HHParams<T>& hp, T bath_s, T intra_s,
                                            XpIn<T>& in, T* A, T* R) {
  const int64_t i = k * a.B + b;
  const T v = in.vk[0] - in.vk[1];
  T G, J;
  hh_advance<T>(v, a.dt, hp, in.sna, in.sk, in.m, in.h, in.n, G, J);
  a.m[i] = in.m;
  a.h[i] = in.h;
  a.n[i] = in.n;
  const T g = G * in.area;
  const T d = (J - G * v) * in.area + in.iamp * intra_s;
#pragma unroll
  for (int c = 0; c < 9; ++c) A[c] = in.mai[c];
  A[0] += g;
  A[1] -= g;
  A[3] -= g;
  A[4] += g;
  const T shell = in.xg0 * (in.vk[1] - in.vk[2]);
  const T bathf = in.xg1 * (in.lead * bath_s - in.vk[2]);
  R[0] = d;
  R[1] = -d - shell;
  R[2] = shell + bathf;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T f = -in.ln[c] * (in.vn[c] - in.vk[c]);
    const T fprev = -in.lp[c] * (in.vk[c] - in.vp[c]);
    R[c] += f - fprev;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void extcell_pair_kernel(XcArgs<T> a) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t b = g >> 1;
  if (b >= a.B) return;  // both lanes of a pair leave together
  const bool up = g & 1;
  const int64_t B = a.B, K = a.K, sp = K / 2;
  const HHParams<T> hp{a.hh[0], a.hh[1], a.hh[2], a.hh[3], a.hh[4], a.hh[5], a.hh[6]};
  const T amp = (a.amp && a.wave && a.lead) ? a.amp[b] : T(0);
  T* W = a.work;
  for (int64_t s = 0; s < a.n_steps; ++s) {
    const T bath_s = (a.amp && a.wave && a.lead) ? a.wave[s] * amp : T(0);
    const T intra_s = (a.iamp && a.iwave) ? a.iwave[s] : T(0);
    T Ainv[9], Rp[3], cprev[3];  // previous block's inverse, rhs, coupling pair
    const int64_t n_own = up ? K - sp : sp;
    XpIn<T> cur, nxt;
    xp_load(a, up ? K - 1 : 0, b, up, cur);
    for (int64_t t = 0; t < n_own; ++t) {
      const int64_t k = up ? K - 1 - t : t;
      const int64_t tn = t + 1 < n_own ? t + 1 : t;
      xp_load(a, up ? K - 1 - tn : tn, b, up, nxt);
      T A[9], R[3];
      xp_assemble(a, k, b, hp, bath_s, intra_s, cur, A, R);
      if (t > 0) {
        // down: couplings (k, k-1) = diag(l_{k-1}), (k-1, k) = diag(u_{k-1});
        // up:   couplings (k, k+1) = diag(u_k),    (k+1, k) = diag(l_k)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
          for (int c = 0; c < 3; ++c) A[r * 3 + c] -= cur.cr[r] * Ainv[r * 3 + c] * cprev[c];
          R[r] -= cur.cr[r] * (Ainv[r * 3] * Rp[0] + Ainv[r * 3 + 1] * Rp[1] + Ainv[r * 3 + 2] * Rp[2]);
        }
      }
      xc_inv3(A, Ainv);
#pragma unroll
      for (int c = 0; c < 9; ++c) W[(k * 12 + c) * B + b] = Ainv[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        W[(k * 12 + 9 + c) * B + b] = R[c];
        Rp[c] = R[c];
        cprev[c] = cur.cn[c];
      }
      cur = nxt;
    }
    // interface: exchange (Ainv, R') of compartments s-1 (down) and s (up)
    T Ao[9], Ro[3];
#pragma unroll
    for (int c = 0; c < 9; ++c) Ao[c] = __shfl_xor(Ainv[c], 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) Ro[c] = __shfl_xor(Rp[c], 1);
    const T* Ad = up ? Ao : Ainv;
    const T* Rd = up ? Ro : Rp;
    const T* Au = up ? Ainv : Ao;
    const T* Ru = up ? Rp : Ro;
    T uu[3], ll[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uu[c] = a.upper[((sp - 1) * 3 + c) * B + b];
      ll[c] = a.lower[((sp - 1) * 3 + c) * B + b];
    }
    T yd[3], yu[3], Cd[9], Cu[9], M[9], Mi[9], rhs[3], xd[3], xu[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      yd[r] = Ad[r * 3] * Rd[0] + Ad[r * 3 + 1] * Rd[1] + Ad[r * 3 + 2] * Rd[2];
      yu[r] = Au[r * 3] * Ru[0] + Au[r * 3 + 1] * Ru[1] + Au[r * 3 + 2] * Ru[2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Cd[r * 3 + c] = Ad[r * 3 + c] * uu[c];
        Cu[r * 3 + c] = Au[r * 3 + c] * ll[c];
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
        M[r * 3 + c] = (r == c ? T(1) : T(0)) -
                       (Cd[r * 3] * Cu[c] + Cd[r * 3 + 1] * Cu[3 + c] + Cd[r * 3 + 2] * Cu[6 + c]);
      rhs[r] = yd[r] - (Cd[r * 3] * yu[0] + Cd[r * 3 + 1] * yu[1] + Cd[r * 3 + 2] * yu[2]);
    }
    xc_inv3(M, Mi);
#pragma unroll
    for (int r = 0; r < 3; ++r) xd[r] = Mi[r * 3] * rhs[0] + Mi[r * 3 + 1] * rhs[1] + Mi[r * 3 + 2] * rhs[2];
#pragma unroll
    for (int r = 0; r < 3; ++r) xu[r] = yu[r] - (Cu[r * 3] * xd[0] + Cu[r * 3 + 1] * xd[1] + Cu[r * 3 + 2] * xd[2]);
    // back-substitution outward from the interface
    T xn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) xn[c] = up ? xu[c] : xd[c];
    // the next (outer) compartment's inverse, eliminated right-hand side,
    // coupling and voltage are fetched a compartment ahead (clamped index)
    const T* outer = up ? a.lower : a.upper;
    T nAi[9], nR[3], nC[3], nV[3];
    auto fetch_back = [&](int64_t t) {
      const int64_t k = up ? K - 1 - t : t;
      const int64_t e = up ? (k > 0 ? k - 1 : 0) : (k < K - 1 ? k : K - 2);
#pragma unroll
      for (int c = 0; c < 9; ++c) nAi[c] = W[(k * 12 + c) * B + b];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        nR[c] = W[(k * 12 + 9 + c) * B + b];
        nC[c] = outer[(e * 3 + c) * B + b];
        nV[c] = a.vc[(k * 3 + c) * B + b];
      }
    };
    fetch_back(n_own - 1);
    for (int64_t t = n_own - 1; t >= 0; --t) {
      const int64_t k = up ? K - 1 - t : t;
      T Ai[9], r[3], x[3], vk[3];
#pragma unroll
      for (int c = 0; c < 9; ++c) Ai[c] = nAi[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        r[c] = nR[c] - nC[c] * xn[c];
        vk[c] = nV[c];
      }
      fetch_back(t > 0 ? t - 1 : 0);
      if (t == n_own - 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) x[c] = xn[c];
      } else {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          x[c] = Ai[c * 3] * r[0] + Ai[c * 3 + 1] * r[1] + Ai[c * 3 + 2] * r[2];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        a.vc[(k * 3 + c) * B + b] = vk[c] + x[c];
        xn[c] = x[c];
      }
    }
    if (a.rec && a.rec_every > 0 && (s + 1) % a.rec_every == 0) {
      const int64_t rr = (s + 1) / a.rec_every - 1;
      for (int64_t t = 0; t < n_own; ++t) {
        const int64_t k = up ? K - 1 - t : t;
        a.rec[(rr * K + k) * B + b] = a.vc[(k * 3) * B + b] - a.vc[(k * 3 + 1) * B + b];
      }
    }
  }
}

}  // namespace dn

template <typename T>
static int extcell_impl(const dn::XcArgs<T>& a, int64_t work_elems, hipStream_t stream) {
  if (a.B < 1 || a.K < 1 || a.n_steps < 0 || a.B > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  if (!a.maind || !a.area || !a.xg || !a.hh || !a.vc || !a.m || !a.h || !a.n || !a.work)
    return DN_EINVAL;
  if (a.K > 1 && (!a.lower || !a.upper)) return DN_EINVAL;
  if (work_elems < 12 * a.B * a.K) return DN_EINVAL;
  if (a.rec && a.rec_every <= 0) return DN_EINVAL;
  if (a.n_steps == 0) return DN_OK;
  if (a.K >= 2 && a.B <= (int64_t)INT32_MAX * 128)
    hipLaunchKernelGGL((dn::extcell_pair_kernel<T>), dim3((unsigned)((2 * a.B + 255) / 256)),
                       dim3(256), 0, stream, a);
  else
    hipLaunchKernelGGL((dn::extcell_run_kernel<T>), dim3((unsigned)((a.B + 255) / 256)),
                       dim3(256), 0, stream, a);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int64_t dn_extcell_work_elems(int64_t B, int64_t K) {
  return (B < 0 || K < 0) ? -1 : 12 * B * K;
}

#define DN_EXTCELL_EXPORT(T, SFX)                                                            \
  DN_EXPORT int dn_extcell_run_##SFX(                                                        \
      int64_t B, int64_t K, const T* maind, const T* lower, const T* upper, const T* area,   \
      const T* xg, const T* sna, const T* sk, const T* hh, const T* lead, const T* amp,      \
      const T* wave, const T* iamp, const T* i_wave, T dt, int64_t n_steps, T* vc, T* m,     \
      T* h, T* n, T* rec, int64_t rec_every, T* work, int64_t work_elems,                    \
      hipStream_t stream) {                                                                  \
    dn::XcArgs<T> a{maind, lower, upper, area, xg, sna, sk, lead, amp, wave, iamp, i_wave,  \
                    hh, vc, m, h, n, work, rec, B, K, n_steps, rec_every, dt};              \
    return extcell_impl<T>(a, work_elems, stream);                                           \
  }

DN_EXTCELL_EXPORT(float, f32)
DN_EXTCELL_EXPORT(double, f64)
