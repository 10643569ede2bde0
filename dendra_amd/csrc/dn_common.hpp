// dn_common.hpp -- shared device code for the gfx950 cable hot path.
//
// Hodgkin-Huxley rates, cnexp gate update and the exprelr ("vtrap") series,
// written for 64-lane CDNA4 wavefronts.  Float32 follows the reference's
// float32 branches (series degrees and switch points), float64 its float64
// branches:
//   rates               dendra/models/mod/HH.py:24-56
//   affine current      dendra/models/mod/HH.py:109-116
//   cnexp step          dendra/models/mechanisms/_cnexp.py:46-132
//   exprelr / vtrap     dendra/models/mechanisms/ops.py:13-128
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DN_OK 0
#define DN_EINVAL -1
#define DN_ELAUNCH -2
#define DN_ERANGE -3
#define DN_EUNSUPPORTED -4

#define DN_EXPORT extern "C" __attribute__((visibility("default")))

namespace dn {

// x where c holds, else +0, as a bit mask: the load of x stays unconditional
// (a select lets the compiler put the load under a branch of its own and
// wait for it there, one load at a time)
__device__ __forceinline__ float keep_if(float x, bool c) {
  return __uint_as_float(__float_as_uint(x) & (c ? ~0u : 0u));
}
__device__ __forceinline__ double keep_if(double x, bool c) {
  return __longlong_as_double(__double_as_longlong(x) & (c ? ~0ll : 0ll));
}

// Launch-contract status word (caller-owned device int32, nullable).  A
// persistent kernel that meets work outside what its launch promised sets
// one of these bits and leaves that work unadvanced.
enum : int32_t {
  kStatusCableRows = 1,   // a cable task has more rows than the kernel instance
  kStatusTreeGroup = 2,   // a tree group exceeds the block's threads or node capacity
  kStatusTreeSize = 4,    // a tree exceeds the launch's node capacity (kmax)
};
__device__ __forceinline__ void flag_status(int32_t* status, int32_t bit) {
  if (status) atomicOr(status, bit);
}

// ---------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float fexp(float x) { return __expf(x); }
__device__ __forceinline__ double fexp(double x) { return exp(x); }
// 1-ulp hardware reciprocal in float32; exact division in float64.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float fabs_(float x) { return fabsf(x); }
__device__ __forceinline__ double fabs_(double x) { return fabs(x); }

template <typename T> struct HHParams {
  T gnabar, gkbar, gl, ena, ek, el, q10;
};

// x / (exp(x/10) - 1) with the local even Bernoulli series near x = 0.
// The series is evaluated only when some lane of the wavefront needs it; the
// per-lane selection is unchanged, so results do not depend on the skip.
__device__ __forceinline__ float vtrap10(float x) {
  const float q = x * 0.1f;
  const bool near0 = fabsf(q) <= 1.0f;
  const float direct = x * frcp(fexp(q) - 1.0f);
  if (__ballot(near0) == 0ull) return direct;
  const float q2 = q * q;
  float e = fmaf(q2, -691.0f / 1307674368000.0f, 1.0f / 47900160.0f);
  e = fmaf(e, q2, -1.0f / 1209600.0f);
  e = fmaf(e, q2, 1.0f / 30240.0f);
  e = fmaf(e, q2, -1.0f / 720.0f);
  e = fmaf(e, q2, 1.0f / 12.0f);
  const float series = (10.0f - x * 0.5f) + x * (q * e);
  return near0 ? series : direct;
}
__device__ __forceinline__ double vtrap10(double x) {
  const double q = x / 10.0;
  const bool near0 = fabs(q) <= 0.5;
  const double direct = x / (exp(q) - 1.0);
  if (__ballot(near0) == 0ull) return direct;
  const double q2 = q * q;
  double e = q2 * (43867.0 / 5109094217170944000.0) - 3617.0 / 10670622842880000.0;
  e = e * q2 + 1.0 / 74724249600.0;
  e = e * q2 - 691.0 / 1307674368000.0;
  e = e * q2 + 1.0 / 47900160.0;
  e = e * q2 - 1.0 / 1209600.0;
  e = e * q2 + 1.0 / 30240.0;
  e = e * q2 - 1.0 / 720.0;
  e = e * q2 + 1.0 / 12.0;
  const double series = (10.0 - x / 2.0) + x * (q * e);
  return near0 ? series : direct;
}

// phi1(z) = expm1(z)/z on |z| <= 0.5 (Taylor, degree 9 / 16).
__device__ __forceinline__ float phi1(float z) {
  float v = fmaf(z, 1.0f / 3628800.0f, 1.0f / 362880.0f);
  v = fmaf(v, z, 1.0f / 40320.0f);
  v = fmaf(v, z, 1.0f / 5040.0f);
  v = fmaf(v, z, 1.0f / 720.0f);
  v = fmaf(v, z, 1.0f / 120.0f);
  v = fmaf(v, z, 1.0f / 24.0f);
  v = fmaf(v, z, 1.0f / 6.0f);
  v = fmaf(v, z, 0.5f);
  v = fmaf(v, z, 1.0f);
  return v;
}
__device__ __forceinline__ double phi1(double z) {
  double v = z * (1.0 / 355687428096000.0) + 1.0 / 20922789888000.0;
  v = v * z + 1.0 / 1307674368000.0;
  v = v * z + 1.0 / 87178291200.0;
  v = v * z + 1.0 / 6227020800.0;
  v = v * z + 1.0 / 479001600.0;
  v = v * z + 1.0 / 39916800.0;
  v = v * z + 1.0 / 3628800.0;
  v = v * z + 1.0 / 362880.0;
  v = v * z + 1.0 / 40320.0;
  v = v * z + 1.0 / 5040.0;
  v = v * z + 1.0 / 720.0;
  v = v * z + 1.0 / 120.0;
  v = v * z + 1.0 / 24.0;
  v = v * z + 1.0 / 6.0;
  v = v * z + 1.0 / 2.0;
  v = v * z + 1.0;
  return v;
}

// Exact affine step of x' = xa - xb x over dt (ms): the Taylor form for
// |z| <= 0.5, the exponential relaxation beyond.  The relaxation is
// evaluated only when some lane of the wavefront needs it (the per-lane
// selection is unchanged).
template <typename T>
__device__ __forceinline__ T cnexp(T x, T xa, T xb, T dt) {
  const T z = -xb * dt;
  const T inc = x + dt * (xa - xb * x) * phi1(z);
  const bool far = fabs_(z) > T(0.5);
  if (__ballot(far) == 0ull) return inc;
  const T eq = xa * frcp(xb);
  const T rel = eq + (x - eq) * fexp(z);
  return far ? rel : inc;
}

// Gate advance at v_old, then the affine current terms from the new gates.
// sna/sk scale gNa/gK per compartment (1 when the model has no scales).
template <typename T>
__device__ __forceinline__ void hh_advance(T v, T dt, const HHParams<T>& p, T sna,
                                           T sk, T& m, T& h, T& n, T& G, T& J) {
  T am = T(0.1) * vtrap10(-(v + T(40)));
  T bm = T(4) * fexp((v + T(65)) * T(-1.0 / 18.0));
  T ah = T(0.07) * fexp((v + T(65)) * T(-1.0 / 20.0));
  T bh = frcp(fexp((v + T(35)) * T(-0.1)) + T(1));
  T an = T(0.01) * vtrap10(-(v + T(55)));
  T bn = T(0.125) * fexp((v + T(65)) * T(-1.0 / 80.0));
  m = cnexp(m, p.q10 * am, p.q10 * (am + bm), dt);
  h = cnexp(h, p.q10 * ah, p.q10 * (ah + bh), dt);
  n = cnexp(n, p.q10 * an, p.q10 * (an + bn), dt);
  T gna = p.gnabar * sna * (m * m * m) * h;
  T n2 = n * n;
  T gk = p.gkbar * sk * (n2 * n2);
  G = p.gl + gna + gk;
  J = p.gl * p.el + gna * p.ena + gk * p.ek;
}

// float32 gate step of x' = q10 (a - s x) with s = a + b, given dq = q10 dt:
// x + dq (a - s x) phi1(z), z = -s dq.  phi1 is the Taylor form for
// |z| <= 0.5 -- cut after z^4 when every lane of the wavefront has
// |z| <= 1/16 (the dropped tail is below |z|^5/720 <= 1.3e-9, under half an
// ulp of phi1 ~ 1) -- and (e^z - 1)/z beyond, which makes the step the
// exponential relaxation a/s + (x - a/s) e^z written without a/s.
// The quotient is evaluated only when some lane of the wavefront has
// |z| > 0.5, which implies the full series branch; the per-lane selection is
// unchanged, so results do not depend on either skip.
__device__ __forceinline__ float gate_step(float x, float a, float s, float dq) {
  const float z = -s * dq;
  float ph;
  if (__ballot(fabsf(z) > 0.0625f) == 0ull) {
    ph = fmaf(z, 1.0f / 120.0f, 1.0f / 24.0f);
    ph = fmaf(ph, z, 1.0f / 6.0f);
    ph = fmaf(ph, z, 0.5f);
    ph = fmaf(ph, z, 1.0f);
  } else {
    ph = phi1(z);
    if (__ballot(fabsf(z) > 0.5f) != 0ull) {
      const float pf = (fexp(z) - 1.0f) * frcp(z);
      ph = fabsf(z) > 0.5f ? pf : ph;
    }
  }
  return fmaf(dq * fmaf(-s, x, a), ph, x);
}

// B(q) = q / (e^q - 1) given e = e^q: the even Bernoulli series for
// |q| <= 1 (evaluated only when some lane needs it), the quotient beyond.
// 0.1 vtrap10(x) = B(x/10).
__device__ __forceinline__ float bern(float q, float e) {
  const bool near0 = fabsf(q) <= 1.0f;
  const float direct = q * frcp(e - 1.0f);
  if (__ballot(near0) == 0ull) return direct;
  const float q2 = q * q;
  float c = fmaf(q2, -691.0f / 1307674368000.0f, 1.0f / 47900160.0f);
  c = fmaf(c, q2, -1.0f / 1209600.0f);
  c = fmaf(c, q2, 1.0f / 30240.0f);
  c = fmaf(c, q2, -1.0f / 720.0f);
  c = fmaf(c, q2, 1.0f / 12.0f);
  const float series = fmaf(q2, c, fmaf(q, -0.5f, 1.0f));
  return near0 ? series : direct;
}

// float32: the six rate exponentials reduce to three, on centred arguments
// (v + 40, v + 55 and v + 65 are exact in float32 near rest):
//   e40 = exp(-(v+40)/10),  exp(-(v+55)/10) = e40 e^-1.5,
//   exp(-(v+35)/10) = e40 e^0.5,
//   e80 = exp(-(v+65)/80),  bn = e80/8,  ah = 0.07 e80^4,
//   bm = 4 exp(-(v+65)/18),
// and q10 enters once, through dq = q10 dt.  Each identity is exact in real
// arithmetic; the float32 results move by a few ulp, with no systematic
// bias at rest (the argument constants stay the reference's 1/10, 1/18,
// 1/80 before the product with log2 e inside fexp).
template <>
__device__ __forceinline__ void hh_advance<float>(float v, float dt, const HHParams<float>& p,
                                                  float sna, float sk, float& m, float& h,
                                                  float& n, float& G, float& J) {
  const float xm = -(v + 40.0f), xn = -(v + 55.0f), w = v + 65.0f;
  const float qm = xm * 0.1f, qn = xn * 0.1f;
  // Exponent arguments are capped (e^80 = 5.5e34; e80 at e^20, so that
  // ah = 0.07 e80^4 stays below e^80): below about -1660 mV bm, and below
  // about -1840 mV ah, would overflow binary32 and turn the gate update into
  // inf - inf or inf / inf (NaN).  A rate that large takes its gate to the
  // equilibrium within one step with or without the cap (e^{-rate dt} is 0),
  // so the cap changes no result above those voltages and keeps the state
  // finite below them, as binary64 stays.
  const float e40 = fexp(fminf(qm, 80.0f));
  const float e80 = fexp(fminf(w * (-1.0f / 80.0f), 20.0f));
  const float bm = 4.0f * fexp(fminf(w * (-1.0f / 18.0f), 80.0f));
  const float e80s = e80 * e80;
  const float ah = 0.07f * (e80s * e80s);
  const float bn = 0.125f * e80;
  const float bh = frcp(fmaf(e40, 1.6487212707001282f, 1.0f));
  const float am = bern(qm, e40);
  const float an = 0.1f * bern(qn, e40 * 0.22313016014842982f);
  const float dq = p.q10 * dt;
  m = gate_step(m, am, am + bm, dq);
  h = gate_step(h, ah, ah + bh, dq);
  n = gate_step(n, an, an + bn, dq);
  float gna = p.gnabar * sna * (m * m * m) * h;
  float n2 = n * n;
  float gk = p.gkbar * sk * (n2 * n2);
  G = p.gl + gna + gk;
  J = p.gl * p.el + gna * p.ena + gk * p.ek;
}

// ---------------------------------------------------------------------------
// cross-lane moves inside a segment of W lanes (W a power of two <= 64).
// Out-of-segment sources return the caller's own value, which the
// eliminations below rely on (the corresponding couplings are exactly 0).
// Where one instruction does the move it is used: DPP row shifts inside
// 16-lane rows (W <= 16, with a select at the segment edge for W < 16), the
// DPP whole-wavefront shift by one lane (W = 32 with an edge select, W = 64),
// and the half swaps v_permlane16_swap (stride 16 at W = 32) and
// v_permlane32_swap (stride 32 at W = 64).  Lanes without a DPP source keep
// the old operand, which is the caller's own value.  The other strides go
// through ds_bpermute.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int bperm_from(int src_lane, int x) {
  return __builtin_amdgcn_ds_bpermute(src_lane << 2, x);
}
template <int W>
__device__ __forceinline__ int seg_up_i(int x, int s) {
  const int lane = __lane_id();
  const int seg_lane = lane & (W - 1);
  if constexpr (W <= 16) {
    int y;
    if (s == 1) y = __builtin_amdgcn_update_dpp(x, x, 0x111, 0xF, 0xF, false);
    else if (s == 2) y = __builtin_amdgcn_update_dpp(x, x, 0x112, 0xF, 0xF, false);
    else if (s == 4) y = __builtin_amdgcn_update_dpp(x, x, 0x114, 0xF, 0xF, false);
    else if (s == 8) y = __builtin_amdgcn_update_dpp(x, x, 0x118, 0xF, 0xF, false);
    else return bperm_from(seg_lane >= s ? lane - s : lane, x);
    if constexpr (W < 16) y = seg_lane >= s ? y : x;
    return y;
  } else {
    if (s == 1) {
      const int y = __builtin_amdgcn_update_dpp(x, x, 0x138, 0xF, 0xF, false);
      if constexpr (W == 32) return seg_lane != 0 ? y : x;
      return y;
    }
    if constexpr (W == 32)
      if (s == 16) return __builtin_amdgcn_permlane16_swap(x, x, false, false)[0];
    if constexpr (W == 64)
      if (s == 32) return __builtin_amdgcn_permlane32_swap(x, x, false, false)[0];
    return bperm_from(seg_lane >= s ? lane - s : lane, x);
  }
}
template <int W>
__device__ __forceinline__ int seg_down_i(int x, int s) {
  const int lane = __lane_id();
  const int seg_lane = lane & (W - 1);
  if constexpr (W <= 16) {
    int y;
    if (s == 1) y = __builtin_amdgcn_update_dpp(x, x, 0x101, 0xF, 0xF, false);
    else if (s == 2) y = __builtin_amdgcn_update_dpp(x, x, 0x102, 0xF, 0xF, false);
    else if (s == 4) y = __builtin_amdgcn_update_dpp(x, x, 0x104, 0xF, 0xF, false);
    else if (s == 8) y = __builtin_amdgcn_update_dpp(x, x, 0x108, 0xF, 0xF, false);
    else return bperm_from(seg_lane + s < W ? lane + s : lane, x);
    if constexpr (W < 16) y = seg_lane + s < W ? y : x;
    return y;
  } else {
    if (s == 1) {
      const int y = __builtin_amdgcn_update_dpp(x, x, 0x130, 0xF, 0xF, false);
      if constexpr (W == 32) return seg_lane != W - 1 ? y : x;
      return y;
    }
    if constexpr (W == 32)
      if (s == 16) return __builtin_amdgcn_permlane16_swap(x, x, false, false)[1];
    if constexpr (W == 64)
      if (s == 32) return __builtin_amdgcn_permlane32_swap(x, x, false, false)[1];
    return bperm_from(seg_lane + s < W ? lane + s : lane, x);
  }
}
template <int W>
__device__ __forceinline__ float shfl_up_seg(float x, int s) {
  return __int_as_float(seg_up_i<W>(__float_as_int(x), s));
}
template <int W>
__device__ __forceinline__ float shfl_down_seg(float x, int s) {
  return __int_as_float(seg_down_i<W>(__float_as_int(x), s));
}
template <int W>
__device__ __forceinline__ double shfl_up_seg(double x, int s) {
  int2 u = *reinterpret_cast<int2*>(&x);
  u.x = seg_up_i<W>(u.x, s);
  u.y = seg_up_i<W>(u.y, s);
  return *reinterpret_cast<double*>(&u);
}
template <int W>
__device__ __forceinline__ double shfl_down_seg(double x, int s) {
  int2 u = *reinterpret_cast<int2*>(&x);
  u.x = seg_down_i<W>(u.x, s);
  u.y = seg_down_i<W>(u.y, s);
  return *reinterpret_cast<double*>(&u);
}

// ---------------------------------------------------------------------------
// Parallel cyclic reduction of one tridiagonal row per lane across a
// W-lane segment:  A x_{l-1} + B x_l + C x_{l+1} = D.  Rows at the segment
// ends must carry A = 0 (first lane) and C = 0 (last lane).  Returns x_l.
// ---------------------------------------------------------------------------
template <typename T, int W>
__device__ __forceinline__ T pcr_lanes(T A, T B, T C, T D) {
#pragma unroll
  for (int s = 1; s < W; s <<= 1) {
    T Al = shfl_up_seg<W>(A, s), Bl = shfl_up_seg<W>(B, s);
    T Cl = shfl_up_seg<W>(C, s), Dl = shfl_up_seg<W>(D, s);
    T Ar = shfl_down_seg<W>(A, s), Br = shfl_down_seg<W>(B, s);
    T Cr = shfl_down_seg<W>(C, s), Dr = shfl_down_seg<W>(D, s);
    T al = -A * frcp(Bl);
    T be = -C * frcp(Br);
    T nA = al * Al;
    T nB = B + al * Cl + be * Ar;
    T nC = be * Cr;
    T nD = D + al * Dl + be * Dr;
    A = nA; B = nB; C = nC; D = nD;
  }
  return D * frcp(B);
}

// Zeroes a record before its fields are set, so its padding is defined.
template <typename R>
__device__ __forceinline__ void zero_record(R& r) {
  __builtin_memset(&r, 0, sizeof(R));
}

// Stores a record with every byte of its slot, padding included, in 16-byte
// (or 8-byte) vector stores: a scattered store that leaves bytes of a cache
// sector unwritten makes the memory side merge it with the old contents.
template <typename R>
__device__ __forceinline__ void store_whole(R* dst, const R& r) {
  static_assert(sizeof(R) % 8 == 0, "records are whole 8-byte words");
  if constexpr (sizeof(R) % 16 == 0) {
    uint4 q[sizeof(R) / 16];
    __builtin_memcpy(q, &r, sizeof(R));
#pragma unroll
    for (unsigned i = 0; i < sizeof(R) / 16; ++i) reinterpret_cast<uint4*>(dst)[i] = q[i];
  } else {
    uint2 q[sizeof(R) / 8];
    __builtin_memcpy(q, &r, sizeof(R));
#pragma unroll
    for (unsigned i = 0; i < sizeof(R) / 8; ++i) reinterpret_cast<uint2*>(dst)[i] = q[i];
  }
}

// pos[perm[p]] = p for p < n: the inverse of a permutation of [0, n), such
// as each connection's position in a target-major CSR (perm = in_edge).
// Entries of perm outside [0, n) are skipped.
template <int = 0>
__global__ __launch_bounds__(256) void inverse_permutation_kernel(
    int32_t* __restrict__ pos, const int32_t* __restrict__ perm, int64_t n) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t e = perm[p];
  if ((uint32_t)e < (uint64_t)n) pos[e] = (int32_t)p;
}

}  // namespace dn
