// dn_tridiag.hip -- batched tridiagonal solves for gfx950.
//
// Drop-in for the reference's thomas_solve_cuda_t / pcr_solve_cuda_t
// (dendra/models/integrators/triton/t_kernel_thomas.py:83-130,218-232 and
// pcr_kernel_thomas.py:99-195,224-226): a (B,K-1) sub-diagonal, b (B,K)
// diagonal, c (B,K-1) super-diagonal, d (B,K) right-hand side, x (B,K)
// output.  Inputs are never modified.  The transposed (adjoint) solve of
// ThomasSolve.backward (t_kernel_thomas.py:147-189) is the same call with
// a and c exchanged.
//
// Kernels:
//  * thomas_lds: one lane per system for short systems (K <= 32 in float32,
//    16 in float64), coefficients staged in LDS by coalesced loads, the
//    sweep in registers.  Best for many short systems; no workspace.
//  * thomas_lane: one lane per system for longer systems; its elimination
//    workspace lives in x (dp) and in a caller scratch of B*K values (cp).
//  * wave_hybrid: one 64-lane wavefront per system.  Each lane eliminates a
//    contiguous chunk of nr rows (local Thomas with a spike column), the 64
//    two-row interfaces are reduced to one row per lane by a cyclic-reduction
//    step and solved by wavefront PCR through ds_bpermute, then every lane
//    back-substitutes its chunk.  The system is staged in LDS (coalesced loads)
//    when 4*64*nr values fit a CU's 160 KiB, else it runs in the caller's
//    global scratch.
#include "dn_wave.hpp"

#include <mutex>

namespace dn {

// LDS of one CU (gfx950): a system of up to 10,240 float32 (5,120 float64)
// rows stages whole in the workgroup's LDS
constexpr size_t kTridiagLdsMax = 160 * 1024;

template <typename T>
__global__ __launch_bounds__(64) void tridiag_wave_kernel(
    const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
    const T* __restrict__ d, T* __restrict__ x, int64_t K, int nr, int use_lds,
    T* __restrict__ scratch) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int64_t r = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t KP = 64 * (int64_t)nr;
  T* base = use_lds ? reinterpret_cast<T*>(dn_smem_raw) : scratch + r * 4 * KP;
  T *sa = base, *sb = base + KP, *sc = base + 2 * KP, *sd = base + 3 * KP;
  const T* ar = a + r * (K - 1);
  const T* br = b + r * K;
  const T* cr = c + r * (K - 1);
  const T* dr = d + r * K;
  // eight rows per lane and pass: 32 loads in flight before their stores
  constexpr int U = 8;
  for (int64_t j0 = lane; j0 < KP; j0 += 64 * U) {
    T va[U], vb[U], vc[U], vd[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t j = j0 + 64 * u;
      const bool in = j < K;
      va[u] = (in && j >= 1) ? ar[j - 1] : T(0);
      vb[u] = in ? br[j] : T(1);
      vc[u] = (j + 1 < K) ? cr[j] : T(0);
      vd[u] = in ? dr[j] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t j = j0 + 64 * u;
      if (j < KP) {
        sa[j] = va[u];
        sb[j] = vb[u];
        sc[j] = vc[u];
        sd[j] = vd[u];
      }
    }
  }
  __syncthreads();
  wave_hybrid_solve<T>(sa, sb, sc, sd, nr);
  __syncthreads();
  T* xr = x + r * K;
  for (int64_t j = lane; j < K; j += 64) xr[j] = sd[j];
}

// Lane-per-system Thomas for short systems (K <= KMAX) with the
// coefficients staged in LDS: the 256 systems of a workgroup are contiguous
// in every array, so each array is read into LDS with consecutive lanes on
// consecutive elements (coalesced), each lane takes its system's row from
// LDS (odd row pitch: consecutive lanes land on different banks) into
// registers, sweeps it there (the sweep of tridiag/pcr.py:125-183: forward
// elimination with c' and d' held in registers, then back-substitution),
// and the solution leaves through LDS by coalesced stores.  No workspace.
template <typename T, int KMAX>
__global__ __launch_bounds__(256) void tridiag_thomas_lds_kernel(
    const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
    const T* __restrict__ d, T* __restrict__ x, int64_t B, int K) {
  __shared__ T st[256 * (KMAX + 1)];
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 256;
  const int ns = (int)(B - r0 < 256 ? B - r0 : 256);  // systems in this workgroup
  const int P = K | 1;                                // odd row pitch >= K
  // coalesced copy of n values per system (first system r0) into LDS rows:
  // element e = tid + 256 q is value i of system s, stepped without
  // division; eight loads are in flight before their LDS stores
  auto stage = [&](const T* src, int n) {
    const int tot = ns * n;
    const T* g = src + r0 * n;
    const int ds = 256 / n, di = 256 % n;
    int sy = tid / n, i = tid % n;
    constexpr int U = 8;
    for (int e0 = tid; e0 < tot; e0 += 256 * U) {
      T v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + 256 * u;
        v[u] = e < tot ? g[e] : T(0);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e0 + 256 * u < tot) st[sy * P + i] = v[u];
        sy += ds;
        i += di;
        if (i >= n) {
          i -= n;
          ++sy;
        }
      }
    }
  };
  T ra[KMAX], rb[KMAX], rc[KMAX], rd[KMAX];
  const T* row = st + tid * P;
  stage(b, K);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < KMAX; ++i) rb[i] = (tid < ns && i < K) ? row[i] : T(1);
  __syncthreads();
  if (K > 1) {
    stage(a, K - 1);  // a[i-1] couples row i to row i-1
    __syncthreads();
#pragma unroll
    for (int i = 1; i < KMAX; ++i) ra[i] = (tid < ns && i < K) ? row[i - 1] : T(0);
    __syncthreads();
    stage(c, K - 1);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < KMAX - 1; ++i) rc[i] = (tid < ns && i + 1 < K) ? row[i] : T(0);
    __syncthreads();
  } else {
#pragma unroll
    for (int i = 0; i < KMAX; ++i) ra[i] = rc[i] = T(0);
  }
  ra[0] = T(0);
  rc[KMAX - 1] = T(0);
  stage(d, K);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < KMAX; ++i) rd[i] = (tid < ns && i < K) ? row[i] : T(0);
  // forward elimination: rc <- c', rd <- d'
  {
    const T den = rb[0];
    rc[0] = rc[0] / den;
    rd[0] = rd[0] / den;
  }
#pragma unroll
  for (int i = 1; i < KMAX; ++i) {
    if (i < K) {
      const T den = rb[i] - ra[i] * rc[i - 1];
      if (i < K - 1) rc[i] = rc[i] / den;
      rd[i] = (rd[i] - ra[i] * rd[i - 1]) / den;
    }
  }
  // back-substitution
#pragma unroll
  for (int i = KMAX - 2; i >= 0; --i)
    if (i < K - 1) rd[i] = rd[i] - rc[i] * rd[i + 1];
  __syncthreads();  // every lane has read its d row
#pragma unroll
  for (int i = 0; i < KMAX; ++i)
    if (tid < ns && i < K) st[tid * P + i] = rd[i];
  __syncthreads();
  T* g = x + r0 * K;
  const int ds = 256 / K, di = 256 % K;
  int sy = tid / K, i = tid % K;
  for (int e = tid; e < ns * K; e += 256) {
    g[e] = st[sy * P + i];
    sy += ds;
    i += di;
    if (i >= K) {
      i -= K;
      ++sy;
    }
  }
}

// Largest K of the LDS-staged lane kernel (registers: four rows of KMAX).
template <typename T> constexpr int thomas_lds_kmax() { return sizeof(T) == 4 ? 32 : 16; }

// Lane-per-system Thomas for longer systems (global workspace): restates the
// sweep of tridiag/pcr.py:125-183.
template <typename T>
__global__ __launch_bounds__(256) void tridiag_thomas_lane_kernel(
    const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
    const T* __restrict__ d, T* __restrict__ x, int64_t B, int64_t K,
    T* __restrict__ cp) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  const T* ar = a + r * (K - 1);
  const T* br = b + r * K;
  const T* cr = c + r * (K - 1);
  const T* dr = d + r * K;
  T* xr = x + r * K;
  T* cpr = cp + r * K;
  if (K == 1) { xr[0] = dr[0] / br[0]; return; }
  T den = br[0];
  T cprev = cr[0] / den;
  T dprev = dr[0] / den;
  cpr[0] = cprev;
  xr[0] = dprev;
  for (int64_t i = 1; i < K; ++i) {
    T ai = ar[i - 1];
    T dn_ = br[i] - ai * cprev;
    if (i < K - 1) cprev = cr[i] / dn_;
    dprev = (dr[i] - ai * dprev) / dn_;
    cpr[i] = cprev;
    xr[i] = dprev;
  }
  T xn = xr[K - 1];
  for (int64_t i = K - 2; i >= 0; --i) {
    xn = xr[i] - cpr[i] * xn;
    xr[i] = xn;
  }
}

template <typename T>
static int tridiag_launch(const T* a, const T* b, const T* c, const T* d, T* x,
                          int64_t B, int64_t K, int method, T* scratch,
                          int64_t scratch_elems, hipStream_t stream) {
  if (!b || !d || !x || B < 1 || K < 1) return DN_EINVAL;
  if (K > 1 && (!a || !c)) return DN_EINVAL;
  if (method == 0) method = (K <= 32 && B >= 4096) ? 1 : 2;
  if (method == 1) {
    const int64_t blocks = (B + 255) / 256;
    if (blocks > 0x7fffffff) return DN_ERANGE;
    if (K <= thomas_lds_kmax<T>()) {
      if (K <= 8)
        hipLaunchKernelGGL((tridiag_thomas_lds_kernel<T, 8>), dim3((unsigned)blocks),
                           dim3(256), 0, stream, a, b, c, d, x, B, (int)K);
      else if (K <= 16)
        hipLaunchKernelGGL((tridiag_thomas_lds_kernel<T, 16>), dim3((unsigned)blocks),
                           dim3(256), 0, stream, a, b, c, d, x, B, (int)K);
      else
        hipLaunchKernelGGL((tridiag_thomas_lds_kernel<T, thomas_lds_kmax<T>()>),
                           dim3((unsigned)blocks), dim3(256), 0, stream, a, b, c, d, x, B,
                           (int)K);
    } else {
      if (!scratch || scratch_elems < B * K) return DN_ERANGE;
      hipLaunchKernelGGL(tridiag_thomas_lane_kernel<T>, dim3((unsigned)blocks), dim3(256), 0,
                         stream, a, b, c, d, x, B, K, scratch);
    }
  } else {
    int nr = (int)((K + 63) / 64);
    int64_t KP = 64 * (int64_t)nr;
    size_t lds = (size_t)4 * KP * sizeof(T);
    int use_lds = lds <= kTridiagLdsMax;
    if (!use_lds && (!scratch || scratch_elems < B * 4 * KP)) return DN_ERANGE;
    if (B > 0x7fffffff) return DN_ERANGE;
    hipLaunchKernelGGL(tridiag_wave_kernel<T>, dim3((unsigned)B), dim3(64),
                       use_lds ? lds : 0, stream, a, b, c, d, x, K, nr, use_lds,
                       scratch);
  }
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace dn

// Scratch elements the caller must provide (0 when none is needed).
DN_EXPORT int64_t dn_tridiag_scratch_elems(int64_t B, int64_t K, int method,
                                           int elem_bytes) {
  if (method == 0) method = (K <= 32 && B >= 4096) ? 1 : 2;
  if (method == 1) return K <= (elem_bytes == 4 ? 32 : 16) ? 0 : B * K;
  int64_t KP = 64 * ((K + 63) / 64);
  return (size_t)4 * KP * elem_bytes <= dn::kTridiagLdsMax ? 0 : B * 4 * KP;
}

DN_EXPORT int dn_tridiag_solve_f32(const float* a, const float* b, const float* c,
                                   const float* d, float* x, int64_t B, int64_t K,
                                   int method, float* scratch, int64_t scratch_elems,
                                   hipStream_t stream) {
  return dn::tridiag_launch<float>(a, b, c, d, x, B, K, method, scratch,
                                   scratch_elems, stream);
}

DN_EXPORT int dn_tridiag_solve_f64(const double* a, const double* b, const double* c,
                                   const double* d, double* x, int64_t B, int64_t K,
                                   int method, double* scratch, int64_t scratch_elems,
                                   hipStream_t stream) {
  return dn::tridiag_launch<double>(a, b, c, d, x, B, K, method, scratch,
                                    scratch_elems, stream);
}

// ---------------------------------------------------------------------------
// Allocation-free variant (float32): the elimination workspace is a fixed
// device array owned by the library instead of caller scratch.  Each GPU
// holds its own copy for the life of the process; its address is resolved
// once per device.  Every call on one GPU shares the array, so calls that
// overlap on one GPU (different streams, or threads) race on it.  The
// kernels write every workspace value before reading it, so nothing needs
// clearing between calls.
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t kNoAllocScratchElems = (int64_t)2 << 20;  // 8 MiB of float32
constexpr int kNoAllocMaxDevices = 64;

}  // namespace

static __device__ __align__(256) float scratch[kNoAllocScratchElems];

namespace {

std::mutex g_scratch_mu;
float* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device (the host-side name is
// only a stand-in), resolved on the first call on that device.
int scratch_on_current_device(float** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
    return DN_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DN_ELAUNCH;
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return DN_OK;
}

}  // namespace

DN_EXPORT int64_t dn_tridiag_noalloc_max_scratch_elems(void) { return kNoAllocScratchElems; }

DN_EXPORT int dn_tridiag_solve_noalloc_f32(const float* a, const float* b, const float* c,
                                           const float* d, float* x, int64_t B, int64_t K,
                                           int method, hipStream_t stream) {
  if (B < 1 || K < 1 || method < 0 || method > 2) return DN_EINVAL;
  const int64_t need = dn_tridiag_scratch_elems(B, K, method, (int)sizeof(float));
  if (need > kNoAllocScratchElems) return DN_ERANGE;
  float* work = nullptr;
  if (need > 0) {
    const int st = scratch_on_current_device(&work);
    if (st != DN_OK) return st;
  }
  return dn::tridiag_launch<float>(a, b, c, d, x, B, K, method, work,
                                   need > 0 ? kNoAllocScratchElems : 0, stream);
}
