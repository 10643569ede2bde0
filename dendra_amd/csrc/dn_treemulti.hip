// dn_treemulti.hip -- packed multi-morphology tree (Hines) solves.
//
// Drop-in for the reference's dhs_solve_multi_cuda and its gradient kernel
// (dendra/models/integrators/triton/dhs_kernel_multi.py:16-206,831-888;
// caller _dhs_multi._step, integrators/tree.py:2284-2300).  Row-padded planes
// d, a, b (B_total, K_stride) in solver order hold many morphologies: the
// concatenated topology tables P_cat / ORDER_cat / LAYER_PTR_cat are local
// to each morphology, and warp-plan entry w covers rows
// row_base[w] .. row_base[w] + row_count[w] - 1 of one morphology (table
// offsets p_off[w], order_off[w], lptr_off[w], L[w] layers).  Per row the
// matrix is A = diag(d + a + sum_children a) - offdiag(a); elimination runs
// over the morphology's layers (at most L_max) with the just-in-time
// diagonal and the cancellation-free Schur update (g/d_i) old_d_i, then
// back-substitution v_i = (b_i + g_i v_p) / d_i.
//
// One workgroup per plan entry: its rows share the morphology, so their
// wavefronts walk the same layers and the workgroup barriers line up; each
// wavefront owns a row (looping when a plan entry has more rows than the
// workgroup has wavefronts).  Sibling updates of a layer chunk are summed by
// the lane that sees the parent first, in lane order (no atomics, bitwise
// reproducible).  The row's eliminated diagonal lives in a caller workspace
// and its right-hand side / solution in the output plane.  Everything the
// plan does not cover -- columns at or beyond a morphology's node count,
// rows of no plan entry -- is written by nobody and so stays as the caller
// initialised it (zero).
#include "dn_common.hpp"

namespace dn {

struct MultiPlan {
  const int64_t *P, *O, *LP;         // concatenated topology tables
  int64_t n_cat, n_lcat;
  const int64_t *p_off, *o_off, *l_off, *nl, *row0, *rows;  // per plan entry
  int64_t B, Kst, L_max;
};

// Validated view of plan entry w (block-uniform); false if it is malformed.
struct Entry {
  int64_t poff, ooff, loff, L, row0, rows, K;
};

__device__ __forceinline__ bool plan_entry(const MultiPlan& M, int64_t w, Entry& e) {
  e.poff = M.p_off[w];
  e.ooff = M.o_off[w];
  e.loff = M.l_off[w];
  e.L = M.nl[w];
  e.row0 = M.row0[w];
  e.rows = M.rows[w];
  if (e.L < 1 || e.loff < 0 || e.loff + e.L >= M.n_lcat) return false;
  if (e.L > M.L_max) e.L = M.L_max;  // layers beyond L_max are not visited
  e.K = M.LP[e.loff + M.nl[w]];
  if (e.K < 1 || e.K > M.Kst || e.poff < 0 || e.poff + e.K > M.n_cat || e.ooff < 0 ||
      e.ooff + e.K > M.n_cat)
    return false;
  if (e.row0 < 0 || e.rows < 0 || e.row0 + e.rows > M.B) return false;
  return true;
}

template <typename T>
__global__ __launch_bounds__(1024) void dhs_multi_kernel(MultiPlan M, const T* __restrict__ dmem,
                                                         const T* __restrict__ ageom,
                                                         const T* __restrict__ rhs,
                                                         T* __restrict__ D, T* __restrict__ X) {
  __shared__ T cD[16][64], cB[16][64];
  __shared__ int32_t cP[16][64];
  Entry e;
  if (!plan_entry(M, blockIdx.x, e)) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int64_t* P = M.P + e.poff;
  const int64_t* O = M.O + e.ooff;
  const int64_t* LP = M.LP + e.loff;
  auto clampK = [&](int64_t v) { return v < 0 ? (int64_t)0 : (v > e.K ? e.K : v); };
  for (int64_t rb = 0; rb < e.rows; rb += nw) {
    const int64_t r = rb + wave;
    const bool live = r < e.rows;
    const int64_t base = (e.row0 + (live ? r : 0)) * M.Kst;
    const T* g = ageom + base;
    T* Dr = D + base;
    T* Br = X + base;
    if (live)
      for (int64_t j = lane; j < e.K; j += 64) {
        Dr[j] = dmem[base + j];
        Br[j] = rhs[base + j];
      }
    __syncthreads();
    for (int64_t l = 0; l < e.L; ++l) {
      const int64_t s = clampK(LP[l]), t = clampK(LP[l + 1]);
      for (int64_t o0 = s; o0 < t; o0 += 64) {
        const int64_t o = o0 + lane;
        const int n_act = (int)(t - o0 < 64 ? t - o0 : 64);
        int64_t i = -1, p = -1;
        if (live && o < t) {
          const int64_t q = O[o];
          i = (q >= 0 && q < e.K) ? q : -1;
        }
        T fd = T(0), fb = T(0);
        if (i >= 0) {
          const T old = Dr[i], gi = g[i];
          const T piv = old + gi;
          Dr[i] = piv;
          const int64_t q = P[i];
          p = (q >= 0 && q < e.K) ? q : -1;
          const T fac = gi / piv;
          fd = fac * old;
          fb = fac * Br[i];
        }
        cD[wave][lane] = fd;
        cB[wave][lane] = fb;
        cP[wave][lane] = (int32_t)p;
        __syncthreads();
        if (p >= 0) {
          bool leader = true;
          for (int j = 0; j < lane; ++j)
            if (cP[wave][j] == p) { leader = false; break; }
          if (leader) {
            T sd = T(0), sb = T(0);
            for (int j = lane; j < n_act; ++j)
              if (cP[wave][j] == p) { sd += cD[wave][j]; sb += cB[wave][j]; }
            Dr[p] += sd;
            Br[p] += sb;
          }
        }
        __syncthreads();
      }
    }
    for (int64_t l = e.L - 1; l >= 0; --l) {
      const int64_t s = clampK(LP[l]), t = clampK(LP[l + 1]);
      for (int64_t o0 = s; o0 < t; o0 += 64) {
        const int64_t o = o0 + lane;
        if (live && o < t) {
          const int64_t i = O[o];
          if (i >= 0 && i < e.K) {
            const int64_t q = P[i];
            const T vp = (q >= 0 && q < e.K) ? Br[q] : T(0);
            Br[i] = (Br[i] + g[i] * vp) / Dr[i];
          }
        }
        __syncthreads();
      }
    }
  }
}

// Gradient with respect to a_geom, for the solution V and the adjoint
// solution G of a row: -(g_i - g_p)(v_i - v_p) at a node with a parent,
// -g_i v_i at the root (dhs_kernel_multi.py:140-206).
template <typename T>
__global__ __launch_bounds__(256) void dhs_multi_grad_a_kernel(MultiPlan M,
                                                               const T* __restrict__ V,
                                                               const T* __restrict__ G,
                                                               T* __restrict__ dA) {
  Entry e;
  if (!plan_entry(M, blockIdx.x, e)) return;
  const int64_t* P = M.P + e.poff;
  for (int64_t r = 0; r < e.rows; ++r) {
    const int64_t base = (e.row0 + r) * M.Kst;
    for (int64_t j = threadIdx.x; j < e.K; j += blockDim.x) {
      const int64_t q = P[j];
      const T gi = G[base + j], vi = V[base + j];
      dA[base + j] = (q >= 0 && q < e.K) ? -((gi - G[base + q]) * (vi - V[base + q]))
                                         : -(gi * vi);
    }
  }
}

}  // namespace dn

namespace {

#define DN_MULTI_PLAN_ARGS                                                              \
  const int64_t *P_cat, const int64_t *order_cat, int64_t n_cat,                        \
      const int64_t *lptr_cat, int64_t n_lcat, const int64_t *warp_p_off,               \
      const int64_t *warp_order_off, const int64_t *warp_lptr_off, const int64_t *warp_l, \
      const int64_t *warp_row_base, const int64_t *warp_row_count, int64_t n_warps,     \
      int64_t B_total, int64_t K_stride, int64_t L_max

int make_plan(dn::MultiPlan& M, DN_MULTI_PLAN_ARGS) {
  if (!P_cat || !order_cat || !lptr_cat || !warp_p_off || !warp_order_off || !warp_lptr_off ||
      !warp_l || !warp_row_base || !warp_row_count)
    return DN_EINVAL;
  if (n_cat < 1 || n_lcat < 2 || n_warps < 1 || B_total < 1 || K_stride < 1 || L_max < 1)
    return DN_EINVAL;
  if (n_warps > INT32_MAX) return DN_ERANGE;
  M = dn::MultiPlan{P_cat, order_cat, lptr_cat, n_cat, n_lcat, warp_p_off, warp_order_off,
                    warp_lptr_off, warp_l, warp_row_base, warp_row_count, B_total, K_stride,
                    L_max};
  return DN_OK;
}

template <typename T>
int multi_solve_impl(const T* d, const T* a, const T* b, DN_MULTI_PLAN_ARGS, int max_rows,
                     T* D_work, T* x, hipStream_t stream) {
  dn::MultiPlan M;
  const int st = make_plan(M, P_cat, order_cat, n_cat, lptr_cat, n_lcat, warp_p_off,
                           warp_order_off, warp_lptr_off, warp_l, warp_row_base,
                           warp_row_count, n_warps, B_total, K_stride, L_max);
  if (st != DN_OK) return st;
  if (!d || !a || !b || !D_work || !x || max_rows < 1) return DN_EINVAL;
  const int waves = max_rows < 16 ? max_rows : 16;
  hipLaunchKernelGGL((dn::dhs_multi_kernel<T>), dim3((unsigned)n_warps), dim3(64 * waves), 0,
                     stream, M, d, a, b, D_work, x);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
int multi_grad_impl(const T* V, const T* G, DN_MULTI_PLAN_ARGS, T* dA, hipStream_t stream) {
  dn::MultiPlan M;
  const int st = make_plan(M, P_cat, order_cat, n_cat, lptr_cat, n_lcat, warp_p_off,
                           warp_order_off, warp_lptr_off, warp_l, warp_row_base,
                           warp_row_count, n_warps, B_total, K_stride, L_max);
  if (st != DN_OK) return st;
  if (!V || !G || !dA) return DN_EINVAL;
  hipLaunchKernelGGL((dn::dhs_multi_grad_a_kernel<T>), dim3((unsigned)n_warps), dim3(256), 0,
                     stream, M, V, G, dA);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace

#define DN_MULTI_PLAN_PASS                                                               \
  P_cat, order_cat, n_cat, lptr_cat, n_lcat, warp_p_off, warp_order_off, warp_lptr_off, \
      warp_l, warp_row_base, warp_row_count, n_warps, B_total, K_stride, L_max

DN_EXPORT int dn_tree_dhs_multi_f32(const float* d, const float* a, const float* b,
                                    DN_MULTI_PLAN_ARGS, int max_rows, float* D_work,
                                    float* x, hipStream_t stream) {
  return multi_solve_impl<float>(d, a, b, DN_MULTI_PLAN_PASS, max_rows, D_work, x, stream);
}
DN_EXPORT int dn_tree_dhs_multi_f64(const double* d, const double* a, const double* b,
                                    DN_MULTI_PLAN_ARGS, int max_rows, double* D_work,
                                    double* x, hipStream_t stream) {
  return multi_solve_impl<double>(d, a, b, DN_MULTI_PLAN_PASS, max_rows, D_work, x, stream);
}
DN_EXPORT int dn_tree_dhs_multi_grad_a_f32(const float* V, const float* G, DN_MULTI_PLAN_ARGS,
                                           float* dA, hipStream_t stream) {
  return multi_grad_impl<float>(V, G, DN_MULTI_PLAN_PASS, dA, stream);
}
DN_EXPORT int dn_tree_dhs_multi_grad_a_f64(const double* V, const double* G,
                                           DN_MULTI_PLAN_ARGS, double* dA, hipStream_t stream) {
  return multi_grad_impl<double>(V, G, DN_MULTI_PLAN_PASS, dA, stream);
}
