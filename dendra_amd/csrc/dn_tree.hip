// dn_tree.hip -- batched tree (Hines) solves for gfx950.
//
// Drop-in for the reference's dhs_solve_cuda
// (dendra/models/integrators/triton/dhs_kernel.py:493-651): d_mem, a_geom,
// b are (B,K) in solver order, the topology (parent, order, layer_ptr) is
// shared by the batch, x (B,K) receives the solution.  The matrix is
//   A = diag(d_mem + a + sum_children a) - offdiag(a),  a[root] = 0,
// eliminated children-first over the caller's layers with the reference's
// just-in-time diagonal and cancellation-free Schur update (g/d_i) * old_d_i
// (dhs_kernel.py:52-77), then back-substituted v_i = (b_i + g_i v_p) / d_i.
//
// The reference accumulates sibling updates with float atomics, so its sums
// depend on scheduling.  Here one 64-lane wavefront owns a system: within a
// layer chunk the lane that sees a parent first sums every sibling's update
// in lane order, so the result is bitwise reproducible run to run.
#include "dn_common.hpp"

namespace dn {

template <typename T>
__global__ __launch_bounds__(64) void dhs_layer_kernel(
    const T* __restrict__ dmem, const T* __restrict__ ageom, const T* __restrict__ rhs,
    const int64_t* __restrict__ parent, const int64_t* __restrict__ order,
    const int64_t* __restrict__ layer_ptr, T* __restrict__ x, int64_t K, int64_t L,
    int use_lds, T* __restrict__ scratch) {
  extern __shared__ unsigned char dn_smem_raw[];
  __shared__ T cD[64], cB[64];
  __shared__ int64_t cP[64];
  const int64_t r = blockIdx.x;
  const int lane = threadIdx.x;
  T* base = use_lds ? reinterpret_cast<T*>(dn_smem_raw) : scratch + r * 4 * K;
  T *D = base, *Bv = base + K, *G = base + 2 * K, *V = base + 3 * K;
  for (int64_t j = lane; j < K; j += 64) {
    D[j] = dmem[r * K + j];
    Bv[j] = rhs[r * K + j];
    G[j] = ageom[r * K + j];
  }
  __syncthreads();
  for (int64_t l = 0; l < L; ++l) {
    const int64_t s = layer_ptr[l], e = layer_ptr[l + 1];
    for (int64_t base_o = s; base_o < e; base_o += 64) {
      const int64_t o = base_o + lane;
      int64_t i = o < e ? order[o] : -1;
      if (i >= K) i = -1;
      int64_t p = -1;
      T fd = T(0), fb = T(0);
      if (i >= 0) {
        T old = D[i];
        T g = G[i];
        T piv = old + g;
        D[i] = piv;
        p = parent[i];
        if (p >= K) p = -1;
        T fac = g / piv;
        fd = fac * old;
        fb = fac * Bv[i];
      }
      cD[lane] = fd;
      cB[lane] = fb;
      cP[lane] = p;
      __syncthreads();
      if (p >= 0) {
        bool leader = true;
        for (int j = 0; j < lane; ++j)
          if (cP[j] == p) { leader = false; break; }
        if (leader) {
          T sd = T(0), sb = T(0);
          for (int j = lane; j < 64; ++j)
            if (cP[j] == p) { sd += cD[j]; sb += cB[j]; }
          D[p] += sd;
          Bv[p] += sb;
        }
      }
      __syncthreads();
    }
  }
  for (int64_t l = L - 1; l >= 0; --l) {
    const int64_t s = layer_ptr[l], e = layer_ptr[l + 1];
    for (int64_t base_o = s; base_o < e; base_o += 64) {
      const int64_t o = base_o + lane;
      int64_t i = o < e ? order[o] : -1;
      if (i >= 0 && i < K) {
        int64_t p = parent[i];
        T vp = (p >= 0 && p < K) ? V[p] : T(0);
        V[i] = (Bv[i] + G[i] * vp) / D[i];
      }
      __syncthreads();
    }
  }
  for (int64_t j = lane; j < K; j += 64) x[r * K + j] = V[j];
}

template <typename T>
static int dhs_launch(const T* d, const T* a, const T* b, const int64_t* parent,
                      const int64_t* order, const int64_t* layer_ptr, T* x, int64_t B,
                      int64_t K, int64_t L, T* scratch, int64_t scratch_elems,
                      hipStream_t stream) {
  if (!d || !a || !b || !parent || !order || !layer_ptr || !x) return DN_EINVAL;
  if (B < 1 || K < 1 || L < 1 || B > 0x7fffffff) return DN_EINVAL;
  size_t lds = (size_t)4 * K * sizeof(T);
  int use_lds = lds <= 48 * 1024;
  if (!use_lds && (!scratch || scratch_elems < 4 * B * K)) return DN_ERANGE;
  hipLaunchKernelGGL(dhs_layer_kernel<T>, dim3((unsigned)B), dim3(64),
                     use_lds ? lds : 0, stream, d, a, b, parent, order, layer_ptr, x, K,
                     L, use_lds, scratch);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace dn

DN_EXPORT int64_t dn_tree_dhs_scratch_elems(int64_t B, int64_t K, int elem_bytes) {
  return (size_t)4 * K * elem_bytes <= 48 * 1024 ? 0 : 4 * B * K;
}

DN_EXPORT int dn_tree_dhs_f32(const float* d, const float* a, const float* b,
                              const int64_t* parent, const int64_t* order,
                              const int64_t* layer_ptr, float* x, int64_t B, int64_t K,
                              int64_t L, float* scratch, int64_t scratch_elems,
                              hipStream_t stream) {
  return dn::dhs_launch<float>(d, a, b, parent, order, layer_ptr, x, B, K, L, scratch,
                               scratch_elems, stream);
}

DN_EXPORT int dn_tree_dhs_f64(const double* d, const double* a, const double* b,
                              const int64_t* parent, const int64_t* order,
                              const int64_t* layer_ptr, double* x, int64_t B, int64_t K,
                              int64_t L, double* scratch, int64_t scratch_elems,
                              hipStream_t stream) {
  return dn::dhs_launch<double>(d, a, b, parent, order, layer_ptr, x, B, K, L, scratch,
                                scratch_elems, stream);
}
