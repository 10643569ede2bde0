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
// Several systems share a workgroup, with the topology staged once in LDS.
#include "dn_common.hpp"

namespace dn {

// One wavefront per system, `blockDim.x / 64` systems per workgroup.  The
// topology is shared by the batch, so every wavefront walks the same layers
// and the workgroup barriers line up; it is staged once per workgroup in LDS
// as int32 (topo_lds) when it fits.  Per system the state D, B, G lives in
// LDS (state_lds) or in the caller's scratch; the back-substituted V
// overwrites B (a node's B is read only by itself, before its V is written).
template <typename T>
__global__ __launch_bounds__(1024) void dhs_layer_kernel(
    const T* __restrict__ dmem, const T* __restrict__ ageom, const T* __restrict__ rhs,
    const int64_t* __restrict__ parent, const int64_t* __restrict__ order,
    const int64_t* __restrict__ layer_ptr, T* __restrict__ x, int64_t B, int64_t K,
    int64_t L, int topo_lds, int state_lds, T* __restrict__ scratch) {
  extern __shared__ __align__(16) unsigned char dn_smem_raw[];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int G = blockDim.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * G + wave;
  const bool live = r < B;
  unsigned char* cur = dn_smem_raw;
  int32_t *tpar = nullptr, *tord = nullptr, *tlp = nullptr;
  if (topo_lds) {
    tpar = reinterpret_cast<int32_t*>(cur);
    tord = tpar + K;
    tlp = tord + K;
    cur += ((size_t)4 * (2 * K + L + 1) + 15) / 16 * 16;
    for (int64_t j = threadIdx.x; j < K; j += blockDim.x) {
      const int64_t p = parent[j];
      const int64_t o = order[j];
      tpar[j] = (p >= 0 && p < K) ? (int32_t)p : -1;
      tord[j] = (o >= 0 && o < K) ? (int32_t)o : -1;
    }
    for (int64_t j = threadIdx.x; j <= L; j += blockDim.x) {
      const int64_t v = layer_ptr[j];
      tlp[j] = (int32_t)(v < 0 ? 0 : (v > K ? K : v));
    }
  }
  T* cD = reinterpret_cast<T*>(cur) + wave * 128;
  T* cB = cD + 64;
  int32_t* cP = reinterpret_cast<int32_t*>(reinterpret_cast<T*>(cur) + G * 128) + wave * 64;
  cur += (size_t)G * 64 * (2 * sizeof(T) + 4);
  cur = dn_smem_raw + (size_t)(cur - dn_smem_raw + 15) / 16 * 16;
  T* base = state_lds ? reinterpret_cast<T*>(cur) + (size_t)wave * 3 * K
                      : scratch + (live ? r : 0) * 3 * K;
  T *D = base, *Bv = base + K, *Gv = base + 2 * K;
  auto bound = [&](int64_t l) -> int64_t {
    if (topo_lds) return tlp[l];
    const int64_t v = layer_ptr[l];
    return v < 0 ? 0 : (v > K ? K : v);
  };
  auto node = [&](int64_t o) -> int64_t {
    if (topo_lds) return tord[o];
    const int64_t i = order[o];
    return (i >= 0 && i < K) ? i : -1;
  };
  auto par = [&](int64_t i) -> int64_t {
    if (topo_lds) return tpar[i];
    const int64_t p = parent[i];
    return (p >= 0 && p < K) ? p : -1;
  };
  if (live)
    for (int64_t j = lane; j < K; j += 64) {
      D[j] = dmem[r * K + j];
      Bv[j] = rhs[r * K + j];
      Gv[j] = ageom[r * K + j];
    }
  __syncthreads();
  for (int64_t l = 0; l < L; ++l) {
    const int64_t s = bound(l), e = bound(l + 1);
    for (int64_t base_o = s; base_o < e; base_o += 64) {
      const int64_t o = base_o + lane;
      const int n_act = (int)(e - base_o < 64 ? e - base_o : 64);
      const int64_t i = (live && o < e) ? node(o) : -1;
      int64_t p = -1;
      T fd = T(0), fb = T(0);
      if (i >= 0) {
        T old = D[i];
        T g = Gv[i];
        T piv = old + g;
        D[i] = piv;
        p = par(i);
        T fac = g / piv;
        fd = fac * old;
        fb = fac * Bv[i];
      }
      cD[lane] = fd;
      cB[lane] = fb;
      cP[lane] = (int32_t)p;
      __syncthreads();
      if (p >= 0) {
        bool leader = true;
        for (int j = 0; j < lane; ++j)
          if (cP[j] == p) { leader = false; break; }
        if (leader) {
          T sd = T(0), sb = T(0);
          for (int j = lane; j < n_act; ++j)
            if (cP[j] == p) { sd += cD[j]; sb += cB[j]; }
          D[p] += sd;
          Bv[p] += sb;
        }
      }
      __syncthreads();
    }
  }
  for (int64_t l = L - 1; l >= 0; --l) {
    const int64_t s = bound(l), e = bound(l + 1);
    for (int64_t base_o = s; base_o < e; base_o += 64) {
      const int64_t o = base_o + lane;
      const int64_t i = (live && o < e) ? node(o) : -1;
      if (i >= 0) {
        const int64_t p = par(i);
        const T vp = p >= 0 ? Bv[p] : T(0);
        Bv[i] = (Bv[i] + Gv[i] * vp) / D[i];
      }
      __syncthreads();
    }
  }
  if (live)
    for (int64_t j = lane; j < K; j += 64) x[r * K + j] = Bv[j];
}

constexpr size_t kDhsLdsMax = 160 * 1024;

struct DhsShape {
  int systems_per_block;
  int topo_lds, state_lds;
  size_t lds;
};

// Systems per workgroup and what is staged in LDS: the most resident
// wavefronts per CU (160 KiB of LDS), fewest systems per workgroup on ties.
template <typename T>
static DhsShape dhs_shape(int64_t B, int64_t K, int64_t L) {
  const size_t topo = ((size_t)4 * (2 * K + L + 1) + 15) / 16 * 16;
  const size_t comm = (size_t)64 * (2 * sizeof(T) + 4);
  const size_t state = (size_t)3 * K * sizeof(T);
  auto bytes = [&](int g, int tl, int sl) {
    return (tl ? topo : 0) + ((g * comm + 15) / 16 * 16) + (sl ? g * state : 0);
  };
  DhsShape best{1, 0, 0, bytes(1, 0, 0)};
  const int gmax = (int)(B < 16 ? B : 16);
  // state in LDS first (it carries every elimination step), then topology
  const int modes[4][2] = {{1, 1}, {0, 1}, {1, 0}, {0, 0}};
  for (const auto& m : modes) {
    int best_waves = 0;
    for (int g = 1; g <= gmax; ++g) {
      const size_t b = bytes(g, m[0], m[1]);
      if (b > kDhsLdsMax) break;
      const int waves = (int)(kDhsLdsMax / b) * g;
      if (waves > best_waves) {
        best_waves = waves;
        best = DhsShape{g, m[0], m[1], b};
      }
    }
    if (best_waves) return best;
  }
  return best;
}

template <typename T>
static int dhs_launch(const T* d, const T* a, const T* b, const int64_t* parent,
                      const int64_t* order, const int64_t* layer_ptr, T* x, int64_t B,
                      int64_t K, int64_t L, T* scratch, int64_t scratch_elems,
                      hipStream_t stream) {
  if (!d || !a || !b || !parent || !order || !layer_ptr || !x) return DN_EINVAL;
  if (B < 1 || K < 1 || L < 1 || B > 0x7fffffff || K > 0x3fffffff || L > 0x3fffffff)
    return DN_EINVAL;
  const DhsShape sh = dhs_shape<T>(B, K, L);
  if (!sh.state_lds && (!scratch || scratch_elems < 4 * B * K)) return DN_ERANGE;
  const int64_t blocks = (B + sh.systems_per_block - 1) / sh.systems_per_block;
  hipLaunchKernelGGL(dhs_layer_kernel<T>, dim3((unsigned)blocks),
                     dim3(64 * sh.systems_per_block), sh.lds, stream, d, a, b, parent, order,
                     layer_ptr, x, B, K, L, sh.topo_lds, sh.state_lds, scratch);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace dn

DN_EXPORT int64_t dn_tree_dhs_scratch_elems(int64_t B, int64_t K, int elem_bytes) {
  // state staged in LDS whenever one system's D, B, G and the lane buffers fit
  const size_t one = (size_t)3 * K * elem_bytes + (size_t)64 * (2 * elem_bytes + 4);
  return one <= 160 * 1024 ? 0 : 4 * B * K;
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
