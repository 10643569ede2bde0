// dn_contract.hip -- batched tree solve by contraction.
//
// Drop-in for the reference's tree-contraction solve
// (dendra/models/integrators/triton/tree_contraction.py:748-770, factory
// integrators/__init__.py:73-110): for every system b of a batch sharing one
// tree topology, solve
//   (d_i + g_i + sum_{children c} g_c) x_i - g_i x_{parent(i)}
//                                   - sum_c g_c x_c = rhs_i
// where g_i is the conductance of the edge from node i to its parent (at
// the root, g acts as a conductance to ground).  This is the Hines matrix of
// _dhs._step (integrators/tree.py:1306-1327).
//
// Two-level contraction, one workgroup per system:
//   1. chain condensation: the host cuts the tree into sections (chains of at
//      most S+1 nodes in which every node but the last has one child).  Each
//      section eliminates its nodes but the tail, head first, all sections
//      at once; the chain becomes one edge from its tail to the parent
//      section's tail, plus additions to that tail's self term and
//      right-hand side.
//   2. the tree of section tails is contracted in host-planned rounds.
//      Each round eliminates an independent set: every current leaf (rake)
//      and alternate one-child nodes whose child is not a leaf (compress).
//      Survivors then pull the contributions of their eliminated neighbours
//      in a fixed order.  A path of n sections takes about log2(n) rounds.
//   3. back-substitution: the root, then the rounds in reverse, then every
//      section's interior at once.
// With g > 0 and d >= 0 every self term and edge is a sum of positive
// products (no cancellation on the diagonal); sums run in a fixed order, so
// results are bitwise reproducible.  When they fit beside the section
// tables, the system's d, g and rhs are staged in LDS by coalesced loads
// (the sections then gather from LDS), x leaves through LDS by coalesced
// stores, and the interior recovery coefficients stay in LDS.
#include "dn_common.hpp"

namespace dn {

template <typename T> struct ContractArgs {
  const T *d, *g, *rhs;
  T* x;
  T* work;  // 3 K per system: interior recovery coefficients
  int64_t K;
  int n_sec, n_rounds;
  const int32_t *perm;  // section-order position -> caller node
  const int32_t *sec_head, *sec_len, *sec_child, *sec_nchild, *sec_parent;
  const int32_t *elim_ptr, *elim_sec, *elim_par, *elim_child;
  const int32_t *recv_ptr, *recv_sec, *src_ptr, *src;
};

// STAGE: 0 all node data in global memory, 1 d, g, rhs and x through LDS,
// 2 also the interior recovery coefficients in LDS.
template <typename T, int STAGE>
__global__ __launch_bounds__(256) void tree_contract_kernel(ContractArgs<T> A) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int ns = A.n_sec;
  T* S = reinterpret_cast<T*>(dn_smem_raw);
  // per section: self term, rhs and parent edge of its tail; what it hands
  // its parent (PM, PF) and, when compressed, its child (CM, CF, CE); the
  // tail's solution.  After elimination SM/SF/SE hold the recovery
  // coefficients x_j = SF + SM x_parent + SE x_child.
  T *SM = S, *SF = S + ns, *SE = S + 2 * ns, *PM = S + 3 * ns, *PF = S + 4 * ns,
    *CM = S + 5 * ns, *CF = S + 6 * ns, *CE = S + 7 * ns, *XS = S + 8 * ns;
  const int64_t K = A.K;
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const T* d = A.d + b * K;
  const T* g = A.g + b * K;
  const T* rhs = A.rhs + b * K;
  T* x = A.x + b * K;
  T* xo = x;  // where the solution is written (LDS when staged)
  if constexpr (STAGE > 0) {
    T* dl = S + 9 * ns;
    T* gl = dl + K;
    T* rl = gl + K;
    for (int64_t i = tid; i < K; i += nt) {
      dl[i] = d[i];
      gl[i] = g[i];
      rl[i] = rhs[i];
    }
    __syncthreads();
    d = dl;
    g = gl;
    rhs = rl;
    xo = dl;  // d is dead after the chain condensation
  }
  T* wa = STAGE == 2 ? S + 9 * ns + 3 * K : A.work + b * 3 * K;
  T* wb = wa + K;
  T* wc = wb + K;

  // 1. chain condensation, head first
  for (int s = tid; s < ns; s += nt) {
    const int hd = A.sec_head[s], len = A.sec_len[s];
    const int o0 = A.perm[hd];
    T e = g[o0], Mc = d[o0], Fc = rhs[o0], dM = T(0), dF = T(0);
    for (int r = 0; r + 1 < len; ++r) {
      const int o1 = A.perm[hd + r + 1];
      const T gn = g[o1];
      const T inv = T(1) / (Mc + e + gn);
      const T be = e * inv, ce = gn * inv;
      dM += be * Mc;
      dF += be * Fc;
      // x_r = a_r + b_r x_(parent tail) + c_r x_(r+1)
      wa[hd + r] = Fc * inv;
      wb[hd + r] = be;
      wc[hd + r] = ce;
      Mc = d[o1] + ce * Mc;
      Fc = rhs[o1] + ce * Fc;
      e *= ce;
    }
    SM[s] = Mc;
    SF[s] = Fc;
    SE[s] = e;
    PM[s] = dM;
    PF[s] = dF;
  }
  __syncthreads();
  // the child sections' condensed heads act on this section's tail
  for (int s = tid; s < ns; s += nt) {
    const int c0 = A.sec_child[s], nc = A.sec_nchild[s];
    T m = SM[s], f = SF[s];
    for (int c = c0; c < c0 + nc; ++c) {
      m += PM[c];
      f += PF[c];
    }
    SM[s] = m;
    SF[s] = f;
  }
  __syncthreads();

  // 2. contraction rounds over the section tree
  for (int r = 0; r < A.n_rounds; ++r) {
    for (int i = A.elim_ptr[r] + tid; i < A.elim_ptr[r + 1]; i += nt) {
      const int j = A.elim_sec[i], c = A.elim_child[i];
      const T ej = SE[j], Mj = SM[j], Fj = SF[j];
      if (c < 0) {  // rake: a leaf folds into its parent
        const T inv = T(1) / (Mj + ej);
        const T fac = ej * inv;
        PM[j] = fac * Mj;
        PF[j] = fac * Fj;
        SF[j] = Fj * inv;
        SM[j] = fac;
        SE[j] = T(0);
      } else {  // compress: parent and child become neighbours
        const T ec = SE[c];
        const T inv = T(1) / (Mj + ej + ec);
        const T bj = ej * inv, cj = ec * inv;
        PM[j] = bj * Mj;
        PF[j] = bj * Fj;
        CM[j] = cj * Mj;
        CF[j] = cj * Fj;
        CE[j] = ej * cj;
        SF[j] = Fj * inv;
        SM[j] = bj;
        SE[j] = cj;
      }
    }
    __syncthreads();
    for (int i = A.recv_ptr[r] + tid; i < A.recv_ptr[r + 1]; i += nt) {
      const int q = A.recv_sec[i];
      T m = SM[q], f = SF[q], e = SE[q];
      for (int k = A.src_ptr[i]; k < A.src_ptr[i + 1]; ++k) {
        const int sj = A.src[k];
        if (sj >= 0) {  // an eliminated child
          m += PM[sj];
          f += PF[sj];
        } else {  // the compressed parent: new parent edge
          const int j = -1 - sj;
          m += CM[j];
          f += CF[j];
          e = CE[j];
        }
      }
      SM[q] = m;
      SF[q] = f;
      SE[q] = e;
    }
    __syncthreads();
  }

  // 3. back-substitution: root (section 0), rounds in reverse, interiors
  if (tid == 0) XS[0] = SF[0] / (SM[0] + SE[0]);
  __syncthreads();
  for (int r = A.n_rounds - 1; r >= 0; --r) {
    for (int i = A.elim_ptr[r] + tid; i < A.elim_ptr[r + 1]; i += nt) {
      const int j = A.elim_sec[i], p = A.elim_par[i], c = A.elim_child[i];
      T xj = SF[j] + SM[j] * XS[p];
      if (c >= 0) xj += SE[j] * XS[c];
      XS[j] = xj;
    }
    __syncthreads();
  }
  for (int s = tid; s < ns; s += nt) {
    const int hd = A.sec_head[s], len = A.sec_len[s], ps = A.sec_parent[s];
    const T xp = ps >= 0 ? XS[ps] : T(0);
    T xv = XS[s];
    xo[A.perm[hd + len - 1]] = xv;
    for (int r = len - 2; r >= 0; --r) {
      xv = wa[hd + r] + wb[hd + r] * xp + wc[hd + r] * xv;
      xo[A.perm[hd + r]] = xv;
    }
  }
  if constexpr (STAGE > 0) {
    __syncthreads();
    for (int64_t i = tid; i < K; i += nt) x[i] = xo[i];
  }
}

constexpr size_t kContractLdsMax = 160 * 1024;

inline size_t contract_lds_bytes(int n_sec, size_t elem) { return 9 * (size_t)n_sec * elem; }
// with d, g and rhs staged (stage 1), and the recovery coefficients (stage 2)
inline size_t contract_stage_lds_bytes(int n_sec, int64_t K, size_t elem, int stage) {
  return (9 * (size_t)n_sec + 3 * (size_t)stage * (size_t)K) * elem;
}

}  // namespace dn

template <typename T>
static int contract_impl(const T* d, const T* g, const T* rhs, T* x, int64_t B, int64_t K,
                         const int32_t* perm, int n_sec, const int32_t* sec_head,
                         const int32_t* sec_len, const int32_t* sec_child,
                         const int32_t* sec_nchild, const int32_t* sec_parent, int n_rounds,
                         const int32_t* elim_ptr, const int32_t* elim_sec,
                         const int32_t* elim_par, const int32_t* elim_child,
                         const int32_t* recv_ptr, const int32_t* recv_sec,
                         const int32_t* src_ptr, const int32_t* src, T* work,
                         int64_t work_elems, hipStream_t stream) {
  if (!d || !g || !rhs || !x || !perm || !sec_head || !sec_len || !sec_child ||
      !sec_nchild || !sec_parent || !work)
    return DN_EINVAL;
  if (B < 1 || K < 1 || K > INT32_MAX || n_sec < 1 || n_sec > K || n_rounds < 0)
    return DN_EINVAL;
  if (n_rounds > 0 && (!elim_ptr || !elim_sec || !elim_par || !elim_child || !recv_ptr ||
                       !recv_sec || !src_ptr || !src))
    return DN_EINVAL;
  if (B > INT32_MAX || work_elems < 3 * B * K) return DN_EINVAL;
  const size_t lds = dn::contract_lds_bytes(n_sec, sizeof(T));
  if (lds > dn::kContractLdsMax) return DN_ERANGE;
  dn::ContractArgs<T> A;
  A.d = d; A.g = g; A.rhs = rhs; A.x = x; A.work = work; A.K = K;
  A.n_sec = n_sec; A.n_rounds = n_rounds; A.perm = perm;
  A.sec_head = sec_head; A.sec_len = sec_len; A.sec_child = sec_child;
  A.sec_nchild = sec_nchild; A.sec_parent = sec_parent;
  A.elim_ptr = elim_ptr; A.elim_sec = elim_sec; A.elim_par = elim_par;
  A.elim_child = elim_child; A.recv_ptr = recv_ptr; A.recv_sec = recv_sec;
  A.src_ptr = src_ptr; A.src = src;
  // stage as much of the system in LDS as keeps two workgroups on a CU
  const size_t lds2 = dn::contract_stage_lds_bytes(n_sec, K, sizeof(T), 2);
  const size_t lds1 = dn::contract_stage_lds_bytes(n_sec, K, sizeof(T), 1);
  if (lds2 <= dn::kContractLdsMax / 2)
    hipLaunchKernelGGL((dn::tree_contract_kernel<T, 2>), dim3((unsigned)B), dim3(256), lds2,
                       stream, A);
  else if (lds1 <= dn::kContractLdsMax / 2)
    hipLaunchKernelGGL((dn::tree_contract_kernel<T, 1>), dim3((unsigned)B), dim3(256), lds1,
                       stream, A);
  else
    hipLaunchKernelGGL((dn::tree_contract_kernel<T, 0>), dim3((unsigned)B), dim3(256), lds,
                       stream, A);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// Workspace elements (of the solve's dtype) for B systems of K nodes.
DN_EXPORT int64_t dn_tree_contract_work_elems(int64_t B, int64_t K) {
  return (B < 0 || K < 0) ? -1 : 3 * B * K;
}

// Largest section count one system may have (LDS holds nine values per
// section).
DN_EXPORT int dn_tree_contract_max_sections(int elem_bytes) {
  if (elem_bytes != 4 && elem_bytes != 8) return DN_EINVAL;
  return (int)(dn::kContractLdsMax / (9 * (size_t)elem_bytes));
}

#define DN_CONTRACT_PLAN_ARGS                                                          \
  const int32_t *perm, int n_sec, const int32_t *sec_head, const int32_t *sec_len,    \
      const int32_t *sec_child, const int32_t *sec_nchild, const int32_t *sec_parent, \
      int n_rounds, const int32_t *elim_ptr, const int32_t *elim_sec,                 \
      const int32_t *elim_par, const int32_t *elim_child, const int32_t *recv_ptr,    \
      const int32_t *recv_sec, const int32_t *src_ptr, const int32_t *src
#define DN_CONTRACT_PLAN_PASS                                                          \
  perm, n_sec, sec_head, sec_len, sec_child, sec_nchild, sec_parent, n_rounds,        \
      elim_ptr, elim_sec, elim_par, elim_child, recv_ptr, recv_sec, src_ptr, src

DN_EXPORT int dn_tree_contract_f32(const float* d, const float* g, const float* rhs,
                                   float* x, int64_t B, int64_t K, DN_CONTRACT_PLAN_ARGS,
                                   float* work, int64_t work_elems, hipStream_t stream) {
  return contract_impl<float>(d, g, rhs, x, B, K, DN_CONTRACT_PLAN_PASS, work, work_elems,
                              stream);
}

DN_EXPORT int dn_tree_contract_f64(const double* d, const double* g, const double* rhs,
                                   double* x, int64_t B, int64_t K, DN_CONTRACT_PLAN_ARGS,
                                   double* work, int64_t work_elems, hipStream_t stream) {
  return contract_impl<double>(d, g, rhs, x, B, K, DN_CONTRACT_PLAN_PASS, work, work_elems,
                               stream);
}
