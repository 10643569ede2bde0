// dn_treerun.hip -- fused HH + backward-Euler steps for branched trees.
//
// Restates the reference's _dhs._step (dendra/models/integrators/tree.py:
// 1202-1357) for a ragged batch of trees, in current units (mA):
//   gates at v_old, G and J from the new gates,
//   f    = (J - G v) area - sum over edges of g_c (v_c - v_p) (child -, parent +)
//          + drive w(t)                 (Ve edge currents + intra, separable)
//   main = cm_dt + G area,  DHS solve of A dv = f,  v += dv
// with A = diag(main + a + sum_children a) - offdiag(a), a child-indexed
// (tree.py:477-612), eliminated with the just-in-time diagonal and the
// (g/d) old_d Schur form of triton/dhs_kernel.py:52-77.
//
// Solver layout ("section chains").  The host renumbers every tree so that
// each maximal unbranched chain (a section: every node but the last has
// exactly one child) is contiguous, head first, and sections are ordered by
// their depth in the section tree.  One workgroup owns one tree:
//   * elimination runs section level by section level, deepest first; a
//     thread walks a whole section tail -> head in registers, pulling the
//     child sections' contributions at the tail in CSR order;
//   * back-substitution runs root level first, head -> tail.
// A barrier separates levels, so a step costs 2 x (section levels) barriers
// instead of 2 x (compartment depth) for a per-node level schedule, and
// every sum has a fixed order (bitwise reproducible, no atomics).
#include "dn_common.hpp"

namespace dn {

constexpr int kTreeSlots = 8;  // nodes per thread in the persistent kernel

struct TreePlan {
  int n_trees;
  const int64_t* node_off;   // n_trees + 1
  const int32_t* parent;     // local parent (solver order), -1 at the root
  const int32_t* ch_ptr;     // per node, into ch_idx (global), total + 1
  const int32_t* ch_idx;     // local child indices
  const int32_t* sec_off;    // n_trees + 1 -> sec_head/sec_tail
  const int32_t* sec_head;   // local node index
  const int32_t* sec_tail;
  const int32_t* lvl_off;    // n_trees + 1 -> lvl_ptr
  const int32_t* lvl_ptr;    // per tree: n_levels + 1 local section offsets
};

template <typename T> struct TreeRunArgs {
  TreePlan pl;
  T *v, *m, *h, *n;
  const T *area, *cmdt, *g, *drive, *sna, *sk;
  const T* wave;
  int64_t step0, n_steps;
  T dt;
  HHParams<T> hh;
  T* rec;
  int64_t rec_every, total;
};

// LDS arrays of one tree.
template <typename T> struct TreeLds {
  T* vs;  // v_old, then accumulated diagonal (before the node's own edge)
  T* dv;  // main, then the inverse pivot
  T* bv;  // rhs, then the eliminated rhs, then dv
  T* gs;  // child-indexed axial conductance
};

// Elimination + back-substitution of one tree held in LDS.  Requires
// vs = v_old (unused here), dv = main, bv = f, gs = g on entry; leaves the
// voltage increment in bv.  All threads of the block must call it.
template <typename T>
__device__ void tree_chain_solve(const TreePlan& pl, int tr, TreeLds<T> L) {
  const int tid = threadIdx.x, nt = blockDim.x;
  const int32_t* lp = pl.lvl_ptr + pl.lvl_off[tr];
  const int n_lvl = pl.lvl_off[tr + 1] - pl.lvl_off[tr] - 1;
  const int32_t so = pl.sec_off[tr];
  const int64_t no = pl.node_off[tr];
  const int32_t* par = pl.parent + no;
  const int32_t* cptr = pl.ch_ptr + no;
  for (int l = n_lvl - 1; l >= 0; --l) {
    for (int s = lp[l] + tid; s < lp[l + 1]; s += nt) {
      const int hd = pl.sec_head[so + s], tl = pl.sec_tail[so + s];
      T D = L.dv[tl], B = L.bv[tl];
      for (int32_t q = cptr[tl]; q < cptr[tl + 1]; ++q) {
        const int c = pl.ch_idx[q];
        const T fac = L.gs[c] * L.dv[c];
        D += fac * L.vs[c];
        B += fac * L.bv[c];
      }
      for (int k = tl; k >= hd; --k) {
        const T gk = L.gs[k];
        const T inv = frcp(D + gk);
        L.dv[k] = inv;
        L.vs[k] = D;
        L.bv[k] = B;
        if (k > hd) {
          const T fac = gk * inv;
          D = L.dv[k - 1] + fac * D;
          B = L.bv[k - 1] + fac * B;
        }
      }
    }
    __syncthreads();
  }
  for (int l = 0; l < n_lvl; ++l) {
    for (int s = lp[l] + tid; s < lp[l + 1]; s += nt) {
      const int hd = pl.sec_head[so + s], tl = pl.sec_tail[so + s];
      const int p = par[hd];
      T x = (L.bv[hd] + L.gs[hd] * (p >= 0 ? L.bv[p] : T(0))) * L.dv[hd];
      L.bv[hd] = x;
      for (int k = hd + 1; k <= tl; ++k) {
        x = (L.bv[k] + L.gs[k] * x) * L.dv[k];
        L.bv[k] = x;
      }
    }
    __syncthreads();
  }
}

// Persistent run: one block per tree, node i of the tree in slot i / nt of
// thread i % nt.  K <= kTreeSlots * blockDim.x.
template <typename T, bool SCALES>
__global__ __launch_bounds__(512) void hh_tree_run_kernel(TreeRunArgs<T> A,
                                                           const int32_t* order,
                                                           int kmax) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int tr = order ? order[blockIdx.x] : (int)blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t no = A.pl.node_off[tr];
  const int K = (int)(A.pl.node_off[tr + 1] - no);
  TreeLds<T> L;
  T* base = reinterpret_cast<T*>(dn_smem_raw);
  L.vs = base; L.dv = base + kmax; L.bv = base + 2 * kmax; L.gs = base + 3 * kmax;
  const int32_t* par = A.pl.parent + no;
  const int32_t* cptr = A.pl.ch_ptr + no;
  T v[kTreeSlots], m[kTreeSlots], h[kTreeSlots], n[kTreeSlots];
  T ar[kTreeSlots], cd[kTreeSlots], dr[kTreeSlots], sna[kTreeSlots], sk[kTreeSlots];
#pragma unroll
  for (int s = 0; s < kTreeSlots; ++s) {
    const int i = tid + s * nt;
    const bool in = i < K;
    const int64_t idx = no + i;
    v[s] = in ? A.v[idx] : T(0);
    m[s] = in ? A.m[idx] : T(0);
    h[s] = in ? A.h[idx] : T(0);
    n[s] = in ? A.n[idx] : T(0);
    ar[s] = in ? A.area[idx] : T(0);
    cd[s] = in ? A.cmdt[idx] : T(1);
    dr[s] = (in && A.drive) ? A.drive[idx] : T(0);
    sna[s] = (SCALES && in && A.sna) ? A.sna[idx] : T(1);
    sk[s] = (SCALES && in && A.sk) ? A.sk[idx] : T(1);
    if (in) L.gs[i] = A.g[idx];
  }
  const T dt = A.dt;
  for (int64_t st = 0; st < A.n_steps; ++st) {
    const T w = A.wave ? A.wave[A.step0 + st] : T(0);
#pragma unroll
    for (int s = 0; s < kTreeSlots; ++s) {
      const int i = tid + s * nt;
      if (i < K) {
        T G, J;
        hh_advance<T>(v[s], dt, A.hh, sna[s], sk[s], m[s], h[s], n[s], G, J);
        L.vs[i] = v[s];
        L.dv[i] = cd[s] + G * ar[s];
        L.bv[i] = (J - G * v[s]) * ar[s] + dr[s] * w;
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kTreeSlots; ++s) {
      const int i = tid + s * nt;
      if (i < K) {
        const int p = par[i];
        T ax = p >= 0 ? -L.gs[i] * (v[s] - L.vs[p]) : T(0);
        for (int32_t q = cptr[i]; q < cptr[i + 1]; ++q) {
          const int c = A.pl.ch_idx[q];
          ax += L.gs[c] * (L.vs[c] - v[s]);
        }
        L.bv[i] += ax;
      }
    }
    __syncthreads();
    tree_chain_solve<T>(A.pl, tr, L);
#pragma unroll
    for (int s = 0; s < kTreeSlots; ++s) {
      const int i = tid + s * nt;
      if (i < K) v[s] += L.bv[i];
    }
    if (A.rec && ((A.step0 + st + 1) % A.rec_every) == 0) {
      const int64_t row = (A.step0 + st + 1) / A.rec_every - 1;
#pragma unroll
      for (int s = 0; s < kTreeSlots; ++s) {
        const int i = tid + s * nt;
        if (i < K) A.rec[row * A.total + no + i] = v[s];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < kTreeSlots; ++s) {
    const int i = tid + s * nt;
    if (i < K) {
      const int64_t idx = no + i;
      A.v[idx] = v[s];
      A.m[idx] = m[s];
      A.h[idx] = h[s];
      A.n[idx] = n[s];
    }
  }
}

// ---------------------------------------------------------------------------
// plain per-step path: assemble (one thread per node) + chain solve per tree
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void hh_tree_assemble_kernel(
    int64_t total, const int32_t* __restrict__ node_tree,
    const int64_t* __restrict__ node_off, const int32_t* __restrict__ parent,
    const int32_t* __restrict__ ch_ptr, const int32_t* __restrict__ ch_idx,
    T* __restrict__ v_, T* __restrict__ m_, T* __restrict__ h_, T* __restrict__ n_,
    const T* __restrict__ area, const T* __restrict__ cmdt, const T* __restrict__ g,
    const T* __restrict__ drive, const T* __restrict__ sna, const T* __restrict__ sk,
    T w, T dt, HHParams<T> hh, T* __restrict__ mainv, T* __restrict__ rhs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t no = node_off[node_tree[i]];
  const T v = v_[i];
  T m = m_[i], h = h_[i], n = n_[i], G, J;
  hh_advance<T>(v, dt, hh, sna ? sna[i] : T(1), sk ? sk[i] : T(1), m, h, n, G, J);
  m_[i] = m; h_[i] = h; n_[i] = n;
  const int p = parent[i];
  T ax = p >= 0 ? -g[i] * (v - v_[no + p]) : T(0);
  for (int32_t q = ch_ptr[i]; q < ch_ptr[i + 1]; ++q) {
    const int64_t c = no + ch_idx[q];
    ax += g[c] * (v_[c] - v);
  }
  mainv[i] = cmdt[i] + G * area[i];
  rhs[i] = (J - G * v) * area[i] + ax + (drive ? drive[i] * w : T(0));
}

template <typename T>
__global__ __launch_bounds__(256) void hh_tree_solve_kernel(
    TreePlan pl, const T* __restrict__ g, const T* __restrict__ mainv,
    const T* __restrict__ rhs, T* __restrict__ v, int kmax) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int tr = blockIdx.x;
  const int64_t no = pl.node_off[tr];
  const int K = (int)(pl.node_off[tr + 1] - no);
  TreeLds<T> L;
  T* base = reinterpret_cast<T*>(dn_smem_raw);
  L.vs = base; L.dv = base + kmax; L.bv = base + 2 * kmax; L.gs = base + 3 * kmax;
  for (int i = threadIdx.x; i < K; i += blockDim.x) {
    L.dv[i] = mainv[no + i];
    L.bv[i] = rhs[no + i];
    L.gs[i] = g[no + i];
  }
  __syncthreads();
  tree_chain_solve<T>(pl, tr, L);
  for (int i = threadIdx.x; i < K; i += blockDim.x) v[no + i] += L.bv[i];
}

static bool plan_ok(const TreePlan& p) {
  return p.n_trees >= 1 && p.node_off && p.parent && p.ch_ptr && p.ch_idx &&
         p.sec_off && p.sec_head && p.sec_tail && p.lvl_off && p.lvl_ptr;
}

}  // namespace dn

#define DN_TREE_PLAN_ARGS                                                            \
  int n_trees, const int64_t *node_off, const int32_t *parent, const int32_t *ch_ptr, \
      const int32_t *ch_idx, const int32_t *sec_off, const int32_t *sec_head,         \
      const int32_t *sec_tail, const int32_t *lvl_off, const int32_t *lvl_ptr

#define DN_TREE_PLAN_FILL(P)                                                          \
  P.n_trees = n_trees; P.node_off = node_off; P.parent = parent; P.ch_ptr = ch_ptr;   \
  P.ch_idx = ch_idx; P.sec_off = sec_off; P.sec_head = sec_head;                      \
  P.sec_tail = sec_tail; P.lvl_off = lvl_off; P.lvl_ptr = lvl_ptr;

template <typename T>
static int tree_run_impl(dn::TreePlan P, const int32_t* order, int kmax, int threads,
                         T* v, T* m, T* h, T* n, const T* area, const T* cmdt,
                         const T* g, const T* drive, const T* sna, const T* sk,
                         const T* wave, int64_t step0, int64_t n_steps, T dt,
                         const T* hh, T* rec, int64_t rec_every, int64_t total,
                         hipStream_t stream) {
  if (!dn::plan_ok(P) || !v || !m || !h || !n || !area || !cmdt || !g || !hh)
    return DN_EINVAL;
  if (threads < 64 || threads > 512 || (threads & 63)) return DN_EINVAL;
  if (kmax < 1 || kmax > dn::kTreeSlots * threads) return DN_ERANGE;
  if (rec && rec_every < 1) return DN_EINVAL;
  size_t lds = (size_t)4 * kmax * sizeof(T);
  if (lds > 160 * 1024) return DN_ERANGE;
  if (n_steps == 0) return DN_OK;
  dn::TreeRunArgs<T> A;
  A.pl = P; A.v = v; A.m = m; A.h = h; A.n = n; A.area = area; A.cmdt = cmdt;
  A.g = g; A.drive = drive; A.sna = sna; A.sk = sk; A.wave = wave; A.step0 = step0;
  A.n_steps = n_steps; A.dt = dt;
  A.hh = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};
  A.rec = rec; A.rec_every = rec_every; A.total = total;
  if (sna || sk)
    hipLaunchKernelGGL((dn::hh_tree_run_kernel<T, true>), dim3(P.n_trees), dim3(threads),
                       lds, stream, A, order, kmax);
  else
    hipLaunchKernelGGL((dn::hh_tree_run_kernel<T, false>), dim3(P.n_trees),
                       dim3(threads), lds, stream, A, order, kmax);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// Persistent multi-step run of a planned ragged tree batch (trees listed in
// `order`, or all trees when order is null).  kmax >= every tree's size in the
// launch and kmax <= 8 * threads.
DN_EXPORT int dn_hh_tree_run_f32(DN_TREE_PLAN_ARGS, const int32_t* order, int n_launch,
                                 int kmax, int threads, float* v, float* m, float* h,
                                 float* n, const float* area, const float* cmdt,
                                 const float* g, const float* drive, const float* sna,
                                 const float* sk, const float* wave, int64_t step0,
                                 int64_t n_steps, float dt, const float* hh, float* rec,
                                 int64_t rec_every, int64_t total, hipStream_t stream) {
  dn::TreePlan P;
  DN_TREE_PLAN_FILL(P)
  if (order) P.n_trees = n_launch;
  return tree_run_impl<float>(P, order, kmax, threads, v, m, h, n, area, cmdt, g, drive,
                              sna, sk, wave, step0, n_steps, dt, hh, rec, rec_every,
                              total, stream);
}

DN_EXPORT int dn_hh_tree_run_f64(DN_TREE_PLAN_ARGS, const int32_t* order, int n_launch,
                                 int kmax, int threads, double* v, double* m, double* h,
                                 double* n, const double* area, const double* cmdt,
                                 const double* g, const double* drive, const double* sna,
                                 const double* sk, const double* wave, int64_t step0,
                                 int64_t n_steps, double dt, const double* hh,
                                 double* rec, int64_t rec_every, int64_t total,
                                 hipStream_t stream) {
  dn::TreePlan P;
  DN_TREE_PLAN_FILL(P)
  if (order) P.n_trees = n_launch;
  return tree_run_impl<double>(P, order, kmax, threads, v, m, h, n, area, cmdt, g,
                               drive, sna, sk, wave, step0, n_steps, dt, hh, rec,
                               rec_every, total, stream);
}

template <typename T>
static int tree_step_impl(dn::TreePlan P, const int32_t* node_tree, int kmax,
                          int64_t total, T* v, T* m, T* h, T* n, const T* area,
                          const T* cmdt, const T* g, const T* drive, const T* sna,
                          const T* sk, T w, T dt, const T* hh, T* mainv, T* rhs,
                          hipStream_t stream) {
  if (!dn::plan_ok(P) || !node_tree || !v || !m || !h || !n || !area || !cmdt || !g ||
      !hh || !mainv || !rhs)
    return DN_EINVAL;
  if (total < 1 || kmax < 1) return DN_EINVAL;
  size_t lds = (size_t)4 * kmax * sizeof(T);
  if (lds > 160 * 1024) return DN_ERANGE;
  dn::HHParams<T> p = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};
  int blocks = (int)((total + 255) / 256);
  hipLaunchKernelGGL(dn::hh_tree_assemble_kernel<T>, dim3(blocks), dim3(256), 0, stream,
                     total, node_tree, P.node_off, P.parent, P.ch_ptr, P.ch_idx, v, m, h,
                     n, area, cmdt, g, drive, sna, sk, w, dt, p, mainv, rhs);
  hipLaunchKernelGGL(dn::hh_tree_solve_kernel<T>, dim3(P.n_trees), dim3(256), lds,
                     stream, P, g, mainv, rhs, v, kmax);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// One plain time step (assemble + chain-solve launches) of the tree batch.
DN_EXPORT int dn_hh_tree_step_f32(DN_TREE_PLAN_ARGS, const int32_t* node_tree, int kmax,
                                  int64_t total, float* v, float* m, float* h, float* n,
                                  const float* area, const float* cmdt, const float* g,
                                  const float* drive, const float* sna, const float* sk,
                                  float w, float dt, const float* hh, float* mainv,
                                  float* rhs, hipStream_t stream) {
  dn::TreePlan P;
  DN_TREE_PLAN_FILL(P)
  return tree_step_impl<float>(P, node_tree, kmax, total, v, m, h, n, area, cmdt, g,
                               drive, sna, sk, w, dt, hh, mainv, rhs, stream);
}

DN_EXPORT int dn_hh_tree_step_f64(DN_TREE_PLAN_ARGS, const int32_t* node_tree, int kmax,
                                  int64_t total, double* v, double* m, double* h,
                                  double* n, const double* area, const double* cmdt,
                                  const double* g, const double* drive, const double* sna,
                                  const double* sk, double w, double dt, const double* hh,
                                  double* mainv, double* rhs, hipStream_t stream) {
  dn::TreePlan P;
  DN_TREE_PLAN_FILL(P)
  return tree_step_impl<double>(P, node_tree, kmax, total, v, m, h, n, area, cmdt, g,
                                drive, sna, sk, w, dt, hh, mainv, rhs, stream);
}

DN_EXPORT int dn_hh_tree_slots(void) { return dn::kTreeSlots; }
