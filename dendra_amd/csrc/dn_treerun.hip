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
// nodes per section (the host splits longer chains): 13 keeps the section
// kernel's register spill small at 128 VGPRs while the bench-size groups
// still fit two 512-thread blocks per CU in one round
constexpr int kSecMax = 13;

struct TreePlan {
  int n_trees;
  const int64_t* node_off;   // n_trees + 1
  const int32_t* parent;     // local parent (solver order), -1 at the root
  const int32_t* ch_ptr;     // per node, into ch_idx (global), total + 1
  const int32_t* ch_idx;     // local child indices
  const int32_t* sec_off;    // n_trees + 1 -> per-section arrays
  const int32_t* sec_head;   // local node index of the section head
  const int32_t* sec_len;    // nodes in the section, 1..kSecMax
  const int32_t* sec_child;  // local index of the first child section
  const int32_t* sec_nchild; // number of child sections (contiguous)
  const int32_t* sec_ptail;  // local parent node of the head (-1 at the root)
  const int32_t* sec_parent; // local parent section (-1 at the root)
  const int32_t* lvl_off;    // n_trees + 1 -> lvl_ptr
  const int32_t* lvl_ptr;    // per tree: n_levels + 1 local section offsets
  int smax, lmax;            // largest section and level count of any tree
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
  int32_t* status;  // launch-contract status word (nullable)
};

// LDS image of one tree: four node arrays and the section table.
template <typename T> struct TreeLds {
  T* vs;       // v_old, then accumulated diagonal (before the node's own edge)
  T* dv;       // main, then the inverse pivot
  T* bv;       // rhs, then the eliminated rhs, then dv
  T* gs;       // child-indexed axial conductance
  int32_t* s0; // head | len << 16
  int32_t* s1; // first child section | n children << 16
  int32_t* s2; // parent node of the head, -1 at the root
  int32_t* lp; // level pointers
  int n_sec, n_lvl;
};

template <typename T>
__device__ __forceinline__ TreeLds<T> tree_lds_carve(unsigned char* raw, int kmax,
                                                     int smax, int lmax) {
  TreeLds<T> L;
  T* base = reinterpret_cast<T*>(raw);
  L.vs = base; L.dv = base + kmax; L.bv = base + 2 * kmax; L.gs = base + 3 * kmax;
  int32_t* ib = reinterpret_cast<int32_t*>(base + 4 * kmax);
  L.s0 = ib; L.s1 = ib + smax; L.s2 = ib + 2 * smax; L.lp = ib + 3 * smax;
  (void)lmax;
  return L;
}

__host__ __device__ constexpr size_t tree_lds_bytes(int kmax, int smax, int lmax,
                                                    size_t elem) {
  return 4 * (size_t)kmax * elem + (3 * (size_t)smax + (size_t)lmax + 1) * 4;
}

// Stage the section table and level pointers of tree `tr` (block-wide).
template <typename T>
__device__ void tree_stage_topology(const TreePlan& pl, int tr, TreeLds<T>& L) {
  const int32_t so = pl.sec_off[tr];
  L.n_sec = pl.sec_off[tr + 1] - so;
  const int32_t lo = pl.lvl_off[tr];
  L.n_lvl = pl.lvl_off[tr + 1] - lo - 1;
  for (int s = threadIdx.x; s < L.n_sec; s += blockDim.x) {
    L.s0[s] = pl.sec_head[so + s] | (pl.sec_len[so + s] << 16);
    L.s1[s] = pl.sec_child[so + s] | (pl.sec_nchild[so + s] << 16);
    L.s2[s] = pl.sec_ptail[so + s];
  }
  for (int l = threadIdx.x; l <= L.n_lvl; l += blockDim.x) L.lp[l] = pl.lvl_ptr[lo + l];
}

// Add the axial edge currents -g_k (v_k - v_p) + sum_children g_c (v_c - v_k)
// to bv, one section per thread (all sections in parallel).  vs holds v_old.
template <typename T>
__device__ void tree_axial(TreeLds<T>& L) {
  for (int s = threadIdx.x; s < L.n_sec; s += blockDim.x) {
    const int w0 = L.s0[s], w1 = L.s1[s];
    const int hd = w0 & 0xFFFF, len = w0 >> 16;
    const int cf = w1 & 0xFFFF, nc = w1 >> 16;
    const int pt = L.s2[s];
    T vp = pt >= 0 ? L.vs[pt] : T(0);
    T vr[kSecMax], gr[kSecMax];
#pragma unroll
    for (int r = 0; r < kSecMax; ++r)
      if (r < len) { vr[r] = L.vs[hd + r]; gr[r] = L.gs[hd + r]; }
    T F[kSecMax];
#pragma unroll
    for (int r = 0; r < kSecMax; ++r) {
      if (r < len) {
        F[r] = gr[r] * (vr[r] - vp);
        vp = vr[r];
      }
    }
    T vt = vp;  // tail voltage
    T tail_in = T(0);
    for (int c = cf; c < cf + nc; ++c) {
      const int hc = L.s0[c] & 0xFFFF;
      tail_in += L.gs[hc] * (L.vs[hc] - vt);
    }
#pragma unroll
    for (int r = 0; r < kSecMax; ++r) {
      if (r < len) {
        T inflow = (r + 1 < len) ? F[r + 1 < kSecMax ? r + 1 : r] : tail_in;
        L.bv[hd + r] += inflow - F[r];
      }
    }
  }
}

// Elimination (deepest level first) and back-substitution (root level
// first) over section chains held in registers.  Entry: dv = main, bv = f,
// gs = g.  Exit: bv = dv (the voltage increment).  Block-wide.
template <typename T>
__device__ void tree_chain_solve(TreeLds<T>& L) {
  for (int l = L.n_lvl - 1; l >= 0; --l) {
    for (int s = L.lp[l] + threadIdx.x; s < L.lp[l + 1]; s += blockDim.x) {
      const int w0 = L.s0[s], w1 = L.s1[s];
      const int hd = w0 & 0xFFFF, len = w0 >> 16;
      const int cf = w1 & 0xFFFF, nc = w1 >> 16;
      T Dr[kSecMax], Br[kSecMax], gr[kSecMax];
#pragma unroll
      for (int r = 0; r < kSecMax; ++r)
        if (r < len) { Dr[r] = L.dv[hd + r]; Br[r] = L.bv[hd + r]; gr[r] = L.gs[hd + r]; }
      T pullD = T(0), pullB = T(0);
      for (int c = cf; c < cf + nc; ++c) {
        const int hc = L.s0[c] & 0xFFFF;
        const T fac = L.gs[hc] * L.dv[hc];
        pullD += fac * L.vs[hc];
        pullB += fac * L.bv[hc];
      }
      T D = T(0), B = T(0), fac = T(0);
#pragma unroll
      for (int r = kSecMax - 1; r >= 0; --r) {
        if (r < len) {
          if (r == len - 1) {
            D = Dr[r] + pullD;
            B = Br[r] + pullB;
          } else {
            D = Dr[r] + fac * D;
            B = Br[r] + fac * B;
          }
          const T inv = frcp(D + gr[r]);
          L.dv[hd + r] = inv;
          L.vs[hd + r] = D;
          L.bv[hd + r] = B;
          fac = gr[r] * inv;
        }
      }
    }
    __syncthreads();
  }
  for (int l = 0; l < L.n_lvl; ++l) {
    for (int s = L.lp[l] + threadIdx.x; s < L.lp[l + 1]; s += blockDim.x) {
      const int w0 = L.s0[s];
      const int hd = w0 & 0xFFFF, len = w0 >> 16;
      const int pt = L.s2[s];
      T ir[kSecMax], Br[kSecMax], gr[kSecMax];
#pragma unroll
      for (int r = 0; r < kSecMax; ++r)
        if (r < len) { ir[r] = L.dv[hd + r]; Br[r] = L.bv[hd + r]; gr[r] = L.gs[hd + r]; }
      T x = pt >= 0 ? L.bv[pt] : T(0);
#pragma unroll
      for (int r = 0; r < kSecMax; ++r) {
        if (r < len) {
          x = (Br[r] + gr[r] * x) * ir[r];
          L.bv[hd + r] = x;
        }
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
                                                           int kmax, int smax, int lmax) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int tr = order ? order[blockIdx.x] : (int)blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t no = A.pl.node_off[tr];
  const int K = (int)(A.pl.node_off[tr + 1] - no);
  const int n_sec = A.pl.sec_off[tr + 1] - A.pl.sec_off[tr];
  const int n_lvl = A.pl.lvl_off[tr + 1] - A.pl.lvl_off[tr] - 1;
  if (K > kmax || n_sec > smax || n_lvl > lmax) {  // launch contract violated
    if (tid == 0) flag_status(A.status, kStatusTreeSize);
    return;
  }
  TreeLds<T> L = tree_lds_carve<T>(dn_smem_raw, kmax, smax, lmax);
  tree_stage_topology<T>(A.pl, tr, L);
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
    tree_axial<T>(L);
    __syncthreads();
    tree_chain_solve<T>(L);
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
// Persistent run, one thread per section, several trees per block.
// Block b runs the trees order[grp_off[b] .. grp_off[b+1]) whose section
// counts sum to at most blockDim.x; thread s owns block-local section s.  The
// thread keeps the V, m, h, n of its section's nodes (<= kSecMax, contiguous)
// in registers for the whole run, re-reads their membrane coefficients from
// L2 each step, and performs their gating and their share of the axial
// assembly.  The solve is a two-level contraction of each tree:
//   1. compress: every section eliminates its nodes but the tail, head
//      first, all sections at once (the chain becomes one edge from its tail
//      to the parent section's tail);
//   2. the tree of section tails (the section tree's shape, one node per
//      section) is eliminated level by level, deepest first, and
//      back-substituted root level first: one reciprocal per level;
//   3. the interior nodes of every section are back-substituted at once.
// The dependent path per step is one section chain plus two passes over the
// section levels, instead of every section chain on a root-to-leaf path.
// LDS carries:
//   gs      axial conductance of every node of the block's trees
//   ars     membrane area of every node, and cms its cm/dt when node_cap
//           leaves room for it (cm_lds; else cm/dt is read each step)
//   vh/vt   head and tail voltage (axial currents across section joins)
//   cD/cB   a section's contribution to its parent tail's self term and
//           right-hand side (compressed chain + Schur term of its tail)
//   xt      the tails' solutions.
// The level loops run to the deepest tree of the block; a thread acts when
// the loop reaches its own section's level.
// ---------------------------------------------------------------------------
// Four waves per SIMD (two 512-thread blocks per CU) beat three waves without
// spills: the level phases are latency-bound, so residency matters more than
// the dozen registers the bound spills.
template <typename T, bool SCALES, bool CM_LDS>
__global__ __launch_bounds__(512, 4) void hh_tree_sec_kernel(TreeRunArgs<T> A,
                                                          const int32_t* order,
                                                          const int32_t* grp_off,
                                                          int node_cap) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int nt = blockDim.x;
  T* gs = reinterpret_cast<T*>(dn_smem_raw);
  T* ars = gs + node_cap;  // per-node membrane area, staged once
  T* cms = ars + node_cap;  // per-node cm/dt when CM_LDS
  T *vh = cms + (CM_LDS ? node_cap : 0), *vt = vh + nt, *cD = vh + 2 * nt,
    *cB = vh + 3 * nt, *xt = vh + 4 * nt, *gh = vh + 5 * nt, *tM = vh + 6 * nt,
    *tF = vh + 7 * nt;
  const int s = threadIdx.x;
  const int g0 = grp_off[blockIdx.x], g1 = grp_off[blockIdx.x + 1];
  // Sections of the group's trees, numbered tree by tree in breadth-first
  // order ("BFS slot"), get two numberings:
  //  * level-major rank (level, tree, section): the section-boundary tables
  //    in LDS (vh, vt, gh, cD, cB, xt, tM, tF) are indexed by it, and thread
  //    j plays tail node j in the level phases, so a level's elimination and
  //    back-substitution run on one or two wavefronts;
  //  * length-major thread (longest section first, level-major among equal
  //    lengths): the thread that holds the section's nodes in registers for
  //    the per-node phases (gating, axial currents, chain compression and
  //    recovery), whose loops run as long as the longest section of the
  //    wavefront -- sorting by length keeps the lanes of a wavefront busy.
  // locate() resolves a BFS slot to its tree and section.
  int sec_base = 0, node_base = 0, nlvl_max = 0;
  for (int k = g0; k < g1; ++k) {
    const int t = order[k];
    sec_base += A.pl.sec_off[t + 1] - A.pl.sec_off[t];
    node_base += (int)(A.pl.node_off[t + 1] - A.pl.node_off[t]);
    const int nl = A.pl.lvl_off[t + 1] - A.pl.lvl_off[t] - 1;
    nlvl_max = nl > nlvl_max ? nl : nlvl_max;
  }
  if (sec_base > nt || node_base > node_cap) {  // launch contract violated
    if (s == 0) flag_status(A.status, kStatusTreeGroup);
    return;
  }
  struct Loc { int tree, k, sec_base, node_base, ls, lvl; };
  auto locate = [&](int slot) {
    Loc L = {-1, -1, 0, 0, 0, -1};
    int sb = 0, nb = 0;
    for (int k = g0; k < g1; ++k) {
      const int t = order[k];
      const int ns = A.pl.sec_off[t + 1] - A.pl.sec_off[t];
      if (slot >= sb && slot < sb + ns) {
        L.tree = t; L.k = k; L.sec_base = sb; L.node_base = nb; L.ls = slot - sb;
        const int32_t* lp = A.pl.lvl_ptr + A.pl.lvl_off[t];
        const int nl = A.pl.lvl_off[t + 1] - A.pl.lvl_off[t] - 1;
        for (int l = 0; l < nl; ++l)
          if (L.ls >= lp[l] && L.ls < lp[l + 1]) L.lvl = l;
      }
      sb += ns;
      nb += (int)(A.pl.node_off[t + 1] - A.pl.node_off[t]);
    }
    return L;
  };
  // level-major rank of this thread's BFS slot
  int32_t* new_of = reinterpret_cast<int32_t*>(vh);   // BFS slot -> rank
  int32_t* who = reinterpret_cast<int32_t*>(vt);      // rank -> BFS slot
  int32_t* len_of = reinterpret_cast<int32_t*>(cD);   // rank -> section length
  int32_t* owned = reinterpret_cast<int32_t*>(cB);    // thread -> rank it holds
  who[s] = -1;
  len_of[s] = 0;
  owned[s] = -1;
  __syncthreads();
  {
    const Loc L = locate(s);
    if (L.tree >= 0) {
      int rank = 0;
      for (int k = g0; k < g1; ++k) {
        const int t = order[k];
        const int32_t* lp = A.pl.lvl_ptr + A.pl.lvl_off[t];
        const int nl = A.pl.lvl_off[t + 1] - A.pl.lvl_off[t] - 1;
        rank += lp[L.lvl < nl ? L.lvl : nl];                       // levels above
        if (k < L.k && L.lvl < nl) rank += lp[L.lvl + 1] - lp[L.lvl];  // earlier trees
      }
      rank += L.ls - A.pl.lvl_ptr[A.pl.lvl_off[L.tree] + L.lvl];
      new_of[s] = rank;
      who[rank] = s;
      len_of[rank] = A.pl.sec_len[A.pl.sec_off[L.tree] + L.ls];
    }
  }
  __syncthreads();
  // tail node s of the level phases (rank s)
  int cf_t = 0, nc_t = 0, psec_t = -1, my_lvl = -1;
  {
    const int bfs = who[s];
    const Loc L = bfs >= 0 ? locate(bfs) : Loc{-1, -1, 0, 0, 0, -1};
    if (L.tree >= 0) {
      const int32_t so = A.pl.sec_off[L.tree];
      nc_t = A.pl.sec_nchild[so + L.ls];
      // children are consecutive at the next level, so they stay consecutive
      cf_t = nc_t > 0 ? new_of[A.pl.sec_child[so + L.ls] + L.sec_base] : 0;
      const int ps = A.pl.sec_parent[so + L.ls];
      psec_t = ps >= 0 ? new_of[ps + L.sec_base] : -1;
      my_lvl = L.lvl;
      // length-major thread of rank s: sections longer than this one, then
      // sections of equal length and lower rank (ranks without a section
      // have length 0 and come after every section)
      const int ln = len_of[s];
      int pos = 0;
      for (int k = 0; k < nt; ++k) {
        const int lk = len_of[k];
        pos += (lk > ln) || (lk == ln && k < s);
      }
      owned[pos] = s;
    }
  }
  __syncthreads();
  // the section this thread holds in registers (rank q)
  const int q = owned[s];
  const Loc L = q >= 0 ? locate(who[q]) : Loc{-1, -1, 0, 0, 0, -1};
  const int my_tree = L.tree, tree_node_base = L.node_base;
  const bool own = my_tree >= 0;
  int hd = 0, len = 0, cf = 0, nc = 0, psec = -1;
  int64_t no = 0;
  if (own) {
    const int32_t so = A.pl.sec_off[my_tree];
    no = A.pl.node_off[my_tree];
    hd = A.pl.sec_head[so + L.ls];
    len = A.pl.sec_len[so + L.ls];
    nc = A.pl.sec_nchild[so + L.ls];
    cf = nc > 0 ? new_of[A.pl.sec_child[so + L.ls] + L.sec_base] : 0;
    const int ps = A.pl.sec_parent[so + L.ls];
    psec = ps >= 0 ? new_of[ps + L.sec_base] : -1;
  }
  // a section longer than the kernel holds: the group is left unadvanced
  if (__syncthreads_or(own && len > kSecMax)) {
    if (s == 0) flag_status(A.status, kStatusTreeGroup);
    return;
  }
  // stage the group's axial conductances into the block's node table
  node_base = 0;
  for (int k = g0; k < g1; ++k) {
    const int t = order[k];
    const int64_t to = A.pl.node_off[t];
    const int nn = (int)(A.pl.node_off[t + 1] - to);
    for (int i = s; i < nn; i += nt) {
      gs[node_base + i] = A.g[to + i];
      ars[node_base + i] = A.area[to + i];
      if (CM_LDS) cms[node_base + i] = A.cmdt[to + i];
    }
    node_base += nn;
  }
  const int gb = tree_node_base + hd;  // block-local index of the section head
  const T* cdp = A.cmdt + no + hd;
  const T* drp = A.drive ? A.drive + no + hd : nullptr;
  T v[kSecMax], m[kSecMax], h[kSecMax], n[kSecMax];
#pragma unroll
  for (int r = 0; r < kSecMax; ++r) {
    const bool in = r < len;
    const int64_t idx = no + hd + r;
    v[r] = in ? A.v[idx] : T(0);
    m[r] = in ? A.m[idx] : T(0);
    h[r] = in ? A.h[idx] : T(0);
    n[r] = in ? A.n[idx] : T(0);
  }
  __syncthreads();
  if (own) gh[q] = gs[gb];
  const T dt = A.dt;
  for (int64_t st = 0; st < A.n_steps; ++st) {
    const T w = A.wave ? A.wave[A.step0 + st] : T(0);
    // the drive is reloaded only on steps whose waveform value is nonzero
    // (w is the same for the whole workgroup)
    const bool drw = drp && w != T(0);
    T vtail = v[0];
#pragma unroll
    for (int r = 1; r < kSecMax; ++r)
      if (r < len) vtail = v[r];
    if (own) { vh[q] = v[0]; vt[q] = vtail; }
    T M[kSecMax], F[kSecMax];
#pragma unroll
    for (int r = 0; r < kSecMax; ++r) {
      if (r < len) {
        const T sna = SCALES && A.sna ? A.sna[no + hd + r] : T(1);
        const T sk = SCALES && A.sk ? A.sk[no + hd + r] : T(1);
        T G, J;
        hh_advance<T>(v[r], dt, A.hh, sna, sk, m[r], h[r], n[r], G, J);
        const T ar = ars[gb + r];
        M[r] = (CM_LDS ? cms[gb + r] : cdp[r]) + G * ar;
        F[r] = (J - G * v[r]) * ar + (drw ? drp[r] * w : T(0));
      }
    }
    __syncthreads();
    // axial currents: -g_k (v_k - v_parent) + sum_children g_c (v_c - v_k)
    if (own) {
      T vprev = psec >= 0 ? vt[psec] : T(0);
      T tail_in = T(0);
      for (int c = cf; c < cf + nc; ++c) tail_in += gh[c] * (vh[c] - vtail);
#pragma unroll
      for (int r = 0; r < kSecMax; ++r) {
        if (r < len) {
          const T fl = gs[gb + r] * (v[r] - vprev);
          F[r] -= fl;
          if (r > 0) F[r - 1 >= 0 ? r - 1 : 0] += fl;
          vprev = v[r];
        }
      }
#pragma unroll
      for (int r = 0; r < kSecMax; ++r)
        if (r == len - 1) F[r] += tail_in;
    }
    // Chain compression, every section at once.  The matrix is
    // diag_i = M_i + (edges of i), off-diagonal -g on each edge, with g_r the
    // edge from node r to its parent (node r-1, or the parent section's tail
    // for the head).  Eliminating the section's nodes head first, all but
    // its tail, leaves
    //   * one equivalent edge e from the tail to the parent section's tail,
    //   * (dM, dF) added to the parent tail's self term and right-hand side,
    //   * per eliminated node r: x_r = a_r + b_r x_ptail + c_r x_{r+1},
    //     kept in F[r], M[r] and gq[r+1] (each is dead once node r is done).
    // Every term of a self term or an edge stays a sum of positive products.
    // All conductances are read before the dependent sweep (reads past the
    // section end stay inside the block's LDS and are unused).
    T gq[kSecMax];
#pragma unroll
    for (int r = 0; r < kSecMax; ++r) gq[r] = gs[gb + r];
    T e = gq[0], dM = T(0), dF = T(0), Mc = T(0), Fc = T(0);
    if (len > 0) { Mc = M[0]; Fc = F[0]; }
#pragma unroll
    for (int r = 0; r + 1 < kSecMax; ++r) {
      if (r + 1 < len) {
        const T gn = gq[r + 1];
        const T inv = frcp(Mc + e + gn);
        const T be = e * inv, ce = gn * inv;
        dM += be * Mc;
        dF += be * Fc;
        F[r] = Fc * inv;
        M[r] = be;
        gq[r + 1] = ce;
        Mc = M[r + 1] + ce * Mc;
        Fc = F[r + 1] + ce * Fc;
        e *= ce;
      }
    }
    // The sections' tails form a tree with the section tree's shape: node j
    // has self term Mc + sum over child sections c of (dM_c + fac_c D_c),
    // right-hand side Fc + sum of (dF_c + fac_c B_c), edge e to its parent.
    // The holder of each section hands (Mc, Fc, e, dM, dF) to tail node j =
    // its rank: tM, tF, xt (e until the tails are solved), cD and cB (dM and
    // dF until node j adds its Schur term).
    if (own) {
      tM[q] = Mc;
      tF[q] = Fc;
      xt[q] = e;
      cD[q] = dM;
      cB[q] = dF;
    }
    __syncthreads();
    // Elimination, deepest level first; one node per section per level.  A
    // tail's own terms are read once, before the level loop; at its level it
    // reads its children's terms (two at a time, so the loads overlap) and
    // adds them in child order.
    T fac = T(0), Binv = T(0);
    T D0 = T(0), B0 = T(0), et = T(0), dM0 = T(0), dF0 = T(0);
    if (my_lvl >= 0) {
      D0 = tM[s];
      B0 = tF[s];
      et = xt[s];
      dM0 = cD[s];
      dF0 = cB[s];
    }
    for (int l = nlvl_max - 1; l >= 0; --l) {
      if (my_lvl == l) {
        T D = D0, B = B0;
        const int ce = cf_t + nc_t;
        int c = cf_t;
        for (; c + 1 < ce; c += 2) {
          const T d0 = cD[c], d1 = cD[c + 1], b0 = cB[c], b1 = cB[c + 1];
          D += d0;
          D += d1;
          B += b0;
          B += b1;
        }
        if (c < ce) {
          D += cD[c];
          B += cB[c];
        }
        const T inv = frcp(D + et);
        fac = et * inv;
        Binv = B * inv;
        cD[s] = dM0 + fac * D;
        cB[s] = dF0 + fac * B;
      }
      __syncthreads();
    }
    // tails, root level first
    for (int l = 0; l < nlvl_max; ++l) {
      if (my_lvl == l) xt[s] = Binv + fac * (psec_t >= 0 ? xt[psec_t] : T(0));
      __syncthreads();
    }
    // interior nodes of every section at once, tail first
    if (own) {
      const T xp = psec >= 0 ? xt[psec] : T(0);
      T x = xt[q];
#pragma unroll
      for (int r = kSecMax - 1; r >= 0; --r) {
        if (r < len) {
          if (r < len - 1) x = F[r] + M[r] * xp + gq[r + 1 < kSecMax ? r + 1 : r] * x;
          v[r] += x;
        }
      }
    }
    if (A.rec && ((A.step0 + st + 1) % A.rec_every) == 0) {
      const int64_t row = (A.step0 + st + 1) / A.rec_every - 1;
#pragma unroll
      for (int r = 0; r < kSecMax; ++r)
        if (r < len) A.rec[row * A.total + no + hd + r] = v[r];
    }
  }
#pragma unroll
  for (int r = 0; r < kSecMax; ++r) {
    if (r < len) {
      const int64_t idx = no + hd + r;
      A.v[idx] = v[r];
      A.m[idx] = m[r];
      A.h[idx] = h[r];
      A.n[idx] = n[r];
    }
  }
}

__host__ __device__ constexpr size_t tree_sec_lds_bytes(int threads, int node_cap,
                                                        size_t elem, int cm_lds = 0) {
  return ((2 + (cm_lds ? 1 : 0)) * (size_t)node_cap + 8 * (size_t)threads) * elem;
}
// cm/dt joins the LDS node table while two blocks of the section kernel still
// fit a CU (80 KiB each)
__host__ __device__ constexpr int tree_sec_cm_lds(int threads, int node_cap, size_t elem) {
  return tree_sec_lds_bytes(threads, node_cap, elem, 1) <= 80 * 1024 ? 1 : 0;
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
    T w, T dt, HHParams<T> hh, const T* __restrict__ ve, const T* __restrict__ intra,
    const T* __restrict__ gx, const T* __restrict__ gex, T* __restrict__ mainv,
    T* __restrict__ rhs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int64_t no = node_off[node_tree[i]];
  const T v = v_[i];
  T m = m_[i], h = h_[i], n = n_[i], G, J;
  hh_advance<T>(v, dt, hh, sna ? sna[i] : T(1), sk ? sk[i] : T(1), m, h, n, G, J);
  m_[i] = m; h_[i] = h; n_[i] = n;
  if (gx) {  // further affine membrane terms (point processes): g, g e per unit area
    G += gx[i];
    J += gex[i];
  }
  const int p = parent[i];
  T ax = p >= 0 ? -g[i] * (v - v_[no + p]) : T(0);
  for (int32_t q = ch_ptr[i]; q < ch_ptr[i + 1]; ++q) {
    const int64_t c = no + ch_idx[q];
    ax += g[c] * (v_[c] - v);
  }
  if (ve) {  // edge currents of this step's extracellular potential
    const T e = ve[i];
    T ex = p >= 0 ? -g[i] * (e - ve[no + p]) : T(0);
    for (int32_t q = ch_ptr[i]; q < ch_ptr[i + 1]; ++q) {
      const int64_t c = no + ch_idx[q];
      ex += g[c] * (ve[c] - e);
    }
    ax += ex;
  }
  mainv[i] = cmdt[i] + G * area[i];
  rhs[i] = (J - G * v) * area[i] + ax + (drive ? drive[i] * w : T(0)) +
           (intra ? intra[i] : T(0));
}

template <typename T>
__global__ __launch_bounds__(256) void hh_tree_solve_kernel(
    TreePlan pl, const T* __restrict__ g, const T* __restrict__ mainv,
    const T* __restrict__ rhs, T* __restrict__ v, int kmax, int smax, int lmax) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int tr = blockIdx.x;
  const int64_t no = pl.node_off[tr];
  const int K = (int)(pl.node_off[tr + 1] - no);
  // a tree beyond the launch's LDS image is left unsolved (block-uniform exit)
  if (K > kmax || pl.sec_off[tr + 1] - pl.sec_off[tr] > smax ||
      pl.lvl_off[tr + 1] - pl.lvl_off[tr] - 1 > lmax)
    return;
  TreeLds<T> L = tree_lds_carve<T>(dn_smem_raw, kmax, smax, lmax);
  tree_stage_topology<T>(pl, tr, L);
  for (int i = threadIdx.x; i < K; i += blockDim.x) {
    L.dv[i] = mainv[no + i];
    L.bv[i] = rhs[no + i];
    L.gs[i] = g[no + i];
  }
  __syncthreads();
  tree_chain_solve<T>(L);
  for (int i = threadIdx.x; i < K; i += blockDim.x) v[no + i] += L.bv[i];
}

static bool plan_ok(const TreePlan& p) {
  return p.n_trees >= 1 && p.node_off && p.parent && p.ch_ptr && p.ch_idx &&
         p.sec_off && p.sec_head && p.sec_len && p.sec_child && p.sec_nchild &&
         p.sec_ptail && p.sec_parent && p.lvl_off && p.lvl_ptr && p.smax >= 1 &&
         p.lmax >= 1;
}

}  // namespace dn

#define DN_TREE_PLAN_ARGS                                                            \
  int n_trees, const int64_t *node_off, const int32_t *parent, const int32_t *ch_ptr, \
      const int32_t *ch_idx, const int32_t *sec_off, const int32_t *sec_head,         \
      const int32_t *sec_len, const int32_t *sec_child, const int32_t *sec_nchild,    \
      const int32_t *sec_ptail, const int32_t *sec_parent, const int32_t *lvl_off,    \
      const int32_t *lvl_ptr, int smax, int lmax

#define DN_TREE_PLAN_FILL(P)                                                          \
  P.n_trees = n_trees; P.node_off = node_off; P.parent = parent; P.ch_ptr = ch_ptr;   \
  P.ch_idx = ch_idx; P.sec_off = sec_off; P.sec_head = sec_head;                      \
  P.sec_len = sec_len; P.sec_child = sec_child; P.sec_nchild = sec_nchild;            \
  P.sec_ptail = sec_ptail; P.sec_parent = sec_parent; P.lvl_off = lvl_off;          \
  P.lvl_ptr = lvl_ptr;                                                                \
  P.smax = smax; P.lmax = lmax;

template <typename T>
static int tree_run_impl(dn::TreePlan P, const int32_t* order, const int32_t* grp_off,
                         int n_groups, int layout, int kmax, int threads,
                         T* v, T* m, T* h, T* n, const T* area, const T* cmdt,
                         const T* g, const T* drive, const T* sna, const T* sk,
                         const T* wave, int64_t step0, int64_t n_steps, T dt,
                         const T* hh, T* rec, int64_t rec_every, int64_t total,
                         int32_t* status, hipStream_t stream) {
  if (!dn::plan_ok(P) || !v || !m || !h || !n || !area || !cmdt || !g || !hh)
    return DN_EINVAL;
  if (threads < 64 || threads > 512 || (threads & 63)) return DN_EINVAL;
  if (layout == 0 && (kmax < 1 || kmax > dn::kTreeSlots * threads)) return DN_ERANGE;
  if (layout == 1 && (!grp_off || kmax < 1)) return DN_EINVAL;
  if (layout != 0 && layout != 1) return DN_EINVAL;
  if (rec && rec_every < 1) return DN_EINVAL;
  const int cm_lds = layout == 1 ? dn::tree_sec_cm_lds(threads, kmax, sizeof(T)) : 0;
  size_t lds = layout == 0 ? dn::tree_lds_bytes(kmax, P.smax, P.lmax, sizeof(T))
                           : dn::tree_sec_lds_bytes(threads, kmax, sizeof(T), cm_lds);
  if (lds > 160 * 1024) return DN_ERANGE;
  if (n_steps == 0) return DN_OK;
  dn::TreeRunArgs<T> A;
  A.pl = P; A.v = v; A.m = m; A.h = h; A.n = n; A.area = area; A.cmdt = cmdt;
  A.g = g; A.drive = drive; A.sna = sna; A.sk = sk; A.wave = wave; A.step0 = step0;
  A.n_steps = n_steps; A.dt = dt;
  A.hh = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};
  A.rec = rec; A.rec_every = rec_every; A.total = total; A.status = status;
  if (layout == 1) {
#define DN_TREE_SEC(SC, CM)                                                                  \
  hipLaunchKernelGGL((dn::hh_tree_sec_kernel<T, SC, CM>), dim3(n_groups), dim3(threads), lds, \
                     stream, A, order, grp_off, kmax)
    const bool sc = sna || sk;
    if (sc && cm_lds) DN_TREE_SEC(true, true);
    else if (sc) DN_TREE_SEC(true, false);
    else if (cm_lds) DN_TREE_SEC(false, true);
    else DN_TREE_SEC(false, false);
#undef DN_TREE_SEC
  } else if (sna || sk)
    hipLaunchKernelGGL((dn::hh_tree_run_kernel<T, true>), dim3(P.n_trees), dim3(threads),
                       lds, stream, A, order, kmax, P.smax, P.lmax);
  else
    hipLaunchKernelGGL((dn::hh_tree_run_kernel<T, false>), dim3(P.n_trees),
                       dim3(threads), lds, stream, A, order, kmax, P.smax, P.lmax);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// Persistent multi-step run of a planned ragged tree batch (trees listed in
// `order`, or all trees when order is null).  layout 0: one tree per block,
// node slots, requires kmax >= every tree's size and kmax <= 8 * threads.
// layout 1: one thread per section, block b runs the trees
// order[grp_off[b]..grp_off[b+1]) whose section counts sum to <= threads and
// whose node counts sum to <= kmax (the LDS node capacity).
DN_EXPORT int dn_hh_tree_run_f32(DN_TREE_PLAN_ARGS, const int32_t* order, int n_launch,
                                 const int32_t* grp_off, int n_groups, int layout,
                                 int kmax, int threads, float* v, float* m, float* h,
                                 float* n, const float* area, const float* cmdt,
                                 const float* g, const float* drive, const float* sna,
                                 const float* sk, const float* wave, int64_t step0,
                                 int64_t n_steps, float dt, const float* hh, float* rec,
                                 int64_t rec_every, int64_t total, int32_t* status,
                                 hipStream_t stream) {
  dn::TreePlan P;
  DN_TREE_PLAN_FILL(P)
  if (order) P.n_trees = n_launch;
  return tree_run_impl<float>(P, order, grp_off, n_groups, layout, kmax, threads, v, m, h, n, area, cmdt, g, drive,
                              sna, sk, wave, step0, n_steps, dt, hh, rec, rec_every,
                              total, status, stream);
}

DN_EXPORT int dn_hh_tree_run_f64(DN_TREE_PLAN_ARGS, const int32_t* order, int n_launch,
                                 const int32_t* grp_off, int n_groups, int layout,
                                 int kmax, int threads, double* v, double* m, double* h,
                                 double* n, const double* area, const double* cmdt,
                                 const double* g, const double* drive, const double* sna,
                                 const double* sk, const double* wave, int64_t step0,
                                 int64_t n_steps, double dt, const double* hh,
                                 double* rec, int64_t rec_every, int64_t total,
                                 int32_t* status, hipStream_t stream) {
  dn::TreePlan P;
  DN_TREE_PLAN_FILL(P)
  if (order) P.n_trees = n_launch;
  return tree_run_impl<double>(P, order, grp_off, n_groups, layout, kmax, threads, v, m, h, n, area, cmdt, g,
                               drive, sna, sk, wave, step0, n_steps, dt, hh, rec,
                               rec_every, total, status, stream);
}

template <typename T>
static int tree_step_impl(dn::TreePlan P, const int32_t* node_tree, int kmax,
                          int64_t total, T* v, T* m, T* h, T* n, const T* area,
                          const T* cmdt, const T* g, const T* drive, const T* sna,
                          const T* sk, T w, T dt, const T* hh, const T* ve,
                          const T* intra, const T* gx, const T* gex, T* mainv, T* rhs,
                          hipStream_t stream) {
  if (!gx != !gex) return DN_EINVAL;
  if (!dn::plan_ok(P) || !node_tree || !v || !m || !h || !n || !area || !cmdt || !g ||
      !hh || !mainv || !rhs)
    return DN_EINVAL;
  if (total < 1 || kmax < 1) return DN_EINVAL;
  size_t lds = dn::tree_lds_bytes(kmax, P.smax, P.lmax, sizeof(T));
  if (lds > 160 * 1024) return DN_ERANGE;
  dn::HHParams<T> p = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};
  int blocks = (int)((total + 255) / 256);
  hipLaunchKernelGGL(dn::hh_tree_assemble_kernel<T>, dim3(blocks), dim3(256), 0, stream,
                     total, node_tree, P.node_off, P.parent, P.ch_ptr, P.ch_idx, v, m, h,
                     n, area, cmdt, g, drive, sna, sk, w, dt, p, ve, intra, gx, gex, mainv,
                     rhs);
  hipLaunchKernelGGL(dn::hh_tree_solve_kernel<T>, dim3(P.n_trees), dim3(256), lds,
                     stream, P, g, mainv, rhs, v, kmax, P.smax, P.lmax);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// One plain time step (assemble + chain-solve launches) of the tree batch.
DN_EXPORT int dn_hh_tree_step_f32(DN_TREE_PLAN_ARGS, const int32_t* node_tree, int kmax,
                                  int64_t total, float* v, float* m, float* h, float* n,
                                  const float* area, const float* cmdt, const float* g,
                                  const float* drive, const float* sna, const float* sk,
                                  float w, float dt, const float* hh, const float* ve,
                                  const float* intra, const float* gx, const float* gex,
                                  float* mainv, float* rhs, hipStream_t stream) {
  dn::TreePlan P;
  DN_TREE_PLAN_FILL(P)
  return tree_step_impl<float>(P, node_tree, kmax, total, v, m, h, n, area, cmdt, g,
                               drive, sna, sk, w, dt, hh, ve, intra, gx, gex, mainv, rhs,
                               stream);
}

DN_EXPORT int dn_hh_tree_step_f64(DN_TREE_PLAN_ARGS, const int32_t* node_tree, int kmax,
                                  int64_t total, double* v, double* m, double* h,
                                  double* n, const double* area, const double* cmdt,
                                  const double* g, const double* drive, const double* sna,
                                  const double* sk, double w, double dt, const double* hh,
                                  const double* ve, const double* intra, const double* gx,
                                  const double* gex, double* mainv, double* rhs,
                                  hipStream_t stream) {
  dn::TreePlan P;
  DN_TREE_PLAN_FILL(P)
  return tree_step_impl<double>(P, node_tree, kmax, total, v, m, h, n, area, cmdt, g,
                                drive, sna, sk, w, dt, hh, ve, intra, gx, gex, mainv, rhs,
                               stream);
}

DN_EXPORT int dn_hh_tree_slots(void) { return dn::kTreeSlots; }
DN_EXPORT int dn_hh_tree_sec_max(void) { return dn::kSecMax; }
