// dn_block.hip -- 3x3 block solvers of the double-layer extracellular
// (ExtCell) model.
//
// Drop-ins for the reference's block solvers:
//   * block Thomas  thomas_bt3_solve (dendra/models/integrators/triton/
//     bt_kernel.py:192-240): tridiagonal with 3x3 main blocks and diagonal
//     coupling blocks, lower (B,K-1,3) at (k,k-1), upper (B,K-1,3) at (k,k+1);
//   * block SPD     solve_bt_spd_fwd_consume_impl / ..._solve_with_chol_impl
//     (triton/bt_spd_kernel.py:575-700): the same system by block Cholesky,
//     returning the factors for reuse by the adjoint;
//   * block tree    _dhs_bt3_adjoint (triton/dhs_kernel_bt.py): Hines matrix
//     with 3x3 node blocks and per-component edge conductances,
//       A_ii = main_i + diag(e_i) + sum_children diag(e_c),
//       A_ip = A_pi = -diag(e_i),  (the root's e is diagonal only).
//
// Block Thomas and block Cholesky: one lane per system (a wavefront solves
// 64 systems); the elimination coefficients live in a caller workspace laid
// out [row][component][system], so every workspace access of a wavefront is
// one coalesced transaction.  Block tree: one wavefront per system over the
// DHS layers (or one lane per system for path-like trees).
#include "dn_common.hpp"

namespace dn {

// inverse of a row-major 3x3 by its adjugate
template <typename T>
__device__ __forceinline__ void inv3(const T* a, T* r) {
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
__device__ __forceinline__ void mv3(const T* m, const T* v, T* r) {
  r[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  r[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  r[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}

// ---------------------------------------------------------------------------
// block Thomas: M'_0 = M_0, M'_k = M_k - diag(l) M'^-1_{k-1} diag(u),
// d'_k = d_k - diag(l) M'^-1_{k-1} d'_{k-1};  x_k = M'^-1_k (d'_k - diag(u_k) x_{k+1})
// work: [K][12][B] = inverse (9) and d' (3) per row
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void bt3_thomas_kernel(const T* __restrict__ lower,
                                                         const T* __restrict__ main,
                                                         const T* __restrict__ upper,
                                                         const T* __restrict__ rhs,
                                                         T* __restrict__ x,
                                                         T* __restrict__ work, int64_t B,
                                                         int64_t K) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const T* L = lower + s * (K - 1) * 3;
  const T* M = main + s * K * 9;
  const T* U = upper + s * (K - 1) * 3;
  const T* D = rhs + s * K * 3;
  T* X = x + s * K * 3;
  T* W = work + s;
  auto w = [&](int64_t k, int c) -> T& { return W[(k * 12 + c) * B]; };
  T inv[9], d[3], a[9];
  for (int c = 0; c < 9; ++c) a[c] = M[c];
  inv3(a, inv);
  for (int c = 0; c < 3; ++c) d[c] = D[c];
  for (int c = 0; c < 9; ++c) w(0, c) = inv[c];
  for (int c = 0; c < 3; ++c) w(0, 9 + c) = d[c];
  // step k's coupling, block and right-hand side are fetched during step
  // k - 1 (index clamped into the system: the loads are unconditional)
  T nl[3], nu[3], nm[9], nd[3];
  auto fetch = [&](int64_t k) {
    for (int c = 0; c < 3; ++c) nl[c] = L[(k - 1) * 3 + c];
    for (int c = 0; c < 3; ++c) nu[c] = U[(k - 1) * 3 + c];
    for (int c = 0; c < 9; ++c) nm[c] = M[k * 9 + c];
    for (int c = 0; c < 3; ++c) nd[c] = D[k * 3 + c];
  };
  if (K > 1) fetch(1);
  for (int64_t k = 1; k < K; ++k) {
    T l[3], u[3], m[9], dk[3];
    for (int c = 0; c < 3; ++c) l[c] = nl[c];
    for (int c = 0; c < 3; ++c) u[c] = nu[c];
    for (int c = 0; c < 9; ++c) m[c] = nm[c];
    for (int c = 0; c < 3; ++c) dk[c] = nd[c];
    fetch(k + 1 < K ? k + 1 : K - 1);
    T wd[3];
    mv3(inv, d, wd);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) a[r * 3 + c] = m[r * 3 + c] - l[r] * inv[r * 3 + c] * u[c];
    for (int r = 0; r < 3; ++r) d[r] = dk[r] - l[r] * wd[r];
    inv3(a, inv);
    for (int c = 0; c < 9; ++c) w(k, c) = inv[c];
    for (int c = 0; c < 3; ++c) w(k, 9 + c) = d[c];
  }
  T xn[3];
  mv3(inv, d, xn);
  for (int c = 0; c < 3; ++c) X[(K - 1) * 3 + c] = xn[c];
  T niv[9], nw[3], nU[3];
  auto fetch_back = [&](int64_t k) {
    for (int c = 0; c < 9; ++c) niv[c] = w(k, c);
    for (int c = 0; c < 3; ++c) nw[c] = w(k, 9 + c);
    for (int c = 0; c < 3; ++c) nU[c] = U[k * 3 + c];
  };
  if (K > 1) fetch_back(K - 2);
  for (int64_t k = K - 2; k >= 0; --k) {
    T t[3], iv[9];
    for (int c = 0; c < 9; ++c) iv[c] = niv[c];
    for (int c = 0; c < 3; ++c) t[c] = nw[c] - nU[c] * xn[c];
    fetch_back(k > 0 ? k - 1 : 0);
    mv3(iv, t, xn);
    for (int c = 0; c < 3; ++c) X[k * 3 + c] = xn[c];
  }
}

// ---------------------------------------------------------------------------
// block SPD: the same recurrence with every M'_k factored by Cholesky.
// A block that is not positive definite after symmetrisation is shifted
// along its diagonal: first by its Gershgorin deficit plus a margin
// (max(eps_abs, eps_rel * sum |a_rr|)), then by 10 more margins; info
// records a block that fails all three.  chol: [B][K][9] row-major lower
// factor; work: [K][3][B] = d'.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T mid(T a, T b) {
  const T big = std::numeric_limits<T>::max() / T(2);
  return (fabs_(a) <= big && fabs_(b) <= big) ? T(0.5) * (a + b) : T(0.5) * a + T(0.5) * b;
}

template <typename T>
__device__ __forceinline__ bool chol3(const T* a, T* l) {
  const bool c0 = a[0] > T(0);
  const T l00 = sqrt(c0 ? a[0] : T(1));
  const T l10 = a[3] / l00, l20 = a[6] / l00;
  const T s11 = a[4] - l10 * l10;
  const bool c1 = s11 > T(0);
  const T l11 = sqrt(c1 ? s11 : T(1));
  const T l21 = (a[7] - l10 * l20) / l11;
  const T s22 = a[8] - l20 * l20 - l21 * l21;
  const bool c2 = s22 > T(0);
  const T l22 = sqrt(c2 ? s22 : T(1));
  l[0] = l00; l[1] = T(0); l[2] = T(0);
  l[3] = l10; l[4] = l11; l[5] = T(0);
  l[6] = l20; l[7] = l21; l[8] = l22;
  return c0 && c1 && c2;
}

template <typename T>
__device__ __forceinline__ bool chol3_shifted(T* a, T* l, T eps_rel, T eps_abs) {
  const T s01 = mid(a[1], a[3]), s02 = mid(a[2], a[6]), s12 = mid(a[5], a[7]);
  a[1] = a[3] = s01;
  a[2] = a[6] = s02;
  a[5] = a[7] = s12;
  if (chol3(a, l)) return true;
  const T tr = fabs_(a[0]) + fabs_(a[4]) + fabs_(a[8]);
  const T g0 = a[0] - fabs_(a[1]) - fabs_(a[2]);
  const T g1 = a[4] - fabs_(a[3]) - fabs_(a[5]);
  const T g2 = a[8] - fabs_(a[6]) - fabs_(a[7]);
  T gmin = g0 < g1 ? g0 : g1;
  gmin = gmin < g2 ? gmin : g2;
  const T margin = eps_abs > eps_rel * tr ? eps_abs : eps_rel * tr;
  T tau = gmin > T(0) ? margin : margin - gmin;
  for (int attempt = 0; attempt < 2; ++attempt) {
    T b[9];
    for (int c = 0; c < 9; ++c) b[c] = a[c];
    b[0] += tau; b[4] += tau; b[8] += tau;
    if (chol3(b, l)) return true;
    tau += T(10) * margin;
  }
  return false;
}

// solve L L^T x = b
template <typename T>
__device__ __forceinline__ void chol_solve3(const T* l, const T* b, T* x) {
  const T y0 = b[0] / l[0];
  const T y1 = (b[1] - l[3] * y0) / l[4];
  const T y2 = (b[2] - l[6] * y0 - l[7] * y1) / l[8];
  x[2] = y2 / l[8];
  x[1] = (y1 - l[7] * x[2]) / l[4];
  x[0] = (y0 - l[3] * x[1] - l[6] * x[2]) / l[0];
}

template <typename T>
__global__ __launch_bounds__(256) void bt3_spd_kernel(
    const T* __restrict__ lower, const T* __restrict__ main, const T* __restrict__ upper,
    const T* __restrict__ rhs, T* __restrict__ x, T* __restrict__ chol,
    int32_t* __restrict__ info, T* __restrict__ work, int64_t B, int64_t K, T eps_rel,
    T eps_abs, int factor) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const T* L = lower + s * (K - 1) * 3;
  const T* U = upper + s * (K - 1) * 3;
  const T* D = rhs + s * K * 3;
  T* X = x + s * K * 3;
  T* C = chol + s * K * 9;
  T* W = work + s;
  auto w = [&](int64_t k, int c) -> T& { return W[(k * 3 + c) * B]; };
  T lk[9], d[3];
  int bad = 0;
  if (factor) {
    const T* M = main + s * K * 9;
    T a[9];
    for (int c = 0; c < 9; ++c) a[c] = M[c];
    bad += !chol3_shifted(a, lk, eps_rel, eps_abs);
    for (int c = 0; c < 9; ++c) C[c] = lk[c];
    for (int c = 0; c < 3; ++c) d[c] = D[c];
    for (int c = 0; c < 3; ++c) w(0, c) = d[c];
    // step k's inputs are fetched during step k - 1 (clamped index:
    // unconditional loads)
    T nl[3], nu[3], nm[9], nd[3];
    auto fetch = [&](int64_t k) {
      for (int c = 0; c < 3; ++c) nl[c] = L[(k - 1) * 3 + c];
      for (int c = 0; c < 3; ++c) nu[c] = U[(k - 1) * 3 + c];
      for (int c = 0; c < 9; ++c) nm[c] = M[k * 9 + c];
      for (int c = 0; c < 3; ++c) nd[c] = D[k * 3 + c];
    };
    if (K > 1) fetch(1);
    for (int64_t k = 1; k < K; ++k) {
      T l[3], u[3], m[9], dk[3];
      for (int c = 0; c < 3; ++c) l[c] = nl[c];
      for (int c = 0; c < 3; ++c) u[c] = nu[c];
      for (int c = 0; c < 9; ++c) m[c] = nm[c];
      for (int c = 0; c < 3; ++c) dk[c] = nd[c];
      fetch(k + 1 < K ? k + 1 : K - 1);
      // columns of M'^-1_{k-1} diag(u), and M'^-1_{k-1} d'_{k-1}
      T col[3][3], wd[3];
      for (int j = 0; j < 3; ++j) {
        const T e[3] = {j == 0 ? u[0] : T(0), j == 1 ? u[1] : T(0), j == 2 ? u[2] : T(0)};
        chol_solve3(lk, e, col[j]);
      }
      chol_solve3(lk, d, wd);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r * 3 + c] = m[r * 3 + c] - l[r] * col[c][r];
      for (int r = 0; r < 3; ++r) d[r] = dk[r] - l[r] * wd[r];
      bad += !chol3_shifted(a, lk, eps_rel, eps_abs);
      for (int c = 0; c < 9; ++c) C[k * 9 + c] = lk[c];
      for (int c = 0; c < 3; ++c) w(k, c) = d[c];
    }
  } else {  // reuse the factors: forward sweep of the right-hand side only
    for (int c = 0; c < 9; ++c) lk[c] = C[c];
    for (int c = 0; c < 3; ++c) d[c] = D[c];
    for (int c = 0; c < 3; ++c) w(0, c) = d[c];
    T nl[3], nd[3], nc[9];
    auto fetch = [&](int64_t k) {
      for (int c = 0; c < 3; ++c) nl[c] = L[(k - 1) * 3 + c];
      for (int c = 0; c < 3; ++c) nd[c] = D[k * 3 + c];
      for (int c = 0; c < 9; ++c) nc[c] = C[k * 9 + c];
    };
    if (K > 1) fetch(1);
    for (int64_t k = 1; k < K; ++k) {
      T l[3], dk[3], ck[9];
      for (int c = 0; c < 3; ++c) l[c] = nl[c];
      for (int c = 0; c < 3; ++c) dk[c] = nd[c];
      for (int c = 0; c < 9; ++c) ck[c] = nc[c];
      fetch(k + 1 < K ? k + 1 : K - 1);
      T wd[3];
      chol_solve3(lk, d, wd);
      for (int r = 0; r < 3; ++r) d[r] = dk[r] - l[r] * wd[r];
      for (int c = 0; c < 9; ++c) lk[c] = ck[c];
      for (int c = 0; c < 3; ++c) w(k, c) = d[c];
    }
  }
  T xn[3];
  chol_solve3(lk, d, xn);
  for (int c = 0; c < 3; ++c) X[(K - 1) * 3 + c] = xn[c];
  T nc[9], nw[3], nU[3];
  auto fetch_back = [&](int64_t k) {
    for (int c = 0; c < 9; ++c) nc[c] = C[k * 9 + c];
    for (int c = 0; c < 3; ++c) nw[c] = w(k, c);
    for (int c = 0; c < 3; ++c) nU[c] = U[k * 3 + c];
  };
  if (K > 1) fetch_back(K - 2);
  for (int64_t k = K - 2; k >= 0; --k) {
    T t[3];
    for (int c = 0; c < 9; ++c) lk[c] = nc[c];
    for (int c = 0; c < 3; ++c) t[c] = nw[c] - nU[c] * xn[c];
    fetch_back(k > 0 ? k - 1 : 0);
    chol_solve3(lk, t, xn);
    for (int c = 0; c < 3; ++c) X[k * 3 + c] = xn[c];
  }
  if (info) info[s] = bad;
}

// ---------------------------------------------------------------------------
// block tree (Hines) solve over the caller's elimination order (children
// before parents).  work: [K][12][B] = accumulated diagonal block, then its
// inverse (9), and the eliminated right-hand side (3).
// ---------------------------------------------------------------------------
// Optional membrane term of the ExtCell step: the solved block is
// main + g [[1, -1, 0], [-1, 1, 0], [0, 0, 0]], added here in the order the
// assembly used to write it (so the caller stores one g per node, not D).
template <typename T>
__device__ __forceinline__ void add_membrane(T* a, T g) {
  a[0] += g;
  a[1] -= g;
  a[3] -= g;
  a[4] += g;
}

template <typename T>
__global__ __launch_bounds__(256) void bt3_tree_kernel(const T* __restrict__ main,
                                                       const T* __restrict__ edge,
                                                       const T* __restrict__ rhs,
                                                       const int64_t* __restrict__ parent,
                                                       const int64_t* __restrict__ order,
                                                       T* __restrict__ x,
                                                       T* __restrict__ work, int64_t B,
                                                       int64_t K, const T* __restrict__ gmem) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B) return;
  const T* M = main + s * K * 9;
  const T* E = edge + s * K * 3;
  const T* R = rhs + s * K * 3;
  T* X = x + s * K * 3;
  T* W = work + s;
  auto w = [&](int64_t k, int c) -> T& { return W[(k * 12 + c) * B]; };
  for (int64_t i = 0; i < K; ++i) {
    for (int c = 0; c < 9; ++c) w(i, c) = M[i * 9 + c];
    if (gmem) {
      T a[9];
      for (int c = 0; c < 9; ++c) a[c] = w(i, c);
      add_membrane(a, gmem[s * K + i]);
      for (int c = 0; c < 9; ++c) w(i, c) = a[c];
    }
    for (int c = 0; c < 3; ++c) {
      w(i, c * 4) += E[i * 3 + c];
      w(i, 9 + c) = R[i * 3 + c];
    }
  }
  // out-of-range parents count as roots and out-of-range order entries are
  // skipped, so a malformed plan cannot address outside the system
  auto par = [&](int64_t i) -> int64_t {
    const int64_t p = parent[i];
    return (p >= 0 && p < K) ? p : -1;
  };
  for (int64_t i = 0; i < K; ++i) {
    const int64_t p = par(i);
    if (p >= 0)
      for (int c = 0; c < 3; ++c) w(p, c * 4) += E[i * 3 + c];
  }
  for (int64_t n = 0; n < K; ++n) {
    const int64_t i = order[n];
    if (i < 0 || i >= K) continue;
    T a[9], iv[9];
    for (int c = 0; c < 9; ++c) a[c] = w(i, c);
    inv3(a, iv);
    for (int c = 0; c < 9; ++c) w(i, c) = iv[c];
    const int64_t p = par(i);
    if (p >= 0) {
      const T e[3] = {E[i * 3], E[i * 3 + 1], E[i * 3 + 2]};
      T b[3], y[3];
      for (int c = 0; c < 3; ++c) b[c] = w(i, 9 + c);
      mv3(iv, b, y);
      for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) w(p, r * 3 + c) -= e[r] * iv[r * 3 + c] * e[c];
        w(p, 9 + r) += e[r] * y[r];
      }
    }
  }
  for (int64_t n = K - 1; n >= 0; --n) {
    const int64_t i = order[n];
    if (i < 0 || i >= K) continue;
    const int64_t p = par(i);
    T b[3], iv[9], xi[3];
    for (int c = 0; c < 3; ++c)
      b[c] = w(i, 9 + c) + (p >= 0 ? E[i * 3 + c] * X[p * 3 + c] : T(0));
    for (int c = 0; c < 9; ++c) iv[c] = w(i, c);
    mv3(iv, b, xi);
    for (int c = 0; c < 3; ++c) X[i * 3 + c] = xi[c];
  }
}

// Layer-parallel block tree solve: one wavefront per system, the nodes of
// each DHS layer chunk (layer_ptr over `order`, all of one depth) in
// parallel.  Each node pulls its children's eliminated terms through the
// children CSR (ascending child index), so no two lanes write one node:
//   A_i = main_i + diag(e_i) + sum_c [diag(e_c) - diag(e_c) Ainv_c diag(e_c)]
//   b_i = rhs_i + sum_c e_c * (Ainv_c b_c)
// then top-down x_i = Ainv_i (b_i + e_i * x_p).  work: [B][K][12] = Ainv (9)
// and the eliminated right-hand side (3) per node.
//
// RING: when every layer is one whole depth level of at most 64 nodes (each
// node's parent in the next layer, checked at staging), a node's children
// are exactly the previous layer, so their eliminated terms pass through a
// two-layer LDS ring indexed by position in the layer instead of the
// workspace, each lane owns at most one node per layer, and the next layer's
// inputs are fetched while the current one is eliminated.  A ring slot holds
// e_c-scaled rows of Ainv_c (9), e_c (3) and Ainv_c b_c (3), from which the
// parent forms the same Schur terms as the workspace form (the two agree to
// rounding).  Layers that fail the check take the workspace form.
constexpr int kRingW = 64;     // nodes per layer the ring holds
constexpr int kRingSlot = 15;  // values per node in the elimination ring
template <typename T, bool RING>
__global__ __launch_bounds__(256) void bt3_tree_layer_kernel(
    const T* __restrict__ main, const T* __restrict__ edge, const T* __restrict__ rhs,
    const int64_t* __restrict__ parent, const int64_t* __restrict__ order,
    const int64_t* __restrict__ layer_ptr, int n_layers, const int32_t* __restrict__ child_ptr,
    const int32_t* __restrict__ child_idx, T* __restrict__ x, T* __restrict__ work,
    int64_t B, int64_t K, int topo_lds, const T* __restrict__ gmem) {
  // G systems per workgroup (one wavefront each) walk the same layers, so the
  // workgroup barriers line up; the shared topology is staged once in LDS as
  // int32 (order, layer bounds, children CSR, parent) when it fits.
  extern __shared__ __align__(16) unsigned char dn_smem_raw[];
  const int lane = threadIdx.x & 63;
  const int G = blockDim.x >> 6;
  const int64_t s = (int64_t)blockIdx.x * G + (threadIdx.x >> 6);
  const bool live = s < B;
  int32_t* t_order = reinterpret_cast<int32_t*>(dn_smem_raw);
  int32_t* t_par = t_order + K;
  int32_t* t_cp = t_par + K;
  int32_t* t_lp = t_cp + K + 1;
  int32_t* t_ci = t_lp + n_layers + 1;
  if (topo_lds) {
    for (int64_t j = threadIdx.x; j < K; j += blockDim.x) {
      const int64_t o = order[j], p = parent[j];
      t_order[j] = (o >= 0 && o < K) ? (int32_t)o : -1;
      t_par[j] = (p >= 0 && p < K) ? (int32_t)p : -1;
    }
    for (int64_t j = threadIdx.x; j <= K; j += blockDim.x) t_cp[j] = child_ptr[j];
    for (int64_t j = threadIdx.x; j <= n_layers; j += blockDim.x) {
      const int64_t v = layer_ptr[j];
      t_lp[j] = (int32_t)(v < 0 ? 0 : (v > K ? K : v));
    }
    int32_t nc = child_ptr[K];
    if (RING && nc > K) nc = (int32_t)K;  // the ring check rejects such a CSR
    for (int64_t j = threadIdx.x; j < nc; j += blockDim.x) t_ci[j] = child_idx[j];
    __syncthreads();
  }
  auto bound = [&](int l) -> int64_t {
    if (topo_lds) return t_lp[l];
    const int64_t v = layer_ptr[l];
    return v < 0 ? 0 : (v > K ? K : v);
  };
  auto node = [&](int64_t n) -> int64_t {
    if (topo_lds) return t_order[n];
    const int64_t i = order[n];
    return (i >= 0 && i < K) ? i : -1;
  };
  auto par = [&](int64_t i) -> int64_t {
    if (topo_lds) return t_par[i];
    const int64_t p = parent[i];
    return (p >= 0 && p < K) ? p : -1;
  };
  const int64_t so = live ? s : 0;
  const T* M = main + so * K * 9;
  const T* E = edge + so * K * 3;
  const T* R = rhs + so * K * 3;
  const T* Gm = gmem ? gmem + so * K : nullptr;
  T* X = x + so * K * 3;
  T* W = work + so * K * 12;
  if (RING && topo_lds) {
    // inverse of order (t_inv, after K entries reserved for the children
    // list), then the checks that make the ring exact
    int32_t* t_inv = t_ci + K;
    for (int64_t j = threadIdx.x; j < K; j += blockDim.x) t_inv[j] = -1;
    __syncthreads();
    for (int64_t n = threadIdx.x; n < K; n += blockDim.x) {
      const int32_t i = t_order[n];
      if (i >= 0) t_inv[i] = (int32_t)n;
    }
    __syncthreads();
    int ok = 1;
    if (threadIdx.x == 0 && (t_lp[0] != 0 || t_lp[n_layers] != K || t_cp[0] != 0 ||
                             t_cp[K] < 0 || t_cp[K] > K))
      ok = 0;
    for (int l = threadIdx.x; l < n_layers; l += blockDim.x)
      if (t_lp[l + 1] < t_lp[l] || t_lp[l + 1] - t_lp[l] > kRingW) ok = 0;
    for (int64_t i = threadIdx.x; i < K; i += blockDim.x) {
      const int32_t n = t_inv[i];
      if (n < 0 || t_order[n] != i) { ok = 0; continue; }
      const int32_t c0 = t_cp[i], c1 = t_cp[i + 1];
      if (c0 < 0 || c1 < c0 || c1 > K) { ok = 0; continue; }
      for (int32_t j = c0; j < c1; ++j) {
        const int32_t ch = t_ci[j];
        if (ch < 0 || ch >= K || t_par[ch] != i) ok = 0;
      }
      const int32_t p = t_par[i];
      if (p < 0) continue;
      // layer of n: the last l with t_lp[l] <= n
      int lo = 0, hi = n_layers;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (t_lp[mid] <= n) lo = mid; else hi = mid;
      }
      const int32_t np = t_inv[p];
      if (lo + 1 >= n_layers || np < t_lp[lo + 1] || np >= t_lp[lo + 2]) ok = 0;
    }
    if (__syncthreads_and(ok)) {
      // Children and parents renamed to their positions in their layers
      // (the workspace form, which needs the node indices, is not taken
      // from here on).
      for (int64_t i = threadIdx.x; i < K; i += blockDim.x) {
        const int32_t n = t_inv[i];
        int lo = 0, hi = n_layers;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (t_lp[mid] <= n) lo = mid; else hi = mid;
        }
        for (int32_t j = t_cp[i]; j < t_cp[i + 1]; ++j) t_ci[j] = t_inv[t_ci[j]] - t_lp[lo - 1];
        const int32_t p = t_par[i];
        t_par[i] = p >= 0 ? t_inv[p] - t_lp[lo + 1] : -1;
      }
      __syncthreads();
      // rings of this wavefront: elimination [2][kRingW][kRingSlot], x [2][kRingW][3],
      // after the staged topology (5K + n_layers + 2 int32, rounded to 16 B)
      const size_t ring_off = ((size_t)4 * (5 * K + n_layers + 2) + 15) & ~(size_t)15;
      T* rS = reinterpret_cast<T*>(dn_smem_raw + ring_off) +
              (threadIdx.x >> 6) * (2 * kRingW * (kRingSlot + 3));
      T* rX = rS + 2 * kRingW * kRingSlot;
      // node of this lane in layer l, or -1 (branch-free: the read is
      // clamped into the staged order and its value dropped when unused)
      auto node_of = [&](int l) -> int32_t {
        const int lc = l < 0 ? 0 : (l >= n_layers ? n_layers - 1 : l);
        const int32_t lo = t_lp[lc], w = t_lp[lc + 1] - lo;
        const int32_t at = lo + (lane < w ? lane : (w > 0 ? w - 1 : 0));
        const int32_t v = t_order[at < K ? at : K - 1];
        return (l >= 0 && l < n_layers && lane < w) ? v : -1;
      };
      // Bottom-up.  Two register sets alternate (the loop is unrolled by
      // two), so the next layer's loads stay in flight while this layer is
      // eliminated.
      // Loads are issued by every lane (a lane without a node reads node 0 of
      // its system, or of system 0 past the batch, and ignores it): with no
      // branch around them the compiler can leave them in flight.
      auto load_fwd = [&](int32_t i, T (&a)[9], T (&e)[3], T (&b)[3]) {
        const int64_t k = i >= 0 ? i : 0;
        for (int c = 0; c < 9; ++c) a[c] = M[k * 9 + c];
        if (Gm) add_membrane(a, Gm[k]);
        for (int c = 0; c < 3; ++c) { e[c] = E[k * 3 + c]; b[c] = R[k * 3 + c]; }
      };
      auto elim = [&](int l, int32_t i, T (&a)[9], const T (&e)[3], T (&b)[3]) {
        if (live && i >= 0) {
          for (int c = 0; c < 3; ++c) a[c * 4] += e[c];
          const T* prev = rS + ((l + 1) & 1) * (kRingW * kRingSlot);
          for (int32_t j = t_cp[i]; j < t_cp[i + 1]; ++j) {
            const T* q = prev + t_ci[j] * kRingSlot;
            for (int r = 0; r < 3; ++r) {
              a[r * 4] += q[9 + r];
              for (int c = 0; c < 3; ++c) a[r * 3 + c] -= q[r * 3 + c] * q[9 + c];
              b[r] += q[9 + r] * q[12 + r];
            }
          }
          T iv[9], y[3];
          inv3(a, iv);
          for (int c = 0; c < 9; ++c) W[(int64_t)i * 12 + c] = iv[c];
          for (int c = 0; c < 3; ++c) W[(int64_t)i * 12 + 9 + c] = b[c];
          mv3(iv, b, y);
          T* q = rS + (l & 1) * (kRingW * kRingSlot) + lane * kRingSlot;
          for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) q[r * 3 + c] = e[r] * iv[r * 3 + c];
            q[9 + r] = e[r];
            q[12 + r] = y[r];
          }
        }
        // The rings belong to this wavefront alone and its LDS operations
        // complete in issue order: a wavefront-scope fence orders them for
        // the compiler without waiting for the prefetched global loads.
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      };
      {
        T a0[9], e0[3], b0[3], a1[9], e1[3], b1[3];
        int32_t i0 = node_of(0), i1;
        load_fwd(i0, a0, e0, b0);
        for (int l = 0; l < n_layers; l += 2) {
          i1 = node_of(l + 1);
          load_fwd(i1, a1, e1, b1);
          elim(l, i0, a0, e0, b0);
          if (l + 1 >= n_layers) break;
          i0 = node_of(l + 2);
          load_fwd(i0, a0, e0, b0);
          elim(l + 1, i1, a1, e1, b1);
        }
      }
      // Top-down, the same alternation over (Ainv, b, e).
      auto load_bwd = [&](int32_t i, T (&w)[12], T (&e)[3]) {
        const int64_t k = i >= 0 ? i : 0;
        for (int c = 0; c < 12; ++c) w[c] = W[k * 12 + c];
        for (int c = 0; c < 3; ++c) e[c] = E[k * 3 + c];
      };
      auto back = [&](int l, int32_t i, const T (&w)[12], const T (&e)[3]) {
        if (live && i >= 0) {
          const int32_t p = t_par[i];  // the parent's position in layer l + 1
          const T* xp = rX + ((l + 1) & 1) * (kRingW * 3) + (p >= 0 ? p : 0) * 3;
          T bb[3], xi[3];
          for (int c = 0; c < 3; ++c) bb[c] = w[9 + c] + (p >= 0 ? e[c] * xp[c] : T(0));
          mv3(w, bb, xi);
          for (int c = 0; c < 3; ++c) {
            X[(int64_t)i * 3 + c] = xi[c];
            rX[(l & 1) * (kRingW * 3) + lane * 3 + c] = xi[c];
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      };
      {
        T w0[12], e0[3], w1[12], e1[3];
        int32_t i0 = node_of(n_layers - 1), i1;
        load_bwd(i0, w0, e0);
        for (int l = n_layers - 1; l >= 0; l -= 2) {
          i1 = node_of(l - 1);
          load_bwd(i1, w1, e1);
          back(l, i0, w0, e0);
          if (l - 1 < 0) break;
          i0 = node_of(l - 2);
          load_bwd(i0, w0, e0);
          back(l - 1, i1, w1, e1);
        }
      }
      return;
    }
  }
  for (int l = 0; l < n_layers; ++l) {
    int64_t lo = bound(l), hi = bound(l + 1);
    if (hi < lo) hi = lo;
    for (int64_t n = lo + lane; live && n < hi; n += 64) {
      const int64_t i = node(n);
      if (i < 0) continue;
      T a[9], b[3];
      for (int c = 0; c < 9; ++c) a[c] = M[i * 9 + c];
      if (Gm) add_membrane(a, Gm[i]);
      for (int c = 0; c < 3; ++c) {
        a[c * 4] += E[i * 3 + c];
        b[c] = R[i * 3 + c];
      }
      const int32_t c0 = topo_lds ? t_cp[i] : child_ptr[i];
      const int32_t c1 = topo_lds ? t_cp[i + 1] : child_ptr[i + 1];
      for (int32_t j = c0; j < c1; ++j) {
        const int64_t ch = topo_lds ? t_ci[j] : child_idx[j];
        const T* Wc = W + ch * 12;
        const T e[3] = {E[ch * 3], E[ch * 3 + 1], E[ch * 3 + 2]};
        T iv[9], y[3];
        for (int c = 0; c < 9; ++c) iv[c] = Wc[c];
        mv3(iv, Wc + 9, y);
        for (int r = 0; r < 3; ++r) {
          a[r * 4] += e[r];
          for (int c = 0; c < 3; ++c) a[r * 3 + c] -= e[r] * iv[r * 3 + c] * e[c];
          b[r] += e[r] * y[r];
        }
      }
      T iv[9];
      inv3(a, iv);
      for (int c = 0; c < 9; ++c) W[i * 12 + c] = iv[c];
      for (int c = 0; c < 3; ++c) W[i * 12 + 9 + c] = b[c];
    }
    __syncthreads();
  }
  for (int l = n_layers - 1; l >= 0; --l) {
    int64_t lo = bound(l), hi = bound(l + 1);
    if (hi < lo) hi = lo;
    for (int64_t n = lo + lane; live && n < hi; n += 64) {
      const int64_t i = node(n);
      if (i < 0) continue;
      const int64_t p = par(i);
      T b[3], xi[3];
      for (int c = 0; c < 3; ++c)
        b[c] = W[i * 12 + 9 + c] + (p >= 0 ? E[i * 3 + c] * X[p * 3 + c] : T(0));
      mv3(W + i * 12, b, xi);
      for (int c = 0; c < 3; ++c) X[i * 3 + c] = xi[c];
    }
    __syncthreads();
  }
}


// ---------------------------------------------------------------------------
// block tree solve by contraction (the 3x3-block form of the scalar tree
// contraction): one workgroup per system over a host plan of sections
// (chains of at most S+1 nodes, head at the top, tail where the child
// sections attach) and rake/compress rounds over the section tree.
//   A_ii = main_i (+ membrane) + diag(e_i) + sum_children diag(e_c),
//   A_ip = A_pi = -diag(e_i)   (main blocks need not be symmetric, so the
//   two couplings U = A_{j,p}, L = A_{p,j} of every eliminated pair are
//   carried separately).
// 1. each section eliminates its nodes but the tail, head first; the tail
//    keeps a self block, a right-hand side and couplings to the parent
//    section's tail, which receives PM, PF; interior recovery
//    x_q = a_q + B_q x_ptail + C_q x_{q+1} goes to the workspace
//    ([B][K][21], by section position);
// 2. rounds: a rake folds a leaf section into its parent, a compress joins
//    its parent and child (new couplings CU, CL); survivors pull their
//    eliminated neighbours' terms in a fixed order;
// 3. back-substitution: the root, the rounds in reverse, every interior.
// A path of n sections takes about log2(n) rounds, so the dependent chain
// is the longest section plus twice the rounds, not the tree's depth.
// ---------------------------------------------------------------------------
template <typename T> struct BcArgs {
  const T *main, *edge, *rhs, *gmem;
  T* x;
  T* work;  // interior recovery terms, [B][K][21] by section position
  int64_t K;
  int n_sec, n_rounds;
  const int32_t *perm, *sec_head, *sec_len, *sec_child, *sec_nchild, *sec_parent;
  const int32_t *elim_ptr, *elim_sec, *elim_par, *elim_child;
  const int32_t *recv_ptr, *recv_sec, *src_ptr, *src;
};
constexpr int kBcSlots = 45;  // LDS values per section
constexpr int kBcThreads = 128;

template <typename T>
__device__ __forceinline__ void mm3(const T* a, const T* b, T* c) {
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k)
      c[r * 3 + k] = a[r * 3] * b[k] + a[r * 3 + 1] * b[3 + k] + a[r * 3 + 2] * b[6 + k];
}

// node block of position q: main (+ membrane) + diag(own edge)
template <typename T>
__device__ __forceinline__ void bc_node(const BcArgs<T>& A, const T* M, const T* E,
                                        const T* G, int k, T* S, T* e) {
  for (int c = 0; c < 9; ++c) S[c] = M[(int64_t)k * 9 + c];
  if (G) add_membrane(S, G[k]);
  for (int c = 0; c < 3; ++c) {
    e[c] = E[(int64_t)k * 3 + c];
    S[c * 4] += e[c];
  }
}

template <typename T>
__global__ __launch_bounds__(kBcThreads) void bt3_contract_kernel(BcArgs<T> A) {
  extern __shared__ __align__(16) unsigned char dn_smem_raw[];
  const int ns = A.n_sec;
  T* L0 = reinterpret_cast<T*>(dn_smem_raw);
  // per section: self block SM, right-hand side SF, couplings SU = A_{j,p}
  // and SL = A_{p,j} to the parent section's tail, what it hands its parent
  // (PM, PF), its tail's solution XS.  A section eliminated in a round keeps
  // its recovery terms x_j = RF + RM x_p + RE x_c in SF, SU, SL (no longer
  // read once it is eliminated), so RF/RM/RE alias them.
  T *SM = L0, *SF = SM + 9 * ns, *SU = SF + 3 * ns, *SL = SU + 9 * ns, *PM = SL + 9 * ns,
    *PF = PM + 9 * ns, *XS = PF + 3 * ns;
  T *RF = SF, *RM = SU, *RE = SL;
  const int64_t K = A.K, b = blockIdx.x;
  const int tid = threadIdx.x, nt = blockDim.x;
  const T* M = A.main + b * K * 9;
  const T* E = A.edge + b * K * 3;
  const T* R = A.rhs + b * K * 3;
  const T* G = A.gmem ? A.gmem + b * K : nullptr;
  T* X = A.x + b * K * 3;
  // interior recovery terms by section position, 21 contiguous values per
  // node (each lane's writes and reads of a node fall in one or two lines;
  // component-major layouts across the lanes measured slower)
  T* W = A.work + b * K * 21;

  // 1. chain condensation, head first
  for (int s = tid; s < ns; s += nt) {
    const int hd = A.sec_head[s], len = A.sec_len[s];
    T S[9], e[3], F[3], U[9], Lc[9], dS[9], dF[3];
    int k = A.perm[hd];
    bc_node(A, M, E, G, k, S, e);
    for (int c = 0; c < 3; ++c) F[c] = R[(int64_t)k * 3 + c];
    for (int c = 0; c < 9; ++c) U[c] = Lc[c] = dS[c] = T(0);
    for (int c = 0; c < 3; ++c) {
      U[c * 4] = Lc[c * 4] = -e[c];
      dF[c] = T(0);
    }
    // the next node's block, edge and right-hand side are fetched a node
    // ahead (index clamped into the section: the loads are unconditional)
    const int last = hd + len - 1;
    T pS[9], pe[3], pr[3];
    auto fetch = [&](int q) {
      const int n = A.perm[q];
      bc_node(A, M, E, G, n, pS, pe);
      for (int c = 0; c < 3; ++c) pr[c] = R[(int64_t)n * 3 + c];
    };
    fetch(hd + 1 < last ? hd + 1 : last);
    for (int q = hd; q < last; ++q) {
      T Sn[9], en[3], rn[3];
      for (int c = 0; c < 9; ++c) Sn[c] = pS[c];
      for (int c = 0; c < 3; ++c) {
        en[c] = pe[c];
        rn[c] = pr[c];
      }
      fetch(q + 2 < last ? q + 2 : last);      for (int c = 0; c < 3; ++c) S[c * 4] += en[c];  // the chain child's edge
      T Si[9], a[3], SiU[9], LS[9], t[9];
      inv3(S, Si);
      mv3(Si, F, a);
      mm3(Si, U, SiU);
      mm3(Lc, Si, LS);
      // recovery x_q = a + B x_ptail + C x_n with B = -Si U, C = Si diag(en)
      T* w = W + (int64_t)q * 21;
      for (int c = 0; c < 3; ++c) w[c] = a[c];
      for (int c = 0; c < 9; ++c) w[3 + c] = -SiU[c];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) w[12 + r * 3 + c] = Si[r * 3 + c] * en[c];
      // the parent tail: dS -= L Si U, dF -= L Si F; new A_{pt,n} = L Si diag(en)
      mm3(LS, U, t);
      for (int c = 0; c < 9; ++c) dS[c] -= t[c];
      for (int r = 0; r < 3; ++r)
        dF[r] -= LS[r * 3] * F[0] + LS[r * 3 + 1] * F[1] + LS[r * 3 + 2] * F[2];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Lc[r * 3 + c] = LS[r * 3 + c] * en[c];
      // the next node: S_n -= diag(en) Si diag(en), F_n += en (Si F),
      // U_n = diag(en) Si U
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          Sn[r * 3 + c] -= en[r] * Si[r * 3 + c] * en[c];
          U[r * 3 + c] = en[r] * SiU[r * 3 + c];
        }
      for (int c = 0; c < 9; ++c) S[c] = Sn[c];
      for (int c = 0; c < 3; ++c) F[c] = rn[c] + en[c] * a[c];
    }
    // the tail's children are the heads of the child sections
    const int c0 = A.sec_child[s], nc = A.sec_nchild[s];
    for (int c = c0; c < c0 + nc; ++c) {
      const int h = A.perm[A.sec_head[c]];
      for (int r = 0; r < 3; ++r) S[r * 4] += E[(int64_t)h * 3 + r];
    }
    for (int c = 0; c < 9; ++c) {
      SM[s * 9 + c] = S[c];
      SU[s * 9 + c] = U[c];
      SL[s * 9 + c] = Lc[c];
      PM[s * 9 + c] = dS[c];
    }
    for (int c = 0; c < 3; ++c) {
      SF[s * 3 + c] = F[c];
      PF[s * 3 + c] = dF[c];
    }
  }
  __syncthreads();
  for (int s = tid; s < ns; s += nt) {  // the child sections' heads act on the tail
    const int c0 = A.sec_child[s], nc = A.sec_nchild[s];
    for (int c = c0; c < c0 + nc; ++c) {
      for (int v = 0; v < 9; ++v) SM[s * 9 + v] += PM[c * 9 + v];
      for (int v = 0; v < 3; ++v) SF[s * 3 + v] += PF[c * 3 + v];
    }
  }
  __syncthreads();

  // 2. contraction rounds over the section tree.  A compressed section's
  // thread also updates its child directly: the child survives the round,
  // only this thread touches it before the barrier, and the order of the
  // additions to it is fixed by the plan (this one first, then the pulls).
  for (int r = 0; r < A.n_rounds; ++r) {
    for (int i = A.elim_ptr[r] + tid; i < A.elim_ptr[r + 1]; i += nt) {
      const int j = A.elim_sec[i], c = A.elim_child[i];
      T Si[9], LS[9], t[9], f[3], Uj[9], Fj[3];
      inv3(SM + j * 9, Si);
      mm3(SL + j * 9, Si, LS);
      for (int v = 0; v < 9; ++v) Uj[v] = SU[j * 9 + v];
      for (int v = 0; v < 3; ++v) Fj[v] = SF[j * 3 + v];
      mm3(LS, Uj, t);
      for (int v = 0; v < 9; ++v) PM[j * 9 + v] = -t[v];
      mv3(LS, Fj, f);
      for (int v = 0; v < 3; ++v) PF[j * 3 + v] = -f[v];
      T re[9];
      if (c < 0) {  // rake
        for (int v = 0; v < 9; ++v) re[v] = T(0);
      } else {  // compress: A_{c,j} = U_c, A_{j,c} = L_c
        T US[9], Lc[9], Uc[9];
        for (int v = 0; v < 9; ++v) {
          Lc[v] = SL[c * 9 + v];
          Uc[v] = SU[c * 9 + v];
        }
        mm3(Si, Lc, t);
        for (int v = 0; v < 9; ++v) re[v] = -t[v];
        mm3(Uc, Si, US);
        mm3(US, Lc, t);
        for (int v = 0; v < 9; ++v) SM[c * 9 + v] -= t[v];
        mv3(US, Fj, f);
        for (int v = 0; v < 3; ++v) SF[c * 3 + v] -= f[v];
        mm3(US, Uj, t);
        for (int v = 0; v < 9; ++v) SU[c * 9 + v] = -t[v];  // new A_{c,p}
        mm3(LS, Lc, t);
        for (int v = 0; v < 9; ++v) SL[c * 9 + v] = -t[v];  // new A_{p,c}
      }
      // recovery terms into the eliminated section's own slots
      mv3(Si, Fj, f);
      mm3(Si, Uj, t);
      for (int v = 0; v < 3; ++v) RF[j * 3 + v] = f[v];
      for (int v = 0; v < 9; ++v) {
        RM[j * 9 + v] = -t[v];
        RE[j * 9 + v] = re[v];
      }
    }
    __syncthreads();
    for (int i = A.recv_ptr[r] + tid; i < A.recv_ptr[r + 1]; i += nt) {
      const int q = A.recv_sec[i];
      T m[9], f[3];
      for (int v = 0; v < 9; ++v) m[v] = SM[q * 9 + v];
      for (int v = 0; v < 3; ++v) f[v] = SF[q * 3 + v];
      for (int k = A.src_ptr[i]; k < A.src_ptr[i + 1]; ++k) {
        const int sj = A.src[k];
        if (sj >= 0) {  // an eliminated child (a compressed parent updated q already)
          for (int v = 0; v < 9; ++v) m[v] += PM[sj * 9 + v];
          for (int v = 0; v < 3; ++v) f[v] += PF[sj * 3 + v];
        }
      }
      for (int v = 0; v < 9; ++v) SM[q * 9 + v] = m[v];
      for (int v = 0; v < 3; ++v) SF[q * 3 + v] = f[v];
    }
    __syncthreads();
  }

  // 3. back-substitution: root (section 0), rounds in reverse, interiors
  if (tid == 0) {
    T Si[9], xr[3];
    inv3(SM, Si);
    mv3(Si, SF, xr);
    for (int v = 0; v < 3; ++v) XS[v] = xr[v];
  }
  __syncthreads();
  for (int r = A.n_rounds - 1; r >= 0; --r) {
    for (int i = A.elim_ptr[r] + tid; i < A.elim_ptr[r + 1]; i += nt) {
      const int j = A.elim_sec[i], p = A.elim_par[i], c = A.elim_child[i];
      T xj[3], t[3];
      mv3(RM + j * 9, XS + p * 3, t);
      for (int v = 0; v < 3; ++v) xj[v] = RF[j * 3 + v] + t[v];
      if (c >= 0) {
        mv3(RE + j * 9, XS + c * 3, t);
        for (int v = 0; v < 3; ++v) xj[v] += t[v];
      }
      for (int v = 0; v < 3; ++v) XS[j * 3 + v] = xj[v];
    }
    __syncthreads();
  }
  for (int s = tid; s < ns; s += nt) {
    const int hd = A.sec_head[s], len = A.sec_len[s], ps = A.sec_parent[s];
    T xp[3], xv[3];
    for (int v = 0; v < 3; ++v) {
      xp[v] = ps >= 0 ? XS[ps * 3 + v] : T(0);
      xv[v] = XS[s * 3 + v];
    }
    int k = A.perm[hd + len - 1];
    for (int v = 0; v < 3; ++v) X[(int64_t)k * 3 + v] = xv[v];
    // each interior's recovery terms are fetched a node ahead (clamped)
    T pw[21];
    auto fetch_w = [&](int q) {
      for (int v = 0; v < 21; ++v) pw[v] = W[(int64_t)q * 21 + v];
    };
    if (len > 1) fetch_w(hd + len - 2);
    for (int q = hd + len - 2; q >= hd; --q) {
      T w[21], t1[3], t2[3];
      for (int v = 0; v < 21; ++v) w[v] = pw[v];
      fetch_w(q > hd ? q - 1 : hd);
      mv3(w + 3, xp, t1);
      mv3(w + 12, xv, t2);
      for (int v = 0; v < 3; ++v) xv[v] = w[v] + t1[v] + t2[v];
      k = A.perm[q];
      for (int v = 0; v < 3; ++v) X[(int64_t)k * 3 + v] = xv[v];
    }
  }
}

}  // namespace dn

// one wavefront per workgroup for the lane-per-system sweeps: a few thousand
// systems then spread over as many compute units as they have wavefronts
constexpr int kSweepThreads = 64;
template <typename T>
static unsigned blocks_for(int64_t B) {
  return (unsigned)((B + kSweepThreads - 1) / kSweepThreads);
}

// LDS of the block contraction: kBcSlots values per section.
constexpr size_t kBcLdsMax = 160 * 1024;

// most sections one system may have (K does not bound the LDS)
DN_EXPORT int dn_bt3_contract_max_sections(int elem_bytes, int64_t K) {
  if ((elem_bytes != 4 && elem_bytes != 8) || K < 0) return -1;
  return (int)(kBcLdsMax / ((size_t)dn::kBcSlots * elem_bytes));
}

DN_EXPORT int64_t dn_bt3_contract_work_elems(int64_t B, int64_t K) {
  if (B < 0 || K < 0) return -1;
  return 21 * B * K;
}

template <typename T>
static int bt3_contract_impl(const dn::BcArgs<T>& a, int64_t B, int64_t work_elems,
                             hipStream_t stream) {
  if (!a.main || !a.edge || !a.rhs || !a.x || !a.work || B < 1 || a.K < 1 || a.n_sec < 1 ||
      a.n_rounds < 0 || !a.perm || !a.sec_head || !a.sec_len || !a.sec_child ||
      !a.sec_nchild || !a.sec_parent || !a.elim_ptr || !a.elim_sec || !a.elim_par ||
      !a.elim_child || !a.recv_ptr || !a.recv_sec || !a.src_ptr || !a.src)
    return DN_EINVAL;
  if (work_elems < 21 * B * a.K || B > 0x7fffffff || a.K > 0x7fffffff) return DN_EINVAL;
  const size_t lds = (size_t)dn::kBcSlots * a.n_sec * sizeof(T);
  if (lds > kBcLdsMax) return DN_ERANGE;
  hipLaunchKernelGGL((dn::bt3_contract_kernel<T>), dim3((unsigned)B), dim3(dn::kBcThreads),
                     lds, stream, a);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

#define DN_BT3_CONTRACT_EXPORT(T, SFX)                                                       \
  DN_EXPORT int dn_bt3_tree_contract_##SFX(                                                  \
      const T* main, const T* edge, const T* rhs, const T* gmem, T* x, int64_t B, int64_t K, \
      const int32_t* perm, int n_sec, const int32_t* sec_head, const int32_t* sec_len,       \
      const int32_t* sec_child, const int32_t* sec_nchild, const int32_t* sec_parent,        \
      int n_rounds, const int32_t* elim_ptr, const int32_t* elim_sec,                        \
      const int32_t* elim_par, const int32_t* elim_child, const int32_t* recv_ptr,           \
      const int32_t* recv_sec, const int32_t* src_ptr, const int32_t* src, T* work,          \
      int64_t work_elems, hipStream_t stream) {                                              \
    dn::BcArgs<T> a{main,     edge,     rhs,      gmem,       x,          work,              \
                    K,        n_sec,    n_rounds, perm,       sec_head,   sec_len,           \
                    sec_child, sec_nchild, sec_parent, elim_ptr, elim_sec, elim_par,          \
                    elim_child, recv_ptr, recv_sec, src_ptr, src};                           \
    return bt3_contract_impl<T>(a, B, work_elems, stream);                                   \
  }
DN_BT3_CONTRACT_EXPORT(float, f32)
DN_BT3_CONTRACT_EXPORT(double, f64)

DN_EXPORT int64_t dn_block_work_elems(int kind, int64_t B, int64_t K) {
  if (B < 0 || K < 0) return -1;
  return (kind == 1 ? 3 : 12) * B * K;  // 0: Thomas, 1: SPD, 2: tree
}

template <typename T>
static int bt3_thomas_impl(const T* lower, const T* main, const T* upper, const T* rhs,
                           T* x, int64_t B, int64_t K, T* work, int64_t work_elems,
                           hipStream_t stream) {
  if (!main || !rhs || !x || !work || B < 1 || K < 1) return DN_EINVAL;
  if (K > 1 && (!lower || !upper)) return DN_EINVAL;
  if (work_elems < 12 * B * K || B > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  hipLaunchKernelGGL((dn::bt3_thomas_kernel<T>), dim3(blocks_for<T>(B)), dim3(kSweepThreads), 0,
                     stream, lower, main, upper, rhs, x, work, B, K);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
static int bt3_spd_impl(const T* lower, const T* main, const T* upper, const T* rhs, T* x,
                        T* chol, int32_t* info, int64_t B, int64_t K, T* work,
                        int64_t work_elems, int factor, T eps_rel, T eps_abs,
                        hipStream_t stream) {
  if (!rhs || !x || !chol || !work || B < 1 || K < 1) return DN_EINVAL;
  if (factor && !main) return DN_EINVAL;
  if (K > 1 && (!lower || !upper)) return DN_EINVAL;
  if (work_elems < 3 * B * K || B > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  hipLaunchKernelGGL((dn::bt3_spd_kernel<T>), dim3(blocks_for<T>(B)), dim3(kSweepThreads), 0, stream,
                     lower, main, upper, rhs, x, chol, info, work, B, K, eps_rel, eps_abs,
                     factor);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
static int bt3_tree_impl(const T* main, const T* edge, const T* rhs, const int64_t* parent,
                         const int64_t* order, const int64_t* layer_ptr, int64_t n_layers,
                         const int32_t* child_ptr, const int32_t* child_idx, T* x, int64_t B,
                         int64_t K, T* work, int64_t work_elems, hipStream_t stream,
                         const T* gmem = nullptr) {
  if (!main || !edge || !rhs || !parent || !order || !x || !work || B < 1 || K < 1)
    return DN_EINVAL;
  if (work_elems < 12 * B * K || B > (int64_t)INT32_MAX * 256 || n_layers < 0 ||
      n_layers > INT32_MAX)
    return DN_EINVAL;
  // Layer-parallel form when the layers average at least two nodes and the
  // children CSR is given; else one lane per system walks `order`.
  if (layer_ptr && child_ptr && child_idx && n_layers >= 1 && 2 * n_layers <= K &&
      B <= INT32_MAX) {
    // systems per workgroup: enough workgroups (~512) to fill the chip first
    const int G = B >= 2048 ? 4 : (B >= 1024 ? 2 : 1);
    const size_t topo = (size_t)4 * (3 * K + 1 + n_layers + 1 + (K > 0 ? K : 1));
    const int topo_lds = topo <= 96 * 1024;
    // ring form: the topology plus the inverse order, then per wavefront the
    // two-layer elimination and solution rings (the kernel decides at
    // staging whether the layers allow it)
    const size_t ring_topo = ((topo + (size_t)4 * K) + 15) & ~(size_t)15;
    const size_t ring = ring_topo + (size_t)G * 2 * dn::kRingW * (dn::kRingSlot + 3) * sizeof(T);
    // the ring form's wavefronts never wait on each other after staging, so
    // four systems share a workgroup (and its staging) from B >= 256 on
    const int Gr = B >= 256 ? 4 : G;
    const size_t ring4 = ring_topo + (size_t)Gr * 2 * dn::kRingW * (dn::kRingSlot + 3) * sizeof(T);
    if (ring4 <= 96 * 1024) {
      hipLaunchKernelGGL((dn::bt3_tree_layer_kernel<T, true>), dim3((unsigned)((B + Gr - 1) / Gr)),
                         dim3(64 * Gr), ring4, stream, main, edge, rhs, parent, order, layer_ptr,
                         (int)n_layers, child_ptr, child_idx, x, work, B, K, 1, gmem);
    } else if (ring <= 96 * 1024) {
      hipLaunchKernelGGL((dn::bt3_tree_layer_kernel<T, true>), dim3((unsigned)((B + G - 1) / G)),
                         dim3(64 * G), ring, stream, main, edge, rhs, parent, order, layer_ptr,
                         (int)n_layers, child_ptr, child_idx, x, work, B, K, 1, gmem);
    } else {
      hipLaunchKernelGGL((dn::bt3_tree_layer_kernel<T, false>),
                         dim3((unsigned)((B + G - 1) / G)), dim3(64 * G), topo_lds ? topo : 0,
                         stream, main, edge, rhs, parent, order, layer_ptr, (int)n_layers,
                         child_ptr, child_idx, x, work, B, K, topo_lds, gmem);
    }
  } else {
    hipLaunchKernelGGL((dn::bt3_tree_kernel<T>), dim3(blocks_for<T>(B)), dim3(kSweepThreads), 0,
                       stream, main, edge, rhs, parent, order, x, work, B, K, gmem);
  }
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int dn_bt3_thomas_f32(const float* lower, const float* main, const float* upper,
                                const float* rhs, float* x, int64_t B, int64_t K,
                                float* work, int64_t work_elems, hipStream_t stream) {
  return bt3_thomas_impl<float>(lower, main, upper, rhs, x, B, K, work, work_elems, stream);
}
DN_EXPORT int dn_bt3_thomas_f64(const double* lower, const double* main,
                                const double* upper, const double* rhs, double* x,
                                int64_t B, int64_t K, double* work, int64_t work_elems,
                                hipStream_t stream) {
  return bt3_thomas_impl<double>(lower, main, upper, rhs, x, B, K, work, work_elems, stream);
}
DN_EXPORT int dn_bt3_spd_f32(const float* lower, const float* main, const float* upper,
                             const float* rhs, float* x, float* chol, int32_t* info,
                             int64_t B, int64_t K, float* work, int64_t work_elems,
                             int factor, hipStream_t stream) {
  return bt3_spd_impl<float>(lower, main, upper, rhs, x, chol, info, B, K, work,
                             work_elems, factor, 1e-6f, 1e-18f, stream);
}
DN_EXPORT int dn_bt3_spd_f64(const double* lower, const double* main, const double* upper,
                             const double* rhs, double* x, double* chol, int32_t* info,
                             int64_t B, int64_t K, double* work, int64_t work_elems,
                             int factor, hipStream_t stream) {
  return bt3_spd_impl<double>(lower, main, upper, rhs, x, chol, info, B, K, work,
                              work_elems, factor, 1e-12, 1e-30, stream);
}
DN_EXPORT int dn_bt3_tree_f32(const float* main, const float* edge, const float* rhs,
                              const int64_t* parent, const int64_t* order,
                              const int64_t* layer_ptr, int64_t n_layers,
                              const int32_t* child_ptr, const int32_t* child_idx, float* x,
                              int64_t B, int64_t K, float* work, int64_t work_elems,
                              hipStream_t stream) {
  return bt3_tree_impl<float>(main, edge, rhs, parent, order, layer_ptr, n_layers, child_ptr,
                            child_idx, x, B, K, work, work_elems, stream);
}
DN_EXPORT int dn_bt3_tree_f64(const double* main, const double* edge, const double* rhs,
                              const int64_t* parent, const int64_t* order,
                              const int64_t* layer_ptr, int64_t n_layers,
                              const int32_t* child_ptr, const int32_t* child_idx, double* x,
                              int64_t B, int64_t K, double* work, int64_t work_elems,
                              hipStream_t stream) {
  return bt3_tree_impl<double>(main, edge, rhs, parent, order, layer_ptr, n_layers, child_ptr,
                            child_idx, x, B, K, work, work_elems, stream);
}

// ---------------------------------------------------------------------------
// ExtCell trees: the block tree backward-Euler step of _dhs_bt._step
// (dendra/models/integrators/tree_bt.py:504-654) for B HH trees of one
// topology, in node order.  Per step and node i (unknowns vc = (vi, ve0, ve1),
// v = vi - ve0):
//   gates by cnexp at v -> G, J;  g = G area, d = (J - G v) area + intra
//   rhs_i = (d, -d - xg0 (ve0 - ve1), xg0 (ve0 - ve1) + xg1 (e_bath - ve1))
//           - f_i + sum over children c (ascending) of f_c,
//   f_i = e_i * (vc_i - vc_parent)   (e_i * vc_i at a root)
//   D_i = main_i + [[g, -g, 0], [-g, g, 0], [0, 0, 0]]
// then the block tree solve (bt3_tree_layer_kernel / bt3_tree_kernel) of
// D, e, rhs for the increment, and vc += increment.  The assembly stores
// only g; the solve adds it to the constant main block as it loads it.
// ---------------------------------------------------------------------------
namespace dn {

template <typename T>
struct XtArgs {
  const T *main, *edge, *area, *xg, *sna, *sk, *hh, *lead, *amp, *wave, *iamp, *iwave;
  const int32_t *parent, *child_ptr, *child_idx;
  T *vc, *m, *h, *n, *D, *R;
  int64_t B, K;
  T dt;
};

template <typename T>
__global__ __launch_bounds__(256) void extcell_tree_assemble_kernel(XtArgs<T> a, int64_t s) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.B * a.K) return;
  const int64_t b = j / a.K, i = j - b * a.K;
  const HHParams<T> hp{a.hh[0], a.hh[1], a.hh[2], a.hh[3], a.hh[4], a.hh[5], a.hh[6]};
  const T* V = a.vc + b * a.K * 3;
  const T* E = a.edge + b * a.K * 3;
  const T vi[3] = {V[i * 3], V[i * 3 + 1], V[i * 3 + 2]};
  const T v = vi[0] - vi[1];
  T m = a.m[j], h = a.h[j], n = a.n[j], G, J;
  hh_advance<T>(v, a.dt, hp, a.sna ? a.sna[j] : T(1), a.sk ? a.sk[j] : T(1), m, h, n, G, J);
  a.m[j] = m;
  a.h[j] = h;
  a.n[j] = n;
  const T area = a.area[j];
  const T g = G * area;
  const T intra = (a.iamp && a.iwave) ? a.iamp[j] * a.iwave[s] : T(0);
  const T d = (J - G * v) * area + intra;
  const T bath = (a.lead && a.amp && a.wave) ? a.lead[j] * (a.amp[b] * a.wave[s]) : T(0);
  const T shell = a.xg[j * 2] * (vi[1] - vi[2]);
  const T bathf = a.xg[j * 2 + 1] * (bath - vi[2]);
  T r[3] = {d, -d - shell, shell + bathf};
  const int32_t p = a.parent[i];
#pragma unroll
  for (int c = 0; c < 3; ++c)
    r[c] -= E[i * 3 + c] * (p >= 0 ? vi[c] - V[p * 3 + c] : vi[c]);
  for (int32_t q = a.child_ptr[i]; q < a.child_ptr[i + 1]; ++q) {
    const int64_t ch = a.child_idx[q];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[c] += E[ch * 3 + c] * (V[ch * 3 + c] - vi[c]);
  }
  a.D[j] = g;  // the solve adds g to the constant main block
#pragma unroll
  for (int c = 0; c < 3; ++c) a.R[j * 3 + c] = r[c];
}

// one thread per node: vc += x, and the membrane potential vi - ve0 when recorded
template <typename T>
__global__ __launch_bounds__(256) void extcell_tree_update_kernel(T* __restrict__ vc,
                                                                  const T* __restrict__ x,
                                                                  int64_t n_nodes,
                                                                  T* __restrict__ rec) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_nodes) return;
  T u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    u[c] = vc[j * 3 + c] + x[j * 3 + c];
    vc[j * 3 + c] = u[c];
  }
  if (rec) rec[j] = u[0] - u[1];
}

}  // namespace dn

template <typename T>
static int extcell_tree_impl(const dn::XtArgs<T>& a0, const int32_t* parent32,
                             const int64_t* parent, const int64_t* order,
                             const int64_t* layer_ptr, int64_t n_layers, int64_t n_steps,
                             T* rec, int64_t rec_every, T* work, int64_t work_elems,
                             hipStream_t stream, const dn::BcArgs<T>* plan = nullptr) {
  const int64_t B = a0.B, K = a0.K;
  if (B < 1 || K < 1 || n_steps < 0 || B * K > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  if (!a0.main || !a0.edge || !a0.area || !a0.xg || !a0.hh || !a0.vc || !a0.m || !a0.h ||
      !a0.n || !parent32 || !a0.child_ptr || !work)
    return DN_EINVAL;
  if (!plan && (!parent || !order)) return DN_EINVAL;
  // D (membrane g; 9 B K reserved), R, X, then the solve's workspace: 12 B K
  // for the layer form, 21 B K for the contraction
  const int64_t solve_elems = (plan ? 21 : 12) * B * K;
  if (work_elems < 15 * B * K + solve_elems || (rec && rec_every <= 0)) return DN_EINVAL;
  dn::XtArgs<T> a = a0;
  a.parent = parent32;
  a.D = work;
  a.R = work + 9 * B * K;
  T* X = work + 12 * B * K;
  T* solve_work = work + 15 * B * K;
  const dim3 nodes((unsigned)((B * K + 255) / 256));
  for (int64_t s = 0; s < n_steps; ++s) {
    hipLaunchKernelGGL((dn::extcell_tree_assemble_kernel<T>), nodes, dim3(256), 0, stream, a,
                       s);
    int st;
    if (plan) {
      dn::BcArgs<T> c = *plan;
      c.main = a.main;
      c.edge = a.edge;
      c.rhs = a.R;
      c.gmem = a.D;
      c.x = X;
      c.work = solve_work;
      c.K = K;
      st = bt3_contract_impl<T>(c, B, solve_elems, stream);
    } else {
      st = bt3_tree_impl<T>(a.main, a.edge, a.R, parent, order, layer_ptr, n_layers,
                            a.child_ptr, a.child_idx, X, B, K, solve_work, 12 * B * K, stream,
                            a.D);
    }
    if (st != DN_OK) return st;
    T* r = (rec && (s + 1) % rec_every == 0) ? rec + ((s + 1) / rec_every - 1) * B * K : nullptr;
    hipLaunchKernelGGL((dn::extcell_tree_update_kernel<T>), nodes, dim3(256), 0, stream, a.vc,
                       (const T*)X, B * K, r);
  }
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int64_t dn_extcell_tree_work_elems(int64_t B, int64_t K) {
  return (B < 0 || K < 0) ? -1 : 27 * B * K;
}
DN_EXPORT int64_t dn_extcell_tree_contract_work_elems(int64_t B, int64_t K) {
  return (B < 0 || K < 0) ? -1 : 36 * B * K;
}

#define DN_EXTCELL_TREE_EXPORT(T, SFX)                                                       \
  DN_EXPORT int dn_extcell_tree_run_##SFX(                                                   \
      int64_t B, int64_t K, const T* main, const T* edge, const T* area, const T* xg,        \
      const int32_t* parent32, const int64_t* parent, const int64_t* order,                  \
      const int64_t* layer_ptr, int64_t n_layers, const int32_t* child_ptr,                  \
      const int32_t* child_idx, const T* sna, const T* sk, const T* hh, const T* lead,       \
      const T* amp, const T* wave, const T* iamp, const T* i_wave, T dt, int64_t n_steps,    \
      T* vc, T* m, T* h, T* n, T* rec, int64_t rec_every, T* work, int64_t work_elems,       \
      hipStream_t stream) {                                                                  \
    dn::XtArgs<T> a{main, edge, area, xg,      sna, sk,      hh, lead, amp, wave, iamp,     \
                    i_wave, nullptr, child_ptr, child_idx, vc, m, h, n, nullptr, nullptr,  \
                    B, K, dt};                                                               \
    return extcell_tree_impl<T>(a, parent32, parent, order, layer_ptr, n_layers, n_steps,    \
                                rec, rec_every, work, work_elems, stream);                   \
  }                                                                                          \
  DN_EXPORT int dn_extcell_tree_run_contract_##SFX(                                          \
      int64_t B, int64_t K, const T* main, const T* edge, const T* area, const T* xg,        \
      const int32_t* parent32, const int32_t* child_ptr, const int32_t* child_idx,           \
      const int32_t* perm, int n_sec, const int32_t* sec_head, const int32_t* sec_len,       \
      const int32_t* sec_child, const int32_t* sec_nchild, const int32_t* sec_parent,        \
      int n_rounds, const int32_t* elim_ptr, const int32_t* elim_sec,                        \
      const int32_t* elim_par, const int32_t* elim_child, const int32_t* recv_ptr,           \
      const int32_t* recv_sec, const int32_t* src_ptr, const int32_t* src, const T* sna,     \
      const T* sk, const T* hh, const T* lead, const T* amp, const T* wave, const T* iamp,   \
      const T* i_wave, T dt, int64_t n_steps, T* vc, T* m, T* h, T* n, T* rec,               \
      int64_t rec_every, T* work, int64_t work_elems, hipStream_t stream) {                  \
    dn::XtArgs<T> a{main, edge, area, xg,      sna, sk,      hh, lead, amp, wave, iamp,     \
                    i_wave, nullptr, child_ptr, child_idx, vc, m, h, n, nullptr, nullptr,  \
                    B, K, dt};                                                               \
    dn::BcArgs<T> plan{nullptr,   nullptr,    nullptr,    nullptr,  nullptr,  nullptr,       \
                       K,         n_sec,      n_rounds,   perm,     sec_head, sec_len,       \
                       sec_child, sec_nchild, sec_parent, elim_ptr, elim_sec, elim_par,      \
                       elim_child, recv_ptr,  recv_sec,   src_ptr,  src};                    \
    return extcell_tree_impl<T>(a, parent32, nullptr, nullptr, nullptr, 0, n_steps, rec,     \
                                rec_every, work, work_elems, stream, &plan);                 \
  }

DN_EXTCELL_TREE_EXPORT(float, f32)
DN_EXTCELL_TREE_EXPORT(double, f64)
