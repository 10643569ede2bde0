// dn_history.hip -- source-history transport of NetCon training.
//
// GPU form of the reference's pure source-history kernels
// (dendra/models/networks/_source_history.py:83-186), used by NetCon's
// source-history training backend (networks/netcon.py:3660-3725):
//
//   project: today's payload from a history ring hist[depth][n_col] of
//     per-source gates, floating (hist[row][source]) or packed
//     ((hist[row][word] & mask) != 0 as 0/1).  Connection e with target t
//     reads its source at rows (slot - off[e][j]) mod depth, j < taps, and
//     contributes
//       taps 1 (fixed delays)  v = w g0
//       taps 2                 v = w (g0 (1 - a) + g1 a)
//       taps 3                 v = w ((g0 f0 + g1 f1) + g2 f2)
//     with every product and sum rounded separately, as the reference's
//     elementwise torch ops round them.  out[t] = 0 + v_e1 + v_e2 + ... in
//     ascending connection order (the sequential index_add), one lane per
//     target walking the target-major CSR, so the payload is bitwise the
//     reference's.
//   project vjp: per connection dw, the tap-weight cotangents and, for a
//     floating history, the history cotangent scattered with atomics.
//   record: overwrite row `slot` with the current gates (floating), or with
//     the packed words sum_s mask_s [gate_s != 0] (integer sums, exact in any
//     order).
#include "dn_common.hpp"

namespace dn {

__device__ __forceinline__ int hist_slot(const int64_t* slot, int64_t depth) {
  const int64_t s = *slot % depth;
  return (int)(s < 0 ? s + depth : s);
}

// row of tap j: (slot - off) mod depth with slot, off in [0, depth)
__device__ __forceinline__ int64_t hist_row(int slot, int32_t off, int64_t depth) {
  const int64_t r = (int64_t)slot - (int64_t)off;
  return r < 0 ? r + depth : r;
}

template <typename T, bool PACKED>
__device__ __forceinline__ T hist_gate(const void* hist, int64_t n_col, int64_t row,
                                       int32_t col, const int64_t* mask, int64_t e) {
  if (PACKED) {
    const int64_t word = static_cast<const int64_t*>(hist)[row * n_col + col];
    return (word & mask[e]) != 0 ? T(1) : T(0);
  }
  return static_cast<const T*>(hist)[row * n_col + col];
}

// A workgroup of 256 threads owns kProjTargets targets, whose connections
// are one contiguous range of the target-major CSR.  The range is taken in
// chunks of kProjChunk: all threads read a chunk's records (in CSR order,
// coalesced) or gather its connections' inputs, gather their history words,
// and leave each connection's value v in LDS (kProjPer connections per
// thread, every load unconditional so that all are in flight), then one lane
// per target adds its own values of the chunk in ascending connection order.  The sums
// are those of the sequential lane-per-target walk, bit for bit.
// 8 targets per workgroup: 2625 workgroups for 21,000 targets, so every SIMD
// holds several (32 targets: 657 workgroups, 2.6 wavefronts per SIMD, and
// 103 against 89-99 us for 4.2M connections)
constexpr int kProjTargets = 8;
constexpr int kProjThreads = 256;
constexpr int kProjPer = 8;
constexpr int kProjChunk = kProjThreads * kProjPer;

// One connection's history reads packed into one record: its column, the
// tap offsets off_j (the rows (slot - off_j) mod depth follow from the
// slot, so the records stay valid while the slot advances), the tap
// coefficients (two taps: 1 - a and a, rounded as the reference rounds
// them; three taps: f_j), w, and for a packed history its bit mask.  Read in
// connection order (coalesced) and written at the connection's position in
// the target-major CSR, so the target-major pass reads the records in order
// and gathers only the taps' history words, instead of five scattered
// arrays.
template <typename T, int TAPS>
struct alignas(16) ProjRec {
  int32_t col;
  int32_t off[TAPS];
  T c[TAPS];
  T w;
};
template <typename T, int TAPS>
struct alignas(16) ProjRecMask {
  int64_t mask;
  int32_t col;
  int32_t off[TAPS];
  T c[TAPS];
  T w;
};
template <typename T, bool PACKED, int TAPS>
using ProjRecT = typename std::conditional<PACKED, ProjRecMask<T, TAPS>, ProjRec<T, TAPS>>::type;

template <typename T, bool PACKED, int TAPS>
__global__ __launch_bounds__(256) void history_pack_rec_kernel(
    ProjRecT<T, PACKED, TAPS>* __restrict__ rec, int64_t n_conn, const int32_t* __restrict__ col,
    const int64_t* __restrict__ mask, const int32_t* __restrict__ off,
    const T* __restrict__ w, const T* __restrict__ frac, const int32_t* __restrict__ pos) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_conn) return;
  const int32_t d = pos[e];
  ProjRecT<T, PACKED, TAPS> r;
  zero_record(r);
  if constexpr (PACKED) r.mask = mask[e];
  r.col = col[e];
#pragma unroll
  for (int j = 0; j < TAPS; ++j) r.off[j] = off[e * TAPS + j];
  if (TAPS == 1) {
    r.c[0] = T(1);
  } else if (TAPS == 2) {
    const T a = frac[e];
    r.c[0] = T(1) - a;
    r.c[TAPS - 1] = a;
  } else {
#pragma unroll
    for (int j = 0; j < TAPS; ++j) r.c[j] = frac[e * 3 + j];
  }
  r.w = w[e];
  if ((uint32_t)d < (uint64_t)n_conn) store_whole(rec + d, r);
}

// v = w s with s = g0 (one tap), g0 c0 + g1 c1 (two), (g0 c0 + g1 c1) + g2 c2
// (three), every product and sum rounded.
template <typename T, int TAPS>
__device__ __forceinline__ T proj_value(const T* gv, const T* c, T w) {
#pragma clang fp contract(off)
  T s;
  if (TAPS == 1) {
    s = gv[0];
  } else if (TAPS == 2) {
    const T lo = gv[0] * c[0];
    const T hi = gv[TAPS - 1] * c[TAPS - 1];
    s = lo + hi;
  } else {
    const T p0 = gv[0] * c[0];
    const T p1 = gv[1] * c[1];
    const T p2 = gv[TAPS - 1] * c[TAPS - 1];
    s = (p0 + p1) + p2;
  }
  return w * s;
}

// Column-major copy of a floating history: ht[c * depth + r] = hist[r * n_col + c],
// so the 2-3 adjacent tap rows a connection reads share one cache line
// (64 x 64 tiles through LDS, both sides coalesced).
template <typename T>
__global__ __launch_bounds__(256) void history_transpose_kernel(const T* __restrict__ hist,
                                                                T* __restrict__ ht,
                                                                int64_t depth, int64_t n_col) {
  __shared__ T tile[64][65];
  const int64_t c0 = (int64_t)blockIdx.x * 64, r0 = (int64_t)blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // loads clamped into the ring (unconditional, so all 16 are in flight;
  // entries past its edge repeat its last row or column and are not stored)
  const int64_t c = c0 + tx < n_col ? c0 + tx : n_col - 1;
  T v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int64_t r = r0 + ty + 4 * q < depth ? r0 + ty + 4 * q : depth - 1;
    v[q] = hist[r * n_col + c];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) tile[ty + 4 * q][tx] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = ty + 4 * q;
    const int64_t cc = c0 + i, r = r0 + tx;
    if (r < depth && cc < n_col) ht[cc * depth + r] = tile[tx][i];
  }
}

// hist element (row, col) at row * rs + col * cs: rs = n_col, cs = 1 for the
// ring as given, rs = 1, cs = depth for its column-major copy.
template <typename T, bool PACKED, int TAPS, bool REC>
__global__ __launch_bounds__(kProjThreads) void history_project_kernel(
    const void* __restrict__ hist, int64_t depth, int64_t rs, int64_t cs,
    const int64_t* __restrict__ slot_p, const int32_t* __restrict__ in_ptr,
    const int32_t* __restrict__ in_edge, const int32_t* __restrict__ col,
    const int64_t* __restrict__ mask, const int32_t* __restrict__ off,
    const T* __restrict__ w, const T* __restrict__ frac,
    const ProjRecT<T, PACKED, TAPS>* __restrict__ rec, T* __restrict__ out, int64_t width,
    int64_t n_rec) {
#pragma clang fp contract(off)
  __shared__ T vs[kProjChunk];
  const int tid = threadIdx.x;
  const int64_t t0 = (int64_t)blockIdx.x * kProjTargets;
  const int nt = (int)(width - t0 < kProjTargets ? width - t0 : kProjTargets);
  const int slot = hist_slot(slot_p, depth);
  const int i0 = in_ptr[t0], i1 = in_ptr[t0 + nt];
  int e0 = 0, e1 = 0;
  if (tid < nt) {
    e0 = in_ptr[t0 + tid];
    e1 = in_ptr[t0 + tid + 1];
  }
  T acc = T(0);
  for (int c0 = i0; c0 < i1; c0 += kProjChunk) {
    const int n = i1 - c0 < kProjChunk ? i1 - c0 : kProjChunk;
    // indices past the chunk repeat its last connection: every load is
    // unconditional and in bounds
    int64_t e[kProjPer];
#pragma unroll
    for (int u = 0; u < kProjPer; ++u) {
      const int k = tid + u * kProjThreads;
      e[u] = c0 + (k < n ? k : n - 1);
      // records: positions past the n_rec records stay in bounds
      if (REC) e[u] = e[u] < n_rec ? e[u] : n_rec - 1;
      else e[u] = in_edge[e[u]];
    }
    int32_t cu[kProjPer], ru[kProjPer][TAPS];
    T cf[kProjPer][TAPS], wu[kProjPer];
    int64_t mu[kProjPer];
    if (REC) {
      ProjRecT<T, PACKED, TAPS> rv[kProjPer];
#pragma unroll
      for (int u = 0; u < kProjPer; ++u) rv[u] = rec[e[u]];
#pragma unroll
      for (int u = 0; u < kProjPer; ++u) {
        cu[u] = rv[u].col;
        wu[u] = rv[u].w;
        if constexpr (PACKED) mu[u] = rv[u].mask;
#pragma unroll
        for (int j = 0; j < TAPS; ++j) {
          ru[u][j] = (int32_t)hist_row(slot, rv[u].off[j], depth);
          cf[u][j] = rv[u].c[j];
        }
      }
      // two taps: every record in hand before the history reads, so the
      // reads go out together (left to itself the compiler interleaves them
      // with the record loads and waits between them: 115 against 103 us
      // for 4.2M connections; three taps are scheduled together as they are)
      if constexpr (TAPS == 2) {
#pragma unroll
        for (int u = 0; u < kProjPer; ++u) {
          asm volatile("" ::"v"(cu[u]), "v"(ru[u][0]));
          asm volatile("" ::"v"(ru[u][1]));
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < kProjPer; ++u) {
        cu[u] = col[e[u]];
        wu[u] = w[e[u]];
        if (PACKED) mu[u] = mask[e[u]];
#pragma unroll
        for (int j = 0; j < TAPS; ++j)
          ru[u][j] = (int32_t)hist_row(slot, off[e[u] * TAPS + j], depth);
        if (TAPS == 1) {
          cf[u][0] = T(1);
        } else if (TAPS == 2) {
          const T a = frac[e[u]];
          cf[u][0] = T(1) - a;
          cf[u][TAPS - 1] = a;
        } else {
#pragma unroll
          for (int j = 0; j < TAPS; ++j) cf[u][j] = frac[e[u] * 3 + j];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kProjPer; ++u) {
      const int k = tid + u * kProjThreads;
      T gv[TAPS];
#pragma unroll
      for (int j = 0; j < TAPS; ++j) {
        if constexpr (PACKED) {
          const int64_t word = static_cast<const int64_t*>(hist)[ru[u][j] * rs + cu[u] * cs];
          gv[j] = (word & mu[u]) != 0 ? T(1) : T(0);
        } else {
          gv[j] = static_cast<const T*>(hist)[ru[u][j] * rs + cu[u] * cs];
        }
      }
      // past the chunk: connection n - 1's own value into its own slot
      vs[k < n ? k : n - 1] = proj_value<T, TAPS>(gv, cf[u], wu[u]);
    }
    __syncthreads();
    if (tid < nt) {
      const int a = (e0 > c0 ? e0 : c0) - c0;
      const int b = (e1 < c0 + n ? e1 : c0 + n) - c0;
      for (int k = a; k < b; ++k) acc = acc + vs[k];
    }
    __syncthreads();
  }
  if (tid < nt) out[t0 + tid] = acc;
}

template <typename T, bool PACKED, int TAPS>
__global__ __launch_bounds__(256) void history_project_vjp_kernel(
    const T* __restrict__ G, const void* __restrict__ hist, int64_t depth, int64_t n_col,
    const int64_t* __restrict__ slot_p, int64_t n_conn, const int32_t* __restrict__ post,
    const int32_t* __restrict__ col, const int64_t* __restrict__ mask,
    const int32_t* __restrict__ off, const T* __restrict__ w, const T* __restrict__ frac,
    T* __restrict__ dw, T* __restrict__ dfrac, T* __restrict__ dhist) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_conn) return;
  const int slot = hist_slot(slot_p, depth);
  const T gt = G[post[e]];
  const int32_t c = col[e];
  int64_t rows[TAPS];
  T g[TAPS], coef[TAPS];
#pragma unroll
  for (int j = 0; j < TAPS; ++j) {
    rows[j] = hist_row(slot, off[e * TAPS + j], depth);
    g[j] = hist_gate<T, PACKED>(hist, n_col, rows[j], c, mask, e);
  }
  T s;
  if (TAPS == 1) {
    coef[0] = T(1);
    s = g[0];
  } else if (TAPS == 2) {
    const T a = frac[e];
    coef[0] = T(1) - a;
    coef[1] = a;
    s = g[0] * coef[0] + g[1] * coef[1];
  } else {
    s = T(0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      coef[j] = frac[e * 3 + j];
      s += g[j] * coef[j];
    }
  }
  dw[e] = gt * s;
  const T ds = gt * w[e];
  if (TAPS == 2) dfrac[e] = ds * (g[1] - g[0]);
  if (TAPS == 3)
#pragma unroll
    for (int j = 0; j < 3; ++j) dfrac[e * 3 + j] = ds * g[j];
  if (!PACKED && dhist)
#pragma unroll
    for (int j = 0; j < TAPS; ++j) atomicAdd(dhist + rows[j] * n_col + c, ds * coef[j]);
}

template <typename T>
__global__ __launch_bounds__(256) void history_record_kernel(T* __restrict__ hist,
                                                             int64_t depth, int64_t n_col,
                                                             const int64_t* __restrict__ slot_p,
                                                             const T* __restrict__ gates) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_col) return;
  hist[(int64_t)hist_slot(slot_p, depth) * n_col + s] = gates[s];
}

__global__ __launch_bounds__(256) void history_clear_row_kernel(
    int64_t* __restrict__ hist, int64_t depth, int64_t n_words,
    const int64_t* __restrict__ slot_p) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_words) return;
  hist[(int64_t)hist_slot(slot_p, depth) * n_words + k] = 0;
}

__global__ __launch_bounds__(256) void history_pack_kernel(
    int64_t* __restrict__ hist, int64_t depth, int64_t n_words,
    const int64_t* __restrict__ slot_p, int64_t n_src, const uint8_t* __restrict__ active,
    const int32_t* __restrict__ word, const int64_t* __restrict__ mask) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_src || !active[s]) return;
  int64_t* row = hist + (int64_t)hist_slot(slot_p, depth) * n_words;
  atomicAdd(reinterpret_cast<unsigned long long*>(row + word[s]),
            static_cast<unsigned long long>(mask[s]));
}

}  // namespace dn

static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <typename T, bool PACKED, int TAPS>
static void project_launch_taps(const void* hist, int64_t depth, int64_t n_col,
                                const int64_t* slot, const int32_t* in_ptr,
                                const int32_t* in_edge, const int32_t* col, const int64_t* mask,
                                const int32_t* off, const T* w, const T* frac, T* out,
                                int64_t width, int64_t n_conn, void* work,
                                const int32_t* pos, const void* ht, bool reuse,
                                hipStream_t stream) {
  const dim3 g((unsigned)((width + dn::kProjTargets - 1) / dn::kProjTargets)),
      b(dn::kProjThreads);
  using Rec = dn::ProjRecT<T, PACKED, TAPS>;
  Rec* rec = static_cast<Rec*>(work);
  if (rec) {
    if (!reuse)
      hipLaunchKernelGGL((dn::history_pack_rec_kernel<T, PACKED, TAPS>), grid_for(n_conn),
                         dim3(256), 0, stream, rec, n_conn, col, mask, off, w, frac, pos);
    // the column-major copy when there is one
    const void* src = ht ? ht : hist;
    const int64_t rs = ht ? 1 : n_col, cs = ht ? depth : 1;
    hipLaunchKernelGGL((dn::history_project_kernel<T, PACKED, TAPS, true>), g, b, 0, stream,
                       src, depth, rs, cs, slot, in_ptr, in_edge, col, mask, off, w, frac, rec,
                       out, width, n_conn);
  } else {
    hipLaunchKernelGGL((dn::history_project_kernel<T, PACKED, TAPS, false>), g, b, 0, stream,
                       hist, depth, n_col, (int64_t)1, slot, in_ptr, in_edge, col, mask, off, w,
                       frac, rec, out, width, n_conn);
  }
}

template <typename T, bool PACKED>
static void project_launch(int taps, const void* hist, int64_t depth, int64_t n_col,
                           const int64_t* slot, const int32_t* in_ptr, const int32_t* in_edge,
                           const int32_t* col, const int64_t* mask, const int32_t* off,
                           const T* w, const T* frac, T* out, int64_t width, int64_t n_conn,
                           void* work, const int32_t* pos, const void* ht, bool reuse,
                           hipStream_t stream) {
  if (taps == 1)
    project_launch_taps<T, PACKED, 1>(hist, depth, n_col, slot, in_ptr, in_edge, col, mask, off,
                                      w, frac, out, width, n_conn, work, pos, ht, reuse, stream);
  else if (taps == 2)
    project_launch_taps<T, PACKED, 2>(hist, depth, n_col, slot, in_ptr, in_edge, col, mask, off,
                                      w, frac, out, width, n_conn, work, pos, ht, reuse, stream);
  else
    project_launch_taps<T, PACKED, 3>(hist, depth, n_col, slot, in_ptr, in_edge, col, mask, off,
                                      w, frac, out, width, n_conn, work, pos, ht, reuse, stream);
}

static int64_t project_rec_bytes(int taps, int packed, int f64) {
  auto sz = [&](auto tag) -> int64_t {
    using T = decltype(tag);
    if (packed)
      return taps == 1   ? sizeof(dn::ProjRecMask<T, 1>)
             : taps == 2 ? sizeof(dn::ProjRecMask<T, 2>)
                         : sizeof(dn::ProjRecMask<T, 3>);
    return taps == 1   ? sizeof(dn::ProjRec<T, 1>)
           : taps == 2 ? sizeof(dn::ProjRec<T, 2>)
                       : sizeof(dn::ProjRec<T, 3>);
  };
  return f64 ? sz(double(0)) : sz(float(0));
}

// Workspace layout: n_conn records, n_conn int32 CSR positions, then (from a
// 256-byte boundary) the column-major copy of a floating history.
static int64_t project_copy_offset(int64_t n_conn, int taps, int packed, int f64) {
  return (n_conn * (project_rec_bytes(taps, packed, f64) + 4) + 255) / 256 * 256;
}
static bool project_copy_fits(int64_t depth, int64_t n_col) {
  return (depth + 63) / 64 <= 65535 && (n_col + 63) / 64 <= (int64_t)INT32_MAX;
}

// work non-null: n_conn records are packed first, each at its connection's
// position in the target-major CSR (work holds the records, then n_conn int32
// positions, computed there when pos is null), and the target-major pass
// reads them in order.  in_edge must then be a permutation of [0, n_conn).
// A floating history is first copied column-major into the workspace when it
// has room for it, so that each connection's tap rows share a cache line.
// reuse != 0: work already holds the records of an earlier call with the same
// in_edge, pos, col, mask, off, w and frac (hist and slot may differ), and
// they are read as they are.
template <typename T>
static int project_impl(const void* hist, int packed, int64_t depth, int64_t n_col,
                        const int64_t* slot, int taps, const int32_t* in_ptr,
                        const int32_t* in_edge, const int32_t* pos, const int32_t* col,
                        const int64_t* mask, const int32_t* off, const T* w, const T* frac,
                        T* out, int64_t width, int64_t n_conn, void* work, int64_t work_bytes,
                        int reuse, hipStream_t stream) {
  if (!slot || !in_ptr || !out || depth < 1 || n_col < 0 || width < 0) return DN_EINVAL;
  if (taps < 1 || taps > 3 || width > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  if (depth > INT32_MAX) return DN_EINVAL;
  if (work && (n_conn < 0 || n_conn > (int64_t)INT32_MAX ||
               work_bytes < project_copy_offset(n_conn, taps, packed, sizeof(T) == 8)))
    return DN_EINVAL;
  if (width == 0) return DN_OK;
  if (!hist || !in_edge || !col || !off || !w || (taps > 1 && !frac) || (packed && !mask))
    return DN_EINVAL;
  void* rec = (work && n_conn > 0) ? work : nullptr;
  if (rec && !pos && !reuse) {
    int32_t* own = reinterpret_cast<int32_t*>(
        static_cast<char*>(work) + n_conn * project_rec_bytes(taps, packed, sizeof(T) == 8));
    hipLaunchKernelGGL(dn::inverse_permutation_kernel<>, grid_for(n_conn), dim3(256), 0,
                       stream, own, in_edge, n_conn);
    pos = own;
  }
  const void* ht = nullptr;
  if (rec && !packed && project_copy_fits(depth, n_col)) {
    const int64_t at = project_copy_offset(n_conn, taps, packed, sizeof(T) == 8);
    if (work_bytes >= at + depth * n_col * (int64_t)sizeof(T) && n_col > 0) {
      T* copy = reinterpret_cast<T*>(static_cast<char*>(work) + at);
      hipLaunchKernelGGL(dn::history_transpose_kernel<T>,
                         dim3((unsigned)((n_col + 63) / 64), (unsigned)((depth + 63) / 64)),
                         dim3(256), 0, stream, static_cast<const T*>(hist), copy, depth, n_col);
      ht = copy;
    }
  }
  if (packed)
    project_launch<T, true>(taps, hist, depth, n_col, slot, in_ptr, in_edge, col, mask, off,
                            w, frac, out, width, n_conn, rec, pos, ht, reuse != 0, stream);
  else
    project_launch<T, false>(taps, hist, depth, n_col, slot, in_ptr, in_edge, col, mask, off,
                             w, frac, out, width, n_conn, rec, pos, ht, reuse != 0, stream);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T, bool PACKED>
static void vjp_launch(int taps, const T* G, const void* hist, int64_t depth, int64_t n_col,
                       const int64_t* slot, int64_t n_conn, const int32_t* post,
                       const int32_t* col, const int64_t* mask, const int32_t* off, const T* w,
                       const T* frac, T* dw, T* dfrac, T* dhist, hipStream_t stream) {
  const dim3 g = grid_for(n_conn), b(256);
  if (taps == 1)
    hipLaunchKernelGGL((dn::history_project_vjp_kernel<T, PACKED, 1>), g, b, 0, stream, G,
                       hist, depth, n_col, slot, n_conn, post, col, mask, off, w, frac, dw,
                       dfrac, dhist);
  else if (taps == 2)
    hipLaunchKernelGGL((dn::history_project_vjp_kernel<T, PACKED, 2>), g, b, 0, stream, G,
                       hist, depth, n_col, slot, n_conn, post, col, mask, off, w, frac, dw,
                       dfrac, dhist);
  else
    hipLaunchKernelGGL((dn::history_project_vjp_kernel<T, PACKED, 3>), g, b, 0, stream, G,
                       hist, depth, n_col, slot, n_conn, post, col, mask, off, w, frac, dw,
                       dfrac, dhist);
}

template <typename T>
static int vjp_impl(const T* G, const void* hist, int packed, int64_t depth, int64_t n_col,
                    const int64_t* slot, int taps, int64_t n_conn, const int32_t* post,
                    const int32_t* col, const int64_t* mask, const int32_t* off, const T* w,
                    const T* frac, T* dw, T* dfrac, T* dhist, hipStream_t stream) {
  if (!slot || depth < 1 || n_col < 0 || n_conn < 0 || taps < 1 || taps > 3)
    return DN_EINVAL;
  if (n_conn == 0) return DN_OK;
  if (!G || !hist || !post || !col || !off || !w || !dw || (packed && !mask)) return DN_EINVAL;
  if (taps > 1 && (!frac || !dfrac)) return DN_EINVAL;
  if (n_conn > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  if (packed)
    vjp_launch<T, true>(taps, G, hist, depth, n_col, slot, n_conn, post, col, mask, off, w,
                        frac, dw, dfrac, nullptr, stream);
  else
    vjp_launch<T, false>(taps, G, hist, depth, n_col, slot, n_conn, post, col, mask, off, w,
                         frac, dw, dfrac, dhist, stream);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
static int record_impl(T* hist, int64_t depth, int64_t n_col, const int64_t* slot,
                       const T* gates, hipStream_t stream) {
  if (!slot || depth < 1 || n_col < 0 || n_col > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  if (n_col == 0) return DN_OK;  // an empty history may have no storage
  if (!hist || !gates) return DN_EINVAL;
  hipLaunchKernelGGL((dn::history_record_kernel<T>), grid_for(n_col), dim3(256), 0, stream,
                     hist, depth, n_col, slot, gates);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

#define DN_HISTORY_EXPORTS(T, SFX)                                                          \
  DN_EXPORT int dn_history_project_##SFX(                                                    \
      const void* hist, int packed, int64_t depth, int64_t n_col, const int64_t* slot,       \
      int taps, const int32_t* in_ptr, const int32_t* in_edge, const int32_t* col,           \
      const int64_t* mask, const int32_t* off, const T* w, const T* frac, T* out,            \
      int64_t width, hipStream_t stream) {                                                   \
    return project_impl<T>(hist, packed, depth, n_col, slot, taps, in_ptr, in_edge, nullptr, \
                           col, mask, off, w, frac, out, width, -1, nullptr, 0, 0, stream);  \
  }                                                                                          \
  DN_EXPORT int dn_history_project_ws_##SFX(                                                 \
      const void* hist, int packed, int64_t depth, int64_t n_col, const int64_t* slot,       \
      int taps, const int32_t* in_ptr, const int32_t* in_edge, const int32_t* pos,           \
      const int32_t* col, const int64_t* mask, const int32_t* off, const T* w,               \
      const T* frac, T* out, int64_t width, int64_t n_conn, void* work, int64_t work_bytes,  \
      int reuse, hipStream_t stream) {                                                       \
    if (!work) return DN_EINVAL;                                                             \
    return project_impl<T>(hist, packed, depth, n_col, slot, taps, in_ptr, in_edge, pos,     \
                           col, mask, off, w, frac, out, width, n_conn, work, work_bytes,    \
                           reuse, stream);                                                   \
  }                                                                                          \
  DN_EXPORT int dn_history_project_vjp_##SFX(                                                \
      const T* grad_out, const void* hist, int packed, int64_t depth, int64_t n_col,         \
      const int64_t* slot, int taps, int64_t n_conn, const int32_t* post,                    \
      const int32_t* col, const int64_t* mask, const int32_t* off, const T* w,               \
      const T* frac, T* dw, T* dfrac, T* dhist, hipStream_t stream) {                        \
    return vjp_impl<T>(grad_out, hist, packed, depth, n_col, slot, taps, n_conn, post, col,  \
                       mask, off, w, frac, dw, dfrac, dhist, stream);                        \
  }                                                                                          \
  DN_EXPORT int dn_history_record_##SFX(T* hist, int64_t depth, int64_t n_col,              \
                                        const int64_t* slot, const T* gates,                 \
                                        hipStream_t stream) {                                \
    return record_impl<T>(hist, depth, n_col, slot, gates, stream);                          \
  }

DN_EXPORT int64_t dn_history_work_bytes(int64_t n_conn, int taps, int packed, int f64,
                                        int64_t depth, int64_t n_col) {
  if (n_conn < 0 || taps < 1 || taps > 3 || depth < 1 || n_col < 0) return -1;
  const int64_t at = project_copy_offset(n_conn, taps, packed != 0, f64 != 0);
  if (packed || !project_copy_fits(depth, n_col)) return at;
  return at + depth * n_col * (f64 ? 8 : 4);
}

DN_HISTORY_EXPORTS(float, f32)
DN_HISTORY_EXPORTS(double, f64)

DN_EXPORT int dn_history_record_packed(int64_t* hist, int64_t depth, int64_t n_words,
                                       const int64_t* slot, int64_t n_src,
                                       const uint8_t* active, const int32_t* word,
                                       const int64_t* mask, hipStream_t stream) {
  if (!slot || depth < 1 || n_words < 0 || n_src < 0) return DN_EINVAL;
  if (n_words > (int64_t)INT32_MAX * 256 || n_src > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  if (n_words == 0 && n_src == 0) return DN_OK;
  if (!hist || n_words == 0) return DN_EINVAL;
  if (n_words > 0)
    hipLaunchKernelGGL(dn::history_clear_row_kernel, grid_for(n_words), dim3(256), 0, stream,
                       hist, depth, n_words, slot);
  if (n_src > 0) {
    if (!active || !word || !mask) return DN_EINVAL;
    hipLaunchKernelGGL(dn::history_pack_kernel, grid_for(n_src), dim3(256), 0, stream, hist,
                       depth, n_words, slot, n_src, active, word, mask);
  }
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}
