// dn_twotap.hip -- NetCon delivery with fractional (interpolated) delays.
//
// GPU form of the reference's dense interpolated-delay transition
// (dendra/models/networks/_event_delivery.py:200-310) and of its ordered
// native two-tap enqueue (networks/_two_tap_enqueue_cpu.cpp:26-125).  On a
// delay ring q[depth][width], row `slot` is cleared, then every connection e
// with target t and weighted gate wg = w_e g_e adds to column t:
//   two taps   (lam = max(delay/dt, 1), k = floor(lam), f = lam - k)
//              wg (1 - f) at row slot + k, wg f at row slot + k + 1,
//              all lower taps first, then all upper taps, in connection order;
//   three taps wg f_j at row slot + k - 1 + j (j = 0, 1, 2; the fractions are
//              the normalised Gaussian weights), connection-major, tap-minor
//              (the order of the reference's flattened scatter).
//   one tap    wg at row slot + delay: the fixed integer-delay transition
//              dense_fixed_delay_transition (_event_delivery.py:125-199),
//              whose scatter_add accumulates in connection order.
// Rows wrap modulo depth.  Each ring cell receives its contributions in the
// reference's order, so the result is bitwise the reference's: one lane per
// target walks that target's connections (target-major CSR, ascending
// connection index), and products and sums are rounded separately (no
// contraction into FMA).  `base` is the first tap's row offset modulo depth
// (the delay for one tap, k for two taps, k - 1 for three).
#include "dn_common.hpp"

namespace dn {

// The ring slot is read from device memory (a one-element int64 tensor, as
// the reference passes it), reduced modulo depth.
__device__ __forceinline__ int ring_slot(const int64_t* slot, int depth) {
  const int64_t s = *slot % depth;
  return (int)(s < 0 ? s + depth : s);
}

__device__ __forceinline__ int64_t ring_row(int slot, int32_t base, int j, int depth) {
  return ((unsigned)slot + (unsigned)base + (unsigned)j) % (unsigned)depth;
}

template <typename T, int TAPS>
__global__ __launch_bounds__(256) void interp_delay_enqueue_kernel(
    T* __restrict__ q, int depth, int64_t width, const int64_t* __restrict__ slot_p,
    const int32_t* __restrict__ in_ptr, const int32_t* __restrict__ in_edge,
    const T* __restrict__ w, const T* __restrict__ g, const T* __restrict__ f,
    const int32_t* __restrict__ base) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= width) return;
  const int slot = ring_slot(slot_p, depth);
  q[(int64_t)slot * width + t] = T(0);
  const int e0 = in_ptr[t], e1 = in_ptr[t + 1];
  if (TAPS == 1) {
    for (int i = e0; i < e1; ++i) {
      const int e = in_edge[i];
      T* cell = q + ring_row(slot, base[e], 0, depth) * width + t;
      *cell = *cell + w[e] * g[e];
    }
  } else if (TAPS == 2) {
    for (int j = 0; j < 2; ++j) {
      for (int i = e0; i < e1; ++i) {
        const int e = in_edge[i];
        const T wg = w[e] * g[e];
        const T tap = j == 0 ? T(1) - f[e] : f[e];
        T* cell = q + ring_row(slot, base[e], j, depth) * width + t;
        *cell = *cell + wg * tap;
      }
    }
  } else {
    for (int i = e0; i < e1; ++i) {
      const int e = in_edge[i];
      const T wg = w[e] * g[e];
      for (int j = 0; j < 3; ++j) {
        T* cell = q + ring_row(slot, base[e], j, depth) * width + t;
        *cell = *cell + wg * f[(int64_t)e * 3 + j];
      }
    }
  }
}

// Tiled form: a workgroup of 256 threads owns kRingTileTargets target
// columns for all depth rows.  The tile of the input ring is staged in LDS
// (coalesced rows) with today's row cleared.  The tile's connections (one
// contiguous range of the target-major CSR) are then taken in chunks of
// kRingChunk: all 256 threads read a chunk's records (packed in CSR order,
// so the reads are coalesced) or, without records, gather its w, g, f and
// base (rounds of kRingPer independent loads per thread and array in
// flight), and leave each connection's tap values and first row in LDS;
// then each wavefront adds one target's items of the chunk to the target's
// LDS column in item order (ring_apply_wave).  Every ring cell therefore
// receives the same values in the same order as the lane-per-target form
// above.  Two taps need all lower taps
// before all upper taps: a tile whose connections fit in one chunk applies
// both from one gather, a longer one makes one pass over its chunks per tap.
// The tile is written back at the end, so the returned copy of the ring is
// fused, and qin may equal qout (each workgroup reads its own columns whole
// before it writes them).
// 4 targets (one per wavefront) and 1024-connection chunks keep a
// workgroup's LDS near 20 KiB for float32 at depth 201, so 8 workgroups
// share a CU and the apply phases of 32 wavefronts overlap (16 targets and
// 4096-connection chunks: 75 against 47 us for 4.2M connections, two taps).
constexpr int kRingTileTargets = 4;
constexpr int kRingThreads = 256;
constexpr int kRingPer = 4;                             // loads in flight per thread
constexpr int kRingBatch = kRingThreads * kRingPer;     // connections per gather round
constexpr int kRingChunk = kRingBatch;                  // connections staged in LDS

template <typename T, int TAPS>
struct RingStage {
  T v[TAPS][kRingChunk];    // tap values of the chunk's connections
  int32_t row[kRingChunk];  // first tap's ring row
};

// One connection's tap values and first row as one aligned record (8 or 16 B
// for float32, 16 or 32 B for float64), written at the connection's
// position in the target-major CSR, so the tile kernel reads a chunk's
// records in order instead of gathering four scattered words per connection.
template <typename T, int TAPS>
struct alignas(sizeof(T) * TAPS + 4 <= 8 ? 8 : (sizeof(T) * TAPS + 4 <= 16 ? 16 : 32))
    RingRec {
  T v[TAPS];
  int32_t row;
};

// The tap values of connection e, rounded as the reference rounds them: one
// tap w g; two taps (w g)(1 - f) and (w g) f; three taps (w g) f_j.
template <typename T, int TAPS>
__device__ __forceinline__ void ring_tap_values(T wv, T gv, const T* fv, T* v) {
#pragma clang fp contract(off)
  const T wg = wv * gv;
  if (TAPS == 1) {
    v[0] = wg;
  } else if (TAPS == 2) {
    v[0] = wg * (T(1) - fv[0]);
    v[1] = wg * fv[0];
  } else {
#pragma unroll
    for (int j = 0; j < TAPS; ++j) v[j] = wg * fv[j];
  }
}

// Records read in connection order (coalesced) and written at the
// connection's CSR position pos[e], so the tile pass reads them in order.
template <typename T, int TAPS>
__global__ __launch_bounds__(256) void interp_delay_pack_kernel(
    RingRec<T, TAPS>* __restrict__ rec, int64_t n_edge, int depth,
    const int64_t* __restrict__ slot_p, const T* __restrict__ w, const T* __restrict__ g,
    const T* __restrict__ f, const int32_t* __restrict__ base,
    const int32_t* __restrict__ pos) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_edge) return;
  const int32_t d = pos[e];
  const int slot = ring_slot(slot_p, depth);
  T fv[3] = {T(0), T(0), T(0)};
  if (TAPS == 2) fv[0] = f[e];
  if (TAPS == 3)
#pragma unroll
    for (int j = 0; j < 3; ++j) fv[j] = f[e * 3 + j];
  RingRec<T, TAPS> r;
  zero_record(r);
  ring_tap_values<T, TAPS>(w[e], g[e], fv, r.v);
  r.row = (int32_t)ring_row(slot, base[e], 0, depth);
  if ((uint32_t)d < (uint64_t)n_edge) store_whole(rec + d, r);
}

// One wavefront adds one target's staged tap values to its LDS column in
// the reference's order: items m in [TS a, TS b) are connection k = m / TS,
// tap j = m % TS, value v[j * kRingChunk + k] into row rows[k] + shift + j
// (mod depth).  The items go 64 at a time, one per lane; within a window the
// lanes whose row no earlier pending lane shares add in one round (the
// smallest lane of each row wins an LDS atomic min on that row's claim
// word, adds, and frees the word), then the next round, and so on, so every
// cell receives its items in item order -- bitwise the sequential walk --
// and a window costs as many rounds as its most repeated row has items.
// claim: depth words of this wavefront, all 0xffffffff between calls.
template <typename T, int TS, int TT>
__device__ __forceinline__ void ring_apply_wave(T* col, const int32_t* rows, const T* v, int a,
                                                int b, int shift, int depth,
                                                uint32_t* claim, int ln) {
#pragma clang fp contract(off)
  for (int m0 = TS * a; m0 < TS * b; m0 += 64) {
    const int m = m0 + ln;
    bool pending = m < TS * b;
    const int mm = pending ? m : TS * b - 1;
    const int k = TS == 1 ? mm : mm / TS;
    const int j = TS == 1 ? 0 : mm - k * TS;
    int r = rows[k] + shift + j;  // < depth + 2: two wraps cover depth 1
    r = r >= depth ? r - depth : r;
    r = r >= depth ? r - depth : r;
    const T add = v[j * kRingChunk + k];
    while (__any(pending)) {
      if (pending) atomicMin(&claim[r], (uint32_t)ln);
      if (pending && claim[r] == (uint32_t)ln) {
        col[r * TT] = col[r * TT] + add;
        claim[r] = 0xffffffffu;
        pending = false;
      }
    }
  }
}

template <typename T, int TAPS, bool REC>
__global__ __launch_bounds__(kRingThreads) void interp_delay_tile_kernel(
    const T* qin, T* qout, int depth, int64_t width, const int64_t* __restrict__ slot_p,
    const int32_t* __restrict__ in_ptr, const int32_t* __restrict__ in_edge,
    const T* __restrict__ w, const T* __restrict__ g, const T* __restrict__ f,
    const int32_t* __restrict__ base, const RingRec<T, TAPS>* __restrict__ rec, int n_rec) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char dn_smem_raw[];
  RingStage<T, TAPS>& st = *reinterpret_cast<RingStage<T, TAPS>*>(dn_smem_raw);
  T* tile = reinterpret_cast<T*>(dn_smem_raw + sizeof(RingStage<T, TAPS>));
  constexpr int TT = kRingTileTargets;
  constexpr int RSTEP = kRingThreads / TT;  // rows staged per pass of the workgroup
  const int tid = threadIdx.x;
  const int64_t t0 = (int64_t)blockIdx.x * TT;
  const int nt = (int)(width - t0 < TT ? width - t0 : TT);
  const int slot = ring_slot(slot_p, depth);
  {
    // loads clamped into the ring (no branches, so all eight are in flight)
    const int c = tid % TT;
    const bool live = c < nt;
    const T* src = qin + t0 + (live ? c : 0);
    for (int r0 = tid / TT; r0 < depth; r0 += 8 * RSTEP) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = r0 + u * RSTEP < depth ? r0 + u * RSTEP : depth - 1;
        v[u] = keep_if(src[(int64_t)r * width], live && r != slot);
      }
      // rows past the ring repeat row depth - 1 with its own value
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = r0 + u * RSTEP < depth ? r0 + u * RSTEP : depth - 1;
        tile[r * TT + c] = v[u];
      }
    }
  }
  const int i0 = in_ptr[t0], i1 = in_ptr[t0 + nt];
  const int lane = tid & 63;
  // each wavefront's claim words (one per ring row), free = 0xffffffff
  uint32_t* claim = reinterpret_cast<uint32_t*>(tile + depth * TT) + (tid >> 6) * depth;
  for (int r = lane; r < depth; r += 64) claim[r] = 0xffffffffu;
  const int passes = (TAPS == 2 && i1 - i0 > kRingChunk) ? 2 : 1;
  for (int pass = 0; pass < passes; ++pass) {
    for (int c0 = i0; c0 < i1; c0 += kRingChunk) {
      const int n = i1 - c0 < kRingChunk ? i1 - c0 : kRingChunk;
      for (int b0 = 0; b0 < n; b0 += kRingBatch) {
        // indices past the chunk repeat its last connection: every load is
        // unconditional and in bounds, so all of them are in flight at once
        int e[kRingPer];
#pragma unroll
        for (int u = 0; u < kRingPer; ++u) {
          const int k = b0 + tid + u * kRingThreads;
          e[u] = c0 + (k < n ? k : n - 1);
          // records: CSR positions past the n_rec records (a CSR that does
          // not cover exactly the records' connections) stay in bounds
          if (REC) e[u] = e[u] < n_rec ? e[u] : n_rec - 1;
          else e[u] = in_edge[e[u]];
        }
        T val[kRingPer][TAPS];
        int32_t row[kRingPer];
        if (REC) {
          RingRec<T, TAPS> rv[kRingPer];
#pragma unroll
          for (int u = 0; u < kRingPer; ++u) rv[u] = rec[e[u]];
#pragma unroll
          for (int u = 0; u < kRingPer; ++u) {
#pragma unroll
            for (int j = 0; j < TAPS; ++j) val[u][j] = rv[u].v[j];
            row[u] = rv[u].row;
          }
        } else {
          T wv[kRingPer], gv[kRingPer], fv[kRingPer][3];
          int32_t bv[kRingPer];
#pragma unroll
          for (int u = 0; u < kRingPer; ++u) {
            wv[u] = w[e[u]];
            gv[u] = g[e[u]];
            bv[u] = base[e[u]];
            if (TAPS == 2) fv[u][0] = f[e[u]];
            if (TAPS == 3)
#pragma unroll
              for (int j = 0; j < 3; ++j) fv[u][j] = f[(int64_t)e[u] * 3 + j];
          }
#pragma unroll
          for (int u = 0; u < kRingPer; ++u) {
            ring_tap_values<T, TAPS>(wv[u], gv[u], fv[u], val[u]);
            row[u] = (int32_t)ring_row(slot, bv[u], 0, depth);
          }
        }
#pragma unroll
        for (int u = 0; u < kRingPer; ++u) {
          // past the chunk: connection n - 1's own values into its own slot
          const int k = b0 + tid + u * kRingThreads < n ? b0 + tid + u * kRingThreads : n - 1;
          if (TAPS == 2 && passes == 2) {
            st.v[0][k] = pass == 0 ? val[u][0] : val[u][TAPS - 1];
          } else {
#pragma unroll
            for (int j = 0; j < TAPS; ++j) st.v[j][k] = val[u][j];
          }
          st.row[k] = row[u];
        }
      }
      __syncthreads();
      // one wavefront per target (targets wave, wave + 4, ...), its lanes
      // over the target's items
      for (int ti = tid >> 6; ti < nt; ti += kRingThreads / 64) {
        const int ea = in_ptr[t0 + ti], eb = in_ptr[t0 + ti + 1];
        const int a = (ea > c0 ? ea : c0) - c0;
        const int b = (eb < c0 + n ? eb : c0 + n) - c0;
        if (a >= b) continue;
        T* col = tile + ti;
        if (TAPS == 3) {
          ring_apply_wave<T, 3, TT>(col, st.row, &st.v[0][0], a, b, 0, depth, claim, lane);
        } else if (TAPS == 2 && passes == 1) {
          ring_apply_wave<T, 1, TT>(col, st.row, st.v[0], a, b, 0, depth, claim, lane);
          ring_apply_wave<T, 1, TT>(col, st.row, st.v[TAPS - 1], a, b, 1, depth, claim, lane);
        } else {
          ring_apply_wave<T, 1, TT>(col, st.row, st.v[0], a, b, pass, depth, claim, lane);
        }
      }
      __syncthreads();
    }
  }
  {
    const int c = tid % TT;
    if (c < nt)
      for (int r = tid / TT; r < depth; r += RSTEP) qout[(int64_t)r * width + t0 + c] = tile[r * TT + c];
  }
}

// Per-connection cotangents for a ring cotangent G.  One tap: dw = g G,
// dg = w G at the connection's cell.  Two taps:
//   dw = g ((1 - f) G_0 + f G_1),  dg = w (...),  df = w g (G_1 - G_0);
// three taps: dw = g sum_j f_j G_j, dg = w sum_j f_j G_j, df_j = w g G_j.
template <typename T, int TAPS>
__global__ __launch_bounds__(256) void interp_delay_vjp_kernel(
    const T* __restrict__ G, int depth, int64_t width, const int64_t* __restrict__ slot_p,
    int64_t n_edge, const int32_t* __restrict__ post, const T* __restrict__ w,
    const T* __restrict__ g, const T* __restrict__ f, const int32_t* __restrict__ base,
    T* __restrict__ dw, T* __restrict__ dg, T* __restrict__ df) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_edge) return;
  const int slot = ring_slot(slot_p, depth);
  const int64_t t = post[e];
  T mix;
  if (TAPS == 1) {
    mix = G[ring_row(slot, base[e], 0, depth) * width + t];
  } else if (TAPS == 2) {
    const T lo = G[ring_row(slot, base[e], 0, depth) * width + t];
    const T hi = G[ring_row(slot, base[e], 1, depth) * width + t];
    mix = (T(1) - f[e]) * lo + f[e] * hi;
    df[e] = w[e] * g[e] * (hi - lo);
  } else {
    mix = T(0);
    for (int j = 0; j < 3; ++j) {
      const T gj = G[ring_row(slot, base[e], j, depth) * width + t];
      mix += f[e * 3 + j] * gj;
      df[e * 3 + j] = w[e] * g[e] * gj;
    }
  }
  dw[e] = g[e] * mix;
  dg[e] = w[e] * mix;
}

}  // namespace dn

// LDS of one ring-tile workgroup: the chunk stage (8-16 KiB for float32,
// 12-28 KiB for float64), depth rows of the tile's columns and each
// wavefront's depth claim words; deeper rings take the global-memory form.
constexpr size_t kRingLdsMax = 160 * 1024;

template <typename T, int TAPS>
static void ring_tiles_launch(const T* src, T* q, int64_t depth, int64_t width,
                              const int64_t* slot, const int32_t* in_ptr,
                              const int32_t* in_edge, const T* w, const T* g, const T* f,
                              const int32_t* base, const void* rec, int64_t n_rec,
                              size_t lds, hipStream_t stream) {
  const dim3 tiles((unsigned)((width + dn::kRingTileTargets - 1) / dn::kRingTileTargets));
  const dim3 block(dn::kRingThreads);
  const auto* r = static_cast<const dn::RingRec<T, TAPS>*>(rec);
  if (rec)
    hipLaunchKernelGGL((dn::interp_delay_tile_kernel<T, TAPS, true>), tiles, block, lds,
                       stream, src, q, (int)depth, width, slot, in_ptr, in_edge, w, g, f,
                       base, r, (int)n_rec);
  else
    hipLaunchKernelGGL((dn::interp_delay_tile_kernel<T, TAPS, false>), tiles, block, lds,
                       stream, src, q, (int)depth, width, slot, in_ptr, in_edge, w, g, f,
                       base, r, (int)n_rec);
}

template <typename T, int TAPS>
static void ring_pack_launch(void* rec, int64_t n_edge, int64_t depth, const int64_t* slot,
                             const T* w, const T* g, const T* f, const int32_t* base,
                             const int32_t* pos, hipStream_t stream) {
  hipLaunchKernelGGL((dn::interp_delay_pack_kernel<T, TAPS>),
                     dim3((unsigned)((n_edge + 255) / 256)), dim3(256), 0, stream,
                     static_cast<dn::RingRec<T, TAPS>*>(rec), n_edge, (int)depth, slot, w, g, f,
                     base, pos);
}

template <typename T>
static size_t ring_rec_bytes(int taps) {
  return taps == 1 ? sizeof(dn::RingRec<T, 1>)
                   : taps == 2 ? sizeof(dn::RingRec<T, 2>) : sizeof(dn::RingRec<T, 3>);
}

// n_edge >= 0 with work: the records are packed first, each at its
// connection's position in the target-major CSR (work holds n_edge records,
// then n_edge int32 positions, computed there when pos is null), and the
// tile kernel reads them in order.  in_edge must be a permutation of
// [0, n_edge) (in_ptr[width] == n_edge).
template <typename T>
static int interp_enqueue_impl(const T* qin, T* q, int64_t depth, int64_t width,
                               const int64_t* slot, int taps, const int32_t* in_ptr,
                               const int32_t* in_edge, const int32_t* pos, const T* w,
                               const T* g, const T* f, const int32_t* base, int64_t n_edge,
                               void* work, int64_t work_bytes, hipStream_t stream) {
  if (!q || !slot || !in_ptr || depth < 1 || depth > (1 << 30) || width < 1) return DN_EINVAL;
  if (taps < 1 || taps > 3 || width > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  if (work && (n_edge < 0 || n_edge > (int64_t)INT32_MAX ||
               work_bytes < n_edge * ((int64_t)ring_rec_bytes<T>(taps) + 4)))
    return DN_EINVAL;
  const size_t stage = taps == 1   ? sizeof(dn::RingStage<T, 1>)
                       : taps == 2 ? sizeof(dn::RingStage<T, 2>)
                                   : sizeof(dn::RingStage<T, 3>);
  const size_t lds = stage + (size_t)depth * dn::kRingTileTargets * sizeof(T) +
                     (size_t)depth * 4 * (dn::kRingThreads / 64);
  if (lds <= kRingLdsMax) {
    const T* src = qin ? qin : q;
    const void* rec = (work && n_edge > 0) ? work : nullptr;
    if (rec) {
      if (!pos) {
        int32_t* own = reinterpret_cast<int32_t*>(static_cast<char*>(work) +
                                                  n_edge * (int64_t)ring_rec_bytes<T>(taps));
        hipLaunchKernelGGL(dn::inverse_permutation_kernel<>, dim3((unsigned)((n_edge + 255) / 256)),
                           dim3(256), 0, stream, own, in_edge, n_edge);
        pos = own;
      }
      if (taps == 1) ring_pack_launch<T, 1>(work, n_edge, depth, slot, w, g, f, base, pos, stream);
      else if (taps == 2) ring_pack_launch<T, 2>(work, n_edge, depth, slot, w, g, f, base, pos, stream);
      else ring_pack_launch<T, 3>(work, n_edge, depth, slot, w, g, f, base, pos, stream);
    }
    if (taps == 1)
      ring_tiles_launch<T, 1>(src, q, depth, width, slot, in_ptr, in_edge, w, g, f, base, rec,
                              n_edge, lds, stream);
    else if (taps == 2)
      ring_tiles_launch<T, 2>(src, q, depth, width, slot, in_ptr, in_edge, w, g, f, base, rec,
                              n_edge, lds, stream);
    else
      ring_tiles_launch<T, 3>(src, q, depth, width, slot, in_ptr, in_edge, w, g, f, base, rec,
                              n_edge, lds, stream);
    return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
  }
  if (qin && qin != q &&
      hipMemcpyAsync(q, qin, sizeof(T) * depth * width, hipMemcpyDeviceToDevice, stream) !=
          hipSuccess)
    return DN_ELAUNCH;
  const dim3 grid((unsigned)((width + 255) / 256));
  if (taps == 1)
    hipLaunchKernelGGL((dn::interp_delay_enqueue_kernel<T, 1>), grid, dim3(256), 0, stream,
                       q, (int)depth, width, slot, in_ptr, in_edge, w, g, f, base);
  else if (taps == 2)
    hipLaunchKernelGGL((dn::interp_delay_enqueue_kernel<T, 2>), grid, dim3(256), 0, stream,
                       q, (int)depth, width, slot, in_ptr, in_edge, w, g, f, base);
  else
    hipLaunchKernelGGL((dn::interp_delay_enqueue_kernel<T, 3>), grid, dim3(256), 0, stream,
                       q, (int)depth, width, slot, in_ptr, in_edge, w, g, f, base);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
static int interp_vjp_impl(const T* G, int64_t depth, int64_t width, const int64_t* slot,
                           int taps, int64_t n_edge, const int32_t* post, const T* w,
                           const T* g, const T* f, const int32_t* base, T* dw, T* dg, T* df,
                           hipStream_t stream) {
  if (!G || !slot || depth < 1 || depth > (1 << 30) || width < 1 || n_edge < 0)
    return DN_EINVAL;
  if (taps < 1 || taps > 3) return DN_EINVAL;
  if (n_edge == 0) return DN_OK;
  if (!post || !w || !g || !base || !dw || !dg) return DN_EINVAL;
  if (taps > 1 && (!f || !df)) return DN_EINVAL;
  if (n_edge > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  const dim3 grid((unsigned)((n_edge + 255) / 256));
  if (taps == 1)
    hipLaunchKernelGGL((dn::interp_delay_vjp_kernel<T, 1>), grid, dim3(256), 0, stream, G,
                       (int)depth, width, slot, n_edge, post, w, g, f, base, dw, dg, df);
  else if (taps == 2)
    hipLaunchKernelGGL((dn::interp_delay_vjp_kernel<T, 2>), grid, dim3(256), 0, stream, G,
                       (int)depth, width, slot, n_edge, post, w, g, f, base, dw, dg, df);
  else
    hipLaunchKernelGGL((dn::interp_delay_vjp_kernel<T, 3>), grid, dim3(256), 0, stream, G,
                       (int)depth, width, slot, n_edge, post, w, g, f, base, dw, dg, df);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int dn_interp_delay_enqueue_f32(const float* qin, float* q, int64_t depth, int64_t width,
                                          const int64_t* slot, int taps,
                                          const int32_t* in_ptr, const int32_t* in_edge,
                                          const float* w, const float* g, const float* f,
                                          const int32_t* base, hipStream_t stream) {
  return interp_enqueue_impl<float>(qin, q, depth, width, slot, taps, in_ptr, in_edge, nullptr,
                                    w, g, f, base, -1, nullptr, 0, stream);
}
DN_EXPORT int dn_interp_delay_enqueue_f64(const double* qin, double* q, int64_t depth, int64_t width,
                                          const int64_t* slot, int taps,
                                          const int32_t* in_ptr, const int32_t* in_edge,
                                          const double* w, const double* g, const double* f,
                                          const int32_t* base, hipStream_t stream) {
  return interp_enqueue_impl<double>(qin, q, depth, width, slot, taps, in_ptr, in_edge,
                                     nullptr, w, g, f, base, -1, nullptr, 0, stream);
}
DN_EXPORT int64_t dn_interp_delay_work_bytes(int64_t n_edge, int taps, int f64) {
  if (n_edge < 0 || taps < 1 || taps > 3) return -1;
  return n_edge * (int64_t)((f64 ? ring_rec_bytes<double>(taps) : ring_rec_bytes<float>(taps)) + 4);
}
DN_EXPORT int dn_interp_delay_enqueue_ws_f32(const float* qin, float* q, int64_t depth,
                                             int64_t width, const int64_t* slot, int taps,
                                             const int32_t* in_ptr, const int32_t* in_edge,
                                             const int32_t* pos, const float* w, const float* g,
                                             const float* f, const int32_t* base, int64_t n_edge,
                                             void* work, int64_t work_bytes, hipStream_t stream) {
  if (!work) return DN_EINVAL;
  return interp_enqueue_impl<float>(qin, q, depth, width, slot, taps, in_ptr, in_edge, pos, w,
                                    g, f, base, n_edge, work, work_bytes, stream);
}
DN_EXPORT int dn_interp_delay_enqueue_ws_f64(const double* qin, double* q, int64_t depth,
                                             int64_t width, const int64_t* slot, int taps,
                                             const int32_t* in_ptr, const int32_t* in_edge,
                                             const int32_t* pos, const double* w, const double* g,
                                             const double* f, const int32_t* base, int64_t n_edge,
                                             void* work, int64_t work_bytes, hipStream_t stream) {
  if (!work) return DN_EINVAL;
  return interp_enqueue_impl<double>(qin, q, depth, width, slot, taps, in_ptr, in_edge, pos,
                                     w, g, f, base, n_edge, work, work_bytes, stream);
}
DN_EXPORT int dn_interp_delay_vjp_f32(const float* G, int64_t depth, int64_t width,
                                      const int64_t* slot, int taps, int64_t n_edge,
                                      const int32_t* post, const float* w, const float* g,
                                      const float* f, const int32_t* base, float* dw,
                                      float* dg, float* df, hipStream_t stream) {
  return interp_vjp_impl<float>(G, depth, width, slot, taps, n_edge, post, w, g, f, base, dw,
                                dg, df, stream);
}
DN_EXPORT int dn_interp_delay_vjp_f64(const double* G, int64_t depth, int64_t width,
                                      const int64_t* slot, int taps, int64_t n_edge,
                                      const int32_t* post, const double* w, const double* g,
                                      const double* f, const int32_t* base, double* dw,
                                      double* dg, double* df, hipStream_t stream) {
  return interp_vjp_impl<double>(G, depth, width, slot, taps, n_edge, post, w, g, f, base,
                                 dw, dg, df, stream);
}
