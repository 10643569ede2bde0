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

// Tiled form: a workgroup of 64 lanes owns 64 target columns for all depth
// rows.  The tile of the input ring is staged in LDS (coalesced), today's row
// cleared, each lane applies its target's contributions in the same order as
// above to the LDS cells, and the tile is written to the output ring.  The
// copy the pure transition returns is thus fused, and the per-cell updates
// are LDS read-modify-writes, so the connection loads can run ahead.
template <typename T, int TAPS>
__global__ __launch_bounds__(64) void interp_delay_tile_kernel(
    const T* __restrict__ qin, T* __restrict__ qout, int depth, int64_t width,
    const int64_t* __restrict__ slot_p, const int32_t* __restrict__ in_ptr,
    const int32_t* __restrict__ in_edge, const T* __restrict__ w, const T* __restrict__ g,
    const T* __restrict__ f, const int32_t* __restrict__ base) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char dn_smem_raw[];
  T* tile = reinterpret_cast<T*>(dn_smem_raw);
  const int lane = threadIdx.x;
  const int64_t t0 = (int64_t)blockIdx.x * 64;
  const int64_t t = t0 + lane;
  const bool live = t < width;
  const int slot = ring_slot(slot_p, depth);
  for (int r = 0; r < depth; ++r)
    tile[r * 64 + lane] = (live && r != slot) ? qin[(int64_t)r * width + t] : T(0);
  if (live) {
    const int e0 = in_ptr[t], e1 = in_ptr[t + 1];
    if (TAPS == 1) {
#pragma unroll 4
      for (int i = e0; i < e1; ++i) {
        const int e = in_edge[i];
        T* cell = tile + ring_row(slot, base[e], 0, depth) * 64 + lane;
        *cell = *cell + w[e] * g[e];
      }
    } else if (TAPS == 2) {
      for (int j = 0; j < 2; ++j) {
#pragma unroll 4
        for (int i = e0; i < e1; ++i) {
          const int e = in_edge[i];
          const T wg = w[e] * g[e];
          const T tap = j == 0 ? T(1) - f[e] : f[e];
          T* cell = tile + ring_row(slot, base[e], j, depth) * 64 + lane;
          *cell = *cell + wg * tap;
        }
      }
    } else {
#pragma unroll 4
      for (int i = e0; i < e1; ++i) {
        const int e = in_edge[i];
        const T wg = w[e] * g[e];
        for (int j = 0; j < 3; ++j) {
          T* cell = tile + ring_row(slot, base[e], j, depth) * 64 + lane;
          *cell = *cell + wg * f[(int64_t)e * 3 + j];
        }
      }
    }
    for (int r = 0; r < depth; ++r) qout[(int64_t)r * width + t] = tile[r * 64 + lane];
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

template <typename T>
static int interp_enqueue_impl(const T* qin, T* q, int64_t depth, int64_t width,
                               const int64_t* slot, int taps, const int32_t* in_ptr,
                               const int32_t* in_edge, const T* w, const T* g, const T* f,
                               const int32_t* base, hipStream_t stream) {
  if (!q || !slot || !in_ptr || depth < 1 || depth > (1 << 30) || width < 1) return DN_EINVAL;
  if (taps < 1 || taps > 3 || width > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  const size_t lds = (size_t)depth * 64 * sizeof(T);
  if (qin && qin != q && lds <= 64 * 1024) {
    const dim3 tiles((unsigned)((width + 63) / 64));
    if (taps == 1)
      hipLaunchKernelGGL((dn::interp_delay_tile_kernel<T, 1>), tiles, dim3(64), lds, stream,
                         qin, q, (int)depth, width, slot, in_ptr, in_edge, w, g, f, base);
    else if (taps == 2)
      hipLaunchKernelGGL((dn::interp_delay_tile_kernel<T, 2>), tiles, dim3(64), lds, stream,
                         qin, q, (int)depth, width, slot, in_ptr, in_edge, w, g, f, base);
    else
      hipLaunchKernelGGL((dn::interp_delay_tile_kernel<T, 3>), tiles, dim3(64), lds, stream,
                         qin, q, (int)depth, width, slot, in_ptr, in_edge, w, g, f, base);
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
  return interp_enqueue_impl<float>(qin, q, depth, width, slot, taps, in_ptr, in_edge, w, g,
                                    f, base, stream);
}
DN_EXPORT int dn_interp_delay_enqueue_f64(const double* qin, double* q, int64_t depth, int64_t width,
                                          const int64_t* slot, int taps,
                                          const int32_t* in_ptr, const int32_t* in_edge,
                                          const double* w, const double* g, const double* f,
                                          const int32_t* base, hipStream_t stream) {
  return interp_enqueue_impl<double>(qin, q, depth, width, slot, taps, in_ptr, in_edge, w,
                                     g, f, base, stream);
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
