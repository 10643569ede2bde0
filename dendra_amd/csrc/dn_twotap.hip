// dn_twotap.hip -- NetCon delivery with fractional (two-tap) delays.
//
// GPU form of the reference's ordered two-tap enqueue
// (dendra/models/networks/_two_tap_enqueue_cpu.cpp:26-125, used by
// dense_interpolated_delay_transition, _event_delivery.py:226-310): on a
// delay ring q[depth][width], clear row `slot`, then for every connection e
// with target t, weight w, gate g, row offset k = floor(delay/dt) >= 1 and
// fraction f = delay/dt - k:
//   q[(slot + k)     % depth][t] += (w g) (1 - f)     (all lower taps, in e order)
//   q[(slot + k + 1) % depth][t] += (w g) f           (then all upper taps)
// Each ring cell receives its contributions in the reference's order, so the
// result is bitwise the reference's: one lane per target walks that target's
// connections (target-major CSR, ascending connection index) twice, and
// products and sums are rounded separately (no contraction into FMA).
#include "dn_common.hpp"

namespace dn {

// The ring slot is read from device memory (a one-element int64 tensor, as
// the reference passes it), reduced modulo depth.
__device__ __forceinline__ int ring_slot(const int64_t* slot, int depth) {
  const int64_t s = *slot % depth;
  return (int)(s < 0 ? s + depth : s);
}

template <typename T>
__global__ __launch_bounds__(256) void two_tap_enqueue_kernel(
    T* __restrict__ q, int depth, int64_t width, const int64_t* __restrict__ slot_p,
    const int32_t* __restrict__ in_ptr, const int32_t* __restrict__ in_edge,
    const T* __restrict__ w, const T* __restrict__ g, const T* __restrict__ f,
    const int32_t* __restrict__ off) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= width) return;
  const int slot = ring_slot(slot_p, depth);
  q[(int64_t)slot * width + t] = T(0);
  const int e0 = in_ptr[t], e1 = in_ptr[t + 1];
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = e0; i < e1; ++i) {
      const int e = in_edge[i];
      const int64_t row = ((unsigned)slot + (unsigned)off[e] + pass) % (unsigned)depth;
      const T wg = w[e] * g[e];
      const T tap = pass == 0 ? T(1) - f[e] : f[e];
      T* cell = q + row * width + t;
      *cell = *cell + wg * tap;
    }
  }
}

// Per-connection cotangents of the enqueue for a ring cotangent G:
//   dw = g ((1 - f) G_lo + f G_hi),  dg = w (...),  df = w g (G_hi - G_lo)
template <typename T>
__global__ __launch_bounds__(256) void two_tap_vjp_kernel(
    const T* __restrict__ G, int depth, int64_t width, const int64_t* __restrict__ slot_p,
    int64_t n_edge, const int32_t* __restrict__ post, const T* __restrict__ w,
    const T* __restrict__ g, const T* __restrict__ f, const int32_t* __restrict__ off,
    T* __restrict__ dw, T* __restrict__ dg, T* __restrict__ df) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_edge) return;
  const int slot = ring_slot(slot_p, depth);
  const int64_t t = post[e];
  const int64_t r0 = ((unsigned)slot + (unsigned)off[e]) % (unsigned)depth;
  const int64_t r1 = ((unsigned)slot + (unsigned)off[e] + 1u) % (unsigned)depth;
  const T lo = G[r0 * width + t], hi = G[r1 * width + t];
  const T mix = (T(1) - f[e]) * lo + f[e] * hi;
  dw[e] = g[e] * mix;
  dg[e] = w[e] * mix;
  df[e] = w[e] * g[e] * (hi - lo);
}

}  // namespace dn

template <typename T>
static int two_tap_enqueue_impl(T* q, int64_t depth, int64_t width, const int64_t* slot,
                                const int32_t* in_ptr, const int32_t* in_edge,
                                const T* w, const T* g, const T* f, const int32_t* off,
                                hipStream_t stream) {
  if (!q || !slot || !in_ptr || depth < 1 || depth > (1 << 30) || width < 1) return DN_EINVAL;
  if (width > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  hipLaunchKernelGGL((dn::two_tap_enqueue_kernel<T>), dim3((unsigned)((width + 255) / 256)),
                     dim3(256), 0, stream, q, (int)depth, width, slot, in_ptr, in_edge, w, g,
                     f, off);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
static int two_tap_vjp_impl(const T* G, int64_t depth, int64_t width, const int64_t* slot,
                            int64_t n_edge, const int32_t* post, const T* w, const T* g,
                            const T* f, const int32_t* off, T* dw, T* dg, T* df,
                            hipStream_t stream) {
  if (!G || !slot || depth < 1 || depth > (1 << 30) || width < 1 || n_edge < 0)
    return DN_EINVAL;
  if (n_edge == 0) return DN_OK;
  if (!post || !w || !g || !f || !off || !dw || !dg || !df) return DN_EINVAL;
  if (n_edge > (int64_t)INT32_MAX * 256) return DN_EINVAL;
  hipLaunchKernelGGL((dn::two_tap_vjp_kernel<T>), dim3((unsigned)((n_edge + 255) / 256)),
                     dim3(256), 0, stream, G, (int)depth, width, slot, n_edge, post, w, g,
                     f, off, dw, dg, df);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int dn_two_tap_enqueue_f32(float* q, int64_t depth, int64_t width, const int64_t* slot,
                                     const int32_t* in_ptr, const int32_t* in_edge,
                                     const float* w, const float* g, const float* f,
                                     const int32_t* off, hipStream_t stream) {
  return two_tap_enqueue_impl<float>(q, depth, width, slot, in_ptr, in_edge, w, g, f, off,
                                     stream);
}
DN_EXPORT int dn_two_tap_enqueue_f64(double* q, int64_t depth, int64_t width, const int64_t* slot,
                                     const int32_t* in_ptr, const int32_t* in_edge,
                                     const double* w, const double* g, const double* f,
                                     const int32_t* off, hipStream_t stream) {
  return two_tap_enqueue_impl<double>(q, depth, width, slot, in_ptr, in_edge, w, g, f, off,
                                      stream);
}
DN_EXPORT int dn_two_tap_vjp_f32(const float* G, int64_t depth, int64_t width, const int64_t* slot,
                                 int64_t n_edge, const int32_t* post, const float* w,
                                 const float* g, const float* f, const int32_t* off,
                                 float* dw, float* dg, float* df, hipStream_t stream) {
  return two_tap_vjp_impl<float>(G, depth, width, slot, n_edge, post, w, g, f, off, dw, dg,
                                 df, stream);
}
DN_EXPORT int dn_two_tap_vjp_f64(const double* G, int64_t depth, int64_t width, const int64_t* slot,
                                 int64_t n_edge, const int32_t* post, const double* w,
                                 const double* g, const double* f, const int32_t* off,
                                 double* dw, double* dg, double* df, hipStream_t stream) {
  return two_tap_vjp_impl<double>(G, depth, width, slot, n_edge, post, w, g, f, off, dw, dg,
                                  df, stream);
}
