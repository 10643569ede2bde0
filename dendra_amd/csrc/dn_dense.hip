// dn_dense.hip -- dense NetCon delay-ring helpers.
//
// GPU forms of two pieces of the reference's dense event transport
// (dendra/models/networks/_event_delivery.py):
//
//   event ring step  the optional int32 event-count ring [depth][n_conn] of
//                    dense_fixed_delay_transition (:175-190): a copy of the
//                    input ring with row `slot` cleared, then every connection
//                    e adds count[e] at row (slot + delay[e]) mod depth of its
//                    own column e (no two connections share a cell);
//   ordered scatter  _enqueue_dense_values (:61-105), q.flat[dest[i]] +=
//                    value[i], with every cell receiving its values in index
//                    order i -- the order of PyTorch's sequential CPU
//                    scatter_add, so the float result is bitwise that one.
//                    The (dest, value) pairs are sorted stably by dest (hipcub
//                    radix sort), then one lane per run of equal dest adds the
//                    run in order into the cell.  Destinations outside the
//                    ring are dropped.
#include "dn_common.hpp"

#include <hipcub/hipcub.hpp>

namespace dn {

// One lane per connection column: copy the column with row `slot` cleared,
// then add the connection's count at its delay row (same lane, in order).
__global__ __launch_bounds__(256) void dense_event_step_kernel(
    const int32_t* __restrict__ qin, int32_t* __restrict__ qout, int64_t depth,
    int64_t n_conn, const int64_t* __restrict__ slot_p, const int64_t* __restrict__ delay,
    const int32_t* __restrict__ counts) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_conn) return;
  int64_t slot = *slot_p % depth;
  if (slot < 0) slot += depth;
  for (int64_t r = 0; r < depth; ++r)
    qout[r * n_conn + e] = r == slot ? 0 : qin[r * n_conn + e];
  int64_t row = (slot + delay[e]) % depth;
  if (row < 0) row += depth;
  qout[row * n_conn + e] += counts[e];
}

__global__ __launch_bounds__(256) void scatter_keys_kernel(const int64_t* __restrict__ dest,
                                                           int64_t n, int64_t n_cells,
                                                           uint32_t* __restrict__ key) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t d = dest[i];
  key[i] = (d < 0 || d >= n_cells) ? (uint32_t)n_cells : (uint32_t)d;
}

template <typename T>
__global__ __launch_bounds__(256) void scatter_runs_kernel(const uint32_t* __restrict__ key,
                                                           const T* __restrict__ val,
                                                           int64_t n, int64_t n_cells,
                                                           T* __restrict__ q) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = key[i];
  if ((int64_t)k >= n_cells || (i > 0 && key[i - 1] == k)) return;
  T acc = q[k];
  for (int64_t j = i; j < n && key[j] == k; ++j) acc = acc + val[j];
  q[k] = acc;
}

}  // namespace dn

namespace {

constexpr int64_t kAlign = 256;
int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

int key_bits(int64_t n_cells) {
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) <= n_cells) ++b;
  return b;
}

template <typename T>
size_t sort_bytes(int64_t n, int bits) {
  size_t bytes = 0;
  if (n > 0)
    hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr,
                                       (uint32_t*)nullptr, (const T*)nullptr, (T*)nullptr,
                                       (int)n, 0, bits);
  return bytes;
}

// workspace: keys, sorted keys, sorted values, radix sort temporary storage
template <typename T>
int64_t scatter_work_bytes(int64_t n, int64_t n_cells) {
  const int64_t m = n > 0 ? n : 1;
  return 2 * align_up(4 * m) + align_up((int64_t)sizeof(T) * m) +
         align_up((int64_t)sort_bytes<T>(n, key_bits(n_cells)));
}

template <typename T>
int ordered_scatter_impl(T* q, int64_t n_cells, const int64_t* dest, const T* values,
                         int64_t n, void* work, int64_t work_bytes, hipStream_t stream) {
  if (n < 0 || n_cells < 0 || n_cells >= INT32_MAX || n >= INT32_MAX) return DN_EINVAL;
  if (n == 0 || n_cells == 0) return DN_OK;
  if (!q || !dest || !values || !work) return DN_EINVAL;
  if (work_bytes < scatter_work_bytes<T>(n, n_cells)) return DN_EINVAL;
  char* p = static_cast<char*>(work);
  uint32_t* key = reinterpret_cast<uint32_t*>(p);
  p += align_up(4 * n);
  uint32_t* key_sorted = reinterpret_cast<uint32_t*>(p);
  p += align_up(4 * n);
  T* val_sorted = reinterpret_cast<T*>(p);
  p += align_up((int64_t)sizeof(T) * n);
  const int bits = key_bits(n_cells);
  size_t bytes = sort_bytes<T>(n, bits);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(dn::scatter_keys_kernel, dim3(blocks), dim3(256), 0, stream, dest, n,
                     n_cells, key);
  if (hipcub::DeviceRadixSort::SortPairs(p, bytes, key, key_sorted, values, val_sorted,
                                         (int)n, 0, bits, stream) != hipSuccess)
    return DN_ELAUNCH;
  hipLaunchKernelGGL((dn::scatter_runs_kernel<T>), dim3(blocks), dim3(256), 0, stream,
                     key_sorted, val_sorted, n, n_cells, q);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace

DN_EXPORT int dn_dense_event_step(const int32_t* qin, int32_t* qout, int64_t depth,
                                  int64_t n_conn, const int64_t* slot, const int64_t* delay,
                                  const int32_t* counts, hipStream_t stream) {
  if (depth < 1 || n_conn < 0) return DN_EINVAL;
  if (n_conn == 0) return DN_OK;
  if (!qin || !qout || !slot || !delay || !counts) return DN_EINVAL;
  const int64_t blocks = (n_conn + 255) / 256;
  if (blocks > INT32_MAX) return DN_ERANGE;
  hipLaunchKernelGGL(dn::dense_event_step_kernel, dim3((unsigned)blocks), dim3(256), 0,
                     stream, qin, qout, depth, n_conn, slot, delay, counts);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int64_t dn_ordered_scatter_work_bytes(int64_t n, int64_t n_cells, int f64) {
  if (n < 0 || n_cells < 0) return -1;
  return f64 ? scatter_work_bytes<double>(n, n_cells) : scatter_work_bytes<float>(n, n_cells);
}

DN_EXPORT int dn_ordered_scatter_add_f32(float* q, int64_t n_cells, const int64_t* dest,
                                         const float* values, int64_t n, void* work,
                                         int64_t work_bytes, hipStream_t stream) {
  return ordered_scatter_impl<float>(q, n_cells, dest, values, n, work, work_bytes, stream);
}

DN_EXPORT int dn_ordered_scatter_add_f64(double* q, int64_t n_cells, const int64_t* dest,
                                         const double* values, int64_t n, void* work,
                                         int64_t work_bytes, hipStream_t stream) {
  return ordered_scatter_impl<double>(q, n_cells, dest, values, n, work, work_bytes, stream);
}
