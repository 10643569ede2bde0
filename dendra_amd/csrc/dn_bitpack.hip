// dn_bitpack.hip -- packed source-spike history and dense delivery.
//
// GPU form of the reference's NetCon bitpack extension
// (dendra/models/networks/netcon_bitpack_kernel.cu:20-126, host wrappers
// netcon_bitpack_kernel.cpp:32-149, loader netcon_bitpack_ops.py:559-700):
//
//   pack     row `cur` of history[D][W] gets word w = the spikes of sources
//            63w .. 63w+62 in bits 0..62 (bit 63 stays clear);
//   deliver  out[post_e] += weight_e for every connection e whose source bit
//            (word_e, mask_e) is set in history row (cur - delay_e) mod D.
//
// Packing is one wavefront per word: lane l < 63 tests source 63w + l and a
// wavefront ballot forms the word in one instruction, so no LDS reduction or
// atomics are needed.  Delivery streams the connection arrays once (one
// connection per lane; 36-44 bytes per connection are the whole cost, the
// history rows stay in L2) and adds a hit with a float atomic, as the
// reference does: hits that share a target in one call land in an unspecified
// order.  Out-of-range words, targets and rows are skipped rather than read
// or written.
#include "dn_common.hpp"

namespace dn {

constexpr int kBitsPerWord = 63;

__global__ __launch_bounds__(256) void bitpack_pack_kernel(const uint8_t* __restrict__ spikes,
                                                           int64_t n_source,
                                                           int64_t* __restrict__ history,
                                                           int64_t depth, int64_t n_words,
                                                           const int64_t* __restrict__ cur_p) {
  const int64_t word = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (word >= n_words) return;
  const int64_t src = word * kBitsPerWord + lane;
  const bool on = lane < kBitsPerWord && src < n_source && spikes[src] != 0;
  const uint64_t bits = __ballot(on);
  int64_t cur = *cur_p % depth;
  if (cur < 0) cur += depth;
  if (lane == 0) history[cur * n_words + word] = (int64_t)bits;
}

template <typename T, bool UNIFORM>
__global__ __launch_bounds__(256) void bitpack_deliver_kernel(
    const int64_t* __restrict__ history, int64_t depth, int64_t n_words,
    const int64_t* __restrict__ cur_p, const int64_t* __restrict__ delay,
    int64_t uniform_delay, const int64_t* __restrict__ word_idx,
    const int64_t* __restrict__ bit_mask, const int64_t* __restrict__ post,
    const T* __restrict__ weight, T* __restrict__ out, int64_t n_out, int64_t n_conn) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_conn) return;
  const int64_t d = UNIFORM ? uniform_delay : delay[e];
  int64_t row = (*cur_p - d) % depth;
  if (row < 0) row += depth;
  const int64_t w = word_idx[e];
  if (w < 0 || w >= n_words) return;
  const uint64_t word = (uint64_t)history[row * n_words + w];
  if ((word & (uint64_t)bit_mask[e]) == 0) return;
  const int64_t t = post[e];
  if (t < 0 || t >= n_out) return;
  atomicAdd(out + t, weight[e]);
}

}  // namespace dn

DN_EXPORT int dn_bitpack_pack(const uint8_t* spikes, int64_t n_source, int64_t* history,
                              int64_t depth, int64_t n_words, const int64_t* cur,
                              hipStream_t stream) {
  if (n_source < 0 || depth < 0 || n_words < 0) return DN_EINVAL;
  if (n_words != (n_source + dn::kBitsPerWord - 1) / dn::kBitsPerWord) return DN_EINVAL;
  if (n_words == 0) return DN_OK;
  if (!spikes || !history || !cur || depth < 1) return DN_EINVAL;
  const int64_t blocks = (n_words + 3) / 4;
  if (blocks > INT32_MAX) return DN_ERANGE;
  hipLaunchKernelGGL(dn::bitpack_pack_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                     spikes, n_source, history, depth, n_words, cur);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
static int bitpack_deliver_impl(const int64_t* history, int64_t depth, int64_t n_words,
                                const int64_t* cur, const int64_t* delay, int64_t uniform_delay,
                                const int64_t* word_idx, const int64_t* bit_mask,
                                const int64_t* post, const T* weight, T* out, int64_t n_out,
                                int64_t n_conn, hipStream_t stream) {
  if (n_conn < 0 || depth < 0 || n_words < 0 || n_out < 0) return DN_EINVAL;
  if (n_conn == 0 || depth == 0 || n_words == 0) return DN_OK;
  if (!history || !cur || !word_idx || !bit_mask || !post || !weight || !out)
    return DN_EINVAL;
  const int64_t blocks = (n_conn + 255) / 256;
  if (blocks > INT32_MAX) return DN_ERANGE;
  if (delay)
    hipLaunchKernelGGL((dn::bitpack_deliver_kernel<T, false>), dim3((unsigned)blocks),
                       dim3(256), 0, stream, history, depth, n_words, cur, delay, 0, word_idx,
                       bit_mask, post, weight, out, n_out, n_conn);
  else {
    // the reference's uniform-delay rule: below 1 -> 1, at or beyond depth -> mod depth
    int64_t d = uniform_delay < 1 ? 1 : uniform_delay;
    if (d >= depth) d %= depth;
    hipLaunchKernelGGL((dn::bitpack_deliver_kernel<T, true>), dim3((unsigned)blocks),
                       dim3(256), 0, stream, history, depth, n_words, cur, nullptr, d,
                       word_idx, bit_mask, post, weight, out, n_out, n_conn);
  }
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int dn_bitpack_deliver_f32(const int64_t* history, int64_t depth, int64_t n_words,
                                     const int64_t* cur, const int64_t* delay,
                                     int64_t uniform_delay, const int64_t* word_idx,
                                     const int64_t* bit_mask, const int64_t* post,
                                     const float* weight, float* out, int64_t n_out,
                                     int64_t n_conn, hipStream_t stream) {
  return bitpack_deliver_impl<float>(history, depth, n_words, cur, delay, uniform_delay,
                                     word_idx, bit_mask, post, weight, out, n_out, n_conn,
                                     stream);
}

DN_EXPORT int dn_bitpack_deliver_f64(const int64_t* history, int64_t depth, int64_t n_words,
                                     const int64_t* cur, const int64_t* delay,
                                     int64_t uniform_delay, const int64_t* word_idx,
                                     const int64_t* bit_mask, const int64_t* post,
                                     const double* weight, double* out, int64_t n_out,
                                     int64_t n_conn, hipStream_t stream) {
  return bitpack_deliver_impl<double>(history, depth, n_words, cur, delay, uniform_delay,
                                      word_idx, bit_mask, post, weight, out, n_out, n_conn,
                                      stream);
}
