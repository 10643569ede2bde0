// dn_calendar.hip -- bounded pending calendar of the sparse NetCon backend.
//
// GPU form of the reference's pure pending calendar
// (dendra/models/networks/_pending_calendar.py:16-113), the storage of the
// "sparse_calendar" delay backend (networks/netcon.py:3309-3450, 5223-5300).
// A calendar is (slots, indices, values)[capacity] with a compact valid
// prefix of `size` entries in enqueue order and a sticky overflow flag.
//
//   append: the active candidates, in order, go to positions size + rank
//     (rank = exclusive count of earlier active candidates), unless
//     size + count > capacity or the calendar already overflowed: then
//     nothing is written and overflow is set (append is all or nothing).
//   pop: the valid entries with slot == `slot` are due.  Their values are
//     added to payload[index] per index in position order (the sequential
//     scatter_add), and the other valid entries are compacted, in order, to
//     the front of fresh arrays whose tail is zero.
//
// Both are stream compactions: per 1024-entry tile a count, one workgroup
// scans the tile counts, and each tile writes its entries at
// tile offset + local rank.  The ordered per-index reduction of a pop sorts
// (key, value) pairs stably by key (hipcub radix sort), key = index for due
// entries and `width` for the rest, then one lane per run of equal keys adds
// the run's values in order.
#include "dn_common.hpp"

#include <hipcub/hipcub.hpp>

namespace dn {

constexpr int kCalThreads = 256;
constexpr int kCalItems = 4;
constexpr int kCalTile = kCalThreads * kCalItems;

__device__ __forceinline__ int64_t cal_block_exclusive(int64_t v, int64_t* buf,
                                                       int64_t* total) {
  const int tid = threadIdx.x;
  buf[tid] = v;
  __syncthreads();
  for (int off = 1; off < kCalThreads; off <<= 1) {
    const int64_t t = tid >= off ? buf[tid - off] : 0;
    __syncthreads();
    buf[tid] += t;
    __syncthreads();
  }
  const int64_t incl = buf[tid];
  *total = buf[kCalThreads - 1];
  __syncthreads();
  return incl - v;
}

// per tile: number of set flags
__global__ __launch_bounds__(kCalThreads) void cal_tile_count_kernel(
    const uint8_t* __restrict__ flags, int64_t n, int64_t* __restrict__ tile_sum) {
  __shared__ int64_t buf[kCalThreads];
  const int64_t base = (int64_t)blockIdx.x * kCalTile + threadIdx.x * kCalItems;
  int64_t c = 0;
  for (int k = 0; k < kCalItems; ++k)
    if (base + k < n && flags[base + k]) ++c;
  int64_t total;
  cal_block_exclusive(c, buf, &total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// pop: keep flags (valid, not due), sort keys and the per-tile keep counts
__global__ __launch_bounds__(kCalThreads) void cal_pop_flags_kernel(
    const int64_t* __restrict__ slots, const int64_t* __restrict__ indices, int64_t cap,
    const int64_t* __restrict__ size_p, const int64_t* __restrict__ slot_p, int64_t width,
    uint8_t* __restrict__ keep, int32_t* __restrict__ key, int64_t* __restrict__ tile_sum) {
  __shared__ int64_t buf[kCalThreads];
  const int64_t size = *size_p, slot = *slot_p;
  const int64_t base = (int64_t)blockIdx.x * kCalTile + threadIdx.x * kCalItems;
  int64_t c = 0;
  for (int k = 0; k < kCalItems; ++k) {
    const int64_t p = base + k;
    if (p >= cap) break;
    const bool valid = p < size;
    const bool due = valid && slots[p] == slot;
    const int64_t idx = indices[p];
    key[p] = (due && idx >= 0 && idx < width) ? (int32_t)idx : (int32_t)width;
    keep[p] = valid && !due;
    c += keep[p];
  }
  int64_t total;
  cal_block_exclusive(c, buf, &total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the tile counts; then the new size (and,
// for an append, the sticky overflow decision)
__global__ __launch_bounds__(1024) void cal_tile_scan_kernel(
    const int64_t* __restrict__ tile_sum, int64_t n_tiles, int64_t* __restrict__ tile_off,
    int append, int64_t cap, const int64_t* __restrict__ old_size,
    const uint8_t* __restrict__ old_ovf, int64_t* __restrict__ new_size,
    uint8_t* __restrict__ new_ovf) {
  __shared__ int64_t buf[1024];
  __shared__ int64_t carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int64_t b = 0; b < n_tiles; b += 1024) {
    const int64_t v = b + tid < n_tiles ? tile_sum[b + tid] : 0;
    buf[tid] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int64_t t = tid >= off ? buf[tid - off] : 0;
      __syncthreads();
      buf[tid] += t;
      __syncthreads();
    }
    if (b + tid < n_tiles) tile_off[b + tid] = carry + buf[tid] - v;
    __syncthreads();
    if (tid == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (tid == 0) {
    const int64_t count = carry;
    if (append) {
      const int64_t s = *old_size;
      const bool ovf = (*old_ovf != 0) || (s + count > cap);
      *new_size = ovf ? s : s + count;
      *new_ovf = ovf ? 1 : 0;
    } else {
      *new_size = count;
    }
  }
}

// append: write the active candidates at old size + rank unless overflowed
template <typename T>
__global__ __launch_bounds__(kCalThreads) void cal_append_write_kernel(
    const uint8_t* __restrict__ active, int64_t n, const int64_t* __restrict__ tile_off,
    const int64_t* __restrict__ old_size, const uint8_t* __restrict__ new_ovf,
    const int64_t* __restrict__ c_slots, const int64_t* __restrict__ c_indices,
    const T* __restrict__ c_values, int64_t* __restrict__ slots, int64_t* __restrict__ indices,
    T* __restrict__ values, int64_t* __restrict__ dest) {
  __shared__ int64_t buf[kCalThreads];
  const int64_t base = (int64_t)blockIdx.x * kCalTile + threadIdx.x * kCalItems;
  int64_t c = 0;
  for (int k = 0; k < kCalItems; ++k)
    if (base + k < n && active[base + k]) ++c;
  int64_t total;
  int64_t rank = cal_block_exclusive(c, buf, &total) + tile_off[blockIdx.x] + *old_size;
  const bool ovf = *new_ovf != 0;
  for (int k = 0; k < kCalItems; ++k) {
    const int64_t i = base + k;
    if (i >= n) break;
    int64_t d = -1;
    if (active[i]) {
      if (!ovf) {
        slots[rank] = c_slots[i];
        indices[rank] = c_indices[i];
        values[rank] = c_values[i];
        d = rank;
      }
      ++rank;
    }
    if (dest) dest[i] = d;
  }
}

// pop: compact the kept entries into the fresh (zeroed) arrays
template <typename T>
__global__ __launch_bounds__(kCalThreads) void cal_pop_compact_kernel(
    const uint8_t* __restrict__ keep, int64_t cap, const int64_t* __restrict__ tile_off,
    const int64_t* __restrict__ slots, const int64_t* __restrict__ indices,
    const T* __restrict__ values, int64_t* __restrict__ n_slots,
    int64_t* __restrict__ n_indices, T* __restrict__ n_values, int64_t* __restrict__ dest,
    const int64_t* __restrict__ new_size) {
  __shared__ int64_t buf[kCalThreads];
  const int64_t base = (int64_t)blockIdx.x * kCalTile + threadIdx.x * kCalItems;
  // the output positions past the kept entries are zeroed here, so the
  // caller need not clear n_* first
  const int64_t kept = *new_size;
  for (int k = 0; k < kCalItems; ++k) {
    const int64_t q = base + k;
    if (q < cap && q >= kept) {
      n_slots[q] = 0;
      n_indices[q] = 0;
      n_values[q] = T(0);
    }
  }
  int64_t c = 0;
  for (int k = 0; k < kCalItems; ++k)
    if (base + k < cap && keep[base + k]) ++c;
  int64_t total;
  int64_t rank = cal_block_exclusive(c, buf, &total) + tile_off[blockIdx.x];
  for (int k = 0; k < kCalItems; ++k) {
    const int64_t p = base + k;
    if (p >= cap) break;
    int64_t d = -1;
    if (keep[p]) {
      n_slots[rank] = slots[p];
      n_indices[rank] = indices[p];
      n_values[rank] = values[p];
      d = rank++;
    }
    if (dest) dest[p] = d;
  }
}

// pop: one lane per run of equal sorted keys adds the run in order
template <typename T>
__global__ __launch_bounds__(256) void cal_reduce_kernel(const int32_t* __restrict__ key,
                                                         const T* __restrict__ val, int64_t cap,
                                                         int64_t width, T* __restrict__ payload) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cap) return;
  const int32_t k = key[i];
  if (k >= width || (i > 0 && key[i - 1] == k)) return;
  T acc = payload[k];
  for (int64_t j = i; j < cap && key[j] == k; ++j) acc = acc + val[j];
  payload[k] = acc;
}

// pop without a sort (the workspace of dn_calendar_pop_work_bytes): each due
// entry counts itself at its index and files its position in the index's hit
// row (up to kCalHit); the first hit lists the index.  Then one thread per
// listed index adds the index's values in position order (selection over
// its short row; a scan of the keys past kCalHit hits), so every payload
// entry sees the sequential scatter_add's sums, and clears the count.
constexpr int kCalHit = 16;

__global__ __launch_bounds__(256) void cal_hit_clear_kernel(int32_t* __restrict__ hcnt,
                                                            int64_t width,
                                                            int32_t* __restrict__ tcount) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = i; t < width; t += n) hcnt[t] = 0;
  if (i == 0) *tcount = 0;
}

__global__ __launch_bounds__(kCalThreads) void cal_mark_kernel(
    const int32_t* __restrict__ key, int64_t cap, int64_t width, int32_t* __restrict__ hcnt,
    int32_t* __restrict__ hpos, int32_t* __restrict__ tlist, int32_t* __restrict__ tcount) {
  const int64_t base = (int64_t)blockIdx.x * kCalTile + threadIdx.x * kCalItems;
#pragma unroll
  for (int k = 0; k < kCalItems; ++k) {
    const int64_t p = base + k;
    if (p >= cap) break;
    const int32_t t = key[p];
    if (t >= width) continue;
    const int s = atomicAdd(&hcnt[t], 1);
    if (s < kCalHit) hpos[(int64_t)t * kCalHit + s] = (int32_t)p;
    if (s == 0) tlist[atomicAdd(tcount, 1)] = t;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void cal_ordered_add_kernel(
    const int32_t* __restrict__ key, const T* __restrict__ val, int64_t cap,
    int32_t* __restrict__ hcnt, const int32_t* __restrict__ hpos,
    const int32_t* __restrict__ tlist, const int32_t* __restrict__ tcount,
    T* __restrict__ payload) {
  const int n = *tcount;
  const int nth = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += nth) {
    const int t = tlist[i];
    const int cnt = hcnt[t];
    T acc = payload[t];
    if (cnt <= kCalHit) {
      const int4* row = reinterpret_cast<const int4*>(hpos + (int64_t)t * kCalHit);
      int e[kCalHit];
#pragma unroll
      for (int q = 0; q < kCalHit / 4; ++q) {
        const int4 v = row[q];
        e[4 * q] = v.x;
        e[4 * q + 1] = v.y;
        e[4 * q + 2] = v.z;
        e[4 * q + 3] = v.w;
      }
#pragma unroll
      for (int q = 0; q < kCalHit; ++q) {
        const int m = -(int)(q < cnt);  // all ones for a filed hit
        e[q] = (e[q] & m) | (0x7fffffff & ~m);
      }
      unsigned used = 0;
      for (int o = 0; o < cnt; ++o) {
        int best = 0x7fffffff, bi = 0;
#pragma unroll
        for (int q = 0; q < kCalHit; ++q)
          if (!((used >> q) & 1u) && e[q] < best) {
            best = e[q];
            bi = q;
          }
        used |= 1u << bi;
        acc = acc + val[best];
      }
    } else {
      for (int64_t p = 0; p < cap; ++p)
        if (key[p] == t) acc = acc + val[p];
    }
    payload[t] = acc;
    hcnt[t] = 0;
  }
}

}  // namespace dn

namespace {

constexpr int64_t kAlign = 256;

int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

int end_bit_for(int64_t width) {
  int b = 1;
  while (b < 31 && ((int64_t)1 << b) <= width) ++b;
  return b;
}

template <typename T>
size_t sort_temp_bytes(int64_t cap) {
  size_t bytes = 0;
  if (cap > 0)
    hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr,
                                       (uint32_t*)nullptr, (const T*)nullptr, (T*)nullptr,
                                       (int)cap, 0, 31);
  return bytes;
}

struct CalWork {
  int64_t* tile_sum;
  int64_t* tile_off;
  uint8_t* flags;
  int32_t* key;
  int32_t* key_sorted;
  void* val_sorted;
  void* sort_temp;
  size_t sort_bytes;
};

// workspace: tile sums and offsets, then (pop) flags, keys, sorted pairs and
// the radix sort's temporary storage
template <typename T>
int64_t cal_work_bytes(int64_t n, int64_t cap) {
  const int64_t m = n > cap ? n : cap;
  const int64_t tiles = (m + dn::kCalTile - 1) / dn::kCalTile;
  int64_t b = 2 * align_up(8 * (tiles > 0 ? tiles : 1));
  b += align_up(cap > 0 ? cap : 1) + 2 * align_up(4 * (cap > 0 ? cap : 1)) +
       align_up((int64_t)sizeof(T) * (cap > 0 ? cap : 1));
  b += align_up((int64_t)sort_temp_bytes<T>(cap));
  return b;
}

template <typename T>
CalWork carve(void* work, int64_t n, int64_t cap) {
  const int64_t m = n > cap ? n : cap;
  const int64_t tiles = (m + dn::kCalTile - 1) / dn::kCalTile;
  char* p = static_cast<char*>(work);
  CalWork w;
  w.tile_sum = reinterpret_cast<int64_t*>(p);
  p += align_up(8 * (tiles > 0 ? tiles : 1));
  w.tile_off = reinterpret_cast<int64_t*>(p);
  p += align_up(8 * (tiles > 0 ? tiles : 1));
  w.flags = reinterpret_cast<uint8_t*>(p);
  p += align_up(cap > 0 ? cap : 1);
  w.key = reinterpret_cast<int32_t*>(p);
  p += align_up(4 * (cap > 0 ? cap : 1));
  w.key_sorted = reinterpret_cast<int32_t*>(p);
  p += align_up(4 * (cap > 0 ? cap : 1));
  w.val_sorted = p;
  p += align_up((int64_t)sizeof(T) * (cap > 0 ? cap : 1));
  w.sort_temp = p;
  w.sort_bytes = sort_temp_bytes<T>(cap);
  return w;
}

constexpr int64_t kCalMax = (int64_t)1 << 30;

// the pop's hit lists after the sort workspace: counts, list count, rows of
// kCalHit positions, the index list
int64_t cal_hit_bytes(int64_t width) {
  const int64_t w = width > 0 ? width : 1;
  return 2 * align_up(4 * w) + align_up(4) + align_up(4 * w * dn::kCalHit);
}

template <typename T>
int append_impl(int64_t* slots, int64_t* indices, T* values, int64_t cap,
                const int64_t* old_size, const uint8_t* old_ovf, int64_t* new_size,
                uint8_t* new_ovf, int64_t n, const int64_t* c_slots, const int64_t* c_indices,
                const T* c_values, const uint8_t* active, int64_t* dest, void* work,
                int64_t work_bytes, hipStream_t stream) {
  if (!old_size || !old_ovf || !new_size || !new_ovf || cap < 0 || n < 0 || cap > kCalMax ||
      n > kCalMax)
    return DN_EINVAL;
  if (!work || work_bytes < cal_work_bytes<T>(n, cap)) return DN_EINVAL;
  if (n > 0 && (!c_slots || !c_indices || !c_values || !active)) return DN_EINVAL;
  if (cap > 0 && (!slots || !indices || !values)) return DN_EINVAL;
  const CalWork w = carve<T>(work, n, cap);
  const int64_t tiles = (n + dn::kCalTile - 1) / dn::kCalTile;
  if (tiles > 0)
    hipLaunchKernelGGL(dn::cal_tile_count_kernel, dim3((unsigned)tiles), dim3(dn::kCalThreads),
                       0, stream, active, n, w.tile_sum);
  hipLaunchKernelGGL(dn::cal_tile_scan_kernel, dim3(1), dim3(1024), 0, stream, w.tile_sum,
                     tiles, w.tile_off, 1, cap, old_size, old_ovf, new_size, new_ovf);
  if (tiles > 0)
    hipLaunchKernelGGL((dn::cal_append_write_kernel<T>), dim3((unsigned)tiles),
                       dim3(dn::kCalThreads), 0, stream, active, n, w.tile_off, old_size,
                       new_ovf, c_slots, c_indices, c_values, slots, indices, values, dest);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
int pop_impl(const int64_t* slots, const int64_t* indices, const T* values, int64_t cap,
             const int64_t* size, const int64_t* slot, int64_t width, int64_t* n_slots,
             int64_t* n_indices, T* n_values, int64_t* new_size, T* payload, int32_t* keys_out,
             int64_t* keep_dest, void* work, int64_t work_bytes, hipStream_t stream) {
  if (!size || !slot || !new_size || cap < 0 || cap > kCalMax || width < 0 ||
      width >= INT32_MAX)
    return DN_EINVAL;
  if (!work || work_bytes < cal_work_bytes<T>(0, cap)) return DN_EINVAL;
  if (cap > 0 && (!slots || !indices || !values || !n_slots || !n_indices || !n_values))
    return DN_EINVAL;
  if (width > 0 && !payload) return DN_EINVAL;
  const CalWork w = carve<T>(work, 0, cap);
  const int64_t tiles = (cap + dn::kCalTile - 1) / dn::kCalTile;
  int32_t* key = keys_out ? keys_out : w.key;
  if (tiles > 0)
    hipLaunchKernelGGL(dn::cal_pop_flags_kernel, dim3((unsigned)tiles), dim3(dn::kCalThreads),
                       0, stream, slots, indices, cap, size, slot, width, w.flags, key,
                       w.tile_sum);
  hipLaunchKernelGGL(dn::cal_tile_scan_kernel, dim3(1), dim3(1024), 0, stream, w.tile_sum,
                     tiles, w.tile_off, 0, cap, nullptr, nullptr, new_size, nullptr);
  if (tiles == 0) return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
  hipLaunchKernelGGL((dn::cal_pop_compact_kernel<T>), dim3((unsigned)tiles),
                     dim3(dn::kCalThreads), 0, stream, w.flags, cap, w.tile_off, slots, indices,
                     values, n_slots, n_indices, n_values, keep_dest, new_size);
  if (width > 0 && work_bytes >= cal_work_bytes<T>(0, cap) + cal_hit_bytes(width)) {
    char* h = static_cast<char*>(work) + cal_work_bytes<T>(0, cap);
    int32_t* hcnt = reinterpret_cast<int32_t*>(h);
    h += align_up(4 * width);
    int32_t* tlist = reinterpret_cast<int32_t*>(h);
    h += align_up(4 * width);
    int32_t* tcount = reinterpret_cast<int32_t*>(h);
    h += align_up(4);
    int32_t* hpos = reinterpret_cast<int32_t*>(h);
    const unsigned clear_blocks = (unsigned)((width + 255) / 256 < 1024 ? (width + 255) / 256 : 1024);
    hipLaunchKernelGGL(dn::cal_hit_clear_kernel, dim3(clear_blocks), dim3(256), 0, stream, hcnt,
                       width, tcount);
    hipLaunchKernelGGL(dn::cal_mark_kernel, dim3((unsigned)tiles), dim3(dn::kCalThreads), 0,
                       stream, key, cap, width, hcnt, hpos, tlist, tcount);
    const int64_t lim = width < cap ? width : cap;
    const unsigned add_blocks = (unsigned)((lim + 255) / 256 < 1024 ? (lim + 255) / 256 : 1024);
    hipLaunchKernelGGL((dn::cal_ordered_add_kernel<T>), dim3(add_blocks), dim3(256), 0, stream,
                       key, values, cap, hcnt, hpos, tlist, tcount, payload);
  } else if (width > 0) {
    // keys are non-negative (index or width): sorted as unsigned
    const uint32_t* ukey = reinterpret_cast<const uint32_t*>(key);
    uint32_t* usorted = reinterpret_cast<uint32_t*>(w.key_sorted);
    const int end_bit = end_bit_for(width);
    size_t need = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, need, ukey, usorted, values,
                                       static_cast<T*>(w.val_sorted), (int)cap, 0, end_bit);
    if (need > w.sort_bytes) return DN_ERANGE;
    size_t bytes = w.sort_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(w.sort_temp, bytes, ukey, usorted, values,
                                           static_cast<T*>(w.val_sorted), (int)cap, 0, end_bit,
                                           stream) != hipSuccess)
      return DN_ELAUNCH;
    hipLaunchKernelGGL((dn::cal_reduce_kernel<T>), dim3((unsigned)((cap + 255) / 256)),
                       dim3(256), 0, stream, w.key_sorted, static_cast<const T*>(w.val_sorted),
                       cap, width, payload);
  }
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace

DN_EXPORT int64_t dn_calendar_pop_work_bytes(int64_t cap, int64_t width, int kind) {
  if (cap < 0 || cap > kCalMax || width < 0 || width >= INT32_MAX) return -1;
  switch (kind) {
    case 0: return cal_work_bytes<float>(0, cap) + cal_hit_bytes(width);
    case 1: return cal_work_bytes<double>(0, cap) + cal_hit_bytes(width);
    case 2: return cal_work_bytes<int32_t>(0, cap) + cal_hit_bytes(width);
    default: return -1;
  }
}

DN_EXPORT int64_t dn_calendar_work_bytes(int64_t n, int64_t cap, int kind) {
  if (n < 0 || cap < 0 || n > kCalMax || cap > kCalMax) return -1;
  switch (kind) {
    case 0: return cal_work_bytes<float>(n, cap);
    case 1: return cal_work_bytes<double>(n, cap);
    case 2: return cal_work_bytes<int32_t>(n, cap);
    default: return -1;
  }
}

#define DN_CALENDAR_EXPORTS(T, SFX)                                                          \
  DN_EXPORT int dn_calendar_append_##SFX(                                                    \
      int64_t* slots, int64_t* indices, T* values, int64_t cap, const int64_t* old_size,     \
      const uint8_t* old_ovf, int64_t* new_size, uint8_t* new_ovf, int64_t n,                \
      const int64_t* c_slots, const int64_t* c_indices, const T* c_values,                   \
      const uint8_t* active, int64_t* dest, void* work, int64_t work_bytes,                  \
      hipStream_t stream) {                                                                  \
    return append_impl<T>(slots, indices, values, cap, old_size, old_ovf, new_size, new_ovf, \
                          n, c_slots, c_indices, c_values, active, dest, work, work_bytes,   \
                          stream);                                                           \
  }                                                                                          \
  DN_EXPORT int dn_calendar_pop_##SFX(                                                       \
      const int64_t* slots, const int64_t* indices, const T* values, int64_t cap,            \
      const int64_t* size, const int64_t* slot, int64_t width, int64_t* n_slots,             \
      int64_t* n_indices, T* n_values, int64_t* new_size, T* payload, int32_t* keys_out,     \
      int64_t* keep_dest, void* work, int64_t work_bytes, hipStream_t stream) {              \
    return pop_impl<T>(slots, indices, values, cap, size, slot, width, n_slots, n_indices,   \
                       n_values, new_size, payload, keys_out, keep_dest, work, work_bytes,   \
                       stream);                                                              \
  }

DN_CALENDAR_EXPORTS(float, f32)
DN_CALENDAR_EXPORTS(double, f64)
DN_CALENDAR_EXPORTS(int32_t, i32)
