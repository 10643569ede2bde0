// dn_net.hip -- event-driven network step: spike detection, delay-ring
// delivery and the population update with exp2syn synapses.
//
// Semantics follow the reference's dense NetCon inference path
// (dendra/models/networks/netcon.py:5183-5205, _event_delivery.py:30-199),
// stepped in Network.step order (networks/net.py:335-377):
//   1. NetStim sources fire (here: counter-RNG Poisson, p = rate * dt),
//   2. the ring row `cur` is delivered to the synapses
//      (exp2syn.net_receive_values: A += w f, B += w f, mod/EXP2SYN.py:154-156)
//      and cleared; cell sources are detected on the previous step's voltage
//      (rising edge v >= theta, networks/spiking.py:7-10,43-73); every
//      connection e of a spiking source adds w_e to ring[(cur + delay_e) % D]
//      [post_e], in connection order; cur advances,
//   3. the populations advance: synapse decay (cnexp exact exponential),
//      HH gates, currents including i_syn = (B - A)(v - e) (nA), solve.
//
// Delivery is O(events): a spiking source walks its connections (by-source
// CSR) and files each connection index at its target (a per-target hit list
// of up to kHitMax entries, slots from a per-target counter); one thread per
// hit target then adds its hits in connection order (the list sorted), which
// is the order of the reference's float32 scatter_add, bit for bit.  A
// target with more hits in one step than the list holds scans its incoming
// segment (connections are numbered target-major) against the step's spike
// bitmask instead.  All per-step state (cur slot, step index, list lengths) lives
// in device memory, so a sequence of steps can be captured as a HIP graph.
#include "dn_wave.hpp"

namespace dn {

// splitmix64 counter RNG (identical in the host generators and the oracle)
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t stream, uint64_t i) {
  const uint64_t key = mix64((seed & 0xFFFFFFFFull) | ((stream & 0xFFFFFFFFull) << 32));
  return (double)(mix64(key + i) >> 11) * 0x1.0p-53;
}

struct NetArgs {
  // cells (all unbranched cables; point cells have K = 1), flat in plan order
  int n_cells;
  const int64_t* off;      // n_cells + 1
  const int32_t* soma;     // flat index of each cell's synapse/detector site
  float *v, *m, *h, *n;
  const float *gL, *gRp, *scale, *sna, *sk, *area;
  float dt;
  HHParams<float> hh;
  // synapses: one exp2syn per cell at its soma
  float *A, *B;
  float factor, decay1, decay2, e_syn;
  // sources: cells [0, n_cells) then Poisson sources [n_cells, n_src)
  int n_src;
  uint8_t* latch;
  float theta;
  uint64_t seed;
  uint32_t bg_stream;
  double bg_p;
  // connectivity
  const int32_t* out_ptr;  // by source, n_src + 1
  const int32_t* out_tgt;
  const int32_t* out_e;    // connection index of each by-source entry
  const int32_t* in_ptr;   // by target (target-major numbering), n_cells + 1
  const int32_t* pre;      // per connection
  const int32_t* delay;    // steps, 1 <= delay < depth
  const float* w;
  int depth;
  float* ring;             // [depth][n_cells]
  // per-step scratch
  uint32_t* spk_bits;      // ceil(n_src / 32) words
  int32_t* spk_list;       // n_src
  int32_t* hit_cnt;        // n_cells: hits filed this step (zero between steps)
  int32_t* hit_e;          // [n_cells][kHitMax] connection indices of the hits
  int32_t* tgt_list;       // n_cells
  int32_t* ctr;            // [0] cur, [1] step, [2] n_spikes, [3] n_active
  // optional records
  int32_t* ev_log;         // (step, source) pairs
  int32_t* ev_count;
  int64_t ev_cap;
  float* delivered;        // [n_rec_steps][n_cells] payload at delivery
  int64_t rec_steps;
  const uint8_t* tape;     // delivery-only mode: [n_steps][n_src] source spikes
};

// Rising-edge detection of source i at this step (cells on the voltage
// before their update, Poisson sources by the counter RNG, or a tape); a
// spike sets the source's bit and appends it to the step's list.
__device__ __forceinline__ void net_detect(const NetArgs& a, int i, int step) {
  bool spike;
  if (a.tape) {
    spike = a.tape[(int64_t)step * a.n_src + i] != 0;
  } else if (i < a.n_cells) {
    const bool above = a.v[a.soma[i]] >= a.theta;
    spike = above && !a.latch[i];
    a.latch[i] = above;
  } else {
    spike = uniform01(a.seed, a.bg_stream,
                      (uint64_t)step * (uint64_t)(a.n_src - a.n_cells) +
                          (uint64_t)(i - a.n_cells)) < a.bg_p;
  }
  if (spike) {
    atomicOr(&a.spk_bits[i >> 5], 1u << (i & 31));
    const int k = atomicAdd(&a.ctr[2], 1);
    a.spk_list[k] = i;
    if (a.ev_log) {
      const int q = atomicAdd(a.ev_count, 1);
      if (q < a.ev_cap) {
        a.ev_log[2 * (int64_t)q] = step;
        a.ev_log[2 * (int64_t)q + 1] = i;
      }
    }
  }
}

// Delivery of ring row cur to target i: the payload is read, the cell
// cleared and (optionally) recorded; returns the payload.
__device__ __forceinline__ float net_take(const NetArgs& a, int i, int cur, int step) {
  float* row = a.ring + (int64_t)cur * a.n_cells;
  const float pay = row[i];
  row[i] = 0.0f;
  if (a.delivered && step < a.rec_steps) a.delivered[(int64_t)step * a.n_cells + i] = pay;
  return pay;
}

// Ka: deliver ring row cur to the synapses and detect this step's sources.
__global__ __launch_bounds__(256) void net_deliver_detect(NetArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int cur = a.ctr[0];
  const int step = a.ctr[1];
  if (i < a.n_cells) {
    const float pay = net_take(a, i, cur, step);
    if (a.A) {
      const float wf = pay * a.factor;
      a.A[i] = a.A[i] + wf;
      a.B[i] = a.B[i] + wf;
    }
  }
  if (i < a.n_src) net_detect(a, i, step);
}

constexpr int kHitMax = 16;  // hits a target's list holds per step

// Kb: every spiking source files its connections at their targets; the first
// hit of a target appends it to the active list.  One wavefront per spiking
// source.
__global__ __launch_bounds__(256) void net_mark(NetArgs a) {
  const int lane = threadIdx.x & 63;
  const int wv = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int n_wv = (gridDim.x * blockDim.x) >> 6;
  const int ns = a.ctr[2];
  for (int k = wv; k < ns; k += n_wv) {
    const int src = a.spk_list[k];
    const int q1 = a.out_ptr[src + 1];
    // four entries per lane and pass: their loads, then their counter
    // increments, are in flight together
    for (int qb = a.out_ptr[src] + lane; qb < q1; qb += 4 * 64) {
      int t[4], e[4], slot[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int q = qb + 64 * u;
        t[u] = q < q1 ? a.out_tgt[q] : -1;
        e[u] = q < q1 ? a.out_e[q] : 0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) slot[u] = t[u] >= 0 ? atomicAdd(&a.hit_cnt[t[u]], 1) : -1;
      // the entries' first hits take one run of the target list per lane
      // (one counter increment, not one per entry)
      int nf = 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) nf += (t[u] >= 0 && slot[u] == 0);
      int at = nf ? atomicAdd(&a.ctr[3], nf) : 0;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (t[u] < 0) continue;
        if (slot[u] < kHitMax) a.hit_e[(int64_t)t[u] * kHitMax + slot[u]] = e[u];
        if (slot[u] == 0) a.tgt_list[at++] = t[u];
      }
    }
  }
}

// Kc: one thread per hit target adds its hits in connection order: the
// filed connection indices taken smallest first, or, when more hits came
// than the list holds, a scan of the target's incoming segment against the
// step's spike bits.  The target's counter is cleared for the next step.
__device__ __forceinline__ void net_enqueue_body(const NetArgs& a) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nth = gridDim.x * blockDim.x;
  const int na = a.ctr[3];
  const int cur = a.ctr[0];
  for (int k = tid; k < na; k += nth) {
    const int t = a.tgt_list[k];
    // the count and the whole hit row are loaded together (the row read
    // whole and masked by the count, so no load waits on the count)
    const int cnt = a.hit_cnt[t];
    const int4* row = reinterpret_cast<const int4*>(a.hit_e + (int64_t)t * kHitMax);
    int4 rv[kHitMax / 4];
#pragma unroll
    for (int i = 0; i < kHitMax / 4; ++i) rv[i] = row[i];
    // keeps the row's loads beside the count's (else they sink into the
    // branch that uses them and start only after the count arrives)
    asm volatile("" ::"v"(cnt), "v"(rv[0].x), "v"(rv[0].y), "v"(rv[0].z), "v"(rv[0].w),
                 "v"(rv[1].x), "v"(rv[1].y), "v"(rv[1].z), "v"(rv[1].w), "v"(rv[2].x),
                 "v"(rv[2].y), "v"(rv[2].z), "v"(rv[2].w), "v"(rv[3].x), "v"(rv[3].y),
                 "v"(rv[3].z), "v"(rv[3].w));
    auto add = [&](int e) {
      int slot = cur + a.delay[e];
      if (slot >= a.depth) slot -= a.depth;
      float* q = a.ring + (int64_t)slot * a.n_cells + t;
      *q = *q + a.w[e];
    };
    if (cnt <= kHitMax) {
      int e[kHitMax];
#pragma unroll
      for (int i = 0; i < kHitMax / 4; ++i) {
        e[4 * i] = rv[i].x;
        e[4 * i + 1] = rv[i].y;
        e[4 * i + 2] = rv[i].z;
        e[4 * i + 3] = rv[i].w;
      }
#pragma unroll
      for (int i = 0; i < kHitMax; ++i) {
        const int m = -(int)(i < cnt);  // all ones for a filed hit
        e[i] = (e[i] & m) | (0x7fffffff & ~m);
      }
      unsigned used = 0;
      for (int o = 0; o < cnt; ++o) {
        int best = 0x7fffffff, bi = 0;
#pragma unroll
        for (int i = 0; i < kHitMax; ++i)
          if (!((used >> i) & 1u) && e[i] < best) {
            best = e[i];
            bi = i;
          }
        used |= 1u << bi;
        add(best);
      }
    } else {
      for (int ee = a.in_ptr[t]; ee < a.in_ptr[t + 1]; ++ee) {
        const int s = a.pre[ee];
        if ((a.spk_bits[s >> 5] >> (s & 31)) & 1u) add(ee);
      }
    }
    a.hit_cnt[t] = 0;
  }
}

__global__ __launch_bounds__(256) void net_enqueue(NetArgs a) { net_enqueue_body(a); }

// Ke: advance the step counters and clear the step's spike bits.
__device__ __forceinline__ void net_advance_body(const NetArgs& a) {
  const int ns = a.ctr[2];
  for (int k = threadIdx.x; k < ns; k += blockDim.x) {
    const int s = a.spk_list[k];
    a.spk_bits[s >> 5] = 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int c = a.ctr[0] + 1;
    a.ctr[0] = c >= a.depth ? 0 : c;
    a.ctr[1] += 1;
    a.ctr[2] = 0;
    a.ctr[3] = 0;
  }
}

__global__ void net_advance(NetArgs a) { net_advance_body(a); }

// Kc + Ke: the enqueue, then the workgroup that finishes last (a counter in
// ctr[4] that each workgroup increments once all its threads are done; no
// workgroup waits on it) advances the counters and clears the spike bits.
// The last workgroup reads only what earlier launches wrote (the spike list
// and counts), every other workgroup has read the counters before its
// increment, and the ring and flag writes reach later launches through the
// launch boundary, so no device-wide fence is needed (one per workgroup
// would write back the L2 of its XCD).
__global__ __launch_bounds__(256) void net_enqueue_advance(NetArgs a) {
  net_enqueue_body(a);
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(&a.ctr[4], 1) == (int)gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  net_advance_body(a);
  if (threadIdx.x == 0) a.ctr[4] = 0;
}

// Ka + Kd: one launch delivers today's ring row to every synapse, detects
// every source on the voltage before its update, and advances the cells.
// Blocks [0, pb) own the point cells (one thread each), blocks [pb, pb + cb)
// the multi-compartment cells (one wavefront each), the rest the Poisson
// sources.  The synapse update is the two kernels' arithmetic: A = (A + w f)
// decay1, B = (B + w f) decay2.
template <int NR>
__global__ __launch_bounds__(256) void net_step_cells(NetArgs a, int n_point, int pb, int cb) {
  const int cur = a.ctr[0];
  const int step = a.ctr[1];
  if ((int)blockIdx.x < pb) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_point) return;
    const float wf = net_take(a, c, cur, step) * a.factor;
    net_detect(a, c, step);
    const int64_t i = a.off[c];
    const float A = (a.A[c] + wf) * a.decay1, B = (a.B[c] + wf) * a.decay2;
    a.A[c] = A;
    a.B[c] = B;
    const float v = a.v[i];
    float m = a.m[i], h = a.h[i], n = a.n[i], G, J;
    hh_advance<float>(v, a.dt, a.hh, a.sna ? a.sna[i] : 1.0f, a.sk ? a.sk[i] : 1.0f, m, h,
                      n, G, J);
    a.m[i] = m; a.h[i] = h; a.n[i] = n;
    const float gs = (B - A) * 1e-6f / a.area[i];  // uS -> S/cm^2
    G += gs;
    J += gs * a.e_syn;
    const float sc = a.scale[i];
    const float dts = a.dt * 1e-3f;
    const float f = (J - G * v) * sc;
    a.v[i] = v + dts * f / (1.0f + dts * G * sc);
    return;
  }
  if ((int)blockIdx.x >= pb + cb) {
    const int i = a.n_cells + (blockIdx.x - pb - cb) * blockDim.x + threadIdx.x;
    if (i < a.n_src) net_detect(a, i, step);
    return;
  }
  const int lane = threadIdx.x & 63;
  const int c = n_point + (((blockIdx.x - pb) * blockDim.x + threadIdx.x) >> 6);
  if (c >= a.n_cells) return;
  const int64_t o = a.off[c];
  const int K = (int)(a.off[c + 1] - o);
  const int nr = (K + 63) / 64;
  __shared__ float sh[4][4 * 64 * NR];
  float* base = sh[threadIdx.x >> 6];
  float *sa = base, *sb = base + 64 * NR, *sc_ = base + 128 * NR, *sd = base + 192 * NR;
  const float dts = a.dt * 1e-3f;
  float As = 0.0f, Bs = 0.0f;
  if (lane == 0) {
    const float wf = net_take(a, c, cur, step) * a.factor;
    net_detect(a, c, step);  // before any lane writes the cell's voltage
    As = (a.A[c] + wf) * a.decay1;
    Bs = (a.B[c] + wf) * a.decay2;
    a.A[c] = As;
    a.B[c] = Bs;
  }
  const float gsyn_uS = __shfl(Bs - As, 0);
  for (int j = lane; j < 64 * nr; j += 64) {
    if (j < K) {
      const int64_t i = o + j;
      const float v = a.v[i];
      float m = a.m[i], h = a.h[i], n = a.n[i], G, J;
      hh_advance<float>(v, a.dt, a.hh, a.sna ? a.sna[i] : 1.0f, a.sk ? a.sk[i] : 1.0f, m,
                        h, n, G, J);
      a.m[i] = m; a.h[i] = h; a.n[i] = n;
      if (o + j == a.soma[c]) {
        const float gs = gsyn_uS * 1e-6f / a.area[i];
        G += gs;
        J += gs * a.e_syn;
      }
      const float gl = a.gL[i], gr = a.gRp[i], s = a.scale[i];
      const float vr = j + 1 < K ? a.v[i + 1] : v;
      const float vl = j > 0 ? a.v[i - 1] : v;
      sa[j] = -dts * gr;
      sb[j] = 1.0f + dts * (gl + gr + G * s);
      sc_[j] = -dts * gl;
      sd[j] = dts * ((J - G * v) * s + gl * (vr - v) - gr * (v - vl));
    } else {
      sa[j] = 0.0f; sb[j] = 1.0f; sc_[j] = 0.0f; sd[j] = 0.0f;
    }
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  wave_hybrid_solve<float>(sa, sb, sc_, sd, nr);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  __builtin_amdgcn_wave_barrier();
  for (int j = lane; j < K; j += 64) a.v[o + j] += sd[j];
}

constexpr int kNetCableNR = 4;  // multi-compartment cells up to 256 compartments

// Three launches per step: cells and sources (Ka + Kd), marking (Kb), and
// the enqueue with the counter advance (Kc + Ke).  The cells' update writes
// only cell state, which marking and enqueueing never read, so it may run
// before them.
static void launch_step(const NetArgs& a, int n_point, int n_cable, int mark_blocks,
                        hipStream_t st) {
  const int pb = (n_point + 255) / 256, cb = (n_cable + 3) / 4;
  const int sb = (a.n_src - a.n_cells + 255) / 256;
  hipLaunchKernelGGL(net_step_cells<kNetCableNR>, dim3(pb + cb + sb), dim3(256), 0, st, a,
                     n_point, pb, cb);
  hipLaunchKernelGGL(net_mark, dim3(mark_blocks), dim3(256), 0, st, a);
  hipLaunchKernelGGL(net_enqueue_advance, dim3(mark_blocks), dim3(256), 0, st, a);
}

}  // namespace dn

// Advance the network n_steps.  Cells [0, n_point) are point cells (K = 1),
// cells [n_point, n_cells) are multi-compartment (K <= 256).  hh (host):
// {gnabar, gkbar, gl, ena, ek, el, q10}; syn (host): {factor, tau1_eff,
// tau2, e}.  ctr (8 ints) holds {cur, step, 0, 0, 0, ...} on entry.  graph_steps > 0 captures
// that many steps once as a HIP graph and replays it (n_steps must be a
// multiple); 0 launches every kernel directly.
DN_EXPORT int dn_net_run_f32(
    int n_cells, int n_point, const int64_t* off, const int32_t* soma, float* v, float* m,
    float* h, float* n, const float* gL, const float* gRp, const float* scale,
    const float* sna, const float* sk, const float* area, float dt, const float* hh,
    float* A, float* B, const float* syn, int n_src, uint8_t* latch, float theta,
    uint64_t seed, uint32_t bg_stream, double bg_p, const int32_t* out_ptr,
    const int32_t* out_tgt, const int32_t* out_e, const int32_t* in_ptr, const int32_t* pre,
    const int32_t* delay, const float* w, int depth, float* ring, uint32_t* spk_bits,
    int32_t* spk_list, int32_t* hit_cnt, int32_t* hit_e, int32_t* tgt_list, int32_t* ctr,
    int32_t* ev_log, int32_t* ev_count, int64_t ev_cap, float* delivered,
    int64_t rec_steps, int64_t n_steps, int graph_steps, hipStream_t stream) {
  if (n_cells < 1 || n_point < 0 || n_point > n_cells || n_src < n_cells || !off || !soma ||
      !v || !m || !h || !n || !gL || !gRp || !scale || !area || !hh || !A || !B || !syn ||
      !latch || !out_ptr || !out_tgt || !out_e || !in_ptr || !pre || !delay || !w ||
      depth < 2 || !ring || !spk_bits || !spk_list || !hit_cnt || !hit_e || !tgt_list || !ctr)
    return DN_EINVAL;
  if (reinterpret_cast<uintptr_t>(hit_e) % 16) return DN_EINVAL;  // rows read as int4
  if (graph_steps < 0 || (graph_steps > 0 && n_steps % graph_steps != 0)) return DN_EINVAL;
  dn::NetArgs a;
  a.n_cells = n_cells; a.off = off; a.soma = soma; a.v = v; a.m = m; a.h = h; a.n = n;
  a.gL = gL; a.gRp = gRp; a.scale = scale; a.sna = sna; a.sk = sk; a.area = area;
  a.dt = dt; a.hh = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};
  a.A = A; a.B = B; a.factor = syn[0];
  a.decay1 = expf(-dt / syn[1]);
  a.decay2 = expf(-dt / syn[2]);
  a.e_syn = syn[3];
  a.n_src = n_src; a.latch = latch; a.theta = theta; a.seed = seed;
  a.bg_stream = bg_stream; a.bg_p = bg_p; a.out_ptr = out_ptr; a.out_tgt = out_tgt;
  a.out_e = out_e; a.in_ptr = in_ptr; a.pre = pre; a.delay = delay; a.w = w;
  a.depth = depth; a.ring = ring; a.spk_bits = spk_bits; a.spk_list = spk_list;
  a.hit_cnt = hit_cnt; a.hit_e = hit_e; a.tgt_list = tgt_list; a.ctr = ctr; a.ev_log = ev_log; a.ev_count = ev_count;
  a.ev_cap = ev_cap; a.delivered = delivered; a.rec_steps = rec_steps;
  a.tape = nullptr;
  const int n_point_cells = n_point, n_cable = n_cells - n_point;
  const int mark_blocks = 256;
  if (graph_steps == 0) {
    for (int64_t s = 0; s < n_steps; ++s)
      dn::launch_step(a, n_point_cells, n_cable, mark_blocks, stream);
    return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
  }
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return DN_ELAUNCH;
  for (int s = 0; s < graph_steps; ++s)
    dn::launch_step(a, n_point_cells, n_cable, mark_blocks, stream);
  if (hipStreamEndCapture(stream, &graph) != hipSuccess) return DN_ELAUNCH;
  if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
    hipGraphDestroy(graph);
    return DN_ELAUNCH;
  }
  int rc = DN_OK;
  for (int64_t s = 0; s < n_steps; s += graph_steps)
    if (hipGraphLaunch(exec, stream) != hipSuccess) { rc = DN_ELAUNCH; break; }
  // the executable graph may be destroyed while launches are pending
  hipGraphExecDestroy(exec);
  hipGraphDestroy(graph);
  return rc;
}

// Delivery only, driven by a spike tape: replaces the reference's bitpacked
// source history + build_delivery native path
// (networks/netcon_bitpack_kernel.cu:20-126, netcon_bitpack_kernel.cpp:32-149)
// and the dense ring it reconstructs.  For each of n_steps steps: the ring
// row `cur` is written to delivered[step][*] and cleared, then every
// connection of a source marked in tape[step][*] adds w to
// ring[(cur + delay) % depth][post] in connection order.  Connections must be
// numbered target-major (in_ptr) with the by-source CSR out_ptr/out_tgt.
DN_EXPORT int dn_net_deliver_tape_f32(
    int n_targets, int n_src, const int32_t* out_ptr, const int32_t* out_tgt,
    const int32_t* out_e, const int32_t* in_ptr, const int32_t* pre, const int32_t* delay,
    const float* w, int depth, float* ring, uint32_t* spk_bits, int32_t* spk_list,
    int32_t* hit_cnt, int32_t* hit_e, int32_t* tgt_list, int32_t* ctr, const uint8_t* tape,
    float* delivered, int64_t n_steps, hipStream_t stream) {
  if (n_targets < 1 || n_src < 1 || !out_ptr || !out_tgt || !out_e || !in_ptr || !pre ||
      !delay || !w || depth < 2 || !ring || !spk_bits || !spk_list || !hit_cnt || !hit_e ||
      !tgt_list || !ctr || !tape || !delivered)
    return DN_EINVAL;
  if (reinterpret_cast<uintptr_t>(hit_e) % 16) return DN_EINVAL;  // rows read as int4
  dn::NetArgs a = {};
  a.n_cells = n_targets;
  a.n_src = n_src;
  a.out_ptr = out_ptr; a.out_tgt = out_tgt; a.out_e = out_e; a.in_ptr = in_ptr;
  a.pre = pre; a.delay = delay; a.w = w; a.depth = depth; a.ring = ring;
  a.spk_bits = spk_bits; a.spk_list = spk_list; a.hit_cnt = hit_cnt; a.hit_e = hit_e;
  a.tgt_list = tgt_list; a.ctr = ctr;
  a.delivered = delivered; a.rec_steps = n_steps; a.tape = tape;
  const int nblk = ((n_src > n_targets ? n_src : n_targets) + 255) / 256;
  for (int64_t s = 0; s < n_steps; ++s) {
    hipLaunchKernelGGL(dn::net_deliver_detect, dim3(nblk), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(dn::net_mark, dim3(256), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(dn::net_enqueue, dim3(256), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(dn::net_advance, dim3(1), dim3(256), 0, stream, a);
  }
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}
