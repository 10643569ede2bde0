/*
 * dn_oracle_net.c -- CPU restatement of Dendra's event delivery (oracle).
 *
 * TEST INFRASTRUCTURE ONLY (see dn_oracle.c).
 *
 * Dense delay-ring semantics of the reference's NetCon inference path:
 *   NetCon.advance_non_diff_dense_{uniform,mixed}  networks/netcon.py:5183-5205
 *   dense_due / _clear_dense_slot / _enqueue_dense_values
 *                                         networks/_event_delivery.py:30-199
 *   threshold_transition (rising edge, equality counts as above)
 *                                         networks/spiking.py:7-10,43-73
 *   delay discretisation round-half-even, clamp >= 1, depth = max + 1
 *                                         networks/netcon.py:2694-2709,3930
 *   exp2syn.net_receive_values: A += w*factor, B += w*factor
 *                                         mod/EXP2SYN.py:154-156
 * Per step: read row `cur` (the payload), clear it, detect sources, then
 * scatter-add w_e * gate_e into row (cur + delay_e) mod depth at post_e in
 * connection order (PyTorch's CPU scatter_add accumulates sequentially), then
 * cur = (cur + 1) mod depth.  All ring arithmetic is float32, as in the
 * reference's float32 networks; adding the +0.0 of a silent connection is an
 * identity, so only spiking connections are visited.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* Round-half-even delay in steps, clamped to >= 1 (netcon.py:3930,2776-2778). */
ORC_API void orc_delay_steps(const double* delay_ms, double dt, int64_t n,
                             int64_t* steps) {
  for (int64_t i = 0; i < n; ++i) {
    double q = delay_ms[i] / dt;
    double r = nearbyint(q); /* default rounding mode: half to even */
    int64_t s = (int64_t)r;
    steps[i] = s < 1 ? 1 : s;
  }
}

/* Replay a spike tape through the dense ring.
 * spikes: [n_steps][n_src] bytes (nonzero = source spiked this step)
 * delivered: [n_steps][n_syn] payload read at each step before clearing
 * ring (in/out): [depth][n_syn], *cur (in/out) */
ORC_API int orc_ring_replay_f32(int64_t n_src, int64_t n_syn, int64_t n_conn,
                                const int32_t* pre, const int32_t* post,
                                const int32_t* delay, const float* w, int64_t depth,
                                int64_t n_steps, const uint8_t* spikes,
                                float* delivered, float* ring, int64_t* cur) {
  for (int64_t e = 0; e < n_conn; ++e) {
    if (pre[e] < 0 || pre[e] >= n_src || post[e] < 0 || post[e] >= n_syn) return -1;
    if (delay[e] < 1 || delay[e] >= depth) return -1;
  }
  int64_t c = *cur;
  for (int64_t s = 0; s < n_steps; ++s) {
    float* row = ring + c * n_syn;
    memcpy(delivered + s * n_syn, row, sizeof(float) * (size_t)n_syn);
    memset(row, 0, sizeof(float) * (size_t)n_syn);
    const uint8_t* sp = spikes + s * n_src;
    for (int64_t e = 0; e < n_conn; ++e) {
      if (!sp[pre[e]]) continue;
      int64_t slot = (c + delay[e]) % depth;
      float* q = ring + slot * n_syn + post[e];
      *q = *q + w[e];
    }
    c = (c + 1) % depth;
  }
  *cur = c;
  return 0;
}

/* Rising-edge detector over one step: spike = (v >= theta) & !latch. */
ORC_API void orc_detect_f32(int64_t n, const float* v, float theta, uint8_t* latch,
                            uint8_t* spike) {
  for (int64_t i = 0; i < n; ++i) {
    uint8_t above = v[i] >= theta;
    spike[i] = above && !latch[i];
    latch[i] = above;
  }
}

/* Ring replay with per-connection gate values (value = w_e * gate_e), the
 * general form of _enqueue_dense_values used by scheduled events
 * (netcon.py schedule -> dense_fixed_delay_transition).  gates: [n_steps][n_conn]. */
ORC_API int orc_ring_replay_gates_f32(int64_t n_syn, int64_t n_conn, const int32_t* post,
                                      const int32_t* delay, const float* w, int64_t depth,
                                      int64_t n_steps, const float* gates,
                                      float* delivered, float* ring, int64_t* cur) {
  for (int64_t e = 0; e < n_conn; ++e) {
    if (post[e] < 0 || post[e] >= n_syn) return -1;
    if (delay[e] < 1 || delay[e] >= depth) return -1;
  }
  int64_t c = *cur;
  for (int64_t s = 0; s < n_steps; ++s) {
    float* row = ring + c * n_syn;
    memcpy(delivered + s * n_syn, row, sizeof(float) * (size_t)n_syn);
    memset(row, 0, sizeof(float) * (size_t)n_syn);
    const float* g = gates + s * n_conn;
    for (int64_t e = 0; e < n_conn; ++e) {
      float val = w[e] * g[e];
      if (g[e] == 0.0f) continue; /* adding +-0 leaves the slot unchanged */
      int64_t slot = (c + delay[e]) % depth;
      float* q = ring + slot * n_syn + post[e];
      *q = *q + val;
    }
    c = (c + 1) % depth;
  }
  *cur = c;
  return 0;
}

/* ---- splitmix64 counter RNG (same stream as the host and the GPU) ---- */
static uint64_t orc_mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
ORC_API double orc_uniform(uint64_t seed, uint64_t stream, uint64_t i) {
  uint64_t key = orc_mix64((seed & 0xFFFFFFFFull) | ((stream & 0xFFFFFFFFull) << 32));
  return (double)(orc_mix64(key + i) >> 11) * (1.0 / 9007199254740992.0);
}

/* HH pieces shared with dn_oracle.c */
void orc_hh_advance(double v, double dt, const double* hhp, double sna, double sk,
                    double* m, double* h, double* n, double* G, double* J);
void orc_thomas(const double* a, const double* b, const double* c, const double* d,
                double* x, int64_t B, int64_t K);

#define ORC_PI_NET 3.14159265358979323846

/*
 * Binary64 network run (the oracle of the GPU network step).  Cells are
 * unbranched cables in the _bwd_euler_ub form (point cells: K = 1), one
 * exp2syn per cell at flat index soma[c] (mod/EXP2SYN.py: A, B decay by
 * exp(-dt/tau), i = (B - A)(v - e) nA).  Sources [0, n_cells) are the cells'
 * somata (rising edge at theta on the previous step's voltage), sources
 * [n_cells, n_src) Poisson with p = bg_p per step from the counter RNG.
 * Connections are processed in index order.  Step order as net.py:335-377.
 * spikes (out, nullable): [n_steps][n_src] bytes.
 */
ORC_API int orc_network_run(int64_t n_cells, const int64_t* off, const int64_t* soma,
                            const double* diam, const double* dx, const double* rhoa,
                            const double* cm, const double* sna, const double* sk,
                            const double* hhp, const double* syn, double theta,
                            int64_t n_src, uint64_t seed, uint32_t bg_stream, double bg_p,
                            int64_t n_conn, const int32_t* pre, const int32_t* post,
                            const int32_t* delay, const double* w, int64_t depth,
                            double dt, int64_t n_steps, double* V, double* M, double* H,
                            double* N, double* A, double* B, uint8_t* latch,
                            uint8_t* spikes) {
  int64_t total = off[n_cells];
  double dt_s = dt * 1e-3;
  double factor = syn[0], d1 = exp(-dt / syn[1]), d2 = exp(-dt / syn[2]), esyn = syn[3];
  double* gL = calloc((size_t)total, sizeof(double));
  double* gR = calloc((size_t)total, sizeof(double));
  double* scale = calloc((size_t)total, sizeof(double));
  double* area = calloc((size_t)total, sizeof(double));
  double* ring = calloc((size_t)(depth * n_cells), sizeof(double));
  uint8_t* sp = calloc((size_t)n_src, 1);
  int64_t kmax = 1;
  for (int64_t c = 0; c < n_cells; ++c)
    if (off[c + 1] - off[c] > kmax) kmax = off[c + 1] - off[c];
  double* wk = malloc(sizeof(double) * (size_t)kmax * 6);
  if (!gL || !gR || !scale || !area || !ring || !sp || !wk) return -2;
  for (int64_t c = 0; c < n_cells; ++c) {
    int64_t o = off[c], K = off[c + 1] - o;
    for (int64_t j = 0; j < K; ++j) {
      area[o + j] = diam[o + j] * 1e-4 * ORC_PI_NET * dx[o + j] * 1e-4;
      scale[o + j] = area[o + j] / (1e-6 * cm[o + j] * area[o + j]);
    }
    for (int64_t j = 0; j + 1 < K; ++j) {
      double r0 = 1e-4 * diam[o + j] / 2, r1 = 1e-4 * diam[o + j + 1] / 2;
      double R0 = rhoa[o + j] * 1e-4 * dx[o + j] / (ORC_PI_NET * r0 * r0);
      double R1 = rhoa[o + j + 1] * 1e-4 * dx[o + j + 1] / (ORC_PI_NET * r1 * r1);
      double g = 2.0 / (R0 + R1);
      gL[o + j] = g / (1e-6 * cm[o + j] * area[o + j]);
      gR[o + j + 1] = g / (1e-6 * cm[o + j + 1] * area[o + j + 1]);
    }
  }
  int64_t cur = 0;
  int64_t n_bg = n_src - n_cells;
  for (int64_t s = 0; s < n_steps; ++s) {
    /* NetCon advance: deliver, clear, detect, enqueue */
    double* row = ring + cur * n_cells;
    for (int64_t c = 0; c < n_cells; ++c) {
      double wf = row[c] * factor;
      A[c] += wf;
      B[c] += wf;
      row[c] = 0.0;
    }
    for (int64_t c = 0; c < n_cells; ++c) {
      uint8_t above = V[soma[c]] >= theta;
      sp[c] = above && !latch[c];
      latch[c] = above;
    }
    for (int64_t b = 0; b < n_bg; ++b)
      sp[n_cells + b] = orc_uniform(seed, bg_stream, (uint64_t)s * (uint64_t)n_bg + b) < bg_p;
    if (spikes) memcpy(spikes + s * n_src, sp, (size_t)n_src);
    for (int64_t e = 0; e < n_conn; ++e) {
      if (!sp[pre[e]]) continue;
      int64_t slot = (cur + delay[e]) % depth;
      ring[slot * n_cells + post[e]] += w[e];
    }
    cur = (cur + 1) % depth;
    /* populations */
    for (int64_t c = 0; c < n_cells; ++c) {
      A[c] *= d1;
      B[c] *= d2;
      int64_t o = off[c], K = off[c + 1] - o;
      double *lo = wk, *mn = wk + kmax, *up = wk + 2 * kmax, *rh = wk + 3 * kmax,
             *dv = wk + 4 * kmax, *vv = wk + 5 * kmax;
      for (int64_t j = 0; j < K; ++j) vv[j] = V[o + j];
      for (int64_t j = 0; j < K; ++j) {
        double G, J;
        orc_hh_advance(vv[j], dt, hhp, sna ? sna[o + j] : 1.0, sk ? sk[o + j] : 1.0,
                       &M[o + j], &H[o + j], &N[o + j], &G, &J);
        if (o + j == soma[c]) {
          double gs = (B[c] - A[c]) * 1e-6 / area[o + j];
          G += gs;
          J += gs * esyn;
        }
        double f = (J - G * vv[j]) * scale[o + j];
        if (j + 1 < K) f += gL[o + j] * (vv[j + 1] - vv[j]);
        if (j > 0) f -= gR[o + j] * (vv[j] - vv[j - 1]);
        mn[j] = 1.0 + dt_s * (gL[o + j] + gR[o + j] + G * scale[o + j]);
        rh[j] = dt_s * f;
        if (j + 1 < K) up[j] = -dt_s * gL[o + j];
        if (j > 0) lo[j - 1] = -dt_s * gR[o + j];
      }
      if (K == 1)
        dv[0] = rh[0] / mn[0];
      else
        orc_thomas(lo, mn, up, rh, dv, 1, K);
      for (int64_t j = 0; j < K; ++j) V[o + j] += dv[j];
    }
  }
  free(gL); free(gR); free(scale); free(area); free(ring); free(sp); free(wk);
  return 0;
}
