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
