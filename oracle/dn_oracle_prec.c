/*
 * dn_oracle_prec.c -- the reference's float32 cable step, restated in C
 * binary32 arithmetic (oracle; TEST INFRASTRUCTURE ONLY, see dn_oracle.c).
 *
 * Two runs of the unbranched HH cable of orc_cable_run (same equations,
 * same step order, integrators/implicit.py:569-664, core.py:769-834):
 *   orc_cable_run_f32        every operation in binary32: the float32
 *                            branches of exprelr (degree 12 on |q| <= 1,
 *                            mechanisms/ops.py:13-45,106-119) and of the
 *                            cnexp phi1 (degree 9, _cnexp.py:46-93), glibc's
 *                            binary32 expf, assembly and Thomas in binary32;
 *   orc_cable_run_f32gates   the gating (rates, cnexp, G and J) in binary32
 *                            as above, the voltage, the assembly and the
 *                            Thomas elimination in binary64.
 * They exist to measure how far a float32 state is from the binary64 oracle
 * on a given workload; the tests compare both against orc_cable_run.
 * Geometry and workspace are formed in binary64 and rounded once, as a
 * float32 caller receives them.  Built with -ffp-contract=off: every
 * operation is rounded separately.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))
#define PREC_PI 3.14159265358979323846

/* q / expm1(q) even part, float32 branch of ops.py:13-45 (degree 12). */
static float exprel_even_f(float q) {
  float q2 = q * q;
  float v = q2 * (float)(-691.0 / 1307674368000.0) + (float)(1.0 / 47900160.0);
  v = v * q2 - (float)(1.0 / 1209600.0);
  v = v * q2 + (float)(1.0 / 30240.0);
  v = v * q2 - (float)(1.0 / 720.0);
  v = v * q2 + (float)(1.0 / 12.0);
  return v;
}

/* exprelr, float32 branch (ops.py:106-119): series on |q| <= 1. */
static float vtrap_f(float x, float y) {
  float q = x / y;
  if (fabsf(q) <= 1.0f) return (y - x / 2.0f) + x * (q * exprel_even_f(q));
  return x / (expf(q) - 1.0f);
}

/* phi1 = expm1(z)/z on |z| <= 0.5, float32 Horner of degree 9 (_cnexp.py:83-93). */
static float phi1_f(float z) {
  float v = z * (float)(1.0 / 3628800.0) + (float)(1.0 / 362880.0);
  v = v * z + (float)(1.0 / 40320.0);
  v = v * z + (float)(1.0 / 5040.0);
  v = v * z + (float)(1.0 / 720.0);
  v = v * z + (float)(1.0 / 120.0);
  v = v * z + (float)(1.0 / 24.0);
  v = v * z + (float)(1.0 / 6.0);
  v = v * z + 0.5f;
  v = v * z + 1.0f;
  return v;
}

/* _linear_exact_step (_cnexp.py:96-132) for x' = xa - xb x. */
static float cnexp_f(float x, float xa, float xb, float dt) {
  float z = -xb * dt;
  if (fabsf(z) <= 0.5f) return x + dt * (xa - xb * x) * phi1_f(z);
  float eq = xa / xb;
  return eq + (x - eq) * expf(z);
}

/* Gate advance and affine terms (HH.py:24-56,109-116) in binary32. */
static void hh_advance_f(float v, float dt, const float* p, float sna, float sk,
                         float* m, float* h, float* n, float* G, float* J) {
  float q10 = p[6];
  float am = 0.1f * vtrap_f(-(v + 40.0f), 10.0f);
  float bm = 4.0f * expf(-(v + 65.0f) / 18.0f);
  float ah = 0.07f * expf(-(v + 65.0f) / 20.0f);
  float bh = 1.0f / (expf(-(v + 35.0f) / 10.0f) + 1.0f);
  float an = 0.01f * vtrap_f(-(v + 55.0f), 10.0f);
  float bn = 0.125f * expf(-(v + 65.0f) / 80.0f);
  *m = cnexp_f(*m, q10 * am, q10 * (am + bm), dt);
  *h = cnexp_f(*h, q10 * ah, q10 * (ah + bh), dt);
  *n = cnexp_f(*n, q10 * an, q10 * (an + bn), dt);
  float gna = p[0] * sna * ((*m) * (*m) * (*m)) * (*h);
  float n2 = (*n) * (*n);
  float gk = p[1] * sk * (n2 * n2);
  *G = p[2] + gna + gk;
  *J = p[2] * p[5] + gna * p[3] + gk * p[4];
}

#define PREC_NAME orc_cable_run_f32
#define VT float
#include "dn_oracle_prec.inc"
#undef PREC_NAME
#undef VT

#define PREC_NAME orc_cable_run_f32gates
#define VT double
#include "dn_oracle_prec.inc"
#undef PREC_NAME
#undef VT

/*
 * orc_tree_run_f32: the reference's float32 tree step (integrators/
 * tree.py:1202-1357 with the float32 mechanism branches above and the DHS
 * elimination of triton/dhs_kernel.py:14-112) in binary32, the tree
 * counterpart of orc_cable_run_f32: same arguments and order as orc_tree_run
 * (dn_oracle.c) minus the intracellular current.  Geometry in binary64,
 * rounded once.
 */
ORC_API int orc_tree_run_f32(int64_t n_trees, const int64_t* off, const int64_t* parent,
                             const double* diam, const double* len, const double* ra,
                             const double* cm, const double* sna, const double* sk,
                             const double* hhp, const double* lead, const double* ve_amp,
                             const double* ve_wave, double dt, int64_t n_steps, double* V,
                             double* M, double* H, double* N, double* rec,
                             int64_t rec_every, int n_threads) {
  float pf[7];
  for (int i = 0; i < 7; ++i) pf[i] = (float)hhp[i];
  int64_t total = off[n_trees];
  const double dt_s = dt * 1.0e-3;
  const float dtf = (float)dt;
  int bad = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
  for (int64_t t = 0; t < n_trees; ++t) {
    int64_t o = off[t], K = off[t + 1] - off[t];
    const int64_t* par = parent + o;
    float* w = (float*)malloc(sizeof(float) * (size_t)K * 13);
    int64_t* iw = (int64_t*)malloc(sizeof(int64_t) * (size_t)(2 * K + 2));
    if (!w || !iw) { bad |= 1; free(w); free(iw); continue; }
    float *area = w, *cmdt = w + K, *ag = w + 2 * K, *f = w + 3 * K, *mainv = w + 4 * K,
          *ve = w + 5 * K, *dv = w + 6 * K, *D = w + 7 * K, *Bv = w + 8 * K, *v = w + 9 * K,
          *m = w + 10 * K, *h = w + 11 * K, *n = w + 12 * K;
    int64_t *depth = iw, *order = iw + K;
    int64_t maxd = 0;
    for (int64_t i = 0; i < K; ++i) {
      double a = PREC_PI * diam[o + i] * len[o + i] * 1.0e-8;
      area[i] = (float)a;
      cmdt[i] = (float)(1.0e-6 * cm[o + i] * a / dt_s);
      v[i] = (float)V[o + i];
      m[i] = (float)M[o + i];
      h[i] = (float)H[o + i];
      n[i] = (float)N[o + i];
      if (par[i] < 0) { ag[i] = 0.0f; depth[i] = 0; continue; }
      if (par[i] >= i) bad |= 2;
      int64_t q = par[i];
      double ri = 0.5 * diam[o + i] * 1.0e-4, rp = 0.5 * diam[o + q] * 1.0e-4;
      double Ri = ra[o + i] * (len[o + i] * 1.0e-4 / 2.0) / (PREC_PI * ri * ri);
      double Rp = ra[o + q] * (len[o + q] * 1.0e-4 / 2.0) / (PREC_PI * rp * rp);
      ag[i] = (float)(1.0 / (Ri + Rp));
      depth[i] = depth[q] + 1;
      if (depth[i] > maxd) maxd = depth[i];
    }
    if (bad) { free(w); free(iw); continue; }
    int64_t pos = 0;
    for (int64_t dd = maxd; dd >= 0; --dd)
      for (int64_t i = 0; i < K; ++i)
        if (depth[i] == dd) order[pos++] = i;
    for (int64_t s = 0; s < n_steps; ++s) {
      for (int64_t i = 0; i < K; ++i) {
        float G, J;
        hh_advance_f(v[i], dtf, pf, sna ? (float)sna[o + i] : 1.0f,
                     sk ? (float)sk[o + i] : 1.0f, &m[i], &h[i], &n[i], &G, &J);
        f[i] = (J - G * v[i]) * area[i];
        mainv[i] = cmdt[i] + G * area[i];
        ve[i] = (lead && ve_wave) ? (float)lead[o + i] * (float)ve_amp[t] * (float)ve_wave[s]
                                  : 0.0f;
      }
      for (int64_t i = 0; i < K; ++i) {
        int64_t q = par[i];
        if (q < 0) continue;
        float flow = ag[i] * (v[i] - v[q]);
        f[i] -= flow;
        f[q] += flow;
      }
      if (lead && ve_wave) {
        for (int64_t i = 0; i < K; ++i) {
          int64_t q = par[i];
          if (q < 0) continue;
          float flow = ag[i] * (ve[i] - ve[q]);
          f[i] -= flow;
          f[q] += flow;
        }
      }
      memcpy(D, mainv, sizeof(float) * (size_t)K);
      memcpy(Bv, f, sizeof(float) * (size_t)K);
      for (int64_t k = 0; k < K; ++k) {
        int64_t i = order[k];
        float old_d = D[i];
        D[i] = old_d + ag[i];
        int64_t q = par[i];
        if (q >= 0) {
          float fac = ag[i] / D[i];
          D[q] += fac * old_d;
          Bv[q] += fac * Bv[i];
        }
      }
      for (int64_t k = K - 1; k >= 0; --k) {
        int64_t i = order[k];
        int64_t q = par[i];
        dv[i] = (Bv[i] + ag[i] * (q >= 0 ? dv[q] : 0.0f)) / D[i];
      }
      for (int64_t i = 0; i < K; ++i) v[i] += dv[i];
      if (rec && rec_every > 0 && (s + 1) % rec_every == 0) {
        int64_t r = (s + 1) / rec_every - 1;
        for (int64_t i = 0; i < K; ++i) rec[r * total + o + i] = (double)v[i];
      }
    }
    for (int64_t i = 0; i < K; ++i) {
      V[o + i] = v[i];
      M[o + i] = m[i];
      H[o + i] = h[i];
      N[o + i] = n[i];
    }
    free(w);
    free(iw);
  }
  return bad ? -1 : 0;
}
