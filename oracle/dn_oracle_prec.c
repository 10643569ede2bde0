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
