/*
 * dn_oracle.c -- CPU restatement of Dendra's per-timestep cable hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path (dendra_amd) never does.
 *
 * Everything here is binary64 unless a function name ends in _f32, in which
 * case it reproduces the reference's float32 arithmetic order exactly (used
 * for the bit-exact event-delivery checks).  Each routine cites the reference
 * source it restates (paths relative to the reference repository root).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  FMA
 * contraction is disabled so the float32 delivery arithmetic rounds after
 * every operation, as PyTorch's CPU scatter_add does.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Tridiagonal solves                                                         */
/* ------------------------------------------------------------------------- */

/* Functional Thomas sweep of dendra/models/integrators/tridiag/pcr.py:125-183
 * (pcr_solve_t, the reference CPU cable solver).  a,c: (B,K-1); b,d,x: (B,K). */
ORC_API void orc_thomas(const double* a, const double* b, const double* c,
                        const double* d, double* x, int64_t B, int64_t K) {
  double* cp = (double*)malloc(sizeof(double) * (size_t)(K > 1 ? K : 1));
  double* dp = (double*)malloc(sizeof(double) * (size_t)K);
  for (int64_t r = 0; r < B; ++r) {
    const double* ar = a + r * (K - 1);
    const double* br = b + r * K;
    const double* cr = c + r * (K - 1);
    const double* dr = d + r * K;
    double* xr = x + r * K;
    if (K == 1) { xr[0] = dr[0] / br[0]; continue; }
    double denom0 = br[0];
    cp[0] = cr[0] / denom0;
    dp[0] = dr[0] / denom0;
    for (int64_t i = 1; i < K; ++i) {
      double denom = br[i] - ar[i - 1] * cp[i - 1];
      if (i < K - 1) cp[i] = cr[i] / denom;
      dp[i] = (dr[i] - ar[i - 1] * dp[i - 1]) / denom;
    }
    xr[K - 1] = dp[K - 1];
    for (int64_t i = K - 2; i >= 0; --i) xr[i] = dp[i] - cp[i] * xr[i + 1];
  }
  free(cp);
  free(dp);
}

/* Parallel cyclic reduction of tridiag/pcr.py:4-96 (pcr_solve_parallel_t):
 * ceil(log2 K) stages, zero-padded shifted bands, x = d/b at the end. */
ORC_API void orc_pcr(const double* a, const double* b, const double* c,
                     const double* d, double* x, int64_t B, int64_t K) {
  size_t n = (size_t)K;
  double* aw = (double*)malloc(sizeof(double) * n * 8);
  double *bw = aw + n, *cw = aw + 2 * n, *dw = aw + 3 * n;
  double *an = aw + 4 * n, *bn = aw + 5 * n, *cn = aw + 6 * n, *dn = aw + 7 * n;
  for (int64_t r = 0; r < B; ++r) {
    const double* ar = a + r * (K - 1);
    const double* cr = c + r * (K - 1);
    if (K == 1) { x[r] = d[r] / b[r]; continue; }
    for (int64_t i = 0; i < K; ++i) {
      aw[i] = i == 0 ? 0.0 : ar[i - 1];
      cw[i] = i == K - 1 ? 0.0 : cr[i];
      bw[i] = b[r * K + i];
      dw[i] = d[r * K + i];
    }
    for (int64_t s = 1; s < K; s <<= 1) {
      for (int64_t i = 0; i < K; ++i) {
        double alpha = i >= s ? -aw[i] / bw[i - s] : 0.0;
        double beta = i + s < K ? -cw[i] / bw[i + s] : 0.0;
        double al = i >= s ? aw[i - s] : 0.0, cl = i >= s ? cw[i - s] : 0.0,
               dl = i >= s ? dw[i - s] : 0.0;
        double ar2 = i + s < K ? aw[i + s] : 0.0, cr2 = i + s < K ? cw[i + s] : 0.0,
               dr2 = i + s < K ? dw[i + s] : 0.0;
        an[i] = alpha * al;
        bn[i] = bw[i] + alpha * cl + beta * ar2;
        cn[i] = beta * cr2;
        dn[i] = dw[i] + alpha * dl + beta * dr2;
      }
      memcpy(aw, an, sizeof(double) * n);
      memcpy(bw, bn, sizeof(double) * n);
      memcpy(cw, cn, sizeof(double) * n);
      memcpy(dw, dn, sizeof(double) * n);
    }
    for (int64_t i = 0; i < K; ++i) x[r * K + i] = dw[i] / bw[i];
  }
  free(aw);
}

/* DHS (Hines) elimination of dendra/models/integrators/triton/dhs_kernel.py:14-112.
 * The matrix is diag(d + a + sum_children a) - offdiag(a), a[root] = 0,
 * children-before-parent elimination over the caller's layers, then
 * back-substitution in reverse.  The Schur update keeps the reference's
 * cancellation-free form (g/d_i) * old_d_i.  Topology shared over the batch. */
ORC_API void orc_dhs(const double* dmem, const double* ageom, const double* rhs,
                     const int64_t* parent, const int64_t* order,
                     const int64_t* layer_ptr, double* x, int64_t B, int64_t K,
                     int64_t L) {
  double* D = (double*)malloc(sizeof(double) * (size_t)K);
  double* Bv = (double*)malloc(sizeof(double) * (size_t)K);
  for (int64_t r = 0; r < B; ++r) {
    const double* a = ageom + r * K;
    memcpy(D, dmem + r * K, sizeof(double) * (size_t)K);
    memcpy(Bv, rhs + r * K, sizeof(double) * (size_t)K);
    double* v = x + r * K;
    for (int64_t l = 0; l < L; ++l) {
      for (int64_t o = layer_ptr[l]; o < layer_ptr[l + 1]; ++o) {
        int64_t i = order[o];
        double old_d = D[i];
        double g = a[i];
        D[i] = old_d + g;
        int64_t p = parent[i];
        if (p >= 0) {
          double fac = g / D[i];
          D[p] += fac * old_d;
          Bv[p] += fac * Bv[i];
        }
      }
    }
    for (int64_t l = L - 1; l >= 0; --l) {
      for (int64_t o = layer_ptr[l]; o < layer_ptr[l + 1]; ++o) {
        int64_t i = order[o];
        int64_t p = parent[i];
        double vp = p >= 0 ? v[p] : 0.0;
        v[i] = (Bv[i] + a[i] * vp) / D[i];
      }
    }
  }
  free(D);
  free(Bv);
}

/* ------------------------------------------------------------------------- */
/* Mechanism arithmetic                                                       */
/* ------------------------------------------------------------------------- */

/* Even Bernoulli part of q/expm1(q), binary64 branch of
 * dendra/models/mechanisms/ops.py:13-45 (degree 18, |q| <= 0.5). */
static double recip_exprel_even(double q) {
  double q2 = q * q;
  double v = q2 * (43867.0 / 5109094217170944000.0) - 3617.0 / 10670622842880000.0;
  v = v * q2 + 1.0 / 74724249600.0;
  v = v * q2 - 691.0 / 1307674368000.0;
  v = v * q2 + 1.0 / 47900160.0;
  v = v * q2 - 1.0 / 1209600.0;
  v = v * q2 + 1.0 / 30240.0;
  v = v * q2 - 1.0 / 720.0;
  v = v * q2 + 1.0 / 12.0;
  return v;
}

/* exprelr / vtrap: x / (exp(x/y) - 1), ops.py:47-128 (binary64 branch). */
ORC_API double orc_vtrap(double x, double y) {
  double q = x / y;
  if (fabs(q) <= 0.5) {
    double even = recip_exprel_even(q);
    return (y - x / 2.0) + x * (q * even);
  }
  return x / (exp(q) - 1.0);
}

/* phi1(z) = expm1(z)/z for |z| <= 0.5, binary64 Horner of
 * dendra/models/mechanisms/_cnexp.py:46-93 (Taylor degree 16). */
ORC_API double orc_phi1(double z) {
  double v = z * (1.0 / 355687428096000.0) + 1.0 / 20922789888000.0;
  v = v * z + 1.0 / 1307674368000.0;
  v = v * z + 1.0 / 87178291200.0;
  v = v * z + 1.0 / 6227020800.0;
  v = v * z + 1.0 / 479001600.0;
  v = v * z + 1.0 / 39916800.0;
  v = v * z + 1.0 / 3628800.0;
  v = v * z + 1.0 / 362880.0;
  v = v * z + 1.0 / 40320.0;
  v = v * z + 1.0 / 5040.0;
  v = v * z + 1.0 / 720.0;
  v = v * z + 1.0 / 120.0;
  v = v * z + 1.0 / 24.0;
  v = v * z + 1.0 / 6.0;
  v = v * z + 1.0 / 2.0;
  v = v * z + 1.0;
  return v;
}

/* _linear_exact_step of _cnexp.py:96-132 for x' = xa - xb*x (slope -xb,
 * source xa; ode.py:70-71 emits it for every HH gate). dt in ms. */
ORC_API double orc_cnexp(double x, double xa, double xb, double dt) {
  double coef = -xb;
  double z = coef * dt;
  if (fabs(z) <= 0.5) {
    double rate = xa - xb * x;
    return x + dt * rate * orc_phi1(z);
  }
  double eq = -xa / coef;
  return eq + (x - eq) * exp(z);
}

typedef struct {
  double gnabar, gkbar, gl, ena, ek, el, q10;
} orc_hh_t;

/* HH rate constants, dendra/models/mod/HH.py:24-40. */
static void hh_rates(double v, double q10, double* r) {
  double am = 0.1 * orc_vtrap(-(v + 40.0), 10.0);
  double bm = 4.0 * exp(-(v + 65.0) / 18.0);
  double ah = 0.07 * exp(-(v + 65.0) / 20.0);
  double bh = 1.0 / (exp(-(v + 35.0) / 10.0) + 1.0);
  double an = 0.01 * orc_vtrap(-(v + 55.0), 10.0);
  double bn = 0.125 * exp(-(v + 65.0) / 80.0);
  r[0] = q10 * am;
  r[1] = q10 * (am + bm);
  r[2] = q10 * ah;
  r[3] = q10 * (ah + bh);
  r[4] = q10 * an;
  r[5] = q10 * (an + bn);
}

/* Steady-state initialisation without q10, HH.py:42-56. */
ORC_API void orc_hh_steady(double v, double* m, double* h, double* n) {
  double am = 0.1 * orc_vtrap(-(v + 40.0), 10.0);
  double bm = 4.0 * exp(-(v + 65.0) / 18.0);
  double ah = 0.07 * exp(-(v + 65.0) / 20.0);
  double bh = 1.0 / (exp(-(v + 35.0) / 10.0) + 1.0);
  double an = 0.01 * orc_vtrap(-(v + 55.0), 10.0);
  double bn = 0.125 * exp(-(v + 65.0) / 80.0);
  *m = am / (am + bm);
  *h = ah / (ah + bh);
  *n = an / (an + bn);
}

/* Gate advance (advance_pre_current, _handler.py:2217-2229) followed by the
 * affine current terms (G, J) of HH.py:109-116 from the new gates. */
static inline void hh_advance(double v, double dt, const orc_hh_t* p,
                              double sna, double sk, double* m, double* h,
                              double* n, double* G, double* J) {
  double r[6];
  hh_rates(v, p->q10, r);
  *m = orc_cnexp(*m, r[0], r[1], dt);
  *h = orc_cnexp(*h, r[2], r[3], dt);
  *n = orc_cnexp(*n, r[4], r[5], dt);
  double gna = p->gnabar * sna * ((*m) * (*m) * (*m)) * (*h);
  double n2 = (*n) * (*n);
  double gk = p->gkbar * sk * (n2 * n2);
  *G = p->gl + gna + gk;
  *J = p->gl * p->el + gna * p->ena + gk * p->ek;
}

ORC_API void orc_hh_advance(double v, double dt, const double* hhp, double sna,
                            double sk, double* m, double* h, double* n,
                            double* G, double* J) {
  orc_hh_t p = {hhp[0], hhp[1], hhp[2], hhp[3], hhp[4], hhp[5], hhp[6]};
  hh_advance(v, dt, &p, sna, sk, m, h, n, G, J);
}

/* exp2syn weight normalisation, dendra/models/mod/EXP2SYN.py:112-145. */
ORC_API double orc_exp2syn_factor(double tau1, double tau2) {
  double ratio = tau1 / tau2;
  if (ratio < 1e-9) ratio = 1e-9;
  if (ratio > 0.9999) ratio = 0.9999;
  double logden = log1p(-ratio) + ratio / (1.0 - ratio) * log(ratio);
  return exp(-logden);
}

/* ------------------------------------------------------------------------- */
/* Geometry and workspaces                                                    */
/* ------------------------------------------------------------------------- */

#define ORC_PI 3.14159265358979323846

/* Lateral cylinder area in cm^2, integrators/cable.py:10-15. */
static inline double cyl_area(double d_um, double l_um) {
  return d_um * 1.0e-4 * ORC_PI * l_um * 1.0e-4;
}

/* Whole-compartment axial resistance (ohm), integrators/cable.py:18-29. */
static inline double cyl_resistance(double d_um, double l_um, double rhoa) {
  double r = 1.0e-4 * d_um / 2.0;
  double lc = 1.0e-4 * l_um;
  return rhoa * lc / (ORC_PI * r * r);
}

/* ------------------------------------------------------------------------- */
/* Unbranched cable run: backward Euler, increment form                      */
/* ------------------------------------------------------------------------- */

/*
 * Ragged batch of unbranched HH cables.  Compartment j of cell c lives at flat
 * index off[c] + j.  Per step (integrators/core.py:769-834 order):
 *   gates by cnexp at v_old -> G, J -> f = (J - G v) * scale + axial(v)
 *   + axial(ve) + intra * cm_inv -> rhs = dt_s f, main = 1 - dt_s (diag_base
 *   - G scale) -> Thomas -> v += dv       (implicit.py:569-664)
 * Workspace as implicit.py:484-522 with the half-cylinder edge conductance
 * of cable.py:18-29.  Extracellular potential at step i: lead * ve_amp[c] *
 * ve_wave[i] (mV; isotropic point source, fields/analytic.py:59-62).  Intra
 * current at step i: iamp[j] * i_wave[i] (mA).
 * rec (optional): V after every rec_every-th step, [n_rec][total].
 */
ORC_API int orc_cable_run(int64_t n_cells, const int64_t* off, const double* diam,
                          const double* dx, const double* rhoa, const double* cm,
                          const double* sna, const double* sk, const double* hhp,
                          const double* lead, const double* ve_amp,
                          const double* ve_wave, const double* iamp,
                          const double* i_wave, double dt, int64_t n_steps,
                          double* V, double* M, double* H, double* N, double* rec,
                          int64_t rec_every, int n_threads) {
  orc_hh_t p = {hhp[0], hhp[1], hhp[2], hhp[3], hhp[4], hhp[5], hhp[6]};
  int64_t total = off[n_cells];
  double dt_s = dt * 1.0e-3;
  int bad = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
  for (int64_t c = 0; c < n_cells; ++c) {
    int64_t o = off[c], K = off[c + 1] - off[c];
    double* w = (double*)malloc(sizeof(double) * (size_t)K * 12);
    if (!w) { bad |= 1; continue; }
    double *cinv = w, *scale = w + K, *gL = w + 2 * K, *gR = w + 3 * K,
           *diag = w + 4 * K, *lower = w + 5 * K, *upper = w + 6 * K,
           *f = w + 7 * K, *mainv = w + 8 * K, *rhs = w + 9 * K,
           *dv = w + 10 * K, *ve = w + 11 * K;
    for (int64_t j = 0; j < K; ++j) {
      double area = cyl_area(diam[o + j], dx[o + j]);
      double cap = 1.0e-6 * cm[o + j] * area;
      cinv[j] = 1.0 / cap;
      scale[j] = area * cinv[j];
      diag[j] = 0.0;
    }
    for (int64_t j = 0; j + 1 < K; ++j) {
      double R0 = cyl_resistance(diam[o + j], dx[o + j], rhoa[o + j]);
      double R1 = cyl_resistance(diam[o + j + 1], dx[o + j + 1], rhoa[o + j + 1]);
      double g = 2.0 / (R0 + R1);
      double capL = 1.0 / cinv[j], capR = 1.0 / cinv[j + 1];
      gL[j] = g / capL;
      gR[j] = g / capR;
      lower[j] = -dt_s * gR[j];
      upper[j] = -dt_s * gL[j];
    }
    for (int64_t j = 0; j < K; ++j) {
      double left = j + 1 < K ? -gL[j] : 0.0;
      double right = j > 0 ? -gR[j - 1] : 0.0;
      diag[j] = left + right;
    }
    for (int64_t s = 0; s < n_steps; ++s) {
      for (int64_t j = 0; j < K; ++j) {
        double G, J;
        double v = V[o + j];
        hh_advance(v, dt, &p, sna ? sna[o + j] : 1.0, sk ? sk[o + j] : 1.0,
                   &M[o + j], &H[o + j], &N[o + j], &G, &J);
        f[j] = (J - G * v) * scale[j];
        mainv[j] = 1.0 - dt_s * (diag[j] - G * scale[j]);
        ve[j] = (lead && ve_wave) ? lead[o + j] * ve_amp[c] * ve_wave[s] : 0.0;
      }
      if (K > 1) {
        for (int64_t j = 0; j < K; ++j) {
          double a = 0.0;
          if (j + 1 < K) a += gL[j] * (V[o + j + 1] - V[o + j]);
          if (j > 0) a -= gR[j - 1] * (V[o + j] - V[o + j - 1]);
          if (lead && ve_wave) {
            if (j + 1 < K) a += gL[j] * (ve[j + 1] - ve[j]);
            if (j > 0) a -= gR[j - 1] * (ve[j] - ve[j - 1]);
          }
          f[j] += a;
        }
      }
      if (iamp && i_wave)
        for (int64_t j = 0; j < K; ++j) f[j] += iamp[o + j] * i_wave[s] * cinv[j];
      for (int64_t j = 0; j < K; ++j) rhs[j] = dt_s * f[j];
      if (K == 1)
        dv[0] = rhs[0] / mainv[0];
      else
        orc_thomas(lower, mainv, upper, rhs, dv, 1, K);
      for (int64_t j = 0; j < K; ++j) V[o + j] += dv[j];
      if (rec && rec_every > 0 && (s + 1) % rec_every == 0) {
        int64_t r = (s + 1) / rec_every - 1;
        for (int64_t j = 0; j < K; ++j) rec[r * total + o + j] = V[o + j];
      }
    }
    free(w);
  }
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Branched trees: current-form backward Euler with DHS                       */
/* ------------------------------------------------------------------------- */

/*
 * Ragged batch of HH trees.  parent[] holds local indices (root -1) with
 * parent < child inside each tree (the topological numbering that
 * graph_to_parent_and_axial establishes, integrators/tree.py:477-612).
 * Child-indexed axial conductance a_i = 1 / (R_half_i + R_half_p) with
 * R_half = Ra (L/2) / (pi r^2) (tree.py:560-585).  Per step (tree.py:1202-1357):
 *   gates -> f = (J - G v) area - edge flows g (v_c - v_p) (child -, parent +)
 *   likewise for ve, + intra (mA) -> main = cm_dt + G area -> DHS -> v += dv.
 * Elimination runs deepest-first by BFS depth (build_dhs_layers, tree.py:615-673).
 */
ORC_API int orc_tree_run(int64_t n_trees, const int64_t* off, const int64_t* parent,
                         const double* diam, const double* len, const double* ra,
                         const double* cm, const double* sna, const double* sk,
                         const double* hhp, const double* lead, const double* ve_amp,
                         const double* ve_wave, const double* iamp,
                         const double* i_wave, double dt, int64_t n_steps, double* V,
                         double* M, double* H, double* N, double* rec,
                         int64_t rec_every, int n_threads) {
  orc_hh_t p = {hhp[0], hhp[1], hhp[2], hhp[3], hhp[4], hhp[5], hhp[6]};
  int64_t total = off[n_trees];
  double dt_s = dt * 1.0e-3;
  int bad = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
  for (int64_t t = 0; t < n_trees; ++t) {
    int64_t o = off[t], K = off[t + 1] - off[t];
    const int64_t* par = parent + o;
    double* w = (double*)malloc(sizeof(double) * (size_t)K * 9);
    int64_t* iw = (int64_t*)malloc(sizeof(int64_t) * (size_t)(3 * K + 2));
    if (!w || !iw) { bad |= 1; free(w); free(iw); continue; }
    double *area = w, *cmdt = w + K, *ag = w + 2 * K, *f = w + 3 * K,
           *mainv = w + 4 * K, *ve = w + 5 * K, *dv = w + 6 * K, *D = w + 7 * K,
           *Bv = w + 8 * K;
    int64_t *depth = iw, *order = iw + K;
    int64_t maxd = 0;
    for (int64_t i = 0; i < K; ++i) {
      area[i] = cyl_area(diam[o + i], len[o + i]);
      cmdt[i] = 1.0e-6 * cm[o + i] * area[i] / dt_s;
      if (par[i] < 0) { ag[i] = 0.0; depth[i] = 0; continue; }
      if (par[i] >= i) bad |= 2;
      int64_t q = par[i];
      double ri = 0.5 * diam[o + i] * 1.0e-4, rp = 0.5 * diam[o + q] * 1.0e-4;
      double Ri = ra[o + i] * (len[o + i] * 1.0e-4 / 2.0) / (ORC_PI * ri * ri);
      double Rp = ra[o + q] * (len[o + q] * 1.0e-4 / 2.0) / (ORC_PI * rp * rp);
      ag[i] = 1.0 / (Ri + Rp);
      depth[i] = depth[q] + 1;
      if (depth[i] > maxd) maxd = depth[i];
    }
    if (bad) { free(w); free(iw); continue; }
    /* deepest -> root bucket order */
    int64_t pos = 0;
    for (int64_t dd = maxd; dd >= 0; --dd)
      for (int64_t i = 0; i < K; ++i)
        if (depth[i] == dd) order[pos++] = i;
    for (int64_t s = 0; s < n_steps; ++s) {
      for (int64_t i = 0; i < K; ++i) {
        double G, J;
        double v = V[o + i];
        hh_advance(v, dt, &p, sna ? sna[o + i] : 1.0, sk ? sk[o + i] : 1.0,
                   &M[o + i], &H[o + i], &N[o + i], &G, &J);
        f[i] = (J - G * v) * area[i];
        mainv[i] = cmdt[i] + G * area[i];
        ve[i] = (lead && ve_wave) ? lead[o + i] * ve_amp[t] * ve_wave[s] : 0.0;
      }
      for (int64_t i = 0; i < K; ++i) {
        int64_t q = par[i];
        if (q < 0) continue;
        double flow = ag[i] * (V[o + i] - V[o + q]);
        f[i] -= flow;
        f[q] += flow;
      }
      if (lead && ve_wave) {
        for (int64_t i = 0; i < K; ++i) {
          int64_t q = par[i];
          if (q < 0) continue;
          double flow = ag[i] * (ve[i] - ve[q]);
          f[i] -= flow;
          f[q] += flow;
        }
      }
      if (iamp && i_wave)
        for (int64_t i = 0; i < K; ++i) f[i] += iamp[o + i] * i_wave[s];
      /* DHS with the reference's just-in-time diagonal and Schur form. */
      memcpy(D, mainv, sizeof(double) * (size_t)K);
      memcpy(Bv, f, sizeof(double) * (size_t)K);
      for (int64_t k = 0; k < K; ++k) {
        int64_t i = order[k];
        double old_d = D[i];
        D[i] = old_d + ag[i];
        int64_t q = par[i];
        if (q >= 0) {
          double fac = ag[i] / D[i];
          D[q] += fac * old_d;
          Bv[q] += fac * Bv[i];
        }
      }
      for (int64_t k = K - 1; k >= 0; --k) {
        int64_t i = order[k];
        int64_t q = par[i];
        dv[i] = (Bv[i] + ag[i] * (q >= 0 ? dv[q] : 0.0)) / D[i];
      }
      for (int64_t i = 0; i < K; ++i) V[o + i] += dv[i];
      if (rec && rec_every > 0 && (s + 1) % rec_every == 0) {
        int64_t r = (s + 1) / rec_every - 1;
        for (int64_t i = 0; i < K; ++i) rec[r * total + o + i] = V[o + i];
      }
    }
    free(w);
    free(iw);
  }
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Activation thresholds by bisection                                         */
/* ------------------------------------------------------------------------- */

/* One trial: the cable at rest, Ve_j(t) = lead_j * amp * wave[s] applied for
 * n_steps backward-Euler steps (as orc_cable_run).  Returns 1 as soon as a
 * checked compartment has v >= vthr after a step (the Active callback,
 * models/callbacks.py:1946-1990; _threshold_callbacks.py:50-53), else 0. */
static int cable_trial(int64_t K, const double* gL, const double* gR,
                       const double* scale, const double* lower, const double* upper,
                       const double* diag, const orc_hh_t* p, const double* lead,
                       double amp, const double* wave, int64_t n_steps, double dt,
                       double v_init, const int64_t* chk, int n_chk, double vthr,
                       double* V, double* M, double* H, double* N, double* f,
                       double* mainv, double* rhs, double* dv, double* ve) {
  double dt_s = dt * 1.0e-3, m0, h0, n0;
  orc_hh_steady(v_init, &m0, &h0, &n0);
  for (int64_t j = 0; j < K; ++j) { V[j] = v_init; M[j] = m0; H[j] = h0; N[j] = n0; }
  for (int64_t s = 0; s < n_steps; ++s) {
    for (int64_t j = 0; j < K; ++j) {
      double G, J, v = V[j];
      hh_advance(v, dt, p, 1.0, 1.0, &M[j], &H[j], &N[j], &G, &J);
      f[j] = (J - G * v) * scale[j];
      mainv[j] = 1.0 - dt_s * (diag[j] - G * scale[j]);
      ve[j] = lead[j] * amp * wave[s];
    }
    for (int64_t j = 0; j < K; ++j) {
      double a = 0.0;
      if (j + 1 < K) a += gL[j] * (V[j + 1] - V[j]) + gL[j] * (ve[j + 1] - ve[j]);
      if (j > 0) a -= gR[j - 1] * (V[j] - V[j - 1]) + gR[j - 1] * (ve[j] - ve[j - 1]);
      rhs[j] = dt_s * (f[j] + a);
    }
    if (K == 1) dv[0] = rhs[0] / mainv[0];
    else orc_thomas(lower, mainv, upper, rhs, dv, 1, K);
    int hit = 0;
    for (int64_t j = 0; j < K; ++j) V[j] += dv[j];
    for (int c = 0; c < n_chk; ++c) hit |= V[chk[c]] >= vthr;
    if (hit) return 1;
  }
  return 0;
}

/*
 * Thresholder.calculate_thresholds (instruments/thresholder.py:873-979) per
 * fiber: _fix_bounds without block (:835-871) multiplies ub by fix_up while
 * the trial at ub is inactive (at most max_fix times; status 1 if it never
 * activates), then arithmetic bisection (:981-993) while (ub - lb) / ub >=
 * rtol (:476-478), at most max_bisect trials (status 2 if not met).
 * ub/lb: initial bounds in, final bounds out.  chk: checked compartments of
 * every fiber (local indices).  trials: simulations run per fiber.
 */
ORC_API int orc_threshold(int64_t n_fibers, const int64_t* off, const double* diam,
                          const double* dx, const double* rhoa, const double* cm,
                          const double* hhp, const double* lead, const double* wave,
                          int64_t n_steps, double dt, double v_init, const int64_t* chk,
                          int n_chk, double vthr, double rtol, double fix_up, int max_fix,
                          int max_bisect, double* ub, double* lb, int32_t* status,
                          int32_t* trials, int n_threads) {
  orc_hh_t p = {hhp[0], hhp[1], hhp[2], hhp[3], hhp[4], hhp[5], hhp[6]};
  double dt_s = dt * 1.0e-3;
  int bad = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
  for (int64_t c = 0; c < n_fibers; ++c) {
    int64_t o = off[c], K = off[c + 1] - off[c];
    double* w = (double*)malloc(sizeof(double) * (size_t)K * 16);
    if (!w) { bad |= 1; continue; }
    double *scale = w, *gL = w + K, *gR = w + 2 * K, *diag = w + 3 * K,
           *lower = w + 4 * K, *upper = w + 5 * K, *V = w + 6 * K, *M = w + 7 * K,
           *H = w + 8 * K, *N = w + 9 * K, *f = w + 10 * K, *mainv = w + 11 * K,
           *rhs = w + 12 * K, *dv = w + 13 * K, *ve = w + 14 * K, *cinv = w + 15 * K;
    for (int64_t j = 0; j < K; ++j) {
      double area = cyl_area(diam[o + j], dx[o + j]);
      cinv[j] = 1.0 / (1.0e-6 * cm[o + j] * area);
      scale[j] = area * cinv[j];
      gL[j] = gR[j] = 0.0;
    }
    for (int64_t j = 0; j + 1 < K; ++j) {
      double g = 2.0 / (cyl_resistance(diam[o + j], dx[o + j], rhoa[o + j]) +
                        cyl_resistance(diam[o + j + 1], dx[o + j + 1], rhoa[o + j + 1]));
      gL[j] = g * cinv[j];
      gR[j] = g * cinv[j + 1];
      lower[j] = -dt_s * gR[j];
      upper[j] = -dt_s * gL[j];
    }
    for (int64_t j = 0; j < K; ++j)
      diag[j] = (j + 1 < K ? -gL[j] : 0.0) + (j > 0 ? -gR[j - 1] : 0.0);
    double u = ub[c], l = lb[c];
    int st = 0, ntr = 0, tries = 0, active = 0;
#define TRIAL(amp)                                                                 \
  (++ntr, cable_trial(K, gL, gR, scale, lower, upper, diag, &p, lead + o, (amp), wave, \
                      n_steps, dt, v_init, chk, n_chk, vthr, V, M, H, N, f, mainv, rhs, \
                      dv, ve))
    for (;;) {
      active = TRIAL(u);
      if (active) break;
      if (tries >= max_fix) { st = 1; break; }
      u *= fix_up;
      ++tries;
    }
    if (st == 0) {
      tries = 0;
      while ((u - l) / u >= rtol) {
        if (tries >= max_bisect) { st = 2; break; }
        double amp = (u + l) / 2.0;
        if (TRIAL(amp)) u = amp;
        else l = amp;
        ++tries;
      }
    }
#undef TRIAL
    ub[c] = u;
    lb[c] = l;
    status[c] = st;
    trials[c] = ntr;
    free(w);
  }
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Double-layer extracellular (ExtCell) cables: 3x3 block backward Euler      */
/* ------------------------------------------------------------------------- */

/* Solve the 3x3 system a x = b by Gaussian elimination with partial pivoting
 * (a and b are overwritten). */
static void solve3_pivot(double* a, double* b, double* x) {
  int p[3] = {0, 1, 2};
  for (int c = 0; c < 3; ++c) {
    int best = c;
    for (int r = c + 1; r < 3; ++r)
      if (fabs(a[p[r] * 3 + c]) > fabs(a[p[best] * 3 + c])) best = r;
    int t = p[c]; p[c] = p[best]; p[best] = t;
    for (int r = c + 1; r < 3; ++r) {
      double f = a[p[r] * 3 + c] / a[p[c] * 3 + c];
      for (int k = c; k < 3; ++k) a[p[r] * 3 + k] -= f * a[p[c] * 3 + k];
      b[p[r]] -= f * b[p[c]];
    }
  }
  for (int c = 2; c >= 0; --c) {
    double s = b[p[c]];
    for (int k = c + 1; k < 3; ++k) s -= a[p[c] * 3 + k] * x[k];
    x[c] = s / a[p[c] * 3 + c];
  }
}

/*
 * Batch of B HH cables of K compartments with two extracellular layers,
 * _bwd_euler_bt._step (dendra/models/integrators/implicit.py:1021-1136).
 * Per step, for compartment k of cell b (unknowns vi, ve0, ve1 = vc[k][0..2]):
 *   v = vi - ve0; gates by cnexp at v -> G, J (HH.py:109-116)
 *   g = G area, d = (J - G v) area + intra
 *   A_k = maind_k + [[g, -g, 0], [-g, g, 0], [0, 0, 0]]
 *   R_k = (d, -d - xg0 (ve0 - ve1), xg0 (ve0 - ve1) + xg1 (e_bath - ve1))
 *         (_block_voltage_residual, implicit.py:1139-1147)
 *       + f_k - f_{k-1},  f_k = -lower_k (vc_{k+1} - vc_k)   (axial flux)
 *   block tridiagonal (lower at (k, k-1), upper at (k-1, k), both diagonal)
 *   solve for the increment; vc += increment.
 * maind (B,K,9), lower/upper (B,K-1,3), area (B,K), xg (B,K,2) are the
 * prepared workspace of _prepare_workspace (implicit.py:800-872); bath
 * potential e = lead[b][k] amp[b] wave[s] (mV), intra = iamp[b][k] i_wave[s]
 * (mA).  vc (B,K,3), m, h, n (B,K) advance in place; rec (optional):
 * v = vi - ve0 every rec_every steps, [n_rec][B*K].
 */
ORC_API int orc_extcell_run(int64_t B, int64_t K, const double* maind, const double* lower,
                            const double* upper, const double* area, const double* xg,
                            const double* sna, const double* sk, const double* hhp,
                            const double* lead, const double* amp, const double* wave,
                            const double* iamp, const double* i_wave, double dt,
                            int64_t n_steps, double* vc, double* M, double* H, double* N,
                            double* rec, int64_t rec_every, int n_threads) {
  orc_hh_t p = {hhp[0], hhp[1], hhp[2], hhp[3], hhp[4], hhp[5], hhp[6]};
  int bad = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
  for (int64_t b = 0; b < B; ++b) {
    double* w = (double*)malloc(sizeof(double) * (size_t)K * 24);
    if (!w) { bad |= 1; continue; }
    double *Ak = w, *Rk = w + 9 * K, *Xk = w + 12 * K, *Fk = w + 15 * K;
    double* V = vc + b * K * 3;
    for (int64_t s = 0; s < n_steps; ++s) {
      for (int64_t k = 0; k < K; ++k) {
        const int64_t i = b * K + k;
        double v = V[k * 3] - V[k * 3 + 1];
        double G, J;
        hh_advance(v, dt, &p, sna ? sna[i] : 1.0, sk ? sk[i] : 1.0, &M[i], &H[i], &N[i], &G,
                   &J);
        double g = G * area[i];
        double d = (J - G * v) * area[i];
        if (iamp && i_wave) d += iamp[i] * i_wave[s];
        for (int c = 0; c < 9; ++c) Ak[k * 9 + c] = maind[i * 9 + c];
        Ak[k * 9 + 0] += g;
        Ak[k * 9 + 1] -= g;
        Ak[k * 9 + 3] -= g;
        Ak[k * 9 + 4] += g;
        double bath = (lead && amp && wave) ? lead[i] * amp[b] * wave[s] : 0.0;
        double shell = xg[i * 2] * (V[k * 3 + 1] - V[k * 3 + 2]);
        double bathf = xg[i * 2 + 1] * (bath - V[k * 3 + 2]);
        Rk[k * 3 + 0] = d;
        Rk[k * 3 + 1] = -d - shell;
        Rk[k * 3 + 2] = shell + bathf;
      }
      for (int64_t k = 0; k + 1 < K; ++k)
        for (int c = 0; c < 3; ++c)
          Fk[k * 3 + c] = -lower[(b * (K - 1) + k) * 3 + c] * (V[(k + 1) * 3 + c] - V[k * 3 + c]);
      for (int64_t k = 0; k < K; ++k)
        for (int c = 0; c < 3; ++c) {
          if (k + 1 < K) Rk[k * 3 + c] += Fk[k * 3 + c];
          if (k > 0) Rk[k * 3 + c] -= Fk[(k - 1) * 3 + c];
        }
      /* block Thomas: A'_k = A_k - diag(l) A'^-1_{k-1} diag(u), one 3x3 solve
       * per column of A'^-1_{k-1} diag(u) and one for the right-hand side */
      for (int64_t k = 1; k < K; ++k) {
        const double* l = lower + (b * (K - 1) + k - 1) * 3;
        const double* u = upper + (b * (K - 1) + k - 1) * 3;
        double Y[9], y[3];
        for (int c = 0; c < 3; ++c) {
          double a[9], rhs[3] = {0.0, 0.0, 0.0}, col[3];
          memcpy(a, Ak + (k - 1) * 9, sizeof a);
          rhs[c] = u[c];
          solve3_pivot(a, rhs, col);
          for (int r = 0; r < 3; ++r) Y[r * 3 + c] = col[r];
        }
        {
          double a[9], rhs[3];
          memcpy(a, Ak + (k - 1) * 9, sizeof a);
          memcpy(rhs, Rk + (k - 1) * 3, sizeof rhs);
          solve3_pivot(a, rhs, y);
        }
        for (int r = 0; r < 3; ++r) {
          for (int c = 0; c < 3; ++c) Ak[k * 9 + r * 3 + c] -= l[r] * Y[r * 3 + c];
          Rk[k * 3 + r] -= l[r] * y[r];
        }
      }
      for (int64_t k = K - 1; k >= 0; --k) {
        double a[9], rhs[3];
        memcpy(a, Ak + k * 9, sizeof a);
        for (int c = 0; c < 3; ++c)
          rhs[c] = Rk[k * 3 + c] -
                   (k + 1 < K ? upper[(b * (K - 1) + k) * 3 + c] * Xk[(k + 1) * 3 + c] : 0.0);
        solve3_pivot(a, rhs, Xk + k * 3);
      }
      for (int64_t k = 0; k < K * 3; ++k) V[k] += Xk[k];
      if (rec && rec_every > 0 && (s + 1) % rec_every == 0) {
        int64_t r = (s + 1) / rec_every - 1;
        for (int64_t k = 0; k < K; ++k) rec[r * B * K + b * K + k] = V[k * 3] - V[k * 3 + 1];
      }
    }
    free(w);
  }
  return bad ? -1 : 0;
}
