// dn_cable.hip -- fused HH + backward-Euler steps for unbranched cables.
//
// Restates, for a ragged batch of cables held structure-of-arrays in one
// flat compartment axis, the per-step transition of the reference's
// _bwd_euler_ub (dendra/models/integrators/implicit.py:569-664) in the step
// order of Integrator._step (dendra/models/integrators/core.py:769-834):
//   gates by cnexp at v_old, G and J from the new gates,
//   f    = (J - G v) scale + gL (v_{j+1} - v_j) - gRp (v_j - v_{j-1}) + drive w(t)
//   main = 1 + dt_s (gL + gRp + G scale)            (= 1 - dt_s (diag_base - G scale))
//   solve(lower = -dt_s gRp, main, upper = -dt_s gL, rhs = dt_s f),  v += dv
// gL_j = g_j / C_j (edge to the right), gRp_j = g_{j-1} / C_j (edge to the
// left), both zero at the cable ends (implicit.py:484-522).  The
// extracellular (axial operator applied to Ve, implicit.py:618-638) and
// intracellular (intra / C, implicit.py:640-641) drives are linear in a
// separable waveform, so the host folds them into one activating-function
// coefficient per compartment, `drive`, scaled each step by w(t).
//
// Two execution paths:
//  * persistent (hh_cable_run_kernel): each 64-lane wavefront owns 64/P
//    cables for the whole run; every lane holds nr <= RMAX consecutive
//    compartments of its cable in registers (state and coefficients), so a
//    step touches no memory except the waveform scalar.  The solve is the
//    partition hybrid: local Thomas with a spike column per lane, one cyclic
//    reduction step, log2(P) PCR stages over ds_bpermute, back-substitution.
//    P (lanes per cable) is the Thomas/PCR crossover: P = 1 would be pure
//    Thomas, nr = 1 pure PCR.
//  * plain per-step (hh_cable_assemble_kernel + hh_cable_solve_kernel): the
//    reference's structure (gating+assembly, then solve), state in HBM.
#include "dn_wave.hpp"

namespace dn {

constexpr int kCableRMax = 13;

template <typename T> struct CableRunArgs {
  int n_tasks;
  const int32_t* task_cell0;  // first cell of the task (plan order)
  const int32_t* task_ncell;  // cells in the task, <= 64 / P
  const int32_t* task_nr;     // rows per lane
  const int32_t* task_p;      // lanes per cable (4, 8, 16, 32, 64)
  const int64_t* off;         // n_cells + 1 compartment offsets
  T *v, *m, *h, *n;           // state (in/out)
  const T *gL, *gRp, *scale, *drive, *sna, *sk;
  const T* wave;              // waveform tape (nullable)
  int64_t step0, n_steps;
  T dt;
  HHParams<T> hh;
  T* rec;                     // optional V record, [n_rec][total]
  int64_t rec_every, total;
  int32_t* status;            // launch-contract status word (nullable)
};

// Per-row read-only coefficients staged once in LDS, row-major over lanes so
// that each row is one conflict-free 16-byte (32-byte for float64) read.
// They are stored multiplied by dt * 1e-3 (ms -> s with the membrane terms
// per unit area), so the row's matrix entries need no further scaling.
template <typename T> struct alignas(4 * sizeof(T)) RowCoef { T gl, grp, sc, dr; };
template <typename T> struct alignas(2 * sizeof(T)) RowScale { T sna, sk; };

template <typename T, int P, int RMAX, bool SCALES>
__device__ __forceinline__ void cable_run_wave(const CableRunArgs<T>& A, int task,
                                               RowCoef<T>* __restrict__ cf,
                                               RowScale<T>* __restrict__ cs) {
  const int lane = __lane_id();
  const int sub = lane / P, pl = lane % P;
  const int nr = A.task_nr[task];
  const bool has_cell = sub < A.task_ncell[task];
  const int cell = A.task_cell0[task] + sub;
  int64_t o = 0;
  int K = 0;
  if (has_cell) {
    o = A.off[cell];
    K = (int)(A.off[cell + 1] - o);
  }
  const int j0 = pl * nr;
  const T dt = A.dt;
  const T dts = dt * T(1e-3);
  T v[RMAX], m[RMAX], h[RMAX], n[RMAX];
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    const int j = j0 + r;
    const bool in = r < nr && j < K;
    const int64_t idx = o + j;
    v[r] = in ? A.v[idx] : T(0);
    m[r] = in ? A.m[idx] : T(0);
    h[r] = in ? A.h[idx] : T(0);
    n[r] = in ? A.n[idx] : T(0);
    if (r < nr) {
      RowCoef<T> c;
      c.gl = in ? dts * A.gL[idx] : T(0);
      c.grp = in ? dts * A.gRp[idx] : T(0);
      c.sc = in ? dts * A.scale[idx] : T(0);
      c.dr = (in && A.drive) ? dts * A.drive[idx] : T(0);
      cf[r * 64 + lane] = c;
      if (SCALES) {
        RowScale<T> q;
        q.sna = (in && A.sna) ? A.sna[idx] : T(1);
        q.sk = (in && A.sk) ? A.sk[idx] : T(1);
        cs[r * 64 + lane] = q;
      }
    }
  }
  for (int64_t s = 0; s < A.n_steps; ++s) {
    const T w = A.wave ? A.wave[A.step0 + s] : T(0);
    T vlast = v[0];
#pragma unroll
    for (int r = 1; r < RMAX; ++r)
      if (r < nr) vlast = v[r];
    const T vprev = shfl_up_seg<P>(vlast, 1);
    const T vnext = shfl_down_seg<P>(v[0], 1);
    T Al[RMAX], Cg[RMAX], Dd[RMAX];
    T a0 = T(0), b0 = T(1), c0 = T(0), d0 = T(0);
    T al = T(0), ga = T(0), de = T(0);
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      if (r < nr) {
        const RowCoef<T> c = cf[r * 64 + lane];
        T sna = T(1), sk = T(1);
        if (SCALES) {
          const RowScale<T> q = cs[r * 64 + lane];
          sna = q.sna;
          sk = q.sk;
        }
        T G, J;
        hh_advance<T>(v[r], dt, A.hh, sna, sk, m[r], h[r], n[r], G, J);
        const T vl = (r == 0) ? vprev : v[r > 0 ? r - 1 : 0];
        T vr = vnext;
        if (r + 1 < RMAX) vr = (r + 1 < nr) ? v[r + 1 < RMAX ? r + 1 : r] : vnext;
        const T dd = (J - G * v[r]) * c.sc + c.gl * (vr - v[r]) -
                     c.grp * (v[r] - vl) + c.dr * w;
        const T bb = T(1) + c.gl + c.grp + G * c.sc;
        const T aa = -c.grp;
        const T cc = -c.gl;
        if (r == 0) {
          a0 = aa; b0 = bb; c0 = cc; d0 = dd;
        } else if (r == 1) {
          const T iv = frcp(bb);
          al = aa * iv; ga = cc * iv; de = dd * iv;
        } else {
          const T iv = frcp(bb - aa * ga);
          al = -aa * al * iv;
          ga = cc * iv;
          de = (dd - aa * de) * iv;
        }
        Al[r] = al; Cg[r] = ga; Dd[r] = de;
      }
    }
    if (nr == 1) {
      v[0] += pcr_lanes<T, P>(a0, b0, c0, d0);
    } else {
      const T aL = al, cL = ga, dL = de;
      T alN = T(0), gaN = T(0), deN = T(0);
#pragma unroll
      for (int k = RMAX - 2; k >= 1; --k) {
        if (k <= nr - 2) {
          if (k == nr - 2) {
            alN = Al[k]; gaN = Cg[k]; deN = Dd[k];
          } else {
            const T g = Cg[k];
            alN = Al[k] - g * alN;
            deN = Dd[k] - g * deN;
            gaN = -g * gaN;
            Al[k] = alN; Cg[k] = gaN; Dd[k] = deN;
          }
        }
      }
      T bF = b0, cF = c0, dF = d0;
      const T aF = a0;
      if (nr >= 3) {
        bF = b0 - c0 * Al[1];
        cF = -c0 * Cg[1];
        dF = d0 - c0 * Dd[1];
      }
      const T aF1 = shfl_down_seg<P>(aF, 1), bF1 = shfl_down_seg<P>(bF, 1);
      const T cF1 = shfl_down_seg<P>(cF, 1), dF1 = shfl_down_seg<P>(dF, 1);
      const T alpha = aL * frcp(bF);
      const T beta = cL * frcp(bF1);
      const T xL = pcr_lanes<T, P>(-alpha * aF, T(1) - alpha * cF - beta * aF1,
                                   -beta * cF1, dL - alpha * dF - beta * dF1);
      const T xLp = shfl_up_seg<P>(xL, 1);
      const T xF = (dF - aF * xLp - cF * xL) * frcp(bF);
      v[0] += xF;
#pragma unroll
      for (int k = 1; k < RMAX; ++k) {
        if (k < nr - 1)
          v[k] += Dd[k] - Al[k] * xF - Cg[k] * xL;
        else if (k == nr - 1)
          v[k] += xL;
      }
    }
    if (A.rec && ((A.step0 + s + 1) % A.rec_every) == 0) {
      const int64_t row = (A.step0 + s + 1) / A.rec_every - 1;
#pragma unroll
      for (int r = 0; r < RMAX; ++r)
        if (r < nr && j0 + r < K) A.rec[row * A.total + o + j0 + r] = v[r];
    }
  }
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    const int j = j0 + r;
    if (r < nr && j < K) {
      const int64_t idx = o + j;
      A.v[idx] = v[r];
      A.m[idx] = m[r];
      A.h[idx] = h[r];
      A.n[idx] = n[r];
    }
  }
}

// RMAX (rows per lane held in registers) is a compile-time bound of the
// launch: a launch whose tasks all have few rows uses a smaller instance,
// with fewer VGPRs and less LDS per wavefront, so more wavefronts share a CU.
template <typename T, bool SCALES, int RMAX>
__global__ __launch_bounds__(256) void hh_cable_run_kernel(CableRunArgs<T> A) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int wv = threadIdx.x >> 6;
  const int task = blockIdx.x * (blockDim.x >> 6) + wv;
  if (task >= A.n_tasks) return;
  if (A.task_nr[task] > RMAX) {  // launch contract violated: report, leave the cells
    flag_status(A.status, kStatusCableRows);
    return;
  }
  RowCoef<T>* cf = reinterpret_cast<RowCoef<T>*>(dn_smem_raw) + wv * RMAX * 64;
  RowScale<T>* cs = reinterpret_cast<RowScale<T>*>(
      reinterpret_cast<RowCoef<T>*>(dn_smem_raw) + (blockDim.x >> 6) * RMAX * 64) +
      wv * RMAX * 64;
  switch (A.task_p[task]) {
    case 4: cable_run_wave<T, 4, RMAX, SCALES>(A, task, cf, cs); break;
    case 8: cable_run_wave<T, 8, RMAX, SCALES>(A, task, cf, cs); break;
    case 16: cable_run_wave<T, 16, RMAX, SCALES>(A, task, cf, cs); break;
    case 32: cable_run_wave<T, 32, RMAX, SCALES>(A, task, cf, cs); break;
    default: cable_run_wave<T, 64, RMAX, SCALES>(A, task, cf, cs); break;
  }
}

// ---------------------------------------------------------------------------
// plain per-step path
// ---------------------------------------------------------------------------

// Gating + assembly, one thread per compartment.  Writes main and rhs.
template <typename T>
__global__ __launch_bounds__(256) void hh_cable_assemble_kernel(
    int64_t total, T* __restrict__ v_, T* __restrict__ m_, T* __restrict__ h_,
    T* __restrict__ n_, const T* __restrict__ gL, const T* __restrict__ gRp,
    const T* __restrict__ scale, const T* __restrict__ drive,
    const T* __restrict__ sna, const T* __restrict__ sk, T w, T dt, HHParams<T> hh,
    const T* __restrict__ ve, const T* __restrict__ intra, const T* __restrict__ icap,
    const T* __restrict__ gx, const T* __restrict__ gex, T* __restrict__ mainv,
    T* __restrict__ rhs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const T v = v_[i];
  T m = m_[i], h = h_[i], n = n_[i], G, J;
  hh_advance<T>(v, dt, hh, sna ? sna[i] : T(1), sk ? sk[i] : T(1), m, h, n, G, J);
  m_[i] = m; h_[i] = h; n_[i] = n;
  if (gx) {  // further affine membrane terms (point processes): g, g e per unit area
    G += gx[i];
    J += gex[i];
  }
  const T g_l = gL[i], g_r = gRp[i], sc = scale[i];
  // Neighbours across a cable boundary carry a zero conductance.
  const T vr = i + 1 < total ? v_[i + 1] : v;
  const T vl = i > 0 ? v_[i - 1] : v;
  T f = (J - G * v) * sc + g_l * (vr - v) - g_r * (v - vl) +
        (drive ? drive[i] * w : T(0));
  if (ve) {  // extracellular potential of this step: axial drive of its differences
    const T e = ve[i];
    const T er = i + 1 < total ? ve[i + 1] : e;
    const T el = i > 0 ? ve[i - 1] : e;
    f += g_l * (er - e) - g_r * (e - el);
  }
  if (intra) f += intra[i] * icap[i];  // injected current (mA) over C
  const T dts = dt * T(1e-3);
  mainv[i] = T(1) + dts * (g_l + g_r + G * sc);
  rhs[i] = dts * f;
}

// One wavefront per cable: stage the bands in LDS, hybrid solve, v += dv.
template <typename T>
__global__ __launch_bounds__(64) void hh_cable_solve_kernel(
    const int64_t* __restrict__ off, const T* __restrict__ gL,
    const T* __restrict__ gRp, const T* __restrict__ mainv, const T* __restrict__ rhs,
    T* __restrict__ v, T dt, int max_nr) {
  extern __shared__ unsigned char dn_smem_raw[];
  const int64_t cell = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t o = off[cell];
  const int64_t K = off[cell + 1] - o;
  const int nr = (int)((K + 63) / 64);
  if (nr > max_nr) return;  // beyond the launch's LDS image: left unsolved
  const int64_t KP = 64 * (int64_t)nr;
  const int64_t KPmax = 64 * (int64_t)max_nr;
  T* sa = reinterpret_cast<T*>(dn_smem_raw);
  T *sb = sa + KPmax, *sc = sa + 2 * KPmax, *sd = sa + 3 * KPmax;
  const T dts = dt * T(1e-3);
  for (int64_t j = lane; j < KP; j += 64) {
    const bool in = j < K;
    sa[j] = in ? -dts * gRp[o + j] : T(0);
    sb[j] = in ? mainv[o + j] : T(1);
    sc[j] = in ? -dts * gL[o + j] : T(0);
    sd[j] = in ? rhs[o + j] : T(0);
  }
  __syncthreads();
  wave_hybrid_solve<T>(sa, sb, sc, sd, nr);
  __syncthreads();
  for (int64_t j = lane; j < K; j += 64) v[o + j] += sd[j];
}

template <typename T, int RMAX>
static void cable_run_dispatch(const CableRunArgs<T>& A, int blocks, int waves_per_block,
                               bool scales, hipStream_t stream) {
  const size_t lds = (size_t)waves_per_block * RMAX * 64 *
                     (sizeof(RowCoef<T>) + (scales ? sizeof(RowScale<T>) : 0));
  if (scales)
    hipLaunchKernelGGL((hh_cable_run_kernel<T, true, RMAX>), dim3(blocks),
                       dim3(64 * waves_per_block), lds, stream, A);
  else
    hipLaunchKernelGGL((hh_cable_run_kernel<T, false, RMAX>), dim3(blocks),
                       dim3(64 * waves_per_block), lds, stream, A);
}

template <typename T>
static int cable_run_launch(const CableRunArgs<T>& A, int waves_per_block, int max_rows,
                            hipStream_t stream) {
  if (A.n_tasks < 0 || !A.task_cell0 || !A.task_ncell || !A.task_nr || !A.task_p ||
      !A.off || !A.v || !A.m || !A.h || !A.n || !A.gL || !A.gRp || !A.scale)
    return DN_EINVAL;
  if (A.rec && A.rec_every < 1) return DN_EINVAL;
  if (waves_per_block < 1 || waves_per_block > 4) return DN_EINVAL;
  if (max_rows < 1 || max_rows > kCableRMax) return DN_ERANGE;
  if (A.n_tasks == 0 || A.n_steps == 0) return DN_OK;
  int blocks = (A.n_tasks + waves_per_block - 1) / waves_per_block;
  const bool scales = A.sna || A.sk;
  if (max_rows <= 8)
    cable_run_dispatch<T, 8>(A, blocks, waves_per_block, scales, stream);
  else if (max_rows <= 10)
    cable_run_dispatch<T, 10>(A, blocks, waves_per_block, scales, stream);
  else
    cable_run_dispatch<T, kCableRMax>(A, blocks, waves_per_block, scales, stream);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

template <typename T>
static int cable_step_launch(int64_t n_cells, const int64_t* off, int64_t total,
                             int max_nr, T* v, T* m, T* h, T* n, const T* gL,
                             const T* gRp, const T* scale, const T* drive, const T* sna,
                             const T* sk, T w, T dt, HHParams<T> hh, const T* ve,
                             const T* intra, const T* icap, const T* gx, const T* gex,
                             T* mainv, T* rhs, hipStream_t stream) {
  if (!off || !v || !m || !h || !n || !gL || !gRp || !scale || !mainv || !rhs)
    return DN_EINVAL;
  if ((intra && !icap) || (!gx != !gex)) return DN_EINVAL;
  if (n_cells < 1 || total < 1 || max_nr < 1 || n_cells > 0x7fffffff)
    return DN_EINVAL;
  size_t lds = (size_t)4 * 64 * max_nr * sizeof(T);
  if (lds > 64 * 1024) return DN_ERANGE;
  int blocks = (int)((total + 255) / 256);
  hipLaunchKernelGGL(hh_cable_assemble_kernel<T>, dim3(blocks), dim3(256), 0, stream,
                     total, v, m, h, n, gL, gRp, scale, drive, sna, sk, w, dt, hh, ve, intra,
                     icap, gx, gex, mainv, rhs);
  hipLaunchKernelGGL(hh_cable_solve_kernel<T>, dim3((unsigned)n_cells), dim3(64), lds,
                     stream, off, gL, gRp, mainv, rhs, v, dt, max_nr);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

}  // namespace dn

#define DN_CABLE_RUN_ARGS(T)                                                        \
  int n_tasks, const int32_t *task_cell0, const int32_t *task_ncell,                \
      const int32_t *task_nr, const int32_t *task_p, const int64_t *off, T *v, T *m, \
      T *h, T *n, const T *gL, const T *gRp, const T *scale, const T *drive,         \
      const T *sna, const T *sk, const T *wave, int64_t step0, int64_t n_steps,      \
      T dt, const T *hh, T *rec, int64_t rec_every, int64_t total,                   \
      int waves_per_block, int max_rows, int32_t *status, hipStream_t stream

#define DN_CABLE_RUN_BODY(T)                                                         \
  dn::CableRunArgs<T> A;                                                            \
  A.n_tasks = n_tasks; A.task_cell0 = task_cell0; A.task_ncell = task_ncell;        \
  A.task_nr = task_nr; A.task_p = task_p; A.off = off;                              \
  A.v = v; A.m = m; A.h = h; A.n = n;                                               \
  A.gL = gL; A.gRp = gRp; A.scale = scale; A.drive = drive; A.sna = sna; A.sk = sk; \
  A.wave = wave; A.step0 = step0; A.n_steps = n_steps; A.dt = dt;                   \
  if (!hh) return DN_EINVAL;                                                        \
  A.hh = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};                         \
  A.rec = rec; A.rec_every = rec_every; A.total = total; A.status = status;         \
  return dn::cable_run_launch<T>(A, waves_per_block, max_rows, stream);

// Persistent multi-step run of a planned ragged cable batch.  hh holds
// {gnabar, gkbar, gl, ena, ek, el, q10}.  Every task row of the plan must
// satisfy task_nr <= 13 and task_p in {4, 8, 16, 32, 64}.
DN_EXPORT int dn_hh_cable_run_f32(DN_CABLE_RUN_ARGS(float)) { DN_CABLE_RUN_BODY(float) }
DN_EXPORT int dn_hh_cable_run_f64(DN_CABLE_RUN_ARGS(double)) { DN_CABLE_RUN_BODY(double) }

DN_EXPORT int dn_hh_cable_rmax(void) { return dn::kCableRMax; }

// One plain time step (assemble + solve launches) of the same batch.
DN_EXPORT int dn_hh_cable_step_f32(int64_t n_cells, const int64_t* off, int64_t total,
                                   int max_nr, float* v, float* m, float* h, float* n,
                                   const float* gL, const float* gRp, const float* scale,
                                   const float* drive, const float* sna, const float* sk,
                                   float w, float dt, const float* hh, const float* ve,
                                   const float* intra, const float* icap, const float* gx,
                                   const float* gex, float* mainv,
                                   float* rhs, hipStream_t stream) {
  if (!hh) return DN_EINVAL;
  dn::HHParams<float> p = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};
  return dn::cable_step_launch<float>(n_cells, off, total, max_nr, v, m, h, n, gL, gRp,
                                      scale, drive, sna, sk, w, dt, p, ve, intra, icap,
                                      gx, gex, mainv, rhs, stream);
}

DN_EXPORT int dn_hh_cable_step_f64(int64_t n_cells, const int64_t* off, int64_t total,
                                   int max_nr, double* v, double* m, double* h,
                                   double* n, const double* gL, const double* gRp,
                                   const double* scale, const double* drive,
                                   const double* sna, const double* sk, double w,
                                   double dt, const double* hh, const double* ve,
                                   const double* intra, const double* icap, const double* gx,
                                   const double* gex, double* mainv,
                                   double* rhs, hipStream_t stream) {
  if (!hh) return DN_EINVAL;
  dn::HHParams<double> p = {hh[0], hh[1], hh[2], hh[3], hh[4], hh[5], hh[6]};
  return dn::cable_step_launch<double>(n_cells, off, total, max_nr, v, m, h, n, gL,
                                       gRp, scale, drive, sna, sk, w, dt, p, ve, intra,
                                       icap, gx, gex, mainv, rhs, stream);
}
