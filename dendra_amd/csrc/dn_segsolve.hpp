// dn_segsolve.hpp -- partition hybrid tridiagonal solve on register arrays,
// shared by the fitting/adjoint kernels and the threshold kernel.
#pragma once
#include "dn_common.hpp"

namespace dn {

// ---------------------------------------------------------------------------
// partition hybrid solve on register arrays: rows r < nr of every lane of a
// P-lane segment form one tridiagonal system (row j = lane_in_seg * nr + r);
// a couples to row j-1, c to row j+1.  The solution replaces d.
// Local Thomas with a spike column, one cyclic-reduction step to one row per
// lane, PCR over the P lanes, back-substitution.
// ---------------------------------------------------------------------------

// Local elimination of a lane's rows (nr >= 2).  After the down and up
// sweeps the interior rows read
//   x_k = Dd[k] - Al[k] x_0 - Cg[k] x_last   (1 <= k <= nr-2),
// with x_0 the lane's first row and x_last its last row; the last row reads
// x_last = dL - aL x_0 - cL x_next (x_next: the next lane's first row), and
// the first row, with the interior rows eliminated,
// aF x_prev_last + bF x_0 + cF x_last = dF (x_prev_last: the previous lane's
// last row).
template <typename T, int R>
struct SegLocal {
  T Al[R], Cg[R], Dd[R];
  T aL, cL, dL;
  T aF, bF, cF, dF;
};

template <typename T, int R>
__device__ __forceinline__ void seg_local(int nr, const T (&a)[R], const T (&b)[R],
                                          const T (&c)[R], const T (&d)[R],
                                          SegLocal<T, R>& L) {
  // (the array stores are unconditional selects: stores of different
  // elements in the two arms of a branch get merged into one store through a
  // pointer select, which keeps the arrays out of registers)
  T al = T(0), ga = T(0), de = T(0);
#pragma unroll
  for (int r = 1; r < R; ++r) {
    const bool in = r < nr;
    if (in) {
      if (r == 1) {
        const T iv = frcp(b[1]);
        al = a[1] * iv;
        ga = c[1] * iv;
        de = d[1] * iv;
      } else {
        const T iv = frcp(b[r] - a[r] * ga);
        al = -a[r] * al * iv;
        ga = c[r] * iv;
        de = (d[r] - a[r] * de) * iv;
      }
    }
    L.Al[r] = in ? al : T(0);
    L.Cg[r] = in ? ga : T(0);
    L.Dd[r] = in ? de : T(0);
  }
  L.aL = al;
  L.cL = ga;
  L.dL = de;
  T alN = T(0), gaN = T(0), deN = T(0);
#pragma unroll
  for (int k = R - 2; k >= 1; --k) {
    if (k <= nr - 2) {
      if (k == nr - 2) {
        alN = L.Al[k];
        gaN = L.Cg[k];
        deN = L.Dd[k];
      } else {
        const T g = L.Cg[k];
        alN = L.Al[k] - g * alN;
        deN = L.Dd[k] - g * deN;
        gaN = -g * gaN;
        L.Al[k] = alN;
        L.Cg[k] = gaN;
        L.Dd[k] = deN;
      }
    }
  }
  L.bF = b[0];
  L.cF = c[0];
  L.dF = d[0];
  L.aF = a[0];
  if (nr >= 3) {
    L.bF = b[0] - c[0] * L.Al[1];
    L.cF = -c[0] * L.Cg[1];
    L.dF = d[0] - c[0] * L.Dd[1];
  }
}

// Back-substitution of a lane's rows from its first (xF) and last (xL) rows.
template <typename T, int R>
__device__ __forceinline__ void seg_back(int nr, const SegLocal<T, R>& L, T xF, T xL,
                                         T (&d)[R]) {
  d[0] = xF;
#pragma unroll
  for (int k = 1; k < R; ++k) {
    if (k < nr - 1)
      d[k] = L.Dd[k] - L.Al[k] * xF - L.Cg[k] * xL;
    else if (k == nr - 1)
      d[k] = xL;
  }
}

template <typename T, int P, int R>
__device__ __forceinline__ void seg_solve(int nr, const T (&a)[R], const T (&b)[R],
                                          const T (&c)[R], T (&d)[R]) {
  if (nr == 1) {
    d[0] = pcr_lanes<T, P>(a[0], b[0], c[0], d[0]);
    return;
  }
  SegLocal<T, R> L;
  seg_local<T, R>(nr, a, b, c, d, L);
  const T aF1 = shfl_down_seg<P>(L.aF, 1), bF1 = shfl_down_seg<P>(L.bF, 1);
  const T cF1 = shfl_down_seg<P>(L.cF, 1), dF1 = shfl_down_seg<P>(L.dF, 1);
  const T alpha = L.aL * frcp(L.bF);
  const T beta = L.cL * frcp(bF1);
  const T xL = pcr_lanes<T, P>(-alpha * L.aF, T(1) - alpha * L.cF - beta * aF1,
                               -beta * cF1, L.dL - alpha * L.dF - beta * dF1);
  const T xLp = shfl_up_seg<P>(xL, 1);
  const T xF = (L.dF - L.aF * xLp - L.cF * xL) * frcp(L.bF);
  seg_back<T, R>(nr, L, xF, xL, d);
}

// ---------------------------------------------------------------------------
// The same solve for one system spread over the NW wavefronts of a
// workgroup: thread tid = 64 w + lane holds rows tid * nr .. tid * nr + nr-1.
// Inside a wavefront every cross-lane move is the in-wave one; what crosses
// a wavefront boundary goes through LDS (`xs`, 12 NW values of T, one
// barrier per exchange; a caller that solves once per step alternates two
// such areas between steps, so that a wavefront never overwrites values
// another one has yet to read):
//  * the next lane's first row for the cyclic-reduction step (lane 63 reads
//    lane 0 of the next wavefront);
//  * the reduced one-row-per-lane system is solved per wavefront by PCR with
//    its boundary couplings cut (A = 0 on lane 0, C = 0 on lane 63) and three
//    right-hand sides: the reduced D and the unit vectors of lanes 0 and 63,
//    giving u, p, q with x = u - A_0 x_prev p - C_63 x_next q, where x_prev
//    (the previous wavefront's lane-63 unknown) and x_next (the next
//    wavefront's lane-0 unknown) solve the 2 (NW-1)-unknown interface system
//      y_b + A_0^b p_63^b y_{b-1} + C_63^b q_63^b z_b = u_63^b
//      z_b + A_0^{b+1} p_0^{b+1} y_b + C_63^{b+1} q_0^{b+1} z_{b+1} = u_0^{b+1}
//    (y_b: lane 63 of wavefront b, z_b: lane 0 of wavefront b+1), solved
//    redundantly by every lane by Gaussian elimination (the reduced system
//    is a Schur complement of the diagonally dominant step matrix, so no
//    pivoting is needed).
// NW = 1 is seg_solve over 64 lanes.
// ---------------------------------------------------------------------------

// Three-right-hand-side PCR over the 64 lanes of a wavefront (A = 0 on lane
// 0 and C = 0 on lane 63); returns the three solutions in D0, D1, D2.
template <typename T>
__device__ __forceinline__ void pcr3_wave(T A, T B, T C, T& D0, T& D1, T& D2) {
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    const T Al = shfl_up_seg<64>(A, s), Bl = shfl_up_seg<64>(B, s);
    const T Cl = shfl_up_seg<64>(C, s);
    const T D0l = shfl_up_seg<64>(D0, s), D1l = shfl_up_seg<64>(D1, s);
    const T D2l = shfl_up_seg<64>(D2, s);
    const T Ar = shfl_down_seg<64>(A, s), Br = shfl_down_seg<64>(B, s);
    const T Cr = shfl_down_seg<64>(C, s);
    const T D0r = shfl_down_seg<64>(D0, s), D1r = shfl_down_seg<64>(D1, s);
    const T D2r = shfl_down_seg<64>(D2, s);
    const T al = -A * frcp(Bl);
    const T be = -C * frcp(Br);
    const T nA = al * Al;
    const T nB = B + al * Cl + be * Ar;
    const T nC = be * Cr;
    D0 = D0 + al * D0l + be * D0r;
    D1 = D1 + al * D1l + be * D1r;
    D2 = D2 + al * D2l + be * D2r;
    A = nA;
    B = nB;
    C = nC;
  }
  const T ib = frcp(B);
  D0 *= ib;
  D1 *= ib;
  D2 *= ib;
}

// Solve of the one-row-per-thread system A x_{t-1} + B x_t + C x_{t+1} = D
// over the 64 NW threads of the workgroup.  Returns x_t; xprev receives
// x_{t-1} (the caller's own x on thread 0).
template <typename T, int NW>
__device__ __forceinline__ T pcr_block(T A, T B, T C, T D, T* xs, T& xprev) {
  const int lane = __lane_id();
  const int w = threadIdx.x >> 6;
  const T Aw = lane == 0 ? T(0) : A;
  const T Cw = lane == 63 ? T(0) : C;
  T u = D, p = lane == 0 ? T(1) : T(0), q = lane == 63 ? T(1) : T(0);
  pcr3_wave<T>(Aw, B, Cw, u, p, q);
  T* mine = xs + 8 * w;
  if (lane == 0) {
    mine[0] = u;
    mine[1] = p;
    mine[2] = q;
    mine[3] = A;
  }
  if (lane == 63) {
    mine[4] = u;
    mine[5] = p;
    mine[6] = q;
    mine[7] = C;
  }
  __syncthreads();
  constexpr int N = 2 * (NW - 1);
  T M[N][N], rhs[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = 0; j < N; ++j) M[i][j] = T(0);
  }
#pragma unroll
  for (int bnd = 0; bnd < NW - 1; ++bnd) {
    const T* lo = xs + 8 * bnd;        // wavefront b
    const T* hi = xs + 8 * (bnd + 1);  // wavefront b + 1
    const int y = 2 * bnd, z = 2 * bnd + 1;
    // y_b row: wavefront b's lane 63
    M[y][y] = T(1);
    if (bnd > 0) M[y][y - 2] = lo[3] * lo[5];  // A_0^b p_63^b
    M[y][z] = lo[7] * lo[6];                   // C_63^b q_63^b
    rhs[y] = lo[4];
    // z_b row: wavefront b+1's lane 0
    M[z][z] = T(1);
    M[z][y] = hi[3] * hi[1];                          // A_0^{b+1} p_0^{b+1}
    if (bnd + 1 < NW - 1) M[z][z + 2] = hi[7] * hi[2];  // C_63^{b+1} q_0^{b+1}
    rhs[z] = hi[0];
  }
  // banded elimination: every row couples at most two unknowns to either
  // side (y_b to y_{b-1}, z_b to z_{b+1}), and the fill stays in that band
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const T iv = frcp(M[k][k]);
#pragma unroll
    for (int i = k + 1; i < N && i <= k + 2; ++i) {
      const T f = M[i][k] * iv;
#pragma unroll
      for (int j = k + 1; j < N && j <= k + 2; ++j) M[i][j] -= f * M[k][j];
      rhs[i] -= f * rhs[k];
    }
  }
#pragma unroll
  for (int k = N - 1; k >= 0; --k) {
    T s = rhs[k];
#pragma unroll
    for (int j = k + 1; j < N && j <= k + 2; ++j) s -= M[k][j] * rhs[j];
    rhs[k] = s * frcp(M[k][k]);
  }
  // x_prev of this wavefront (y_{w-1}) and x_next (z_w)
  T yprev = T(0), znext = T(0);
#pragma unroll
  for (int bnd = 0; bnd < NW - 1; ++bnd) {
    if (bnd == w - 1) yprev = rhs[2 * bnd];
    if (bnd == w) znext = rhs[2 * bnd + 1];
  }
  const T x = u - mine[3] * yprev * p - mine[7] * znext * q;
  const T xin = shfl_up_seg<64>(x, 1);
  xprev = (lane == 0 && w > 0) ? yprev : xin;
  return x;
}

template <typename T, int R, int NW>
__device__ __forceinline__ void blk_solve(int nr, const T (&a)[R], const T (&b)[R],
                                          const T (&c)[R], T (&d)[R], T* xs) {
  if constexpr (NW == 1) {
    seg_solve<T, 64, R>(nr, a, b, c, d);
  } else {
    const int lane = __lane_id();
    const int w = threadIdx.x >> 6;
    T xprev;
    if (nr == 1) {
      d[0] = pcr_block<T, NW>(a[0], b[0], c[0], d[0], xs + 4 * NW, xprev);
      return;
    }
    SegLocal<T, R> L;
    seg_local<T, R>(nr, a, b, c, d, L);
    // the next thread's first row; lane 0 publishes it for the previous wavefront
    T aF1 = shfl_down_seg<64>(L.aF, 1), bF1 = shfl_down_seg<64>(L.bF, 1);
    T cF1 = shfl_down_seg<64>(L.cF, 1), dF1 = shfl_down_seg<64>(L.dF, 1);
    if (lane == 0) {
      xs[4 * w] = L.aF;
      xs[4 * w + 1] = L.bF;
      xs[4 * w + 2] = L.cF;
      xs[4 * w + 3] = L.dF;
    }
    __syncthreads();
    if (lane == 63 && w < NW - 1) {
      aF1 = xs[4 * (w + 1)];
      bF1 = xs[4 * (w + 1) + 1];
      cF1 = xs[4 * (w + 1) + 2];
      dF1 = xs[4 * (w + 1) + 3];
    }
    const T alpha = L.aL * frcp(L.bF);
    const T beta = L.cL * frcp(bF1);
    const T xL = pcr_block<T, NW>(-alpha * L.aF, T(1) - alpha * L.cF - beta * aF1,
                                  -beta * cF1, L.dL - alpha * L.dF - beta * dF1,
                                  xs + 4 * NW, xprev);
    const T xF = (L.dF - L.aF * xprev - L.cF * xL) * frcp(L.bF);
    seg_back<T, R>(nr, L, xF, xL, d);
  }
}

// Values of x on the neighbouring threads of a workgroup-wide segment:
// prev = xlast of thread t-1, next = xfirst of thread t+1 (the caller's own
// values on the first and last threads).  One barrier when NW > 1 (`xs`
// holds 2 NW values).
template <typename T, int NW>
__device__ __forceinline__ void blk_neighbours(T xlast, T xfirst, T* xs, T& prev, T& next) {
  prev = shfl_up_seg<64>(xlast, 1);
  next = shfl_down_seg<64>(xfirst, 1);
  if constexpr (NW > 1) {
    const int lane = __lane_id();
    const int w = threadIdx.x >> 6;
    if (lane == 63) xs[2 * w] = xlast;
    if (lane == 0) xs[2 * w + 1] = xfirst;
    __syncthreads();
    if (lane == 0 && w > 0) prev = xs[2 * (w - 1)];
    if (lane == 63 && w < NW - 1) next = xs[2 * (w + 1) + 1];
  }
}

}  // namespace dn
