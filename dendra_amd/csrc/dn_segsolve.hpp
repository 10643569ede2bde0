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
template <typename T, int P, int R>
__device__ __forceinline__ void seg_solve(int nr, const T (&a)[R], const T (&b)[R],
                                          const T (&c)[R], T (&d)[R]) {
  if (nr == 1) {
    d[0] = pcr_lanes<T, P>(a[0], b[0], c[0], d[0]);
    return;
  }
  T Al[R], Cg[R], Dd[R];
  T al = T(0), ga = T(0), de = T(0);
#pragma unroll
  for (int r = 1; r < R; ++r) {
    if (r < nr) {
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
      Al[r] = al;
      Cg[r] = ga;
      Dd[r] = de;
    } else {
      Al[r] = Cg[r] = Dd[r] = T(0);
    }
  }
  const T aL = al, cL = ga, dL = de;
  T alN = T(0), gaN = T(0), deN = T(0);
#pragma unroll
  for (int k = R - 2; k >= 1; --k) {
    if (k <= nr - 2) {
      if (k == nr - 2) {
        alN = Al[k];
        gaN = Cg[k];
        deN = Dd[k];
      } else {
        const T g = Cg[k];
        alN = Al[k] - g * alN;
        deN = Dd[k] - g * deN;
        gaN = -g * gaN;
        Al[k] = alN;
        Cg[k] = gaN;
        Dd[k] = deN;
      }
    }
  }
  T bF = b[0], cF = c[0], dF = d[0];
  const T aF = a[0];
  if (nr >= 3) {
    bF = b[0] - c[0] * Al[1];
    cF = -c[0] * Cg[1];
    dF = d[0] - c[0] * Dd[1];
  }
  const T aF1 = shfl_down_seg<P>(aF, 1), bF1 = shfl_down_seg<P>(bF, 1);
  const T cF1 = shfl_down_seg<P>(cF, 1), dF1 = shfl_down_seg<P>(dF, 1);
  const T alpha = aL * frcp(bF);
  const T beta = cL * frcp(bF1);
  const T xL = pcr_lanes<T, P>(-alpha * aF, T(1) - alpha * cF - beta * aF1, -beta * cF1,
                               dL - alpha * dF - beta * dF1);
  const T xLp = shfl_up_seg<P>(xL, 1);
  const T xF = (dF - aF * xLp - cF * xL) * frcp(bF);
  d[0] = xF;
#pragma unroll
  for (int k = 1; k < R; ++k) {
    if (k < nr - 1)
      d[k] = Dd[k] - Al[k] * xF - Cg[k] * xL;
    else if (k == nr - 1)
      d[k] = xL;
  }
}

}  // namespace dn
