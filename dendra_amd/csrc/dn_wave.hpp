// dn_wave.hpp -- one-wavefront hybrid (partition Thomas + PCR) tridiagonal
// solve over padded arrays in LDS or global memory.  Shared by the batched
// drop-in solver and the plain per-step cable path.
#pragma once
#include "dn_common.hpp"

namespace dn {

// In-place hybrid solve of one system held in padded arrays of 64*nr rows:
// a[j] couples row j to j-1 (a[0] = 0), c[j] couples to j+1 (c[last] = 0).
// Rows >= K are identity rows.  On return d holds x.  Called by all 64 lanes.
template <typename T>
__device__ void wave_hybrid_solve(T* a, T* b, T* c, T* d, int nr) {
  const int lane = __lane_id();
  if (nr == 1) {
    T x = pcr_lanes<T, 64>(a[lane], b[lane], c[lane], d[lane]);
    d[lane] = x;
    return;
  }
  T* A = a + lane * nr;
  T* Bm = b + lane * nr;
  T* C = c + lane * nr;
  T* D = d + lane * nr;
  const T a0 = A[0], b0 = Bm[0], c0 = C[0], d0 = D[0];
  // forward: rows k>=1 become  al_k x_0 + x_k + ga_k x_{k+1} = de_k
  T inv = frcp(Bm[1]);
  T al = A[1] * inv, ga = C[1] * inv, de = D[1] * inv;
  A[1] = al; C[1] = ga; D[1] = de;
  for (int k = 2; k < nr; ++k) {
    T ak = A[k];
    T iv = frcp(Bm[k] - ak * ga);
    al = -ak * al * iv;
    ga = C[k] * iv;
    de = (D[k] - ak * de) * iv;
    A[k] = al; C[k] = ga; D[k] = de;
  }
  const T aL = al, cL = ga, dL = de;  // bL = 1
  // backward: rows k in [1, nr-2] become x_k = de'' - al'' x_0 - ga'' x_last
  T alN = al, gaN = ga, deN = de;  // values of row k+1 (row nr-1 is special)
  bool first = true;
  for (int k = nr - 2; k >= 1; --k) {
    T alk = A[k], gak = C[k], dek = D[k];
    if (!first) {
      alk = alk - gak * alN;
      dek = dek - gak * deN;
      gak = -gak * gaN;
    }
    first = false;
    A[k] = alk; C[k] = gak; D[k] = dek;
    alN = alk; gaN = gak; deN = dek;
  }
  T aF = a0, bF = b0, cF = c0, dF = d0;
  if (nr >= 3) {
    bF = b0 - c0 * A[1];
    cF = -c0 * C[1];
    dF = d0 - c0 * D[1];
  }
  // cyclic-reduction step: eliminate the first-row unknowns of every lane.
  T aF1 = shfl_down_seg<64>(aF, 1), bF1 = shfl_down_seg<64>(bF, 1);
  T cF1 = shfl_down_seg<64>(cF, 1), dF1 = shfl_down_seg<64>(dF, 1);
  T alpha = aL * frcp(bF);
  T beta = cL * frcp(bF1);
  T RA = -alpha * aF;
  T RB = T(1) - alpha * cF - beta * aF1;
  T RC = -beta * cF1;
  T RD = dL - alpha * dF - beta * dF1;
  T xL = pcr_lanes<T, 64>(RA, RB, RC, RD);
  T xLp = shfl_up_seg<64>(xL, 1);
  T xF = (dF - aF * xLp - cF * xL) * frcp(bF);
  D[0] = xF;
  for (int k = 1; k < nr - 1; ++k) D[k] = D[k] - A[k] * xF - C[k] * xL;
  D[nr - 1] = xL;
}

}  // namespace dn
