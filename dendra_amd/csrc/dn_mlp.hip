// dn_mlp.hip -- threshold surrogate: BF16 MFMA multilayer perceptron,
// training (forward, backward, Adam) and ensemble inference on gfx950.
//
// Network (per ensemble member z): x[128] (102 features, zero padded) ->
// 512 -> 512 -> 512 (GELU, exact erf form) -> 1.  Weights are stored as
// float32 masters (Adam state) and rounded to BF16 copies W[out][in] and
// W^T[in][out] for the matrix cores; every product accumulates in float32.
// Activations cross layers as BF16 in both layouts ([batch][feature] and
// [feature][batch]) so that each product is "NT": both operands contiguous
// along the reduction index, the natural operand layout of
// v_mfma_f32_32x32x16_bf16 (lane l holds row l%32, k = 8*(l/32)..+7; result
// register r of lane l is row 8*(r/4) + 4*(l/32) + r%4, column l%32).
//
//   forward    Z_l = H_{l-1} W_l^T + b_l,  H_l = gelu(Z_l)      (gemm, EPI_GELU)
//   output     y = H_3 w_4 + b_4,  dy = 2 (y - t) / B              (mlp_out_kernel)
//   backward   dZ_3 = dy w_4^T * gelu'(Z_3)                         (mlp_out_bwd_kernel)
//              dZ_{l-1} = (dZ_l W_l) * gelu'(Z_{l-1})               (gemm, EPI_DGELU)
//              dW_l = dZ_l^T H_{l-1}  (split over the batch)        (gemm, EPI_F32 + reduce)
//              db_l = sum_b dZ_l,  dw_4 = H_3^T dy                  (row_dot_kernel)
//   update     Adam on float32 masters, BF16 copies rewritten       (adam_kernel)
#include <hip/hip_bf16.h>

#include "dn_common.hpp"

namespace dn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTM = 128, kTN = 128, kTK = 64;
constexpr int kAS = kTK + 8;    // LDS row pitch (144 B: 16-byte reads of 16 rows hit distinct banks)
constexpr int kTS = kTM + 8;    // LDS row pitch of the transposed output tile
constexpr int kStage = 2 * kTM * kAS;  // one stage: A and B tiles

enum { EPI_GELU = 0, EPI_DGELU = 1, EPI_F32 = 2 };

struct GemmArgs {
  int M, N, K, splits;
  const bf16* A; int64_t lda, sa;     // A[z][m*lda + k]
  const bf16* B; int64_t ldb, sb;     // B[z][n*ldb + k]
  const float* bias; int64_t sbias;   // EPI_GELU: bias[z][n]
  float* Z; int64_t ldz, sz;          // EPI_GELU: pre-activation; EPI_F32: C
  bf16* H; int64_t ldh, sh;           // BF16 result [m][n]
  bf16* HT; int64_t ldht, sht;        // BF16 result transposed [n][m]
  const float* Zp; int64_t ldzp, szp; // EPI_DGELU: pre-activation multiplier input
};

// gelu(z) = z Phi(z) with Phi(z) = erfc(-z / sqrt(2)) / 2: erfc keeps full
// relative accuracy for negative z, where 1 + erf(z / sqrt(2)) cancels.
__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * erfcf(-z * 0.70710678118654752f);
}
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * erfcf(-z * 0.70710678118654752f) +
         z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt_kernel(GemmArgs G) {
  // two stages of (A, B) tiles; the epilogue reuses the space for the
  // transposed output tile
  __shared__ __align__(16) bf16 smem[2 * kStage];
  static_assert(2 * kStage >= kTN * kTS, "transpose tile must fit");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
  const int z = blockIdx.z / G.splits, split = blockIdx.z % G.splits;
  const int kspan = G.K / G.splits;
  const int kbeg = split * kspan;
  const int ntile = kspan / kTK;
  const bf16* A = G.A + z * G.sa;
  const bf16* B = G.B + z * G.sb;
  f32x16 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
  // each thread moves 4 16-byte chunks of A and 4 of B per tile:
  // chunk c = tid + 256 q -> row c / 8, k offset (c % 8) * 8; the chunks are
  // named registers (no array) so that they never leave the register file
  const int crow = tid >> 3, ckc = (tid & 7) * 8;
  const bf16* gA = A + (int64_t)(m0 + crow) * G.lda + ckc;
  const bf16* gB = B + (int64_t)(n0 + crow) * G.ldb + ckc;
  const int64_t sa32 = 32 * G.lda, sb32 = 32 * G.ldb;
  uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define DN_GEMM_FETCH(K0)                                                             \
  {                                                                                   \
    const int k_ = (K0);                                                              \
    ra0 = *reinterpret_cast<const uint4*>(gA + k_);                                   \
    ra1 = *reinterpret_cast<const uint4*>(gA + sa32 + k_);                            \
    ra2 = *reinterpret_cast<const uint4*>(gA + 2 * sa32 + k_);                        \
    ra3 = *reinterpret_cast<const uint4*>(gA + 3 * sa32 + k_);                        \
    rb0 = *reinterpret_cast<const uint4*>(gB + k_);                                   \
    rb1 = *reinterpret_cast<const uint4*>(gB + sb32 + k_);                            \
    rb2 = *reinterpret_cast<const uint4*>(gB + 2 * sb32 + k_);                        \
    rb3 = *reinterpret_cast<const uint4*>(gB + 3 * sb32 + k_);                        \
  }
#define DN_GEMM_STASH(STAGE)                                                          \
  {                                                                                   \
    bf16* sA_ = smem + (STAGE) * kStage + crow * kAS + ckc;                           \
    bf16* sB_ = sA_ + kTM * kAS;                                                      \
    *reinterpret_cast<uint4*>(sA_) = ra0;                                             \
    *reinterpret_cast<uint4*>(sA_ + 32 * kAS) = ra1;                                  \
    *reinterpret_cast<uint4*>(sA_ + 64 * kAS) = ra2;                                  \
    *reinterpret_cast<uint4*>(sA_ + 96 * kAS) = ra3;                                  \
    *reinterpret_cast<uint4*>(sB_) = rb0;                                             \
    *reinterpret_cast<uint4*>(sB_ + 32 * kAS) = rb1;                                  \
    *reinterpret_cast<uint4*>(sB_ + 64 * kAS) = rb2;                                  \
    *reinterpret_cast<uint4*>(sB_ + 96 * kAS) = rb3;                                  \
  }
  DN_GEMM_FETCH(kbeg)
  DN_GEMM_STASH(0)
  __syncthreads();
  for (int t = 0; t < ntile; ++t) {
    if (t + 1 < ntile) {  // in flight during the products
      DN_GEMM_FETCH(kbeg + (t + 1) * kTK)
    }
    const bf16* sA = smem + (t & 1) * kStage;
    const bf16* sB = sA + kTM * kAS;
#pragma unroll
    for (int ks = 0; ks < kTK; ks += 16) {
      bf16x8 af[2], bfr[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        af[u] = *reinterpret_cast<const bf16x8*>(
            &sA[(wm * 64 + u * 32 + (lane & 31)) * kAS + ks + 8 * (lane >> 5)]);
        bfr[u] = *reinterpret_cast<const bf16x8*>(
            &sB[(wn * 64 + u * 32 + (lane & 31)) * kAS + ks + 8 * (lane >> 5)]);
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[a], bfr[b], acc[a][b], 0, 0,
                                                              0);
    }
    // the other stage was last read before the previous barrier
    if (t + 1 < ntile) {
      DN_GEMM_STASH((t + 1) & 1)
    }
    __syncthreads();
  }
#undef DN_GEMM_FETCH
#undef DN_GEMM_STASH
  // epilogue: element (i, j) of the block tile held by (a, b, r, lane)
  const bool want_t = (EPI != EPI_F32) && G.HT;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = wm * 64 + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int j = wn * 64 + b * 32 + (lane & 31);
        const int64_t m = m0 + i, n = n0 + j;
        float v = acc[a][b][r];
        if (EPI == EPI_F32) {
          G.Z[(int64_t)blockIdx.z * G.sz + m * G.ldz + n] = v;
          continue;
        }
        if (EPI == EPI_GELU) {
          v += G.bias[z * G.sbias + n];
          if (G.Z) G.Z[z * G.sz + m * G.ldz + n] = v;
          v = gelu(v);
        } else {
          v *= gelu_d(G.Zp[z * G.szp + m * G.ldzp + n]);
        }
        const bf16 hv = (bf16)v;
        if (G.H) G.H[z * G.sh + m * G.ldh + n] = hv;
        if (want_t) smem[j * kTS + i] = hv;
      }
  if (want_t) {
    __syncthreads();
#pragma unroll
    for (int c2 = 0; c2 < 8; ++c2) {
      const int c = tid + c2 * 256;
      const int jr = c >> 4, mc = (c & 15) * 8;
      *reinterpret_cast<uint4*>(&G.HT[z * G.sht + (int64_t)(n0 + jr) * G.ldht + m0 + mc]) =
          *reinterpret_cast<const uint4*>(&smem[jr * kTS + mc]);
    }
  }
}

// Output unit: y = H3 . w4 + b4 (one wavefront per sample).  With targets,
// also r = y - t and dy = 2 r / B.
__global__ __launch_bounds__(256) void mlp_out_kernel(int Bn, int Hd, const bf16* H3,
                                                      int64_t sh, const bf16* w4,
                                                      int64_t sw, const float* b4,
                                                      const float* t, int64_t st, float* y,
                                                      float* dy, int64_t sy, int nz) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (wave >= Bn * nz) return;
  const int z = wave / Bn, b = wave % Bn;
  const bf16* h = H3 + z * sh + (int64_t)b * Hd;
  const bf16* w = w4 + z * sw;
  float s = 0.0f;
  for (int k = lane * 8; k < Hd; k += 512) {
    const bf16x8 hv = *reinterpret_cast<const bf16x8*>(h + k);
    const bf16x8 wv = *reinterpret_cast<const bf16x8*>(w + k);
#pragma unroll
    for (int e = 0; e < 8; ++e) s += (float)hv[e] * (float)wv[e];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) {
    const float yv = s + b4[z];
    y[z * sy + b] = yv;
    if (t) dy[z * sy + b] = 2.0f * (yv - t[z * st + b]) / (float)Bn;
  }
}

// dZ3[b][k] = dy[b] w4[k] gelu'(Z3[b][k]) in [b][k] and [k][b] layouts
// (64 x 64 tiles through LDS).
__global__ __launch_bounds__(256) void mlp_out_bwd_kernel(int Bn, int Hd, const float* dy,
                                                          int64_t sy, const bf16* w4,
                                                          int64_t sw, const float* Z3,
                                                          int64_t sz, bf16* dZ, bf16* dZT,
                                                          int64_t sd) {
  __shared__ __align__(16) bf16 tile[64 * 72];
  const int z = blockIdx.z;
  const int b0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int bi = e >> 6, ki = e & 63;
    const int64_t b = b0 + bi, k = k0 + ki;
    const float g = dy[z * sy + b] * (float)w4[z * sw + k] *
                    gelu_d(Z3[z * sz + b * Hd + k]);
    const bf16 gv = (bf16)g;
    dZ[z * sd + b * Hd + k] = gv;
    tile[ki * 72 + bi] = gv;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < 64 * 8; c += 256) {
    const int ki = c >> 3, bc = (c & 7) * 8;
    *reinterpret_cast<uint4*>(&dZT[z * sd + (int64_t)(k0 + ki) * Bn + b0 + bc]) =
        *reinterpret_cast<const uint4*>(&tile[ki * 72 + bc]);
  }
}

// out[z][row] = sum_b X[z][row][b] * (w ? w[z][b] : 1), one wavefront per row,
// fixed summation order.
template <typename TX>
__global__ __launch_bounds__(256) void row_dot_kernel(int rows, int cols, const TX* X,
                                                      int64_t sx, const float* w, int64_t swz,
                                                      float* out, int64_t so, int nz) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (wave >= rows * nz) return;
  const int z = wave / rows, row = wave % rows;
  const TX* x = X + z * sx + (int64_t)row * cols;
  const float* wz = w ? w + z * swz : nullptr;
  float s = 0.0f;
  if (sizeof(TX) == 2 && (cols % 8) == 0 &&
      (reinterpret_cast<uintptr_t>(x) % 16) == 0) {
    // 8 BF16 values (16 bytes) per lane and load
    for (int c = lane * 8; c < cols; c += 512) {
      const bf16x8 xv = *reinterpret_cast<const bf16x8*>(x + c);
#pragma unroll
      for (int e = 0; e < 8; ++e) s += (float)xv[e] * (wz ? wz[c + e] : 1.0f);
    }
  } else {
    for (int c = lane; c < cols; c += 64) s += (float)x[c] * (wz ? wz[c] : 1.0f);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[z * so + row] = s;
}

// dW[z][i] = sum_s P[z][s][i]
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int64_t n, int splits,
                                                            const float* P, float* out,
                                                            int nz) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nz) return;
  const int64_t z = t / n, i = t % n;
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += P[(z * splits + k) * n + i];
  out[t] = s;
}

// Adam on float32 masters [z][rows][cols]; BF16 copies [z][rows][cols] and
// transposed [z][cols][rows] (either may be null).
__global__ __launch_bounds__(256) void adam_kernel(int64_t n, int rows, int cols, float* p,
                                                   const float* g, float* m, float* v,
                                                   float lr_t, float b1, float b2, float eps,
                                                   bf16* pb, bf16* pbt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gi = g[t];
  const float mi = b1 * m[t] + (1.0f - b1) * gi;
  const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  const float pi = p[t] - lr_t * mi / (sqrtf(vi) + eps);
  p[t] = pi;
  const int64_t per = (int64_t)rows * cols;
  const int64_t z = t / per, e = t % per, r = e / cols, c = e % cols;
  if (pb) pb[t] = (bf16)pi;
  if (pbt) pbt[z * per + c * rows + r] = (bf16)pi;
}

// Features: x[0..100] = profile / max|profile|, x[101] = (d - dc) / ds, rest 0;
// lmag = log(max|profile|).  With bounds, flag = 1 when any feature or lmag
// lies outside [lo - margin, hi + margin].
__global__ __launch_bounds__(256) void featurize_kernel(int64_t n, int P, const float* prof,
                                                        const float* diam, float dc, float ds,
                                                        bf16* X, int ldx, float* lmag,
                                                        const float* lo, const float* hi,
                                                        float margin, int32_t* flag) {
  const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (s >= n) return;
  const float* p = prof + s * P;
  float mx = 0.0f;
  for (int j = lane; j < P; j += 64) mx = fmaxf(mx, fabsf(p[j]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  const float inv = mx > 0.0f ? 1.0f / mx : 0.0f;
  const float lm = logf(mx);
  bool out = false;
  for (int j = lane; j < ldx; j += 64) {
    float f = 0.0f;
    if (j < P) f = p[j] * inv;
    else if (j == P) f = (diam[s] - dc) / ds;
    X[s * ldx + j] = (bf16)f;
    if (lo && j <= P) out = out || f < lo[j] - margin || f > hi[j] + margin;
  }
  if (lo && lane == 0) out = out || lm < lo[P + 1] - margin || lm > hi[P + 1] + margin;
  const bool any_out = __ballot(out) != 0ull;
  if (lane == 0) {
    lmag[s] = lm;
    if (flag) flag[s] = any_out ? 1 : 0;
  }
}

// Minibatch gather: X0[z][b][:] = X[idx[z][b]][:], X0T[z][:][b], t0[z][b] = t[idx].
__global__ __launch_bounds__(256) void gather_kernel(int Bn, int F, const int32_t* idx,
                                                     const bf16* X, const float* t,
                                                     bf16* X0, bf16* X0T, float* t0) {
  __shared__ __align__(16) bf16 tile[128 * 72];
  const int z = blockIdx.z, b0 = blockIdx.x * 64;
  const int32_t* id = idx + (int64_t)z * Bn;
  for (int e = threadIdx.x; e < 64 * F; e += 256) {
    const int bi = e / F, f = e % F;
    const bf16 v = X[(int64_t)id[b0 + bi] * F + f];
    X0[((int64_t)z * Bn + b0 + bi) * F + f] = v;
    tile[f * 72 + bi] = v;
  }
  if (threadIdx.x < 64) t0[(int64_t)z * Bn + b0 + threadIdx.x] = t[id[b0 + threadIdx.x]];
  __syncthreads();
  for (int c = threadIdx.x; c < F * 8; c += 256) {
    const int f = c >> 3, bc = (c & 7) * 8;
    *reinterpret_cast<uint4*>(&X0T[((int64_t)z * F + f) * Bn + b0 + bc]) =
        *reinterpret_cast<const uint4*>(&tile[f * 72 + bc]);
  }
}

// Ensemble: thr_z = exp(pred_z * ystd + ymean - lmag); out = median of the
// members; flag |= (max - min) / median > spread (flag may be preset).
__global__ __launch_bounds__(256) void ensemble_kernel(int64_t n, int nz, const float* pred,
                                                       int64_t sp, const float* lmag,
                                                       float ymean, float ystd, float spread,
                                                       float* thr, int32_t* flag) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float v[8];
  for (int z = 0; z < nz; ++z) v[z] = __expf(pred[z * sp + s] * ystd + ymean - lmag[s]);
  for (int i = 1; i < nz; ++i)
    for (int j = i; j > 0 && v[j - 1] > v[j]; --j) {
      const float tmp = v[j];
      v[j] = v[j - 1];
      v[j - 1] = tmp;
    }
  const float med = (nz & 1) ? v[nz / 2] : 0.5f * (v[nz / 2 - 1] + v[nz / 2]);
  thr[s] = med;
  if ((v[nz - 1] - v[0]) > spread * med) flag[s] = 1;
}

}  // namespace dn

using dn::bf16;

static inline bool gemm_shape_ok(int M, int N, int K, int splits) {
  return M > 0 && N > 0 && K > 0 && splits > 0 && M % dn::kTM == 0 && N % dn::kTN == 0 &&
         K % (dn::kTK * splits) == 0;
}

// C = A B^T for nz independent products (A: [M][K] rows, B: [N][K] rows, both
// BF16, float32 accumulation), with one of three epilogues:
//   epi 0: Z = C + bias (optional store), H = gelu(Z) in BF16, HT = H^T (optional)
//   epi 1: H = C * gelu'(Zp) in BF16, HT = H^T (optional)
//   epi 2: Z[z*splits + s] = partial C over the s-th of `splits` K ranges
// M, N multiples of 128, K a multiple of 32 * splits; strides in elements.
DN_EXPORT int dn_mlp_gemm(int epi, int nz, int M, int N, int K, int splits,
                          const void* A, int64_t lda, int64_t sa, const void* B, int64_t ldb,
                          int64_t sb, const float* bias, int64_t sbias, float* Z,
                          int64_t ldz, int64_t sz, void* H, int64_t ldh, int64_t sh,
                          void* HT, int64_t ldht, int64_t sht, const float* Zp,
                          int64_t ldzp, int64_t szp, hipStream_t stream) {
  if (!gemm_shape_ok(M, N, K, splits) || nz < 1 || !A || !B || lda % 8 || ldb % 8 ||
      (HT && (ldht % 8)))
    return DN_EINVAL;
  if (epi == 0 && (!bias || splits != 1 || (!H && !HT && !Z))) return DN_EINVAL;
  if (epi == 1 && (!Zp || splits != 1 || (!H && !HT))) return DN_EINVAL;
  if (epi == 2 && !Z) return DN_EINVAL;
  if (epi < 0 || epi > 2) return DN_EINVAL;
  if ((int64_t)nz * splits > 65535) return DN_ERANGE;
  dn::GemmArgs G;
  G.M = M; G.N = N; G.K = K; G.splits = splits;
  G.A = static_cast<const bf16*>(A); G.lda = lda; G.sa = sa;
  G.B = static_cast<const bf16*>(B); G.ldb = ldb; G.sb = sb;
  G.bias = bias; G.sbias = sbias; G.Z = Z; G.ldz = ldz; G.sz = sz;
  G.H = static_cast<bf16*>(H); G.ldh = ldh; G.sh = sh;
  G.HT = static_cast<bf16*>(HT); G.ldht = ldht; G.sht = sht;
  G.Zp = Zp; G.ldzp = ldzp; G.szp = szp;
  const dim3 grid(N / dn::kTN, M / dn::kTM, nz * splits), block(256);
  if (epi == 0) hipLaunchKernelGGL(dn::gemm_nt_kernel<dn::EPI_GELU>, grid, block, 0, stream, G);
  else if (epi == 1)
    hipLaunchKernelGGL(dn::gemm_nt_kernel<dn::EPI_DGELU>, grid, block, 0, stream, G);
  else hipLaunchKernelGGL(dn::gemm_nt_kernel<dn::EPI_F32>, grid, block, 0, stream, G);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// y[z][b] = H3[z][b][:] . w4[z][:] + b4[z]; with t: dy = 2 (y - t) / Bn.
DN_EXPORT int dn_mlp_out(int nz, int Bn, int Hd, const void* H3, int64_t sh, const void* w4,
                         int64_t sw, const float* b4, const float* t, int64_t st, float* y,
                         float* dy, int64_t sy, hipStream_t stream) {
  if (nz < 1 || Bn < 1 || Hd < 8 || Hd % 8 || !H3 || !w4 || !b4 || !y || (t && !dy))
    return DN_EINVAL;
  const int64_t threads = (int64_t)nz * Bn * 64;
  hipLaunchKernelGGL(dn::mlp_out_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256),
                     0, stream, Bn, Hd, static_cast<const bf16*>(H3), sh,
                     static_cast<const bf16*>(w4), sw, b4, t, st, y, dy, sy, nz);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// dZ3 = dy w4^T * gelu'(Z3) as [z][b][k] and [z][k][b] BF16; Bn, Hd multiples of 64.
DN_EXPORT int dn_mlp_out_bwd(int nz, int Bn, int Hd, const float* dy, int64_t sy,
                             const void* w4, int64_t sw, const float* Z3, int64_t sz,
                             void* dZ, void* dZT, int64_t sd, hipStream_t stream) {
  if (nz < 1 || Bn < 64 || Hd < 64 || Bn % 64 || Hd % 64 || !dy || !w4 || !Z3 || !dZ || !dZT)
    return DN_EINVAL;
  hipLaunchKernelGGL(dn::mlp_out_bwd_kernel, dim3(Hd / 64, Bn / 64, nz), dim3(256), 0, stream,
                     Bn, Hd, dy, sy, static_cast<const bf16*>(w4), sw, Z3, sz,
                     static_cast<bf16*>(dZ), static_cast<bf16*>(dZT), sd);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// out[z][r] = sum_c X[z][r][c] * (w ? w[z][c] : 1); x_bf16 selects BF16 or float32 X.
DN_EXPORT int dn_mlp_row_dot(int nz, int rows, int cols, int x_bf16, const void* X,
                             int64_t sx, const float* w, int64_t sw, float* out, int64_t so,
                             hipStream_t stream) {
  if (nz < 1 || rows < 1 || cols < 1 || !X || !out) return DN_EINVAL;
  const int64_t threads = (int64_t)nz * rows * 64;
  const dim3 grid((unsigned)((threads + 255) / 256));
  if (x_bf16)
    hipLaunchKernelGGL(dn::row_dot_kernel<bf16>, grid, dim3(256), 0, stream, rows, cols,
                       static_cast<const bf16*>(X), sx, w, sw, out, so, nz);
  else
    hipLaunchKernelGGL(dn::row_dot_kernel<float>, grid, dim3(256), 0, stream, rows, cols,
                       static_cast<const float*>(X), sx, w, sw, out, so, nz);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int dn_mlp_splitk_reduce(int nz, int64_t n, int splits, const float* P, float* out,
                                   hipStream_t stream) {
  if (nz < 1 || n < 1 || splits < 1 || !P || !out) return DN_EINVAL;
  const int64_t t = n * nz;
  hipLaunchKernelGGL(dn::splitk_reduce_kernel, dim3((unsigned)((t + 255) / 256)), dim3(256),
                     0, stream, n, splits, P, out, nz);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// One Adam step over n = nz*rows*cols float32 parameters; lr_t is the
// bias-corrected step lr * sqrt(1 - b2^t) / (1 - b1^t).
DN_EXPORT int dn_mlp_adam(int64_t n, int rows, int cols, float* p, const float* g, float* m,
                          float* v, float lr_t, float b1, float b2, float eps, void* pb,
                          void* pbt, hipStream_t stream) {
  if (n < 1 || rows < 1 || cols < 1 || n % ((int64_t)rows * cols) || !p || !g || !m || !v)
    return DN_EINVAL;
  hipLaunchKernelGGL(dn::adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                     n, rows, cols, p, g, m, v, lr_t, b1, b2, eps, static_cast<bf16*>(pb),
                     static_cast<bf16*>(pbt));
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int dn_mlp_featurize(int64_t n, int P, const float* prof, const float* diam,
                               float dc, float ds, void* X, int ldx, float* lmag,
                               const float* lo, const float* hi, float margin, int32_t* flag,
                               hipStream_t stream) {
  if (n < 1 || P < 1 || ldx < P + 1 || ldx % 8 || !prof || !diam || !X || !lmag ||
      (lo && (!hi || !flag)))
    return DN_EINVAL;
  const int64_t threads = n * 64;
  hipLaunchKernelGGL(dn::featurize_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256),
                     0, stream, n, P, prof, diam, dc, ds, static_cast<bf16*>(X), ldx, lmag, lo,
                     hi, margin, flag);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// Bn a multiple of 64, F <= 128 and a multiple of 8.
DN_EXPORT int dn_mlp_gather(int nz, int Bn, int F, const int32_t* idx, const void* X,
                            const float* t, void* X0, void* X0T, float* t0,
                            hipStream_t stream) {
  if (nz < 1 || Bn < 64 || Bn % 64 || F < 8 || F > 128 || F % 8 || !idx || !X || !t || !X0 ||
      !X0T || !t0)
    return DN_EINVAL;
  hipLaunchKernelGGL(dn::gather_kernel, dim3(Bn / 64, 1, nz), dim3(256), 0, stream, Bn, F, idx,
                     static_cast<const bf16*>(X), t, static_cast<bf16*>(X0),
                     static_cast<bf16*>(X0T), t0);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

DN_EXPORT int dn_mlp_ensemble(int64_t n, int nz, const float* pred, int64_t sp,
                              const float* lmag, float ymean, float ystd, float spread,
                              float* thr, int32_t* flag, hipStream_t stream) {
  if (n < 1 || nz < 1 || nz > 8 || !pred || !lmag || !thr || !flag) return DN_EINVAL;
  hipLaunchKernelGGL(dn::ensemble_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     stream, n, nz, pred, sp, lmag, ymean, ystd, spread, thr, flag);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}
