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

// gelu(z) = z Phi(z), Phi(z) = erfc(-z / sqrt 2) / 2.  erfc(x), x >= 0, by
// the Chebyshev fit of Press et al., Numerical Recipes, "erfcc":
//   t = 1 / (1 + x/2),  erfc(x) = t exp(-x^2 + P(t)),
// P of degree 9, fractional error below 1.2e-7 for every x, so Phi keeps
// float32 relative accuracy in both tails (for z < 0 it is erfc(|z|/sqrt 2)/2
// itself; for z > 0, 1 minus it).  About 20 VALU operations: the epilogues
// that evaluate it on every activation cost as much as the matrix products
// that produce them, so the evaluation count matters.
__device__ __forceinline__ float phi_cdf(float z) {
  const float x = fabsf(z) * 0.70710678118654752f;
  const float t = frcp(fmaf(0.5f, x, 1.0f));
  float p = 0.17087277f;
  p = fmaf(p, t, -0.82215223f);
  p = fmaf(p, t, 1.48851587f);
  p = fmaf(p, t, -1.13520398f);
  p = fmaf(p, t, 0.27886807f);
  p = fmaf(p, t, -0.18628806f);
  p = fmaf(p, t, 0.09678418f);
  p = fmaf(p, t, 0.37409196f);
  p = fmaf(p, t, 1.00002368f);
  p = fmaf(p, t, -1.26551223f);
  const float h = 0.5f * t * __expf(fmaf(-x, x, p));  // Phi(-|z|)
  return z >= 0.0f ? 1.0f - h : h;
}
__device__ __forceinline__ float gelu(float z) { return z * phi_cdf(z); }
__device__ __forceinline__ float gelu_d(float z) {
  return phi_cdf(z) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt_kernel(GemmArgs G) {
  // two stages of (A, B) tiles; the epilogue reuses the space for the
  // transposed output tile
  __shared__ __align__(16) bf16 smem[2 * kStage];
  static_assert(2 * kStage >= kTN * kTS, "transpose tile must fit");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // Workgroups are dispatched to the 8 XCDs round-robin by linear id, and
  // each XCD has its own L2.  Tiles are renumbered so that each XCD takes one
  // contiguous run of them: the tiles of one (member, split) product, which
  // share their A rows and B columns, then meet in the same L2 instead of
  // being fetched once per XCD.
  const int nbx = gridDim.x, nby = gridDim.y;
  const int total = nbx * nby * gridDim.z;
  const int lin = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
  const int tile = total % 8 == 0 ? (lin % 8) * (total / 8) + lin / 8 : lin;
  const int bx = tile % nbx, by = (tile / nbx) % nby, bz = tile / (nbx * nby);
  const int m0 = by * kTM, n0 = bx * kTN;
  const int z = bz / G.splits, split = bz % G.splits;
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
          G.Z[(int64_t)bz * G.sz + m * G.ldz + n] = v;
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

// ---------------------------------------------------------------------------
// Fused forward of the whole network for a tile of rows of one member
// (inference, and the forward of a training step).  The tile's activations
// stay in LDS from layer to layer (128 x 512 BF16 = 130 KB with padding, or
// 64 rows in half that); each wavefront computes 64 (128) output features of
// all 128 (64) rows (4 x 2 or 2 x 4 blocks of 32 x 32, float32 accumulators
// in registers) with the
// weights read from L2 in MFMA fragment order ("packed": the 16 bytes lane l
// of a 32x32x16 product needs are contiguous with its neighbours'), four
// k-steps ahead.  Per layer: products over K, barrier, epilogue (bias, GELU,
// BF16 into LDS over the consumed input; in training also the float32
// pre-activation Z and the transposed BF16 activation H^T the backward
// products read), barrier.  The output unit sums each row from LDS.
// Packed weights: Wp[z][nb][ks][lane][8], element e of lane l in block
// (nb, ks) = W[z][32 nb + (l & 31)][16 ks + 8 (l >> 5) + e].
// ---------------------------------------------------------------------------
constexpr int kFH = 512;         // hidden width
constexpr int kFX = 128;         // padded input width
constexpr int kFP = kFH + 8;     // LDS row pitch of a hidden layer (1040 B)
constexpr int kFPX = kFX + 8;    // LDS row pitch of the input tile (272 B)

struct FwdArgs {
  int n;                          // rows per member, a multiple of the tile rows
  const bf16* X; int64_t sx;      // [z][n][128]; sx = 0: one input for all members
  const bf16* Wp[3]; int64_t sw[3];
  const float* b[3];              // [z][512]
  const bf16* w4; const float* b4;  // [z][512], [z]
  float* y; int64_t sy;           // [z][n]
  float* Z[3];                    // training: [z][n][512] float32 pre-activations
  bf16* HT[3];                    // training: [z][512][n] BF16 activations
  const float* t; float* dy;      // training: targets [z][n], dy = 2 (y - t) / n
  float* part;                    // training: [z][tiles][5][512] per-tile partial sums
};

// Per-tile partial sums of the small gradients (fixed order; the tiles are
// summed in order by partials_reduce_kernel): slot 0..2 db1..db3 (column
// sums of dZ_l), 3 dw4 = H3^T dy, 4 db4 = sum dy (element 0).
constexpr int kParts = 5;

// Tile shape: RB row blocks and CB column blocks of 32 per wavefront, so the
// workgroup has 32 RB rows and 16 / CB wavefronts (64 features x 128 rows
// with RB = 4, CB = 2: one workgroup per CU; 128 features x 64 rows with
// RB = 2, CB = 4: two per CU, whose epilogues and products can overlap).
// Weights are prefetched PD = 8 / CB k-steps ahead.
template <int RB, int CB>
__device__ __forceinline__ void tile_products(f32x16 (&acc)[RB][CB], const bf16* sA, int pitch,
                                              int KS, const bf16* wp, int lane) {
  constexpr int PD = 8 / CB;
#pragma unroll
  for (int a = 0; a < RB; ++a)
#pragma unroll
    for (int b = 0; b < CB; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;
  // column block cb of the caller's wavefront is packed block (its first) + cb
  bf16x8 bq[PD][CB];
#pragma unroll
  for (int q = 0; q < PD; ++q)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
      bq[q][cb] = *reinterpret_cast<const bf16x8*>(wp + ((int64_t)cb * KS + q) * 512);
  const bf16* arow = sA + (lane & 31) * pitch + 8 * (lane >> 5);
  for (int ks0 = 0; ks0 < KS; ks0 += PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      const int ks = ks0 + q;
      bf16x8 b[CB];
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) b[cb] = bq[q][cb];
      if (ks + PD < KS) {  // PD k-steps ahead
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
          bq[q][cb] = *reinterpret_cast<const bf16x8*>(wp + ((int64_t)cb * KS + ks + PD) * 512);
      }
      bf16x8 a[RB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
        a[rb] = *reinterpret_cast<const bf16x8*>(arow + rb * 32 * pitch + ks * 16);
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
          acc[rb][cb] =
              __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[rb], b[cb], acc[rb][cb], 0, 0, 0);
    }
  }
}

// Transposed store of an LDS tile [ROWS][512] (pitch kFP) to T[j][m0 + i]
// (row length n), with the column sums over the tile's rows into psum[j]
// (psum may be null).  Chunk c: feature j = c / (ROWS / 8), rows
// 8 (c % (ROWS / 8)) .. + 7; the ROWS / 8 chunks of one feature sit on
// consecutive lanes and are summed by a butterfly in fixed order.
template <int ROWS, int NT>
__device__ __forceinline__ void store_transposed(const bf16* sT, bf16* T, int64_t n,
                                                 float* psum, int tid) {
  constexpr int G = ROWS / 8;
#pragma unroll 1
  for (int q = 0; q < kFH * ROWS / 8 / NT; ++q) {
    const int c = tid + NT * q;
    const int j = c / G, i0 = (c % G) * 8;
    bf16x8 v;
    float sm = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      v[e] = sT[(i0 + e) * kFP + j];
      sm += (float)v[e];
    }
    *reinterpret_cast<bf16x8*>(&T[(int64_t)j * n + i0]) = v;
    if (psum) {
#pragma unroll
      for (int o = 1; o < G; o <<= 1) sm += __shfl_xor(sm, o);
      if ((c % G) == 0) psum[j] = sm;
    }
  }
}

template <bool TRAIN, int RB, int CB>
__global__ __launch_bounds__(1024 / CB, 2) void mlp_fwd_kernel(FwdArgs F) {
  constexpr int ROWS = 32 * RB, NT = 1024 / CB;
  __shared__ __align__(16) bf16 sH[ROWS * kFP];
  __shared__ float sDy[TRAIN ? ROWS : 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int z = blockIdx.y;
  const int64_t m0 = (int64_t)blockIdx.x * ROWS;
  {
    const bf16* X = F.X + z * F.sx + m0 * kFX;
#pragma unroll
    for (int q = 0; q < ROWS * kFX / 8 / NT; ++q) {
      const int c = tid + NT * q;
      const int r = c >> 4, k = (c & 15) * 8;
      *reinterpret_cast<uint4*>(&sH[r * kFPX + k]) =
          *reinterpret_cast<const uint4*>(&X[(int64_t)r * kFX + k]);
    }
  }
  __syncthreads();
#pragma unroll 1
  for (int l = 0; l < 3; ++l) {
    const int KS = (l == 0 ? kFX : kFH) / 16;
    const int pitch = l == 0 ? kFPX : kFP;
    f32x16 acc[RB][CB];
    tile_products<RB, CB>(acc, sH, pitch, KS,
                          F.Wp[l] + z * F.sw[l] + (int64_t)(CB * wave) * KS * 512 + lane * 8,
                          lane);
    __syncthreads();  // every wavefront is done reading the layer input
    const float* bias = F.b[l] + z * kFH;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const int j = wave * 32 * CB + cb * 32 + (lane & 31);
        const float bj = bias[j];
        float* zrow = nullptr;
        if (TRAIN) {
          zrow = F.Z[l] + ((int64_t)z * F.n + m0 + rb * 32 + 4 * (lane >> 5)) * kFH + j;
          // the block's store addresses are formed from here on, not hoisted
          // out of the layer loop (where they would be held for all blocks)
          asm volatile("" : "+v"(zrow));
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          const float v = acc[rb][cb][r] + bj;
          // training keeps gelu'(Z), the only use the backward makes of Z
          if (TRAIN) zrow[(8 * (r >> 2) + (r & 3)) * kFH] = gelu_d(v);
          sH[i * kFP + j] = (bf16)gelu(v);
        }
      }
    __syncthreads();
    if (TRAIN)
      store_transposed<ROWS, NT>(sH, F.HT[l] + (int64_t)z * kFH * F.n + m0, F.n, nullptr,
                                 tid);
  }
  // output unit: one wavefront per row at a time, 8 features per lane and a
  // butterfly -- the summation order of dn_mlp_out, so y and dy are the
  // layer products' bit for bit
  {
    const bf16x8 wv = *reinterpret_cast<const bf16x8*>(F.w4 + z * kFH + lane * 8);
    for (int i = wave; i < ROWS; i += NT / 64) {
      const bf16x8 hv = *reinterpret_cast<const bf16x8*>(sH + i * kFP + lane * 8);
      float s = 0.0f;
#pragma unroll
      for (int e = 0; e < 8; ++e) s += (float)hv[e] * (float)wv[e];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) {
        const float yv = s + F.b4[z];
        F.y[z * F.sy + m0 + i] = yv;
        if (TRAIN) {
          const float d = 2.0f * (yv - F.t[z * F.sy + m0 + i]) / (float)F.n;
          F.dy[z * F.sy + m0 + i] = d;
          sDy[i] = d;
        }
      }
    }
  }
  if (TRAIN) {
    // partial dw4 = H3^T dy and db4 = sum dy over the tile's rows
    __syncthreads();
    float* part = F.part + ((int64_t)z * (F.n / ROWS) + blockIdx.x) * kParts * kFH;
    for (int j = tid; j < kFH; j += NT) {
      float s = 0.0f;
      for (int i = 0; i < ROWS; ++i) s += (float)sH[i * kFP + j] * sDy[i];
      part[3 * kFH + j] = s;
    }
    if (tid == 0) {
      float s = 0.0f;
      for (int i = 0; i < ROWS; ++i) s += sDy[i];
      part[4 * kFH] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Fused backward data products for a tile of rows of one member:
//   dZ3 = dy w4^T * gelu'(Z3),  dZ2 = (dZ3 W3) * gelu'(Z2),
//   dZ1 = (dZ2 W2) * gelu'(Z1)
// (BF16 results, as the layer products round them), each dZ held in LDS as
// the next product's A operand, W_l^T read packed from L2; stores dZ_l^T
// [z][512][n] for the weight-gradient products and the tile's column sums
// of dZ_l (partial bias gradients).
// ---------------------------------------------------------------------------
struct BwdArgs {
  int n;
  const float* dy; int64_t sy;    // [z][n]
  const bf16* w4;                 // [z][512]
  const float* Z[3];              // [z][n][512] gelu'(Z_l) from the fused forward
  const bf16* Wtp[2]; int64_t sw; // packed W2^T, W3^T, member stride
  bf16* dZT[3];                   // [z][512][n]
  float* part;                    // [z][tiles][5][512]
};

template <int RB, int CB>
__global__ __launch_bounds__(1024 / CB, 2) void mlp_bwd_kernel(BwdArgs A) {
  constexpr int ROWS = 32 * RB, NT = 1024 / CB;
  __shared__ __align__(16) bf16 sD[ROWS * kFP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int z = blockIdx.y;
  const int64_t m0 = (int64_t)blockIdx.x * ROWS;
  float* part = A.part + ((int64_t)z * (A.n / ROWS) + blockIdx.x) * kParts * kFH;
  // dZ3: four consecutive features per thread and chunk
  {
    const float* Z3 = A.Z[2] + ((int64_t)z * A.n + m0) * kFH;
    const bf16* w4 = A.w4 + z * kFH;
#pragma unroll 4
    for (int q = 0; q < ROWS * kFH / 4 / NT; ++q) {
      const int c = tid + NT * q;
      const int i = c >> 7, j = (c & 127) * 4;
      const float4 zv = *reinterpret_cast<const float4*>(&Z3[(int64_t)i * kFH + j]);
      const float d = A.dy[z * A.sy + m0 + i];
      bf16* o = &sD[i * kFP + j];
      o[0] = (bf16)(d * (float)w4[j] * zv.x);
      o[1] = (bf16)(d * (float)w4[j + 1] * zv.y);
      o[2] = (bf16)(d * (float)w4[j + 2] * zv.z);
      o[3] = (bf16)(d * (float)w4[j + 3] * zv.w);
    }
  }
  __syncthreads();
  store_transposed<ROWS, NT>(sD, A.dZT[2] + (int64_t)z * kFH * A.n + m0, A.n, part + 2 * kFH,
                             tid);
#pragma unroll 1
  for (int l = 2; l >= 1; --l) {  // dZ_l (index l-1) from dZ_{l+1} and W_{l+1}
    f32x16 acc[RB][CB];
    tile_products<RB, CB>(acc, sD, kFP, kFH / 16,
                          A.Wtp[l - 1] + z * A.sw + (int64_t)(CB * wave) * (kFH / 16) * 512 +
                              lane * 8,
                          lane);
    __syncthreads();  // every wavefront is done reading dZ_{l+1}
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const int j = wave * 32 * CB + cb * 32 + (lane & 31);
        const float* zrow =
            A.Z[l - 1] + ((int64_t)z * A.n + m0 + rb * 32 + 4 * (lane >> 5)) * kFH + j;
        asm volatile("" : "+v"(zrow));
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int i = rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          sD[i * kFP + j] = (bf16)(acc[rb][cb][r] * zrow[(8 * (r >> 2) + (r & 3)) * kFH]);
        }
        // one block's Z loads in flight at a time (register pressure)
        __builtin_amdgcn_sched_barrier(0);
      }
    __syncthreads();
    // (the next product only reads sD; its epilogue writes after a barrier)
    store_transposed<ROWS, NT>(sD, A.dZT[l - 1] + (int64_t)z * kFH * A.n + m0, A.n,
                               part + (l - 1) * kFH, tid);
  }
}

// grad[z][s][j] = sum over tiles of part[z][tile][s][j], tiles in order; the
// five sums go to b1, b2, b3, w4 (each [z][512]) and b4 ([z], element 0).
__global__ __launch_bounds__(256) void partials_reduce_kernel(int tiles, const float* part,
                                                              float* gb1, float* gb2,
                                                              float* gb3, float* gw4,
                                                              float* gb4) {
  const int z = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // slot * 512 + j
  if (e >= kParts * kFH) return;
  const int slot = e / kFH, j = e % kFH;
  if (slot == 4 && j > 0) return;
  const float* p = part + (int64_t)z * tiles * kParts * kFH + e;
  float s = 0.0f;
  // loads in batches of 16 (independent), sums in tile order
  int t = 0;
  for (; t + 16 <= tiles; t += 16) {
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = p[(int64_t)(t + u) * kParts * kFH];
#pragma unroll
    for (int u = 0; u < 16; ++u) s += v[u];
  }
  for (; t < tiles; ++t) s += p[(int64_t)t * kParts * kFH];
  float* out = slot == 0 ? gb1 : slot == 1 ? gb2 : slot == 2 ? gb3 : slot == 3 ? gw4 : gb4;
  out[z * (slot == 4 ? 1 : kFH) + j] = s;
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

// Adam on float32 masters [z][rows][cols]; BF16 copies [z][rows][cols],
// transposed [z][cols][rows], and the packed MFMA fragment orders of W and
// W^T (any may be null).
__device__ __forceinline__ void adam_elem(int64_t t, int rows, int cols, float* p,
                                          const float* g, float* m, float* v, float lr_t,
                                          float b1, float b2, float eps, bf16* pb, bf16* pbt,
                                          bf16* pbp, bf16* pbtp) {
  const float gi = g[t];
  const float mi = b1 * m[t] + (1.0f - b1) * gi;
  const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  const float pi = p[t] - lr_t * mi / (sqrtf(vi) + eps);
  p[t] = pi;
  // 32-bit index arithmetic (the entry points reject n >= 2^31)
  const uint32_t per = (uint32_t)rows * (uint32_t)cols, tu = (uint32_t)t;
  const uint32_t z = tu / per, e = tu - z * per, r = e / (uint32_t)cols,
                 c = e - r * (uint32_t)cols;
  if (pb) pb[t] = (bf16)pi;
  if (pbt) pbt[z * per + c * rows + r] = (bf16)pi;
  if (pbp) {  // MFMA fragment order of the fused forward (rows % 32 == 0, cols % 16 == 0)
    const uint32_t nb = r >> 5, ks = c >> 4, ln = ((c >> 3) & 1) * 32 + (r & 31);
    pbp[z * per + ((nb * (cols >> 4) + ks) * 64 + ln) * 8 + (c & 7)] = (bf16)pi;
  }
  if (pbtp) {  // the same order for W^T [cols][rows] (cols % 32 == 0, rows % 16 == 0)
    const uint32_t nb = c >> 5, ks = r >> 4, ln = ((r >> 3) & 1) * 32 + (c & 31);
    pbtp[z * per + ((nb * (rows >> 4) + ks) * 64 + ln) * 8 + (r & 7)] = (bf16)pi;
  }
}

__global__ __launch_bounds__(256) void adam_kernel(int64_t n, int rows, int cols, float* p,
                                                   const float* g, float* m, float* v,
                                                   float lr_t, float b1, float b2, float eps,
                                                   bf16* pb, bf16* pbt, bf16* pbp,
                                                   bf16* pbtp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  adam_elem(t, rows, cols, p, g, m, v, lr_t, b1, b2, eps, pb, pbt, pbp, pbtp);
}

// Every parameter tensor of a step in one launch: tensor k = blockIdx.y.
constexpr int kAdamMax = 8;
struct AdamTable {
  int64_t n[kAdamMax];
  int rows[kAdamMax], cols[kAdamMax];
  float* p[kAdamMax];
  const float* g[kAdamMax];
  float* m[kAdamMax];
  float* v[kAdamMax];
  bf16* pb[kAdamMax];
  bf16* pbt[kAdamMax];
  bf16* pbp[kAdamMax];
  bf16* pbtp[kAdamMax];
};

__global__ __launch_bounds__(256) void adam_multi_kernel(AdamTable T, float lr_t, float b1,
                                                         float b2, float eps) {
  const int k = blockIdx.y;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T.n[k]) return;
  adam_elem(t, T.rows[k], T.cols[k], T.p[k], T.g[k], T.m[k], T.v[k], lr_t, b1, b2, eps,
            T.pb[k], T.pbt[k], T.pbp[k], T.pbtp[k]);
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
  const int FC = F / 8;  // 16-byte chunks per row
  for (int e = threadIdx.x; e < 64 * FC; e += 256) {
    const int bi = e / FC, f = (e % FC) * 8;
    const bf16x8 v = *reinterpret_cast<const bf16x8*>(&X[(int64_t)id[b0 + bi] * F + f]);
    *reinterpret_cast<bf16x8*>(&X0[((int64_t)z * Bn + b0 + bi) * F + f]) = v;
#pragma unroll
    for (int u = 0; u < 8; ++u) tile[(f + u) * 72 + bi] = v[u];
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

// Fused forward (see mlp_fwd_kernel): tile = 128 or 64 rows per workgroup
// (training: 64, the tile of dn_mlp_backward), n rows per member (a multiple
// of tile), 128 input features, three hidden layers of 512, packed weights
// Wp1..Wp3 with member strides sw1, sw23 (elements).  Inference: Z, HT, t,
// dy, part null.  Training: Z[l] [z][n][512] float32, HT[l] [z][512][n]
// BF16, t and dy [z][n] (stride sy), part [z][n/64][5][512] (slots 3, 4:
// the tile's dw4 and db4 sums), all required.
DN_EXPORT int dn_mlp_forward(int nz, int n, int tile, const void* X, int64_t sx,
                             const void* Wp1, const void* Wp2, const void* Wp3, int64_t sw1,
                             int64_t sw23, const float* b1, const float* b2, const float* b3,
                             const void* w4, const float* b4, float* y, int64_t sy,
                             float* Z1, float* Z2, float* Z3, void* HT1, void* HT2,
                             void* HT3, const float* t, float* dy, float* part,
                             hipStream_t stream) {
  if (nz < 1 || (tile != 64 && tile != 128) || n < tile || n % tile || !X || !Wp1 || !Wp2 ||
      !Wp3 || !b1 || !b2 || !b3 || !w4 || !b4 || !y || sy < n)
    return DN_EINVAL;
  const bool train = Z1 || Z2 || Z3 || HT1 || HT2 || HT3 || t || dy || part;
  if (train && !(Z1 && Z2 && Z3 && HT1 && HT2 && HT3 && t && dy && part && tile == 64))
    return DN_EINVAL;
  if (nz > 65535 || n / tile > 0x7fffffff) return DN_ERANGE;
  dn::FwdArgs F = {};
  F.n = n; F.X = static_cast<const bf16*>(X); F.sx = sx;
  F.Wp[0] = static_cast<const bf16*>(Wp1); F.Wp[1] = static_cast<const bf16*>(Wp2);
  F.Wp[2] = static_cast<const bf16*>(Wp3);
  F.sw[0] = sw1; F.sw[1] = sw23; F.sw[2] = sw23;
  F.b[0] = b1; F.b[1] = b2; F.b[2] = b3;
  F.w4 = static_cast<const bf16*>(w4); F.b4 = b4; F.y = y; F.sy = sy;
  F.Z[0] = Z1; F.Z[1] = Z2; F.Z[2] = Z3;
  F.HT[0] = static_cast<bf16*>(HT1); F.HT[1] = static_cast<bf16*>(HT2);
  F.HT[2] = static_cast<bf16*>(HT3);
  F.t = t; F.dy = dy; F.part = part;
  const dim3 grid(n / tile, nz);
  if (tile == 128) {
    hipLaunchKernelGGL((dn::mlp_fwd_kernel<false, 4, 2>), grid, dim3(512), 0, stream, F);
  } else {
    if (train) hipLaunchKernelGGL((dn::mlp_fwd_kernel<true, 2, 4>), grid, dim3(256), 0, stream, F);
    else hipLaunchKernelGGL((dn::mlp_fwd_kernel<false, 2, 4>), grid, dim3(256), 0, stream, F);
  }
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// Fused backward data products (see mlp_bwd_kernel), tile = 64 rows:
// dy [z][n] (stride sy), w4 [z][512] BF16, Z1..Z3 [z][n][512], packed W2^T
// and W3^T (member stride sw), dZ1T..dZ3T [z][512][n] BF16, part the
// forward's partial buffer [z][n/tile][5][512] (slots 0..2 written here).
DN_EXPORT int dn_mlp_backward(int nz, int n, int tile, const float* dy, int64_t sy,
                              const void* w4, const float* Z1, const float* Z2,
                              const float* Z3, const void* Wt2p, const void* Wt3p, int64_t sw,
                              void* dZ1T, void* dZ2T, void* dZ3T, float* part,
                              hipStream_t stream) {
  if (nz < 1 || tile != 64 || n < tile || n % tile || !dy || sy < n || !w4 || !Z1 || !Z2 ||
      !Z3 || !Wt2p || !Wt3p || !dZ1T || !dZ2T || !dZ3T || !part)
    return DN_EINVAL;
  if (nz > 65535) return DN_ERANGE;
  dn::BwdArgs A = {};
  A.n = n; A.dy = dy; A.sy = sy; A.w4 = static_cast<const bf16*>(w4);
  A.Z[0] = Z1; A.Z[1] = Z2; A.Z[2] = Z3;
  A.Wtp[0] = static_cast<const bf16*>(Wt2p); A.Wtp[1] = static_cast<const bf16*>(Wt3p);
  A.sw = sw;
  A.dZT[0] = static_cast<bf16*>(dZ1T); A.dZT[1] = static_cast<bf16*>(dZ2T);
  A.dZT[2] = static_cast<bf16*>(dZ3T);
  A.part = part;
  hipLaunchKernelGGL((dn::mlp_bwd_kernel<2, 4>), dim3(n / tile, nz), dim3(256), 0, stream, A);
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// Sums the per-tile partials [z][tiles][5][512] in tile order into the
// gradients of b1, b2, b3, w4 ([z][512] each) and b4 ([z]).
DN_EXPORT int dn_mlp_partials_reduce(int nz, int tiles, const float* part, float* gb1,
                                     float* gb2, float* gb3, float* gw4, float* gb4,
                                     hipStream_t stream) {
  if (nz < 1 || tiles < 1 || !part || !gb1 || !gb2 || !gb3 || !gw4 || !gb4) return DN_EINVAL;
  if (nz > 65535) return DN_ERANGE;
  hipLaunchKernelGGL(dn::partials_reduce_kernel, dim3((dn::kParts * dn::kFH + 255) / 256, nz),
                     dim3(256), 0, stream, tiles, part, gb1, gb2, gb3, gw4, gb4);
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
                          void* pbt, void* pbp, void* pbtp, hipStream_t stream) {
  if (n < 1 || rows < 1 || cols < 1 || n % ((int64_t)rows * cols) || !p || !g || !m || !v ||
      (pbp && (rows % 32 || cols % 16)) || (pbtp && (cols % 32 || rows % 16)))
    return DN_EINVAL;
  if (n > INT32_MAX) return DN_ERANGE;  // 32-bit element indices
  hipLaunchKernelGGL(dn::adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                     n, rows, cols, p, g, m, v, lr_t, b1, b2, eps, static_cast<bf16*>(pb),
                     static_cast<bf16*>(pbt), static_cast<bf16*>(pbp), static_cast<bf16*>(pbtp));
  return hipGetLastError() == hipSuccess ? DN_OK : DN_ELAUNCH;
}

// Adam over `count` (<= 8) tensors in one launch; per tensor k the
// arguments of dn_mlp_adam, passed as host arrays (n[k], rows[k], cols[k],
// and the device addresses p[k], g[k], m[k], v[k], pb[k], pbt[k], pbp[k],
// pbtp[k]; the copies may be 0).
DN_EXPORT int dn_mlp_adam_multi(int count, const int64_t* n, const int* rows, const int* cols,
                                const uint64_t* p, const uint64_t* g, const uint64_t* m,
                                const uint64_t* v, const uint64_t* pb, const uint64_t* pbt,
                                const uint64_t* pbp, const uint64_t* pbtp, float lr_t, float b1,
                                float b2, float eps, hipStream_t stream) {
  if (count < 1 || count > dn::kAdamMax || !n || !rows || !cols || !p || !g || !m || !v ||
      !pb || !pbt || !pbp || !pbtp)
    return DN_EINVAL;
  dn::AdamTable T = {};
  int64_t nmax = 0;
  for (int k = 0; k < count; ++k) {
    if (n[k] < 1 || rows[k] < 1 || cols[k] < 1 || n[k] % ((int64_t)rows[k] * cols[k]) ||
        !p[k] || !g[k] || !m[k] || !v[k] || (pbp[k] && (rows[k] % 32 || cols[k] % 16)) ||
        (pbtp[k] && (cols[k] % 32 || rows[k] % 16)))
      return DN_EINVAL;
    if (n[k] > INT32_MAX) return DN_ERANGE;  // 32-bit element indices
    T.n[k] = n[k]; T.rows[k] = rows[k]; T.cols[k] = cols[k];
    T.p[k] = reinterpret_cast<float*>(p[k]); T.g[k] = reinterpret_cast<const float*>(g[k]);
    T.m[k] = reinterpret_cast<float*>(m[k]); T.v[k] = reinterpret_cast<float*>(v[k]);
    T.pb[k] = reinterpret_cast<bf16*>(pb[k]); T.pbt[k] = reinterpret_cast<bf16*>(pbt[k]);
    T.pbp[k] = reinterpret_cast<bf16*>(pbp[k]); T.pbtp[k] = reinterpret_cast<bf16*>(pbtp[k]);
    nmax = n[k] > nmax ? n[k] : nmax;
  }
  const int64_t blocks = (nmax + 255) / 256;
  if (blocks > 0x7fffffff) return DN_ERANGE;
  hipLaunchKernelGGL(dn::adam_multi_kernel, dim3((unsigned)blocks, count), dim3(256), 0, stream,
                     T, lr_t, b1, b2, eps);
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
