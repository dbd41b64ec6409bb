// Device kernels of the rate surrogate: a small MLP (GELU hidden layers)
// trained and evaluated with BF16 operands on the CDNA4 matrix cores and FP32
// accumulation.
//
// GEMM: C[m][n] = sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak] and
// B(k,n) = B[k*sbk + n*sbn] in bf16, so one kernel serves the forward pass
// (X W^T), the input gradient (dZ W) and the weight gradient (dZ^T H) by
// strides alone.  A 256-thread block computes a 64x64 tile of C: the A and B
// tiles (64x32 each) are staged through LDS as [row][k] (B transposed on the
// way in), so a lane reads its 8 consecutive k values with one 16-byte LDS
// read whichever operand layout the caller has; the four waves (2x2) each own
// a 32x32 sub-tile and issue two v_mfma_f32_32x32x16_bf16 per 32-wide k step.
// Operand lane layout of that instruction (checked on the hardware): lane l
// holds A[l%32][8*(l/32) + q] and B[8*(l/32) + q][l%32], q = 0..7; accumulator
// register r of lane l is C[8*(r/4) + 4*(l/32) + r%4][l%32].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsm {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum Epi : int32_t {
  EPI_F32 = 0,       // out_f32[m][n] = acc                        (weight gradient)
  EPI_BIAS_GELU = 1, // z = acc + bias[n]; out_f32 = z (if given); out_bf16 = gelu(z)
  EPI_BIAS = 2,      // out_f32[m][n] = acc + bias[n]              (output layer)
  EPI_GELU_BWD = 3,  // g = acc * gelu'(zprev[m][n]); out_bf16 = g; out_f32 = g (if given)
};

struct GemmArgs {
  const bf16 *A;
  int64_t sam, sak;
  const bf16 *B;
  int64_t sbk, sbn;
  int32_t M, N, K;
  int32_t epi;
  int32_t n_valid;         // columns stored (n < n_valid)
  int32_t vec_a, vec_b;    // 1: 16-byte loads along the contiguous dimension are legal
  const float *bias;
  float *out_f32;
  int64_t ld_f32;
  bf16 *out_bf16;
  int64_t ld_bf16;
  const float *zprev;
  int64_t ld_z;
};

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

// Fused forward of a whole network whose padded widths are all <= 256
// (k_mlp_fused).  Each wave keeps 32 rows in registers through every layer:
// the product is taken transposed, Z^T = W H^T, with the weights as the
// MFMA's A operand and the activations as its B operand, so lane l always
// holds data row l % 32 -- the accumulator layout of one layer is, up to a
// fixed permutation of the K index inside each 16-wide k step, the B-operand
// layout of the next.  Output unit n = 32 nt + 8 (r/4) + 4 (l/32) + r%4 of
// accumulator register r of n-tile nt becomes element q = r % 8 of k step
// 2 nt + r / 8 of the next layer, i.e. MFMA k position 8 (l/32) + q of that
// step holds unit 8 (q/4) + 4 (l/32) + q % 4 of it.  The weight copy the
// kernel reads (fused layout) stores each layer's K axis in that order
// (layers after the first), cut into slices of 64 k, each slice [n][64] with
// the eight 16-byte chunks of row n at chunk position c ^ (n % 8), so the 32
// rows one A-operand read touches spread over the LDS banks.
constexpr int FUSED_MAXW = 256, FUSED_MAXL = 8, FUSED_ROWS = 256, FUSED_THREADS = 512;

// MFMA k position of input unit k of a layer after the first: swap the
// second and third group of four within each 16 (an involution)
__host__ __device__ constexpr int32_t fused_kpos(int32_t k) {
  const int32_t p = k & 15, g = p >> 2;
  return (k & ~15) | ((g == 1 ? 2 : g == 2 ? 1 : g) << 2) | (p & 3);
}

// element offset of weight (o, k) of layer i within that layer's fused-layout
// block; n_out_pad = padded output width
__host__ __device__ constexpr int64_t fused_woff(int32_t layer, int32_t o, int32_t k, int32_t n_out_pad) {
  const int32_t P = layer == 0 ? k : fused_kpos(k);
  const int32_t sl = P >> 6, pl = P & 63, c = pl >> 3, e = pl & 7;
  return (int64_t)sl * n_out_pad * 64 + (int64_t)o * 64 + ((c ^ (o & 7)) << 3) + e;
}

struct FusedArgs {
  const float *x;          // [n][in]
  float *y;                // [n][out]
  int64_t n;
  int32_t L;
  int32_t sz[FUSED_MAXL + 1];   // widths
  int32_t sp[FUSED_MAXL + 1];   // padded widths (multiples of 64, <= 256)
  const bf16 *w[FUSED_MAXL];    // fused-layout weights of each layer
  const float *b[FUSED_MAXL];   // biases (f32, sz[i+1] each)
};

// erf by Abramowitz & Stegun 7.1.26 (Hastings' rational approximation):
// erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2),
// t = 1 / (1 + p x), x >= 0, odd extension; absolute error <= 1.5e-7, far
// below the bf16 rounding of the activations it feeds (2^-9 relative).
// Branch-free: one reciprocal, one exp2, five fused multiply-adds.
__device__ __forceinline__ float erf_as(float x) {
  const float ax = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.0f));
  float pl = fmaf(1.061405429f, t, -1.453152027f);
  pl = fmaf(pl, t, 1.421413741f);
  pl = fmaf(pl, t, -0.284496736f);
  pl = fmaf(pl, t, 0.254829592f);
  pl *= t;
  const float e = __builtin_amdgcn_exp2f(-1.4426950408889634f * ax * ax);
  return copysignf(fmaf(-pl, e, 1.0f), x);
}

// GELU (erf form) and its derivative.  gelu(z) = z (1 + erf(z / sqrt 2)) / 2
// with erf_as written out: for x = |z| / sqrt 2, 1 - erf(x) = t P(t) e^(-x^2),
// so gelu(z) = max(z, 0) - |z| t P(t) e^(-z^2 / 2) / 2 for either sign of z
// (the odd extension); the 1/2 and the 1/sqrt 2 are folded into the
// constants: the same approximation in fewer operations.
__device__ __forceinline__ float gelu(float z) {
  const float az = fabsf(z);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.23164189f, az, 1.0f));  // 0.3275911 / sqrt 2
  float pl = fmaf(0.5307027145f, t, -0.7265760135f);                   // A&S a5 / 2, a4 / 2
  pl = fmaf(pl, t, 0.7107068705f);                                     // a3 / 2
  pl = fmaf(pl, t, -0.142248368f);                                     // a2 / 2
  pl = fmaf(pl, t, 0.127414796f);                                      // a1 / 2
  const float e = __builtin_amdgcn_exp2f(-0.72134752044448170f * z * z);  // e^(-z^2 / 2)
  return fmaf(-(az * t) * pl, e, fmaxf(z, 0.0f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erf_as(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * __builtin_amdgcn_exp2f(-0.72134752044448170f * z * z);
}

// Stage a 64 (rows) x 32 (k) tile of X(r, k) = X[r*sr + k*sk] into LDS as
// S[r][k]; rows >= R or k >= K are zero.  vec: the contiguous dimension
// (sk == 1 or sr == 1) allows aligned 8-element loads.
__device__ __forceinline__ void stage_tile(bf16 (*S)[LDK], const bf16 *X, int64_t sr, int64_t sk,
                                           int32_t r0, int32_t k0, int32_t R, int32_t K,
                                           int32_t vec) {
  const int t = threadIdx.x;
  const bf16 zero = (bf16)0.0f;
  if (sk == 1) {  // k-contiguous: thread -> (row t/4, 8 k values)
    const int r = t >> 2, kc = (t & 3) * 8;
    const int32_t gr = r0 + r, gk = k0 + kc;
    if (vec && gr < R && gk + 8 <= K) {
      const bf16x8 v = *reinterpret_cast<const bf16x8 *>(X + (int64_t)gr * sr + gk);
#pragma unroll
      for (int i = 0; i < 8; ++i) S[r][kc + i] = v[i];
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i)
        S[r][kc + i] = (gr < R && gk + i < K) ? X[(int64_t)gr * sr + gk + i] : zero;
    }
  } else {  // row-contiguous (sr == 1) or general: thread -> (k t/8, 8 rows)
    const int kk = t >> 3, rc = (t & 7) * 8;
    const int32_t gk = k0 + kk, gr = r0 + rc;
    if (vec && sr == 1 && gk < K && gr + 8 <= R) {
      const bf16x8 v = *reinterpret_cast<const bf16x8 *>(X + (int64_t)gk * sk + gr);
#pragma unroll
      for (int i = 0; i < 8; ++i) S[rc + i][kk] = v[i];
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i)
        S[rc + i][kk] = (gk < K && gr + i < R) ? X[(int64_t)(gr + i) * sr + (int64_t)gk * sk] : zero;
    }
  }
}

__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ bf16 As[BM][LDK];
  __shared__ bf16 Bs[BN][LDK];
  const int32_t m0 = (int32_t)blockIdx.y * BM, n0 = (int32_t)blockIdx.x * BN;
  const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int wm = wave >> 1, wn = wave & 1;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
  for (int32_t k0 = 0; k0 < g.K; k0 += BK) {
    stage_tile(As, g.A, g.sam, g.sak, m0, k0, g.M, g.K, g.vec_a);
    // B(k, n): rows of the staged tile are n, so pass (sr, sk) = (sbn, sbk)
    stage_tile(Bs, g.B, g.sbn, g.sbk, n0, k0, g.N, g.K, g.vec_b);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm * 32 + (lane & 31)][ks + 8 * (lane >> 5)]);
      const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 32 + (lane & 31)][ks + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int32_t n = n0 + wn * 32 + (lane & 31);
  if (n >= g.N || n >= g.n_valid) return;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int32_t m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    if (m >= g.M) continue;
    const float v = acc[r];
    switch (g.epi) {
      case EPI_F32:
        g.out_f32[(int64_t)m * g.ld_f32 + n] = v;
        break;
      case EPI_BIAS_GELU: {
        const float z = v + g.bias[n];
        if (g.out_f32) g.out_f32[(int64_t)m * g.ld_f32 + n] = z;
        g.out_bf16[(int64_t)m * g.ld_bf16 + n] = (bf16)gelu(z);
        break;
      }
      case EPI_BIAS:
        g.out_f32[(int64_t)m * g.ld_f32 + n] = v + g.bias[n];
        break;
      default: {  // EPI_GELU_BWD
        const float d = v * gelu_grad(g.zprev[(int64_t)m * g.ld_z + n]);
        g.out_bf16[(int64_t)m * g.ld_bf16 + n] = (bf16)d;
        if (g.out_f32) g.out_f32[(int64_t)m * g.ld_f32 + n] = d;
        break;
      }
    }
  }
}

// The whole forward, one persistent block per compute unit: tiles of
// FUSED_ROWS rows (32 per wave); per layer the weight slices (64 k each)
// stream through two LDS buffers -- the loads of the next slice are in
// flight while the waves multiply the current one -- and the activations
// never leave the registers.  The sums of the layer-by-layer path: each
// output is bias + the sum over k in ascending 16-wide steps of bf16
// products (only the order of the terms inside a step differs, the matrix
// core's own), GELU then bf16 between layers.  Measured on 2^20 rows of the
// 54-256-256-256-50 net: about a third of the time each goes to the MFMAs,
// the GELU epilogue and the LDS traffic of the weight operands, which the
// barrier per slice keeps from overlapping.
__global__ void __launch_bounds__(FUSED_THREADS) k_mlp_fused(FusedArgs a) {
  __shared__ __align__(16) bf16 wsl[2][FUSED_MAXW * 64];
  __shared__ __align__(16) float sbias[FUSED_MAXL][FUSED_MAXW];  // zero past each layer's width
  const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int h = lane >> 5, lr = lane & 31;
  for (int i = 0; i < a.L; ++i)
    for (int n = (int)threadIdx.x; n < FUSED_MAXW; n += FUSED_THREADS)
      sbias[i][n] = n < a.sz[i + 1] ? a.b[i][n] : 0.0f;
  const int64_t tiles = (a.n + FUSED_ROWS - 1) / FUSED_ROWS;
  // slice g of the sequence (layer, slice) repeated per tile of this block;
  // layer i has sp[i] / 64 slices
  int per_tile = 0;
  for (int i = 0; i < a.L; ++i) per_tile += a.sp[i] >> 6;
  const int64_t my_tiles = (int64_t)blockIdx.x < tiles ? (tiles - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  const int64_t total = my_tiles * per_tile;
  // The copy of slice (layer, sl) goes through registers: loaded while the
  // waves multiply the current slice, written to the other LDS buffer after
  // it (a copy straight into LDS would make every LDS read of the current
  // slice wait for it).  Thread t holds 16-byte chunks t, t + 512, ... of
  // the slice (np * 8 chunks, np / 64 per thread).
  static_assert(FUSED_MAXW / 64 == 4, "four staged chunks per thread");
  uint4 st0 = {0u, 0u, 0u, 0u}, st1 = st0, st2 = st0, st3 = st0;
  int stg_n = 0;  // chunks per thread of the staged slice (wave-uniform)
#define FUSED_STAGE_LOAD(layer, sl)                                                           \
  do {                                                                                        \
    const int np_ = a.sp[(layer) + 1];                                                        \
    const uint4 *src_ = reinterpret_cast<const uint4 *>(a.w[layer] + (size_t)(sl) * np_ * 64) + \
                        threadIdx.x;                                                          \
    stg_n = np_ >> 6;                                                                         \
    st0 = src_[0];                                                                            \
    if (stg_n > 1) st1 = src_[FUSED_THREADS];                                                 \
    if (stg_n > 2) st2 = src_[2 * FUSED_THREADS];                                             \
    if (stg_n > 3) st3 = src_[3 * FUSED_THREADS];                                             \
  } while (0)
#define FUSED_STAGE_STORE(buf)                                                  \
  do {                                                                          \
    uint4 *dst_ = reinterpret_cast<uint4 *>(wsl[buf]) + threadIdx.x;            \
    dst_[0] = st0;                                                              \
    if (stg_n > 1) dst_[FUSED_THREADS] = st1;                                   \
    if (stg_n > 2) dst_[2 * FUSED_THREADS] = st2;                               \
    if (stg_n > 3) dst_[3 * FUSED_THREADS] = st3;                               \
  } while (0)
  int64_t g = 0;
  int cl = 0, cs = 0;  // layer and slice of sequence element g (per tile)
  if (total > 0) {
    FUSED_STAGE_LOAD(0, 0);
    FUSED_STAGE_STORE(0);
  }
  __syncthreads();
  for (int64_t t = 0; t < my_tiles; ++t) {
    const int64_t row = ((int64_t)blockIdx.x + t * gridDim.x) * FUSED_ROWS + wave * 32 + lr;
    const bool live = row < a.n;
    bf16x8 act[FUSED_MAXW / 16];
    // layer 0 input: k step s, lane half h holds x[row][16 s + 8 h + q];
    // four k steps at a time, every load issued before any is used (the
    // addresses are clamped into the array, the values selected after)
    // the input width, opaque to the optimiser per tile: the per-element
    // index clamps then stay here instead of being hoisted out of the tile
    // loop into registers
    int32_t szx = a.sz[0];
    __asm__ volatile("" : "+s"(szx));
    const float *xr = a.x + (live ? row : a.n - 1) * szx;
#pragma unroll
    for (int s4 = 0; s4 < FUSED_MAXW / 64; ++s4) {
      if (s4 < (a.sp[0] >> 6)) {
        float v[4][8];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int32_t k = 64 * s4 + 16 * j + 8 * h + q;
            v[j][q] = xr[k < szx ? k : szx - 1];
          }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int32_t k = 64 * s4 + 16 * j + 8 * h + q;
            act[4 * s4 + j][q] = (bf16)((live && k < szx) ? v[j][q] : 0.0f);
          }
      }
    }
    for (int layer = 0; layer < a.L; ++layer) {
      const int NT = a.sp[layer + 1] / 32;
      f32x16 acc[FUSED_MAXW / 32];
#pragma unroll
      for (int nt = 0; nt < FUSED_MAXW / 32; ++nt)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[nt][r] = 0.0f;
#pragma unroll
      for (int sl = 0; sl < FUSED_MAXW / 64; ++sl) {
        if (sl >= (a.sp[layer] >> 6)) break;
        // the next slice of the sequence, into registers for now
        const bool more = g + 1 < total;
        if (more) {
          int nl = cl, ns = cs + 1;
          if (ns >= (a.sp[nl] >> 6)) { ns = 0; nl = nl + 1 < a.L ? nl + 1 : 0; }
          FUSED_STAGE_LOAD(nl, ns);
        }
        const bf16 *ws = wsl[g & 1];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          const int c = 2 * ks + h;  // 16-byte chunk of the row
          // every n-tile, branch-free (the reads of a k step then issue
          // ahead of its MFMAs): tiles past the layer's width multiply
          // stale LDS rows into accumulators nothing reads
#pragma unroll
          for (int nt = 0; nt < FUSED_MAXW / 32; ++nt) {
            const int n = nt * 32 + lr;
            const bf16x8 wa = *reinterpret_cast<const bf16x8 *>(ws + n * 64 + ((c ^ (n & 7)) << 3));
            acc[nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wa, act[sl * 4 + ks], acc[nt], 0, 0, 0);
          }
        }
        if (more) FUSED_STAGE_STORE((int)((g + 1) & 1));
        __syncthreads();
        ++g;
        if (++cs >= (a.sp[cl] >> 6)) { cs = 0; cl = cl + 1 < a.L ? cl + 1 : 0; }
      }
      const float *bias = sbias[layer];
      const int32_t nv = a.sz[layer + 1];
      if (layer < a.L - 1) {
        // bias, GELU, bf16: unit 32 nt + 8 j + 4 h + i -> element r % 8 of k step 2 nt + r / 8
#pragma unroll
        for (int nt = 0; nt < FUSED_MAXW / 32; ++nt) {
          if (nt < NT) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int32_t n0 = nt * 32 + 8 * j + 4 * h;
              const float4 b4 = *reinterpret_cast<const float4 *>(bias + n0);
              const float bb[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                const int r = 4 * j + i;
                // a padded unit has zero weights and bias: gelu(0) = 0
                act[2 * nt + (r >> 3)][r & 7] = (bf16)gelu(acc[nt][r] + bb[i]);
              }
            }
          }
          // one n-tile at a time: keeps the scheduler from hoisting every
          // tile's bias reads (and their registers) to the top
          __builtin_amdgcn_sched_barrier(0);
        }
      } else if (live) {
        float *yr = a.y + row * a.sz[a.L];
#pragma unroll
        for (int nt = 0; nt < FUSED_MAXW / 32; ++nt) {
          if (nt < NT) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int32_t n0 = nt * 32 + 8 * j + 4 * h;
              const float4 b4 = *reinterpret_cast<const float4 *>(bias + n0);
              const float bb[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
              for (int i = 0; i < 4; ++i)
                if (n0 + i < nv) yr[n0 + i] = acc[nt][4 * j + i] + bb[i];
            }
          }
        }
      }
    }
  }
}

#undef FUSED_STAGE_LOAD
#undef FUSED_STAGE_STORE

// x f32 [*][in] -> bf16 [rows][ld] zero-padded; output row r is input row
// idx[*bsel * rows + r] when idx is given (minibatch *bsel of a permutation),
// else row r
__global__ void k_pack_rows(const float *x, const int64_t *idx, const int64_t *bsel, int64_t rows,
                            int32_t in, int32_t ld, bf16 *out) {
  const int64_t total = rows * ld;
  const int64_t base = (idx && bsel) ? *bsel * rows : 0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / ld;
    const int32_t c = (int32_t)(e - r * ld);
    const int64_t src = idx ? idx[base + r] : r;
    out[e] = c < in ? (bf16)x[src * in + c] : (bf16)0.0f;
  }
}

// rows idx[*bsel * rows + r] of a f32 [*][cols] matrix (minibatch targets)
__global__ void k_gather_rows_f32(const float *x, const int64_t *idx, const int64_t *bsel, int64_t rows,
                                  int32_t cols, float *out) {
  const int64_t total = rows * cols;
  const int64_t base = *bsel * rows;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / cols;
    out[e] = x[idx[base + r] * cols + (e - r * cols)];
  }
}

// Mean-squared-error over rows x n_out outputs: y f32 [rows][ld_y], t f32
// [rows][n_out].  dz = 2 (y - t) / (rows * n_out) in f32 and bf16 ([rows][ld],
// padded columns zero); the loss partial sums go to loss[0] (one atomic per block).
__global__ void __launch_bounds__(256) k_mse(const float *y, int64_t ld_y, const float *t, int32_t rows,
                                             int32_t n_out, int32_t ld, float *dz, bf16 *dz16,
                                             float *loss) {
  __shared__ float red[256];
  float part = 0.0f;
  const float scale = 2.0f / ((float)rows * (float)n_out);
  const int64_t total = (int64_t)rows * ld;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / ld;
    const int32_t c = (int32_t)(e - r * ld);
    float d = 0.0f;
    if (c < n_out) {
      const float diff = y[r * ld_y + c] - t[r * n_out + c];
      part += diff * diff;
      d = scale * diff;
    }
    dz[e] = d;
    dz16[e] = (bf16)d;
  }
  red[threadIdx.x] = part;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] / ((float)rows * (float)n_out));
}

// db[n] = sum over rows of dz[m][n] (f32), one block per column group
__global__ void __launch_bounds__(256) k_colsum(const float *dz, int32_t rows, int32_t cols, int64_t ld,
                                                float *db) {
  __shared__ float red[4][64];
  const int c = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63);
  const int q = (int)(threadIdx.x >> 6);
  float s = 0.0f;
  if (c < cols)
    for (int32_t m = q; m < rows; m += 4) s += dz[(int64_t)m * ld + c];
  red[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q == 0 && c < cols) db[c] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                                  red[3][threadIdx.x];
}

// Adam (Kingma & Ba 2015) with bias correction over the flat f32 master
// parameters; step = the 1-based step index held on the device; writes the
// bf16 copies of the weights (w16_map[i] >= 0: index into w16; wf_map, when
// given, likewise into the fused kernel's copy wf16).
__global__ void k_adam(float *p, const float *grad, float *m1, float *m2, int64_t n, const int64_t *step,
                       float lr, float b1, float b2, float eps, float wd, const int64_t *w16_map,
                       bf16 *w16, const int64_t *wf_map, bf16 *wf16) {
  const int64_t t = *step;
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float g = grad[i] + wd * p[i];
    const float a = b1 * m1[i] + (1.0f - b1) * g;
    const float b = b2 * m2[i] + (1.0f - b2) * g * g;
    m1[i] = a;
    m2[i] = b;
    const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
    const int64_t j = w16_map[i];
    if (j >= 0) w16[j] = (bf16)v;
    if (wf_map) {  // the fused kernel's copy
      const int64_t jf = wf_map[i];
      if (jf >= 0) wf16[jf] = (bf16)v;
    }
  }
}

__global__ void k_incr(int64_t *c) { *c += 1; }

// copy f32 [rows][ld_in] -> f32 [rows][cols] (unpadded outputs)
__global__ void k_unpad(const float *x, int64_t ld_in, int64_t rows, int32_t cols, float *y) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / cols;
    y[e] = x[r * ld_in + (e - r * cols)];
  }
}

}  // namespace fsm
