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

// exact (erf) GELU and its derivative
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
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
// bf16 copies of the weights (w16_map[i] >= 0: index into w16).
__global__ void k_adam(float *p, const float *grad, float *m1, float *m2, int64_t n, const int64_t *step,
                       float lr, float b1, float b2, float eps, float wd, const int64_t *w16_map,
                       bf16 *w16) {
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
