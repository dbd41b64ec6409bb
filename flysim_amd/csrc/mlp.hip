// Host side of the rate surrogate (include/flysim_mlp.h): parameter storage,
// forward / backward / Adam launch sequences and the C ABI.
//
// Flat parameter layout (host API, gradients and the CPU oracle): for each
// layer i, W_i[out][in] row-major then b_i[out], in f32.  The device keeps
// that f32 master copy for Adam plus bf16 copies of the weights padded to
// multiples of 64 in both dimensions ([out_pad][in_pad], zero padding), the
// operands of the matrix-core GEMMs.  Activations are bf16 [rows][width_pad];
// padding columns stay zero because no kernel ever writes them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/flysim_mlp.h"
#include "host/host_common.hpp"
#include "mlp_kernels.hpp"

namespace {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIPCHECK(x)                                                                    \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess)                                                              \
      throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));                  \
  } while (0)

struct Buf {
  void *p = nullptr;
  size_t bytes = 0;
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  ~Buf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // zero-filled allocation (padding columns rely on it)
  void alloc(size_t b) {
    release();
    if (b == 0) b = 256;
    HIPCHECK(hipMalloc(&p, b));
    HIPCHECK(hipMemset(p, 0, b));
    bytes = b;
  }
  void ensure(size_t b) {
    if (bytes < b) alloc(b);
  }
  template <class X> X *as() const { return static_cast<X *>(p); }
};

using fsm::bf16;

// Allocation-free forward (fs_mlp_predict_noalloc): the packed input rows and
// two ping-pong hidden activations of one chunk of rows live in this fixed
// device array.  Each GPU the library runs on holds its own copy until the
// process ends; host code uses only the per-device address resolved below,
// never the array's name.
constexpr size_t NOALLOC_BYTES = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];
// fs_mlp_create admits widths up to 8192, so one 64-row chunk always fits
static_assert(64 * 2 * (8192 + 2 * 8192) <= NOALLOC_BYTES, "scratch below one chunk of the widest net");

// Device address of `scratch` on the current device, resolved once per device
// (at fs_mlp_create, with that device current).
void *noalloc_scratch_current_device() {
  static std::mutex mu;
  static std::vector<void *> ptrs;
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((size_t)dev >= ptrs.size()) ptrs.resize((size_t)dev + 1, nullptr);
  if (!ptrs[dev]) {
    void *p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptrs[dev] = p;
  }
  return ptrs[dev];
}

inline int32_t pad64(int32_t x) { return (x + 63) / 64 * 64; }

struct Mlp {
  int32_t L = 0;                  // layers (weight matrices)
  std::vector<int32_t> s, sp;     // widths and padded widths, L+1 each
  std::vector<int64_t> offW, offb, offW16;
  int64_t n_params = 0, n_w16 = 0;
  Buf p32, grad, m1, m2, w16, w16map, ctr, loss;
  // whole-network forward in one kernel (every padded width <= 256): its
  // copy of the weights in the fused layout, kept beside w16 by set_params
  // and Adam
  bool fused = false;
  Buf wf16, wfmap;
  int32_t n_cu = 0;
  // training set and minibatch buffers
  Buf xall, yall, perm;
  int64_t n_data = 0;
  int32_t rows_cap = 0;           // rows of the activation buffers below
  Buf x16, tgt, yout;
  std::vector<std::unique_ptr<Buf>> z, h, dz32, dz16;  // per layer
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int32_t graph_rows = -1;
  fs_adam graph_opt{};
  hipEvent_t ev[2] = {nullptr, nullptr};
  void *noalloc = nullptr;        // this device's address of `scratch`

  Mlp(const int32_t *sizes, int32_t n_layers, uint64_t seed) {
    if (n_layers < 1 || n_layers > 16) throw std::invalid_argument("1..16 layers");
    L = n_layers;
    s.assign(sizes, sizes + L + 1);
    for (int32_t v : s)
      if (v < 1 || v > 8192) throw std::invalid_argument("layer width out of range 1..8192");
    sp.resize(L + 1);
    for (int i = 0; i <= L; ++i) sp[i] = pad64(s[i]);
    offW.resize(L);
    offb.resize(L);
    offW16.resize(L);
    for (int i = 0; i < L; ++i) {
      offW[i] = n_params;
      n_params += (int64_t)s[i + 1] * s[i];
      offb[i] = n_params;
      n_params += s[i + 1];
      offW16[i] = n_w16;
      n_w16 += (int64_t)sp[i + 1] * sp[i];
    }
    HIPCHECK(hipStreamCreate(&stream));
    for (auto &e : ev) HIPCHECK(hipEventCreate(&e));
    noalloc = noalloc_scratch_current_device();
    p32.alloc(n_params * 4);
    grad.alloc(n_params * 4);
    m1.alloc(n_params * 4);
    m2.alloc(n_params * 4);
    w16.alloc(n_w16 * 2);
    ctr.alloc(2 * 8);
    loss.alloc(4);
    std::vector<int64_t> map(n_params, -1);
    for (int i = 0; i < L; ++i)
      for (int32_t o = 0; o < s[i + 1]; ++o)
        for (int32_t k = 0; k < s[i]; ++k)
          map[offW[i] + (int64_t)o * s[i] + k] = offW16[i] + (int64_t)o * sp[i] + k;
    w16map.alloc(n_params * 8);
    HIPCHECK(hipMemcpy(w16map.p, map.data(), n_params * 8, hipMemcpyHostToDevice));
    fused = L <= fsm::FUSED_MAXL &&
            *std::max_element(sp.begin(), sp.end()) <= fsm::FUSED_MAXW;
    if (fused) {
      std::vector<int64_t> fmap(n_params, -1);
      for (int i = 0; i < L; ++i)
        for (int32_t o = 0; o < s[i + 1]; ++o)
          for (int32_t k = 0; k < s[i]; ++k)
            fmap[offW[i] + (int64_t)o * s[i] + k] = offW16[i] + fsm::fused_woff(i, o, k, sp[i + 1]);
      wfmap.alloc(n_params * 8);
      HIPCHECK(hipMemcpy(wfmap.p, fmap.data(), n_params * 8, hipMemcpyHostToDevice));
      wf16.alloc(n_w16 * 2);
      int dev = 0;
      HIPCHECK(hipGetDevice(&dev));
      HIPCHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    // He-uniform initialisation of the weights (bound sqrt(6/fan_in)), zero biases
    std::mt19937_64 rng(seed);
    std::vector<float> init(n_params, 0.0f);
    for (int i = 0; i < L; ++i) {
      const double bound = std::sqrt(6.0 / s[i]);
      std::uniform_real_distribution<double> U(-bound, bound);
      for (int64_t e = 0; e < (int64_t)s[i + 1] * s[i]; ++e) init[offW[i] + e] = (float)U(rng);
    }
    set_params(init.data());
  }

  ~Mlp() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void set_params(const float *host) {
    HIPCHECK(hipMemcpy(p32.p, host, n_params * 4, hipMemcpyHostToDevice));
    std::vector<bf16> w(n_w16, (bf16)0.0f);
    for (int i = 0; i < L; ++i)
      for (int32_t o = 0; o < s[i + 1]; ++o)
        for (int32_t k = 0; k < s[i]; ++k)
          w[offW16[i] + (int64_t)o * sp[i] + k] = (bf16)host[offW[i] + (int64_t)o * s[i] + k];
    HIPCHECK(hipMemcpy(w16.p, w.data(), n_w16 * 2, hipMemcpyHostToDevice));
    if (fused) {
      std::vector<bf16> wf(n_w16, (bf16)0.0f);
      for (int i = 0; i < L; ++i)
        for (int32_t o = 0; o < s[i + 1]; ++o)
          for (int32_t k = 0; k < s[i]; ++k)
            wf[offW16[i] + fsm::fused_woff(i, o, k, sp[i + 1])] = (bf16)host[offW[i] + (int64_t)o * s[i] + k];
      HIPCHECK(hipMemcpy(wf16.p, wf.data(), n_w16 * 2, hipMemcpyHostToDevice));
    }
    reset_optimizer();
  }

  void reset_optimizer() {
    HIPCHECK(hipMemset(m1.p, 0, n_params * 4));
    HIPCHECK(hipMemset(m2.p, 0, n_params * 4));
    HIPCHECK(hipMemset(ctr.p, 0, 16));
  }

  void get_params(float *host) const { HIPCHECK(hipMemcpy(host, p32.p, n_params * 4, hipMemcpyDeviceToHost)); }

  void ensure_rows(int32_t rows) {
    if (rows <= rows_cap) return;
    rows_cap = rows;
    x16.alloc((size_t)rows * sp[0] * 2);
    tgt.alloc((size_t)rows * s[L] * 4);
    yout.alloc((size_t)rows * sp[L] * 4);
    z.clear(); h.clear(); dz32.clear(); dz16.clear();
    for (int i = 0; i < L; ++i) {
      z.emplace_back(new Buf); h.emplace_back(new Buf); dz32.emplace_back(new Buf); dz16.emplace_back(new Buf);
      z[i]->alloc((size_t)rows * sp[i + 1] * 4);
      h[i]->alloc((size_t)rows * sp[i + 1] * 2);
      dz32[i]->alloc((size_t)rows * sp[i + 1] * 4);
      dz16[i]->alloc((size_t)rows * sp[i + 1] * 2);
    }
    graph_rows = -1;
  }

  static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

  void gemm(const bf16 *A, int64_t sam, int64_t sak, const bf16 *B, int64_t sbk, int64_t sbn, int32_t M,
            int32_t N, int32_t K, int32_t epi, int32_t n_valid, const float *bias, float *of, int64_t ldf,
            bf16 *ob, int64_t ldb, const float *zp, int64_t ldz) {
    fsm::GemmArgs g{};
    g.A = A; g.sam = sam; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbn = sbn;
    g.M = M; g.N = N; g.K = K;
    g.epi = epi; g.n_valid = n_valid;
    // 16-byte loads along the unit-stride dimension need an aligned base and
    // a multiple-of-8 leading stride
    g.vec_a = aligned16(A) && ((sak == 1 && sam % 8 == 0) || (sam == 1 && sak % 8 == 0));
    g.vec_b = aligned16(B) && ((sbk == 1 && sbn % 8 == 0) || (sbn == 1 && sbk % 8 == 0));
    g.bias = bias; g.out_f32 = of; g.ld_f32 = ldf; g.out_bf16 = ob; g.ld_bf16 = ldb;
    g.zprev = zp; g.ld_z = ldz;
    if (M <= 0 || N <= 0) return;
    const dim3 grid((N + fsm::BN - 1) / fsm::BN, (M + fsm::BM - 1) / fsm::BM);
    if (grid.y > 65535u) throw std::invalid_argument("GEMM row count too large for one launch");
    hipLaunchKernelGGL(fsm::k_gemm, grid, dim3(256), 0, stream, g);
    HIPCHECK(hipGetLastError());
  }

  const bf16 *W16(int i) const { return w16.as<bf16>() + offW16[i]; }
  float *P32(int64_t off) const { return p32.as<float>() + off; }

  // forward of `rows` packed rows in x16; training keeps pre-activations
  void forward(int32_t rows, bool keep_z, float *y_out, int64_t ld_y) {
    const bf16 *in = x16.as<bf16>();
    for (int i = 0; i < L; ++i) {
      const bool last = i == L - 1;
      if (!last)
        gemm(in, sp[i], 1, W16(i), 1, sp[i], rows, sp[i + 1], sp[i], fsm::EPI_BIAS_GELU, s[i + 1],
             P32(offb[i]), keep_z ? z[i]->as<float>() : nullptr, sp[i + 1], h[i]->as<bf16>(), sp[i + 1],
             nullptr, 0);
      else
        gemm(in, sp[i], 1, W16(i), 1, sp[i], rows, sp[i + 1], sp[i], fsm::EPI_BIAS, s[i + 1], P32(offb[i]),
             y_out, ld_y, nullptr, 0, nullptr, 0);
      in = h[i]->as<bf16>();
    }
  }

  // the whole forward of n rows, x f32 [n][in] -> y f32 [n][out] (device
  // memory), one persistent block per compute unit; nothing allocated
  void forward_fused(const float *x, int64_t n, float *y) {
    fsm::FusedArgs a{};
    a.x = x;
    a.y = y;
    a.n = n;
    a.L = L;
    for (int i = 0; i <= L; ++i) {
      a.sz[i] = s[i];
      a.sp[i] = sp[i];
    }
    for (int i = 0; i < L; ++i) {
      a.w[i] = wf16.as<bf16>() + offW16[i];
      a.b[i] = P32(offb[i]);
    }
    const int64_t tiles = (n + fsm::FUSED_ROWS - 1) / fsm::FUSED_ROWS;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, std::max(1, n_cu)));
    hipLaunchKernelGGL(fsm::k_mlp_fused, dim3(grid), dim3(fsm::FUSED_THREADS), 0, stream, a);
    HIPCHECK(hipGetLastError());
  }

  // loss and gradients of the rows in x16 / tgt (gradient into `grad`)
  void loss_backward(int32_t rows) {
    float *y = yout.as<float>();
    forward(rows, true, y, sp[L]);
    const int blocks = std::max(1, std::min(1024, (int)(((int64_t)rows * sp[L] + 255) / 256)));
    hipLaunchKernelGGL(fsm::k_mse, dim3(blocks), dim3(256), 0, stream, y, (int64_t)sp[L], tgt.as<float>(),
                       rows, s[L], sp[L], dz32[L - 1]->as<float>(), dz16[L - 1]->as<bf16>(),
                       loss.as<float>());
    HIPCHECK(hipGetLastError());
    for (int i = L - 1; i >= 0; --i) {
      const bf16 *hin = i == 0 ? x16.as<bf16>() : h[i - 1]->as<bf16>();
      // dW_i[o][k] = sum_b dZ_i[b][o] * H_{i-1}[b][k]
      gemm(dz16[i]->as<bf16>(), 1, sp[i + 1], hin, sp[i], 1, s[i + 1], s[i], rows, fsm::EPI_F32, s[i],
           nullptr, grad.as<float>() + offW[i], s[i], nullptr, 0, nullptr, 0);
      hipLaunchKernelGGL(fsm::k_colsum, dim3((s[i + 1] + 63) / 64), dim3(256), 0, stream,
                         dz32[i]->as<float>(), rows, s[i + 1], (int64_t)sp[i + 1],
                         grad.as<float>() + offb[i]);
      HIPCHECK(hipGetLastError());
      if (i > 0)  // dZ_{i-1} = (dZ_i W_i) * gelu'(Z_{i-1})
        gemm(dz16[i]->as<bf16>(), sp[i + 1], 1, W16(i), sp[i], 1, rows, sp[i], sp[i + 1], fsm::EPI_GELU_BWD,
             s[i], nullptr, dz32[i - 1]->as<float>(), sp[i], dz16[i - 1]->as<bf16>(), sp[i],
             z[i - 1]->as<float>(), sp[i]);
    }
  }

  void pack_host_rows(const float *x, int32_t rows) {
    Buf tmp;
    tmp.alloc((size_t)rows * s[0] * 4);
    HIPCHECK(hipMemcpy(tmp.p, x, (size_t)rows * s[0] * 4, hipMemcpyHostToDevice));
    const int blocks = std::max(1, std::min(4096, (int)(((int64_t)rows * sp[0] + 255) / 256)));
    hipLaunchKernelGGL(fsm::k_pack_rows, dim3(blocks), dim3(256), 0, stream, tmp.as<float>(),
                       (const int64_t *)nullptr, (const int64_t *)nullptr, (int64_t)rows, s[0], sp[0],
                       x16.as<bf16>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream));
  }

  void loss_grad(const float *x, const float *y, int32_t rows, float *loss_out, float *grad_out) {
    if (rows < 1) throw std::invalid_argument("rows must be >= 1");
    ensure_rows(rows);
    pack_host_rows(x, rows);
    HIPCHECK(hipMemcpy(tgt.p, y, (size_t)rows * s[L] * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(loss.p, 0, 4));
    loss_backward(rows);
    HIPCHECK(hipStreamSynchronize(stream));
    if (loss_out) HIPCHECK(hipMemcpy(loss_out, loss.p, 4, hipMemcpyDeviceToHost));
    if (grad_out) HIPCHECK(hipMemcpy(grad_out, grad.p, n_params * 4, hipMemcpyDeviceToHost));
  }

  // rows of x (host or device f32 [n][in]) -> y (f32 [n][out]), in chunks
  // (the fused kernel takes device rows in one launch)
  double predict(const float *x, int64_t n, float *y, bool dev) {
    if (fused && dev) {
      HIPCHECK(hipEventRecord(ev[0], stream));
      forward_fused(x, n, y);
      HIPCHECK(hipEventRecord(ev[1], stream));
      HIPCHECK(hipEventSynchronize(ev[1]));
      float t = 0;
      HIPCHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
      return t;
    }
    const int32_t chunk = (int32_t)std::min<int64_t>(n, 1 << 18);
    if (!fused) ensure_rows(std::max<int32_t>(chunk, 1));
    Buf xin, yo;
    if (!dev) {
      xin.alloc((size_t)chunk * s[0] * 4);
      yo.alloc((size_t)chunk * s[L] * 4);
    }
    double ms = 0.0;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
      const int32_t rows = (int32_t)std::min<int64_t>(chunk, n - r0);
      const float *xs = dev ? x + r0 * s[0] : xin.as<float>();
      float *ys = dev ? y + r0 * s[L] : yo.as<float>();
      if (!dev) HIPCHECK(hipMemcpy(xin.p, x + r0 * s[0], (size_t)rows * s[0] * 4, hipMemcpyHostToDevice));
      HIPCHECK(hipEventRecord(ev[0], stream));
      if (fused) {
        forward_fused(xs, rows, ys);
      } else {
        const int blocks = std::max(1, std::min(8192, (int)(((int64_t)rows * sp[0] + 255) / 256)));
        hipLaunchKernelGGL(fsm::k_pack_rows, dim3(blocks), dim3(256), 0, stream, xs, (const int64_t *)nullptr,
                           (const int64_t *)nullptr, (int64_t)rows, s[0], sp[0], x16.as<bf16>());
        HIPCHECK(hipGetLastError());
        forward(rows, false, ys, s[L]);
      }
      HIPCHECK(hipEventRecord(ev[1], stream));
      HIPCHECK(hipEventSynchronize(ev[1]));
      float t = 0;
      HIPCHECK(hipEventElapsedTime(&t, ev[0], ev[1]));
      ms += t;
      if (!dev) HIPCHECK(hipMemcpy(y + r0 * s[L], yo.p, (size_t)rows * s[L] * 4, hipMemcpyDeviceToHost));
    }
    return ms;
  }

  // Rows of one chunk of the allocation-free forward: the packed input
  // (in_pad bf16) and two hidden buffers (widest hidden layer, bf16) per row
  // in NOALLOC_BYTES, a multiple of the 64-row GEMM tile; 0 if 64 rows do
  // not fit.
  int64_t noalloc_rows() const {
    int64_t maxw = 0;
    for (int i = 1; i < L; ++i) maxw = std::max<int64_t>(maxw, sp[i]);
    const int64_t per_row = 2 * ((int64_t)sp[0] + 2 * maxw);
    return (int64_t)NOALLOC_BYTES / per_row / 64 * 64;
  }

  // x (f32 [n][in]) and y (f32 [n][out]) in device memory; no allocation:
  // every intermediate lives in this device's `scratch`.  The GEMMs read only
  // the unpadded inner dimension (K = width), so nothing depends on what an
  // earlier call left in the padding columns.
  void predict_noalloc(const float *x, int64_t n, float *y) {
    if (fused) {  // no intermediate in memory at all
      forward_fused(x, n, y);
      HIPCHECK(hipStreamSynchronize(stream));
      return;
    }
    const int64_t R = noalloc_rows();
    if (R < 64)
      throw std::length_error("layer widths need more than the 8 MiB allocation-free scratch "
                              "for one 64-row chunk");
    int64_t maxw = 0;
    for (int i = 1; i < L; ++i) maxw = std::max<int64_t>(maxw, sp[i]);
    bf16 *x16s = static_cast<bf16 *>(noalloc);
    bf16 *hbuf[2] = {x16s + R * sp[0], x16s + R * sp[0] + R * maxw};
    for (int64_t r0 = 0; r0 < n; r0 += R) {
      const int32_t rows = (int32_t)std::min<int64_t>(R, n - r0);
      const int blocks = std::max(1, std::min(8192, (int)(((int64_t)rows * sp[0] + 255) / 256)));
      hipLaunchKernelGGL(fsm::k_pack_rows, dim3(blocks), dim3(256), 0, stream, x + r0 * s[0],
                         (const int64_t *)nullptr, (const int64_t *)nullptr, (int64_t)rows, s[0], sp[0],
                         x16s);
      HIPCHECK(hipGetLastError());
      const bf16 *in = x16s;
      for (int i = 0; i < L; ++i) {
        if (i < L - 1) {
          bf16 *out = hbuf[i & 1];
          gemm(in, sp[i], 1, W16(i), 1, sp[i], rows, sp[i + 1], s[i], fsm::EPI_BIAS_GELU, s[i + 1],
               P32(offb[i]), nullptr, 0, out, sp[i + 1], nullptr, 0);
          in = out;
        } else {
          gemm(in, sp[i], 1, W16(i), 1, sp[i], rows, sp[i + 1], s[i], fsm::EPI_BIAS, s[i + 1],
               P32(offb[i]), y + r0 * s[L], s[L], nullptr, 0, nullptr, 0);
        }
      }
    }
    HIPCHECK(hipStreamSynchronize(stream));
  }

  void set_data(const float *x, const float *y, int64_t n) {
    if (n < 1) throw std::invalid_argument("empty training set");
    xall.alloc((size_t)n * s[0] * 4);
    yall.alloc((size_t)n * s[L] * 4);
    HIPCHECK(hipMemcpy(xall.p, x, (size_t)n * s[0] * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(yall.p, y, (size_t)n * s[L] * 4, hipMemcpyHostToDevice));
    n_data = n;
    perm.alloc((size_t)n * 8);
    graph_rows = -1;
  }

  // one optimiser step on minibatch ctr[1] of the current permutation
  void launch_train_step(int32_t rows, const fs_adam &o) {
    int64_t *c = ctr.as<int64_t>();
    const int b1 = std::max(1, std::min(4096, (int)(((int64_t)rows * sp[0] + 255) / 256)));
    hipLaunchKernelGGL(fsm::k_pack_rows, dim3(b1), dim3(256), 0, stream, xall.as<float>(),
                       perm.as<int64_t>(), (const int64_t *)(c + 1), (int64_t)rows, s[0], sp[0], x16.as<bf16>());
    HIPCHECK(hipGetLastError());
    const int b2 = std::max(1, std::min(4096, (int)(((int64_t)rows * s[L] + 255) / 256)));
    hipLaunchKernelGGL(fsm::k_gather_rows_f32, dim3(b2), dim3(256), 0, stream, yall.as<float>(),
                       perm.as<int64_t>(), (const int64_t *)(c + 1), (int64_t)rows, s[L], tgt.as<float>());
    HIPCHECK(hipGetLastError());
    loss_backward(rows);
    hipLaunchKernelGGL(fsm::k_incr, dim3(1), dim3(1), 0, stream, c);  // Adam step t (1-based)
    HIPCHECK(hipGetLastError());
    const int b3 = std::max(1, std::min(2048, (int)((n_params + 255) / 256)));
    hipLaunchKernelGGL(fsm::k_adam, dim3(b3), dim3(256), 0, stream, p32.as<float>(), grad.as<float>(),
                       m1.as<float>(), m2.as<float>(), n_params, (const int64_t *)c, o.lr, o.beta1, o.beta2,
                       o.eps, o.weight_decay, w16map.as<int64_t>(), w16.as<bf16>(),
                       fused ? wfmap.as<int64_t>() : (const int64_t *)nullptr,
                       fused ? wf16.as<bf16>() : (bf16 *)nullptr);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(fsm::k_incr, dim3(1), dim3(1), 0, stream, c + 1);  // next minibatch
    HIPCHECK(hipGetLastError());
  }

  void train(int32_t epochs, int32_t batch, const fs_adam &o, uint64_t shuffle_seed, int32_t shuffle,
             float *epoch_loss) {
    if (!n_data) throw std::invalid_argument("no training set (fs_mlp_set_data)");
    if (batch < 1 || batch > n_data) throw std::invalid_argument("batch must be 1..n_data");
    ensure_rows(batch);
    const int64_t steps = n_data / batch;  // a partial last minibatch is dropped
    if (graph_rows != batch || std::memcmp(&graph_opt, &o, sizeof o) != 0) {
      if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
      if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
      HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      launch_train_step(batch, o);
      HIPCHECK(hipStreamEndCapture(stream, &graph));
      HIPCHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
      graph_rows = batch;
      graph_opt = o;
    }
    std::vector<int64_t> idx(n_data);
    std::iota(idx.begin(), idx.end(), 0);
    std::mt19937_64 rng(shuffle_seed);
    for (int32_t ep = 0; ep < epochs; ++ep) {
      if (shuffle) std::shuffle(idx.begin(), idx.end(), rng);
      HIPCHECK(hipMemcpyAsync(perm.p, idx.data(), n_data * 8, hipMemcpyHostToDevice, stream));
      HIPCHECK(hipMemsetAsync(ctr.as<int64_t>() + 1, 0, 8, stream));
      HIPCHECK(hipMemsetAsync(loss.p, 0, 4, stream));
      for (int64_t st = 0; st < steps; ++st) HIPCHECK(hipGraphLaunch(gexec, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      float l = 0;
      HIPCHECK(hipMemcpy(&l, loss.p, 4, hipMemcpyDeviceToHost));
      if (epoch_loss) epoch_loss[ep] = l / (float)steps;
    }
  }
};

}  // namespace

struct fs_mlp {
  std::unique_ptr<Mlp> m;
};

#define MLP_CALL(...)                                                   \
  try {                                                                 \
    __VA_ARGS__;                                                        \
    return FS_OK;                                                       \
  } catch (const HipError &ex) {                                        \
    return fs::set_error(FS_ERR_HIP, ex.what());                        \
  } catch (const std::length_error &ex) {                               \
    return fs::set_error(FS_ERR_CAPACITY, ex.what());                   \
  } catch (const std::bad_alloc &) {                                    \
    return fs::set_error(FS_ERR_CAPACITY, "host allocation failed");    \
  } catch (const std::exception &ex) {                                  \
    return fs::set_error(FS_ERR_ARG, ex.what());                        \
  }

extern "C" {

int fs_mlp_create(const int32_t *sizes, int32_t n_layers, uint64_t seed, fs_mlp **out) {
  if (!sizes || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  MLP_CALL({
    auto h = std::make_unique<fs_mlp>();
    h->m.reset(new Mlp(sizes, n_layers, seed));
    *out = h.release();
  })
}

int fs_mlp_destroy(fs_mlp *h) {
  delete h;
  return FS_OK;
}

int fs_mlp_n_params(fs_mlp *h, int64_t *n) {
  if (!h || !n) return fs::set_error(FS_ERR_ARG, "null argument");
  *n = h->m->n_params;
  return FS_OK;
}

int fs_mlp_get_params(fs_mlp *h, float *out) {
  if (!h || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  MLP_CALL(h->m->get_params(out))
}

int fs_mlp_set_params(fs_mlp *h, const float *in) {
  if (!h || !in) return fs::set_error(FS_ERR_ARG, "null argument");
  MLP_CALL(h->m->set_params(in))
}

int fs_mlp_loss_grad(fs_mlp *h, const float *x, const float *y, int32_t rows, float *loss, float *grad) {
  if (!h || !x || !y) return fs::set_error(FS_ERR_ARG, "null argument");
  MLP_CALL(h->m->loss_grad(x, y, rows, loss, grad))
}

int fs_mlp_predict(fs_mlp *h, const float *x, int64_t n, float *y, int32_t device_ptrs, double *ms) {
  if (!h || !x || !y || n < 0) return fs::set_error(FS_ERR_ARG, "bad argument");
  MLP_CALL({
    const double t = n ? h->m->predict(x, n, y, device_ptrs != 0) : 0.0;
    if (ms) *ms = t;
  })
}

int fs_mlp_predict_noalloc(fs_mlp *h, const float *x, int64_t n, float *y) {
  if (!h || !x || !y || n < 0) return fs::set_error(FS_ERR_ARG, "bad argument");
  MLP_CALL(if (n) h->m->predict_noalloc(x, n, y))
}

int fs_mlp_set_data(fs_mlp *h, const float *x, const float *y, int64_t n) {
  if (!h || !x || !y) return fs::set_error(FS_ERR_ARG, "null argument");
  MLP_CALL(h->m->set_data(x, y, n))
}

int fs_mlp_train(fs_mlp *h, int32_t epochs, int32_t batch, const fs_adam *opt, uint64_t shuffle_seed,
                 int32_t shuffle, float *epoch_loss) {
  if (!h || !opt || epochs < 0) return fs::set_error(FS_ERR_ARG, "bad argument");
  MLP_CALL(h->m->train(epochs, batch, *opt, shuffle_seed, shuffle, epoch_loss))
}

int fs_device_alloc(int64_t bytes, void **ptr) {
  if (!ptr || bytes < 0) return fs::set_error(FS_ERR_ARG, "bad argument");
  MLP_CALL(HIPCHECK(hipMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1))))
}

int fs_device_free(void *ptr) {
  MLP_CALL(if (ptr) HIPCHECK(hipFree(ptr)))
}

int fs_device_copy(void *dst, const void *src, int64_t bytes, int32_t dir) {
  if (!dst || !src || bytes < 0 || dir < 0 || dir > 2) return fs::set_error(FS_ERR_ARG, "bad argument");
  static const hipMemcpyKind kind[3] = {hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
                                        hipMemcpyDeviceToDevice};
  MLP_CALL(HIPCHECK(hipMemcpy(dst, src, (size_t)bytes, kind[dir])))
}

}  // extern "C"
