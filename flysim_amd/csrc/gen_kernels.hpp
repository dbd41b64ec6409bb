// Device generation of batched random E/I networks (many instances of
// uneven size in one handle), so that billions of synapses never pass
// through host memory.
//
// Instance k has n_k neurons, the first n_exc_k excitatory.  Every ordered
// pair (j, p), j != p, of the instance is a synapse with probability p_conn,
// decided by a counter-based hash of (instance key, j, p): splitmix64
// (Steele, Lea, Flood, OOPSLA 2014) of key ^ (j << 32 | p); its high 32 bits
// against p_conn * 2^32 decide the pair and its low 24 bits give the
// efficacy g = g_lo + (g_hi - g_lo) * u, u = low24 / 2^24 (float, one
// rounding per operation).  E->X pairs carry AMPA (g) and NMDA
// (g * nmda_ratio * gain_nmda), I->X pairs GABA (g * gain_gaba); every
// synapse takes the same synaptic delay.  Row scans are wave-wide: one wave per CSR row tests 64 candidates
// per ballot, so rows come out in ascending column order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsg {

__host__ __device__ inline uint64_t mix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct GenInst {
  int64_t first;      // global id of neuron 0
  int32_t n, n_exc;
  uint64_t key;       // mix64(seed)
  float gain_nmda, gain_gaba;
};

struct GenArgs {
  const GenInst *inst;
  const uint16_t *inst_of;  // [N] instance of each neuron
  int32_t N;
  uint32_t thr;             // p_conn * 2^32
  float g_lo, g_span, nmda_ratio;
  uint32_t delay;           // synaptic delay of every synapse (steps, < 64)
  // outputs (fill pass): CSR by pre (AMPA from E, GABA from I) and CSR by
  // post (NMDA from E); count pass writes row lengths to rp[row + 1]
  uint64_t *rp_p, *rp_n;
  uint32_t *meta_p, *meta_n, *nmda_out;
  float *w_p, *w_n;
  int32_t fill;
};

__device__ __forceinline__ bool pair_of(uint64_t key, uint32_t j, uint32_t p, uint32_t thr, float &u) {
  const uint64_t h = mix64(key ^ (((uint64_t)j << 32) | p));
  u = (float)(uint32_t)(h & 0xFFFFFFu) * (1.0f / 16777216.0f);
  return (uint32_t)(h >> 32) < thr;
}

__device__ __forceinline__ float eff(const GenArgs &a, float u) {
  return __fadd_rn(a.g_lo, __fmul_rn(a.g_span, u));
}

// one wave per global neuron r: its push row (as presynaptic neuron) and its
// NMDA gather row (as postsynaptic neuron)
__global__ void __launch_bounds__(256) k_gen_rows(GenArgs a) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = (int)(threadIdx.x & 63);
  if (r >= a.N) return;  // whole waves exit together
  const GenInst I = a.inst[a.inst_of[r]];
  const uint32_t loc = (uint32_t)(r - I.first);
  const bool exc = (int32_t)loc < I.n_exc;
  const uint64_t lanes_below = (1ull << lane) - 1ull;
  // ---- push row: targets p of presynaptic loc
  {
    uint64_t pos = a.fill ? a.rp_p[r] : 0;
    for (int32_t b0 = 0; b0 < I.n; b0 += 64) {
      const uint32_t p = (uint32_t)(b0 + lane);
      float u = 0.0f;
      const bool ok = (int32_t)p < I.n && p != loc && pair_of(I.key, loc, p, a.thr, u);
      const uint64_t m = __ballot(ok);
      if (a.fill && ok) {
        const uint64_t e = pos + (uint64_t)__popcll(m & lanes_below);
        const float g = eff(a, u);
        a.meta_p[e] = (uint32_t)(I.first + p) | (a.delay << 24) | ((exc ? 0u : 1u) << 30);
        a.w_p[e] = exc ? g : __fmul_rn(g, I.gain_gaba);
      }
      pos += (uint64_t)__popcll(m);
    }
    if (!a.fill && lane == 0) a.rp_p[r + 1] = pos;
    if (!a.fill && lane == 0) a.nmda_out[r] = exc ? (uint32_t)pos : 0u;
  }
  // ---- NMDA row: excitatory presynaptic j of postsynaptic loc
  {
    uint64_t pos = a.fill ? a.rp_n[r] : 0;
    for (int32_t b0 = 0; b0 < I.n_exc; b0 += 64) {
      const uint32_t j = (uint32_t)(b0 + lane);
      float u = 0.0f;
      const bool ok = (int32_t)j < I.n_exc && j != loc && pair_of(I.key, j, loc, a.thr, u);
      const uint64_t m = __ballot(ok);
      if (a.fill && ok) {
        const uint64_t e = pos + (uint64_t)__popcll(m & lanes_below);
        a.meta_n[e] = (uint32_t)(I.first + j) | (a.delay << 24);
        a.w_n[e] = __fmul_rn(__fmul_rn(eff(a, u), a.nmda_ratio), I.gain_nmda);
      }
      pos += (uint64_t)__popcll(m);
    }
    if (!a.fill && lane == 0) a.rp_n[r + 1] = pos;
  }
}

}  // namespace fsg
