// Device kernels of the Flysim engine (gfx950, wave64).
//
// One simulated step t is two launches on the engine's stream:
//   k_update  : per neuron, elementwise over SoA state -- ring-slot intake of
//               the linear conductances, external Poisson gates (counter RNG
//               or replayed bits), conductance LIF update with NMDA Mg block
//               and the refractory state machine, presynaptic gate update,
//               and wave-ballot compaction of the step's spike list;
//   k_deliver : CSR push of the compacted spike list into per-delay ring
//               buffers (one wave per spike, float/double atomics) and, in
//               the same grid, the NMDA pull-gather over CSR-by-post rows
//               (a group of L lanes per row with a shuffle reduction).
// The equations restate code/EqsGnl2.cpp:152-337 (VmEul, firing, BgInit) and
// :404-502 / :866-1002 (gates and sums).  For the linear receptors the
// per-synapse gates of the reference are replaced by the exact push identity
//   G_p(t) = df_p G_p(t-1) + sum_{j spiking at t-d} g_jp (1 + (1-df_p) s_j(t-d-1))
// where s_j is the presynaptic gate shared by all of j's synapses; NMDA keeps
// the per-step gather G_N(p) = sum_j g_jp s_j(t-d_jp) because its gate
// saturates (the (1-s) term) and has no push form.
//
// Replicas: B copies of one network share its synapse tables.  Their state is
// interleaved, global neuron n = local * B + replica, so a wave of B >= 64
// consecutive neurons is one local neuron in 64 copies: the CSR row of that
// neuron is read once (same address in every lane) and the presynaptic state
// it points at is read as contiguous B-wide vectors.  Per copy: recurrent
// gains by receptor, an external efficacy gain, external event probabilities
// per (population, receptor) and the key of the counter-based random stream.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsk {

// event counters are spread over NCTR slots (block index) against contention
constexpr int NCTR = 64;
// The spike record of a batch is split into NSUB sub-lists: block b appends
// to sub-list b % NSUB, so the claim of each block's range is an atomic on one
// of NSUB counters instead of one counter for the whole grid.  Sub-list s owns
// a fixed region of the record sized for every neuron of its blocks spiking
// every step of the batch.
constexpr int NSUB = 64;

// state word: refractory count in bits 0..23, flags above
constexpr uint32_t CNT_MASK = 0x00FFFFFFu;
constexpr uint32_t F_FIRE = 1u << 24, F_SPIKE = 1u << 25, F_LPFIRE = 1u << 26, F_LPSPIKE = 1u << 27,
                   F_OVF = 1u << 28, F_UDF = 1u << 29;
constexpr uint32_t ACT_AMPA = 1, ACT_GABA = 2, ACT_NMDA = 4, ACT_ACH = 32, ACT_GCL = 64,
                   ACT_INJ = 128;
constexpr int32_t STP_STD = 1, STP_FULL = 2;
// linear receptor slots li = 0..3 map to receptor ids A=0, G=1, H=3, L=4
__host__ __device__ constexpr int lin_rid(int li) { return li < 2 ? li : li + 1; }
__host__ __device__ constexpr uint32_t lin_act(int li) {
  return li == 0 ? ACT_AMPA : li == 1 ? ACT_GABA : li == 2 ? ACT_ACH : ACT_GCL;
}

struct ZigTables;

// Per-population parameters in the precision of the run.
template <typename T>
struct DevPop {
  T vth, leak_drive, cm, vreset, df, ef, dvmax, mg;
  T vrev[5];
  T dfst[5];            // postsynaptic gate decay by receptor id
  T dfxt_n, efas_n;     // postsynaptic NMDA (external gate) x decay, efst*as
  T gext[5];
  T pre_df;             // presynaptic linear gate decay (targets' dfst)
  T pre_dfxt, pre_dfst, pre_efas;  // presynaptic NMDA gate (targets' params)
  T inj_mean, inj_std;
  T ltp_dfLP, ltp_pos, ltp_neg;  // long-term plasticity (single synapses)
  int32_t ltp_spidy;
  // short-term plasticity of the population's outgoing synapses
  T stp_pv, stp_dfD, stp_efD, stp_aF, stp_dfF, stp_efF, stp_aCap, stp_dfCap, stp_aCar,
      stp_dfCar, stp_efCar;
  int32_t rfped, spidy;
  uint32_t act;             // only ACT_INJ is read on the device
  uint32_t has_lin_out, has_nmda_out;
  // random stream of the population's network instance: Philox key
  // seed ^ replica key ^ (key_hi:key_lo), counter = local neuron - first
  uint32_t key_lo, key_hi;
  int32_t first;
  // replay mode: the Gaussian injection of neuron n at step t is call
  // t * zig_a + n + zig_b (mod 2^32) of the reference's ziggurat stream
  uint32_t zig_a, zig_b;
};

template <typename T>
struct Net {
  int32_t N;
  int32_t D;            // ring depth (linear receptors)
  int32_t DN;           // NMDA gate history depth
  int32_t lin_mask;     // bit li set: linear receptor slot li has synapses
  int32_t ext_mask;     // bit rid set: some population has external drive on rid
  int32_t has_nmda_syn;
  int32_t ext_mode;     // 0 philox, 1 replay
  int32_t stp_mode;     // 0 none, STP_STD depression only, STP_FULL depression + facilitation
  int32_t ltp_mode;     // 1: LTP traces and learning synapses (two-launch path)
  int32_t words;        // ceil(N/32)
  int32_t B;            // replicas (interleaved: n = local * B + replica)
  int32_t N0;           // neurons of one replica (N = N0 * B)
  int32_t P0;           // populations of one replica
  const uint32_t *npop; // [N0] population of each local neuron
  const DevPop<T> *pops;// [P0]
  const float *pbr;     // [5][P0][B] external event probability per step; < 0: the
                        // receptor's activity bit is cleared (no draw, gate frozen)
  const float *gainr;   // [6][B] recurrent gain by receptor id, then external efficacy gain
  const uint64_t *skey; // [B] random-stream key of each replica (xor into the seed)
  T *V;
  uint32_t *cs;
  T *G;                 // [4][N] running linear conductances
  T *ring;              // [4][D][N]
  T *spre;              // [N] linear presynaptic gate
  T *xn;                // [N] NMDA presynaptic x
  T *shist;             // [N0][DN][B] NMDA presynaptic s history (slot-contiguous per neuron)
  T *gn;                // [N] gathered NMDA conductance of the last step
  T *elin;              // [4][N] external linear gates
  T *enx, *ens;         // [N] external NMDA gate
  T *sD, *sF, *sCap, *sCar;  // [N] presynaptic short-term plasticity state (stp_mode != 0)
  // gap junctions (two-launch path only): both ends of each junction
  const uint64_t *rp_g; // [N0+1]
  const uint32_t *col_g;// other end (local neuron)
  const float *w_g;     // conductance (nS)
  T *igap;              // [N] gap current for the next step (null: no gap junctions)
  // long-term plasticity (ltp_mode): per-neuron LP trace and LTP-spike delay
  // counter, and the learning synapses (the reference's single-synapse path)
  T *lp;
  uint32_t *cntlp;
  int64_t n_learn;
  const uint32_t *l_pre, *l_post;
  const uint8_t *l_rec;
  T *l_w;               // current weights (nS)
  T *gl;                // [5][N] the learning synapses' conductance of the last Spike phase
  T *irep;              // [5][N] synaptic current by receptor of the last non-refractory
                        // step (pA), written only when recording is on (else null)
  // push CSR (linear)
  const uint64_t *rp_p;
  const uint32_t *meta_p;   // col | delay<<24 | li<<30
  const float *w_p;
  // gather CSR (NMDA)
  const uint64_t *rp_n;
  const uint32_t *meta_n;   // pre | delay<<24
  const float *w_n;
  const uint32_t *nmda_out; // [N] NMDA out-degree (event accounting)
  // replayed external bits
  const uint32_t *ext_bits;
  int64_t ext_first, ext_steps;
  uint64_t seed;
  // replay mode: the reference's ziggurat tables for the Gaussian injection
  const ZigTables *zig;
  // spike record of the current batch
  uint32_t *rec_ids;
  T *rec_a;
  uint32_t *rec_k;          // windowed path: step of each record entry within the batch
  uint32_t *total;          // [NSUB] spikes recorded in this batch, per sub-list
  uint32_t *cum;            // [K+1][NSUB] per-step offsets within each sub-list
  uint32_t *overflow;       // [1]
  uint32_t *lcnt;           // [3][NSUB] fused path: spike counts of the lists of steps k%3
  uint32_t *lbase;          // [3][NSUB] fused path: offsets of the list of step k%3
  const uint32_t *sub_off;  // [NSUB] region of each sub-list in rec_ids / rec_a
  const uint32_t *sub_cap;  // [NSUB] its capacity
  uint64_t rec_cap;
  const int64_t *t0;        // [1] first step of the batch
  const int32_t *tmod;      // [2] t0 % D, t0 % DN (slot arithmetic stays 32-bit)
  unsigned long long *pop_cnt;  // [P0][B]
  unsigned long long *counters; // [3][NCTR]: spikes, linear pushes, NMDA out-synapses of
                                // spikers, each spread over NCTR slots by block index
  int32_t gather_lanes;
  int32_t rec_on;           // 1: the batch's spikes go to the record (readback asked for)
  const uint64_t *wblk;     // windowed path: block b covers neurons [first, first + count),
                            // first | count << 32, at most blockDim / gather_lanes of them
  const uint64_t *wgat;     // windowed path, one copy: [2 * b], [2 * b + 1] = block b's range
                            // of gather CSR entries
  const uint32_t *wpop;     // windowed path: population of block b's neurons (0xFFFFFFFF: several)
  int32_t gather_chunk;     // windowed path, one copy: cap on the entries per staged chunk (0: none)
  int32_t gather_split;     // windowed path, one copy: sub-lanes per (row, step) pair of the gather
  int32_t inj_on;           // some population has Gaussian injection on (windowed path's LDS)
  int32_t push_waves;       // waves of k_deliver dedicated to the push
  int32_t push_blocks;      // fused step: leading blocks that push; the rest gather + update
};

// ------------------------------------------------------------------ RNG
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based.
__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
__device__ __forceinline__ float u24(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// The reference's Gaussian injection (ziggurat::Igauss, code/StiExt.h:214-218,
// one stream per solver thread): every call advances the LCG
// jcong' = 69069 jcong + 1234567 once (code/StiExt.h:222-229) and seeds SHR3
// with it, so call k depends only on the seed and k.  lcg_jump gives the
// state after m calls directly, and every neuron-step draws independently.
struct ZigTables {
  uint32_t kn[128];
  float fn[128], wn[128];
  // the LCG advanced b * 256^j calls: x -> ja[j][b] x + jc[j][b]
  uint32_t ja[4][256], jc[4][256];
};

// the LCG state after m calls from seed: one table step per byte of m (the
// powers of one affine map commute)
__device__ __forceinline__ uint32_t lcg_jump(uint32_t seed, uint32_t m, const ZigTables *z) {
  uint32_t s = seed;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t b = (m >> (8 * j)) & 255u;
    s = z->ja[j][b] * s + z->jc[j][b];
  }
  return s;
}

// SHR3 step: returns the input plus the shifted state (shr3_seeded / r4_uni,
// code/StiExt.h:231-259)
__device__ __forceinline__ uint32_t shr3_next(uint32_t &s) {
  const uint32_t in = s;
  s ^= s << 13;
  s ^= s >> 17;
  s ^= s << 5;
  return in + s;
}

// r4_uni: fmod(0.5 + float(u) / 2^32, 1.0) in double, returned as float
__device__ __forceinline__ float zig_uni(uint32_t &s) {
  double d = 0.5 + (double)(float)shr3_next(s) / 65536.0 / 65536.0;
  if (d >= 1.0) d -= 1.0;  // fmod(d, 1.0) for d in [0.5, 1.5]
  return (float)d;
}

// logf, rounded from the double logarithm
__device__ __forceinline__ float zig_logf(float u) { return (float)log((double)u); }

// |hz| < kn[iz], the reference's fabs(hz) < kn[iz] in exact integers
__device__ __forceinline__ bool zig_inside(int32_t hz, uint32_t kn) {
  const uint32_t a = hz < 0 ? 0u - (uint32_t)hz : (uint32_t)hz;
  return a < kn;
}

// r4_nor (code/StiExt.h:262-319), Marsaglia and Tsang's 128-strip ziggurat
// with r = 3.442620, in the reference's float / double mix and without
// contraction.  The rejection loops end after 256 rounds (the expected count
// is about 1.01), where the reference would keep drawing.  Out of line: the
// wedges and the tail (about 1 % of the draws); zig_normal takes the rest.
__device__ __noinline__ float zig_normal_slow(int32_t hz, uint32_t iz, uint32_t jsr,
                                              const ZigTables *z) {
#pragma clang fp contract(off)
  const float r = 3.442620f;
  for (int round = 0; round < 256; ++round) {
    if (iz == 0) {
      float x = 0.0f;
      for (int j = 0; j < 256; ++j) {
        x = (float)(-0.2904764 * (double)zig_logf(zig_uni(jsr)));
        const float y = -zig_logf(zig_uni(jsr));
        if (x * x <= y + y) break;
      }
      return hz <= 0 ? -r - x : r + x;
    }
    const float x = (float)hz * z->wn[iz];
    const float f = z->fn[iz] + zig_uni(jsr) * (z->fn[iz - 1] - z->fn[iz]);
    if ((double)f < exp(-0.5 * (double)x * (double)x)) return x;
    hz = (int32_t)shr3_next(jsr);
    iz = (uint32_t)hz & 127u;
    if (zig_inside(hz, z->kn[iz])) return (float)hz * z->wn[iz];
  }
  return (float)hz * z->wn[iz];
}

__device__ __forceinline__ float zig_normal(uint32_t jsr, const ZigTables *z) {
  const int32_t hz = (int32_t)shr3_next(jsr);
  const uint32_t iz = (uint32_t)hz & 127u;
  if (zig_inside(hz, z->kn[iz])) return (float)hz * z->wn[iz];
  return zig_normal_slow(hz, iz, jsr, z);
}

// Igauss for call k of the stream seeded with seed: mean + sd * r4_nor
__device__ __forceinline__ float zig_igauss(uint32_t seed, uint32_t k, float mean, float sd,
                                            const ZigTables *z) {
#pragma clang fp contract(off)
  const float v = zig_normal(lcg_jump(seed, k + 1u, z), z);
  return mean + sd * v;
}

// Load through the constant address space: for tables no kernel writes
// (CSR, population parameters) a wave-uniform address then becomes a scalar
// load instead of a 64-lane vector load of one address.
template <typename X>
__device__ __forceinline__ X ldc(const X *p) {
  return *(const __attribute__((address_space(4))) X *)p;
}

// Barrier for data exchanged through LDS only: its fences order LDS accesses
// (s_waitcnt lgkmcnt) and leave the block's global loads, stores and atomics
// in flight, where __syncthreads would wait for every one of them.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// (t0 % M + k) mod M for a step offset k >= -M of the batch starting at t0
// M (the ring depth D or the gate history depth DN) is a power of two, so
// the remainder is a mask (any k: two's complement wraps the same way).
__device__ __forceinline__ int32_t slot_of(int32_t t0mod, int32_t k, int32_t M) {
  return (int32_t)((uint32_t)(t0mod + k) & (uint32_t)(M - 1));
}

// Position of the NMDA gate of local neuron nl, replica rep, history slot
// `slot` (DN, a power of two, slots per neuron): the history is slot-contiguous
// per neuron, [N0][DN][B], so the consecutive slots one window reads for a
// presynaptic neuron share one or two cache lines (and B copies of one neuron
// are adjacent).
template <typename T>
__device__ __forceinline__ size_t hidx(const Net<T> &net, uint32_t nl, uint32_t rep, uint32_t slot) {
  return ((size_t)nl * (uint32_t)net.DN + slot) * (uint32_t)net.B + rep;
}

// ------------------------------------------------------------------ update
// Per-neuron state between steps.  The one-step kernels load it, run one
// step and store it; the windowed kernel keeps it in registers for all the
// steps of a launch.  Only the fields the network uses are loaded/stored.
template <typename T>
struct NState {
  T v;
  uint32_t cs;
  T G[4];          // running linear conductances (lin_mask)
  T spre;          // linear presynaptic gate (has_lin_out)
  T xn, sn;        // NMDA presynaptic x and the last step's s (has_nmda_out)
  T elin[4];       // external linear gates (ext_mask)
  T enx, ens;      // external NMDA gate
  T sD, sF, sCap, sCar;  // short-term plasticity (stp_mode)
};

template <typename T>
using PopC = const __attribute__((address_space(4))) DevPop<T>;

// the state of neuron n as of the end of step k - 1 (hprev: gate history
// slot of step k - 1).  The presynaptic gates are loaded whether or not the
// neuron's population has outgoing synapses of their kind (unused values go
// back unchanged), so no load waits on the population parameters.
template <typename T>
__device__ __forceinline__ void load_state(const Net<T> &net, int32_t nl, int32_t rep, int32_t n,
                                           int32_t hprev, NState<T> &S) {
  const size_t N = (size_t)net.N;
  S.v = net.V[n];
  S.cs = net.cs[n];
#pragma unroll
  for (int li = 0; li < 4; ++li) {
    if (net.lin_mask & (1 << li)) S.G[li] = net.G[li * N + n];
    if (net.ext_mask & (1 << lin_rid(li))) S.elin[li] = net.elin[li * N + n];
  }
  S.spre = net.spre[n];
  S.xn = net.xn[n];
  S.sn = net.shist[hidx(net, (uint32_t)nl, (uint32_t)rep, (uint32_t)hprev)];
  if (net.ext_mask & (1 << 2)) {
    S.enx = net.enx[n];
    S.ens = net.ens[n];
  }
  if (net.stp_mode) {
    S.sD = net.sD[n];
    if (net.stp_mode == STP_FULL) {
      S.sF = net.sF[n];
      S.sCap = net.sCap[n];
      S.sCar = net.sCar[n];
    }
  }
}

template <typename T>
__device__ __forceinline__ void store_state(const Net<T> &net, int32_t n, const NState<T> &S) {
  const size_t N = (size_t)net.N;
  net.V[n] = S.v;
  net.cs[n] = S.cs;
#pragma unroll
  for (int li = 0; li < 4; ++li) {
    if (net.lin_mask & (1 << li)) net.G[li * N + n] = S.G[li];
    if (net.ext_mask & (1 << lin_rid(li))) net.elin[li * N + n] = S.elin[li];
  }
  net.spre[n] = S.spre;
  net.xn[n] = S.xn;  // s goes to the history every step
  if (net.ext_mask & (1 << 2)) {
    net.enx[n] = S.enx;
    net.ens[n] = S.ens;
  }
  if (net.stp_mode) {
    net.sD[n] = S.sD;
    if (net.stp_mode == STP_FULL) {
      net.sF[n] = S.sF;
      net.sCap[n] = S.sCap;
      net.sCar[n] = S.sCar;
    }
  }
}

// The linear conductance increments delivered for step k - 1 (its ring
// slot), cleared for reuse.
template <typename T>
__device__ __forceinline__ void ring_take(const Net<T> &net, int32_t n, int32_t slot, T inc[4]) {
#pragma unroll
  for (int li = 0; li < 4; ++li) {
    inc[li] = T(0);
    if (!(net.lin_mask & (1 << li))) continue;
    T *rs = net.ring + ((size_t)li * net.D + slot) * (size_t)net.N + n;
    inc[li] = *rs;
    *rs = T(0);
  }
}

// Per-neuron inputs that stay fixed over a batch of steps: external event
// probabilities (< 0: the receptor's activity bit is cleared), external
// efficacy gain and the counter-based stream's key and counter (neuron
// within its network instance; key = seed ^ replica key ^ instance key, so a
// replica or a batched instance draws the same stream as the same network
// run alone with that key).
template <typename T>
struct NConst {
  float pb[5];
  T eg;
  uint64_t key;
  uint32_t ctr;
  uint32_t nl, rep;  // local neuron and replica
};

template <typename T>
__device__ __forceinline__ void load_const(const Net<T> &net, PopC<T> &P, int32_t nl, int32_t rep,
                                           uint32_t p0, NConst<T> &C) {
  const float *pbp = net.pbr + (size_t)p0 * net.B + rep;
  const size_t pbs = (size_t)net.P0 * net.B;  // receptor stride of pbr
#pragma unroll
  for (int r = 0; r < 5; ++r) C.pb[r] = (net.ext_mask & (1 << r)) ? pbp[r * pbs] : -1.0f;
  C.eg = T(net.gainr[5 * net.B + rep]);
  C.key = net.seed ^ net.skey[rep] ^ (((uint64_t)P.key_hi << 32) | P.key_lo);
  C.ctr = (uint32_t)(nl - P.first);
  C.nl = (uint32_t)nl;
  C.rep = (uint32_t)rep;
}

// load_const in two halves for the windowed kernel: the loads alone (no
// arithmetic on what they return, so issuing them never waits), then the
// derived fields, once the values have landed.
struct NConstRaw {
  float pb[5];
  float eg;
  uint64_t skey;
  uint32_t key_lo, key_hi;
  int32_t first;
};

template <typename T>
__device__ __forceinline__ void load_const_raw(const Net<T> &net, PopC<T> &P, int32_t rep, uint32_t p0,
                                               NConstRaw &R) {
  const float *pbp = net.pbr + (size_t)p0 * net.B + rep;
  const size_t pbs = (size_t)net.P0 * net.B;  // receptor stride of pbr
#pragma unroll
  for (int r = 0; r < 5; ++r) R.pb[r] = (net.ext_mask & (1 << r)) ? pbp[r * pbs] : -1.0f;
  R.eg = net.gainr[5 * net.B + rep];
  R.skey = net.skey[rep];
  R.key_lo = P.key_lo;
  R.key_hi = P.key_hi;
  R.first = P.first;
}

// the fields load_const derives, from load_const_raw's values
template <typename T>
__device__ __forceinline__ void finish_const(const Net<T> &net, const NConstRaw &R, int32_t nl, int32_t rep,
                                             NConst<T> &C) {
#pragma unroll
  for (int r = 0; r < 5; ++r) C.pb[r] = R.pb[r];
  C.eg = T(R.eg);
  C.key = net.seed ^ R.skey ^ (((uint64_t)R.key_hi << 32) | R.key_lo);
  C.ctr = (uint32_t)(nl - R.first);
  C.nl = (uint32_t)nl;
  C.rep = (uint32_t)rep;
}

// External event outcomes and the Gaussian injection draw of neuron n at step
// t, the stochastic part of BgInit (code/EqsGnl2.cpp:284-337): bit rid of
// dbits is set when receptor rid's external gate receives an event this step
// (order A, G, NMDA, ACh, GluCl of the receptor ids; receptors whose activity
// bit is cleared, pb < 0, draw nothing); injv is the injection draw -- replay
// mode: the reference's Igauss value (nA) from its ziggurat stream;
// counter-based mode: a standard normal (Box-Muller on the second Philox
// block) that the update scales by the population's mean and deviation.
template <typename T, bool REPLAY>
__device__ __forceinline__ void neuron_draws(const Net<T> &net, PopC<T> &P, const NConst<T> &C,
                                             int32_t n, int64_t t, uint32_t &dbits, float &injv) {
  dbits = 0u;
  injv = 0.0f;
  const uint32_t act = P.act;
  if (REPLAY) {
    const int64_t s = t - net.ext_first;
    if (s >= 0 && s < net.ext_steps) {
#pragma unroll
      for (int r = 0; r < 5; ++r)
        if ((net.ext_mask & (1 << r)) && C.pb[r] >= 0.0f)
          dbits |= ((net.ext_bits[((size_t)s * 5 + r) * net.words + (n >> 5)] >> (n & 31)) & 1u) << r;
    }
    if (act & ACT_INJ) {
      // the reference's own ziggurat stream (single-thread order: neurons in
      // index order, only those with injection on draw)
      const uint32_t k = (uint32_t)t * P.zig_a + (uint32_t)n + P.zig_b;
      injv = zig_igauss((uint32_t)net.seed, k, (float)P.inj_mean, (float)P.inj_std, net.zig);
    }
    return;
  }
  if (!net.ext_mask && !(act & ACT_INJ)) return;
  // Philox counter (neuron within its instance, step), key C.key
  const uint64_t key = C.key;
  uint32_t rnd[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  uint32_t c[4] = {C.ctr, (uint32_t)t, (uint32_t)((uint64_t)t >> 32), 0u};
  philox10(c, (uint32_t)key, (uint32_t)(key >> 32));
  rnd[0] = c[0]; rnd[1] = c[1]; rnd[2] = c[2]; rnd[3] = c[3];
  if ((net.ext_mask & (1 << 4)) || (act & ACT_INJ)) {
    uint32_t e[4] = {C.ctr, (uint32_t)t, (uint32_t)((uint64_t)t >> 32), 1u};
    philox10(e, (uint32_t)key, (uint32_t)(key >> 32));
    rnd[4] = e[0]; rnd[5] = e[1]; rnd[6] = e[2]; rnd[7] = e[3];
  }
#pragma unroll
  for (int r = 0; r < 5; ++r)
    if ((net.ext_mask & (1 << r)) && C.pb[r] >= 0.0f) dbits |= (C.pb[r] > u24(rnd[r]) ? 1u : 0u) << r;
  if (act & ACT_INJ) {
    const float u1 = ((float)(rnd[6] >> 8) + 1.0f) * (1.0f / 16777216.0f);
    const float u2 = u24(rnd[7]);
    injv = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
  }
}

// One neuron's step t (the reference's Activation, then its presynaptic
// gates and plasticity) on state S: (1) linear conductances of the previous
// Spike phase, (2) external gates, (3) membrane, (4) gates, (4b) STP/STD.
// gn: NMDA conductance of the previous Spike phase; inc: ring_take's output;
// dbits / injv: neuron_draws' output for (n, t).  Writes the NMDA gate
// history slot hslot (step t); everything else stays in S.
// WIN: the windowed path's instance, which never runs with LTP or gap
// junctions (the host gives such networks no window), so their branches are
// compiled out and leave the registers to the window's state; it leaves the
// step's NMDA gate in S.sn for its caller to stage instead of storing it to
// the history.  PT: the parameter block's type (global constant memory or an
// LDS copy).
template <typename T, bool REPLAY, bool WIN = false, typename PT = PopC<T>>
__device__ __forceinline__ void neuron_core(const Net<T> &net, PT &P, const NConst<T> &C,
                                            int32_t n, int64_t t, int32_t hslot, T gn,
                                            const T inc[4], uint32_t dbits, float injv, NState<T> &S,
                                            bool &spk, T &a_push) {
  const int32_t N = net.N;
  uint32_t cs = S.cs;
  uint32_t cnt = cs & CNT_MASK;
  T v = S.v;
  // (1) G = df*G + slot
  T g[5] = {T(0), T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int li = 0; li < 4; ++li) {
    if (!(net.lin_mask & (1 << li))) continue;
    const int rid = lin_rid(li);
    const T gv = P.dfst[rid] * S.G[li] + inc[li];
    S.G[li] = gv;
    g[rid] = gv;
  }
  if (net.has_nmda_syn) g[2] = gn;
  if (!WIN && net.ltp_mode) {  // single synapses, summed by k_learn with the weights before its update
#pragma unroll
    for (int r = 0; r < 5; ++r) {
      T *gp = net.gl + (size_t)r * N + n;
      g[r] = g[r] + *gp;
      *gp = T(0);
    }
  }
  // (2) external Poisson gates, order A, G, ACh, GluCl, NMDA (BgInit)
  const uint32_t act = P.act;
  const T eg = C.eg;
#pragma unroll
  for (int li = 0; li < 4; ++li) {
    const int rid = lin_rid(li);
    if (!(net.ext_mask & (1 << rid))) continue;
    if (C.pb[rid] < 0.0f) continue;
    T e = S.elin[li];
    e = ((dbits >> rid) & 1u) ? e + T(1) : e * P.dfst[rid];
    S.elin[li] = e;
    g[rid] = g[rid] + P.gext[rid] * eg * e;
  }
  if (net.ext_mask & (1 << 2)) {
    if (C.pb[2] >= 0.0f) {
      T x = S.enx, s = S.ens;
      x = ((dbits >> 2) & 1u) ? x + T(1) : x * P.dfxt_n;
      s = P.dfst[2] * s + P.efas_n * x * (T(1) - s);
      S.enx = x;
      S.ens = s;
      g[2] = g[2] + P.gext[2] * eg * s;
    }
  }
  // Isyn: the gap current of the previous Spike phase, plus the Gaussian
  // injection drawn now (BgInit adds to Isyn, code/EqsGnl2.cpp:335)
  T isyn = (!WIN && net.igap) ? net.igap[n] : T(0);
  if (act & ACT_INJ) {
    if (REPLAY) isyn = isyn + T(injv);
    else isyn = isyn + (P.inj_mean + P.inj_std * T(injv));
  }
  // (3) membrane (VmEul / firing), as selects: the Euler step is evaluated
  // for every lane and kept where the neuron was not firing; firing() (the
  // refractory machine) applies where it was, or where the step crossed the
  // threshold.  The values are those of the branches of the reference.
  // Receptors with no conductance source in this network contribute -0 (the
  // reference's 0 * (E - V) is +-0; adding either leaves a sum unchanged).
  const bool f0 = (cs & F_FIRE) != 0;
  auto live = [&](int r) -> bool {
    if (!WIN && net.ltp_mode) return true;
    if (net.ext_mask & (1 << r)) return true;
    if (r == 2) return net.has_nmda_syn != 0;
    return (net.lin_mask & (1 << (r < 2 ? r : r - 1))) != 0;
  };
  const T i0 = live(0) ? g[0] * (P.vrev[0] - v) : T(-0.0);
  const T i1 = live(1) ? g[1] * (P.vrev[1] - v) : T(-0.0);
  T i2 = T(-0.0);
  if (live(2)) {
    if constexpr (sizeof(T) == 4) {
      // FP32 mode: hardware exp2 and reciprocal (about 1 ulp each) in place
      // of the correctly rounded expf and divisions
      const float ex = __expf(-0.062f * v);
      i2 = g[2] * (P.vrev[2] - v) * __builtin_amdgcn_rcpf(1.0f + P.mg * ex * (1.0f / 3.57f));
    } else {
      const T ex = exp(T(-0.062f) * v);
      i2 = g[2] * (P.vrev[2] - v) / (T(1) + P.mg * ex / T(3.57f));
    }
  }
  const T i3 = live(3) ? g[3] * (P.vrev[3] - v) : T(-0.0);
  const T i4 = live(4) ? g[4] * (P.vrev[4] - v) : T(-0.0);
  if (net.irep && !f0) {  // Nets[].Irep (code/EqsGnl2.cpp:196-200)
    T *ir = net.irep + n;
    ir[0] = g[0] * (P.vrev[0] - v);
    ir[(size_t)N] = g[1] * (P.vrev[1] - v);
    ir[2 * (size_t)N] = live(2) ? i2 : g[2] * (P.vrev[2] - v);
    ir[3 * (size_t)N] = g[3] * (P.vrev[3] - v);
    ir[4 * (size_t)N] = g[4] * (P.vrev[4] - v);
  }
  T I = i0 + i1 + i2 + i3 + i4 + isyn * T(1e3f) + P.leak_drive;
  if constexpr (sizeof(T) == 4) I = (1e-3f * I) * __builtin_amdgcn_rcpf(P.cm);
  else I = T(1e-3f) * I / P.cm;
  const bool ovf = I > P.dvmax, udf = !ovf && I < -P.dvmax;
  I = ovf ? P.dvmax : (udf ? -P.dvmax : I);
  const T vm = P.df * v + P.ef * I;
  const uint32_t csm = (cs & ~(F_OVF | F_UDF)) | (ovf ? F_OVF : 0u) | (udf ? F_UDF : 0u);
  const bool cross = !f0 && vm > P.vth;
  // firing() on the state it starts from: the old one if firing, else the
  // Euler step's with F_FIRE set
  const uint32_t csf = f0 ? cs : (csm | F_FIRE);
  const bool inref = (int32_t)cnt < P.rfped;
  const T vfire = inref ? (cnt == 0 ? T(0) : P.vreset) : (f0 ? v : vm);
  const uint32_t csfire = inref ? (((int32_t)cnt == P.spidy) ? (csf | F_SPIKE) : (csf & ~F_SPIKE)) : (csf & ~F_FIRE);
  const bool fire = f0 || cross;
  v = fire ? vfire : vm;
  cs = fire ? csfire : csm;
  cnt = fire ? (inref ? cnt + 1u : 0u) : cnt;
  if (!WIN && net.ltp_mode && cross) cs |= F_LPFIRE;
  spk = (cs & F_SPIKE) != 0;
  // (4) presynaptic gates (the reference's per-synapse gate rule, shared);
  // a spike adds 1, or D*F under STP and D under STD, the presynaptic
  // state before this step's plasticity update (code/EqsGnl2.cpp:505-863)
  T one = T(1);
  if (net.stp_mode) {
    one = S.sD;
    if (net.stp_mode == STP_FULL) one = one * S.sF;
  }
  a_push = T(0);
  if (P.has_lin_out) {
    const T s = S.spre;
    a_push = one + (T(1) - P.pre_df) * s;
    S.spre = spk ? s + one : s * P.pre_df;
  }
  if (P.has_nmda_out) {
    T x = S.xn;
    const T s0 = S.sn;
    x = spk ? x + one : x * P.pre_dfxt;
    S.xn = x;
    const T s1 = P.pre_dfst * s0 + P.pre_efas * x * (T(1) - s0);
    S.sn = s1;
    if (!WIN) net.shist[hidx(net, C.nl, C.rep, (uint32_t)hslot)] = s1;
  }
  // (4b) EqsGnl2::STP / STD (code/EqsGnl2.cpp:1048-1090), after the gates
  if (net.stp_mode == STP_FULL) {
    T d = S.sD, f = S.sF, cap = S.sCap, car = S.sCar;
    if (spk) {
      car = car + P.stp_aCar;
      cap = cap + P.stp_aCap;
      f = f + (cap + car) * (T(1) - f) * P.stp_aF;
      d = d * (T(1) - P.stp_pv * f);
    } else {
      car = P.stp_dfCar * car + P.stp_efCar;
      cap = P.stp_dfCap * cap;
      f = P.stp_dfF * f + P.stp_efF * (cap + car) * (T(1) - f) * P.stp_aF;
      d = P.stp_dfD * d + P.stp_efD;
    }
    S.sD = d; S.sF = f; S.sCap = cap; S.sCar = car;
  } else if (net.stp_mode == STP_STD) {
    const T d = S.sD;
    S.sD = spk ? d * (T(1) - P.stp_pv) : P.stp_dfD * d + P.stp_efD;
  }
  if (!WIN && net.ltp_mode) {
    // firLP: the LTP spike follows the onset after SpiLPDy steps
    // (code/EqsGnl2.cpp:254-272); ltpEul: LP trace (:274-280)
    uint32_t c = net.cntlp[n];
    if (cs & F_LPFIRE) {
      if ((int32_t)c == P.ltp_spidy) {
        cs |= F_LPSPIKE;
        cs &= ~F_LPFIRE;
        c = 0;
      } else {
        cs &= ~F_LPSPIKE;
        c++;
      }
    } else {
      cs &= ~F_LPSPIKE;
    }
    net.cntlp[n] = c;
    net.lp[n] = (cs & F_LPSPIKE) ? net.lp[n] + T(1) : P.ltp_dfLP * net.lp[n];
  }
  S.v = v;
  S.cs = (cs & ~CNT_MASK) | (cnt & CNT_MASK);
}

// Local neuron, replica and population of global neuron n.
template <bool UNI>
__device__ __forceinline__ void locate(const uint32_t *npop, int32_t B, int32_t n, int32_t &nl,
                                       int32_t &rep, uint32_t &p0) {
  nl = n / B;
  if constexpr (UNI) nl = __builtin_amdgcn_readfirstlane(nl);
  rep = n - nl * B;
  p0 = ldc(npop + nl);
  if constexpr (UNI) p0 = __builtin_amdgcn_readfirstlane(p0);
}

// One step of neuron n from and back to memory (n >= N: inactive lane,
// spk = false); gn_given: the NMDA conductance comes from the caller (fused
// step) instead of the gn array written by the separate gather.
template <typename T, bool REPLAY, bool UNI>
__device__ __forceinline__ void neuron_step_mem(const Net<T> &net, int32_t n, int64_t t, int32_t k,
                                                bool gn_given, T gn_in, bool &spk, T &a_push,
                                                uint32_t &pid) {
  spk = false;
  a_push = T(0);
  pid = 0;  // population x replica index of the spike counters
  if (n >= net.N) return;
  int32_t nl, rep;
  uint32_t p0;
  locate<UNI>(net.npop, net.B, n, nl, rep, p0);
  pid = p0 * (uint32_t)net.B + (uint32_t)rep;
  PopC<T> &P = *(PopC<T> *)(net.pops + p0);
  NState<T> S;
  load_state(net, nl, rep, n, slot_of(net.tmod[1], k - 1, net.DN), S);
  NConst<T> C;
  load_const(net, P, nl, rep, p0, C);
  T inc[4];
  ring_take(net, n, slot_of(net.tmod[0], k - 1, net.D), inc);
  const T gn = net.has_nmda_syn ? (gn_given ? gn_in : net.gn[n]) : T(0);
  uint32_t dbits;
  float injv;
  neuron_draws<T, REPLAY>(net, P, C, n, t, dbits, injv);
  neuron_core<T, REPLAY>(net, P, C, n, t, slot_of(net.tmod[1], k, net.DN), gn, inc, dbits, injv, S, spk,
                         a_push);
  store_state(net, n, S);
}

// One neuron's step t (the reference's Activation) followed by the wave's
// spike compaction; every thread of the block must reach it (ballots, block
// barriers); list_ctr[NSUB] and list_base are the sub-list counters and the
// offset of this step's list within sub-list `sub` (the update block's index
// modulo NSUB).  With gn_given the NMDA conductance comes from the caller
// (fused step) instead of the gn array written by the separate gather.
// UNI: the caller guarantees that all 64 lanes of the wave hold the same
// local neuron (replicas in whole waves); its index, population and
// parameters are then wave-uniform and come through scalar loads.
template <typename T, bool REPLAY, bool UNI>
__device__ __forceinline__ void update_neuron(const Net<T> &net, int32_t n, int64_t t, int32_t k,
                                              bool gn_given, T gn_in, uint32_t *list_ctr,
                                              uint32_t list_base, int sub) {
  bool spk;
  T a_push;
  uint32_t pid;
  neuron_step_mem<T, REPLAY, UNI>(net, n, t, k, gn_given, gn_in, spk, a_push, pid);
  // (5) spike-list compaction: wave ballot, then one block-level LDS count,
  // so one global atomic per block claims the block's range of the list
  // (a global atomic per wave serialises on one address at high rates)
  __shared__ uint32_t s_cnt, s_base;
  if (threadIdx.x == 0) s_cnt = 0u;
  __syncthreads();
  const uint64_t mask = __ballot(spk);
  const int lane = (int)__lane_id();
  const int leader = mask ? __ffsll((unsigned long long)mask) - 1 : 0;
  const uint32_t cntw = (uint32_t)__popcll(mask);
  uint32_t woff = 0u;
  if (mask && lane == leader) woff = atomicAdd(&s_cnt, cntw);
  woff = __shfl(woff, leader);
  __syncthreads();
  if (threadIdx.x == 0 && s_cnt) {
    s_base = list_base + atomicAdd(list_ctr + sub, s_cnt);
    atomicAdd(&net.counters[blockIdx.x & (NCTR - 1)], (unsigned long long)s_cnt);
  }
  __syncthreads();
  if (spk) {
    const uint32_t pos = s_base + woff + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (pos < net.sub_cap[sub]) {
      net.rec_ids[net.sub_off[sub] + pos] = (uint32_t)n;
      net.rec_a[net.sub_off[sub] + pos] = a_push;
    } else {
      atomicOr(net.overflow, 1u);
    }
    // per-population spike count, aggregated when the wave shares one population
    const uint32_t p = pid;
    const uint32_t p0 = __shfl(p, leader);
    const bool same = __ballot(spk && p != p0) == 0ull;
    if (same) {
      if (lane == leader) atomicAdd(&net.pop_cnt[p0], (unsigned long long)cntw);
    } else {
      atomicAdd(&net.pop_cnt[p], 1ull);
    }
  }
}

template <typename T, bool REPLAY, bool UNI>
__global__ void __launch_bounds__(1024) k_update(Net<T> net, int32_t k) {
  update_neuron<T, REPLAY, UNI>(net, (int32_t)(blockIdx.x * blockDim.x + threadIdx.x), *net.t0 + k, k,
                                false, T(0), net.total, 0u, (int)(blockIdx.x % NSUB));
}

// ------------------------------------------------------------------ delivery
template <typename T>
__device__ __forceinline__ void atomic_add_t(T *p, T v) {
  unsafeAtomicAdd(p, v);
}

// Push the spikes of one step (emitted at step ts, ring slot s0) into the
// ring slots of their synapses' delays; one wave per spike.  range(s, a, b)
// gives the step's spikes [a, b) within sub-list s; the waves are spread over
// the NSUB sub-lists (nwaves / NSUB waves each, or several sub-lists per wave
// when there are fewer waves than sub-lists).  Lane 0 of each wave
// accumulates the event counters.
template <typename T, typename Range>
__device__ __forceinline__ void push_spikes(const Net<T> &net, Range range, int32_t s0, int gwave,
                                            int nwaves, int lane, unsigned long long &events,
                                            unsigned long long &nout) {
  const int32_t N = net.N, B = net.B;
  // s0: ring slot of the emission step; delay d lands in s0 + d wrapped once
  // (d <= maxd < D)
  auto one = [&](int s, uint32_t first, uint32_t stride) {
    uint32_t a = 0, b = 0;
    range(s, a, b);
    const uint32_t off = net.sub_off[s];
    for (uint32_t i = a + first; i < b; i += stride) {
      // i is the same in every lane, so the spike and its row are wave-uniform
      const uint32_t j = (uint32_t)__builtin_amdgcn_readfirstlane((int)net.rec_ids[off + i]);
      const uint32_t jl = j / (uint32_t)B, rep = j - jl * (uint32_t)B;
      const T amp = net.rec_a[off + i];
      const T ga = amp * T(net.gainr[0 * B + rep]), gg = amp * T(net.gainr[1 * B + rep]),
              gh = amp * T(net.gainr[3 * B + rep]), gl = amp * T(net.gainr[4 * B + rep]);
      const uint64_t r0 = ldc(net.rp_p + jl), r1 = ldc(net.rp_p + jl + 1);
      if (lane == 0) { events += r1 - r0; nout += ldc(net.nmda_out + jl); }
      for (uint64_t e = r0 + (uint64_t)lane; e < r1; e += 64) {
        const uint32_t m = ldc(net.meta_p + e);
        const uint32_t col = (m & 0x00FFFFFFu) * (uint32_t)B + rep;
        const uint32_t d = (m >> 24) & 63u;
        const uint32_t li = m >> 30;
        const T sc = li == 0 ? ga : li == 1 ? gg : li == 2 ? gh : gl;
        int32_t slot = s0 + (int32_t)d;
        slot -= slot >= net.D ? net.D : 0;
        atomic_add_t(net.ring + ((size_t)li * net.D + slot) * N + col, T(ldc(net.w_p + e)) * sc);
      }
    }
  };
  // Replicas in whole waves: spikes recorded by one wave are copies of one
  // local neuron, contiguous in the record.  A wave takes 64 record entries
  // at a time, and the lanes holding copies of the same neuron walk its row
  // together: each atomic instruction then covers up to 64 consecutive
  // copies of one target (a few cache lines) instead of one copy of up to 64
  // targets (a line each, shared with the other copies' waves).  Groups of
  // fewer than 8 copies keep the one-wave-per-spike walk.
  auto grouped = [&](int s, uint32_t first, uint32_t stride) {
    uint32_t a = 0, b = 0;
    range(s, a, b);
    const uint32_t off = net.sub_off[s];
    for (uint32_t c = a + 64u * first; c < b; c += 64u * stride) {
      const uint32_t i = c + (uint32_t)lane;
      bool act = i < b;
      uint32_t jl = 0, rep = 0;
      T amp = T(0);
      if (act) {
        const uint32_t j = net.rec_ids[off + i];
        jl = j / (uint32_t)B;
        rep = j - jl * (uint32_t)B;
        amp = net.rec_a[off + i];
      }
      uint64_t left = __ballot(act);
      while (left) {
        const int ld = __ffsll((unsigned long long)left) - 1;
        const uint32_t J = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)jl, ld));
        const bool mine = act && jl == J;
        const uint64_t mm = __ballot(mine);
        left &= ~mm;
        act = act && !mine;
        const uint32_t cnt = (uint32_t)__popcll(mm);
        const uint64_t r0 = ldc(net.rp_p + J), r1 = ldc(net.rp_p + J + 1);
        if (lane == 0) { events += (r1 - r0) * cnt; nout += (uint64_t)ldc(net.nmda_out + J) * cnt; }
        if (cnt >= 8) {
          const T sc[4] = {amp * T(net.gainr[0 * B + rep]), amp * T(net.gainr[1 * B + rep]),
                           amp * T(net.gainr[3 * B + rep]), amp * T(net.gainr[4 * B + rep])};
          for (uint64_t e = r0; e < r1; ++e) {
            const uint32_t m = ldc(net.meta_p + e);
            const float w = ldc(net.w_p + e);
            const uint32_t li = m >> 30;
            int32_t slot = s0 + (int32_t)((m >> 24) & 63u);
            slot -= slot >= net.D ? net.D : 0;
            if (mine)
              atomic_add_t(net.ring + ((size_t)li * net.D + slot) * N +
                               (size_t)(m & 0x00FFFFFFu) * B + rep,
                           T(w) * sc[li]);
          }
        } else {
          uint64_t g = mm;
          while (g) {
            const int q = __ffsll((unsigned long long)g) - 1;
            g &= g - 1;
            const uint32_t rq = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)rep, q));
            const T aq = __shfl(amp, q);
            const T sq[4] = {aq * T(net.gainr[0 * B + rq]), aq * T(net.gainr[1 * B + rq]),
                             aq * T(net.gainr[3 * B + rq]), aq * T(net.gainr[4 * B + rq])};
            for (uint64_t e = r0 + (uint64_t)lane; e < r1; e += 64) {
              const uint32_t m = ldc(net.meta_p + e);
              const uint32_t li = m >> 30;
              int32_t slot = s0 + (int32_t)((m >> 24) & 63u);
              slot -= slot >= net.D ? net.D : 0;
              atomic_add_t(net.ring + ((size_t)li * net.D + slot) * N +
                               (size_t)(m & 0x00FFFFFFu) * B + rq,
                           T(ldc(net.w_p + e)) * sq[li]);
            }
          }
        }
      }
    }
  };
  const bool wave_reps = (B & 63) == 0;
  if (nwaves >= NSUB) {
    const int per = nwaves / NSUB;
    if (gwave < per * NSUB) {
      if (wave_reps) grouped(gwave % NSUB, (uint32_t)(gwave / NSUB), (uint32_t)per);
      else one(gwave % NSUB, (uint32_t)(gwave / NSUB), (uint32_t)per);
    }
  } else {
    for (int s = gwave; s < NSUB; s += nwaves) {
      if (wave_reps) grouped(s, 0u, 1u);
      else one(s, 0u, 1u);
    }
  }
}

// Block-level reduction of the push counters (synaptic events, NMDA
// out-synapses of the spikers): one global atomic per block that pushed.
template <typename T>
__device__ __forceinline__ void block_add_events(const Net<T> &net, unsigned long long events,
                                                 unsigned long long nout) {
  __shared__ unsigned long long s_ev[2];
  if (threadIdx.x == 0) { s_ev[0] = 0; s_ev[1] = 0; }
  lds_barrier();
  if (events) atomicAdd(&s_ev[0], events);
  if (nout) atomicAdd(&s_ev[1], nout);
  lds_barrier();
  if (threadIdx.x == 0) {
    const int slot = (int)(blockIdx.x & (NCTR - 1));
    if (s_ev[0]) atomicAdd(&net.counters[NCTR + slot], s_ev[0]);
    if (s_ev[1]) atomicAdd(&net.counters[2 * NCTR + slot], s_ev[1]);
  }
}

// NMDA conductance of one row for one step: sum_j w * s_j(ts - d) over
// CSR-by-post row p.  L lanes of a group stride the row (lane sl takes entries
// r0 + sl + i*L); each lane sums its entries in entry order with fused
// multiply-adds, the sum is scaled by the replica's NMDA gain, and a shuffle
// reduction over the group follows (result valid in every lane of the group).
// gather_window evaluates the same expression for several steps at once, so
// every path gives the same bits.  Entries are fetched four at a time ahead
// of their sums.  All lanes of the group must call it.  h0: history slot of
// ts.  UNI: all lanes of the wave gather copies of one local row (L == 1,
// replicas in whole waves), so the row's columns and weights are scalar loads.
template <typename T, bool UNI>
__device__ __forceinline__ T gather_row(const Net<T> &net, int64_t p, int32_t h0, int L, int sl) {
  T acc = T(0);
  if (p < net.N) {
    const uint32_t B = (uint32_t)net.B, dmask = (uint32_t)net.DN - 1u;
    uint32_t pl = (uint32_t)p / B;
    if constexpr (UNI) pl = __builtin_amdgcn_readfirstlane(pl);
    const uint32_t rep = (uint32_t)p - pl * B;
    const uint64_t r0 = ldc(net.rp_n + pl), r1 = ldc(net.rp_n + pl + 1);
    const size_t rs = (size_t)net.DN * B;  // history of one presynaptic neuron
    const T *sh = net.shist + rep;
    auto gate = [&](uint32_t m) -> T {
      return sh[(size_t)(m & 0x00FFFFFFu) * rs + (size_t)(((uint32_t)h0 - (m >> 24)) & dmask) * B];
    };
    const uint64_t step = (uint64_t)L;
    uint64_t e = r0 + (uint64_t)sl;
    for (; e + 3 * step < r1; e += 4 * step) {
      const uint32_t m0 = ldc(net.meta_n + e), m1 = ldc(net.meta_n + e + step),
                     m2 = ldc(net.meta_n + e + 2 * step), m3 = ldc(net.meta_n + e + 3 * step);
      const float w0 = ldc(net.w_n + e), w1 = ldc(net.w_n + e + step), w2 = ldc(net.w_n + e + 2 * step),
                  w3 = ldc(net.w_n + e + 3 * step);
      const T g0 = gate(m0), g1 = gate(m1), g2 = gate(m2), g3 = gate(m3);
      acc = fma(T(w0), g0, acc);
      acc = fma(T(w1), g1, acc);
      acc = fma(T(w2), g2, acc);
      acc = fma(T(w3), g3, acc);
    }
    for (; e < r1; e += step) acc = fma(T(ldc(net.w_n + e)), gate(ldc(net.meta_n + e)), acc);
    acc *= T(net.gainr[2 * B + rep]);
  }
  for (int off = L >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
  return acc;
}

// One fused launch per step, valid when every synapse has delay >= 1 step:
// (a) push the spikes of step t-1 (they land in ring slots >= t), (b) gather
// the NMDA conductance of step t-1 (gates of steps <= t-2) with a group of
// L = gather_lanes lanes per neuron, (c) the group's first lane updates the
// neuron for step t and appends its spike to the list of step k.  The per-step
// lists rotate over three counters (k%3): launch k reads the final count of
// list k-1, appends to list k (zeroed by launch k-1) and zeroes list k+1, so no
// block waits for another within a launch.  The first push_blocks blocks do
// (a), the others (b) and (c) with N*L threads in all, so small networks still
// present L times more waves to hide memory latency and the push runs beside
// the gather instead of ahead of it.
template <typename T, bool REPLAY, bool UNI>
__global__ void __launch_bounds__(1024) k_step(Net<T> net, int32_t k) {
  const int64_t t = *net.t0 + k;
  const int pm = (k + 2) % 3, pc = k % 3, pn = (k + 1) % 3;
  const uint32_t *pbase = net.lbase + pm * NSUB, *pcnt = net.lcnt + pm * NSUB;
  if (blockIdx.x == 0 && threadIdx.x < NSUB) {
    const int s = (int)threadIdx.x;
    const uint32_t base = pbase[s] + pcnt[s];
    net.lbase[pc * NSUB + s] = base;
    net.lcnt[pn * NSUB + s] = 0u;
    net.cum[(size_t)k * NSUB + s] = base < net.sub_cap[s] ? base : net.sub_cap[s];
  }
  const int wave = (int)(threadIdx.x >> 6);
  const int lane = (int)(threadIdx.x & 63);
  const int wpb = (int)(blockDim.x >> 6);
  if ((int)blockIdx.x < net.push_blocks) {
    // leading blocks: push the spikes of step t-1, in parallel with the
    // gather + update blocks (dispatched first, so they never trail them)
    unsigned long long my_events = 0, my_nout = 0;
    auto prev = [&](int s, uint32_t &a, uint32_t &b) {
      const uint32_t cap = net.sub_cap[s], x = pbase[s], y = pbase[s] + pcnt[s];
      a = x < cap ? x : cap;
      b = y < cap ? y : cap;
    };
    push_spikes(net, prev, slot_of(net.tmod[0], k - 1, net.D), (int)blockIdx.x * wpb + wave,
                net.push_blocks * wpb, lane, my_events, my_nout);
    block_add_events(net, my_events, my_nout);
    return;
  }
  const int ub = (int)blockIdx.x - net.push_blocks;  // gather + update block index
  const int L = net.gather_lanes;
  const int64_t gt = (int64_t)ub * blockDim.x + threadIdx.x;
  const int64_t nn = gt / L;
  const int sl = (int)(gt % L);
  T gn = T(0);
  if (net.has_nmda_syn) gn = gather_row<T, UNI>(net, nn, slot_of(net.tmod[1], k - 1, net.DN), L, sl);
  int32_t n = (sl == 0 && nn < net.N) ? (int32_t)nn : net.N;  // N: inactive lane
  if (L > 1) {
    // the block's blockDim/L neurons are updated by its first lanes, so the
    // waves past them skip the update instead of running it for a few lanes
    __shared__ T s_gn[1024];
    if (sl == 0) s_gn[threadIdx.x / L] = gn;
    __syncthreads();
    const int npb = (int)blockDim.x / L;
    const int64_t nu = (int64_t)ub * npb + threadIdx.x;
    const bool up = (int)threadIdx.x < npb && nu < net.N;
    n = up ? (int32_t)nu : net.N;
    gn = up ? s_gn[threadIdx.x] : T(0);
  }
  const int sub = ub % NSUB;
  update_neuron<T, REPLAY, UNI>(net, n, t, k, true, gn, net.lcnt + pc * NSUB,
                                pbase[sub] + pcnt[sub], sub);
}

// NMDA conductances of post row p for the W steps k0 .. k0+W-1 of a window,
// into s_gn[s * npb + i] (i: the row's neuron within the block).  The L lanes
// of the row are TW = min(L, 16) time lanes times L / TW entry sub-lanes:
// sub-lane q takes the row's entries q, q + L/TW, ... and time lane tl the
// steps tl, tl + TW, ... (four per pass).  A term's gates of consecutive
// steps are consecutive history slots of one presynaptic neuron, so the time
// lanes of a row read one or two cache lines per term together, and its
// column and weight are one broadcast load.  Each step's terms are summed in
// entry order per sub-lane with fused multiply-adds, scaled by the replica's
// NMDA gain and reduced over the sub-lanes by a shuffle tree (gather_row's
// expression with L / TW lanes).  Four entries are in flight per lane
// hb: history slot of step k0 - 1.  All lanes of the row must call it.
template <typename T, bool UNI>
__device__ __forceinline__ void gather_window(const Net<T> &net, int64_t p, int32_t hb, int32_t W, int L,
                                              int sl, T *s_gn, int npb, int i) {
  constexpr int NA = 4;  // steps per lane per pass
  constexpr int U = 4;   // entries per round
  const uint32_t B = (uint32_t)net.B, dmask = (uint32_t)net.DN - 1u;
  const int TW = L < 16 ? L : 16;
  const int G = L / TW;
  const int q = sl / TW, tl = sl - q * TW;
  const bool live = p < net.N;
  uint64_t r0 = 0, r1 = 0;
  uint32_t rep = 0;
  T gain = T(0);
  if (live) {
    uint32_t pl = (uint32_t)p / B;
    if constexpr (UNI) pl = __builtin_amdgcn_readfirstlane(pl);
    rep = (uint32_t)p - pl * B;
    r0 = ldc(net.rp_n + pl);
    r1 = ldc(net.rp_n + pl + 1);
    gain = T(net.gainr[2 * B + rep]);
  }
  const size_t rs = (size_t)net.DN * B;
  const T *sh = net.shist + rep;
  const uint64_t st = (uint64_t)G;
  for (int32_t c0 = 0; c0 < W; c0 += TW * NA) {
    // this lane's steps: c0 + tl + a * TW, a < na
    int na = 0;
#pragma unroll
    for (int a = 0; a < NA; ++a) na += (c0 + tl + a * TW < W) ? 1 : 0;
    const uint32_t sb = (uint32_t)(hb + c0 + tl);  // slot of step k0 + c0 + tl - 1 (before the delay)
    T acc[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] = T(0);
    uint64_t e = r0 + (uint64_t)q;
    uint32_t m[U];
    float w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t eu = e + (uint64_t)u * st;
      m[u] = eu < r1 ? ldc(net.meta_n + eu) : 0u;
      w[u] = eu < r1 ? ldc(net.w_n + eu) : 0.0f;
    }
    while (e < r1) {
      // the gates of this round's entries, then the next round's entries
      T x[U][NA];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool ok = e + (uint64_t)u * st < r1;
        const T *g = sh + (size_t)(m[u] & 0x00FFFFFFu) * rs;
        const uint32_t bs = sb - (m[u] >> 24);
#pragma unroll
        for (int a = 0; a < NA; ++a)
          x[u][a] = (ok && a < na) ? g[(size_t)((bs + (uint32_t)(a * TW)) & dmask) * B] : T(0);
      }
      const uint64_t en = e + (uint64_t)U * st;
      uint32_t mn[U];
      float wn[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint64_t eu = en + (uint64_t)u * st;
        mn[u] = eu < r1 ? ldc(net.meta_n + eu) : 0u;
        wn[u] = eu < r1 ? ldc(net.w_n + eu) : 0.0f;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (e + (uint64_t)u * st < r1) {
          const T tw = T(w[u]);
#pragma unroll
          for (int a = 0; a < NA; ++a) acc[a] = fma(tw, x[u][a], acc[a]);
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        m[u] = mn[u];
        w[u] = wn[u];
      }
      e = en;
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      T v = acc[a] * gain;
      for (int off = L >> 1; off >= TW; off >>= 1) v += __shfl_xor(v, off, L);
      if (q == 0 && a < na && live) s_gn[(c0 + tl + a * TW) * npb + i] = v;
    }
  }
}

// Copy of 4 bytes per lane from global memory straight into LDS (the LDS
// DMA of the vector memory unit): lane i of the wave writes lds_wave[i], so
// lds_wave must be the same for every lane of the wave.  Completion is
// counted by vmcnt, like a load.
__device__ __forceinline__ void lds_dma4(const void *src, void *lds_wave) {
  __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)src,
                                   (void __attribute__((address_space(3))) *)lds_wave, 4, 0, 0);
}

// Waits until this wave's memory operations, LDS DMA included, are done,
// then for the whole block.
__device__ __forceinline__ void dma_barrier() {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
}

// Most steps a window of gather_block may span: each thread keeps one
// partial sum per step of the window in registers.
constexpr int FLAT_WMAX = 16;

// x / W for a window length W <= 64 as __umulhi(x, kDivW.m[W]), exact for
// x < 2^26 (m = floor(2^32 / W) + 1; W = 1 needs x itself): the kernels'
// divisions by W become one table read.
struct DivWTable {
  uint32_t m[65];
  constexpr DivWTable() : m() {
    for (int w = 2; w <= 64; ++w) m[w] = 0xFFFFFFFFu / (uint32_t)w + 1u;
  }
};
__device__ constexpr DivWTable kDivW{};

// NMDA sums of the block's rows nb .. nb+nbc-1 (one network copy) for the W
// steps of a window, into s_gn[s * npb + r].  The block's entries are
// consecutive in the gather CSR; they are taken in chunks of CH: (1) their
// columns and weights are copied into LDS; (2) every thread copies (entry,
// step) gates of the chunk, entry-major, into LDS -- the W gates of one
// entry are W consecutive history slots of its presynaptic neuron, so
// consecutive lanes read consecutive words -- as LDS DMA, which takes no
// registers, so each wave keeps all its copies of the chunk in flight; (3)
// L = gather_split consecutive threads (a power of two, at most blockDim /
// npb) share each (row, step) pair: sub-lane q adds the row's entries q,
// q + L, ... (counted from the row's start) of the chunk in entry order with
// fused multiply-adds, carrying its partial sum across chunks in a register.
// At the end each partial sum is scaled by the NMDA gain and the L of them
// are reduced by a shuffle tree: gather_row's expression with L lanes, so
// the two-launch and fused paths with gather_lanes = L give the same bits,
// and a block's time follows its entry count rather than its longest row,
// long rows being split L ways (short rows keep L = 1).  FP64
// gates travel as their low and high words (s_x: [2][CH * W] words).  hb:
// history slot of step k0 - 1; W <= FLAT_WMAX.  All threads must call it.
// ov() runs once, while the first chunk's gates are in flight (work that
// needs nothing loaded after the columns' barrier, and no staging array).
template <typename T, typename Ov>
__device__ __forceinline__ void gather_block(const Net<T> &net, int32_t nb, int nbc, uint64_t e0, uint64_t e1,
                                             int32_t hb, int32_t W, T *s_gn, int npb, uint64_t *s_rpn,
                                             uint32_t *s_m, float *s_w, uint32_t *s_x, int CH, Ov &&ov) {
  const uint32_t tid = threadIdx.x, bd = blockDim.x, wb = tid & ~63u;
  // L sub-lanes per (row, step) pair: item f = pair * L + sub-lane, thread
  // tid holds items tid, tid + bd, ... (at most W of them, L <= bd / npb)
  const uint32_t L = (uint32_t)net.gather_split;
  const uint32_t lgL = (uint32_t)__builtin_ctz(L);  // L is a power of two
  const uint32_t q = tid & (L - 1u);
  const uint32_t n_items = (uint32_t)npb * (uint32_t)W * L;
  const uint32_t nj = (bd & (bd - 1u)) == 0 ? (n_items + bd - 1u) >> __builtin_ctz(bd)
                                            : (n_items + bd - 1u) / bd;
  const uint32_t DN = (uint32_t)net.DN, dmask = DN - 1u;
  // x / W as umulhi(x, mw), exact for x < 2^26 (W == 1: x itself)
  const uint32_t mw = kDivW.m[W];
  auto divw = [&](uint32_t x) -> uint32_t { return W == 1 ? x : __umulhi(x, mw); };
  T acc[FLAT_WMAX];
#pragma unroll
  for (int i = 0; i < FLAT_WMAX; ++i) acc[i] = T(0);
  const uint32_t *sh = (const uint32_t *)net.shist;
  constexpr uint32_t TW = sizeof(T) / 4;  // words per gate
  auto gate = [&](uint32_t f) -> T {
    if constexpr (TW == 2) return __hiloint2double((int)s_x[(size_t)CH * W + f], (int)s_x[f]);
    else return __uint_as_float(s_x[f]);
  };
  for (uint64_t c0 = e0; c0 < e1; c0 += (uint64_t)CH) {
    const uint32_t nc = (uint32_t)(e1 - c0 < (uint64_t)CH ? e1 - c0 : (uint64_t)CH);
    const uint32_t nf = nc * (uint32_t)W;
    for (uint32_t e = tid, ew = wb; ew < nc; e += bd, ew += bd) {
      if (e < nc) {
        lds_dma4(net.meta_n + c0 + e, s_m + ew);
        lds_dma4(net.w_n + c0 + e, s_w + ew);
      }
    }
    if (c0 == e0)  // the rows' starts, fetched beside the first chunk
      for (uint32_t r = tid; r <= (uint32_t)nbc; r += bd) s_rpn[r] = ldc(net.rp_n + nb + r);
    dma_barrier();
    for (uint32_t f = tid, fw = wb; fw < nf; f += bd, fw += bd) {
      if (f < nf) {
        const uint32_t e = divw(f), s = f - e * (uint32_t)W;
        const uint32_t m = s_m[e];
        const uint32_t *g = sh + ((size_t)(m & 0x00FFFFFFu) * DN + (((uint32_t)hb + s - (m >> 24)) & dmask)) * TW;
        lds_dma4(g, s_x + fw);
        if constexpr (TW == 2) lds_dma4(g + 1, s_x + (size_t)CH * W + fw);
      }
    }
    if (c0 == e0) ov();
    dma_barrier();
#pragma unroll
    for (int i = 0; i < FLAT_WMAX; ++i) {
      if ((uint32_t)i >= nj) break;
      const uint32_t f = tid + (uint32_t)i * bd;
      const uint32_t pr = f >> lgL;  // (row, step) pair
      const uint32_t r = divw(pr), s = pr - r * (uint32_t)W;
      if (f >= n_items || r >= (uint32_t)nbc) continue;
      const uint64_t a = s_rpn[r], b = s_rpn[r + 1];
      const uint64_t lo = a > c0 ? a : c0, hi = b < c0 + nc ? b : c0 + nc;  // the row's part of the chunk
      if (lo >= hi) continue;
      // this sub-lane's first entry of the part: (e - a) % L == q
      uint32_t e = (uint32_t)(lo - c0) + ((q - (uint32_t)(lo - a)) & (L - 1u));
      const uint32_t eh = (uint32_t)(hi - c0);
      T v = acc[i];
      // four entries' LDS reads ahead of their (ordered) multiply-adds
      for (; e + 3u * L < eh; e += 4u * L) {
        const float w0 = s_w[e], w1 = s_w[e + L], w2 = s_w[e + 2u * L], w3 = s_w[e + 3u * L];
        const T x0 = gate(e * (uint32_t)W + s), x1 = gate((e + L) * (uint32_t)W + s),
                x2 = gate((e + 2u * L) * (uint32_t)W + s), x3 = gate((e + 3u * L) * (uint32_t)W + s);
        v = fma(T(w0), x0, v);
        v = fma(T(w1), x1, v);
        v = fma(T(w2), x2, v);
        v = fma(T(w3), x3, v);
      }
      for (; e < eh; e += L) v = fma(T(s_w[e]), gate(e * (uint32_t)W + s), v);
      acc[i] = v;
    }
    lds_barrier();
  }
  if (e1 <= e0) ov();  // no entries: no chunk ran it
  const T gain = T(net.gainr[2]);
#pragma unroll
  for (int i = 0; i < FLAT_WMAX; ++i) {
    if ((uint32_t)i >= nj) break;
    T v = acc[i] * gain;
    for (uint32_t off = L >> 1; off > 0; off >>= 1) v += __shfl_xor(v, (int)off, (int)L);
    const uint32_t f = tid + (uint32_t)i * bd;
    const uint32_t pr = f >> lgL;
    const uint32_t r = divw(pr), s = pr - r * (uint32_t)W;
    if (q == 0 && f < n_items && r < (uint32_t)nbc) s_gn[s * npb + r] = v;
  }
}

// Entries per chunk of gather_block: W gates of each staged in LDS, 32 KiB,
// or cap entries when 0 < cap is smaller (smaller chunks take less LDS, so
// more blocks share a compute unit and hide each other's staging latency).
template <typename T>
__host__ __device__ constexpr int window_chunk(int W, int cap) {
  const int full = (int)(32768 / sizeof(T)) / W < 1024 ? (int)(32768 / sizeof(T)) / W : 1024;
  return cap > 0 && cap < full ? cap : full;
}

// Bytes of dynamic LDS k_window takes for npb neurons per block, W steps,
// nlu linear receptors in use and one network copy (flat, gather_block's
// staging) or several (the carve-up at the top of k_window): the arrays of
// phases A and B, then a region phase A uses for the gather's staging and
// phases B and C for the spike staging.
template <typename T>
__host__ __device__ constexpr size_t window_lds_bytes(int npb, int W, int nlu, bool flat, int chunk_cap,
                                                     bool inj) {
  const size_t nst = (size_t)((npb + 63) / 64) * 64 * W;
  const size_t fixed = ((size_t)8 * npb + sizeof(T) * ((size_t)W * npb * (1 + nlu)) + (size_t)4 * npb +
                        (inj ? (size_t)4 * W * npb : 0) + (size_t)W * npb + 7) & ~(size_t)7;
  const size_t bc = sizeof(T) * ((size_t)W * npb + nst) + (size_t)4 * nst;
  const int ch = window_chunk<T>(W, chunk_cap);
  const size_t stage = flat ? (size_t)8 * (npb + 1) + sizeof(T) * (size_t)ch * W + (size_t)8 * ch : 0;
  return fixed + (bc > stage ? bc : stage);
}

// Temporal blocking over the smallest synaptic delay: one launch runs the W
// consecutive steps k0 .. k0+W-1 of the batch, W <= 1 + the smallest delay of
// any synapse.  A spike of step u >= k0 lands in ring slot u + d >= k0 + W - 1,
// read only by updates of steps >= k0 + W, and its NMDA gate enters gathers of
// steps >= u + d + 1 >= k0 + W: every input a window reads was written before
// the launch.  So a block runs three phases, exchanging data through LDS:
//   A  prefetch, all threads: the NMDA sums of all W steps (gather_window, L
//      lanes per row), the W ring slots of each linear receptor, and the
//      external event outcomes and injection draws of all W steps
//      (neuron_draws);
//   B  update, one lane per neuron: W steps of neuron_core from registers and
//      LDS; the population parameters come through the scalar cache (each
//      wave runs the update once per distinct population among its lanes,
//      usually one), so the loop issues no vector memory access and never
//      waits on one; each wave stages its spikes (neuron, step, push
//      amplitude) and the step's NMDA gates in LDS;
//   C  push, all waves: one wave per staged spike walks its CSR-by-pre row
//      into the ring slots of the spike's step plus each synapse's delay; the
//      ring slots read in A are cleared, the NMDA gates go to the history, and
//      the spike record (when readback is on) and counters are written.
// The barriers between the phases wait for LDS only (lds_barrier): global
// stores and atomics stay in flight until the kernel ends.  The ring
// (D >= W + max delay + 1 slots) and the gate history (DN >= W + max NMDA
// delay + 1) are deep enough that no write of the launch lands in a slot some
// block still has to read.  UNI: replicas in whole waves with L == 1 (the
// row's tables are scalar loads).
// PROF 1: the network's receptor layout is AMPA and GABA linear (lin_mask 3),
// external drive on AMPA only (ext_mask 1) and no short-term plasticity; the
// instance folds those fields to constants, dropping the other receptors'
// code from the update loop.  PROF 0: any layout.
// DUP: every wave issues the prologue's loads (waves without a neuron of
// their own load the block's first neuron's values, unused) with no branch
// around them, for blocks of few waves per neuron group (gather_lanes <= 8);
// else only the waves that use them, behind a wave-uniform branch (which
// the compiler schedules less well, but which spares blocks of many lanes
// per neuron dozens of duplicate loads).
template <typename T, bool REPLAY, bool UNI, int PROF, bool DUP>
__global__ void __launch_bounds__(512) k_window(Net<T> net_, int32_t k0, int32_t W) {
  Net<T> net = net_;
  if constexpr (PROF == 1) {
    net.lin_mask = 3;
    net.ext_mask = 1;
    net.stp_mode = 0;
  }
  extern __shared__ __align__(16) unsigned char wsm[];
  const int L = net.gather_lanes;
  const int npb = (int)(blockDim.x >> __builtin_ctz((unsigned)L));  // neurons a block can hold (L: a power of two)
  const uint64_t blk = net.wblk[blockIdx.x];
  const int32_t nb = (int32_t)(uint32_t)blk;         // first neuron of the block
  const int nbc = (int)(uint32_t)(blk >> 32);        // its neurons (<= npb)
  const int32_t N = net.N, B = net.B;
  const int tid = (int)threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wpb = (int)(blockDim.x >> 6);
  const int nlu = __builtin_popcount((uint32_t)net.lin_mask);
  const int nst = ((npb + 63) / 64) * 64 * W;  // spike staging entries (64 W per update wave)
  const bool flat = !UNI && B == 1;           // gather_block (else gather_window)
  // LDS carve-up (window_lds_bytes)
  uint64_t *s_prow = (uint64_t *)wsm;                         // [npb] push row start
  T *s_gn = (T *)(s_prow + npb);                              // [W][npb] NMDA sums
  T *s_inc = s_gn + (size_t)W * npb;                          // [nlu][W][npb] ring slots
  uint32_t *s_pdeg = (uint32_t *)(s_inc + (size_t)nlu * W * npb);  // [npb] push row length
  const bool inj = net.inj_on != 0;
  float *s_inj = (float *)(s_pdeg + npb);                     // [W][npb] injection draws (inj only)
  uint8_t *s_draw = (uint8_t *)(s_inj + (inj ? (size_t)W * npb : 0));  // [W][npb] event outcomes
  unsigned char *s_u = (unsigned char *)(s_draw + (size_t)W * npb);
  s_u += (uintptr_t)(-(intptr_t)s_u) & 7u;                    // 8-byte aligned shared region
  // phases B and C
  T *s_hs = (T *)s_u;                                         // [W][npb] NMDA gates of each step
  T *s_amp = s_hs + (size_t)W * npb;                          // [nst] staged push amplitudes
  uint32_t *s_spk = (uint32_t *)(s_amp + nst);                // [nst] step << 24 | neuron
  // phase A, gather_block's staging
  // window_chunk<T>(W, gather_chunk) with its division by W from the table
  int CH = W == 1 ? (int)(32768 / sizeof(T)) : (int)__umulhi((uint32_t)(32768 / sizeof(T)), kDivW.m[W]);
  CH = CH < 1024 ? CH : 1024;
  if (net.gather_chunk > 0 && net.gather_chunk < CH) CH = net.gather_chunk;
  uint64_t *s_rpn = (uint64_t *)s_u;                          // [npb + 1] gather row starts
  uint32_t *s_x = (uint32_t *)(s_rpn + npb + 1);              // [CH][W] staged gates (words)
  uint32_t *s_m = s_x + (size_t)CH * W * (sizeof(T) / 4);     // [CH] staged columns | delays
  float *s_w = (float *)(s_m + CH);                           // [CH] staged weights
  __shared__ uint32_t s_wcnt[8], s_base;
  if (tid < 8) s_wcnt[tid] = 0u;
  const int64_t tb = *net.t0;
  const int32_t tm0 = net.tmod[0], tm1 = net.tmod[1];

  // update lanes: npb consecutive threads, one neuron each, starting at a
  // wave that rotates with the block index, so the update waves of blocks
  // sharing a compute unit do not all land on one SIMD
  const int ubase = (wpb & (wpb - 1)) == 0 ? (int)(blockIdx.x & (unsigned)(wpb - 1)) * 64
                                           : (int)(blockIdx.x % (unsigned)wpb) * 64;
  const int ul = tid - ubase + (tid < ubase ? (int)blockDim.x : 0);  // update-lane index
  const int uw = ul >> 6;                                            // update wave ordinal
  const bool act = ul < nbc;
  const int32_t n = act ? nb + ul : N;
  // Every lane of a wave that holds update lanes (with DUP: of every wave)
  // loads (a lane without a neuron loads the block's first neuron's values
  // and never uses them): a
  // load under a per-lane condition would leave the other lanes' registers
  // to be merged with defaults afterwards, and that merge waits for the
  // load before the gather can start; a wave-uniform branch merges nothing.
  // For the same reason the push rows reach LDS only after the gather.
  const bool uwave = DUP || uw * 64 < npb;  // wave-uniform
  const int32_t nld = act ? n : nb;
  // (one network copy, or a power-of-two number of copies, needs no division)
  const bool b2 = (B & (B - 1)) == 0;
  const int lgB = __builtin_ctz((unsigned)B);
  const int32_t nl = b2 ? nld >> lgB : nld / B, rep = nld - nl * B;
  // The populations of the update and ring/draw neurons come first, while
  // nothing else is in flight: the block table's when its neurons share one
  // (the windowed blocks are runs of one population: no lookup at all),
  // else both lookups together, so the parameter loads below wait for one
  // round trip at most.
  const bool npb2 = (npb & (npb - 1)) == 0;  // (power-of-two blocks need no division)
  const int ir = npb2 ? tid & (npb - 1) : tid % npb;
  const int s0 = npb2 ? tid >> __builtin_ctz((unsigned)npb) : tid / npb;
  const bool rl = ir < nbc;  // this thread's neuron of the ring / draw work is in the block
  const int32_t nr = nb + (rl ? ir : 0);  // (loads by whole waves, as below)
  const int32_t nlr = b2 ? nr >> lgB : nr / B, repr = nr - nlr * B;
  const uint32_t bpop = ldc(net.wpop + blockIdx.x);
  uint32_t p0, pr;
  if (bpop != 0xFFFFFFFFu) {
    p0 = bpop;
    pr = bpop;
  } else {
    p0 = ldc(net.npop + nl);
    pr = ldc(net.npop + nlr);
  }
  NState<T> S;
  NConst<T> C;
  NConstRaw Rc;  // C's loads; its derived fields follow after the gather
  uint64_t prow0, prow1;
  if (uwave) {
    load_const_raw(net, *(PopC<T> *)(net.pops + p0), rep, p0, Rc);
    load_state(net, nl, rep, nld, slot_of(tm1, k0 - 1, net.DN), S);
    prow0 = ldc(net.rp_p + nl);
    prow1 = ldc(net.rp_p + nl + 1);
  }  // (no defaults for the other waves: they never read these, and a
     // default would be merged with the loaded values, and so wait for them)

  // ---- A: prefetch the window's inputs.  Thread -> (neuron ir, steps s0,
  // s0 + L, ...) for the ring slots and the draws (blockDim = npb * L): their
  // loads are issued first and consumed after the NMDA gather, whose latency
  // they share.
  // the wave's first thread has a step in the window (wave-uniform)
  const bool rwave = DUP || (int)(((unsigned)tid & ~63u) / (unsigned)npb) < W;
  constexpr int RQ = 3;      // ring slots per receptor held in registers across the gather
  T rv[4][RQ];
  NConst<T> Cr;
  NConstRaw Rr;  // Cr's loads; its derived fields are formed in the staging
  if (rwave) {
#pragma unroll
    for (int li = 0; li < 4; ++li) {
#pragma unroll
      for (int j = 0; j < RQ; ++j) {
        // slots past the window read the step before it (a valid slot, unused)
        const int32_t s = s0 + j * L < W ? s0 + j * L : 0;
        rv[li][j] = (net.lin_mask & (1 << li))
                        ? net.ring[((size_t)li * net.D + slot_of(tm0, k0 + s - 1, net.D)) * (size_t)N + nr]
                        : T(0);
      }
    }
    load_const_raw(net, *(PopC<T> *)(net.pops + pr), repr, pr, Rr);
  }  // (a wave with no step of the window stages no ring slot or draw)
  // Phase A's LDS side: the push rows, the ring slots and the draws (the
  // draws' population parameters through the pointer popr).  For a block
  // of one population, popr is wave-uniform (scalar loads, nothing that
  // waits on the vector memory counter), so with short rows (DUP: few lanes
  // per row, where the staging is a sizeable part of the launch) it runs
  // while the gather's gates are in flight; otherwise after the gather.
  auto stage = [&](uint32_t popr) __attribute__((always_inline)) {
    if (act) {
      s_prow[ul] = prow0;
      s_pdeg[ul] = (uint32_t)(prow1 - prow0);
    }
    if (rl) {
      int c = 0;
#pragma unroll
      for (int li = 0; li < 4; ++li) {
        if (!(net.lin_mask & (1 << li))) continue;
        int j = 0;
        for (int32_t s = s0; s < W; s += L, ++j) {
          const T v = j < RQ ? (j == 0 ? rv[li][0] : j == 1 ? rv[li][1] : rv[li][2])
                             : net.ring[((size_t)li * net.D + slot_of(tm0, k0 + s - 1, net.D)) * (size_t)N + nr];
          s_inc[((size_t)c * W + s) * npb + ir] = v;
        }
        ++c;
      }
      finish_const(net, Rr, nlr, repr, Cr);
      for (int32_t s = s0; s < W; s += L) {
        uint32_t db;
        float iv;
        neuron_draws<T, REPLAY>(net, *(PopC<T> *)(net.pops + popr), Cr, nr, tb + k0 + s, db, iv);
        s_draw[s * npb + ir] = (uint8_t)db;
        if (inj) s_inj[s * npb + ir] = iv;
      }
    }
  };
  const bool upop = DUP && bpop != 0xFFFFFFFFu;  // block-uniform
  if (net.has_nmda_syn && flat) {
    gather_block<T>(net, nb, nbc, ldc(net.wgat + 2 * blockIdx.x), ldc(net.wgat + 2 * blockIdx.x + 1),
                    slot_of(tm1, k0 - 1, net.DN), W, s_gn, npb, s_rpn, s_m, s_w, s_x, CH, [&]() {
                      if (upop) stage((uint32_t)__builtin_amdgcn_readfirstlane((int)bpop));
                    });
    if (!upop) stage(pr);
  } else {
    if (net.has_nmda_syn)
      gather_window<T, UNI>(net, tid / L < nbc ? (int64_t)nb + tid / L : (int64_t)N,
                            slot_of(tm1, k0 - 1, net.DN), W, L, tid % L, s_gn, npb, tid / L);
    stage(pr);
  }
  if (uwave) finish_const(net, Rc, nl, rep, C);
  lds_barrier();

  // ---- B: the W updates of the block's neurons, spikes staged per wave;
  // one pass per distinct population among the wave's lanes (usually one),
  // with that population's parameters copied into registers by scalar loads
  uint32_t hn = 0u;  // the lane's population has NMDA outputs
  if (uw * 64 < npb) {
    uint32_t wcnt = 0u;  // spikes this wave staged (wave-uniform)
    uint32_t mine = 0u;  // spikes of this lane's neuron
    const uint32_t wbase = (uint32_t)uw * 64u * (uint32_t)W;
    uint64_t todo = __ballot(act);
    while (todo) {
      const int ld = __ffsll((unsigned long long)todo) - 1;
      const uint32_t pc = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(p0, ld));
      const bool here = act && p0 == pc;
      todo &= ~__ballot(here);
      const DevPop<T> P = net.pops[pc];
      if (here) hn = P.has_nmda_out;
      // step s's inputs from LDS, read one step ahead of their use
      struct In {
        T inc[4];
        T gn;
        uint32_t db;
        float iv;
      };
      auto fetch = [&](int32_t s, In &x) {
        int c = 0;
#pragma unroll
        for (int li = 0; li < 4; ++li) {
          x.inc[li] = T(0);
          if (!(net.lin_mask & (1 << li))) continue;
          x.inc[li] = s_inc[((size_t)c * W + s) * npb + ul];
          ++c;
        }
        x.gn = net.has_nmda_syn ? s_gn[s * npb + ul] : T(0);
        x.db = s_draw[s * npb + ul];
        x.iv = inj ? s_inj[s * npb + ul] : 0.0f;
      };
      In cur;
      if (here) fetch(0, cur);
      for (int32_t s = 0; s < W; ++s) {
        const int32_t k = k0 + s;
        bool spk = false;
        T a_push = T(0);
        if (here) {
          In nxt = cur;
          if (s + 1 < W) fetch(s + 1, nxt);
          neuron_core<T, REPLAY, true, const DevPop<T>>(net, P, C, n, tb + k, 0, cur.gn, cur.inc, cur.db, cur.iv,
                                                        S, spk, a_push);
          s_hs[s * npb + ul] = S.sn;
          cur = nxt;
        }
        const uint64_t mask = __ballot(spk);
        if (mask) {
          if (spk) {
            const uint32_t pos = wbase + wcnt + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
            s_spk[pos] = ((uint32_t)s << 24) | (uint32_t)n;
            s_amp[pos] = a_push;
            ++mine;
          }
          wcnt += (uint32_t)__popcll(mask);
        }
      }
    }
    if (act) store_state(net, n, S);
    // per-population spike counts, one atomic when the wave's spikers share a population
    if (wcnt) {
      const uint32_t pid = p0 * (uint32_t)B + (uint32_t)rep;
      const uint64_t who = __ballot(mine != 0u);
      const int leader = __ffsll((unsigned long long)who) - 1;
      const uint32_t q0 = __shfl(pid, leader);
      if (__ballot(mine != 0u && pid != q0) == 0ull) {
        if (lane == leader) atomicAdd(&net.pop_cnt[q0], (unsigned long long)wcnt);
      } else if (mine) {
        atomicAdd(&net.pop_cnt[pid], (unsigned long long)mine);
      }
    }
    if (lane == 0) s_wcnt[uw] = wcnt;
  }
  lds_barrier();

  // ---- C: push the staged spikes, one wave per spike
  const int nwu = (npb + 63) / 64;
  uint32_t cnt = 0u;
  for (int w = 0; w < nwu; ++w) cnt += s_wcnt[w];
  auto entry = [&](uint32_t q) -> uint32_t {  // position of spike q of the concatenated lists
    int w = 0;
    while (q >= s_wcnt[w]) q -= s_wcnt[w++];
    return (uint32_t)w * 64u * (uint32_t)W + q;
  };
  unsigned long long events = 0, nout = 0;
  for (uint32_t q = (uint32_t)wave; q < cnt; q += (uint32_t)wpb) {
    const uint32_t ent = entry(q);
    const uint32_t pk = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_spk[ent]);
    const uint32_t j = pk & 0x00FFFFFFu, s = pk >> 24;
    const T amp = s_amp[ent];
    const uint32_t jl = b2 ? j >> lgB : j / (uint32_t)B, jr = j - jl * (uint32_t)B;
    const uint64_t r0 = s_prow[j - (uint32_t)nb];
    const uint64_t r1 = r0 + s_pdeg[j - (uint32_t)nb];
    const T ga = amp * T(net.gainr[0 * B + jr]), gg = amp * T(net.gainr[1 * B + jr]),
            gh = amp * T(net.gainr[3 * B + jr]), gl = amp * T(net.gainr[4 * B + jr]);
    if (lane == 0) {
      events += r1 - r0;
      nout += ldc(net.nmda_out + jl);
    }
    const int32_t sp = slot_of(tm0, k0 + (int32_t)s, net.D);  // ring slot of the emission step
    for (uint64_t e = r0 + (uint64_t)lane; e < r1; e += 64) {
      const uint32_t m = ldc(net.meta_p + e);
      const float w = ldc(net.w_p + e);
      const uint32_t col = (m & 0x00FFFFFFu) * (uint32_t)B + jr;
      const uint32_t li = m >> 30;
      const T sc = li == 0 ? ga : li == 1 ? gg : li == 2 ? gh : gl;
      int32_t slot = sp + (int32_t)((m >> 24) & 63u);
      slot -= slot >= net.D ? net.D : 0;
      atomic_add_t(net.ring + ((size_t)li * net.D + slot) * N + col, T(w) * sc);
    }
  }
  // the ring slots read in A, cleared for reuse (no push of this launch lands there)
  if (rl) {
#pragma unroll
    for (int li = 0; li < 4; ++li) {
      if (!(net.lin_mask & (1 << li))) continue;
      for (int32_t s = s0; s < W; s += L)
        net.ring[((size_t)li * net.D + slot_of(tm0, k0 + s - 1, net.D)) * (size_t)N + nr] = T(0);
    }
  }
  // the window's NMDA gates into the history (consecutive slots of the neuron)
  if (act && hn)
    for (int32_t s = 0; s < W; ++s)
      net.shist[hidx(net, (uint32_t)nl, (uint32_t)rep, (uint32_t)slot_of(tm1, k0 + s, net.DN))] =
          s_hs[s * npb + ul];
  if (net.rec_on) {
    // the block's range of its sub-list: (neuron, step within the batch)
    const int sub = (int)(blockIdx.x % NSUB);
    if (tid == 0) s_base = cnt ? atomicAdd(net.total + sub, cnt) : 0u;
    lds_barrier();
    const uint32_t off = net.sub_off[sub], cap = net.sub_cap[sub];
    for (uint32_t q = (uint32_t)tid; q < cnt; q += blockDim.x) {
      const uint32_t pk = s_spk[entry(q)];
      const uint32_t r = s_base + q;
      if (r < cap) {
        net.rec_ids[off + r] = pk & 0x00FFFFFFu;
        net.rec_k[off + r] = (uint32_t)k0 + (pk >> 24);
      } else {
        atomicOr(net.overflow, 1u);
      }
    }
  }
  if (tid == 0 && cnt) atomicAdd(&net.counters[blockIdx.x & (NCTR - 1)], (unsigned long long)cnt);
  block_add_events(net, events, nout);
}

// Push of the last step of a fused batch of K launches; publishes the end
// offset of the batch's record.
template <typename T>
__global__ void __launch_bounds__(1024) k_push_tail(Net<T> net, int32_t K) {
  const int pl = (K - 1) % 3;
  const uint32_t *lb = net.lbase + pl * NSUB, *lc = net.lcnt + pl * NSUB;
  auto last = [&](int s, uint32_t &a, uint32_t &b) {
    const uint32_t cap = net.sub_cap[s], x = lb[s], y = lb[s] + lc[s];
    a = x < cap ? x : cap;
    b = y < cap ? y : cap;
  };
  if (blockIdx.x == 0 && threadIdx.x < NSUB) {
    uint32_t a, b;
    last((int)threadIdx.x, a, b);
    net.cum[(size_t)K * NSUB + threadIdx.x] = b;
  }
  const int wave = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const int wpb = (int)(blockDim.x >> 6);
  unsigned long long ev = 0, no = 0;
  push_spikes(net, last, slot_of(net.tmod[0], K - 1, net.D), (int)blockIdx.x * wpb + wave,
              (int)gridDim.x * wpb, lane, ev, no);
  block_add_events(net, ev, no);
}

// NMDA gather of step ts into gn (state readback of the fused path).
template <typename T>
__global__ void __launch_bounds__(1024) k_gather_only(Net<T> net, int32_t h0) {
  const int lane = (int)(threadIdx.x & 63);
  const int64_t p = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const T v = gather_row<T, false>(net, p, h0, 64, lane);
  if (lane == 0 && p < net.N) net.gn[p] = v;
}

template <typename T, bool UNI>
__global__ void __launch_bounds__(1024) k_deliver(Net<T> net, int32_t k) {
  const int64_t t = *net.t0 + k;
  const int32_t N = net.N;
  const int wave = (int)(threadIdx.x >> 6);
  const int lane = (int)(threadIdx.x & 63);
  const int waves_per_block = (int)(blockDim.x >> 6);
  const int gwave = (int)blockIdx.x * waves_per_block + wave;
  const int total_waves = (int)gridDim.x * waves_per_block;
  unsigned long long my_events = 0, my_nout = 0;
  // spikes of step t: [cum[k][s], total[s]) of each sub-list
  auto cur = [&](int s, uint32_t &a, uint32_t &b) {
    const uint32_t cap = net.sub_cap[s], y = net.total[s];
    a = net.cum[(size_t)k * NSUB + s];
    b = y < cap ? y : cap;
  };
  if (blockIdx.x == 0 && threadIdx.x < NSUB) {
    uint32_t a, b;
    cur((int)threadIdx.x, a, b);
    net.cum[(size_t)(k + 1) * NSUB + threadIdx.x] = b;
  }
  if (gwave < net.push_waves) {
    // ---- push: one wave per spike of step t
    push_spikes(net, cur, slot_of(net.tmod[0], k, net.D), gwave, net.push_waves, lane, my_events,
                my_nout);
  } else if (net.has_nmda_syn) {
    // ---- NMDA gather over CSR-by-post, L lanes per row
    const int L = net.gather_lanes;
    const int32_t h0 = slot_of(net.tmod[1], k, net.DN);
    const int gw = gwave - net.push_waves;
    const int gwaves = total_waves - net.push_waves;
    const int groups_per_wave = 64 / L;
    const int sub = lane / L, sl = lane % L;
    const int64_t stride = (int64_t)gwaves * groups_per_wave;
    for (int64_t p = (int64_t)gw * groups_per_wave + sub; p - sub < N; p += stride) {
      const T acc = gather_row<T, UNI>(net, p, h0, L, sl);
      if (p < N && sl == 0) net.gn[p] = acc;
    }
  }
  block_add_events(net, my_events, my_nout);
}

// batch prologue: first step of the batch, record buffer reset
// (the fused path's list rotation starts with lists 2 (step -1) and 0 empty)
__global__ void k_batch_begin(int64_t *t0, int64_t value, uint32_t *total, uint32_t *cum,
                              uint32_t *lcnt, uint32_t *lbase, int32_t *tmod, int32_t D, int32_t DN) {
  const int s = (int)threadIdx.x;  // NSUB threads
  if (s == 0) {
    *t0 = value;
    tmod[0] = (int32_t)(value % D);
    tmod[1] = (int32_t)(value % DN);
  }
  if (s < NSUB) {
    total[s] = 0u;
    cum[s] = 0u;
    lcnt[s] = 0u;
    lcnt[2 * NSUB + s] = 0u;
    lbase[2 * NSUB + s] = 0u;
  }
}

// Learning synapses (LTP, code/EqsGnl2.cpp:1005-1044) after the update of
// step k of the batch, one thread per synapse: add w * gate of the
// presynaptic neuron (linear: its shared gate; NMDA: its s of step k) to the
// post's conductance of this Spike phase, then apply the weight rules with
// this step's spikes and LP traces: a presynaptic spike adds LP(post) *
// NegISI (or scales by 1 + LP(post) * NegISI when NegISI <= 0), a
// postsynaptic spike likewise with LP(pre) and PosISI; both from the
// presynaptic neuron's parameters.
template <typename T>
__global__ void __launch_bounds__(256) k_learn(Net<T> net, int32_t k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= net.n_learn) return;
  const int32_t hslot = slot_of(net.tmod[1], k, net.DN);  // gate history slot of step k
  const uint32_t a = net.l_pre[i], b = net.l_post[i];
  const int r = net.l_rec[i];
  const size_t N = (size_t)net.N;
  T w = net.l_w[i];
  const T gate = r == 2 ? net.shist[hidx(net, a, 0u, (uint32_t)hslot)] : net.spre[a];
  atomic_add_t(net.gl + (size_t)r * N + b, w * gate);
  const DevPop<T> &Pa = net.pops[net.npop[a]];
  if (net.cs[a] & F_SPIKE) {
    if (Pa.ltp_neg > T(0)) w = w + net.lp[b] * Pa.ltp_neg;
    else w = w * (T(1) + net.lp[b] * Pa.ltp_neg);
  }
  if (net.cs[b] & F_SPIKE) {
    if (Pa.ltp_pos > T(0)) w = w + net.lp[a] * Pa.ltp_pos;
    else w = w * (T(1) + net.lp[a] * Pa.ltp_pos);
  }
  net.l_w[i] = w;
}

// Gap-junction currents (ThdGapSyn, code/EqsGnl2.cpp:705-725) from the
// potentials of the step just updated: each neuron sums g (V_other - V_self)
// over both ends of its junctions; the next update adds it to Isyn.
template <typename T>
__global__ void __launch_bounds__(256) k_gap(Net<T> net) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= net.N) return;
  const uint32_t B = (uint32_t)net.B;
  const uint32_t nl = (uint32_t)n / B, rep = (uint32_t)n - nl * B;
  const T v = net.V[n];
  T acc = T(0);
  for (uint64_t e = net.rp_g[nl]; e < net.rp_g[nl + 1]; ++e)
    acc += T(net.w_g[e]) * (net.V[(size_t)net.col_g[e] * B + rep] - v);
  net.igap[n] = acc;
}

template <typename T>
__global__ void k_fill(T *p, T v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// reference-equivalent synaptic conductance after the Spike phase of the
// last step (G_r(t) = df G_r(t-1) + pending slot t), for state readback
template <typename T>
__global__ void k_read_g(Net<T> net, int64_t t, int li, T *out) {
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= net.N) return;
  const DevPop<T> &P = net.pops[net.npop[n / net.B]];
  const int32_t slot = (int32_t)(((t % net.D) + net.D) % net.D);
  out[n] = P.dfst[lin_rid(li)] * net.G[(size_t)li * net.N + n] +
           net.ring[((size_t)li * net.D + slot) * net.N + n];
}

// NMDA presynaptic gate of every neuron at history slot `slot` (state readback)
template <typename T>
__global__ void k_read_hist(Net<T> net, int32_t slot, T *out) {
  const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= net.N) return;
  const uint32_t nl = (uint32_t)n / (uint32_t)net.B;
  out[n] = net.shist[hidx(net, nl, (uint32_t)n - nl * (uint32_t)net.B, (uint32_t)slot)];
}

}  // namespace fsk
