// Host side of the Flysim engine: network upload, CSR construction, the
// batched step loop (hipGraph-captured), spike readback and the C ABI of
// include/flysim.h for the engine entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/flysim.h"
#include "engine_kernels.hpp"
#include "gen_kernels.hpp"
#include "host/host_common.hpp"

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

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 256;
    HIPCHECK(hipMalloc(&p, b));
    bytes = b;
  }
  template <class X> X *as() const { return static_cast<X *>(p); }
};

struct EngineBase {
  virtual ~EngineBase() {}
  virtual void reset(uint64_t seed) = 0;
  virtual void set_pop(int32_t pop, const fs_pop_params &p) = 0;
  virtual fs_pop_params pop(int32_t i) const = 0;
  virtual void set_ext(int32_t pop, int32_t receptor, float pb, bool on) = 0;
  virtual void set_replica(int32_t rep, const fs_replica_params &p) = 0;
  virtual void load_ext(const uint32_t *bits, int64_t first, int64_t n) = 0;
  virtual void run(int64_t n, int64_t *off, uint32_t *ids, int64_t cap) = 0;
  virtual void get_state(int32_t f, double *out) = 0;
  virtual void pop_counts(uint64_t *out) = 0;
  virtual void times(double *u, double *p, double *g, int64_t *n) = 0;
  virtual void counters(int64_t *out) = 0;
  virtual void stats(int64_t *out) const = 0;
  virtual bool fused_active() const = 0;
  virtual int32_t window_active() const = 0;
  virtual void record_currents(bool on) = 0;
  fs_schedule sched{};
  bool timing = false;
  int64_t step = 0;
  float dt = 0.1f;
  int32_t n_pop = 0;     // populations over all replicas (replica-major)
  int32_t replicas = 1;
};

template <typename T>
struct Engine final : EngineBase {
  fsk::Net<T> net{};
  int32_t N = 0, P = 0;   // N over all replicas; P of one replica
  int32_t N0 = 0, B = 1;
  int32_t precision = FS_FP32;
  bool record = true;
  std::vector<fs_pop_params> hpop;
  std::vector<fsk::DevPop<T>> dpop;
  std::vector<float> hpbr;      // [5][P][B]
  std::vector<float> hgain;     // [6][B]
  std::vector<uint64_t> hskey;  // [B]
  // device buffers
  DevBuf b_npop, b_pops, b_V, b_cs, b_G, b_ring, b_spre, b_xn, b_shist, b_gn, b_elin, b_enx,
      b_ens, b_stp, b_irep, b_rp_g, b_col_g, b_w_g, b_igap, b_lp, b_cntlp, b_lpre, b_lpost, b_lrec,
      b_lw, b_lw0, b_gl, b_rp_p, b_meta_p, b_w_p, b_rp_n, b_meta_n, b_w_n, b_ext, b_rec_ids, b_rec_a,
      b_total, b_rec_k, b_cum, b_overflow, b_t0, b_popcnt, b_counters, b_tmp, b_nmda_out, b_lcnt, b_lbase, b_suboff, b_subcap,
      b_pbr, b_gain, b_skey, b_tmod, b_zig, b_wblk, b_wgat, b_wpop;
  std::vector<uint32_t> hnpop;   // population of each local neuron
  std::vector<uint64_t> hwblk;   // windowed path: first | count << 32 of each block
  std::vector<uint64_t> hrp_n;   // gather CSR row starts of one copy (block balancing)
  // replay mode's Gaussian injection: population ranges and the position in
  // the reference's ziggurat stream (zig_calls calls before step zig_t0,
  // zig_a calls per step from there)
  std::vector<int64_t> pfirst, pcount;
  bool pops_contiguous = true;
  uint32_t zig_calls = 0, zig_a = 0;
  int64_t zig_t0 = 0;
  bool zig_dirty = false;  // injection switched since the last refresh
  bool fused_ok = false;  // every synapse has delay >= 1: one launch per step is valid
  bool fused() const { return fused_ok && sched.fused && !window(); }
  // windowed path: steps per launch, at most 1 + the smallest synaptic delay
  // and at most what the block's LDS holds (window_lds_bytes <= 64 KiB)
  int32_t win_max = 1;
  int32_t mind = 0, maxdp = 0, maxdn = 0;  // smallest delay; largest linear / NMDA delay (steps)
  size_t lds_max = 65536;  // dynamic LDS a block may take (the device's limit)
  // The windowed kernel may take nearly the device's whole LDS per block
  // (160 KiB on gfx950) when a window of many neurons needs it.
  void init_lds_limit() {
    int dev = 0, v = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // the kernel's own static LDS (a few counters) comes out of the same limit
    lds_max = std::max<size_t>(65536, (size_t)v - 1024);
    const int bytes = (int)lds_max;
#define FS_LDS(R, U, PR, DUP)                                                                     \
  HIPCHECK(hipFuncSetAttribute((const void *)fsk::k_window<T, R, U, PR, DUP>,                    \
                               hipFuncAttributeMaxDynamicSharedMemorySize, bytes))
    FS_LDS(false, false, 0, false); FS_LDS(false, false, 1, false); FS_LDS(true, false, 0, false);
    FS_LDS(true, false, 1, false); FS_LDS(false, false, 0, true); FS_LDS(false, false, 1, true);
    FS_LDS(true, false, 0, true); FS_LDS(true, false, 1, true); FS_LDS(false, true, 0, true);
    FS_LDS(false, true, 1, true); FS_LDS(true, true, 0, true); FS_LDS(true, true, 1, true);
#undef FS_LDS
  }
  bool inj_any() const {
    for (const auto &p : hpop)
      if (p.act & FS_ACT_INJ) return true;
    return false;
  }
  int32_t window() const {
    if (sched.window <= 0) return 0;
    const int npb = std::max(1, sched.update_block / sched.gather_lanes);
    const int nlu = __builtin_popcount((uint32_t)net.lin_mask);
    int32_t w = std::min(sched.window, win_max);
    if (B == 1) w = std::min(w, fsk::FLAT_WMAX);  // gather_block's partial sums per thread
    const int gc = sched.gather_chunk;
    const bool inj = inj_any();
    while (w > 1 && fsk::window_lds_bytes<T>(npb, w, nlu, B == 1, gc, inj) > lds_max) --w;
    if (w >= 1 && fsk::window_lds_bytes<T>(npb, w, nlu, B == 1, gc, inj) > lds_max) return 0;
    return w;
  }
  bool fused_active() const override { return fused(); }
  int32_t window_active() const override { return window(); }
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int32_t graph_k = -1;
  bool graph_dirty = true;
  bool tables_dirty = false;  // per-replica tables changed since the last upload
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double t_upd = 0, t_del = 0, t_graph = 0;
  int64_t n_graph_steps = 0;
  int64_t n_timed = 0;
  int64_t syn_lin = 0, syn_nmda = 0;
  std::vector<uint32_t> cum_h, hsub_off, hsub_cap, stage, stage_k;
  std::vector<uint32_t> ids_h;
  unsigned long long ctr_last[4] = {0, 0, 0, 0};

  ~Engine() override {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto &e : ev) if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  static fsk::DevPop<T> to_dev(const fs_pop_params &p, const fsk::DevPop<T> &pre) {
    fsk::DevPop<T> d = pre;  // keeps the presynaptic fields
    d.vth = T(p.vth);
    // gl*El*1e3f as the reference evaluates it each step (float, left to right)
    if (sizeof(T) == 4) d.leak_drive = T(p.gl * p.el * 1e3f);
    else d.leak_drive = T(p.gl) * T(p.el) * T(1e3f);
    d.cm = T(p.cm);
    d.vreset = T(p.vreset);
    d.df = T(p.df);
    d.ef = T(p.ef);
    d.dvmax = T(p.dvmax);
    d.mg = T(p.mg_n);
    for (int r = 0; r < 5; ++r) {
      d.vrev[r] = T(p.vrev[r]);
      d.dfst[r] = T(p.dfst[r]);
      d.gext[r] = T(p.gext[r]);
    }
    d.dfxt_n = T(p.dfxt_n);
    if (sizeof(T) == 4) d.efas_n = T(p.efst_n * p.as_n);
    else d.efas_n = T(p.efst_n) * T(p.as_n);
    d.inj_mean = T(p.inj_mean);
    d.inj_std = T(p.inj_std);
    d.ltp_dfLP = T(p.ltp_dfLP);
    d.ltp_pos = T(p.ltp_pos);
    d.ltp_neg = T(p.ltp_neg);
    d.ltp_spidy = p.ltp_spidy;
    d.stp_pv = T(p.stp_pv);
    d.stp_dfD = T(p.stp_dfD);
    d.stp_efD = T(p.stp_efD);
    d.stp_aF = T(p.stp_aF);
    d.stp_dfF = T(p.stp_dfF);
    d.stp_efF = T(p.stp_efF);
    d.stp_aCap = T(p.stp_aCap);
    d.stp_dfCap = T(p.stp_dfCap);
    d.stp_aCar = T(p.stp_aCar);
    d.stp_dfCar = T(p.stp_dfCar);
    d.stp_efCar = T(p.stp_efCar);
    d.rfped = p.rfped;
    d.spidy = p.spidy;
    d.act = p.act;
    return d;
  }

  // r4_nor_setup (code/StiExt.h:321-351): the 128 strips of Marsaglia and
  // Tsang's ziggurat, computed in double and stored as the reference stores them
  static fsk::ZigTables zig_tables() {
    fsk::ZigTables z{};
    const double m1 = 2147483648.0, vn = 9.91256303526217e-3;
    double dn = 3.442619855899, tn = dn;
    const double q = vn / std::exp(-0.5 * dn * dn);
    z.kn[0] = (uint32_t)((dn / q) * m1);
    z.kn[1] = 0;
    z.wn[0] = (float)(q / m1);
    z.wn[127] = (float)(dn / m1);
    z.fn[0] = 1.0f;
    z.fn[127] = (float)std::exp(-0.5 * dn * dn);
    for (int i = 126; i >= 1; --i) {
      dn = std::sqrt(-2.0 * std::log(vn / dn + std::exp(-0.5 * dn * dn)));
      z.kn[i + 1] = (uint32_t)((dn / tn) * m1);
      tn = dn;
      z.fn[i] = (float)std::exp(-0.5 * dn * dn);
      z.wn[i] = (float)(dn / m1);
    }
    // jump tables of jcong' = 69069 jcong + 1234567 (code/StiExt.h:222-229):
    // entry [j][b] advances b * 256^j calls
    uint32_t a = 69069u, c = 1234567u;  // one call of 256^j
    for (int j = 0; j < 4; ++j) {
      uint32_t pa = 1u, pc = 0u;
      for (int b = 0; b < 256; ++b) {
        z.ja[j][b] = pa;
        z.jc[j][b] = pc;
        pc = a * pc + c;
        pa = a * pa;
      }
      a = pa;  // 256 calls of 256^j
      c = pc;
    }
    return z;
  }

  // Replay mode: each step the reference draws one ziggurat call per neuron
  // with injection on, in index order, from one stream seeded at Run
  // (code/EqsGnl2.cpp:335, 1098-1105; code/NetSimGnl2.cpp:323).  Called when
  // the set of injected populations changes (and at reset): the calls up to
  // the current step are banked, and each population gets the offsets that
  // make neuron n's call at step t equal t * zig_a + n + zig_b (mod 2^32).
  void zig_refresh() {
    zig_dirty = false;
    if (net.ext_mode != FS_EXT_REPLAY) return;
    zig_calls += (uint32_t)(step - zig_t0) * zig_a;
    zig_t0 = step;
    std::vector<int32_t> on;
    for (int32_t i = 0; i < P; ++i)
      if ((hpop[i].act & FS_ACT_INJ) && pcount[i]) on.push_back(i);
    if (!on.empty() && (B != 1 || !pops_contiguous))
      throw std::invalid_argument(
          "replay-mode Gaussian injection needs one instance whose populations are contiguous neuron ranges");
    std::sort(on.begin(), on.end(), [&](int32_t x, int32_t y) { return pfirst[x] < pfirst[y]; });
    uint32_t a = 0;
    for (int32_t i : on) a += (uint32_t)pcount[i];
    uint32_t rank = 0;
    for (int32_t i = 0; i < P; ++i) dpop[i].zig_a = a, dpop[i].zig_b = 0;
    for (int32_t i : on) {
      dpop[i].zig_b = zig_calls - (uint32_t)zig_t0 * a + rank - (uint32_t)pfirst[i];
      rank += (uint32_t)pcount[i];
    }
    zig_a = a;
    upload_pops();
  }

  void upload_pops() {
    HIPCHECK(hipMemcpyAsync(b_pops.p, dpop.data(), sizeof(fsk::DevPop<T>) * P,
                            hipMemcpyHostToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }

  void upload_replica_tables() {
    HIPCHECK(hipMemcpyAsync(b_pbr.p, hpbr.data(), hpbr.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(b_gain.p, hgain.data(), hgain.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(b_skey.p, hskey.data(), hskey.size() * 8, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));
  }
  float &pbr(int r, int32_t p0, int32_t rep) { return hpbr[((size_t)r * P + p0) * B + rep]; }
  // pb of population p0 in every replica from its parameter block (a cleared
  // activity bit is stored as -1)
  void pbr_from_pop(int32_t p0, int32_t rep, const fs_pop_params &p) {
    static const uint32_t bit[5] = {FS_ACT_AMPA, FS_ACT_GABA, FS_ACT_NMDA, FS_ACT_ACH, FS_ACT_GCL};
    for (int r = 0; r < 5; ++r) pbr(r, p0, rep) = (p.act & bit[r]) ? p.pb[r] : -1.0f;
  }

  // receptor r gets the external-input code path once any population of any
  // replica has pb > 0 (a graph argument: changing it re-captures)
  void note_pb(int r, float pb) {
    if (pb > 0.0f && !(net.ext_mask & (1 << r))) {
      net.ext_mask |= 1 << r;
      graph_dirty = true;
    }
  }
  void update_ext_mask() {
    int32_t m = 0;
    for (size_t i = 0; i < hpbr.size(); ++i)
      if (hpbr[i] > 0.0f) m |= 1 << (int)(i / ((size_t)P * B));
    if ((m | net.ext_mask) != net.ext_mask) {
      net.ext_mask |= m;
      graph_dirty = true;
    }
  }

  Engine(const fs_net_desc &d, int32_t n_rep) {
    N0 = d.n_neuron;
    P = d.n_pop;
    B = n_rep;
    replicas = B;
    dt = d.dt;
    precision = d.precision;
    record = d.record_spikes != 0;
    if (N0 <= 0 || P <= 0) throw std::invalid_argument("network needs neurons and populations");
    if (B < 1) throw std::invalid_argument("replicas must be >= 1");
    if ((int64_t)N0 * B >= (1 << 24)) throw std::invalid_argument("at most 2^24-1 neurons per device");
    N = N0 * B;
    n_pop = P * B;
    HIPCHECK(hipStreamCreate(&stream));
    init_lds_limit();
    for (auto &e : ev) HIPCHECK(hipEventCreate(&e));
    hpop.assign(d.pops, d.pops + P);
    for (int32_t i = 0; i < N0; ++i)
      if (d.neuron_pop[i] >= (uint32_t)P) throw std::invalid_argument("neuron_pop out of range");
    // ---------------- synapses: classify, validate, CSR by pre / by post
    std::vector<uint64_t> cnt_p(N0 + 1, 0), cnt_n(N0 + 1, 0), cnt_g(N0 + 1, 0);
    // LTP: the populations' switch (global in the reference, -LTP); the
    // plastic synapses then leave the push / gather tables for the learning list
    int32_t ltp = 0;
    for (int32_t i = 0; i < P; ++i) ltp |= (hpop[i].act & FS_ACT_LTP) ? 1 : 0;
    for (int32_t i = 0; i < P; ++i)
      if (((hpop[i].act & FS_ACT_LTP) ? 1 : 0) != ltp)
        throw std::invalid_argument("populations disagree on the LTP switch");
    if (ltp && B > 1) throw std::invalid_argument("LTP is not supported with replicas");
    auto learns = [&](int64_t s) { return ltp && d.plastic && d.plastic[s] && d.receptor[s] != FS_GAP; };
    int64_t n_learn = 0;
    int32_t maxd_p = 0, maxd_n = 0, mind_all = 1 << 30;
    int32_t lin_mask = 0;
    // presynaptic gate parameters per population (targets' receptor params)
    std::vector<int> pre_set(P, 0), pre_set_n(P, 0);
    std::vector<float> pre_df(P, 0.f), pre_dfxt(P, 0.f), pre_dfst(P, 0.f), pre_ef(P, 0.f),
        pre_as(P, 0.f);
    for (int64_t s = 0; s < d.n_syn; ++s) {
      const uint32_t a = d.pre[s], b = d.post[s];
      const uint8_t r = d.receptor[s];
      const int32_t dl = d.delay ? d.delay[s] : 0;
      if (a >= (uint32_t)N0 || b >= (uint32_t)N0) throw std::invalid_argument("synapse endpoint out of range");
      if (r > FS_GCL && r != FS_GAP) throw std::invalid_argument("synapse receptor id out of range");
      if (r == FS_GAP) {  // electrical: both endpoints, no delay
        if (dl != 0) throw std::invalid_argument("gap junctions take no delay");
        cnt_g[a + 1]++;
        cnt_g[b + 1]++;
        continue;
      }
      const uint32_t pp = d.neuron_pop[a];
      const fs_pop_params &q = hpop[d.neuron_pop[b]];
      const bool lr = learns(s);
      if (lr && dl != 0) throw std::invalid_argument("plastic synapses take no delay");
      n_learn += lr ? 1 : 0;
      mind_all = std::min(mind_all, dl);
      if (r == FS_NMDA) {
        if (dl > 255) throw std::invalid_argument("NMDA delay above 255 steps");
        maxd_n = std::max(maxd_n, dl);
        if (!lr) cnt_n[b + 1]++;
        if (!pre_set_n[pp]) {
          pre_set_n[pp] = 1;
          pre_dfxt[pp] = q.dfxt_n; pre_dfst[pp] = q.dfst[FS_NMDA];
          pre_ef[pp] = q.efst_n; pre_as[pp] = q.as_n;
        } else if (pre_dfxt[pp] != q.dfxt_n || pre_dfst[pp] != q.dfst[FS_NMDA] ||
                   pre_ef[pp] != q.efst_n || pre_as[pp] != q.as_n) {
          throw std::invalid_argument(
              "a population projects NMDA synapses onto populations with different NMDA time "
              "constants; one presynaptic NMDA gate per population is supported");
        }
      } else {
        if (dl > 63) throw std::invalid_argument("linear-receptor delay above 63 steps");
        maxd_p = std::max(maxd_p, dl);
        if (!lr) cnt_p[a + 1]++;
        const int li = r < 2 ? r : r - 1;
        lin_mask |= 1 << li;
        if (!pre_set[pp]) { pre_set[pp] = 1; pre_df[pp] = q.dfst[r]; }
        else if (pre_df[pp] != q.dfst[r])
          throw std::invalid_argument(
              "a population projects linear synapses whose postsynaptic decay constants differ; "
              "one presynaptic linear gate per population is supported");
      }
    }
    for (int32_t i = 0; i < N0; ++i) {
      cnt_p[i + 1] += cnt_p[i];
      cnt_n[i + 1] += cnt_n[i];
      cnt_g[i + 1] += cnt_g[i];
    }
    const int64_t gap_ends = (int64_t)cnt_g[N0];  // two per gap junction
    std::vector<uint32_t> col_g(gap_ends);
    std::vector<float> w_g(gap_ends);
    std::vector<uint32_t> l_pre, l_post;
    std::vector<uint8_t> l_rec;
    std::vector<T> l_w;
    l_pre.reserve(n_learn); l_post.reserve(n_learn); l_rec.reserve(n_learn); l_w.reserve(n_learn);
    syn_lin = (int64_t)cnt_p[N0];
    syn_nmda = (int64_t)cnt_n[N0];
    std::vector<uint32_t> meta_p(syn_lin), meta_n(syn_nmda);
    std::vector<float> w_p(syn_lin), w_n(syn_nmda);
    {
      std::vector<uint64_t> fp(cnt_p.begin(), cnt_p.end() - 1), fn(cnt_n.begin(), cnt_n.end() - 1),
          fg(cnt_g.begin(), cnt_g.end() - 1);
      for (int64_t s = 0; s < d.n_syn; ++s) {
        const uint32_t a = d.pre[s], b = d.post[s];
        const uint8_t r = d.receptor[s];
        const uint32_t dl = d.delay ? d.delay[s] : 0u;
        if (learns(s)) {
          l_pre.push_back(a);
          l_post.push_back(b);
          l_rec.push_back(r);
          l_w.push_back(T(d.w[s]));
        } else if (r == FS_GAP) {  // each end sums g (V_other - V_self)
          const uint64_t e1 = fg[b]++, e2 = fg[a]++;
          col_g[e1] = a;
          w_g[e1] = d.w[s];
          col_g[e2] = b;
          w_g[e2] = d.w[s];
        } else if (r == FS_NMDA) {
          const uint64_t e = fn[b]++;
          meta_n[e] = a | (dl << 24);
          w_n[e] = d.w[s];
        } else {
          const uint32_t li = r < 2 ? r : r - 1u;
          const uint64_t e = fp[a]++;
          meta_p[e] = b | (dl << 24) | (li << 30);
          w_p[e] = d.w[s];
        }
      }
    }
    std::vector<uint32_t> nout(N0, 0u);
    for (int64_t s = 0; s < d.n_syn; ++s)
      if (d.receptor[s] == FS_NMDA && !learns(s)) nout[d.pre[s]]++;
    PreGates pg{pre_set, pre_set_n, pre_df, pre_dfxt, pre_dfst, pre_ef, pre_as};
    init_pops(pg);
    const int64_t n_chem = d.n_syn - gap_ends / 2;
    win_max = n_chem > 0 ? std::min(mind_all + 1, 64) : 64;
    mind = n_chem > 0 ? mind_all : 0;
    maxdp = maxd_p;
    maxdn = maxd_n;
    // gap currents use every neuron's potential of the step before: one
    // step per launch, after the whole update (the two-launch path)
    if (gap_ends || ltp) win_max = 0;
    init_dims(maxd_p, maxd_n, n_chem > 0 && mind_all >= 1 && !gap_ends && !ltp, lin_mask, d.ext_mode);
    // the NMDA gather tables may be empty while plastic NMDA synapses still
    // need the presynaptic gate history
    net.has_nmda_syn = syn_nmda > 0 || std::count(l_rec.begin(), l_rec.end(), (uint8_t)FS_NMDA) > 0;
    const size_t n0 = (size_t)N0;
    b_rp_p.alloc((n0 + 1) * 8);
    HIPCHECK(hipMemcpy(b_rp_p.p, cnt_p.data(), (n0 + 1) * 8, hipMemcpyHostToDevice));
    b_meta_p.alloc(syn_lin * 4);
    b_w_p.alloc(syn_lin * 4);
    if (syn_lin) {
      HIPCHECK(hipMemcpy(b_meta_p.p, meta_p.data(), syn_lin * 4, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(b_w_p.p, w_p.data(), syn_lin * 4, hipMemcpyHostToDevice));
    }
    b_rp_n.alloc((n0 + 1) * 8);
    HIPCHECK(hipMemcpy(b_rp_n.p, cnt_n.data(), (n0 + 1) * 8, hipMemcpyHostToDevice));
    hrp_n.assign(cnt_n.begin(), cnt_n.begin() + (n0 + 1));
    b_meta_n.alloc(syn_nmda * 4);
    b_w_n.alloc(syn_nmda * 4);
    if (syn_nmda) {
      HIPCHECK(hipMemcpy(b_meta_n.p, meta_n.data(), syn_nmda * 4, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(b_w_n.p, w_n.data(), syn_nmda * 4, hipMemcpyHostToDevice));
    }
    b_nmda_out.alloc(n0 * 4);
    HIPCHECK(hipMemcpy(b_nmda_out.p, nout.data(), n0 * 4, hipMemcpyHostToDevice));
    finish(d.neuron_pop);
    if (ltp) {
      net.ltp_mode = 1;
      b_lp.alloc((size_t)N * sizeof(T));
      b_cntlp.alloc((size_t)N * 4);
      b_gl.alloc((size_t)5 * N * sizeof(T));
      const size_t L = l_pre.size();
      b_lpre.alloc(std::max<size_t>(L, 1) * 4);
      b_lpost.alloc(std::max<size_t>(L, 1) * 4);
      b_lrec.alloc(std::max<size_t>(L, 1));
      b_lw.alloc(std::max<size_t>(L, 1) * sizeof(T));
      b_lw0.alloc(std::max<size_t>(L, 1) * sizeof(T));
      if (L) {
        HIPCHECK(hipMemcpy(b_lpre.p, l_pre.data(), L * 4, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(b_lpost.p, l_post.data(), L * 4, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(b_lrec.p, l_rec.data(), L, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(b_lw0.p, l_w.data(), L * sizeof(T), hipMemcpyHostToDevice));
      }
      net.lp = b_lp.as<T>();
      net.cntlp = b_cntlp.as<uint32_t>();
      net.gl = b_gl.as<T>();
      net.n_learn = (int64_t)L;
      net.l_pre = b_lpre.as<uint32_t>();
      net.l_post = b_lpost.as<uint32_t>();
      net.l_rec = b_lrec.as<uint8_t>();
      net.l_w = b_lw.as<T>();
      graph_dirty = true;
      reset(0);
    }
    if (gap_ends) {
      b_rp_g.alloc((n0 + 1) * 8);
      HIPCHECK(hipMemcpy(b_rp_g.p, cnt_g.data(), (n0 + 1) * 8, hipMemcpyHostToDevice));
      b_col_g.alloc(gap_ends * 4);
      b_w_g.alloc(gap_ends * 4);
      HIPCHECK(hipMemcpy(b_col_g.p, col_g.data(), gap_ends * 4, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(b_w_g.p, w_g.data(), gap_ends * 4, hipMemcpyHostToDevice));
      b_igap.alloc((size_t)N * sizeof(T));
      net.rp_g = b_rp_g.as<uint64_t>();
      net.col_g = b_col_g.as<uint32_t>();
      net.w_g = b_w_g.as<float>();
      net.igap = b_igap.as<T>();
      graph_dirty = true;
      reset(0);
    }
  }

  // Batched random E/I instances, synapse tables generated on the device
  // (gen_kernels.hpp).  Stream key of instance k: mix64(seed_k ^ salt).
  explicit Engine(const fs_ei_batch_desc &d) {
    const int32_t K = d.n_inst;
    if (K < 1 || K > 65535) throw std::invalid_argument("n_inst must be 1..65535");
    if (!d.n_neuron || !d.seed || !d.ext_hz || !d.gain_nmda || !d.gain_gaba)
      throw std::invalid_argument("incomplete batch description");
    if (!(d.p_conn >= 0.0f && d.p_conn <= 1.0f)) throw std::invalid_argument("p_conn must be in [0,1]");
    if (!(d.exc_frac >= 0.0f && d.exc_frac <= 1.0f)) throw std::invalid_argument("exc_frac must be in [0,1]");
    if (d.delay_steps < 0 || d.delay_steps > 63) throw std::invalid_argument("delay_steps must be 0..63");
    int64_t tot = 0;
    for (int32_t k = 0; k < K; ++k) {
      if (d.n_neuron[k] < 1) throw std::invalid_argument("every instance needs neurons");
      tot += d.n_neuron[k];
    }
    if (tot >= (1 << 24)) throw std::invalid_argument("at most 2^24-1 neurons per device");
    // the presynaptic gates need one decay constant per receptor across E and I targets
    if (d.exc.dfst[FS_AMPA] != d.inh.dfst[FS_AMPA] || d.exc.dfst[FS_GABA] != d.inh.dfst[FS_GABA] ||
        d.exc.dfst[FS_NMDA] != d.inh.dfst[FS_NMDA] || d.exc.dfxt_n != d.inh.dfxt_n ||
        d.exc.efst_n != d.inh.efst_n || d.exc.as_n != d.inh.as_n)
      throw std::invalid_argument("E and I populations must share receptor time constants");
    N0 = N = (int32_t)tot;
    B = replicas = 1;
    P = n_pop = 2 * K;
    dt = d.dt;
    precision = d.precision;
    record = d.record_spikes != 0;
    HIPCHECK(hipStreamCreate(&stream));
    init_lds_limit();
    for (auto &e : ev) HIPCHECK(hipEventCreate(&e));
    std::vector<fsg::GenInst> gi(K);
    std::vector<uint32_t> npop(N);
    std::vector<uint16_t> inst_of(N);
    hpop.resize(P);
    int64_t first = 0;
    for (int32_t k = 0; k < K; ++k) {
      const int32_t n = d.n_neuron[k];
      const int32_t ne = (int32_t)std::floor((double)d.exc_frac * n + 0.5);
      gi[k] = fsg::GenInst{first, n, ne, fsg::mix64(d.seed[k]), d.gain_nmda[k], d.gain_gaba[k]};
      for (int32_t i = 0; i < n; ++i) {
        npop[first + i] = (uint32_t)(2 * k + (i < ne ? 0 : 1));
        inst_of[first + i] = (uint16_t)k;
      }
      const float pb = (float)(d.ext_hz[k] * d.dt * 1e-3f);
      for (int e = 0; e < 2; ++e) {
        fs_pop_params q = e == 0 ? d.exc : d.inh;
        if (d.ext_hz[k] == 0.0f) q.act &= ~(uint32_t)FS_ACT_AMPA;
        else { q.act |= FS_ACT_AMPA; q.pb[FS_AMPA] = pb; }
        hpop[2 * k + e] = q;
      }
      first += n;
    }
    // synapse tables: count pass, row offsets on the host, fill pass
    const size_t n0 = (size_t)N0;
    DevBuf b_gi, b_inst;
    b_gi.alloc(sizeof(fsg::GenInst) * K);
    HIPCHECK(hipMemcpy(b_gi.p, gi.data(), sizeof(fsg::GenInst) * K, hipMemcpyHostToDevice));
    b_inst.alloc(n0 * 2);
    HIPCHECK(hipMemcpy(b_inst.p, inst_of.data(), n0 * 2, hipMemcpyHostToDevice));
    b_rp_p.alloc((n0 + 1) * 8);
    b_rp_n.alloc((n0 + 1) * 8);
    b_nmda_out.alloc(n0 * 4);
    HIPCHECK(hipMemset(b_rp_p.p, 0, (n0 + 1) * 8));
    HIPCHECK(hipMemset(b_rp_n.p, 0, (n0 + 1) * 8));
    fsg::GenArgs ga{};
    ga.inst = b_gi.as<fsg::GenInst>();
    ga.inst_of = b_inst.as<uint16_t>();
    ga.N = N0;
    ga.thr = (uint32_t)std::min(4294967295.0, std::floor((double)d.p_conn * 4294967296.0));
    ga.g_lo = d.g_lo;
    ga.g_span = d.g_hi - d.g_lo;
    ga.nmda_ratio = d.nmda_ratio;
    ga.delay = (uint32_t)d.delay_steps;
    ga.rp_p = b_rp_p.as<uint64_t>();
    ga.rp_n = b_rp_n.as<uint64_t>();
    ga.nmda_out = b_nmda_out.as<uint32_t>();
    const dim3 grid((unsigned)((n0 * 64 + 255) / 256));
    ga.fill = 0;
    hipLaunchKernelGGL(fsg::k_gen_rows, grid, dim3(256), 0, stream, ga);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream));
    for (int w = 0; w < 2; ++w) {
      DevBuf &rp = w == 0 ? b_rp_p : b_rp_n;
      std::vector<uint64_t> h(n0 + 1);
      HIPCHECK(hipMemcpy(h.data(), rp.p, (n0 + 1) * 8, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n0; ++i) h[i + 1] += h[i];
      HIPCHECK(hipMemcpy(rp.p, h.data(), (n0 + 1) * 8, hipMemcpyHostToDevice));
      (w == 0 ? syn_lin : syn_nmda) = (int64_t)h[n0];
      if (w == 1) hrp_n.swap(h);
    }
    b_meta_p.alloc((size_t)syn_lin * 4);
    b_w_p.alloc((size_t)syn_lin * 4);
    b_meta_n.alloc((size_t)syn_nmda * 4);
    b_w_n.alloc((size_t)syn_nmda * 4);
    ga.meta_p = b_meta_p.as<uint32_t>();
    ga.w_p = b_w_p.as<float>();
    ga.meta_n = b_meta_n.as<uint32_t>();
    ga.w_n = b_w_n.as<float>();
    ga.fill = 1;
    hipLaunchKernelGGL(fsg::k_gen_rows, grid, dim3(256), 0, stream, ga);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(stream));
    // presynaptic gates: E drive AMPA + NMDA targets, I drive GABA targets
    PreGates pg;
    pg.set.assign(P, 1);
    pg.set_n.assign(P, 0);
    pg.df.assign(P, 0.f); pg.dfxt.assign(P, 0.f); pg.dfst.assign(P, 0.f);
    pg.ef.assign(P, 0.f); pg.as.assign(P, 0.f);
    for (int32_t k = 0; k < K; ++k) {
      pg.df[2 * k] = d.exc.dfst[FS_AMPA];
      pg.set_n[2 * k] = 1;
      pg.dfxt[2 * k] = d.exc.dfxt_n;
      pg.dfst[2 * k] = d.exc.dfst[FS_NMDA];
      pg.ef[2 * k] = d.exc.efst_n;
      pg.as[2 * k] = d.exc.as_n;
      pg.df[2 * k + 1] = d.exc.dfst[FS_GABA];
    }
    init_pops(pg);
    for (int32_t k = 0; k < K; ++k) {
      const uint64_t skey = fsg::mix64(d.seed[k] ^ 0x5851F42D4C957F2Dull);
      for (int e = 0; e < 2; ++e) {
        dpop[2 * k + e].key_lo = (uint32_t)skey;
        dpop[2 * k + e].key_hi = (uint32_t)(skey >> 32);
        dpop[2 * k + e].first = (int32_t)gi[k].first;
      }
    }
    // one synaptic delay for every synapse: windows of 1 + delay steps
    const int32_t dl = d.delay_steps;
    win_max = std::min(dl + 1, 64);
    mind = maxdp = maxdn = dl;
    init_dims(dl, dl, dl >= 1, (1 << 0) | (1 << 1), d.ext_mode);
    finish(npop.data());
  }

  // presynaptic gate parameters per population (its targets' receptor constants)
  struct PreGates {
    std::vector<int> set, set_n;
    std::vector<float> df, dfxt, dfst, ef, as;
  };

  void init_pops(const PreGates &g) {
    dpop.resize(P);
    for (int32_t i = 0; i < P; ++i) {
      fsk::DevPop<T> base{};
      base.has_lin_out = g.set[i];
      base.has_nmda_out = g.set_n[i];
      base.pre_df = T(g.df[i]);
      base.pre_dfxt = T(g.dfxt[i]);
      base.pre_dfst = T(g.dfst[i]);
      if (sizeof(T) == 4) base.pre_efas = T(g.ef[i] * g.as[i]);
      else base.pre_efas = T(g.ef[i]) * T(g.as[i]);
      dpop[i] = to_dev(hpop[i], base);
    }
  }

  void init_dims(int32_t maxd_p, int32_t maxd_n, bool fusable, int32_t lin_mask, int32_t ext_mode) {
    net.N = N;
    // the windowed path writes up to win_max steps ahead of the slots other
    // blocks still read (the fused step needs DN >= max NMDA delay + 2)
    // (a power of two of slots, like the gate history: slot arithmetic by mask)
    int32_t d = 1;
    while (d < maxd_p + 1 + win_max) d <<= 1;
    net.D = d;
    // the gate history: a power of two of slots per neuron (slot arithmetic by mask)
    int32_t dn = 1;
    while (dn < maxd_n + 1 + win_max) dn <<= 1;
    net.DN = dn;
    fused_ok = fusable;
    net.lin_mask = lin_mask;
    net.has_nmda_syn = syn_nmda > 0;
    net.ext_mode = ext_mode;
    net.stp_mode = stp_mode_of();
    net.words = (N + 31) / 32;
    net.B = B;
    net.N0 = N0;
    net.P0 = P;
  }

  // The plasticity switch is global in the reference (-STP / -STD set the
  // solver's ActCtr, code/main.cpp:229-230): every population carries it.
  int32_t stp_mode_of() const {
    const uint32_t m = FS_ACT_STP | FS_ACT_STD;
    const uint32_t a0 = P ? hpop[0].act & m : 0u;
    for (int32_t i = 1; i < P; ++i)
      if ((hpop[i].act & m) != a0)
        throw std::invalid_argument("populations disagree on the STP / STD switch");
    if (a0 & FS_ACT_STP) return fsk::STP_FULL;  // STP wins over STD (code/EqsGnl2.cpp:1109-1128)
    if (a0 & FS_ACT_STD) return fsk::STP_STD;
    return 0;
  }

  // state, tables and launch set-up once the synapse tables are on the device
  void finish(const uint32_t *neuron_pop) {
    const size_t n = (size_t)N, n0 = (size_t)N0;
    b_npop.alloc(n0 * 4);
    HIPCHECK(hipMemcpy(b_npop.p, neuron_pop, n0 * 4, hipMemcpyHostToDevice));
    hnpop.assign(neuron_pop, neuron_pop + n0);
    b_pops.alloc(sizeof(fsk::DevPop<T>) * P);
    upload_pops();
    pfirst.assign(P, -1);
    pcount.assign(P, 0);
    for (size_t k = 0; k < n0; ++k) {
      const uint32_t p = neuron_pop[k];
      if (pfirst[p] < 0) pfirst[p] = (int64_t)k;
      pcount[p]++;
    }
    for (size_t k = 0; k < n0; ++k)
      if ((int64_t)k >= pfirst[neuron_pop[k]] + pcount[neuron_pop[k]]) pops_contiguous = false;
    const fsk::ZigTables zt = zig_tables();
    b_zig.alloc(sizeof(zt));
    HIPCHECK(hipMemcpy(b_zig.p, &zt, sizeof(zt), hipMemcpyHostToDevice));
    net.zig = b_zig.as<fsk::ZigTables>();
    // per-replica tables: pb from the population blocks, unit gains, and
    // stream keys spaced by the 64-bit golden ratio (replica 0 keeps the seed)
    hpbr.assign((size_t)5 * P * B, 0.0f);
    for (int32_t rep = 0; rep < B; ++rep)
      for (int32_t i = 0; i < P; ++i) pbr_from_pop(i, rep, hpop[i]);
    hgain.assign((size_t)6 * B, 1.0f);
    hskey.resize(B);
    for (int32_t rep = 0; rep < B; ++rep) hskey[rep] = (uint64_t)rep * 0x9E3779B97F4A7C15ull;
    b_pbr.alloc(hpbr.size() * 4);
    b_gain.alloc(hgain.size() * 4);
    b_skey.alloc(hskey.size() * 8);
    upload_replica_tables();
    update_ext_mask();
    b_V.alloc(n * sizeof(T));
    b_cs.alloc(n * 4);
    const int nlin = 4;
    b_G.alloc((size_t)nlin * n * sizeof(T));
    b_ring.alloc((size_t)nlin * net.D * n * sizeof(T));
    b_spre.alloc(n * sizeof(T));
    b_xn.alloc(n * sizeof(T));
    b_shist.alloc((size_t)net.DN * n * sizeof(T));
    b_gn.alloc(n * sizeof(T));
    b_elin.alloc((size_t)nlin * n * sizeof(T));
    b_enx.alloc(n * sizeof(T));
    b_ens.alloc(n * sizeof(T));
    if (net.stp_mode) b_stp.alloc(4 * n * sizeof(T));
    b_total.alloc(fsk::NSUB * 4);
    b_overflow.alloc(4);
    b_lcnt.alloc(3 * fsk::NSUB * 4);
    b_lbase.alloc(3 * fsk::NSUB * 4);
    b_suboff.alloc(fsk::NSUB * 4);
    b_subcap.alloc(fsk::NSUB * 4);
    b_t0.alloc(8);
    b_tmod.alloc(8);
    b_popcnt.alloc((size_t)P * B * 8);
    b_counters.alloc(3 * fsk::NCTR * 8);
    b_tmp.alloc(n * sizeof(T));
    net.npop = b_npop.as<uint32_t>();
    net.pops = b_pops.as<fsk::DevPop<T>>();
    net.V = b_V.as<T>();
    net.cs = b_cs.as<uint32_t>();
    net.G = b_G.as<T>();
    net.ring = b_ring.as<T>();
    net.spre = b_spre.as<T>();
    net.xn = b_xn.as<T>();
    net.shist = b_shist.as<T>();
    net.gn = b_gn.as<T>();
    net.elin = b_elin.as<T>();
    net.enx = b_enx.as<T>();
    net.ens = b_ens.as<T>();
    if (net.stp_mode) {
      net.sD = b_stp.as<T>();
      net.sF = net.sD + n;
      net.sCap = net.sF + n;
      net.sCar = net.sCap + n;
    }
    net.rp_p = b_rp_p.as<uint64_t>();
    net.meta_p = b_meta_p.as<uint32_t>();
    net.w_p = b_w_p.as<float>();
    net.rp_n = b_rp_n.as<uint64_t>();
    net.meta_n = b_meta_n.as<uint32_t>();
    net.w_n = b_w_n.as<float>();
    net.nmda_out = b_nmda_out.as<uint32_t>();
    net.total = b_total.as<uint32_t>();
    net.overflow = b_overflow.as<uint32_t>();
    net.lcnt = b_lcnt.as<uint32_t>();
    net.lbase = b_lbase.as<uint32_t>();
    net.t0 = b_t0.as<int64_t>();
    net.tmod = b_tmod.as<int32_t>();
    net.pop_cnt = b_popcnt.as<unsigned long long>();
    net.counters = b_counters.as<unsigned long long>();
    net.pbr = b_pbr.as<float>();
    net.gainr = b_gain.as<float>();
    net.skey = b_skey.as<uint64_t>();
    net.ext_bits = nullptr;
    net.ext_first = 0;
    net.ext_steps = 0;
    net.irep = nullptr;
    net.rp_g = nullptr;
    net.col_g = nullptr;
    net.w_g = nullptr;
    net.igap = nullptr;
    net.ltp_mode = 0;
    net.lp = nullptr;
    net.cntlp = nullptr;
    net.gl = nullptr;
    net.n_learn = 0;
    net.l_pre = net.l_post = nullptr;
    net.l_rec = nullptr;
    net.l_w = nullptr;
    // default schedule (the autotuner may replace it)
    sched.update_block = 256;
    sched.gather_lanes = 8;
    sched.push_block = 256;
    // steps per graph: 100, or fewer where the spike record of the
    // one-step-per-launch paths (every neuron every step at worst) would
    // pass 2^26 entries
    sched.steps_per_graph = (int32_t)std::max<int64_t>(1, std::min<int64_t>(100, ((int64_t)1 << 26) / N));
    sched.use_graph = 1;
    sched.fused = 1;
    // the windowed path when synapses have delays (several steps per
    // launch), with graphs of whole windows
    sched.window = win_max >= 2 ? 64 : 0;
    if (sched.window) {
      // up to 100 windows per graph: each batch ends in a host round trip
      // (overflow check, spike readback), so long batches amortise it; with
      // spike recording the record (worst case every neuron every step)
      // stays within 2^26 entries, down to one window per graph for the
      // largest networks
      const int64_t w = std::min(sched.window, win_max);
      const int64_t per =
          record ? std::max<int64_t>(1, std::min<int64_t>(100, ((int64_t)1 << 26) / ((int64_t)N * w))) : 100;
      sched.steps_per_graph = (int32_t)(per * w);
    }
    apply_schedule();
    reset(0);
  }

  void apply_schedule() {
    const int ub = sched.update_block;
    if (ub < 64 || ub > 1024 || (ub & 63)) throw std::invalid_argument("update_block must be 64..1024, multiple of 64");
    const int L = sched.gather_lanes;
    if (L < 1 || L > 64 || (L & (L - 1))) throw std::invalid_argument("gather_lanes must be a power of two <= 64");
    if (sched.push_block < 64 || sched.push_block > 1024 || (sched.push_block & 63))
      throw std::invalid_argument("push_block must be 64..1024, multiple of 64");
    if (sched.steps_per_graph < 1 || sched.steps_per_graph > 100000)
      throw std::invalid_argument("steps_per_graph out of range");
    if (window() && ub > 512) throw std::invalid_argument("the windowed path takes update_block <= 512");
    if (sched.gather_chunk < 0 || (sched.gather_chunk > 0 && sched.gather_chunk < 64))
      throw std::invalid_argument("gather_chunk must be 0 or >= 64");
    if (sched.xcd_order < 0 || sched.xcd_order > 1) throw std::invalid_argument("xcd_order must be 0 or 1");
    // The windowed path writes the record only for readback, so a handle
    // created without spike recording needs none there; the other paths
    // push from it (it is their per-step spike list).
    const bool rec_needed = record || !window();
    if (rec_needed && (uint64_t)N * (uint64_t)sched.steps_per_graph > ((uint64_t)1 << 27))
      throw std::invalid_argument("steps_per_graph too large for the spike record of this network "
                                  "(neurons x steps_per_graph > 2^27)");
    if (sched.gather_split < 0 || sched.gather_split > L || (sched.gather_split & (sched.gather_split - 1)))
      throw std::invalid_argument("gather_split must be 0 or a power of two <= gather_lanes");
    net.gather_lanes = L;
    net.gather_chunk = sched.gather_chunk;
    net.gather_split = std::max(1, sched.gather_split);
    const int K = sched.steps_per_graph;
    // record buffer: every neuron could hold SPIKE_F every step (sticky
    // case).  Block b of the update grid appends to sub-list b % NSUB, whose
    // region holds its blocks' neurons for K steps.
    const int64_t npb = (fused() || window()) ? ub / L : ub;  // neurons per block
    std::vector<uint64_t> cov(fsk::NSUB, 0);
    if (window()) {
      build_window_blocks((int32_t)npb);
      for (size_t b = 0; b < hwblk.size(); ++b) cov[b % fsk::NSUB] += hwblk[b] >> 32;
    } else {
      const int64_t nblk = ((int64_t)N + npb - 1) / npb;
      for (int64_t b = 0; b < nblk; ++b) cov[b % fsk::NSUB] += std::min<int64_t>(npb, N - b * npb);
    }
    hsub_off.assign(fsk::NSUB, 0u);
    hsub_cap.assign(fsk::NSUB, 0u);
    uint64_t acc = 0;
    for (int s = 0; s < fsk::NSUB; ++s) {
      const uint64_t c = rec_needed ? cov[s] * (uint64_t)K : 0;
      hsub_off[s] = (uint32_t)acc;
      hsub_cap[s] = (uint32_t)c;
      acc += c;
    }
    if (b_rec_ids.bytes < acc * 4 || b_rec_a.bytes < acc * sizeof(T) || b_rec_k.bytes < acc * 4) {
      b_rec_ids.alloc(acc * 4);
      b_rec_a.alloc(acc * sizeof(T));
      b_rec_k.alloc(acc * 4);
    }
    HIPCHECK(hipMemcpy(b_suboff.p, hsub_off.data(), fsk::NSUB * 4, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b_subcap.p, hsub_cap.data(), fsk::NSUB * 4, hipMemcpyHostToDevice));
    net.sub_off = b_suboff.as<uint32_t>();
    net.sub_cap = b_subcap.as<uint32_t>();
    net.rec_cap = acc;
    net.rec_ids = b_rec_ids.as<uint32_t>();
    net.rec_a = b_rec_a.as<T>();
    net.rec_k = b_rec_k.as<uint32_t>();
    const size_t cum_bytes = (size_t)(K + 1) * fsk::NSUB * 4;
    if (b_cum.bytes < cum_bytes) b_cum.alloc(cum_bytes);
    net.cum = b_cum.as<uint32_t>();
    graph_dirty = true;
  }

  int update_grid() const { return (N + sched.update_block - 1) / sched.update_block; }

  // Blocks of the windowed path: runs of neurons of one population (global
  // order), cut into pieces of at most npb, so every block's update lanes
  // share one parameter block and update in one pass; where that would take
  // more than 1.25 times the blocks of plain npb-neuron pieces (many small
  // populations), plain pieces (the update then runs once per population
  // present in a wave).  With one network copy a block's NMDA gather time
  // follows its entry count, so a piece also ends before it would hold more
  // than the whole chunks of gather_block that the mean plain piece needs
  // (a single row may).  wgat: each block's range of gather entries.
  void build_window_blocks(int32_t npb) {
    std::vector<uint64_t> runs;
    const int64_t plain = ((int64_t)N + npb - 1) / npb;
    const bool bal = B == 1 && hrp_n.size() == (size_t)N0 + 1 && hrp_n[N0] > 0;
    uint64_t cap = 0;
    if (bal) {
      const double mean = (double)hrp_n[N0] / (double)plain;
      const double ch = (double)fsk::window_chunk<T>(std::max(1, window()), sched.gather_chunk);
      cap = (uint64_t)(ch * std::max(1.0, std::ceil(mean / ch)));
    }
    int64_t g = 0, pieces = 0;
    while (g < N) {
      const uint32_t p = hnpop[g / B];
      int64_t e = g + 1;
      while (e < N && hnpop[e / B] == p) ++e;
      pieces += (e - g + npb - 1) / npb;
      for (int64_t a = g; a < e;) {
        int64_t c = std::min<int64_t>(npb, e - a);
        if (bal)
          for (int64_t k = 1; k < c; ++k)
            if (hrp_n[a + k + 1] - hrp_n[a] > cap) { c = k; break; }
        runs.push_back((uint64_t)a | ((uint64_t)c << 32));
        a += c;
      }
      g = e;
    }
    if (pieces * 4 > plain * 5) {
      runs.clear();
      for (int64_t a = 0; a < N; a += npb)
        runs.push_back((uint64_t)a | ((uint64_t)std::min<int64_t>(npb, N - a) << 32));
    }
    if (sched.xcd_order && runs.size() >= 16) {
      // block i runs on XCD i % 8: give XCD x the x-th eighth of the pieces
      // in order, so the blocks one XCD runs at a time are neighbours
      const size_t nbk = runs.size(), qx = nbk / 8, rx = nbk % 8;
      std::vector<uint64_t> perm(nbk);
      for (size_t i = 0; i < nbk; ++i) {
        const size_t x = i % 8, j = i / 8;
        perm[i] = runs[x * qx + std::min(x, rx) + j];
      }
      runs.swap(perm);
    }
    hwblk.swap(runs);
    b_wblk.alloc(hwblk.size() * 8);
    HIPCHECK(hipMemcpy(b_wblk.p, hwblk.data(), hwblk.size() * 8, hipMemcpyHostToDevice));
    net.wblk = b_wblk.as<uint64_t>();
    // each block's population, or 0xFFFFFFFF where its neurons span several
    std::vector<uint32_t> bp(hwblk.size());
    for (size_t b = 0; b < hwblk.size(); ++b) {
      const int64_t a = (int64_t)(hwblk[b] & 0xFFFFFFFFu), c = (int64_t)(hwblk[b] >> 32);
      uint32_t p = hnpop[a / B];
      for (int64_t g = a + 1; g < a + c && p != 0xFFFFFFFFu; ++g)
        if (hnpop[g / B] != p) p = 0xFFFFFFFFu;
      bp[b] = p;
    }
    b_wpop.alloc(bp.size() * 4);
    HIPCHECK(hipMemcpy(b_wpop.p, bp.data(), bp.size() * 4, hipMemcpyHostToDevice));
    net.wpop = b_wpop.as<uint32_t>();
    net.wgat = nullptr;
    if (B == 1 && hrp_n.size() == (size_t)N0 + 1) {
      std::vector<uint64_t> ge(2 * hwblk.size());
      for (size_t b = 0; b < hwblk.size(); ++b) {
        const uint64_t a = hwblk[b] & 0xFFFFFFFFu, c = hwblk[b] >> 32;
        ge[2 * b] = hrp_n[a];
        ge[2 * b + 1] = hrp_n[a + c];
      }
      b_wgat.alloc(ge.size() * 8);
      HIPCHECK(hipMemcpy(b_wgat.p, ge.data(), ge.size() * 8, hipMemcpyHostToDevice));
      net.wgat = b_wgat.as<uint64_t>();
    }
  }
  // fused step: a group of gather_lanes lanes per neuron
  int step_grid() const {
    const int64_t threads = (int64_t)N * sched.gather_lanes;
    return (int)((threads + sched.update_block - 1) / sched.update_block);
  }
  // fused step: push blocks (update_block threads each) ahead of the
  // gather + update blocks, enough waves for the typical spike count
  int step_push_blocks() const {
    const int wpb = sched.update_block / 64;
    const int pw = std::max(64, std::min(4096, N / 64));
    return (pw + wpb - 1) / wpb;
  }
  int deliver_grid(int *push_waves) const {
    const int wpb = sched.push_block / 64;
    // push waves: enough for the typical spike count of a step (spikes are
    // strided over them); gather waves: enough rows for all CUs
    int pw = std::max(64, std::min(4096, N / 64));
    int gw = 0;
    if (net.has_nmda_syn) {
      const int groups_per_wave = 64 / sched.gather_lanes;
      gw = std::max(wpb, std::min(16384, (N + groups_per_wave - 1) / groups_per_wave));
    }
    pw = (pw + wpb - 1) / wpb * wpb;
    gw = (gw + wpb - 1) / wpb * wpb;
    *push_waves = pw;
    return (pw + gw) / wpb;
  }

  void launch_step(int32_t k, bool timed) {
    int pw;
    const int dgrid = deliver_grid(&pw);
    net.push_waves = pw;
    if (timed) HIPCHECK(hipEventRecord(ev[0], stream));
    // replicas in whole waves: one local neuron per wave (scalar-load variants)
    const bool uni = (B % 64) == 0;
    const bool uni_row = uni && sched.gather_lanes == 1;
    const bool rp = net.ext_mode == FS_EXT_REPLAY;
    if (fused()) {
      net.push_blocks = step_push_blocks();
      const dim3 g(net.push_blocks + step_grid()), b(sched.update_block);
      if (rp && uni_row) hipLaunchKernelGGL((fsk::k_step<T, true, true>), g, b, 0, stream, net, k);
      else if (rp) hipLaunchKernelGGL((fsk::k_step<T, true, false>), g, b, 0, stream, net, k);
      else if (uni_row) hipLaunchKernelGGL((fsk::k_step<T, false, true>), g, b, 0, stream, net, k);
      else hipLaunchKernelGGL((fsk::k_step<T, false, false>), g, b, 0, stream, net, k);
      HIPCHECK(hipGetLastError());
      if (timed) {
        HIPCHECK(hipEventRecord(ev[1], stream));
        HIPCHECK(hipEventRecord(ev[2], stream));
      }
      return;
    }
    {
      const dim3 g(update_grid()), b(sched.update_block);
      if (rp && uni) hipLaunchKernelGGL((fsk::k_update<T, true, true>), g, b, 0, stream, net, k);
      else if (rp) hipLaunchKernelGGL((fsk::k_update<T, true, false>), g, b, 0, stream, net, k);
      else if (uni) hipLaunchKernelGGL((fsk::k_update<T, false, true>), g, b, 0, stream, net, k);
      else hipLaunchKernelGGL((fsk::k_update<T, false, false>), g, b, 0, stream, net, k);
    }
    HIPCHECK(hipGetLastError());
    if (net.igap) {  // gap currents from this step's potentials, for the next step
      hipLaunchKernelGGL((fsk::k_gap<T>), dim3((N + 255) / 256), dim3(256), 0, stream, net);
      HIPCHECK(hipGetLastError());
    }
    if (net.n_learn) {  // single synapses: sums with the old weights, then the LTP rules
      hipLaunchKernelGGL((fsk::k_learn<T>), dim3((unsigned)((net.n_learn + 255) / 256)), dim3(256), 0,
                         stream, net, k);
      HIPCHECK(hipGetLastError());
    }
    if (timed) HIPCHECK(hipEventRecord(ev[1], stream));
    if (uni_row)
      hipLaunchKernelGGL((fsk::k_deliver<T, true>), dim3(dgrid), dim3(sched.push_block), 0, stream, net, k);
    else
      hipLaunchKernelGGL((fsk::k_deliver<T, false>), dim3(dgrid), dim3(sched.push_block), 0, stream, net, k);
    HIPCHECK(hipGetLastError());
    if (timed) HIPCHECK(hipEventRecord(ev[2], stream));
  }

  // windowed path: steps k0 .. k0+w-1 in one launch
  void launch_window(int32_t k0, int32_t w, bool timed) {
    if (timed) HIPCHECK(hipEventRecord(ev[0], stream));
    const dim3 g((unsigned)hwblk.size()), b(sched.update_block);
    net.inj_on = inj_any() ? 1 : 0;
    const size_t lds = fsk::window_lds_bytes<T>(sched.update_block / sched.gather_lanes, w,
                                                __builtin_popcount((uint32_t)net.lin_mask), B == 1,
                                                sched.gather_chunk, net.inj_on != 0);
    // replicas in whole waves with one lane per row: the row tables are wave-uniform
    const bool uni = (B % 64) == 0 && sched.gather_lanes == 1;
    // the AMPA + GABA (+ NMDA) layout with external AMPA drive and no STP
    const bool prof = net.lin_mask == 3 && net.ext_mask == 1 && net.stp_mode == 0;
    // every wave issues the prologue's loads when that duplicates them at
    // most 8 times (the replicas' instance, one lane per row, always)
    const bool dup = sched.gather_lanes <= 8;
#define FS_WIN(R, U, DUP)                                                                          \
  do {                                                                                             \
    if (prof) hipLaunchKernelGGL((fsk::k_window<T, R, U, 1, DUP>), g, b, lds, stream, net, k0, w);   \
    else hipLaunchKernelGGL((fsk::k_window<T, R, U, 0, DUP>), g, b, lds, stream, net, k0, w);        \
  } while (0)
    if (net.ext_mode == FS_EXT_REPLAY) {
      if (uni) FS_WIN(true, true, true);
      else if (dup) FS_WIN(true, false, true);
      else FS_WIN(true, false, false);
    } else {
      if (uni) FS_WIN(false, true, true);
      else if (dup) FS_WIN(false, false, true);
      else FS_WIN(false, false, false);
    }
#undef FS_WIN
    HIPCHECK(hipGetLastError());
    if (timed) {
      HIPCHECK(hipEventRecord(ev[1], stream));
      HIPCHECK(hipEventRecord(ev[2], stream));
    }
  }

  // fused path: the spikes of the last step of a batch of K launches are
  // pushed by this launch
  void launch_tail(int32_t K) {
    int pw;
    (void)deliver_grid(&pw);
    const int wpb = sched.push_block / 64;
    hipLaunchKernelGGL((fsk::k_push_tail<T>), dim3((pw + wpb - 1) / wpb), dim3(sched.push_block), 0,
                       stream, net, K);
    HIPCHECK(hipGetLastError());
  }

  void build_graph(int32_t K) {
    if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
    if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
    HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    if (const int32_t W = window()) {
      for (int32_t k0 = W; k0 < K; k0 += W) launch_window(k0, std::min(W, K - k0), false);
    } else {
      for (int32_t k = 1; k < K; ++k) launch_step(k, false);
      if (fused()) launch_tail(K);
    }
    HIPCHECK(hipStreamEndCapture(stream, &graph));
    HIPCHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    graph_k = K;
    graph_dirty = false;
  }

  // Irep recording (IsynSeparatedSum output): a [5][N] buffer the update
  // writes each non-refractory step
  void record_currents(bool on) override {
    if (on && !net.irep) {
      b_irep.alloc((size_t)5 * N * sizeof(T));
      HIPCHECK(hipMemset(b_irep.p, 0, (size_t)5 * N * sizeof(T)));
      net.irep = b_irep.as<T>();
    } else if (!on) {
      net.irep = nullptr;
    }
    graph_dirty = true;  // the buffer is a kernel argument
  }

  void reset(uint64_t seed) override {
    const size_t n = (size_t)N;
    auto fill = [&](void *p, size_t count, T v) {
      if (!count) return;
      hipLaunchKernelGGL((fsk::k_fill<T>), dim3(1024), dim3(256), 0, stream, (T *)p, v, count);
      HIPCHECK(hipGetLastError());
    };
    fill(b_V.p, n, T(-70.0f));  // VRESTING (code/EqsGnl2.cpp:347)
    HIPCHECK(hipMemsetAsync(b_cs.p, 0, n * 4, stream));
    fill(b_G.p, 4 * n, T(0));
    fill(b_ring.p, (size_t)4 * net.D * n, T(0));
    fill(b_spre.p, n, T(0));
    fill(b_xn.p, n, T(0));
    fill(b_shist.p, (size_t)net.DN * n, T(0));
    fill(b_gn.p, n, T(0));
    fill(b_elin.p, 4 * n, T(0));
    fill(b_enx.p, n, T(0));
    fill(b_ens.p, n, T(0));
    if (net.irep) fill(net.irep, (size_t)5 * n, T(0));  // Irep = 0 (code/NetSimGnl2_BulPp.cpp:204-207)
    if (net.igap) fill(net.igap, n, T(0));              // Isyn = 0 (code/EqsGnl2.cpp:346)
    if (net.ltp_mode) {  // VarLtp (code/EqsGnl2.cpp:365-369); weights back to their initial values
      fill(net.lp, n, T(0));
      fill(net.gl, (size_t)5 * n, T(0));
      HIPCHECK(hipMemsetAsync(net.cntlp, 0, n * 4, stream));
      if (net.n_learn)
        HIPCHECK(hipMemcpyAsync(net.l_w, b_lw0.p, (size_t)net.n_learn * sizeof(T),
                                hipMemcpyDeviceToDevice, stream));
    }
    if (net.stp_mode) {  // VarStp (code/EqsGnl2.cpp:353-362): D = F = 1, Cap = 0, Car = 20 uM
      fill(net.sD, n, T(1));
      fill(net.sF, n, T(1));
      fill(net.sCap, n, T(0));
      fill(net.sCar, n, T(20.0f));
    }
    HIPCHECK(hipMemsetAsync(b_popcnt.p, 0, (size_t)P * B * 8, stream));
    HIPCHECK(hipMemsetAsync(b_counters.p, 0, 3 * fsk::NCTR * 8, stream));
    HIPCHECK(hipMemsetAsync(b_overflow.p, 0, 4, stream));
    HIPCHECK(hipMemsetAsync(b_lcnt.p, 0, 3 * fsk::NSUB * 4, stream));
    HIPCHECK(hipMemsetAsync(b_lbase.p, 0, 3 * fsk::NSUB * 4, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    net.seed = seed;
    graph_dirty = true;  // the seed is a kernel argument
    step = 0;
    zig_calls = 0;
    zig_a = 0;
    zig_t0 = 0;
    zig_refresh();
    std::memset(ctr_last, 0, sizeof(ctr_last));
  }

  // Population g = replica * P + p0.  The parameters of p0 are shared by
  // every replica; the external probabilities and activity bits are set for
  // replica g / P only.
  void set_pop(int32_t g, const fs_pop_params &p) override {
    if (g < 0 || g >= P * B) throw std::invalid_argument("population index out of range");
    const int32_t rep = g / P, i = g - rep * P;
    const bool inj_switch = ((hpop[i].act ^ p.act) & FS_ACT_INJ) != 0;
    if (inj_switch && (p.act & FS_ACT_INJ) && net.ext_mode == FS_EXT_REPLAY && (B != 1 || !pops_contiguous))
      throw std::invalid_argument(
          "replay-mode Gaussian injection needs one instance whose populations are contiguous neuron ranges");
    hpop[i] = p;
    dpop[i] = to_dev(p, dpop[i]);
    HIPCHECK(hipMemcpyAsync(b_pops.as<fsk::DevPop<T>>() + i, &dpop[i], sizeof(fsk::DevPop<T>),
                            hipMemcpyHostToDevice, stream));
    if (inj_switch) {
      zig_dirty = true;    // refreshed once before the next step (events come in bulk)
      graph_dirty = true;  // the windowed kernel's LDS holds injection draws only while some is on
    }
    pbr_from_pop(i, rep, p);
    for (int r = 0; r < 5; ++r) note_pb(r, pbr(r, i, rep));
    tables_dirty = true;
  }
  fs_pop_params pop(int32_t g) const override {
    if (g < 0 || g >= P * B) throw std::invalid_argument("population index out of range");
    const int32_t rep = g / P, i = g - rep * P;
    fs_pop_params p = hpop[i];
    static const uint32_t bit[5] = {FS_ACT_AMPA, FS_ACT_GABA, FS_ACT_NMDA, FS_ACT_ACH, FS_ACT_GCL};
    for (int r = 0; r < 5; ++r) {
      const float v = hpbr[((size_t)r * P + i) * B + rep];
      if (v < 0.0f) p.act &= ~bit[r];
      else { p.act |= bit[r]; p.pb[r] = v; }
    }
    return p;
  }
  // ChangeExtFreq of one receptor of population g (code/NetSimGnl2.cpp:60-75):
  // on = false clears the activity bit, else pb is the new probability
  void set_ext(int32_t g, int32_t r, float pb, bool on) override {
    if (g < 0 || g >= P * B) throw std::invalid_argument("population index out of range");
    const int32_t rep = g / P, i = g - rep * P;
    pbr(r, i, rep) = on ? pb : -1.0f;
    if (B == 1) {  // keep the parameter block in step for fs_get/set round trips
      static const uint32_t bit[5] = {FS_ACT_AMPA, FS_ACT_GABA, FS_ACT_NMDA, FS_ACT_ACH, FS_ACT_GCL};
      if (on) { hpop[i].act |= bit[r]; hpop[i].pb[r] = pb; }
      else hpop[i].act &= ~bit[r];
    }
    if (on) note_pb(r, pb);
    tables_dirty = true;
  }
  void set_replica(int32_t rep, const fs_replica_params &p) override {
    if (rep < 0 || rep >= B) throw std::invalid_argument("replica index out of range");
    for (int r = 0; r < 5; ++r) hgain[(size_t)r * B + rep] = p.gain[r];
    hgain[(size_t)5 * B + rep] = p.ext_gain;
    hskey[rep] = p.stream_key;
    tables_dirty = true;
  }

  void load_ext(const uint32_t *bits, int64_t first, int64_t nsteps) override {
    if (net.ext_mode != FS_EXT_REPLAY) throw std::invalid_argument("engine not created in replay mode");
    const size_t bytes = (size_t)nsteps * 5 * net.words * 4;
    b_ext.alloc(bytes);
    HIPCHECK(hipMemcpy(b_ext.p, bits, bytes, hipMemcpyHostToDevice));
    net.ext_bits = b_ext.as<uint32_t>();
    net.ext_first = first;
    net.ext_steps = nsteps;
    graph_dirty = true;
  }

  void run(int64_t nsteps, int64_t *off, uint32_t *ids, int64_t cap) override {
    if (net.ext_mode == FS_EXT_REPLAY) {
      if (!net.ext_bits || step < net.ext_first || step + nsteps > net.ext_first + net.ext_steps)
        throw std::invalid_argument("replayed external trains do not cover the requested steps");
    }
    if (tables_dirty) {
      upload_replica_tables();
      tables_dirty = false;
    }
    if (zig_dirty) zig_refresh();
    int64_t written = 0;
    if (off) off[0] = 0;
    // the windowed kernel writes the spike record only when it is read back
    const int32_t rec_on = (off || ids) ? 1 : 0;
    if (rec_on && !record)
      throw std::invalid_argument("spike readback needs a handle created with record_spikes = 1");
    if (net.rec_on != rec_on) {
      net.rec_on = rec_on;
      graph_dirty = true;  // a kernel argument
    }
    t_upd = t_del = t_graph = 0;
    n_timed = 0;
    n_graph_steps = 0;
    HIPCHECK(hipMemsetAsync(b_counters.p, 0, 3 * fsk::NCTR * 8, stream));
    int64_t done = 0;
    while (done < nsteps) {
      const int32_t K = (int32_t)std::min<int64_t>(sched.steps_per_graph, nsteps - done);
      hipLaunchKernelGGL(fsk::k_batch_begin, dim3(1), dim3(fsk::NSUB), 0, stream, b_t0.as<int64_t>(), step,
                         b_total.as<uint32_t>(), b_cum.as<uint32_t>(), b_lcnt.as<uint32_t>(),
                         b_lbase.as<uint32_t>(), b_tmod.as<int32_t>(), net.D, net.DN);
      HIPCHECK(hipGetLastError());
      const int32_t W = window();
      const int32_t w0 = W ? std::min(W, K) : 1;  // steps of the eager first launch
      if (W) launch_window(0, w0, timing);
      else launch_step(0, timing);
      if (K > w0) {
        if (sched.use_graph) {
          if (graph_dirty || graph_k != K) build_graph(K);
          if (timing) HIPCHECK(hipEventRecord(ev[3], stream));
          HIPCHECK(hipGraphLaunch(gexec, stream));
          if (timing) HIPCHECK(hipEventRecord(ev[4], stream));
        } else if (W) {
          for (int32_t k0 = W; k0 < K; k0 += W) launch_window(k0, std::min(W, K - k0), false);
        } else {
          for (int32_t k = 1; k < K; ++k) launch_step(k, false);
          if (fused()) launch_tail(K);
        }
      } else if (fused()) {
        launch_tail(1);
      }
      HIPCHECK(hipStreamSynchronize(stream));
      if (timing) {
        float a = 0, b = 0;
        HIPCHECK(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIPCHECK(hipEventElapsedTime(&b, ev[1], ev[2]));
        t_upd += a;
        t_del += b;
        n_timed++;
        if (K > w0 && sched.use_graph) {
          float c = 0;
          HIPCHECK(hipEventElapsedTime(&c, ev[3], ev[4]));
          t_graph += c;
          n_graph_steps += K - w0;
        }
      }
      // from here on the batch has run: a failure still advances the step
      // index, so that it stays the index of the device state
      auto batch_fail = [&](const char *msg) {
        step += K;
        throw std::length_error(msg);
      };
      uint32_t ovf = 0;
      HIPCHECK(hipMemcpy(&ovf, b_overflow.p, 4, hipMemcpyDeviceToHost));
      if (ovf) batch_fail("device spike record overflow; lower steps_per_graph");
      if ((off || ids) && W) {
        // windowed path: sub-list s holds total[s] entries (id, step) in
        // arrival order; bucket them by step, ascending ids within a step
        const int S = fsk::NSUB;
        std::vector<uint32_t> tot_s(S);
        HIPCHECK(hipMemcpy(tot_s.data(), b_total.p, S * 4, hipMemcpyDeviceToHost));
        uint64_t tot = 0;
        for (int s = 0; s < S; ++s) tot += std::min(tot_s[s], hsub_cap[s]);
        stage.resize(tot);
        stage_k.resize(tot);
        uint64_t at = 0;
        for (int s = 0; s < S; ++s) {
          const uint32_t ns = std::min(tot_s[s], hsub_cap[s]);
          if (!ns) continue;
          HIPCHECK(hipMemcpy(stage.data() + at, b_rec_ids.as<uint32_t>() + hsub_off[s], (size_t)ns * 4,
                             hipMemcpyDeviceToHost));
          HIPCHECK(hipMemcpy(stage_k.data() + at, b_rec_k.as<uint32_t>() + hsub_off[s], (size_t)ns * 4,
                             hipMemcpyDeviceToHost));
          at += ns;
        }
        std::vector<int64_t> kst(K + 1, 0);
        for (uint64_t i = 0; i < tot; ++i) {
          if (stage_k[i] >= (uint32_t)K) batch_fail("corrupt spike record");
          kst[stage_k[i] + 1]++;
        }
        for (int32_t k = 0; k < K; ++k) kst[k + 1] += kst[k];
        if (ids) {
          if (written + (int64_t)tot > cap) batch_fail("spike id buffer too small");
          std::vector<int64_t> fill(kst.begin(), kst.end() - 1);
          for (uint64_t i = 0; i < tot; ++i) ids[written + fill[stage_k[i]]++] = stage[i];
          for (int32_t k = 0; k < K; ++k) std::sort(ids + written + kst[k], ids + written + kst[k + 1]);
        }
        if (off)
          for (int32_t k = 0; k < K; ++k) off[done + k + 1] = written + kst[k + 1];
        written += (int64_t)tot;
      } else if (off || ids) {
        // cum[k][s]: offset of step k's spikes within sub-list s
        const int S = fsk::NSUB;
        cum_h.resize((size_t)(K + 1) * S);
        HIPCHECK(hipMemcpy(cum_h.data(), b_cum.p, cum_h.size() * 4, hipMemcpyDeviceToHost));
        uint64_t tot = 0;
        for (int s = 0; s < S; ++s) tot += cum_h[(size_t)K * S + s];
        if (ids) {
          if (written + (int64_t)tot > cap) batch_fail("spike id buffer too small");
          stage.resize(tot);
          std::vector<uint64_t> st(S + 1, 0);
          for (int s = 0; s < S; ++s) {
            const uint32_t ns = cum_h[(size_t)K * S + s];
            st[s + 1] = st[s] + ns;
            if (ns)
              HIPCHECK(hipMemcpy(stage.data() + st[s], b_rec_ids.as<uint32_t>() + hsub_off[s],
                                 (size_t)ns * 4, hipMemcpyDeviceToHost));
          }
          int64_t pos = written;
          for (int32_t k = 0; k < K; ++k) {
            const int64_t p0 = pos;
            for (int s = 0; s < S; ++s) {
              const uint32_t a = cum_h[(size_t)k * S + s], b = cum_h[(size_t)(k + 1) * S + s];
              std::memcpy(ids + pos, stage.data() + st[s] + a, (size_t)(b - a) * 4);
              pos += b - a;
            }
            std::sort(ids + p0, ids + pos);
          }
        }
        if (off) {
          int64_t pos = written;
          for (int32_t k = 0; k < K; ++k) {
            for (int s = 0; s < S; ++s) pos += cum_h[(size_t)(k + 1) * S + s] - cum_h[(size_t)k * S + s];
            off[done + k + 1] = pos;
          }
        }
        written += (int64_t)tot;
      }
      step += K;
      done += K;
    }
    {
      std::vector<unsigned long long> c(3 * fsk::NCTR);
      HIPCHECK(hipMemcpy(c.data(), b_counters.p, c.size() * 8, hipMemcpyDeviceToHost));
      for (int i = 0; i < 3; ++i) {
        ctr_last[i] = 0;
        for (int s = 0; s < fsk::NCTR; ++s) ctr_last[i] += c[(size_t)i * fsk::NCTR + s];
      }
    }
    // every step gathers each NMDA synapse's term once
    ctr_last[3] = (unsigned long long)syn_nmda * (unsigned long long)B * (unsigned long long)nsteps;
  }

  void get_state(int32_t f, double *out) override {
    const size_t n = (size_t)N;
    std::vector<T> tmp(n);
    auto d2h = [&](const void *p) {
      HIPCHECK(hipMemcpy(tmp.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    };
    switch (f) {
      case 0: d2h(b_V.p); break;
      case 1: case 2: {
        std::vector<uint32_t> c(n);
        HIPCHECK(hipMemcpy(c.data(), b_cs.p, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
          if (f == 1) out[i] = (double)(c[i] & fsk::CNT_MASK);
          else {
            uint32_t s = 0;  // reference NState bits (code/Inc.h:225-231)
            if (c[i] & fsk::F_FIRE) s |= 1;
            if (c[i] & fsk::F_SPIKE) s |= 2;
            if (c[i] & fsk::F_LPFIRE) s |= 4;
            if (c[i] & fsk::F_LPSPIKE) s |= 8;
            if (c[i] & fsk::F_OVF) s |= 16;
            if (c[i] & fsk::F_UDF) s |= 32;
            out[i] = (double)s;
          }
        }
        break;
      }
      case 3: case 4: case 6: case 7: {
        const int rid = f - 3;
        const int li = rid < 2 ? rid : rid - 1;
        hipLaunchKernelGGL((fsk::k_read_g<T>), dim3((N + 255) / 256), dim3(256), 0, stream, net,
                           step - 1, li, b_tmp.as<T>());
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(stream));
        d2h(b_tmp.p);
        break;
      }
      case 5:
        if ((fused() || window()) && net.has_nmda_syn) {  // these paths keep gN in registers
          hipLaunchKernelGGL((fsk::k_gather_only<T>), dim3((N + 3) / 4), dim3(256), 0, stream, net,
                             (int32_t)(((step - 1) % net.DN + net.DN) % net.DN));
          HIPCHECK(hipGetLastError());
          HIPCHECK(hipStreamSynchronize(stream));
        }
        d2h(b_gn.p);
        break;
      case 8: d2h(b_spre.p); break;
      case 9: d2h(b_xn.p); break;
      case 10: {
        const int32_t s = (int32_t)(((step - 1) % net.DN + net.DN) % net.DN);
        hipLaunchKernelGGL((fsk::k_read_hist<T>), dim3((N + 255) / 256), dim3(256), 0, stream, net, s,
                           b_tmp.as<T>());
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(stream));
        d2h(b_tmp.p);
        break;
      }
      case 11: case 12: case 13: case 14: d2h(b_elin.as<T>() + (size_t)(f - 11) * n); break;
      case 15: d2h(b_enx.p); break;
      case 16: d2h(b_ens.p); break;
      case 17: case 18:
        if (!net.stp_mode) throw std::invalid_argument("no STP / STD state: the switch is off");
        d2h(f == 17 ? net.sD : net.sF);
        break;
      case 19: case 20: case 21: case 22: case 23:
        if (!net.irep) throw std::invalid_argument("receptor currents are not recorded (fs_record_currents)");
        d2h(net.irep + (size_t)(f - 19) * n);
        break;
      default: throw std::invalid_argument("unknown state field");
    }
  }

  // device layout [P][B] (coalesced atomics), returned replica-major
  void pop_counts(uint64_t *out) override {
    std::vector<uint64_t> c((size_t)P * B);
    HIPCHECK(hipMemcpy(c.data(), b_popcnt.p, c.size() * 8, hipMemcpyDeviceToHost));
    for (int32_t rep = 0; rep < B; ++rep)
      for (int32_t i = 0; i < P; ++i) out[(size_t)rep * P + i] = c[(size_t)i * B + rep];
  }
  void times(double *u, double *p, double *g, int64_t *n) override {
    if (u) *u = n_timed ? t_upd / n_timed : 0.0;
    if (p) *p = n_timed ? t_del / n_timed : 0.0;
    if (g) *g = n_graph_steps ? t_graph / n_graph_steps : 0.0;
    if (n) *n = n_timed;
  }
  void counters(int64_t *out) override {
    for (int i = 0; i < 4; ++i) out[i] = (int64_t)ctr_last[i];
  }
  void stats(int64_t *out) const override {
    const int64_t v[10] = {N0, P, B, syn_lin, syn_nmda, mind, maxdp, maxdn, precision, net.ext_mode};
    for (int i = 0; i < 10; ++i) out[i] = v[i];
  }
};

}  // namespace

struct fs_handle {
  std::unique_ptr<EngineBase> e;
};

#define ENGINE_CALL(...)                                               \
  try {                                                                 \
    __VA_ARGS__;                                                             \
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

int fs_create(const fs_net_desc *d, fs_handle **out) {
  if (!d || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  if (!d->neuron_pop || !d->pops || (d->n_syn > 0 && (!d->pre || !d->post || !d->w || !d->receptor)))
    return fs::set_error(FS_ERR_ARG, "incomplete network description");
  if (!(d->dt > 0.0f)) return fs::set_error(FS_ERR_ARG, "dt must be positive");
  ENGINE_CALL({
    auto h = std::make_unique<fs_handle>();
    if (d->precision == FS_FP64) h->e.reset(new Engine<double>(*d, 1));
    else if (d->precision == FS_FP32) h->e.reset(new Engine<float>(*d, 1));
    else throw std::invalid_argument("precision must be FS_FP32 or FS_FP64");
    *out = h.release();
  })
}

int fs_create_ei_batch(const fs_ei_batch_desc *d, fs_handle **out) {
  if (!d || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  if (!(d->dt > 0.0f)) return fs::set_error(FS_ERR_ARG, "dt must be positive");
  ENGINE_CALL({
    auto h = std::make_unique<fs_handle>();
    if (d->precision == FS_FP64) h->e.reset(new Engine<double>(*d));
    else if (d->precision == FS_FP32) h->e.reset(new Engine<float>(*d));
    else throw std::invalid_argument("precision must be FS_FP32 or FS_FP64");
    *out = h.release();
  })
}

int fs_create_replicated(const fs_net_desc *d, int32_t replicas, fs_handle **out) {
  if (!d || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  if (!d->neuron_pop || !d->pops || (d->n_syn > 0 && (!d->pre || !d->post || !d->w || !d->receptor)))
    return fs::set_error(FS_ERR_ARG, "incomplete network description");
  if (!(d->dt > 0.0f)) return fs::set_error(FS_ERR_ARG, "dt must be positive");
  if (replicas < 1) return fs::set_error(FS_ERR_ARG, "replicas must be >= 1");
  ENGINE_CALL({
    auto h = std::make_unique<fs_handle>();
    if (d->precision == FS_FP64) h->e.reset(new Engine<double>(*d, replicas));
    else if (d->precision == FS_FP32) h->e.reset(new Engine<float>(*d, replicas));
    else throw std::invalid_argument("precision must be FS_FP32 or FS_FP64");
    *out = h.release();
  })
}

int fs_destroy(fs_handle *h) {
  delete h;
  return FS_OK;
}

int fs_reset(fs_handle *h, uint64_t seed) {
  if (!h) return fs::set_error(FS_ERR_ARG, "null handle");
  ENGINE_CALL(h->e->reset(seed))
}

int fs_set_ext_rate(fs_handle *h, int32_t pop, int32_t receptor, float hz) {
  if (!h) return fs::set_error(FS_ERR_ARG, "null handle");
  if (receptor < 0 || receptor > FS_GCL) return fs::set_error(FS_ERR_ARG, "receptor out of range");
  ENGINE_CALL({
    if (hz == 0.0f) h->e->set_ext(pop, receptor, 0.0f, false);
    else h->e->set_ext(pop, receptor, (float)(hz * h->e->dt * 1e-3f), true);
  })
}

int fs_set_ext_rates(fs_handle *h, int32_t receptor, const float *hz, int32_t n) {
  if (!h || !hz) return fs::set_error(FS_ERR_ARG, "null argument");
  if (receptor < 0 || receptor > FS_GCL) return fs::set_error(FS_ERR_ARG, "receptor out of range");
  if (n != h->e->n_pop) return fs::set_error(FS_ERR_ARG, "hz must hold one rate per population");
  ENGINE_CALL({
    for (int32_t g = 0; g < n; ++g) {
      if (hz[g] == 0.0f) h->e->set_ext(g, receptor, 0.0f, false);
      else h->e->set_ext(g, receptor, (float)(hz[g] * h->e->dt * 1e-3f), true);
    }
  })
}

int fs_set_replica(fs_handle *h, int32_t replica, const fs_replica_params *p) {
  if (!h || !p) return fs::set_error(FS_ERR_ARG, "null argument");
  ENGINE_CALL(h->e->set_replica(replica, *p))
}

int fs_set_injection(fs_handle *h, int32_t pop, float mean, float sd) {
  if (!h) return fs::set_error(FS_ERR_ARG, "null handle");
  ENGINE_CALL({
    fs_pop_params p = h->e->pop(pop);
    p.inj_mean = mean;
    p.inj_std = sd * (1.0f + (float)std::log((1.0f - std::exp(-0.1f / p.taum)) / (1.0f - p.df)));
    if (mean == 0.0f && sd == 0.0f) p.act &= ~(uint32_t)FS_ACT_INJ;
    else p.act |= FS_ACT_INJ;
    h->e->set_pop(pop, p);
  })
}

int fs_set_pop_params(fs_handle *h, int32_t pop, const fs_pop_params *p) {
  if (!h || !p) return fs::set_error(FS_ERR_ARG, "null argument");
  ENGINE_CALL(h->e->set_pop(pop, *p))
}

int fs_load_ext_trains(fs_handle *h, const uint32_t *bits, int64_t first, int64_t n) {
  if (!h || !bits || n <= 0 || first < 0) return fs::set_error(FS_ERR_ARG, "bad argument");
  ENGINE_CALL(h->e->load_ext(bits, first, n))
}

int fs_run(fs_handle *h, int64_t n, int64_t *off, uint32_t *ids, int64_t cap) {
  if (!h || n < 0) return fs::set_error(FS_ERR_ARG, "bad argument");
  ENGINE_CALL(h->e->run(n, off, ids, cap))
}

int fs_step_index(fs_handle *h, int64_t *s) {
  if (!h || !s) return fs::set_error(FS_ERR_ARG, "null argument");
  *s = h->e->step;
  return FS_OK;
}

int fs_get_state(fs_handle *h, int32_t f, double *out) {
  if (!h || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  ENGINE_CALL(h->e->get_state(f, out))
}

int fs_pop_spike_counts(fs_handle *h, uint64_t *out) {
  if (!h || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  ENGINE_CALL(h->e->pop_counts(out))
}

int fs_get_schedule(fs_handle *h, fs_schedule *s) {
  if (!h || !s) return fs::set_error(FS_ERR_ARG, "null argument");
  *s = h->e->sched;
  s->fused = h->e->fused_active() ? 1 : 0;
  s->window = h->e->window_active();
  return FS_OK;
}

int fs_set_schedule(fs_handle *h, const fs_schedule *s) {
  if (!h || !s) return fs::set_error(FS_ERR_ARG, "null argument");
  ENGINE_CALL({
    const fs_schedule old = h->e->sched;
    h->e->sched = *s;
    try {
      static_cast<void>(0);
      if (auto *e32 = dynamic_cast<Engine<float> *>(h->e.get())) e32->apply_schedule();
      else if (auto *e64 = dynamic_cast<Engine<double> *>(h->e.get())) e64->apply_schedule();
    } catch (...) {
      // back to the old schedule, re-applied: a failure part-way may have
      // rebuilt tables (window blocks) for the rejected one
      h->e->sched = old;
      if (auto *e32 = dynamic_cast<Engine<float> *>(h->e.get())) e32->apply_schedule();
      else if (auto *e64 = dynamic_cast<Engine<double> *>(h->e.get())) e64->apply_schedule();
      throw;
    }
  })
}

int fs_record_currents(fs_handle *h, int32_t on) {
  if (!h) return fs::set_error(FS_ERR_ARG, "null handle");
  ENGINE_CALL(h->e->record_currents(on != 0))
}

int fs_enable_timing(fs_handle *h, int32_t on) {
  if (!h) return fs::set_error(FS_ERR_ARG, "null handle");
  h->e->timing = on != 0;
  return FS_OK;
}

int fs_kernel_times(fs_handle *h, double *u, double *p, double *g, int64_t *n) {
  if (!h) return fs::set_error(FS_ERR_ARG, "null handle");
  h->e->times(u, p, g, n);
  return FS_OK;
}

int fs_network_stats(fs_handle *h, int64_t out[10]) {
  if (!h || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  h->e->stats(out);
  return FS_OK;
}

int fs_device_info(char *name, int32_t name_cap, int32_t *compute_units) {
  if (!name || name_cap < 1) return fs::set_error(FS_ERR_ARG, "bad argument");
  ENGINE_CALL({
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t p{};
    HIPCHECK(hipGetDeviceProperties(&p, dev));
    std::snprintf(name, (size_t)name_cap, "%s %s", p.name, p.gcnArchName);
    if (compute_units) *compute_units = p.multiProcessorCount;
  })
}

int fs_run_counters(fs_handle *h, int64_t out[4]) {
  if (!h || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  h->e->counters(out);
  return FS_OK;
}

}  // extern "C"
