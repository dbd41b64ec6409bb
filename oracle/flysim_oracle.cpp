// ORACLE -- test infrastructure only; never called by the product.
// CPU restatement of the reference Flysim GNL2 time step.  Every function
// names the reference lines it restates; the header lists the parity pins.
// Built with -ffp-contract=off so float expressions round exactly as the
// reference's x86-64 SSE build (which has no FMA contraction) does.
#include "flysim_oracle.h"

#include <atomic>
#include <chrono>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>
#include <xmmintrin.h>

namespace {

// NState bits (code/Inc.h:225-231) and ActCtr bits (code/Inc.h:253-280).
constexpr uint32_t FIRE_F = 1, SPIKE_F = 2, LP_FIRE_F = 4, LP_SPIKE_F = 8, MEMDV_OVF = 16,
                   MEMDV_UDF = 32;
constexpr uint32_t AMPA_ACT = 1, GABA_ACT = 2, NMDA_ACT = 4, ACH_ACT = 32,
                   GCL_ACT = 64, INJ_ACT = 128, LTP_ACT = 8192, STD_ACT = 16384, STP_ACT = 32768;
// receptor ids of this library: A=0 G=1 N=2 H(ACh)=3 L(GluCl)=4
constexpr int RA = 0, RG = 1, RN = 2, RH = 3, RL = 4;
// parser keywords used by ChangeExtFreq events (code/Inc.h:35-47)
constexpr int KW_AMPA = 0, KW_GABA = 1, KW_NMDA = 2, KW_ACH = 11, KW_GCL = 12;

// Per-neuron parameters exactly as SetPp derives them, in float
// (code/NetSimGnl2_BulPp.cpp:159-278).
struct NeuPar {
  float vth, el, gl, cm, vreset, df, ef, dvmax;
  int32_t rfped, spidy;
  float vrev[5], dfst[5];
  float dfxt_n, efst_n, as_n, mg_n;
  float pb[5];    // reference pb[] order: A, G, ACh, GluCl, NMDA
  float gext[5];  // same order
  float taum;
  // short-term plasticity (ParStp2, code/NetSimGnl2_BulPp.cpp:185-200)
  float pv, dfD, efD, aF, dfF, efF, aCap, dfCap, aCar, dfCar, efCar;
  // long-term plasticity (ParLtp2, code/NetSimGnl2_BulPp.cpp:180-183)
  float dfLP, lpPos, lpNeg;
  int32_t lpSpidy;
};

NeuPar derive(const or_pop &p, float dt) {
  NeuPar q;
  q.vth = p.vth;
  q.el = p.el;
  q.gl = p.cm / p.taum;
  q.cm = p.cm;
  q.vreset = p.vreset;
  q.df = std::exp(-dt / p.taum);
  q.ef = p.taum * (1.0f - q.df);
  q.rfped = p.rfped;
  q.spidy = p.spidy;
  q.taum = p.taum;
  // receptor defaults for undeclared receptors (:208-226)
  q.vrev[RA] = 0.0f;  q.dfst[RA] = 0.1f;
  q.vrev[RG] = -90.0f; q.dfst[RG] = 0.1f;
  q.dvmax = (q.vth - q.vrev[RG]) / q.ef;
  q.as_n = 0.6332f; q.dfxt_n = 0.1f; q.dfst[RN] = 0.1f; q.efst_n = 0.01f;
  q.vrev[RN] = 0.0f; q.mg_n = 1.0f;
  q.vrev[RH] = 0.0f;  q.dfst[RH] = 0.1f;
  q.vrev[RL] = -90.0f; q.dfst[RL] = 0.1f;
  for (int k = 0; k < 5; ++k) { q.pb[k] = 0.0f; q.gext[k] = 0.0f; }
  static const int pbslot[5] = {0, 1, 4, 2, 3};  // receptor id -> pb[] slot
  for (int r = 0; r < 5; ++r) {
    if (!p.declared[r]) continue;
    const float tau = p.tau[r];
    if (r == RG) q.dvmax = (p.vth - p.vrev[r]) / q.ef;
    q.vrev[r] = p.vrev[r];
    q.dfst[r] = std::exp(-dt / tau);
    if (r == RN) {
      q.as_n = 0.6332f;
      q.dfxt_n = std::exp(-dt / p.tau_xt[r]);
      q.efst_n = tau * (1.0f - q.dfst[r]);
      q.mg_n = 1.0f;
    }
    q.pb[pbslot[r]] = p.freq_ext[r] * dt * 1e-3f;
    q.gext[pbslot[r]] = p.ext_eff[r] * p.ext_con[r];
  }
  q.pv = p.stp[0];
  q.dfD = std::exp(-dt / p.stp[1]);
  q.efD = 1.0f - q.dfD;
  q.aF = p.stp[2];
  q.dfF = std::exp(-dt / p.stp[3]);
  q.efF = p.stp[3] * (1.0f - q.dfF);
  q.aCap = p.stp[4];
  q.dfCap = std::exp(-dt / p.stp[5]);
  q.aCar = p.stp[6];
  q.dfCar = std::exp(-dt / p.stp[7]);
  q.efCar = (1.0f - q.dfCar) * 20.0f * 1e-6f;  // CarBase = 20 uM (code/Inc.h:27)
  q.dfLP = std::exp(-dt / p.ltp[0]);
  q.lpSpidy = (int32_t)p.ltp[1];
  q.lpPos = p.ltp[2];
  q.lpNeg = p.ltp[3];
  return q;
}

// generate_canonical<float,24>(mt19937): one 32-bit draw, converted to float
// (round to nearest), scaled by 2^-32, clamped below 1
// (/usr/include/c++/11/bits/random.tcc:3348-3379).
inline float canonical_f(std::mt19937 &g) {
  const float x = static_cast<float>(static_cast<unsigned long>(g()));
  float r = x / 4294967296.0f;
  if (r >= 1.0f) r = std::nextafter(1.0f, 0.0f);
  return r;
}

// Marsaglia-Tsang 128-strip ziggurat as the reference uses it
// (code/StiExt.h:191-353): LCG-seeded SHR3.
struct Zig {
  uint32_t jcong = 0, jsr = 0;
  uint32_t kn[128];
  float fn[128], wn[128];
  Zig() {
    double dn = 3.442619855899, tn = dn, q;
    const double m1 = 2147483648.0, vn = 9.91256303526217E-03;
    q = vn / std::exp(-0.5 * dn * dn);
    kn[0] = (uint32_t)((dn / q) * m1);
    kn[1] = 0;
    wn[0] = (float)(q / m1);
    wn[127] = (float)(dn / m1);
    fn[0] = 1.0f;
    fn[127] = (float)(std::exp(-0.5 * dn * dn));
    for (int i = 126; 1 <= i; i--) {
      dn = std::sqrt(-2.0 * std::log(vn / dn + std::exp(-0.5 * dn * dn)));
      kn[i + 1] = (uint32_t)((dn / tn) * m1);
      tn = dn;
      fn[i] = (float)(std::exp(-0.5 * dn * dn));
      wn[i] = (float)(dn / m1);
    }
  }
  static uint32_t shr3(uint32_t &s) {
    uint32_t in = s;
    s ^= s << 13; s ^= s >> 17; s ^= s << 5;
    return in + s;
  }
  static float uni(uint32_t &s) {
    uint32_t in = s;
    s ^= s << 13; s ^= s >> 17; s ^= s << 5;
    return (float)std::fmod(0.5 + (float)(in + s) / 65536.0 / 65536.0, 1.0);
  }
  float nor(uint32_t &s) {
    const float r = 3.442620f;
    int hz = (int)shr3(s);
    uint32_t iz = hz & 127;
    if (std::fabs((double)hz) < kn[iz]) return (float)hz * wn[iz];
    for (;;) {
      float x, y;
      if (iz == 0) {
        for (;;) {
          x = -0.2904764 * std::log(uni(s));
          y = -std::log(uni(s));
          if (x * x <= y + y) break;
        }
        return hz <= 0 ? -r - x : r + x;
      }
      x = (float)hz * wn[iz];
      if (fn[iz] + uni(s) * (fn[iz - 1] - fn[iz]) < std::exp(-0.5 * x * x)) return x;
      hz = (int)shr3(s);
      iz = hz & 127;
      if (std::fabs((double)hz) < kn[iz]) return (float)hz * wn[iz];
    }
  }
  float igauss(float mean, float sd) {
    jcong = 69069u * jcong + 1234567u;
    jsr = jcong;
    return mean + sd * nor(jsr);
  }
};

struct SimBase {
  virtual ~SimBase() {}
  virtual void reset(uint32_t seed) = 0;
  virtual int64_t run(int64_t n, int64_t *ss, int32_t *sn, int64_t cap) = 0;
  virtual void get(int f, double *o) const = 0;
  virtual void get_gates(int f, double *o) const = 0;
  virtual void get_params(int f, double *o) const = 0;
  virtual double time_run(int64_t n, int threads) = 0;
  const uint32_t *replay = nullptr;
  int64_t replay_steps = 0;
  uint32_t *record = nullptr;
  int64_t record_steps = 0;
  int64_t step = 0;
};

template <typename T>
struct Sim final : SimBase {
  // copied network
  int N, P, G, R;
  std::vector<int32_t> npop;
  std::vector<or_pop> pops;
  std::vector<uint32_t> gpre, gpost0, rpost, rbeg, term;
  std::vector<uint8_t> grec, rrec;
  std::vector<float> gg;
  uint32_t act0;
  float dt;
  std::vector<float> et, ev1, ev2;
  std::vector<int32_t> etype, epop, esub;
  std::vector<int64_t> pop_first, pop_n;
  // per neuron
  std::vector<NeuPar> par;
  std::vector<uint32_t> act;
  std::vector<float> pb0[5];      // conf pb for the Run() reset
  std::vector<float> inj_mean, inj_std;
  std::vector<T> V, Isyn, g[5], est[4], ext_xt, ext_st;
  std::vector<int32_t> cnt;
  std::vector<uint32_t> ns;
  std::vector<T> gst, gxt;
  std::vector<T> sD, sF, sCap, sCar;  // presynaptic STP state (VarStp)
  std::vector<uint32_t> gap_pre, gap_post, gap_by_pre;
  std::vector<uint8_t> ggroup;
  std::vector<T> gw;           // gate weights (single synapses change under LTP)
  std::vector<T> lp;           // NetVLTP.LP
  std::vector<int32_t> cntlp;  // NetVLTP.CountLP
  std::vector<float> gap_g;
  // per-gate copies of the postsynaptic decay constants (RPs[PosID])
  std::vector<float> gdf, gdfxt, gefst, gas;
  std::vector<uint8_t> gnm;
  std::mt19937 rd;
  Zig zig;
  unsigned iEvt = 0;

  explicit Sim(const or_net *n) {
    N = n->n_neuron; P = n->n_pop; G = n->n_gate; R = n->n_row;
    npop.assign(n->neuron_pop, n->neuron_pop + N);
    pops.assign(n->pops, n->pops + P);
    gpre.assign(n->gate_pre, n->gate_pre + G);
    gpost0.assign(n->gate_post0, n->gate_post0 + G);
    grec.assign(n->gate_rec, n->gate_rec + G);
    gg.assign(n->gate_g, n->gate_g + G);
    rpost.assign(n->row_post, n->row_post + R);
    rrec.assign(n->row_rec, n->row_rec + R);
    rbeg.assign(n->row_beg, n->row_beg + R + 1);
    term.assign(n->term_gate, n->term_gate + (R ? n->row_beg[R] : 0));
    act0 = n->act_ctr;
    dt = n->dt;
    ggroup.assign(n->gate_group, n->gate_group + G);
    gap_pre.assign(n->gap_pre, n->gap_pre + n->n_gap);
    gap_post.assign(n->gap_post, n->gap_post + n->n_gap);
    gap_g.assign(n->gap_g, n->gap_g + n->n_gap);
    // PreMapCnd1: synapse indices grouped by increasing PreID, in SynCnd1
    // order within a neuron (BulDPreNeuMap, code/Eqs.h:1168-1202)
    gap_by_pre.resize(gap_pre.size());
    for (size_t q = 0; q < gap_by_pre.size(); ++q) gap_by_pre[q] = (uint32_t)q;
    std::stable_sort(gap_by_pre.begin(), gap_by_pre.end(),
                     [&](uint32_t a, uint32_t b) { return gap_pre[a] < gap_pre[b]; });
    et.assign(n->evt_time_s, n->evt_time_s + n->n_evt);
    etype.assign(n->evt_type, n->evt_type + n->n_evt);
    epop.assign(n->evt_pop, n->evt_pop + n->n_evt);
    esub.assign(n->evt_sub, n->evt_sub + n->n_evt);
    ev1.assign(n->evt_var1, n->evt_var1 + n->n_evt);
    ev2.assign(n->evt_var2, n->evt_var2 + n->n_evt);
    pop_first.assign(P, -1);
    pop_n.assign(P, 0);
    for (int i = 0; i < N; ++i) {
      int p = npop[i];
      if (pop_first[p] < 0) pop_first[p] = i;
      pop_n[p]++;
    }
    par.resize(N);
    act.assign(N, act0);
    for (auto &v : pb0) v.resize(N);
    inj_mean.assign(N, 0.0f);
    inj_std.assign(N, 0.0f);
    for (int i = 0; i < N; ++i) {
      par[i] = derive(pops[npop[i]], dt);
      for (int k = 0; k < 5; ++k) pb0[k][i] = par[i].pb[k];
      // MNoiseMean = MNoiseSTD = 0 -> INJ off (code/NetSimGnl2_BulPp.cpp:142-148)
      act[i] &= ~INJ_ACT;
    }
    V.resize(N); Isyn.resize(N); cnt.resize(N); ns.resize(N);
    for (auto &v : g) v.resize(N);
    for (auto &v : est) v.resize(N);
    ext_xt.resize(N); ext_st.resize(N);
    sD.assign(N, T(1)); sF.assign(N, T(1)); sCap.assign(N, T(0)); sCar.assign(N, T(20));
    lp.assign(N, T(0)); cntlp.assign(N, 0);
    gw.resize(G);
    for (int k = 0; k < G; ++k) gw[k] = T(gg[k]);
    gst.assign(G, T(0)); gxt.assign(G, T(0));
    gdf.resize(G); gdfxt.resize(G); gefst.resize(G); gas.resize(G); gnm.resize(G);
    for (int k = 0; k < G; ++k) {
      const NeuPar &q = par[gpost0[k]];
      gnm[k] = grec[k] == RN;
      gdf[k] = q.dfst[grec[k]];
      gdfxt[k] = q.dfxt_n;
      gefst[k] = q.efst_n;
      gas[k] = q.as_n;
    }
  }

  // VarInit (code/EqsGnl2.cpp:339-401) + pb reset and reseeding of Run()
  // (code/NetSimGnl2.cpp:277-323).
  void reset(uint32_t seed) override {
    for (int i = 0; i < N; ++i) {
      Isyn[i] = T(0);
      V[i] = T(-70.0f);  // VRESTING, not RestPot (code/Inc.h:25)
      cnt[i] = 0; ns[i] = 0;
      for (auto &v : g) v[i] = T(0);
      for (auto &v : est) v[i] = T(0);
      ext_xt[i] = T(0); ext_st[i] = T(0);
      for (int k = 0; k < 5; ++k) par[i].pb[k] = pb0[k][i];
      // VarStp (code/EqsGnl2.cpp:353-362); under STD only D is initialised
      // and F stays 1 here (the reference leaves it unset)
      sD[i] = T(1); sF[i] = T(1); sCap[i] = T(0); sCar[i] = T(20.0f);
      lp[i] = T(0); cntlp[i] = 0;  // VarLtp (code/EqsGnl2.cpp:365-369)
    }
    for (int k = 0; k < G; ++k) gw[k] = T(gg[k]);
    std::fill(gst.begin(), gst.end(), T(0));
    std::fill(gxt.begin(), gxt.end(), T(0));
    rd.seed(seed);
    zig.jcong = seed;
    iEvt = 0;
    step = 0;
  }

  // DoEvt (code/NetSimGnl2.cpp:52-141)
  void do_evt(float evtime) {
    for (unsigned i = iEvt; i < et.size(); ++i) {
      const float tms = et[i] * 1e3f;
      if (tms >= evtime && tms < evtime + dt) {
        iEvt = i;
        const int p = epop[i];
        if (p < 0 || p >= P) continue;
        for (int64_t m = pop_first[p]; m < pop_first[p] + pop_n[p]; ++m) {
          if (etype[i] == 1) {
            uint32_t bit; int slot;
            switch (esub[i]) {
              case KW_AMPA: bit = AMPA_ACT; slot = 0; break;
              case KW_GABA: bit = GABA_ACT; slot = 1; break;
              case KW_ACH: bit = ACH_ACT; slot = 2; break;
              case KW_GCL: bit = GCL_ACT; slot = 3; break;
              case KW_NMDA: bit = NMDA_ACT; slot = 4; break;
              default: bit = 0; slot = -1;
            }
            if (slot < 0) continue;
            if (ev1[i] == 0.0f) act[m] &= ~bit;
            else { act[m] |= bit; par[m].pb[slot] = (float)(ev1[i] * dt * 1e-3f); }
          } else if (etype[i] == 2) {
            inj_mean[m] = ev1[i];
            inj_std[m] = ev2[i] * (1.0f + (float)std::log((1.0f - std::exp(-0.1f / pops[p].taum)) /
                                                          (1.0f - par[m].df)));
            if (ev1[i] == 0.0f && ev2[i] == 0.0f) act[m] &= ~INJ_ACT;
            else act[m] |= INJ_ACT;
          }
        }
      } else if (tms >= evtime + dt) {
        break;
      }
    }
  }

  inline bool draw(int r, int slot, int i) {
    // one draw of the external stream for receptor r of neuron i
    if (replay) {
      const int64_t words = (N + 31) / 32;
      if (step >= replay_steps) return false;
      return (replay[(step * 5 + r) * words + i / 32] >> (i % 32)) & 1u;
    }
    const bool b = par[i].pb[slot] > canonical_f(rd);
    if (record && step < record_steps) {
      const int64_t words = (N + 31) / 32;
      if (b) record[(step * 5 + r) * words + i / 32] |= 1u << (i % 32);
    }
    return b;
  }

  // BgInit (code/EqsGnl2.cpp:284-337)
  inline void bginit(int i) {
    const uint32_t a = act[i];
    const NeuPar &q = par[i];
    static const int lin_r[4] = {RA, RG, RH, RL};
    static const uint32_t lin_bit[4] = {AMPA_ACT, GABA_ACT, ACH_ACT, GCL_ACT};
    for (int k = 0; k < 4; ++k) {
      if (!(a & lin_bit[k])) continue;
      const int r = lin_r[k];
      if (draw(r, k, i)) est[k][i] = est[k][i] + T(1);
      else est[k][i] = est[k][i] * T(q.dfst[r]);
      g[r][i] = g[r][i] + T(q.gext[k]) * est[k][i];
    }
    if (a & NMDA_ACT) {
      if (draw(RN, 4, i)) ext_xt[i] = ext_xt[i] + T(1);
      else ext_xt[i] = ext_xt[i] * T(q.dfxt_n);
      ext_st[i] = T(q.dfst[RN]) * ext_st[i] +
                  T(q.efst_n) * T(q.as_n) * ext_xt[i] * (T(1) - ext_st[i]);
      g[RN][i] = g[RN][i] + T(q.gext[4]) * ext_st[i];
    }
    if (a & INJ_ACT) Isyn[i] = Isyn[i] + T(zig.igauss(inj_mean[i], inj_std[i]));
  }

  static inline T texp(T x);

  // firing (code/EqsGnl2.cpp:229-251)
  inline void firing(int i) {
    if (cnt[i] < par[i].rfped) {
      V[i] = cnt[i] == 0 ? T(0) : T(par[i].vreset);
      if (cnt[i] == par[i].spidy) ns[i] |= SPIKE_F;
      else ns[i] &= ~SPIKE_F;
      cnt[i]++;
    } else {
      cnt[i] = 0;
      ns[i] &= ~FIRE_F;
    }
  }

  // VmEul (code/EqsGnl2.cpp:152-225)
  inline void vmeul(int i) {
    const NeuPar &q = par[i];
    if (ns[i] & FIRE_F) {
      firing(i);
    } else {
      const T v = V[i];
      const T i0 = g[RA][i] * (T(q.vrev[RA]) - v);
      const T i1 = g[RG][i] * (T(q.vrev[RG]) - v);
      const T i2 = g[RN][i] * (T(q.vrev[RN]) - v) /
                   (T(1) + T(q.mg_n) * texp(T(-0.062f) * v) / T(3.57f));
      const T i3 = g[RH][i] * (T(q.vrev[RH]) - v);
      const T i4 = g[RL][i] * (T(q.vrev[RL]) - v);
      T I = i0 + i1 + i2 + i3 + i4 + Isyn[i] * T(1e3f) + T(q.gl) * T(q.el) * T(1e3f);
      I = T(1e-3f) * I / T(q.cm);
      const T dvm = T(q.dvmax);
      if (I > dvm) { I = dvm; ns[i] |= MEMDV_OVF; ns[i] &= ~MEMDV_UDF; }
      else if (I < -dvm) { I = -dvm; ns[i] &= ~MEMDV_OVF; ns[i] |= MEMDV_UDF; }
      else { ns[i] &= ~MEMDV_OVF; ns[i] &= ~MEMDV_UDF; }
      V[i] = T(q.df) * V[i] + T(q.ef) * I;
      if (V[i] > T(q.vth)) {
        ns[i] |= FIRE_F;
        firing(i);
        if (act0 & LTP_ACT) ns[i] |= LP_FIRE_F;
      }
    }
    if (act0 & LTP_ACT) {
      // firLP: the LTP spike follows the onset after SpiLPDy steps (:254-272)
      if (ns[i] & LP_FIRE_F) {
        if (cntlp[i] == q.lpSpidy) {
          ns[i] |= LP_SPIKE_F;
          ns[i] &= ~LP_FIRE_F;
          cntlp[i] = 0;
        } else {
          ns[i] &= ~LP_SPIKE_F;
          cntlp[i]++;
        }
      } else {
        ns[i] &= ~LP_SPIKE_F;
      }
      // ltpEul (:274-280)
      if (ns[i] & LP_SPIKE_F) lp[i] = lp[i] + T(1);
      else lp[i] = T(q.dfLP) * lp[i];
    }
    for (auto &v : g) v[i] = T(0);
    Isyn[i] = T(0);
  }

  // ThdPGatVar + ThdSynapse gate rule (code/EqsGnl2.cpp:404-502, 866-932),
  // with STP (D*F, :505-603, :728-794) or STD (D, :605-703, :797-863) in
  // place of the unit increment on a presynaptic spike
  inline T increment(uint32_t pre) const {
    if (act0 & STP_ACT) return sD[pre] * sF[pre];
    if (act0 & STD_ACT) return sD[pre];
    return T(1);
  }
  inline void gate_update(int k) {
    const bool spk = ns[gpre[k]] & SPIKE_F;
    if (gnm[k]) {
      if (spk) gxt[k] = gxt[k] + increment(gpre[k]);
      else gxt[k] = gxt[k] * T(gdfxt[k]);
      gst[k] = T(gdf[k]) * gst[k] + T(gefst[k]) * T(gas[k]) * gxt[k] * (T(1) - gst[k]);
    } else {
      if (spk) gst[k] = gst[k] + increment(gpre[k]);
      else gst[k] = gst[k] * T(gdf[k]);
    }
  }

  // ThdGapSyn (code/EqsGnl2.cpp:705-725): each post's current gathers
  // g (V_pre - V_post) over its synapses (SynCnd1 order, grouped by post),
  // then each pre's current loses g (V_pre - V_post) over its synapses
  // (PreMapCnd1 order); the currents enter the next step's VmEul
  inline void gap_update() {
    const size_t S = gap_pre.size();
    for (size_t q = 0; q < S;) {
      const uint32_t p = gap_post[q];
      T acc = Isyn[p];
      const T vp = V[p];
      for (; q < S && gap_post[q] == p; ++q) acc = acc + T(gap_g[q]) * (V[gap_pre[q]] - vp);
      Isyn[p] = acc;
    }
    for (size_t q = 0; q < S;) {
      const uint32_t a = gap_pre[gap_by_pre[q]];
      T acc = Isyn[a];
      const T va = V[a];
      for (; q < S && gap_pre[gap_by_pre[q]] == a; ++q) {
        const uint32_t s = gap_by_pre[q];
        acc = acc - T(gap_g[s]) * (va - V[gap_post[s]]);
      }
      Isyn[a] = acc;
    }
  }

  // EqsGnl2::LTP (code/EqsGnl2.cpp:1005-1044): single synapses only (the
  // SynFst1 segments and SynSlo1); PreID is what the gate's presynaptic
  // index holds (the population index on this path)
  inline void ltp_update() {
    for (int k = 0; k < G; ++k) {
      if (ggroup[k]) continue;
      const uint32_t a = gpre[k], b = gpost0[k];
      const NeuPar &qa = par[a];
      if (ns[a] & SPIKE_F) {
        if (qa.lpNeg > 0.0f) gw[k] = gw[k] + lp[b] * T(qa.lpNeg);
        else gw[k] = gw[k] * (T(1) + lp[b] * T(qa.lpNeg));
      }
      if (ns[b] & SPIKE_F) {
        if (qa.lpPos > 0.0f) gw[k] = gw[k] + lp[a] * T(qa.lpPos);
        else gw[k] = gw[k] * (T(1) + lp[a] * T(qa.lpPos));
      }
    }
  }

  // EqsGnl2::STP / STD (code/EqsGnl2.cpp:1048-1090), after all gates
  inline void stp_update(int i) {
    const NeuPar &q = par[i];
    const bool spk = ns[i] & SPIKE_F;
    if (act0 & STP_ACT) {
      if (spk) {
        sCar[i] = sCar[i] + T(q.aCar);
        sCap[i] = sCap[i] + T(q.aCap);
        sF[i] = sF[i] + (sCap[i] + sCar[i]) * (T(1) - sF[i]) * T(q.aF);
        sD[i] = sD[i] * (T(1) - T(q.pv) * sF[i]);
      } else {
        sCar[i] = T(q.dfCar) * sCar[i] + T(q.efCar);
        sCap[i] = T(q.dfCap) * sCap[i];
        sF[i] = T(q.dfF) * sF[i] + T(q.efF) * (sCap[i] + sCar[i]) * (T(1) - sF[i]) * T(q.aF);
        sD[i] = T(q.dfD) * sD[i] + T(q.efD);
      }
    } else if (act0 & STD_ACT) {
      if (spk) sD[i] = sD[i] * (T(1) - T(q.pv));
      else sD[i] = T(q.dfD) * sD[i] + T(q.efD);
    }
  }

  // ThdPSynapse / ThdSynapse pull-sum in row order (code/EqsGnl2.cpp:937-1002)
  inline void row_sum(int r) {
    const int rec = rrec[r];
    const uint32_t p = rpost[r];
    T acc = g[rec][p];
    for (uint32_t t = rbeg[r]; t < rbeg[r + 1]; ++t) {
      const uint32_t k = term[t];
      acc = acc + gst[k] * gw[k];
    }
    g[rec][p] = acc;
  }

  void one_step() {
    const float evtime = (float)(unsigned)step * dt;
    do_evt(evtime);
    for (int i = 0; i < N; ++i) { bginit(i); vmeul(i); }  // Activation
    for (int k = 0; k < G; ++k) gate_update(k);             // Spike
    for (int r = 0; r < R; ++r) row_sum(r);
    gap_update();  // unconditional in Spike (code/EqsGnl2.cpp:1125)
    if (act0 & (STP_ACT | STD_ACT))
      for (int i = 0; i < N; ++i) stp_update(i);
    if (act0 & LTP_ACT) ltp_update();
  }

  int64_t run(int64_t n, int64_t *ss, int32_t *sn, int64_t cap) override {
    int64_t w = 0;
    bool over = false;
    for (int64_t s = 0; s < n; ++s) {
      one_step();
      for (int i = 0; i < N; ++i) {
        if (ns[i] & SPIKE_F) {
          if (w < cap) { ss[w] = step; sn[w] = i; }
          else over = true;
          ++w;
        }
      }
      ++step;
    }
    return over ? -1 : w;
  }

  void get(int f, double *o) const override {
    for (int i = 0; i < N; ++i) {
      double v = 0;
      switch (f) {
        case 0: v = (double)V[i]; break;
        case 1: v = cnt[i]; break;
        case 2: v = ns[i]; break;
        case 3: case 4: case 5: case 6: case 7: v = (double)g[f - 3][i]; break;
        case 8: v = (double)Isyn[i]; break;
        case 9: case 10: case 11: case 12: v = (double)est[f - 9][i]; break;
        case 14: v = (double)ext_xt[i]; break;
        case 15: v = (double)ext_st[i]; break;
        case 16: v = (double)sD[i]; break;
        case 17: v = (double)sF[i]; break;
        default: v = 0;
      }
      o[i] = v;
    }
  }
  void get_gates(int f, double *o) const override {
    for (int k = 0; k < G; ++k) o[k] = (double)(f == 0 ? gst[k] : gxt[k]);
  }
  void get_params(int f, double *o) const override {
    for (int i = 0; i < N; ++i) {
      const NeuPar &q = par[i];
      const float v[11] = {q.df, q.ef, q.gl, q.dvmax, q.dfst[RA], q.dfst[RG], q.dfst[RN],
                           q.dfxt_n, q.efst_n, q.pb[0], q.gext[0]};
      o[i] = (f >= 0 && f < 11) ? (double)v[f] : 0.0;
    }
  }

  // Timed CPU path: the reference's algorithm class (O(S) pull every step)
  // with its std::thread strided partition and a barrier between phases
  // (code/NetSimGnl2.cpp:160-229).  Each worker draws from its own engine,
  // so the run is not the reference's stream; it is for timing only.
  double time_run(int64_t n, int threads) override {
    if (threads < 1) threads = 1;
    std::atomic<int> arrive{0};
    std::atomic<int64_t> gen{0};
    auto barrier = [&](int64_t &local_gen) {
      local_gen++;
      if (arrive.fetch_add(1) + 1 == threads) {
        arrive.store(0);
        gen.store(local_gen);
      } else {
        while (gen.load() < local_gen) std::this_thread::yield();
      }
    };
    const int64_t step0 = step;
    auto work = [&](int tid) {
      // flush-to-zero / denormals-are-zero: decaying gates would otherwise
      // enter the x86 subnormal slow path (the reference pays that cost;
      // the timed baseline is given the faster mode)
      _mm_setcsr(_mm_getcsr() | 0x8040u);
      std::mt19937 eng(1234u + tid);
      int64_t lg = 0;
      for (int64_t s = 0; s < n; ++s) {
        if (tid == 0) do_evt((float)(unsigned)(step0 + s) * dt);
        barrier(lg);
        for (int i = tid; i < N; i += threads) {
          const uint32_t a = act[i];
          const NeuPar &q = par[i];
          if (a & AMPA_ACT) {
            if (q.pb[0] > canonical_f(eng)) est[0][i] = est[0][i] + T(1);
            else est[0][i] = est[0][i] * T(q.dfst[RA]);
            g[RA][i] = g[RA][i] + T(q.gext[0]) * est[0][i];
          }
          for (int k = 1; k < 4; ++k) {
            static const uint32_t bits[4] = {AMPA_ACT, GABA_ACT, ACH_ACT, GCL_ACT};
            if (a & bits[k]) (void)canonical_f(eng);
          }
          if (a & NMDA_ACT) (void)canonical_f(eng);
          vmeul(i);
        }
        barrier(lg);
        for (int k = tid; k < G; k += threads) gate_update(k);
        barrier(lg);
        for (int r = tid; r < R; r += threads) row_sum(r);
        barrier(lg);
      }
    };
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int t = 1; t < threads; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &t : th) t.join();
    auto t1 = std::chrono::steady_clock::now();
    step += n;
    return std::chrono::duration<double>(t1 - t0).count();
  }
};

template <> inline float Sim<float>::texp(float x) { return std::exp(x); }
template <> inline double Sim<double>::texp(double x) { return std::exp(x); }

}  // namespace

struct or_sim {
  SimBase *impl;
};

extern "C" {

or_sim *or_create(const or_net *net, int precision) {
  or_sim *s = new or_sim;
  if (precision == 1) s->impl = new Sim<double>(net);
  else s->impl = new Sim<float>(net);
  s->impl->reset(0);
  return s;
}

void or_destroy(or_sim *s) {
  if (!s) return;
  delete s->impl;
  delete s;
}

void or_reset(or_sim *s, uint32_t seed) { s->impl->reset(seed); }

void or_set_ext_bits(or_sim *s, const uint32_t *bits, int64_t n_steps) {
  s->impl->replay = bits;
  s->impl->replay_steps = n_steps;
}

void or_record_ext_bits(or_sim *s, uint32_t *bits, int64_t n_steps) {
  s->impl->record = bits;
  s->impl->record_steps = n_steps;
}

int64_t or_run(or_sim *s, int64_t n, int64_t *ss, int32_t *sn, int64_t cap) {
  return s->impl->run(n, ss, sn, cap);
}

int64_t or_step_index(const or_sim *s) { return s->impl->step; }

void or_get(const or_sim *s, int f, double *o) { s->impl->get(f, o); }
void or_get_gates(const or_sim *s, int f, double *o) { s->impl->get_gates(f, o); }
void or_get_params(const or_sim *s, int f, double *o) { s->impl->get_params(f, o); }

void or_mt19937_uniform(uint32_t seed, int64_t n, float *out) {
  std::mt19937 g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = canonical_f(g);
}

double or_time_run(or_sim *s, int64_t n, int threads) { return s->impl->time_run(n, threads); }

}  // extern "C"
