// Network builder; see builder.hpp.
#include "builder.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <random>
#include <stdexcept>

namespace fs {

uint8_t receptor_of_keyword(uint8_t kw) {
  switch (kw) {
    case KW_AMPA: return R_AMPA;
    case KW_GABA: return R_GABA;
    case KW_NMDA: return R_NMDA;
    case KW_ACH: return R_ACH;
    case KW_GCL: return R_GCL;
    case KW_GAP: return R_GAP;
    default: return R_NONE;  // Sine / CUBA: GetAct gives no synapse class
  }
}

static void to_pop_conf(const PopDecl &p, fs_pop_conf &c) {
  c = fs_pop_conf();
  c.cm = p.cm; c.taum = p.taum; c.vth = p.vth; c.el = p.el; c.vreset = p.vreset;
  c.refractory_ms = 0; c.spike_delay_ms = 0;  // steps are carried separately
  for (const ReceptorDecl &r : p.recs) {      // later declarations override earlier ones
    const uint8_t id = receptor_of_keyword(r.kw);
    if (id >= 5) continue;
    c.declared[id] = 1;
    c.tau[id] = r.tau; c.tau_xt[id] = r.tau_xt; c.vrev[id] = r.vrev;
    c.freq_ext[id] = r.fext; c.ext_eff[id] = r.ext_eff; c.ext_con[id] = r.ext_con;
  }
  const float stp[8] = {p.stp.pv, p.stp.tD, p.stp.aF, p.stp.tF, p.stp.aCap, p.stp.tCap,
                        p.stp.aCar, p.stp.tCar};
  for (int k = 0; k < 8; ++k) c.stp[k] = stp[k];
  c.ltp[0] = p.ltp.tLP;
  c.ltp[1] = (float)p.ltp.SpiLPDy;
  c.ltp[2] = p.ltp.PosISI;
  c.ltp[3] = p.ltp.NegISI;
}

// Parameter derivation of code/NetSimGnl2_BulPp.cpp:159-278, in float.
static void derive(const fs_pop_conf &c, int32_t rfped, int32_t spidy, float dt, uint32_t act,
                   fs_pop_params &q) {
  q = fs_pop_params();
  q.vth = c.vth;
  q.el = c.el;
  q.gl = c.cm / c.taum;          // LeakyConductance is ignored (:168)
  q.cm = c.cm;
  q.vreset = c.vreset;
  q.df = std::exp(-dt / c.taum);
  q.ef = c.taum * (1.0f - q.df);
  q.rfped = rfped;
  q.spidy = spidy;
  q.taum = c.taum;
  q.vrev[R_AMPA] = 0.0f;   q.dfst[R_AMPA] = 0.1f;
  q.vrev[R_GABA] = -90.0f; q.dfst[R_GABA] = 0.1f;
  q.dvmax = (q.vth - q.vrev[R_GABA]) / q.ef;
  q.as_n = 0.6332f; q.dfxt_n = 0.1f; q.dfst[R_NMDA] = 0.1f; q.efst_n = 0.01f;
  q.vrev[R_NMDA] = 0.0f; q.mg_n = 1.0f;
  q.vrev[R_ACH] = 0.0f;    q.dfst[R_ACH] = 0.1f;
  q.vrev[R_GCL] = -90.0f;  q.dfst[R_GCL] = 0.1f;
  for (int r = 0; r < 5; ++r) {
    if (!c.declared[r]) continue;
    if (r == R_GABA) q.dvmax = (c.vth - c.vrev[r]) / q.ef;
    q.vrev[r] = c.vrev[r];
    q.dfst[r] = std::exp(-dt / c.tau[r]);
    if (r == R_NMDA) {
      q.as_n = 0.6332f;
      q.dfxt_n = std::exp(-dt / c.tau_xt[r]);
      q.efst_n = c.tau[r] * (1.0f - q.dfst[r]);
      q.mg_n = 1.0f;
    }
    q.pb[r] = c.freq_ext[r] * dt * 1e-3f;
    q.gext[r] = c.ext_eff[r] * c.ext_con[r];
  }
  // short-term plasticity (:185-200); CarBase = 20 uM (code/Inc.h:27)
  q.stp_pv = c.stp[0];
  q.stp_dfD = std::exp(-dt / c.stp[1]);
  q.stp_efD = 1.0f - q.stp_dfD;
  q.stp_aF = c.stp[2];
  q.stp_dfF = std::exp(-dt / c.stp[3]);
  q.stp_efF = c.stp[3] * (1.0f - q.stp_dfF);
  q.stp_aCap = c.stp[4];
  q.stp_dfCap = std::exp(-dt / c.stp[5]);
  q.stp_aCar = c.stp[6];
  q.stp_dfCar = std::exp(-dt / c.stp[7]);
  q.stp_efCar = (1.0f - q.stp_dfCar) * 20.0f * 1e-6f;
  // long-term plasticity (:180-183)
  q.ltp_dfLP = std::exp(-dt / c.ltp[0]);
  q.ltp_spidy = (int32_t)c.ltp[1];
  q.ltp_pos = c.ltp[2];
  q.ltp_neg = c.ltp[3];
  // MNoiseMean = MNoiseSTD = 0: injection off (:142-148)
  q.act = act & ~(uint32_t)FS_ACT_INJ;
  q.inj_mean = 0.0f;
  q.inj_std = 0.0f;
}

fs_pop_params derive_pop_params(const PopDecl &p, float dt, uint32_t act_ctr) {
  fs_pop_conf c;
  to_pop_conf(p, c);
  fs_pop_params q;
  derive(c, p.rfped, p.spidy, dt, act_ctr, q);
  return q;
}

namespace {

struct Idx { uint32_t pre, post; };
struct Single { uint32_t pre, post; float g; float st; };

// BulBiasDNeuMap (code/Eqs.h:1120-1165): rows of equal consecutive PosID.
template <class C>
void rows_of(const std::vector<C> &v, std::vector<std::pair<uint32_t, uint32_t>> &rows) {
  rows.clear();
  size_t i = 0;
  while (i < v.size()) {
    size_t j = i;
    while (j < v.size() && v[j].post == v[i].post) ++j;
    rows.emplace_back((uint32_t)i, (uint32_t)j);
    i = j;
  }
}

}  // namespace

void build_network(const Config &cfg, const BuildOptions &opt, BuiltNet &out) {
  const float dt = cfg.dt;
  const size_t NP = cfg.pops.size();
  out = BuiltNet();
  // neuron index ranges (code/NetSimGnl2_BulPp.cpp:40-47)
  std::vector<uint32_t> nidx(NP);
  uint32_t q = 0;
  for (size_t i = 0; i < NP; ++i) { nidx[i] = q; q += cfg.pops[i].n; }
  out.n_neuron = q;
  out.pop_first = nidx;
  out.pop_n.resize(NP);
  out.neuron_pop.resize(q);
  for (size_t i = 0; i < NP; ++i) {
    out.pop_n[i] = cfg.pops[i].n;
    for (uint32_t k = nidx[i]; k < nidx[i] + cfg.pops[i].n; ++k) out.neuron_pop[k] = (uint32_t)i;
  }
  out.pops.resize(NP);
  for (size_t i = 0; i < NP; ++i) out.pops[i] = derive_pop_params(cfg.pops[i], dt, opt.act_ctr);

  // seeds (:80-127)
  const int RpSize = std::max(1, opt.repeats);
  out.rseeds.assign(RpSize, 0);
  // UDFRDSEED_ACT is STD_ACT's bit (code/Inc.h:275,280): -STD alone also
  // selects the command-line seed, 0 when -udfsed is not given
  if (opt.udf_seed || (opt.act_ctr & FS_ACT_STD)) out.rseeds[0] = opt.udf_seed_value;
  else if (!cfg.seeds.empty()) out.rseeds[0] = cfg.seeds[0];
  else out.rseeds[0] = std::random_device()();
  std::mt19937 rd(out.rseeds[0]);
  std::uniform_int_distribution<unsigned int> dist(0, RAND_MAX);
  for (int i = 1; i < opt.threads; ++i) (void)dist(rd);
  for (int i = 1; i < RpSize; ++i) {
    if ((size_t)i < cfg.seeds.size()) out.rseeds[i] = cfg.seeds[i];
    else {
      unsigned s;
      bool found = true;
      while (found) {
        found = false;
        s = dist(rd);
        for (int j = 0; j < i; ++j) if (s == out.rseeds[j]) { found = true; break; }
      }
      out.rseeds[i] = s;
    }
  }

  // PosBuild (code/NetSim.h:431-454)
  auto pos_build = [&](uint32_t k, uint32_t num, uint32_t star, uint32_t posnum, bool selfen,
                       std::vector<uint32_t> &res) {
    std::vector<uint32_t> con(posnum);
    const uint32_t jstart = k - star;
    bool selfc = false;
    for (uint32_t jj = 0; jj < posnum; ++jj) {
      con[jj] = star + jj;
      if (jstart == jj && !selfen) selfc = true;
    }
    if (selfc) con.erase(con.begin() + jstart);
    std::shuffle(con.begin(), con.end(), rd);
    if (num > con.size()) throw std::runtime_error("connectivity exceeds target population size");
    std::sort(con.begin(), con.begin() + num);
    res.assign(con.begin(), con.begin() + num);
  };
  // PosBuild1 (code/NetSim.h:403-429)
  auto pos_build1 = [&](uint32_t k, uint32_t star, uint32_t posnum, bool selfen) -> uint32_t {
    if (posnum == 1) return star;
    std::vector<uint32_t> con(posnum);
    const uint32_t jstart = k - star;
    bool selfc = false;
    for (uint32_t jj = 0; jj < posnum; ++jj) {
      con[jj] = star + jj;
      if (jstart == jj && !selfen) selfc = true;
    }
    if (selfc) con.erase(con.begin() + jstart);
    std::shuffle(con.begin(), con.end(), rd);
    return con[0];
  };

  // group gates (:303-566): one gate per (pre neuron x link) with >= 2 targets
  struct Gate { uint32_t pre, post0; float g; };
  std::vector<Gate> pgat[6];
  std::vector<Idx> tpsyn[6];
  for (size_t i = 0; i < NP; ++i) {
    const PopDecl &P = cfg.pops[i];
    for (uint32_t k = nidx[i]; k < nidx[i] + P.n; ++k) {
      for (const LinkDecl &L : P.links) {
        const uint8_t r = receptor_of_keyword(L.rec_kw);
        const uint32_t tp = L.target;
        const uint32_t tn = cfg.pops.at(tp).n;
        uint32_t num = (uint32_t)std::ceil((double)tn * L.conn);
        if (i == tp && tn == num && !P.self_con) num--;
        std::vector<uint32_t> pos;
        pos_build(k, num, nidx[tp], tn, P.self_con, pos);
        const float g = L.mean_eff * L.weight;
        if (num < 2 || r == R_NONE) continue;
        const uint32_t gi = (uint32_t)pgat[r].size();
        for (uint32_t p : pos) tpsyn[r].push_back({gi, p});
        pgat[r].push_back({k, pos[0], g});
      }
    }
  }
  // single synapses (:571-695); PreID is the *population* index (:645)
  std::vector<Single> syn1[6];
  for (size_t i = 0; i < NP; ++i) {
    const PopDecl &P = cfg.pops[i];
    for (uint32_t k = nidx[i]; k < nidx[i] + P.n; ++k) {
      for (const LinkDecl &L : P.links) {
        const uint32_t tp = L.target;
        const uint32_t tn = cfg.pops.at(tp).n;
        const uint32_t num = (uint32_t)std::ceil((double)tn * L.conn);
        if (num != 1) continue;
        const uint8_t r = receptor_of_keyword(L.rec_kw);
        if (r == R_NONE) continue;
        const uint32_t post = pos_build1(k, nidx[tp], tn, P.self_con);
        syn1[r].push_back({(uint32_t)i, post, L.mean_eff * L.weight, 0.0f});
      }
    }
  }
  for (auto &v : tpsyn)
    std::sort(v.begin(), v.end(), [](Idx a, Idx b) { return a.post < b.post; });
  for (auto &v : syn1)
    std::sort(v.begin(), v.end(), [](const Single &a, const Single &b) { return a.post < b.post; });

  out.gap_synapses = tpsyn[R_GAP].size() + syn1[R_GAP].size();
  // the solver couples only single-synapse gap junctions (ThdGapSyn,
  // code/EqsGnl2.cpp:705-725); group gap links are built but never used
  for (const Single &s : syn1[R_GAP]) {
    out.gap_pre.push_back(s.pre);
    out.gap_post.push_back(s.post);
    out.gap_g.push_back(s.g);
  }
  out.total_synapses = 0;
  for (int r = 0; r < 6; ++r) out.total_synapses += tpsyn[r].size() + syn1[r].size();

  // ---------------- reference layout
  static const uint8_t order[5] = {R_AMPA, R_GABA, R_ACH, R_GCL, R_NMDA};
  std::vector<uint32_t> gbase(6, 0);
  for (uint8_t r : order) {
    gbase[r] = (uint32_t)out.gate_pre.size();
    for (const Gate &g : pgat[r]) {
      out.gate_pre.push_back(g.pre);
      out.gate_post0.push_back(g.post0);
      out.gate_rec.push_back(r);
      out.gate_group.push_back(1);
      out.gate_g.push_back(g.g);
    }
  }
  out.row_beg.push_back(0);
  std::vector<std::pair<uint32_t, uint32_t>> rows;
  for (uint8_t r : order) {
    rows_of(tpsyn[r], rows);
    for (auto &rw : rows) {
      out.row_post.push_back(tpsyn[r][rw.first].post);
      out.row_rec.push_back(r);
      for (uint32_t t = rw.first; t < rw.second; ++t) out.term.push_back(gbase[r] + tpsyn[r][t].pre);
      out.row_beg.push_back((uint32_t)out.term.size());
    }
  }
  for (uint8_t r : order) {
    const uint32_t base = (uint32_t)out.gate_pre.size();
    for (const Single &s : syn1[r]) {
      out.gate_pre.push_back(s.pre);
      out.gate_post0.push_back(s.post);
      out.gate_rec.push_back(r);
      out.gate_group.push_back(0);
      out.gate_g.push_back(s.g);
    }
    rows_of(syn1[r], rows);
    for (auto &rw : rows) {
      out.row_post.push_back(syn1[r][rw.first].post);
      out.row_rec.push_back(r);
      for (uint32_t t = rw.first; t < rw.second; ++t) out.term.push_back(base + t);
      out.row_beg.push_back((uint32_t)out.term.size());
    }
  }
  // ---------------- engine synapse list: every (gate, target) pair
  for (uint8_t r : order) {
    for (const Idx &e : tpsyn[r]) {
      const Gate &g = pgat[r][e.pre];
      out.pre.push_back(g.pre);
      out.post.push_back(e.post);
      out.w.push_back(g.g);
      out.rec.push_back(r);
      out.plastic.push_back(0);
    }
    for (const Single &s : syn1[r]) {
      out.pre.push_back(s.pre);
      out.post.push_back(s.post);
      out.w.push_back(s.g);
      out.rec.push_back(r);
      out.plastic.push_back(1);
    }
  }
  for (size_t q = 0; q < out.gap_pre.size(); ++q) {
    out.pre.push_back(out.gap_pre[q]);
    out.post.push_back(out.gap_post[q]);
    out.w.push_back(out.gap_g[q]);
    out.rec.push_back(R_GAP);
    out.plastic.push_back(0);
  }
}

}  // namespace fs

extern "C" int fs_derive_params(const fs_pop_conf *conf, float dt, uint32_t act, fs_pop_params *out) {
  if (!conf || !out || !(dt > 0.0f)) return FS_ERR_ARG;
  const int32_t rfped = (int32_t)(std::floor(conf->refractory_ms / dt) + dt);
  const int32_t spidy = (int32_t)std::floor(conf->spike_delay_ms / dt);
  fs::derive(*conf, rfped, spidy, dt, act, *out);
  return FS_OK;
}
