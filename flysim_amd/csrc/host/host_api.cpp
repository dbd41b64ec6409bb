// C ABI of the drop-in host side: network.conf / network.pro loading, the
// built network, the reference's internal layout and the reference's
// external-input draw stream.
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "logs.hpp"
#include "builder.hpp"
#include "conf_state.hpp"
#include "events.hpp"
#include "host_common.hpp"
#include "netconf.hpp"
#include "writers.hpp"


namespace fs {

uint32_t act_from_opts(const fs_cli_opts &o) {
  // code/main.cpp:220-231 and :263; UDFRDSEED_ACT shares STD_ACT's bit
  uint32_t a = FS_ACT_AMPA | FS_ACT_GABA | FS_ACT_NMDA | FS_ACT_ACH | FS_ACT_GCL | FS_ACT_GAP |
               FS_ACT_LIF;
  if (o.udf_seed) a |= FS_ACT_STD;
  if (o.stp) a |= FS_ACT_STP;
  if (o.std_) a |= FS_ACT_STD;
  if (o.ltp) a |= FS_ACT_LTP;
  return a;
}

void pop_conf_of(const PopDecl &p, float dt, fs_pop_conf &c) {
  c = fs_pop_conf();
  c.cm = p.cm; c.taum = p.taum; c.vth = p.vth; c.el = p.el; c.vreset = p.vreset;
  // steps are stored; keep the ms fields consistent for fs_derive_params users
  c.refractory_ms = (float)p.rfped * dt;
  c.spike_delay_ms = (float)p.spidy * dt;
  for (const ReceptorDecl &r : p.recs) {
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

}  // namespace fs

extern "C" {

int fs_conf_load(const char *conf_path, const char *pro_path, const fs_cli_opts *opts,
                 fs_conf **out) {
  FS_TRY_BEGIN
  if (!conf_path || !pro_path || !out) return fs::set_error(FS_ERR_ARG, "null argument");
  fs_cli_opts o{};
  o.dt = 0.1f; o.threads = 1; o.repeats = 1;
  if (opts) o = *opts;
  if (!(o.dt > 0.0f)) o.dt = 0.1f;
  auto *c = new fs_conf();
  c->cfg.dt = o.dt;
  if (o.out_mem) { c->cfg.comm_mem = true; c->cfg.out_mem = o.out_mem; }
  if (o.out_spi) { c->cfg.comm_spi = true; c->cfg.out_spi = o.out_spi; }
  if (o.out_fr) { c->cfg.comm_fr = true; c->cfg.out_fr = o.out_fr; }
  try {
    fs::load_config(c->cfg, conf_path, pro_path);
  } catch (const std::invalid_argument &e) {
    delete c;
    return fs::set_error(FS_ERR_PARSE, std::string("malformed number: ") + e.what());
  } catch (const std::out_of_range &e) {
    delete c;
    return fs::set_error(FS_ERR_PARSE, std::string("number out of range: ") + e.what());
  } catch (const std::runtime_error &e) {
    delete c;
    const std::string m = e.what();
    return fs::set_error(m.find("open file") != std::string::npos ? FS_ERR_IO : FS_ERR_PARSE, m);
  }
  c->opt.threads = o.threads < 1 ? 1 : o.threads;
  c->opt.repeats = o.repeats < 1 ? 1 : o.repeats;
  c->opt.udf_seed = o.udf_seed != 0;
  c->opt.udf_seed_value = o.udf_seed_value;
  c->opt.act_ctr = fs::act_from_opts(o);
  try {
    fs::build_network(c->cfg, c->opt, c->net);
  } catch (const std::exception &e) {
    delete c;
    return fs::set_error(FS_ERR_ARG, std::string("network build: ") + e.what());
  }
  c->pop_conf.resize(c->cfg.pops.size());
  for (size_t i = 0; i < c->cfg.pops.size(); ++i) {
    fs::pop_conf_of(c->cfg.pops[i], o.dt, c->pop_conf[i]);
    c->pop_rfped.push_back(c->cfg.pops[i].rfped);
    c->pop_spidy.push_back(c->cfg.pops[i].spidy);
  }
  for (const fs::Event &e : c->cfg.events) {
    c->evt_t.push_back(e.t_s);
    c->evt_v1.push_back(e.var1);
    c->evt_v2.push_back(e.var2);
    c->evt_type.push_back((int32_t)e.type);
    c->evt_pop.push_back((int32_t)e.pop);
    c->evt_sub.push_back((int32_t)e.sub);
  }
  *out = c;
  return FS_OK;
  FS_TRY_END
}

int fs_conf_free(fs_conf *c) {
  delete c;
  return FS_OK;
}

int fs_conf_desc(fs_conf *c, fs_net_desc *d) {
  if (!c || !d) return fs::set_error(FS_ERR_ARG, "null argument");
  std::memset(d, 0, sizeof(*d));
  d->n_neuron = (int32_t)c->net.n_neuron;
  d->n_pop = (int32_t)c->net.pops.size();
  d->neuron_pop = c->net.neuron_pop.data();
  d->pops = c->net.pops.data();
  d->n_syn = (int64_t)c->net.pre.size();
  d->pre = c->net.pre.data();
  d->post = c->net.post.data();
  d->w = c->net.w.data();
  d->receptor = c->net.rec.data();
  d->delay = nullptr;
  d->plastic = c->net.plastic.data();
  d->dt = c->cfg.dt;
  d->precision = FS_FP32;
  d->ext_mode = FS_EXT_PHILOX;
  d->record_spikes = 1;
  return FS_OK;
}

int fs_conf_sizes(fs_conf *c, int64_t o[7]) {
  if (!c || !o) return fs::set_error(FS_ERR_ARG, "null argument");
  o[0] = c->net.n_neuron;
  o[1] = (int64_t)c->net.pops.size();
  o[2] = (int64_t)c->net.pre.size();
  o[3] = (int64_t)c->net.total_synapses;
  o[4] = (int64_t)c->cfg.events.size();
  o[5] = (int64_t)c->cfg.outs.size();
  o[6] = (int64_t)c->net.rseeds.size();
  return FS_OK;
}

int fs_conf_seed(fs_conf *c, int32_t r, uint32_t *seed) {
  if (!c || !seed || r < 0 || r >= (int32_t)c->net.rseeds.size())
    return fs::set_error(FS_ERR_ARG, "bad repeat index");
  *seed = c->net.rseeds[r];
  return FS_OK;
}

int fs_conf_ext_trains(fs_conf *c, int32_t repeat, int64_t n_steps, uint32_t *bits) {
  FS_TRY_BEGIN
  if (!c || !bits || n_steps < 0 || repeat < 0 || repeat >= (int32_t)c->net.rseeds.size())
    return fs::set_error(FS_ERR_ARG, "bad argument");
  const uint32_t N = c->net.n_neuron;
  const int64_t words = (N + 31) / 32;
  std::memset(bits, 0, sizeof(uint32_t) * (size_t)(n_steps * 5 * words));
  std::vector<fs_pop_params> pops = c->net.pops;
  // Run(): rdgen reseeded from rseeds[Repeat-1]; thsize-1 draws for the
  // other threads' ziggurat seeds (code/NetSimGnl2.cpp:313-330)
  std::mt19937 rd(c->net.rseeds[repeat]);
  std::uniform_int_distribution<unsigned int> dist(0, RAND_MAX);
  for (int i = 1; i < c->opt.threads; ++i) (void)dist(rd);
  std::uniform_real_distribution<float> uni(0.0f, 1.0f);
  fs::EventCursor cur(c->cfg.events, c->cfg.dt);
  // receptor draw order of BgInit (code/EqsGnl2.cpp:288-332)
  static const int rid[5] = {FS_AMPA, FS_GABA, FS_ACH, FS_GCL, FS_NMDA};
  static const uint32_t rbit[5] = {FS_ACT_AMPA, FS_ACT_GABA, FS_ACT_ACH, FS_ACT_GCL, FS_ACT_NMDA};
  for (int64_t s = 0; s < n_steps; ++s) {
    cur.apply((uint64_t)s, pops, nullptr);
    uint32_t *plane = bits + s * 5 * words;
    for (uint32_t n = 0; n < N; ++n) {
      const fs_pop_params &p = pops[c->net.neuron_pop[n]];
      for (int k = 0; k < 5; ++k) {
        if (!(p.act & rbit[k])) continue;
        const float u = uni(rd);
        if (p.pb[rid[k]] > u) plane[rid[k] * words + n / 32] |= 1u << (n % 32);
      }
    }
  }
  return FS_OK;
  FS_TRY_END
}

int fs_conf_write_outputs(fs_conf *c, int32_t repeat, int64_t n_steps, const int64_t *offsets,
                          const uint32_t *ids, const char *dir) {
  FS_TRY_BEGIN
  if (!c || !offsets || (!ids && n_steps > 0 && offsets[n_steps] > 0) || n_steps < 0)
    return fs::set_error(FS_ERR_ARG, "bad argument");
  fs::OutputWriters w(c->cfg, c->net, repeat, dir ? dir : "");
  if (w.needs())
    return fs::set_error(FS_ERR_ARG, "MemPot / DeltaMemPotLimit / IsynSeparatedSum / GatingVariable "
                                     "outputs need per-step state (use the CLI)");
  const fs::StepState none;
  for (int64_t k = 0; k < n_steps; ++k) {
    for (int64_t q = offsets[k]; q < offsets[k + 1]; ++q)
      if (ids[q] >= c->net.n_neuron) return fs::set_error(FS_ERR_ARG, "spike id out of range");
    w.step((uint64_t)k, ids + offsets[k], offsets[k + 1] - offsets[k], none);
  }
  return FS_OK;
  FS_TRY_END
}

int fs_conf_write_logs(fs_conf *c, const char *dir) {
  FS_TRY_BEGIN
  if (!c) return fs::set_error(FS_ERR_ARG, "null handle");
  std::string d = dir ? dir : "";
  if (!d.empty() && d.back() != '/') d += '/';
  fs::write_confs_info(c->cfg, d + "ConfsInfo.log");
  fs::write_network_log(c->cfg, c->net, c->opt, d + "network.log");
  return FS_OK;
  FS_TRY_END
}

int fs_conf_ref_layout(fs_conf *c, fs_ref_layout *o) {
  if (!c || !o) return fs::set_error(FS_ERR_ARG, "null argument");
  std::memset(o, 0, sizeof(*o));
  const fs::BuiltNet &n = c->net;
  o->n_gate = (int64_t)n.gate_pre.size();
  o->n_row = (int64_t)n.row_post.size();
  o->gate_pre = n.gate_pre.data();
  o->gate_post0 = n.gate_post0.data();
  o->gate_rec = n.gate_rec.data();
  o->gate_group = n.gate_group.data();
  o->gate_g = n.gate_g.data();
  o->row_post = n.row_post.data();
  o->row_beg = n.row_beg.data();
  o->term = n.term.data();
  o->row_rec = n.row_rec.data();
  o->pop_conf = c->pop_conf.data();
  o->pop_rfped = c->pop_rfped.data();
  o->pop_spidy = c->pop_spidy.data();
  o->neuron_pop = n.neuron_pop.data();
  o->n_evt = (int64_t)c->evt_t.size();
  o->evt_time_s = c->evt_t.data();
  o->evt_var1 = c->evt_v1.data();
  o->evt_var2 = c->evt_v2.data();
  o->evt_type = c->evt_type.data();
  o->evt_pop = c->evt_pop.data();
  o->evt_sub = c->evt_sub.data();
  o->act_ctr = c->opt.act_ctr;
  o->n_gap = (int64_t)n.gap_pre.size();
  o->gap_pre = n.gap_pre.data();
  o->gap_post = n.gap_post.data();
  o->gap_g = n.gap_g.data();
  return FS_OK;
}

}  // extern "C"
