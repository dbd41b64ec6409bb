// Drop-in command line of the Flysim core: argument handling of
// code/main.cpp:26-256, the step driver of code/NetSimGnl2.cpp:231-358
// (events, repeats, output files) with the solver replaced by the GPU
// engine, and the output writers of code/NetSimGnl2_info.cpp:548-699.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "builder.hpp"
#include "conf_state.hpp"
#include "events.hpp"
#include "host_common.hpp"
#include "netconf.hpp"


namespace fs {

namespace {

void print_help() {
  std::cout << "MI355X-native Flysim core (drop-in for flysim.out)\n\n"
            << "flysim_amd -option parameters\n\n"
            << "-pro network.pro      # protocol file: default=network.pro\n"
            << "-conf network.conf    # configuration file: default=network.conf\n"
            << "-om my_membrane.dat   # output membrane potential file\n"
            << "-os my_spike.dat      # output spikes file\n"
            << "-or my_rate.dat       # output firing rate (window 50ms, print 100ms)\n"
            << "-rp 1                 # repeats\n"
            << "-t 4                  # thread count (kept for connectome parity)\n"
            << "-dt 0.1               # time step (ms)\n"
            << "-STP / -STD / -LTP    # plasticity switches (not supported by the GPU engine)\n"
            << "-nmodel GNL2          # neuron model (GNL2 only)\n"
            << "-udfsed N             # user-defined random seed\n"
            << "-fp64                 # double-precision state (default float)\n"
            << "-ext replay|philox    # external input: the reference's mt19937 stream\n"
            << "                      # (replay) or the counter-based generator (default)\n";
}

// One output file of the protocol (OutCtr, code/parser.h:98-120).
struct Writer {
  OutCtrl oc;
  std::ofstream f;
  std::vector<uint32_t> spirwin;  // FiringRate window ring (code/NetSim.cpp:39-40)
};

}  // namespace

int cli_main(int argc, const char *const *argv) {
  std::string conf = "network.conf", pro = "network.pro";
  std::string om = "MemPot.dat", os_ = "Spikes.dat", orr = "FRate.dat";
  bool om_en = false, os_en = false, or_en = false, stp = false, stdp = false, ltp = false;
  bool udf = false, fp64 = false, replay = false, help = false;
  uint32_t seed = 0;
  int rp = 1, th = 1;
  float dt = 0.1f;
  for (int i = 1; i < argc; ++i) {
    const std::string s = argv[i];
    auto val = [&](std::string &dst) {
      if (i + 1 < argc) dst = argv[++i];
      else help = true;
    };
    std::string v;
    if (s == "-pro") val(pro);
    else if (s == "-conf") val(conf);
    else if (s == "-om") { val(om); om_en = true; }
    else if (s == "-os") { val(os_); os_en = true; }
    else if (s == "-or") { val(orr); or_en = true; }
    else if (s == "-rp") { val(v); rp = std::atoi(v.c_str()); }
    else if (s == "-t") { val(v); th = std::atoi(v.c_str()); }
    else if (s == "-s" || s == "-nmodel") { val(v); }  // parsed, ignored (code/main.cpp:161-193)
    else if (s == "-dt") { val(v); dt = (float)std::atof(v.c_str()); }
    else if (s == "-udfsed") { val(v); seed = (uint32_t)std::atoi(v.c_str()); udf = true; }
    else if (s == "-STP") stp = true;
    else if (s == "-STD") stdp = true;
    else if (s == "-LTP") ltp = true;
    else if (s == "-fp64") fp64 = true;
    else if (s == "-ext") { val(v); replay = (v == "replay"); }
    else { help = true; }
    if (help) break;
  }
  if (help) { print_help(); return FS_OK; }
  if (stp || stdp || ltp || udf) {
    // -udfsed shares STD_ACT's bit in the reference (code/Inc.h:275,280)
    return set_error(FS_ERR_ARG, "short/long-term plasticity (-STP/-STD/-LTP, and -udfsed which "
                                 "implies STD) is not supported by the GPU engine");
  }
  fs_cli_opts o{};
  o.dt = dt; o.threads = th; o.repeats = rp; o.stp = stp; o.std_ = stdp; o.ltp = ltp;
  o.udf_seed = udf; o.udf_seed_value = seed;
  o.out_mem = om_en ? om.c_str() : nullptr;
  o.out_spi = os_en ? os_.c_str() : nullptr;
  o.out_fr = or_en ? orr.c_str() : nullptr;
  fs_conf *c = nullptr;
  int rc = fs_conf_load(conf.c_str(), pro.c_str(), &o, &c);
  if (rc != FS_OK) return rc;
  struct Guard { fs_conf *c; ~Guard() { fs_conf_free(c); } } guard{c};
  for (const std::string &w : c->cfg.warnings) std::cerr << "warning: " << w << "\n";
  if (c->net.gap_synapses)
    return set_error(FS_ERR_ARG, "gap junctions are not supported by the GPU engine");
  if (c->cfg.events.empty()) return set_error(FS_ERR_ARG, "the protocol has no events (no end time)");
  const Event &last = c->cfg.events.back();
  if (last.nolimit) return set_error(FS_ERR_ARG, "EventTime NoLimit is not supported");
  const uint64_t endtime = (uint64_t)std::floor(last.t_s * 1e3f / dt);  // code/NetSimGnl2.cpp:270
  const unsigned one_ms = (unsigned)(1u / dt);
  std::cout << "\ncapable mode of sim07_21 leaky integrate and fire model\n"
            << "time step:" << dt << ", total neurons=" << c->net.n_neuron
            << ", total synapses=" << c->net.total_synapses << "\n";
  fs_net_desc d;
  fs_conf_desc(c, &d);
  d.precision = fp64 ? FS_FP64 : FS_FP32;
  d.ext_mode = replay ? FS_EXT_REPLAY : FS_EXT_PHILOX;
  fs_handle *h = nullptr;
  rc = fs_create(&d, &h);
  if (rc != FS_OK) return rc;
  struct HGuard { fs_handle *h; ~HGuard() { fs_destroy(h); } } hguard{h};
  const uint32_t N = c->net.n_neuron;
  bool need_v = false;
  for (const OutCtrl &oc : c->cfg.outs) need_v |= oc.type_id == OUT_MEMPOT;
  for (const OutCtrl &oc : c->cfg.outs)
    if (oc.type_id != OUT_SPIKE && oc.type_id != OUT_FRATE && oc.type_id != OUT_MEMPOT)
      std::cerr << "warning: output type " << oc.type << " is not written by the GPU engine\n";

  for (int r = 1; r <= std::max(1, rp); ++r) {
    std::cout << "Iteration:" << r << "\n";
    std::vector<fs_pop_params> pops = c->net.pops;  // Run() resets pb (code/NetSimGnl2.cpp:279-297)
    for (size_t i = 0; i < pops.size(); ++i) {
      fs_pop_params p = pops[i];
      rc = fs_set_pop_params(h, (int32_t)i, &p);
      if (rc != FS_OK) return rc;
    }
    rc = fs_reset(h, c->net.rseeds[r - 1]);
    if (rc != FS_OK) return rc;
    std::vector<uint32_t> bits;
    if (replay) {
      const int64_t words = (N + 31) / 32;
      bits.assign((size_t)endtime * 5 * words, 0u);
      rc = fs_conf_ext_trains(c, r - 1, (int64_t)endtime, bits.data());
      if (rc != FS_OK) return rc;
      if (endtime) {
        rc = fs_load_ext_trains(h, bits.data(), 0, (int64_t)endtime);
        if (rc != FS_OK) return rc;
      }
    }
    // event schedule: the steps at which DoEvt changes a population
    EventCursor cur(c->cfg.events, dt);
    std::vector<uint64_t> ev_steps;
    std::vector<std::vector<std::pair<uint32_t, fs_pop_params>>> ev_changes;
    {
      std::vector<fs_pop_params> sim = pops;
      std::vector<uint32_t> ch;
      for (uint64_t s = 0; s < endtime; ++s) {
        ch.clear();
        if (cur.apply(s, sim, &ch)) {
          ev_steps.push_back(s);
          std::vector<std::pair<uint32_t, fs_pop_params>> v;
          std::sort(ch.begin(), ch.end());
          ch.erase(std::unique(ch.begin(), ch.end()), ch.end());
          for (uint32_t p : ch) v.emplace_back(p, sim[p]);
          ev_changes.push_back(v);
        }
      }
    }
    // output files (code/NetSim.cpp:32-66)
    std::vector<Writer> ws(c->cfg.outs.size());
    for (size_t i = 0; i < ws.size(); ++i) {
      ws[i].oc = c->cfg.outs[i];
      std::string fn = ws[i].oc.file;
      if (r > 1) fn += "_" + std::to_string((long long)r);
      if (ws[i].oc.type_id == OUT_SPIKE || ws[i].oc.type_id == OUT_FRATE ||
          ws[i].oc.type_id == OUT_MEMPOT) {
        ws[i].f.open(fn);
        ws[i].f << std::noshowpoint << std::setprecision(7);
      }
      if (ws[i].oc.type_id == OUT_FRATE)
        ws[i].spirwin.assign(ws[i].oc.pops.size() * (ws[i].oc.rwin + 2), 0u);
    }
    // step loop in batches between event steps
    std::vector<uint8_t> spiked(N, 0);
    std::vector<double> vbuf(need_v ? N : 0);
    uint64_t s = 0;
    size_t ei = 0;
    const int64_t batch_max = need_v ? 1 : 1000;
    std::vector<int64_t> off;
    std::vector<uint32_t> ids;
    while (s < endtime) {
      if (ei < ev_steps.size() && ev_steps[ei] == s) {
        for (auto &pc : ev_changes[ei]) {
          rc = fs_set_pop_params(h, (int32_t)pc.first, &pc.second);
          if (rc != FS_OK) return rc;
        }
        ++ei;
      }
      uint64_t e = std::min<uint64_t>(endtime, s + batch_max);
      if (ei < ev_steps.size() && ev_steps[ei] < e) e = ev_steps[ei];
      const int64_t K = (int64_t)(e - s);
      off.assign(K + 1, 0);
      ids.resize((size_t)N * K);
      rc = fs_run(h, K, off.data(), ids.data(), (int64_t)ids.size());
      if (rc != FS_OK) return rc;
      if (need_v) {
        rc = fs_get_state(h, 0, vbuf.data());
        if (rc != FS_OK) return rc;
      }
      for (int64_t k = 0; k < K; ++k) {
        const uint64_t step = s + (uint64_t)k;
        const float evtime = (float)(unsigned)step * dt;
        for (int64_t q = off[k]; q < off[k + 1]; ++q) spiked[ids[q]] = 1;
        // FileOut (code/NetSimGnl2_info.cpp:548-699)
        for (Writer &w : ws) {
          const OutCtrl &oc = w.oc;
          if (oc.type_id == OUT_SPIKE) {
            std::ostringstream o;
            o << std::noshowpoint << std::setprecision(7);
            unsigned n = 0, kk = 0;
            for (uint32_t m : oc.pops) {
              for (uint32_t nid = c->net.pop_first[m]; nid < c->net.pop_first[m] + c->net.pop_n[m]; ++nid) {
                if (spiked[nid]) {
                  o << evtime * 1e-3f << " " << ((oc.status & OF_ALIGN) ? n : nid) << "\n";
                  ++kk;
                }
                ++n;
              }
            }
            if (kk) w.f << o.str() << std::flush;
          } else if (oc.type_id == OUT_FRATE) {
            const size_t L = oc.pops.size();
            for (size_t j = 0; j < L; ++j) {
              const uint32_t m = oc.pops[j];
              for (uint32_t nid = c->net.pop_first[m]; nid < c->net.pop_first[m] + c->net.pop_n[m]; ++nid)
                if (spiked[nid]) w.spirwin[L * oc.rwin + j]++;
            }
            if (step % one_ms == 0) {
              const uint64_t s1 = step / one_ms;
              const unsigned m = (unsigned)(s1 % oc.rwin);
              if (s1 % oc.fpnt == 0) {
                std::ostringstream o;
                o << std::noshowpoint << std::setprecision(7) << evtime * 1e-3f;
                if (oc.status & OF_SUMALL) {
                  unsigned mm = 0, nn = 0;
                  for (size_t j = 0; j < L; ++j) {
                    mm += w.spirwin[L * (oc.rwin + 1) + j];
                    nn += c->cfg.pops[j].n;  // the reference indexes ConfSta by list position
                  }
                  o << " " << 1e3f * mm / oc.rwin / nn;
                } else {
                  for (size_t j = 0; j < L; ++j)
                    o << " " << (1e3f * w.spirwin[L * (oc.rwin + 1) + j] / oc.rwin / c->cfg.pops[j].n);
                }
                w.f << o.str() << std::endl;
              }
              for (size_t j = 0; j < L; ++j) {
                w.spirwin[L * (oc.rwin + 1) + j] -= w.spirwin[j * oc.rwin + m];
                w.spirwin[L * (oc.rwin + 1) + j] += w.spirwin[L * oc.rwin + j];
                w.spirwin[j * oc.rwin + m] = w.spirwin[L * oc.rwin + j];
                w.spirwin[L * oc.rwin + j] = 0;
              }
            }
          } else if (oc.type_id == OUT_MEMPOT) {
            std::ostringstream o;
            o << std::noshowpoint << std::setprecision(7) << evtime * 1e-3f;
            for (uint32_t m : oc.pops)
              for (uint32_t nid = c->net.pop_first[m]; nid < c->net.pop_first[m] + c->net.pop_n[m]; ++nid)
                o << " " << (float)vbuf[nid] * 1e-3f;
            w.f << o.str() << std::endl;
          }
        }
        for (int64_t q = off[k]; q < off[k + 1]; ++q) spiked[ids[q]] = 0;
      }
      s = e;
    }
    const float evlast = (float)(unsigned)(endtime ? endtime - 1 : 0) * dt;
    std::cout << "Expertment time=" << std::setprecision(7) << evlast * 1e-3f << "s 100%\n";
  }
  return FS_OK;
}

}  // namespace fs

extern "C" int fs_cli_main(int argc, const char *const *argv) {
  FS_TRY_BEGIN
  return fs::cli_main(argc, argv);
  FS_TRY_END
}
