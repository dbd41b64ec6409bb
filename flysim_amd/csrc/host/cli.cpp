// Drop-in command line of the Flysim core: argument handling of
// code/main.cpp:26-256 and the step driver of code/NetSimGnl2.cpp:231-358
// (events, repeats, output files) with the solver replaced by the GPU engine.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "builder.hpp"
#include "conf_state.hpp"
#include "events.hpp"
#include "host_common.hpp"
#include "netconf.hpp"
#include "writers.hpp"

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
            << "-STP / -STD           # short-term plasticity (facilitation + depression / depression)\n"
            << "-LTP                  # long-term plasticity of single synapses\n"
            << "-nmodel GNL2          # neuron model (GNL2 only)\n"
            << "-udfsed N             # user-defined random seed (switches STD on, as the reference does)\n"
            << "-fp64                 # double-precision state (default float)\n"
            << "-ext replay|philox    # external input: the reference's mt19937 stream\n"
            << "                      # (replay) or the counter-based generator (default)\n";
}

// Steps at which DoEvt changes some population, with the new parameters.
struct Schedule {
  std::vector<uint64_t> steps;
  std::vector<std::vector<std::pair<uint32_t, fs_pop_params>>> changes;
};

Schedule event_schedule(const Config &cfg, std::vector<fs_pop_params> pops, uint64_t endtime) {
  Schedule s;
  EventCursor cur(cfg.events, cfg.dt);
  std::vector<uint32_t> ch;
  for (uint64_t t = 0; t < endtime; ++t) {
    ch.clear();
    if (cur.apply(t, pops, &ch)) {
      std::sort(ch.begin(), ch.end());
      ch.erase(std::unique(ch.begin(), ch.end()), ch.end());
      s.steps.push_back(t);
      s.changes.emplace_back();
      for (uint32_t p : ch) s.changes.back().emplace_back(p, pops[p]);
    }
  }
  return s;
}

}  // namespace

int cli_main(int argc, const char *const *argv) {
  std::string conf = "network.conf", pro = "network.pro";
  std::string om = "MemPot.dat", os_ = "Spikes.dat", orr = "FRate.dat";
  bool om_en = false, os_en = false, or_en = false, stp = false, stdp = false, ltp = false;
  bool udf = false, fp64 = false, replay = false, help = false;
  uint32_t seed = 0;
  int rp = 1, th = 1;
  float dt = 0.1f;
  for (int i = 1; i < argc && !help; ++i) {
    const std::string s = argv[i];
    std::string v;
    auto val = [&](std::string &dst) {
      if (i + 1 < argc) dst = argv[++i];
      else help = true;
    };
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
    else help = true;
  }
  if (help) { print_help(); return FS_OK; }
  fs_cli_opts o{};
  o.stp = stp;
  o.std_ = stdp;
  o.ltp = ltp;
  o.udf_seed = udf;
  o.udf_seed_value = seed;
  o.dt = dt; o.threads = th; o.repeats = rp;
  o.out_mem = om_en ? om.c_str() : nullptr;
  o.out_spi = os_en ? os_.c_str() : nullptr;
  o.out_fr = or_en ? orr.c_str() : nullptr;
  fs_conf *c = nullptr;
  int rc = fs_conf_load(conf.c_str(), pro.c_str(), &o, &c);
  if (rc != FS_OK) return rc;
  std::unique_ptr<fs_conf, int (*)(fs_conf *)> cg(c, fs_conf_free);
  for (const std::string &w : c->cfg.warnings) std::cerr << "warning: " << w << "\n";
  if (c->cfg.events.empty()) return set_error(FS_ERR_ARG, "the protocol has no events (no end time)");
  const Event &last = c->cfg.events.back();
  if (last.nolimit) return set_error(FS_ERR_ARG, "EventTime NoLimit is not supported");
  const uint64_t endtime = (uint64_t)std::floor(last.t_s * 1e3f / dt);  // code/NetSimGnl2.cpp:270
  rc = fs_conf_write_logs(c, "");  // ConfsInfo.log and network.log (code/main.cpp:247, NetSimGnl2.cpp:154)
  if (rc != FS_OK) return rc;
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
  std::unique_ptr<fs_handle, int (*)(fs_handle *)> hg(h, fs_destroy);
  const uint32_t N = c->net.n_neuron;
  // activity bits persist across repeats; pb is reset from the conf (code/NetSimGnl2.cpp:279-297)
  std::vector<fs_pop_params> pops = c->net.pops;
  for (int r = 1; r <= std::max(1, rp); ++r) {
    std::cout << "Iteration:" << r << "\n";
    for (size_t i = 0; i < pops.size(); ++i)
      for (int k = 0; k < 5; ++k) pops[i].pb[k] = c->net.pops[i].pb[k];
    for (size_t i = 0; i < pops.size(); ++i) {
      rc = fs_set_pop_params(h, (int32_t)i, &pops[i]);
      if (rc != FS_OK) return rc;
    }
    rc = fs_reset(h, c->net.rseeds[r - 1]);
    if (rc != FS_OK) return rc;
    if (replay && endtime) {
      const int64_t words = (N + 31) / 32;
      std::vector<uint32_t> bits((size_t)endtime * 5 * words, 0u);
      rc = fs_conf_ext_trains(c, r - 1, (int64_t)endtime, bits.data());
      if (rc != FS_OK) return rc;
      rc = fs_load_ext_trains(h, bits.data(), 0, (int64_t)endtime);
      if (rc != FS_OK) return rc;
    }
    const Schedule sch = event_schedule(c->cfg, pops, endtime);
    OutputWriters out(c->cfg, c->net, r, "");
    for (const std::string &u : out.unsupported())
      std::cerr << "warning: output type " << u << " is not written by the GPU engine\n";
    // outputs that print per-step state run one step per fs_run
    const unsigned need = out.needs();
    if (need & OutputWriters::NEED_IREP) {
      rc = fs_record_currents(h, 1);
      if (rc != FS_OK) return rc;
    }
    std::vector<double> vbuf((need & OutputWriters::NEED_V) ? N : 0);
    std::vector<double> fbuf((need & OutputWriters::NEED_FLAGS) ? N : 0);
    std::vector<double> gbuf((need & OutputWriters::NEED_GATES) ? 3 * (size_t)N : 0);
    std::vector<double> ibuf((need & OutputWriters::NEED_IREP) ? 5 * (size_t)N : 0);
    StepState st;
    if (!vbuf.empty()) st.v = vbuf.data();
    if (!fbuf.empty()) st.nstate = fbuf.data();
    if (!gbuf.empty()) {
      st.spre = gbuf.data();
      st.xn = gbuf.data() + N;
      st.sn = gbuf.data() + 2 * (size_t)N;
    }
    for (int r = 0; r < 5 && !ibuf.empty(); ++r) st.irep[r] = ibuf.data() + (size_t)r * N;
    const int64_t batch_max = need ? 1 : 1000;
    std::vector<int64_t> off;
    std::vector<uint32_t> ids;
    uint64_t s = 0;
    size_t ei = 0;
    while (s < endtime) {
      if (ei < sch.steps.size() && sch.steps[ei] == s) {
        for (const auto &pc : sch.changes[ei]) {
          pops[pc.first] = pc.second;
          rc = fs_set_pop_params(h, (int32_t)pc.first, &pc.second);
          if (rc != FS_OK) return rc;
        }
        ++ei;
      }
      uint64_t e = std::min<uint64_t>(endtime, s + batch_max);
      if (ei < sch.steps.size() && sch.steps[ei] < e) e = sch.steps[ei];
      const int64_t K = (int64_t)(e - s);
      off.assign(K + 1, 0);
      ids.resize((size_t)N * K);
      rc = fs_run(h, K, off.data(), ids.data(), (int64_t)ids.size());
      if (rc != FS_OK) return rc;
      // fs_get_state fields: 0 V, 2 NState, 8/9/10 presynaptic gates, 19.. Irep
      if (st.v && (rc = fs_get_state(h, 0, vbuf.data())) != FS_OK) return rc;
      if (st.nstate && (rc = fs_get_state(h, 2, fbuf.data())) != FS_OK) return rc;
      if (st.spre) {
        for (int f = 0; f < 3; ++f)
          if ((rc = fs_get_state(h, 8 + f, gbuf.data() + (size_t)f * N)) != FS_OK) return rc;
      }
      for (int r = 0; r < 5 && st.irep[0]; ++r)
        if ((rc = fs_get_state(h, 19 + r, ibuf.data() + (size_t)r * N)) != FS_OK) return rc;
      for (int64_t k = 0; k < K; ++k)
        out.step(s + (uint64_t)k, ids.data() + off[k], off[k + 1] - off[k], st);
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
