// Output files of the drop-in protocol, byte-compatible with the reference's
// FileOut (code/NetSimGnl2_info.cpp:548-699) and FileOpen
// (code/NetSim.cpp:45-66): Spike, FiringRate and MemPot.
#pragma once
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include "builder.hpp"
#include "netconf.hpp"

namespace fs {

class OutputWriters {
 public:
  // Opens every Spike / FiringRate / MemPot file of cfg.outs (name suffixed
  // with "_<repeat>" when repeat > 1) under `dir` (empty: current directory).
  OutputWriters(const Config &cfg, const BuiltNet &net, int repeat, const std::string &dir);
  bool needs_potential() const { return need_v_; }
  // One step's FileOut: `ids` are the neurons with SPIKE_F set (any order),
  // `v` the membrane potentials (mV) when needs_potential().
  void step(uint64_t step, const uint32_t *ids, int64_t n_ids, const double *v);
  std::vector<std::string> unsupported() const { return unsupported_; }

 private:
  struct W {
    OutCtrl oc;
    std::ofstream f;
    std::vector<uint32_t> spirwin;
  };
  const Config &cfg_;
  const BuiltNet &net_;
  std::vector<W> ws_;
  std::vector<uint8_t> spiked_;
  bool need_v_ = false;
  unsigned one_ms_;
  std::vector<std::string> unsupported_;
};

}  // namespace fs
