// Output files of the drop-in protocol, in the formats of the reference's
// FileOut (code/NetSimGnl2_info.cpp:548-758) and FileOpen
// (code/NetSim.cpp:45-66): Spike, FiringRate, MemPot, DeltaMemPotLimit,
// IsynSeparatedSum and GatingVariable.
#pragma once
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include "builder.hpp"
#include "netconf.hpp"

namespace fs {

// Per-neuron state of one step that some outputs print (double[N] each, as
// fs_get_state writes them; null when not needed).
struct StepState {
  const double *v = nullptr;       // membrane potential (mV)
  const double *nstate = nullptr;  // NState bits
  const double *spre = nullptr;    // linear presynaptic gate
  const double *xn = nullptr;      // NMDA presynaptic x
  const double *sn = nullptr;      // NMDA presynaptic s
  const double *irep[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // Irep by receptor (pA)
};

class OutputWriters {
 public:
  enum : unsigned { NEED_V = 1, NEED_FLAGS = 2, NEED_GATES = 4, NEED_IREP = 8 };
  // Opens every supported output file of cfg.outs (name suffixed with
  // "_<repeat>" when repeat > 1) under `dir` (empty: current directory).
  OutputWriters(const Config &cfg, const BuiltNet &net, int repeat, const std::string &dir);
  // NEED_* bits: the per-step state the files need besides the spikes
  unsigned needs() const { return need_; }
  // One step's FileOut: `ids` are the neurons with SPIKE_F set (any order).
  void step(uint64_t step, const uint32_t *ids, int64_t n_ids, const StepState &st);
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
  unsigned need_ = 0;
  unsigned one_ms_;
  std::vector<std::string> unsupported_;
  // GatingVariable: single-synapse gates whose PreID is neuron k, by
  // receptor (code/NetSimGnl2_info.cpp:701-752 walks them per neuron)
  std::vector<uint32_t> gate_cnt_;  // [n_neuron][5]
};

}  // namespace fs
