// Network builder: turns a parsed Config into the engine's synapse list and
// per-population parameter table, realising the connectome exactly as the
// reference's NetSimGnl2::SetPp / PostSetPp do (code/NetSimGnl2_BulPp.cpp:
// 27-727): the same seed selection, the same mt19937 consumption, the same
// std::shuffle target choice (code/NetSim.h:403-454) and the same group-gate /
// single-synapse split.  It also emits the reference's internal layout
// (gates and post-synaptic pull-sum rows in execution order), which is what
// a CPU restatement of the reference needs to reproduce its summation order.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "netconf.hpp"
#include "../../../include/flysim.h"

namespace fs {

// receptor ids of the engine (RVs member order, code/neuron_affi.h:648-656)
enum : uint8_t { R_AMPA = 0, R_GABA = 1, R_NMDA = 2, R_ACH = 3, R_GCL = 4, R_GAP = 5, R_NONE = 255 };
uint8_t receptor_of_keyword(uint8_t kw);

struct BuildOptions {
  int threads = 1;          // -t (consumes mt19937 draws before the connectome)
  int repeats = 1;          // -rp
  bool udf_seed = false;    // -udfsed given
  uint32_t udf_seed_value = 0;
  uint32_t act_ctr = 0;     // global ActCtr
};

struct BuiltNet {
  uint32_t n_neuron = 0;
  std::vector<uint32_t> neuron_pop, pop_first, pop_n;
  std::vector<fs_pop_params> pops;   // derived parameters per population
  // engine synapse list (any order)
  std::vector<uint32_t> pre, post;
  std::vector<float> w;
  std::vector<uint8_t> rec;
  std::vector<uint8_t> plastic;      // 1: single-synapse path (LTP learns it)
  // reference layout: gates (group gates first, then single synapses)
  std::vector<uint32_t> gate_pre, gate_post0;
  std::vector<uint8_t> gate_rec, gate_group;
  std::vector<float> gate_g;
  // pull-sum rows in execution order (group rows, then single rows)
  std::vector<uint32_t> row_post, row_beg, term;
  std::vector<uint8_t> row_rec;
  // gap junctions of the single-synapse path (SynCnd1 after its sort by
  // PosID, code/NetSimGnl2_BulPp.cpp:683-712); PreID is the population index
  std::vector<uint32_t> gap_pre, gap_post;
  std::vector<float> gap_g;
  std::vector<uint32_t> rseeds;      // one per repeat
  uint64_t total_synapses = 0;       // TolSynp as the reference reports it
  uint64_t gap_synapses = 0;
};

// Derived per-population parameters (code/NetSimGnl2_BulPp.cpp:159-278).
fs_pop_params derive_pop_params(const PopDecl &p, float dt, uint32_t act_ctr);

void build_network(const Config &cfg, const BuildOptions &opt, BuiltNet &out);

}  // namespace fs
