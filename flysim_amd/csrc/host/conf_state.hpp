// Loaded drop-in configuration: parsed files, built network and the flat
// arrays handed out through the C ABI.
#pragma once
#include <vector>

#include "builder.hpp"
#include "netconf.hpp"

struct fs_conf {
  fs::Config cfg;
  fs::BuiltNet net;
  fs::BuildOptions opt;
  std::vector<fs_pop_conf> pop_conf;
  std::vector<int32_t> pop_rfped, pop_spidy;
  std::vector<float> evt_t, evt_v1, evt_v2;
  std::vector<int32_t> evt_type, evt_pop, evt_sub;
};
