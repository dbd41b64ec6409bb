#pragma once
#include <string>

#include "builder.hpp"
#include "netconf.hpp"

namespace fs {

// ConfsInfo.log: the parsed configuration (code/parser.cpp:1488-1625)
void write_confs_info(const Config &cfg, const std::string &path);

// network.log: the built network before the first step
// (code/NetSimGnl2_info.cpp:93-541)
void write_network_log(const Config &cfg, const BuiltNet &net, const BuildOptions &opt,
                       const std::string &path);

}  // namespace fs
