// Drop-in parser for the Flysim network.conf / network.pro text formats.
//
// Behaviour follows the reference token FSMs (code/parser.cpp:51-294 CSta,
// :297-907 conf, :911-1029 macros, :1032-1407 protocol, :1424-1464 command
// line outputs): same delimiters, same keywords, same defaults, the same
// float conversions (RefractoryPeriod -> int(floor(x/dt)+dt)), the same
// silent fallbacks (unknown population name -> population 0) and the same
// event expansion and ordering.  Values are kept in float, as the reference
// keeps them.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace fs {

// parser keywords (code/Inc.h:35-47)
enum : uint8_t {
  KW_AMPA = 0, KW_GABA = 1, KW_NMDA = 2, KW_GAP = 3, KW_CUBA = 4, KW_SINE = 5,
  RSM_SYNCOND = 6, RSM_GATVAR = 7, RSM_MEMPOT = 8, RSM_RNDSED = 9, RSM_ALL = 10,
  KW_ACH = 11, KW_GCL = 12
};
// event types (code/Inc.h:176-177)
enum : uint32_t { EVT_NONE = 0, EVT_EXTFREQ = 1, EVT_CURINJ = 2 };
// output type ids (code/Inc.h:183-193)
enum : int {
  OUT_NONE = 0, OUT_MEMPOT = 1, OUT_SPIKE = 2, OUT_FRATE = 3, OUT_GATVAR = 5,
  OUT_ISYNSEP = 6, OUT_SAVE = 7, OUT_SYNWGT = 9, OUT_DVOVF = 10
};
enum : uint32_t { OF_ALIGN = 1, OF_SUMALL = 2 };

struct ReceptorDecl {
  uint8_t kw = KW_AMPA;
  float tau = 5.0f, tau_xt = 2.0f, vrev = 0.0f, fext = 0.0f, ext_eff = 2.1f, ext_con = 1.0f;
};

struct LinkDecl {
  uint32_t target = 0;
  uint8_t rec_kw = KW_AMPA;
  float conn = 1.0f, mean_eff = 0.0f, weight = 1.0f;
};

struct StpDecl {  // code/parser.cpp:229-243
  float pv = 0.6f, tD = 300.0f, aF = 0.0035f, tF = 7000.0f, aCap = 250.0f, tCap = 2.0f,
        aCar = 0.1f, tCar = 2000.0f;
};

struct LtpDecl {  // code/parser.cpp:246-253
  float tLP = 50.0f;
  int32_t SpiLPDy = 0;
  float PosISI = 1e-5f, NegISI = 1e-5f;
};

struct PopDecl {
  std::string name;
  uint32_t n = 0;
  bool self_con = false;
  float taum = 2.0f, vth = -50.0f, el = -70.0f, gl = 2.5f, cm = 1.0f, vreset = -50.0f;
  int32_t rfped = 0, spidy = 0;
  StpDecl stp;
  LtpDecl ltp;
  std::vector<ReceptorDecl> recs;
  std::vector<LinkDecl> links;
};

struct Event {
  float t_s = 0.0f;  // TimEv (seconds)
  bool nolimit = false;
  uint32_t type = EVT_NONE;
  uint32_t pop = 0;
  uint8_t sub = KW_AMPA;
  float var1 = 0.0f, var2 = 0.0f;
};

struct OutCtrl {
  std::string type = "Non", file = "Non";
  int type_id = OUT_NONE;
  std::vector<uint32_t> pops;
  uint32_t rwin = 50, fpnt = 100, status = 0;
  float t_s = 0.0f;
};

struct Group {
  std::string name;
  std::vector<uint32_t> pops;
};

struct Config {
  float dt = 0.1f;
  std::vector<PopDecl> pops;
  std::vector<Event> events;     // expanded and sorted (code/parser.cpp:1368-1395)
  std::vector<OutCtrl> outs;
  std::vector<Group> groups;
  std::vector<uint32_t> seeds;   // RandomSeeds macro
  // command-line output overrides (code/main.cpp:133-149, parser.cpp:1424-1464)
  bool comm_mem = false, comm_spi = false, comm_fr = false;
  std::string out_mem = "MemPot.dat", out_spi = "Spikes.dat", out_fr = "FRate.dat";
  std::vector<std::string> warnings;
};

// Full drop-in pipeline of code/main.cpp:233-247: CSta, parse(conf), Comm,
// DefMacro, parse(pro).  Throws std::runtime_error on unreadable files or
// malformed numbers (where the reference prints and continues, or aborts with
// an uncaught std::stoi/stof exception).
void load_config(Config &cfg, const std::string &conf_path, const std::string &pro_path);

// Tokeniser used by all FSMs: delimiters " :;,'=\t\r\n" (code/parser.cpp:28).
std::vector<std::string> tokenize(const std::string &text);

}  // namespace fs
