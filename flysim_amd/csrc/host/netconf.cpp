// Drop-in network.conf / network.pro parser; see netconf.hpp.
#include "netconf.hpp"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace fs {

namespace {

// FSM states; numbering is private to this file.
enum State {
  S_INITIAL, S_POPNAME, S_POP, S_N, S_TAUM, S_VTH, S_GL, S_CM, S_VREST, S_VRESET,
  S_RFPED, S_SPIDY, S_SELFCON, S_STP_PV, S_STP_TD, S_STP_AF, S_STP_TF, S_STP_ACAP,
  S_STP_TCAP, S_STP_ACAR, S_STP_TCAR, S_LTP_TLP, S_LTP_SPILPDY, S_LTP_POSISI,
  S_LTP_NEGISI, S_SKIPVAL, S_REPNAME, S_REC, S_REC_TAU, S_REC_TAUXT, S_REC_VREV,
  S_REC_FEXT, S_REC_EXTEFF, S_REC_EXTCON, S_LINKNAME, S_LINK, S_LINK_CONN,
  S_LINK_REC, S_LINK_EFF, S_LINK_WGT,
  // macro
  S_DEFMAC, S_MAC_GPNAME, S_MAC_GPMBS, S_MAC_SEEDS,
  // protocol
  S_TIMEV, S_TIMEVDEF, S_TIMTYP, S_TIMLBL, S_TIMPP, S_TIMSUB, S_TIMVAR1, S_TIMVAR2,
  S_OUTCTR, S_OUTNAME, S_OUTTYPE, S_OUTRWIN, S_OUTFPNT, S_OUTTIMEV, S_OUTPOPU,
  S_OUTSUMALL, S_OUTALIGN
};

std::string read_file(const std::string &path, const char *what) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error(std::string(what) + " open file error: " + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

bool rec_keyword(const std::string &s, uint8_t &kw) {
  // code/parser.cpp:731-746 (conf) and :842-857 (link target receptor)
  if (s == "AMPA") kw = KW_AMPA;
  else if (s == "GABA") kw = KW_GABA;
  else if (s == "NMDA") kw = KW_NMDA;
  else if (s == "GAP") kw = KW_GAP;
  else if (s == "Sine") kw = KW_SINE;
  else if (s == "Ach") kw = KW_ACH;
  else if (s == "GluCl") kw = KW_GCL;
  else return false;
  return true;
}

float to_f(const std::string &s) { return std::stof(s); }

}  // namespace

namespace {

// The reference reads its input through istream::get(ch, 65536, EOF): at most
// 65535 bytes per call, stopping before a 0xFF byte, eofbit set when the byte
// after the last one read is the end of the file, failbit when nothing was
// read (libstdc++ basic_istream::get(char_type*, streamsize, char_type)).
struct ChunkReader {
  const std::string &data;
  size_t off = 0;
  bool eof = false, fail = false;
  explicit ChunkReader(const std::string &d) : data(d) {}
  std::string get() {
    if (eof || fail) { fail = true; return std::string(); }
    size_t i = off;
    while (i < data.size() && i - off < 65535 && (unsigned char)data[i] != 0xFFu) ++i;
    std::string r = data.substr(off, i - off);
    if (i == data.size()) eof = true;
    if (r.empty()) fail = true;
    off = i;
    return r;
  }
};

}  // namespace

std::vector<std::string> tokenize(const std::string &data) {
  // Token stream of the reference's read loop (code/parser.cpp:70-171, the
  // same loop in every FSM pass), including its chunk-boundary behaviour:
  // a chunk that starts with a delimiter yields one empty token, and a token
  // cut by a chunk boundary is re-joined with the next chunk.
  static const std::string delims = " :;,'=\t\r\n";
  const std::string::size_type npos = std::string::npos;
  std::vector<std::string> out;
  ChunkReader fin(data);
  std::string str = fin.get(), s1;
  std::string::size_type startpos = 0, pos = str.find_first_of(delims, 0);
  while (pos != npos || startpos != npos) {
    s1 = str.substr(startpos, pos - startpos);
    startpos = str.find_first_not_of(delims, pos);
    pos = str.find_first_of(delims, startpos);
    out.push_back(s1);
    if (fin.eof) {
    } else if (startpos == npos) {
      if (fin.fail) throw std::runtime_error("input contains a 0xFF byte (the reference loops forever)");
      str = fin.get();
      startpos = 0;
      pos = str.find_first_of(delims, startpos);
    } else if (pos == npos) {
      if (fin.fail) throw std::runtime_error("input contains a 0xFF byte (the reference loops forever)");
      s1 = str.substr(startpos, str.size() - 1);
      str = s1 + fin.get();
      startpos = 0;
      pos = str.find_first_of(delims, startpos);
    }
  }
  return out;
}

void load_config(Config &cfg, const std::string &conf_path, const std::string &pro_path) {
  const float dt = cfg.dt;
  int state = S_INITIAL;  // shared by all passes, as in the reference Parser
  // ---------------------------------------------------------------- CSta
  const std::vector<std::string> ct = tokenize(read_file(conf_path, ".conf"));
  struct Sta { std::string name; uint32_t nrec = 0, nlink = 0; bool selfcon = false; };
  std::vector<Sta> sta;
  for (const std::string &s1 : ct) {
    switch (state) {
      case S_INITIAL:
        if (s1 == "NeuralPopulation") { sta.emplace_back(); state = S_POPNAME; }
        break;
      case S_POPNAME: sta.back().name = s1; state = S_POP; break;
      case S_POP:
        if (s1 == "Receptor") { sta.back().nrec++; state = S_REC; }
        else if (s1 == "TargetPopulation") { sta.back().nlink++; state = S_LINK; }
        else if (s1 == "SelfConnection") state = S_SELFCON;
        else if (s1 == "EndNeuralPopulation") state = S_INITIAL;
        break;
      case S_REC: if (s1 == "EndReceptor") state = S_POP; break;
      case S_LINK: if (s1 == "EndTargetPopulation") state = S_POP; break;
      case S_SELFCON: sta.back().selfcon = (s1 == "true"); state = S_POP; break;
      default: break;
    }
  }
  // defaults (code/parser.cpp:178-253)
  cfg.pops.assign(sta.size(), PopDecl());
  for (size_t i = 0; i < sta.size(); ++i) {
    PopDecl &p = cfg.pops[i];
    p.name = sta[i].name;
    p.self_con = sta[i].selfcon;
    p.recs.assign(sta[i].nrec, ReceptorDecl());
    p.links.assign(sta[i].nlink, LinkDecl());
    p.taum = 2.0f; p.vth = -50.0f; p.el = -70.0f; p.gl = 2.5f; p.cm = 1.0f; p.vreset = -50.0f;
    p.rfped = (int32_t)(1.8f / dt + dt);
    p.spidy = (int32_t)0.0f;
  }
  auto find_pop = [&](const std::string &s, uint32_t &idx) {
    for (size_t i = 0; i < cfg.pops.size(); ++i)
      if (cfg.pops[i].name == s) { idx = (uint32_t)i; return true; }
    return false;
  };
  // ---------------------------------------------------------------- parse(conf)
  {
    size_t Ni = 0, Ri = 0, Li = 0;
    uint8_t kw = KW_AMPA;
    auto pop = [&]() -> PopDecl & {
      if (Ni >= cfg.pops.size()) throw std::runtime_error(".conf: population index out of range");
      return cfg.pops[Ni];
    };
    for (const std::string &s1 : ct) {
      switch (state) {
        case S_INITIAL: if (s1 == "NeuralPopulation") state = S_POPNAME; break;
        case S_POPNAME: state = S_POP; break;
        case S_POP:
          if (s1 == "N") state = S_N;
          else if (s1 == "Taum") state = S_TAUM;
          else if (s1 == "Threshold") state = S_VTH;
          else if (s1 == "LeakyConductance") state = S_GL;
          else if (s1 == "C") state = S_CM;
          else if (s1 == "RevPot" || s1 == "RestPot") state = S_VREST;
          else if (s1 == "ResetPot") state = S_VRESET;
          else if (s1 == "RefractoryPeriod") state = S_RFPED;
          else if (s1 == "SpikeDly") state = S_SPIDY;
          else if (s1 == "STP_pv") state = S_STP_PV;
          else if (s1 == "STP_tD") state = S_STP_TD;
          else if (s1 == "STP_aF") state = S_STP_AF;
          else if (s1 == "STP_tF") state = S_STP_TF;
          else if (s1 == "STP_aCap") state = S_STP_ACAP;
          else if (s1 == "STP_tCap") state = S_STP_TCAP;
          else if (s1 == "STP_aCar") state = S_STP_ACAR;
          else if (s1 == "STP_tCar") state = S_STP_TCAR;
          else if (s1 == "LTP_tLP") state = S_LTP_TLP;
          else if (s1 == "LTP_SpiLPDy") state = S_LTP_SPILPDY;
          else if (s1 == "LTP_PosISI") state = S_LTP_POSISI;
          else if (s1 == "LTP_NegISI") state = S_LTP_NEGISI;
          else if (s1 == "VthSodium" || s1.compare(0, 3, "HH_") == 0) {
            // Hodgkin-Huxley / sodium keys of other models: value parsed and
            // discarded (the GNL2 model never reads ConfHH).
            static const char *hh[] = {"HH_Ek", "HH_gk", "HH_Ena", "HH_gna", "HH_AnA", "HH_AnB",
                                       "HH_AnC", "HH_AnD", "HH_BnA", "HH_BnB", "HH_BnC", "HH_BnD",
                                       "HH_AmA", "HH_AmB", "HH_AmC", "HH_AmD", "HH_BmA", "HH_BmB",
                                       "HH_BmC", "HH_BmD", "HH_AhA", "HH_AhB", "HH_AhC", "HH_AhD",
                                       "HH_BhA", "HH_BhB", "HH_BhC", "HH_BhD"};
            bool known = s1 == "VthSodium";
            for (const char *k : hh) known = known || s1 == k;
            if (known) state = S_SKIPVAL;
          } else if (s1 == "Receptor") state = S_REPNAME;
          else if (s1 == "TargetPopulation") state = S_LINKNAME;
          else if (s1 == "EndNeuralPopulation") { Ni++; Ri = 0; Li = 0; state = S_INITIAL; }
          break;
        case S_N: pop().n = (uint32_t)std::stoi(s1); state = S_POP; break;
        case S_TAUM: pop().taum = to_f(s1); state = S_POP; break;
        case S_VTH: pop().vth = to_f(s1); state = S_POP; break;
        case S_GL: pop().gl = to_f(s1); state = S_POP; break;  // parsed, unused (gl=C/Taum)
        case S_CM: pop().cm = to_f(s1); state = S_POP; break;
        case S_VRESET: pop().vreset = to_f(s1); state = S_POP; break;
        case S_VREST: pop().el = to_f(s1); state = S_POP; break;
        case S_RFPED: pop().rfped = (int32_t)(std::floor(to_f(s1) / dt) + dt); state = S_POP; break;
        case S_SPIDY: pop().spidy = (int32_t)std::floor(to_f(s1) / dt); state = S_POP; break;
        case S_STP_PV: pop().stp.pv = to_f(s1); state = S_POP; break;
        case S_STP_TD: pop().stp.tD = to_f(s1); state = S_POP; break;
        case S_STP_AF: pop().stp.aF = to_f(s1); state = S_POP; break;
        case S_STP_TF: pop().stp.tF = to_f(s1); state = S_POP; break;
        case S_STP_ACAP: pop().stp.aCap = to_f(s1); state = S_POP; break;
        case S_STP_TCAP: pop().stp.tCap = to_f(s1); state = S_POP; break;
        case S_STP_ACAR: pop().stp.aCar = to_f(s1); state = S_POP; break;
        case S_STP_TCAR: pop().stp.tCar = to_f(s1); state = S_POP; break;
        case S_LTP_TLP: pop().ltp.tLP = to_f(s1); state = S_POP; break;
        case S_LTP_SPILPDY: pop().ltp.SpiLPDy = (int32_t)std::floor(to_f(s1) / dt); state = S_POP; break;
        case S_LTP_POSISI: pop().ltp.PosISI = to_f(s1); state = S_POP; break;
        case S_LTP_NEGISI: pop().ltp.NegISI = to_f(s1); state = S_POP; break;
        case S_SKIPVAL: (void)to_f(s1); state = S_POP; break;
        case S_REPNAME:
          rec_keyword(s1, kw);  // unknown name keeps the previous keyword
          if (Ri >= pop().recs.size()) throw std::runtime_error(".conf: receptor index out of range");
          pop().recs[Ri].kw = kw;
          state = S_REC;
          break;
        case S_REC:
          if (s1 == "Tau") state = S_REC_TAU;
          else if (s1 == "TauXT") state = S_REC_TAUXT;
          else if (s1 == "RevPot") state = S_REC_VREV;
          else if (s1 == "FreqExt") state = S_REC_FEXT;
          else if (s1 == "MeanExtEff") state = S_REC_EXTEFF;
          else if (s1 == "MeanExtCon") state = S_REC_EXTCON;
          else if (s1 == "EndReceptor") { Ri++; state = S_POP; }
          break;
        case S_REC_TAU: pop().recs.at(Ri).tau = to_f(s1); state = S_REC; break;
        case S_REC_TAUXT: pop().recs.at(Ri).tau_xt = to_f(s1); state = S_REC; break;
        case S_REC_VREV: pop().recs.at(Ri).vrev = to_f(s1); state = S_REC; break;
        case S_REC_FEXT: pop().recs.at(Ri).fext = to_f(s1); state = S_REC; break;
        case S_REC_EXTEFF: pop().recs.at(Ri).ext_eff = to_f(s1); state = S_REC; break;
        case S_REC_EXTCON: pop().recs.at(Ri).ext_con = to_f(s1); state = S_REC; break;
        case S_LINKNAME: {
          uint32_t idx;
          if (find_pop(s1, idx)) pop().links.at(Li).target = idx;
          else cfg.warnings.push_back("unknown target population '" + s1 + "' maps to population 0");
          state = S_LINK;
          break;
        }
        case S_LINK:
          if (s1 == "Connectivity") state = S_LINK_CONN;
          else if (s1 == "TargetReceptor") state = S_LINK_REC;
          else if (s1 == "MeanEff") state = S_LINK_EFF;
          else if (s1 == "weight") state = S_LINK_WGT;
          else if (s1 == "EndTargetPopulation") { Li++; state = S_POP; }
          break;
        case S_LINK_CONN: {
          float c = to_f(s1);
          if (c > 1.0f) c = 1.0f;
          pop().links.at(Li).conn = c;
          state = S_LINK;
          break;
        }
        case S_LINK_REC:
          rec_keyword(s1, kw);
          pop().links.at(Li).rec_kw = kw;
          state = S_LINK;
          break;
        case S_LINK_EFF: pop().links.at(Li).mean_eff = to_f(s1); state = S_LINK; break;
        case S_LINK_WGT: pop().links.at(Li).weight = to_f(s1); state = S_LINK; break;
        default: break;
      }
    }
  }
  // ---------------------------------------------------------------- Comm
  auto all_pops = [&]() {
    std::vector<uint32_t> v(cfg.pops.size());
    for (size_t i = 0; i < v.size(); ++i) v[i] = (uint32_t)i;
    return v;
  };
  if (cfg.comm_mem) { OutCtrl o; o.type = "MemPot"; o.type_id = OUT_MEMPOT; o.file = cfg.out_mem; o.pops = all_pops(); cfg.outs.push_back(o); }
  if (cfg.comm_spi) { OutCtrl o; o.type = "Spike"; o.type_id = OUT_SPIKE; o.file = cfg.out_spi; o.pops = all_pops(); cfg.outs.push_back(o); }
  if (cfg.comm_fr) { OutCtrl o; o.type = "FiringRate"; o.type_id = OUT_FRATE; o.file = cfg.out_fr; o.pops = all_pops(); cfg.outs.push_back(o); }
  // ---------------------------------------------------------------- DefMacro
  const std::vector<std::string> pt = tokenize(read_file(pro_path, ".pro"));
  {
    Group g;
    for (const std::string &s1 : pt) {
      switch (state) {
        case S_INITIAL: if (s1 == "DefineMacro") state = S_DEFMAC; break;
        case S_DEFMAC:
          if (s1 == "GroupName") state = S_MAC_GPNAME;
          else if (s1 == "GroupMembers") state = S_MAC_GPMBS;
          else if (s1 == "RandomSeeds") state = S_MAC_SEEDS;
          else if (s1 == "EndDefineMacro") state = S_INITIAL;
          break;
        case S_MAC_GPNAME: g.name = s1; state = S_DEFMAC; break;
        case S_MAC_SEEDS:
          if (s1 == "EndRandomSeeds") state = S_DEFMAC;
          else {
            const unsigned v = (unsigned)std::stoi(s1);
            if (v <= (unsigned)RAND_MAX) cfg.seeds.push_back(v);
          }
          break;
        case S_MAC_GPMBS:
          if (s1 == "EndGroupMembers") { cfg.groups.push_back(g); g.pops.clear(); state = S_DEFMAC; }
          else { uint32_t idx; if (find_pop(s1, idx)) g.pops.push_back(idx); }
          break;
        default: break;
      }
    }
  }
  // ---------------------------------------------------------------- parse(pro)
  {
    std::vector<Event> unsol;
    Event ev;
    OutCtrl of;
    uint8_t kw = KW_AMPA;
    const uint32_t NPOP = (uint32_t)cfg.pops.size();
    for (const std::string &s1 : pt) {
      switch (state) {
        case S_INITIAL:
          if (s1 == "EventTime") state = S_TIMEV;
          else if (s1 == "OutControl") state = S_OUTCTR;
          break;
        case S_TIMEV:
          if (s1 == "NoLimit") { ev.nolimit = true; ev.t_s = FLT_MAX; }
          else { ev.nolimit = false; ev.t_s = to_f(s1) * 1e-3f; }
          state = S_TIMEVDEF;
          break;
        case S_TIMEVDEF:
          if (s1 == "Type") state = S_TIMTYP;
          else if (s1 == "Label" || s1 == "InFile") state = S_TIMLBL;
          else if (s1 == "Population") state = S_TIMPP;
          else if (s1 == "Receptor" || s1 == "SubType") state = S_TIMSUB;
          else if (s1 == "FreqExt" || s1 == "GaussMean") state = S_TIMVAR1;
          else if (s1 == "GaussSTD") state = S_TIMVAR2;
          else if (s1 == "EndEvent") { unsol.push_back(ev); ev = Event(); state = S_INITIAL; }
          break;
        case S_TIMTYP:
          if (s1 == "ChangeExtFreq") ev.type = EVT_EXTFREQ;
          else if (s1 == "ChangeMembraneNoise" || s1 == "CurrentInjec") ev.type = EVT_CURINJ;
          state = S_TIMEVDEF;
          break;
        case S_TIMLBL: state = S_TIMEVDEF; break;
        case S_TIMPP: {
          uint32_t idx;
          if (find_pop(s1, idx)) ev.pop = idx;
          for (size_t i = 0; i < cfg.groups.size(); ++i)
            if (cfg.groups[i].name == s1) { ev.pop = (uint32_t)(i + NPOP); break; }
          if (s1 == "AllPopulation") ev.pop = UINT_MAX;
          state = S_TIMEVDEF;
          break;
        }
        case S_TIMSUB:
          if (!rec_keyword(s1, kw)) {
            if (s1 == "ResumeSynCdt") kw = RSM_SYNCOND;
            else if (s1 == "ResumeGatingVariable") kw = RSM_GATVAR;
            else if (s1 == "ResumeMemPot") kw = RSM_MEMPOT;
            else if (s1 == "ResumeRndSeed") kw = RSM_RNDSED;
            else if (s1 == "ResumeAll") kw = RSM_ALL;
          }
          ev.sub = kw;
          state = S_TIMEVDEF;
          break;
        case S_TIMVAR1: ev.var1 = to_f(s1); state = S_TIMEVDEF; break;
        case S_TIMVAR2: ev.var2 = to_f(s1); state = S_TIMEVDEF; break;
        case S_OUTCTR:
          if (s1 == "FileName") state = S_OUTNAME;
          else if (s1 == "Type") state = S_OUTTYPE;
          else if (s1 == "FiringRateWindow") state = S_OUTRWIN;
          else if (s1 == "PrintStep") state = S_OUTFPNT;
          else if (s1 == "SumAllPopulation") state = S_OUTSUMALL;
          else if (s1 == "Align") state = S_OUTALIGN;
          else if (s1 == "EventTime") state = S_OUTTIMEV;
          else if (s1 == "population") state = S_OUTPOPU;
          else if (s1 == "EndOutputFile") {
            if (of.type != "Non") cfg.outs.push_back(of);
            of = OutCtrl();
            state = S_OUTCTR;
          } else if (s1 == "EndOutControl") state = S_INITIAL;
          break;
        case S_OUTNAME: of.file = s1; state = S_OUTCTR; break;
        case S_OUTTYPE:
          if (s1 == "MemPot" && !cfg.comm_mem) { of.type = s1; of.type_id = OUT_MEMPOT; }
          else if (s1 == "Spike" && !cfg.comm_spi) { of.type = s1; of.type_id = OUT_SPIKE; }
          else if (s1 == "FiringRate" && !cfg.comm_fr) { of.type = s1; of.type_id = OUT_FRATE; }
          else if (s1 == "GatingVariable") { of.type = s1; of.type_id = OUT_GATVAR; }
          else if (s1 == "IsynSeparatedSum") { of.type = s1; of.type_id = OUT_ISYNSEP; }
          else if (s1 == "SaveAllStates") { of.type = s1; of.type_id = OUT_SAVE; }
          else if (s1 == "SynapticWeight") { of.type = s1; of.type_id = OUT_SYNWGT; }
          else if (s1 == "DeltaMemPotLimit") { of.type = s1; of.type_id = OUT_DVOVF; }
          state = S_OUTCTR;
          break;
        case S_OUTRWIN: of.rwin = (uint32_t)std::ceil(to_f(s1)); state = S_OUTCTR; break;
        case S_OUTFPNT: of.fpnt = (uint32_t)std::ceil(to_f(s1)); state = S_OUTCTR; break;
        case S_OUTSUMALL:
          if (s1 == "true") of.status |= OF_SUMALL; else of.status &= ~OF_SUMALL;
          state = S_OUTCTR;
          break;
        case S_OUTALIGN:
          if (s1 == "true") of.status |= OF_ALIGN; else of.status &= ~OF_ALIGN;
          state = S_OUTCTR;
          break;
        case S_OUTTIMEV: of.t_s = to_f(s1) * 1e-3f; state = S_OUTCTR; break;
        case S_OUTPOPU:
          if (s1 == "AllPopulation") {
            for (uint32_t i = 0; i < NPOP; ++i) of.pops.push_back(i);
          } else {
            uint32_t idx;
            if (find_pop(s1, idx)) of.pops.push_back(idx);
            for (const Group &g : cfg.groups)
              if (g.name == s1) { of.pops.insert(of.pops.end(), g.pops.begin(), g.pops.end()); break; }
          }
          state = S_OUTCTR;
          break;
        default: break;
      }
    }
    // macro expansion of events (code/parser.cpp:1368-1392)
    for (const Event &e : unsol) {
      if (e.pop < NPOP) cfg.events.push_back(e);
      else if (e.pop == UINT_MAX) {
        for (uint32_t m = 0; m < NPOP; ++m) { Event x = e; x.pop = m; cfg.events.push_back(x); }
      } else {
        for (size_t j = 0; j < cfg.groups.size(); ++j)
          if (e.pop == NPOP + j)
            for (uint32_t m : cfg.groups[j].pops) { Event x = e; x.pop = m; cfg.events.push_back(x); }
      }
    }
    // the reference orders events with std::sort by time (not stable)
    std::sort(cfg.events.begin(), cfg.events.end(),
              [](const Event &a, const Event &b) { return a.t_s < b.t_s; });
  }
}

}  // namespace fs
