// The two description files the reference writes before it simulates:
// ConfsInfo.log, the parsed configuration (Parser::ConfsInfo,
// code/parser.cpp:1488-1625), and network.log, the built network
// (NetSimGnl2::logDataWrite, code/NetSimGnl2_info.cpp:93-541): per-neuron
// derived parameters, the synapse arrays in solver order, the expanded event
// list and the output controls with their column headers.  Line formats are
// the reference's (default ostream float formatting, the same separators and
// line breaks), so tools that read these files read ours.
#include "logs.hpp"

#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace fs {

namespace {

const char *rec_name(uint8_t kw) {  // NeuRepName / TarReceptor spellings
  switch (kw) {
    case KW_AMPA: return "AMPA";
    case KW_GABA: return "GABA";
    case KW_NMDA: return "NMDA";
    case KW_ACH: return "Ach";
    case KW_GCL: return "GluCl";
    default: return "";
  }
}

// Parser defaults of the Hodgkin-Huxley block (code/parser.cpp:256-293); the
// LIF model never reads them, but ConfsInfo prints them
void write_hh(std::ostream &o) {
  const float vr = -70.0f;  // VRESTING
  o << "Ek=" << 0.0f << " gk=" << 0.0f << " Ena=" << 0.0f << " gna=" << 0.0f << "\n"
    << "AnA=" << 0.01f << " AnB=" << 10.0f + vr << " AnC=" << 10.0f << " AnD=" << -1.0f << "\n"
    << "BnA=" << 0.125f << " BnB=" << 0.0f + vr << " BnC=" << 80.0f << " BnD=" << 0.0f << "\n"
    << "AmA=" << 0.1f << " AmB=" << 25.0f + vr << " AmC=" << 10.0f << " AmD=" << -1.0f << "\n"
    << "BmA=" << 4.0f << " BmB=" << 0.0f + vr << " BmC=" << 18.0f << " BmD=" << 0.0f << "\n"
    << "AhA=" << 0.07f << " AhB=" << 0.0f + vr << " AhC=" << 20.0f << " AhD=" << 0.0f << "\n"
    << "BhA=" << 1.0f << " BhB=" << 30.0f + vr << " BhC=" << 10.0f << " BhD=" << 1.0f << "\n";
}

// one single-synapse array section: "Name:count;pre,post,g;..." with a line
// break after every element whose array index is a multiple of 10
void single_section(std::ostream &o, const char *name, const BuiltNet &net,
                    const std::vector<uint32_t> &gates, uint64_t index0) {
  if (gates.empty()) return;
  o << name << ":Num;PreID,PosID,g;PreID,PosID,g...\n" << name << ":" << gates.size();
  uint64_t i = index0;
  for (uint32_t g : gates) {
    o << ";" << net.gate_pre[g] << "," << net.gate_post0[g] << "," << net.gate_g[g];
    if (i % 10 == 0) o << "\n";
    ++i;
  }
  o << std::endl;
}

}  // namespace

void write_confs_info(const Config &cfg, const std::string &path) {
  std::ofstream o(path);
  if (!o) throw std::runtime_error("cannot write " + path);
  o << "configuration information:----------------------------------------\n";
  for (size_t i = 0; i < cfg.pops.size(); ++i) {
    const PopDecl &p = cfg.pops[i];
    o << "population:" << i << "\n"
      << "NeuPopuName:" << p.name << ", N=" << p.n << " Taum=" << p.taum << " Vth=" << p.vth
      << " El=" << p.el << " g=" << p.gl << " Cm=" << p.cm << " Vreset=" << p.vreset << "\n";
    write_hh(o);
    o << "pv=" << p.stp.pv << " tD=" << p.stp.tD << " aF=" << p.stp.aF << " tF=" << p.stp.tF
      << " aCap=" << p.stp.aCap << " tCap=" << p.stp.tCap << " aCar=" << p.stp.aCar
      << " tCar=" << p.stp.tCar << "\n";
    o << "tLP=" << p.ltp.tLP << " SpiLPDy=" << p.ltp.SpiLPDy << " PosISI=" << p.ltp.PosISI
      << " NegISI=" << p.ltp.NegISI << "\n";
    for (const ReceptorDecl &r : p.recs) {
      o << "NeuRepName:" << rec_name(r.kw) << ", Tau=" << r.tau;
      if (r.kw == KW_NMDA) o << ", TauXT=" << r.tau_xt;
      o << " Vrv=" << r.vrev << " FExt=" << r.fext << " MnExtEff=" << r.ext_eff
        << " MnExtCon=" << r.ext_con << "\n";
    }
    o << "\n";
    for (const LinkDecl &l : p.links) {
      o << "TarPopuName:" << cfg.pops[l.target].name << " TarReceptor:"
        << (l.rec_kw == KW_GAP ? "GAP" : rec_name(l.rec_kw)) << ", Connectivity=" << l.conn
        << " MeanEff=" << l.mean_eff << " weight=" << l.weight << "\n";
    }
    o << "------------------------------------------------------------------" << std::endl;
  }
}

void write_network_log(const Config &cfg, const BuiltNet &net, const BuildOptions &opt,
                       const std::string &path) {
  std::ofstream o(path);
  if (!o) throw std::runtime_error("cannot write " + path);
  // SInfo (code/NetSimGnl2_info.cpp:24-58)
  o << "\ncapable mode of sim07_21 leaky integrate and fire model\ntime step:" << cfg.dt
    << ", total neurons=" << net.n_neuron << ", total synapses=" << net.total_synapses << "\n"
    << "\n";
  // UDFRDSEED_ACT shares STD_ACT's bit (code/Inc.h:275,280)
  o << ((opt.act_ctr & FS_ACT_STD) ? "User define random seed enable!\n"
                                   : "internal random seed enable!\n");
  o << "repeats=" << opt.repeats << ",each repeat random seed=";
  for (int r = 0; r < opt.repeats; ++r) o << " " << net.rseeds[r];
  o << "\n\npopulation information:-------------------------------------------\n";
  const float dt = cfg.dt;
  for (uint32_t i = 0; i < net.n_neuron; ++i) {
    const fs_pop_params &p = net.pops[net.neuron_pop[i]];
    o << "-------------------------------------------\nneuron:" << i << "\n";
    // every neuron carries the global ActCtr (code/NetSimGnl2_BulPp.cpp:163)
    o << "ActCtr=" << opt.act_ctr << " Vth=" << p.vth << "mv El=" << p.el << "mv g=" << p.gl
      << "uS Cm=" << p.cm << "nF Vreset=" << p.vreset << "mv df=" << p.df << " ef=" << p.ef
      << " dVmax=" << p.dvmax << "mv\n";
    o << "AMPA df=" << p.dfst[0] << " Vrev=" << p.vrev[0] << "mV GABA df=" << p.dfst[1]
      << " Vrev=" << p.vrev[1] << "mV\n"
      << "NMDA dfxt=" << p.dfxt_n << " dfst=" << p.dfst[2] << " efst=" << p.efst_n
      << " Vrev=" << p.vrev[2] << "mV as=" << p.as_n << " mg=" << p.mg_n << "\n"
      << "ACH df=" << p.dfst[3] << " Vrev=" << p.vrev[3] << "mV GCL df=" << p.dfst[4]
      << " Vrev=" << p.vrev[4] << "mV\n";
    if (opt.act_ctr & FS_ACT_LIF)
      o << "LIF: RfPed=" << p.rfped * dt << "ms SpiDy=" << p.spidy * dt << "ms\n";
    if (opt.act_ctr & FS_ACT_STP)
      o << "STP: pv=" << p.stp_pv << " dfD=" << p.stp_dfD << " efD=" << p.stp_efD
        << " aF=" << p.stp_aF << " dfF=" << p.stp_dfF << " efF=" << p.stp_efF
        << " aCap=" << p.stp_aCap << " dfCap=" << p.stp_dfCap << " aCar=" << p.stp_aCar
        << " dfCar=" << p.stp_dfCar << " efCar=" << p.stp_efCar << "\n";
    if (opt.act_ctr & FS_ACT_LTP)
      o << "LTP: df=" << p.ltp_dfLP << " SpiLPDy=" << p.ltp_spidy << " PosISI=" << p.ltp_pos
        << " NegISI=" << p.ltp_neg << "\n";
    // BgSt mean / std start at MNoiseMean = MNoiseSTD = 0 (code/Inc.h:29-30);
    // SetPp then rescales the std of BgSt[population index], not of the
    // neuron (code/NetSimGnl2_BulPp.cpp:177): a negative factor, so neurons
    // 0 .. P-1 log -0
    float sd = 0.0f;
    if (i < cfg.pops.size())
      sd = 0.0f * (1.0f + (float)std::log(1.0f - std::exp(-0.1f / cfg.pops[i].taum)) /
                              (1.0f - net.pops[i].df));
    o << "noise parameters:\nInjection current:mean=" << 0.0f << " stddev=" << sd << "\n";
    // BgSt pb / efficacy order: AMPA, GABA, ACh, GluCl, NMDA
    o << "Ext AMPA:prob=" << p.pb[0] << ",g=" << p.gext[0] << "\n"
      << "Ext GABA:prob=" << p.pb[1] << ",g=" << p.gext[1] << "\n"
      << "Ext Ach:prob=" << p.pb[3] << ",g=" << p.gext[3] << "\n"
      << "Ext GluCl:prob=" << p.pb[4] << ",g=" << p.gext[4] << "\n"
      << "Ext NMDA:prob=" << p.pb[2] << ",g=" << p.gext[2] << "\n" << std::flush;
  }
  o << "\nsynapse connection information----------------------------------\n";
  // single synapses: SynFst1 holds AMPA | GABA | ACh | GluCl segments, SynSlo1
  // NMDA, SynCnd1 gap junctions; the builder keeps each in solver order
  std::vector<uint32_t> single[5];
  for (uint32_t g = 0; g < net.gate_pre.size(); ++g)
    if (!net.gate_group[g]) single[net.gate_rec[g]].push_back(g);
  uint64_t base = 0;
  const char *fst[4] = {"SynAMPA1", "SynGABA1", "SynACH1", "SynGCL1"};
  const int fst_rec[4] = {R_AMPA, R_GABA, R_ACH, R_GCL};
  for (int k = 0; k < 4; ++k) {
    single_section(o, fst[k], net, single[fst_rec[k]], base);
    base += single[fst_rec[k]].size();
  }
  single_section(o, "SynNMDA1", net, single[R_NMDA], 0);
  if (!net.gap_pre.empty()) {
    o << "SynCnd1:Num;PreID,PosID,g;PreID,PosID,g...\nSynCnd1:" << net.gap_pre.size();
    for (size_t i = 0; i < net.gap_pre.size(); ++i) {
      o << ";" << net.gap_pre[i] << "," << net.gap_post[i] << "," << net.gap_g[i];
      if (i % 10 == 0) o << "\n";
    }
    o << std::endl;
  }
  // group gates: the gather rows of each receptor (NeuMapP* / PSyn*), terms
  // numbered within the receptor's own PSyn array
  const char *grp[5] = {"PSynAMPA", "PSynGABA", "PSynNMDA", "PSynACH", "PSynGCL"};
  const int grp_rec[5] = {R_AMPA, R_GABA, R_NMDA, R_ACH, R_GCL};
  for (int k = 0; k < 5; ++k) {
    uint64_t count = 0;
    for (size_t r = 0; r < net.row_post.size(); ++r)
      if (net.row_rec[r] == grp_rec[k])
        for (uint32_t t = net.row_beg[r]; t < net.row_beg[r + 1]; ++t)
          if (net.gate_group[net.term[t]]) ++count;
    if (!count) continue;
    o << grp[k] << ":" << grp[k] << "size;PreID,PosID,g;PreID,PosID,g...\n"
      << grp[k] << ":" << count << ";";
    uint64_t i = 0;
    for (size_t r = 0; r < net.row_post.size(); ++r) {
      if (net.row_rec[r] != grp_rec[k]) continue;
      for (uint32_t t = net.row_beg[r]; t < net.row_beg[r + 1]; ++t) {
        const uint32_t g = net.term[t];
        if (!net.gate_group[g]) continue;
        o << net.gate_pre[g] << "," << net.row_post[r] << "," << net.gate_g[g] << ";";
        if (i % 10 == 0) o << "\n";
        ++i;
      }
    }
    o << std::endl;
  }
  // TmEvt::TexInfo (code/parser_affi.cpp:30-48)
  o << "\ntime event information:-------------------------------------------\n";
  for (const Event &e : cfg.events) {
    std::ostringstream s;
    s << "Event Time=" << e.t_s;
    if (e.type == EVT_EXTFREQ) s << ", type:ChangeExtFreq";
    else if (e.type == EVT_CURINJ) s << ", type:CurrentInjec";
    s << ", Target population:" << e.pop << ", Receptor:" << (unsigned)e.sub
      << ", Variable1=" << e.var1 << ", Variable2=" << e.var2 << "\n";
    o << s.str() << std::flush;
  }
  // OutCtr::TexInfo (code/parser_affi.cpp:75-92) and the column headers
  o << "\noutput files information:-----------------------------------------\n";
  static const char *lin_tag[4] = {"_AMAP_st ", "_GABA_st ", "_ACH_st ", "_GCL_st "};
  for (const OutCtrl &oc : cfg.outs) {
    o << "Output File Name:" << oc.file << ", Type:" << oc.type << "\nPopulation:";
    for (size_t i = 0; i < oc.pops.size(); ++i) {
      o << oc.pops[i] << ", ";
      if (i % 10 == 0) o << "\n";
    }
    o << "\n";
    if (oc.type_id == OUT_ISYNSEP) {
      o << "1:Time";
      unsigned kk = 2;
      for (uint32_t m : oc.pops)
        for (uint32_t k = net.pop_first[m]; k < net.pop_first[m] + net.pop_n[m]; ++k) {
          o << " NID=" << k;
          for (const char *r : {"AMPA", "GABA", "NMDA", "ACH", "GCL"}) o << " " << kk++ << ":" << r << " ";
        }
      o << "\n";
    } else if (oc.type_id == OUT_MEMPOT) {
      o << "1:Time";
      unsigned kk = 2;
      for (uint32_t m : oc.pops)
        for (uint32_t k = net.pop_first[m]; k < net.pop_first[m] + net.pop_n[m]; ++k)
          o << " " << kk++ << ":NID=" << k;
      o << "\n";
    } else if (oc.type_id == OUT_FRATE) {
      o << "FiringRateWinodw=" << oc.rwin << "ms PrintStep=" << oc.fpnt << "ms\n1:Time";
      if (oc.status & OF_SUMALL) o << "2:total firing rate";
      else
        for (size_t j = 0, kk = 2; j < oc.pops.size(); ++j) o << " " << kk++ << ":NID=" << j;
      o << "\n";
    } else if (oc.type_id == OUT_GATVAR) {
      // single synapses grouped by presynaptic id, each group in array order
      // (BulDPreNeuMap, code/Eqs.h:1168-1199)
      o << "1:Time ";
      unsigned kk = 2;
      static const int lin_rec[4] = {R_AMPA, R_GABA, R_ACH, R_GCL};
      for (uint32_t m : oc.pops)
        for (uint32_t k = net.pop_first[m]; k < net.pop_first[m] + net.pop_n[m]; ++k) {
          for (int q = 0; q < 4; ++q)
            for (uint32_t g : single[lin_rec[q]])
              if (net.gate_pre[g] == k)
                o << kk++ << ":NID=" << k << "_PosID=" << net.gate_post0[g] << lin_tag[q];
          for (uint32_t g : single[R_NMDA])
            if (net.gate_pre[g] == k) {
              o << kk++ << ":NID=" << k << "_PosID=" << net.gate_post0[g] << "_NMDA_xt ";
              o << kk++ << ":NID=" << k << "_PosID=" << net.gate_post0[g] << "_NMDA_st ";
            }
        }
    }
    o << std::endl;
  }
}

}  // namespace fs
