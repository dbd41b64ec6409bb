// Output writers; see writers.hpp.
#include "writers.hpp"

#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace fs {

OutputWriters::OutputWriters(const Config &cfg, const BuiltNet &net, int repeat,
                             const std::string &dir)
    : cfg_(cfg), net_(net), spiked_(net.n_neuron, 0) {
  one_ms_ = (unsigned)(1u / cfg.dt);  // OneMsStep=(unsigned int)1/dt (code/NetSimGnl2.cpp:272)
  if (one_ms_ == 0) one_ms_ = 1;
  for (const OutCtrl &oc : cfg.outs) {
    // SaveAllStates and SynapticWeight files are opened like every output
    // file but stay empty: FileOut has nothing for them
    // (code/NetSim.cpp:47-62, code/NetSimGnl2_info.cpp:760-761)
    const bool empty_file = oc.type_id == OUT_SAVE || oc.type_id == OUT_SYNWGT;
    if (!empty_file && oc.type_id != OUT_SPIKE && oc.type_id != OUT_FRATE &&
        oc.type_id != OUT_MEMPOT && oc.type_id != OUT_DVOVF && oc.type_id != OUT_ISYNSEP &&
        oc.type_id != OUT_GATVAR) {
      unsupported_.push_back(oc.type);
      continue;
    }
    ws_.emplace_back();
    W &w = ws_.back();
    w.oc = oc;
    std::string fn = oc.file;
    if (repeat > 1) fn += "_" + std::to_string((long long)repeat);
    if (!dir.empty()) fn = dir + "/" + fn;
    w.f.open(fn);
    if (!w.f) throw std::runtime_error("cannot open output file " + fn);
    w.f << std::noshowpoint << std::setprecision(7);
    if (oc.type_id == OUT_FRATE) {
      if (oc.rwin == 0 || oc.fpnt == 0) throw std::runtime_error("FiringRate window/print step must be positive");
      w.spirwin.assign(oc.pops.size() * (oc.rwin + 2), 0u);
    }
    if (oc.type_id == OUT_MEMPOT) need_ |= NEED_V;
    if (oc.type_id == OUT_DVOVF) need_ |= NEED_FLAGS;
    if (oc.type_id == OUT_ISYNSEP) need_ |= NEED_IREP;
    if (oc.type_id == OUT_GATVAR) need_ |= NEED_GATES;
  }
  if (need_ & NEED_GATES) {
    gate_cnt_.assign((size_t)net.n_neuron * 5, 0u);
    for (size_t g = 0; g < net.gate_pre.size(); ++g)
      if (!net.gate_group[g] && net.gate_pre[g] < net.n_neuron && net.gate_rec[g] < 5)
        gate_cnt_[(size_t)net.gate_pre[g] * 5 + net.gate_rec[g]]++;
  }
}

void OutputWriters::step(uint64_t step, const uint32_t *ids, int64_t n_ids, const StepState &st) {
  const double *v = st.v;
  const float evtime = (float)(unsigned)step * cfg_.dt;  // EvTime=i*dt
  for (int64_t q = 0; q < n_ids; ++q) spiked_[ids[q]] = 1;
  for (W &w : ws_) {
    const OutCtrl &oc = w.oc;
    if (oc.type_id == OUT_SPIKE) {  // code/NetSimGnl2_info.cpp:624-648
      // lines are formatted in a default-formatted string stream (6
      // significant digits), as the reference's ostr_tmp member is; the
      // file stream's own setprecision(7) never applies to them
      std::ostringstream o;
      unsigned n = 0, kk = 0;
      for (uint32_t m : oc.pops) {
        for (uint32_t nid = net_.pop_first[m]; nid < net_.pop_first[m] + net_.pop_n[m]; ++nid) {
          if (spiked_[nid]) {
            o << evtime * 1e-3f << " " << ((oc.status & OF_ALIGN) ? n : nid) << "\n";
            ++kk;
          }
          ++n;
        }
      }
      if (kk) w.f << o.str() << std::flush;
    } else if (oc.type_id == OUT_FRATE) {  // code/NetSimGnl2_info.cpp:650-699
      const size_t L = oc.pops.size();
      for (size_t j = 0; j < L; ++j) {
        const uint32_t m = oc.pops[j];
        for (uint32_t nid = net_.pop_first[m]; nid < net_.pop_first[m] + net_.pop_n[m]; ++nid)
          if (spiked_[nid]) w.spirwin[L * oc.rwin + j]++;
      }
      if (step % one_ms_ == 0) {
        const uint64_t s1 = step / one_ms_;
        const unsigned m = (unsigned)(s1 % oc.rwin);
        if (s1 % oc.fpnt == 0) {
          std::ostringstream o;
          o << evtime * 1e-3f;
          if (oc.status & OF_SUMALL) {
            unsigned mm = 0, nn = 0;
            for (size_t j = 0; j < L; ++j) {
              mm += w.spirwin[L * (oc.rwin + 1) + j];
              nn += cfg_.pops.at(j).n;  // ConfSta indexed by list position, as the reference does
            }
            o << " " << 1e3f * mm / oc.rwin / nn;
          } else {
            for (size_t j = 0; j < L; ++j)
              o << " " << (1e3f * w.spirwin[L * (oc.rwin + 1) + j] / oc.rwin / cfg_.pops.at(j).n);
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
    } else if (oc.type_id == OUT_MEMPOT) {  // code/NetSimGnl2_info.cpp:575-587
      if (!v) throw std::runtime_error("MemPot output needs membrane potentials");
      std::ostringstream o;
      o << evtime * 1e-3f;
      for (uint32_t m : oc.pops)
        for (uint32_t nid = net_.pop_first[m]; nid < net_.pop_first[m] + net_.pop_n[m]; ++nid)
          o << " " << (float)v[nid] * 1e-3f;
      w.f << o.str() << std::endl;
    } else if (oc.type_id == OUT_DVOVF) {  // code/NetSimGnl2_info.cpp:589-622
      if (!st.nstate) throw std::runtime_error("DeltaMemPotLimit output needs the NState flags");
      std::ostringstream o;
      unsigned n = 0, kk = 0;
      for (uint32_t m : oc.pops) {
        for (uint32_t nid = net_.pop_first[m]; nid < net_.pop_first[m] + net_.pop_n[m]; ++nid) {
          const unsigned f = (unsigned)st.nstate[nid];
          const unsigned id = ((oc.status & OF_ALIGN) ? n : nid) + 1;
          if (f & 16u) {  // MEMDV_OVF
            o << evtime * 1e-3f << " " << id << "\n";
            ++kk;
          } else if (f & 32u) {  // MEMDV_UDF
            o << evtime * 1e-3f << " -" << id << "\n";
            ++kk;
          }
          ++n;
        }
      }
      if (kk) w.f << o.str() << std::flush;
    } else if (oc.type_id == OUT_ISYNSEP) {  // code/NetSimGnl2_info.cpp:559-573
      for (int r = 0; r < 5; ++r)
        if (!st.irep[r]) throw std::runtime_error("IsynSeparatedSum output needs the receptor currents");
      std::ostringstream o;
      o << evtime * 1e-3f;
      for (uint32_t m : oc.pops)
        for (uint32_t nid = net_.pop_first[m]; nid < net_.pop_first[m] + net_.pop_n[m]; ++nid)
          for (int r = 0; r < 5; ++r) o << " " << (float)st.irep[r][nid] * 1e-12f;
      w.f << o.str() << std::endl;
    } else if (oc.type_id == OUT_GATVAR) {  // code/NetSimGnl2_info.cpp:701-752
      if (!st.spre || !st.xn || !st.sn)
        throw std::runtime_error("GatingVariable output needs the presynaptic gates");
      // a single synapse's gate equals its presynaptic neuron's gate (one
      // decay per presynaptic population), printed once per synapse in the
      // receptor order AMPA, GABA, ACh, GluCl, then NMDA x and s
      std::ostringstream o;
      o << evtime * 1e-3f << " ";
      static const int lin[4] = {0, 1, 3, 4};
      for (uint32_t m : oc.pops)
        for (uint32_t nid = net_.pop_first[m]; nid < net_.pop_first[m] + net_.pop_n[m]; ++nid) {
          const uint32_t *c = &gate_cnt_[(size_t)nid * 5];
          for (int r : lin)
            for (uint32_t q = 0; q < c[r]; ++q) o << (float)st.spre[nid] << " ";
          for (uint32_t q = 0; q < c[2]; ++q) o << (float)st.xn[nid] << " " << (float)st.sn[nid] << " ";
        }
      w.f << o.str() << std::endl;
    }
  }
  for (int64_t q = 0; q < n_ids; ++q) spiked_[ids[q]] = 0;
}

}  // namespace fs
