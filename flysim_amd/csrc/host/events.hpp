// Event schedule of the drop-in protocol: applies ChangeExtFreq and
// ChangeMembraneNoise/CurrentInjec events to per-population parameters at
// the start of a step, exactly as NetSimGnl2::DoEvt (code/NetSimGnl2.cpp:
// 52-141) applies them to its per-neuron copies.
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

#include "builder.hpp"
#include "netconf.hpp"

namespace fs {

class EventCursor {
 public:
  EventCursor(const std::vector<Event> &ev, float dt) : ev_(ev), dt_(dt) {}
  void rewind() { i_ = 0; }
  // Apply the events of step `step`; returns true if any population changed.
  // `changed` receives the indices of the populations that changed.
  bool apply(uint64_t step, std::vector<fs_pop_params> &pops, std::vector<uint32_t> *changed) {
    const float evtime = (float)(unsigned)step * dt_;
    bool any = false;
    for (size_t i = i_; i < ev_.size(); ++i) {
      const float tms = ev_[i].t_s * 1e3f;
      if (tms >= evtime && tms < evtime + dt_) {
        i_ = i;
        const Event &e = ev_[i];
        if (e.pop >= pops.size()) continue;
        fs_pop_params &p = pops[e.pop];
        if (e.type == EVT_EXTFREQ) {
          uint32_t bit;
          int rid;
          switch (e.sub) {
            case KW_AMPA: bit = FS_ACT_AMPA; rid = FS_AMPA; break;
            case KW_GABA: bit = FS_ACT_GABA; rid = FS_GABA; break;
            case KW_ACH: bit = FS_ACT_ACH; rid = FS_ACH; break;
            case KW_GCL: bit = FS_ACT_GCL; rid = FS_GCL; break;
            case KW_NMDA: bit = FS_ACT_NMDA; rid = FS_NMDA; break;
            default: continue;
          }
          if (e.var1 == 0.0f) p.act &= ~bit;
          else { p.act |= bit; p.pb[rid] = (float)(e.var1 * dt_ * 1e-3f); }
        } else if (e.type == EVT_CURINJ) {
          p.inj_mean = e.var1;
          p.inj_std = e.var2 * (1.0f + (float)std::log((1.0f - std::exp(-0.1f / p.taum)) / (1.0f - p.df)));
          if (e.var1 == 0.0f && e.var2 == 0.0f) p.act &= ~(uint32_t)FS_ACT_INJ;
          else p.act |= FS_ACT_INJ;
        } else {
          continue;
        }
        any = true;
        if (changed) changed->push_back(e.pop);
      } else if (tms >= evtime + dt_) {
        break;
      }
    }
    return any;
  }
 private:
  const std::vector<Event> &ev_;
  float dt_;
  size_t i_ = 0;
};

}  // namespace fs
