/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement of the reference Flysim GNL2 step (the per-0.1 ms update:
 * external Poisson input, conductance LIF membrane update with refractory
 * state machine, per-synapse gate update and postsynaptic pull-sum), written
 * from the reference's equations, single-threaded in the reference's
 * operation order, and templated on float (bit-faithful FP32 restatement) and
 * double (the FP64 oracle).  Only the test-suite, the package smoke check and
 * the CPU-baseline leg of the benchmark load this library, as the checker or
 * as the timed CPU path; the product never calls it.
 *
 * Reference files followed (paths relative to the reference root):
 *   code/EqsGnl2.cpp:152-251   VmEul, firing
 *   code/EqsGnl2.cpp:284-337   BgInit (external Poisson gates)
 *   code/EqsGnl2.cpp:339-401   VarInit
 *   code/EqsGnl2.cpp:404-502   ThdSynapse (single-synapse gates + pull-sum)
 *   code/EqsGnl2.cpp:866-1002  ThdPGatVar / ThdPSynapse (group gates + gather)
 *   code/EqsGnl2.cpp:505-863, 1048-1090  STP / STD gates and presynaptic state
 *   code/EqsGnl2.cpp:705-725  ThdGapSyn (gap junction currents)
 *   code/EqsGnl2.cpp:200-216, 254-280, 1005-1044  LTP (traces, delayed LTP spike,
 *                              single-synapse weight rules)
 *   code/EqsGnl2.cpp:1098-1131 Activation / Spike dispatch
 *   code/NetSimGnl2.cpp:52-141 DoEvt,  :231-330 SimOne / Run seeding
 *   code/NetSimGnl2_BulPp.cpp:159-278 per-neuron parameter derivation
 *   /usr/include/c++/11/bits/random.tcc:3348-3379 generate_canonical<float>
 *
 * Parity pins (see tests/): the exact raster of the reference built from
 * source on the 1k E/I network (golden fixtures), the closed forms of the
 * membrane decay / conductance steady state / NMDA fixed point / refractory
 * freeze, and libstdc++'s own uniform_real_distribution<float>.
 */
#ifndef FLYSIM_ORACLE_H
#define FLYSIM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Raw population parameters as parsed from network.conf (float, as the
 * reference stores them in NetConf / NeuRep, code/parser.h:9-66). */
typedef struct {
  float cm, taum, vth, el, vreset;
  int32_t rfped, spidy;           /* already converted to steps */
  /* receptor params, index AMPA=0 GABA=1 NMDA=2 ACH=3 GCL=4 */
  int32_t declared[5];
  float tau[5], tau_xt[5], vrev[5], freq_ext[5], ext_eff[5], ext_con[5];
  /* STP_pv, STP_tD, STP_aF, STP_tF, STP_aCap, STP_tCap, STP_aCar, STP_tCar */
  float stp[8];
  /* LTP_tLP (ms), LTP_SpiLPDy (steps), LTP_PosISI, LTP_NegISI */
  float ltp[4];
} or_pop;

typedef struct {
  int32_t n_neuron, n_pop;
  const int32_t *neuron_pop;     /* [n_neuron] */
  const or_pop *pops;            /* [n_pop] */
  /* gates: one per (pre neuron x link) group gate or per single synapse */
  int32_t n_gate;
  const uint32_t *gate_pre;      /* neuron whose SPIKE_F drives the gate */
  const uint8_t *gate_rec;       /* receptor 0..4 */
  const float *gate_g;           /* efficacy nS */
  const uint32_t *gate_post0;    /* post neuron whose receptor decay is used */
  /* pull-sum rows in reference execution order; term_gate lists gate ids */
  int32_t n_row;
  const uint32_t *row_post;
  const uint8_t *row_rec;
  const uint32_t *row_beg;       /* [n_row+1] */
  const uint32_t *term_gate;
  uint32_t act_ctr;              /* global ActCtr (code/main.cpp:263) */
  float dt;                      /* ms */
  /* events: sorted by time as in code/parser.cpp:1395 */
  int32_t n_evt;
  const float *evt_time_s;       /* TimEv in seconds (float) */
  const int32_t *evt_type;       /* 1 = ChangeExtFreq, 2 = CurrentInjec, 0 = other */
  const int32_t *evt_pop;
  const int32_t *evt_sub;        /* receptor keyword KW_* for ChangeExtFreq */
  const float *evt_var1, *evt_var2;
  const uint8_t *gate_group;     /* [n_gate] 1: group gate, 0: single synapse (LTP learns these) */
  /* single-synapse gap junctions in SynCnd1 order (sorted by post) */
  int32_t n_gap;
  const uint32_t *gap_pre, *gap_post;
  const float *gap_g;
} or_net;

typedef struct or_sim or_sim;

/* precision: 0 = float (reference arithmetic), 1 = double */
or_sim *or_create(const or_net *net, int precision);
void or_destroy(or_sim *s);
/* VarInit + pb reset + mt19937 reseed (code/NetSimGnl2.cpp:266-330). */
void or_reset(or_sim *s, uint32_t seed);
/* Replay mode: external gate bits instead of the mt19937 stream.
 * bits[(step*5 + r)*words + n/32] bit n%32 ; words = ceil(N/32). */
void or_set_ext_bits(or_sim *s, const uint32_t *bits, int64_t n_steps);
/* Record mode: store the draw outcomes (pb > u) of the internal stream. */
void or_record_ext_bits(or_sim *s, uint32_t *bits, int64_t n_steps);
/* Run n steps; spikes (SPIKE_F set) appended as (step, nid) pairs in step,
 * nid order, up to cap pairs.  Returns the number of pairs written, or -1 if
 * cap was exceeded. */
int64_t or_run(or_sim *s, int64_t n_steps, int64_t *spk_step, int32_t *spk_nid,
               int64_t cap);
int64_t or_step_index(const or_sim *s);
/* fields: 0 V, 1 Count, 2 NState, 3..7 g A,G,N,H,L (RVs), 8 Isyn,
 * 9..13 ext st A,G,H,L,  14 ext NMDA xt, 15 ext NMDA st.  Output as double. */
void or_get(const or_sim *s, int field, double *out);
/* gate state: 0 st, 1 xt */
void or_get_gates(const or_sim *s, int field, double *out);
/* derived per-neuron parameters (code/NetSimGnl2_BulPp.cpp:159-278):
 * 0 df 1 ef 2 gl 3 dVmax 4 dfst_A 5 dfst_G 6 dfst_N 7 dfxt_N 8 efst_N
 * 9 pb_A 10 gext_A */
void or_get_params(const or_sim *s, int field, double *out);

/* Reference uniform draw: generate_canonical<float,24>(mt19937) */
void or_mt19937_uniform(uint32_t seed, int64_t n, float *out);

/* Thread-parallel pull-sum CPU path (same algorithm class as the reference's
 * std::thread strided partition) used only as the timed CPU baseline. */
double or_time_run(or_sim *s, int64_t n_steps, int threads);

#ifdef __cplusplus
}
#endif
#endif
