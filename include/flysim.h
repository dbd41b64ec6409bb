/*
 * flysim.h -- C ABI of the MI355X-native Flysim simulation core.
 *
 * Plain pointers and sizes only; every function returns an int status
 * (FS_OK = 0, negative on error) and fs_last_error() gives the message of the
 * last failing call on the calling thread.  No exceptions cross this ABI.
 *
 * The reference has no interface between its parser, network builder and
 * solver: they are one object (NetSimGnl2 : NetSim, EqsGnl2,
 * code/NetSim.h:367), and the solver is driven by direct calls and public
 * arrays.  Each entry point below names the reference code it replaces.
 */
#ifndef FLYSIM_H
#define FLYSIM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FS_OK = 0,
  FS_ERR_ARG = -1,       /* bad argument / unsupported network */
  FS_ERR_HIP = -2,       /* HIP runtime error */
  FS_ERR_CAPACITY = -3,  /* caller buffer or device record buffer too small */
  FS_ERR_IO = -4,        /* file could not be read or written */
  FS_ERR_PARSE = -5,     /* malformed network.conf / network.pro */
  FS_ERR_STATE = -6      /* call out of sequence */
};

/* receptor ids (member order of the reference's RVs, code/neuron_affi.h:648);
 * FS_GAP marks an electrical synapse (gap junction, code/EqsGnl2.cpp:705-725):
 * symmetric, g (nS) times the voltage difference of the step before enters
 * both neurons' synaptic current; no delay */
enum { FS_AMPA = 0, FS_GABA = 1, FS_NMDA = 2, FS_ACH = 3, FS_GCL = 4, FS_NREC = 5, FS_GAP = 5 };

/* activity bits, same values as the reference's ActCtr (code/Inc.h:253-280) */
enum {
  FS_ACT_AMPA = 1, FS_ACT_GABA = 2, FS_ACT_NMDA = 4, FS_ACT_GAP = 8, FS_ACT_ACH = 32,
  FS_ACT_GCL = 64, FS_ACT_INJ = 128, FS_ACT_LIF = 256, FS_ACT_LTP = 8192,
  FS_ACT_STD = 16384, FS_ACT_STP = 32768
};

enum { FS_FP32 = 0, FS_FP64 = 1 };
/* Random streams.  FS_EXT_PHILOX: counter-based (Philox4x32-10) Poisson draws
 * and Box-Muller Gaussian injection.  FS_EXT_REPLAY: the reference's streams
 * -- external draws loaded with fs_load_ext_trains (its mt19937 outcomes,
 * code/EqsGnl2.cpp:284-332) and Gaussian injection from its ziggurat stream
 * seeded at fs_reset (code/StiExt.h:214-319, code/NetSimGnl2.cpp:323), in
 * single-thread draw order. */
enum { FS_EXT_PHILOX = 0, FS_EXT_REPLAY = 1 };

/* Derived per-population parameters, the values the reference stores per
 * neuron after SetPp (code/NetSimGnl2_BulPp.cpp:159-278); fs_derive_params
 * computes them from conf-level values exactly as the reference does. */
typedef struct fs_pop_params {
  float vth, el, gl, cm, vreset, df, ef, dvmax;
  int32_t rfped, spidy;
  float vrev[5];          /* by receptor id */
  float dfst[5];          /* gate decay per step by receptor id (NMDA: s) */
  float dfxt_n, efst_n, as_n, mg_n;
  float pb[5];            /* external Bernoulli probability per step */
  float gext[5];          /* external efficacy (nS) */
  uint32_t act;           /* FS_ACT_* bits of this population */
  float inj_mean, inj_std;/* Gaussian current injection (nA) */
  float taum;             /* ms */
  /* short-term plasticity of the population's outgoing synapses, used when
   * act has FS_ACT_STP or FS_ACT_STD (code/NetSimGnl2_BulPp.cpp:185-200) */
  float stp_pv, stp_dfD, stp_efD, stp_aF, stp_dfF, stp_efF, stp_aCap, stp_dfCap;
  float stp_aCar, stp_dfCar, stp_efCar;
  /* long-term plasticity of the population's single synapses, used when act
   * has FS_ACT_LTP (code/NetSimGnl2_BulPp.cpp:180-183) */
  float ltp_dfLP, ltp_pos, ltp_neg;
  int32_t ltp_spidy;      /* back-propagation delay of the LTP spike, steps */
} fs_pop_params;

/* Conf-level population values (NetConf + NeuRep, code/parser.h:9-66). */
typedef struct fs_pop_conf {
  float cm, taum, vth, el, vreset;
  float refractory_ms, spike_delay_ms;
  int32_t declared[5];    /* receptor declared, by receptor id */
  float tau[5], tau_xt[5], vrev[5], freq_ext[5], ext_eff[5], ext_con[5];
  /* STP_pv, STP_tD, STP_aF, STP_tF, STP_aCap, STP_tCap, STP_aCar, STP_tCar;
   * reference defaults 0.6, 300, 0.0035, 7000, 250, 2, 0.1, 2000
   * (code/parser.cpp:229-243) */
  float stp[8];
  /* LTP_tLP (ms), LTP_SpiLPDy (steps), LTP_PosISI, LTP_NegISI; reference
   * defaults 50, 0, 1e-5, 1e-5 (code/parser.cpp:246-253) */
  float ltp[4];
} fs_pop_conf;

/* Network handed to fs_create.  Neurons may be ordered arbitrarily; synapses
 * are a COO list in any order.  Independent network instances (parameter
 * sweeps) are simply disjoint neuron ranges of one description. */
typedef struct fs_net_desc {
  int32_t n_neuron, n_pop;
  const uint32_t *neuron_pop;     /* [n_neuron] */
  const fs_pop_params *pops;      /* [n_pop] */
  int64_t n_syn;
  const uint32_t *pre, *post;     /* [n_syn] */
  const float *w;                 /* [n_syn] efficacy, nS */
  const uint8_t *receptor;        /* [n_syn] FS_AMPA..FS_GCL, FS_GAP */
  const uint8_t *delay;           /* [n_syn] extra delay in steps, NULL = 0 */
  const uint8_t *plastic;         /* [n_syn] 1: the weight follows the LTP rules when the
                                     populations carry FS_ACT_LTP (the reference's
                                     single-synapse path, code/EqsGnl2.cpp:1005-1044);
                                     NULL = none */
  float dt;                       /* ms */
  int32_t precision;              /* FS_FP32 | FS_FP64 */
  int32_t ext_mode;               /* FS_EXT_PHILOX | FS_EXT_REPLAY */
  int32_t record_spikes;          /* 1: keep per-step spike lists for readback; 0: fs_run
                                     takes no spike buffers (FS_ERR_ARG), and the windowed
                                     path keeps no spike record at all */
} fs_net_desc;

typedef struct fs_handle fs_handle;

/* Replaces the parameter derivation of NetSimGnl2::SetPp
 * (code/NetSimGnl2_BulPp.cpp:159-278); float arithmetic as the reference. */
int fs_derive_params(const fs_pop_conf *conf, float dt, uint32_t act, fs_pop_params *out);

/* Replaces NetSimGnl2::SetPp + PostSetPp (code/NetSimGnl2_BulPp.cpp:27-727):
 * copies the description, builds CSR-by-pre (linear receptors, push) and
 * CSR-by-post (NMDA, gather) on the device.  Uses device 0 (the current HIP
 * device of the calling thread). */
int fs_create(const fs_net_desc *desc, fs_handle **out);

/* Parameter sweeps over one network (no counterpart in the reference, which
 * runs one network per process): `replicas` copies of the network of desc in
 * one handle, sharing its synapse tables.  Neuron ids are interleaved,
 * global = local * replicas + replica (fs_run ids, fs_get_state order);
 * population ids are replica-major, global = replica * n_pop + pop
 * (fs_set_ext_rate, fs_set_pop_params, fs_pop_spike_counts).  Population
 * parameters other than the external rates and activity bits are shared by
 * all copies.  replicas = 1 is fs_create. */
int fs_create_replicated(const fs_net_desc *desc, int32_t replicas, fs_handle **out);

/* Batched random E/I networks generated on the device (no counterpart in the
 * reference, which runs one network per process and builds it on the host,
 * code/NetSimGnl2_BulPp.cpp:303-727): n_inst independent instances of
 * uneven size in one handle, each n[k] neurons (the first
 * round(exc_frac * n[k]) excitatory), every ordered pair j != p of an
 * instance a synapse with probability p_conn, efficacy g ~ U(g_lo, g_hi) nS;
 * E->X pairs carry AMPA (g) and NMDA (g * nmda_ratio * gain_nmda[k]), I->X
 * pairs GABA (g * gain_gaba[k]); every synapse takes delay_steps steps of
 * synaptic delay (0..63) on top of the SpikeDly of the population
 * parameters.  Pairs and efficacies come from a counter-based hash
 * of (seed[k], j, p), which also keys the instance's external input stream,
 * so an instance runs identically alone or in any batch.  Populations:
 * 2k = E, 2k+1 = I of instance k, external AMPA at ext_hz[k]. */
typedef struct fs_ei_batch_desc {
  int32_t n_inst;
  const int32_t *n_neuron;              /* [n_inst] */
  const uint64_t *seed;                 /* [n_inst] */
  const float *ext_hz;                  /* [n_inst] */
  const float *gain_nmda, *gain_gaba;   /* [n_inst] */
  float exc_frac, p_conn, g_lo, g_hi, nmda_ratio;
  fs_pop_params exc, inh;               /* derived parameters (pb[AMPA] replaced) */
  float dt;
  int32_t precision;
  int32_t ext_mode;
  int32_t record_spikes;
  int32_t delay_steps;                  /* synaptic delay of every synapse (steps) */
} fs_ei_batch_desc;
int fs_create_ei_batch(const fs_ei_batch_desc *desc, fs_handle **out);

/* Per-copy knobs of a replicated handle: gain[r] multiplies every recurrent
 * efficacy of receptor r, ext_gain every external efficacy, and stream_key
 * is xor-ed into the reset seed to key the copy's counter-based stream (the
 * copy then draws what a single network with that key draws).  Defaults:
 * gains 1, stream_key = replica * 0x9E3779B97F4A7C15. */
typedef struct fs_replica_params {
  float gain[5];
  float ext_gain;
  uint64_t stream_key;
} fs_replica_params;
int fs_set_replica(fs_handle *h, int32_t replica, const fs_replica_params *p);

/* Replaces EqsGnl2::VarInit + the reseeding of NetSimGnl2::Run
 * (code/EqsGnl2.cpp:339-401, code/NetSimGnl2.cpp:277-323). */
int fs_reset(fs_handle *h, uint64_t seed);

/* ChangeExtFreq semantics of NetSimGnl2::DoEvt (code/NetSimGnl2.cpp:60-119):
 * hz == 0 clears the receptor's activity bit (no draws, external gate frozen),
 * otherwise sets it and pb = (float)(hz*dt*1e-3f). */
int fs_set_ext_rate(fs_handle *h, int32_t pop, int32_t receptor, float hz);
/* The same for every population at once: hz[n_pop] (sweep drivers). */
int fs_set_ext_rates(fs_handle *h, int32_t receptor, const float *hz, int32_t n_pop);

/* ChangeMembraneNoise / CurrentInjec semantics (code/NetSimGnl2.cpp:120-133).
 * In FS_EXT_REPLAY mode the handle must hold one instance whose populations
 * are contiguous neuron ranges (FS_ERR_ARG otherwise). */
int fs_set_injection(fs_handle *h, int32_t pop, float mean, float std_dev);

/* Overwrite one population's derived parameters (sweeps, gains). */
int fs_set_pop_params(fs_handle *h, int32_t pop, const fs_pop_params *p);

/* Replay mode: external gate outcomes, one bit per (step, receptor, neuron):
 * bits[((step - first_step)*5 + receptor)*ceil(N/32) + n/32] >> (n%32) & 1,
 * covering steps [first_step, first_step + n_steps).  Replaces the mt19937
 * draws of EqsGnl2::BgInit (code/EqsGnl2.cpp:284-332). */
int fs_load_ext_trains(fs_handle *h, const uint32_t *bits, int64_t first_step, int64_t n_steps);

/* Replaces NetSimGnl2::SimOne's loop body for n_steps steps
 * (code/NetSimGnl2.cpp:231-264: Activation + Spike; DoEvt and FileOut stay
 * with the caller).  Spikes (SPIKE_F) of step s are returned in
 * ids[offsets[s-s0] .. offsets[s-s0+1]) in ascending neuron order when
 * offsets/ids are given (offsets has n_steps+1 entries).  The steps run in
 * batches of the schedule's steps_per_graph; FS_ERR_CAPACITY (ids_cap
 * exceeded, device record overflow) comes after the batch that overflowed
 * has run and its spikes are lost: fs_step_index reports how far the state
 * went.  ids_cap >= n_neuron * n_steps never overflows. */
int fs_run(fs_handle *h, int64_t n_steps, int64_t *offsets, uint32_t *ids, int64_t ids_cap);

/* Number of the next step to run. */
int fs_step_index(fs_handle *h, int64_t *step);

/* State readback (the reference's public arrays NetVBse/NetVLIF/RVs):
 * 0 V, 1 refractory count, 2 NState flags, 3..7 synaptic conductance of the
 * last Spike phase by receptor A,G,N,H,L, 8 linear presynaptic gate,
 * 9 NMDA presynaptic x, 10 NMDA presynaptic s, 11..14 external gate of
 * AMPA, GABA, ACh, GluCl, 15/16 external NMDA x/s, 17/18 presynaptic STP
 * depression D / facilitation F (VarStp, with -STP or -STD only), 19..23
 * synaptic current by receptor A,G,N,H,L of the last non-refractory step in
 * pA (Nets[].Irep, with fs_record_currents on).  Written as double[N]. */
int fs_get_state(fs_handle *h, int32_t field, double *out);

/* Keep the per-receptor synaptic currents of each neuron's last
 * non-refractory step (Nets[].Irep, code/EqsGnl2.cpp:196-200; printed by the
 * IsynSeparatedSum output, code/NetSimGnl2_info.cpp:559-573) in a device
 * buffer of 5 x n_neuron values; off by default.  Zeroed by fs_reset. */
int fs_record_currents(fs_handle *h, int32_t on);

/* Per-population spike counts accumulated since fs_reset (uint64[n_pop]). */
int fs_pop_spike_counts(fs_handle *h, uint64_t *out);

/* Launch schedule (autotuner output): block size of the neuron update,
 * lanes per row of the NMDA gather, block size of the delivery, steps per
 * captured graph, graph on/off, the one-launch-per-step fused path (used
 * only when every synapse has a delay of at least one step), and window:
 * steps per launch of the temporal-blocking path (0 = off; clamped to 1 +
 * the smallest synaptic delay in steps, and to 64).  The windowed path takes
 * precedence over the fused one.  fs_get_schedule reports the path in
 * effect: fused = 1 only while the fused path runs, window = the steps per
 * launch actually used (0 when off).  xcd_order (windowed path): 1 hands
 * consecutive blocks of neurons to one XCD (the chip's
 * workgroup dispatcher deals blocks to its 8 XCDs in turn), so the blocks
 * an XCD runs together share presynaptic neurons in its L2.  gather_chunk
 * (windowed path, one network copy): at most this many NMDA entries per
 * LDS-staged chunk of a block's gather (0: as many as 32 KiB of staged gates
 * hold; else >= 64), trading LDS per block for blocks per compute unit.
 * gather_split (windowed path, one network copy): lanes that share the
 * entries of one (row, step) pair of that gather (a power of two, at most
 * gather_lanes; 0 or 1: one lane per pair), for networks with long rows. */
typedef struct fs_schedule {
  int32_t update_block, gather_lanes, push_block, steps_per_graph, use_graph, fused;
  int32_t window;
  int32_t xcd_order;
  int32_t gather_chunk;
  int32_t gather_split;
} fs_schedule;
int fs_get_schedule(fs_handle *h, fs_schedule *s);
int fs_set_schedule(fs_handle *h, const fs_schedule *s);

/* Kernel timing with HIP events on the engine's stream over the last fs_run
 * (enabled by fs_enable_timing): average duration in ms of the first step of
 * each captured batch, split into the update (or fused step) kernel and the
 * delivery kernel, then the average time per step of the graph-replayed
 * steps (events around each graph launch divided by its steps; for the fused
 * path that is the per-launch time of its one kernel), and the number of
 * timed batches. */
int fs_enable_timing(fs_handle *h, int32_t on);
int fs_kernel_times(fs_handle *h, double *update_ms, double *push_ms, double *gather_ms,
                    int64_t *n_timed);

/* Algorithmic counters of the last fs_run: out[0] spikes, out[1] linear
 * receptor-synapses pushed, out[2] NMDA receptor-synapses of the spiking
 * neurons (out[1]+out[2] = synaptic events), out[3] NMDA terms gathered. */
int fs_run_counters(fs_handle *h, int64_t out[4]);

/* Shape of the network on the handle, for schedule caches (no counterpart in
 * the reference): out[0] neurons of one copy, [1] populations of one copy,
 * [2] copies (replicas), [3] linear receptor-synapses, [4] NMDA
 * receptor-synapses (of one copy), [5] smallest synaptic delay, [6] largest
 * linear delay, [7] largest NMDA delay (steps), [8] precision, [9] input
 * mode. */
int fs_network_stats(fs_handle *h, int64_t out[10]);

/* Name and architecture of the current HIP device, and its compute units
 * (no counterpart in the reference; schedule caches key on it). */
int fs_device_info(char *name, int32_t name_cap, int32_t *compute_units);

int fs_destroy(fs_handle *h);
const char *fs_last_error(void);

/* ---------------------------------------------------------------------
 * Drop-in host side: network.conf / network.pro.
 * ------------------------------------------------------------------- */
typedef struct fs_conf fs_conf;

typedef struct fs_cli_opts {
  float dt;               /* -dt (default 0.1) */
  int32_t threads;        /* -t  (only consumes draws, as in the reference) */
  int32_t repeats;        /* -rp */
  int32_t stp, std_, ltp; /* -STP / -STD / -LTP */
  int32_t udf_seed;       /* -udfsed given */
  uint32_t udf_seed_value;
  const char *out_mem, *out_spi, *out_fr;   /* -om / -os / -or, NULL = unset */
} fs_cli_opts;

/* Replaces Parser::CSta/parse/Comm/DefMacro/parse(pro) + NetSimGnl2::SetPp
 * (code/main.cpp:233-248). */
int fs_conf_load(const char *conf_path, const char *pro_path, const fs_cli_opts *opts,
                 fs_conf **out);
int fs_conf_free(fs_conf *c);
/* Engine description of the loaded network (pointers stay owned by c). */
int fs_conf_desc(fs_conf *c, fs_net_desc *out);
/* Sizes: neurons, populations, synapses, total synapses as the reference
 * reports them, events, outputs, repeats. */
int fs_conf_sizes(fs_conf *c, int64_t out[7]);
/* Seed of repeat r (code/NetSimGnl2_BulPp.cpp:83-127). */
int fs_conf_seed(fs_conf *c, int32_t repeat, uint32_t *seed);
/* External gate outcomes of the reference's mt19937 stream for steps
 * [0, n_steps) of one repeat, in the fs_load_ext_trains layout
 * (code/EqsGnl2.cpp:284-332 with the seeding of code/NetSimGnl2.cpp:313-330
 * and the ActCtr changes of DoEvt). */
int fs_conf_ext_trains(fs_conf *c, int32_t repeat, int64_t n_steps, uint32_t *bits);

/* ConfsInfo.log and network.log into dir ("" = current directory), as the
 * reference writes them before it simulates: Parser::ConfsInfo
 * (code/parser.cpp:1488-1625, called from code/main.cpp:247) and
 * NetSimGnl2::logDataWrite (code/NetSimGnl2_info.cpp:93-541, called from
 * RunInit, code/NetSimGnl2.cpp:154). */
int fs_conf_write_logs(fs_conf *c, const char *dir);

/* Writes the Spike / FiringRate output files of the protocol (FileOut,
 * code/NetSimGnl2_info.cpp:624-699) for a given spike raster, in the
 * fs_run offsets/ids layout, into directory dir (NULL: current directory). */
int fs_conf_write_outputs(fs_conf *c, int32_t repeat, int64_t n_steps, const int64_t *offsets,
                          const uint32_t *ids, const char *dir);

/* Reference internal layout (gates and pull-sum rows in execution order),
 * for CPU restatements of the reference's summation order. */
typedef struct fs_ref_layout {
  int64_t n_gate, n_row;
  const uint32_t *gate_pre, *gate_post0;
  const uint8_t *gate_rec, *gate_group;
  const float *gate_g;
  const uint32_t *row_post, *row_beg, *term;
  const uint8_t *row_rec;
  const fs_pop_conf *pop_conf;    /* [n_pop] */
  const int32_t *pop_rfped, *pop_spidy;  /* [n_pop] refractory / spike delay in steps */
  const uint32_t *neuron_pop;     /* [n_neuron] */
  int64_t n_evt;
  const float *evt_time_s, *evt_var1, *evt_var2;
  const int32_t *evt_type, *evt_pop, *evt_sub;
  uint32_t act_ctr;
  /* single-synapse gap junctions in SynCnd1 order (sorted by post) */
  int64_t n_gap;
  const uint32_t *gap_pre, *gap_post;
  const float *gap_g;
} fs_ref_layout;
int fs_conf_ref_layout(fs_conf *c, fs_ref_layout *out);

/* The drop-in command line (code/main.cpp:26-256): parses argv like the
 * reference, runs every repeat on the GPU and writes the Spike / FiringRate
 * / MemPot output files byte-compatibly. */
int fs_cli_main(int argc, const char *const *argv);

#ifdef __cplusplus
}
#endif
#endif /* FLYSIM_H */
