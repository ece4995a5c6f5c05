/*
 * bsim_amd.h -- C ABI of the MI355X simulation-step library (libbsim_amd.so).
 *
 * Plain C types only.  Every function returns 0 / a non-negative count on
 * success and a negative BSIM_E* code on failure; bsim_last_error() then
 * describes the failure (thread-local).  No function exits the process.
 *
 * Two groups of entry points:
 *
 * 1. Network construction on the host -- what the reference's Network class
 *    stores when a program calls createPopulation()/addNeuron()/connect():
 *      bsim_net_add_lif       <- Network::createPopulation<LIF_brian>,
 *                                Population::addNeuron (src/net/Network.h:120-144,
 *                                src/base/Population.h:158-168)
 *      bsim_net_add_array     <- createPopulation<Array_spikes>
 *                                (src/neuron/array/ArrayNeuron.cpp:70-77)
 *      bsim_net_add_poisson   <- createPopulation<Poisson_spikes>
 *                                (src/neuron/poisson/PoissonNeuron.cpp:13-63), with
 *                                seeded draws instead of the clock-seeded engine
 *      bsim_net_add_bernoulli <- (no counterpart: the reference's Poisson_spikes
 *                                is seeded from the clock, PoissonNeuron.cpp:25-26)
 *      bsim_net_connect       <- Network::connect(NeuronBase*, NeuronBase*, w, d,
 *                                type, tau) applied to arrays of synapses
 *                                (src/net/Network.cpp:118-150)
 *
 * 2. The device simulation -- the reference's GPU plug-in table and runtime:
 *      bsim_sim_create  <- Network::reset + Network::buildNetwork +
 *                          copyNetworkToGPU + alloc_buffers + getBlockSize
 *                          (src/net/Network.cpp:286-306, src/net/BuildNetwork.cpp:36-205,
 *                          src/gpu_utils/gpu_net_copy.cu:8-109, src/gpu_utils/GBuffers.cu:9-56,
 *                          src/gpu_utils/gpu_func.cu:90-127; per type cudaAllocLIFE /
 *                          cudaAllocStatic, src/utils/macros.h:11-32)
 *      bsim_sim_step    <- one cycle of SingleGPUSimulator::run's loop:
 *                          cudaUpdateLIFE + cudaUpdateStatic + update_time
 *                          (src/sim/SingleGPUSimulator.cu:146-245, src/gpu_utils/gpu_func.cu:44-51,
 *                          67-73, src/gpu_utils/runtime.cu:81-92), n cycles per call
 *      bsim_sim_read_*  <- cudaFetchLIFE / fetchNetworkFromGPU, the GSim.log raster and
 *                          g_v.data trace, gFireCount ("count" of FireInfo)
 *                          (src/neuron/lif/GLIFENeurons.cu:29-40, src/gpu_utils/gpu_net_copy.cu:
 *                          111-132, src/sim/SingleGPUSimulator.cu:176-296)
 *      bsim_sim_destroy <- free_buffers + freeGPUNetwork (GBuffers.cu:64-78,
 *                          gpu_net_copy.cu:134-191)
 */
#ifndef BSIM_AMD_H
#define BSIM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSIM_ABI_VERSION 1

enum {
    BSIM_OK = 0,
    BSIM_E_ARG = -1,        /* invalid argument */
    BSIM_E_NETWORK = -2,    /* network not supported by the GPU step (message says why) */
    BSIM_E_NOMEM = -3,      /* host or device allocation failed */
    BSIM_E_HIP = -4,        /* HIP runtime error */
    BSIM_E_STATE = -5,      /* call not valid in the object's current state */
    BSIM_E_CAPACITY = -6    /* caller buffer or recording window too small */
};

typedef struct bsim_network bsim_network;
typedef struct bsim_sim bsim_sim;

int32_t bsim_abi_version(void);
const char *bsim_last_error(void);

/* ---- network construction (host) ---------------------------------------- */

bsim_network *bsim_net_create(void);
void bsim_net_destroy(bsim_network *net);

/* Appends n LIF_brian neurons; params is n rows of 12 floats:
 * v_init v_rest v_reset cm tau_m tau_refrac tau_syn_E tau_syn_I v_thresh
 * i_offset, then the composite's own tau_syn_E tau_syn_I.  Returns the global
 * id of the first neuron added (ids are creation order). */
int32_t bsim_net_add_lif(bsim_network *net, int32_t n, const float *params);

/* Appends n array sources; neuron i fires at cycles
 * cycles[offsets[i] .. offsets[i+1]) (ascending, offsets[0] == 0).
 * taus (may be NULL) holds the composite tau_syn_E/I per neuron. */
int32_t bsim_net_add_array(bsim_network *net, int32_t n, const int64_t *offsets,
                           const int32_t *cycles, const float *taus);

/* Appends n seeded Bernoulli sources firing at cycle t iff
 * hash(seed, id, t) < p[i], id = the neuron's global id. */
int32_t bsim_net_add_bernoulli(bsim_network *net, int32_t n, const uint32_t *p, uint64_t seed);

/* The same with explicit keys: source i fires at cycle t iff
 * hash(seed, first_key + i, t) < p[i].  A network appended to a batch with
 * its own seed and keys 0.. draws exactly what it draws alone.  (No
 * counterpart in the reference, whose Poisson sources are clock-seeded.) */
int32_t bsim_net_add_bernoulli_keyed(bsim_network *net, int32_t n, const uint32_t *p, uint64_t seed,
                                     uint32_t first_key);

/* Appends n Poisson sources with the reference's CPU semantics
 * (PoissonNeuron.cpp:36-63): source i first fires at cycle
 * int(start[i] / dt) + X, then 1 + X cycles after each spike, while the
 * firing cycle is below int((start[i] + duration[i]) / dt); every X is an
 * independent Poisson(rate[i]) draw (rate = the mean number of cycles
 * between spikes minus one, as the reference's poisson_distribution takes
 * it; 0..1e6).  Draws are seeded: X = the inverse CDF of Poisson(rate[i]) at
 * hash(seed, id, t) (t = the spike's cycle; 0xFFFFFFFF for the first draw),
 * id = the neuron's global id.  taus (may be NULL) holds the composite
 * tau_syn_E/I per neuron.  The reference's refract argument has no effect on
 * its CPU path and is not taken. */
int32_t bsim_net_add_poisson(bsim_network *net, int32_t n, const float *rate, const float *start,
                             const float *duration, const float *taus, uint64_t seed);

/* Appends count synapses in the given order (= creation order).  d is the
 * delay in seconds; stype (NULL: all excitatory) and tau (NULL: all 0) as in
 * Network::connect.  Returns the number of synapses now in the network. */
int64_t bsim_net_connect(bsim_network *net, int64_t count, const int32_t *pre,
                         const int32_t *post, const float *w, const float *d,
                         const int8_t *stype, const float *tau);

int32_t bsim_net_num_neurons(const bsim_network *net);
/* Kind of every neuron (BSIM_KIND_*: 0 LIF_brian, 1 Array_spikes, 2 Bernoulli,
 * 3 Poisson_spikes) in creation order; returns the count.  The reference
 * groups neurons by type (Network::buildNetwork, src/net/BuildNetwork.cpp:
 * 36-94); its -DLOG_DATA traces cover the LIF type. */
int32_t bsim_net_kinds(const bsim_network *net, int8_t *kind);
int64_t bsim_net_num_synapses(const bsim_network *net);

/* ---- device simulation ---------------------------------------------------- */

typedef struct bsim_sim_config {
    int32_t device;        /* HIP device ordinal */
    int32_t tile;          /* neurons per tile (multiple of 192, at most 1536); 0 = automatic:
                              192 for delays up to 8 steps, 384 up to 24, else 768 */
    int32_t rec_steps;     /* cycles of raster/trace kept on the device (0 = none) */
    int32_t n_rec;         /* neurons whose V is traced (ids may repeat: one column each) */
    const int32_t *rec_ids;
} bsim_sim_config;

/* Resets the network (LIF constants, synapse delays), builds the device
 * layout, uploads it and allocates the step buffers.  The network may be
 * destroyed afterwards. */
bsim_sim *bsim_sim_create(const bsim_network *net, float dt, const bsim_sim_config *cfg);

/* Advances n cycles (asynchronous on the simulation's stream). */
int32_t bsim_sim_step(bsim_sim *sim, int32_t n);

/* Performs phase (ii) of the last cycle, which the next cycle would consume;
 * run(T) does this once at the end so that every delivery of the T cycles is
 * executed, as in the reference.  The neuron state is not changed. */
int32_t bsim_sim_finish(bsim_sim *sim);

int32_t bsim_sim_sync(bsim_sim *sim);

/* Cycles simulated so far. */
int32_t bsim_sim_cycle(const bsim_sim *sim);

/* Raster of the recorded cycles as ascending fired ids per cycle: counts has
 * one entry per recorded cycle (min(cycle, rec_steps)); ids receives the
 * concatenation.  Returns the number of ids, or BSIM_E_CAPACITY (with the
 * needed size in *needed) if cap is too small. */
int64_t bsim_sim_read_raster(bsim_sim *sim, int32_t *counts, int32_t *ids, int64_t cap,
                             int64_t *needed);

/* V trace of the recorded neurons: [recorded cycles][n_rec]. */
int32_t bsim_sim_read_trace(bsim_sim *sim, float *v);

/* State traces of the traced neurons, as the reference's -DLOG_DATA run
 * writes them per cycle (src/sim/SingleGPUSimulator.cu:150-216): i_E and i_I
 * after the update (g_ie.data, g_ii.data) and the E / I synaptic input the
 * update consumed (g_input_e.data, g_input_i.data).  Enable before the first
 * cycle of a simulation created with traced neurons; read [recorded
 * cycles][n_rec] per array (NULL arrays skipped). */
int32_t bsim_sim_enable_state_trace(bsim_sim *sim);
int32_t bsim_sim_read_state_trace(bsim_sim *sim, float *ie, float *ii, float *in_e, float *in_i);

/* Neuron state (any pointer may be NULL): V, i_E, i_I, refractory counter,
 * spike count. */
int32_t bsim_sim_read_state(bsim_sim *sim, float *v, float *ie, float *ii, int32_t *refrac,
                            int32_t *fire_count);

/* Advances n cycles one launch at a time with a pair of HIP events around
 * each of the cycle's two kernels on the simulation's stream (the step
 * kernel, then the active-row list kernel), synchronises, and returns their
 * mean durations in microseconds.  For measurement; bsim_sim_step is the
 * production path (graph-replayed, no events). */
int32_t bsim_sim_time_steps(bsim_sim *sim, int32_t n, double *step_us, double *rows_us);

/* Turns on the per-neuron input logs of the reference GPU simulator's
 * FireInfo (src/sim/SingleGPUSimulator.cu:284-296): "Y" -- a LIF neuron's
 * E + I input summed over the cycles it neither was refractory nor fired
 * (gXInput, src/neuron/lif/lif.cu:116) -- and "X" -- the delay slots of the
 * neuron's spikes visited by the deliveries (gLayerInput,
 * src/synapse/static/static.cu:42, :82, :106).  Counting starts with the next
 * cycle; call before the first bsim_sim_step to cover the whole run. */
int32_t bsim_sim_enable_input_log(bsim_sim *sim);
/* Copies the logs (n values each; NULL arrays skipped). */
int32_t bsim_sim_read_input_log(bsim_sim *sim, float *in_sum, int32_t *slot_count);

/* Synaptic deliveries performed so far (sum over tiles). */
int64_t bsim_sim_events(bsim_sim *sim);

typedef struct bsim_sim_info {
    int32_t n, tile, ntiles, dmax;
    int64_t s;
    int64_t device_bytes;
    int64_t n_src_entries;
    int64_t off_entries;
} bsim_sim_info;

int32_t bsim_sim_get_info(const bsim_sim *sim, bsim_sim_info *info);

/* The simulation's HIP stream (hipStream_t) for callers that time or order
 * work around the step. */
void *bsim_sim_stream(bsim_sim *sim);

/* Tile schedule (no counterpart in the reference, whose per-type launches
 * take the neurons and synapses in id order with a fixed grid,
 * SingleGPUSimulator.cu:100, 165-171): the order in which the
 * step's workgroups take the tiles.  Every rank_every cycles the row-list
 * launch ranks the tiles by their work in the step launch before it,
 * heaviest first (the default, 2); 0 keeps the tiles in their last order.
 * Results do not depend on it.  Takes effect from the next step. */
int32_t bsim_sim_set_schedule(bsim_sim *sim, int32_t rank_every);

void bsim_sim_destroy(bsim_sim *sim);

/* ---- workload generation and export (host) -------------------------------- */

/* Parameters of the standard feed-forward workload (see bsim_net_gen_standard). */
typedef struct bsim_gen_params {
    int32_t npop, n, fanout;      /* populations, neurons each, out-synapses per neuron */
    float w0, inh_frac;           /* exc. w ~ U[0, 0.7 w0), inh. w ~ U[-0.5 w0, 0) */
    int32_t dmax;                 /* delays U{1..dmax} steps */
    float remote_frac;            /* share of populations with a remote projection */
    float drive_rate_hz;          /* Bernoulli drive of population 0 (0 = none) */
    int32_t n_drive;
    float drive_w;
    int32_t drive_fanout;
    uint64_t seed;
    float tau_refrac, ioff_lo, ioff_hi, tau_syn_E, tau_syn_I, v_thresh;
    float dt;
} bsim_gen_params;

/* Builds the standard feed-forward chain into an empty network: optional
 * drive sources, then npop LIF_brian populations (tau_m 20 ms, R 1,
 * v_rest = v_reset = 0, i_offset ~ U[ioff_lo, ioff_hi)); population i
 * projects to i+1 and, for a remote_frac share, to one remote population,
 * splitting its fan-out; targets by stratified sampling; everything drawn
 * from a counter-based hash of the seed.  The reference's benchmark program
 * builds the same topology from all-to-all projections
 * (test/gpu/standard_test.cpp:92-139).  Returns the number of synapses. */
int64_t bsim_net_gen_standard(bsim_network *net, const bsim_gen_params *p);

/* One network of a parameter sweep: a feed-forward chain of npop LIF_brian
 * populations (tau_m 20 ms, R 1, v_rest = v_reset = 0, i_offset ~ U[ioff_lo,
 * ioff_hi)); population i projects to i+1 with about fan_in synapses per
 * target (per pre round(fan_in * n_dst / n_src) distinct targets by
 * stratified sampling); exc. w ~ U[0, 0.7 w0), inh. (share inh_frac)
 * w ~ U[-0.5 w0, 0); delays U{1..dmax} steps; n_drive Bernoulli sources at
 * drive_rate_hz feed population 0 with about drive_fan_in synapses per target
 * of weight drive_w. */
typedef struct bsim_chain_params {
    int32_t npop;                 /* 1..8 */
    int32_t sizes[8];
    int32_t fan_in;
    float w0, inh_frac;
    int32_t dmax;
    float drive_rate_hz;
    int32_t n_drive, drive_fan_in;
    float drive_w;
    uint64_t seed;
    float tau_refrac, ioff_lo, ioff_hi, tau_syn_E, tau_syn_I, v_thresh;
    float dt;
} bsim_chain_params;

/* One simulation of count sweep networks built without holding them all on
 * the host: the networks are generated, laid out and uploaded in parts of at
 * most part_synapses synapses (at least one network each), the parts joined
 * on the device into one layout -- each part padded with inert neurons to a
 * whole number of tiles -- and stepped as one network by bsim_sim_step.  No
 * counterpart in the reference (its single-GPU simulator runs one network);
 * it serves parameter sweeps whose networks do not fit one host build.
 * first_ids[c] receives network c's first neuron id (the padding shifts
 * them from what bsim_net_gen_chains would give).  cfg as for
 * bsim_sim_create; recorded ids refer to the joined network. */
bsim_sim *bsim_sim_create_chains(const bsim_chain_params *p, int32_t count, float dt,
                                 const bsim_sim_config *cfg, int64_t part_synapses,
                                 int32_t *first_ids);

/* Appends one sweep network to net (which may hold others: a batch) and
 * returns the id of its first neuron.  Its draws, Bernoulli keys included,
 * depend only on p, so the network behaves in a batch exactly as alone. */
int32_t bsim_net_gen_chain(bsim_network *net, const bsim_chain_params *p);

/* Appends count sweep networks, the same network as count calls of
 * bsim_net_gen_chain in that order, with the host arrays sized once and all
 * chains generated in parallel; first_ids[c] (if not NULL) receives chain c's
 * first neuron id.  Returns the id of the first chain's first neuron. */
int32_t bsim_net_gen_chains(bsim_network *net, const bsim_chain_params *p, int32_t count,
                            int32_t *first_ids);

/* Copies out the network as stored (creation order); NULL arrays skipped.
 * lif receives n * 12 floats. */
int32_t bsim_net_export_neurons(const bsim_network *net, int8_t *kind, float *lif,
                                uint32_t *bern_p, uint64_t *seed);
int64_t bsim_net_export_synapses(const bsim_network *net, int64_t start, int64_t count,
                                 int32_t *pre, int32_t *post, float *w, float *d, int8_t *stype);

/* ---- host-side inspection (no device needed) ------------------------------ */

/* LIF constants as LIFENeuron::reset computes them: per neuron Cm CE CI C_E
 * C_I v_tmp v_thresh v_reset v_init (9 floats) and refrac_time. */
int32_t bsim_net_lif_constants(const bsim_network *net, float dt, float *out9,
                               int32_t *refrac_time);

/* Device layout computed on the host, for inspection: the local targets /
 * weights as stored on the device, and per tile (ntiles + 1 entries) the
 * first synapse of the tile's group -- a run of consecutive tiles with the
 * same source ranges whose synapses are stored in (range, pre, delay, tile,
 * part, creation) order.  Pass NULL arrays to query sizes through *ntiles /
 * *s. */
int32_t bsim_net_layout(const bsim_network *net, float dt, int32_t tile, int32_t *ntiles,
                        int64_t *s, int32_t *dmax, int64_t *tile_syn, uint16_t *syn_tgt,
                        float *syn_w);

#ifdef __cplusplus
}
#endif
#endif
