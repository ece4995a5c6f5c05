/*
 * bsim_oracle.h -- CPU oracle for the BSim simulation step (TEST INFRASTRUCTURE).
 *
 * This header and the C files beside it restate, in plain C, the algorithm of
 * BSim's single-threaded CPU simulator (STSim) for LIF_brian neurons, static
 * synapses and spike-array inputs.  It exists only to check the HIP product:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are its only
 * callers.  The product library never links, loads or calls it.
 *
 * Reference semantics followed (paths relative to the BSim repository):
 *   - LIF constants:       src/neuron/lif/LIFENeuron.cpp:68-112  (reset)
 *   - LIF step:            src/neuron/lif/LIFENeuron.cpp:114-160  (update)
 *   - input split E/I:     src/neuron/lif/LIFENeuron.cpp:52-61  (recv)
 *   - synapse delay/weight src/synapse/static/StaticSynapse.cpp:43-70
 *   - tau choice per type  src/base/CompositeNeuron.h:62-79
 *   - step order:          src/net/Network.cpp:308-330 (neurons in creation
 *                          order, then every synapse in creation order)
 *   - array input source:  src/neuron/array/ArrayNeuron.cpp:67-80
 *   - cycle count:         src/sim/SingleThreadSimulator.cpp:23 (round(T/dt))
 *
 * Two entry points per precision:
 *   bso_run_*      exact restatement: the synapses that deliver at a step are
 *                  applied in synapse-creation order (a sort per step), so it
 *                  holds for any network, whatever order it was built in.
 *   bso_run_csr_*  the event-driven CSR form of the same step used as the CPU
 *                  baseline: rows sorted by (pre, delay), targets partitioned
 *                  into ranges processed by OpenMP threads.  It is bit-identical
 *                  to bso_run_* whenever, for every target, creation order of
 *                  its incoming synapses is ascending in (pre, delay, position),
 *                  which bso_csr_order_ok() checks.
 */
#ifndef BSIM_ORACLE_H
#define BSIM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { BSO_LIF = 0, BSO_ARRAY = 1, BSO_BERNOULLI = 2, BSO_POISSON = 3 };

/* Number of REAL values per LIF neuron in the parameter block:
 * v_init v_rest v_reset cm tau_m tau_refrac tau_syn_E tau_syn_I v_thresh
 * i_offset, then the LIF_brian composite's own tau_syn_E tau_syn_I
 * (CompositeNeuron.h:43-46), used to pick a synapse's tau. */
#define BSO_LIF_NPARAM 12

#define BSO_DECLARE_NET(REAL, SUF)                                            \
typedef struct bso_net_##SUF {                                                \
    int32_t n;               /* neurons, in creation order                */ \
    const int8_t *kind;      /* BSO_LIF / ARRAY / BERNOULLI / POISSON       */ \
    const REAL *lif;         /* n * BSO_LIF_NPARAM (rows of non-LIF unused) */ \
    const int64_t *arr_off;  /* n + 1 offsets into arr_t (may be NULL)      */ \
    const int32_t *arr_t;    /* ascending fire cycles per array neuron      */ \
    const uint32_t *bern_p;  /* per neuron: fires iff hash < bern_p         */ \
    uint64_t seed;           /* Bernoulli source seed                       */ \
    int64_t s;               /* synapses, in creation order                 */ \
    const int32_t *pre;                                                       \
    const int32_t *post;                                                      \
    const REAL *w;                                                            \
    const REAL *d;           /* delay in seconds                            */ \
    const int8_t *stype;     /* 0 excitatory, 1 inhibitory (may be NULL)    */ \
    const REAL *tau;         /* explicit tau (0 -> composite); may be NULL  */ \
    const REAL *pois;        /* n * 3: rate, startTime, duration of each    */ \
                             /* Poisson source (other rows unused; may be   */ \
                             /* NULL without Poisson sources)               */ \
} bso_net_##SUF;                                                              \
                                                                              \
typedef struct bso_out_##SUF {                                                \
    /* raster: ascending fired ids per step, concatenated */                  \
    int32_t *raster_count;   /* [steps]            (may be NULL)            */ \
    int32_t *raster_ids;     /* [raster_cap]       (may be NULL)            */ \
    int64_t raster_cap;                                                       \
    int64_t raster_len;      /* written                                     */ \
    /* per-step traces of selected neurons (V after the step's update) */     \
    int32_t n_rec;                                                            \
    const int32_t *rec_ids;                                                   \
    REAL *rec_v;             /* [steps * n_rec]    (may be NULL)            */ \
    /* per neuron at the end */                                               \
    int32_t *fire_count;     /* [n]                (may be NULL)            */ \
    REAL *final_v;           /* [n]                (may be NULL)            */ \
    REAL *final_ie;          /* [n]                (may be NULL)            */ \
    REAL *final_ii;          /* [n]                (may be NULL)            */ \
    int32_t *final_refrac;   /* [n]                (may be NULL)            */ \
    int64_t events;          /* synaptic deliveries performed               */ \
    int32_t time_from;       /* cycle from which wall time is measured      */ \
    double wall_s;           /* seconds spent in cycles >= time_from        */ \
    int64_t events_from;     /* deliveries in cycles >= time_from           */ \
    /* GPU simulator logs (FireInfo "Y", "X"; may be NULL) */                 \
    REAL *input_sum;         /* [n] LIF input E + I summed over the cycles   */ \
                             /* it neither was refractory nor fired         */ \
                             /* (src/neuron/lif/lif.cu:116)                  */ \
    int32_t *slot_count;     /* [n] delay slots of the neuron's spikes that  */ \
                             /* the deliveries visited                       */ \
                             /* (src/synapse/static/static.cu:42, :82, :106) */ \
    /* per-step state traces of the selected neurons (may be NULL): i_E and  */ \
    /* i_I after the update, and the E / I input the update consumed -- the  */ \
    /* values LIFENeuron::update pushes into SimInfo (LIFENeuron.cpp:144-148) */ \
    REAL *rec_ie;            /* [steps * n_rec]                             */ \
    REAL *rec_ii;                                                             \
    REAL *rec_in_e;                                                           \
    REAL *rec_in_i;                                                           \
} bso_out_##SUF;                                                              \
                                                                              \
int bso_lif_constants_##SUF(const bso_net_##SUF *net, REAL dt, REAL *out9,   \
                            int32_t *refrac_time);                            \
int bso_synapse_reset_##SUF(const bso_net_##SUF *net, REAL dt, REAL *w_out,  \
                            int32_t *delay_steps);                            \
int bso_run_##SUF(const bso_net_##SUF *net, REAL dt, int32_t steps,          \
                  bso_out_##SUF *out);                                        \
int bso_run_csr_##SUF(const bso_net_##SUF *net, REAL dt, int32_t steps,      \
                      int32_t tile, int32_t threads, bso_out_##SUF *out);

BSO_DECLARE_NET(float, f32)
BSO_DECLARE_NET(double, f64)

/* 1 if every target's incoming synapses appear in creation order when sorted
 * by (pre, delay_steps, creation index); 0 otherwise. */
int bso_csr_order_ok(int32_t n, int64_t s, const int32_t *pre,
                     const int32_t *post, const int32_t *delay_steps);

/* Deterministic Bernoulli source hash (splitmix64 finalizer over seed, id and
 * cycle); a source fires at cycle t iff bso_hash(seed, id, t) < p. */
uint32_t bso_hash(uint64_t seed, uint32_t id, uint32_t t);

/* Poisson sources (src/neuron/poisson/PoissonNeuron.cpp:36-63, drawn from
 * seeded uniforms instead of a clock-seeded engine): X ~ Poisson(rate) is the
 * smallest j with u < floor(P(X <= j) * 2^32), u = bso_hash(seed, id, t).
 * bso_poisson_table returns a malloc'd table {J, thresholds[0 .. J-1]}. */
uint32_t *bso_poisson_table(double rate);
uint32_t bso_poisson_pick(const uint32_t *tab, uint32_t u);

#ifdef __cplusplus
}
#endif
#endif
