/*
 * network.h -- host-side network store and the builder of the device layout.
 *
 * HostNet holds a network as the user built it through BSim's Network API
 * (src/net/Network.h:120-209, src/net/Network.cpp:118-150): neurons in
 * creation order with their own parameters, synapses in connect() order.
 * Unlike the reference (one heap object per synapse, StaticSynapse.h:29-36)
 * synapses are a compact structure of arrays (17 bytes each).
 *
 * build_layout() performs what Network::reset + Network::buildNetwork +
 * copyNetworkToGPU's host half do in the reference (Network.cpp:286-306,
 * BuildNetwork.cpp:36-205): per-neuron LIF constants, per-synapse delay steps
 * and weight, and the tile-major (pre, delay) CSR described in layout.h, in
 * O(S log S) per tile instead of the reference's O(S * D) scan.
 */
#ifndef BSIM_NETWORK_H
#define BSIM_NETWORK_H

#include <stdint.h>

#include <string>
#include <vector>

#include "layout.h"

namespace bsim {

constexpr int kLifParams = 12;

struct HostNet {
    std::vector<int8_t> kind;
    std::vector<float> lif;            /* n * kLifParams */
    std::vector<int64_t> arr_off{0};   /* n + 1, into arr_t */
    std::vector<int32_t> arr_t;
    std::vector<uint32_t> bern_p;
    std::vector<uint64_t> bern_seed;   /* per neuron: seed and key of its Bernoulli */
    std::vector<uint32_t> bern_key;    /* draws, hash(seed, key, t) (0 for others) */
    uint64_t seed = 0;                 /* seed of the last add_bernoulli (export) */
    struct PoissonSource {             /* PoissonNeuron(rate, refract, startTime, duration) */
        int32_t id;
        float rate, start, end;        /* end = startTime + duration, as the reference stores it */
    };
    std::vector<PoissonSource> pois;   /* Poisson sources, ascending ids */
    std::vector<int32_t> pre, post;
    std::vector<float> w, d;
    std::vector<int8_t> stype;
    std::vector<float> tau;            /* empty: every synapse has tau_in = 0 */

    int32_t n() const { return (int32_t)kind.size(); }
    int64_t s() const { return (int64_t)pre.size(); }
};

struct Layout {
    int32_t n = 0, tile = 0, ntiles = 0, dmax = 1;
    int64_t s = 0;
    std::vector<uint32_t> ncls;              /* kind << 30 | LIF parameter set */
    std::vector<bsim_lifc> lifc;             /* the distinct LIF parameter sets */
    std::vector<float> v_tmp, v_init;
    std::vector<int32_t> arr_cur0, arr_end, arr_t;
    std::vector<uint32_t> bern_p;
    std::vector<uint64_t> bern_seed;
    std::vector<uint32_t> bern_key;
    std::vector<uint32_t> pois_tab;          /* inverse-CDF tables of the Poisson sources' rates */
    std::vector<int64_t> tile_syn;           /* [ntiles + 1] tile-major synapse counts (prefix) */
    std::vector<uint16_t> syn_tgt;
    std::vector<float> syn_w;
    std::vector<int32_t> tile_rng;
    std::vector<bsim_range> rng;
    std::vector<uint32_t> off_tab;
    std::vector<int32_t> src_lo, src_hi;     /* distinct source ranges */
    std::vector<int64_t> src_list_base;      /* list capacity (hi - lo) * D each */
    int64_t src_list_len = 0;
    int32_t ngroups = 0, nslots = 0;         /* target groups; step workgroups */
    std::vector<int64_t> tile_base;          /* per tile: first synapse of its group */
    std::vector<uint32_t> tile_nsyn;         /* per tile: synapses of its group */
    std::vector<uint8_t> tile_res;           /* per tile: XCD residue of its group */
    std::vector<int64_t> chunk_tiles_off;    /* per chunk: the tiles reading it */
    std::vector<int32_t> chunk_tiles;
    std::vector<int32_t> chunk_rng;          /* range entry of each (chunk, tile) */
    std::vector<uint32_t> chunk_wrow;
};

/* LIFENeuron::reset (src/neuron/lif/LIFENeuron.cpp:68-112) for one neuron,
 * in float, in the reference's operation order.  out: Cm CE CI C_E C_I v_tmp
 * v_thresh v_reset v_init. */
void lif_constants(const float *p, float dt, float out[9], int32_t *refrac_time);

/* StaticSynapse::reset (src/synapse/static/StaticSynapse.cpp:43-54) with the
 * tau picked as CompositeNeuron::createSynapse does (CompositeNeuron.h:62-79).
 * Returns false where the reference asserts (non-integral delay, zero tau). */
bool synapse_reset(float w, float delay, float tau, float dt, float *w_out, int32_t *ds);

/* Chooses the tile size for n neurons on a device with `cus` compute units. */
int32_t default_tile(int32_t dmax);
/* Largest round(delay / dt) over the synapses (at least 1). */
int32_t max_delay_steps(const HostNet &net, float dt);

/* Returns 0 or a negative code with a message in err.  The delay ring is at
 * least dmax_min steps deep (a part of a larger layout uses the whole
 * network's depth). */
int build_layout(const HostNet &net, float dt, int32_t tile, Layout &out, std::string &err,
                 int32_t dmax_min = 0);

/* Appends layout B -- built, with the same tile and delay ring depth, for a
 * network placed after A's neurons, with no synapse between the two -- to A:
 * B's neuron ids, tiles, synapse offsets, source ranges, row-table offsets,
 * chunks and tables are shifted past A's, which must end on a tile
 * boundary.  B is emptied.  Returns 0 or a negative code with err set. */
int append_layout(Layout &A, Layout &B, std::string &err);

/* Appends the inverse-CDF table of Poisson(rate) to tab (format in layout.h). */
void poisson_table(float rate, std::vector<uint32_t> &tab);
/* X ~ Poisson from the uniform u with a table made by poisson_table. */
uint32_t poisson_pick(const uint32_t *tab, uint32_t u);
/* The sources' counter-based hash: splitmix64's finalizer over
 * seed ^ (key << 32 | t), upper 32 bits. */
uint32_t source_hash(uint64_t seed, uint32_t key, uint32_t t);

}  // namespace bsim

#endif
