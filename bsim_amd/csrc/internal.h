/*
 * internal.h -- definitions shared by the host translation units of the
 * library (not part of the C ABI).
 */
#ifndef BSIM_INTERNAL_H
#define BSIM_INTERNAL_H

#include <functional>
#include <string>

#include "network.h"

struct bsim_network {
    bsim::HostNet net;
};

/* message of the last failure on this thread (bsim_last_error) */
extern thread_local std::string g_bsim_err;

struct bsim_chain_params;

namespace bsim {

/* Neurons, synapses and largest delay (steps) of count sweep networks;
 * -1 when a parameter set is invalid. */
int chains_size(const bsim_chain_params *p, int32_t count, int64_t *neurons, int64_t *synapses,
                int32_t *dmax);

/* The layout of count sweep networks as one simulation, built part by part:
 * chains are generated into a host network of at most part_synapses
 * synapses (at least one chain), padded with inert neurons to whole tiles,
 * laid out with the given tile and delay ring depth, handed to upload
 * (part, first synapse index of the part in the whole layout) -- which takes
 * its synapses -- and appended to out without them.  first_ids[c] receives
 * chain c's first neuron id.  0, or -1 with err set (empty when upload
 * failed and set the error itself). */
int build_chains_layout(const bsim_chain_params *p, int32_t count, float dt, int32_t tile,
                        int32_t dmax, int64_t part_synapses, Layout &out, int32_t *first_ids,
                        std::string &err,
                        const std::function<int(const Layout &, int64_t)> &upload);

}  // namespace bsim

#endif
