/*
 * capi.cpp -- the C ABI declared in include/bsim_amd.h: host network store,
 * device upload, the per-cycle launch sequence and read-back.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/bsim_amd.h"
#include "layout.h"
#include "internal.h"
#include "network.h"

extern "C" hipError_t bsim_launch_rows(const bsim_dev *d, hipStream_t stream);
extern "C" hipError_t bsim_launch_step(const bsim_dev *d, int32_t t, int32_t flags,
                                       hipStream_t stream);

thread_local std::string g_bsim_err;

static int fail(int code, const std::string &msg)
{
    g_bsim_err = msg;
    return code;
}

struct bsim_sim {
    int device = 0;
    hipStream_t stream = nullptr;
    bsim_dev d{};
    std::vector<void *> allocs;
    int64_t device_bytes = 0;
    int32_t cycle = 0;
    int32_t finished_at = -1;
    int64_t *ev_fin = nullptr;   /* deliveries of the deliver-only launch */
    hipGraph_t graph = nullptr;  /* kGraphCycles clocked launches */
    hipGraphExec_t gexec = nullptr;
    int32_t rec_steps = 0, n_rec = 0;
    int64_t n_src = 0, n_off = 0, s = 0;
    std::vector<int32_t> rec_col;  /* trace column of each requested id (ids may repeat) */
    float *log_in = nullptr;       /* input logs (bsim_sim_enable_input_log) */
    int32_t *log_slot = nullptr;
};

extern "C" {

int32_t bsim_abi_version(void) { return BSIM_ABI_VERSION; }

const char *bsim_last_error(void) { return g_bsim_err.c_str(); }

/* ---- network construction ------------------------------------------------ */

bsim_network *bsim_net_create(void)
{
    try {
        return new bsim_network();
    } catch (...) {
        g_bsim_err = "out of host memory";
        return nullptr;
    }
}

void bsim_net_destroy(bsim_network *net) { delete net; }

int32_t bsim_net_add_lif(bsim_network *net, int32_t n, const float *params)
{
    if (!net || n < 0 || (n > 0 && !params)) return fail(BSIM_E_ARG, "bsim_net_add_lif: bad argument");
    auto &h = net->net;
    if ((int64_t)h.n() + n > 0x7FFFFFFF) return fail(BSIM_E_ARG, "too many neurons");
    const int32_t first = h.n();
    try {
        h.kind.insert(h.kind.end(), (size_t)n, (int8_t)BSIM_KIND_LIF);
        h.lif.insert(h.lif.end(), params, params + (size_t)n * bsim::kLifParams);
        h.arr_off.insert(h.arr_off.end(), (size_t)n, h.arr_off.back());
        h.bern_p.insert(h.bern_p.end(), (size_t)n, 0u);
        h.bern_seed.insert(h.bern_seed.end(), (size_t)n, 0ull);
        h.bern_key.insert(h.bern_key.end(), (size_t)n, 0u);
    } catch (...) {
        return fail(BSIM_E_NOMEM, "out of host memory");
    }
    return first;
}

int32_t bsim_net_add_array(bsim_network *net, int32_t n, const int64_t *offsets,
                           const int32_t *cycles, const float *taus)
{
    if (!net || n < 0 || (n > 0 && !offsets) || (offsets && offsets[0] != 0))
        return fail(BSIM_E_ARG, "bsim_net_add_array: bad argument");
    auto &h = net->net;
    const int32_t first = h.n();
    try {
        for (int32_t i = 0; i < n; i++) {
            if (offsets[i + 1] < offsets[i]) return fail(BSIM_E_ARG, "offsets must not decrease");
            for (int64_t j = offsets[i] + 1; j < offsets[i + 1]; j++)
                if (cycles[j] < cycles[j - 1]) return fail(BSIM_E_ARG, "spike cycles must ascend");
        }
        h.kind.insert(h.kind.end(), (size_t)n, (int8_t)BSIM_KIND_ARRAY);
        for (int32_t i = 0; i < n; i++) {
            float row[bsim::kLifParams] = {0};
            if (taus) {
                row[10] = taus[2 * i];
                row[11] = taus[2 * i + 1];
            } else {
                row[10] = 1e-3f;   // CompositeNeuron's default taus (CompositeNeuron.h:21)
                row[11] = 1e-3f;
            }
            h.lif.insert(h.lif.end(), row, row + bsim::kLifParams);
            const int64_t base = h.arr_off.back();
            h.arr_t.insert(h.arr_t.end(), cycles + offsets[i], cycles + offsets[i + 1]);
            h.arr_off.push_back(base + offsets[i + 1] - offsets[i]);
        }
        h.bern_p.insert(h.bern_p.end(), (size_t)n, 0u);
        h.bern_seed.insert(h.bern_seed.end(), (size_t)n, 0ull);
        h.bern_key.insert(h.bern_key.end(), (size_t)n, 0u);
    } catch (...) {
        return fail(BSIM_E_NOMEM, "out of host memory");
    }
    return first;
}

int32_t bsim_net_add_bernoulli(bsim_network *net, int32_t n, const uint32_t *p, uint64_t seed)
{
    if (!net) return fail(BSIM_E_ARG, "bsim_net_add_bernoulli: bad argument");
    return bsim_net_add_bernoulli_keyed(net, n, p, seed, (uint32_t)net->net.n());
}

int32_t bsim_net_add_bernoulli_keyed(bsim_network *net, int32_t n, const uint32_t *p, uint64_t seed,
                                     uint32_t first_key)
{
    if (!net || n < 0 || (n > 0 && !p)) return fail(BSIM_E_ARG, "bsim_net_add_bernoulli: bad argument");
    auto &h = net->net;
    const int32_t first = h.n();
    try {
        h.kind.insert(h.kind.end(), (size_t)n, (int8_t)BSIM_KIND_BERNOULLI);
        for (int32_t i = 0; i < n; i++) {
            float row[bsim::kLifParams] = {0};
            row[10] = 1e-3f;
            row[11] = 1e-3f;
            h.lif.insert(h.lif.end(), row, row + bsim::kLifParams);
        }
        h.arr_off.insert(h.arr_off.end(), (size_t)n, h.arr_off.back());
        h.bern_p.insert(h.bern_p.end(), p, p + n);
        h.bern_seed.insert(h.bern_seed.end(), (size_t)n, seed);
        for (int32_t i = 0; i < n; i++) h.bern_key.push_back(first_key + (uint32_t)i);
        h.seed = seed;
    } catch (...) {
        return fail(BSIM_E_NOMEM, "out of host memory");
    }
    return first;
}

int32_t bsim_net_add_poisson(bsim_network *net, int32_t n, const float *rate, const float *start,
                             const float *duration, const float *taus, uint64_t seed)
{
    if (!net || n < 0 || (n > 0 && (!rate || !start || !duration)))
        return fail(BSIM_E_ARG, "bsim_net_add_poisson: bad argument");
    auto &h = net->net;
    if ((int64_t)h.n() + n > 0x7FFFFFFF) return fail(BSIM_E_ARG, "too many neurons");
    for (int32_t i = 0; i < n; i++) {
        if (!(rate[i] >= 0.0f && rate[i] <= BSIM_POISSON_MAX_RATE))
            return fail(BSIM_E_ARG, "Poisson rate (mean cycles between spikes, minus one) must be in [0, 1e6]");
        if (!(start[i] >= 0.0f && duration[i] >= 0.0f && std::isfinite(start[i] + duration[i])))
            return fail(BSIM_E_ARG, "Poisson start time and duration must be finite and >= 0");
    }
    const int32_t first = h.n();
    try {
        h.kind.insert(h.kind.end(), (size_t)n, (int8_t)BSIM_KIND_POISSON);
        for (int32_t i = 0; i < n; i++) {
            float row[bsim::kLifParams] = {0};
            row[10] = taus ? taus[2 * i] : 1e-3f;       // CompositeNeuron's default taus
            row[11] = taus ? taus[2 * i + 1] : 1e-3f;   // (CompositeNeuron.h:21)
            h.lif.insert(h.lif.end(), row, row + bsim::kLifParams);
            h.pois.push_back({first + i, rate[i], start[i], start[i] + duration[i]});
        }
        h.arr_off.insert(h.arr_off.end(), (size_t)n, h.arr_off.back());
        h.bern_p.insert(h.bern_p.end(), (size_t)n, 0u);
        h.bern_seed.insert(h.bern_seed.end(), (size_t)n, seed);
        for (int32_t i = 0; i < n; i++) h.bern_key.push_back((uint32_t)(first + i));
        h.seed = seed;
    } catch (...) {
        return fail(BSIM_E_NOMEM, "out of host memory");
    }
    return first;
}

int64_t bsim_net_connect(bsim_network *net, int64_t count, const int32_t *pre,
                         const int32_t *post, const float *w, const float *d,
                         const int8_t *stype, const float *tau)
{
    if (!net || count < 0 || (count > 0 && (!pre || !post || !w || !d)))
        return fail(BSIM_E_ARG, "bsim_net_connect: bad argument");
    auto &h = net->net;
    try {
        h.pre.insert(h.pre.end(), pre, pre + count);
        h.post.insert(h.post.end(), post, post + count);
        h.w.insert(h.w.end(), w, w + count);
        h.d.insert(h.d.end(), d, d + count);
        if (stype)
            h.stype.insert(h.stype.end(), stype, stype + count);
        else
            h.stype.insert(h.stype.end(), (size_t)count, (int8_t)0);
        if (tau) {
            if (h.tau.size() < h.pre.size() - (size_t)count) h.tau.resize(h.pre.size() - (size_t)count, 0.0f);
            h.tau.insert(h.tau.end(), tau, tau + count);
        } else if (!h.tau.empty()) {
            h.tau.insert(h.tau.end(), (size_t)count, 0.0f);
        }
    } catch (...) {
        return fail(BSIM_E_NOMEM, "out of host memory");
    }
    return h.s();
}

int32_t bsim_net_num_neurons(const bsim_network *net) { return net ? net->net.n() : BSIM_E_ARG; }

int64_t bsim_net_num_synapses(const bsim_network *net) { return net ? net->net.s() : BSIM_E_ARG; }

int32_t bsim_net_lif_constants(const bsim_network *net, float dt, float *out9, int32_t *refrac_time)
{
    if (!net || !out9 || !refrac_time) return fail(BSIM_E_ARG, "bad argument");
    const auto &h = net->net;
    for (int32_t i = 0; i < h.n(); i++) {
        if (h.kind[i] == BSIM_KIND_LIF) {
            bsim::lif_constants(&h.lif[(size_t)i * bsim::kLifParams], dt, out9 + (size_t)i * 9,
                                &refrac_time[i]);
        } else {
            memset(out9 + (size_t)i * 9, 0, 9 * sizeof(float));
            refrac_time[i] = 0;
        }
    }
    return 0;
}

int32_t bsim_net_layout(const bsim_network *net, float dt, int32_t tile, int32_t *ntiles,
                        int64_t *s, int32_t *dmax, int64_t *tile_syn, uint16_t *syn_tgt,
                        float *syn_w)
{
    if (!net) return fail(BSIM_E_ARG, "bad argument");
    bsim::Layout L;
    std::string err;
    try {
        if (bsim::build_layout(net->net, dt, tile, L, err) != 0) return fail(BSIM_E_NETWORK, err);
    } catch (...) {
        return fail(BSIM_E_NOMEM, "out of host memory");
    }
    if (ntiles) *ntiles = L.ntiles;
    if (s) *s = L.s;
    if (dmax) *dmax = L.dmax;
    if (tile_syn) {
        for (int32_t t = 0; t < L.ntiles; t++) tile_syn[t] = L.tile_base[t];
        tile_syn[L.ntiles] = L.s;
    }
    if (syn_tgt) memcpy(syn_tgt, L.syn_tgt.data(), L.syn_tgt.size() * sizeof(uint16_t));
    if (syn_w) memcpy(syn_w, L.syn_w.data(), L.syn_w.size() * sizeof(float));
    return 0;
}

/* ---- device simulation ---------------------------------------------------- */

}  // extern "C"

/* Releases a simulation: waits for its stream, destroys the captured graph
 * before the memory its launches use, frees, destroys the stream.  Every step
 * is attempted; the first failure is kept as the thread's error message. */
static void sim_free(bsim_sim *sim)
{
    if (!sim) return;
    hipError_t first = hipSuccess;
    auto keep = [&first](hipError_t e) {
        if (e != hipSuccess && first == hipSuccess) first = e;
    };
    keep(hipSetDevice(sim->device));
    if (sim->stream) keep(hipStreamSynchronize(sim->stream));
    if (sim->gexec) keep(hipGraphExecDestroy(sim->gexec));
    if (sim->graph) keep(hipGraphDestroy(sim->graph));
    sim->gexec = nullptr;
    sim->graph = nullptr;
    for (void *p : sim->allocs) keep(hipFree(p));
    sim->allocs.clear();
    if (sim->stream) keep(hipStreamDestroy(sim->stream));
    if (first != hipSuccess) fail(BSIM_E_HIP, std::string("bsim_sim_destroy: ") + hipGetErrorString(first));
    delete sim;
}

/* Makes the simulation's device current (0, or the failure code). */
static int32_t use_device(const bsim_sim *sim)
{
    const hipError_t e = hipSetDevice(sim->device);
    return e == hipSuccess ? 0 : fail(BSIM_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
}

template <class T>
static T *dev_alloc(bsim_sim *sim, size_t count, hipError_t &e)
{
    void *p = nullptr;
    const size_t bytes = (count ? count : 1) * sizeof(T);
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();   /* the failure is reported here, not by a later launch check */
        return nullptr;
    }
    sim->allocs.push_back(p);
    sim->device_bytes += (int64_t)bytes;
    return (T *)p;
}

template <class T>
static T *dev_upload(bsim_sim *sim, const std::vector<T> &v, hipError_t &e)
{
    T *p = dev_alloc<T>(sim, v.size(), e);
    if (!p) return nullptr;
    if (!v.empty()) e = hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, sim->stream);
    if (e != hipSuccess) return nullptr;
    e = hipStreamSynchronize(sim->stream);
    return e == hipSuccess ? p : nullptr;
}

extern "C" {

}  // extern "C"

/* Checks the simulation config and makes its device current (0, or a
 * failure code). */
/* Tile schedule: the row-list launch ranks the tiles heaviest first every
 * kRankEvery cycles, by their work in the step launch before it (the volleys
 * of the standard network move slowly against 2 cycles: on C1, every 2nd
 * cycle ran 114.3 us per cycle, every cycle 114.8, every 4th 116.4);
 * bsim_sim_set_schedule changes it. */
static const int32_t kRankEvery = 2;

static int32_t check_config(const bsim_sim_config &c, int32_t N, const char *who)
{
    if (c.rec_steps < 0 || c.n_rec < 0 || (c.n_rec > 0 && !c.rec_ids))
        return fail(BSIM_E_ARG, std::string(who) + ": bad recording configuration");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(BSIM_E_HIP, "no HIP device available");
    if (c.device < 0 || c.device >= ndev) return fail(BSIM_E_ARG, "device ordinal out of range");
    const hipError_t e = hipSetDevice(c.device);
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    for (int32_t r = 0; r < c.n_rec; r++)
        if (c.rec_ids[r] < 0 || c.rec_ids[r] >= N) return fail(BSIM_E_ARG, "recorded neuron id out of range");
    return 0;
}

static bsim_sim *sim_from_layout(bsim::Layout &L, const bsim_sim_config &c, bsim_syn *dev_syn,
                                 int64_t dev_syn_bytes);

extern "C" {

bsim_sim *bsim_sim_create(const bsim_network *net, float dt, const bsim_sim_config *cfg)
{
    if (!net || !(dt > 0)) {
        fail(BSIM_E_ARG, "bsim_sim_create: bad argument");
        return nullptr;
    }
    bsim_sim_config c{};
    if (cfg) c = *cfg;
    const int32_t N = net->net.n();
    if (check_config(c, N, "bsim_sim_create") != 0) return nullptr;
    const int32_t tile = c.tile > 0 ? c.tile : bsim::default_tile(bsim::max_delay_steps(net->net, dt));
    bsim::Layout L;
    std::string err;
    try {
        if (bsim::build_layout(net->net, dt, tile, L, err) != 0) {
            fail(BSIM_E_NETWORK, err);
            return nullptr;
        }
    } catch (const std::bad_alloc &) {
        fail(BSIM_E_NOMEM, "out of host memory while building the layout");
        return nullptr;
    }
    return sim_from_layout(L, c, nullptr, 0);
}

bsim_sim *bsim_sim_create_chains(const bsim_chain_params *p, int32_t count, float dt,
                                 const bsim_sim_config *cfg, int64_t part_synapses,
                                 int32_t *first_ids)
{
    if (!p || count <= 0 || !(dt > 0) || part_synapses <= 0) {
        fail(BSIM_E_ARG, "bsim_sim_create_chains: bad argument");
        return nullptr;
    }
    bsim_sim_config c{};
    if (cfg) c = *cfg;
    int64_t nn = 0, ns = 0;
    int32_t dmax = 1;
    if (bsim::chains_size(p, count, &nn, &ns, &dmax) != 0) {
        fail(BSIM_E_ARG, "bsim_sim_create_chains: invalid chain parameters");
        return nullptr;
    }
    const int32_t tile = c.tile > 0 ? c.tile : bsim::default_tile(dmax);
    /* recorded ids are checked against the laid-out network below */
    if (check_config(c, 0x7FFFFFFF, "bsim_sim_create_chains") != 0) return nullptr;
    /* the packed synapses go straight to the device, part by part */
    bsim_syn *dsyn = nullptr;
    const int64_t sbytes = std::max<int64_t>(ns, 1) * (int64_t)sizeof(bsim_syn);
    hipError_t e = hipMalloc((void **)&dsyn, (size_t)sbytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(e == hipErrorOutOfMemory ? BSIM_E_NOMEM : BSIM_E_HIP,
             std::string("bsim_sim_create_chains: ") + hipGetErrorString(e));
        return nullptr;
    }
    bsim::Layout L;
    std::string err;
    int rc = 0;
    try {
        auto upload = [dsyn](const bsim::Layout &part, int64_t s0) -> int {
            std::vector<bsim_syn> packed(part.syn_tgt.size());
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)packed.size(); i++) {
                packed[i].tgt = part.syn_tgt[i];
                memcpy(&packed[i].w, &part.syn_w[i], 4);
            }
            if (packed.empty()) return 0;
            const hipError_t x = hipMemcpy(dsyn + s0, packed.data(), packed.size() * sizeof(bsim_syn),
                                           hipMemcpyHostToDevice);
            return x == hipSuccess ? 0 : fail(BSIM_E_HIP, std::string("synapse upload: ") + hipGetErrorString(x));
        };
        rc = bsim::build_chains_layout(p, count, dt, tile, dmax, part_synapses, L, first_ids, err, upload);
    } catch (const std::bad_alloc &) {
        rc = -1;
        err = "out of host memory while building the layout";
    }
    if (rc != 0) {
        if (!err.empty()) fail(BSIM_E_NETWORK, err);
        (void)hipFree(dsyn);
        return nullptr;
    }
    for (int32_t r = 0; r < c.n_rec; r++)
        if (c.rec_ids[r] < 0 || c.rec_ids[r] >= L.n) {
            (void)hipFree(dsyn);
            fail(BSIM_E_ARG, "recorded neuron id out of range");
            return nullptr;
        }
    return sim_from_layout(L, c, dsyn, sbytes);
}

}  // extern "C"

/* Uploads a layout and allocates the step state; dev_syn, when not null, is
 * the packed synapse array already on the device (adopted by the
 * simulation, freed with it). */
static bsim_sim *sim_from_layout(bsim::Layout &L, const bsim_sim_config &c, bsim_syn *dev_syn,
                                 int64_t dev_syn_bytes)
{
    hipError_t e = hipSuccess;
    bsim_sim *sim = new (std::nothrow) bsim_sim();
    if (!sim) {
        if (dev_syn) (void)hipFree(dev_syn);
        fail(BSIM_E_NOMEM, "out of host memory");
        return nullptr;
    }
    if (dev_syn) {
        sim->allocs.push_back(dev_syn);
        sim->device_bytes += dev_syn_bytes;
    }
    sim->device = c.device;
    e = hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        fail(BSIM_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        sim->stream = nullptr;
        sim_free(sim);
        return nullptr;
    }
    bsim_dev &d = sim->d;
    d.n = L.n;
    d.tile = L.tile;
    d.ntiles = L.ntiles;
    d.dmax = L.dmax;
    d.nwords = (L.n + 63) / 64;
    sim->s = L.s;
    sim->n_src = (int64_t)L.rng.size();
    sim->n_off = (int64_t)L.off_tab.size();
#define UP(field, vec)                                   \
    do {                                                 \
        d.field = dev_upload(sim, vec, e);               \
        if (e != hipSuccess) goto hip_fail;              \
    } while (0)
    UP(ncls, L.ncls);
    UP(lifc, L.lifc);
    UP(v_tmp, L.v_tmp);
    UP(arr_end, L.arr_end);
    UP(arr_t, L.arr_t);
    UP(bern_p, L.bern_p);
    UP(bern_seed, L.bern_seed);
    UP(bern_key, L.bern_key);
    UP(pois_tab, L.pois_tab);
    UP(tile_base, L.tile_base);
    UP(tile_nsyn, L.tile_nsyn);
    UP(tile_res, L.tile_res);
    d.nslots = L.nslots;
    if (dev_syn) {
        d.syn = dev_syn;
    } else {
        /* targets and weights packed into one 8-byte word per synapse */
        std::vector<bsim_syn> packed(L.syn_tgt.size());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)packed.size(); i++) {
            packed[i].tgt = L.syn_tgt[i];
            memcpy(&packed[i].w, &L.syn_w[i], 4);
        }
        std::vector<uint16_t>().swap(L.syn_tgt);
        std::vector<float>().swap(L.syn_w);
        UP(syn, packed);
    }
    UP(tile_rng, L.tile_rng);
    UP(rng, L.rng);
    UP(off_tab, L.off_tab);
    UP(src_lo, L.src_lo);
    UP(src_hi, L.src_hi);
    UP(src_list_base, L.src_list_base);
    UP(chunk_tiles_off, L.chunk_tiles_off);
    UP(chunk_rng, L.chunk_rng);
    d.nsrc = (int32_t)L.src_lo.size();
    /* state */
    UP(v, L.v_init);
    UP(arr_cur, L.arr_cur0);
#undef UP
    {
        const size_t n = (size_t)(L.n > 0 ? L.n : 1);
        const size_t nt = (size_t)(L.ntiles > 0 ? L.ntiles : 1);
        const size_t nw = (size_t)(d.nwords > 0 ? d.nwords : 1);
        if (!(d.ie = dev_alloc<float>(sim, n, e))) goto hip_fail;
        if (!(d.ii = dev_alloc<float>(sim, n, e))) goto hip_fail;
        if (!(d.refrac = dev_alloc<int32_t>(sim, n, e))) goto hip_fail;
        if (!(d.fire_count = dev_alloc<int32_t>(sim, n, e))) goto hip_fail;
        if (!(d.hist = dev_alloc<uint64_t>(sim, 2 * n, e))) goto hip_fail;
        if (!(d.active = dev_alloc<uint64_t>(sim, 2 * nw, e))) goto hip_fail;
        if (!(d.ev_count = dev_alloc<int64_t>(sim, nt, e))) goto hip_fail;
        if (!(sim->ev_fin = dev_alloc<int64_t>(sim, nt, e))) goto hip_fail;
        if (!(d.done = dev_alloc<uint64_t>(sim, 8 * BSIM_DONE_STRIDE, e))) goto hip_fail;
        if ((e = hipMemsetAsync(d.done, 0, 8 * BSIM_DONE_STRIDE * sizeof(uint64_t), sim->stream)) != hipSuccess)
            goto hip_fail;
        {
            /* until the first ranking: the tiles of each residue in order */
            std::vector<int32_t> perm((size_t)(L.nslots > 0 ? L.nslots : 1), -1);
            std::vector<int32_t> rank(8, 0);
            for (int32_t t = 0; t < L.ntiles; t++) {
                const int x = L.tile_res[t];
                perm[(size_t)8 * rank[x]++ + x] = t;
            }
            d.perm = dev_upload(sim, perm, e);
            if (e != hipSuccess) goto hip_fail;
        }
        d.rank_every = kRankEvery;
        d.done_inv = d.nslots >= 8 ? 8.0 / (double)d.nslots : 0.0;
        if (!(d.tile_load = dev_alloc<uint64_t>(sim, nt, e))) goto hip_fail;
        if ((e = hipMemsetAsync(d.tile_load, 0, nt * sizeof(uint64_t), sim->stream)) != hipSuccess) goto hip_fail;
        if (!(d.tile_cls = dev_alloc<uint8_t>(sim, nt, e))) goto hip_fail;
        const size_t nre = (size_t)(L.rng.empty() ? 1 : L.rng.size());
        if (!(d.rng_count = dev_alloc<uint32_t>(sim, nre, e))) goto hip_fail;
        if ((e = hipMemsetAsync(d.rng_count, 0, nre * sizeof(uint32_t), sim->stream)) != hipSuccess) goto hip_fail;
        if (!(d.src_list = dev_alloc<uint32_t>(sim, (size_t)(L.src_list_len > 0 ? L.src_list_len : 1), e)))
            goto hip_fail;
        if ((e = hipMemsetAsync(d.ie, 0, n * 4, sim->stream)) != hipSuccess) goto hip_fail;
        if ((e = hipMemsetAsync(d.ii, 0, n * 4, sim->stream)) != hipSuccess) goto hip_fail;
        if ((e = hipMemsetAsync(d.refrac, 0, n * 4, sim->stream)) != hipSuccess) goto hip_fail;
        if ((e = hipMemsetAsync(d.fire_count, 0, n * 4, sim->stream)) != hipSuccess) goto hip_fail;
        if ((e = hipMemsetAsync(d.hist, 0, 2 * n * 8, sim->stream)) != hipSuccess) goto hip_fail;
        if ((e = hipMemsetAsync(d.active, 0, 2 * nw * 8, sim->stream)) != hipSuccess) goto hip_fail;
        if ((e = hipMemsetAsync(d.ev_count, 0, nt * 8, sim->stream)) != hipSuccess) goto hip_fail;
        if ((e = hipMemsetAsync(sim->ev_fin, 0, nt * 8, sim->stream)) != hipSuccess) goto hip_fail;
        if (c.rec_steps > 0) {
            d.rec_steps = c.rec_steps;
            d.rec_t0 = 0;
            const size_t rb = (size_t)c.rec_steps * (size_t)(d.nwords > 0 ? d.nwords : 1);
            if (!(d.raster = dev_alloc<uint64_t>(sim, rb, e))) goto hip_fail;
            if ((e = hipMemsetAsync(d.raster, 0, rb * 8, sim->stream)) != hipSuccess) goto hip_fail;
            if (c.n_rec > 0) {
                /* one device column per distinct neuron; repeated ids share it */
                std::vector<int32_t> idx(L.n, -1);
                int32_t ncol = 0;
                sim->rec_col.resize((size_t)c.n_rec);
                for (int32_t r = 0; r < c.n_rec; r++) {
                    int32_t &x = idx[c.rec_ids[r]];
                    if (x < 0) x = ncol++;
                    sim->rec_col[r] = x;
                }
                d.rec_index = dev_upload(sim, idx, e);
                if (e != hipSuccess) goto hip_fail;
                d.n_rec = ncol;
                const size_t vb = (size_t)c.rec_steps * (size_t)ncol;
                if (!(d.rec_v = dev_alloc<float>(sim, vb, e))) goto hip_fail;
                if ((e = hipMemsetAsync(d.rec_v, 0, vb * 4, sim->stream)) != hipSuccess) goto hip_fail;
            }
        }
        sim->rec_steps = c.rec_steps;
        sim->n_rec = c.rec_steps > 0 ? c.n_rec : 0;
        e = hipStreamSynchronize(sim->stream);
        if (e != hipSuccess) goto hip_fail;
    }
    return sim;
hip_fail:
    fail(e == hipErrorOutOfMemory ? BSIM_E_NOMEM : BSIM_E_HIP,
         std::string("device setup failed: ") + hipGetErrorString(e));
    sim_free(sim);
    return nullptr;
}

/* Cycles per captured graph: runs of this many cycles are replayed from one
 * hipGraph of identical clocked launches (the step kernel reads and advances
 * the device-side cycle counter; the row-list kernel after it builds the
 * next cycle's active-row lists and ranks the tiles). */
static const int32_t kGraphCycles = 64;

static hipError_t build_graph(bsim_sim *sim)
{
    hipError_t e = hipStreamBeginCapture(sim->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = hipSuccess;
    for (int32_t i = 0; i < kGraphCycles && le == hipSuccess; i++) {
        le = bsim_launch_step(&sim->d, -1, 0, sim->stream);
        if (le == hipSuccess) le = bsim_launch_rows(&sim->d, sim->stream);
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(sim->stream, &g);
    if (le != hipSuccess) e = le;
    if (e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return e;
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return e;
    }
    sim->graph = g;
    sim->gexec = x;
    return hipSuccess;
}

extern "C" {

int32_t bsim_sim_step(bsim_sim *sim, int32_t n)
{
    if (!sim || n < 0) return fail(BSIM_E_ARG, "bsim_sim_step: bad argument");
    if ((int64_t)sim->cycle + n > 0x7FFFFFFF) return fail(BSIM_E_ARG, "bsim_sim_step: cycle overflow");
    if (int32_t rc = use_device(sim)) return rc;
    hipError_t e = hipSuccess;
    if (n >= kGraphCycles && !sim->gexec) {
        e = build_graph(sim);
        if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("step graph: ") + hipGetErrorString(e));
    }
    while (n >= kGraphCycles) {
        e = hipGraphLaunch(sim->gexec, sim->stream);
        if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("step graph launch: ") + hipGetErrorString(e));
        sim->cycle += kGraphCycles;
        n -= kGraphCycles;
    }
    for (; n > 0; n--) {
        e = bsim_launch_step(&sim->d, -1, 0, sim->stream);
        if (e == hipSuccess) e = bsim_launch_rows(&sim->d, sim->stream);
        if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("step launch: ") + hipGetErrorString(e));
        sim->cycle++;
    }
    return 0;
}

int32_t bsim_sim_time_steps(bsim_sim *sim, int32_t n, double *step_us, double *rows_us)
{
    if (!sim || n <= 0) return fail(BSIM_E_ARG, "bsim_sim_time_steps: bad argument");
    if (int32_t rc = use_device(sim)) return rc;
    std::vector<hipEvent_t> ev((size_t)n * 3, nullptr);
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < ev.size() && e == hipSuccess; i++) e = hipEventCreate(&ev[i]);
    for (int32_t i = 0; i < n && e == hipSuccess; i++) {
        e = hipEventRecord(ev[3 * i], sim->stream);
        if (e == hipSuccess) e = bsim_launch_step(&sim->d, -1, 0, sim->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[3 * i + 1], sim->stream);
        if (e == hipSuccess) e = bsim_launch_rows(&sim->d, sim->stream);
        if (e == hipSuccess) e = hipEventRecord(ev[3 * i + 2], sim->stream);
        if (e == hipSuccess) sim->cycle++;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(sim->stream);
    double a = 0.0, b = 0.0;
    for (int32_t i = 0; i < n && e == hipSuccess; i++) {
        float x = 0.f, y = 0.f;
        e = hipEventElapsedTime(&x, ev[3 * i], ev[3 * i + 1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&y, ev[3 * i + 1], ev[3 * i + 2]);
        a += x;
        b += y;
    }
    for (hipEvent_t x : ev)
        if (x) hipEventDestroy(x);
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("time_steps: ") + hipGetErrorString(e));
    if (step_us) *step_us = a * 1e3 / n;
    if (rows_us) *rows_us = b * 1e3 / n;
    return 0;
}

int32_t bsim_sim_finish(bsim_sim *sim)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (sim->cycle == 0 || sim->finished_at == sim->cycle) return 0;
    if (int32_t rc = use_device(sim)) return rc;
    bsim_dev d = sim->d;
    d.ev_count = sim->ev_fin;
    hipError_t e = hipMemsetAsync(sim->ev_fin, 0, (size_t)(d.ntiles > 0 ? d.ntiles : 1) * 8, sim->stream);
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("finish: ") + hipGetErrorString(e));
    e = bsim_launch_step(&d, sim->cycle, 1, sim->stream);
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("finish launch: ") + hipGetErrorString(e));
    sim->finished_at = sim->cycle;
    return 0;
}

int32_t bsim_sim_sync(bsim_sim *sim)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (int32_t rc = use_device(sim)) return rc;
    hipError_t e = hipStreamSynchronize(sim->stream);
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("stream: ") + hipGetErrorString(e));
    return 0;
}

int32_t bsim_sim_cycle(const bsim_sim *sim) { return sim ? sim->cycle : BSIM_E_ARG; }

}  // extern "C"

template <class T>
static hipError_t fetch(bsim_sim *sim, T *dst, const T *src, size_t count)
{
    if (!dst || !count) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, sim->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(sim->stream);
}

extern "C" {

int64_t bsim_sim_read_raster(bsim_sim *sim, int32_t *counts, int32_t *ids, int64_t cap,
                             int64_t *needed)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (!sim->d.raster) return fail(BSIM_E_STATE, "raster recording was not enabled");
    if (int32_t rc = use_device(sim)) return rc;
    const int32_t steps = sim->cycle < sim->rec_steps ? sim->cycle : sim->rec_steps;
    std::vector<uint64_t> bits((size_t)steps * sim->d.nwords);
    hipError_t e = fetch(sim, bits.data(), sim->d.raster, bits.size());
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("raster copy: ") + hipGetErrorString(e));
    int64_t total = 0;
    for (uint64_t b : bits) total += __builtin_popcountll(b);
    if (needed) *needed = total;
    if (total > cap || !counts || (total > 0 && !ids)) return fail(BSIM_E_CAPACITY, "raster buffer too small");
    int64_t o = 0;
    for (int32_t t = 0; t < steps; t++) {
        int32_t c = 0;
        for (int32_t wd = 0; wd < sim->d.nwords; wd++) {
            uint64_t b = bits[(size_t)t * sim->d.nwords + wd];
            while (b) {
                const int bit = __builtin_ctzll(b);
                b &= b - 1;
                ids[o++] = wd * 64 + bit;
                c++;
            }
        }
        counts[t] = c;
    }
    return o;
}

int32_t bsim_sim_read_trace(bsim_sim *sim, float *v)
{
    if (!sim || !v) return fail(BSIM_E_ARG, "bad argument");
    if (!sim->d.rec_v) return fail(BSIM_E_STATE, "no neuron is traced");
    if (int32_t rc = use_device(sim)) return rc;
    const int32_t steps = sim->cycle < sim->rec_steps ? sim->cycle : sim->rec_steps;
    const int32_t ncol = sim->d.n_rec;
    std::vector<float> cols((size_t)steps * (size_t)ncol);
    hipError_t e = fetch(sim, cols.data(), sim->d.rec_v, cols.size());
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("trace copy: ") + hipGetErrorString(e));
    for (int32_t t = 0; t < steps; t++)
        for (int32_t r = 0; r < sim->n_rec; r++)
            v[(size_t)t * sim->n_rec + r] = cols[(size_t)t * ncol + sim->rec_col[r]];
    return steps;
}

int32_t bsim_sim_enable_state_trace(bsim_sim *sim)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (!sim->d.rec_v) return fail(BSIM_E_STATE, "no neuron is traced");
    if (sim->d.rec_ie) return 0;
    if (sim->cycle > 0) return fail(BSIM_E_STATE, "state traces must be enabled before the first cycle");
    hipError_t e = hipSetDevice(sim->device);
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    const size_t vb = (size_t)sim->rec_steps * (size_t)sim->d.n_rec;
    float *buf[4];
    for (int i = 0; i < 4; i++) {
        buf[i] = dev_alloc<float>(sim, vb, e);
        if (!buf[i]) return fail(e == hipErrorOutOfMemory ? BSIM_E_NOMEM : BSIM_E_HIP,
                                 std::string("state trace: ") + hipGetErrorString(e));
        e = hipMemsetAsync(buf[i], 0, vb * sizeof(float), sim->stream);
        if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("state trace: ") + hipGetErrorString(e));
    }
    sim->d.rec_ie = buf[0];
    sim->d.rec_ii = buf[1];
    sim->d.rec_in_e = buf[2];
    sim->d.rec_in_i = buf[3];
    /* the captured launches hold the kernel arguments: capture again */
    if (sim->gexec) {
        if ((e = hipStreamSynchronize(sim->stream)) != hipSuccess)
            return fail(BSIM_E_HIP, std::string("state trace: ") + hipGetErrorString(e));
        if ((e = hipGraphExecDestroy(sim->gexec)) != hipSuccess ||
            (e = hipGraphDestroy(sim->graph)) != hipSuccess)
            return fail(BSIM_E_HIP, std::string("state trace: ") + hipGetErrorString(e));
        sim->gexec = nullptr;
        sim->graph = nullptr;
    }
    return 0;
}

int32_t bsim_sim_read_state_trace(bsim_sim *sim, float *ie, float *ii, float *in_e, float *in_i)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (!sim->d.rec_ie) return fail(BSIM_E_STATE, "state traces not enabled (bsim_sim_enable_state_trace)");
    hipError_t e = hipSetDevice(sim->device);
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    const int32_t steps = sim->cycle < sim->rec_steps ? sim->cycle : sim->rec_steps;
    const int32_t ncol = sim->d.n_rec;
    float *out[4] = {ie, ii, in_e, in_i};
    const float *src[4] = {sim->d.rec_ie, sim->d.rec_ii, sim->d.rec_in_e, sim->d.rec_in_i};
    std::vector<float> cols((size_t)steps * (size_t)ncol);
    for (int i = 0; i < 4; i++) {
        if (!out[i]) continue;
        e = fetch(sim, cols.data(), src[i], cols.size());
        if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("state trace copy: ") + hipGetErrorString(e));
        for (int32_t t = 0; t < steps; t++)
            for (int32_t r = 0; r < sim->n_rec; r++)
                out[i][(size_t)t * sim->n_rec + r] = cols[(size_t)t * ncol + sim->rec_col[r]];
    }
    return steps;
}

int32_t bsim_sim_read_state(bsim_sim *sim, float *v, float *ie, float *ii, int32_t *refrac,
                            int32_t *fire_count)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (int32_t rc = use_device(sim)) return rc;
    const size_t n = (size_t)sim->d.n;
    hipError_t e;
    if ((e = fetch(sim, v, sim->d.v, n)) != hipSuccess ||
        (e = fetch(sim, ie, sim->d.ie, n)) != hipSuccess ||
        (e = fetch(sim, ii, sim->d.ii, n)) != hipSuccess ||
        (e = fetch(sim, refrac, sim->d.refrac, n)) != hipSuccess ||
        (e = fetch(sim, fire_count, sim->d.fire_count, n)) != hipSuccess)
        return fail(BSIM_E_HIP, std::string("state copy: ") + hipGetErrorString(e));
    return 0;
}

int32_t bsim_sim_enable_input_log(bsim_sim *sim)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (sim->log_in) return 0;
    if (int32_t rc = use_device(sim)) return rc;
    const size_t n = (size_t)(sim->d.n > 0 ? sim->d.n : 1);
    hipError_t e = hipSuccess;
    float *in = dev_alloc<float>(sim, n, e);
    if (e != hipSuccess) return fail(BSIM_E_NOMEM, std::string("input log: ") + hipGetErrorString(e));
    int32_t *slot = dev_alloc<int32_t>(sim, n, e);
    if (e != hipSuccess) return fail(BSIM_E_NOMEM, std::string("input log: ") + hipGetErrorString(e));
    if ((e = hipMemsetAsync(in, 0, n * sizeof(float), sim->stream)) != hipSuccess ||
        (e = hipMemsetAsync(slot, 0, n * sizeof(int32_t), sim->stream)) != hipSuccess)
        return fail(BSIM_E_HIP, std::string("input log: ") + hipGetErrorString(e));
    sim->log_in = in;
    sim->log_slot = slot;
    sim->d.in_sum = in;
    /* without synapses the reference's delay ring has no slot (MAX_DELAY 0) */
    sim->d.slot_count = sim->s > 0 ? slot : nullptr;
    /* the captured launches hold the kernel arguments: capture again */
    if (sim->gexec) {
        hipError_t x;
        if ((x = hipStreamSynchronize(sim->stream)) != hipSuccess ||
            (x = hipGraphExecDestroy(sim->gexec)) != hipSuccess ||
            (x = hipGraphDestroy(sim->graph)) != hipSuccess)
            return fail(BSIM_E_HIP, std::string("input log: ") + hipGetErrorString(x));
        sim->gexec = nullptr;
        sim->graph = nullptr;
    }
    return 0;
}

int32_t bsim_sim_read_input_log(bsim_sim *sim, float *in_sum, int32_t *slot_count)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (!sim->log_in) return fail(BSIM_E_STATE, "input log not enabled (bsim_sim_enable_input_log)");
    if (int32_t rc = use_device(sim)) return rc;
    const size_t n = (size_t)sim->d.n;
    hipError_t e;
    if ((e = fetch(sim, in_sum, sim->log_in, n)) != hipSuccess ||
        (e = fetch(sim, slot_count, sim->log_slot, n)) != hipSuccess)
        return fail(BSIM_E_HIP, std::string("input log copy: ") + hipGetErrorString(e));
    return 0;
}

int64_t bsim_sim_events(bsim_sim *sim)
{
    if (!sim) return fail(BSIM_E_ARG, "bad argument");
    if (int32_t rc = use_device(sim)) return rc;
    const size_t nt = (size_t)sim->d.ntiles;
    std::vector<int64_t> a(nt), b(nt);
    hipError_t e = fetch(sim, a.data(), sim->d.ev_count, nt);
    if (e == hipSuccess && sim->finished_at == sim->cycle) e = fetch(sim, b.data(), sim->ev_fin, nt);
    if (e != hipSuccess) return fail(BSIM_E_HIP, std::string("event count copy: ") + hipGetErrorString(e));
    int64_t s = 0;
    for (size_t i = 0; i < nt; i++) s += a[i] + b[i];
    return s;
}

int32_t bsim_sim_get_info(const bsim_sim *sim, bsim_sim_info *info)
{
    if (!sim || !info) return fail(BSIM_E_ARG, "bad argument");
    info->n = sim->d.n;
    info->tile = sim->d.tile;
    info->ntiles = sim->d.ntiles;
    info->dmax = sim->d.dmax;
    info->s = sim->s;
    info->device_bytes = sim->device_bytes;
    info->n_src_entries = sim->n_src;
    info->off_entries = sim->n_off;
    return 0;
}

void *bsim_sim_stream(bsim_sim *sim) { return sim ? (void *)sim->stream : nullptr; }

int32_t bsim_sim_set_schedule(bsim_sim *sim, int32_t rank_every)
{
    if (!sim || rank_every < 0) return fail(BSIM_E_ARG, "bsim_sim_set_schedule: bad argument");
    if (int32_t rc = use_device(sim)) return rc;
    if (sim->d.rank_every == rank_every) return 0;
    hipError_t e;
    /* the captured launches hold the kernel arguments: capture again; the
     * loads of a tile ranked later start from zero */
    if ((e = hipStreamSynchronize(sim->stream)) != hipSuccess ||
        (e = hipMemsetAsync(sim->d.tile_load, 0, (size_t)(sim->d.ntiles > 0 ? sim->d.ntiles : 1) * 8,
                            sim->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(sim->stream)) != hipSuccess)
        return fail(BSIM_E_HIP, std::string("bsim_sim_set_schedule: ") + hipGetErrorString(e));
    if (sim->gexec) {
        if ((e = hipGraphExecDestroy(sim->gexec)) != hipSuccess || (e = hipGraphDestroy(sim->graph)) != hipSuccess)
            return fail(BSIM_E_HIP, std::string("bsim_sim_set_schedule: ") + hipGetErrorString(e));
        sim->gexec = nullptr;
        sim->graph = nullptr;
    }
    sim->d.rank_every = rank_every;
    return 0;
}

void bsim_sim_destroy(bsim_sim *sim) { sim_free(sim); }

}  // extern "C"
