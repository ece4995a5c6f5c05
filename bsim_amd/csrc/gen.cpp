/*
 * gen.cpp -- seeded generator of the standard feed-forward workloads, and
 * export of a host network's arrays.
 *
 * The reference's benchmark network (test/gpu/standard_test.cpp:92-139) is a
 * chain of populations with some projections to remote populations, built
 * with all-to-all connect() calls and a time-seeded RNG (src/utils/utils.h:
 * 34-47).  This generator builds the sparse form used by the workloads: a
 * chain of npop LIF populations where population i projects to i+1 and a
 * remote_frac share also to one remote population, each pre neuron drawing
 * its targets by stratified sampling (one target per equal integer stratum,
 * hence distinct), with weights, signs and delays from a counter-based hash
 * (splitmix64 finalizer) so that the network is a pure function of the
 * parameters, whatever the thread count.  Synapses are created source
 * population by source population, pre-major: every target receives its
 * inputs in ascending pre order, at most one synapse per (pre, projection).
 *
 * bsim_net_gen_chain appends one network of a parameter sweep (a chain of
 * populations of given sizes with a given fan-in) to a network; all of its
 * draws are keyed by chain-local indices, so a chain built into a batch is
 * the same network, neuron for neuron and draw for draw, as built alone.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <functional>
#include <string>
#include <vector>

#include "../../include/bsim_amd.h"
#include "layout.h"
#include "internal.h"
#include "network.h"


namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* hash of (seed, stream, counter); stream = tag << 32 | index */
inline uint64_t h3(uint64_t seed, uint64_t stream, uint64_t counter)
{
    return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) ^ counter);
}

inline double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

enum : uint64_t { TAG_TGT = 1, TAG_INH = 2, TAG_W = 3, TAG_D = 4, TAG_IOFF = 5, TAG_REM = 6, TAG_REMDST = 7 };

struct Proj {
    int32_t src, dst, f;   /* population indices (0 = drive if present), fan-out */
};

}  // namespace

extern "C" int64_t bsim_net_gen_standard(bsim_network *netp, const bsim_gen_params *gp)
{
    if (!netp || !gp || gp->npop <= 0 || gp->n <= 0 || gp->fanout < 0 || gp->dmax < 1 ||
        gp->dmax > BSIM_MAX_DELAY || !(gp->dt > 0)) {
        g_bsim_err = "bsim_net_gen_standard: bad parameters";
        return BSIM_E_ARG;
    }
    if (netp->net.n() != 0 || netp->net.s() != 0) {
        g_bsim_err = "bsim_net_gen_standard: network must be empty";
        return BSIM_E_STATE;
    }
    const bsim_gen_params p = *gp;
    bsim::HostNet &h = netp->net;
    std::vector<std::pair<int32_t, int32_t>> pops;   /* (start, count) */
    int64_t nn = 0;
    const bool drive = p.drive_rate_hz > 0 && p.n_drive > 0;
    try {
        if (drive) {
            double pr = std::min(1.0, (double)p.drive_rate_hz * (double)p.dt);
            double thr = std::min(4294967295.0, std::round(pr * 4294967296.0));
            h.kind.insert(h.kind.end(), (size_t)p.n_drive, (int8_t)BSIM_KIND_BERNOULLI);
            for (int32_t i = 0; i < p.n_drive; i++) {
                float row[bsim::kLifParams] = {0};
                row[10] = 1e-3f;
                row[11] = 1e-3f;
                h.lif.insert(h.lif.end(), row, row + bsim::kLifParams);
            }
            h.bern_p.insert(h.bern_p.end(), (size_t)p.n_drive, (uint32_t)thr);
            h.bern_seed.insert(h.bern_seed.end(), (size_t)p.n_drive, p.seed);
            for (int32_t i = 0; i < p.n_drive; i++) h.bern_key.push_back((uint32_t)i);
            h.arr_off.insert(h.arr_off.end(), (size_t)p.n_drive, 0);
            h.seed = p.seed;
            pops.push_back({0, p.n_drive});
            nn = p.n_drive;
        }
        const int32_t first_lif = (int32_t)pops.size();
        if (nn + (int64_t)p.npop * p.n > 0x7FFFFFFF) {
            g_bsim_err = "too many neurons";
            return BSIM_E_ARG;
        }
        const size_t base = h.kind.size();
        h.kind.insert(h.kind.end(), (size_t)p.npop * p.n, (int8_t)BSIM_KIND_LIF);
        h.lif.resize((base + (size_t)p.npop * p.n) * bsim::kLifParams);
        h.bern_p.insert(h.bern_p.end(), (size_t)p.npop * p.n, 0u);
        h.bern_seed.insert(h.bern_seed.end(), (size_t)p.npop * p.n, 0ull);
        h.bern_key.insert(h.bern_key.end(), (size_t)p.npop * p.n, 0u);
        h.arr_off.insert(h.arr_off.end(), (size_t)p.npop * p.n, 0);
#pragma omp parallel for schedule(static)
        for (int64_t g = (int64_t)nn; g < nn + (int64_t)p.npop * p.n; g++) {
            float *row = &h.lif[(size_t)g * bsim::kLifParams];
            float ioff = 0.0f;
            if (p.ioff_hi > p.ioff_lo)
                ioff = (float)((double)p.ioff_lo +
                               ((double)p.ioff_hi - (double)p.ioff_lo) *
                                   u01(h3(p.seed, TAG_IOFF << 32, (uint64_t)g)));
            const float r[bsim::kLifParams] = {0.0f, 0.0f, 0.0f, 20e-3f, 20e-3f, p.tau_refrac,
                                               p.tau_syn_E, p.tau_syn_I, p.v_thresh, ioff,
                                               p.tau_syn_E, p.tau_syn_I};
            std::copy(r, r + bsim::kLifParams, row);
        }
        for (int32_t i = 0; i < p.npop; i++) pops.push_back({(int32_t)(nn + (int64_t)i * p.n), p.n});

        /* projections, in ascending source population */
        std::vector<int32_t> remote(p.npop, -1);
        for (int32_t i = 0; i < p.npop; i++) {
            if (p.npop > 2 && u01(h3(p.seed, TAG_REM << 32, (uint64_t)i)) < (double)p.remote_frac) {
                /* uniform over the populations other than i and i+1 */
                const int32_t nch = p.npop - ((i + 1 < p.npop) ? 2 : 1);
                int32_t c = (int32_t)(u01(h3(p.seed, TAG_REMDST << 32, (uint64_t)i)) * nch);
                if (c >= nch) c = nch - 1;
                int32_t j = 0, seen = 0;
                for (j = 0; j < p.npop; j++) {
                    if (j == i || j == i + 1) continue;
                    if (seen == c) break;
                    seen++;
                }
                remote[i] = j;
            }
        }
        std::vector<Proj> proj;
        if (drive) proj.push_back({0, first_lif, std::min(p.drive_fanout, p.n)});
        for (int32_t i = 0; i < p.npop; i++) {
            const int32_t src = first_lif + i;
            if (i + 1 < p.npop) {
                const int32_t f = remote[i] >= 0 ? p.fanout - p.fanout / 2 : p.fanout;
                proj.push_back({src, src + 1, std::min(f, p.n)});
            }
            if (remote[i] >= 0) proj.push_back({src, first_lif + remote[i], std::min(p.fanout / 2, p.n)});
        }
        int64_t total = 0;
        for (const Proj &q : proj) total += (int64_t)pops[q.src].second * q.f;
        h.pre.resize(total);
        h.post.resize(total);
        h.w.resize(total);
        h.d.resize(total);
        h.stype.resize(total);
        int64_t o = 0;
        for (size_t pi = 0; pi < proj.size(); pi++) {
            const Proj q = proj[pi];
            const int32_t s0 = pops[q.src].first, sn = pops[q.src].second;
            const int32_t d0 = pops[q.dst].first, dn = pops[q.dst].second;
            const int32_t f = q.f;
            const bool is_drive = drive && q.src == 0;
            const uint64_t sid = (uint64_t)pi;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < sn; i++) {
                for (int32_t k = 0; k < f; k++) {
                    const int64_t e = o + (int64_t)i * f + k;
                    const uint64_t c = (uint64_t)i * (uint64_t)f + (uint64_t)k;
                    const int64_t lo = ((int64_t)k * dn) / f, hi = ((int64_t)(k + 1) * dn) / f;
                    int64_t off = lo + (int64_t)(u01(h3(p.seed, (TAG_TGT << 32) | sid, c)) * (double)(hi - lo));
                    if (off >= hi) off = hi - 1;
                    h.pre[e] = s0 + i;
                    h.post[e] = d0 + (int32_t)off;
                    float w;
                    int8_t inh = 0;
                    if (is_drive) {
                        w = p.drive_w;
                    } else {
                        inh = u01(h3(p.seed, (TAG_INH << 32) | sid, c)) < (double)p.inh_frac;
                        const double u = u01(h3(p.seed, (TAG_W << 32) | sid, c));
                        w = inh ? (float)(-0.5 * (double)p.w0 * u) : (float)(0.7 * (double)p.w0 * u);
                    }
                    int32_t dk = 1 + (int32_t)(u01(h3(p.seed, (TAG_D << 32) | sid, c)) * p.dmax);
                    if (dk > p.dmax) dk = p.dmax;
                    h.w[e] = w;
                    h.d[e] = (float)dk * p.dt;
                    h.stype[e] = inh;
                }
            }
            o += (int64_t)sn * f;
        }
    } catch (...) {
        g_bsim_err = "out of host memory";
        return BSIM_E_NOMEM;
    }
    return h.s();
}

namespace {

/* One chain's place in the network and its projections. */
struct ChainPlan {
    int64_t first = 0;       /* first neuron id */
    int64_t nn = 0;          /* neurons */
    int64_t syn0 = 0;        /* first synapse index */
    int64_t nsyn = 0;
    bool drive = false;
    std::vector<std::pair<int64_t, int32_t>> pops;   /* (global start, count) */
    std::vector<Proj> proj;
    std::vector<int64_t> proj_off;                   /* first synapse of each projection */
};

bool chain_valid(const bsim_chain_params &p)
{
    if (p.npop < 1 || p.npop > 8 || p.fan_in < 0 || p.dmax < 1 || p.dmax > BSIM_MAX_DELAY ||
        !(p.dt > 0) || p.n_drive < 0 || p.drive_fan_in < 0)
        return false;
    for (int32_t i = 0; i < p.npop; i++)
        if (p.sizes[i] < 1) return false;
    return true;
}

/* Populations and projections of chain p placed at (first, syn0):
 * drive -> 0 and i -> i+1, with a per-pre fan-out of
 * round(fan_in * n_dst / n_src) distinct targets (clamped to [1, n_dst]). */
ChainPlan plan_chain(const bsim_chain_params &p, int64_t first, int64_t syn0)
{
    ChainPlan c;
    c.first = first;
    c.syn0 = syn0;
    c.nn = p.n_drive;
    for (int32_t i = 0; i < p.npop; i++) c.nn += p.sizes[i];
    c.drive = p.n_drive > 0 && p.drive_rate_hz > 0 && p.drive_fan_in > 0;
    if (p.n_drive > 0) c.pops.push_back({first, p.n_drive});
    int64_t at = first + p.n_drive;
    for (int32_t i = 0; i < p.npop; i++) {
        c.pops.push_back({at, p.sizes[i]});
        at += p.sizes[i];
    }
    const int32_t pl = p.n_drive > 0 ? 1 : 0;   /* index of population 0 in pops */
    if (c.drive) {
        const double k = std::round((double)p.drive_fan_in * p.sizes[0] / (double)p.n_drive);
        c.proj.push_back({0, pl, (int32_t)std::max(1.0, std::min(k, (double)p.sizes[0]))});
    }
    for (int32_t i = 0; i + 1 < p.npop && p.fan_in > 0; i++) {
        const double k = std::round((double)p.fan_in * p.sizes[i + 1] / (double)p.sizes[i]);
        c.proj.push_back({pl + i, pl + i + 1, (int32_t)std::max(1.0, std::min(k, (double)p.sizes[i + 1]))});
    }
    int64_t o = syn0;
    for (const Proj &q : c.proj) {
        c.proj_off.push_back(o);
        o += (int64_t)c.pops[q.src].second * q.f;
    }
    c.nsyn = o - syn0;
    return c;
}

/* Neuron rows of chain p (arrays already sized); draws keyed by chain-local
 * indices. */
void write_chain_neurons(bsim::HostNet &h, const bsim_chain_params &p, const ChainPlan &c,
                         int64_t arr_end)
{
    const double pr = std::min(1.0, (double)p.drive_rate_hz * (double)p.dt);
    const uint32_t thr = (uint32_t)std::min(4294967295.0, std::round(pr * 4294967296.0));
    for (int64_t j = 0; j < c.nn; j++) {
        const size_t g = (size_t)(c.first + j);
        float *row = &h.lif[g * bsim::kLifParams];
        h.arr_off[g + 1] = arr_end;
        if (j < p.n_drive) {   /* drive sources: keys 0.. of the chain's own seed */
            h.kind[g] = (int8_t)BSIM_KIND_BERNOULLI;
            std::fill(row, row + bsim::kLifParams, 0.0f);
            row[10] = 1e-3f;
            row[11] = 1e-3f;
            h.bern_p[g] = c.drive ? thr : 0u;
            h.bern_seed[g] = p.seed;
            h.bern_key[g] = (uint32_t)j;
            continue;
        }
        const int64_t li = j - p.n_drive;   /* LIF index within the chain */
        float ioff = 0.0f;
        if (p.ioff_hi > p.ioff_lo)
            ioff = (float)((double)p.ioff_lo + ((double)p.ioff_hi - (double)p.ioff_lo) *
                                                   u01(h3(p.seed, TAG_IOFF << 32, (uint64_t)li)));
        const float r[bsim::kLifParams] = {0.0f, 0.0f, 0.0f, 20e-3f, 20e-3f, p.tau_refrac,
                                           p.tau_syn_E, p.tau_syn_I, p.v_thresh, ioff,
                                           p.tau_syn_E, p.tau_syn_I};
        std::copy(r, r + bsim::kLifParams, row);
        h.kind[g] = (int8_t)BSIM_KIND_LIF;
        h.bern_p[g] = 0u;
        h.bern_seed[g] = 0ull;
        h.bern_key[g] = 0u;
    }
}

/* The f synapses of pre i of projection pi of chain p, pre-major, targets by
 * stratified sampling; every draw keyed by (projection, local counter). */
void write_pre_synapses(bsim::HostNet &h, const bsim_chain_params &p, const ChainPlan &c,
                        size_t pi, int32_t i)
{
    const Proj q = c.proj[pi];
    const int64_t s0 = c.pops[q.src].first;
    const int64_t d0 = c.pops[q.dst].first;
    const int32_t dn = c.pops[q.dst].second;
    const int32_t f = q.f;
    const bool is_drive = c.drive && pi == 0;
    const uint64_t sid = (uint64_t)pi;
    for (int32_t k = 0; k < f; k++) {
        const int64_t e = c.proj_off[pi] + (int64_t)i * f + k;
        const uint64_t cnt = (uint64_t)i * (uint64_t)f + (uint64_t)k;
        const int64_t lo = ((int64_t)k * dn) / f, hi = ((int64_t)(k + 1) * dn) / f;
        int64_t off = lo + (int64_t)(u01(h3(p.seed, (TAG_TGT << 32) | sid, cnt)) * (double)(hi - lo));
        if (off >= hi) off = hi - 1;
        h.pre[e] = (int32_t)(s0 + i);
        h.post[e] = (int32_t)(d0 + off);
        float w;
        int8_t inh = 0;
        if (is_drive) {
            w = p.drive_w;
        } else {
            inh = u01(h3(p.seed, (TAG_INH << 32) | sid, cnt)) < (double)p.inh_frac;
            const double u = u01(h3(p.seed, (TAG_W << 32) | sid, cnt));
            w = inh ? (float)(-0.5 * (double)p.w0 * u) : (float)(0.7 * (double)p.w0 * u);
        }
        int32_t dk = 1 + (int32_t)(u01(h3(p.seed, (TAG_D << 32) | sid, cnt)) * p.dmax);
        if (dk > p.dmax) dk = p.dmax;
        h.w[e] = w;
        h.d[e] = (float)dk * p.dt;
        h.stype[e] = inh;
    }
}

}  // namespace

extern "C" int32_t bsim_net_gen_chains(bsim_network *netp, const bsim_chain_params *params,
                                       int32_t count, int32_t *first_ids)
{
    if (!netp || !params || count < 0) {
        g_bsim_err = "bsim_net_gen_chains: bad arguments";
        return BSIM_E_ARG;
    }
    for (int32_t c = 0; c < count; c++)
        if (!chain_valid(params[c])) {
            g_bsim_err = "bsim_net_gen_chain: bad parameters (chain " + std::to_string(c) + ")";
            return BSIM_E_ARG;
        }
    bsim::HostNet &h = netp->net;
    try {
        std::vector<ChainPlan> plan((size_t)count);
        int64_t n0 = h.n(), s0 = h.s();
        int64_t n = n0, s = s0;
        for (int32_t c = 0; c < count; c++) {
            plan[c] = plan_chain(params[c], n, s);
            n += plan[c].nn;
            s += plan[c].nsyn;
            if (n > 0x7FFFFFFF) {
                g_bsim_err = "too many neurons";
                return BSIM_E_ARG;
            }
        }
        if (count == 0) return (int32_t)n0;
        /* every array sized once; all chains filled in parallel */
        const int64_t arr_end = h.arr_off.back();
        h.kind.resize(n);
        h.lif.resize((size_t)n * bsim::kLifParams);
        h.bern_p.resize(n);
        h.bern_seed.resize(n);
        h.bern_key.resize(n);
        h.arr_off.resize((size_t)n + 1);
        h.pre.resize(s);
        h.post.resize(s);
        h.w.resize(s);
        h.d.resize(s);
        h.stype.resize(s);
        if (!h.tau.empty()) h.tau.resize(s, 0.0f);
#pragma omp parallel for schedule(dynamic, 1)
        for (int32_t c = 0; c < count; c++) write_chain_neurons(h, params[c], plan[c], arr_end);
        /* work items: (chain, projection), each a run of pres */
        std::vector<std::pair<int32_t, int32_t>> item;
        std::vector<int64_t> item_pre0(1, 0);
        for (int32_t c = 0; c < count; c++)
            for (size_t pi = 0; pi < plan[c].proj.size(); pi++) {
                item.push_back({c, (int32_t)pi});
                item_pre0.push_back(item_pre0.back() + plan[c].pops[plan[c].proj[pi].src].second);
            }
        const int64_t npre = item_pre0.back();
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t x = 0; x < npre; x++) {
            const size_t it = (size_t)(std::upper_bound(item_pre0.begin(), item_pre0.end(), x) -
                                       item_pre0.begin()) - 1;
            const int32_t c = item[it].first, pi = item[it].second;
            write_pre_synapses(h, params[c], plan[c], (size_t)pi, (int32_t)(x - item_pre0[it]));
        }
        for (int32_t c = 0; c < count; c++)
            if (params[c].n_drive > 0) h.seed = params[c].seed;
        if (first_ids)
            for (int32_t c = 0; c < count; c++) first_ids[c] = (int32_t)plan[c].first;
        return (int32_t)n0;
    } catch (...) {
        g_bsim_err = "out of host memory";
        return BSIM_E_NOMEM;
    }
}

namespace bsim {

int chains_size(const bsim_chain_params *p, int32_t count, int64_t *neurons, int64_t *synapses,
                int32_t *dmax)
{
    int64_t n = 0, s = 0;
    int32_t d = 1;
    for (int32_t c = 0; c < count; c++) {
        if (!chain_valid(p[c])) return -1;
        const ChainPlan pl = plan_chain(p[c], 0, 0);
        n += pl.nn;
        s += pl.nsyn;
        d = std::max(d, p[c].dmax);
    }
    *neurons = n;
    *synapses = s;
    *dmax = d;
    return 0;
}

int build_chains_layout(const bsim_chain_params *p, int32_t count, float dt, int32_t tile,
                        int32_t dmax, int64_t part_synapses, Layout &out, int32_t *first_ids,
                        std::string &err,
                        const std::function<int(const Layout &, int64_t)> &upload)
{
    out = Layout();
    /* per-part progress on stderr when BSIM_PROGRESS is set (large sweeps
     * take minutes to build) */
    const char *pe = getenv("BSIM_PROGRESS");
    const bool progress = pe && pe[0] && pe[0] != '0';
    int32_t c0 = 0;
    while (c0 < count) {
        /* the next part: chains while their synapses fit part_synapses (at
         * least one chain) */
        int32_t c1 = c0;
        int64_t syn = 0;
        while (c1 < count) {
            const int64_t cs = plan_chain(p[c1], 0, 0).nsyn;
            if (c1 > c0 && syn + cs > part_synapses) break;
            syn += cs;
            c1++;
        }
        bsim_network part;
        std::vector<int32_t> first((size_t)(c1 - c0));
        if (bsim_net_gen_chains(&part, p + c0, c1 - c0, first.data()) < 0) {
            err = g_bsim_err;
            return -1;
        }
        /* inert neurons (array sources without spikes) pad the part to a
         * whole number of tiles, so that the next part starts on a tile */
        HostNet &h = part.net;
        const int32_t pad = (int32_t)((tile - h.n() % tile) % tile);
        for (int32_t i = 0; i < pad; i++) {
            h.kind.push_back((int8_t)BSIM_KIND_ARRAY);
            h.lif.insert(h.lif.end(), (size_t)kLifParams, 0.0f);
            h.arr_off.push_back(h.arr_off.back());
            h.bern_p.push_back(0u);
            h.bern_seed.push_back(0ull);
            h.bern_key.push_back(0u);
        }
        if ((int64_t)out.n + h.n() > 0x7FFFFFFF) {
            err = "bsim_sim_create_chains: more than 2^31 - 1 neurons";
            return -1;
        }
        Layout pl;
        if (build_layout(h, dt, tile, pl, err, dmax) != 0) return -1;
        part = bsim_network();          /* the host network is no longer needed */
        if (upload(pl, out.s) != 0) {
            err.clear();                /* the upload set the error message */
            return -1;
        }
        std::vector<uint16_t>().swap(pl.syn_tgt);
        std::vector<float>().swap(pl.syn_w);
        const int32_t base = out.n;
        if (append_layout(out, pl, err) != 0) return -1;
        if (first_ids)
            for (int32_t c = c0; c < c1; c++) first_ids[c] = base + first[(size_t)(c - c0)];
        if (progress)
            fprintf(stderr, "[bsim] chains %d..%d of %d laid out and uploaded: %d neurons, %lld synapses so far\n",
                    c0, c1 - 1, count, out.n, (long long)out.s);
        c0 = c1;
    }
    return 0;
}

}  // namespace bsim

extern "C" int32_t bsim_net_gen_chain(bsim_network *netp, const bsim_chain_params *gp)
{
    if (!netp || !gp || !chain_valid(*gp)) {
        g_bsim_err = "bsim_net_gen_chain: bad parameters";
        return BSIM_E_ARG;
    }
    int32_t first = 0;
    const int32_t r = bsim_net_gen_chains(netp, gp, 1, &first);
    return r < 0 ? r : first;
}

extern "C" int32_t bsim_net_export_neurons(const bsim_network *netp, int8_t *kind, float *lif,
                                           uint32_t *bern_p, uint64_t *seed)
{
    if (!netp) {
        g_bsim_err = "bad argument";
        return BSIM_E_ARG;
    }
    const bsim::HostNet &h = netp->net;
    if (!h.pois.empty()) {
        g_bsim_err = "bsim_net_export_neurons: networks with Poisson sources are not exported";
        return BSIM_E_ARG;
    }
    if (kind) std::copy(h.kind.begin(), h.kind.end(), kind);
    if (lif) std::copy(h.lif.begin(), h.lif.end(), lif);
    if (bern_p) std::copy(h.bern_p.begin(), h.bern_p.end(), bern_p);
    if (seed) *seed = h.seed;
    return h.n();
}

extern "C" int32_t bsim_net_kinds(const bsim_network *netp, int8_t *kind)
{
    if (!netp || !kind) {
        g_bsim_err = "bad argument";
        return BSIM_E_ARG;
    }
    std::copy(netp->net.kind.begin(), netp->net.kind.end(), kind);
    return netp->net.n();
}

extern "C" int64_t bsim_net_export_synapses(const bsim_network *netp, int64_t start, int64_t count,
                                            int32_t *pre, int32_t *post, float *w, float *d,
                                            int8_t *stype)
{
    if (!netp || start < 0 || count < 0 || start + count > netp->net.s()) {
        g_bsim_err = "bsim_net_export_synapses: range outside the network";
        return BSIM_E_ARG;
    }
    const bsim::HostNet &h = netp->net;
    if (pre) std::copy(h.pre.begin() + start, h.pre.begin() + start + count, pre);
    if (post) std::copy(h.post.begin() + start, h.post.begin() + start + count, post);
    if (w) std::copy(h.w.begin() + start, h.w.begin() + start + count, w);
    if (d) std::copy(h.d.begin() + start, h.d.begin() + start + count, d);
    if (stype) std::copy(h.stype.begin() + start, h.stype.begin() + start + count, stype);
    return count;
}
