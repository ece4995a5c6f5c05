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

#include <algorithm>
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

extern "C" int32_t bsim_net_gen_chain(bsim_network *netp, const bsim_chain_params *gp)
{
    if (!netp || !gp || gp->npop < 1 || gp->npop > 8 || gp->fan_in < 0 || gp->dmax < 1 ||
        gp->dmax > BSIM_MAX_DELAY || !(gp->dt > 0) || gp->n_drive < 0 || gp->drive_fan_in < 0) {
        g_bsim_err = "bsim_net_gen_chain: bad parameters";
        return BSIM_E_ARG;
    }
    const bsim_chain_params p = *gp;
    bsim::HostNet &h = netp->net;
    const int64_t first = h.n();
    int64_t nn = p.n_drive;
    for (int32_t i = 0; i < p.npop; i++) {
        if (p.sizes[i] < 1) {
            g_bsim_err = "bsim_net_gen_chain: empty population";
            return BSIM_E_ARG;
        }
        nn += p.sizes[i];
    }
    if (first + nn > 0x7FFFFFFF) {
        g_bsim_err = "too many neurons";
        return BSIM_E_ARG;
    }
    const bool drive = p.n_drive > 0 && p.drive_rate_hz > 0 && p.drive_fan_in > 0;
    try {
        /* neurons: drive sources (keys 0.. of the chain's own seed), then the
         * populations; every draw is keyed by chain-local indices */
        std::vector<std::pair<int64_t, int32_t>> pops;   /* (global start, count) */
        if (p.n_drive > 0) {
            const double pr = std::min(1.0, (double)p.drive_rate_hz * (double)p.dt);
            const uint32_t thr = (uint32_t)std::min(4294967295.0, std::round(pr * 4294967296.0));
            h.kind.insert(h.kind.end(), (size_t)p.n_drive, (int8_t)BSIM_KIND_BERNOULLI);
            for (int32_t i = 0; i < p.n_drive; i++) {
                float row[bsim::kLifParams] = {0};
                row[10] = 1e-3f;
                row[11] = 1e-3f;
                h.lif.insert(h.lif.end(), row, row + bsim::kLifParams);
                h.bern_key.push_back((uint32_t)i);
            }
            h.bern_p.insert(h.bern_p.end(), (size_t)p.n_drive, drive ? thr : 0u);
            h.bern_seed.insert(h.bern_seed.end(), (size_t)p.n_drive, p.seed);
            h.arr_off.insert(h.arr_off.end(), (size_t)p.n_drive, h.arr_off.back());
            h.seed = p.seed;
        }
        const int64_t lif0 = first + p.n_drive;
        const int64_t nlif = nn - p.n_drive;
        const size_t base = h.kind.size();
        h.kind.insert(h.kind.end(), (size_t)nlif, (int8_t)BSIM_KIND_LIF);
        h.lif.resize((base + (size_t)nlif) * bsim::kLifParams);
        h.bern_p.insert(h.bern_p.end(), (size_t)nlif, 0u);
        h.bern_seed.insert(h.bern_seed.end(), (size_t)nlif, 0ull);
        h.bern_key.insert(h.bern_key.end(), (size_t)nlif, 0u);
        h.arr_off.insert(h.arr_off.end(), (size_t)nlif, h.arr_off.back());
#pragma omp parallel for schedule(static)
        for (int64_t g = 0; g < nlif; g++) {
            float *row = &h.lif[(base + (size_t)g) * bsim::kLifParams];
            float ioff = 0.0f;
            if (p.ioff_hi > p.ioff_lo)
                ioff = (float)((double)p.ioff_lo + ((double)p.ioff_hi - (double)p.ioff_lo) *
                                                       u01(h3(p.seed, TAG_IOFF << 32, (uint64_t)g)));
            const float r[bsim::kLifParams] = {0.0f, 0.0f, 0.0f, 20e-3f, 20e-3f, p.tau_refrac,
                                               p.tau_syn_E, p.tau_syn_I, p.v_thresh, ioff,
                                               p.tau_syn_E, p.tau_syn_I};
            std::copy(r, r + bsim::kLifParams, row);
        }
        if (p.n_drive > 0) pops.push_back({first, p.n_drive});
        int64_t at = lif0;
        for (int32_t i = 0; i < p.npop; i++) {
            pops.push_back({at, p.sizes[i]});
            at += p.sizes[i];
        }
        const int32_t pl = p.n_drive > 0 ? 1 : 0;   /* index of population 0 in pops */

        /* projections drive -> 0 and i -> i+1; per pre a fan-out of
         * round(fan_in * n_dst / n_src) distinct targets by stratified
         * sampling, pre-major (ascending pre for every target) */
        std::vector<Proj> proj;
        if (drive) {
            const double k = std::round((double)p.drive_fan_in * p.sizes[0] / (double)p.n_drive);
            proj.push_back({0, pl, (int32_t)std::max(1.0, std::min(k, (double)p.sizes[0]))});
        }
        for (int32_t i = 0; i + 1 < p.npop && p.fan_in > 0; i++) {
            const double k = std::round((double)p.fan_in * p.sizes[i + 1] / (double)p.sizes[i]);
            proj.push_back({pl + i, pl + i + 1, (int32_t)std::max(1.0, std::min(k, (double)p.sizes[i + 1]))});
        }
        int64_t total = 0;
        for (const Proj &q : proj) total += (int64_t)pops[q.src].second * q.f;
        const int64_t o0 = h.s();
        h.pre.resize(o0 + total);
        h.post.resize(o0 + total);
        h.w.resize(o0 + total);
        h.d.resize(o0 + total);
        h.stype.resize(o0 + total);
        if (!h.tau.empty()) h.tau.resize(o0 + total, 0.0f);
        int64_t o = o0;
        for (size_t pi = 0; pi < proj.size(); pi++) {
            const Proj q = proj[pi];
            const int64_t s0 = pops[q.src].first;
            const int32_t sn = pops[q.src].second;
            const int64_t d0 = pops[q.dst].first;
            const int32_t dn = pops[q.dst].second;
            const int32_t f = q.f;
            const bool is_drive = drive && pi == 0;
            const uint64_t sid = (uint64_t)pi;
#pragma omp parallel for schedule(static)
            for (int32_t i = 0; i < sn; i++) {
                for (int32_t k = 0; k < f; k++) {
                    const int64_t e = o + (int64_t)i * f + k;
                    const uint64_t c = (uint64_t)i * (uint64_t)f + (uint64_t)k;
                    const int64_t lo = ((int64_t)k * dn) / f, hi = ((int64_t)(k + 1) * dn) / f;
                    int64_t off = lo + (int64_t)(u01(h3(p.seed, (TAG_TGT << 32) | sid, c)) * (double)(hi - lo));
                    if (off >= hi) off = hi - 1;
                    h.pre[e] = (int32_t)(s0 + i);
                    h.post[e] = (int32_t)(d0 + off);
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
    return (int32_t)first;
}

extern "C" int32_t bsim_net_export_neurons(const bsim_network *netp, int8_t *kind, float *lif,
                                           uint32_t *bern_p, uint64_t *seed)
{
    if (!netp) {
        g_bsim_err = "bad argument";
        return BSIM_E_ARG;
    }
    const bsim::HostNet &h = netp->net;
    if (kind) std::copy(h.kind.begin(), h.kind.end(), kind);
    if (lif) std::copy(h.lif.begin(), h.lif.end(), lif);
    if (bern_p) std::copy(h.bern_p.begin(), h.bern_p.end(), bern_p);
    if (seed) *seed = h.seed;
    return h.n();
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
