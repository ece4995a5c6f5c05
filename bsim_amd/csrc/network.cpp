/*
 * network.cpp -- LIF/synapse reset and the tile-major (pre, delay) CSR build.
 * Compiled with -ffp-contract=off and without fast-math: the constants are
 * computed in float in the reference's operation order, so they carry the
 * reference's bits (expf from the host libm, as libstdc++'s exp(float)).
 */
#include "network.h"

#include <math.h>

#include <cmath>
#include <omp.h>

#include <algorithm>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <cstring>
#include <numeric>

namespace bsim {

static const float kZero = 1e-10f;  // `const real ZERO = 1e-10;` (src/utils/constant.h:20)

/* Two pres with synapses into a tile share a source range when fewer than
 * this many pres without synapses into it lie between them. */
static const int32_t kRangeGap = 256;
static const int32_t kGroupTiles = 32;   /* tiles per target group */
/* Runs of ascending (pre, delay) order per target, in creation order, that a
 * tile's layout keeps apart (segments, see build_layout). */
static const int32_t kMaxSeg = 32;

void lif_constants(const float *p, float dt, float out[9], int32_t *refrac_time)
{
    const float v_init = p[0], v_rest = p[1], v_reset = p[2], cm = p[3];
    const float tau_m = p[4], tau_refrac = p[5], tau_E = p[6], tau_I = p[7];
    const float v_thresh = p[8], i_offset = p[9];
    float rm = 1.0f;
    if (fabsf(cm) > kZero) {
        rm = tau_m / cm;
    }
    const float Cm = tau_m > 0 ? expf(-dt / tau_m) : 0.0f;
    const float CE = tau_E > 0 ? expf(-dt / tau_E) : 0.0f;
    const float CI = tau_I > 0 ? expf(-dt / tau_I) : 0.0f;
    float v_tmp = i_offset * rm + v_rest;
    v_tmp *= (1 - Cm);
    float C_E = rm * tau_E / (tau_E - tau_m);
    float C_I = rm * tau_I / (tau_I - tau_m);
    C_E = C_E * (CE - Cm);
    C_I = C_I * (CI - Cm);
    out[0] = Cm;
    out[1] = CE;
    out[2] = CI;
    out[3] = C_E;
    out[4] = C_I;
    out[5] = v_tmp;
    out[6] = v_thresh;
    out[7] = v_reset;
    out[8] = v_init;
    *refrac_time = (int32_t)(tau_refrac / dt);
}

bool synapse_reset(float w, float delay, float tau, float dt, float *w_out, int32_t *ds)
{
    const int32_t steps = (int32_t)roundf(delay / dt);
    const float resid = delay - dt * steps;
    if (!(fabsf(tau) > kZero || fabsf(resid) < kZero)) {
        return false;
    }
    if (fabsf(tau) > kZero && fabsf(resid) > kZero) {
        const float c1 = expf(-(resid) / tau);
        w = w * c1;
    }
    *w_out = w;
    *ds = steps;
    return true;
}

int32_t max_delay_steps(const HostNet &net, float dt)
{
    const int64_t S = net.s();
    int32_t dmax = 1;
#pragma omp parallel for schedule(static) reduction(max : dmax)
    for (int64_t k = 0; k < S; k++) {
        const int32_t steps = (int32_t)roundf(net.d[k] / dt);
        if (steps > dmax) dmax = steps;
    }
    return dmax;
}

int32_t default_tile(int32_t dmax)
{
    /* Each active (pre, delay) row costs one row-table vector per tile it
     * reaches; a row's synapses spread over more delays leave fewer of them
     * per tile, so deep delay rings take larger tiles (fewer, fuller rows per
     * tile) while shallow ones keep the smallest tile (most workgroups to
     * balance uneven activity).  Measured: D = 6 best at 192, D = 6..20
     * sweeps at 192..384, D = 50 at 768. */
    if (dmax <= 8) return BSIM_TILE_ALIGN;
    if (dmax <= 24) return 2 * BSIM_TILE_ALIGN;
    return 4 * BSIM_TILE_ALIGN;
}

/* Weight and delay steps of synapse k (StaticSynapse::reset); 0, or a
 * negative code with the message in *err when err is not null. */
static int reset_synapse(const HostNet &net, int64_t k, int32_t N, float dt, float *w_out,
                         int32_t *steps, std::string *err)
{
    const int32_t a = net.pre[k], b = net.post[k];
    if (a < 0 || a >= N || b < 0 || b >= N) {
        if (err) *err = "synapse " + std::to_string(k) + " has an endpoint outside the network";
        return -2;
    }
    const float tau_in = net.tau.empty() ? 0.0f : net.tau[k];
    float tau;
    if (fabsf(tau_in) > kZero) {
        tau = tau_in;
    } else {
        const float *pd = &net.lif[(size_t)b * kLifParams];
        tau = net.stype[k] ? pd[11] : pd[10];
    }
    if (!synapse_reset(net.w[k], net.d[k], tau, dt, w_out, steps)) {
        if (err)
            *err = "synapse " + std::to_string(k) +
                   ": delay is not a whole number of steps and its tau is zero";
        return -3;
    }
    if (*steps < 1 || *steps > BSIM_MAX_DELAY) {
        if (err)
            *err = "synapse " + std::to_string(k) + ": delay of " + std::to_string(*steps) +
                   " steps is outside 1.." + std::to_string(BSIM_MAX_DELAY);
        return -4;
    }
    return 0;
}

int build_layout(const HostNet &net, float dt, int32_t tile, Layout &L, std::string &err,
                 int32_t dmax_min)
{
    const int32_t N = net.n();
    const int64_t S = net.s();
    if (tile <= 0 || tile % BSIM_TILE_ALIGN != 0 || tile > BSIM_MAX_TILE) {
        err = "tile size must be a positive multiple of " + std::to_string(BSIM_TILE_ALIGN) +
              " up to " + std::to_string(BSIM_MAX_TILE);
        return -1;
    }
    if (tile > 65536) {
        err = "tile size exceeds the 16-bit local target index";
        return -1;
    }
    L = Layout();
    L.n = N;
    L.s = S;
    L.tile = tile;
    L.ntiles = N > 0 ? (N + tile - 1) / tile : 0;
    L.ncls.assign(N, 0u);
    L.lifc.clear();
    L.v_tmp.assign(N, 0.f);
    L.v_init.assign(N, 0.f);
    L.arr_cur0.assign(N, 0);
    L.arr_end.assign(N, 0);
    L.bern_p.assign(N, 0u);
    L.bern_seed.assign(N, 0ull);
    L.bern_key.assign(N, 0u);
    /* LIF parameter sets: the constants shared by neurons are stored once
     * (consecutive neurons usually share them: the previous neuron's set is
     * tried first) */
    std::unordered_map<std::string, uint32_t> set_of;
    uint32_t prev_set = 0xFFFFFFFFu;
    for (int32_t i = 0; i < N; i++) {
        if (net.kind[i] == BSIM_KIND_LIF) {
            float c[9];
            bsim_lifc k;
            lif_constants(&net.lif[(size_t)i * kLifParams], dt, c, &k.refrac_time);
            k.Cm = c[0]; k.CE = c[1]; k.CI = c[2]; k.C_E = c[3]; k.C_I = c[4];
            k.v_thresh = c[6]; k.v_reset = c[7];
            L.v_tmp[i] = c[5];
            L.v_init[i] = c[8];
            uint32_t set;
            if (prev_set != 0xFFFFFFFFu && memcmp(&L.lifc[prev_set], &k, sizeof(k)) == 0) {
                set = prev_set;
            } else {
                const std::string key(reinterpret_cast<const char *>(&k), sizeof(k));
                auto it = set_of.find(key);
                if (it == set_of.end()) {
                    if (L.lifc.size() > BSIM_CLS_SET_MASK) {
                        err = "too many distinct LIF parameter sets";
                        return -1;
                    }
                    it = set_of.emplace(key, (uint32_t)L.lifc.size()).first;
                    L.lifc.push_back(k);
                }
                set = it->second;
            }
            prev_set = set;
            L.ncls[i] = ((uint32_t)BSIM_KIND_LIF << 30) | set;
        } else if (net.kind[i] == BSIM_KIND_ARRAY) {
            L.ncls[i] = (uint32_t)BSIM_KIND_ARRAY << 30;
            L.arr_cur0[i] = (int32_t)net.arr_off[i];
            L.arr_end[i] = (int32_t)net.arr_off[i + 1];
        } else if (net.kind[i] == BSIM_KIND_BERNOULLI) {
            L.ncls[i] = (uint32_t)BSIM_KIND_BERNOULLI << 30;
            L.bern_p[i] = net.bern_p[i];
            L.bern_seed[i] = net.bern_seed[i];
            L.bern_key[i] = net.bern_key[i];
        } else if (net.kind[i] == BSIM_KIND_POISSON) {
            L.ncls[i] = (uint32_t)BSIM_KIND_POISSON << 30;
            L.bern_seed[i] = net.bern_seed[i];
            L.bern_key[i] = net.bern_key[i];
        } else {
            err = "unknown neuron kind";
            return -1;
        }
    }
    /* Poisson sources: PoissonNeuron::reset (src/neuron/poisson/PoissonNeuron.cpp:36-45)
     * -- first firing cycle int(startTime / dt) + X, end cycle int(endTime / dt);
     * the first draw uses t = 0xFFFFFFFF, a cycle the simulation never reaches.
     * One table per distinct rate. */
    {
        std::map<uint32_t, uint32_t> tab_of;     /* rate bits -> table offset */
        for (const auto &p : net.pois) {
            uint32_t bits;
            memcpy(&bits, &p.rate, 4);
            auto it = tab_of.find(bits);
            if (it == tab_of.end()) {
                if (L.pois_tab.size() > 0xFFFFFFFFull) {
                    err = "too many distinct Poisson rates";
                    return -1;
                }
                it = tab_of.emplace(bits, (uint32_t)L.pois_tab.size()).first;
                poisson_table(p.rate, L.pois_tab);
            }
            const int32_t i = p.id;
            L.bern_p[i] = it->second;
            const float c0 = p.start / dt, c1 = p.end / dt;
            if (!(c0 < 2.0e9f) || !(c1 < 2.0e9f)) {
                err = "Poisson source start or end beyond 2e9 cycles";
                return -1;
            }
            const uint32_t x = poisson_pick(&L.pois_tab[it->second],
                                            source_hash(net.bern_seed[i], net.bern_key[i], 0xFFFFFFFFu));
            L.arr_cur0[i] = (int32_t)std::min<int64_t>((int64_t)(int32_t)c0 + x, 0x7FFFFFFF);
            L.arr_end[i] = (int32_t)c1;
        }
    }
    if (net.arr_t.size() > 0x7FFFFFFF) {
        err = "too many array-source spike times";
        return -1;
    }
    L.arr_t = net.arr_t;

    /* synapse reset: weight and delay steps (parallel; on failure the first
     * failing synapse is re-run serially for its message) */
    std::vector<float> w(S);
    std::vector<uint8_t> ds(S);
    int32_t dmax = 1;
    int64_t first_bad = S;
#pragma omp parallel for schedule(static) reduction(max : dmax) reduction(min : first_bad)
    for (int64_t k = 0; k < S; k++) {
        int32_t steps;
        if (reset_synapse(net, k, N, dt, &w[k], &steps, nullptr) != 0) {
            if (k < first_bad) first_bad = k;
            continue;
        }
        ds[k] = (uint8_t)steps;
        if (steps > dmax) dmax = steps;
    }
    if (first_bad < S) {
        float wk;
        int32_t st;
        return reset_synapse(net, first_bad, N, dt, &wk, &st, &err);
    }
    if (dmax_min > BSIM_MAX_DELAY) {
        err = "delay ring deeper than " + std::to_string(BSIM_MAX_DELAY) + " steps";
        return -4;
    }
    dmax = std::max(dmax, dmax_min);
    L.dmax = dmax;

    /* bucket synapses by target tile, keeping creation order: a parallel
     * stable counting sort (per-thread counts, prefix over (tile, thread)) */
    const int32_t NT = L.ntiles;
    L.tile_syn.assign((size_t)NT + 1, 0);
    /* creation index of each synapse in tile-bucket order (not value-
     * initialised: every entry is written below) */
    std::unique_ptr<int64_t[]> order_buf(new int64_t[(size_t)std::max<int64_t>(S, 1)]);
    int64_t *order = order_buf.get();
    /* post / tile without a division: a double reciprocal, corrected */
    const double inv_tile = 1.0 / (double)tile;
    auto tile_of = [tile, inv_tile](int32_t post) {
        int64_t t = (int64_t)((double)post * inv_tile);
        if (t * tile > post) t--;
        else if ((t + 1) * tile <= post) t++;
        return (int32_t)t;
    };
    {
        const int nmax = omp_get_max_threads();
        std::vector<int64_t> cnt((size_t)nmax * NT, 0);
        int nth = 1;
#pragma omp parallel
        {
            const int id = omp_get_thread_num(), T = omp_get_num_threads();
#pragma omp single
            nth = T;
            const int64_t k0 = S * id / T, k1 = S * (id + 1) / T;
            int64_t *c = cnt.data() + (size_t)id * NT;
            for (int64_t k = k0; k < k1; k++) c[tile_of(net.post[k])]++;
#pragma omp barrier
#pragma omp single
            {
                int64_t run = 0;
                for (int32_t t = 0; t < NT; t++) {
                    L.tile_syn[t] = run;
                    for (int j = 0; j < nth; j++) {
                        const int64_t x = cnt[(size_t)j * NT + t];
                        cnt[(size_t)j * NT + t] = run;
                        run += x;
                    }
                }
                L.tile_syn[NT] = run;
            }
            for (int64_t k = k0; k < k1; k++) order[c[tile_of(net.post[k])]++] = k;
        }
    }
    for (int32_t t = 0; t < NT; t++) {
        if (L.tile_syn[t + 1] - L.tile_syn[t] >= ((int64_t)1 << 32)) {
            err = "a tile receives 2^32 or more synapses; use a smaller tile";
            return -6;
        }
    }
    L.syn_tgt.resize(S);
    L.syn_w.resize(S);
    L.tile_rng.assign((size_t)NT + 1, 0);
    std::vector<std::vector<bsim_range>> trng(NT);
    std::vector<std::vector<uint32_t>> toff(NT);
    const int32_t D = dmax;
    const int32_t NW = BSIM_NWAVE;
    const int32_t qw = tile / NW;   /* targets per part */
    int fail = 0;
    int64_t first_unordered = S;

    /* Summation order.  The reference sums a target's inputs of a cycle in
     * synapse creation order (Network.cpp:324-327).  Each target's incoming
     * synapses, taken in creation order, are cut into runs at every step where
     * (pre, delay) descends; the run index of a synapse is its segment (0 for
     * a target whose inputs were created pre by pre; 1 from the second
     * `connect` call on when a target is fed first by a later population and
     * then by an earlier one, as test/gpu/standard_test.cpp:116-121 does).
     * The tile's synapses are sorted by (segment, pre, delay, part, creation)
     * and its source ranges are formed per segment, so the stream order of a
     * target's inputs -- segment by segment, (pre, delay) ascending inside
     * one -- is its creation order, for any network.  A segment of a tile
     * costs its own source ranges and row-table vectors, so a tile whose
     * targets have more than kMaxSeg segments is rejected (the first synapse
     * that opens segment kMaxSeg is reported). */
#pragma omp parallel for schedule(dynamic, 1) reduction(min : first_unordered) reduction(max : fail)
    for (int32_t t = 0; t < NT; t++) {
        const int64_t lo = L.tile_syn[t], hi = L.tile_syn[t + 1], m = hi - lo;
        /* row key: segment (5 bits) | pre (31 bits) | delay (6 bits) |
         * part (2 bits); ties keep the creation rank */
        struct Ent { uint64_t key; uint32_t rank; };
        std::vector<Ent> ent(m);
        std::vector<int64_t> last(tile, -1);
        std::vector<uint8_t> seg(tile, 0);
        for (int64_t e = 0; e < m; e++) {
            const int64_t k = order[lo + e];
            const int32_t lt = net.post[k] - t * tile;
            const int64_t okey = ((int64_t)net.pre[k] << 8) | ds[k];
            if (okey < last[lt]) {
                if (seg[lt] + 1 >= kMaxSeg) {
                    if (k < first_unordered) first_unordered = k;
                } else {
                    seg[lt]++;
                }
            }
            last[lt] = okey;
            const uint32_t q = (uint32_t)(lt / qw);
            ent[e].key = ((uint64_t)seg[lt] << 39) | ((uint64_t)(uint32_t)net.pre[k] << 8) |
                         ((uint64_t)ds[k] << 2) | q;
            ent[e].rank = (uint32_t)e;
        }
        std::sort(ent.begin(), ent.end(), [](const Ent &x, const Ent &y) {
            return x.key < y.key || (x.key == y.key && x.rank < y.rank);
        });
        for (int64_t e = 0; e < m; e++) {
            const int64_t k = order[lo + ent[e].rank];
            L.syn_tgt[lo + e] = (uint16_t)(net.post[k] - t * tile);
            L.syn_w[lo + e] = w[k];
        }
        /* source ranges, per segment: runs of pres whose gaps are at most
         * kRangeGap */
        int64_t e = 0;
        while (e < m) {
            const uint64_t sg = ent[e].key >> 39;
            const int32_t p0 = (int32_t)((ent[e].key >> 8) & 0x7FFFFFFFu);
            int64_t f = e;
            int32_t plast = p0;
            while (f < m && (ent[f].key >> 39) == sg) {
                const int32_t pf = (int32_t)((ent[f].key >> 8) & 0x7FFFFFFFu);
                if (pf - plast > kRangeGap || pf - p0 >= BSIM_MAX_RANGE) break;
                plast = pf;
                f++;
            }
            bsim_range r;
            r.lo = p0;
            r.hi = plast + 1;
            r.off_base = (int64_t)toff[t].size();
            r.uid = 0;
            r.wrow = 0;
            r.stride = 0;
            r.pad = 0;
            const int64_t cnt = (int64_t)(r.hi - r.lo) * D * NW + 1;
            const size_t base = toff[t].size();
            toff[t].resize(base + cnt);
            int64_t pos = e;
            size_t x = base;
            for (int32_t p = r.lo; p < r.hi; p++) {
                for (int32_t d = 1; d <= D; d++) {
                    for (int32_t q = 0; q < NW; q++) {
                        const uint64_t key = (sg << 39) | ((uint64_t)(uint32_t)p << 8) |
                                             ((uint64_t)d << 2) | (uint64_t)q;
                        toff[t][x++] = (uint32_t)pos;
                        while (pos < f && ent[pos].key == key) pos++;
                    }
                }
            }
            toff[t][base + cnt - 1] = (uint32_t)pos;
            if (pos != f) fail = 1;
            trng[t].push_back(r);
            e = f;
        }
    }
    if (first_unordered < S) {
        const int64_t k = first_unordered;
        err = "synapse " + std::to_string(k) + " (" + std::to_string(net.pre[k]) + " -> " +
              std::to_string(net.post[k]) + ") would open run " + std::to_string(kMaxSeg + 1) +
              " of ascending (pre, delay) order among its target's inputs in creation "
              "order; the GPU layout keeps at most " + std::to_string(kMaxSeg) +
              " such runs per target (each costs its own row tables)";
        return -5;
    }
    if (fail) {
        err = "internal error: CSR row walk mismatch";
        return -7;
    }

    /* Target groups: runs of at most kGroupTiles consecutive tiles with the
     * same source ranges.  A group's synapses (the same block of the synapse
     * arrays as its tiles' in tile-major order) are stored pre-major:
     * (range, pre, delay, tile, part, creation), and its row table per range
     * holds the part starts of every tile for each (pre, delay) row
     * consecutively, so that the tiles of a group read shared cache lines. */
    std::vector<int32_t> gstart;
    for (int32_t t = 0; t < NT; t++) {
        bool same = t > 0 && t - gstart.back() < kGroupTiles && trng[t].size() == trng[t - 1].size();
        for (size_t i = 0; same && i < trng[t].size(); i++)
            same = trng[t][i].lo == trng[t - 1][i].lo && trng[t][i].hi == trng[t - 1][i].hi;
        if (!same) gstart.push_back(t);
    }
    const int32_t NG = (int32_t)gstart.size();
    gstart.push_back(NT);
    for (int32_t g = 0; g < NG; g++) {
        if (L.tile_syn[gstart[g + 1]] - L.tile_syn[gstart[g]] >= ((int64_t)1 << 32)) {
            err = "a tile group receives 2^32 or more synapses; use a smaller tile";
            return -6;
        }
    }
    /* group row tables: per group and range, entry
     * ((p - lo) * D + d - 1) * nk * RW + k * RW + q (RW = NW + 1) holds the
     * start of part q of tile k's sub-rows of row (p, d), entry q = NW its
     * end: one aligned 16-byte vector per tile and row.  Offsets relative to
     * the group's first synapse. */
    const int32_t RW = NW + 1;
    static_assert(BSIM_NWAVE + 1 == 4, "a tile's row-table entry is one 16-byte vector");
    std::vector<std::vector<int64_t>> gtab_base(NG);
    std::vector<int64_t> gtab_off((size_t)NG + 1, 0);
    for (int32_t g = 0; g < NG; g++) {
        const int32_t nk = gstart[g + 1] - gstart[g];
        int64_t n = 0;
        for (const auto &r : trng[gstart[g]]) {
            gtab_base[g].push_back(n);
            n += (int64_t)(r.hi - r.lo) * D * nk * RW;
        }
        gtab_off[g + 1] = gtab_off[g] + n;
    }
    L.off_tab.assign((size_t)gtab_off[NG], 0u);
    std::vector<uint16_t> gtgt(S);
    std::vector<float> gw(S);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t g = 0; g < NG; g++) {
        const int32_t k0 = gstart[g], nk = gstart[g + 1] - k0;
        const int64_t gs = L.tile_syn[k0];
        uint32_t *tab = L.off_tab.data() + gtab_off[g];
        uint32_t pos = 0;
        for (size_t ri = 0; ri < trng[k0].size(); ri++) {
            const bsim_range &r0 = trng[k0][ri];
            uint32_t *gt = tab + gtab_base[g][ri];
            size_t x = 0;
            for (int32_t p = r0.lo; p < r0.hi; p++)
                for (int32_t d = 1; d <= D; d++)
                    for (int32_t kk = 0; kk < nk; kk++) {
                        const int32_t t = k0 + kk;
                        const size_t e = (size_t)trng[t][ri].off_base +
                                         ((size_t)(p - r0.lo) * D + (size_t)(d - 1)) * NW;
                        for (int32_t q = 0; q < NW; q++) {
                            gt[x++] = pos;
                            const uint32_t a = toff[t][e + q], b = toff[t][e + q + 1];
                            for (uint32_t i = a; i < b; i++) {
                                gtgt[gs + pos] = L.syn_tgt[L.tile_syn[t] + i];
                                gw[gs + pos] = L.syn_w[L.tile_syn[t] + i];
                                pos++;
                            }
                        }
                        gt[x++] = pos;
                    }
        }
    }
    L.syn_tgt.swap(gtgt);
    L.syn_w.swap(gw);
    std::vector<uint16_t>().swap(gtgt);
    std::vector<float>().swap(gw);
    L.ngroups = NG;
    L.tile_base.resize(NT);
    L.tile_nsyn.resize(NT);
    L.tile_res.resize(NT);
    /* XCD residue of each group (workgroups are dispatched to the 8 XCDs
     * round-robin by workgroup id): groups are spread over the residues to
     * balance tile counts, so that a group's tiles share one L2 */
    {
        std::vector<int64_t> rcount(8, 0);
        for (int32_t g = 0; g < NG; g++) {
            const int32_t nk = gstart[g + 1] - gstart[g];
            const int x = (int)(std::min_element(rcount.begin(), rcount.end()) - rcount.begin());
            rcount[x] += nk;
            for (int32_t t = gstart[g]; t < gstart[g + 1]; t++) L.tile_res[t] = (uint8_t)x;
        }
        L.nslots = (int32_t)(8 * *std::max_element(rcount.begin(), rcount.end()));
    }

    size_t nr = 0;
    for (int32_t t = 0; t < NT; t++) nr += trng[t].size();
    L.rng.reserve(nr);
    /* distinct source ranges, each cut into chunks of BSIM_CHUNK pres; a
     * tile's range becomes one entry per chunk */
    std::map<std::pair<int32_t, int32_t>, int32_t> uid;
    std::vector<int32_t> tile_nrng(NT, 0);
    for (int32_t g = 0; g < NG; g++) {
        const int32_t k0 = gstart[g], nk = gstart[g + 1] - k0;
        for (int32_t t = k0; t < k0 + nk; t++) {
            L.tile_base[t] = L.tile_syn[k0];
            L.tile_nsyn[t] = (uint32_t)(L.tile_syn[k0 + nk] - L.tile_syn[k0]);
            for (size_t ri = 0; ri < trng[t].size(); ri++) {
                const bsim_range &r = trng[t][ri];
                auto it = uid.find({r.lo, r.hi});
                if (it == uid.end()) {
                    it = uid.emplace(std::make_pair(r.lo, r.hi), (int32_t)L.src_lo.size()).first;
                    for (int32_t c = r.lo; c < r.hi; c += BSIM_CHUNK) {
                        const int32_t ce = std::min(r.hi, c + BSIM_CHUNK);
                        L.src_lo.push_back(c);
                        L.src_hi.push_back(ce);
                        L.src_list_base.push_back(L.src_list_len);
                        L.src_list_len += (int64_t)(ce - c) * D;
                    }
                }
                int32_t u = it->second;
                for (int32_t c = r.lo; c < r.hi; c += BSIM_CHUNK, u++) {
                    bsim_range rc;
                    rc.lo = c;
                    rc.hi = std::min(r.hi, c + BSIM_CHUNK);
                    rc.stride = (uint32_t)(nk * RW);
                    rc.off_base = gtab_off[g] + gtab_base[g][ri] +
                                  (int64_t)(c - r.lo) * D * rc.stride + (int64_t)(t - k0) * RW;
                    rc.uid = u;
                    rc.lbase = L.src_list_base[(size_t)u];
                    rc.pad = 0;
                    /* synapses of the tile from the chunk, per (pre, delay) row */
                    const size_t e0 = (size_t)(r.off_base + (int64_t)(c - r.lo) * D * NW);
                    const size_t e1 = e0 + (size_t)(rc.hi - rc.lo) * D * NW;
                    const double per_row = (double)(toff[t][e1] - toff[t][e0]) / ((double)(rc.hi - rc.lo) * D);
                    rc.wrow = (uint32_t)std::min(per_row * 16.0 + 0.5, 4.0e9);
                    L.rng.push_back(rc);
                    tile_nrng[t]++;
                }
            }
        }
    }
    for (int32_t t = 0; t < NT; t++) {
        std::vector<bsim_range>().swap(trng[t]);
        std::vector<uint32_t>().swap(toff[t]);
    }
    for (int32_t t = 0; t < NT; t++) L.tile_rng[t + 1] = L.tile_rng[t] + tile_nrng[t];
    /* reverse index: for each chunk, the tiles that read it */
    const size_t nch = L.src_lo.size();
    L.chunk_tiles_off.assign(nch + 1, 0);
    for (const auto &r : L.rng) L.chunk_tiles_off[(size_t)r.uid + 1]++;
    for (size_t u = 0; u < nch; u++) L.chunk_tiles_off[u + 1] += L.chunk_tiles_off[u];
    L.chunk_tiles.resize(L.rng.size());
    L.chunk_rng.resize(L.rng.size());
    L.chunk_wrow.resize(L.rng.size());
    {
        std::vector<int64_t> fill(L.chunk_tiles_off.begin(), L.chunk_tiles_off.end() - 1);
        for (int32_t t = 0; t < NT; t++)
            for (int32_t ri = L.tile_rng[t]; ri < L.tile_rng[t + 1]; ri++) {
                const int64_t x = fill[L.rng[ri].uid]++;
                L.chunk_tiles[x] = t;
                L.chunk_rng[x] = ri;
                L.chunk_wrow[x] = L.rng[ri].wrow;
            }
    }
    return 0;
}

int append_layout(Layout &A, Layout &B, std::string &err)
{
    if (A.ntiles > 0 && (A.tile != B.tile || A.dmax != B.dmax)) {
        err = "append_layout: parts differ in tile size or delay ring depth";
        return -1;
    }
    if (A.ntiles > 0 && A.n != A.ntiles * A.tile) {
        err = "append_layout: the first part does not end on a tile boundary";
        return -1;
    }
    if ((int64_t)A.n + B.n > 0x7FFFFFFF) {
        err = "append_layout: more than 2^31 - 1 neurons";
        return -1;
    }
    if (A.ntiles == 0) {
        const int32_t pn = A.n;
        A = std::move(B);
        if (pn != 0) {
            err = "append_layout: neurons without tiles";
            return -1;
        }
        return 0;
    }
    const int32_t n0 = A.n, t0 = A.ntiles;
    const int64_t s0 = A.s;
    const uint32_t c0 = (uint32_t)A.lifc.size();
    const int64_t a0 = (int64_t)A.arr_t.size();
    const uint32_t p0 = (uint32_t)A.pois_tab.size();
    const int64_t r0 = (int64_t)A.rng.size();
    const int64_t o0 = (int64_t)A.off_tab.size();
    const int32_t u0 = (int32_t)A.src_lo.size();
    const int64_t l0 = A.src_list_len;
    const int64_t k0 = (int64_t)A.chunk_tiles.size();
    if (a0 + (int64_t)B.arr_t.size() > 0x7FFFFFFF || (uint64_t)p0 + B.pois_tab.size() > 0xFFFFFFFFull ||
        (uint64_t)c0 + B.lifc.size() > BSIM_CLS_SET_MASK) {
        err = "append_layout: source tables too large";
        return -1;
    }
    for (int32_t i = 0; i < B.n; i++) {
        const uint32_t kind = B.ncls[i] >> 30;
        uint32_t cls = B.ncls[i];
        int32_t cur = B.arr_cur0[i], end = B.arr_end[i];
        uint32_t bp = B.bern_p[i];
        if (kind == BSIM_KIND_LIF) {
            cls += c0;
        } else if (kind == BSIM_KIND_ARRAY) {
            cur += (int32_t)a0;
            end += (int32_t)a0;
        } else if (kind == BSIM_KIND_POISSON) {
            bp += p0;   /* arr_cur0 / arr_end of a Poisson source are cycles */
        }
        A.ncls.push_back(cls);
        A.arr_cur0.push_back(cur);
        A.arr_end.push_back(end);
        A.bern_p.push_back(bp);
    }
    A.lifc.insert(A.lifc.end(), B.lifc.begin(), B.lifc.end());
    A.v_tmp.insert(A.v_tmp.end(), B.v_tmp.begin(), B.v_tmp.end());
    A.v_init.insert(A.v_init.end(), B.v_init.begin(), B.v_init.end());
    A.arr_t.insert(A.arr_t.end(), B.arr_t.begin(), B.arr_t.end());
    A.bern_seed.insert(A.bern_seed.end(), B.bern_seed.begin(), B.bern_seed.end());
    A.bern_key.insert(A.bern_key.end(), B.bern_key.begin(), B.bern_key.end());
    A.pois_tab.insert(A.pois_tab.end(), B.pois_tab.begin(), B.pois_tab.end());
    /* tiles */
    A.tile_syn.pop_back();
    for (size_t t = 0; t < B.tile_syn.size(); t++) A.tile_syn.push_back(B.tile_syn[t] + s0);
    for (int32_t t = 0; t < B.ntiles; t++) {
        A.tile_base.push_back(B.tile_base[t] + s0);
        A.tile_nsyn.push_back(B.tile_nsyn[t]);
        A.tile_res.push_back(B.tile_res[t]);
    }
    A.syn_tgt.insert(A.syn_tgt.end(), B.syn_tgt.begin(), B.syn_tgt.end());
    A.syn_w.insert(A.syn_w.end(), B.syn_w.begin(), B.syn_w.end());
    /* source ranges, row tables (group-relative offsets: copied), chunks */
    A.tile_rng.pop_back();
    for (size_t t = 0; t < B.tile_rng.size(); t++) A.tile_rng.push_back(B.tile_rng[t] + (int32_t)r0);
    for (bsim_range r : B.rng) {
        r.lo += n0;
        r.hi += n0;
        r.off_base += o0;
        r.uid += u0;
        r.lbase += l0;
        A.rng.push_back(r);
    }
    A.off_tab.insert(A.off_tab.end(), B.off_tab.begin(), B.off_tab.end());
    for (size_t u = 0; u < B.src_lo.size(); u++) {
        A.src_lo.push_back(B.src_lo[u] + n0);
        A.src_hi.push_back(B.src_hi[u] + n0);
        A.src_list_base.push_back(B.src_list_base[u] + l0);
    }
    A.src_list_len += B.src_list_len;
    A.chunk_tiles_off.pop_back();
    for (size_t u = 0; u < B.chunk_tiles_off.size(); u++) A.chunk_tiles_off.push_back(B.chunk_tiles_off[u] + k0);
    for (int32_t t : B.chunk_tiles) A.chunk_tiles.push_back(t + t0);
    for (int32_t x : B.chunk_rng) A.chunk_rng.push_back(x + (int32_t)r0);
    A.chunk_wrow.insert(A.chunk_wrow.end(), B.chunk_wrow.begin(), B.chunk_wrow.end());
    A.n += B.n;
    A.ntiles += B.ntiles;
    A.s += B.s;
    A.ngroups += B.ngroups;
    std::vector<int64_t> rcount(8, 0);
    for (uint8_t x : A.tile_res) rcount[x]++;
    A.nslots = (int32_t)(8 * *std::max_element(rcount.begin(), rcount.end()));
    B = Layout();
    return 0;
}

void poisson_table(float rate, std::vector<uint32_t> &tab)
{
    /* P(X = j) = exp(-lam + j log lam - lgamma(j + 1)), summed in double; the
     * table ends where the remaining mass is below 2^-32 for any rate */
    const double lam = (double)rate;
    const uint32_t J = (uint32_t)std::ceil(lam + 12.0 * std::sqrt(lam) + 16.0);
    const size_t at = tab.size();
    tab.resize(at + 1 + J);
    tab[at] = J;
    double cdf = 0.0;
    for (uint32_t j = 0; j < J; j++) {
        const double lp = (lam > 0.0) ? -lam + (double)j * std::log(lam) - std::lgamma((double)j + 1.0)
                                      : (j == 0 ? 0.0 : -INFINITY);
        cdf += std::exp(lp);
        const double q = std::floor(cdf * 4294967296.0);
        tab[at + 1 + j] = q >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)q;
    }
}

uint32_t poisson_pick(const uint32_t *tab, uint32_t u)
{
    uint32_t lo = 0, hi = tab[0] - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (u < tab[1 + mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

uint32_t source_hash(uint64_t seed, uint32_t key, uint32_t t)
{
    uint64_t z = seed ^ (((uint64_t)key << 32) | (uint64_t)t);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 32);
}

}  // namespace bsim
