/*
 * host_layout_check.cpp -- host-side checks of the layout builder, built
 * from the library's host sources with g++ (no HIP), normally under
 * AddressSanitizer and UBSan (tests/cpp/Makefile):
 *
 *   1. sweep networks laid out in parts and joined (build_chains_layout:
 *      parts padded to whole tiles, append_layout) give exactly the layout
 *      of the same padded network built at once -- every array, synapse
 *      words included (XCD residues aside: they are balanced per part);
 *   2. networks with every creation order the layout accepts (shuffled
 *      inputs, duplicate synapses, several runs per target) lay out without
 *      a memory error.
 *
 * usage: host_layout_check        (prints "host_layout_check ok" or fails)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <functional>
#include <random>
#include <string>
#include <vector>

#include "../../include/bsim_amd.h"
#include "../../bsim_amd/csrc/internal.h"
#include "../../bsim_amd/csrc/network.h"

thread_local std::string g_bsim_err;

static int fails = 0;
#define CHECK(c, ...)                                        \
    do {                                                     \
        if (!(c)) {                                          \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                    \
            fprintf(stderr, "\n");                           \
            fails++;                                         \
        }                                                    \
    } while (0)

template <class T>
static bool same(const std::vector<T> &a, const std::vector<T> &b)
{
    return a.size() == b.size() && (a.empty() || memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

static std::vector<bsim_chain_params> chains(int n, uint32_t seed)
{
    std::mt19937 r(seed);
    std::vector<bsim_chain_params> v(n);
    for (auto &p : v) {
        memset(&p, 0, sizeof(p));
        p.npop = 2 + (int)(r() % 3);
        for (int i = 0; i < p.npop; i++) p.sizes[i] = 150 + (int)(r() % 900);
        p.fan_in = 20 + (int)(r() % 100);
        p.w0 = 0.6f;
        p.inh_frac = 0.2f;
        p.dmax = 6 + (int)(r() % 12);
        p.drive_rate_hz = 20.0f;
        p.n_drive = 100;
        p.drive_fan_in = 20;
        p.drive_w = 1.0f;
        p.seed = 1000 + r();
        p.tau_refrac = 2e-3f;
        p.ioff_lo = 14.5f;
        p.ioff_hi = 16.5f;
        p.tau_syn_E = 5e-3f;
        p.tau_syn_I = 10e-3f;
        p.v_thresh = 15.0f;
        p.dt = 1e-4f;
    }
    return v;
}

static void check_parts(int nchains, int64_t part_syn, int32_t tile, uint32_t seed)
{
    const float dt = 1e-4f;
    std::vector<bsim_chain_params> p = chains(nchains, seed);
    int64_t nn = 0, ns = 0;
    int32_t dmax = 1;
    CHECK(bsim::chains_size(p.data(), nchains, &nn, &ns, &dmax) == 0, "chains_size");
    /* joined from parts; the uploaded synapses collected on the host */
    std::vector<uint16_t> tgt((size_t)ns);
    std::vector<float> w((size_t)ns);
    std::vector<int64_t> part_ends;
    auto upload = [&](const bsim::Layout &pl, int64_t s0) -> int {
        CHECK(s0 + (int64_t)pl.syn_tgt.size() <= ns, "upload beyond the synapse count");
        std::copy(pl.syn_tgt.begin(), pl.syn_tgt.end(), tgt.begin() + s0);
        std::copy(pl.syn_w.begin(), pl.syn_w.end(), w.begin() + s0);
        part_ends.push_back(s0 + (int64_t)pl.syn_tgt.size());
        return 0;
    };
    bsim::Layout J;
    std::vector<int32_t> first((size_t)nchains);
    std::string err;
    CHECK(bsim::build_chains_layout(p.data(), nchains, dt, tile, dmax, part_syn, J, first.data(), err,
                                    upload) == 0, "build_chains_layout: %s", err.c_str());
    CHECK(part_ends.size() >= 2, "expected several parts, got %zu", part_ends.size());
    J.syn_tgt = tgt;
    J.syn_w = w;
    /* the same padded network built at once */
    bsim_network one;
    int32_t c0 = 0;
    size_t part = 0;
    while (c0 < nchains) {
        int32_t c1 = c0;
        int64_t syn = 0;
        while (c1 < nchains) {
            int64_t a, b;
            int32_t d;
            bsim::chains_size(&p[c1], 1, &a, &b, &d);
            if (c1 > c0 && syn + b > part_syn) break;
            syn += b;
            c1++;
        }
        std::vector<int32_t> f((size_t)(c1 - c0));
        bsim_net_gen_chains(&one, &p[c0], c1 - c0, f.data());
        for (int32_t c = c0; c < c1; c++) CHECK(first[c] == f[c - c0], "first id of chain %d", c);
        bsim::HostNet &h = one.net;
        const int32_t pad = (int32_t)((tile - h.n() % tile) % tile);
        for (int32_t i = 0; i < pad; i++) {
            h.kind.push_back((int8_t)BSIM_KIND_ARRAY);
            h.lif.insert(h.lif.end(), (size_t)bsim::kLifParams, 0.0f);
            h.arr_off.push_back(h.arr_off.back());
            h.bern_p.push_back(0u);
            h.bern_seed.push_back(0ull);
            h.bern_key.push_back(0u);
        }
        c0 = c1;
        part++;
    }
    CHECK(part == part_ends.size(), "part count");
    bsim::Layout O;
    CHECK(bsim::build_layout(one.net, dt, tile, O, err, dmax) == 0, "build_layout: %s", err.c_str());
    CHECK(J.n == O.n && J.ntiles == O.ntiles && J.s == O.s && J.dmax == O.dmax && J.tile == O.tile,
          "sizes: n %d/%d tiles %d/%d s %lld/%lld", J.n, O.n, J.ntiles, O.ntiles, (long long)J.s,
          (long long)O.s);
    CHECK(J.ngroups == O.ngroups, "groups %d/%d", J.ngroups, O.ngroups);
#define SAME(f) CHECK(same(J.f, O.f), "array " #f " differs")
    CHECK(J.ncls.size() == O.ncls.size(), "neuron count");
    for (size_t i = 0; i < J.ncls.size() && i < O.ncls.size(); i++)   /* kinds; sets below */
        CHECK((J.ncls[i] >> 30) == (O.ncls[i] >> 30), "kind of neuron %zu", i);
    SAME(v_tmp); SAME(v_init); SAME(arr_cur0); SAME(arr_end); SAME(arr_t);
    SAME(bern_p); SAME(bern_seed); SAME(bern_key); SAME(pois_tab);
    SAME(tile_syn); SAME(syn_tgt); SAME(syn_w); SAME(tile_rng); SAME(off_tab);
    SAME(src_lo); SAME(src_hi); SAME(src_list_base); SAME(tile_base); SAME(tile_nsyn);
    SAME(chunk_tiles_off); SAME(chunk_tiles); SAME(chunk_rng); SAME(chunk_wrow);
    CHECK(J.src_list_len == O.src_list_len, "src_list_len");
    CHECK(J.rng.size() == O.rng.size(), "range entries");
    for (size_t i = 0; i < J.rng.size() && i < O.rng.size(); i++) {
        const bsim_range &a = J.rng[i], &b = O.rng[i];
        CHECK(a.lo == b.lo && a.hi == b.hi && a.off_base == b.off_base && a.lbase == b.lbase &&
                  a.uid == b.uid && a.wrow == b.wrow && a.stride == b.stride,
              "range entry %zu", i);
    }
    /* parameter sets: the joined layout keeps each part's sets; the sets of
     * every LIF neuron must hold the same constants */
    for (int32_t i = 0; i < J.n && i < O.n; i++) {
        if ((J.ncls[i] >> 30) != BSIM_KIND_LIF) continue;
        const bsim_lifc &a = J.lifc[J.ncls[i] & BSIM_CLS_SET_MASK];
        const bsim_lifc &b = O.lifc[O.ncls[i] & BSIM_CLS_SET_MASK];
        CHECK(memcmp(&a, &b, sizeof(a)) == 0, "LIF constants of neuron %d", i);
    }
    int64_t res[8] = {0};
    for (uint8_t x : J.tile_res) res[x]++;
    int64_t mx = 0;
    for (int64_t x : res) mx = x > mx ? x : mx;
    CHECK(J.nslots == 8 * mx, "nslots");
}

static void check_orders(uint32_t seed)
{
    std::mt19937 r(seed);
    for (int rep = 0; rep < 20; rep++) {
        bsim_network net;
        bsim::HostNet &h = net.net;
        const int n = 50 + (int)(r() % 400);
        for (int i = 0; i < n; i++) {
            h.kind.push_back((int8_t)BSIM_KIND_LIF);
            const float row[bsim::kLifParams] = {0, 0, 0, 0.02f, 0.02f, 1e-3f, 5e-3f, 8e-3f, 1.0f, 0.5f,
                                                 5e-3f, 8e-3f};
            h.lif.insert(h.lif.end(), row, row + bsim::kLifParams);
            h.arr_off.push_back(0);
            h.bern_p.push_back(0u);
            h.bern_seed.push_back(0ull);
            h.bern_key.push_back(0u);
        }
        const int m = (int)(r() % 20000);
        for (int k = 0; k < m; k++) {
            h.pre.push_back((int32_t)(r() % n));
            h.post.push_back((int32_t)(r() % n));
            h.w.push_back(((int)(r() % 200) - 80) * 1e-3f);
            h.d.push_back((float)(1 + r() % 9) * 1e-4f);
            h.stype.push_back((int8_t)(r() % 2));
        }
        for (int32_t tile : {192, 384}) {
            bsim::Layout L;
            std::string err;
            const int rc = bsim::build_layout(h, 1e-4f, tile, L, err);
            /* random orders may exceed the runs a target may have; either way
             * no memory error, and an accepted layout covers every synapse */
            if (rc == 0) CHECK(L.s == m && (int64_t)L.syn_tgt.size() == m, "synapse count");
            else CHECK(err.find("runs per target") != std::string::npos, "unexpected error: %s", err.c_str());
        }
    }
}

int main()
{
    check_parts(9, 60000, 192, 1);
    check_parts(14, 150000, 384, 2);
    check_parts(5, 1, 192, 3);   /* one chain per part */
    check_orders(4);
    if (fails) {
        fprintf(stderr, "host_layout_check: %d failure(s)\n", fails);
        return 1;
    }
    printf("host_layout_check ok\n");
    return 0;
}
