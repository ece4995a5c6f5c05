/*
 * step.hip -- the simulation step on CDNA4 (gfx950).
 *
 * One launch of bsim_step_kernel advances the network by one cycle.  Each
 * 256-thread workgroup (four wavefronts) owns one tile of TILE consecutive
 * neurons (layout.h) and performs, for that tile:
 *
 *   (ii) phase (ii) of cycle t-1: every synapse into the tile whose pre fired
 *        d cycles earlier adds its weight to the target's E or I accumulator
 *        (LDS), in the order the reference's CPU simulator applies them
 *        (src/net/Network.cpp:324-327, src/neuron/lif/LIFENeuron.cpp:107-116);
 *   (i)  the LIF_brian update of cycle t (src/neuron/lif/LIFENeuron.cpp:169-215),
 *        which consumes exactly those accumulators, the spike history update
 *        and the ordered compaction of the tile's fired / active neurons into
 *        64-bit words by wavefront ballot.
 *
 * Phase (ii) pipeline for a tile:
 *   items     the tile's active pres, ascending: the "active" bitmap words of
 *             its source ranges are scanned by all 256 threads (popcount +
 *             block prefix) -- one round trip to memory;
 *   segments  for each item the history word gives its active delays; each
 *             (pre, delay) row of the tile is one segment [start, start+len)
 *             of the tile's synapse arrays, read from the row table -- two
 *             round trips;
 *   events    the concatenated segments are streamed in chunks of 512 events;
 *             the loads of chunk c+1 are in flight while chunk c is applied;
 *   apply     the events of a chunk are split, in stream order, into four
 *             queues by target quarter; wavefront q applies queue q to its own
 *             accumulators 64 events at a time.  Lanes whose accumulator also
 *             occurs in a lower lane wait: an LDS atomic-min on a per-
 *             accumulator tag elects the lowest lane, which adds; the others
 *             are carried, in order, to the front of the next batch.
 * Every accumulator thus receives its adds one at a time in stream order,
 * (pre, delay, creation) -- the creation order the host verified -- so the
 * sums are the reference's sequential fp32 sums, bit for bit.
 *
 * Arithmetic: compiled with -ffp-contract=off (no FMA contraction) and
 * without denormal flushing, so V = ((Cm*V + v_tmp) + iE*C_E) + iI*C_I rounds
 * exactly as the reference's CPU code.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

#define WAVE 64
#define NWAVE 4
#define NTHREAD (WAVE * NWAVE)
#define ITEMCAP 512      /* items staged per pass */
#define SEGCAP 512       /* segments staged per pass */
#define CHUNK 512        /* events per chunk: 2 per thread */
#define SUBW (CHUNK / NWAVE)   /* events of a chunk loaded by one wavefront */
#define BCAP 1024        /* boundary table: positions per part = BCAP * SUBW */
#define TAG_TOP 0x03FFFFFFu

/* fixed LDS bytes besides the 16 bytes per target of acc + tag */
#define LDS_FIXED (ITEMCAP * 4 + SEGCAP * 12 + BCAP * 2 + NWAVE * CHUNK * 8 + \
                   16 * 4 * 4 + 16 * 4 * 4 + NWAVE * WAVE * 8 + 64)

__device__ __forceinline__ uint32_t bern_hash(uint64_t seed, uint32_t id, uint32_t t)
{
    uint64_t z = seed ^ (((uint64_t)id << 32) | (uint64_t)t);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 32);
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

/* number of set bits of m below this lane */
__device__ __forceinline__ uint32_t rank_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t readlane(uint32_t x, int32_t l)
{
    return __builtin_amdgcn_readlane(x, l);
}

/* inclusive prefix sum over the wavefront */
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    const int lane = threadIdx.x & (WAVE - 1);
    uint32_t v = x;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(v, off, WAVE);
        if (lane >= off) v += y;
    }
    return v;
}

/* exclusive prefix sum over the workgroup; *total = sum over all threads.
 * scr: NWAVE uint32 of LDS.  Contains two barriers. */
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *total, uint32_t *scr)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int w = threadIdx.x / WAVE;
    const uint32_t v = wave_incl_scan(x);
    if (lane == WAVE - 1) scr[w] = v;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NWAVE; i++) {
        const uint32_t s = scr[i];
        before += (i < w) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return before + v - x;
}

/* Per-wavefront accumulation state (registers). */
struct Acc {
    uint32_t round, ncarry, ckey;
    float cw;
};

/* Applies up to 64 events (lanes below a.ncarry hold carried events, lanes
 * flagged `fresh` new ones) to the accumulators; events that must wait are
 * left, in order, in lanes [0, a.ncarry). */
__device__ __forceinline__ void apply_batch(float *acc, uint32_t *tag, uint2 *carry_buf, bool fresh,
                                            uint32_t nkey, float nw, Acc &a)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const bool valid = fresh || lane < a.ncarry;
    const uint32_t key = fresh ? nkey : a.ckey;
    const float w = fresh ? nw : a.cw;
    const uint32_t tv = ((TAG_TOP - a.round) << 6) | lane;
    a.round++;
    if (valid) atomicMin(&tag[key], tv);
    __builtin_amdgcn_wave_barrier();
    bool leader = false;
    if (valid) leader = (tag[key] == tv);
    if (leader) {
        const float s = acc[key];
        acc[key] = s + w;
    }
    __builtin_amdgcn_wave_barrier();
    const uint64_t lose = ballot64(valid && !leader);
    if (lose) {
        if (valid && !leader) carry_buf[rank_below(lose)] = make_uint2(key, __float_as_uint(w));
        __builtin_amdgcn_wave_barrier();
        a.ncarry = (uint32_t)__popcll(lose);
        if (lane < a.ncarry) {
            const uint2 c = carry_buf[lane];
            a.ckey = c.x;
            a.cw = __uint_as_float(c.y);
        }
        __builtin_amdgcn_wave_barrier();
    } else {
        a.ncarry = 0;
    }
}

struct Shared {
    float *acc;          /* [2T]: E at 2i, I at 2i+1 */
    uint32_t *tag;       /* [2T] */
    uint32_t *items;     /* [ITEMCAP] pre ids */
    uint32_t *sst, *sln, *spos;   /* [SEGCAP] segment start / length / stream position */
    uint16_t *bseg;      /* [BCAP] segment holding position part0 + b * SUBW */
    uint2 *queue;        /* [NWAVE][CHUNK] (key, weight bits) */
    uint32_t *gcnt;      /* [16][4] events per (load group, owner) */
    uint32_t *gbase;     /* [16][4] exclusive prefix over load groups */
    uint2 *carry;        /* [NWAVE][WAVE] */
    uint32_t *scr;       /* [16] scratch */
};

/* Event loader state of one chunk: two events per thread. */
struct Chunk {
    uint32_t tl[2];
    float w[2];
    bool v[2];
};

/* Issues the loads of chunk c of the current part: wavefront w loads stream
 * positions [pos0 + w * SUBW, + SUBW), in two 64-lane groups. */
__device__ __forceinline__ void load_chunk(const Shared &S, Chunk &ch, uint32_t pos0, uint32_t pend,
                                           uint32_t part0, int32_t nseg, const uint16_t *tgt,
                                           const float *wts)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t w = threadIdx.x / WAVE;
    const uint32_t base = pos0 + w * SUBW;
    ch.v[0] = ch.v[1] = false;
    ch.tl[0] = ch.tl[1] = 0;
    ch.w[0] = ch.w[1] = 0.0f;
    if (base >= pend) return;                       /* uniform per wavefront */
    int32_t wb = (int32_t)S.bseg[(base - part0) / SUBW];   /* first segment of the window */
    uint32_t st = 0, ps = 0, ln = 0;
    if (wb + (int32_t)lane < nseg) {
        st = S.sst[wb + lane];
        ps = S.spos[wb + lane];
        ln = S.sln[wb + lane];
    }
    int32_t jj = 0;
    uint32_t addr[2] = {0, 0};
#pragma unroll
    for (int g = 0; g < 2; g++) {
        const uint32_t qlo = base + g * WAVE;
        const uint32_t qhi = min(qlo + WAVE, pend);
        const uint32_t q = qlo + lane;
        const bool v = qlo < pend && q < qhi;
        ch.v[g] = v;
        if (qlo >= pend) continue;                  /* uniform */
        while (true) {
            if (jj >= WAVE) {                       /* slide the register window */
                wb += WAVE;
                jj = 0;
                st = ps = ln = 0;
                if (wb + (int32_t)lane < nseg) {
                    st = S.sst[wb + lane];
                    ps = S.spos[wb + lane];
                    ln = S.sln[wb + lane];
                }
            }
            if (wb + jj >= nseg) break;
            const uint32_t e = readlane(ps, jj);
            if (e >= qhi) break;
            const uint32_t l = readlane(ln, jj);
            if (v && q >= e && q < e + l) addr[g] = readlane(st, jj) + (q - e);
            if (e + l > qhi) break;                 /* continues into the next group */
            jj++;
        }
    }
#pragma unroll
    for (int g = 0; g < 2; g++) {
        if (ch.v[g]) {
            ch.tl[g] = tgt[addr[g]];
            ch.w[g] = wts[addr[g]];
        }
    }
}

/* Splits a loaded chunk into the four owner queues in stream order and has
 * every wavefront apply its queue.  Owner of local target tl: tl / tw. */
__device__ __forceinline__ void apply_chunk(const Shared &S, const Chunk &ch, uint32_t tw, Acc &a)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & (WAVE - 1);
    const uint32_t w = tid / WAVE;
    uint32_t own[2], key[2];
    uint64_t b0[2], b1[2], b2[2], b3[2];   /* per owner ballots (named: no indexed array) */
#pragma unroll
    for (int g = 0; g < 2; g++) {
        own[g] = ch.v[g] ? ch.tl[g] / tw : NWAVE;
        key[g] = 2u * ch.tl[g] + ((ch.w[g] >= 0.0f) ? 0u : 1u);
        b0[g] = ballot64(own[g] == 0u);
        b1[g] = ballot64(own[g] == 1u);
        b2[g] = ballot64(own[g] == 2u);
        b3[g] = ballot64(own[g] == 3u);
        if (lane < NWAVE) {
            /* lane o writes the count of owner o for load group (w, g) */
            const uint64_t b = lane == 0 ? b0[g] : lane == 1 ? b1[g] : lane == 2 ? b2[g] : b3[g];
            S.gcnt[(w * 2 + g) * NWAVE + lane] = (uint32_t)__popcll(b);
        }
    }
    __syncthreads();
    if (tid < 8 * NWAVE) {
        /* thread (G, o): prefix of owner o over load groups before G */
        const uint32_t G = tid / NWAVE, o = tid % NWAVE;
        uint32_t s = 0;
        for (uint32_t x = 0; x < G; x++) s += S.gcnt[x * NWAVE + o];
        S.gbase[G * NWAVE + o] = s;
        if (G == 7) S.scr[8 + o] = s + S.gcnt[7 * NWAVE + o];   /* queue length */
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < 2; g++) {
        if (ch.v[g]) {
            const uint32_t o = own[g];
            const uint64_t b = o == 0 ? b0[g] : o == 1 ? b1[g] : o == 2 ? b2[g] : b3[g];
            const uint32_t pos = S.gbase[(w * 2 + g) * NWAVE + o] + rank_below(b);
            S.queue[o * CHUNK + pos] = make_uint2(key[g], __float_as_uint(ch.w[g]));
        }
    }
    __syncthreads();
    /* wavefront w applies queue w */
    const uint32_t qlen = S.scr[8 + w];
    const uint2 *qu = S.queue + w * CHUNK;
    uint2 *carry = S.carry + w * WAVE;
    uint32_t qp = 0;
    while (qp < qlen) {
        const uint32_t take = min(WAVE - a.ncarry, qlen - qp);
        const bool fresh = lane >= a.ncarry && lane < a.ncarry + take;
        uint32_t nk = 0;
        float nwt = 0.0f;
        if (fresh) {
            const uint2 e = qu[qp + lane - a.ncarry];
            nk = e.x;
            nwt = __uint_as_float(e.y);
        }
        qp += take;
        apply_batch(S.acc, S.tag, carry, fresh, nk, nwt, a);
    }
}

/* Streams the events of the staged segments [0, nseg) and applies them. */
__device__ void flush_segments(const Shared &S, int32_t nseg, uint32_t tw, const uint16_t *tgt,
                               const float *wts, Acc &a, int64_t &events)
{
    const uint32_t tid = threadIdx.x;
    /* stream positions: exclusive prefix of the segment lengths */
    uint32_t len[2], lp[2], cnt = 0;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int32_t s = 2 * tid + j;
        len[j] = s < nseg ? S.sln[s] : 0u;
        lp[j] = cnt;
        cnt += len[j];
    }
    uint32_t E;
    const uint32_t ex = block_excl_scan(cnt, &E, S.scr);
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int32_t s = 2 * tid + j;
        if (s < nseg) S.spos[s] = ex + lp[j];
    }
    events += E;
    for (uint32_t part0 = 0; part0 < E; part0 += BCAP * SUBW) {
        const uint32_t pend = min(E, part0 + BCAP * SUBW);
        /* boundary table: segment holding each position part0 + b * SUBW */
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int32_t s = 2 * tid + j;
            if (s < nseg && len[j] > 0) {
                const uint32_t p0 = ex + lp[j], p1 = p0 + len[j];
                const uint32_t lo = max(p0, part0), hi = min(p1, pend);
                if (lo < hi) {
                    uint32_t b = (lo - part0 + SUBW - 1) / SUBW;
                    for (; part0 + b * SUBW < hi; b++) S.bseg[b] = (uint16_t)s;
                }
            }
        }
        __syncthreads();
        /* two chunk buffers: the loads of the next chunk are issued before
         * the current one is applied */
        Chunk A, B;
        load_chunk(S, A, part0, pend, part0, nseg, tgt, wts);
        uint32_t c0 = part0;
        while (true) {
            if (c0 + CHUNK < pend) load_chunk(S, B, c0 + CHUNK, pend, part0, nseg, tgt, wts);
            apply_chunk(S, A, tw, a);          /* barriers inside */
            c0 += CHUNK;
            if (c0 >= pend) break;
            if (c0 + CHUNK < pend) load_chunk(S, A, c0 + CHUNK, pend, part0, nseg, tgt, wts);
            apply_chunk(S, B, tw, a);
            c0 += CHUNK;
            if (c0 >= pend) break;
        }
        __syncthreads();
    }
}

/* Turns the staged items [0, nitems) of source range r into segments and
 * streams them.  Items are pre ids, ascending. */
__device__ void flush_items(const Shared &S, const bsim_dev &d, int32_t nitems, const bsim_range &r,
                            int32_t par, uint32_t tw, const uint16_t *tgt, const float *wts, Acc &a,
                            int64_t &events)
{
    const uint32_t tid = threadIdx.x;
    const int32_t D = d.dmax;
    const uint64_t dmask = (D >= 63) ? (~0ull >> 1) : ((1ull << D) - 1ull);
    const uint64_t *hist = d.hist + (int64_t)par * d.n;
    uint32_t p[2];
    uint64_t m[2];
    uint32_t nb[2], lp[2], cnt = 0;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int32_t i = 2 * tid + j;
        p[j] = i < nitems ? S.items[i] : 0u;
        m[j] = i < nitems ? ((hist[p[j]] >> 1) & dmask) : 0ull;
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        nb[j] = (uint32_t)__popcll(m[j]);
        lp[j] = cnt;
        cnt += nb[j];
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan(cnt, &tot, S.scr);
    for (uint32_t emitted = 0; emitted < tot; emitted += SEGCAP) {
        const uint32_t wend = min(tot, emitted + SEGCAP);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            uint64_t mm = m[j];
            uint32_t rank = ex + lp[j];
            while (mm) {
                const int32_t bit = __builtin_ctzll(mm);
                mm &= mm - 1;
                if (rank >= emitted && rank < wend) {
                    const int64_t row = r.off_base + (int64_t)(p[j] - (uint32_t)r.lo) * D + bit;
                    const uint32_t o0 = d.off_tab[row], o1 = d.off_tab[row + 1];
                    S.sst[rank - emitted] = o0;
                    S.sln[rank - emitted] = o1 - o0;
                }
                rank++;
            }
        }
        __syncthreads();
        flush_segments(S, (int32_t)(wend - emitted), tw, tgt, wts, a, events);
    }
}

__global__ __launch_bounds__(NTHREAD) void bsim_step_kernel(bsim_dev d, int32_t t, int32_t flags)
{
    extern __shared__ __align__(16) uint32_t lds[];
    const int32_t T = d.tile;
    const int32_t k = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & (WAVE - 1);
    const uint32_t w = tid / WAVE;
    const uint32_t tw = (uint32_t)T / NWAVE;

    Shared S;
    S.acc = reinterpret_cast<float *>(lds);
    S.tag = lds + 2 * T;
    uint32_t *p = S.tag + 2 * T;
    S.queue = reinterpret_cast<uint2 *>(p);           p += NWAVE * CHUNK * 2;
    S.carry = reinterpret_cast<uint2 *>(p);           p += NWAVE * WAVE * 2;
    S.items = p;                                      p += ITEMCAP;
    S.sst = p;                                        p += SEGCAP;
    S.sln = p;                                        p += SEGCAP;
    S.spos = p;                                       p += SEGCAP;
    S.gcnt = p;                                       p += 16 * NWAVE;
    S.gbase = p;                                      p += 16 * NWAVE;
    S.scr = p;                                        p += 16;
    S.bseg = reinterpret_cast<uint16_t *>(p);

    for (int32_t i = tid; i < 2 * T; i += NTHREAD) {
        S.acc[i] = 0.0f;
        S.tag[i] = 0xFFFFFFFFu;
    }
    __syncthreads();

    /* ---- phase (ii) of cycle t-1 ---------------------------------------- */
    if (flags & 1) {
        const int32_t par = (t - 1) & 1;
        const int64_t syn0 = d.tile_syn[k];
        const uint16_t *tgt = d.syn_tgt + syn0;
        const float *wts = d.syn_w + syn0;
        const uint64_t *act = d.active + (int64_t)par * d.nwords;
        Acc a;
        a.round = 0;
        a.ncarry = 0;
        a.ckey = 0;
        a.cw = 0.0f;
        int64_t events = 0;
        for (int32_t ri = d.tile_rng[k]; ri < d.tile_rng[k + 1]; ri++) {
            const bsim_range r = d.rng[ri];
            const int32_t wlo = r.lo >> 6, whi = (r.hi + 63) >> 6;
            int32_t nitems = 0;
            for (int32_t wb = wlo; wb < whi; wb += NTHREAD) {
                const int32_t wi = wb + (int32_t)tid;
                uint64_t word = 0;
                if (wi < whi) {
                    word = act[wi];
                    if (wi == wlo && (r.lo & 63)) word &= ~0ull << (r.lo & 63);
                    if (wi == whi - 1 && (r.hi & 63)) word &= ~0ull >> (64 - (r.hi & 63));
                }
                uint32_t tot;
                const uint32_t ex = block_excl_scan((uint32_t)__popcll(word), &tot, S.scr);
                for (uint32_t emitted = 0; emitted < tot;) {
                    const uint32_t n = min((uint32_t)(ITEMCAP - nitems), tot - emitted);
                    uint64_t mm = word;
                    uint32_t rank = ex;
                    while (mm) {
                        const int32_t bit = __builtin_ctzll(mm);
                        mm &= mm - 1;
                        if (rank >= emitted && rank < emitted + n)
                            S.items[nitems + (rank - emitted)] = (uint32_t)(wi * 64 + bit);
                        rank++;
                    }
                    __syncthreads();
                    nitems += (int32_t)n;
                    emitted += n;
                    if (nitems == ITEMCAP) {
                        flush_items(S, d, nitems, r, par, tw, tgt, wts, a, events);
                        nitems = 0;
                    }
                }
            }
            if (nitems > 0) flush_items(S, d, nitems, r, par, tw, tgt, wts, a, events);
        }
        /* drain the carried events */
        while (__builtin_amdgcn_readfirstlane(a.ncarry) > 0)
            apply_batch(S.acc, S.tag, S.carry + w * WAVE, false, 0u, 0.0f, a);
        if (tid == 0) d.ev_count[k] += events;
        __syncthreads();
    }

    /* ---- phase (i) of cycle t -------------------------------------------- */
    if (flags & 2) {
        const int32_t D = d.dmax;
        const uint64_t dmask = (D >= 63) ? (~0ull >> 1) : ((1ull << D) - 1ull);
        const int32_t base = k * T;
        const uint64_t *hist_in = d.hist + (int64_t)((t - 1) & 1) * d.n;
        uint64_t *hist_out = d.hist + (int64_t)(t & 1) * d.n;
        uint64_t *act_out = d.active + (int64_t)(t & 1) * d.nwords;
        const bool rec = d.raster != nullptr && t >= d.rec_t0 && t < d.rec_t0 + d.rec_steps;
        for (int32_t i0 = (int32_t)w * WAVE; i0 < T; i0 += NTHREAD) {
            const int32_t nid = base + i0 + (int32_t)lane;
            const bool valid = nid < d.n;
            bool fired = false;
            uint64_t h = 0;
            if (valid) {
                const int8_t kind = d.kind[nid];
                if (kind == BSIM_KIND_LIF) {
                    const float inE = S.acc[2 * (i0 + lane)];
                    const float inI = S.acc[2 * (i0 + lane) + 1];
                    const int32_t rs = d.refrac[nid];
                    if (rs > 0) {
                        d.refrac[nid] = rs - 1;
                    } else {
                        float v = d.v[nid], ie = d.ie[nid], ii = d.ii[nid];
                        v = d.Cm[nid] * v + d.v_tmp[nid] + ie * d.C_E[nid] + ii * d.C_I[nid];
                        ie = d.CE[nid] * ie;
                        ii = d.CI[nid] * ii;
                        if (v >= d.v_thresh[nid]) {
                            fired = true;
                            d.refrac[nid] = d.refrac_time[nid] - 1;
                            v = d.v_reset[nid];
                        } else {
                            ie += inE;
                            ii += inI;
                        }
                        d.v[nid] = v;
                        d.ie[nid] = ie;
                        d.ii[nid] = ii;
                    }
                } else if (kind == BSIM_KIND_ARRAY) {
                    const int32_t c = d.arr_cur[nid];
                    if (c < d.arr_end[nid] && d.arr_t[c] <= t) {
                        fired = true;
                        d.arr_cur[nid] = c + 1;
                    }
                } else {
                    fired = bern_hash(d.seed, (uint32_t)nid, (uint32_t)t) < d.bern_p[nid];
                }
                h = (hist_in[nid] << 1) | (fired ? 1ull : 0ull);
                hist_out[nid] = h;
                if (fired) d.fire_count[nid] += 1;
                if (rec && d.rec_index) {
                    const int32_t ri = d.rec_index[nid];
                    if (ri >= 0) d.rec_v[(int64_t)(t - d.rec_t0) * d.n_rec + ri] = d.v[nid];
                }
            }
            const uint64_t fm = ballot64(fired);
            const uint64_t am = ballot64(valid && ((h >> 1) & dmask) != 0);
            if (lane == 0 && base + i0 < d.n) {
                act_out[(base + i0) / 64] = am;
                if (rec) d.raster[(int64_t)(t - d.rec_t0) * d.nwords + (base + i0) / 64] = fm;
            }
        }
    }
}

extern "C" size_t bsim_step_lds_bytes(int32_t tile)
{
    return (size_t)tile * 16 + LDS_FIXED;
}

extern "C" hipError_t bsim_launch_step(const bsim_dev *d, int32_t t, int32_t flags,
                                       hipStream_t stream)
{
    if (d->ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(bsim_step_kernel, dim3(d->ntiles), dim3(NTHREAD),
                       bsim_step_lds_bytes(d->tile), stream, *d, t, flags);
    return hipGetLastError();
}
