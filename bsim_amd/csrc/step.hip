/*
 * step.hip -- the simulation step on CDNA4 (gfx950).
 *
 * A cycle is two launches.  bsim_step_kernel: one workgroup of BSIM_NWAVE + 1
 * wavefronts per tile of TILE consecutive neurons (layout.h) performs
 *
 *   (ii) phase (ii) of cycle t-1: every synapse into the tile whose pre fired
 *        d cycles earlier adds its weight to the target's E or I accumulator,
 *        in the order the reference's CPU simulator applies them
 *        (src/net/Network.cpp:324-327, src/neuron/lif/LIFENeuron.cpp:52-61);
 *   (i)  the LIF_brian update of cycle t (src/neuron/lif/LIFENeuron.cpp:114-160),
 *        which consumes exactly those accumulators, the spike history update
 *        and the ordered compaction of the tile's fired / active neurons into
 *        64-bit words by wavefront ballot.
 *
 * bsim_rows_kernel then builds, for every source chunk, the list of its
 * active (pre, delay) rows for the next phase (ii), shared by all tiles that
 * receive from the chunk; one more workgroup ranks the tiles, heaviest first,
 * by the work the step launch recorded.
 *
 * Phase (ii) for a tile, in windows of up to RCAP rows of the concatenated
 * lists of its source chunks:
 *   stage   the producer wavefront (the last one) reads the list entries and
 *           the row-table entries of window i+1 -- the BSIM_NWAVE + 1 part
 *           boundaries of each row -- into LDS while the others stream window
 *           i; its loads stall no streaming wavefront;
 *   stream  streaming wavefront q walks the staged rows and streams its own
 *           part of each: the concatenated sub-rows form a stream of
 *           positions; per step of 256 positions the sub-rows (64 rows at a
 *           time in the lanes) write the synapse index of each of their
 *           positions into an LDS address list -- from TILE 384 up, 128 rows
 *           at a time write only their first position in the step and a bit
 *           of a position mask, from which each position finds its sub-row
 *           -- and lane i of group g loads position 64 g + i; the loads of
 *           step s+1 are in flight while step s is applied;
 *   apply   the owner of local target x is lane x mod 64 of wavefront q, which
 *           keeps x's E and I accumulators in registers.  For each 64-event
 *           group, the event in lane i sets bit i of its owner's mask by an
 *           LDS atomic OR (order-free); every lane walks its mask upwards --
 *           stream order -- and adds.  No two lanes touch one accumulator, so
 *           the sums need no atomics, no conflict rounds and no barriers
 *           between wavefronts; each wavefront then updates its own part.
 * Every accumulator thus receives its adds one at a time in stream order,
 * (pre, delay, creation) -- the creation order the host verified -- so the
 * sums are the reference's sequential fp32 sums, bit for bit.  The kernel is
 * instantiated for S = TILE / (64 * BSIM_NWAVE) accumulator slots per lane.
 *
 * Arithmetic: compiled with -ffp-contract=off (no FMA contraction) and
 * without denormal flushing, so V = ((Cm*V + v_tmp) + iE*C_E) + iI*C_I rounds
 * exactly as the reference's CPU code.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

#define WAVE 64
#define NWAVE BSIM_NWAVE              /* streaming wavefronts = target quarters */
#define NTHREAD (WAVE * (NWAVE + 1))  /* plus one producer wavefront */
#define ROWW (NWAVE + 1)              /* uint32 per staged row: part starts and end */
#define NGRP 4                        /* 64-position groups per streaming step */
#define SSTEP (NGRP * WAVE)
#define MAXR 128                      /* source ranges per batch */
#define ROWS_THREADS 1024             /* row-list kernel: 4 pres per thread */
#define HEAVY_WGS 256                 /* workgroups run at raised wave priority */
static_assert(BSIM_CHUNK == 4 * ROWS_THREADS, "a source chunk is one row-list workgroup");

#ifndef MARK_MIN_S
#define MARK_MIN_S 2                  /* slots per lane from which sub-rows are expanded by marks */
#endif
static_assert(ROWW == 4 && NWAVE == 3, "a staged row is one 16-byte vector of 3 part starts and the end");

/* Rows per staged window (RCAP): 384 at one accumulator slot per lane (TILE
 * 192, where LDS limits a CU to five workgroups), 512 from two slots up (four
 * workgroups per CU either way, bounded by registers), so that a heavy tile
 * passes fewer window barriers.  RPL rows per producer lane, RW words of
 * packed range indices. */
template <int S_>
struct Win {
    static constexpr int RCAP = S_ >= 2 ? 512 : 384;
    static constexpr int RPL = RCAP / WAVE;
    static constexpr int RW = (RPL + 3) / 4;
};

/* LDS bytes of a workgroup: lbase, obase | stage | rows | tot | alist | cum | stride | omask | pmask */
__host__ __device__ constexpr size_t lds_bytes(int rcap)
{
    return 16 * MAXR + 8 * NWAVE * SSTEP + 4 * (2 * rcap * ROWW + 2 * NWAVE + NWAVE * SSTEP) +
           4 * (MAXR + 2) + 4 * MAXR + 8 * NWAVE * NGRP * WAVE + 8 * NWAVE * NGRP;
}

__device__ __forceinline__ uint32_t bern_hash(uint64_t seed, uint32_t id, uint32_t t)
{
    uint64_t z = seed ^ (((uint64_t)id << 32) | (uint64_t)t);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 32);
}

/* X ~ Poisson from the uniform u by binary search of the rate's inverse-CDF
 * table (layout.h) */
__device__ __forceinline__ uint32_t poisson_pick(const uint32_t *tab, uint32_t u)
{
    uint32_t lo = 0, hi = tab[0] - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (u < tab[1 + mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ uint32_t readlane(uint32_t x, int32_t l)
{
    return __builtin_amdgcn_readlane(x, l);
}

/* A value that is the same in every lane (an LDS word all lanes read, the
 * wavefront index), moved to a scalar register: branches and loops on it are
 * then scalar branches, not exec-masked ones.  Without this the compiler
 * cannot tell that such a loop exits at once for the whole wavefront; it
 * builds the loop with lane masks, merges its exit paths, and waits for all
 * loads in flight where the paths join. */
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint32_t wave_index() { return uniform(threadIdx.x / WAVE); }

/* s_waitcnt vmcnt(0) (GFX9 encoding: vmcnt 0, expcnt and lgkmcnt at their
 * maxima, i.e. not waited for): this wavefront's vector memory loads have
 * completed */
__device__ __forceinline__ void wait_all_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

/* orders this wavefront's LDS accesses (lanes exchange data through LDS) */
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* DPP controls (GFX9 encoding): shift right within a row of 16 lanes, and
 * broadcast lane 15 / 31 into the following rows */
#define DPP_ROW_SHR(n) (0x110 + (n))
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

/* The device clock: the cycle of the coming clocked step launch.  Every
 * workgroup of a clocked launch adds one to the completion counter of its XCD
 * residue x = workgroup % 8 when it ends (an atomic whose result nobody waits
 * for); each residue holds nslots / 8 workgroups, so its counter grows by
 * that much per launch.  While any workgroup of launch t runs, its own
 * residue's counter lies in [t nslots / 8, (t + 1) nslots / 8) -- it has not
 * added its one yet -- so the quotient is t; a row-list launch, after its
 * step launch has completed, reads t + 1 from any counter.  Plain (scalar,
 * cached) loads: a launch sees the counts of the launches before it, and a
 * value cached during the launch is still within its interval.  (A single
 * counter whose last incrementer advanced the cycle made every workgroup wait
 * on one contended address: 48 against 20 us per launch in a cycle without
 * deliveries.) */
__device__ __forceinline__ int32_t device_cycle(const bsim_dev &d, uint32_t x)
{
    /* count / (nslots / 8) without an integer division: the quotient's
     * fractional part is a multiple of 8 / nslots, so half of that added
     * before truncation absorbs the rounding of the product (below 2^-20 for
     * cycles below 2^31, against a half step of at least 2^-19 for grids of
     * up to 2^18 workgroups per residue) */
    const double c = (double)d.done[BSIM_DONE_STRIDE * x];
    return (int32_t)(c * d.done_inv + 0.5 * d.done_inv);
}

/* inclusive prefix sum over the wavefront: four in-row shifts, then the row
 * totals are carried across rows by two broadcasts; lanes without a source
 * read 0 */
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x)
{
    int v = (int)x;
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(1), 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(2), 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(4), 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_SHR(8), 0xf, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST15, 0xa, 0xf, false);
    v += __builtin_amdgcn_update_dpp(0, v, DPP_ROW_BCAST31, 0xc, 0xf, false);
    return (uint32_t)v;
}

struct Shared {
    uint32_t *rows;      /* [2][RCAP][ROWW] part boundaries of the staged rows, double-buffered */
    uint32_t *tot;       /* [2][NWAVE] stream length of each quarter in the window */
    uint2 *stage;        /* [NWAVE][SSTEP] (local target in quarter, weight bits) */
    uint64_t *omask;     /* [NWAVE][NGRP][WAVE] events of a group owned by each lane */
    uint64_t *pmask;     /* [NWAVE][NGRP] step positions where a sub-row starts (mark expansion) */
    uint32_t *alist;     /* [NWAVE][SSTEP] synapse index of each position of a step */
    uint32_t *cum;       /* [MAXR + 1] rows before each source range of the batch */
    uint32_t *stride;    /* [MAXR] row-table entries per row of each source range */
    int64_t *lbase;      /* [MAXR] first list entry of each source range */
    int64_t *obase;      /* [MAXR] row-table base of each source range */
};

/* A window of 64 R staged rows in the lanes of a wavefront: lane j holds
 * rows rw + R j .. rw + R j + R - 1 of the wavefront's sub-row stream. */
template <int R>
struct RowWin {
    int32_t rw;         /* first row (uniform) */
    uint32_t pos[R];    /* stream position of each sub-row */
    uint32_t len[R];    /* sub-row length */
    uint32_t base[R];   /* synapse index of the sub-row minus pos */
    uint32_t end;       /* stream position after the window (uniform) */
};

template <int R>
__device__ __forceinline__ void load_win(const uint32_t *rows, RowWin<R> &W, int32_t rw, uint32_t pos0,
                                         int32_t nrows)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t q = wave_index();
    uint32_t s0[R], sum = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int32_t j = rw + R * (int32_t)lane + r;
        uint32_t a = 0, c = 0;
        if (j < nrows) {
            if constexpr (R == 1) {
                a = rows[j * ROWW + q];
                c = rows[j * ROWW + q + 1];
            } else {                      /* one 16-byte read per row */
                const uint4 v = reinterpret_cast<const uint4 *>(rows)[j];
                a = (q == 0) ? v.x : (q == 1) ? v.y : v.z;
                c = (q == 0) ? v.y : (q == 1) ? v.z : v.w;
            }
        }
        s0[r] = a;
        W.len[r] = c - a;
        sum += W.len[r];
    }
    const uint32_t incl = wave_incl_scan(sum);
    uint32_t p = pos0 + incl - sum;
#pragma unroll
    for (int r = 0; r < R; r++) {
        W.pos[r] = p;
        W.base[r] = s0[r] - p;
        p += W.len[r];
    }
    W.rw = rw;
    W.end = pos0 + readlane(incl, WAVE - 1);
}

/* Loaded events of one streaming step: NGRP groups of 64. */
struct Step {
    uint32_t tl[NGRP];
    float w[NGRP];
    bool v[NGRP];
};

/* Loads the events of stream positions [b, b + SSTEP) into lane i of group g
 * (position b + 64 g + i), sliding the row window over the staged rows as
 * needed.  Two ways to find each position's synapse index:
 *   R == 1  every sub-row that intersects the step writes the index of each
 *           of its positions into the wavefront's address list (one pass per
 *           four positions of the window's longest sub-row);
 *   R >= 2  (marks) every such sub-row writes only its first position in the
 *           step: its synapse base into the address list and a bit into the
 *           step's position mask; each position then takes the base of the
 *           last mark at or before it (the highest set mask bit below it).
 *           Costs one pass per window whatever the sub-row lengths.
 * The loads are issued unconditionally (lanes without an event read the
 * tile's first synapse), so that the compiler can count them exactly and the
 * apply of the previous step waits only for its own loads. */
template <int R>
__device__ __forceinline__ void load_step(const Shared &S, const uint32_t *rows, RowWin<R> &W, uint32_t b,
                                          uint32_t tot, int32_t nrows, uint32_t nsyn,
                                          const bsim_syn *syn, Step &st)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t q = wave_index();
    uint32_t *al = S.alist + q * SSTEP;
    uint64_t *pm = S.pmask + q * NGRP;
    const uint32_t e = min(b + SSTEP, tot);
    if (b < tot) {                                            /* uniform */
        while (true) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const uint32_t lo = max(W.pos[r], b), hi = min(W.pos[r] + W.len[r], e);
                if constexpr (R == 1) {
                    const uint32_t n = hi > lo ? hi - lo : 0u;
                    const uint32_t a0 = W.base[r] + lo;
                    uint32_t *dst = al + (lo - b);
                    for (uint32_t x = 0; ballot64(x < n) != 0ull; x += 4) {
                        if (x < n) dst[x] = a0 + x;
                        if (x + 1 < n) dst[x + 1] = a0 + x + 1;
                        if (x + 2 < n) dst[x + 2] = a0 + x + 2;
                        if (x + 3 < n) dst[x + 3] = a0 + x + 3;
                    }
                } else if (hi > lo) {
                    const uint32_t x = lo - b;
                    al[x] = W.base[r];
                    atomicOr((unsigned long long *)&pm[x / WAVE], 1ull << (x & (WAVE - 1)));
                }
            }
            if (W.end >= e) break;                            /* uniform */
            if (W.rw + R * WAVE >= nrows) break;
            load_win<R>(rows, W, W.rw + R * WAVE, W.end, nrows);
        }
    }
    wave_sync();
    uint64_t M[NGRP];
    if constexpr (R > 1) {
#pragma unroll
        for (int g = 0; g < NGRP; g++) M[g] = pm[g];
        wave_sync();
        if (lane < NGRP) pm[lane] = 0ull;
    }
    const uint64_t le = ~0ull >> (WAVE - 1 - lane);           /* bits 0..lane */
    uint32_t last = 0;          /* the step's position 0 is always marked */
#pragma unroll
    for (int g = 0; g < NGRP; g++) {
        const uint32_t pr = g * WAVE + lane;
        const bool v = b + pr < e;
        uint32_t addr;
        if constexpr (R == 1) {
            addr = v ? al[pr] : 0u;
        } else {
            const uint64_t mg = M[g] & le;
            const uint32_t idx = mg ? (uint32_t)(g * WAVE + WAVE - 1 - __clzll(mg)) : last;
            last = M[g] ? (uint32_t)(g * WAVE + WAVE - 1 - __clzll(M[g])) : last;
            addr = v ? al[idx] + b + pr : 0u;
        }
        addr = (addr < nsyn) ? addr : 0u;
        st.v[g] = v;
        const bsim_syn sv = syn[addr];
        st.tl[g] = sv.tgt;
        st.w[g] = __uint_as_float(sv.w);
    }
    wave_sync();
}

/* E and I accumulators of the S local targets lane, lane + 64, ... of the
 * wavefront's quarter */
template <int S_>
struct Acc {
    float e[S_], i[S_];
};

/* The state of a streaming wavefront's neurons that phase (i) reads, loaded
 * before phase (ii) so that its latency is hidden behind the delivery (no
 * other workgroup writes these words during the launch: a neuron belongs to
 * one tile, and the history words read are the previous cycle's copy).  Kept
 * in registers up to PREFETCH_MAX_S slots per lane. */
#ifndef PREFETCH_MAX_S
#define PREFETCH_MAX_S 1
#endif
template <int S_>
struct NState {
    uint32_t cls[S_];
    int32_t refrac[S_];
    float v[S_], ie[S_], ii[S_], vtmp[S_];
    uint64_t hist[S_];
};

template <int S_>
__device__ __forceinline__ void load_nstate(const bsim_dev &d, int32_t base, const uint64_t *hist_in,
                                            NState<S_> &ns)
{
    const int32_t lane = (int32_t)(threadIdx.x & (WAVE - 1));
#pragma unroll
    for (int j = 0; j < S_; j++) {
        const int32_t nid = base + j * WAVE + lane;
        const bool valid = nid < d.n;
        const int32_t x = valid ? nid : 0;
        ns.cls[j] = d.ncls[x];
        ns.refrac[j] = d.refrac[x];
        ns.v[j] = d.v[x];
        ns.ie[j] = d.ie[x];
        ns.ii[j] = d.ii[x];
        ns.vtmp[j] = d.v_tmp[x];
        ns.hist[j] = hist_in[x];
    }
}

/* Adds one owned event to the accumulator of its slot (E if w >= 0, else I)
 * when `valid`.  Written as selects over statically indexed accumulators, so
 * that they stay in registers for every S_ and no branch separates the staged
 * reads of a round. */
template <int S_>
__device__ __forceinline__ void add_event(Acc<S_> &acc, uint2 ev, bool valid)
{
    const float x = __uint_as_float(ev.y);
    const uint32_t slot = ev.x >> 6;
    const bool exc = x >= 0.0f;
#pragma unroll
    for (int j = 0; j < S_; j++) {
        /* with one slot per lane an owned event's local target is below 64:
         * its slot is 0 and needs no test */
        const bool mine = valid && (S_ == 1 || slot == (uint32_t)j);
        const float ae = acc.e[j] + x, ai = acc.i[j] + x;
        acc.e[j] = (mine && exc) ? ae : acc.e[j];
        acc.i[j] = (mine && !exc) ? ai : acc.i[j];
    }
}

/* Applies one streaming step: every lane applies, in stream order, the events
 * whose target it owns.  For each group, the event in lane i sets bit i of its
 * owner's mask with an LDS atomic OR (order-free); the owner walks its mask
 * upwards, i.e. in stream order. */
template <int S_>
__device__ __forceinline__ void apply_step(const Shared &S, const Step &st, uint32_t qbase, Acc<S_> &acc)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t q = wave_index();
    uint2 *stg = S.stage + q * SSTEP;
    uint64_t *om = S.omask + q * (NGRP * WAVE);
    const uint64_t mybit = 1ull << lane;
#pragma unroll
    for (int g = 0; g < NGRP; g++) {
        const uint32_t li = st.tl[g] - qbase;     /* local target within the quarter */
        stg[g * WAVE + lane] = make_uint2(li, __float_as_uint(st.w[g]));
        if (st.v[g]) atomicOr((unsigned long long *)&om[g * WAVE + (li & (WAVE - 1))], (unsigned long long)mybit);
    }
    wave_sync();
    uint64_t M[NGRP];
#pragma unroll
    for (int g = 0; g < NGRP; g++) {
        M[g] = om[g * WAVE + lane];
        om[g * WAVE + lane] = 0ull;
    }
#pragma unroll
    for (int g = 0; g < NGRP; g++) {
        uint64_t m = M[g];
        while (m) {
            /* up to 2 owned events per round: their staged reads are issued
             * together, the adds stay in stream order */
            const uint32_t i0 = (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            uint32_t i1 = i0;
            const bool h1 = m != 0ull;
            if (h1) { i1 = (uint32_t)__builtin_ctzll(m); m &= m - 1; }
            const uint2 e0 = stg[g * WAVE + i0];
            const uint2 e1 = stg[g * WAVE + i1];
            add_event<S_>(acc, e0, true);
            add_event<S_>(acc, e1, h1);
        }
    }
    wave_sync();
}

/* Streams this wavefront's quarter of the staged rows [0, nrows).  The loads
 * of step s+1 are in flight while step s is applied.  No load is issued past
 * the last step: a load left in flight at the end of the stream would still be
 * pending when the next window starts (the workgroup barrier waits for LDS
 * only), and its destination registers, reused there, would make the compiler
 * wait for every load in flight at the top of each step -- the step pipeline
 * would then hold one step's loads at a time. */
template <int S_>
__device__ void stream_rows(const Shared &S, const uint32_t *rows, int32_t nrows, uint32_t tot,
                            uint32_t qbase, uint32_t nsyn, const bsim_syn *syn,
                            Acc<S_> &acc)
{
    if (tot == 0) return;                                     /* uniform */
    constexpr int R = (S_ >= MARK_MIN_S) ? 2 : 1;
    RowWin<R> W;
    load_win<R>(rows, W, 0, 0, nrows);
    Step A, B;
    load_step(S, rows, W, 0, tot, nrows, nsyn, syn, A);
    uint32_t b = 0;
    while (true) {
        if (b + SSTEP >= tot) {                               /* uniform: A is the last step */
            apply_step<S_>(S, A, qbase, acc);
            break;
        }
        load_step(S, rows, W, b + SSTEP, tot, nrows, nsyn, syn, B);
        apply_step<S_>(S, A, qbase, acc);
        b += SSTEP;
        if (b + SSTEP >= tot) {                               /* uniform: B is the last step */
            apply_step<S_>(S, B, qbase, acc);
            break;
        }
        load_step(S, rows, W, b + SSTEP, tot, nrows, nsyn, syn, A);
        apply_step<S_>(S, B, qbase, acc);
        b += SSTEP;
    }
}

/* ---- the producer wavefront: staging of row windows ---------------------- */

/* Row j of a window of the current range batch belongs to the source range
 * whose cumulative row interval holds it: the largest i with cum[i] <= j
 * (ranges without active rows repeat the count and are passed over). */
__device__ __forceinline__ int32_t range_of(const Shared &S, uint32_t j, int32_t nrb)
{
    int32_t a = 0, b = nrb - 1;
    while (a < b) {
        const int32_t m = (a + b + 1) >> 1;
        if (S.cum[m] <= j) a = m; else b = m - 1;
    }
    return a;
}

/* The source range index of each of a lane's rows (< MAXR = 128) is found
 * once per window, when its list entries are read, and kept packed four to a
 * register (RW words) until the window is staged. */
static_assert(MAXR <= 256, "range indices are packed in bytes");
template <int RW>
__device__ __forceinline__ int32_t range_at(const uint32_t (&R)[RW], int x)
{
    return (int32_t)((R[x / 4] >> (8 * (x % 4))) & 0xFFu);
}

/* Active-row list entries of window `win` (lane holds rows lane + 64 x) and
 * the source range of each row (R).  The rows of a lane ascend from window
 * to window, so their ranges are found by walking forward from the range of
 * the lane's last row in the previous window (ri, carried by the caller; a
 * search for the first window). */
template <int RCAP, int RPL = RCAP / WAVE, int RW = (RPL + 3) / 4>
__device__ __forceinline__ void prod_entries(const Shared &S, const bsim_dev &d, int32_t win,
                                             uint32_t nrows_tot, int32_t nrb, uint32_t (&E)[RPL],
                                             uint32_t (&R)[RW], int32_t &ri)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    if (win == 0) ri = range_of(S, lane, nrb);
#pragma unroll
    for (int w = 0; w < RW; w++) R[w] = 0u;
#pragma unroll
    for (int x = 0; x < RPL; x++) {
        const uint32_t j = (uint32_t)win * RCAP + lane + WAVE * x;
        E[x] = 0u;
        if (j < nrows_tot) {
            while (ri + 1 < nrb && S.cum[ri + 1] <= j) ri++;
            E[x] = d.src_list[S.lbase[ri] + (j - S.cum[ri])];
        }
        R[x / 4] |= (uint32_t)ri << (8 * (x % 4));      /* ri < nrb <= MAXR */
    }
}

/* Stages window `win` into LDS buffer `buf`: the row-table entry of each of
 * its rows (from the list entries E) -- the part starts and the end of the
 * tile's sub-rows, one aligned 16-byte load -- and the stream length of each
 * part.  The rows of a lane ascend, so their source range is found by walking
 * forward.  Every lane uses all of its E and Tb registers, including those of
 * rows past the window's end (zeros; the slots of the buffer they fill are
 * never read), so that no load of the producer is still pending, as far as
 * the compiler can tell, when the registers are reused by the streaming code
 * that shares the program. */
template <int RCAP, int RPL = RCAP / WAVE, int RW = (RPL + 3) / 4>
__device__ __forceinline__ void prod_stage(const Shared &S, const bsim_dev &d, int32_t buf, int32_t win,
                                           uint32_t nrows_tot, const uint32_t (&E)[RPL],
                                           const uint32_t (&R)[RW])
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    uint4 *rows = reinterpret_cast<uint4 *>(S.rows + buf * (RCAP * ROWW));
    uint4 Tb[RPL];
#pragma unroll
    for (int x = 0; x < RPL; x++) {
        const uint32_t j = (uint32_t)win * RCAP + lane + WAVE * x;
        const int32_t ri = range_at<RW>(R, x);
        const int64_t e0 = S.obase[ri] + (int64_t)E[x] * S.stride[ri];
        Tb[x] = make_uint4(0u, 0u, 0u, 0u);
        if (j < nrows_tot) Tb[x] = *reinterpret_cast<const uint4 *>(d.off_tab + e0);
    }
    uint32_t ql[NWAVE];
#pragma unroll
    for (int q = 0; q < NWAVE; q++) ql[q] = 0u;
#pragma unroll
    for (int x = 0; x < RPL; x++) {
        const uint32_t jj = lane + WAVE * x;
        rows[jj] = Tb[x];
        ql[0] += Tb[x].y - Tb[x].x;
        ql[1] += Tb[x].z - Tb[x].y;
        ql[2] += Tb[x].w - Tb[x].z;
    }
#pragma unroll
    for (int q = 0; q < NWAVE; q++) {
        const uint32_t sum = readlane(wave_incl_scan(ql[q]), WAVE - 1);
        if (lane == 0) S.tot[buf * NWAVE + q] = sum;
    }
}

/* Phase (ii) of one batch of up to MAXR source ranges of the tile: wavefront
 * NWAVE stages window i+1 of the concatenated active-row lists while
 * wavefronts 0..NWAVE-1 stream window i. */
template <int S_>
__device__ void deliver_batch(const Shared &S, const bsim_dev &d, int32_t r0, int32_t nrb,
                              uint32_t qbase, uint32_t nsyn, const bsim_syn *syn,
                              int64_t &events, uint32_t &rows, Acc<S_> &acc)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t w = wave_index();
    const bool prod = w == NWAVE;
    if (prod) {
        /* range metadata and the cumulative row counts of the batch */
        uint32_t c[MAXR / WAVE];
#pragma unroll
        for (int x = 0; x < MAXR / WAVE; x++) {
            const int32_t i = (int32_t)lane * (MAXR / WAVE) + x;
            c[x] = 0u;
            if (i < nrb) {
                const bsim_range r = d.rng[r0 + i];
                c[x] = d.rng_count[r0 + i];
                S.lbase[i] = r.lbase;
                S.obase[i] = r.off_base;
                S.stride[i] = r.stride;
            }
        }
        uint32_t sum = 0;
#pragma unroll
        for (int x = 0; x < MAXR / WAVE; x++) sum += c[x];
        const uint32_t incl = wave_incl_scan(sum);
        uint32_t run = incl - sum;
#pragma unroll
        for (int x = 0; x < MAXR / WAVE; x++) {
            const int32_t i = (int32_t)lane * (MAXR / WAVE) + x;
            if (i < nrb) S.cum[i] = run;
            run += c[x];
        }
        if (lane == WAVE - 1) S.cum[nrb] = incl;
    }
    __syncthreads();
    constexpr int RCAP = Win<S_>::RCAP, RPL = Win<S_>::RPL, RW = Win<S_>::RW;
    const uint32_t nrows_tot = uniform(S.cum[nrb]);
    rows += nrows_tot;
    const int32_t nwin = (int32_t)((nrows_tot + RCAP - 1) / RCAP);
    if (nwin == 0) return;                                    /* uniform */
    /* The producer issues the list entries of the window after next (En)
     * before it stages the next one (from E), and ends each turn with all of
     * its loads complete: the staging waits for its row-table loads, issued
     * after those entries, so the wait costs nothing, and the compiler, which
     * sees one program for all wavefronts, then knows that no load of the
     * producer is pending in the streaming code (otherwise it makes the
     * streaming wavefronts wait for loads that only the producer issued). */
    uint32_t E[RPL], En[RPL];
    uint32_t R[RW], Rn[RW];
    int32_t ri = 0;
    if (prod) {
        prod_entries<RCAP>(S, d, 0, nrows_tot, nrb, E, R, ri);
        if (nwin > 1) prod_entries<RCAP>(S, d, 1, nrows_tot, nrb, En, Rn, ri);
        prod_stage<RCAP>(S, d, 0, 0, nrows_tot, E, R);
        wait_all_loads();
    }
    __syncthreads();
    for (int32_t i = 0; i < nwin; i++) {
        const int32_t buf = i & 1;
        if (prod) {
            if (i + 1 < nwin) {
#pragma unroll
                for (int x = 0; x < RPL; x++) E[x] = En[x];
#pragma unroll
                for (int x = 0; x < RW; x++) R[x] = Rn[x];
                if (i + 2 < nwin) prod_entries<RCAP>(S, d, i + 2, nrows_tot, nrb, En, Rn, ri);
                prod_stage<RCAP>(S, d, buf ^ 1, i + 1, nrows_tot, E, R);
            }
            wait_all_loads();
        } else {
            const int32_t nrows = (int32_t)min((uint32_t)RCAP, nrows_tot - (uint32_t)i * RCAP);
            const uint32_t qtot = uniform(S.tot[buf * NWAVE + w]);
            events += qtot;
            stream_rows<S_>(S, S.rows + buf * (RCAP * ROWW), nrows, qtot, qbase, nsyn, syn, acc);
        }
        __syncthreads();
    }
}

template <int S_>
__global__ __launch_bounds__(NTHREAD) void bsim_step_kernel(bsim_dev d, int32_t t, int32_t flags)
{
    extern __shared__ __align__(16) uint32_t lds[];
    const bool clocked = t < 0;
    if (clocked) {
        t = device_cycle(d, blockIdx.x % 8u);
        flags = (t > 0 ? 1 : 0) | 2;
    }
    /* the heaviest tiles come first (bsim_rows_kernel ranks them): their
     * wavefronts get issue priority on their compute unit, so that they are
     * not slowed down by the light tiles that fill the unit around them */
    if (blockIdx.x < HEAVY_WGS) __builtin_amdgcn_s_setprio(3);
    const int32_t T = d.tile;
    const int32_t k = d.perm[blockIdx.x];       /* -1: an empty slot */
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & (WAVE - 1);
    const uint32_t w = wave_index();
    const uint32_t qbase = w * ((uint32_t)T / NWAVE);

    Shared S;
    {
        int64_t *p8 = reinterpret_cast<int64_t *>(lds);
        S.lbase = p8;                                 p8 += MAXR;
        S.obase = p8;                                 p8 += MAXR;
        S.stage = reinterpret_cast<uint2 *>(p8);
        uint32_t *p = reinterpret_cast<uint32_t *>(p8) + NWAVE * SSTEP * 2;
        S.rows = p;                                   p += 2 * Win<S_>::RCAP * ROWW;
        S.tot = p;                                    p += 2 * NWAVE;
        S.alist = p;                                  p += NWAVE * SSTEP;
        S.cum = p;                                    p += MAXR + 2;
        S.stride = p;                                 p += MAXR;
        S.omask = reinterpret_cast<uint64_t *>(p);    /* 8-byte aligned: see lds_bytes */
        S.pmask = S.omask + NWAVE * NGRP * WAVE;
    }

    Acc<S_> acc;
#pragma unroll
    for (int j = 0; j < S_; j++) {
        acc.e[j] = 0.0f;
        acc.i[j] = 0.0f;
    }
    if (w < NWAVE) {
#pragma unroll
        for (int g = 0; g < NGRP; g++) S.omask[(w * NGRP + g) * WAVE + lane] = 0ull;
        if (lane < NGRP) S.pmask[w * NGRP + lane] = 0ull;
    }

    constexpr bool PRE = S_ <= PREFETCH_MAX_S;
    NState<PRE ? S_ : 1> ns;
    if constexpr (PRE) {
        if ((flags & 2) && w < NWAVE && k >= 0)
            load_nstate<S_>(d, k * T + (int32_t)qbase, d.hist + (int64_t)((t - 1) & 1) * d.n, ns);
    }

    /* ---- phase (ii) of cycle t-1 ---------------------------------------- */
    if ((flags & 1) && k >= 0) {
        const int64_t syn0 = d.tile_base[k];
        const uint32_t nsyn = d.tile_nsyn[k];
        const bsim_syn *syn = d.syn + syn0;
        int64_t events = 0;
        uint32_t rows = 0;
        const int32_t r0 = d.tile_rng[k], r1 = d.tile_rng[k + 1];
        for (int32_t rb = r0; rb < r1; rb += MAXR)
            deliver_batch<S_>(S, d, rb, min(MAXR, r1 - rb), qbase, nsyn, syn, events, rows, acc);
        if (lane == 0 && w < NWAVE)
            atomicAdd((unsigned long long *)&d.ev_count[k], (unsigned long long)events);
        /* the tile's work in this launch, when the row-list launch that
         * follows ranks the tiles (it runs for cycle t + 1): its deliveries
         * plus a fixed cost of two deliveries per active row, x 16 */
        const int32_t re = d.rank_every;
        if (clocked && (re == 1 || (re > 1 && (t + 1) % re == 0)) && lane == 0)
            atomicAdd((unsigned long long *)&d.tile_load[k],
                      w < NWAVE ? (unsigned long long)events * 16ull : (unsigned long long)rows * 32ull);
    }

    /* ---- phase (i) of cycle t: wavefront w updates its own quarter ---------- */
    if ((flags & 2) && w < NWAVE && k >= 0) {
        const int32_t D = d.dmax;
        const uint64_t dmask = (D >= 63) ? (~0ull >> 1) : ((1ull << D) - 1ull);
        const int32_t base = k * T + (int32_t)qbase;
        const uint64_t *hist_in = d.hist + (int64_t)((t - 1) & 1) * d.n;
        uint64_t *hist_out = d.hist + (int64_t)(t & 1) * d.n;
        uint64_t *act_out = d.active + (int64_t)(t & 1) * d.nwords;
        const bool rec = d.raster != nullptr && t >= d.rec_t0 && t < d.rec_t0 + d.rec_steps;
#pragma unroll
        for (int j = 0; j < S_; j++) {
            const int32_t n0 = base + j * WAVE;
            const int32_t nid = n0 + (int32_t)lane;
            const bool valid = nid < d.n;
            bool fired = false;
            uint64_t h = 0;
            if (valid) {
                uint32_t cls;
                int32_t rs;
                float v, ie, ii, vtmp;
                uint64_t hin;
                if constexpr (PRE) {
                    cls = ns.cls[j]; rs = ns.refrac[j]; v = ns.v[j]; ie = ns.ie[j]; ii = ns.ii[j];
                    vtmp = ns.vtmp[j]; hin = ns.hist[j];
                } else {
                    cls = d.ncls[nid]; rs = d.refrac[nid]; v = d.v[nid]; ie = d.ie[nid]; ii = d.ii[nid];
                    vtmp = d.v_tmp[nid]; hin = hist_in[nid];
                }
                const uint32_t kind = cls >> 30;
                if (kind == BSIM_KIND_LIF) {
                    const float inE = acc.e[j];
                    const float inI = acc.i[j];
                    if (rs > 0) {
                        d.refrac[nid] = rs - 1;
                    } else {
                        const bsim_lifc c = d.lifc[cls & BSIM_CLS_SET_MASK];
                        v = c.Cm * v + vtmp + ie * c.C_E + ii * c.C_I;
                        ie = c.CE * ie;
                        ii = c.CI * ii;
                        if (v >= c.v_thresh) {
                            fired = true;
                            d.refrac[nid] = c.refrac_time - 1;
                            v = c.v_reset;
                        } else {
                            if (d.in_sum) d.in_sum[nid] = d.in_sum[nid] + (inE + inI);
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
                } else if (kind == BSIM_KIND_POISSON) {
                    /* PoissonNeuron::update (src/neuron/poisson/PoissonNeuron.cpp:52-63):
                     * fire at the drawn cycle before the end cycle, then draw
                     * the next one 1 + X cycles later */
                    const int32_t nx = d.arr_cur[nid];
                    if (t == nx && nx < d.arr_end[nid]) {
                        fired = true;
                        const uint32_t x = poisson_pick(d.pois_tab + d.bern_p[nid],
                                                        bern_hash(d.bern_seed[nid], d.bern_key[nid], (uint32_t)t));
                        d.arr_cur[nid] = (int32_t)min((int64_t)nx + 1 + x, (int64_t)0x7FFFFFFF);
                    }
                } else {
                    fired = bern_hash(d.bern_seed[nid], d.bern_key[nid], (uint32_t)t) < d.bern_p[nid];
                }
                h = (hin << 1) | (fired ? 1ull : 0ull);
                hist_out[nid] = h;
                if (d.slot_count) d.slot_count[nid] += __popcll((h >> 1) & dmask);
                if (fired) d.fire_count[nid] += 1;
                if (rec && d.rec_index) {
                    const int32_t ri = d.rec_index[nid];
                    if (ri >= 0) {
                        const int64_t o = (int64_t)(t - d.rec_t0) * d.n_rec + ri;
                        d.rec_v[o] = kind == BSIM_KIND_LIF ? v : 0.0f;
                        if (d.rec_ie) {
                            d.rec_ie[o] = kind == BSIM_KIND_LIF ? ie : 0.0f;
                            d.rec_ii[o] = kind == BSIM_KIND_LIF ? ii : 0.0f;
                            d.rec_in_e[o] = kind == BSIM_KIND_LIF ? acc.e[j] : 0.0f;
                            d.rec_in_i[o] = kind == BSIM_KIND_LIF ? acc.i[j] : 0.0f;
                        }
                    }
                }
            }
            const uint64_t fm = ballot64(fired);
            const uint64_t am = ballot64(valid && ((h >> 1) & dmask) != 0);
            if (lane == 0 && n0 < d.n) {
                act_out[n0 / 64] = am;
                if (rec) d.raster[(int64_t)(t - d.rec_t0) * d.nwords + n0 / 64] = fm;
            }
        }
    }
    if (clocked && tid == 0)        /* the cycle was read by every wavefront at its start */
        (void)__hip_atomic_fetch_add(d.done + BSIM_DONE_STRIDE * (blockIdx.x % 8u), (uint64_t)1,
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Orders the tiles by their work in the step launch just finished, heaviest
 * first, so that the hardware dispatches the heavy tiles of the coming launch
 * before the light ones (results do not depend on the order; the firing
 * volleys move slowly against one cycle).  A tile's work, accumulated by its
 * own step workgroup, is its deliveries plus a fixed cost of two deliveries
 * per active row.  Counting sort over half-octave classes; within a class the
 * order is arbitrary.  Runs in one workgroup of ROWS_THREADS threads, which
 * also clears the loads for the next ranking. */
#define RANK_CLASSES 128

__device__ __forceinline__ uint32_t load_class(uint64_t x)
{
    if (x < 2) return 0;
    const uint32_t e = 63u - (uint32_t)__builtin_clzll(x);        /* floor(log2 x) >= 1 */
    const uint32_t h = (uint32_t)(x >> (e - 1)) & 1u;              /* next bit */
    return min(2u * e + h, (uint32_t)RANK_CLASSES - 1u);
}

#define RANK_LDS_TILES 16384   /* tile classes kept in LDS up to this many tiles */
#define NRES 8                  /* XCDs: workgroup w runs on XCD (w + c) mod 8 */
#define RANK_REG 4              /* tiles per thread kept in registers, up to RANK_REG * ROWS_THREADS tiles */

__device__ void rank_tiles(const bsim_dev &d)
{
    __shared__ uint32_t cnt[NRES][RANK_CLASSES], fill[NRES][RANK_CLASSES], rtot[NRES];
    __shared__ uint8_t cls_lds[RANK_LDS_TILES];
    const uint32_t tid = threadIdx.x;
    const bool in_lds = d.ntiles <= RANK_LDS_TILES;
    uint8_t *cls = in_lds ? cls_lds : d.tile_cls;
    const bool in_regs = d.ntiles <= RANK_REG * ROWS_THREADS;
    uint32_t xr[RANK_REG], cr[RANK_REG];
    for (uint32_t i = tid; i < NRES * RANK_CLASSES; i += ROWS_THREADS) (&cnt[0][0])[i] = 0u;
    if (in_regs) {              /* all loads of a thread issued together */
        uint64_t ld[RANK_REG];
#pragma unroll
        for (int i = 0; i < RANK_REG; i++) {
            const int32_t k = (int32_t)tid + i * ROWS_THREADS;
            ld[i] = (k < d.ntiles) ? d.tile_load[k] : 0ull;
            xr[i] = (k < d.ntiles) ? d.tile_res[k] : 0u;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < RANK_REG; i++) {
            const int32_t k = (int32_t)tid + i * ROWS_THREADS;
            cr[i] = load_class(ld[i]);
            if (k < d.ntiles) {
                d.tile_load[k] = 0ull;
                atomicAdd(&cnt[xr[i]][cr[i]], 1u);
            }
        }
    } else {
        __syncthreads();
        for (int32_t k = tid; k < d.ntiles; k += ROWS_THREADS) {
            /* written by other workgroups' atomics in L2; this workgroup never read it */
            const uint32_t c = load_class(d.tile_load[k]);
            cls[k] = (uint8_t)c;
            d.tile_load[k] = 0ull;
            atomicAdd(&cnt[d.tile_res[k]][c], 1u);
        }
    }
    __syncthreads();
    /* exclusive prefix of each residue's class counts, heaviest class first:
     * thread t takes residue t / RANK_CLASSES, class RANK_CLASSES - 1 - t mod
     * RANK_CLASSES (two wavefronts per residue) */
    static_assert(NRES * RANK_CLASSES == ROWS_THREADS && RANK_CLASSES == 2 * WAVE, "one thread per class");
    {
        __shared__ uint32_t half[ROWS_THREADS / WAVE];
        const uint32_t x = tid / RANK_CLASSES, c = RANK_CLASSES - 1u - tid % RANK_CLASSES;
        const uint32_t v = cnt[x][c];
        const uint32_t incl = wave_incl_scan(v);
        if ((tid & (WAVE - 1)) == WAVE - 1) half[tid / WAVE] = incl;
        __syncthreads();
        const uint32_t carry = ((tid / WAVE) & 1u) ? half[tid / WAVE - 1] : 0u;
        fill[x][c] = carry + incl - v;
        if (tid % RANK_CLASSES == RANK_CLASSES - 1u) rtot[x] = carry + incl;
    }
    __syncthreads();
    /* tile of residue x at rank i goes to workgroup NRES * i + x */
    if (in_regs) {
#pragma unroll
        for (int i = 0; i < RANK_REG; i++) {
            const int32_t k = (int32_t)tid + i * ROWS_THREADS;
            if (k < d.ntiles) {
                const uint32_t pos = atomicAdd(&fill[xr[i]][cr[i]], 1u);
                d.perm[NRES * pos + xr[i]] = k;
            }
        }
    } else {
        for (int32_t k = tid; k < d.ntiles; k += ROWS_THREADS) {
            const uint32_t x = d.tile_res[k];
            const uint32_t pos = atomicAdd(&fill[x][cls[k]], 1u);
            d.perm[NRES * pos + x] = k;
        }
    }
    /* the remaining slots of each residue are empty */
    const int32_t per = d.nslots / NRES;
    for (int32_t sidx = tid; sidx < d.nslots; sidx += ROWS_THREADS) {
        const int32_t x = sidx % NRES, i = sidx / NRES;
        if (i >= (int32_t)rtot[x] && i < per) d.perm[sidx] = -1;
    }
}

/* The row-list kernel, launched after every update: one workgroup per source
 * chunk (at most BSIM_CHUNK = 4 * ROWS_THREADS pres) builds the chunk's list
 * of active (pre, delay) rows, ascending, from the history words the update
 * just wrote: every thread reads the words of 4 consecutive pres (coalesced),
 * counts their active delays, and a workgroup prefix places the rows.  The
 * lists are read by every tile that receives from the chunk. */
__global__ __launch_bounds__(ROWS_THREADS) void bsim_rows_kernel(bsim_dev d)
{
    __shared__ uint32_t scr[ROWS_THREADS / WAVE];
    const int32_t u = blockIdx.x;
    /* the coming step launch's cycle (the step launch before this one has
     * completed, so it is the same for every workgroup of this launch) */
    const int32_t cyc = device_cycle(d, 0);
    /* one workgroup more than source chunks: it ranks the tiles every
     * d.rank_every cycles, from the work the step launch just recorded,
     * while the others build the lists */
    if (u == d.nsrc) {
        const bool rank = d.rank_every == 1 || (d.rank_every > 1 && cyc % d.rank_every == 0);
        if (rank) rank_tiles(d);
        return;
    }
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & (WAVE - 1);
    const uint32_t w = tid / WAVE;
    const int32_t lo = d.src_lo[u], hi = d.src_hi[u];
    const int32_t par = (cyc - 1) & 1;             /* parity of the last update */
    const uint64_t *hist = d.hist + (int64_t)par * d.n;
    uint32_t *out = d.src_list + d.src_list_base[u];
    const int32_t D = d.dmax;
    const uint64_t dmask = (D >= 63) ? (~0ull >> 1) : ((1ull << D) - 1ull);
    const int32_t p0 = lo + 4 * (int32_t)tid;
    uint64_t m[4];
    uint32_t cnt = 0;
#pragma unroll
    for (int x = 0; x < 4; x++) {
        m[x] = (p0 + x < hi) ? ((hist[p0 + x] >> 1) & dmask) : 0ull;
        cnt += (uint32_t)__popcll(m[x]);
    }
    const uint32_t incl = wave_incl_scan(cnt);
    if (lane == WAVE - 1) scr[w] = incl;
    __syncthreads();
    uint32_t pos = incl - cnt, tot = 0;
#pragma unroll
    for (int i = 0; i < ROWS_THREADS / WAVE; i++) {
        const uint32_t x = scr[i];
        pos += ((uint32_t)i < w) ? x : 0u;
        tot += x;
    }
#pragma unroll
    for (int x = 0; x < 4; x++) {
        const uint32_t rb = (uint32_t)(p0 + x - lo) * (uint32_t)D;
        for (uint64_t mm = m[x]; mm; mm &= mm - 1) out[pos++] = rb + (uint32_t)__builtin_ctzll(mm);
    }
    /* for every tile that reads this chunk: the list's length in the tile's
     * range entry */
    for (int64_t x = d.chunk_tiles_off[u] + tid; x < d.chunk_tiles_off[u + 1]; x += ROWS_THREADS)
        d.rng_count[d.chunk_rng[x]] = tot;
}

extern "C" hipError_t bsim_launch_rows(const bsim_dev *d, hipStream_t stream)
{
    if (d->nsrc <= 0) return hipSuccess;
    (void)hipGetLastError();       /* the check below then reports this launch alone */
    hipLaunchKernelGGL(bsim_rows_kernel, dim3(d->nsrc + 1), dim3(ROWS_THREADS), 0, stream, *d);
    return hipGetLastError();
}

extern "C" size_t bsim_step_lds_bytes(int32_t tile)
{
    return tile >= 2 * NWAVE * WAVE ? lds_bytes(Win<2>::RCAP) : lds_bytes(Win<1>::RCAP);
}

extern "C" hipError_t bsim_launch_step(const bsim_dev *d, int32_t t, int32_t flags,
                                       hipStream_t stream)
{
    if (d->ntiles <= 0) return hipSuccess;
    const dim3 grid(d->nslots), block(NTHREAD);
    const size_t lds = bsim_step_lds_bytes(d->tile);
    (void)hipGetLastError();       /* the check below then reports this launch alone */
    switch (d->tile / (NWAVE * WAVE)) {
    case 1: hipLaunchKernelGGL(bsim_step_kernel<1>, grid, block, lds, stream, *d, t, flags); break;
    case 2: hipLaunchKernelGGL(bsim_step_kernel<2>, grid, block, lds, stream, *d, t, flags); break;
    case 3: hipLaunchKernelGGL(bsim_step_kernel<3>, grid, block, lds, stream, *d, t, flags); break;
    case 4: hipLaunchKernelGGL(bsim_step_kernel<4>, grid, block, lds, stream, *d, t, flags); break;
    case 5: hipLaunchKernelGGL(bsim_step_kernel<5>, grid, block, lds, stream, *d, t, flags); break;
    case 6: hipLaunchKernelGGL(bsim_step_kernel<6>, grid, block, lds, stream, *d, t, flags); break;
    case 7: hipLaunchKernelGGL(bsim_step_kernel<7>, grid, block, lds, stream, *d, t, flags); break;
    case 8: hipLaunchKernelGGL(bsim_step_kernel<8>, grid, block, lds, stream, *d, t, flags); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
