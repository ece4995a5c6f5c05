/*
 * step.hip -- the simulation step on CDNA4 (gfx950).
 *
 * One launch of bsim_step_kernel advances the network by one cycle.  Each
 * workgroup is a single wavefront (64 lanes) that owns one tile of TILE
 * consecutive neurons (layout.h) and does, for that tile:
 *
 *   (ii) delivery of cycle t-1 -- phase (ii) of the previous cycle: every
 *        synapse into the tile whose pre fired d cycles before it adds its
 *        weight to the target's E or I accumulator (LDS), in the order the
 *        reference's CPU simulator applies them (src/net/Network.cpp:324-327,
 *        src/neuron/lif/LIFENeuron.cpp:107-116);
 *   (i)  the LIF_brian update of cycle t (src/neuron/lif/LIFENeuron.cpp:
 *        169-215), which consumes exactly those accumulators, and the ordered
 *        compaction of the tile's recently-fired neurons into its active list
 *        (wavefront ballot + mbcnt prefix), read by the next launch.
 *
 * Delivery order.  A tile's synapses are stored sorted by (pre, delay,
 * creation); the tile walks active pres in ascending id and their active
 * delays in ascending order, so its event stream is in that order, which the
 * host has checked equals creation order for every target.  Events are taken
 * 64 at a time; lanes whose accumulator (target, sign) also appears in a lower
 * lane of the batch must wait: the lowest lane per accumulator is found with
 * an LDS atomic-min on a per-accumulator tag, applies its add, and the others
 * are carried, in order, to the front of the next batch.  Every accumulator
 * therefore receives its adds one at a time in stream order: the sums are the
 * reference's sequential fp32 sums, bit for bit.
 *
 * Arithmetic: compiled with -ffp-contract=off (no FMA contraction) and
 * without denormal flushing, so V = ((Cm*V + v_tmp) + iE*C_E) + iI*C_I rounds
 * exactly as the reference's CPU code.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

#define WAVE 64
#define SEGQ 256          /* segment descriptors staged per pass */
#define TAG_TOP 0x03FFFFFFu

__device__ __forceinline__ uint32_t bern_hash(uint64_t seed, uint32_t id, uint32_t t)
{
    uint64_t z = seed ^ (((uint64_t)id << 32) | (uint64_t)t);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 32);
}

__device__ __forceinline__ uint64_t ballot64(bool p)
{
    return __builtin_amdgcn_ballot_w64(p);
}

/* number of set bits of m below this lane */
__device__ __forceinline__ uint32_t rank_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* exclusive prefix sum over the wavefront; *total = sum of all lanes */
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t *total)
{
    const int lane = threadIdx.x;
    uint32_t v = x;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        uint32_t y = __shfl_up(v, off, WAVE);
        if (lane >= off) v += y;
    }
    *total = __shfl(v, WAVE - 1, WAVE);
    return v - x;
}

__device__ __forceinline__ uint32_t readlane(uint32_t x, int32_t l)
{
    return __builtin_amdgcn_readlane(x, l);
}

/* Applies up to 64 events (one per lane, lanes >= *ncarry hold new ones, lanes
 * below hold carried ones) to the accumulators; leaves the events that must
 * wait in lanes [0, *ncarry) for the next call.  `round` makes tags unique. */
__device__ __forceinline__ void apply_batch(float *acc, uint32_t *tag, uint2 *carry_buf,
                                            bool valid, uint32_t key, float w,
                                            uint32_t &round, uint32_t &ncarry,
                                            uint32_t &ckey, float &cw)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t tv = ((TAG_TOP - round) << 6) | lane;
    round++;
    if (valid) atomicMin(&tag[key], tv);
    __builtin_amdgcn_wave_barrier();
    bool leader = false;
    if (valid) leader = (tag[key] == tv);
    __builtin_amdgcn_wave_barrier();
    if (leader) {
        const float a = acc[key];
        acc[key] = a + w;
    }
    __builtin_amdgcn_wave_barrier();
    const uint64_t lose = ballot64(valid && !leader);
    if (lose) {
        if (valid && !leader) carry_buf[rank_below(lose)] = make_uint2(key, __float_as_uint(w));
        __builtin_amdgcn_wave_barrier();
        ncarry = (uint32_t)__popcll(lose);
        if (lane < ncarry) {
            const uint2 c = carry_buf[lane];
            ckey = c.x;
            cw = __uint_as_float(c.y);
        }
        __builtin_amdgcn_wave_barrier();
    } else {
        ncarry = 0;
    }
}

__global__ __launch_bounds__(WAVE) void bsim_step_kernel(bsim_dev d, int32_t t, int32_t flags)
{
    extern __shared__ uint32_t lds[];
    const int32_t T = d.tile;
    const int32_t k = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    float *acc = reinterpret_cast<float *>(lds);       /* [2T]: E at 2i, I at 2i+1 */
    uint32_t *tag = lds + 2 * T;                       /* [2T] */
    uint32_t *segq = tag + 2 * T;                      /* [SEGQ][2] */
    uint2 *carry_buf = reinterpret_cast<uint2 *>(segq + 2 * SEGQ);  /* [64] */
    const bool do_deliver = (flags & 1) != 0;
    const bool do_update = (flags & 2) != 0;

    for (int32_t i = lane; i < 2 * T; i += WAVE) {
        acc[i] = 0.0f;
        tag[i] = 0xFFFFFFFFu;
    }
    __syncthreads();

    /* ---- phase (ii) of cycle t-1 ---------------------------------------- */
    if (do_deliver) {
        const int32_t par = (t - 1) & 1;
        const int32_t D = d.dmax;
        const uint64_t dmask = (D >= 63) ? (~0ull >> 1) : ((1ull << D) - 1ull);
        const int64_t syn0 = d.tile_syn[k];
        const uint16_t *tgt = d.syn_tgt + syn0;
        const float *wts = d.syn_w + syn0;
        uint32_t round = 0, ncarry = 0, ckey = 0;
        float cw = 0.0f;
        int64_t events = 0;
        const int32_t e0 = d.tile_src[k], e1 = d.tile_src[k + 1];
        for (int32_t e = e0; e < e1; e++) {
            const bsim_src_entry se = d.src[e];
            const int32_t cnt = d.act_cnt[par * d.ntiles + se.src_tile];
            const int32_t *ids = d.act_ids + (int64_t)par * d.n + (int64_t)se.src_tile * T;
            const uint64_t *hs = d.act_hist + (int64_t)par * d.n + (int64_t)se.src_tile * T;
            for (int32_t b0 = 0; b0 < cnt; b0 += WAVE) {
                const int32_t idx = b0 + (int32_t)lane;
                int32_t p = -1;
                uint64_t m = 0;
                if (idx < cnt) {
                    p = ids[idx];
                    if (p >= se.lo && p < se.hi) m = (hs[idx] >> 1) & dmask;
                }
                uint32_t tot;
                const uint32_t pos = wave_excl_scan((uint32_t)__popcll(m), &tot);
                for (uint32_t w0 = 0; w0 < tot; w0 += SEGQ) {
                    /* stage this lane's rows (ascending delay) */
                    uint64_t mm = m;
                    uint32_t g = pos;
                    while (mm) {
                        const int32_t bit = __builtin_ctzll(mm);
                        mm &= mm - 1;
                        if (g >= w0 && g < w0 + SEGQ) {
                            const int64_t r = se.off_base + (int64_t)(p - se.lo) * D + bit;
                            const uint32_t o0 = d.off_tab[r], o1 = d.off_tab[r + 1];
                            segq[2 * (g - w0)] = o0;
                            segq[2 * (g - w0) + 1] = o1 - o0;
                        }
                        g++;
                    }
                    __syncthreads();
                    const uint32_t nq = min((uint32_t)SEGQ, tot - w0);
                    for (uint32_t q0 = 0; q0 < nq; q0 += WAVE) {
                        /* window of up to 64 rows, one per lane */
                        uint32_t sst = 0, sln = 0;
                        if (q0 + lane < nq) {
                            sst = segq[2 * (q0 + lane)];
                            sln = segq[2 * (q0 + lane) + 1];
                        }
                        uint32_t ew;
                        const uint32_t ep = wave_excl_scan(sln, &ew);
                        const int32_t nseg = (int32_t)min((uint32_t)WAVE, nq - q0);
                        events += ew;
                        uint32_t produced = 0;
                        int32_t cur = 0;
                        while (produced < ew) {
                            /* lanes [ncarry, 64) take window events produced.. */
                            const uint32_t room = WAVE - ncarry;
                            const uint32_t take = min(room, ew - produced);
                            const uint32_t q = produced + (lane - ncarry);
                            const bool mine = lane >= ncarry && lane < ncarry + take;
                            while (cur < nseg && readlane(ep, cur) + readlane(sln, cur) <= produced)
                                cur++;
                            uint32_t addr = 0;
                            for (int32_t j = cur; j < nseg; j++) {
                                const uint32_t ej = readlane(ep, j);
                                if (ej >= produced + take) break;
                                const uint32_t lj = readlane(sln, j);
                                if (mine && q >= ej && q < ej + lj) addr = readlane(sst, j) + (q - ej);
                            }
                            bool valid = lane < ncarry;
                            uint32_t key = ckey;
                            float w = cw;
                            if (mine) {
                                const uint32_t tl = tgt[addr];
                                w = wts[addr];
                                key = 2u * tl + ((w >= 0.0f) ? 0u : 1u);
                                valid = true;
                            }
                            produced += take;
                            apply_batch(acc, tag, carry_buf, valid, key, w, round, ncarry, ckey, cw);
                        }
                    }
                    __syncthreads();
                }
            }
        }
        while (ncarry > 0) {
            apply_batch(acc, tag, carry_buf, lane < ncarry, ckey, cw, round, ncarry, ckey, cw);
        }
        if (lane == 0) d.ev_count[k] += events;
        __syncthreads();
    }

    /* ---- phase (i) of cycle t -------------------------------------------- */
    if (do_update) {
        const int32_t par = t & 1;
        const int32_t D = d.dmax;
        const uint64_t dmask = (D >= 63) ? (~0ull >> 1) : ((1ull << D) - 1ull);
        const int32_t base = k * T;
        int32_t *act_ids = d.act_ids + (int64_t)par * d.n + (int64_t)k * T;
        uint64_t *act_hist = d.act_hist + (int64_t)par * d.n + (int64_t)k * T;
        const bool rec = d.raster != nullptr && t >= d.rec_t0 && t < d.rec_t0 + d.rec_steps;
        uint32_t nact = 0;
        for (int32_t i0 = 0; i0 < T; i0 += WAVE) {
            const int32_t nid = base + i0 + (int32_t)lane;
            const bool valid = nid < d.n;
            bool fired = false;
            uint64_t h = 0;
            if (valid) {
                const int8_t kind = d.kind[nid];
                if (kind == BSIM_KIND_LIF) {
                    const float inE = acc[2 * (i0 + lane)];
                    const float inI = acc[2 * (i0 + lane) + 1];
                    int32_t rs = d.refrac[nid];
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
                h = (d.hist[nid] << 1) | (fired ? 1ull : 0ull);
                d.hist[nid] = h;
                if (fired) d.fire_count[nid] += 1;
                if (rec && d.rec_index) {
                    const int32_t ri = d.rec_index[nid];
                    if (ri >= 0) d.rec_v[(int64_t)(t - d.rec_t0) * d.n_rec + ri] = d.v[nid];
                }
            }
            const uint64_t fm = ballot64(fired);
            if (rec && lane == 0 && base + i0 < d.n)
                d.raster[(int64_t)(t - d.rec_t0) * d.nwords + (base + i0) / 64] = fm;
            const bool act = valid && ((h >> 1) & dmask) != 0;
            const uint64_t am = ballot64(act);
            if (act) {
                const uint32_t r = nact + rank_below(am);
                act_ids[r] = nid;
                act_hist[r] = h;
            }
            nact += (uint32_t)__popcll(am);
        }
        if (lane == 0) d.act_cnt[par * d.ntiles + k] = (int32_t)nact;
    }
}

extern "C" size_t bsim_step_lds_bytes(int32_t tile)
{
    return (size_t)tile * 16 + SEGQ * 8 + WAVE * 8;
}

extern "C" hipError_t bsim_launch_step(const bsim_dev *d, int32_t t, int32_t flags,
                                       hipStream_t stream)
{
    if (d->ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(bsim_step_kernel, dim3(d->ntiles), dim3(WAVE),
                       bsim_step_lds_bytes(d->tile), stream, *d, t, flags);
    return hipGetLastError();
}
