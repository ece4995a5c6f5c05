/*
 * layout.h -- device-side data layout of one simulation (host and HIP code).
 *
 * Neurons are numbered in creation order (the order BSim's CPU simulator
 * updates them and reports spikes in).  They are cut into tiles of TILE
 * consecutive neurons (TILE a multiple of BSIM_TILE_ALIGN = 64 * BSIM_NWAVE).
 * A tile is the unit of work of one workgroup: it receives the synapses whose
 * target lies in the tile and updates the tile's neurons.  Streaming
 * wavefront q (q < BSIM_NWAVE) owns part q of the tile, local targets
 * [q * TILE/BSIM_NWAVE, (q + 1) * TILE/BSIM_NWAVE), and within it lane l owns
 * the local targets congruent to l modulo 64.
 *
 * Target groups: runs of up to 32 consecutive tiles with the same source
 * ranges (the pres with synapses into the tile, merged over gaps of at most
 * 256).  A group's synapses form one contiguous block, stored pre-major:
 * sorted by (source range, pre, delay_steps, tile, part of the target,
 * creation index) -- the (pre, delay) CSR of the reference
 * (src/net/BuildNetwork.cpp:96-176) restricted to the group, each row split
 * by tile and part -- as a tile-local target and an fp32 weight packed in
 * one 8-byte word (bsim_syn).  Sorting
 * inside a row by tile and part keeps, for every target, the (pre, delay,
 * creation) order of its inputs; keeping a row's pieces for all tiles of the
 * group adjacent lets those tiles share cache lines.  Per group and source
 * range, the row table holds for each (pre, delay) row and each tile of the
 * group one aligned 16-byte vector (BSIM_NWAVE part starts, then the end of
 * the tile's sub-rows), offsets relative to the group's first synapse:
 * vector ((p - lo) * D + d - 1) * ntiles_in_group + k.  A tile's range
 * entries (bsim_range, one per source chunk of BSIM_CHUNK pres) point at its
 * vector of row (lo, 1) and give the vector stride per row.
 *
 * Workgroups are dispatched to the 8 XCDs round-robin; every group has an XCD
 * residue and its tiles run in workgroups of that residue, so that they share
 * the XCD's L2.
 *
 * Spike history: hist[n] bit j = neuron n fired j cycles ago (bit 0 = this
 * cycle).  After updating cycle t, a tile writes its neurons' history words
 * and one bit per neuron in the "active" bitmap: set iff any of bits 1..D of
 * the history is set, i.e. iff some (n, d) row delivers in phase (ii) of cycle
 * t.  Both are double-buffered by cycle parity, so that a launch reads the
 * previous cycle's copy while it writes the current one.
 */
#ifndef BSIM_LAYOUT_H
#define BSIM_LAYOUT_H

#include <stdint.h>

#define BSIM_MAX_DELAY 63   /* history word: bits 1..D of a uint64 */
#define BSIM_NWAVE 3                        /* streaming wavefronts per tile = target parts */
#define BSIM_TILE_ALIGN (64 * BSIM_NWAVE)   /* parts of whole 64-lane wavefronts */
#define BSIM_MAX_TILE (8 * BSIM_TILE_ALIGN) /* at most 8 accumulator slots per lane */

enum { BSIM_KIND_LIF = 0, BSIM_KIND_ARRAY = 1, BSIM_KIND_BERNOULLI = 2, BSIM_KIND_POISSON = 3 };

/* Poisson sources: draw X ~ Poisson(rate) from a 32-bit uniform u by the
 * table tab = pois_tab + bern_p[i] of its rate: tab[0] = J entries, then
 * tab[1 + j] = floor(P(X <= j) * 2^32) for j < J - 1; X is the smallest j
 * with u < tab[1 + j], or J - 1 when there is none. */
#define BSIM_POISSON_MAX_RATE 1.0e6f

typedef struct bsim_range {
    int32_t lo, hi;     /* pre neurons [lo, hi) of one chunk of a source range of the tile */
    int64_t off_base;   /* off_tab index of the tile's vector of row (lo, 1) */
    int64_t lbase;      /* first entry of the chunk's active-row list in src_list */
    int32_t uid;        /* source chunk: index into src_* */
    uint32_t wrow;      /* expected deliveries of one active row into the tile, x 16 (host statistics) */
    uint32_t stride;    /* off_tab entries per (pre, delay) row of the group */
    uint32_t pad;
} bsim_range;

/* One synapse on the device: the tile-local target and the fp32 weight bits
 * in one aligned 8-byte word, so that an event is one load of one cache line
 * (with separate 2-byte target and 4-byte weight arrays an event was two
 * loads touching two lines). */
typedef struct __attribute__((aligned(8))) bsim_syn {
    uint32_t tgt;
    uint32_t w;
} bsim_syn;

#define BSIM_MAX_RANGE (1 << 26)   /* pres per source range: row index fits 32 bits */
#define BSIM_CHUNK 4096             /* pres per source chunk (one row-list workgroup) */

/* LIF_brian constants shared by the neurons of one parameter set (usually a
 * population): the reference keeps them per neuron (GLIFENeurons.h:7-24);
 * here a neuron carries only its set and its own v_tmp (which holds its
 * i_offset). */
typedef struct __attribute__((aligned(16))) bsim_lifc {
    float Cm, CE, CI, C_E, C_I, v_thresh, v_reset;
    int32_t refrac_time;
} bsim_lifc;

#define BSIM_CLS_SET_MASK 0x3FFFFFFFu   /* ncls: kind << 30 | parameter set */
#define BSIM_DONE_STRIDE 16              /* uint64 per completion counter: one 128-byte line each */

/* Kernel arguments (passed by value). */
typedef struct bsim_dev {
    int32_t n;           /* neurons */
    int32_t tile;        /* TILE */
    int32_t ntiles;
    int32_t dmax;        /* D, 1..BSIM_MAX_DELAY */
    int32_t nwords;      /* (n + 63) / 64 */
    int32_t pad0;
    /* neuron constants: kind and LIF parameter set per neuron, the sets'
     * shared constants, v_tmp per neuron */
    const uint32_t *ncls;       /* kind << 30 | index into lifc (LIF neurons) */
    const bsim_lifc *lifc;
    const float *v_tmp;
    const int32_t *arr_end;     /* array sources: end index into arr_t */
    const int32_t *arr_t;
    const uint32_t *bern_p;
    const uint64_t *bern_seed;  /* Bernoulli draw of neuron i at cycle t: */
    const uint32_t *bern_key;   /*   hash(bern_seed[i], bern_key[i], t) < bern_p[i] */
    const uint32_t *pois_tab;   /* Poisson sources: inverse-CDF tables (bern_p = offset); their
                                   next firing cycle and end cycle live in arr_cur / arr_end */
    /* neuron state */
    float *v, *ie, *ii;
    int32_t *refrac;
    int32_t *arr_cur;
    int32_t *fire_count;
    uint64_t *hist;             /* [2][n] by cycle parity */
    uint64_t *active;           /* [2][nwords] by cycle parity */
    /* synapses */
    int32_t nslots;             /* step workgroups (tiles placed by XCD residue) */
    int32_t pad3;
    const int64_t *tile_base;   /* [ntiles] first synapse of the tile's group */
    const uint32_t *tile_nsyn;  /* [ntiles] synapses of the tile's group */
    const uint8_t *tile_res;    /* [ntiles] XCD residue of the tile's group */
    const bsim_syn *syn;        /* [S] target and weight of each synapse, group-major */
    const int32_t *tile_rng;    /* [ntiles + 1] first source range of each tile */
    const bsim_range *rng;
    const uint32_t *off_tab;    /* group row tables (16-byte aligned vectors) */
    /* source chunks and their active-row lists: the distinct source ranges of
     * all tiles, cut into chunks of at most BSIM_CHUNK pres (a tile's range
     * entries are per chunk).  After the update of cycle t, list u holds the
     * row index (p - lo) * D + d - 1 of every (pre, delay) of chunk u that
     * delivers in phase (ii) of cycle t, ascending; it is shared by every
     * tile that receives from the chunk */
    int32_t nsrc;
    int32_t pad2;
    const int32_t *src_lo, *src_hi;
    const int64_t *src_list_base;
    uint32_t *src_list;
    uint32_t *rng_count;        /* [range entries] rows in the entry's chunk list this cycle */
    /* accounting / recording */
    int64_t *ev_count;          /* [ntiles] deliveries performed */
    uint64_t *done;             /* [8 * BSIM_DONE_STRIDE]: step workgroups finished by clocked
                                   launches, per XCD residue x at done[BSIM_DONE_STRIDE * x];
                                   each grows by nslots / 8 per launch */
    double done_inv;            /* 8 / nslots */
    int32_t *perm;              /* [nslots] tile of each step workgroup (-1: none) */
    uint64_t *tile_load;        /* [ntiles] work of the launch before a ranking (x 16) */
    int32_t rank_every;         /* tiles ranked heaviest first every this many cycles (0: never,
                                   the tiles stay in their initial order) */
    int32_t pad5;
    uint8_t *tile_cls;          /* [ntiles] its half-octave class */
    const int64_t *chunk_tiles_off;  /* [nsrc + 1] tiles that read each chunk: */
    const int32_t *chunk_rng;        /*   the tile's range entry for the chunk */
    uint64_t *raster;           /* [rec_steps][nwords] fired bitmaps, or NULL */
    int32_t rec_steps;
    int32_t rec_t0;             /* cycle of raster/trace row 0 */
    int32_t n_rec;
    int32_t pad1;
    const int32_t *rec_index;   /* [n]: trace column of a neuron or -1 (NULL: none) */
    float *rec_v;               /* [rec_steps][n_rec] */
    /* state traces of the traced neurons (NULL: off), [rec_steps][n_rec] each:
     * i_E and i_I after the update, and the E / I synaptic input the update
     * consumed (the reference's -DLOG_DATA files g_ie.data, g_ii.data,
     * g_input_e.data, g_input_i.data, src/sim/SingleGPUSimulator.cu:150-216) */
    float *rec_ie, *rec_ii, *rec_in_e, *rec_in_i;
    /* input logs (NULL: off).  in_sum[n]: a LIF neuron's E + I input summed
     * over the cycles it neither was refractory nor fired (the reference GPU
     * simulator's gXInput, src/neuron/lif/lif.cu:116); slot_count[n]: the
     * delay slots of the neuron's spikes visited by the deliveries, i.e. the
     * number of bits 1..D of its history summed over cycles (gLayerInput,
     * src/synapse/static/static.cu:42) */
    float *in_sum;
    int32_t *slot_count;
} bsim_dev;

#endif
