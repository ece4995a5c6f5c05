/*
 * layout.h -- device-side data layout of one simulation (host and HIP code).
 *
 * Neurons are numbered in creation order (the order BSim's CPU simulator
 * updates them and reports spikes in).  They are cut into tiles of TILE
 * consecutive neurons (TILE a multiple of 64); a tile is the unit of work of
 * one wavefront in the step kernel: it receives the synapses whose target lies
 * in the tile and updates the tile's neurons.
 *
 * Synapses are stored tile-major: all synapses into tile k, sorted by
 * (pre, delay_steps, creation index) -- the (pre, delay) CSR of the reference
 * (src/net/BuildNetwork.cpp:96-176) restricted to the tile -- as a 16-bit local
 * target and an fp32 weight.  A tile reads its rows through "source entries":
 * one per source tile s that has synapses into it, covering pres [lo, hi) of
 * s, with an offset table of (hi - lo) * D + 1 uint32 row starts relative to
 * the tile's first synapse (row (p, d) = [off[(p-lo)*D + d-1], next)).
 *
 * Spike history: hist[n] bit j = neuron n fired j cycles ago (bit 0 = this
 * cycle).  After updating cycle t a tile publishes the ascending list of its
 * neurons with any of bits 1..D set, with their history word: exactly the
 * pres whose (pre, d) rows deliver in phase (ii) of cycle t (spike at t - d).
 */
#ifndef BSIM_LAYOUT_H
#define BSIM_LAYOUT_H

#include <stdint.h>

#define BSIM_MAX_DELAY 63   /* history word: bits 1..D of a uint64 */
#define BSIM_MAX_TILE 3840  /* LDS: 16 bytes per target + 2.5 KiB, within 64 KiB */

enum { BSIM_KIND_LIF = 0, BSIM_KIND_ARRAY = 1, BSIM_KIND_BERNOULLI = 2 };

typedef struct bsim_src_entry {
    int32_t src_tile;   /* tile whose active list is read */
    int32_t lo, hi;     /* pre range (global ids) within that tile */
    int32_t pad;
    int64_t off_base;   /* first index into off_tab */
} bsim_src_entry;

/* Kernel arguments (passed by value). */
typedef struct bsim_dev {
    int32_t n;           /* neurons */
    int32_t tile;        /* TILE */
    int32_t ntiles;
    int32_t dmax;        /* D, 1..BSIM_MAX_DELAY */
    /* neuron constants (per neuron) */
    const int8_t *kind;
    const float *Cm, *CE, *CI, *C_E, *C_I, *v_tmp, *v_thresh, *v_reset;
    const int32_t *refrac_time;
    const int32_t *arr_end;     /* array sources: end index into arr_t */
    const int32_t *arr_t;
    const uint32_t *bern_p;
    uint64_t seed;
    /* neuron state */
    float *v, *ie, *ii;
    int32_t *refrac;
    int32_t *arr_cur;
    uint64_t *hist;
    int32_t *fire_count;
    /* synapses */
    const int64_t *tile_syn;    /* [ntiles + 1] first synapse of each tile */
    const uint16_t *syn_tgt;
    const float *syn_w;
    const int32_t *tile_src;    /* [ntiles + 1] first source entry of each tile */
    const bsim_src_entry *src;
    const uint32_t *off_tab;
    /* active lists, double-buffered by cycle parity: [2][ntiles], [2][n] */
    int32_t *act_cnt;
    int32_t *act_ids;
    uint64_t *act_hist;
    /* accounting / recording */
    int64_t *ev_count;          /* [ntiles] deliveries performed */
    uint64_t *raster;           /* [rec_steps][nwords] fired bitmaps, or NULL */
    int32_t rec_steps, nwords;
    int32_t rec_t0;             /* cycle of raster/trace row 0 */
    int32_t n_rec;
    const int32_t *rec_index;   /* [n]: trace column of a neuron or -1 (NULL: none) */
    float *rec_v;               /* [rec_steps][n_rec] */
} bsim_dev;

#endif
