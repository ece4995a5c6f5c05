/*
 * stsim.c -- CPU oracle for the BSim simulation step (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * this file; the product library never links or calls it.  Built with
 * -ffp-contract=off and without -ffast-math so that every REAL operation is
 * rounded exactly as written (the reference CPU build: g++ -O3, no FMA).
 *
 * Precision-generic body: stsim_body.inc.  Semantics: see bsim_oracle.h.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "bsim_oracle.h"

#include <time.h>

static double bso_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- small utilities ----------------------------------------------------- */

typedef struct bso_vec {
    int64_t *v;
    int64_t n, cap;
} bso_vec;

static int bso_vec_push(bso_vec *b, int64_t x)
{
    if (b->n == b->cap) {
        int64_t nc = b->cap ? b->cap * 2 : 256;
        int64_t *nv = (int64_t *)realloc(b->v, (size_t)nc * sizeof(int64_t));
        if (!nv) return -1;
        b->v = nv;
        b->cap = nc;
    }
    b->v[b->n++] = x;
    return 0;
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* LSD radix sort of non-negative int64 keys, 16 bits per pass. */
static void bso_sort_i64(int64_t *v, int64_t n)
{
    if (n < 2) return;
    if (n < 4096) {
        qsort(v, (size_t)n, sizeof(int64_t), cmp_i64);
        return;
    }
    int64_t mx = 0;
    for (int64_t i = 0; i < n; i++) if (v[i] > mx) mx = v[i];
    int64_t *tmp = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!tmp) {
        qsort(v, (size_t)n, sizeof(int64_t), cmp_i64);
        return;
    }
    int64_t *src = v, *dst = tmp;
    static int64_t cnt[65537];
    for (int shift = 0; shift < 64 && (mx >> shift) != 0; shift += 16) {
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; i++) cnt[((src[i] >> shift) & 0xFFFF) + 1]++;
        for (int j = 0; j < 65536; j++) cnt[j + 1] += cnt[j];
        for (int64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 0xFFFF]++] = src[i];
        int64_t *t = src; src = dst; dst = t;
    }
    if (src != v) memcpy(v, src, (size_t)n * sizeof(int64_t));
    free(tmp);
}

/* Stable merge sort of synapse indices by (pre, delay_steps, index). */
static int key3_less(int64_t a, int64_t b, const int32_t *pre, const int32_t *ds)
{
    if (pre[a] != pre[b]) return pre[a] < pre[b];
    if (ds[a] != ds[b]) return ds[a] < ds[b];
    return a < b;
}

static void bso_sort_by_key3(int64_t *v, int64_t n, const int32_t *pre, const int32_t *ds)
{
    if (n < 2) return;
    int64_t *tmp = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!tmp) return;
    for (int64_t width = 1; width < n; width *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * width) {
            int64_t mid = lo + width < n ? lo + width : n;
            int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            int64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) tmp[o++] = key3_less(v[j], v[i], pre, ds) ? v[j++] : v[i++];
            while (i < mid) tmp[o++] = v[i++];
            while (j < hi) tmp[o++] = v[j++];
        }
        memcpy(v, tmp, (size_t)n * sizeof(int64_t));
    }
    free(tmp);
}

/* Sorts the synapse indices of one tile (given in creation order) by
 * (pre, delay_steps, creation): keys pre << 32 | delay << 24 | rank, LSD
 * radix sort with 16-bit digits.  Returns -1 if the tile is too large. */
static int bso_sort_tile(int64_t *v, int64_t n, const int32_t *pre, const int32_t *ds)
{
    if (n < 2) return 0;
    if (n >= (1 << 24)) {
        bso_sort_by_key3(v, n, pre, ds);
        return 0;
    }
    uint64_t *a = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    uint64_t *b = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
    int64_t *c = (int64_t *)malloc(65537 * sizeof(int64_t));
    if (!a || !b || !c) {
        free(a); free(b); free(c);
        return -1;
    }
    for (int64_t i = 0; i < n; i++)
        a[i] = ((uint64_t)(uint32_t)pre[v[i]] << 32) | ((uint64_t)(ds[v[i]] & 0xFF) << 24) | (uint64_t)i;
    for (int shift = 0; shift < 64; shift += 16) {
        memset(c, 0, 65537 * sizeof(int64_t));
        for (int64_t i = 0; i < n; i++) c[((a[i] >> shift) & 0xFFFF) + 1]++;
        for (int j = 0; j < 65536; j++) c[j + 1] += c[j];
        for (int64_t i = 0; i < n; i++) b[c[(a[i] >> shift) & 0xFFFF]++] = a[i];
        uint64_t *t = a; a = b; b = t;
    }
    int64_t *o = (int64_t *)b;  /* reuse: 8-byte slots */
    for (int64_t i = 0; i < n; i++) o[i] = v[a[i] & 0xFFFFFF];
    memcpy(v, o, (size_t)n * sizeof(int64_t));
    free(a); free(b); free(c);
    return 0;
}

uint32_t bso_hash(uint64_t seed, uint32_t id, uint32_t t)
{
    uint64_t z = seed ^ (((uint64_t)id << 32) | (uint64_t)t);
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 32);
}

/* The Poisson pmf from its definition, P(X = j) = lam^j e^-lam / j!, in log
 * form; the table stops where the remaining mass is below 2^-32. */
uint32_t *bso_poisson_table(double lam)
{
    uint32_t J = (uint32_t)ceil(lam + 12.0 * sqrt(lam) + 16.0);
    uint32_t *tab = (uint32_t *)malloc(((size_t)J + 1) * sizeof(uint32_t));
    if (!tab) return NULL;
    tab[0] = J;
    double cdf = 0.0;
    for (uint32_t j = 0; j < J; j++) {
        double lp = (lam > 0.0) ? -lam + (double)j * log(lam) - lgamma((double)j + 1.0)
                                : (j == 0 ? 0.0 : -INFINITY);
        cdf += exp(lp);
        double q = floor(cdf * 4294967296.0);
        tab[1 + j] = q >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)q;
    }
    return tab;
}

uint32_t bso_poisson_pick(const uint32_t *tab, uint32_t u)
{
    uint32_t j = 0;                   /* linear search: the restatement, not the fast form */
    while (j + 1 < tab[0] && u >= tab[1 + j]) j++;
    return j;
}

int bso_csr_order_ok(int32_t n, int64_t s, const int32_t *pre, const int32_t *post,
                     const int32_t *delay_steps)
{
    int64_t *last = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!last) return 0;
    for (int32_t i = 0; i < n; i++) last[i] = -1;
    int ok = 1;
    for (int64_t k = 0; k < s; k++) {
        int64_t key = ((int64_t)pre[k] << 20) | (int64_t)(delay_steps[k] & 0xFFFFF);
        if (key < last[post[k]]) { ok = 0; break; }
        last[post[k]] = key;
    }
    free(last);
    return ok;
}

/* ---- precision instantiations -------------------------------------------- */

#define REAL float
#define SUF f32
#define EXP_R expf
#define FABS_R fabsf
#define ROUND_R roundf
#include "stsim_body.inc"
#undef REAL
#undef SUF
#undef EXP_R
#undef FABS_R
#undef ROUND_R

#define REAL double
#define SUF f64
#define EXP_R exp
#define FABS_R fabs
#define ROUND_R round
#include "stsim_body.inc"
#undef REAL
#undef SUF
#undef EXP_R
#undef FABS_R
#undef ROUND_R
