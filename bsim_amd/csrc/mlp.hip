/*
 * mlp.hip -- the rate surrogate: a ReLU multilayer perceptron trained and run
 * on the GPU with bf16 MFMA (v_mfma_f32_16x16x32_bf16, fp32 accumulation).
 *
 * Model: in (padded to 64) -> L hidden layers of H units (ReLU) -> out
 * (padded to a multiple of 64, linear).  Weights are held twice: an fp32
 * master copy updated by Adam and the bf16 copy the GEMMs read.  Activations
 * and activation gradients are stored in bf16 (round to nearest even, done
 * here bit by bit), every product is accumulated in fp32.
 *
 * One training step (bsim_mlp_train):
 *   gather     minibatch rows of the resident data set -> X (bf16), Y (f32)
 *   forward    H_l = bf16(relu(H_{l-1} W_l^T + b_l)),  out = H_L W^T + b (f32)
 *   loss       dZ = bf16(2 (out - Y) / (B * out_dim)); loss and bias-gradient
 *              partials
 *   backward   dZ_{l-1} = bf16(dZ_l W_l) where H_{l-1} > 0, with its
 *              bias-gradient partials; dW_l = dZ_l^T H_{l-1} (split over the
 *              batch, fp32 partials)
 *   adam       sums every gradient's partials in a fixed order, then Adam
 *              (Kingma & Ba, 2015) on the fp32 master, writes the bf16 copy
 *              and its transpose
 * The input, hidden activations and activation gradients of a training step
 * are also written transposed ([unit][row]), so that the weight-gradient
 * GEMM reads both of its operands with the batch (its reduction index)
 * contiguous; likewise the transposed bf16 copy of every weight matrix lets
 * the activation-gradient products read W^T with their reduction index
 * contiguous.
 * Launches per step: one GEMM per layer forward (the first gathers its
 * minibatch rows from the resident set, the last forms the loss and dZ of the
 * output in its epilogue), one GEMM per activation gradient, every weight
 * gradient in one launch (mlp_gemm_partials) and Adam with the transposes in
 * one (mlp_adam_tiles).
 * Every reduction has a fixed order, so results repeat bit for bit.
 *
 * There is no counterpart in the reference; the algorithm is the textbook
 * back-propagation of a fully connected ReLU network and Adam.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/bsim_mlp.h"
#include "internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int WAVE = 64;
constexpr int BM = 64, BN = 64, BK = 64;   /* workgroup tile and k step */
constexpr int LDS_STRIDE = BK + 8;         /* bf16 per LDS tile row (144 B) */
constexpr int GEMM_THREADS = 256;          /* 2 x 2 wavefronts of 32 x 32 */
constexpr int KSPLIT = 4;                  /* batch split of the weight gradients */
constexpr int TRAIN_BATCH_ALIGN = KSPLIT * BK;   /* training batches: whole k steps per split */
constexpr int ROWS_PER_PART = 64;          /* bias-gradient partial: rows per partial */
constexpr int MAX_LAYERS = 10;             /* linear layers (hidden + 1) */

__host__ __device__ inline uint16_t f2bf(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);   /* round to nearest even (finite x) */
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

enum Epi { EPI_RELU_BF16 = 0, EPI_LINEAR_F32 = 1, EPI_DRELU_BF16 = 2, EPI_PARTIAL_F32 = 3, EPI_LOSS = 4 };

/* C[m][n] = sum_k A[m * lda + k] B[n * ldb + k] over this workgroup's k
 * range (both operands k-contiguous). */
struct GemmArgs {
    const uint16_t *A;
    const uint16_t *B;
    int64_t lda, ldb;
    int32_t M, N, K;        /* K: length of one split (blockIdx.z) */
    int32_t pad;
    const float *bias;      /* EPI_RELU_BF16, EPI_LINEAR_F32: [N] */
    uint16_t *out_bf;       /* EPI_RELU_BF16, EPI_DRELU_BF16: [M][ldo] */
    float *out_f;           /* EPI_LINEAR_F32: [M][ldo]; EPI_PARTIAL_F32: [split][M][ldo] */
    int64_t ldo;
    const uint16_t *mask;   /* EPI_DRELU_BF16: keep where mask[m * ldo + n] > 0 */
    float *bpart;           /* EPI_DRELU_BF16: bias-gradient partials [M / 64][N] */
    uint16_t *out_t;        /* EPI_RELU_BF16, EPI_DRELU_BF16, EPI_LOSS: transposed copy [N][ldt] or null */
    int64_t ldt;
    const int32_t *arow;    /* row m of A is row arow[m] of the array A (null: row m) */
    uint16_t *a_t;          /* with K == BK: workgroups of column tile 0 write A^T [K][ldt] */
    const float *y;         /* EPI_LOSS: targets, row m at y[(yrow ? yrow[m] : m) * ny] */
    const int32_t *yrow;
    int32_t ny;             /* EPI_LOSS: valid output columns */
    float scale;            /* EPI_LOSS: dZ = bf16(scale * (C + bias - y)) */
    float *lpart;           /* EPI_LOSS: squared-error sum of each 64 x 64 tile [M / 64][N / 64] */
};

/* One k step of a 64-row operand: thread tid holds rows tid / 8 and
 * tid / 8 + 32, k (tid % 8) * 8 .. + 7. */
struct Tile {
    uint4 v[2];
};

__device__ inline Tile load_tile(const uint16_t *P, int64_t ld, int32_t r0, int32_t k0, uint32_t tid,
                                 const int32_t *rowmap = nullptr)
{
    const uint32_t r = tid >> 3, kc = (tid & 7u) * 8u;
    const int64_t ra = rowmap ? (int64_t)rowmap[r0 + r] : (int64_t)(r0 + r);
    const int64_t rb = rowmap ? (int64_t)rowmap[r0 + r + 32] : (int64_t)(r0 + r + 32);
    Tile t;
    t.v[0] = *(const uint4 *)(P + ra * ld + k0 + kc);
    t.v[1] = *(const uint4 *)(P + rb * ld + k0 + kc);
    return t;
}

__device__ inline void store_tile(uint16_t *T, const Tile &t, uint32_t tid)
{
    const uint32_t r = tid >> 3, kc = (tid & 7u) * 8u;
    *(uint4 *)(T + r * LDS_STRIDE + kc) = t.v[0];
    *(uint4 *)(T + (r + 32) * LDS_STRIDE + kc) = t.v[1];
}

/* Output tile (bx, by) of split bz: C[by * 64 .. + 63][bx * 64 .. + 63]. */
template <int EPI>
__device__ inline void gemm_tile(const GemmArgs &g, int32_t bx, int32_t by, int32_t bz)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[2][BM * LDS_STRIDE];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN * LDS_STRIDE];
    __shared__ float bsum[2][BN];
    __shared__ float lwave[GEMM_THREADS / WAVE];

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & (WAVE - 1), w = tid / WAVE;
    const uint32_t wm = w >> 1, wn = w & 1u;     /* wavefront's 32 x 32 quadrant */
    const int32_t n0 = bx * BN, m0 = by * BM;
    const int32_t kbeg = bz * g.K;
    const int32_t nk = g.K / BK;

    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    /* k step s is multiplied from LDS buffer s & 1; at its start, step s+1
     * (loaded during step s-1) moves from registers to the other buffer and
     * the loads of step s+2 are issued, so that they are in flight during a
     * whole step */
    Tile ra, rb;
    {
        const Tile a0 = load_tile(g.A, g.lda, m0, kbeg, tid, g.arow);
        const Tile b0 = load_tile(g.B, g.ldb, n0, kbeg, tid);
        if (nk > 1) {
            ra = load_tile(g.A, g.lda, m0, kbeg + BK, tid, g.arow);
            rb = load_tile(g.B, g.ldb, n0, kbeg + BK, tid);
        }
        store_tile(As[0], a0, tid);
        store_tile(Bs[0], b0, tid);
    }
    /* the epilogue's other operands for this lane's 16 outputs (rows m0 +
     * 32 wm + 16 i + 4 (lane / 16) + r, column n0 + 32 wn + 16 j + lane %
     * 16): the activation mask (EPI_DRELU_BF16) is loaded while the products
     * run; of the targets (EPI_LOSS) only the row indices are -- the target
     * loads themselves are issued after the products, since loads issued
     * before them hold up every wait of the k loop (the wait counter is in
     * order) and the gathered targets are a chain of two loads */
    uint16_t mk[2][2][4];
    float yv[2][2][4];
    int32_t yr[2][4];
    const uint32_t rq0 = (lane >> 4) * 4u, c16 = lane & 15u;
    if (EPI == EPI_DRELU_BF16) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int32_t row = m0 + (int32_t)(wm * 32 + i * 16 + rq0) + r;
                    const int32_t col = n0 + (int32_t)(wn * 32 + j * 16 + c16);
                    mk[i][j][r] = g.mask[(int64_t)row * g.ldo + col];
                }
    }
    if (EPI == EPI_LOSS) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int32_t row = m0 + (int32_t)(wm * 32 + i * 16 + rq0) + r;
                yr[i][r] = g.yrow ? g.yrow[row] : row;
            }
    }
    __syncthreads();
    /* the first layer of a training step: A's rows are gathered, and the
     * workgroups of column tile 0 also write them transposed (one k step) */
    if (g.a_t && bx == 0 && nk == 1) {
        for (uint32_t x = tid; x < (uint32_t)(BM * BK); x += GEMM_THREADS) {
            const uint32_t k = x / BM, r = x % BM;
            g.a_t[(int64_t)(kbeg + (int32_t)k) * g.ldt + m0 + (int32_t)r] = As[0][r * LDS_STRIDE + k];
        }
    }

    const uint32_t r16 = lane & 15u, kg = (lane >> 4) * 8u;
    for (int32_t s = 0; s < nk; s++) {
        const int cur = s & 1;
        if (s + 1 < nk) {
            store_tile(As[cur ^ 1], ra, tid);
            store_tile(Bs[cur ^ 1], rb, tid);
            if (s + 2 < nk) {
                ra = load_tile(g.A, g.lda, m0, kbeg + (s + 2) * BK, tid, g.arow);
                rb = load_tile(g.B, g.ldb, n0, kbeg + (s + 2) * BK, tid);
            }
        }
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++)
                a[i] = *(const bf16x8 *)(As[cur] + (wm * 32 + i * 16 + r16) * LDS_STRIDE + kk + kg);
#pragma unroll
            for (int j = 0; j < 2; j++)
                b[j] = *(const bf16x8 *)(Bs[cur] + (wn * 32 + j * 16 + r16) * LDS_STRIDE + kk + kg);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    if (EPI == EPI_LOSS) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int32_t col = n0 + (int32_t)(wn * 32 + j * 16 + c16);
                    yv[i][j][r] = col < g.ny ? g.y[(int64_t)yr[i][r] * g.ny + col] : 0.f;
                }
    }
    /* lane holds C[row0 + 4 * (lane / 16) + r][col0 + lane % 16], r = 0..3 */
    const uint32_t rq = (lane >> 4) * 4u;
    float csum[2] = {0.f, 0.f};
    float lsum = 0.f;
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int32_t col = n0 + wn * 32 + j * 16 + r16;
            const int32_t row0 = m0 + wm * 32 + i * 16 + rq;
            uint16_t h[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int32_t row = row0 + r;
                const float v = acc[i][j][r];
                if (EPI == EPI_RELU_BF16) {
                    const float z = v + g.bias[col];
                    h[r] = f2bf(z > 0.f ? z : 0.f);
                    g.out_bf[(int64_t)row * g.ldo + col] = h[r];
                } else if (EPI == EPI_LINEAR_F32) {
                    g.out_f[(int64_t)row * g.ldo + col] = v + g.bias[col];
                } else if (EPI == EPI_DRELU_BF16) {
                    const uint16_t m = mk[i][j][r];
                    h[r] = (m != 0 && !(m & 0x8000u)) ? f2bf(v) : (uint16_t)0;
                    g.out_bf[(int64_t)row * g.ldo + col] = h[r];
                    csum[j] += bf2f(h[r]);
                } else if (EPI == EPI_LOSS) {
                    /* the output (as EPI_LINEAR_F32 rounds it), its error and dZ */
                    const float out = v + g.bias[col];
                    h[r] = 0;
                    if (col < g.ny) {
                        const float diff = out - yv[i][j][r];
                        lsum += diff * diff;
                        h[r] = f2bf(g.scale * diff);
                    }
                    g.out_bf[(int64_t)row * g.ldo + col] = h[r];
                    csum[j] += bf2f(h[r]);
                } else {
                    g.out_f[((int64_t)bz * g.M + row) * g.ldo + col] = v;
                }
            }
            if ((EPI == EPI_RELU_BF16 || EPI == EPI_DRELU_BF16 || EPI == EPI_LOSS) && g.out_t) {
                /* 4 consecutive rows of one column: one 8-byte store */
                const uint2 q = {(uint32_t)h[0] | ((uint32_t)h[1] << 16),
                                 (uint32_t)h[2] | ((uint32_t)h[3] << 16)};
                *(uint2 *)(g.out_t + (int64_t)col * g.ldt + row0) = q;
            }
        }
    }
    if (EPI == EPI_DRELU_BF16 || EPI == EPI_LOSS) {
        /* column sums over the workgroup's 64 rows, in a fixed order */
#pragma unroll
        for (int j = 0; j < 2; j++) {
            float x = csum[j];
            x += __shfl_xor(x, 16);
            x += __shfl_xor(x, 32);
            if (lane < 16) bsum[wm][wn * 32 + j * 16 + lane] = x;
        }
        if (EPI == EPI_LOSS) {
            /* the tile's squared error: lane, wavefront, then the four
             * wavefronts, in a fixed order */
#pragma unroll
            for (int o = 1; o < WAVE; o <<= 1) lsum += __shfl_xor(lsum, o);
            if (lane == 0) lwave[w] = lsum;
        }
        __syncthreads();
        if (tid < BN) g.bpart[(int64_t)by * g.N + n0 + tid] = bsum[0][tid] + bsum[1][tid];
        if (EPI == EPI_LOSS && tid == 0)
            g.lpart[(int64_t)by * (g.N / BN) + bx] = ((lwave[0] + lwave[1]) + lwave[2]) + lwave[3];
    }
}

template <int EPI>
__global__ __launch_bounds__(GEMM_THREADS) void mlp_gemm(GemmArgs g)
{
    gemm_tile<EPI>(g, blockIdx.x, blockIdx.y, blockIdx.z);
}

/* The weight-gradient GEMMs of every layer in one launch: workgroups
 * first[l] .. first[l + 1] - 1 take layer l's tiles, split by split. */
struct GemmGroup {
    GemmArgs g[MAX_LAYERS];
    int32_t first[MAX_LAYERS + 1];
    int32_t n;
};

__global__ __launch_bounds__(GEMM_THREADS) void mlp_gemm_partials(GemmGroup a)
{
    int l = 0;
    while (l + 1 < a.n && (int32_t)blockIdx.x >= a.first[l + 1]) l++;
    const GemmArgs &g = a.g[l];
    const int32_t t = (int32_t)blockIdx.x - a.first[l];
    const int32_t tn = g.N / BN, tmn = tn * (g.M / BM);
    const int32_t r = t % tmn;
    gemm_tile<EPI_PARTIAL_F32>(g, r % tn, r / tn, t / tmn);
}

struct AdamSeg {
    int64_t off, count;   /* into the flat parameter arrays */
    const float *parts;   /* gradient partials, [nparts][count] */
    int32_t nparts;
    int32_t pad;
};

struct AdamArgs {
    AdamSeg seg[2 * MAX_LAYERS];
    int32_t nseg;
    int32_t update;       /* 0: only write the summed gradient */
    float lr, b1, b2, eps, c1, c2;   /* c1 = 1 - b1^t, c2 = 1 - b2^t */
    float *p, *m, *v, *grad;
    uint16_t *pbf;
    const float *lpart;   /* loss partials of the step, summed by wavefront 0 of workgroup 0 */
    int32_t nlpart;
    float lscale;
    float *loss;
};

/* Adam (Kingma & Ba, 2015) for the j-th parameter of segment sg: sums its
 * gradient partials in a fixed order, updates the fp32 master (m, v in fp32)
 * and writes the bf16 copy; returns the new bf16 value (0 when only the
 * gradient is written). */
__device__ inline uint16_t adam_elem(const AdamArgs &a, const AdamSeg &sg, int64_t j)
{
    const int64_t i = sg.off + j;
    float gsum = 0.f;
#pragma unroll 8
    for (int32_t k = 0; k < sg.nparts; k++) gsum += sg.parts[(int64_t)k * sg.count + j];
    if (a.grad) a.grad[i] = gsum;
    if (!a.update) return 0;
    const float m = a.b1 * a.m[i] + (1.f - a.b1) * gsum;
    const float v = a.b2 * a.v[i] + (1.f - a.b2) * (gsum * gsum);
    a.m[i] = m;
    a.v[i] = v;
    const float mhat = m / a.c1, vhat = v / a.c2;
    const float p = a.p[i] - a.lr * (mhat / (sqrtf(vhat) + a.eps));
    a.p[i] = p;
    const uint16_t h = f2bf(p);
    a.pbf[i] = h;
    return h;
}

/* adam_elem for E parameters j[0..E-1] of one segment at once: every load
 * of the E is issued before the first store (the same arithmetic, element by
 * element). */
template <int E>
__device__ inline void adam_elems(const AdamArgs &a, const AdamSeg &sg, const int64_t (&j)[E], uint16_t (&h)[E])
{
    float g[E], mo[E], vo[E], po[E];
#pragma unroll
    for (int e = 0; e < E; e++) g[e] = 0.f;
    for (int32_t k = 0; k < sg.nparts; k++)
#pragma unroll
        for (int e = 0; e < E; e++) g[e] += sg.parts[(int64_t)k * sg.count + j[e]];
    if (a.grad)
#pragma unroll
        for (int e = 0; e < E; e++) a.grad[sg.off + j[e]] = g[e];
    if (!a.update) {
#pragma unroll
        for (int e = 0; e < E; e++) h[e] = 0;
        return;
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
        mo[e] = a.m[sg.off + j[e]];
        vo[e] = a.v[sg.off + j[e]];
        po[e] = a.p[sg.off + j[e]];
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int64_t i = sg.off + j[e];
        const float m = a.b1 * mo[e] + (1.f - a.b1) * g[e];
        const float v = a.b2 * vo[e] + (1.f - a.b2) * (g[e] * g[e]);
        a.m[i] = m;
        a.v[i] = v;
        const float mhat = m / a.c1, vhat = v / a.c2;
        const float p = po[e] - a.lr * (mhat / (sqrtf(vhat) + a.eps));
        a.p[i] = p;
        h[e] = f2bf(p);
        a.pbf[i] = h[e];
    }
}

/* Adam over every parameter, with the transposed bf16 weights written
 * alongside (the copy the activation-gradient products read): workgroups 0 .. wfirst[nlin] - 1 take
 * one 64 x 64 tile of a weight matrix each and transpose its new bf16 values
 * through LDS, the others 256 bias entries each. */
struct AdamTileArgs {
    AdamArgs a;
    int32_t n[MAX_LAYERS], k[MAX_LAYERS];   /* W_l is [n][k] */
    int32_t wfirst[MAX_LAYERS + 1];         /* first tile workgroup of W_l */
    int32_t bfirst[MAX_LAYERS + 1];         /* first bias workgroup of layer l, after the tiles */
    int32_t nlin;
    uint16_t *pbf_t;                        /* W_l^T [k][n] at the offset of W_l */
};

constexpr int ADAM_THREADS = 1024;                   /* 4 parameters of a tile per thread */
constexpr int ADAM_ROWS_STEP = ADAM_THREADS / 64;

__global__ __launch_bounds__(ADAM_THREADS) void mlp_adam_tiles(AdamTileArgs t)
{
    __shared__ uint16_t tile[64][66];
    const AdamArgs &a = t.a;
    if (blockIdx.x == 0 && threadIdx.x < WAVE && a.loss) {
        /* the step's loss from the output tiles' partials: lane l adds
         * partials l, l + 64, ... in order (their loads issued together),
         * then a fixed butterfly combines the lanes; lane 0's total is
         * written (one thread adding them one by one waited for each load
         * in turn, which made this workgroup the last of the launch) */
        float s = 0.f;
#pragma unroll 4
        for (int32_t k = (int32_t)threadIdx.x; k < a.nlpart; k += WAVE) s += a.lpart[k];
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) s += __shfl_xor(s, o);
        if (threadIdx.x == 0) *a.loss = s * a.lscale;
    }
    const int32_t b = (int32_t)blockIdx.x;
    if (b < t.wfirst[t.nlin]) {
        int l = 0;
        while (l + 1 < t.nlin && b >= t.wfirst[l + 1]) l++;
        const int32_t ti = b - t.wfirst[l], tk = t.k[l] / 64;
        const int32_t r0 = (ti / tk) * 64, c0 = (ti % tk) * 64;
        const AdamSeg &sg = a.seg[2 * l];
        const uint32_t x = threadIdx.x & 63u, y = threadIdx.x >> 6;
        constexpr int E = 64 / ADAM_ROWS_STEP;
        int64_t j[E];
        uint16_t h[E];
#pragma unroll
        for (int e = 0; e < E; e++) j[e] = (int64_t)(r0 + (int32_t)y + e * ADAM_ROWS_STEP) * t.k[l] + c0 + x;
        adam_elems<E>(a, sg, j, h);
#pragma unroll
        for (int e = 0; e < E; e++) tile[y + e * ADAM_ROWS_STEP][x] = h[e];
        if (!a.update) return;                    /* uniform */
        __syncthreads();
        uint16_t *wt = t.pbf_t + sg.off;
        for (uint32_t c = y; c < 64; c += ADAM_ROWS_STEP) wt[(int64_t)(c0 + c) * t.n[l] + r0 + x] = tile[x][c];
    } else {
        const int32_t bb = b - t.wfirst[t.nlin];
        int l = 0;
        while (l + 1 < t.nlin && bb >= t.bfirst[l + 1]) l++;
        const AdamSeg &sg = a.seg[2 * l + 1];
        const int64_t j = (int64_t)(bb - t.bfirst[l]) * ADAM_THREADS + threadIdx.x;
        if (j < sg.count) (void)adam_elem(a, sg, j);
    }
}

/* Input conversion: f32 [n][in_dim] -> bf16 [n][kin] (zero padded). */
__global__ __launch_bounds__(256) void mlp_convert_x(const float *x, uint16_t *xb, uint16_t *xb_t,
                                                     int64_t n, int32_t in_dim, int32_t kin)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * kin) return;
    const int64_t r = i / kin;
    const int32_t c = (int32_t)(i % kin);
    const uint16_t v = c < in_dim ? f2bf(x[r * in_dim + c]) : (uint16_t)0;
    xb[i] = v;
    if (xb_t) xb_t[(int64_t)c * n + r] = v;
}

struct TransArgs {
    int64_t off[MAX_LAYERS];    /* weight matrix l at off[l] in both copies */
    int32_t n[MAX_LAYERS], k[MAX_LAYERS];
    int32_t first[MAX_LAYERS + 1];   /* first 64 x 64 tile of matrix l */
    int32_t nlin;
    const uint16_t *w;           /* W_l [n][k] */
    uint16_t *wt;                /* W_l^T [k][n] */
};

/* bf16 weight copies transposed through LDS, one 64 x 64 tile per workgroup. */
__global__ __launch_bounds__(256) void mlp_transpose_w(TransArgs a)
{
    __shared__ uint16_t t[64][66];
    int l = 0;
    while (l + 1 < a.nlin && (int32_t)blockIdx.x >= a.first[l + 1]) l++;
    const int32_t tile = blockIdx.x - a.first[l];
    const int32_t tk = a.k[l] / 64;
    const int32_t r0 = (tile / tk) * 64, c0 = (tile % tk) * 64;
    const uint16_t *w = a.w + a.off[l];
    uint16_t *wt = a.wt + a.off[l];
    const uint32_t x = threadIdx.x & 63u, y = threadIdx.x >> 6;
    for (uint32_t r = y; r < 64; r += 4) t[r][x] = w[(int64_t)(r0 + r) * a.k[l] + c0 + x];
    __syncthreads();
    for (uint32_t c = y; c < 64; c += 4) wt[(int64_t)(c0 + c) * a.n[l] + r0 + x] = t[x][c];
}

template <int EPI>
hipError_t gemm(const GemmArgs &g, int32_t splits, hipStream_t st)
{
    const dim3 grid(g.N / BN, g.M / BM, splits);
    (void)hipGetLastError();   /* the check after the launch reports it alone */
    hipLaunchKernelGGL((mlp_gemm<EPI>), grid, dim3(GEMM_THREADS), 0, st, g);
    return hipGetLastError();
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

}  // namespace

struct bsim_mlp {
    bsim_mlp_config cfg{};
    int32_t nlin = 0;                 /* linear layers */
    int32_t kdim[MAX_LAYERS + 1]{};   /* padded widths: kdim[0] input, kdim[nlin] output */
    int64_t woff[MAX_LAYERS]{}, boff[MAX_LAYERS]{};
    int64_t nparam = 0;
    int32_t step = 0;                 /* Adam steps taken */
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *pbf = nullptr;
    uint16_t *pbf_t = nullptr;        /* weights transposed (W_l^T [k][n] at woff[l]) */
    uint16_t *act[MAX_LAYERS + 1]{};  /* act[0] input, act[l] hidden l (bf16) */
    float *out = nullptr, *ybat = nullptr;
    uint16_t *dz[2]{};                /* activation gradients (ping-pong) */
    uint16_t *act_t[MAX_LAYERS]{};    /* act[l] transposed, [kdim[l]][batch] */
    float *wpart[MAX_LAYERS]{}, *bpart[MAX_LAYERS]{};
    float *lpart = nullptr, *loss = nullptr;
    uint16_t *xall = nullptr;
    float *yall = nullptr;
    int64_t ndata = 0;
    float *xstage = nullptr;          /* predict input staging (f32) */
    uint16_t *dzt[MAX_LAYERS]{};      /* dZ of layer l's output, transposed [kdim[l+1]][batch] */
};

namespace {

int mfail(int code, const std::string &msg)
{
    g_bsim_err = msg;
    return code;
}

int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return BSIM_OK;
    return mfail(BSIM_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
int dalloc(bsim_mlp *m, T **p, int64_t count)
{
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, (size_t)std::max<int64_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();   /* the failure is reported here, not by a later launch check */
        return hip_check(e, "hipMalloc");
    }
    m->allocs.push_back(q);
    *p = (T *)q;
    return BSIM_OK;
}

/* uniform in [-1, 1) from (seed, index) */
double hash_uniform(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

/* Forward pass of a batch.  predict (train false): from act[0], fp32 output
 * in m->out.  train: the hidden activations also written transposed, and the
 * output layer forms the loss partials, dZ of the output (row-major and
 * transposed) and its bias partials against the targets (EPI_LOSS); with
 * rows, layer 0 gathers its inputs from the resident set (writing them
 * transposed) and the targets are the resident ones, else act[0] / act_t[0]
 * and ybat hold them. */
int forward(bsim_mlp *m, int32_t batch, bool train, const int32_t *rows = nullptr)
{
    const int L = m->nlin;
    for (int l = 0; l < L; l++) {
        GemmArgs g{};
        g.A = (l == 0 && rows) ? m->xall : m->act[l];
        g.arow = l == 0 ? rows : nullptr;
        g.a_t = (l == 0 && rows) ? m->act_t[0] : nullptr;
        g.lda = m->kdim[l];
        g.B = m->pbf + m->woff[l];
        g.ldb = m->kdim[l];
        g.M = batch;
        g.N = m->kdim[l + 1];
        g.K = m->kdim[l];
        g.bias = m->p + m->boff[l];
        g.ldo = m->kdim[l + 1];
        g.ldt = batch;
        hipError_t e;
        if (l + 1 < L) {
            g.out_bf = m->act[l + 1];
            g.out_t = train ? m->act_t[l + 1] : nullptr;
            e = gemm<EPI_RELU_BF16>(g, 1, m->stream);
        } else if (!train) {
            g.out_f = m->out;
            e = gemm<EPI_LINEAR_F32>(g, 1, m->stream);
        } else {
            g.out_bf = m->dz[0];
            g.out_t = m->dzt[L - 1];
            g.bpart = m->bpart[L - 1];
            g.lpart = m->lpart;
            g.y = rows ? m->yall : m->ybat;
            g.yrow = rows;
            g.ny = m->cfg.out_dim;
            g.scale = 2.0f / ((float)batch * (float)m->cfg.out_dim);
            e = gemm<EPI_LOSS>(g, 1, m->stream);
        }
        if (e != hipSuccess) return hip_check(e, "forward launch");
    }
    return BSIM_OK;
}

/* Backward pass of a batch after a training forward pass (which formed dZ
 * of the output): gradient partials left in wpart / bpart.  The activation
 * gradients go down the layers one GEMM each, their transposed copies kept
 * per layer, and then every weight gradient is formed in one launch. */
int backward(bsim_mlp *m, int32_t batch)
{
    const int L = m->nlin;
    hipError_t e;
    int cur = 0;
    for (int l = L - 1; l >= 1; l--) {
        /* dZ_{l-1} [batch][kdim[l]] = (dZ W_l) where H > 0:
         * B(n = in unit, k = out unit) = W[k][n] = W^T[n][k] */
        GemmArgs gx{};
        gx.A = m->dz[cur];
        gx.lda = m->kdim[l + 1];
        gx.B = m->pbf_t + m->woff[l];
        gx.ldb = m->kdim[l + 1];
        gx.M = batch;
        gx.N = m->kdim[l];
        gx.K = m->kdim[l + 1];
        gx.out_bf = m->dz[cur ^ 1];
        gx.ldo = m->kdim[l];
        gx.mask = m->act[l];
        gx.bpart = m->bpart[l - 1];
        gx.out_t = m->dzt[l - 1];
        gx.ldt = batch;
        e = gemm<EPI_DRELU_BF16>(gx, 1, m->stream);
        if (e != hipSuccess) return hip_check(e, "activation-gradient launch");
        cur ^= 1;
    }
    /* dW_l [kdim[l+1]][kdim[l]] = dZ_l^T H_l: A(m = out unit, k = row) = dzt[l][m][k],
     * B(n = in unit, k = row) = act_t[l][n][k]; every layer in one launch */
    GemmGroup g{};
    g.n = L;
    for (int l = 0; l < L; l++) {
        GemmArgs &x = g.g[l];
        x.A = m->dzt[l];
        x.lda = batch;
        x.B = m->act_t[l];
        x.ldb = batch;
        x.M = m->kdim[l + 1];
        x.N = m->kdim[l];
        x.K = batch / KSPLIT;
        x.out_f = m->wpart[l];
        x.ldo = m->kdim[l];
        g.first[l + 1] = g.first[l] + (x.M / BM) * (x.N / BN) * KSPLIT;
    }
    (void)hipGetLastError();   /* the check after the launch reports it alone */
    hipLaunchKernelGGL(mlp_gemm_partials, dim3(g.first[L]), dim3(GEMM_THREADS), 0, m->stream, g);
    return hip_check(hipGetLastError(), "weight-gradient launch");
}

int adam(bsim_mlp *m, int32_t batch, bool update, float lr, float *grad_out, float *loss_dev)
{
    AdamArgs a{};
    a.nseg = 2 * m->nlin;
    for (int l = 0; l < m->nlin; l++) {
        a.seg[2 * l] = AdamSeg{m->woff[l], (int64_t)m->kdim[l + 1] * m->kdim[l], m->wpart[l], KSPLIT, 0};
        a.seg[2 * l + 1] = AdamSeg{m->boff[l], m->kdim[l + 1], m->bpart[l], batch / ROWS_PER_PART, 0};
    }
    a.update = update ? 1 : 0;
    a.lr = lr;
    a.b1 = m->cfg.beta1;
    a.b2 = m->cfg.beta2;
    a.eps = m->cfg.eps;
    const int t = m->step + 1;
    a.c1 = (float)(1.0 - std::pow((double)m->cfg.beta1, t));
    a.c2 = (float)(1.0 - std::pow((double)m->cfg.beta2, t));
    a.p = m->p;
    a.m = m->m;
    a.v = m->v;
    a.grad = grad_out;
    a.pbf = m->pbf;
    a.lpart = m->lpart;
    a.nlpart = batch / ROWS_PER_PART * (m->kdim[m->nlin] / 64);
    a.lscale = 1.0f / ((float)batch * (float)m->cfg.out_dim);
    a.loss = loss_dev;
    AdamTileArgs ta{};
    ta.a = a;
    ta.nlin = m->nlin;
    ta.pbf_t = m->pbf_t;
    for (int l = 0; l < m->nlin; l++) {
        ta.n[l] = m->kdim[l + 1];
        ta.k[l] = m->kdim[l];
        ta.wfirst[l + 1] = ta.wfirst[l] + (ta.n[l] / 64) * (ta.k[l] / 64);
        ta.bfirst[l + 1] = ta.bfirst[l] + (ta.n[l] + ADAM_THREADS - 1) / ADAM_THREADS;
    }
    (void)hipGetLastError();   /* the check after the launch reports it alone */
    hipLaunchKernelGGL(mlp_adam_tiles, dim3((uint32_t)(ta.wfirst[m->nlin] + ta.bfirst[m->nlin])),
                       dim3(ADAM_THREADS), 0, m->stream, ta);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_check(e, "adam launch");
    if (update) m->step++;
    return BSIM_OK;
}

int transpose_weights(bsim_mlp *m)
{
    TransArgs a{};
    a.nlin = m->nlin;
    a.first[0] = 0;
    for (int l = 0; l < m->nlin; l++) {
        a.off[l] = m->woff[l];
        a.n[l] = m->kdim[l + 1];
        a.k[l] = m->kdim[l];
        a.first[l + 1] = a.first[l] + (a.n[l] / 64) * (a.k[l] / 64);
    }
    a.w = m->pbf;
    a.wt = m->pbf_t;
    (void)hipGetLastError();   /* the check after the launch reports it alone */
    hipLaunchKernelGGL(mlp_transpose_w, dim3(a.first[m->nlin]), dim3(256), 0, m->stream, a);
    return hip_check(hipGetLastError(), "transpose launch");
}

int check_batch(const bsim_mlp *m, int32_t batch)
{
    if (batch <= 0 || batch % TRAIN_BATCH_ALIGN != 0 || batch > m->cfg.max_batch)
        return mfail(BSIM_E_ARG, "batch must be a positive multiple of " +
                                     std::to_string(TRAIN_BATCH_ALIGN) + " and at most max_batch (" +
                                     std::to_string(m->cfg.max_batch) + ")");
    return BSIM_OK;
}

}  // namespace

extern "C" {

bsim_mlp *bsim_mlp_create(const bsim_mlp_config *cfg)
{
    if (!cfg) {
        g_bsim_err = "null config";
        return nullptr;
    }
    const bsim_mlp_config c = *cfg;
    if (c.in_dim < 1 || c.in_dim > 64 || c.hidden < 64 || c.hidden % 64 != 0 || c.hidden > 4096 ||
        c.layers < 1 || c.layers > MAX_LAYERS - 1 || c.out_dim < 1 || c.out_dim > 4096 ||
        c.max_batch < TRAIN_BATCH_ALIGN || c.max_batch % TRAIN_BATCH_ALIGN != 0 ||
        c.max_batch > (1 << 20) || !(c.beta1 >= 0.f && c.beta1 < 1.f) ||
        !(c.beta2 >= 0.f && c.beta2 < 1.f) || !(c.eps > 0.f)) {
        g_bsim_err = "invalid MLP config: in_dim 1..64, hidden a multiple of 64 up to 4096, "
                     "layers 1..9, out_dim 1..4096, max_batch a multiple of 256 up to 2^20, "
                     "0 <= beta < 1, eps > 0";
        return nullptr;
    }
    bsim_mlp *m = new (std::nothrow) bsim_mlp();
    if (!m) {
        g_bsim_err = "out of host memory";
        return nullptr;
    }
    m->cfg = c;
    m->nlin = c.layers + 1;
    m->kdim[0] = 64;
    for (int l = 1; l < m->nlin; l++) m->kdim[l] = c.hidden;
    m->kdim[m->nlin] = (int32_t)round_up(c.out_dim, 64);
    int64_t off = 0;
    for (int l = 0; l < m->nlin; l++) {
        m->woff[l] = off;
        off += (int64_t)m->kdim[l + 1] * m->kdim[l];
        m->boff[l] = off;
        off += m->kdim[l + 1];
    }
    m->nparam = off;
    auto bad = [&](int) {
        bsim_mlp_destroy(m);
        return (bsim_mlp *)nullptr;
    };
    if (hip_check(hipSetDevice(c.device), "hipSetDevice") != BSIM_OK) return bad(0);
    if (hip_check(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "stream") != BSIM_OK)
        return bad(0);
    const int64_t B = c.max_batch;
    int rc = dalloc(m, &m->p, m->nparam);
    if (!rc) rc = dalloc(m, &m->m, m->nparam);
    if (!rc) rc = dalloc(m, &m->v, m->nparam);
    if (!rc) rc = dalloc(m, &m->grad, m->nparam);
    if (!rc) rc = dalloc(m, &m->pbf, m->nparam);
    if (!rc) rc = dalloc(m, &m->pbf_t, m->nparam);
    for (int l = 0; l < m->nlin && !rc; l++) rc = dalloc(m, &m->act[l], B * m->kdim[l]);
    if (!rc) rc = dalloc(m, &m->out, B * m->kdim[m->nlin]);
    if (!rc) rc = dalloc(m, &m->ybat, B * c.out_dim);
    int32_t wmax = 0;
    for (int l = 0; l <= m->nlin; l++) wmax = std::max(wmax, m->kdim[l]);
    if (!rc) rc = dalloc(m, &m->dz[0], B * wmax);
    if (!rc) rc = dalloc(m, &m->dz[1], B * wmax);
    for (int l = 0; l < m->nlin && !rc; l++) rc = dalloc(m, &m->act_t[l], B * m->kdim[l]);
    for (int l = 0; l < m->nlin && !rc; l++) {
        rc = dalloc(m, &m->wpart[l], (int64_t)KSPLIT * m->kdim[l + 1] * m->kdim[l]);
        if (!rc) rc = dalloc(m, &m->bpart[l], B / ROWS_PER_PART * m->kdim[l + 1]);
        if (!rc) rc = dalloc(m, &m->dzt[l], B * m->kdim[l + 1]);
    }
    if (!rc) rc = dalloc(m, &m->lpart, B / ROWS_PER_PART * (m->kdim[m->nlin] / 64));
    if (!rc) rc = dalloc(m, &m->loss, 1);
    if (!rc) rc = dalloc(m, &m->xstage, B * c.in_dim);
    if (rc != BSIM_OK) return bad(0);
    /* He-uniform weights (bound sqrt(6 / fan_in)), zero biases, from the seed */
    std::vector<float> init(m->nparam, 0.f);
    for (int l = 0; l < m->nlin; l++) {
        const int32_t fan_in = l == 0 ? c.in_dim : c.hidden;
        const int32_t nrow = l + 1 == m->nlin ? c.out_dim : c.hidden;
        const double bound = std::sqrt(6.0 / fan_in);
        for (int32_t r = 0; r < nrow; r++)
            for (int32_t k = 0; k < fan_in; k++) {
                const int64_t i = m->woff[l] + (int64_t)r * m->kdim[l] + k;
                init[i] = (float)(bound * hash_uniform(c.seed, (uint64_t)i));
            }
    }
    if (bsim_mlp_set_params(m, init.data()) != BSIM_OK) return bad(0);
    return m;
}

void bsim_mlp_destroy(bsim_mlp *m)
{
    if (!m) return;
    if (m->stream) hipStreamSynchronize(m->stream);
    for (void *q : m->allocs) hipFree(q);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
}

int64_t bsim_mlp_num_params(const bsim_mlp *m) { return m ? m->nparam : -1; }

int bsim_mlp_layout(const bsim_mlp *m, int32_t *dims, int32_t cap)
{
    if (!m || !dims) return mfail(BSIM_E_ARG, "null argument");
    if (cap < m->nlin + 1) return mfail(BSIM_E_ARG, "dims capacity too small");
    for (int l = 0; l <= m->nlin; l++) dims[l] = m->kdim[l];
    return m->nlin;
}

int bsim_mlp_get_params(bsim_mlp *m, float *out)
{
    if (!m || !out) return mfail(BSIM_E_ARG, "null argument");
    if (int rc = hip_check(hipStreamSynchronize(m->stream), "sync")) return rc;
    return hip_check(hipMemcpy(out, m->p, m->nparam * sizeof(float), hipMemcpyDeviceToHost),
                     "hipMemcpy");
}

int bsim_mlp_set_params(bsim_mlp *m, const float *in)
{
    if (!m || !in) return mfail(BSIM_E_ARG, "null argument");
    if (int rc = hip_check(hipStreamSynchronize(m->stream), "sync")) return rc;
    std::vector<float> p(in, in + m->nparam);
    /* padded positions stay zero */
    for (int l = 0; l < m->nlin; l++) {
        const int32_t fan_in = l == 0 ? m->cfg.in_dim : m->cfg.hidden;
        const int32_t nrow = l + 1 == m->nlin ? m->cfg.out_dim : m->cfg.hidden;
        for (int32_t r = 0; r < m->kdim[l + 1]; r++) {
            for (int32_t k = 0; k < m->kdim[l]; k++)
                if (r >= nrow || k >= fan_in) p[m->woff[l] + (int64_t)r * m->kdim[l] + k] = 0.f;
            if (r >= nrow) p[m->boff[l] + r] = 0.f;
        }
    }
    std::vector<uint16_t> pb(m->nparam);
    for (int64_t i = 0; i < m->nparam; i++) pb[i] = f2bf(p[i]);
    int rc = hip_check(hipMemcpy(m->p, p.data(), m->nparam * 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = hip_check(hipMemcpy(m->pbf, pb.data(), m->nparam * 2, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = transpose_weights(m);
    if (!rc) rc = hip_check(hipStreamSynchronize(m->stream), "sync");
    if (!rc) rc = hip_check(hipMemset(m->m, 0, m->nparam * 4), "hipMemset");
    if (!rc) rc = hip_check(hipMemset(m->v, 0, m->nparam * 4), "hipMemset");
    m->step = 0;
    return rc;
}

int bsim_mlp_set_data(bsim_mlp *m, const float *x, const float *y, int64_t n)
{
    if (!m || !x || !y || n <= 0) return mfail(BSIM_E_ARG, "null argument or empty data set");
    if (int rc = hip_check(hipStreamSynchronize(m->stream), "sync")) return rc;
    if (m->xall) {
        for (auto it = m->allocs.begin(); it != m->allocs.end();)
            if (*it == m->xall || *it == m->yall) {
                hipFree(*it);
                it = m->allocs.erase(it);
            } else {
                ++it;
            }
        m->xall = nullptr;
        m->yall = nullptr;
        m->ndata = 0;
    }
    const int32_t kin = m->kdim[0], din = m->cfg.in_dim, dout = m->cfg.out_dim;
    if (int rc = dalloc(m, &m->xall, n * kin)) return rc;
    if (int rc = dalloc(m, &m->yall, n * dout)) return rc;
    std::vector<uint16_t> xb((size_t)(n * kin), 0);
    for (int64_t r = 0; r < n; r++)
        for (int32_t c = 0; c < din; c++) xb[r * kin + c] = f2bf(x[r * din + c]);
    int rc = hip_check(hipMemcpy(m->xall, xb.data(), xb.size() * 2, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) rc = hip_check(hipMemcpy(m->yall, y, (size_t)(n * dout) * 4, hipMemcpyHostToDevice), "hipMemcpy");
    if (!rc) m->ndata = n;
    return rc;
}

}  // extern "C"

namespace {

int check_train(bsim_mlp *m, const int32_t *rows, int32_t steps, int32_t batch)
{
    if (!m || !rows || steps < 0) return mfail(BSIM_E_ARG, "null argument");
    if (int rc = check_batch(m, batch)) return rc;
    if (m->ndata <= 0) return mfail(BSIM_E_ARG, "no data set (bsim_mlp_set_data)");
    for (int64_t i = 0; i < (int64_t)steps * batch; i++)
        if (rows[i] < 0 || rows[i] >= m->ndata) return mfail(BSIM_E_ARG, "row index out of range");
    return BSIM_OK;
}

/* The training steps of bsim_mlp_train and bsim_mlp_train_noalloc: the
 * minibatch row indices are in rdev (device), step s's loss goes to ldev[s]
 * (device); synchronises the stream. */
int train_steps(bsim_mlp *m, const int32_t *rdev, float *ldev, int32_t steps, int32_t batch, float lr)
{
    int rc = BSIM_OK;
    for (int32_t s = 0; s < steps && !rc; s++) {
        /* the minibatch is gathered by the first layer's GEMM and the loss */
        rc = forward(m, batch, true, rdev + (int64_t)s * batch);
        if (!rc) rc = backward(m, batch);
        if (!rc) rc = adam(m, batch, true, lr, nullptr, ldev + s);   /* writes W^T too */
    }
    const int src = hip_check(hipStreamSynchronize(m->stream), "sync");
    return rc ? rc : src;
}

/* Scratch of bsim_mlp_train_noalloc: the per-step losses (from word 0), then
 * the minibatch row indices (from word round_up(steps, 64)).  A file-scope
 * device array, one copy per GPU, held until the process ends; its address
 * on each device is taken once, with that device current, by
 * hipGetSymbolAddress (the array's name is only a host-side stand-in). */
static __device__ __align__(256) uint32_t scratch[BSIM_MLP_NOALLOC_WORDS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[kMaxDevices] = {};

int scratch_on(int device, uint32_t **out)
{
    if (device < 0 || device >= kMaxDevices) return mfail(BSIM_E_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = static_cast<uint32_t *>(ptr);
    }
    *out = g_scratch_ptr[device];
    return BSIM_OK;
}

}  // namespace

extern "C" {

int bsim_mlp_train(bsim_mlp *m, const int32_t *rows, int32_t steps, int32_t batch, float lr,
                   float *loss)
{
    if (int rc = check_train(m, rows, steps, batch)) return rc;
    if (steps == 0) return BSIM_OK;
    float *ldev = nullptr;
    int32_t *rdev = nullptr;
    int rc = dalloc(m, &ldev, steps);
    if (!rc) rc = dalloc(m, &rdev, (int64_t)steps * batch);
    if (!rc) rc = hip_check(hipMemcpy(rdev, rows, (size_t)steps * batch * 4, hipMemcpyHostToDevice),
                            "hipMemcpy");
    if (!rc) rc = train_steps(m, rdev, ldev, steps, batch, lr);
    if (!rc && loss)
        rc = hip_check(hipMemcpy(loss, ldev, steps * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    for (void *q : {(void *)ldev, (void *)rdev}) {
        if (!q) continue;
        m->allocs.erase(std::find(m->allocs.begin(), m->allocs.end(), q));
        hipFree(q);
    }
    return rc;
}

int bsim_mlp_train_noalloc(bsim_mlp *m, const int32_t *rows, int32_t steps, int32_t batch,
                           float lr, float *loss)
{
    if (int rc = check_train(m, rows, steps, batch)) return rc;
    const int64_t roff = round_up(steps, 64);
    if (roff + (int64_t)steps * batch > (int64_t)BSIM_MLP_NOALLOC_WORDS)
        return mfail(BSIM_E_ARG, "bsim_mlp_train_noalloc: round_up(steps, 64) + steps * batch exceeds "
                                 "the scratch of " + std::to_string(BSIM_MLP_NOALLOC_WORDS) + " words");
    if (steps == 0) return BSIM_OK;
    int rc = hip_check(hipSetDevice(m->cfg.device), "hipSetDevice");
    uint32_t *base = nullptr;
    if (!rc) rc = scratch_on(m->cfg.device, &base);
    if (rc) return rc;
    float *ldev = reinterpret_cast<float *>(base);
    int32_t *rdev = reinterpret_cast<int32_t *>(base + roff);
    /* every word the steps read is written here first; the losses are
     * stored, not accumulated: nothing needs clearing */
    rc = hip_check(hipMemcpyAsync(rdev, rows, (size_t)steps * batch * 4, hipMemcpyHostToDevice,
                                  m->stream), "hipMemcpyAsync");
    if (!rc) rc = train_steps(m, rdev, ldev, steps, batch, lr);
    if (!rc && loss) {
        rc = hip_check(hipMemcpyAsync(loss, ldev, (size_t)steps * 4, hipMemcpyDeviceToHost, m->stream),
                       "hipMemcpyAsync");
        if (!rc) rc = hip_check(hipStreamSynchronize(m->stream), "sync");
    }
    return rc;
}

int bsim_mlp_gradients(bsim_mlp *m, const float *x, const float *y, int32_t batch, float *grad,
                       float *loss)
{
    if (!m || !x || !y || !grad || !loss) return mfail(BSIM_E_ARG, "null argument");
    if (int rc = check_batch(m, batch)) return rc;
    const int32_t din = m->cfg.in_dim, kin = m->kdim[0];
    int rc = hip_check(hipMemcpyAsync(m->xstage, x, (size_t)batch * din * 4, hipMemcpyHostToDevice,
                                      m->stream), "hipMemcpyAsync");
    if (!rc) rc = hip_check(hipMemcpyAsync(m->ybat, y, (size_t)batch * m->cfg.out_dim * 4,
                                           hipMemcpyHostToDevice, m->stream), "hipMemcpyAsync");
    if (rc) return rc;
    const int64_t tot = (int64_t)batch * kin;
    (void)hipGetLastError();   /* the check after the launch reports it alone */
    hipLaunchKernelGGL(mlp_convert_x, dim3((uint32_t)((tot + 255) / 256)), dim3(256), 0, m->stream,
                       m->xstage, m->act[0], m->act_t[0], (int64_t)batch, din, kin);
    rc = hip_check(hipGetLastError(), "convert launch");
    if (!rc) rc = forward(m, batch, true);
    if (!rc) rc = backward(m, batch);
    if (!rc) rc = adam(m, batch, false, 0.f, m->grad, m->loss);
    if (!rc) rc = hip_check(hipStreamSynchronize(m->stream), "sync");
    if (!rc) rc = hip_check(hipMemcpy(grad, m->grad, m->nparam * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (!rc) rc = hip_check(hipMemcpy(loss, m->loss, 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return rc;
}

int bsim_mlp_predict(bsim_mlp *m, const float *x, int64_t n, float *y)
{
    if (!m || !x || !y || n < 0) return mfail(BSIM_E_ARG, "null argument");
    const int32_t din = m->cfg.in_dim, kin = m->kdim[0], dout = m->cfg.out_dim;
    const int32_t nout = m->kdim[m->nlin];
    const int64_t B = m->cfg.max_batch;
    std::vector<float> tmp;
    int rc = BSIM_OK;
    for (int64_t r0 = 0; r0 < n && !rc; r0 += B) {
        const int64_t rows = std::min<int64_t>(B, n - r0);
        const int32_t batch = (int32_t)round_up(rows, 64);
        rc = hip_check(hipMemsetAsync(m->xstage, 0, (size_t)batch * din * 4, m->stream), "memset");
        if (!rc) rc = hip_check(hipMemcpyAsync(m->xstage, x + r0 * din, (size_t)rows * din * 4,
                                               hipMemcpyHostToDevice, m->stream), "hipMemcpyAsync");
        if (rc) break;
        const int64_t tot = (int64_t)batch * kin;
        (void)hipGetLastError();   /* the check after the launch reports it alone */
        hipLaunchKernelGGL(mlp_convert_x, dim3((uint32_t)((tot + 255) / 256)), dim3(256), 0,
                           m->stream, m->xstage, m->act[0], (uint16_t *)nullptr, (int64_t)batch,
                           din, kin);
        rc = hip_check(hipGetLastError(), "convert launch");
        if (!rc) rc = forward(m, batch, false);
        if (rc) break;
        tmp.resize((size_t)rows * nout);
        rc = hip_check(hipMemcpyAsync(tmp.data(), m->out, (size_t)rows * nout * 4,
                                      hipMemcpyDeviceToHost, m->stream), "hipMemcpyAsync");
        if (!rc) rc = hip_check(hipStreamSynchronize(m->stream), "sync");
        for (int64_t r = 0; r < rows && !rc; r++)
            memcpy(y + (r0 + r) * dout, tmp.data() + r * nout, dout * 4);
    }
    return rc;
}

}  // extern "C"
