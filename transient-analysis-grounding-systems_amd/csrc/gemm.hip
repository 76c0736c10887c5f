// Complex GEMM on the MFMA cores of gfx950: C = A B or C -= A B, row-major
// operands read and written through optional row maps (the LU's logical-to-
// physical row permutation). This is the trailing update of the blocked LU
// (A22 -= L21 U12) and the TRSM-as-GEMM of the panel rows.
//
// Kernel shape: 256 threads = 4 waves in a 2 x 2 arrangement; work-group tile
// BM x BN; K staged through double-buffered LDS in chunks of BK, stored
// k-major (one row of the tile per k) with a 128-byte row skew so that the
// four 16-lane groups of an MFMA operand read hit disjoint banks; the next
// chunk is fetched from HBM into registers while the MFMAs of the current one
// run. Complex products use four real MFMAs (re.re, im.im, re.im, im.re) or,
// with M3, three (Gauss: re.re, im.im, (re+im).(re+im)).
#include <algorithm>
#include <atomic>

#include "lu.h"

namespace tags {

template <typename T>
struct Mfma;
// 16x16x4 blocks; lane l supplies A[l%16][l/16] and B[l/16][l%16]. Result
// element v of lane l is D[row(l, v)][l%16] with row = l/16 + 4v for the f64
// instruction and 4(l/16) + v for the f32 one (measured on gfx950).
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 run(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 run(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Largest K of a product with a row map on B, which is staged in LDS (the LU's
// products have K <= 256)
constexpr int GEMM_BROW_LDS = 512;

template <typename T, int BM, int BN, int BK, bool M3, int WGM, int WGN, int AL>
__device__ __forceinline__ void cgemm_block(const GemmArgs<T>& g)
{
    using V4 = typename Mfma<T>::v4;
    constexpr int NT = 64 * WGM * WGN;
    constexpr int WM = BM / WGM, WN = BN / WGN, TM = WM / 16, TN = WN / 16;
    constexpr int SKEW = 128 / (int)sizeof(cx<T>);
    constexpr int SA = BM + SKEW, SB = BN + SKEW;
    // AL = 1: the A tile row-major (k contiguous, row pitch BK + 1), so that
    // the 16 rows of an MFMA operand read and the 8-lane row segments of the
    // staging stores each fall on distinct banks
    constexpr int SAK = BK + 1;
    constexpr int LA = BM * BK / NT, LB = BK * BN / NT;
    static_assert(LA * NT == BM * BK && LB * NT == BK * BN, "tile/thread mismatch");
    static_assert(TM * 16 == WM && TN * 16 == WN, "wave tile");
    __shared__ cx<T> As[2][AL ? BM * SAK : BK * SA];
    __shared__ cx<T> Bs[2][BK * SB];
    __shared__ int s_brow[GEMM_BROW_LDS];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave / WGN, wn = wave % WGN;
    const int tiles_n = (g.N + BN - 1) / BN;
    const int tiles_m = (g.M + BM - 1) / BM;
    const int ntiles = tiles_n * tiles_m;
    // XCD-aware order: dispatch id b runs on XCD b % 8; give each XCD a
    // contiguous range of tiles, walked in windows of 8 tile rows (down the
    // window's rows, then across its columns), so that the work-groups an XCD
    // holds at once share a few row panels of A and column panels of B in its
    // L2 instead of one row panel and many column panels.
    int bid;
    {
        const int per = (ntiles + 7) / 8;
        bid = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    }
    if (bid >= ntiles) return;
    constexpr int GM = 8;
    const int first_m = bid / (GM * tiles_n) * GM;
    const int gm = min(tiles_m - first_m, GM);
    const int rem = bid - first_m * tiles_n;
    const int tm = first_m + rem % gm, tn = rem / gm;
    const int m0 = tm * BM, n0 = tn * BN;

    // global-load slots of this thread (fixed rows of A, fixed columns of B)
    long aoff[LA];
    int acol[LA], arw[LA];
#pragma unroll
    for (int q = 0; q < LA; q++) {
        const int e = tid + NT * q;
        const int r = e / BK, c = e % BK, gr = m0 + r;
        arw[q] = r;
        acol[q] = c;
        aoff[q] = gr < g.M ? (long)(g.arow ? g.arow[gr] : gr) * g.lda : -1;
    }
    // B slots: row r of the chunk, column c (fixed for the whole K loop)
    int bcol[LB], brw[LB];
#pragma unroll
    for (int q = 0; q < LB; q++) {
        const int e = tid + NT * q;
        brw[q] = e / BN;
        bcol[q] = (n0 + e % BN) < g.N ? e % BN : -1;
    }
    // registers of the next chunk, fetched from HBM while the MFMAs of the
    // current one run
    cx<T> ra[LA], rb[LB];
#define TAGS_GLOAD(k0_, RA, RB)                                                           \
    {                                                                                     \
        _Pragma("unroll") for (int q = 0; q < LA; q++)                                    \
        {                                                                                 \
            const int gc = (k0_) + acol[q];                                               \
            const bool ok = aoff[q] >= 0 && gc < g.K;                                     \
            const cx<T> v = g.A[ok ? aoff[q] + gc : 0];                                   \
            RA[q] = ok ? v : cx<T>{0, 0};                                                 \
        }                                                                                 \
        _Pragma("unroll") for (int q = 0; q < LB; q++)                                    \
        {                                                                                 \
            const int gk = (k0_) + brw[q];                                                \
            const bool ok = gk < g.K && bcol[q] >= 0;                                     \
            const int br = g.brow ? s_brow[ok ? gk : 0] : gk;                              \
            const long pr = ok ? (long)br * g.ldb + n0 + bcol[q] : 0;                      \
            const cx<T> v = g.B[pr];                                                      \
            RB[q] = ok ? v : cx<T>{0, 0};                                                 \
        }                                                                                 \
    }
#define TAGS_SSTORE(buf_, RA, RB)                                                         \
    {                                                                                     \
        _Pragma("unroll") for (int q = 0; q < LA; q++)                                    \
            As[buf_][AL ? arw[q] * SAK + acol[q] : acol[q] * SA + arw[q]] = RA[q];       \
        _Pragma("unroll") for (int q = 0; q < LB; q++)                                    \
        {                                                                                 \
            const int e = tid + NT * q;                                                   \
            Bs[buf_][brw[q] * SB + (e % BN)] = RB[q];                                     \
        }                                                                                 \
    }

    V4 acc0[TM][TN], acc1[TM][TN], acc2[M3 ? TM : 1][M3 ? TN : 1];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
            acc0[i][j] = V4{0, 0, 0, 0};
            acc1[i][j] = V4{0, 0, 0, 0};
            if constexpr (M3) acc2[i][j] = V4{0, 0, 0, 0};
        }

    // the MFMAs of the chunk in LDS buffer buf_
    auto mma = [&](int buf_) {
        const cx<T>* as = As[buf_];
        const cx<T>* bs = Bs[buf_];
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            cx<T> a[TM], b[TN];
            const int kr = kk + (lane >> 4);
#pragma unroll
            for (int i = 0; i < TM; i++)
                a[i] = AL ? as[(wm * WM + i * 16 + (lane & 15)) * SAK + kr]
                          : as[kr * SA + wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < TN; j++) b[j] = bs[kr * SB + wn * WN + j * 16 + (lane & 15)];
            if constexpr (M3) {
                T asum[TM], bsum[TN];
#pragma unroll
                for (int i = 0; i < TM; i++) asum[i] = a[i].x + a[i].y;
#pragma unroll
                for (int j = 0; j < TN; j++) bsum[j] = b[j].x + b[j].y;
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++) {
                        acc0[i][j] = Mfma<T>::run(a[i].x, b[j].x, acc0[i][j]);
                        acc1[i][j] = Mfma<T>::run(a[i].y, b[j].y, acc1[i][j]);
                        acc2[i][j] = Mfma<T>::run(asum[i], bsum[j], acc2[i][j]);
                    }
            } else {
#pragma unroll
                for (int i = 0; i < TM; i++)
#pragma unroll
                    for (int j = 0; j < TN; j++) {
                        acc0[i][j] = Mfma<T>::run(a[i].x, b[j].x, acc0[i][j]);
                        acc0[i][j] = Mfma<T>::run(-a[i].y, b[j].y, acc0[i][j]);
                        acc1[i][j] = Mfma<T>::run(a[i].x, b[j].y, acc1[i][j]);
                        acc1[i][j] = Mfma<T>::run(a[i].y, b[j].x, acc1[i][j]);
                    }
            }
        }
    };

    // B's row map for the whole K range (K <= GEMM_BROW_LDS, checked at launch)
    // in LDS: read from HBM inside the K loop, each index would make the next
    // B address wait for every load in flight (loads complete in issue order),
    // the next chunk's A prefetch too
    if (g.brow) {
        for (int k = tid; k < g.K; k += NT) s_brow[k] = g.brow[k];
        __syncthreads();
    }
    TAGS_GLOAD(0, ra, rb);
    TAGS_SSTORE(0, ra, rb);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        const bool more = k0 + BK < g.K;
        if (more) TAGS_GLOAD(k0 + BK, ra, rb);
        mma(buf);
        if (more) TAGS_SSTORE(buf ^ 1, ra, rb);
        __syncthreads();
        buf ^= 1;
    }

#undef TAGS_GLOAD
#undef TAGS_SSTORE
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int row = m0 + wm * WM + i * 16 + Mfma<T>::row(lane, v);
            if (row >= g.M) continue;
            cx<T>* crow = g.C + (long)(g.crow ? g.crow[row] : row) * g.ldc;
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const int col = n0 + wn * WN + j * 16 + (lane & 15);
                if (col >= g.N) continue;
                T re, im;
                if constexpr (M3) {
                    re = acc0[i][j][v] - acc1[i][j][v];
                    im = acc2[i][j][v] - acc0[i][j][v] - acc1[i][j][v];
                } else {
                    re = acc0[i][j][v];
                    im = acc1[i][j][v];
                }
                if constexpr ((AL & 2) != 0) {
                    // non-temporal C: streamed, so that it does not evict the
                    // operand tiles other work-groups still read from L2
                    T* f = (T*)(crow + col);
                    if (g.mode == 0) {
                        __builtin_nontemporal_store(re, f);
                        __builtin_nontemporal_store(im, f + 1);
                    } else {
                        const T cr = __builtin_nontemporal_load(f);
                        const T ci = __builtin_nontemporal_load(f + 1);
                        __builtin_nontemporal_store(cr - re, f);
                        __builtin_nontemporal_store(ci - im, f + 1);
                    }
                } else if (g.mode == 0) {
                    crow[col] = cx<T>{re, im};
                } else {
                    const cx<T> c = crow[col];
                    crow[col] = cx<T>{c.x - re, c.y - im};
                }
            }
        }
}

// The operands of item blockIdx.z of a batched launch. The grid's x extent is
// a multiple of 8, so a work-group's XCD (its linear dispatch id mod 8) is
// still blockIdx.x mod 8 for every item.
template <typename T>
__device__ __forceinline__ GemmArgs<T> gemm_item(const GemmArgs<T>& g0)
{
    GemmArgs<T> g = g0;
    const long z = blockIdx.z;
    if (z) {
        g.A += z * g.sa;
        g.B += z * g.sb;
        g.C += z * g.sc;
        if (g.arow) g.arow += z * g.sar;
        if (g.brow) g.brow += z * g.sbr;
        if (g.crow) g.crow += z * g.scr;
    }
    return g;
}

template <typename T, int BM, int BN, int BK, bool M3, int WGM, int WGN, int AL>
__global__ __launch_bounds__(64 * WGM * WGN) void k_cgemm(GemmArgs<T> g)
{
    cgemm_block<T, BM, BN, BK, M3, WGM, WGN, AL>(gemm_item(g));
}

// The same product under its own name for the LU's trailing update
// A22 -= L21 U12, so that kernel traces and counters attribute it apart from
// the other products (panel rows, triangular solves).
template <typename T, int BM, int BN, int BK, bool M3, int WGM, int WGN, int AL>
__global__ __launch_bounds__(64 * WGM * WGN) void k_trailing_update(GemmArgs<T> g)
{
    cgemm_block<T, BM, BN, BK, M3, WGM, WGN, AL>(gemm_item(g));
}

// ---------------------------------------------------------------------------
// complex64 with 32 x 32 x 2 MFMA blocks: 128 x 128 tiles of 4 waves (2 x 2),
// each wave 64 x 64 (2 x 2 blocks of v_mfma_f32_32x32x2f32), Gauss
// three-multiply product, K staged through double-buffered LDS in chunks of
// 16. Against the 16 x 16 x 4 tiles above, a wave reads half the LDS bytes per
// MFMA cycle and the work-group meets half the barriers per K. Lane l supplies
// A[l % 32][l / 32] and B[l / 32][l % 32]; result value v of lane l is
// D[(v % 4) + 8 (v / 4) + 4 (l / 32)][l % 32] (measured on gfx950).
// At most 256 registers per lane, so two work-groups share a CU.
// ---------------------------------------------------------------------------
constexpr int G32_BM = 128, G32_BN = 128, G32_BK = 16, G32_NT = 256;
constexpr int G32_SA = G32_BM + 4, G32_SB = G32_BN + 4;  // row skew of 4 complex

__device__ __forceinline__ void cgemm32_block(const GemmArgs<float>& g)
{
    typedef float v16 __attribute__((ext_vector_type(16)));
    __shared__ cx<float> As[2][G32_BK * G32_SA];
    __shared__ cx<float> Bs[2][G32_BK * G32_SB];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int tiles_n = (g.N + G32_BN - 1) / G32_BN;
    const int tiles_m = (g.M + G32_BM - 1) / G32_BM;
    const int ntiles = tiles_n * tiles_m;
    int bid;
    {
        // XCD-aware order: dispatch id b runs on XCD b % 8; each XCD takes a
        // contiguous range of tiles so that neighbouring tiles share its L2
        const int per = (ntiles + 7) / 8;
        bid = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    }
    if (bid >= ntiles) return;
    const int m0 = (bid / tiles_n) * G32_BM, n0 = (bid % tiles_n) * G32_BN;

    // global-load slots: A row m0 + (tid >> 1), 8 consecutive k from
    // 8 (tid & 1); B chunk row tid >> 4, 8 consecutive columns from 8 (tid & 15)
    const int ar = tid >> 1, ak = (tid & 1) * 8;
    const long aoff = m0 + ar < g.M ? (long)(g.arow ? g.arow[m0 + ar] : m0 + ar) * g.lda : -1;
    const int bk = tid >> 4, bc = (tid & 15) * 8;
    cx<float> ra[8], rb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int gk = k0 + ak + q;
            const bool ok = aoff >= 0 && gk < g.K;
            const cx<float> v = g.A[ok ? aoff + gk : 0];
            ra[q] = ok ? v : cx<float>{0, 0};
        }
        const int gk = k0 + bk;
        const long brow = gk < g.K ? (long)(g.brow ? g.brow[gk] : gk) * g.ldb : 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int gc = n0 + bc + q;
            const bool ok = gk < g.K && gc < g.N;
            const cx<float> v = g.B[ok ? brow + gc : 0];
            rb[q] = ok ? v : cx<float>{0, 0};
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 8; q++) As[buf][(ak + q) * G32_SA + ar] = ra[q];
#pragma unroll
        for (int q = 0; q < 8; q++) Bs[buf][bk * G32_SB + bc + q] = rb[q];
    };

    v16 acc0[2][2], acc1[2][2], acc2[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                acc0[i][j][v] = 0.f;
                acc1[i][j][v] = 0.f;
                acc2[i][j][v] = 0.f;
            }

    gload(0);
    sstore(0);
    __syncthreads();
    int buf = 0;
    const int kl = lane >> 5, ml = lane & 31;
    for (int k0 = 0; k0 < g.K; k0 += G32_BK) {
        const bool more = k0 + G32_BK < g.K;
        if (more) gload(k0 + G32_BK);
        const cx<float>* as = As[buf] + wm * 64 + ml;
        const cx<float>* bs = Bs[buf] + wn * 64 + ml;
#pragma unroll
        for (int kk = 0; kk < G32_BK; kk += 2) {
            cx<float> a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; i++) a[i] = as[(kk + kl) * G32_SA + i * 32];
#pragma unroll
            for (int j = 0; j < 2; j++) b[j] = bs[(kk + kl) * G32_SB + j * 32];
            float asum[2], bsum[2];
#pragma unroll
            for (int i = 0; i < 2; i++) asum[i] = a[i].x + a[i].y;
#pragma unroll
            for (int j = 0; j < 2; j++) bsum[j] = b[j].x + b[j].y;
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    acc0[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].x, b[j].x, acc0[i][j],
                                                                      0, 0, 0);
                    acc1[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i].y, b[j].y, acc1[i][j],
                                                                      0, 0, 0);
                    acc2[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(asum[i], bsum[j],
                                                                      acc2[i][j], 0, 0, 0);
                }
        }
        if (more) sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }

#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int row = m0 + wm * 64 + i * 32 + (v & 3) + 8 * (v >> 2) + 4 * kl;
            if (row >= g.M) continue;
            cx<float>* crow = g.C + (long)(g.crow ? g.crow[row] : row) * g.ldc;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int col = n0 + wn * 64 + j * 32 + ml;
                if (col >= g.N) continue;
                const float re = acc0[i][j][v] - acc1[i][j][v];
                const float im = acc2[i][j][v] - acc0[i][j][v] - acc1[i][j][v];
                if (g.mode == 0) {
                    crow[col] = cx<float>{re, im};
                } else {
                    const cx<float> c = crow[col];
                    crow[col] = cx<float>{c.x - re, c.y - im};
                }
            }
        }
}

__global__ __launch_bounds__(G32_NT, 2) void k_cgemm32(GemmArgs<float> g)
{
    cgemm32_block(gemm_item(g));
}
// the same product under the trailing update's name (see k_trailing_update)
__global__ __launch_bounds__(G32_NT, 2) void k_trailing_update32(GemmArgs<float> g)
{
    cgemm32_block(gemm_item(g));
}

static int launch_v32(const GemmArgs<float>& g, hipStream_t st, bool trailing)
{
    const long tiles = (long)((g.M + G32_BM - 1) / G32_BM) * ((g.N + G32_BN - 1) / G32_BN);
    const long grid = (tiles + 7) / 8 * 8;
    const dim3 gd((unsigned)grid, 1, (unsigned)g.nb);
    if (trailing)
        hipLaunchKernelGGL(k_trailing_update32, gd, dim3(G32_NT), 0, st, g);
    else
        hipLaunchKernelGGL(k_cgemm32, gd, dim3(G32_NT), 0, st, g);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// complex64 on the bf16 MFMA cores (variants 22..32). Every fp32 operand value
// x is split into three bf16 parts x = x1 + x2 + x3, each the round-to-nearest
// bf16 of what the earlier parts leave: x1 holds the leading 8 significant
// bits, x2 the next 8, x3 the last 8, so that the parts sum to x exactly (for
// |x| well inside the normal range; the differences are exact in fp32). A real
// product x y is formed from the six part products x1y1, x1y2, x2y1, x1y3,
// x2y2, x3y1, accumulated in fp32; the three left out (x2y3, x3y2, x3y3) add up
// to less than 2^-24 |x||y|, the size of one fp32 rounding of the term.
// Complex products take four real ones (re re - im im, re im + im re): 24 bf16
// MFMAs where the three-multiply fp32 form issues 3 fp32 ones, and bf16 runs at
// 16x the fp32 MFMA rate on gfx950, so a complex product takes 24 / 16 = 1.5
// fp32-MFMA issue slots instead of 3.
//
// The operands are first packed (k_pack_x6) into MFMA fragment order: for a
// block of 32 rows of A (32 columns of B) and a slab of 16 k, part p (the
// three parts of the real values, then of the imaginary ones) is 64 lanes x 8
// bf16 = 1 KiB, lane l holding row (column) l % 32 at k = 8 (l / 32) .. + 7,
// zero past the matrix. A wave loads each fragment with one 16-byte load per
// lane straight into MFMA operand registers, the next slab's while the
// current one's MFMAs run; blocks past the product's edge repeat the last
// block and their results are dropped at the store.
// v_mfma_f32_32x32x16_bf16: lane l supplies A[l % 32][8 (l / 32) + e] and
// B[8 (l / 32) + e][l % 32], e = 0..7 (A and B packed alike, so the k order
// inside a slab is the same on both sides, which is all the sum needs); result
// value v of lane l is D[(v % 4) + 8 (v / 4) + 4 (l / 32)][l % 32], as for the
// 32 x 32 f32 blocks above.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int X6_KMAX = 256;  // largest K (the LU's widest step)
constexpr int X6_FRAG = 64;   // 16-byte words per fragment
constexpr int X6_PARTS = 6;

// 16-byte words of one item's pack of R rows (columns) and K
static inline long x6_pack_words(int R, int K)
{
    return (long)((R + 31) / 32) * ((K + 15) / 16) * X6_PARTS * X6_FRAG;
}

long gemm_x6_ws_bytes(int M, int N, int K)
{
    if (M <= 0 || N <= 0 || K <= 0 || K > X6_KMAX) return 0;
    return 16 * (x6_pack_words(M, K) + x6_pack_words(N, K));
}

// round to nearest even bf16 (finite x), as the upper 16 bits of an fp32
__device__ inline unsigned bf16_rn(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ inline void x6_split(float x, unsigned& h1, unsigned& h2, unsigned& h3)
{
    h1 = bf16_rn(x);
    const float r1 = x - __uint_as_float(h1 << 16);
    h2 = bf16_rn(r1);
    const float r2 = r1 - __uint_as_float(h2 << 16);
    h3 = bf16_rn(r2);
}

// SIDE 0, rows of A: X(i, k) = X[(map ? map[i] : i) * ld + k], i < R;
// SIDE 1, columns of B: X(k, i) = X[(map ? map[k] : k) * ld + i], i < R.
// One thread per (block, slab, lane); item blockIdx.y. TR: the trailing
// update's packs (their own name in kernel traces).
template <int SIDE, bool TR>
__global__ __launch_bounds__(256) void k_pack_x6(const c32* X, long ld, const int* map, int R,
                                                  int K, u32x4* out, long sx, long smap,
                                                  long sout)
{
    const int KS = (K + 15) / 16;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)((R + 31) / 32) * KS * 64) return;
    const long z = blockIdx.y;
    X += z * sx;
    if (map) map += z * smap;
    out += z * sout;
    const int lane = (int)(t & 63);
    const long f = t >> 6;  // block * KS + slab
    const int ks = (int)(f % KS), blk = (int)(f / KS);
    const int i = blk * 32 + (lane & 31);
    const int k0 = ks * 16 + 8 * (lane >> 5);
    unsigned w[X6_PARTS][4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        float v[2][2];  // [element][re, im]
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int k = k0 + e + q;
            c32 x{0, 0};
            if (i < R && k < K) {
                const long off = SIDE == 0 ? (long)(map ? map[i] : i) * ld + k
                                           : (long)(map ? map[k] : k) * ld + i;
                x = X[off];
            }
            v[q][0] = x.x;
            v[q][1] = x.y;
        }
#pragma unroll
        for (int c = 0; c < 2; c++) {
            unsigned a1, a2, a3, b1, b2, b3;
            x6_split(v[0][c], a1, a2, a3);
            x6_split(v[1][c], b1, b2, b3);
            w[3 * c + 0][e / 2] = a1 | (b1 << 16);
            w[3 * c + 1][e / 2] = a2 | (b2 << 16);
            w[3 * c + 2][e / 2] = a3 | (b3 << 16);
        }
    }
    u32x4* o = out + f * (X6_PARTS * X6_FRAG) + lane;
#pragma unroll
    for (int p = 0; p < X6_PARTS; p++) {
        u32x4 q;
        q.x = w[p][0];
        q.y = w[p][1];
        q.z = w[p][2];
        q.w = w[p][3];
        o[p * X6_FRAG] = q;
    }
}

struct X6Args {
    int M, N, K;
    const u32x4* pa;  // packed A rows
    const u32x4* pb;  // packed B columns
    long spk;         // item stride of the packs (16-byte words)
    c32* C;
    long ldc;
    const int* crow;
    long sc, scr;
    int mode;  // 0: C = A B; 1: C -= A B
};

__device__ __forceinline__ bf16x8 as_bf16(u32x4 u) { return __builtin_bit_cast(bf16x8, u); }
__device__ __forceinline__ f32x16 mfma_x6(u32x4 a, u32x4 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16(a), as_bf16(b), c, 0, 0, 0);
}

// C = v (mode 0) or C -= v (mode 1) for one complex64 entry
template <bool NTC>
__device__ __forceinline__ void x6_store(c32* p, float re, float im, int mode)
{
    float* f = (float*)p;
    if constexpr (NTC) {
        if (mode == 0) {
            __builtin_nontemporal_store(re, f);
            __builtin_nontemporal_store(im, f + 1);
        } else {
            const float cr = __builtin_nontemporal_load(f), ci = __builtin_nontemporal_load(f + 1);
            __builtin_nontemporal_store(cr - re, f);
            __builtin_nontemporal_store(ci - im, f + 1);
        }
    } else {
        if (mode == 0) {
            *p = c32{re, im};
        } else {
            const c32 c = *p;
            *p = c32{c.x - re, c.y - im};
        }
    }
}

// Work-group of WGM x WGN waves, each TM x TN blocks of 32 x 32; tiles in
// windows of GMW tile rows; NTC: C read and written with non-temporal hints
// (streamed through the caches, so that it does not evict the fragments).
template <int WGM, int WGN, int TM, int TN, int GMW, bool NTC>
__device__ __forceinline__ void x6_block(const X6Args& g)
{
    constexpr int BM = 32 * TM * WGM, BN = 32 * TN * WGN;
    const long z = blockIdx.z;
    const u32x4* pa = g.pa + z * g.spk;
    const u32x4* pb = g.pb + z * g.spk;
    c32* C = g.C + z * g.sc;
    const int* crow = g.crow ? g.crow + z * g.scr : nullptr;
    const int M = g.M, N = g.N, KS = (g.K + 15) / 16;
    // the wave index is uniform: kept in scalar registers, so that the
    // fragment addresses are a scalar base plus the lane's offset
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / WGN, wn = wave % WGN;
    const int tiles_n = (N + BN - 1) / BN, tiles_m = (M + BM - 1) / BM;
    const int ntiles = tiles_n * tiles_m;
    // XCD-aware order, as for the f32 tiles: a contiguous range of tiles per
    // XCD, walked in windows of 8 tile rows
    int bid;
    {
        const int per = (ntiles + 7) / 8;
        bid = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    }
    if (bid >= ntiles) return;
    constexpr int GM = GMW;
    const int first_m = bid / (GM * tiles_n) * GM;
    const int gm = min(tiles_m - first_m, GM);
    const int rem = bid - first_m * tiles_n;
    const int tm = first_m + rem % gm, tn = rem / gm;
    const int m0 = tm * BM + wm * TM * 32, n0 = tn * BN + wn * TN * 32;
    const int nbm = (M + 31) / 32, nbn = (N + 31) / 32;
    const u32x4* fa[TM];
    const u32x4* fb[TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
        fa[i] = pa + (long)min(m0 / 32 + i, nbm - 1) * KS * (X6_PARTS * X6_FRAG) + lane;
#pragma unroll
    for (int j = 0; j < TN; j++)
        fb[j] = pb + (long)min(n0 / 32 + j, nbn - 1) * KS * (X6_PARTS * X6_FRAG) + lane;

    f32x16 re[TM][TN], im[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
#pragma unroll
            for (int v = 0; v < 16; v++) {
                re[i][j][v] = 0.f;
                im[i][j][v] = 0.f;
            }
        }
    u32x4 a0[TM][X6_PARTS], b0[TN][X6_PARTS], a1[TM][X6_PARTS], b1[TN][X6_PARTS];
    auto load = [&](u32x4(&a)[TM][X6_PARTS], u32x4(&b)[TN][X6_PARTS], int ks) {
        const long o = (long)ks * (X6_PARTS * X6_FRAG);
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int p = 0; p < X6_PARTS; p++) a[i][p] = fa[i][o + p * X6_FRAG];
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int p = 0; p < X6_PARTS; p++) b[j][p] = fb[j][o + p * X6_FRAG];
    };
    auto mma = [&](const u32x4(&a)[TM][X6_PARTS], const u32x4(&b)[TN][X6_PARTS]) {
        // -im(B), for the re re - im im half
        u32x4 nbi[TN][3];
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int t = 0; t < 3; t++) nbi[j][t] = b[j][3 + t] ^ 0x80008000u;
        // part pairs (s, t) of the six kept products
        constexpr int PS[6] = {2, 1, 0, 1, 0, 0}, PT[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int s = PS[q], t = PT[q];
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    re[i][j] = mfma_x6(a[i][s], b[j][t], re[i][j]);
                    im[i][j] = mfma_x6(a[i][s], b[j][3 + t], im[i][j]);
                }
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    re[i][j] = mfma_x6(a[i][3 + s], nbi[j][t], re[i][j]);
                    im[i][j] = mfma_x6(a[i][3 + s], b[j][t], im[i][j]);
                }
        }
    };
    load(a0, b0, 0);
    int ks = 0;
    for (; ks + 1 < KS; ks += 2) {
        load(a1, b1, ks + 1);
        mma(a0, b0);
        if (ks + 2 < KS) load(a0, b0, ks + 2);
        mma(a1, b1);
    }
    if (ks < KS) mma(a0, b0);

    const int kl = lane >> 5, ml = lane & 31;
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int row = m0 + i * 32 + (v & 3) + 8 * (v >> 2) + 4 * kl;
            if (row >= M) continue;
            c32* cr = C + (long)(crow ? crow[row] : row) * g.ldc;
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const int col = n0 + j * 32 + ml;
                if (col >= N) continue;
                x6_store<NTC>(cr + col, re[i][j][v], im[i][j][v], g.mode);
            }
        }
}

// The same product with the work-group's fragments of each slab staged once
// in LDS (double-buffered): every thread copies its share of the slab's
// fragments (global -> registers -> LDS, 16 bytes at a time, a wave's words
// contiguous) for the next slab while the MFMAs of the current one run, and
// each wave reads its own fragments from LDS, one barrier per slab. The
// work-group fetches a fragment once instead of once per wave that uses it.
template <int WGM, int WGN, int TM, int TN, int GMW, bool NTC, int DEPTH>
__device__ __forceinline__ void x6l_tile(const X6Args& g, int bid)
{
    constexpr int NT = 64 * WGM * WGN;
    constexpr int BM = 32 * TM * WGM, BN = 32 * TN * WGN;
    constexpr int FA = BM / 32 * X6_PARTS, FB = BN / 32 * X6_PARTS;  // fragments per slab
    constexpr int NF = FA + FB;
    constexpr int NW = NF * X6_FRAG / NT;  // 16-byte words a thread stages
    static_assert(NW * NT == NF * X6_FRAG, "slab words per thread");
    static_assert(NT % 64 == 0 && (NT / 64) * NW == NF, "fragments per wave");
    __shared__ u32x4 lds[2][NF * X6_FRAG];

    const long z = blockIdx.z;
    const u32x4* pa = g.pa + z * g.spk;
    const u32x4* pb = g.pb + z * g.spk;
    c32* C = g.C + z * g.sc;
    const int* crow = g.crow ? g.crow + z * g.scr : nullptr;
    const int M = g.M, N = g.N, KS = (g.K + 15) / 16;
    // the wave index is uniform: kept in scalar registers, so that the
    // fragment addresses are a scalar base plus the lane's offset
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / WGN, wn = wave % WGN;
    const int tiles_n = (N + BN - 1) / BN, tiles_m = (M + BM - 1) / BM;
    constexpr int GM = GMW;
    const int first_m = bid / (GM * tiles_n) * GM;
    const int gm = min(tiles_m - first_m, GM);
    const int rem = bid - first_m * tiles_n;
    const int tm = first_m + rem % gm, tn = rem / gm;
    const int nbm = (M + 31) / 32, nbn = (N + 31) / 32;
    // staging: word q of this thread is word `lane` of fragment wave + (NT/64) q
    // of the slab (A fragments first: block, part; then B's)
    const u32x4* src[NW];
#pragma unroll
    for (int q = 0; q < NW; q++) {
        const int fi = wave + (NT / 64) * q;
        if (fi < FA) {
            const int b = min(tm * (BM / 32) + fi / X6_PARTS, nbm - 1);
            src[q] = pa + (long)b * KS * (X6_PARTS * X6_FRAG) + (fi % X6_PARTS) * X6_FRAG + lane;
        } else {
            const int f2 = fi - FA;
            const int b = min(tn * (BN / 32) + f2 / X6_PARTS, nbn - 1);
            src[q] = pb + (long)b * KS * (X6_PARTS * X6_FRAG) + (f2 % X6_PARTS) * X6_FRAG + lane;
        }
    }
    // two register sets: with DEPTH 2 a slab's words are fetched two slabs
    // before they are stored to LDS, so that a fetch from beyond L2 is covered
    u32x4 stage[2][NW];
    auto gload = [&](int set, int ks) {
        const long o = (long)ks * (X6_PARTS * X6_FRAG);
#pragma unroll
        for (int q = 0; q < NW; q++) stage[set][q] = src[q][o];
    };
    auto sstore = [&](int set, int buf) {
#pragma unroll
        for (int q = 0; q < NW; q++)
            lds[buf][(wave + (NT / 64) * q) * X6_FRAG + lane] = stage[set][q];
    };

    f32x16 re[TM][TN], im[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
#pragma unroll
            for (int v = 0; v < 16; v++) {
                re[i][j][v] = 0.f;
                im[i][j][v] = 0.f;
            }
        }
    auto mma = [&](int buf) {
        const u32x4* la = lds[buf] + (wm * TM * X6_PARTS) * X6_FRAG + lane;
        const u32x4* lb = lds[buf] + (FA + wn * TN * X6_PARTS) * X6_FRAG + lane;
        u32x4 a[TM][X6_PARTS], b[TN][X6_PARTS], nbi[TN][3];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int p = 0; p < X6_PARTS; p++) a[i][p] = la[(i * X6_PARTS + p) * X6_FRAG];
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int p = 0; p < X6_PARTS; p++) b[j][p] = lb[(j * X6_PARTS + p) * X6_FRAG];
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int t = 0; t < 3; t++) nbi[j][t] = b[j][3 + t] ^ 0x80008000u;
        constexpr int PS[6] = {2, 1, 0, 1, 0, 0}, PT[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int s = PS[q], t = PT[q];
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    re[i][j] = mfma_x6(a[i][s], b[j][t], re[i][j]);
                    im[i][j] = mfma_x6(a[i][s], b[j][3 + t], im[i][j]);
                }
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    re[i][j] = mfma_x6(a[i][3 + s], nbi[j][t], re[i][j]);
                    im[i][j] = mfma_x6(a[i][3 + s], b[j][t], im[i][j]);
                }
        }
    };
    // the same products one column block at a time, a scheduling barrier
    // between blocks, so that only one block's B fragments are live (fewer
    // registers, room for the second staging set)
    auto mma_cols = [&](int buf) {
        const u32x4* la = lds[buf] + (wm * TM * X6_PARTS) * X6_FRAG + lane;
        const u32x4* lb = lds[buf] + (FA + wn * TN * X6_PARTS) * X6_FRAG + lane;
        u32x4 a[TM][X6_PARTS];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int p = 0; p < X6_PARTS; p++) a[i][p] = la[(i * X6_PARTS + p) * X6_FRAG];
#pragma unroll
        for (int j = 0; j < TN; j++) {
            u32x4 b[X6_PARTS], nbi[3];
#pragma unroll
            for (int p = 0; p < X6_PARTS; p++) b[p] = lb[(j * X6_PARTS + p) * X6_FRAG];
#pragma unroll
            for (int t = 0; t < 3; t++) nbi[t] = b[3 + t] ^ 0x80008000u;
            constexpr int PS[6] = {2, 1, 0, 1, 0, 0}, PT[6] = {0, 1, 2, 0, 1, 0};
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const int s = PS[q], t = PT[q];
#pragma unroll
                for (int i = 0; i < TM; i++) {
                    re[i][j] = mfma_x6(a[i][s], b[t], re[i][j]);
                    im[i][j] = mfma_x6(a[i][s], b[3 + t], im[i][j]);
                    re[i][j] = mfma_x6(a[i][3 + s], nbi[t], re[i][j]);
                    im[i][j] = mfma_x6(a[i][3 + s], b[t], im[i][j]);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    if constexpr (DEPTH == 1) {
        gload(0, 0);
        sstore(0, 0);
        __syncthreads();
        for (int ks = 0; ks < KS; ks++) {
            const bool more = ks + 1 < KS;
            if (more) gload(0, ks + 1);
            mma(ks & 1);
            if (more) sstore(0, (ks + 1) & 1);
            __syncthreads();
        }
    } else {
        // slab s is fetched into set s % 2 at slab s - 2 and stored to LDS
        // buffer s % 2 at the end of slab s - 1
        gload(0, 0);
        sstore(0, 0);
        if (KS > 1) gload(1, 1);
        if (KS > 2) gload(0, 2);
        __syncthreads();
        for (int ks = 0; ks < KS; ks += 2) {
            // even slab ks in LDS buffer 0; ks + 1 in set 1, ks + 2 in set 0
            mma_cols(0);
            if (ks + 1 < KS) sstore(1, 1);
            if (ks + 3 < KS) gload(1, ks + 3);
            __syncthreads();
            if (ks + 1 >= KS) break;
            mma_cols(1);
            if (ks + 2 < KS) sstore(0, 0);
            if (ks + 4 < KS) gload(0, ks + 4);
            __syncthreads();
        }
    }

    const int m0 = tm * BM + wm * TM * 32, n0 = tn * BN + wn * TN * 32;
    const int kl = lane >> 5, ml = lane & 31;
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int row = m0 + i * 32 + (v & 3) + 8 * (v >> 2) + 4 * kl;
            if (row >= M) continue;
            c32* cr = C + (long)(crow ? crow[row] : row) * g.ldc;
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const int col = n0 + j * 32 + ml;
                if (col >= N) continue;
                x6_store<NTC>(cr + col, re[i][j][v], im[i][j][v], g.mode);
            }
        }
}

// Tiles of a launch: work-group b takes tile b (XCD-aware order: dispatch b
// runs on XCD b % 8, which gets a contiguous range of tiles).
template <int WGM, int WGN, int TM, int TN, int GMW, bool NTC, int DEPTH>
__device__ __forceinline__ void x6l_block(const X6Args& g)
{
    constexpr int BM = 32 * TM * WGM, BN = 32 * TN * WGN;
    const int ntiles = ((g.N + BN - 1) / BN) * ((g.M + BM - 1) / BM);
    const int per = (ntiles + 7) / 8;
    const int vb = blockIdx.x, bid = (vb & 7) * per + (vb >> 3);
    if (bid < ntiles) x6l_tile<WGM, WGN, TM, TN, GMW, NTC, DEPTH>(g, bid);
}

// LDS: staged slabs; otherwise each wave loads its own fragments
template <int WGM, int WGN, int TM, int TN, int LDS, int GMW, bool NTC>
__global__ __launch_bounds__(64 * WGM * WGN, TM * TN >= 4 ? 1 : 2) void k_cgemm_x6(X6Args g)
{
    if constexpr (LDS > 0)
        x6l_block<WGM, WGN, TM, TN, GMW, NTC, LDS>(g);
    else
        x6_block<WGM, WGN, TM, TN, GMW, NTC>(g);
}
// the same product under the trailing update's name (see k_trailing_update)
template <int WGM, int WGN, int TM, int TN, int LDS, int GMW, bool NTC>
__global__ __launch_bounds__(64 * WGM * WGN, TM * TN >= 4 ? 1 : 2) void k_trailing_update_x6(X6Args g)
{
    if constexpr (LDS > 0)
        x6l_block<WGM, WGN, TM, TN, GMW, NTC, LDS>(g);
    else
        x6_block<WGM, WGN, TM, TN, GMW, NTC>(g);
}

constexpr long kX6StreamEntries = 1L << 25;

// Packs both operands into the product's work space, then multiplies.
template <int WGM, int WGN, int TM, int TN, int LDS, int GMW = 8, bool NTC = false>
static int launch_x6(const GemmArgs<float>& g, hipStream_t st, bool trailing)
{
    const long need = gemm_x6_ws_bytes(g.M, g.N, g.K);
    if (need <= 0 || !g.ws || g.ws_item < need || g.ws_item % 16) return TAGS_ERR_ARG;
    const long wa = x6_pack_words(g.M, g.K), wb = x6_pack_words(g.N, g.K);
    u32x4* pa = (u32x4*)g.ws;
    u32x4* pb = pa + wa;
    const long spk = g.ws_item / 16;
    // threads: one per (block, slab, lane) = words / parts
    const long ta = wa / X6_PARTS, tb = wb / X6_PARTS;
    const dim3 ga((unsigned)((ta + 255) / 256), (unsigned)g.nb);
    const dim3 gb((unsigned)((tb + 255) / 256), (unsigned)g.nb);
    if (trailing) {
        hipLaunchKernelGGL((k_pack_x6<0, true>), ga, dim3(256), 0, st, g.A, g.lda, g.arow, g.M,
                           g.K, pa, g.sa, g.sar, spk);
        hipLaunchKernelGGL((k_pack_x6<1, true>), gb, dim3(256), 0, st, g.B, g.ldb, g.brow, g.N,
                           g.K, pb, g.sb, g.sbr, spk);
    } else {
        hipLaunchKernelGGL((k_pack_x6<0, false>), ga, dim3(256), 0, st, g.A, g.lda, g.arow, g.M,
                           g.K, pa, g.sa, g.sar, spk);
        hipLaunchKernelGGL((k_pack_x6<1, false>), gb, dim3(256), 0, st, g.B, g.ldb, g.brow, g.N,
                           g.K, pb, g.sb, g.sbr, spk);
    }
    TAGS_HIP(hipGetLastError());
    constexpr int BM = 32 * TM * WGM, BN = 32 * TN * WGN;
    const X6Args x{g.M, g.N, g.K, pa, pb, spk, g.C, g.ldc, g.crow, g.sc, g.scr, g.mode};
    const long tiles = (long)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
    const long grid = (tiles + 7) / 8 * 8;
    const dim3 gd((unsigned)grid, 1, (unsigned)g.nb);
    // C streamed past the caches only where it is large: a C of fewer than
    // kX6StreamEntries entries stays cached (streaming it cost 20 % at 3000 x
    // 3000 and gained 4 % at 16384 x 16384)
    if (NTC && (long)g.M * g.N < kX6StreamEntries) {
        if (trailing)
            hipLaunchKernelGGL((k_trailing_update_x6<WGM, WGN, TM, TN, LDS, GMW, false>), gd,
                               dim3(64 * WGM * WGN), 0, st, x);
        else
            hipLaunchKernelGGL((k_cgemm_x6<WGM, WGN, TM, TN, LDS, GMW, false>), gd,
                               dim3(64 * WGM * WGN), 0, st, x);
    } else if (trailing) {
        hipLaunchKernelGGL((k_trailing_update_x6<WGM, WGN, TM, TN, LDS, GMW, NTC>), gd,
                           dim3(64 * WGM * WGN), 0, st, x);
    } else {
        hipLaunchKernelGGL((k_cgemm_x6<WGM, WGN, TM, TN, LDS, GMW, NTC>), gd,
                           dim3(64 * WGM * WGN), 0, st, x);
    }
    TAGS_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// variants and selection
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK, bool M3, int WGM = 2, int WGN = 2, int AL = 0>
static int launch_v(const GemmArgs<T>& g, hipStream_t st, bool trailing)
{
    const long tiles = (long)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
    const long grid = (tiles + 7) / 8 * 8;
    const dim3 block(64 * WGM * WGN);
    const dim3 gd((unsigned)grid, 1, (unsigned)g.nb);
    if (trailing)
        hipLaunchKernelGGL((k_trailing_update<T, BM, BN, BK, M3, WGM, WGN, AL>), gd, block, 0, st,
                           g);
    else
        hipLaunchKernelGGL((k_cgemm<T, BM, BN, BK, M3, WGM, WGN, AL>), gd, block, 0, st, g);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// Variant table (tile BM x BN, K chunk BK, four- or three-multiply complex
// product, waves WGM x WGN).
template <typename T>
struct GemmTable;
struct GemmShape {
    int bm, bn;
};
template <>
struct GemmTable<float> {
    static constexpr int count = 34;
    static int launch(int v, const GemmArgs<float>& g, hipStream_t st, bool tr)
    {
        switch (v) {
        case 0: return launch_v<float, 64, 64, 16, false>(g, st, tr);
        case 1: return launch_v<float, 128, 128, 16, false>(g, st, tr);
        case 2: return launch_v<float, 128, 64, 16, false>(g, st, tr);
        case 3: return launch_v<float, 64, 128, 16, false>(g, st, tr);
        case 4: return launch_v<float, 128, 128, 16, true>(g, st, tr);
        case 5: return launch_v<float, 64, 64, 16, true>(g, st, tr);
        case 6: return launch_v<float, 128, 128, 16, true, 2, 4>(g, st, tr);
        case 7: return launch_v<float, 128, 128, 16, true, 4, 2>(g, st, tr);
        case 8: return launch_v<float, 128, 128, 8, true, 2, 4>(g, st, tr);
        case 9: return launch_v<float, 256, 128, 8, true, 4, 4>(g, st, tr);
        case 10: return launch_v<float, 128, 64, 16, true, 4, 2>(g, st, tr);
        case 11: return launch_v<float, 128, 128, 16, true, 4, 4>(g, st, tr);
        case 12: return launch_v<float, 128, 128, 8, true, 4, 4>(g, st, tr);
        case 13: return launch_v<float, 64, 128, 16, true, 2, 4>(g, st, tr);
        case 14: return launch_v<float, 128, 64, 8, true, 4, 2>(g, st, tr);
        case 15: return launch_v<float, 128, 128, 8, true, 4, 2>(g, st, tr);
        case 16: return launch_v<float, 256, 64, 16, true, 8, 2>(g, st, tr);
        case 18: return launch_v32(g, st, tr);
        case 19: return launch_v<float, 128, 64, 8, true, 4, 2, 1>(g, st, tr);
        case 20: return launch_v<float, 128, 64, 16, true, 4, 2, 1>(g, st, tr);
        case 21: return launch_v<float, 128, 128, 8, true, 4, 4, 1>(g, st, tr);
        // split-bf16 (see k_pack_x6): slabs staged in LDS, fetched two slabs ahead
        case 22: return launch_x6<4, 1, 1, 2, 2>(g, st, tr);
        case 23: return launch_x6<4, 1, 1, 2, 2, 8, true>(g, st, tr);
        case 24: return launch_x6<4, 1, 1, 2, 2, 4>(g, st, tr);
        case 25: return launch_x6<2, 2, 1, 2, 2>(g, st, tr);
        // one slab ahead; without LDS (each wave loads its own fragments)
        case 26: return launch_x6<4, 1, 1, 2, 1>(g, st, tr);
        case 27: return launch_x6<4, 1, 1, 2, 0>(g, st, tr);
        case 28: return launch_x6<2, 2, 2, 2, 0>(g, st, tr);
        // 128 x 128 tiles of 8 waves, two slabs ahead (one work-group per CU)
        case 29: return launch_x6<4, 2, 1, 2, 2>(g, st, tr);
        case 30: return launch_x6<4, 2, 1, 2, 2, 8, true>(g, st, tr);
        // 128 x 128 tiles of 16 waves (32 x 32 each), one slab ahead
        case 31: return launch_x6<4, 4, 1, 1, 1>(g, st, tr);
        // 64 x 64 tiles of 4 waves (32 x 32 each), two slabs ahead
        case 32: return launch_x6<2, 2, 1, 1, 2>(g, st, tr);
        // variant 19 with C streamed past the caches
        case 33: return launch_v<float, 128, 64, 8, true, 4, 2, 3>(g, st, tr);
        default: return launch_v<float, 64, 64, 8, true, 2, 2>(g, st, tr);
        }
    }
    static GemmShape shape(int v)
    {
        static const GemmShape t[count] = {{64, 64},   {128, 128}, {128, 64},  {64, 128},
                                           {128, 128}, {64, 64},   {128, 128}, {128, 128},
                                           {128, 128}, {256, 128}, {128, 64},  {128, 128},
                                           {128, 128}, {64, 128},  {128, 64},  {128, 128},
                                           {256, 64},  {64, 64},   {G32_BM, G32_BN},
                                           {128, 64},  {128, 64},  {128, 128},
                                           {128, 64},  {128, 64},  {128, 64},  {64, 128},
                                           {128, 64},  {128, 64},  {128, 128}, {128, 128},
                                           {128, 128}, {128, 128}, {64, 64},   {128, 64}};
        return t[v];
    }
    static int tile_m(int v) { return shape(v).bm; }
    static int tile_n(int v) { return shape(v).bn; }
};
template <>
struct GemmTable<double> {
    static constexpr int count = 16;
    static int launch(int v, const GemmArgs<double>& g, hipStream_t st, bool tr)
    {
        switch (v) {
        case 0: return launch_v<double, 64, 64, 8, false>(g, st, tr);
        case 1: return launch_v<double, 128, 64, 8, false>(g, st, tr);
        case 2: return launch_v<double, 64, 128, 8, false>(g, st, tr);
        case 3: return launch_v<double, 128, 64, 16, false>(g, st, tr);
        case 4: return launch_v<double, 128, 128, 8, false, 2, 4>(g, st, tr);
        case 5: return launch_v<double, 128, 128, 8, false, 4, 2>(g, st, tr);
        case 6: return launch_v<double, 128, 64, 8, false, 4, 2>(g, st, tr);
        case 7: return launch_v<double, 128, 128, 8, false, 4, 4>(g, st, tr);
        case 8: return launch_v<double, 128, 64, 16, false, 4, 2>(g, st, tr);
        case 9: return launch_v<double, 64, 128, 8, false, 2, 4>(g, st, tr);
        case 10: return launch_v<double, 256, 64, 8, false, 4, 2>(g, st, tr);
        case 11: return launch_v<double, 64, 64, 16, false, 2, 2>(g, st, tr);
        case 13: return launch_v<double, 128, 64, 8, false, 4, 2, 1>(g, st, tr);
        case 14: return launch_v<double, 128, 64, 16, false, 4, 2, 1>(g, st, tr);
        case 15: return launch_v<double, 128, 64, 8, false, 4, 2, 3>(g, st, tr);
        default: return launch_v<double, 64, 64, 4, false, 2, 2>(g, st, tr);
        }
    }
    static GemmShape shape(int v)
    {
        static const GemmShape t[count] = {{64, 64},  {128, 64},  {64, 128}, {128, 64},
                                           {128, 128}, {128, 128}, {128, 64}, {128, 128},
                                           {128, 64}, {64, 128},  {256, 64},  {64, 64},
                                           {64, 64},  {128, 64},  {128, 64},  {128, 64}};
        return t[v];
    }
    static int tile_m(int v) { return shape(v).bm; }
    static int tile_n(int v) { return shape(v).bn; }
};

// Selected variant for large and small problems, [double, float]. The
// defaults are the fastest on MI355X at the LU's shapes (K = 128 for
// complex128, 256 for complex64): for large complex128 products 128 x 64
// tiles of 8 waves (32 x 32 per wave, A tile row-major in LDS), for large
// complex64 ones the split-bf16 product (128 x 64 tiles of 4 waves, slabs
// fetched two ahead, tiles in windows of 4 rows); 64 x 64 tiles of 4 waves (K
// chunks of 8) for small ones; tags_autotune may replace them.
static std::atomic<int> g_big_variant[2] = {13, 24};
static std::atomic<int> g_small_variant[2] = {0, 17};

template <typename T>
static constexpr int prec_index()
{
    return sizeof(T) == sizeof(float) ? 1 : 0;
}

template <typename T>
int launch_gemm(const GemmArgs<T>& g, hipStream_t st, bool trailing)
{
    if (g.M <= 0 || g.N <= 0 || g.nb <= 0) return 0;
    if (g.nb > 65535) return TAGS_ERR_ARG;
    if (g.brow && g.K > GEMM_BROW_LDS) return TAGS_ERR_ARG;
    const int p = prec_index<T>();
    // the calling thread's tuned variants, else the process-wide ones
    const Tuning* tu = tuning_current();
    const int big = tu && tu->gemm_big[p] >= 0 ? tu->gemm_big[p]
                                               : g_big_variant[p].load(std::memory_order_relaxed);
    const long tiles_big = (long)((g.M + GemmTable<T>::tile_m(big) - 1) / GemmTable<T>::tile_m(big)) *
                           ((g.N + GemmTable<T>::tile_n(big) - 1) / GemmTable<T>::tile_n(big)) *
                           (g.nb > 0 ? g.nb : 1);
    const int small = tu && tu->gemm_small[p] >= 0
                          ? tu->gemm_small[p]
                          : g_small_variant[p].load(std::memory_order_relaxed);
    int v = tiles_big >= 512 ? big : small;
    // the split-bf16 variants need the caller's work space (and K <= 256);
    // without it the complex64 product takes the three-multiply f32 tiles.
    // They also serve only products with K > 64, the step-width products the
    // LU queues on its main stream (trailing update, columns of a step): the
    // panel's products (K <= 64) run on the look-ahead stream beside them and
    // must not share the work space.
    if (p == 1 && gemm_variant_is_x6(v) &&
        (!g.ws || g.K <= 64 || g.ws_item < gemm_x6_ws_bytes(g.M, g.N, g.K) ||
         gemm_x6_ws_bytes(g.M, g.N, g.K) <= 0))
        v = tiles_big >= 512 ? 19 : 17;
    return GemmTable<T>::launch(v, g, st, trailing);
}

template <typename T>
int launch_gemm_variant(const GemmArgs<T>& g, int variant, hipStream_t st)
{
    if (g.M <= 0 || g.N <= 0) return 0;
    if (variant < 0 || variant >= GemmTable<T>::count) return TAGS_ERR_ARG;
    if (g.brow && g.K > GEMM_BROW_LDS) return TAGS_ERR_ARG;
    return GemmTable<T>::launch(variant, g, st, false);
}

int gemm_set_variants(int precision, int big, int small)
{
    const int p = precision == 1 ? 1 : 0;
    const int cnt = p ? GemmTable<float>::count : GemmTable<double>::count;
    if (big < 0 || big >= cnt || small < 0 || small >= cnt) return TAGS_ERR_ARG;
    g_big_variant[p] = big;
    g_small_variant[p] = small;
    return 0;
}

int gemm_variant_count(int precision)
{
    return precision == 1 ? GemmTable<float>::count : GemmTable<double>::count;
}

template int launch_gemm<double>(const GemmArgs<double>&, hipStream_t, bool);
template int launch_gemm<float>(const GemmArgs<float>&, hipStream_t, bool);
template int launch_gemm_variant<double>(const GemmArgs<double>&, int, hipStream_t);
template int launch_gemm_variant<float>(const GemmArgs<float>&, int, hipStream_t);

}  // namespace tags
