// Blocked complex LU with partial pivoting, triangular solves and FP64
// residuals on gfx950, for complex128 (plain form) and complex64 (factor of
// the mixed-precision form).
//
// Replaces the reference's LAPACK zsysv call (src/linalg.c:113-139). Layout:
// the matrix is row-major with a logical-to-physical row map `perm`. Partial
// pivoting swaps entries of `perm` instead of moving rows, so a pivot never
// costs HBM traffic; every kernel reads and writes rows through the map. After
// the factorisation row perm[i] holds row i of L (unit, strictly lower) and U,
// i.e. P A = L U with (P A)[i] = A[perm[i]].
//
// Per 128-column step j (two 64-column sub-panels a, b):
//   panel(a)            cooperative kernel, one grid barrier per column
//   inv(diag a)         64 x 64 unit-lower and upper inverses in LDS
//   update b            U_ab = inv(L_aa) A_ab; A_b -= L_a U_ab   (MFMA GEMM)
//   panel(b), inv(diag b)
//   U12                 inv(L11) applied to the 128 pivot rows    (MFMA GEMM)
//   trailing update     A22 -= L21 U12, K = 128                   (MFMA GEMM)
#include <algorithm>
#include <cmath>
#include <vector>

#include "lu.h"

namespace tags {

// ---------------------------------------------------------------------------
// MFMA wrappers: 16x16x4 blocks; lane l supplies A[l%16][l/16] and
// B[l/16][l%16]. Result element v of lane l is D[row(l, v)][l%16] with
// row = l/16 + 4v for the f64 instruction and 4(l/16) + v for the f32 one
// (measured on gfx950).
// ---------------------------------------------------------------------------
template <typename T>
struct Mfma;
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

// ---------------------------------------------------------------------------
// Complex GEMM  C (+)= s * A B  with row maps, row-major operands.
// Work-group tile BM x BN, 4 waves in a 2 x 2 arrangement, K staged through
// LDS in chunks of BK; each wave owns (BM/2) x (BN/2) as 16 x 16 MFMA blocks,
// four real MFMAs per complex block product (re*re - im*im, re*im + im*re).
// ---------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK>
__global__ __launch_bounds__(256) void k_cgemm(GemmArgs<T> g)
{
    using V4 = typename Mfma<T>::v4;
    constexpr int TM = BM / 32, TN = BN / 32;
    __shared__ cx<T> As[BM][BK + 1];
    __shared__ cx<T> Bs[BK][BN + 1];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int tiles_n = (g.N + BN - 1) / BN;
    const int tiles_m = (g.M + BM - 1) / BM;
    const int ntiles = tiles_n * tiles_m;
    // XCD-aware order: consecutive dispatch ids land on different XCDs; give
    // each XCD a contiguous range of tiles (row-band) so A rows stay in its L2.
    int bid = blockIdx.x;
    {
        const int per = (ntiles + 7) / 8;
        const int xcd = bid & 7, idx = bid >> 3;
        bid = xcd * per + idx;  // a bijection of [0, 8 * per)
    }
    if (bid >= ntiles) return;
    const int tm = bid / tiles_n, tn = bid % tiles_n;
    const int m0 = tm * BM, n0 = tn * BN;

    V4 are[TM][TN], aim[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
            are[i][j] = V4{0, 0, 0, 0};
            aim[i][j] = V4{0, 0, 0, 0};
        }

    for (int k0 = 0; k0 < g.K; k0 += BK) {
        __syncthreads();
        for (int e = tid; e < BM * BK; e += 256) {
            const int r = e / BK, c = e % BK;
            const int gr = m0 + r, gc = k0 + c;
            cx<T> v = cx<T>{0, 0};
            if (gr < g.M && gc < g.K) {
                const long pr = g.arow ? g.arow[gr] : gr;
                v = g.A[pr * g.lda + gc];
            }
            As[r][c] = v;
        }
        for (int e = tid; e < BK * BN; e += 256) {
            const int r = e / BN, c = e % BN;
            const int gr = k0 + r, gc = n0 + c;
            cx<T> v = cx<T>{0, 0};
            if (gr < g.K && gc < g.N) {
                const long pr = g.brow ? g.brow[gr] : gr;
                v = g.B[pr * g.ldb + gc];
            }
            Bs[r][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            cx<T> a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; i++)
                a[i] = As[wm * (BM / 2) + i * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int j = 0; j < TN; j++)
                b[j] = Bs[kk + (lane >> 4)][wn * (BN / 2) + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    are[i][j] = Mfma<T>::run(a[i].x, b[j].x, are[i][j]);
                    are[i][j] = Mfma<T>::run(-a[i].y, b[j].y, are[i][j]);
                    aim[i][j] = Mfma<T>::run(a[i].x, b[j].y, aim[i][j]);
                    aim[i][j] = Mfma<T>::run(a[i].y, b[j].x, aim[i][j]);
                }
        }
    }
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = m0 + wm * (BM / 2) + i * 16 + Mfma<T>::row(lane, v);
                const int col = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
                if (row < g.M && col < g.N) {
                    const long pr = g.crow ? g.crow[row] : row;
                    cx<T>* p = g.C + pr * g.ldc + col;
                    if (g.mode == 0) {
                        *p = cx<T>{are[i][j][v], aim[i][j][v]};
                    } else {
                        cx<T> c = *p;
                        *p = cx<T>{c.x - are[i][j][v], c.y - aim[i][j][v]};
                    }
                }
            }
}

template <typename T>
int launch_gemm(const GemmArgs<T>& g, hipStream_t st)
{
    if (g.M <= 0 || g.N <= 0) return 0;
    constexpr int BM = 64, BN = 64, BK = 16;
    const long tiles = (long)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
    const long grid = (tiles + 7) / 8 * 8;
    hipLaunchKernelGGL((k_cgemm<T, BM, BN, BK>), dim3((unsigned)grid), dim3(256), 0, st, g);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Cooperative panel factorisation of `w` <= 64 columns starting at column c0
// over logical rows r0 .. r0+M-1. Every work-group holds R rows of the panel in
// LDS. Per column: local pivot search, one 64-bit atomicMax on a packed key
// (39 bits of |re|+|im| | 16 bits inverted row offset | 9 bits work-group),
// grid barrier, broadcast of the winning row through a global candidate slot,
// relabelling of the two logical positions, scaling and rank-1 update.
// ---------------------------------------------------------------------------
__device__ inline bool grid_barrier(unsigned* bar, unsigned target, int* err)
{
    __syncthreads();
    __shared__ int s_fail;
    if (threadIdx.x == 0) {
        s_fail = 0;
        __threadfence();
        atomicAdd(bar, 1u);
        long spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if ((++spins & 1023) == 0) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                    spins > (1L << 24)) {
                    atomicExch(err, 1);
                    s_fail = 1;
                    break;
                }
            }
        }
        __threadfence();
    }
    __syncthreads();
    return s_fail == 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_panel(PanelArgs<T> p)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* rows = reinterpret_cast<cx<T>*>(smem);   // [R][64]
    int* phys = reinterpret_cast<int*>(rows + (size_t)p.R * 64);
    int* lpos = phys + p.R;
    __shared__ cx<T> prow[64];
    __shared__ unsigned long long red[4];
    __shared__ unsigned long long s_key;
    __shared__ int s_best;

    const int P = gridDim.x, wg = blockIdx.x, t = threadIdx.x;
    const int base = p.r0 + wg * p.R;
    const int rend = p.r0 + p.M;
    for (int r = t; r < p.R; r += 256) {
        const int lp = base + r;
        if (lp < rend) {
            phys[r] = p.perm[lp];
            lpos[r] = lp;
        } else {
            phys[r] = -1;
            lpos[r] = -1;
        }
    }
    __syncthreads();
    for (int e = t; e < p.R * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        cx<T> v = cx<T>{0, 0};
        if (phys[r] >= 0 && c < p.w) v = p.A[(long)phys[r] * p.ld + p.c0 + c];
        rows[e] = v;
    }
    __syncthreads();

    bool ok = true;
    for (int c = 0; c < p.w; c++) {
        const int d = p.r0 + c;
        unsigned long long best = 0;
        for (int r = t; r < p.R; r += 256) {
            if (lpos[r] >= d) {
                const double v = (double)cabs1(rows[r * 64 + c]);
                const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
                const unsigned long long key =
                    ((bits >> 25) << 25) |
                    ((unsigned long long)(0xFFFF - (lpos[r] - p.r0)) << 9) | (unsigned)wg;
                best = key > best ? key : best;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(best, off);
            best = o > best ? o : best;
        }
        if ((t & 63) == 0) red[t >> 6] = best;
        __syncthreads();
        if (t == 0) {
            unsigned long long b = red[0];
            for (int i = 1; i < 4; i++) b = red[i] > b ? red[i] : b;
            s_key = b;
            s_best = -1;
        }
        __syncthreads();
        const unsigned long long wkey = s_key;
        if (wkey) {
            const int lpw = p.r0 + 0xFFFF - (int)((wkey >> 9) & 0xFFFF);
            for (int r = t; r < p.R; r += 256)
                if (lpos[r] == lpw) s_best = r;
            __syncthreads();
            cx<T>* dst = p.cand + ((size_t)(c & 1) * P + wg) * 64;
            const int rb = s_best;
            for (int q = t; q < 64; q += 256)
                if (q >= c && q < p.w) dst[q] = rows[rb * 64 + q];
            if (t == 0) atomicMax(&p.keys[c % 3], wkey);
        }
        ok = grid_barrier(p.bar, (unsigned)(c + 1) * P, p.err);
        if (!ok) break;
        if (wg == 0 && t == 0) p.keys[(c + 2) % 3] = 0ull;
        const unsigned long long key =
            __hip_atomic_load(&p.keys[c % 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int pwg = (int)(key & 0x1FF);
        const int plp = p.r0 + 0xFFFF - (int)((key >> 9) & 0xFFFF);
        const cx<T>* src = p.cand + ((size_t)(c & 1) * P + pwg) * 64;
        for (int q = t; q < 64; q += 256)
            if (q >= c && q < p.w) prow[q] = src[q];
        __syncthreads();
        for (int r = t; r < p.R; r += 256) {
            const int l = lpos[r];
            if (l == d && plp != d)
                lpos[r] = plp;
            else if (l == plp)
                lpos[r] = d;
        }
        const cx<T> piv = prow[c];
        const bool zero = (piv.x == T(0) && piv.y == T(0));
        if (zero && wg == 0 && t == 0 && *p.info == 0) *p.info = d + 1;
        __syncthreads();
        if (!zero) {
            const cx<T> inv = crecip(piv);
            for (int r = t; r < p.R; r += 256)
                if (lpos[r] > d) rows[r * 64 + c] = cmul(rows[r * 64 + c], inv);
            __syncthreads();
            const int nc = p.w - c - 1;
            if (nc > 0) {
                for (int e = t; e < p.R * nc; e += 256) {
                    const int r = e / nc, q = c + 1 + e % nc;
                    if (lpos[r] > d)
                        rows[r * 64 + q] = csub(rows[r * 64 + q], cmul(rows[r * 64 + c], prow[q]));
                }
            }
        }
        __syncthreads();
    }
    if (ok) {
        for (int e = t; e < p.R * 64; e += 256) {
            const int r = e >> 6, q = e & 63;
            if (phys[r] >= 0 && q < p.w) p.A[(long)phys[r] * p.ld + p.c0 + q] = rows[e];
        }
        for (int r = t; r < p.R; r += 256)
            if (phys[r] >= 0) p.perm[lpos[r]] = phys[r];
    }
    __syncthreads();
    if (t == 0) {
        __threadfence();
        const unsigned prev = atomicAdd(&p.bar[1], 1u);
        if (prev == (unsigned)P - 1) {
            p.bar[0] = 0;
            p.bar[1] = 0;
            p.keys[0] = p.keys[1] = p.keys[2] = 0ull;
            __threadfence();
        }
    }
}

// ---------------------------------------------------------------------------
// Inverses of the 64 x 64 diagonal block at (c0, c0): unit lower L and upper U,
// inverted in place in LDS (columns right-to-left for L, left-to-right for U),
// written as dense 64 x 64 row-major blocks (identity padding when w < 64).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_diag_inv(const cx<T>* A, long ld, const int* perm,
                                                   int c0, int w, cx<T>* invL, cx<T>* invU)
{
    __shared__ cx<T> B[64][64];
    const int t = threadIdx.x;
    for (int e = t; e < 64 * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        cx<T> v = cx<T>{0, 0};
        if (r < w && c < w)
            v = A[(long)perm[c0 + r] * ld + c0 + c];
        else if (r == c)
            v = cx<T>{1, 0};
        B[r][c] = v;
    }
    __syncthreads();
    // L^{-1}: X[i][j] = -(L[i][j] + sum_{k=j+1}^{i-1} X[i][k] L[k][j]),  i > j
    for (int j = 62; j >= 0; j--) {
        cx<T> s = cx<T>{0, 0};
        const int i = t;
        if (i < 64 && i > j) {
            s = B[i][j];
            for (int k = j + 1; k < i; k++) s = cadd(s, cmul(B[i][k], B[k][j]));
        }
        __syncthreads();
        if (i < 64 && i > j) B[i][j] = cx<T>{-s.x, -s.y};
        __syncthreads();
    }
    // U^{-1}: X[j][j] = 1/u_jj; X[i][j] = -(sum_{k=i}^{j-1} X[i][k] U[k][j]) / u_jj
    for (int j = 0; j < 64; j++) {
        const cx<T> ujj = B[j][j];
        const cx<T> inv = crecip(ujj);
        cx<T> s = cx<T>{0, 0};
        const int i = t;
        if (i < j) {
            for (int k = i; k < j; k++) s = cadd(s, cmul(B[i][k], B[k][j]));
        }
        __syncthreads();
        if (i < j) {
            const cx<T> v = cmul(s, inv);
            B[i][j] = cx<T>{-v.x, -v.y};
        }
        if (i == j) B[j][j] = inv;
        __syncthreads();
    }
    for (int e = t; e < 64 * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        invL[e] = r > c ? B[r][c] : (r == c ? cx<T>{1, 0} : cx<T>{0, 0});
        invU[e] = r <= c ? B[r][c] : cx<T>{0, 0};
    }
}

// dst[r][c] = A[rowmap[r]][c0 + c] for r < rows, c < cols (dst ld = cols)
template <typename T>
__global__ void k_gather_rows(const cx<T>* A, long ld, const int* rowmap, int c0, int rows,
                              int cols, cx<T>* dst)
{
    const int r = blockIdx.y;
    if (r >= rows) return;
    const long pr = rowmap[r];
    for (int c = blockIdx.x * 256 + threadIdx.x; c < cols; c += gridDim.x * 256)
        dst[(long)r * cols + c] = A[pr * ld + c0 + c];
}

__global__ void k_iota(int* p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = i;
}

// ---------------------------------------------------------------------------
// Triangular solves (one right-hand side, logical order) with the inverted
// diagonal blocks; vector type V (the factor precision).
// ---------------------------------------------------------------------------
// y[r0 .. r0+64) = inv * y[r0 .. r0+64)   (inv: 64 x 64 row-major)
template <typename T>
__global__ __launch_bounds__(64) void k_block_apply(const cx<T>* inv, cx<T>* y, int r0, int w)
{
    __shared__ cx<T> v[64];
    const int t = threadIdx.x;
    v[t] = t < w ? y[r0 + t] : cx<T>{0, 0};
    __syncthreads();
    cx<T> s = cx<T>{0, 0};
    for (int k = 0; k < 64; k++) s = cadd(s, cmul(inv[t * 64 + k], v[k]));
    if (t < w) y[r0 + t] = s;
}

// y[i] -= sum_{k<w} A[perm[i]][c0+k] * y[c0+k]  for logical rows i in [lo, hi)
// one wave per row
template <typename T>
__global__ __launch_bounds__(256) void k_panel_gemv(const cx<T>* A, long ld, const int* perm,
                                                     cx<T>* y, int lo, int hi, int c0, int w)
{
    const int i = lo + blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= hi) return;
    cx<T> s = cx<T>{0, 0};
    if (lane < w) s = cmul(A[(long)perm[i] * ld + c0 + lane], y[c0 + lane]);
    for (int off = 32; off > 0; off >>= 1) {
        s.x += __shfl_xor(s.x, off);
        s.y += __shfl_xor(s.y, off);
    }
    if (lane == 0) y[i] = csub(y[i], s);
}

// ---------------------------------------------------------------------------
// FP64 residual r = b - W x (W row-major n x n) and |r|_inf, |x|_inf, |W|_inf
// reductions (non-negative doubles compare as unsigned 64-bit integers).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_residual(const c64* W, long ld, int n, const c64* x,
                                                   const c64* b, c64* r,
                                                   unsigned long long* norms, int want_wnorm)
{
    const int row = blockIdx.x;
    const c64* Wr = W + (long)row * ld;
    double sx = 0, sy = 0, an = 0;
    for (int c = threadIdx.x; c < n; c += 256) {
        const c64 a = Wr[c], v = x[c];
        sx += a.x * v.x - a.y * v.y;
        sy += a.x * v.y + a.y * v.x;
        if (want_wnorm) an += sqrt(a.x * a.x + a.y * a.y);
    }
    __shared__ double red[3][4];
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
        an += __shfl_xor(an, off);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = sx;
        red[1][wv] = sy;
        red[2][wv] = an;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tx = 0, ty = 0, ta = 0;
        for (int i = 0; i < 4; i++) {
            tx += red[0][i];
            ty += red[1][i];
            ta += red[2][i];
        }
        const c64 rr = c64{b[row].x - tx, b[row].y - ty};
        r[row] = rr;
        const double ra = sqrt(rr.x * rr.x + rr.y * rr.y);
        const double xa = sqrt(x[row].x * x[row].x + x[row].y * x[row].y);
        atomicMax(&norms[0], (unsigned long long)__double_as_longlong(ra));
        atomicMax(&norms[1], (unsigned long long)__double_as_longlong(xa));
        if (want_wnorm) atomicMax(&norms[2], (unsigned long long)__double_as_longlong(ta));
    }
}

// row maxima of |W_ij| (for symmetric equilibration): out[i] = 1/sqrt(max_j |W_ij|)
__global__ __launch_bounds__(256) void k_row_scale(const c64* W, long ld, int n, double* d)
{
    const int row = blockIdx.x;
    double mx = 0;
    for (int c = threadIdx.x; c < n; c += 256) {
        const c64 a = W[(long)row * ld + c];
        mx = fmax(mx, sqrt(a.x * a.x + a.y * a.y));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        d[row] = m > 0 ? 1.0 / sqrt(m) : 1.0;
    }
}

// W32[i][j] = d_i W64[i][j] d_j
__global__ __launch_bounds__(256) void k_to_c32_scaled(const c64* W, long ld, int n,
                                                        const double* d, c32* W32, long ld32)
{
    const int row = blockIdx.y;
    const double di = d[row];
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) {
        const c64 a = W[(long)row * ld + c];
        const double s = di * d[c];
        W32[(long)row * ld32 + c] = c32{(float)(a.x * s), (float)(a.y * s)};
    }
}

// t[i] = (float) (d[perm[i]] * r[perm[i]] * inv_scale)
__global__ void k_prep_rhs32(const c64* r, const double* d, const int* perm, int n,
                             double inv_scale, c32* t)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    const double s = d[p] * inv_scale;
    t[i] = c32{(float)(r[p].x * s), (float)(r[p].y * s)};
}

// x[j] += scale * d[j] * z[j]
__global__ void k_update_x(c64* x, const c32* z, const double* d, int n, double scale)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const double s = scale * d[j];
    x[j] = c64{x[j].x + s * (double)z[j].x, x[j].y + s * (double)z[j].y};
}

// t[i] = b[perm[i]]
template <typename T>
__global__ void k_gather_vec(const cx<T>* b, const int* perm, int n, cx<T>* t)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) t[i] = b[perm[i]];
}

__global__ void k_fill_c64(c64* p, long n, c64 v)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// W (row-major view of a column-major symmetric matrix whose lower triangle is
// stored): copy the upper triangle of the view into its lower triangle.
__global__ __launch_bounds__(256) void k_sym_from_upper(c64* W, long ld, int n)
{
    __shared__ c64 tile[32][33];
    const int bi = blockIdx.y, bj = blockIdx.x;  // tile (bi, bj) with bj > bi? use bi <= bj
    if (bj < bi) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = bi * 32 + k, c = bj * 32 + tx;
        if (r < n && c < n) tile[k][tx] = W[(long)r * ld + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int r = bj * 32 + k, c = bi * 32 + tx;  // (r, c) = transpose of (c, r)
        if (r < n && c < n && r > c) W[(long)r * ld + c] = tile[tx][k];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
int LUWork<T>::alloc(int n_, int max_panel_wgs, size_t panel_lds_limit)
{
    num_cu = max_panel_wgs;
    if (n_ <= 0) return TAGS_ERR_ARG;
    if (n_ >= 65536) return TAGS_ERR_TOO_LARGE;
    if (n == n_) return 0;
    release();
    n = n_;
    nblk = (n + 63) / 64;
    TAGS_HIP(hipMalloc(&perm, sizeof(int) * n));
    TAGS_HIP(hipMalloc(&invL, sizeof(cx<T>) * 64 * 64 * nblk));
    TAGS_HIP(hipMalloc(&invU, sizeof(cx<T>) * 64 * 64 * nblk));
    TAGS_HIP(hipMalloc(&tmp, sizeof(cx<T>) * 128 * (size_t)n));
    TAGS_HIP(hipMalloc(&cand, sizeof(cx<T>) * 2 * 512 * 64));
    TAGS_HIP(hipMalloc(&keys, sizeof(unsigned long long) * 4));
    TAGS_HIP(hipMalloc(&bar, sizeof(unsigned) * 4));
    TAGS_HIP(hipMalloc(&flags, sizeof(int) * 4));
    TAGS_HIP(hipMalloc(&vec, sizeof(cx<T>) * n));
    TAGS_HIP(hipMemset(keys, 0, sizeof(unsigned long long) * 4));
    TAGS_HIP(hipMemset(bar, 0, sizeof(unsigned) * 4));
    TAGS_HIP(hipMemset(flags, 0, sizeof(int) * 4));
    max_wgs = max_panel_wgs;
    lds_limit = panel_lds_limit;
    return 0;
}

template <typename T>
void LUWork<T>::release()
{
    hipFree(perm);
    hipFree(invL);
    hipFree(invU);
    hipFree(tmp);
    hipFree(cand);
    hipFree(keys);
    hipFree(bar);
    hipFree(flags);
    hipFree(vec);
    perm = nullptr;
    invL = invU = tmp = cand = vec = nullptr;
    keys = nullptr;
    bar = nullptr;
    flags = nullptr;
    n = 0;
}

template <typename T>
static int gemm(Profiler* pr, hipStream_t st, int M, int N, int K, const cx<T>* A, long lda,
                const int* arow, const cx<T>* B, long ldb, const int* brow, cx<T>* C, long ldc,
                const int* crow, int mode, bool trailing)
{
    GemmArgs<T> g{M, N, K, A, lda, arow, B, ldb, brow, C, ldc, crow, mode};
    Timed tm(pr, FAM_GEMM, st);
    int rc = launch_gemm<T>(g, st);
    tm.end(8.0 * M * N * (double)K, trailing);
    return rc;
}

template <typename T>
static int panel(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld, int c0,
                 int cols)
{
    const int M = w.n - c0;
    // rows per work-group: at least 64, bounded by the LDS a work-group may use
    const size_t per_row = 64 * sizeof(cx<T>) + 2 * sizeof(int);
    const int rmax = (int)((w.lds_limit - 2048) / per_row);
    int P = std::min(w.max_wgs, (M + 63) / 64);
    if (P < 1) P = 1;
    int R = (M + P - 1) / P;
    if (R > rmax) {
        R = rmax;
        P = (M + R - 1) / R;
    }
    {
        // every work-group of a cooperative launch must be resident at once
        int occ = 0;
        TAGS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &occ, (const void*)k_panel<T>, 256, (size_t)R * per_row));
        if (P > occ * w.num_cu || P > 511) return TAGS_ERR_TOO_LARGE;
    }
    PanelArgs<T> p{A, ld, w.perm, c0, M, c0, cols, R, w.keys, w.cand, w.bar, w.flags,
                   w.flags + 1};
    const size_t lds = (size_t)R * per_row;
    void* args[] = {&p};
    Timed tm(pr, FAM_PANEL, st);
    TAGS_HIP(hipLaunchCooperativeKernel((const void*)k_panel<T>, dim3(P), dim3(256), args,
                                        (unsigned)lds, st));
    tm.end();
    return 0;
}

template <typename T>
int LUWork<T>::setup_kernel_attrs()
{
    static bool done = false;
    if (done) return 0;
    // raise the dynamic-LDS cap of the panel kernel; keep the default cap
    // (and a smaller panel chunk) when the runtime refuses
    if (hipFuncSetAttribute((const void*)k_panel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_limit) != hipSuccess) {
        (void)hipGetLastError();
        lds_limit = 65536;
    }
    done = true;
    return 0;
}

template <typename T>
int lu_factor(LUWork<T>& w, cx<T>* A, long ld, hipStream_t st, Profiler* pr)
{
    const int n = w.n;
    int rc = w.setup_kernel_attrs();
    if (rc) return rc;
    hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, st, w.perm, n);
    TAGS_HIP(hipMemsetAsync(w.flags, 0, sizeof(int) * 2, st));
    int* perm = w.perm;
    for (int j = 0; j < n; j += 128) {
        const int jb = std::min(128, n - j);
        const int wa = std::min(64, jb), wb = jb - wa;
        // sub-panel a
        if ((rc = panel<T>(w, pr, st, A, ld, j, wa))) return rc;
        {
            Timed tm(pr, FAM_OTHER, st);
            hipLaunchKernelGGL(k_diag_inv<T>, dim3(1), dim3(256), 0, st, A, ld, perm, j, wa,
                               w.invL + (size_t)(j / 64) * 4096, w.invU + (size_t)(j / 64) * 4096);
            tm.end();
        }
        if (wb > 0) {
            const int c0 = j + 64;
            // U_ab = inv(L_aa) A_ab
            {
                Timed tm(pr, FAM_OTHER, st);
                hipLaunchKernelGGL(k_gather_rows<T>, dim3(1, 64), dim3(256), 0, st, A, ld,
                                   perm + j, c0, 64, wb, w.tmp);
                tm.end();
            }
            if ((rc = gemm<T>(pr, st, 64, wb, 64, w.invL + (size_t)(j / 64) * 4096, 64, nullptr,
                              w.tmp, wb, nullptr, A + c0, ld, perm + j, 0, false)))
                return rc;
            // A_b -= L_a U_ab  (logical rows j+64 .. n)
            if ((rc = gemm<T>(pr, st, n - c0, wb, 64, A + j, ld, perm + c0, A + c0, ld,
                              perm + j, A + c0, ld, perm + c0, 1, false)))
                return rc;
            if ((rc = panel<T>(w, pr, st, A, ld, c0, wb))) return rc;
            Timed tm(pr, FAM_OTHER, st);
            hipLaunchKernelGGL(k_diag_inv<T>, dim3(1), dim3(256), 0, st, A, ld, perm, c0, wb,
                               w.invL + (size_t)(c0 / 64) * 4096,
                               w.invU + (size_t)(c0 / 64) * 4096);
            tm.end();
        }
        const int c1 = j + jb;
        const int M2 = n - c1;
        if (M2 <= 0) break;
        // U12 rows j .. j+64: inv(L_aa) A
        {
            Timed tm(pr, FAM_OTHER, st);
            hipLaunchKernelGGL(k_gather_rows<T>, dim3((M2 + 255) / 256 < 64 ? (M2 + 255) / 256 : 64, wa),
                               dim3(256), 0, st, A, ld, perm + j, c1, wa, M2, w.tmp);
            tm.end();
        }
        if ((rc = gemm<T>(pr, st, wa, M2, wa, w.invL + (size_t)(j / 64) * 4096, 64, nullptr,
                          w.tmp, M2, nullptr, A + c1, ld, perm + j, 0, false)))
            return rc;
        if (wb > 0) {
            const int c0 = j + 64;
            // rows c0 .. c0+wb: A -= L_ba U_a; then inv(L_bb)
            if ((rc = gemm<T>(pr, st, wb, M2, 64, A + j, ld, perm + c0, A + c1, ld, perm + j,
                              A + c1, ld, perm + c0, 1, false)))
                return rc;
            {
                Timed tm(pr, FAM_OTHER, st);
                hipLaunchKernelGGL(k_gather_rows<T>,
                                   dim3((M2 + 255) / 256 < 64 ? (M2 + 255) / 256 : 64, wb),
                                   dim3(256), 0, st, A, ld, perm + c0, c1, wb, M2, w.tmp);
                tm.end();
            }
            if ((rc = gemm<T>(pr, st, wb, M2, wb, w.invL + (size_t)(c0 / 64) * 4096, 64,
                              nullptr, w.tmp, M2, nullptr, A + c1, ld, perm + c0, 0, false)))
                return rc;
        }
        // trailing update A22 -= L21 U12
        if ((rc = gemm<T>(pr, st, M2, M2, jb, A + j, ld, perm + c1, A + c1, ld, perm + j,
                          A + c1, ld, perm + c1, 1, true)))
            return rc;
    }
    return 0;
}

// Solves (L U) x = P b in place on `y` (logical-order input already gathered:
// y[i] = b[perm[i]]); output x in natural order in `y`.
template <typename T>
int lu_solve_inplace(LUWork<T>& w, const cx<T>* A, long ld, cx<T>* y, hipStream_t st,
                     Profiler* pr)
{
    const int n = w.n;
    Timed tm(pr, FAM_SOLVE, st);
    for (int b = 0; b < w.nblk; b++) {
        const int r0 = b * 64, bw = std::min(64, n - r0);
        hipLaunchKernelGGL(k_block_apply<T>, dim3(1), dim3(64), 0, st,
                           w.invL + (size_t)b * 4096, y, r0, bw);
        const int lo = r0 + bw;
        if (lo < n)
            hipLaunchKernelGGL(k_panel_gemv<T>, dim3((n - lo + 3) / 4), dim3(256), 0, st, A, ld,
                               w.perm, y, lo, n, r0, bw);
    }
    for (int b = w.nblk - 1; b >= 0; b--) {
        const int r0 = b * 64, bw = std::min(64, n - r0);
        hipLaunchKernelGGL(k_block_apply<T>, dim3(1), dim3(64), 0, st,
                           w.invU + (size_t)b * 4096, y, r0, bw);
        if (r0 > 0)
            hipLaunchKernelGGL(k_panel_gemv<T>, dim3((r0 + 3) / 4), dim3(256), 0, st, A, ld,
                               w.perm, y, 0, r0, r0, bw);
    }
    TAGS_HIP(hipGetLastError());
    tm.end();
    return 0;
}

template <typename T>
int lu_gather_rhs(LUWork<T>& w, const cx<T>* b, cx<T>* y, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_vec<T>, dim3((w.n + 255) / 256), dim3(256), 0, st, b, w.perm,
                       w.n, y);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int residual64(const c64* W, long ld, int n, const c64* x, const c64* b, c64* r,
               unsigned long long* norms, bool want_wnorm, hipStream_t st, Profiler* pr)
{
    Timed tm(pr, FAM_RESID, st);
    TAGS_HIP(hipMemsetAsync(norms, 0, sizeof(unsigned long long) * (want_wnorm ? 3 : 2), st));
    hipLaunchKernelGGL(k_residual, dim3(n), dim3(256), 0, st, W, ld, n, x, b, r, norms,
                       want_wnorm ? 1 : 0);
    TAGS_HIP(hipGetLastError());
    tm.end();
    return 0;
}

int row_scale(const c64* W, long ld, int n, double* d, hipStream_t st)
{
    hipLaunchKernelGGL(k_row_scale, dim3(n), dim3(256), 0, st, W, ld, n, d);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int to_c32_scaled(const c64* W, long ld, int n, const double* d, c32* W32, long ld32,
                  hipStream_t st)
{
    unsigned gx = (unsigned)std::min(16, (n + 255) / 256);
    hipLaunchKernelGGL(k_to_c32_scaled, dim3(gx, n), dim3(256), 0, st, W, ld, n, d, W32, ld32);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int prep_rhs32(const c64* r, const double* d, const int* perm, int n, double inv_scale, c32* t,
               hipStream_t st)
{
    hipLaunchKernelGGL(k_prep_rhs32, dim3((n + 255) / 256), dim3(256), 0, st, r, d, perm, n,
                       inv_scale, t);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int update_x(c64* x, const c32* z, const double* d, int n, double scale, hipStream_t st)
{
    hipLaunchKernelGGL(k_update_x, dim3((n + 255) / 256), dim3(256), 0, st, x, z, d, n, scale);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int fill_c64(c64* p, long n, c64 v, hipStream_t st)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_fill_c64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int sym_from_upper(c64* W, long ld, int n, hipStream_t st)
{
    const int T = (n + 31) / 32;
    hipLaunchKernelGGL(k_sym_from_upper, dim3(T, T), dim3(256), 0, st, W, ld, n);
    TAGS_HIP(hipGetLastError());
    return 0;
}

template <typename T>
int gemm_public(const GemmArgs<T>& g, hipStream_t st)
{
    return launch_gemm<T>(g, st);
}

template struct LUWork<double>;
template struct LUWork<float>;
template int lu_factor<double>(LUWork<double>&, c64*, long, hipStream_t, Profiler*);
template int lu_factor<float>(LUWork<float>&, c32*, long, hipStream_t, Profiler*);
template int lu_solve_inplace<double>(LUWork<double>&, const c64*, long, c64*, hipStream_t,
                                      Profiler*);
template int lu_solve_inplace<float>(LUWork<float>&, const c32*, long, c32*, hipStream_t,
                                     Profiler*);
template int lu_gather_rhs<double>(LUWork<double>&, const c64*, c64*, hipStream_t);
template int lu_gather_rhs<float>(LUWork<float>&, const c32*, c32*, hipStream_t);
template int gemm_public<double>(const GemmArgs<double>&, hipStream_t);
template int gemm_public<float>(const GemmArgs<float>&, hipStream_t);

}  // namespace tags
