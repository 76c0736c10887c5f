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
// A factorisation follows an LUPlan: a list of column phases, each with the
// logical row range that can hold non-zeros in its columns and an optional
// extra range of trailing columns. Rows and columns that are structurally zero
// for a phase are skipped, which is exact for partial pivoting (a zero row is
// never the largest candidate unless the whole column is zero).
//
// Per step j of 256 columns (lu_set_step_width; sub-panels t of <= 64 columns):
//   per sub-panel t     pivots and L_tt \ U_tt: tournament pivoting (tslu.hip),
//                       partial pivoting over the whole panel (cooperative
//                       kernel, one grid barrier per column), or, for phases
//                       without row exchanges, the diagonal block's own LU;
//                       inverses of both triangles of the diagonal block;
//                       L21 = A21 inv(U_tt)                          (MFMA GEMM)
//                       and the step's later columns: U = inv(L_tt) A,
//                       A_below -= L_t U                              (MFMA GEMM)
//   U12                 blocked forward solve of the step's pivot rows (MFMA GEMM)
//   trailing update     A22 -= L21 U12, K = 256                        (MFMA GEMM)
// The block inverses are kept for the triangular solves.
//
// Look-ahead: step j first updates the columns of step j+1's panel; step j+1's
// panel work then runs on a second stream with its own copy of the row map
// while the rest of step j's update runs (in two launches, whose boundary lets
// waiting panel work-groups start). The panel works on its own copy of the row
// map, so the rest of step j addresses exactly the physical rows it would
// without the panel running; the two touch disjoint columns, and the next
// step starts from the panel's row map once both are done.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

#include "lu.h"

namespace tags {

// ---------------------------------------------------------------------------
// Cooperative panel factorisation of w <= 64 columns starting at column c0 over
// logical rows r0 .. r0+M-1. Each work-group holds R rows of the panel in LDS.
// Per column: local pivot search; the work-group writes its best candidate
// (packed key: 39 bits of |re|+|im|, 16 bits inverted row offset, 9 bits
// work-group id) and that row into its own slot; one grid barrier (release
// arrival, acquire wait); every work-group reads all keys, agrees on the
// winner, reads its row, relabels the two logical positions, scales and does
// the rank-1 update. No contended atomics besides the barrier counter.
// ---------------------------------------------------------------------------
__device__ inline bool grid_barrier(unsigned* bar, unsigned target, int* err)
{
    __syncthreads();
    __shared__ int s_fail;
    if (threadIdx.x == 0) {
        s_fail = 0;
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        long spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
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
    }
    __syncthreads();
    return s_fail == 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_panel(PanelArgs<T> p)
{
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* rows = reinterpret_cast<cx<T>*>(smem);  // [R][64]
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
            if (lpos[r] >= d && lpos[r] < p.cand_end) {
                const double v = (double)cabs1(rows[r * 64 + c]);
                const unsigned long long bits =
                    (unsigned long long)__double_as_longlong(v) & 0x7FFFFFFFFFFFFFFFull;
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
        PanelSlot<T>* mine = p.slots + (size_t)(c & 1) * P + wg;
        const unsigned long long wkey = s_key;
        if (wkey) {
            const int lpw = p.r0 + 0xFFFF - (int)((wkey >> 9) & 0xFFFF);
            for (int r = t; r < p.R; r += 256)
                if (lpos[r] == lpw) s_best = r;
            __syncthreads();
            const int rb = s_best;
            for (int q = t; q < 64; q += 256)
                if (q >= c && q < p.w) mine->row[q] = rows[rb * 64 + q];
        }
        if (t == 0) mine->key = wkey;
        ok = grid_barrier(p.bar, (unsigned)(c + 1) * P, p.err);
        if (!ok) break;
        // agree on the winner: max over all slots' keys
        unsigned long long kb = 0;
        for (int i = t; i < P; i += 256) {
            const unsigned long long k = p.slots[(size_t)(c & 1) * P + i].key;
            kb = k > kb ? k : kb;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(kb, off);
            kb = o > kb ? o : kb;
        }
        if ((t & 63) == 0) red[t >> 6] = kb;
        __syncthreads();
        if (t == 0) {
            unsigned long long b = red[0];
            for (int i = 1; i < 4; i++) b = red[i] > b ? red[i] : b;
            s_key = b;
        }
        __syncthreads();
        const unsigned long long key = s_key;
        const int pwg = (int)(key & 0x1FF);
        const int plp = p.r0 + 0xFFFF - (int)((key >> 9) & 0xFFFF);
        const PanelSlot<T>* src = p.slots + (size_t)(c & 1) * P + pwg;
        for (int q = t; q < 64; q += 256)
            if (q >= c && q < p.w) prow[q] = src->row[q];
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
    // the last work-group out resets the barrier counters for the next launch
    __syncthreads();
    if (t == 0) {
        const unsigned prev = __hip_atomic_fetch_add(&p.bar[1], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        if (prev == (unsigned)P - 1) {
            __hip_atomic_store(&p.bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&p.bar[1], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------------------
// Diagonal block of a pivoted panel (w x w at column c0, logical rows
// c0 .. c0+w, physical pm[c0 + r], already L11 \ U11): the dense 64 x 64
// inverses of its unit-lower and upper triangles (identity padding past w),
// which the panel's L21 = A21 inv(U11), the rows' U12 = inv(L11) A12 and the
// triangular solves use. Two work-groups of 256 threads, one per inverse;
// thread (r, p) holds columns 16p .. 16p+15 of row r in registers (rows and
// columns past w are padded with the identity, which the sweeps leave
// unchanged).
//   inv(L): top-down (row k final at step k, X[r] -= L[r][k] X[k] for r > k);
//   inv(U): bottom-up (row k divided by u_kk at its step, X[r] -= U[r][k] X[k]
//           for r < k).
// One barrier per step (row buffers alternate by parity).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_diag_block(const cx<T>* A, long ld, const int* pm, int c0,
                                                    int w, cx<T>* invL, cx<T>* invU, long sa,
                                                    long sp, long sinv)
{
    const long z_ = blockIdx.z;
    A += z_ * sa;
    pm += z_ * sp;
    invL += z_ * sinv;
    invU += z_ * sinv;
    __shared__ cx<T> s_v[2][64];
    __shared__ int s_pm[64];
    __builtin_amdgcn_s_setprio(1);  // on the panel's critical path
    const int t = threadIdx.x;
    const int r = t >> 2, p = t & 3, lane = t & 63;
    const int quad = lane & ~3;
    if (t < 64) s_pm[t] = t < w ? pm[c0 + t] : 0;
    __syncthreads();
    cx<T> a[16];
    {
        const cx<T>* src = r < w ? A + (long)s_pm[r] * ld + c0 + 16 * p : nullptr;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int c = 16 * p + q;
            a[q] = src ? (c < w ? src[q] : cx<T>{0, 0}) : cx<T>{c == r ? T(1) : T(0), T(0)};
        }
    }
    const bool lower = blockIdx.x == 0;
    cx<T> xv[16];
#pragma unroll
    for (int q = 0; q < 16; q++) xv[q] = cx<T>{16 * p + q == r ? T(1) : T(0), T(0)};
    for (int kb = 0; kb < 4; kb++) {
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            // inv(L): row k = 16 kb + kk (top-down); inv(U): row 63 - k (bottom-up)
            const int kbx = lower ? kb : 3 - kb;
            const int k = lower ? kb * 16 + kk : 63 - (kb * 16 + kk), par = kk & 1;
            if (k >= w) continue;  // identity padding: no change
            // column k of this row (the multiplier L[r][k], or U[r][k] and for
            // row k itself its pivot) from the quad's thread kbx
            const cx<T> ak = lower ? a[kk] : a[15 - kk];
            cx<T> c;
            c.x = __shfl(ak.x, quad | kbx);
            c.y = __shfl(ak.y, quad | kbx);
            if (r == k) {
                if (!lower) {
                    const bool z = c.x == T(0) && c.y == T(0);
                    const cx<T> rd = z ? cx<T>{0, 0} : crecip_fast(c);
#pragma unroll
                    for (int q = 0; q < 16; q++) xv[q] = cmul(xv[q], rd);
                }
#pragma unroll
                for (int q = 0; q < 16; q++) s_v[par][16 * p + q] = xv[q];
            }
            __syncthreads();
            if (lower ? r > k : r < k) {
#pragma unroll
                for (int q = 0; q < 16; q++) xv[q] = cfnma(xv[q], c, s_v[par][16 * p + q]);
            }
        }
    }
    cx<T>* out = lower ? invL : invU;
#pragma unroll
    for (int q = 0; q < 16; q++) out[r * 64 + 16 * p + q] = xv[q];
}

// ---------------------------------------------------------------------------
// Diagonal block of a panel without row exchanges (w x w at column c0, logical
// rows c0 .. c0+w, physical pm[c0 + r]): its LU in place and the same two
// inverses as k_diag_block, all in one sweep of w steps with one barrier each.
// Work-group of 512 threads, two groups of 256 with thread (r, p) holding
// columns 16p .. 16p+15 of row r: group 0 the block and the rows of inv(L),
// group 1 the rows of inv(U). At step k row k of U and row k of inv(L) are
// final; the row's owner puts both, and 1/u_kk, in LDS (buffers alternate by
// parity), then
//   group 0, rows r > k:  l = a[r][k] / u_kk,  a[r] -= l u[k],  y[r] -= l y[k]
//                         (L Y = I row by row; l is stored as L[r][k]);
//   group 1, rows r <= k: x[r][k] = t[r][k] / u_kk,  t[r][q] -= x[r][k] u[k][q]
//                         for q > k (X U = I column by column; rows r > k have
//                         t[r][k] = 0 exactly and are skipped).
// Rows and columns past w are the identity's, and steps past w are skipped
// (w is uniform, so whole steps, barriers included). A zero pivot gives zero
// multipliers in its column and is reported in *info (first such column,
// 1-based, as LAPACK); a caller then falls back to a pivoted factorisation.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(512)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2))) void k_diag_factor(cx<T>* A, long ld, const int* pm, int c0,
                                                     int w, cx<T>* invL, cx<T>* invU, int* info,
                                                     long sa, long sp, long sinv, long sf)
{
    const long z_ = blockIdx.z;
    A += z_ * sa;
    pm += z_ * sp;
    invL += z_ * sinv;
    invU += z_ * sinv;
    info += z_ * sf;
    __shared__ cx<T> s_u[2][64];  // row k of U right of k (zero at and left of k)
    __shared__ cx<T> s_y[2][64];  // row k of inv(L)
    __shared__ cx<T> s_rp[2];     // 1 / u_kk (zero for a zero pivot)
    __shared__ int s_pm[64];
    __shared__ int s_zero;
    __builtin_amdgcn_s_setprio(1);  // on the panel's critical path
    const int tid = threadIdx.x, t = tid & 255, grp = tid >> 8;
    const int r = t >> 2, p = t & 3, lane = t & 63;
    const int quad = lane & ~3;
    if (tid < 64) s_pm[tid] = tid < w ? pm[c0 + tid] : 0;
    if (tid == 0) s_zero = -1;
    __syncthreads();
    // group 0: a = row r of the block; both groups: y = row r of the identity
    cx<T> a[16], y[16];
#pragma unroll
    for (int q = 0; q < 16; q++) y[q] = cx<T>{16 * p + q == r ? T(1) : T(0), T(0)};
    if (grp == 0) {
        const cx<T>* src = r < w ? A + (long)s_pm[r] * ld + c0 + 16 * p : nullptr;
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int c = 16 * p + q;
            a[q] = src ? (c < w ? src[q] : cx<T>{0, 0}) : y[q];
        }
    }
    for (int kb = 0; kb < 4 && 16 * kb < w; kb++) {
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            const int k = kb * 16 + kk, par = k & 1;
            if (k >= w) break;
            if (grp == 0 && r == k) {
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    s_u[par][16 * p + q] = 16 * p + q > k ? a[q] : cx<T>{0, 0};
                    s_y[par][16 * p + q] = y[q];
                }
                if (p == kb) {
                    const bool z = a[kk].x == T(0) && a[kk].y == T(0);
                    s_rp[par] = z ? cx<T>{0, 0} : crecip_fast(a[kk]);
                    if (z && s_zero < 0) s_zero = k;
                }
            }
            __syncthreads();
            const cx<T> rp = s_rp[par];
            if (grp == 0) {
                const cx<T> lo = cmul(a[kk], rp);
                cx<T> l;
                l.x = __shfl(lo.x, quad | kb);
                l.y = __shfl(lo.y, quad | kb);
                if (r > k) {
                    if (p == kb) a[kk] = l;
#pragma unroll
                    for (int q = 0; q < 16; q++) {
                        a[q] = cfnma(a[q], l, s_u[par][16 * p + q]);
                        y[q] = cfnma(y[q], l, s_y[par][16 * p + q]);
                    }
                }
            } else {
                const cx<T> xo = cmul(y[kk], rp);
                cx<T> x;
                x.x = __shfl(xo.x, quad | kb);
                x.y = __shfl(xo.y, quad | kb);
                if (r <= k) {
                    if (p == kb) y[kk] = x;
#pragma unroll
                    for (int q = 0; q < 16; q++) y[q] = cfnma(y[q], x, s_u[par][16 * p + q]);
                }
            }
        }
    }
    if (grp == 0 && r < w) {
        cx<T>* dst = A + (long)s_pm[r] * ld + c0 + 16 * p;
#pragma unroll
        for (int q = 0; q < 16; q++)
            if (16 * p + q < w) dst[q] = a[q];
    }
    if (tid == 0 && s_zero >= 0 && *info == 0) *info = c0 + s_zero + 1;
    cx<T>* out = grp == 0 ? invL : invU;
#pragma unroll
    for (int q = 0; q < 16; q++) out[r * 64 + 16 * p + q] = y[q];
}

// U12 of a step for the columns [cb, cb + cols), in place: the blocked forward
// solve over the step's sub-panels t (rows j + 64 t .. of the logical order,
// physical pm[..]; jb <= 256 rows in all)
//     U_t = inv(L_tt) (A_t - sum_{u<t} L_tu U_u),
// with block t's unit-lower inverse (identity padding past its width) at
// invl + 4096 t. With UP the backward solve with the step's upper triangle
// instead, X_t = inv(U_tt) (A_t - sum_{u>t} U_tu X_u), last block first, with
// the upper inverses at invl. Work-group b owns the CW columns cb + CW b ..:
// its solved blocks stay in LDS for the later rows, the off-diagonal blocks
// and the inverses pass through LDS in 64 x KC chunks, and thread (tr, tc)
// holds rows 4 tr .. 4 tr + 3 at columns tc + 16 q (q < CW / 16) of each
// 64 x CW product.
template <typename T, int CW, int KC, bool UP>
__global__ __launch_bounds__(256) void k_step_u12(cx<T>* A, long ld, const int* pm, int j, int jb,
                                                  const cx<T>* invl, int cb, int cols, long sa,
                                                  long sp, long sinv)
{
    const long z_ = blockIdx.z;
    A += z_ * sa;
    pm += z_ * sp;
    invl += z_ * sinv;
    constexpr int CT = CW / 16;
    __shared__ cx<T> Us[4][64][CW];
    __shared__ cx<T> Ls[64][KC + 1];
    __shared__ int s_pm[256];
    const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
    const int c0 = cb + blockIdx.x * CW, cend = cb + cols;
    const int nsub = (jb + 63) / 64;
    s_pm[tid] = tid < jb ? pm[j + tid] : 0;
    __syncthreads();
    for (int tt = 0; tt < nsub; tt++) {
        const int t = UP ? nsub - 1 - tt : tt;
        const int r0 = 64 * t, wt = min(64, jb - r0);
        cx<T> acc[4][CT];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = 4 * tr + i;
#pragma unroll
            for (int q = 0; q < CT; q++) {
                const int c = c0 + tc + 16 * q;
                acc[i][q] = r < wt && c < cend ? A[(long)s_pm[r0 + r] * ld + c] : cx<T>{0, 0};
            }
        }
        // acc -= L_tu U_u over the sub-panels solved before this one
        for (int uu = 0; uu < (UP ? nsub - 1 - t : t); uu++) {
            const int u = UP ? t + 1 + uu : uu;
            for (int kc = 0; kc < 64; kc += KC) {
                for (int e = tid; e < 64 * KC; e += 256) {
                    const int r = e / KC, kk = e % KC, k = 64 * u + kc + kk;
                    Ls[r][kk] = r < wt && k < jb ? A[(long)s_pm[r0 + r] * ld + j + k]
                                                 : cx<T>{0, 0};
                }
                __syncthreads();
#pragma unroll
                for (int kk = 0; kk < KC; kk++) {
                    cx<T> x[CT];
#pragma unroll
                    for (int q = 0; q < CT; q++) x[q] = Us[u][kc + kk][tc + 16 * q];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const cx<T> l = Ls[4 * tr + i][kk];
#pragma unroll
                        for (int q = 0; q < CT; q++) acc[i][q] = cfnma(acc[i][q], l, x[q]);
                    }
                }
                __syncthreads();
            }
        }
        // U_t = inv(L_tt) acc (acc staged in U_t's slot; the inverse negated
        // on staging so that the fused step adds)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int q = 0; q < CT; q++) Us[t][4 * tr + i][tc + 16 * q] = acc[i][q];
        cx<T> out[4][CT];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int q = 0; q < CT; q++) out[i][q] = cx<T>{0, 0};
        const cx<T>* li = invl + (size_t)t * 4096;
        for (int kc = 0; kc < 64; kc += KC) {
            for (int e = tid; e < 64 * KC; e += 256) {
                const int r = e / KC, kk = e % KC;
                const cx<T> v = li[r * 64 + kc + kk];
                Ls[r][kk] = cx<T>{-v.x, -v.y};
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < KC; kk++) {
                cx<T> x[CT];
#pragma unroll
                for (int q = 0; q < CT; q++) x[q] = Us[t][kc + kk][tc + 16 * q];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const cx<T> l = Ls[4 * tr + i][kk];
#pragma unroll
                    for (int q = 0; q < CT; q++) out[i][q] = cfnma(out[i][q], l, x[q]);
                }
            }
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = 4 * tr + i;
#pragma unroll
            for (int q = 0; q < CT; q++) {
                const int c = c0 + tc + 16 * q;
                Us[t][r][tc + 16 * q] = out[i][q];
                if (r < wt && c < cend) A[(long)s_pm[r0 + r] * ld + c] = out[i][q];
            }
        }
        __syncthreads();
    }
}

// complex64 form of k_step_u12 on the MFMA cores: work-group b owns the 32
// columns cb + 32 b ..; wave w forms rows 16 w .. 16 w + 15 of each 64 x 32
// block product as two 16 x 16 tiles of v_mfma_f32_16x16x4f32 (lane l supplies
// A[l % 16][l / 16] and B[l / 16][l % 16]; result value v of lane l is
// D[4 (l / 16) + v][l % 16]), four real products per complex one. The
// accumulators start from A_t and take the off-diagonal blocks negated, so they
// end as A_t - sum L_tu U_u; that is staged in the block's LDS slot and
// multiplied by the staged inverse. The left operands of a block pass through
// LDS in 64 x 8 chunks, double-buffered: the next chunk is fetched into
// registers while the MFMAs of the current one run.
typedef float u12v4 __attribute__((ext_vector_type(4)));
template <bool UP>
__global__ __launch_bounds__(256) void k_step_u12_mfma(c32* A, long ld, const int* pm, int j,
                                                       int jb, const c32* invl, int cb, int cols,
                                                       long sa, long sp, long sinv)
{
    const long z_ = blockIdx.z;
    A += z_ * sa;
    pm += z_ * sp;
    invl += z_ * sinv;
    constexpr int CW = 32, KC = 8, NCH = 64 / KC, LPT = 64 * KC / 256;
    __shared__ c32 Us[4][64][CW];
    __shared__ c32 Ls[2][64][KC + 1];
    __shared__ int s_pm[256];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int c0 = cb + blockIdx.x * CW, cend = cb + cols;
    const int nsub = (jb + 63) / 64;
    s_pm[tid] = tid < jb ? pm[j + tid] : 0;
    __syncthreads();
    for (int tt = 0; tt < nsub; tt++) {
        const int t = UP ? nsub - 1 - tt : tt;
        const int r0 = 64 * t, wt = min(64, jb - r0);
        const int nu = UP ? nsub - 1 - t : t;  // off-diagonal blocks of this row block
        const int nch = NCH * (nu + 1);        // their chunks, then the inverse's
        const c32* li = invl + (size_t)t * 4096;
        // chunk ci: off-diagonal block u = (ci / NCH)-th, or (last NCH) the inverse
        auto fetch = [&](int ci, c32 (&reg)[LPT]) {
            const int kc = (ci % NCH) * KC, b = ci / NCH;
            const int u = UP ? t + 1 + b : b;
#pragma unroll
            for (int q = 0; q < LPT; q++) {
                const int e = tid + 256 * q, r = e / KC, kk = e % KC;
                if (b < nu) {
                    const int k = 64 * u + kc + kk;
                    reg[q] = r < wt && k < jb ? A[(long)s_pm[r0 + r] * ld + j + k] : c32{0, 0};
                } else {
                    reg[q] = li[r * 64 + kc + kk];
                }
            }
        };
        u12v4 ar[2], ai[2], orr[2], oi[2];
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = 16 * wv + 4 * lk + v, c = c0 + 16 * q + lr;
                const c32 x = row < wt && c < cend ? A[(long)s_pm[r0 + row] * ld + c] : c32{0, 0};
                ar[q][v] = x.x;
                ai[q][v] = x.y;
                orr[q][v] = 0;
                oi[q][v] = 0;
            }
        c32 reg[LPT];
        fetch(0, reg);
        for (int ci = 0; ci < nch; ci++) {
            const int buf = ci & 1;
#pragma unroll
            for (int q = 0; q < LPT; q++) {
                const int e = tid + 256 * q;
                Ls[buf][e / KC][e % KC] = reg[q];
            }
            if (ci + 1 < nch) fetch(ci + 1, reg);
            const int b = ci / NCH, kc = (ci % NCH) * KC;
            if (ci == NCH * nu) {
                // the block's right-hand side A_t - sum L_tu U_u into its slot
#pragma unroll
                for (int q = 0; q < 2; q++)
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        Us[t][16 * wv + 4 * lk + v][16 * q + lr] = c32{ar[q][v], ai[q][v]};
            }
            __syncthreads();
            if (b < nu) {
                const int u = UP ? t + 1 + b : b;
#pragma unroll
                for (int kk = 0; kk < KC; kk += 4) {
                    const c32 l = Ls[buf][16 * wv + lr][kk + lk];
#pragma unroll
                    for (int q = 0; q < 2; q++) {
                        const c32 x = Us[u][kc + kk + lk][16 * q + lr];
                        // acc -= l x
                        ar[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(-l.x, x.x, ar[q], 0, 0, 0);
                        ar[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(l.y, x.y, ar[q], 0, 0, 0);
                        ai[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(-l.x, x.y, ai[q], 0, 0, 0);
                        ai[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(-l.y, x.x, ai[q], 0, 0, 0);
                    }
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < KC; kk += 4) {
                    const c32 l = Ls[buf][16 * wv + lr][kk + lk];
#pragma unroll
                    for (int q = 0; q < 2; q++) {
                        const c32 x = Us[t][kc + kk + lk][16 * q + lr];
                        orr[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(l.x, x.x, orr[q], 0, 0, 0);
                        orr[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(-l.y, x.y, orr[q], 0, 0, 0);
                        oi[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(l.x, x.y, oi[q], 0, 0, 0);
                        oi[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(l.y, x.x, oi[q], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = 16 * wv + 4 * lk + v, c = c0 + 16 * q + lr;
                const c32 o = c32{orr[q][v], oi[q][v]};
                Us[t][row][16 * q + lr] = o;
                if (row < wt && c < cend) A[(long)s_pm[r0 + row] * ld + c] = o;
            }
        __syncthreads();
    }
}

// Node block of the nodal plan (stored order [I_L; I_T; U], m segments, n
// nodes): for node row i (logical 2m + i, physical pm[2m + i]),
//     S[i][k] = W[i][2m + k] - sum_c W[i][c] X[c][k],   c in [0, 2m), k in [0, n),
// where W[i][c] are the row's incidence entries (A^T and B^T, untouched by the
// Z_L and Z_T phases) and X[c] = Z^-1 A row of unknown c, at physical row pm[c]
// in the node columns. One work-group per node row: the row's non-zeros are
// gathered into LDS in column order (ballot prefix sums, so the sums do not
// depend on timing), in batches of at most NODAL_NZ_MAX, and after each batch
// every thread subtracts its output columns' share.
constexpr int NODAL_NZ_MAX = 512;
template <typename T>
__global__ __launch_bounds__(256) void k_nodal_schur(cx<T>* A, long ld, const int* pm, int m,
                                                     int n, long sa, long sp)
{
    const long z_ = blockIdx.z;
    A += z_ * sa;
    pm += z_ * sp;
    __shared__ long s_row[NODAL_NZ_MAX];
    __shared__ cx<T> s_cf[NODAL_NZ_MAX];
    __shared__ int s_wcnt[4];
    __shared__ int s_total;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    cx<T>* row = A + (long)pm[2 * m + blockIdx.x] * ld;
    int c_next = 0;
    for (;;) {
        if (tid == 0) s_total = 0;
        __syncthreads();
        // gather while a whole 256-column pass still fits
        while (c_next < 2 * m && s_total <= NODAL_NZ_MAX - 256) {
            const int c = c_next + tid;
            const cx<T> v = c < 2 * m ? row[c] : cx<T>{0, 0};
            const bool nz = v.x != T(0) || v.y != T(0);
            const unsigned long long bal = __ballot(nz);
            if (lane == 0) s_wcnt[wv] = __popcll(bal);
            __syncthreads();
            int pos = s_total + __popcll(bal & ((1ull << lane) - 1ull));
            for (int q = 0; q < wv; q++) pos += s_wcnt[q];
            if (nz) {
                s_row[pos] = (long)pm[c] * ld + 2 * m;
                s_cf[pos] = v;
            }
            __syncthreads();
            if (tid == 0) s_total += s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
            c_next += 256;
            __syncthreads();
        }
        const int cnt = s_total;
        for (int k = tid; k < n && cnt > 0; k += 256) {
            T sx = 0, sy = 0;
            for (int e = 0; e < cnt; e++) {
                const cx<T> a = s_cf[e], x = A[s_row[e] + k];
                sx = fma(a.x, x.x, sx);
                sx = fma(-a.y, x.y, sx);
                sy = fma(a.x, x.y, sy);
                sy = fma(a.y, x.x, sy);
            }
            const cx<T> o = row[2 * m + k];
            row[2 * m + k] = cx<T>{o.x - sx, o.y - sy};
        }
        __syncthreads();
        if (c_next >= 2 * m) break;
    }
}

// y[i] -= sum_{k<w} A[perm[i]][c0+k] y[c0+k] for logical rows i in [lo, hi), any
// w (the rows and the columns [c0, c0 + w) are disjoint); one work-group per row
template <typename T>
__global__ __launch_bounds__(256) void k_rows_gemv(const cx<T>* A, long ld, const int* perm,
                                                    cx<T>* y, int lo, int hi, int c0, int w,
                                                    long sa, long sp, long sv)
{
    const long z_ = blockIdx.z;
    A += z_ * sa;
    perm += z_ * sp;
    y += z_ * sv;
    __shared__ T red[2][4];
    const int i = lo + blockIdx.x;
    if (i >= hi) return;
    const cx<T>* row = A + (long)perm[i] * ld + c0;
    T sx = 0, sy = 0;
    // unrolled so that several loads per lane are in flight
#pragma unroll 4
    for (int k = threadIdx.x; k < w; k += 256) {
        const cx<T> a = row[k], v = y[c0 + k];
        sx = fma(a.x, v.x, sx);
        sx = fma(-a.y, v.y, sx);
        sy = fma(a.x, v.y, sy);
        sy = fma(a.y, v.x, sy);
    }
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = sx;
        red[1][threadIdx.x >> 6] = sy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const T tx = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        const T ty = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        y[i] = cx<T>{y[i].x - tx, y[i].y - ty};
    }
}

// perm[c0 + i] = order[i] for i < len (a phase's row order)
__global__ void k_set_order(int* perm, int c0, int len, const int* order, long sp)
{
    perm += (long)blockIdx.z * sp;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < len) perm[c0 + i] = order[i];
}

// perm = init (or identity); flags (info, barrier error) cleared
__global__ void k_lu_init(int* perm, const int* init, int n, int* flags, long sp, long sf)
{
    perm += (long)blockIdx.z * sp;
    flags += (long)blockIdx.z * sf;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) perm[i] = init ? init[i] : i;
    if (i < 2) flags[i] = 0;
}

// ---------------------------------------------------------------------------
// Triangular solves (one right-hand side, logical order) with the inverted
// diagonal blocks; vector type V (the factor precision).
// ---------------------------------------------------------------------------
// The triangular solve within one factorisation step (logical rows j .. j+jb,
// jb <= 256, sub-blocks t of 64 with inverses at inv + 4096 t): forward,
//     y_t = inv(L_tt) (y_t - sum_{u<t} L_tu y_u),
// or with UP backward with the upper triangle, last block first. One
// work-group of 4 waves; wave w takes rows 16 w .. 16 w + 15 of a block, its
// lanes reading 64 consecutive columns of a row at a time (wave sums by lane
// shuffles).
// Wave sum by DPP (quad swaps, row rotations, then the row broadcasts, whose
// masked-out rows read 0), read from lane 63: uniform.
template <int CTRL, int ROW_MASK>
__device__ inline float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF,
                                                      false));
}
__device__ inline float wave_sum(float v)
{
    v += dpp_f<0xB1, 0xF>(v);
    v += dpp_f<0x4E, 0xF>(v);
    v += dpp_f<0x124, 0xF>(v);
    v += dpp_f<0x128, 0xF>(v);
    v += dpp_f<0x142, 0xA>(v);
    v += dpp_f<0x143, 0xC>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROW_MASK, 0xF,
                                               false);
    return __longlong_as_double(
        (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline double wave_sum(double v)
{
    v += dpp_d<0xB1, 0xF>(v);
    v += dpp_d<0x4E, 0xF>(v);
    v += dpp_d<0x124, 0xF>(v);
    v += dpp_d<0x128, 0xF>(v);
    v += dpp_d<0x142, 0xA>(v);
    v += dpp_d<0x143, 0xC>(v);
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T>
__device__ inline cx<T> wave_sum_cx(cx<T> v)
{
    return cx<T>{wave_sum(v.x), wave_sum(v.y)};
}
template <typename T, bool UP>
__global__ __launch_bounds__(256) void k_step_trsv(const cx<T>* A, long ld, const int* perm,
                                                   cx<T>* y, int j, int jb, const cx<T>* inv,
                                                   long sa, long sp, long sv, long sinv)
{
    const long z_ = blockIdx.z;
    A += z_ * sa;
    perm += z_ * sp;
    y += z_ * sv;
    inv += z_ * sinv;
    __shared__ cx<T> ys[256];
    __shared__ cx<T> acc[64];
    __shared__ int s_pm[256];
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int nsub = (jb + 63) / 64;
    ys[tid] = tid < jb ? y[j + tid] : cx<T>{0, 0};
    s_pm[tid] = tid < jb ? perm[j + tid] : 0;
    __syncthreads();
    for (int tt = 0; tt < nsub; tt++) {
        const int t = UP ? nsub - 1 - tt : tt;
        const int r0 = 64 * t, wt = min(64, jb - r0);
        const int k0 = UP ? r0 + 64 : 0, k1 = UP ? jb : r0;
        // acc[r] = y_t[r] - sum over the solved blocks of the step
        // (k1 - k0 <= 192: every load of the 16 rows is issued before any sum)
        cx<T> part[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int r = 16 * wv + q;
            const cx<T>* row = A + (long)s_pm[r0 + (r < wt ? r : 0)] * ld + j;
            cx<T> s = cx<T>{0, 0};
#pragma unroll
            for (int kk = 0; kk < 3; kk++) {
                // loads at clamped addresses, the sum taken by a select, so that
                // no load sits behind a branch
                const int k = k0 + lane + 64 * kk;
                const bool ok = r < wt && k < k1;
                const cx<T> p = cmul(row[ok ? k : 0], ys[ok ? k : 0]);
                s = csel(ok, cadd(s, p), s);
            }
            part[q] = s;
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const cx<T> s = wave_sum_cx<T>(part[q]);
            const int r = 16 * wv + q;
            if (lane == 0) acc[r] = r < wt ? csub(ys[r0 + r], s) : cx<T>{0, 0};
        }
        __syncthreads();
        // y_t = inv_t acc (the inverse is padded with the identity past wt)
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int r = 16 * wv + q;
            part[q] = cmul(inv[(size_t)t * 4096 + r * 64 + lane], acc[lane]);
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const cx<T> s = wave_sum_cx<T>(part[q]);
            const int r = 16 * wv + q;
            if (lane == 0 && r < wt) ys[r0 + r] = s;
        }
        __syncthreads();
    }
    if (tid < jb) y[j + tid] = ys[tid];
}

// y[i] -= sum_{k<w} A[perm[i]][c0+k] * y[c0+k]  for logical rows i in [lo, hi),
// w <= 256; one wave per row (the rows and the columns [c0, c0 + w) are
// disjoint); the loads go out unconditionally (clamped addresses, sums by
// select), so none waits behind a branch
template <typename T>
__global__ __launch_bounds__(256) void k_panel_gemv(const cx<T>* A, long ld, const int* perm,
                                                     cx<T>* y, int lo, int hi, int c0, int w,
                                                     long sa, long sp, long sv)
{
    const long z_ = blockIdx.z;
    A += z_ * sa;
    perm += z_ * sp;
    y += z_ * sv;
    const int i = lo + blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= hi) return;
    cx<T> s = cx<T>{0, 0};
    const cx<T>* row = A + (long)perm[i] * ld + c0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = lane + 64 * q;
        const bool ok = c < w;
        const cx<T> p = cmul(row[ok ? c : 0], y[c0 + (ok ? c : 0)]);
        s = csel(ok, cadd(s, p), s);
    }
    s = wave_sum_cx<T>(s);
    if (lane == 0) y[i] = csub(y[i], s);
}

// ---------------------------------------------------------------------------
// Flow triangular solve: the whole forward (unit L) or backward (UP: U) solve
// over the diagonal blocks [b_lo, b_hi) of a factorisation in one launch.
// Work-groups take groups of RB consecutive blocks (64 RB rows) in
// dependency order from one ticket counter per launch (the first block first
// forward, the last first backward; group-major over the items of a batched
// launch), so a group only waits on groups that hold earlier tickets and are
// therefore running or done. A group
//   1. streams the off-diagonal part of its rows, block column by block
//      column in dependency order, from HBM: the loads of a chunk of KU block
//      columns go out first; then it waits until the group that owns the
//      chunk's latest block has published (the chain of the solve makes all
//      earlier blocks final too) and reads their solved values;
//   2. solves its own blocks in order with their 64 x 64 inverses, the
//      second block (RB = 2) taking the first one's contribution; the
//      inverse rows and the coupling entries sit in registers, loaded before
//      the first wait, so that only the hand-off and register arithmetic
//      lie on the chain;
//   3. publishes its solved values (wave 0: agent-scope stores, then one
//      release store of the launch's epoch into its completion word).
// Waits spin on relaxed loads with a bounded count (an acquire fence once the
// word reads done); past the bound the group flags *err, writes NaN and
// publishes, so that every group ends. Sums run in a fixed order: the result
// does not depend on timing. Wave v holds rows 16 (v % 4) .. + 15 of the
// group's block v / 4, one block column per lane.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline cx<T> flow_load(const cx<T>* p)
{
    return cx<T>{__hip_atomic_load(&p->x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                 __hip_atomic_load(&p->y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
}
template <typename T>
__device__ inline void flow_store(cx<T>* p, cx<T> v)
{
    __hip_atomic_store(&p->x, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&p->y, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr long kFlowSpinLimit = 1L << 22;

template <typename T, bool UP, int RB>
__global__ __launch_bounds__(256 * RB) void k_trsv_flow(const cx<T>* A_, long ld, const int* perm_,
                                                        cx<T>* y_, const cx<T>* inv_, TrsvGeom g,
                                                        int b_lo, int b_hi, int nitems,
                                                        unsigned* ctl, int* err_, unsigned epoch,
                                                        long sa, long sp, long sv, long sinv,
                                                        long sctl, long sf, FlowPair pr)
{
    static_assert(RB == 1 || RB == 2, "groups of one or two blocks");
    // block columns per chunk of loads (a lane holds 16 KU values of A)
    constexpr int KU = sizeof(T) == 4 ? 2 : 1;
    __shared__ cx<T> acc[64 * RB];
    __shared__ cx<T> xs[64 * RB];
    __shared__ int s_ticket, s_fail;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int nbk = b_hi - b_lo;
    const int ngrp = (nbk + RB - 1) / RB;
    const int total = ngrp * nitems;
    const int nwg = (int)gridDim.x;
    // One ticket counter for the launch (item 0's), tickets group-major over
    // the items, so the groups a work-group waits for were handed out nitems
    // tickets or more before. Each work-group takes groups in turn until it
    // draws a ticket past the last: a few work-groups stream the later
    // groups' blocks while the chain moves on, instead of one resident
    // work-group per group waiting for its turn.
    for (;;) {
        if (tid == 0) {
            const unsigned t =
                __hip_atomic_fetch_add(&ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // every work-group draws exactly one ticket past the last; the last of
            // those draws resets the counter for the next launch
            if ((int)t == total + nwg - 1)
                __hip_atomic_store(&ctl[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_ticket = (int)t;
            s_fail = 0;
        }
        __syncthreads();
        if (s_ticket >= total) return;
        const int gi = s_ticket / nitems;
        const long zv = s_ticket - (long)gi * nitems;
        // item and range (two ranges per item with pr.on)
        const long z_ = pr.on ? zv >> 1 : zv;
        const bool r2 = pr.on && (zv & 1);
        const cx<T>* A = A_ + z_ * sa + (r2 ? pr.a : 0);
        const int* perm = perm_ + z_ * sp + (r2 ? pr.p : 0);
        cx<T>* y = y_ + z_ * sv + (r2 ? pr.v : 0);
        const cx<T>* inv = inv_ + z_ * sinv + (r2 ? pr.inv : 0);
        int* err = err_ + z_ * sf;
        unsigned* done = ctl + z_ * sctl + 2 + (r2 ? pr.done : 0);
        // this group's blocks [q0, q1) and rows [r_lo, r_hi)
        int q0, q1;
        if (!UP) {
            q0 = b_lo + RB * gi;
            q1 = min(q0 + RB, b_hi);
        } else {
            q1 = b_hi - RB * gi;
            q0 = max(q1 - RB, b_lo);
        }
        const int nsb = q1 - q0;
        int r_lo, w0, r_last, w_last;
        trsv_block(g, q0, r_lo, w0);
        trsv_block(g, q1 - 1, r_last, w_last);
        const int r_hi = r_last + w_last;
        // this wave's rows
        const int sb = wv >> 2;  // block of the group
        int sb_row0 = 0, sb_w = 0;
        if (sb < nsb) trsv_block(g, q0 + sb, sb_row0, sb_w);
        const int rb0 = 16 * (wv & 3);
        const cx<T>* rowp[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const bool ok = sb < nsb && rb0 + i < sb_w;
            rowp[i] = A + (long)perm[ok ? sb_row0 + rb0 + i : r_lo] * ld;
        }
        // what the group's own solve needs from memory, loaded before any wait:
        // the rows of this wave's block inverse, and (RB = 2) this wave's entries
        // in the columns of the group's other block when that block comes first
        cx<T> ivr[16], lw[16];
        const int ob = UP ? nsb - 1 : 0;  // the group's first block in dependency order
        int ob_r0 = 0, ob_w = 0;
        if (nsb > 1) trsv_block(g, q0 + ob, ob_r0, ob_w);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            ivr[i] = sb < nsb ? inv[(size_t)(q0 + sb) * 4096 + (rb0 + i) * 64 + lane] : cx<T>{0, 0};
            const bool okc = sb < nsb && sb != ob && lane < ob_w;
            lw[i] = okc ? rowp[i][ob_r0 + lane] : cx<T>{0, 0};
        }
        cx<T> part[16];
#pragma unroll
        for (int i = 0; i < 16; i++) part[i] = cx<T>{0, 0};
        // this wave's right-hand side entries (lane i < 16: row rb0 + i), read
        // before any wait: only this group writes its rows, at its publish
        cx<T> yrhs{0, 0};
        if (lane < 16 && sb < nsb && rb0 + lane < sb_w) yrhs = y[sb_row0 + rb0 + lane];
        // 1. off-diagonal block columns: forward b_lo .. q0 - 1, backward q1 .. b_hi - 1
        const int u_first = UP ? b_hi - 1 : b_lo;
        const int nu = UP ? b_hi - q1 : q0 - b_lo;
        int known = -1;  // highest group index seen published
        for (int c0 = 0; c0 < nu; c0 += KU) {
            cx<T> a[KU][16];
            int ucol[KU], uw[KU];
#pragma unroll
            for (int c = 0; c < KU; c++) {
                const int u = UP ? u_first - (c0 + c) : u_first + (c0 + c);
                const bool in = c0 + c < nu;
                int r0u = 0, wu = 0;
                if (in) trsv_block(g, u, r0u, wu);
                ucol[c] = r0u;
                uw[c] = wu;
                const bool okc = lane < wu;
#pragma unroll
                for (int i = 0; i < 16; i++) a[c][i] = rowp[i][r0u + (okc ? lane : 0)];
            }
            // wait for the group owning the chunk's latest block
            const int last = min(c0 + KU, nu) - 1;
            const int need = last / RB;
            if (need > known) {
                long spins = 0;
                while (__hip_atomic_load(&done[need], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
                       epoch) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > kFlowSpinLimit) {
                        if (lane == 0) {
                            atomicExch(err, 1);
                            s_fail = 1;
                        }
                        break;
                    }
                }
                // the later groups this group reads that are published already:
                // one acquire fence covers them all (a fence waits for every
                // load in flight, so a group behind the chain's front would
                // otherwise drain its loads once per chunk)
                int seen = need;
                const int lim = (nu - 1) / RB;
                while (seen < lim && __hip_atomic_load(&done[seen + 1], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT) == epoch)
                    seen++;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                known = seen;
            }
#pragma unroll
            for (int c = 0; c < KU; c++) {
                const bool okc = c0 + c < nu && lane < uw[c];
                const cx<T> xv = flow_load(y + ucol[c] + (okc ? lane : 0));
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const cx<T> p = cmul(a[c][i], xv);
                    part[i] = csel(okc, cadd(part[i], p), part[i]);
                }
            }
        }
        // right-hand side less the off-diagonal sums, into LDS
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const cx<T> s = wave_sum_cx<T>(part[i]);
            const int r = rb0 + i;
            if (lane == i && sb < nsb) acc[64 * sb + r] = r < sb_w ? csub(yrhs, s) : cx<T>{0, 0};
        }
        // 2. the group's own blocks, in dependency order (RB <= 2: the later
        // block takes the first one's contribution)
        for (int ss = 0; ss < nsb; ss++) {
            const int tb = UP ? nsb - 1 - ss : ss;
            __syncthreads();
            // x_tb = inv_tb acc_tb (identity padding past the block's width), by
            // the block's own waves
            if (sb == tb) {
                cx<T> pr[16];
#pragma unroll
                for (int i = 0; i < 16; i++) pr[i] = cmul(ivr[i], acc[64 * tb + lane]);
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const cx<T> s = wave_sum_cx<T>(pr[i]);
                    if (lane == 0) xs[64 * tb + rb0 + i] = s;
                }
            }
            if (ss + 1 < nsb) {
                __syncthreads();
                if (sb < nsb && sb != ob) {
                    const bool okc = lane < ob_w;
                    const cx<T> xv = xs[64 * ob + (okc ? lane : 0)];
                    cx<T> pr[16];
#pragma unroll
                    for (int i = 0; i < 16; i++) pr[i] = csel(okc, cmul(lw[i], xv), cx<T>{0, 0});
#pragma unroll
                    for (int i = 0; i < 16; i++) {
                        const cx<T> s = wave_sum_cx<T>(pr[i]);
                        if (lane == 0) acc[64 * sb + rb0 + i] = csub(acc[64 * sb + rb0 + i], s);
                    }
                }
            }
        }
        __syncthreads();
        // 3. publish
        if (wv == 0) {
            const bool bad = s_fail != 0;
            for (int r = lane; r < r_hi - r_lo; r += 64) {
                // rows of the group are contiguous: find the block of row r_lo + r
                int rr = r, tb = 0, tw = w0;
                while (tb + 1 < nsb && rr >= tw) {
                    rr -= tw;
                    tb++;
                    int t0;
                    trsv_block(g, q0 + tb, t0, tw);
                }
                cx<T> v = xs[64 * tb + rr];
                if (bad) v = cx<T>{(T)NAN, (T)NAN};
                flow_store(y + r_lo + r, v);
            }
            if (lane == 0)
                __hip_atomic_store(&done[gi], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        // the group's LDS and ticket are free again only after every wave is done
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// FP64 residual r = b - W x (W row-major n x n) and |r|_inf, |x|_inf, |W|_inf
// reductions (non-negative doubles compare as unsigned 64-bit integers).
// ---------------------------------------------------------------------------
// Double-double accumulation for the refinement residuals: a sum kept as
// hi + lo with every product and addition error carried (Dekker/Knuth exact
// transformations: a b = p + fma(a, b, -p), s = hi + p + e by TwoSum), so that
// b - W x is accurate to about one FP64 rounding of its own size even where
// b and W x cancel. With FP64 residuals alone, refinement stops at a forward
// error of about cond(W) eps; with these it converges to working accuracy.
// Contraction is off here: a fused multiply-add in place of the separate
// product and sum would break the error terms.
struct DD {
    double hi, lo;
};
__device__ inline void dd_mac(DD& d, double a, double b)
{
#pragma clang fp contract(off)
    const double p = a * b;
    const double pe = fma(a, b, -p);
    const double s = d.hi + p;
    const double bb = s - d.hi;
    const double se = (d.hi - (s - bb)) + (p - bb);
    d.hi = s;
    d.lo += pe + se;
}
__device__ inline void dd_add(DD& d, double p)
{
#pragma clang fp contract(off)
    const double s = d.hi + p;
    const double bb = s - d.hi;
    const double se = (d.hi - (s - bb)) + (p - bb);
    d.hi = s;
    d.lo += se;
}
__device__ inline DD dd_plus(DD a, DD b)
{
#pragma clang fp contract(off)
    DD r = a;
    dd_add(r, b.hi);
    r.lo += b.lo;
    return r;
}
__device__ inline DD dd_shfl_xor(DD a, int o)
{
    return DD{__shfl_xor(a.hi, o), __shfl_xor(a.lo, o)};
}
// b - (hi + lo), rounded once
__device__ inline double dd_sub_from(double b, DD d)
{
#pragma clang fp contract(off)
    DD t{b, 0.0};
    dd_add(t, -d.hi);
    return t.hi + (t.lo - d.lo);
}

__global__ __launch_bounds__(256) void k_residual(const c64* W, long ld, int n, const c64* x,
                                                   const c64* b, c64* r,
                                                   unsigned long long* norms, int want_wnorm,
                                                   long sw, long sn, const int* active)
{
    const long z_ = blockIdx.z;
    if (active && !active[z_]) return;
    W += z_ * sw;
    x += z_ * n;
    b += z_ * n;
    r += z_ * n;
    norms += z_ * sn;
    const int row = blockIdx.x;
    const c64* Wr = W + (long)row * ld;
    DD sx{0, 0}, sy{0, 0};
    double an = 0;
    for (int c = threadIdx.x; c < n; c += 256) {
        const c64 a = Wr[c], v = x[c];
        dd_mac(sx, a.x, v.x);
        dd_mac(sx, -a.y, v.y);
        dd_mac(sy, a.x, v.y);
        dd_mac(sy, a.y, v.x);
        if (want_wnorm) an += sqrt(a.x * a.x + a.y * a.y);
    }
    __shared__ DD red[2][4];
    __shared__ double reda[4];
    for (int off = 32; off > 0; off >>= 1) {
        sx = dd_plus(sx, dd_shfl_xor(sx, off));
        sy = dd_plus(sy, dd_shfl_xor(sy, off));
        an += __shfl_xor(an, off);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = sx;
        red[1][wv] = sy;
        reda[wv] = an;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD tx{0, 0}, ty{0, 0};
        double ta = 0;
        for (int i = 0; i < 4; i++) {
            tx = dd_plus(tx, red[0][i]);
            ty = dd_plus(ty, red[1][i]);
            ta += reda[i];
        }
        const c64 rr = c64{dd_sub_from(b[row].x, tx), dd_sub_from(b[row].y, ty)};
        r[row] = rr;
        const double ra = sqrt(rr.x * rr.x + rr.y * rr.y);
        const double xa = sqrt(x[row].x * x[row].x + x[row].y * x[row].y);
        atomicMax(&norms[0], (unsigned long long)__double_as_longlong(ra));
        atomicMax(&norms[1], (unsigned long long)__double_as_longlong(xa));
        if (want_wnorm) atomicMax(&norms[2], (unsigned long long)__double_as_longlong(ta));
    }
}

// Products with a W of HEM structure (HemOp). Work-groups [0, 2m) take one
// impedance row each (the Z_L block's rows, then Z_T's): the row's m entries
// of the block from W, strided over the threads and summed in the order of
// k_residual, then its one or two incidence entries. The remaining
// work-groups take 256 node rows each, one thread per row, summing over the
// node's segment ends in list order. mode 0: r = W x; 1: r = b - W x with
// |r| and |x| max-reductions into norms[0], norms[1]; 2: also |W| (row sums
// of |W_ij|) into norms[2]. The residual modes sum in double-double.
__global__ __launch_bounds__(256) void k_hem_apply(const c64* W, long ld, HemOp op, const c64* x,
                                                    long sx, const c64* b, c64* r, long sr,
                                                    unsigned long long* norms, long sn, int mode,
                                                    long sw, const int* active)
{
    const long z_ = blockIdx.z;
    if (active && !active[z_]) return;
    W += z_ * sw;
    x += z_ * sx;
    if (b) b += z_ * sr;
    r += z_ * sr;
    if (norms) norms += z_ * sn;
    const bool want_w = mode == 2;
    // residuals (modes 1 and 2) in double-double; plain products (mode 0) in FP64
    const bool comp = mode != 0;
    const long m = op.m;
    if ((long)blockIdx.x < 2 * m) {
        const bool tblk = (long)blockIdx.x >= m;
        const long i = tblk ? (long)blockIdx.x - m : (long)blockIdx.x;
        const long off = tblk ? op.zt_off : op.zl_off;
        const long row = off + i;
        const c64* Wr = W + row * ld + off;
        const c64* xz = x + off;
        DD sx_{0, 0}, sy_{0, 0};
        double an = 0;
        // unrolled so that several loads per lane are in flight
        if (comp) {
#pragma unroll 4
            for (long c = threadIdx.x; c < m; c += 256) {
                const c64 a = Wr[c], v = xz[c];
                dd_mac(sx_, a.x, v.x);
                dd_mac(sx_, -a.y, v.y);
                dd_mac(sy_, a.x, v.y);
                dd_mac(sy_, a.y, v.x);
                if (want_w) an += sqrt(a.x * a.x + a.y * a.y);
            }
        } else {
#pragma unroll 4
            for (long c = threadIdx.x; c < m; c += 256) {
                const c64 a = Wr[c], v = xz[c];
                sx_.hi += a.x * v.x - a.y * v.y;
                sy_.hi += a.x * v.y + a.y * v.x;
            }
        }
        __shared__ DD red[2][4];
        __shared__ double reda[4];
        for (int o = 32; o > 0; o >>= 1) {
            sx_ = dd_plus(sx_, dd_shfl_xor(sx_, o));
            sy_ = dd_plus(sy_, dd_shfl_xor(sy_, o));
            an += __shfl_xor(an, o);
        }
        const int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            red[0][wv] = sx_;
            red[1][wv] = sy_;
            reda[wv] = an;
        }
        __syncthreads();
        if (threadIdx.x != 0) return;
        DD tx{0, 0}, ty{0, 0};
        double ta = 0;
        for (int q = 0; q < 4; q++) {
            tx = dd_plus(tx, red[0][q]);
            ty = dd_plus(ty, red[1][q]);
            ta += reda[q];
        }
        // incidence: A (+1 start, -1 end) on Z_L rows, B (1/2, 1/2) on Z_T rows
        // (products by +-1 and 1/2 are exact)
        const int s = op.sn[i], e = op.en[i];
        const double cs = tblk ? 0.5 : 1.0, ce = tblk ? 0.5 : -1.0;
        const c64 xs = x[op.node_off + s];
        dd_add(tx, cs * xs.x);
        dd_add(ty, cs * xs.y);
        ta += fabs(cs);
        if (e != s) {
            const c64 xe = x[op.node_off + e];
            dd_add(tx, ce * xe.x);
            dd_add(ty, ce * xe.y);
            ta += fabs(ce);
        }
        if (mode == 0) {
            r[row] = c64{tx.hi + tx.lo, ty.hi + ty.lo};
            return;
        }
        const c64 rr = c64{dd_sub_from(b[row].x, tx), dd_sub_from(b[row].y, ty)};
        r[row] = rr;
        const double ra = sqrt(rr.x * rr.x + rr.y * rr.y);
        const double xa = sqrt(x[row].x * x[row].x + x[row].y * x[row].y);
        atomicMax(&norms[0], (unsigned long long)__double_as_longlong(ra));
        atomicMax(&norms[1], (unsigned long long)__double_as_longlong(xa));
        if (want_w) atomicMax(&norms[2], (unsigned long long)__double_as_longlong(ta));
        return;
    }
    // node rows: sum over the node's segment ends of A^T (x_L) and B^T (x_T)
    const long u = ((long)blockIdx.x - 2 * m) * 256 + threadIdx.x;
    double ra = 0, xa = 0, ta = 0;
    if (u < op.n) {
        DD tx{0, 0}, ty{0, 0};
        const int p1 = op.node_ptr[u + 1];
        for (int p = op.node_ptr[u]; p < p1; p++) {
            const int se = op.node_seg[p];
            const long s = se >> 1;
            const double cl = (se & 1) ? -1.0 : 1.0;
            const c64 xl = x[op.zl_off + s], xt = x[op.zt_off + s];
            dd_add(tx, cl * xl.x);
            dd_add(tx, 0.5 * xt.x);
            dd_add(ty, cl * xl.y);
            dd_add(ty, 0.5 * xt.y);
            ta += 1.5;
        }
        const long row = op.node_off + u;
        if (mode == 0) {
            r[row] = c64{tx.hi + tx.lo, ty.hi + ty.lo};
        } else {
            const c64 rr = c64{dd_sub_from(b[row].x, tx), dd_sub_from(b[row].y, ty)};
            r[row] = rr;
            ra = sqrt(rr.x * rr.x + rr.y * rr.y);
            xa = sqrt(x[row].x * x[row].x + x[row].y * x[row].y);
        }
    }
    if (mode == 0) return;
    for (int o = 32; o > 0; o >>= 1) {
        ra = fmax(ra, __shfl_xor(ra, o));
        xa = fmax(xa, __shfl_xor(xa, o));
        ta = fmax(ta, __shfl_xor(ta, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&norms[0], (unsigned long long)__double_as_longlong(ra));
        atomicMax(&norms[1], (unsigned long long)__double_as_longlong(xa));
        if (want_w) atomicMax(&norms[2], (unsigned long long)__double_as_longlong(ta));
    }
}

// row maxima of |W_ij| (for symmetric equilibration): out[i] = 1/sqrt(max_j |W_ij|),
// or 1 when equilibration is off
__global__ __launch_bounds__(256) void k_row_scale(const c64* W, long ld, int n, double* d,
                                                    int enabled, long sw)
{
    const long z_ = blockIdx.z;
    W += z_ * sw;
    d += z_ * n;
    const int row = blockIdx.x;
    if (!enabled) {
        if (threadIdx.x == 0) d[row] = 1.0;
        return;
    }
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
                                                        const double* d, c32* W32, long ld32,
                                                        long sw, long sw32)
{
    // four loads per thread go out before any store (HBM-bound: bytes in flight)
    constexpr int U = 4;
    const long z_ = blockIdx.z;
    W += z_ * sw;
    d += z_ * n;
    W32 += z_ * sw32;
    const int row = blockIdx.y;
    const double di = d[row];
    const c64* src = W + (long)row * ld;
    c32* dst = W32 + (long)row * ld32;
    for (int c0 = blockIdx.x * 256 * U + threadIdx.x; c0 < n; c0 += gridDim.x * 256 * U) {
        c64 a[U];
        double sc[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c0 + 256 * u;
            a[u] = c < n ? src[c] : c64{0.0, 0.0};
            sc[u] = c < n ? di * d[c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c = c0 + 256 * u;
            if (c < n) dst[c] = c32{(float)(a[u].x * sc[u]), (float)(a[u].y * sc[u])};
        }
    }
}

// ---------------------------------------------------------------------------
// Grounded node unknowns of the mixed form. The node voltages U (physical rows
// and columns [nb, ne)) are replaced by U_g and V_i = U_i - U_g (i != g):
// W' = T^T W T, T = I plus ones in column g on the node rows, x = T y. Only row
// and column g change: column g of W' is the sum of W's node columns, row g the
// sum of its node rows, and W'[g][g] the sum of its node block. For the
// immittance matrix the incidence rows sum to 0 (A) and 1 (B), so the
// Laplacian-like part of the node block no longer holds the constant vector
// that a low-precision factor cannot resolve at low frequencies.
// ---------------------------------------------------------------------------
__device__ inline c64 block_sum256(c64 s, c64* red)
{
    for (int off = 32; off > 0; off >>= 1) {
        s.x += __shfl_xor(s.x, off);
        s.y += __shfl_xor(s.y, off);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    return cadd(cadd(red[0], red[1]), cadd(red[2], red[3]));
}

// gcol[i] = sum_{c in [nb, ne)} W[i][c], one work-group per row
__global__ __launch_bounds__(256) void k_ground_rowsums(const c64* W, long ld, int nb, int ne,
                                                         c64* gcol)
{
    __shared__ c64 red[4];
    const long row = blockIdx.x;
    c64 s{0.0, 0.0};
    for (int c = nb + (int)threadIdx.x; c < ne; c += 256) s = cadd(s, W[row * ld + c]);
    s = block_sum256(s, red);
    if (threadIdx.x == 0) gcol[row] = s;
}

// part[k][j] = sum of W[r][j] over the k-th of `nch` chunks of the node rows
__global__ __launch_bounds__(256) void k_ground_colsums(const c64* W, long ld, int n, int nb,
                                                         int ne, int nch, c64* part)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int len = (ne - nb + nch - 1) / nch;
    const int r0 = nb + (int)blockIdx.y * len, r1 = min(ne, r0 + len);
    c64 s{0.0, 0.0};
    for (int r = r0; r < r1; r++) s = cadd(s, W[(long)r * ld + j]);
    part[(long)blockIdx.y * n + j] = s;
}

// grow[j] = sum over the chunks (fixed order); gg = sum_{c in [nb, ne)} grow[c]
__global__ __launch_bounds__(256) void k_ground_finish(const c64* part, int n, int nb, int ne,
                                                        int nch, c64* grow, c64* gg)
{
    __shared__ c64 red[4];
    c64 acc{0.0, 0.0};
    for (int j = threadIdx.x; j < n; j += 256) {
        c64 s{0.0, 0.0};
        for (int k = 0; k < nch; k++) s = cadd(s, part[(long)k * n + j]);
        grow[j] = s;
        if (j >= nb && j < ne) acc = cadd(acc, s);
    }
    acc = block_sum256(acc, red);
    if (threadIdx.x == 0) *gg = acc;
}

// W32 = T^T W T (unscaled)
__global__ __launch_bounds__(256) void k_to_c32_grounded(const c64* W, long ld, int n, int g,
                                                          const c64* gcol, const c64* grow,
                                                          const c64* gg, c32* W32, long ld32)
{
    const int row = blockIdx.y;
    const c64* src = row == g ? grow : W + (long)row * ld;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) {
        const c64 a = c == g ? (row == g ? *gg : gcol[row]) : src[c];
        W32[(long)row * ld32 + c] = c32{(float)a.x, (float)a.y};
    }
}

// *out = sum_{i in [nb, ne)} r[i] (one work-group)
__global__ __launch_bounds__(256) void k_ground_rhs(const c64* r, int nb, int ne, c64* out)
{
    __shared__ c64 red[4];
    c64 s{0.0, 0.0};
    for (int i = nb + (int)threadIdx.x; i < ne; i += 256) s = cadd(s, r[i]);
    s = block_sum256(s, red);
    if (threadIdx.x == 0) *out = s;
}

// A scale factor given on the host, or as max |r| in device memory (the bit
// pattern the residual kernel leaves), read by every thread.
__device__ inline double rnorm_of(const unsigned long long* rn)
{
    return __longlong_as_double((long long)*rn);
}

// t[i] = (float) (d[p] * r'[p] / |r|), p = perm[i] (|r| = 1 / inv_scale, or
// taken from rn when rn is given; a zero |r| gives t = 0); with grounding
// (g >= 0) r'[g] = *rg (the sum of the node entries) and r'[p] = r[p] otherwise
__global__ void k_prep_rhs32(const c64* r, const double* d, const int* perm, int n,
                             double inv_scale, const unsigned long long* rn, int g,
                             const c64* rg, c32* t, long sn, const int* active)
{
    const long z_ = blockIdx.z;
    if (active && !active[z_]) return;
    r += z_ * n;
    d += z_ * n;
    perm += z_ * n;
    t += z_ * n;
    if (rn) rn += z_ * sn;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (rn) {
        const double v = rnorm_of(rn);
        inv_scale = v > 0.0 ? 1.0 / v : 0.0;
    }
    const int p = perm[i];
    const double s = d[p] * inv_scale;
    const c64 v = p == g ? *rg : r[p];
    t[i] = c32{(float)(v.x * s), (float)(v.y * s)};
}

// x[j] += scale * d[j] * z'[j] (scale = |r| from rn when given); with grounding
// (g >= 0) z'[j] = z[j] + z[g] on the node range except g itself (x = T y).
// dxn (optional, zeroed by the caller): max_j |scale d[j] z'[j]|, the size of
// the correction, as the bit pattern of a non-negative double.
__global__ __launch_bounds__(256) void k_update_x(c64* x, const c32* z, const double* d, int n,
                                                   double scale, const unsigned long long* rn,
                                                   int g, int nb, int ne,
                                                   unsigned long long* dxn, long sn,
                                                   const int* active)
{
    const long z_ = blockIdx.z;
    if (active && !active[z_]) return;
    x += z_ * n;
    z += z_ * n;
    d += z_ * n;
    if (rn) rn += z_ * sn;
    if (dxn) dxn += z_ * sn;
    const int j = blockIdx.x * 256 + threadIdx.x;
    double a = 0.0;
    if (j < n) {
        if (rn) scale = rnorm_of(rn);
        const double s = scale * d[j];
        double zx = (double)z[j].x, zy = (double)z[j].y;
        if (g >= 0 && j != g && j >= nb && j < ne) {
            zx += (double)z[g].x;
            zy += (double)z[g].y;
        }
        const double dx = s * zx, dy = s * zy;
        x[j] = c64{x[j].x + dx, x[j].y + dy};
        a = sqrt(dx * dx + dy * dy);
    }
    if (!dxn) return;
    for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off));
    if ((threadIdx.x & 63) == 0) atomicMax(dxn, (unsigned long long)__double_as_longlong(a));
}

// t[i] = b[perm[i]]
template <typename T>
__global__ void k_gather_vec(const cx<T>* b, const int* perm, int n, cx<T>* t)
{
    b += (long)blockIdx.z * n;
    perm += (long)blockIdx.z * n;
    t += (long)blockIdx.z * n;
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
int LUWork<T>::alloc(int n_, int max_panel_wgs, size_t panel_lds_limit, int batch_)
{
    num_cu = max_panel_wgs;
    max_wgs = max_panel_wgs;
    lds_limit = panel_lds_limit;
    if (n_ <= 0 || batch_ <= 0) return TAGS_ERR_ARG;
    if (n_ >= 65536 || batch_ > 65535) return TAGS_ERR_TOO_LARGE;
    if (n == n_ && batch >= batch_) return 0;
    // element counts for (n_, batch_); arrays are kept while they are large
    // enough (a work space serves systems of different orders in turn)
    const int nblk_ = (n_ + 63) / 64 + 8;
    TsluTree tr;
    int rc = tslu_tree(n_, tslu_min_leaf_rows(), &tr);
    if (rc) return rc;
    const size_t B = (size_t)batch_;
    size_t need_perm = (size_t)n_ * B, need_inv = 4096 * (size_t)nblk_ * B;
    // split-bf16 packs of every product of the factorisation (M, N <= n_,
    // K <= 256), rounded to 256 bytes per item
    const long x6_item_ = sizeof(T) == sizeof(float)
                              ? (gemm_x6_ws_bytes(n_, n_, 256) + 255) / 256 * 256
                              : 0;
    size_t need_x6 = (size_t)x6_item_ * B;
    size_t need_cand = (size_t)tr.cand_total * B;
    size_t need_cnt = (size_t)std::max(1, tr.cnt_total) * B;
    // flow-solve control words per item: ticket, spare, one per block group,
    // padded to 64 words so that items do not share cache lines
    const long tstride = ((long)nblk_ + 2 + 63) / 64 * 64;
    size_t need_tctl = (size_t)tstride * B;
    if (!s2) {
        TAGS_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        TAGS_HIP(hipEventCreateWithFlags(&ev_left, hipEventDisableTiming));
        TAGS_HIP(hipEventCreateWithFlags(&ev_panel, hipEventDisableTiming));
        TAGS_HIP(hipMalloc(&slots, sizeof(PanelSlot<T>) * 2 * 512));
        TAGS_HIP(hipMalloc(&bar, sizeof(unsigned) * 64));
        TAGS_HIP(hipMemset(bar, 0, sizeof(unsigned) * 64));
    }
    if (need_perm > cap_perm || need_inv > cap_inv || B > cap_items || need_cand > cap_cand ||
        need_cnt > cap_cnt || need_tctl > cap_tctl || need_x6 > cap_x6) {
        // grow to cover every earlier request as well: a work space that
        // serves systems of different orders in turn then reallocates (which
        // waits for the whole device) only while its largest request grows
        need_perm = std::max(need_perm, cap_perm);
        need_inv = std::max(need_inv, cap_inv);
        need_cand = std::max(need_cand, cap_cand);
        need_cnt = std::max(need_cnt, cap_cnt);
        need_tctl = std::max(need_tctl, cap_tctl);
        need_x6 = std::max(need_x6, cap_x6);
        const size_t items = std::max(B, cap_items);
        (void)hipStreamSynchronize(s2);
        (void)hipFree(perm);
        (void)hipFree(perm_alt);
        (void)hipFree(invL);
        (void)hipFree(invU);
        (void)hipFree(flags);
        (void)hipFree(vec);
        (void)hipFree(d_init);
        (void)hipFree(d_orders);
        (void)hipFree(tcand);
        (void)hipFree(tcnt);
        (void)hipFree(tctl);
        (void)hipFree(x6ws);
        (void)hipFree(x6ws2);
        x6ws2 = nullptr;
        perm = perm_alt = flags = d_init = d_orders = tcand = nullptr;
        invL = invU = vec = nullptr;
        tcnt = tctl = nullptr;
        x6ws = nullptr;
        cap_perm = cap_inv = cap_items = cap_cand = cap_cnt = cap_tctl = cap_x6 = 0;
        TAGS_HIP(hipMalloc(&perm, sizeof(int) * need_perm));
        TAGS_HIP(hipMalloc(&perm_alt, sizeof(int) * need_perm));
        TAGS_HIP(hipMalloc(&invL, sizeof(cx<T>) * need_inv));
        TAGS_HIP(hipMalloc(&invU, sizeof(cx<T>) * need_inv));
        TAGS_HIP(hipMalloc(&flags, sizeof(int) * 4 * items));
        TAGS_HIP(hipMalloc(&vec, sizeof(cx<T>) * need_perm));
        TAGS_HIP(hipMalloc(&d_init, sizeof(int) * need_perm));
        TAGS_HIP(hipMalloc(&d_orders, sizeof(int) * need_perm));
        TAGS_HIP(hipMemset(flags, 0, sizeof(int) * 4 * items));
        TAGS_HIP(hipMalloc(&tcand, sizeof(int) * need_cand));
        TAGS_HIP(hipMalloc(&tcnt, sizeof(unsigned) * need_cnt));
        // arrival counters start at zero; every tournament's root resets the
        // ones it used, so they stay zero between launches whatever the order
        TAGS_HIP(hipMemset(tcnt, 0, sizeof(unsigned) * need_cnt));
        // tickets start at zero (each launch's last ticket resets its counter);
        // completion words start below every epoch a launch uses
        TAGS_HIP(hipMalloc(&tctl, sizeof(unsigned) * need_tctl));
        TAGS_HIP(hipMemset(tctl, 0, sizeof(unsigned) * need_tctl));
        if (need_x6) {
            TAGS_HIP(hipMalloc(&x6ws, need_x6));
            TAGS_HIP(hipMalloc(&x6ws2, need_x6));
        }
        epoch = 0;
        cap_perm = need_perm;
        cap_inv = need_inv;
        cap_items = items;
        cap_cand = need_cand;
        cap_cnt = need_cnt;
        cap_tctl = need_tctl;
        cap_x6 = need_x6;
        tctl_stride = tstride;
    }
    if (tstride != tctl_stride) {
        // the items' control words move: their tickets must read zero again
        TAGS_HIP(hipMemset(tctl, 0, sizeof(unsigned) * cap_tctl));
        epoch = 0;
        tctl_stride = tstride;
    }
    n = n_;
    batch = batch_;
    nb = 1;
    sA = 0;
    nblk_cap = nblk_;
    x6_item = x6_item_;
    cand_stride = tr.cand_total;
    cnt_stride = std::max(1, tr.cnt_total);
    h_init.clear();
    h_orders.clear();
    blk.clear();
    nblk = 0;
    return 0;
}

template <typename T>
void LUWork<T>::release()
{
    if (s2) (void)hipStreamSynchronize(s2);
    (void)hipFree(perm);
    (void)hipFree(perm_alt);
    if (s2) (void)hipStreamDestroy(s2);
    if (ev_left) (void)hipEventDestroy(ev_left);
    if (ev_panel) (void)hipEventDestroy(ev_panel);
    perm_alt = nullptr;
    s2 = nullptr;
    ev_left = ev_panel = nullptr;
    (void)hipFree(invL);
    (void)hipFree(invU);
    (void)hipFree(slots);
    (void)hipFree(bar);
    (void)hipFree(flags);
    (void)hipFree(vec);
    (void)hipFree(d_init);
    (void)hipFree(d_orders);
    (void)hipFree(tcand);
    (void)hipFree(tcnt);
    (void)hipFree(tctl);
    (void)hipFree(x6ws);
    (void)hipFree(x6ws2);
    x6ws2 = nullptr;
    tcand = nullptr;
    tcnt = nullptr;
    tctl = nullptr;
    x6ws = nullptr;
    cap_tctl = cap_x6 = 0;
    x6_item = 0;
    d_init = nullptr;
    d_orders = nullptr;
    h_init.clear();
    h_orders.clear();
    perm = nullptr;
    invL = invU = vec = nullptr;
    slots = nullptr;
    bar = nullptr;
    flags = nullptr;
    blk.clear();
    n = nblk = 0;
    batch = nb = 1;
    sA = 0;
    cap_perm = cap_inv = cap_items = cap_cand = cap_cnt = 0;
}

// One product per item of the factorisation's batch: the matrix operands at the
// batch's matrix stride and row maps at its row-map stride; the operand named
// by inv_op (1: A, 2: B) is a block inverse at the inverse stride instead.
// *flops counts one item's flops, the profiler the launch's.
template <typename T>
static int gemm(Profiler* pr, hipStream_t st, int M, int N, int K, const cx<T>* A, long lda,
                const int* arow, const cx<T>* B, long ldb, const int* brow, cx<T>* C, long ldc,
                const int* crow, int mode, bool trailing, double* flops, const BatchStrides& bs,
                int inv_op = 0)
{
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    GemmArgs<T> g{M, N, K, A, lda, arow, B, ldb, brow, C, ldc, crow, mode};
    g.nb = bs.nb;
    g.ws = bs.ws;
    g.ws_item = bs.ws_item;
    g.sa = inv_op == 1 ? bs.inv : bs.a;
    g.sb = inv_op == 2 ? bs.inv : bs.a;
    g.sc = bs.a;
    g.sar = g.sbr = g.scr = bs.p;
    Timed tm(pr, FAM_GEMM, st);
    int rc = launch_gemm<T>(g, st, trailing);
    const double f = 8.0 * M * N * (double)K;
    tm.end(f * bs.nb, trailing);
    if (flops) *flops += f;
    return rc;
}

// rows of the panel held by one work-group (bounded by the LDS it may use)
template <typename T>
static int panel_rows_max(size_t lds_limit)
{
    const size_t per_row = 64 * sizeof(cx<T>) + 2 * sizeof(int);
    return (int)((lds_limit - 4096) / per_row);
}

static std::atomic<int> g_panel_kind{PANEL_TOURNAMENT};

void lu_set_panel_kind(int kind) { g_panel_kind.store(kind == PANEL_PARTIAL ? 1 : 0); }
int lu_panel_kind() { return g_panel_kind.load(); }

static std::atomic<bool> g_lookahead{true};
void lu_set_lookahead(bool on) { g_lookahead.store(on); }

// two launches per step measured fastest on MI355X for the 60 x 60 m grid
static std::atomic<int> g_update_chunks{4};
void lu_set_update_chunks(int n) { g_update_chunks.store(n < 1 ? 1 : n); }

// Columns per step (four sub-panels of 64): the trailing update's K. 256 runs
// the complex64 product ~26 % faster than 128 (the A22 read-modify-write is
// shared by twice the flops).
// By default complex128 factors take 128-column steps (measured faster for
// the FP64 form on the 60 x 60 m grid) and complex64 factors 256.
static std::atomic<int> g_step_width[2] = {{128}, {256}};
void lu_set_step_width(int w)
{
    g_step_width[0].store(w <= 0 ? 128 : w <= 128 ? 128 : 256);
    g_step_width[1].store(w <= 0 ? 256 : w <= 128 ? 128 : 256);
}
int lu_step_width(bool c64) { return g_step_width[c64 ? 1 : 0].load(); }

static thread_local const Tuning* t_tuning = nullptr;
const Tuning* tuning_current() { return t_tuning; }
TuningScope::TuningScope(const Tuning* t) : prev(t_tuning) { t_tuning = t; }
TuningScope::~TuningScope() { t_tuning = prev; }

// the step width and update launches of the calling thread's factorisations
static int step_width_now(bool c64)
{
    const Tuning* t = t_tuning;
    const int p = c64 ? 1 : 0;
    if (t && t->step_width[p] > 0) return t->step_width[p] <= 128 ? 128 : 256;
    return g_step_width[p].load(std::memory_order_relaxed);
}
static int update_chunks_now()
{
    const Tuning* t = t_tuning;
    if (t && t->update_chunks > 0) return t->update_chunks;
    return std::max(1, g_update_chunks.load(std::memory_order_relaxed));
}

template <typename T>
static int panel(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld, int* pm,
                 int c0, int cols, int row_end, int cand_end)
{
    const int M = row_end - c0;
    if (g_panel_kind.load(std::memory_order_relaxed) == PANEL_TOURNAMENT) {
        Timed tm(pr, FAM_PANEL, st);
        const int rc = launch_tslu<T>(A, ld, pm, c0, cols, row_end, cand_end, w.tcand,
                                      w.tcnt, w.flags, st, w.bs());
        tm.end();
        w.flops += 8.0 * (double)M * cols * cols / 2;
        return rc;
    }
    // the cooperative panel holds the whole device: one item at a time only
    if (w.nb != 1) return TAGS_ERR_ARG;
    const size_t per_row = 64 * sizeof(cx<T>) + 2 * sizeof(int);
    const int rmax = panel_rows_max<T>(w.lds_limit);
    // few work-groups keep the grid barrier cheap; ~160 rows each
    int R = std::min(rmax, 160);
    int P = (M + R - 1) / R;
    if (P < 1) P = 1;
    R = (M + P - 1) / P;
    {
        // every work-group of a cooperative launch must be resident at once
        int occ = 0;
        TAGS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_panel<T>, 256,
                                                              (size_t)R * per_row));
        if (P > occ * w.num_cu || P > 511) return TAGS_ERR_TOO_LARGE;
    }
    PanelArgs<T> p{A, ld, pm, c0, M, c0, cols, R, cand_end, w.slots, w.bar, w.flags,
                   w.flags + 1};
    const size_t lds = (size_t)R * per_row;
    void* args[] = {&p};
    Timed tm(pr, FAM_PANEL, st);
    TAGS_HIP(hipLaunchCooperativeKernel((const void*)k_panel<T>, dim3(P), dim3(256), args,
                                        (unsigned)lds, st));
    tm.end();
    w.flops += 8.0 * (double)M * cols * cols / 2;
    return 0;
}

template <typename T>
int LUWork<T>::setup_kernel_attrs()
{
    // raise the dynamic-LDS cap of the panel kernel once per process; keep
    // the default cap (and a smaller panel chunk) when the runtime refuses
    static std::mutex mu;
    static int state = 0;  // 0 not tried, 1 raised, 2 refused
    std::lock_guard<std::mutex> lock(mu);
    if (state == 0) {
        if (hipFuncSetAttribute((const void*)k_panel<T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds_limit) != hipSuccess) {
            (void)hipGetLastError();
            state = 2;
        } else {
            state = 1;
        }
    }
    if (state == 2) lds_limit = 65536;
    return 0;
}

// U12 for logical rows [r0, r0+h) of the column range [cb, ce): rows are
// overwritten with inv(L_block) * rows (block inverse `inv`, h <= 64), as one
// product written in place: M = K = h <= 64 is a single row tile of every GEMM
// variant, so the work-group of each column tile reads those rows of its
// columns in full before it writes them, and no other work-group touches them.
template <typename T>
static int apply_linv_rows(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld,
                           const int* pm, const cx<T>* inv, int r0, int h, int cb, int ce)
{
    const int cols = ce - cb;
    if (cols <= 0) return 0;
    if (h > 64) return TAGS_ERR_ARG;
    return gemm<T>(pr, st, h, cols, h, inv, 64, nullptr, A + cb, ld, pm + r0, A + cb, ld, pm + r0,
                   0, false, &w.flops, w.bs(), 1);
}

// One step of the factorisation (up to the step width in columns).
struct LUStep {
    int j, jb;    // first column, width (<= the step width)
    int phase;    // index into the plan's phases
    int slot;     // index of its first diagonal block in the block list
};

// L21 = A21 inv(U11) for logical rows [c0 + cols, row_end), one product written
// in place: N = cols <= 64 is a single column tile of every GEMM variant, so
// each work-group reads its rows in full before it writes them, and no other
// work-group touches them.
template <typename T>
static int l21_gemm(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld, const int* pm,
                    int c0, int cols, int row_end, const cx<T>* invu)
{
    const int rb = c0 + cols;
    if (row_end <= rb) return 0;
    return gemm<T>(pr, st, row_end - rb, cols, cols, A + c0, ld, pm + rb, invu, 64, nullptr,
                   A + c0, ld, pm + rb, 0, false, &w.flops, w.bs(), 2);
}

// One sub-panel of `cols` columns at c0 whose diagonal block is block `slot`:
//   without row exchanges  k_diag_factor (LU of the diagonal block and the
//                          inverses of its triangles), L21 product;
//   tournament             k_tslu (pivots, L11 \ U11), the inverses, L21 product;
//   partial                cooperative k_panel (pivots, L11 \ U11 and L21), the
//                          inverses.
template <typename T>
static int sub_panel(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld, int* pm,
                     const LUPhase& ph, int c0, int cols, int slot)
{
    cx<T>* invl = w.invL + (size_t)slot * 4096;
    cx<T>* invu = w.invU + (size_t)slot * 4096;
    const bool tournament =
        g_panel_kind.load(std::memory_order_relaxed) == PANEL_TOURNAMENT;
    int rc;
    if (ph.pivot && (rc = panel<T>(w, pr, st, A, ld, pm, c0, cols, ph.row_end, ph.cand_end)))
        return rc;
    {
        Timed tm(pr, ph.pivot ? FAM_OTHER : FAM_PANEL, st);
        const BatchStrides b = w.bs();
        if (ph.pivot)
            hipLaunchKernelGGL((k_diag_block<T>), dim3(2, 1, b.nb), dim3(256), 0, st, A, ld, pm,
                               c0, cols, invl, invu, b.a, b.p, b.inv);
        else
            hipLaunchKernelGGL((k_diag_factor<T>), dim3(1, 1, b.nb), dim3(512), 0, st, A, ld, pm,
                               c0, cols, invl, invu, w.flags, b.a, b.p, b.inv, b.f);
        TAGS_HIP(hipGetLastError());
        tm.end();
    }
    if (!ph.pivot) w.flops += 8.0 * (double)cols * cols * cols / 3;
    if (!ph.pivot || tournament) return l21_gemm<T>(w, pr, st, A, ld, pm, c0, cols, ph.row_end, invu);
    return 0;
}

// Panel work of a step on stream st with row map pm: its sub-panels of <= 64
// columns in order, each followed (right-looking within the step) by the
// update of the step's later columns: U_t,rest = inv(L_tt) A_t,rest and
// A_below,rest -= L_t U_t,rest.
template <typename T>
static int step_panels(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld, int* pm,
                       const LUPhase& ph, const LUStep& s)
{
    const int end = s.j + s.jb;
    int rc;
    for (int t = 0, c0 = s.j; c0 < end; t++, c0 += 64) {
        const int wt = std::min(64, end - c0), c1 = c0 + wt;
        if ((rc = sub_panel<T>(w, pr, st, A, ld, pm, ph, c0, wt, s.slot + t))) return rc;
        if (c1 >= end) break;
        if ((rc = apply_linv_rows<T>(w, pr, st, A, ld, pm, w.invL + (size_t)(s.slot + t) * 4096,
                                     c0, wt, c1, end)))
            return rc;
        if ((rc = gemm<T>(pr, st, ph.row_end - c1, end - c1, wt, A + c0, ld, pm + c1, A + c1, ld,
                          pm + c0, A + c1, ld, pm + c1, 1, false, &w.flops, w.bs())))
            return rc;
    }
    return 0;
}

// U12 of a step's pivot rows for the columns [cb, ce), in one launch of
// k_step_u12 (the blocked forward solve over the step's sub-panels).
template <typename T>
static int step_u12(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld,
                    const int* pm, const LUStep& s, int cb, int ce)
{
    constexpr int CW = sizeof(T) == 4 ? 32 : 16, KC = sizeof(T) == 4 ? 16 : 8;
    const int cols = ce - cb;
    if (cols <= 0) return 0;
    if (s.jb > 256) return TAGS_ERR_ARG;
    double flops = 0;
    for (int r0 = 0; r0 < s.jb; r0 += 64) {
        const int wt = std::min(64, s.jb - r0);
        flops += 8.0 * wt * (double)cols * (r0 + wt);
    }
    Timed tm(pr, FAM_GEMM, st);
    const BatchStrides b = w.bs();
    if constexpr (sizeof(T) == 4)
        hipLaunchKernelGGL(k_step_u12_mfma<false>, dim3((cols + 31) / 32, 1, b.nb), dim3(256), 0,
                           st, A, ld, pm, s.j, s.jb, (const c32*)(w.invL + (size_t)s.slot * 4096),
                           cb, cols, b.a, b.p, b.inv);
    else
        hipLaunchKernelGGL((k_step_u12<T, CW, KC, false>), dim3((cols + CW - 1) / CW, 1, b.nb),
                           dim3(256), 0, st, A, ld, pm, s.j, s.jb,
                           (const cx<T>*)(w.invL + (size_t)s.slot * 4096), cb, cols, b.a, b.p,
                           b.inv);
    TAGS_HIP(hipGetLastError());
    tm.end(flops * b.nb);
    w.flops += flops;
    return 0;
}

// Update of a step for the trailing columns [cb, ce): U12 of the step's pivot
// rows (the blocked forward solve over its sub-panels: rows of sub-panel t
// take the earlier sub-panels' products, then inv(L_tt)), then
// A22 -= L21 U12 over logical rows [j + jb, row_end) with K = jb.
template <typename T>
static int step_update(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld,
                       const int* pm, const LUPhase& ph, const LUStep& s, int cb, int ce)
{
    if (ce <= cb) return 0;
    const int j = s.j, c1 = j + s.jb;
    int rc;
    if ((rc = step_u12<T>(w, pr, st, A, ld, pm, s, cb, ce))) return rc;
    return gemm<T>(pr, st, ph.row_end - c1, ce - cb, s.jb, A + j, ld, pm + c1, A + cb, ld, pm + j,
                   A + cb, ld, pm + c1, 1, true, &w.flops, w.bs());
}

// Backward blocked solve of the extra columns [x0, x0 + nx) in the logical rows
// of the steps of phase p (p < 0: every step) from c_lo on: those rows hold
// Y = L^-1 P C (the factorisation's trailing updates of the extra columns) and
// become X = U^-1 Y. Last step first: the step's own rows in k_step_u12, then
// the rows above it from c_lo by one product with K = the step width.
// Flops: 4 r^2 nx for r rows.
template <typename T>
static int backward_cols(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld,
                         const int* pm, const std::vector<LUStep>& steps, int p, int c_lo,
                         int x0, int nx, void* ws = nullptr)
{
    // the split-bf16 products' work space (another one for a second stream)
    BatchStrides bw = w.bs();
    if (ws) bw.ws = ws;
    constexpr int CW = sizeof(T) == 4 ? 32 : 16, KC = sizeof(T) == 4 ? 16 : 8;
    int rc;
    for (int k = (int)steps.size() - 1; k >= 0; k--) {
        const LUStep& s = steps[k];
        if ((p >= 0 && s.phase != p) || s.j < c_lo) continue;
        if (s.jb > 256) return TAGS_ERR_ARG;
        double flops = 0;
        for (int r0 = 0; r0 < s.jb; r0 += 64) {
            const int wt = std::min(64, s.jb - r0);
            flops += 8.0 * wt * (double)nx * (s.jb - r0);
        }
        {
            Timed tm(pr, FAM_GEMM, st);
            const BatchStrides b = w.bs();
            if constexpr (sizeof(T) == 4)
                hipLaunchKernelGGL(k_step_u12_mfma<true>, dim3((nx + 31) / 32, 1, b.nb),
                                   dim3(256), 0, st, A, ld, pm, s.j, s.jb,
                                   (const c32*)(w.invU + (size_t)s.slot * 4096), x0, nx, b.a,
                                   b.p, b.inv);
            else
                hipLaunchKernelGGL((k_step_u12<T, CW, KC, true>),
                                   dim3((nx + CW - 1) / CW, 1, b.nb), dim3(256), 0, st, A, ld, pm,
                                   s.j, s.jb, (const cx<T>*)(w.invU + (size_t)s.slot * 4096), x0,
                                   nx, b.a, b.p, b.inv);
            TAGS_HIP(hipGetLastError());
            tm.end(flops * b.nb);
            w.flops += flops;
        }
        if (s.j > c_lo &&
            (rc = gemm<T>(pr, st, s.j - c_lo, nx, s.jb, A + s.j, ld, pm + c_lo, A + x0, ld,
                          pm + s.j, A + x0, ld, pm + c_lo, 1, false, &w.flops, bw)))
            return rc;
    }
    return 0;
}

// Nodal plan, before its node phase: the node columns of the Z_L and Z_T rows
// hold Y = L^-1 P A (their phases' trailing updates); complete them to
// X = U^-1 Y (backward_cols over each phase's steps), then write the node
// block's Schur complement (k_nodal_schur). Flops: 4 m^2 n per current block.
template <typename T>
static int nodal_prepare(LUWork<T>& w, Profiler* pr, hipStream_t st, cx<T>* A, long ld,
                         const int* pm, const LUPlan& plan, const std::vector<LUStep>& steps)
{
    const int m = plan.nodal_m, nn = w.n - 2 * m, x0 = 2 * m;
    int rc;
    if (w.s2 && (w.x6ws2 || !w.x6ws)) {
        // the two phases' rows are independent: the Z_T phase's on the
        // look-ahead stream (idle here; its own split-bf16 work space) beside
        // the Z_L phase's
        TAGS_HIP(hipEventRecord(w.ev_left, st));
        TAGS_HIP(hipStreamWaitEvent(w.s2, w.ev_left, 0));
        if ((rc = backward_cols<T>(w, pr, w.s2, A, ld, pm, steps, 1, plan.phases[1].c_begin, x0,
                                   nn, w.x6ws2)))
            return rc;
        if ((rc = backward_cols<T>(w, pr, st, A, ld, pm, steps, 0, plan.phases[0].c_begin, x0,
                                   nn)))
            return rc;
        TAGS_HIP(hipEventRecord(w.ev_panel, w.s2));
        TAGS_HIP(hipStreamWaitEvent(st, w.ev_panel, 0));
    } else {
        for (int p = 0; p < 2; p++)
            if ((rc = backward_cols<T>(w, pr, st, A, ld, pm, steps, p, plan.phases[p].c_begin,
                                       x0, nn)))
                return rc;
    }
    Timed tm(pr, FAM_OTHER, st);
    hipLaunchKernelGGL(k_nodal_schur<T>, dim3(nn, 1, w.nb), dim3(256), 0, st, A, ld, pm, m, nn,
                       w.sA, (long)w.n);
    TAGS_HIP(hipGetLastError());
    tm.end();
    return 0;
}

template <typename T>
int lu_factor(LUWork<T>& w, cx<T>* A, long ld, hipStream_t st, Profiler* pr, const LUPlan* plan,
              int nb, long sA)
{
    const int n = w.n;
    if (nb < 1 || nb > w.batch || (nb > 1 && sA < (long)n * ld)) return TAGS_ERR_ARG;
    w.nb = nb;
    w.sA = nb > 1 ? sA : 0;
    int rc = w.setup_kernel_attrs();
    if (rc) return rc;
    LUPlan dense;
    if (!plan) {
        dense = LUPlan::dense(n);
        plan = &dense;
    }
    // Initial row order and cleared flags are set by an ordinary kernel on the
    // stream: the cooperative panel launches are ordered after kernels, and the
    // initial order is uploaded once with a blocking copy when it changes.
    const int* init = nullptr;
    const int* orders = w.d_orders;
    if (!plan->init_perm.empty() && plan->dev_init) {
        if ((int)plan->init_perm.size() != n) return TAGS_ERR_ARG;
        init = plan->dev_init;
    } else if (!plan->init_perm.empty()) {
        if ((int)plan->init_perm.size() != n) return TAGS_ERR_ARG;
        if (w.h_init != plan->init_perm) {
            TAGS_HIP(hipStreamSynchronize(st));
            TAGS_HIP(hipMemcpy(w.d_init, plan->init_perm.data(), sizeof(int) * n,
                               hipMemcpyHostToDevice));
            w.h_init = plan->init_perm;
        }
        init = w.d_init;
    }
    if (!plan->orders.empty() && plan->dev_orders) {
        if ((int)plan->orders.size() > n) return TAGS_ERR_ARG;
        orders = plan->dev_orders;
    } else if (!plan->orders.empty() && w.h_orders != plan->orders) {
        if ((int)plan->orders.size() > n) return TAGS_ERR_ARG;
        TAGS_HIP(hipStreamSynchronize(st));
        TAGS_HIP(hipMemcpy(w.d_orders, plan->orders.data(), sizeof(int) * plan->orders.size(),
                           hipMemcpyHostToDevice));
        w.h_orders = plan->orders;
    }
    for (const LUPhase& ph : plan->phases)
        if (ph.order_off >= 0 &&
            (ph.pivot || ph.order_off + (n - ph.c_begin) > (int)plan->orders.size()))
            return TAGS_ERR_ARG;
    if (plan->nodal_m > 0 &&
        (plan->phases.size() != 3 || 2 * plan->nodal_m >= n || plan->phases[2].c_begin !=
                                                                    2 * plan->nodal_m))
        return TAGS_ERR_ARG;
    w.nodal_m = plan->nodal_m;
    hipLaunchKernelGGL(k_lu_init, dim3((n + 255) / 256, 1, w.nb), dim3(256), 0, st, w.perm,
                       init, n, w.flags, (long)n, w.bs().f);
    TAGS_HIP(hipGetLastError());
    // a phase's row order, set at its first step on the row map that step uses
    auto set_order = [&](const LUStep& sp, int* pm, hipStream_t sq) -> int {
        const LUPhase& ph = plan->phases[sp.phase];
        if (ph.order_off < 0 || sp.j != ph.c_begin) return 0;
        const int len = n - ph.c_begin;
        hipLaunchKernelGGL(k_set_order, dim3((len + 255) / 256, 1, w.nb), dim3(256), 0, sq, pm,
                           ph.c_begin, len, orders + ph.order_off, (long)n);
        TAGS_HIP(hipGetLastError());
        return 0;
    };
    // flat list of steps over the phases
    std::vector<LUStep> steps;
    w.blk.clear();
    w.flops = 0;
    const int width = step_width_now(sizeof(T) == 4);
    for (int p = 0; p < (int)plan->phases.size(); p++) {
        const LUPhase& ph = plan->phases[p];
        for (int j = ph.c_begin; j < ph.c_end; j += width) {
            const int jb = std::min(width, ph.c_end - j);
            steps.push_back(LUStep{j, jb, p, (int)w.blk.size()});
            for (int c = j; c < j + jb; c += 64) w.blk.push_back(c);
        }
    }
    if ((int)w.blk.size() > w.nblk_cap) return TAGS_ERR_TOO_LARGE;
    if (plan->nodal_m > 0) {
        // the Z_L and Z_T phases share no rows or columns but the node
        // columns, which each updates in its own rows only: alternate their
        // steps, so that each phase's next panel runs beside the other's
        // whole update
        std::vector<LUStep> a, b, rest;
        for (const LUStep& sp : steps) (sp.phase == 0 ? a : sp.phase == 1 ? b : rest).push_back(sp);
        steps.clear();
        for (size_t q = 0; q < std::max(a.size(), b.size()); q++) {
            if (q < a.size()) steps.push_back(a[q]);
            if (q < b.size()) steps.push_back(b[q]);
        }
        steps.insert(steps.end(), rest.begin(), rest.end());
    }
    // Look-ahead: the panel work of step s+1 runs on the second stream with its
    // own copy of the row map while this stream finishes step s's update of
    // the columns the panel does not touch. The update of the next panel's
    // columns goes first; the panel only permutes logical rows at or after its
    // first column, so the rest of step s (pivot rows before it, the same set
    // of physical rows after it, other columns) is unaffected by its changes.
    // The cooperative panel needs the whole device and is not overlapped.
    // Panels without row exchanges leave the row map alone: they share it with
    // the update (unless their step sets a new phase order, which they do on a
    // copy like a pivoting panel).
    const bool lookahead = g_lookahead.load(std::memory_order_relaxed);
    const bool tournament = g_panel_kind.load(std::memory_order_relaxed) == PANEL_TOURNAMENT;
    int* pcur = w.perm;
    int* palt = w.perm_alt;
    if (!steps.empty()) {
        const LUStep& s0 = steps[0];
        if ((rc = set_order(s0, pcur, st))) return rc;
        if ((rc = step_panels<T>(w, pr, st, A, ld, pcur, plan->phases[s0.phase], s0)))
            return rc;
    }
    for (size_t k = 0; k < steps.size(); k++) {
        const LUStep& s = steps[k];
        const LUPhase& ph = plan->phases[s.phase];
        const int c1 = s.j + s.jb;
        // trailing column ranges of the step: [c1, c_end) and the phase's extra range
        int rb[2] = {c1, ph.x_begin}, re[2] = {ph.c_end, ph.x_end};
        if (ph.x_begin == ph.c_end && ph.x_end > ph.x_begin) {
            re[0] = ph.x_end;
            re[1] = rb[1];
        }
        const LUPhase* nph = k + 1 < steps.size() ? &plan->phases[steps[k + 1].phase] : nullptr;
        // the nodal plan's node phase starts after the whole of the current
        // blocks' work (its Schur complement reads their finished rows)
        const bool nodal_next =
            plan->nodal_m > 0 && nph && steps[k + 1].phase == 2 && s.phase != 2;
        const bool overlap = nph && lookahead && (tournament || !nph->pivot) && !nodal_next;
        if (nph && !overlap) {
            // in order: the whole update of this step, then the next panel (the
            // next panel may permute rows beyond this step's row range, which
            // the rest of this update still addresses through the row map)
            for (int q = 0; q < 2; q++)
                if ((rc = step_update<T>(w, pr, st, A, ld, pcur, ph, s, rb[q], re[q])))
                    return rc;
            const LUStep& nx = steps[k + 1];
            if (nodal_next && (rc = nodal_prepare<T>(w, pr, st, A, ld, pcur, *plan, steps)))
                return rc;
            if ((rc = set_order(nx, pcur, st))) return rc;
            if ((rc = step_panels<T>(w, pr, st, A, ld, pcur, plan->phases[nx.phase], nx)))
                return rc;
        } else if (nph) {
            const LUStep& nx = steps[k + 1];
            const int nb = nx.j, ne = nx.j + nx.jb;
            // columns of the next panel first
            for (int q = 0; q < 2; q++) {
                const int cb = std::max(rb[q], nb), ce = std::min(re[q], ne);
                if ((rc = step_update<T>(w, pr, st, A, ld, pcur, ph, s, cb, ce))) return rc;
            }
            // the next panel on the second stream, with its own copy of the row
            // map when it changes the map
            const bool own_map =
                nph->pivot || (nph->order_off >= 0 && nx.j == nph->c_begin);
            TAGS_HIP(hipEventRecord(w.ev_left, st));
            TAGS_HIP(hipStreamWaitEvent(w.s2, w.ev_left, 0));
            int* pnx = pcur;
            if (own_map) {
                TAGS_HIP(hipMemcpyAsync(palt, pcur, sizeof(int) * n * (size_t)w.nb,
                                        hipMemcpyDeviceToDevice, w.s2));
                if ((rc = set_order(nx, palt, w.s2))) return rc;
                pnx = palt;
            }
            if ((rc = step_panels<T>(w, pr, w.s2, A, ld, pnx, *nph, nx))) return rc;
            TAGS_HIP(hipEventRecord(w.ev_panel, w.s2));
            // the remaining columns of this step, in update_chunks_now() launches
            // (each launch boundary lets waiting panel work-groups start: the
            // tournament's need a whole CU), at most 2 when the next panel has
            // no row exchanges (one work-group of 512 threads)
            int nch = update_chunks_now();
            if (!nph->pivot) nch = std::min(nch, 2);
            for (int q = 0; q < 2; q++) {
                if (re[q] <= rb[q]) continue;
                const int lo1 = rb[q], hi1 = std::min(re[q], nb);
                const int lo2 = std::max(rb[q], ne), hi2 = re[q];
                if ((rc = step_update<T>(w, pr, st, A, ld, pcur, ph, s, lo1, hi1)))
                    return rc;
                const int width = hi2 - lo2;
                const int cw = std::max(128, (width / nch + 63) / 64 * 64);
                for (int c = lo2; c < hi2; c += cw)
                    if ((rc = step_update<T>(w, pr, st, A, ld, pcur, ph, s, c,
                                             std::min(hi2, c + cw))))
                        return rc;
            }
            TAGS_HIP(hipStreamWaitEvent(st, w.ev_panel, 0));
            if (own_map) std::swap(pcur, palt);
        } else {
            for (int q = 0; q < 2; q++)
                if ((rc = step_update<T>(w, pr, st, A, ld, pcur, ph, s, rb[q], re[q])))
                    return rc;
        }
    }
    if (pcur != w.perm) std::swap(w.perm, w.perm_alt);
    w.nblk = (int)w.blk.size();
    w.step_j.clear();
    w.step_w.clear();
    w.step_slot.clear();
    for (const LUStep& sp : steps) {
        w.step_j.push_back(sp.j);
        w.step_w.push_back(sp.jb);
        w.step_slot.push_back(sp.slot);
    }
    // the flow solves' view of the blocks: usable when the phases tile the
    // columns in order and reproduce the block list exactly
    TrsvGeom& g = w.geom;
    g = TrsvGeom();
    const int np = (int)plan->phases.size();
    if (np >= 1 && np <= kTrsvMaxPhases && width % 64 == 0) {
        g.np = np;
        g.width = width;
        int nbk = 0;
        bool ok = true;
        for (int p = 0; p < np; p++) {
            const LUPhase& ph = plan->phases[p];
            g.pb[p] = ph.c_begin;
            g.pe[p] = ph.c_end;
            g.first[p] = nbk;
            const int len = ph.c_end - ph.c_begin;
            if (len < 0 || (p > 0 && ph.c_begin != plan->phases[p - 1].c_end)) ok = false;
            nbk += len / width * (width / 64) + (len % width + 63) / 64;
        }
        g.first[np] = nbk;
        ok = ok && nbk == (int)w.blk.size();
        for (int q = 0; ok && q < nbk; q++) {
            int r0, wq;
            trsv_block(g, q, r0, wq);
            ok = r0 == w.blk[q] && wq >= 1 && wq <= 64;
        }
        g.ok = ok ? 1 : 0;
    }
    return 0;
}

static std::atomic<int> g_solve_rows{2};
void lu_set_solve_rows(int rpb) { g_solve_rows.store(rpb == 1 || rpb == 2 ? rpb : 0); }
int lu_solve_rows() { return g_solve_rows.load(); }

template <typename T, bool UP, int RB>
static void flow_launch(LUWork<T>& w, const cx<T>* A, long ld, cx<T>* y, hipStream_t st,
                        int b_lo, int b_hi, unsigned epoch, const FlowPair& pr = FlowPair())
{
    const BatchStrides b = w.bs();
    const int ngrp = (b_hi - b_lo + RB - 1) / RB;
    const int nitems = b.nb * (pr.on ? 2 : 1);
    // 16 work-groups per item and range take the groups in turn (enough to
    // stream ahead of the chain), at most one per group
    const int nwg = std::min(ngrp * nitems, 16 * nitems);
    hipLaunchKernelGGL((k_trsv_flow<T, UP, RB>), dim3(nwg), dim3(256 * RB), 0, st, A, ld,
                       (const int*)w.perm, y, (const cx<T>*)(UP ? w.invU : w.invL), w.geom, b_lo,
                       b_hi, nitems, w.tctl, w.flags + 1, epoch, b.a, b.p, b.v, b.inv,
                       w.tctl_stride, b.f, pr);
}

// The blocks [b_lo, b_hi) of the rows [c_lo, c_hi); false when the range does
// not start and end on block boundaries.
static bool flow_range(const TrsvGeom& g, int c_lo, int c_hi, int* b_lo_, int* b_hi_)
{
    const int nbk = g.first[g.np];
    int b_lo = nbk, b_hi = nbk;
    for (int q = 0; q < nbk; q++) {
        int r0, wq;
        trsv_block(g, q, r0, wq);
        if (b_lo == nbk && r0 >= c_lo) b_lo = q;
        if (r0 >= c_hi) {
            b_hi = q;
            break;
        }
    }
    *b_lo_ = b_lo;
    *b_hi_ = b_hi;
    if (b_lo >= b_hi) return true;
    int r0, wq;
    trsv_block(g, b_lo, r0, wq);
    if (r0 != c_lo) return false;
    trsv_block(g, b_hi - 1, r0, wq);
    return r0 + wq == c_hi;
}

// Flow solves over the blocks of the rows [c_lo, c_hi), forward then
// backward; false when the factorisation's blocks do not allow them. With a
// second range [c_lo2, c_hi2) (c_lo2 >= 0) independent of the first and of the
// same block structure, both run in the same launches (FlowPair); false when
// the structures differ.
template <typename T>
static bool flow_solve(LUWork<T>& w, const cx<T>* A, long ld, cx<T>* y, hipStream_t st,
                       int c_lo, int c_hi, int c_lo2 = -1, int c_hi2 = -1)
{
    const Tuning* tu = tuning_current();
    int rb = tu && tu->solve_rows >= 0 ? tu->solve_rows
                                       : g_solve_rows.load(std::memory_order_relaxed);
    // complex128 groups take one block (two would not fit the registers)
    if (sizeof(T) == 8 && rb > 1) rb = 1;
    if (!rb || !w.geom.ok || !w.tctl) return false;
    const TrsvGeom& g = w.geom;
    int b_lo, b_hi;
    if (!flow_range(g, c_lo, c_hi, &b_lo, &b_hi)) return false;
    FlowPair pr;
    if (c_lo2 >= 0) {
        int b_lo2, b_hi2;
        if (!flow_range(g, c_lo2, c_hi2, &b_lo2, &b_hi2)) return false;
        if (b_hi2 - b_lo2 != b_hi - b_lo || b_lo >= b_hi) return false;
        for (int q = b_lo; q < b_hi; q++) {
            int r0, wq, r2, w2;
            trsv_block(g, q, r0, wq);
            trsv_block(g, q + (b_lo2 - b_lo), r2, w2);
            if (w2 != wq || r2 - r0 != c_lo2 - c_lo) return false;
        }
        const int R = sizeof(T) == 8 ? 1 : rb;
        pr.on = 1;
        pr.a = pr.p = pr.v = c_lo2 - c_lo;
        pr.inv = 4096L * (b_lo2 - b_lo);
        pr.done = (b_hi - b_lo + R - 1) / R;
    }
    if (b_lo >= b_hi) return true;
    for (int dir = 0; dir < 2; dir++) {
        if (++w.epoch == 0) ++w.epoch;
        const unsigned ep = w.epoch;
        constexpr int R2 = sizeof(T) == 8 ? 1 : 2;
        if (dir == 0) {
            if (rb == 1) flow_launch<T, false, 1>(w, A, ld, y, st, b_lo, b_hi, ep, pr);
            else flow_launch<T, false, R2>(w, A, ld, y, st, b_lo, b_hi, ep, pr);
        } else {
            if (rb == 1) flow_launch<T, true, 1>(w, A, ld, y, st, b_lo, b_hi, ep, pr);
            else flow_launch<T, true, R2>(w, A, ld, y, st, b_lo, b_hi, ep, pr);
        }
    }
    return true;
}

// Solves (L U) x = P b in place on `y` (logical-order input already gathered:
// y[i] = b[perm[i]]); output x in natural order in `y`.
// Triangular solves over the factorisation steps within the columns
// [c_lo, c_hi), the off-diagonal products confined to the rows [c_lo, c_hi):
// per step one k_step_trsv (its diagonal block) and one gemv (the rows beyond
// it), forward then backward.
template <typename T>
static void range_solve(LUWork<T>& w, const cx<T>* A, long ld, cx<T>* y, hipStream_t st,
                        int c_lo, int c_hi)
{
    if (flow_solve<T>(w, A, ld, y, st, c_lo, c_hi)) return;
    const BatchStrides b = w.bs();
    const int ns = (int)w.step_j.size();
    for (int k = 0; k < ns; k++) {
        const int j = w.step_j[k], jb = w.step_w[k];
        if (j < c_lo || j >= c_hi) continue;
        hipLaunchKernelGGL((k_step_trsv<T, false>), dim3(1, 1, b.nb), dim3(256), 0, st, A, ld,
                           w.perm, y, j, jb,
                           (const cx<T>*)(w.invL + (size_t)w.step_slot[k] * 4096), b.a, b.p, b.v,
                           b.inv);
        const int lo = j + jb;
        if (lo < c_hi)
            hipLaunchKernelGGL(k_panel_gemv<T>, dim3((c_hi - lo + 3) / 4, 1, b.nb), dim3(256), 0,
                               st, A, ld, w.perm, y, lo, c_hi, j, jb, b.a, b.p, b.v);
    }
    for (int k = ns - 1; k >= 0; k--) {
        const int j = w.step_j[k], jb = w.step_w[k];
        if (j < c_lo || j >= c_hi) continue;
        hipLaunchKernelGGL((k_step_trsv<T, true>), dim3(1, 1, b.nb), dim3(256), 0, st, A, ld,
                           w.perm, y, j, jb,
                           (const cx<T>*)(w.invU + (size_t)w.step_slot[k] * 4096), b.a, b.p, b.v,
                           b.inv);
        if (j > c_lo)
            hipLaunchKernelGGL(k_panel_gemv<T>, dim3((j - c_lo + 3) / 4, 1, b.nb), dim3(256), 0,
                               st, A, ld, w.perm, y, c_lo, j, j, jb, b.a, b.p, b.v);
    }
}

// Solve with the nodal plan's factors (logical order in, unknowns out):
// z = Z_L^-1 and Z_T^-1 of the current parts, the node right-hand side less
// A^T z_L + B^T z_T (the node rows' incidence entries), x_U = S^-1 of that,
// then x_L,T = z - X x_U (X = Z^-1 A in the node columns of the current rows).
template <typename T>
static void nodal_solve(LUWork<T>& w, const cx<T>* A, long ld, cx<T>* y, hipStream_t st)
{
    const int n = w.n, m = w.nodal_m;
    // the Z_L and Z_T blocks are independent: one pair of flow launches when
    // their blocks match
    if (!flow_solve<T>(w, A, ld, y, st, 0, m, m, 2 * m)) {
        range_solve<T>(w, A, ld, y, st, 0, m);
        range_solve<T>(w, A, ld, y, st, m, 2 * m);
    }
    const BatchStrides b = w.bs();
    hipLaunchKernelGGL(k_rows_gemv<T>, dim3(n - 2 * m, 1, b.nb), dim3(256), 0, st, A, ld, w.perm,
                       y, 2 * m, n, 0, 2 * m, b.a, b.p, b.v);
    range_solve<T>(w, A, ld, y, st, 2 * m, n);
    hipLaunchKernelGGL(k_rows_gemv<T>, dim3(2 * m, 1, b.nb), dim3(256), 0, st, A, ld, w.perm, y,
                       0, 2 * m, 2 * m, n - 2 * m, b.a, b.p, b.v);
}

template <typename T>
int lu_solve_inplace(LUWork<T>& w, const cx<T>* A, long ld, cx<T>* y, hipStream_t st,
                     Profiler* pr)
{
    const int n = w.n;
    Timed tm(pr, FAM_SOLVE, st);
    if (w.nodal_m > 0) {
        nodal_solve<T>(w, A, ld, y, st);
        TAGS_HIP(hipGetLastError());
        tm.end();
        return 0;
    }
    range_solve<T>(w, A, ld, y, st, 0, n);
    TAGS_HIP(hipGetLastError());
    tm.end();
    return 0;
}

template <typename T>
int lu_solve_cols(LUWork<T>& w, cx<T>* A, long ld, hipStream_t st, Profiler* pr, int x0, int nx)
{
    if (w.nodal_m > 0 || nx < 0 || x0 < w.n || w.nb != 1) return TAGS_ERR_ARG;
    if (nx == 0) return 0;
    std::vector<LUStep> steps;
    for (size_t k = 0; k < w.step_j.size(); k++)
        steps.push_back(LUStep{w.step_j[k], w.step_w[k], 0, w.step_slot[k]});
    return backward_cols<T>(w, pr, st, A, ld, w.perm, steps, -1, 0, x0, nx);
}

template <typename T>
int lu_gather_rhs(LUWork<T>& w, const cx<T>* b, cx<T>* y, hipStream_t st)
{
    hipLaunchKernelGGL(k_gather_vec<T>, dim3((w.n + 255) / 256, 1, w.nb), dim3(256), 0, st, b,
                       w.perm, w.n, y);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int residual64(const c64* W, long ld, int n, const c64* x, const c64* b, c64* r,
               unsigned long long* norms, bool want_wnorm, hipStream_t st, Profiler* pr,
               bool zero_norms, const VecBatch& vb)
{
    Timed tm(pr, FAM_RESID, st);
    if (zero_norms) {
        if (vb.nb != 1) return TAGS_ERR_ARG;
        TAGS_HIP(hipMemsetAsync(norms, 0, sizeof(unsigned long long) * (want_wnorm ? 3 : 2), st));
    }
    if (vb.hem) {
        const int rc = hem_apply(W, ld, *vb.hem, x, n, b, r, n, norms, vb.sn, want_wnorm ? 2 : 1,
                                 vb.nb, vb.sw, vb.active, st);
        if (rc) return rc;
    } else {
        hipLaunchKernelGGL(k_residual, dim3(n, 1, vb.nb), dim3(256), 0, st, W, ld, n, x, b, r,
                           norms, want_wnorm ? 1 : 0, vb.sw, vb.sn, vb.active);
        TAGS_HIP(hipGetLastError());
    }
    tm.end();
    return 0;
}

int hem_apply(const c64* W, long ld, const HemOp& op, const c64* x, long sx, const c64* b,
              c64* r, long sr, unsigned long long* norms, long sn, int mode, int nb, long sw,
              const int* active, hipStream_t st)
{
    if (op.m < 1 || op.n < 1 || nb < 1 || !op.sn || !op.en || !op.node_ptr || !op.node_seg)
        return TAGS_ERR_ARG;
    if (mode != 0 && (!b || !norms)) return TAGS_ERR_ARG;
    const long g = 2 * op.m + (op.n + 255) / 256;
    hipLaunchKernelGGL(k_hem_apply, dim3((unsigned)g, 1, (unsigned)nb), dim3(256), 0, st, W, ld, op,
                       x, sx, b, r, sr, norms, sn, mode, sw, active);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int row_scale(const c64* W, long ld, int n, double* d, bool enabled, hipStream_t st,
              const VecBatch& vb)
{
    hipLaunchKernelGGL(k_row_scale, dim3(n, 1, vb.nb), dim3(256), 0, st, W, ld, n, d,
                       enabled ? 1 : 0, vb.sw);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int to_c32_scaled(const c64* W, long ld, int n, const double* d, c32* W32, long ld32,
                  hipStream_t st, const VecBatch& vb)
{
    unsigned gx = (unsigned)std::min(8, (n + 1023) / 1024);
    hipLaunchKernelGGL(k_to_c32_scaled, dim3(gx, n, vb.nb), dim3(256), 0, st, W, ld, n, d, W32,
                       ld32, vb.sw, vb.sw32);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int prep_rhs32(const c64* r, const double* d, const int* perm, int n, double inv_scale, c32* t,
               hipStream_t st, const Grounding* gr, c64* rg, const unsigned long long* rn,
               const VecBatch& vb)
{
    const bool on = gr && gr->g >= 0;
    if (on && vb.nb != 1) return TAGS_ERR_ARG;
    if (on) {
        hipLaunchKernelGGL(k_ground_rhs, dim3(1), dim3(256), 0, st, r, gr->nb, gr->ne, rg);
        TAGS_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_prep_rhs32, dim3((n + 255) / 256, 1, vb.nb), dim3(256), 0, st, r, d,
                       perm, n, inv_scale, rn, on ? gr->g : -1, rg, t, vb.sn, vb.active);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int update_x(c64* x, const c32* z, const double* d, int n, double scale, hipStream_t st,
             const Grounding* gr, const unsigned long long* rn, unsigned long long* dxn,
             const VecBatch& vb)
{
    const bool on = gr && gr->g >= 0;
    if (on && vb.nb != 1) return TAGS_ERR_ARG;
    hipLaunchKernelGGL(k_update_x, dim3((n + 255) / 256, 1, vb.nb), dim3(256), 0, st, x, z, d, n,
                       scale, rn, on ? gr->g : -1, on ? gr->nb : 0, on ? gr->ne : 0, dxn, vb.sn,
                       vb.active);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int ground_sums(const c64* W, long ld, int n, const Grounding& gr, c64* gcol, c64* grow, c64* gg,
                c64* part, hipStream_t st)
{
    if (gr.g < gr.nb || gr.g >= gr.ne || gr.nb < 0 || gr.ne > n) return TAGS_ERR_ARG;
    hipLaunchKernelGGL(k_ground_rowsums, dim3(n), dim3(256), 0, st, W, ld, gr.nb, gr.ne, gcol);
    hipLaunchKernelGGL(k_ground_colsums, dim3((n + 255) / 256, kGroundChunks), dim3(256), 0, st,
                       W, ld, n, gr.nb, gr.ne, kGroundChunks, part);
    hipLaunchKernelGGL(k_ground_finish, dim3(1), dim3(256), 0, st, part, n, gr.nb, gr.ne,
                       kGroundChunks, grow, gg);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int to_c32_grounded(const c64* W, long ld, int n, const Grounding& gr, const c64* gcol,
                    const c64* grow, const c64* gg, c32* W32, long ld32, hipStream_t st)
{
    unsigned gx = (unsigned)std::min(16, (n + 255) / 256);
    hipLaunchKernelGGL(k_to_c32_grounded, dim3(gx, n), dim3(256), 0, st, W, ld, n, gr.g, gcol,
                       grow, gg, W32, ld32);
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


template struct LUWork<double>;
template struct LUWork<float>;
template int lu_factor<double>(LUWork<double>&, c64*, long, hipStream_t, Profiler*,
                               const LUPlan*, int, long);
template int lu_factor<float>(LUWork<float>&, c32*, long, hipStream_t, Profiler*, const LUPlan*,
                              int, long);
template int lu_solve_inplace<double>(LUWork<double>&, const c64*, long, c64*, hipStream_t,
                                      Profiler*);
template int lu_solve_inplace<float>(LUWork<float>&, const c32*, long, c32*, hipStream_t,
                                     Profiler*);
template int lu_gather_rhs<double>(LUWork<double>&, const c64*, c64*, hipStream_t);
template int lu_solve_cols<double>(LUWork<double>&, c64*, long, hipStream_t, Profiler*, int, int);
template int lu_solve_cols<float>(LUWork<float>&, c32*, long, hipStream_t, Profiler*, int, int);
template int lu_gather_rhs<float>(LUWork<float>&, const c32*, c32*, hipStream_t);

}  // namespace tags
