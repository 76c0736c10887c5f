// Panel factorisation with tournament pivoting (communication-avoiding LU,
// "CALU"), used in place of column-by-column partial pivoting across the
// whole panel.
//
// A panel is w <= 64 columns starting at column c0. Its candidate pivot rows
// are the logical rows [c0, cand_end); the rows [c0 + w, row_end) receive
// their L entries afterwards.
//
//  k_tslu       one ordinary launch, one work-group per leaf of S candidate
//               rows (512 for complex64, 256 for complex128). Each leaf runs
//               partial pivoting on its rows (in registers) and nominates up
//               to 64 rows, in pivot order. Tree nodes merge F = S / 64
//               children: the LAST child work-group to finish (atomic arrival
//               counter, acquire/release at agent scope) stacks the children's
//               nominees (their current, i.e. unfactored, rows), runs partial
//               pivoting on that stack and nominates again; the others exit.
//               No work-group ever waits for another, so the launch needs no
//               co-residency. The root's factored stack holds L11 \ U11 of the
//               winners in pivot order; it writes them to the matrix, moves the
//               winners to logical rows c0 .. c0+w-1 (row-map entries only) and
//               resets the counters.
//
// L21 = A21 inv(U11) for the rows below is formed afterwards by the caller
// (one MFMA product with the inverse of U11).
//
// Register layout of a work-group's stack (S rows x 64 columns, 16 waves):
// lane l holds rows {i * 64 + l : i < S/64}, wave w the columns 4w .. 4w+3.
// A column's candidates therefore all sit in one wave: its pivot search is a
// wave maximum (DPP) with no barrier. Pivot choice per column: largest
// |re| + |im| (as LAPACK's izamax; the sign bit is cleared, so -0.0 never
// wins), ties to the lowest lane, then the lowest register row.
#include "lu.h"

namespace tags {

constexpr int TS_NT = 1024;        // threads per work-group
constexpr int TS_NW = TS_NT / 64;  // waves
constexpr int TS_CW = 64 / TS_NW;  // panel columns per wave: wave w owns 4w .. 4w+3

// Rows per lane: 8 for complex64 (S = 512 rows per work-group, 8 children per
// tree node), 4 for complex128 (S = 256, 4 children).
template <typename T>
struct TsShape {
    static constexpr int RPT = sizeof(T) == 8 ? 4 : 8;
    static constexpr int S = 64 * RPT;
    static constexpr int F = S / 64;
};

template <typename T>
struct TsluArgs {
    cx<T>* A;
    long ld;
    int* perm;
    int c0, w, cand_end;
    int* cand;      // nominee lists, 64 logical rows per node (-1: none)
    unsigned* cnt;  // arrival counters of the tree nodes (zero between launches)
    int* info;
    TsluTree tree;
    // batched launch: item blockIdx.z at these strides (elements)
    long sa, sp, scand, scnt, sf;
};

__device__ inline cx<float> readlane_cx(cx<float> v, int l)
{
    return cx<float>{__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.x), l)),
                     __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.y), l))};
}
__device__ inline double readlane_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline cx<double> readlane_cx(cx<double> v, int l)
{
    return cx<double>{readlane_d(v.x, l), readlane_d(v.y, l)};
}

// Wave maximum of a per-lane value (DPP row operations, then lane 63); uniform.
__device__ inline float wave_max(float v)
{
#define TS_DPP_MAXF(CTRL, RM)                                                               \
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v),             \
                                                            __float_as_int(v), CTRL, RM,   \
                                                            0xF, false)));
    TS_DPP_MAXF(0xB1, 0xF) TS_DPP_MAXF(0x4E, 0xF) TS_DPP_MAXF(0x124, 0xF)
    TS_DPP_MAXF(0x128, 0xF) TS_DPP_MAXF(0x142, 0xA) TS_DPP_MAXF(0x143, 0xC)
#undef TS_DPP_MAXF
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ inline double wave_max(double v)
{
#define TS_DPP_MAXD(CTRL, RM)                                                               \
    {                                                                                        \
        const long long b = __double_as_longlong(v);                                        \
        const int lo = (int)(unsigned)b, hi = (int)(unsigned)(b >> 32);                     \
        const unsigned olo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, RM, 0xF, false); \
        const unsigned ohi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, RM, 0xF, false); \
        v = fmax(v, __longlong_as_double((long long)(((unsigned long long)ohi << 32) | olo))); \
    }
    TS_DPP_MAXD(0xB1, 0xF) TS_DPP_MAXD(0x4E, 0xF) TS_DPP_MAXD(0x124, 0xF)
    TS_DPP_MAXD(0x128, 0xF) TS_DPP_MAXD(0x142, 0xA) TS_DPP_MAXD(0x143, 0xC)
#undef TS_DPP_MAXD
    return readlane_d(v, 63);
}

// A lane's part of the stack: RPT rows x TS_CW columns (row i * 64 + lane).
template <typename T>
using StackArr = cx<T>[TsShape<T>::RPT][TS_CW];

template <typename T>
__device__ __forceinline__ cx<T> sget(const StackArr<T>& a, int i, int q)
{
    return a[i][q];
}
template <typename T>
__device__ __forceinline__ void sset(StackArr<T>& a, int i, int q, cx<T> c)
{
    a[i][q] = c;
}
// Row `ri` (wave-uniform, chosen at run time) of this lane's registers: a
// chain of selects (a run-time register index would put the stack in memory).
template <typename T>
__device__ __forceinline__ cx<T> spick(const StackArr<T>& a, int ri, int q)
{
    cx<T> v = a[0][q];
#pragma unroll
    for (int i = 1; i < TsShape<T>::RPT; i++) v = csel(i == ri, a[i][q], v);
    return v;
}
template <typename T>
__device__ __forceinline__ void sput(StackArr<T>& a, int ri, int q, cx<T> c)
{
#pragma unroll
    for (int i = 0; i < TsShape<T>::RPT; i++) a[i][q] = csel(i == ri, c, a[i][q]);
}

template <typename T>
struct TsLds {
    cx<T> L[2][TS_CW][TsShape<T>::S];  // multipliers of the last block, by parity
                                       // (column-major: lanes write consecutive rows)
    cx<T> P[TS_NW][TS_CW][TS_CW];      // per wave: its columns of the block's pivot rows
    cx<T> U[TS_NW][TS_CW][TS_CW];      // per wave: its columns of U12 for the block
    int pv[2][TS_CW];                  // pivot stack rows of the block
    int nb[2];                         // pivots found in the block
};

__device__ inline void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Column K of the block held by this wave: partial pivoting within the wave
// (largest |re| + |im| by a wave maximum; ties to the lowest lane, then the
// lowest register row), pivot row by lane reads, rank-1 update of the block's
// columns right of K. The multipliers (zero for rows that are not live) go
// straight to the block's LDS slot. Returns false when no live row is left.
template <typename T, int K>
__device__ __forceinline__ bool cw_step(StackArr<T>& a, unsigned& live, int kn, int j,
                                        int (&pvk)[TS_CW], int& zc, cx<T> (*Lk)[TsShape<T>::S])
{
    constexpr int RPT = TsShape<T>::RPT;
    if (K >= kn) return false;
    const int lane = threadIdx.x & 63;
    T best = T(-1);
    int bi = 0;
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        const T v = ((live >> i) & 1) ? cabs1(sget<T>(a, i, K)) : T(-1);
        bi = v > best ? i : bi;
        best = v > best ? v : best;
    }
    const T vmax = wave_max(best);
    if (vmax < T(0)) return false;
    const unsigned long long hit = __ballot(best == vmax);
    const int wr = __builtin_amdgcn_readfirstlane(__ffsll((long long)hit) - 1);
    const int wi = __builtin_amdgcn_readlane(bi, wr);
    pvk[K] = wi * 64 + wr;
    cx<T> u[TS_CW];
#pragma unroll
    for (int q = K; q < TS_CW; q++) u[q] = readlane_cx(spick<T>(a, wi, q), wr);
    if (lane == wr) live &= ~(1u << wi);
    if (vmax > T(0)) {
        const cx<T> rinv = crecip_fast(u[K]);
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const bool lv = (live >> i) & 1;
            const cx<T> l = csel(lv, cmul(sget<T>(a, i, K), rinv), cx<T>{0, 0});
            sset<T>(a, i, K, csel(lv, l, sget<T>(a, i, K)));
            Lk[K][i * 64 + lane] = l;
#pragma unroll
            for (int q = K + 1; q < TS_CW; q++) sset<T>(a, i, q, cfnma(sget<T>(a, i, q), l, u[q]));
        }
    } else {
        if (zc < 0) zc = j;
#pragma unroll
        for (int i = 0; i < RPT; i++) Lk[K][i * 64 + lane] = cx<T>{0, 0};
    }
    return true;
}

// Partial pivoting on the register stack (S rows x 64 columns; lane l holds
// rows {i * 64 + l}, wave w columns 4w .. 4w+3); s_lrow[li] is the logical row
// of stack row li (-1 empty). Blocked by the 4 columns of a wave: the owning
// wave factors its block alone (its multipliers going to LDS as they are
// formed), and after one barrier every wave to the right forms its part of the
// pivot rows (U12 = inv(L11) A12, 4 x 4 per wave) and applies the rank-4
// update; the next block's wave does this first and goes on while the others
// update. With `keep_pivot_rows` the pivot rows also take their U12 entries
// (the root needs the factored winners; other levels only their indices).
// Pivots (stack rows) go to s_piv in order; returns their number. Columns
// whose pivot is exactly zero are left unscaled (as LAPACK), the first is
// reported in *zero_col.
template <typename T>
__device__ __forceinline__ int tslu_gepp(StackArr<T>& a, const int* s_lrow, int w, int* s_piv,
                                         TsLds<T>& sh, int* zero_col, bool keep_pivot_rows)
{
    constexpr int RPT = TsShape<T>::RPT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned live = 0;
#pragma unroll
    for (int i = 0; i < RPT; i++)
        if (s_lrow[i * 64 + lane] >= 0) live |= 1u << i;
    int npiv = 0;
    const int nblk = (w + TS_CW - 1) / TS_CW;
    for (int jb = 0; jb < nblk; jb++) {
        const int par = jb & 1;
        if (wave == jb) {
            // the owning wave's chain is the critical path: issue it first on
            // its SIMD (instruction priority, this wave only)
            __builtin_amdgcn_s_setprio(3);
            const int kn = min(TS_CW, w - jb * TS_CW);
            int pvk[TS_CW] = {-1, -1, -1, -1};
            int zc = -1, nb = 0;
            bool go = true;
#define TS_STEP(K)                                                           \
    if (go) {                                                                \
        go = cw_step<T, K>(a, live, kn, jb * TS_CW + K, pvk, zc, sh.L[par]); \
        nb += go ? 1 : 0;                                                    \
    }
            static_assert(TS_CW == 4, "column steps below");
            TS_STEP(0) TS_STEP(1) TS_STEP(2) TS_STEP(3)
#undef TS_STEP
            // columns without a pivot (past the panel, or no live row left)
            for (int q = nb; q < TS_CW; q++)
#pragma unroll
                for (int i = 0; i < RPT; i++) sh.L[par][q][i * 64 + lane] = cx<T>{0, 0};
            if (lane == 0) {
#pragma unroll
                for (int t = 0; t < TS_CW; t++) {
                    sh.pv[par][t] = pvk[t];
                    if (t < nb) s_piv[jb * TS_CW + t] = pvk[t];
                }
                sh.nb[par] = nb;
                if (zero_col && zc >= 0 && *zero_col < 0) *zero_col = zc;
            }
            __builtin_amdgcn_s_setprio(1);
        }
        __syncthreads();
        const int nb = sh.nb[par];
        npiv += nb;
        int prs[TS_CW];
#pragma unroll
        for (int t = 0; t < TS_CW; t++) prs[t] = sh.pv[par][t];
        if (wave != jb) {
#pragma unroll
            for (int t = 0; t < TS_CW; t++)
                if (t < nb && lane == (prs[t] & 63)) live &= ~(1u << (prs[t] >> 6));
        }
        if (wave > jb && nb > 0) {
            if (wave == jb + 1) __builtin_amdgcn_s_setprio(2);  // next owner updates first
            // this wave's columns of the block's pivot rows
#pragma unroll
            for (int t = 0; t < TS_CW; t++) {
                if (t < nb) {
                    const int ri = __builtin_amdgcn_readfirstlane(prs[t] >> 6);
                    if (lane == (prs[t] & 63)) {
#pragma unroll
                        for (int q = 0; q < TS_CW; q++) sh.P[wave][t][q] = spick<T>(a, ri, q);
                    }
                }
            }
            wave_lds_fence();
            if (lane < TS_CW) {  // U12 column `lane`: forward substitution with L11
                cx<T> u[TS_CW];
#pragma unroll
                for (int t = 0; t < TS_CW; t++) {
                    if (t < nb) {
                        cx<T> v = sh.P[wave][t][lane];
#pragma unroll
                        for (int s2 = 0; s2 < t; s2++) v = cfnma(v, sh.L[par][s2][prs[t]], u[s2]);
                        u[t] = v;
                        sh.U[wave][t][lane] = v;
                    }
                }
            }
            wave_lds_fence();
            // rank-nb update (multipliers of non-live rows are zero)
#pragma unroll
            for (int t = 0; t < TS_CW; t++) {
                if (t < nb) {
                    cx<T> l[RPT];
#pragma unroll
                    for (int i = 0; i < RPT; i++) l[i] = sh.L[par][t][i * 64 + lane];
#pragma unroll
                    for (int q = 0; q < TS_CW; q++) {
                        const cx<T> u = sh.U[wave][t][q];
#pragma unroll
                        for (int i = 0; i < RPT; i++) sset<T>(a, i, q, cfnma(sget<T>(a, i, q), l[i], u));
                    }
                }
            }
            if (keep_pivot_rows) {
#pragma unroll
                for (int t = 0; t < TS_CW; t++) {
                    if (t < nb) {
                        const int ri = __builtin_amdgcn_readfirstlane(prs[t] >> 6);
                        if (lane == (prs[t] & 63)) {
#pragma unroll
                            for (int q = 0; q < TS_CW; q++) sput<T>(a, ri, q, sh.U[wave][t][q]);
                        }
                    }
                }
            }
            if (wave == jb + 1) __builtin_amdgcn_s_setprio(1);
        }
        if (nb < min(TS_CW, w - jb * TS_CW)) break;  // no live row left
    }
    __syncthreads();
    return npiv;
}

template <typename T>
__global__ __launch_bounds__(TS_NT) void k_tslu(TsluArgs<T> p)
{
    {
        const long z = blockIdx.z;
        p.A += z * p.sa;
        p.perm += z * p.sp;
        p.cand += z * p.scand;
        p.cnt += z * p.scnt;
        p.info += z * p.sf;
    }
    constexpr int TS_RPT = TsShape<T>::RPT, TS_S = TsShape<T>::S, TS_F = TsShape<T>::F;
    __shared__ int s_lrow[TS_S];
    __shared__ int s_piv[64];
    __shared__ int s_rank[TS_S];
    __shared__ TsLds<T> sh;
    __shared__ int s_tflag[64], s_vac[64];
    __shared__ int s_last, s_zero;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // the panel chain is the critical path of the factorisation: its waves issue
    // ahead of trailing-update waves that share their SIMDs
    __builtin_amdgcn_s_setprio(1);
    const TsluTree& tr = p.tree;
    int level = 0, node = blockIdx.x;
    // leaf: consecutive candidate rows
    {
        const int base = p.c0 + node * TS_S;
        for (int li = t; li < TS_S; li += TS_NT)
            s_lrow[li] = base + li < p.cand_end ? base + li : -1;
    }
    if (t == 0) s_zero = -1;
    __syncthreads();
    StackArr<T> a;
    for (;;) {
        // load the stack: current values of the rows in s_lrow
#pragma unroll
        for (int i = 0; i < TS_RPT; i++) {
            const int lr = s_lrow[i * 64 + lane];
            const cx<T>* src =
                lr >= 0 ? p.A + (long)p.perm[lr] * p.ld + p.c0 + wave * TS_CW : nullptr;
#pragma unroll
            for (int q = 0; q < TS_CW; q++)
                sset<T>(a, i, q, (src && wave * TS_CW + q < p.w) ? src[q] : cx<T>{0, 0});
        }
        const bool root = level == tr.nlev - 1;
        const int npiv = tslu_gepp<T>(a, s_lrow, p.w, s_piv, sh, root ? &s_zero : nullptr, root);
        if (root) {
            // winners' factored rows L11 \ U11 into the matrix (physical rows as
            // of the start of the panel), then the row map
            for (int li = t; li < TS_S; li += TS_NT) s_rank[li] = -1;
            __syncthreads();
            if (t < npiv) s_rank[s_piv[t]] = t;
            __syncthreads();
#pragma unroll
            for (int i = 0; i < TS_RPT; i++) {
                const int li = i * 64 + lane;
                if (s_rank[li] >= 0) {
                    cx<T>* dst = p.A + (long)p.perm[s_lrow[li]] * p.ld + p.c0 + wave * TS_CW;
#pragma unroll
                    for (int q = 0; q < TS_CW; q++)
                        if (wave * TS_CW + q < p.w) dst[q] = sget<T>(a, i, q);
                }
            }
            // Row map: logical row c0 + j <- winner j. Target rows that are not
            // winners take, in order, the positions vacated by winners from
            // outside the target range (all reads before any write: one wave).
            if (t < 64) s_tflag[t] = 0;
            __syncthreads();
            const int wpos = t < npiv ? s_lrow[s_piv[t]] : -1;
            if (t < npiv && wpos < p.c0 + npiv) s_tflag[wpos - p.c0] = 1;
            __syncthreads();
            bool disp = false;
            int drank = 0;
            if (t < 64) {  // wave 0
                const unsigned long long below = (1ull << t) - 1ull;
                const bool vac = t < npiv && wpos >= p.c0 + npiv;
                const unsigned long long vmask = __ballot(vac);
                if (vac) s_vac[__popcll(vmask & below)] = wpos;
                disp = t < npiv && s_tflag[t] == 0;
                const unsigned long long dmask = __ballot(disp);
                drank = __popcll(dmask & below);
            }
            __syncthreads();
            if (t < 64) {
                const int old_t = t < npiv ? p.perm[p.c0 + t] : 0;
                const int old_w = t < npiv ? p.perm[wpos] : 0;
                const int dst = disp ? s_vac[drank] : 0;
                if (t < npiv) p.perm[p.c0 + t] = old_w;
                if (disp) p.perm[dst] = old_t;
                if (t == 0) {
                    int zc = s_zero;
                    if (zc < 0 && npiv < p.w) zc = npiv;
                    if (zc >= 0 && *p.info == 0) *p.info = p.c0 + zc + 1;
                }
            }
            // counters back to zero for the next launch (every arrival is done)
            for (int l = 1; l < tr.nlev; l++)
                for (int q = t; q < tr.nodes[l]; q += TS_NT) p.cnt[tr.cnt_off[l] + q] = 0u;
            return;
        }
        // nominate: wave 0 writes the list, then its thread 0 arrives at the
        // parent with one acquire-release atomic (the release orders wave 0's
        // stores before the arrival; the acquire of the last arrival makes the
        // siblings' lists visible; the barrier passes it to the other waves).
        if (t < 64) {
            int* out = p.cand + tr.cand_off[level] + (long)node * 64;
            __hip_atomic_store(out + t, t < npiv ? s_lrow[s_piv[t]] : -1, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            if (t == 0) {
                const int parent = node / TS_F;
                const int nch = min(TS_F, tr.nodes[level] - parent * TS_F);
                const unsigned old =
                    __hip_atomic_fetch_add(&p.cnt[tr.cnt_off[level + 1] + parent], 1u,
                                           __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                s_last = (int)old == nch - 1;
            }
        }
        __syncthreads();
        if (!s_last) return;
        // last child in: stack the nominees of all children of the parent
        level++;
        node /= TS_F;
        {
            const int first = node * TS_F;
            const int nch = min(TS_F, tr.nodes[level - 1] - first);
            const int* in = p.cand + tr.cand_off[level - 1] + (long)first * 64;
            for (int li = t; li < TS_S; li += TS_NT) {
                const int ch = li >> 6;
                s_lrow[li] = ch < nch ? __hip_atomic_load(in + li, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : -1;
            }
        }
        __syncthreads();
    }
}

int tslu_tree(int cand_rows, int leaf_rows, TsluTree* tr)
{
    if (cand_rows <= 0 || leaf_rows < 128) return TAGS_ERR_ARG;
    const int fan = leaf_rows / 64;
    int nodes = (cand_rows + leaf_rows - 1) / leaf_rows;
    int coff = 0, noff = 0, lev = 0;
    for (;;) {
        if (lev >= TSLU_MAX_LEVELS) return TAGS_ERR_TOO_LARGE;
        tr->nodes[lev] = nodes;
        tr->cand_off[lev] = coff;
        tr->cnt_off[lev] = noff;
        coff += nodes * 64;
        if (lev > 0) noff += nodes;
        lev++;
        if (nodes == 1) break;
        nodes = (nodes + fan - 1) / fan;
    }
    tr->nlev = lev;
    tr->cand_total = coff;
    tr->cnt_total = noff;
    return 0;
}

int tslu_min_leaf_rows() { return TsShape<double>::S; }

template <typename T>
int launch_tslu(cx<T>* A, long ld, int* perm, int c0, int w, int row_end, int cand_end,
                int* cand, unsigned* cnt, int* info, hipStream_t st, const BatchStrides& bs)
{
    if (bs.nb <= 0) return 0;
    if (bs.nb > 65535) return TAGS_ERR_ARG;
    if (w <= 0 || w > 64 || cand_end - c0 < w || row_end < cand_end) return TAGS_ERR_ARG;
    TsluArgs<T> p;
    p.A = A;
    p.ld = ld;
    p.perm = perm;
    p.c0 = c0;
    p.w = w;
    p.cand_end = cand_end;
    p.cand = cand;
    p.cnt = cnt;
    p.info = info;
    p.sa = bs.a;
    p.sp = bs.p;
    p.scand = bs.cand;
    p.scnt = bs.cnt;
    p.sf = bs.f;
    int rc = tslu_tree(cand_end - c0, TsShape<T>::S, &p.tree);
    if (rc) return rc;
    hipLaunchKernelGGL(k_tslu<T>, dim3(p.tree.nodes[0], 1, bs.nb), dim3(TS_NT), 0, st, p);
    TAGS_HIP(hipGetLastError());
    return 0;
}

template int launch_tslu<double>(c64*, long, int*, int, int, int, int, int*, unsigned*, int*,
                                 hipStream_t, const BatchStrides&);
template int launch_tslu<float>(c32*, long, int*, int, int, int, int, int*, unsigned*, int*,
                                hipStream_t, const BatchStrides&);

}  // namespace tags
