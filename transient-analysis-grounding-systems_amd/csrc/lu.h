// Dense complex LU / solve building blocks (see lu.hip).
#pragma once

#include <vector>

#include "tags_common.h"

namespace tags {

enum Family { FAM_FILL = 0, FAM_PANEL, FAM_GEMM, FAM_OTHER, FAM_SOLVE, FAM_RESID, FAM_COUNT };

// HIP-event timing of launch groups on one stream, resolved at sync points.
struct Profiler {
    bool on = false;
    struct Rec {
        hipEvent_t a, b;
        int fam;
        double flops;
        bool trailing;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    double ms[FAM_COUNT] = {0};
    double gemm_flops = 0, trailing_ms = 0, trailing_flops = 0;
    long trailing_launches = 0;

    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // a timing that cannot get or record its events is skipped
    hipEvent_t begin(hipStream_t st)
    {
        hipEvent_t a = get();
        if (a && hipEventRecord(a, st) != hipSuccess) {
            pool.push_back(a);
            return nullptr;
        }
        return a;
    }
    void end(hipEvent_t a, hipStream_t st, int fam, double flops, bool trailing)
    {
        if (!a) return;
        hipEvent_t b = get();
        if (!b || hipEventRecord(b, st) != hipSuccess) {
            pool.push_back(a);
            if (b) pool.push_back(b);
            return;
        }
        recs.push_back(Rec{a, b, fam, flops, trailing});
    }
    // call after the stream is synchronised
    void resolve()
    {
        for (auto& r : recs) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) t = 0.f;
            ms[r.fam] += t;
            if (r.fam == FAM_GEMM) gemm_flops += r.flops;
            if (r.trailing) {
                trailing_ms += t;
                trailing_flops += r.flops;
                trailing_launches++;
            }
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        recs.clear();
    }
    void reset()
    {
        for (int i = 0; i < FAM_COUNT; i++) ms[i] = 0;
        gemm_flops = trailing_ms = trailing_flops = 0;
        trailing_launches = 0;
    }
    ~Profiler()
    {
        for (auto& r : recs) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Records start/end events around a launch group when profiling.
struct Timed {
    Profiler* pr;
    int fam;
    hipStream_t st;
    hipEvent_t a = nullptr;
    Timed(Profiler* p, int f, hipStream_t s) : pr(p), fam(f), st(s)
    {
        if (pr && pr->on) a = pr->begin(st);
    }
    void end(double flops = 0.0, bool trailing = false)
    {
        if (pr && pr->on && a) pr->end(a, st, fam, flops, trailing);
        a = nullptr;
    }
};

template <typename T>
struct GemmArgs {
    int M, N, K;
    const cx<T>* A;
    long lda;
    const int* arow;  // A(i, k) = A[(arow ? arow[i] : i) * lda + k]
    const cx<T>* B;
    long ldb;
    const int* brow;
    cx<T>* C;
    long ldc;
    const int* crow;
    int mode;  // 0: C = A B; 1: C -= A B
    // batched launch (blockIdx.z = item z): every operand and row map of item z
    // starts z times its stride (elements) after item 0's
    int nb = 1;
    long sa = 0, sar = 0, sb = 0, sbr = 0, sc = 0, scr = 0;
    // work space of the split-bf16 variant (complex64 only): ws_item bytes per
    // item, at least gemm_x6_ws_bytes(M, N, K)
    void* ws = nullptr;
    long ws_item = 0;
};

// Work space bytes per item of the split-bf16 complex64 product (its operands
// packed in MFMA fragment order; 0 when K exceeds the LU's largest step).
long gemm_x6_ws_bytes(int M, int N, int K);
// complex64 variants of the split-bf16 product (they need the work space)
constexpr bool gemm_variant_is_x6(int v) { return v >= 22 && v <= 32; }

// One work-group's pivot candidate for one column of the panel.
template <typename T>
struct PanelSlot {
    unsigned long long key;
    unsigned long long pad;
    cx<T> row[64];
};

template <typename T>
struct PanelArgs {
    cx<T>* A;
    long ld;
    int* perm;
    int r0, M, c0, w, R;
    int cand_end;         // logical rows >= cand_end are updated but never pivots
    PanelSlot<T>* slots;  // [2][P]
    unsigned* bar;
    int* info;
    int* err;
};

// Reduction tree of the tournament-pivoting panel (tslu.hip): level 0 holds
// the leaves (one per leaf_rows candidate rows, leaf_rows / 64 children per
// node), the last level the root.
constexpr int TSLU_MAX_LEVELS = 12;
struct TsluTree {
    int nlev;
    int nodes[TSLU_MAX_LEVELS];
    int cand_off[TSLU_MAX_LEVELS];  // offset of each level's nominee lists (ints)
    int cnt_off[TSLU_MAX_LEVELS];   // offset of each level's arrival counters
    int cand_total, cnt_total;
};
int tslu_tree(int cand_rows, int leaf_rows, TsluTree* tr);
int tslu_min_leaf_rows();  // smallest leaf of any precision (largest tree)
// Strides (elements) between the items of a batched factorisation or solve:
// blockIdx.z is the item; its matrix, row map, block inverses, vectors, flags
// and tournament lists start z times these after item 0's.
struct BatchStrides {
    int nb = 1;       // items in the launch
    long a = 0;       // matrix
    long p = 0;       // row map (ints)
    long inv = 0;     // 64 x 64 block inverses
    long v = 0;       // vectors
    long f = 0;       // flags (ints)
    long cand = 0;    // tournament nominee lists (ints)
    long cnt = 0;     // tournament arrival counters
    void* ws = nullptr;  // split-bf16 product work space, ws_item bytes per item
    long ws_item = 0;
};
template <typename T>
int launch_tslu(cx<T>* A, long ld, int* perm, int c0, int w, int row_end, int cand_end,
                int* cand, unsigned* cnt, int* info, hipStream_t st,
                const BatchStrides& bs = BatchStrides());

// Panel algorithm: tournament pivoting (default) or partial pivoting over the
// whole panel with a cooperative launch.
enum PanelKind { PANEL_TOURNAMENT = 0, PANEL_PARTIAL = 1 };
void lu_set_panel_kind(int kind);
int lu_panel_kind();
// Look-ahead (next panel on a second stream during the trailing update): on by default
void lu_set_lookahead(bool on);
// Launches the look-ahead trailing update of a step is split into (by columns)
void lu_set_update_chunks(int n);
// Columns per factorisation step: 128 or 256 for both precisions; w <= 0
// restores the defaults (128 for complex128 factors, 256 for complex64)
void lu_set_step_width(int w);
int lu_step_width(bool c64);

// Tile and schedule choices for the factorisations queued from one host
// thread (a sweep's tuned values): while a TuningScope is alive on a thread,
// its fields override the process-wide settings above and the GEMM variants
// (gemm_set_variants) for the launches that thread queues; a field < 0 (0 for
// step_width) leaves the process-wide value. Index [0] is complex128, [1]
// complex64.
struct Tuning {
    int gemm_big[2] = {-1, -1};
    int gemm_small[2] = {-1, -1};
    int step_width[2] = {0, 0};
    int update_chunks = -1;
    int solve_rows = -1;  // lu_set_solve_rows
};
const Tuning* tuning_current();
struct TuningScope {
    const Tuning* prev;
    explicit TuningScope(const Tuning* t);
    ~TuningScope();
    TuningScope(const TuningScope&) = delete;
    TuningScope& operator=(const TuningScope&) = delete;
};

// Column phases of a factorisation (see lu.hip).
struct LUPhase {
    int c_begin, c_end;   // columns factored in this phase
    int row_end;          // logical rows [column, row_end) may be non-zero
    int cand_end;         // pivot candidates are logical rows [column, cand_end)
    int x_begin, x_end;   // extra trailing columns beyond c_end (empty if x_begin >= x_end)
    bool pivot = true;    // false: no row exchanges (the diagonal is the pivot)
    int order_off = -1;   // >= 0: logical rows [c_begin, n) are set at the start of the
                          // phase to orders[order_off ...] (phases without pivoting only)
};

struct LUPlan {
    std::vector<LUPhase> phases;
    std::vector<int> init_perm;  // logical -> physical row order at start (empty: identity)
    // device copies of init_perm and orders made by the plan's owner (not
    // owned here); when set, a factorisation reads them instead of uploading
    // its own copies, which takes a blocking host copy whenever the plan
    // differs from the work space's last one
    const int* dev_init = nullptr;
    const int* dev_orders = nullptr;
    static LUPlan dense(int n)
    {
        LUPlan p;
        p.phases.push_back(LUPhase{0, n, n, n, 0, 0});
        return p;
    }
    // Immittance matrix in the reference unknown order [U (n), I_L (m), I_T (m)]:
    //   [[0, A^T, B^T], [A, -Z_L, 0], [B, 0, -Z_T]].
    // The node rows are exactly zero in the node columns (Y_E = 0) and stay
    // so while those columns are factored, so the first n columns pivot and
    // update over the 2m impedance rows only (logical rows start as
    // [Z_L rows, Z_T rows, node rows]). Same pivots as dense partial pivoting.
    static LUPlan nodes_first(int m, int n)
    {
        LUPlan p;
        const int N = n + 2 * m;
        p.phases.push_back(LUPhase{0, n, 2 * m, 2 * m, n, N});
        p.phases.push_back(LUPhase{n, N, N, N, 0, 0});
        p.init_perm.resize(N);
        for (int i = 0; i < 2 * m; i++) p.init_perm[i] = n + i;
        for (int i = 0; i < n; i++) p.init_perm[2 * m + i] = i;
        return p;
    }
    // Immittance matrix stored with unknowns [I_L (m), I_T (m), U (n)]:
    //   [[-Z_L, 0, A], [0, -Z_T, B], [A^T, B^T, 0]].
    // Logical rows start as [Z_L rows, node rows, Z_T rows]. While the Z_L
    // columns are factored the Z_T rows are exactly zero in them (and receive
    // no update), so they are skipped; the rest is ordinary partial pivoting
    // (exact). With block_pivot the Z_L columns pivot among Z_L rows only, the
    // Z_T columns then stay untouched too (fewest flops, weaker pivoting).
    static LUPlan zl_first(int m, int n, bool block_pivot = false)
    {
        LUPlan p;
        const int N = n + 2 * m;
        if (block_pivot)
            p.phases.push_back(LUPhase{0, m, m + n, m, 2 * m, N});
        else
            p.phases.push_back(LUPhase{0, m, m + n, m + n, m, N});
        p.phases.push_back(LUPhase{m, N, N, N, 0, 0});
        p.init_perm.resize(N);
        for (int i = 0; i < m; i++) p.init_perm[i] = i;
        for (int i = 0; i < n; i++) p.init_perm[m + i] = 2 * m + i;
        for (int i = 0; i < m; i++) p.init_perm[m + n + i] = m + i;
        return p;
    }
    // The same matrix factored without row exchanges (the mixed form's factor of
    // the grounded matrix): the Z_L columns over rows [Z_L, node] (the Z_T rows
    // are zero there and receive no update; the pivot rows are Z_L rows, zero in
    // the Z_T columns, which therefore receive no update either), then the
    // remaining columns over rows reordered to [Z_T, node], so that each
    // column's pivot is its own diagonal entry.
    static LUPlan zl_first_unpivoted(int m, int n)
    {
        LUPlan p = zl_first(m, n, false);
        const int N = n + 2 * m;
        p.phases[0].pivot = false;
        p.phases[0].cand_end = m + n;
        p.phases[0].x_begin = 2 * m;
        p.phases[1].pivot = false;
        p.phases[1].order_off = 0;
        p.orders.resize(N - m);
        for (int i = 0; i < m + n; i++) p.orders[i] = m + i;
        return p;
    }
    // Nodal block elimination of the [I_L; I_T; U] storage order (m > 0: on):
    // Z_L and Z_T each factored within their own rows (with row exchanges among
    // them, or with `pivot` false on their diagonals: both are complex
    // symmetric with a definite real or imaginary part, as is S), the node
    // columns carried as extra columns, so that their trailing updates form
    // L^-1 P A; before the node phase the node columns of those rows are
    // completed to X = Z^-1 A (backward blocked solves), and the node block is
    // set to its Schur complement S = -(A^T X_L + B^T X_T) from the non-zeros of
    // the node rows (the incidence), not by dense products; then S is factored.
    static LUPlan nodal(int m, int n, bool pivot = true)
    {
        LUPlan p;
        const int N = n + 2 * m;
        p.phases.push_back(LUPhase{0, m, m, m, 2 * m, N, pivot});
        p.phases.push_back(LUPhase{m, 2 * m, 2 * m, 2 * m, 2 * m, N, pivot});
        p.phases.push_back(LUPhase{2 * m, N, N, N, 0, 0, pivot});
        p.nodal_m = m;
        return p;
    }
    int nodal_m = 0;
    std::vector<int> orders;  // row orders referenced by phases (order_off)
    bool pivots() const
    {
        for (const LUPhase& ph : phases)
            if (ph.pivot) return true;
        return phases.empty();
    }
};

// Device work space of `batch` LUs of order n in precision T (the items of a
// batched factorisation, each with its own row maps, inverses, vectors, flags
// and tournament lists at the strides of bs()).
// The 64-row diagonal blocks of a factorisation, as the flow solves address
// them: the phases' column ranges [pb, pe) in order, steps of `width` columns
// from each phase's start, sub-blocks of 64 from each step's start. Block q
// (its global index, which is also the slot of its 64 x 64 inverses) starts
// at row0 and is wq <= 64 rows wide. ok = 0 when the phases do not tile the
// columns in order (the solves then take the per-step launches).
constexpr int kTrsvMaxPhases = 8;
struct TrsvGeom {
    int ok = 0;
    int np = 0, width = 256;
    int pb[kTrsvMaxPhases] = {}, pe[kTrsvMaxPhases] = {};
    int first[kTrsvMaxPhases + 1] = {};  // first block of each phase; first[np] = all
};
__host__ __device__ inline void trsv_block(const TrsvGeom& g, int q, int& row0, int& wq)
{
    int p = 0;
    while (p + 1 < g.np && q >= g.first[p + 1]) p++;
    const int l = q - g.first[p];
    const int spb = g.width >> 6;
    const int k = l / spb, t = l - k * spb;
    const int j = g.pb[p] + g.width * k;
    const int jb = g.pe[p] - j < g.width ? g.pe[p] - j : g.width;
    row0 = j + 64 * t;
    wq = jb - 64 * t < 64 ? jb - 64 * t : 64;
}
// Two ranges of one flow launch: with `on`, each item's second range (items
// 2 i and 2 i + 1 of the launch are item i's two ranges) has the same block
// structure as its first, shifted by `a` columns (matrix), `p` rows (row map),
// `v` entries (vector) and `inv` elements (block inverses); its completion
// words follow the first range's `done` groups.
struct FlowPair {
    int on = 0;
    long a = 0, p = 0, v = 0, inv = 0;
    int done = 0;
};
// Triangular solves: flow (one launch per range and direction, work-groups
// of 64 rpb rows in dependency order; rpb 1 or 2) or the per-step launches
// (0)
void lu_set_solve_rows(int rpb);
int lu_solve_rows();

template <typename T>
struct LUWork {
    int n = 0, nblk = 0, nblk_cap = 0;
    int batch = 1;   // items the work space holds
    int nb = 1;      // items of the current factorisation (<= batch)
    long sA = 0;     // matrix stride (elements) of the current factorisation
    long cand_stride = 0, cnt_stride = 0;
    // allocated element counts (kept while large enough for later orders)
    size_t cap_perm = 0, cap_inv = 0, cap_items = 0, cap_cand = 0, cap_cnt = 0;
    int max_wgs = 256;
    int num_cu = 256;
    size_t lds_limit = 65536;
    int* perm = nullptr;
    int* perm_alt = nullptr;      // row map of the look-ahead panel (swapped with perm)
    hipStream_t s2 = nullptr;     // look-ahead stream (default priority)
    hipEvent_t ev_left = nullptr, ev_panel = nullptr;
    cx<T>* invL = nullptr;  // dense 64 x 64 inverses of the diagonal blocks
    cx<T>* invU = nullptr;
    cx<T>* vec = nullptr;
    PanelSlot<T>* slots = nullptr;
    unsigned* bar = nullptr;
    int* flags = nullptr;   // [0] info, [1] barrier error
    int* d_init = nullptr;  // device copy of the plan's initial row order
    int* d_orders = nullptr;  // device copy of the plan's phase row orders (capacity n)
    int* tcand = nullptr;       // tournament nominee lists
    unsigned* tcnt = nullptr;   // tournament arrival counters (kept at zero)
    std::vector<int> blk;   // start column of every diagonal block, in order
    // the steps of the last factorisation: first column, width, first block
    std::vector<int> step_j, step_w, step_slot;
    std::vector<int> h_init, h_orders;  // what d_init / d_orders hold
    double flops = 0;       // real flops of the last factorisation (GEMM + panel)
    int nodal_m = 0;        // the last factorisation's nodal block size (0: none)
    // flow solves: the last factorisation's block geometry; per item a ticket
    // counter, a spare word and one completion word per group of blocks (each
    // launch tags its completions with its own epoch, so nothing is cleared
    // between launches)
    TrsvGeom geom;
    unsigned* tctl = nullptr;
    size_t cap_tctl = 0;
    long tctl_stride = 0;
    unsigned epoch = 0;
    // complex64 only: work space of the split-bf16 products (operands packed
    // in MFMA fragment order), x6_item bytes per item
    void* x6ws = nullptr;
    void* x6ws2 = nullptr;  // the same for products on the look-ahead stream
    size_t cap_x6 = 0;
    long x6_item = 0;

    int alloc(int n_, int max_panel_wgs, size_t panel_lds_limit, int batch_ = 1);
    void release();
    int setup_kernel_attrs();
    // strides of the current call (nb items, matrix stride sA)
    BatchStrides bs() const
    {
        BatchStrides b;
        b.nb = nb;
        b.a = sA;
        b.p = n;
        b.inv = 4096L * nblk_cap;
        b.v = n;
        b.f = 4;
        b.cand = cand_stride;
        b.cnt = cnt_stride;
        b.ws = x6ws;
        b.ws_item = x6_item;
        return b;
    }
    ~LUWork() { release(); }
};

// Factors nb items (matrices A + z sA, z < nb <= w.batch) alike; the solves
// below then work on all nb items (vectors at stride n).
template <typename T>
int lu_factor(LUWork<T>& w, cx<T>* A, long ld, hipStream_t st, Profiler* pr,
              const LUPlan* plan = nullptr, int nb = 1, long sA = 0);
template <typename T>
int lu_solve_inplace(LUWork<T>& w, const cx<T>* A, long ld, cx<T>* y, hipStream_t st,
                     Profiler* pr);
template <typename T>
int lu_gather_rhs(LUWork<T>& w, const cx<T>* b, cx<T>* y, hipStream_t st);
// After a factorisation whose plan carried extra columns [x0, x0 + nx) through
// its trailing updates (they hold L^-1 P C): completes them to U^-1 L^-1 P C =
// A^-1 C, in logical rows (physical row w.perm[i] holds unknown i). Not for the
// nodal plan.
template <typename T>
int lu_solve_cols(LUWork<T>& w, cx<T>* A, long ld, hipStream_t st, Profiler* pr, int x0, int nx);
// trailing selects the kernel name k_trailing_update (same code as k_cgemm)
template <typename T>
int launch_gemm(const GemmArgs<T>& g, hipStream_t st, bool trailing = false);
template <typename T>
int launch_gemm_variant(const GemmArgs<T>& g, int variant, hipStream_t st);
int gemm_set_variants(int precision, int big, int small);
int gemm_variant_count(int precision);

// Structure of an immittance matrix W of the HEM: the dense impedance blocks
// at rows and columns [zl_off, zl_off + m) and [zt_off, zt_off + m) of W, the
// incidence entries from the segments' nodes (A: +1 at the start node, -1 at
// the end node; B: 1/2 at both; one entry where the two nodes coincide), zero
// elsewhere. Products with W through it read only the impedance blocks of W
// (2 m^2 of its N^2 entries). node_ptr / node_seg list each node's segment
// ends in segment order: 2 s for the start of segment s, 2 s + 1 for its end.
struct HemOp {
    long m = 0, n = 0;  // segments, nodes
    long node_off = 0, zl_off = 0, zt_off = 0;
    const int* sn = nullptr;        // start node of each segment (device, m)
    const int* en = nullptr;        // end node (device, m)
    const int* node_ptr = nullptr;  // device, n + 1
    const int* node_seg = nullptr;  // device, node_ptr[n]
};

// norms[0..1] (and [2] with want_wnorm) receive |r|, |x| (|W|) as max-reductions;
// zero_norms false: the caller has already zeroed them (no memset node)
// Batched forms (nb items): W of item z at W + z sw, vectors at stride n, the
// norms of item z at norms + z sn.
struct VecBatch {
    int nb = 1;
    long sw = 0;   // c64 matrix stride
    long sw32 = 0; // c32 matrix stride
    long sn = 0;   // norms stride (u64)
    const int* active = nullptr;  // item z skipped where active[z] == 0 (nullable)
    const HemOp* hem = nullptr;   // W has HEM structure: products read its impedance blocks only
};
int residual64(const c64* W, long ld, int n, const c64* x, const c64* b, c64* r,
               unsigned long long* norms, bool want_wnorm, hipStream_t st, Profiler* pr,
               bool zero_norms = true, const VecBatch& vb = VecBatch());
int row_scale(const c64* W, long ld, int n, double* d, bool enabled, hipStream_t st,
              const VecBatch& vb = VecBatch());
int to_c32_scaled(const c64* W, long ld, int n, const double* d, c32* W32, long ld32,
                  hipStream_t st, const VecBatch& vb = VecBatch());
// Grounded node unknowns of the mixed form (lu.hip): node rows and columns
// [nb, ne) of the physical order, reference node g (g < 0: off).
struct Grounding {
    int g = -1;
    int nb = 0, ne = 0;
};
constexpr int kGroundChunks = 32;  // row chunks of the node column sums
// gcol, grow: n each; gg: 1; part: kGroundChunks * n
int ground_sums(const c64* W, long ld, int n, const Grounding& gr, c64* gcol, c64* grow, c64* gg,
                c64* part, hipStream_t st);
int to_c32_grounded(const c64* W, long ld, int n, const Grounding& gr, const c64* gcol,
                    const c64* grow, const c64* gg, c32* W32, long ld32, hipStream_t st);
// rg: one c64 of device scratch (used with grounding); rn: when given, |r|
// (max norm as left by residual64 in device memory) replaces the host scale
int prep_rhs32(const c64* r, const double* d, const int* perm, int n, double inv_scale, c32* t,
               hipStream_t st, const Grounding* gr = nullptr, c64* rg = nullptr,
               const unsigned long long* rn = nullptr, const VecBatch& vb = VecBatch());
// dxn (optional, zeroed beforehand): receives the max-norm of the correction
int update_x(c64* x, const c32* z, const double* d, int n, double scale, hipStream_t st,
             const Grounding* gr = nullptr, const unsigned long long* rn = nullptr,
             unsigned long long* dxn = nullptr, const VecBatch& vb = VecBatch());
int fill_c64(c64* p, long n, c64 v, hipStream_t st);

// Work space of the FP64 flexible GMRES correction solve (gmres.hip) for cap
// items of order n with up to kmax Arnoldi steps.
struct GmresWork {
    int n = 0, kmax = 0, cap = 0, nchunk = 0;
    long n_cap = 0, part_cap = 0;
    int kmax_cap = 0, item_cap = 0;
    c64 *V = nullptr, *Z = nullptr, *w = nullptr, *H = nullptr, *g = nullptr, *y = nullptr;
    c64 *h1 = nullptr, *h2 = nullptr, *nrm = nullptr, *sn = nullptr, *part = nullptr;
    double *cs = nullptr, *inv = nullptr;
    int alloc(int n_, int kmax_, int cap_);
    void release();
    ~GmresWork() { release(); }
};
// One outer step of GMRES-based refinement for nb items: d ~ W^-1 r by k
// Arnoldi steps right-preconditioned with the complex64 factor (w32, W32,
// row scaling d), then x += d with max |d| into dxn (item stride sn; items
// with active[z] == 0 keep x). W at stride sW (HEM structure when hem is
// given); r, x, t32 at stride n.
int gmres_step(GmresWork& g, int k, int nb, LUWork<float>& w32, const c32* W32, c32* t32,
               const double* d, const c64* W, long sW, const c64* r, c64* x,
               unsigned long long* dxn, long sn, const int* active, hipStream_t st, Profiler* pr,
               const HemOp* hem = nullptr);
// y = W x (mode 0), or r = b - W x with |r|, |x| (mode 1) and |W| (mode 2) as
// max-reductions into norms, for a W of HEM structure; item z at W + z sw,
// x + z sx, b and r + z sr, norms + z sn (items with active[z] == 0 skipped)
int hem_apply(const c64* W, long ld, const HemOp& op, const c64* x, long sx, const c64* b,
              c64* r, long sr, unsigned long long* norms, long sn, int mode, int nb, long sw,
              const int* active, hipStream_t st);
int sym_from_upper(c64* W, long ld, int n, hipStream_t st);

}  // namespace tags
