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
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    hipEvent_t begin(hipStream_t st)
    {
        hipEvent_t a = get();
        hipEventRecord(a, st);
        return a;
    }
    void end(hipEvent_t a, hipStream_t st, int fam, double flops, bool trailing)
    {
        hipEvent_t b = get();
        hipEventRecord(b, st);
        recs.push_back(Rec{a, b, fam, flops, trailing});
    }
    // call after the stream is synchronised
    void resolve()
    {
        for (auto& r : recs) {
            float t = 0.f;
            hipEventElapsedTime(&t, r.a, r.b);
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
            hipEventDestroy(r.a);
            hipEventDestroy(r.b);
        }
        for (auto e : pool) hipEventDestroy(e);
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
};

template <typename T>
struct PanelArgs {
    cx<T>* A;
    long ld;
    int* perm;
    int r0, M, c0, w, R;
    unsigned long long* keys;
    cx<T>* cand;
    unsigned* bar;
    int* info;
    int* err;
};

// Device work space of one LU of order n in precision T.
template <typename T>
struct LUWork {
    int n = 0, nblk = 0;
    int max_wgs = 256;
    int num_cu = 256;
    size_t lds_limit = 65536;
    int* perm = nullptr;
    cx<T>* invL = nullptr;
    cx<T>* invU = nullptr;
    cx<T>* tmp = nullptr;
    cx<T>* cand = nullptr;
    cx<T>* vec = nullptr;
    unsigned long long* keys = nullptr;
    unsigned* bar = nullptr;
    int* flags = nullptr;  // [0] info, [1] barrier error

    int alloc(int n_, int max_panel_wgs, size_t panel_lds_limit);
    void release();
    int setup_kernel_attrs();
    ~LUWork() { release(); }
};

template <typename T>
int lu_factor(LUWork<T>& w, cx<T>* A, long ld, hipStream_t st, Profiler* pr);
template <typename T>
int lu_solve_inplace(LUWork<T>& w, const cx<T>* A, long ld, cx<T>* y, hipStream_t st,
                     Profiler* pr);
template <typename T>
int lu_gather_rhs(LUWork<T>& w, const cx<T>* b, cx<T>* y, hipStream_t st);
template <typename T>
int gemm_public(const GemmArgs<T>& g, hipStream_t st);

int residual64(const c64* W, long ld, int n, const c64* x, const c64* b, c64* r,
               unsigned long long* norms, bool want_wnorm, hipStream_t st, Profiler* pr);
int row_scale(const c64* W, long ld, int n, double* d, hipStream_t st);
int to_c32_scaled(const c64* W, long ld, int n, const double* d, c32* W32, long ld32,
                  hipStream_t st);
int prep_rhs32(const c64* r, const double* d, const int* perm, int n, double inv_scale, c32* t,
               hipStream_t st);
int update_x(c64* x, const c32* z, const double* d, int n, double scale, hipStream_t st);
int fill_c64(c64* p, long n, c64 v, hipStream_t st);
int sym_from_upper(c64* W, long ld, int n, hipStream_t st);

}  // namespace tags
