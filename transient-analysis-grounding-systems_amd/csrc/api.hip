// C ABI of the library: the reference's solve-path API (host pointers, same
// semantics) and the device-resident sweep.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <chrono>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "lu.h"

namespace tags {

struct DevInfo {
    int ok = 0;
    int num_cu = 0;
    size_t lds_limit = 65536;
};

static DevInfo& dev_info()
{
    static DevInfo d;
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (!d.ok) {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess) {
            d.num_cu = p.multiProcessorCount;
            size_t lim = p.sharedMemPerBlock;
            if (p.sharedMemPerBlockOptin > lim) lim = p.sharedMemPerBlockOptin;
            if (lim > 160 * 1024) lim = 160 * 1024;
            d.lds_limit = lim;
            d.ok = 1;
        }
    }
    return d;
}

// Device copy of an electrode array in structure-of-arrays form.
struct DevSegs {
    double* buf = nullptr;  // 8 arrays of m doubles
    long m = 0;
    int upload(const Electrode* e, long m_, hipStream_t st)
    {
        m = m_;
        std::vector<double> h(8 * (size_t)m);
        for (long i = 0; i < m; i++) {
            h[0 * m + i] = e[i].start_point[0];
            h[1 * m + i] = e[i].start_point[1];
            h[2 * m + i] = e[i].start_point[2];
            h[3 * m + i] = e[i].end_point[0];
            h[4 * m + i] = e[i].end_point[1];
            h[5 * m + i] = e[i].end_point[2];
            h[6 * m + i] = e[i].length;
            h[7 * m + i] = e[i].radius;
        }
        TAGS_HIP(hipMalloc(&buf, sizeof(double) * 8 * (size_t)(m > 0 ? m : 1)));
        TAGS_HIP(hipMemcpyAsync(buf, h.data(), sizeof(double) * 8 * m, hipMemcpyHostToDevice, st));
        TAGS_HIP(hipStreamSynchronize(st));
        return 0;
    }
    SegSoA soa() const
    {
        return SegSoA{buf, buf + m, buf + 2 * m, buf + 3 * m, buf + 4 * m, buf + 5 * m,
                      buf + 6 * m, buf + 7 * m};
    }
    ~DevSegs() { hipFree(buf); }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { hipStreamCreate(&s); }
    ~Stream() { hipStreamDestroy(s); }
};

template <typename X>
struct DevBuf {
    X* p = nullptr;
    int alloc(size_t n)
    {
        TAGS_HIP(hipMalloc(&p, sizeof(X) * (n ? n : 1)));
        return 0;
    }
    ~DevBuf() { hipFree(p); }
};

static PairScalars make_scalars(c64 gamma, c64 s, double mur, c64 kappa, c64 ref_l, c64 ref_t,
                                int type)
{
    PairScalars sc;
    sc.gamma = gamma;
    sc.intg_type = type;
    if (type == INTG_NONE || type == INTG_MHEM) {
        sc.zl_dir = sc.zt_dir = sc.zl_img = sc.zt_img = c64{1.0, 0.0};
        return sc;
    }
    // src/electrode.c:317-318 and :397-398
    const c64 iwu = c64{s.x * mur * kMu0 / kFourPi, s.y * mur * kMu0 / kFourPi};
    const c64 fk = c64{kFourPi * kappa.x, kFourPi * kappa.y};
    sc.zl_dir = iwu;
    sc.zt_dir = crecip(fk);
    sc.zl_img = cmul(ref_l, iwu);
    sc.zt_img = cdiv(ref_t, fk);
    return sc;
}

// ---------------------------------------------------------------------------
// Dense solver shared by solve_immittance, tags_dense_solve and the sweep.
// It solves nb items at once (the frequency samples of a sweep batch): item z
// has its matrix at W + z sW, right-hand side and solution at bvec + z n, and
// every launch covers all items (blockIdx.z). Each item's arithmetic is that
// of a one-item solve: refinement steps past the first kSpecIters run only on
// the items still refining (an active mask), and an item that misses the
// acceptance rule is retried or refactored alone.
// ---------------------------------------------------------------------------
struct Solver {
    static constexpr int kSpecIters = 3;   // refinement steps queued before the first check
    static constexpr int kMaxIters = 20;   // refinement steps in all
    // GMRES-based refinement (gmres_k > 0): each step is gmres_k Arnoldi steps
    // preconditioned by the complex64 factor; kGmresSpec steps are queued
    // before the first check, kGmresMaxIters in all
    static constexpr int kGmresSpec = 2;
    static constexpr int kGmresMaxIters = 6;
    // Acceptance needs both a small backward error |b - W x| / (|W| |x|) and a
    // small estimate of the error left in x. The backward error alone does not
    // bound the error of x for an ill-conditioned W (cond_1 ~ 6e8 on the
    // 30 x 30 m grid at 31 kHz: a backward error of 1e-15 still allows ~1e-7).
    // The correction dx_k of step k is about (1 - rho) times the error of
    // x_{k-1}, rho the contraction per step, so the error of x_k is about
    // rho / (1 - rho) |dx_k|; rho is taken as |dx_k| / |dx_{k-1}| when that is
    // below 1/2, otherwise (or at the first step) the estimate is |dx_k| itself.
    static constexpr double kAccept = 1e-15;      // backward error, a few FP64 ulps
    static constexpr double kFwdAccept = 1e-10;   // estimated error / |x|, 100 x below 1e-8
    // refinement stops at the first acceptable iterate, or when a step gains
    // less than a factor kMinGain in both measures
    static constexpr double kMinGain = 0.8;
    // norms per refinement step k at norms[kNS k ..]: |r|, |x|, |W| (k = 0 only)
    // and |dx| (the correction of step k), as bit patterns of non-negative doubles;
    // an item's norms take kNormStride
    static constexpr int kNS = 4;
    static constexpr long kNormStride = kNS * (kMaxIters + 1);
    int n = 0;
    int cap = 0;      // items the buffers hold
    bool cap_mixed = false;
    hipStream_t st = nullptr;
    Profiler* pr = nullptr;
    LUWork<double> w64;
    LUWork<float> w32;
    c32* W32 = nullptr;
    double* dscale = nullptr;
    c64* x = nullptr;
    c64* r = nullptr;
    c64* b = nullptr;
    c32* t32 = nullptr;
    // residual and correction norms; hnorms is their pinned copy
    unsigned long long* norms = nullptr;
    unsigned long long* hnorms = nullptr;
    int* hflags = nullptr;  // pinned copy of the factorisations' flags (4 per item:
                            // info, barrier error)
    int* active = nullptr;  // items still refining (device, and pinned host copy)
    int* hactive = nullptr;
    // grounded node unknowns of the mixed factor (ground.g < 0: off; one item only)
    Grounding ground;
    bool ground_valid = false;  // gcol / grow / gsum hold the sums of the current W's node part
    c64* gcol = nullptr;
    c64* grow = nullptr;
    c64* gsum = nullptr;  // [0] node block sum, [1] node sum of the right-hand side
    c64* gpart = nullptr;
    // the solve in flight (issue_* then finish)
    bool pending_mixed = false;
    int queued = 0;  // refinement steps queued by issue_mixed
    int nb = 1;      // its items
    long sW = 0;     // and their matrix stride
    // Arnoldi steps per refinement step (0: classic refinement, one correction
    // from the complex64 factor per step); set before issue_mixed
    int gmres_k = 0;
    GmresWork gm;
    // the structure of W when it is an immittance matrix of the HEM (nullptr:
    // dense): the refinement's FP64 products then read its impedance blocks only
    const HemOp* hem = nullptr;

    // buffers for `batch` items of order n_ (the FP64 work space for all of them
    // in the plain form, for one in the mixed form, where it is the fallback's);
    // kept while large enough, so that one solver serves systems of different
    // orders in turn
    size_t vec_cap = 0, w32_cap = 0;
    size_t w32_hi = 0;  // largest complex64 copy asked for so far
    int item_cap = 0;
    int ensure(int n_, bool mixed, int batch = 1)
    {
        DevInfo& di = dev_info();
        if (!di.ok) return TAGS_ERR_HIP;
        if (batch < 1 || n_ < 1) return TAGS_ERR_ARG;
        const size_t need = (size_t)n_ * batch;
        if (need > vec_cap || batch > item_cap) {
            // grow to cover every earlier request as well: a solver that
            // serves systems of different orders in turn (a lane of a
            // multi-case solve) then reallocates, which waits for the whole
            // device, only while its largest request still grows
            const size_t vc = std::max(vec_cap, need);
            const int ic = std::max(item_cap, batch);
            free_vecs();
            vec_cap = vc;
            item_cap = ic;
            TAGS_HIP(hipMalloc(&dscale, sizeof(double) * vec_cap));
            TAGS_HIP(hipMalloc(&x, sizeof(c64) * vec_cap));
            TAGS_HIP(hipMalloc(&r, sizeof(c64) * vec_cap));
            TAGS_HIP(hipMalloc(&b, sizeof(c64) * vec_cap));
            TAGS_HIP(hipMalloc(&t32, sizeof(c32) * vec_cap));
            TAGS_HIP(hipMalloc(&norms, sizeof(unsigned long long) * kNormStride * item_cap));
            TAGS_HIP(hipHostMalloc(&hnorms, sizeof(unsigned long long) * kNormStride * item_cap,
                                   0));
            TAGS_HIP(hipHostMalloc(&hflags, sizeof(int) * 4 * item_cap, 0));
            TAGS_HIP(hipMalloc(&active, sizeof(int) * item_cap));
            TAGS_HIP(hipHostMalloc(&hactive, sizeof(int) * item_cap, 0));
        }
        if (n_ != n) {
            // the grounding sums belong to one system
            hipFree(gcol);
            hipFree(grow);
            hipFree(gsum);
            hipFree(gpart);
            gcol = grow = gsum = gpart = nullptr;
            ground_valid = false;
        }
        n = n_;
        cap = batch;
        int rc = w64.alloc(n, di.num_cu, di.lds_limit, mixed ? 1 : batch);
        if (rc) return rc;
        if (mixed) {
            if ((rc = w32.alloc(n, di.num_cu, di.lds_limit, batch))) return rc;
            const size_t need32 = (size_t)n * n * batch;
            w32_hi = std::max(w32_hi, need32);
            if (need32 > w32_cap) {
                hipFree(W32);
                W32 = nullptr;
                w32_cap = 0;
                TAGS_HIP(hipMalloc(&W32, sizeof(c32) * w32_hi));
                w32_cap = w32_hi;
            }
        }
        return 0;
    }
    void free_vecs()
    {
        hipFree(dscale);
        hipFree(x);
        hipFree(r);
        hipFree(b);
        hipFree(t32);
        hipFree(norms);
        if (hnorms) hipHostFree(hnorms);
        if (hflags) hipHostFree(hflags);
        hipFree(active);
        if (hactive) hipHostFree(hactive);
        hipFree(W32);
        hipFree(gcol);
        hipFree(grow);
        hipFree(gsum);
        hipFree(gpart);
        gcol = grow = gsum = gpart = nullptr;
        ground_valid = false;
        W32 = nullptr;
        dscale = nullptr;
        x = r = b = nullptr;
        t32 = nullptr;
        norms = hnorms = nullptr;
        hflags = nullptr;
        active = hactive = nullptr;
        cap = 0;
        vec_cap = w32_cap = 0;
        item_cap = 0;
    }
    ~Solver() { free_vecs(); }

    const LUPlan* plan = nullptr;  // nullptr: dense partial pivoting
    const LUPlan* plan32 = nullptr;  // plan of the mixed form's factor (nullptr: `plan`)
    // a second complex64 plan tried, on the intact W, when the refinement with
    // plan32 fails, before the FP64 factor (nullptr: none)
    const LUPlan* plan32_retry = nullptr;
    bool equilibrate = false;      // symmetric row-max scaling before the FP32 factor
                                   // (not applied with grounding)

    double norm_at(int z, int k, int i) const
    {
        double v;
        std::memcpy(&v, &hnorms[kNormStride * z + kNS * k + i], 8);
        return v;
    }
    double norm_at(int k, int i) const { return norm_at(0, k, i); }
    VecBatch vb(bool masked) const
    {
        VecBatch v;
        v.nb = nb;
        v.sw = sW;
        v.sw32 = (long)n * n;
        v.sn = kNormStride;
        v.active = masked ? active : nullptr;
        v.hem = hem;
        return v;
    }

    // Plain form, first half: factor the nb items of W (destroyed) in FP64, solve
    // into w64.vec and copy x to `bvec`, all queued on the stream.
    int issue64(c64* W, c64* bvec, int nb_ = 1, long sW_ = 0)
    {
        pending_mixed = false;
        if (nb_ < 1 || nb_ > w64.batch) return TAGS_ERR_ARG;
        nb = nb_;
        sW = nb_ > 1 ? sW_ : 0;
        int rc = lu_factor<double>(w64, W, n, st, pr, plan, nb, sW);
        if (rc) return rc;
        TAGS_HIP(hipMemcpyAsync(hflags, w64.flags, sizeof(int) * 4 * nb, hipMemcpyDeviceToHost,
                                st));
        if ((rc = lu_gather_rhs<double>(w64, bvec, w64.vec, st))) return rc;
        if ((rc = lu_solve_inplace<double>(w64, W, n, w64.vec, st, pr))) return rc;
        TAGS_HIP(hipMemcpyAsync(bvec, w64.vec, sizeof(c64) * n * nb, hipMemcpyDeviceToDevice,
                                st));
        return 0;
    }
    // Second half: waits; stats (nullable) and info (nullable) receive each
    // item's; returns 0, an error, or for one item its info (> 0: x not valid).
    int finish64(tags_item_stats* stt, int* info = nullptr)
    {
        TAGS_HIP(hipStreamSynchronize(st));
        int first = 0;
        for (int z = 0; z < nb; z++) {
            if (hflags[4 * z + 1]) return TAGS_ERR_BARRIER;
            const int inf = hflags[4 * z];
            if (info) info[z] = inf;
            if (!first) first = inf;
            if (stt) {
                stt[z].precision_used = TAGS_PREC_FP64;
                stt[z].refine_iters = 0;
                stt[z].info = inf;
                stt[z].backward_err = -1.0;  // W was factored in place (a sweep can verify)
                stt[z].forward_est = -1.0;
            }
        }
        return nb == 1 ? first : 0;
    }
    int solve64(c64* W, c64* bvec, tags_item_stats* stt)
    {
        int rc = issue64(W, bvec);
        return rc ? rc : finish64(stt);
    }

    // One refinement step k >= 1 queued for the items (the active ones when
    // masked): correction from the complex64 factor with the right-hand side
    // scaled by |r_{k-1}| (read on the device), x += correction (its norm into
    // slot k), r_k = b - W x and its norms into slot k (zeroed when the solve
    // was issued).
    int queue_step(c64* W, int k, const Grounding* gr, bool masked)
    {
        int rc;
        const VecBatch v = vb(masked);
        if (gmres_k > 0) {
            // the correction from FGMRES on W d = r, then r = b - W x
            if ((rc = gmres_step(gm, gmres_k, nb, w32, W32, t32, dscale, W, sW, r, x,
                                 norms + kNS * k + 3, kNormStride, masked ? active : nullptr, st,
                                 pr, hem)))
                return rc;
            return residual64(W, n, n, x, b, r, norms + kNS * k, false, st, pr, false, v);
        }
        const unsigned long long* rn = norms + kNS * (k - 1);
        if ((rc = prep_rhs32(r, dscale, w32.perm, n, 1.0, t32, st, gr,
                             gr ? gsum + 1 : nullptr, rn, v)))
            return rc;
        if ((rc = lu_solve_inplace<float>(w32, W32, n, t32, st, pr))) return rc;
        if ((rc = update_x(x, t32, dscale, n, 1.0, st, gr, rn, norms + kNS * k + 3, v)))
            return rc;
        return residual64(W, n, n, x, b, r, norms + kNS * k, false, st, pr, false, v);
    }

    // Mixed form, first half: W (FP64, kept intact unless the FP64 fallback
    // runs) to complex64 (grounded or scaled), its factor, r_0 = b and the
    // first kSpecIters refinement steps, all queued on the stream without a
    // host wait; `bvec` holds b.
    int issue_mixed(c64* W, c64* bvec, int nb_ = 1, long sW_ = 0)
    {
        int rc;
        pending_mixed = true;
        if (nb_ < 1 || nb_ > w32.batch) return TAGS_ERR_ARG;
        nb = nb_;
        sW = nb_ > 1 ? sW_ : 0;
        const bool grounded = ground.g >= 0;
        if (grounded && nb > 1) return TAGS_ERR_ARG;
        if (grounded) gmres_k = 0;
        if (gmres_k > 0 && (rc = gm.alloc(n, gmres_k, cap))) return rc;
        {
            Timed tm(pr, FAM_OTHER, st);
            if (grounded) {
                if (!gcol) {
                    TAGS_HIP(hipMalloc(&gcol, sizeof(c64) * n));
                    TAGS_HIP(hipMalloc(&grow, sizeof(c64) * n));
                    TAGS_HIP(hipMalloc(&gsum, sizeof(c64) * 2));
                    TAGS_HIP(hipMalloc(&gpart, sizeof(c64) * kGroundChunks * (size_t)n));
                    ground_valid = false;
                }
                if (!ground_valid) {
                    if ((rc = ground_sums(W, n, n, ground, gcol, grow, gsum, gpart, st)))
                        return rc;
                    ground_valid = true;
                }
                if ((rc = row_scale(W, n, n, dscale, false, st))) return rc;
                if ((rc = to_c32_grounded(W, n, n, ground, gcol, grow, gsum, W32, n, st)))
                    return rc;
            } else {
                if ((rc = row_scale(W, n, n, dscale, equilibrate, st, vb(false)))) return rc;
                if ((rc = to_c32_scaled(W, n, n, dscale, W32, n, st, vb(false)))) return rc;
            }
            tm.end();
        }
        if ((rc = lu_factor<float>(w32, W32, n, st, pr, plan32 ? plan32 : plan, nb,
                                   (long)n * n)))
            return rc;
        TAGS_HIP(hipMemcpyAsync(hflags, w32.flags, sizeof(int) * 4 * nb, hipMemcpyDeviceToHost,
                                st));
        TAGS_HIP(hipMemcpyAsync(b, bvec, sizeof(c64) * n * nb, hipMemcpyDeviceToDevice, st));
        if ((rc = fill_c64(x, (long)n * nb, c64{0.0, 0.0}, st))) return rc;
        // every step's norms start at zero (max-reductions)
        if ((rc = fill_c64((c64*)norms, kNormStride * nb / 2, c64{0.0, 0.0}, st))) return rc;
        // r_0 = b (x = 0), with |W|
        if ((rc = residual64(W, n, n, x, b, r, norms, true, st, pr, false, vb(false)))) return rc;
        const Grounding* gr = grounded ? &ground : nullptr;
        const int spec = gmres_k > 0 ? kGmresSpec : kSpecIters;
        for (int k = 1; k <= spec; k++)
            if ((rc = queue_step(W, k, gr, false))) return rc;
        queued = spec;
        m_first = true;
        TAGS_HIP(hipMemcpyAsync(hnorms, norms, sizeof(unsigned long long) * kNormStride * nb,
                                hipMemcpyDeviceToHost, st));
        return 0;
    }

    // Second half, in steps that never wait: advance_mixed is called when the
    // stream has finished what was queued (the caller syncs or polls). It
    // checks each item's steps so far by the backward error |b - W x| / (|W|
    // |x|) and the forward-error estimate |dx| / |x|; while the stop rule (see
    // kMinGain) does not hold for some items it queues one more step for
    // them (masked to those items) and returns with *complete false; then it
    // accepts an item's last iterate at a backward error <= kAccept and
    // |dx| / |x| <= kFwdAccept. An item that fails, or whose complex64 factor
    // broke down, starts again alone with plan32_retry's complex64 factor if
    // one is set, else the FP64 factor of its intact W gives its answer (these
    // wait). Solutions in `bvec`; stt (nullable) and info (nullable) receive
    // each item's. Returns 0, an error, or for one item its info.
    std::vector<int> m_info32, m_done, m_go;
    std::vector<double> m_berr, m_fwd;
    int m_step = 0;
    bool m_first = true;
    c64* m_W = nullptr;
    c64* m_bvec = nullptr;
    double berr_at(int z, int k) const
    {
        const double wnorm = norm_at(z, 0, 2);
        const double rn = norm_at(z, k, 0), xn = norm_at(z, k, 1);
        if (rn == 0.0) return 0.0;
        return (wnorm > 0 && xn > 0) ? rn / (wnorm * xn) : 1.0;
    }
    // |dx_k| / |x_k|
    double dx_at(int z, int k) const
    {
        const double dn = norm_at(z, k, 3), xn = norm_at(z, k, 1);
        if (dn == 0.0) return 0.0;
        return xn > 0 ? dn / xn : 1.0;
    }
    // estimated error of x_k relative to |x_k| (see kFwdAccept). FGMRES steps
    // take the correction itself: their contraction varies from one restart
    // to the next (a fast first step says little about the second), so the
    // last step's ratio does not bound the error left after it.
    double fwd_at(int z, int k) const
    {
        const double d = dx_at(z, k);
        if (k < 2 || d == 0.0 || gmres_k > 0) return d;
        const double dp = dx_at(z, k - 1);
        const double rho = dp > 0.0 ? d / dp : 1.0;
        return rho < 0.5 ? d * rho / (1.0 - rho) : d;
    }
    bool keep_going(int z) const
    {
        const int d = m_done[z];
        const double be = m_berr[z], fe = m_fwd[z];
        if (m_info32[z] != 0 || !(be == be) || !(fe == fe) ||
            d >= (gmres_k > 0 ? kGmresMaxIters : kMaxIters))
            return false;
        if (be <= kAccept && fe <= kFwdAccept) return false;
        return d < 2 || be <= kMinGain * berr_at(z, d - 1) || dx_at(z, d) <= kMinGain * dx_at(z, d - 1);
    }
    int advance_mixed(c64* W, c64* bvec, tags_item_stats* stt, int* info, bool* complete)
    {
        *complete = false;
        const int nbi = nb;
        const Grounding* gr = ground.g >= 0 ? &ground : nullptr;
        int ngo = 0;
        if (m_first) {
            // the queued steps
            m_first = false;
            m_W = W;
            m_bvec = bvec;
            m_info32.assign(nbi, 0);
            m_done.assign(nbi, queued);
            m_go.assign(nbi, 0);
            m_berr.assign(nbi, 0.0);
            m_fwd.assign(nbi, 0.0);
            m_step = queued;
            for (int z = 0; z < nbi; z++) {
                if (hflags[4 * z + 1]) return TAGS_ERR_BARRIER;
                m_info32[z] = hflags[4 * z];
            }
            for (int z = 0; z < nbi; z++) {
                if (stt) {
                    for (int i = 0; i < 8; i++) stt[z].refine_hist[i] = -1.0;
                    for (int k = 1; k <= queued && k <= 8; k++)
                        stt[z].refine_hist[k - 1] = berr_at(z, k);
                }
                m_berr[z] = berr_at(z, queued);
                m_fwd[z] = fwd_at(z, queued);
                m_go[z] = keep_going(z);
                ngo += m_go[z];
            }
        } else {
            // the step queued by the last call, for the items still refining
            m_step++;
            for (int z = 0; z < nbi; z++) {
                if (!m_go[z]) continue;
                m_done[z] = m_step;
                m_berr[z] = berr_at(z, m_step);
                m_fwd[z] = fwd_at(z, m_step);
                if (stt && m_step <= 8) stt[z].refine_hist[m_step - 1] = m_berr[z];
                m_go[z] = keep_going(z);
                ngo += m_go[z];
            }
        }
        if (ngo > 0) {
            // one more step, batch-wide but masked to the items still refining
            // (all of which have taken the same number of steps)
            if (nbi > 1) {
                for (int z = 0; z < nbi; z++) hactive[z] = m_go[z];
                TAGS_HIP(hipMemcpyAsync(active, hactive, sizeof(int) * nbi,
                                        hipMemcpyHostToDevice, st));
            }
            int rc = queue_step(W, m_step + 1, gr, nbi > 1);
            if (rc) return rc;
            for (int z = 0; z < nbi; z++)
                if (m_go[z])
                    TAGS_HIP(hipMemcpyAsync(hnorms + kNormStride * z + kNS * (m_step + 1),
                                            norms + kNormStride * z + kNS * (m_step + 1),
                                            sizeof(unsigned long long) * kNS,
                                            hipMemcpyDeviceToHost, st));
            return 0;
        }
        pending_mixed = false;
        const long sWi = sW;
        const std::vector<int> done = m_done;
        TAGS_HIP(hipMemcpyAsync(bvec, x, sizeof(c64) * n * nbi, hipMemcpyDeviceToDevice, st));
        std::vector<int> failed;
        for (int z = 0; z < nbi; z++) {
            const bool ok = m_info32[z] == 0 && m_berr[z] == m_berr[z] && m_berr[z] <= kAccept &&
                            m_fwd[z] <= kFwdAccept;
            if (ok) {
                if (stt) {
                    stt[z].precision_used = TAGS_PREC_MIXED;
                    stt[z].refine_iters = done[z];
                    stt[z].info = 0;
                    stt[z].backward_err = m_berr[z];
                    stt[z].forward_est = m_fwd[z];
                }
                if (info) info[z] = 0;
            } else {
                failed.push_back(z);
            }
        }
        int first = 0;
        // items that failed, one at a time on the intact W of each
        const LUPlan* retry = plan32_retry;
        const LUPlan* p32 = plan32;
        for (int z : failed) {
            c64* Wz = W + (long)z * sWi;
            c64* bz = bvec + (long)z * n;
            tags_item_stats one{};
            tags_item_stats* sz = stt ? &stt[z] : &one;
            int rc;
            // b_z again (bvec holds the failed iterate)
            TAGS_HIP(hipMemcpyAsync(bz, b + (long)z * n, sizeof(c64) * n, hipMemcpyDeviceToDevice,
                                    st));
            if (retry) {
                // the second complex64 plan, with classic refinement; its steps
                // add to this one's
                const int gk = gmres_k;
                plan32 = retry;
                plan32_retry = nullptr;
                gmres_k = 0;
                rc = issue_mixed(Wz, bz);
                if (!rc) rc = finish_mixed(Wz, bz, sz);
                plan32 = p32;
                plan32_retry = retry;
                gmres_k = gk;
                if (rc >= 0) sz->refine_iters += done[z];
            } else {
                // FP64 fallback on the intact W
                rc = issue64(Wz, bz);
                if (!rc) rc = finish64(sz);
                if (rc >= 0) sz->refine_iters = done[z];
            }
            if (rc < 0) return rc;
            if (info) info[z] = rc;
            if (!first) first = rc;
            TAGS_HIP(hipStreamSynchronize(st));
        }
        nb = nbi;
        sW = sWi;
        *complete = true;
        return nbi == 1 ? first : 0;
    }
    // The whole second half, waiting for each step.
    int finish_mixed(c64* W, c64* bvec, tags_item_stats* stt, int* info = nullptr)
    {
        bool complete = false;
        int rc = 0;
        while (!complete) {
            TAGS_HIP(hipStreamSynchronize(st));
            if ((rc = advance_mixed(W, bvec, stt, info, &complete)) < 0) return rc;
        }
        return rc;
    }
    int solve_mixed(c64* W, c64* bvec, tags_item_stats* stt)
    {
        int rc = issue_mixed(W, bvec);
        return rc ? rc : finish_mixed(W, bvec, stt);
    }
};

}  // namespace tags

using namespace tags;

extern "C" {

const char* tags_version(void) { return "tags_hip 0.1 (gfx950)"; }

int tags_device_synchronize(void)
{
    TAGS_HIP(hipDeviceSynchronize());
    return 0;
}

}  // extern "C"

namespace tags {
__global__ void k_fill_pattern(double* p, long n, unsigned seed)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        unsigned h = (unsigned)i * 2654435761u ^ seed;
        h ^= h >> 13;
        h *= 0x5bd1e995u;
        h ^= h >> 15;
        p[i] = (double)(h & 0xffff) / 65536.0 - 0.5;
    }
}
__global__ void k_to_f32(const double* a, float* b, long n)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// The split-bf16 complex64 variants pack their operands into a work space:
// give a stand-alone product one of its own.
template <typename T>
static int gemm_ws_attach(GemmArgs<T>& g, DevBuf<unsigned char>& ws)
{
    if (sizeof(T) != sizeof(float)) return 0;
    const long bytes = gemm_x6_ws_bytes(g.M, g.N, g.K);
    if (bytes <= 0) return 0;
    int rc;
    if ((rc = ws.alloc((size_t)bytes))) return rc;
    g.ws = ws.p;
    g.ws_item = bytes;
    return 0;
}

// Times C -= A B (M x N x K, row-major, a random row map on A and C like the
// LU trailing update) for one variant; returns the mean kernel time in ms.
template <typename T>
static int gemm_bench_t(int M, int N, int K, int variant, int iters, double* ms)
{
    Stream S;
    const long na = (long)M * K, nb = (long)K * N, nc = (long)M * N;
    DevBuf<double> tmp;
    DevBuf<cx<T>> A, B, C;
    DevBuf<int> rows, krows;
    int rc;
    if ((rc = tmp.alloc(2 * std::max(na, std::max(nb, nc)))) || (rc = A.alloc(na)) ||
        (rc = B.alloc(nb)) || (rc = C.alloc(nc)) || (rc = rows.alloc(M)) ||
        (rc = krows.alloc(K)))
        return rc;
    cx<T>* bufs[3] = {A.p, B.p, C.p};
    const long sz[3] = {na, nb, nc};
    for (int b = 0; b < 3; b++) {
        hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)((2 * sz[b] + 255) / 256)), dim3(256), 0,
                           S.s, tmp.p, 2 * sz[b], 17u + b);
        if (sizeof(T) == sizeof(double)) {
            TAGS_HIP(hipMemcpyAsync(bufs[b], tmp.p, sizeof(double) * 2 * sz[b],
                                    hipMemcpyDeviceToDevice, S.s));
        } else {
            hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((2 * sz[b] + 255) / 256)), dim3(256), 0,
                               S.s, tmp.p, (float*)bufs[b], 2 * sz[b]);
        }
    }
    std::vector<int> h(M);
    for (int i = 0; i < M; i++) h[i] = (int)(((long)i * 7919) % M);
    bool perm_ok = (M % 7919) != 0;
    if (!perm_ok)
        for (int i = 0; i < M; i++) h[i] = i;
    TAGS_HIP(hipMemcpyAsync(rows.p, h.data(), sizeof(int) * M, hipMemcpyHostToDevice, S.s));
    // B's rows through a reversal of 0 .. K-1 (the LU's U12 rows are addressed
    // through the row map too)
    std::vector<int> hk(K);
    for (int i = 0; i < K; i++) hk[i] = K - 1 - i;
    TAGS_HIP(hipMemcpyAsync(krows.p, hk.data(), sizeof(int) * K, hipMemcpyHostToDevice, S.s));
    GemmArgs<T> g{M, N, K, A.p, K, rows.p, B.p, N, krows.p, C.p, N, rows.p, 1};
    DevBuf<unsigned char> ws;
    if ((rc = gemm_ws_attach(g, ws))) return rc;
    if ((rc = launch_gemm_variant<T>(g, variant, S.s))) return rc;  // warm-up
    hipEvent_t e0, e1;
    TAGS_HIP(hipEventCreate(&e0));
    TAGS_HIP(hipEventCreate(&e1));
    TAGS_HIP(hipEventRecord(e0, S.s));
    for (int i = 0; i < iters; i++)
        if ((rc = launch_gemm_variant<T>(g, variant, S.s))) return rc;
    TAGS_HIP(hipEventRecord(e1, S.s));
    TAGS_HIP(hipEventSynchronize(e1));
    float t = 0;
    TAGS_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / iters;
    return 0;
}
// C (M x N) = A B or C -= A B with host operands (row-major, lda = K, ldb = ldc
// = N) and optional row maps, through one GEMM variant.
template <typename T>
static int gemm_run_t(int M, int N, int K, int variant, const c64* A, const c64* B, c64* C,
                      const int* arow, const int* brow, const int* crow, int mode)
{
    auto bad_map = [](const int* r, int n) {
        if (!r) return false;
        for (int i = 0; i < n; i++)
            if (r[i] < 0 || r[i] >= n) return true;
        return false;
    };
    if (bad_map(arow, M) || bad_map(brow, K) || bad_map(crow, M)) return TAGS_ERR_ARG;
    const size_t na = (size_t)M * K, nb = (size_t)K * N, nc = (size_t)M * N;
    std::vector<cx<T>> ha(na), hb(nb), hc(nc);
    for (size_t i = 0; i < na; i++) ha[i] = cx<T>{(T)A[i].x, (T)A[i].y};
    for (size_t i = 0; i < nb; i++) hb[i] = cx<T>{(T)B[i].x, (T)B[i].y};
    for (size_t i = 0; i < nc; i++) hc[i] = cx<T>{(T)C[i].x, (T)C[i].y};
    Stream S;
    DevBuf<cx<T>> dA, dB, dC;
    DevBuf<int> m1, m2, m3;
    int rc;
    if ((rc = dA.alloc(na)) || (rc = dB.alloc(nb)) || (rc = dC.alloc(nc)) ||
        (rc = m1.alloc(M)) || (rc = m2.alloc(K)) || (rc = m3.alloc(M)))
        return rc;
    TAGS_HIP(hipMemcpyAsync(dA.p, ha.data(), sizeof(cx<T>) * na, hipMemcpyHostToDevice, S.s));
    TAGS_HIP(hipMemcpyAsync(dB.p, hb.data(), sizeof(cx<T>) * nb, hipMemcpyHostToDevice, S.s));
    TAGS_HIP(hipMemcpyAsync(dC.p, hc.data(), sizeof(cx<T>) * nc, hipMemcpyHostToDevice, S.s));
    if (arow) TAGS_HIP(hipMemcpyAsync(m1.p, arow, sizeof(int) * M, hipMemcpyHostToDevice, S.s));
    if (brow) TAGS_HIP(hipMemcpyAsync(m2.p, brow, sizeof(int) * K, hipMemcpyHostToDevice, S.s));
    if (crow) TAGS_HIP(hipMemcpyAsync(m3.p, crow, sizeof(int) * M, hipMemcpyHostToDevice, S.s));
    GemmArgs<T> g{M, N, K, dA.p, K, arow ? m1.p : nullptr, dB.p, N, brow ? m2.p : nullptr,
                  dC.p, N, crow ? m3.p : nullptr, mode ? 1 : 0};
    DevBuf<unsigned char> ws;
    if ((rc = gemm_ws_attach(g, ws))) return rc;
    if ((rc = launch_gemm_variant<T>(g, variant, S.s))) return rc;
    TAGS_HIP(hipMemcpyAsync(hc.data(), dC.p, sizeof(cx<T>) * nc, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    for (size_t i = 0; i < nc; i++) C[i] = c64{(double)hc[i].x, (double)hc[i].y};
    return 0;
}
}  // namespace tags

extern "C" {

int tags_gemm_bench(int M, int N, int K, int precision, int variant, int iters, double* ms)
{
    if (M <= 0 || N <= 0 || K <= 0 || iters <= 0 || !ms) return TAGS_ERR_ARG;
    return precision == TAGS_PREC_MIXED ? gemm_bench_t<float>(M, N, K, variant, iters, ms)
                                        : gemm_bench_t<double>(M, N, K, variant, iters, ms);
}

int tags_gemm_variants(int precision) { return gemm_variant_count(precision); }

int tags_gemm_run(int M, int N, int K, int precision, int variant, const TAGS_COMPLEX* A,
                  const TAGS_COMPLEX* B, TAGS_COMPLEX* C, const int* arow, const int* brow,
                  const int* crow, int mode)
{
    if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return TAGS_ERR_ARG;
    return precision == TAGS_PREC_MIXED
               ? gemm_run_t<float>(M, N, K, variant, (const c64*)A, (const c64*)B, (c64*)C, arow,
                                   brow, crow, mode)
               : gemm_run_t<double>(M, N, K, variant, (const c64*)A, (const c64*)B, (c64*)C, arow,
                                    brow, crow, mode);
}

int tags_set_panel(int kind)
{
    if (kind != TAGS_PANEL_TOURNAMENT && kind != TAGS_PANEL_PARTIAL) return TAGS_ERR_ARG;
    lu_set_panel_kind(kind);
    return 0;
}

int tags_set_lookahead(int on)
{
    lu_set_lookahead(on != 0);
    return 0;
}

int tags_set_update_chunks(int n)
{
    if (n < 1) return TAGS_ERR_ARG;
    lu_set_update_chunks(n);
    return 0;
}

int tags_set_step_width(int columns)
{
    if (columns != 0 && columns != 128 && columns != 256) return TAGS_ERR_ARG;
    lu_set_step_width(columns);
    return 0;
}

int tags_set_solve_rows(int rows)
{
    if (rows != 0 && rows != 1 && rows != 2) return TAGS_ERR_ARG;
    lu_set_solve_rows(rows);
    return 0;
}

int tags_gemm_select(int precision, int big_variant, int small_variant)
{
    return gemm_set_variants(precision, big_variant, small_variant);
}

int tags_autotune(int precision, int* big_out, int* small_out)
{
    // large: the first trailing updates of the 21,421 system; small: the tail;
    // K: the LU's step width (128 for complex128 factors, 256 for complex64)
    const int nv = gemm_variant_count(precision);
    int best[2] = {0, 0};
    const int K = precision == TAGS_PREC_MIXED ? 256 : 128;
    const int sizes[2][3] = {{8192, 8192, K}, {1024, 1024, K}};
    for (int s = 0; s < 2; s++) {
        double bt = 1e30;
        for (int v = 0; v < nv; v++) {
            double ms = 0;
            int rc = tags_gemm_bench(sizes[s][0], sizes[s][1], sizes[s][2], precision, v,
                                     s == 0 ? 3 : 20, &ms);
            if (rc) return rc;
            if (ms < bt) {
                bt = ms;
                best[s] = v;
            }
        }
    }
    int rc = gemm_set_variants(precision, best[0], best[1]);
    if (big_out) *big_out = best[0];
    if (small_out) *small_out = best[1];
    return rc;
}

// ---------------------------------------------------------------------------
// geometry (host code), src/electrode.c:12-32, 103-127; src/grid.c:9-91
// ---------------------------------------------------------------------------
static double vlen(const double a[3], const double b[3])
{
    double r = 0.0;
    for (int i = 0; i < 3; i++) r += std::pow(a[i] - b[i], 2.0);
    return std::sqrt(r);
}

int populate_electrode(Electrode* e, const double start[3], const double end[3], double radius)
{
    bool eq = true;
    for (int i = 0; i < 3; i++)
        if (std::fabs(start[i] - end[i]) > DBL_EPSILON) eq = false;
    if (eq) {
        std::printf("Error: start_point and end_point are equal.\n");
        return 1;
    }
    if (radius <= 0) {
        std::printf("Error: radius < 0.\n");
        return 2;
    }
    e->radius = radius;
    e->length = vlen(start, end);
    for (int i = 0; i < 3; i++) {
        e->start_point[i] = start[i];
        e->end_point[i] = end[i];
        e->middle_point[i] = (start[i] + end[i]) / 2.0;
    }
    return 0;
}

static void lin(double a, double b, size_t n, double* u)
{
    const double c = (b - a) / (n - 1);
    for (size_t i = 0; i < n - 1; ++i) u[i] = a + i * c;
    u[n - 1] = b;
}

int segment_electrode(Electrode* e, double* nodes, size_t num_segments, const double* start,
                      const double* end, double radius)
{
    if (num_segments < 1) {
        std::printf("Error: number of segments should be greater than 0.\n");
        return 1;
    }
    const size_t nn = num_segments + 1;
    std::vector<double> x(nn), y(nn), z(nn);
    lin(start[0], end[0], nn, x.data());
    lin(start[1], end[1], nn, y.data());
    lin(start[2], end[2], nn, z.data());
    for (size_t i = 0; i < nn; i++) {
        nodes[3 * i] = x[i];
        nodes[3 * i + 1] = y[i];
        nodes[3 * i + 2] = z[i];
    }
    for (size_t i = 0; i < num_segments; i++)
        populate_electrode(&e[i], nodes + 3 * i, nodes + 3 * i + 3, radius);
    return 0;
}

int number_segments(const Grid g)
{
    const unsigned N = g.edge_segments_x, vx = g.vertices_x;
    const unsigned M = g.edge_segments_y, vy = g.vertices_y;
    return (int)(N * vy * (vx - 1) + M * vx * (vy - 1));
}

int number_nodes(const Grid g)
{
    const unsigned N = g.edge_segments_x, vx = g.vertices_x;
    const unsigned M = g.edge_segments_y, vy = g.vertices_y;
    return (int)(vx * vy + vx * (vy - 1) * (M - 1) + vy * (vx - 1) * (N - 1));
}

int electrode_grid(const Grid g, Electrode* e, double* nodes)
{
    const size_t N = g.edge_segments_x, vx = g.vertices_x;
    const size_t M = g.edge_segments_y, vy = g.vertices_y;
    const double lx = g.length_x / (N * (vx - 1));
    const double ly = g.length_y / (M * (vy - 1));
    double sp[3] = {0.0, 0.0, g.depth}, ep[3] = {0.0, 0.0, g.depth};
    size_t ed = 0, nd = 0;
    for (size_t h = 0; h < vy; h++)
        for (size_t n = 0; n < vx - 1; n++)
            for (size_t k = 0; k < N; k++) {
                sp[0] = lx * (n * N + k);
                sp[1] = ly * M * h;
                ep[0] = sp[0] + lx;
                ep[1] = sp[1];
                populate_electrode(&e[ed++], sp, ep, g.radius);
                if (n == 0 && k == 0) {
                    std::memcpy(nodes + 3 * nd, sp, sizeof(sp));
                    nd++;
                }
                std::memcpy(nodes + 3 * nd, ep, sizeof(ep));
                nd++;
            }
    for (size_t gg = 0; gg < vx; gg++)
        for (size_t mm = 0; mm < vy - 1; mm++)
            for (size_t k = 0; k < M; k++) {
                sp[0] = lx * N * gg;
                sp[1] = ly * (mm * M + k);
                ep[0] = sp[0];
                ep[1] = sp[1] + ly;
                populate_electrode(&e[ed++], sp, ep, g.radius);
                if (k < M - 1) {
                    std::memcpy(nodes + 3 * nd, ep, sizeof(ep));
                    nd++;
                }
            }
    return 0;
}

// ---------------------------------------------------------------------------
// integrals
// ---------------------------------------------------------------------------
TAGS_COMPLEX self_integral(const Electrode* sender)
{
    const double ls = sender->length;
    const double k1 = sender->radius / ls;
    const double k2 = std::sqrt(1.0 + k1 * k1);
    return to_c(c64{2.0 * ls * (std::log((k2 + 1.0) / k1) - k2 + k1), 0.0});
}

int integral(const Electrode* sender, const Electrode* receiver, TAGS_COMPLEX gamma,
             size_t max_eval, double req_abs_error, double req_rel_error, int integration_type,
             double result[2], double error[2])
{
    (void)max_eval;
    (void)req_abs_error;
    (void)req_rel_error;
    result[0] = result[1] = error[0] = error[1] = 0.0;
    if (integration_type == INTG_NONE) {
        result[0] = vlen(sender->middle_point, receiver->middle_point);
        return 0;
    }
    if (integration_type != INTG_DOUBLE && integration_type != INTG_MHEM &&
        integration_type != INTG_SINGLE) {
        std::printf("integration failed with code: %i\n", -10);
        return -10;
    }
    const bool mhem = integration_type == INTG_MHEM;
    Stream S;
    DevSegs a, b;
    int rc;
    if ((rc = a.upload(sender, 1, S.s))) return rc;
    if ((rc = b.upload(receiver, 1, S.s))) return rc;
    DevBuf<double> out;
    if ((rc = out.alloc(4))) return rc;
    if ((rc = launch_single_pair(a.soa(), b.soa(), mhem ? c64{0.0, 0.0} : from_c(gamma),
                                 mhem ? INTG_DOUBLE : integration_type, out.p, S.s)))
        return rc;
    double h[4];
    TAGS_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    if (mhem) {
        // the gamma = 0 double integral (real); the reference scales only
        // result[0] by the receiver length, its error stays per unit length
        result[0] = h[0];
        error[0] = receiver->length > 0 ? h[2] / receiver->length : h[2];
        return 0;
    }
    result[0] = h[0];
    result[1] = h[1];
    error[0] = h[2];
    error[1] = h[3];
    return 0;
}

static int pairs_dropin(TAGS_COMPLEX* zl, TAGS_COMPLEX* zt, const Electrode* e,
                        const Electrode* img, size_t m, const PairScalars& sc, bool images)
{
    if (m == 0) return 0;
    if (!zl || !zt || !e) return TAGS_ERR_ARG;
    Stream S;
    DevSegs de, di;
    int rc;
    if ((rc = de.upload(e, (long)m, S.s))) return rc;
    if ((rc = di.upload(img ? img : e, (long)m, S.s))) return rc;
    const bool alias = (void*)zl == (void*)zt;
    const size_t bytes = sizeof(c64) * m * m;
    DevBuf<c64> dl, dt;
    if ((rc = dl.alloc(m * m))) return rc;
    TAGS_HIP(hipMemcpyAsync(dl.p, zl, bytes, hipMemcpyHostToDevice, S.s));
    c64* dtp = dl.p;
    if (!alias) {
        if ((rc = dt.alloc(m * m))) return rc;
        TAGS_HIP(hipMemcpyAsync(dt.p, zt, bytes, hipMemcpyHostToDevice, S.s));
        dtp = dt.p;
    }
    if ((rc = launch_pairs_dropin((double*)dl.p, (double*)dtp, alias, de.soa(), di.soa(),
                                  (long)m, sc, images, S.s)))
        return rc;
    TAGS_HIP(hipMemcpyAsync(zl, dl.p, bytes, hipMemcpyDeviceToHost, S.s));
    if (!alias) TAGS_HIP(hipMemcpyAsync(zt, dtp, bytes, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    return 0;
}

int calculate_impedances(TAGS_COMPLEX* zl, TAGS_COMPLEX* zt, const Electrode* electrodes,
                         size_t num_electrodes, TAGS_COMPLEX gamma, TAGS_COMPLEX s, double mur,
                         TAGS_COMPLEX kappa, size_t max_eval, double req_abs_error,
                         double req_rel_error, int integration_type)
{
    (void)max_eval;
    (void)req_abs_error;
    (void)req_rel_error;
    if (integration_type != INTG_DOUBLE && integration_type != INTG_NONE &&
        integration_type != INTG_MHEM && integration_type != INTG_SINGLE) {
        std::printf("integration failed with code: %i\n", -10);
        return -10;
    }
    PairScalars sc = make_scalars(from_c(gamma), from_c(s), mur, from_c(kappa), c64{1, 0},
                                  c64{1, 0}, integration_type);
    return pairs_dropin(zl, zt, electrodes, nullptr, num_electrodes, sc, false);
}

int impedances_images(TAGS_COMPLEX* zl, TAGS_COMPLEX* zt, const Electrode* electrodes,
                      const Electrode* images, size_t num_electrodes, TAGS_COMPLEX gamma,
                      TAGS_COMPLEX s, double mur, TAGS_COMPLEX kappa, TAGS_COMPLEX ref_l,
                      TAGS_COMPLEX ref_t, size_t max_eval, double req_abs_error,
                      double req_rel_error, int integration_type)
{
    (void)max_eval;
    (void)req_abs_error;
    (void)req_rel_error;
    if (integration_type != INTG_DOUBLE && integration_type != INTG_NONE &&
        integration_type != INTG_MHEM && integration_type != INTG_SINGLE) {
        std::printf("integration failed with code: %i\n", -10);
        return -10;
    }
    if (!images) return TAGS_ERR_ARG;
    PairScalars sc = make_scalars(from_c(gamma), from_c(s), mur, from_c(kappa), from_c(ref_l),
                                  from_c(ref_t), integration_type);
    return pairs_dropin(zl, zt, electrodes, images, num_electrodes, sc, true);
}

// ---------------------------------------------------------------------------
// fill_incidence_imm (src/linalg.c:12-86)
// ---------------------------------------------------------------------------
}  // extern "C"

namespace tags {
// blk1: rows n..N of columns 0..nw (column-major, ld 2m): A (rows s) then B
// (rows m+s); blk2: rows 0..nw of columns n..N (column-major, ld nw).
// counts[0] = #start matches, counts[1] = #end matches (over nodes < nw).
__global__ __launch_bounds__(256) void k_incidence_blocks(const double* __restrict__ sp, long m,
                                                          const double* __restrict__ nodes,
                                                          long nw, c64* blk1, c64* blk2,
                                                          unsigned long long* counts)
{
    const double tol = 1e-5;
    unsigned long long cs = 0, ce = 0;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < nw * m;
         idx += (long)gridDim.x * 256) {
        const long nd = idx / m, s = idx % m;
        const double x = nodes[3 * nd], y = nodes[3 * nd + 1], z = nodes[3 * nd + 2];
        const bool st = fabs(sp[6 * s] - x) <= tol && fabs(sp[6 * s + 1] - y) <= tol &&
                        fabs(sp[6 * s + 2] - z) <= tol;
        const bool ed = !st && fabs(sp[6 * s + 3] - x) <= tol && fabs(sp[6 * s + 4] - y) <= tol &&
                        fabs(sp[6 * s + 5] - z) <= tol;
        const double a = st ? 1.0 : (ed ? -1.0 : 0.0);
        const double b = (st || ed) ? 0.5 : 0.0;
        blk1[s + 2 * m * nd] = c64{a, 0.0};
        blk1[m + s + 2 * m * nd] = c64{b, 0.0};
        blk2[nd + nw * s] = c64{a, 0.0};
        blk2[nd + nw * (m + s)] = c64{b, 0.0};
        cs += st;
        ce += ed;
    }
    for (int off = 32; off > 0; off >>= 1) {
        cs += __shfl_xor(cs, off);
        ce += __shfl_xor(ce, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&counts[0], cs);
        atomicAdd(&counts[1], ce);
    }
}
}  // namespace tags

extern "C" {

static std::vector<double> seg_points(const Electrode* e, size_t m)
{
    std::vector<double> sp(6 * m);
    for (size_t s = 0; s < m; s++)
        for (int c = 0; c < 3; c++) {
            sp[6 * s + c] = e[s].start_point[c];
            sp[6 * s + 3 + c] = e[s].end_point[c];
        }
    return sp;
}

// Node/segment incidence on the GPU, in the reference's node order: nodes up to
// and including the first one no segment touches (nw of them; bad = its index
// or -1), as blk1 (2m x nw column-major: A in rows s, B in rows m + s) and blk2
// (nw x 2m column-major: the transposes); counts = start and end matches.
struct Incidence {
    DevBuf<c64> b1, b2;
    long nw = 0, bad = -1;
    unsigned long long c[2] = {0, 0};
};

static int incidence_blocks(Incidence& inc, const Electrode* electrodes, long m,
                            const double* nodes, long n, hipStream_t st)
{
    int rc;
    std::vector<double> sp = seg_points(electrodes, m);
    DevBuf<double> dsp, dnodes;
    DevBuf<int> dsn, den, dhits;
    if ((rc = dsp.alloc(6 * m)) || (rc = dnodes.alloc(3 * n)) || (rc = dsn.alloc(m)) ||
        (rc = den.alloc(m)) || (rc = dhits.alloc(n + 2 * m)))
        return rc;
    TAGS_HIP(hipMemcpyAsync(dsp.p, sp.data(), sizeof(double) * 6 * m, hipMemcpyHostToDevice, st));
    TAGS_HIP(hipMemcpyAsync(dnodes.p, nodes, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
    TAGS_HIP(hipMemsetAsync(dhits.p, 0, sizeof(int) * (n + 2 * m), st));
    if ((rc = launch_incidence(dsp.p, m, dnodes.p, n, dsn.p, den.p, dhits.p, st))) return rc;
    std::vector<int> hits(n + 2 * m);
    TAGS_HIP(hipMemcpyAsync(hits.data(), dhits.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    TAGS_HIP(hipStreamSynchronize(st));
    // the reference stops after the first node without a segment (src/linalg.c:72-75)
    inc.nw = n;
    inc.bad = -1;
    for (long i = 0; i < n; i++)
        if (hits[i] == 0) {
            inc.bad = i;
            inc.nw = i + 1;
            break;
        }
    DevBuf<unsigned long long> cnt;
    if ((rc = inc.b1.alloc((size_t)2 * m * inc.nw)) || (rc = inc.b2.alloc((size_t)2 * m * inc.nw)) ||
        (rc = cnt.alloc(2)))
        return rc;
    TAGS_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long) * 2, st));
    {
        long tot = inc.nw * m;
        unsigned g = (unsigned)std::min<long>((tot + 255) / 256, 4096);
        hipLaunchKernelGGL(k_incidence_blocks, dim3(g), dim3(256), 0, st, dsp.p, m, dnodes.p,
                           inc.nw, inc.b1.p, inc.b2.p, cnt.p);
        TAGS_HIP(hipGetLastError());
    }
    TAGS_HIP(hipMemcpyAsync(inc.c, cnt.p, sizeof(inc.c), hipMemcpyDeviceToHost, st));
    TAGS_HIP(hipStreamSynchronize(st));
    return 0;
}

// the reference's return codes and messages (src/linalg.c:72-84, :186-198)
static int incidence_verdict(long bad, long suma, double sumb, long m)
{
    if (bad >= 0) {
        std::printf("No electrode is connected to node[%i]\n", (int)bad);
        return 1;
    }
    if (suma != 0) {
        std::printf("incidence matrix A is wrong. sum(A) = %li != 0\n", suma);
        return (int)-suma;
    }
    if (sumb != (double)m) {
        std::printf("incidence matrix B is wrong. sum(B) = %.1f != %.1f\n", sumb, (double)m);
        return (int)(sumb - m);
    }
    return 0;
}

int fill_incidence_imm(TAGS_COMPLEX* wg, const Electrode* electrodes, size_t num_electrodes,
                       const double* nodes, size_t num_nodes)
{
    const long m = (long)num_electrodes, n = (long)num_nodes, N = n + 2 * m;
    // no nodes: the sums stay 0 (B is reported wrong unless m = 0); no
    // segments: node 0 is unconnected
    if (n == 0) return incidence_verdict(-1, 0, 0.0, m);
    if (m == 0) return incidence_verdict(0, 0, 0.0, 0);
    Stream S;
    Incidence inc;
    int rc = incidence_blocks(inc, electrodes, m, nodes, n, S.s);
    if (rc) return rc;
    const long nw = inc.nw;
    // blk1 -> W rows n..N, columns 0..nw ; blk2 -> W rows 0..nw, columns n..N
    c64* W = (c64*)wg;
    TAGS_HIP(hipMemcpy2DAsync(W + n, sizeof(c64) * N, inc.b1.p, sizeof(c64) * 2 * m,
                              sizeof(c64) * 2 * m, nw, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipMemcpy2DAsync(W + N * n, sizeof(c64) * N, inc.b2.p, sizeof(c64) * nw,
                              sizeof(c64) * nw, 2 * m, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    return incidence_verdict(inc.bad, (long)inc.c[0] - (long)inc.c[1],
                             0.5 * (double)(inc.c[0] + inc.c[1]), m);
}

// src/linalg.c:143-199: a and b (m x n column-major, entry (s, node) at
// m * node + s) for the nodes up to the first unconnected one. b is written
// before a, so with a == b the array ends as A; the reference then also sums
// B over a's values (sum(B) = sum(A) = 0) and reports B as wrong.
int fill_incidence_adm(TAGS_COMPLEX* a, TAGS_COMPLEX* b, const Electrode* electrodes,
                       size_t num_electrodes, const double* nodes, size_t num_nodes)
{
    const long m = (long)num_electrodes, n = (long)num_nodes;
    if (n == 0) return incidence_verdict(-1, 0, 0.0, m);
    if (m == 0) return incidence_verdict(0, 0, 0.0, 0);
    Stream S;
    Incidence inc;
    int rc = incidence_blocks(inc, electrodes, m, nodes, n, S.s);
    if (rc) return rc;
    const long nw = inc.nw;
    TAGS_HIP(hipMemcpy2DAsync(b, sizeof(c64) * m, inc.b1.p + m, sizeof(c64) * 2 * m,
                              sizeof(c64) * m, nw, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipMemcpy2DAsync(a, sizeof(c64) * m, inc.b1.p, sizeof(c64) * 2 * m,
                              sizeof(c64) * m, nw, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    const long suma = (long)inc.c[0] - (long)inc.c[1];
    const double sumb = a == b ? (double)suma : 0.5 * (double)(inc.c[0] + inc.c[1]);
    return incidence_verdict(inc.bad, suma, sumb, m);
}

}  // extern "C"

namespace tags {
// Z-region of W (rows/cols n..N), column-major with ld 2m:
// [[-zl, 0], [0, -zt]] with zl[i + m k] at (i, k) (src/linalg.c:96-103).
__global__ void k_zregion(const c64* zl, const c64* zt, long m, c64* blk)
{
    const long M2 = 2 * m;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < M2 * M2;
         idx += (long)gridDim.x * 256) {
        const long col = idx / M2, row = idx % M2;
        c64 v = c64{0.0, 0.0};
        if (row < m && col < m) {
            const c64 z = zl[row + m * col];
            v = c64{-z.x, -z.y};
        } else if (row >= m && col >= m) {
            const c64 z = zt[(row - m) + m * (col - m)];
            v = c64{-z.x, -z.y};
        }
        blk[idx] = v;
    }
}
}  // namespace tags

extern "C" {

int fill_impedance_imm(TAGS_COMPLEX* wg, const TAGS_COMPLEX* zl, const TAGS_COMPLEX* zt,
                       size_t num_electrodes, size_t num_nodes)
{
    const long m = (long)num_electrodes, n = (long)num_nodes, N = n + 2 * m;
    Stream S;
    int rc;
    c64* W = (c64*)wg;
    if (m > 0) {
        DevBuf<c64> dl, dt, blk;
        if ((rc = dl.alloc(m * m)) || (rc = dt.alloc(m * m)) || (rc = blk.alloc(4 * m * m)))
            return rc;
        TAGS_HIP(hipMemcpyAsync(dl.p, zl, sizeof(c64) * m * m, hipMemcpyHostToDevice, S.s));
        TAGS_HIP(hipMemcpyAsync(dt.p, zt, sizeof(c64) * m * m, hipMemcpyHostToDevice, S.s));
        unsigned g = (unsigned)std::min<long>((4 * m * m + 255) / 256, 8192);
        hipLaunchKernelGGL(k_zregion, dim3(g), dim3(256), 0, S.s, dl.p, dt.p, m, blk.p);
        TAGS_HIP(hipGetLastError());
        TAGS_HIP(hipMemcpy2DAsync(W + n + N * n, sizeof(c64) * N, blk.p, sizeof(c64) * 2 * m,
                                  sizeof(c64) * 2 * m, 2 * m, hipMemcpyDeviceToHost, S.s));
        TAGS_HIP(hipStreamSynchronize(S.s));
    }
    for (long k = 0; k < n; k++) std::memset(W + N * k, 0, sizeof(c64) * n);
    return 0;
}

// ---------------------------------------------------------------------------
// solve_immittance (src/linalg.c:113-139)
// ---------------------------------------------------------------------------
static int dense_solve_host(const c64* a, c64* bx, long n, int precision, bool sym_lower_cm,
                            tags_item_stats* stt, const LUPlan* plan = nullptr,
                            const LUPlan* plan32 = nullptr, Grounding ground = Grounding{})
{
    if (n <= 0) return 0;
    if (n >= 65536) return TAGS_ERR_TOO_LARGE;
    Stream S;
    Solver sv;
    sv.st = S.s;
    int rc;
    const bool mixed = precision == TAGS_PREC_MIXED;
    if ((rc = sv.ensure((int)n, mixed))) return rc;
    DevBuf<c64> W, bv;
    if ((rc = W.alloc((size_t)n * n)) || (rc = bv.alloc(n))) return rc;
    TAGS_HIP(hipMemcpyAsync(W.p, a, sizeof(c64) * n * n, hipMemcpyHostToDevice, S.s));
    TAGS_HIP(hipMemcpyAsync(bv.p, bx, sizeof(c64) * n, hipMemcpyHostToDevice, S.s));
    if (sym_lower_cm && (rc = sym_from_upper(W.p, n, (int)n, S.s))) return rc;
    sv.plan = plan;
    sv.plan32 = plan32;
    sv.ground = ground;
    rc = mixed ? sv.solve_mixed(W.p, bv.p, stt) : sv.solve64(W.p, bv.p, stt);
    if (rc) return rc;
    TAGS_HIP(hipMemcpyAsync(bx, bv.p, sizeof(c64) * n, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    return 0;
}

int solve_immittance(TAGS_COMPLEX* wg, TAGS_COMPLEX* ie, size_t num_electrodes, size_t num_nodes)
{
    const long N = (long)(2 * num_electrodes + num_nodes);
    int rc = dense_solve_host((const c64*)wg, (c64*)ie, N, TAGS_PREC_FP64, true, nullptr);
    if (rc > 0) {
        std::printf("The diagonal element of the triangular factor of YN,\n");
        std::printf("U(%i,%i) is zero, so that YN is singular;\n", rc, rc);
        std::printf("the solution could not be computed.\n");
    }
    return rc;
}

int tags_dense_solve(const TAGS_COMPLEX* a, TAGS_COMPLEX* b, size_t n, int precision,
                     tags_item_stats* stats)
{
    return dense_solve_host((const c64*)a, (c64*)b, (long)n, precision, false, stats);
}

int tags_immittance_solve(const TAGS_COMPLEX* a, TAGS_COMPLEX* b, size_t num_electrodes,
                          size_t num_nodes, int precision, int plan, tags_item_stats* stats)
{
    const long m = (long)num_electrodes, nn = (long)num_nodes;
    if (m <= 0 || nn <= 0) return TAGS_ERR_ARG;
    LUPlan p, p32;
    Grounding gr;
    switch (plan) {
    case TAGS_PLAN_DENSE: p = LUPlan::dense((int)(nn + 2 * m)); break;
    case TAGS_PLAN_STRUCTURED: p = LUPlan::nodes_first((int)m, (int)nn); break;
    case TAGS_PLAN_ZL_FIRST: p = LUPlan::zl_first((int)m, (int)nn, false); break;
    case TAGS_PLAN_BLOCK_PIVOT:  // the weak pivoting for the complex64 factor only
        p = LUPlan::zl_first((int)m, (int)nn, false);
        p32 = LUPlan::zl_first((int)m, (int)nn, true);
        break;
    case TAGS_PLAN_NODAL:
        p = LUPlan::zl_first((int)m, (int)nn, false);
        p32 = LUPlan::nodal((int)m, (int)nn);
        break;
    case TAGS_PLAN_HYBRID:  // no frequency here: as ZL_FIRST
        p = LUPlan::zl_first((int)m, (int)nn, false);
        break;
    case TAGS_PLAN_GROUNDED:
    case TAGS_PLAN_GROUNDED_PIVOTED:
        p = LUPlan::zl_first((int)m, (int)nn, false);
        p32 = plan == TAGS_PLAN_GROUNDED ? LUPlan::zl_first_unpivoted((int)m, (int)nn) : p;
        gr = Grounding{(int)(2 * m), (int)(2 * m), (int)(2 * m + nn)};
        break;
    default: return TAGS_ERR_ARG;
    }
    const bool g = plan == TAGS_PLAN_GROUNDED || plan == TAGS_PLAN_GROUNDED_PIVOTED;
    return dense_solve_host((const c64*)a, (c64*)b, nn + 2 * m, precision, false, stats, &p,
                            g || plan == TAGS_PLAN_NODAL || plan == TAGS_PLAN_BLOCK_PIVOT
                                ? &p32
                                : nullptr,
                            gr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// nodal admittance formulation (src/linalg.c:201-350)
// ---------------------------------------------------------------------------
namespace tags {
// dst[c * ldd + r] = scale * src[(map ? map[r] : r) * lds + c] for r < rows,
// c < cols (lower: r >= c only), through a 32 x 32 LDS tile; grid (cols / 32,
// rows / 32) rounded up.
__global__ __launch_bounds__(256) void k_transpose(const c64* __restrict__ src, long lds,
                                                   const int* __restrict__ map, c64* dst,
                                                   long ldd, int rows, int cols, double scale,
                                                   int lower)
{
    __shared__ c64 t[32][33];
    const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
    if (lower && r0 + 31 < c0) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int i = ty; i < 32; i += 8) {
        const int r = r0 + i, c = c0 + tx;
        if (r < rows && c < cols) t[i][tx] = src[(long)(map ? map[r] : r) * lds + c];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int c = c0 + i, r = r0 + tx;
        if (r < rows && c < cols && (!lower || r >= c)) {
            const c64 v = t[tx][i];
            dst[(long)c * ldd + r] = c64{scale * v.x, scale * v.y};
        }
    }
}

// rows [0, m) of the m x m block at column c0 of a row-major matrix set to I
__global__ __launch_bounds__(256) void k_identity(c64* M, long ld, long c0, long m)
{
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < m * m;
         idx += (long)gridDim.x * 256) {
        const long r = idx / m, c = idx % m;
        M[r * ld + c0 + c] = c64{r == c ? 1.0 : 0.0, 0.0};
    }
}

static int transpose(const c64* src, long lds, const int* map, c64* dst, long ldd, long rows,
                     long cols, double scale, bool lower, hipStream_t st)
{
    if (rows <= 0 || cols <= 0) return 0;
    if ((rows + 31) / 32 > 65535) return TAGS_ERR_TOO_LARGE;
    hipLaunchKernelGGL(k_transpose, dim3((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32)),
                       dim3(256), 0, st, src, lds, map, dst, ldd, (int)rows, (int)cols, scale,
                       lower ? 1 : 0);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// One complex symmetric block Z (m x m, its lower triangle in the column-major
// device array dz, as zsytrf/zsytri read it) and a right-hand block C, crm
// holding s C (m x n, row-major on the device, ld n): the row-major work
// matrix [Z | s C | I]
// (ld m + n + m) is factored on its first m columns with row exchanges, its
// trailing updates carrying s C and I through L^-1 P, and the extra columns
// are completed by the backward blocked solve to [s Z^-1 C | Z^-1]. Then
// Z^-1 C goes to dy (m x n column-major) and the lower triangle of Z^-1 over
// dz's lower triangle (its strict upper triangle is left as it was). Returns
// 0, the first zero pivot (1-based) or an error.
static int adm_block(LUWork<double>& w, c64* dz, const c64* crm, double s, long m, long n,
                     c64* dy, hipStream_t st)
{
    if (m >= 65536 || n >= 65536) return TAGS_ERR_TOO_LARGE;
    const long ld = 2 * m + n;
    int rc;
    DevInfo& di = dev_info();
    if (!di.ok) return TAGS_ERR_HIP;
    if ((rc = w.alloc((int)m, di.num_cu, di.lds_limit))) return rc;
    DevBuf<c64> M;
    if ((rc = M.alloc((size_t)m * ld))) return rc;
    TAGS_HIP(hipMemcpy2DAsync(M.p, sizeof(c64) * ld, dz, sizeof(c64) * m, sizeof(c64) * m, m,
                              hipMemcpyDeviceToDevice, st));
    if ((rc = sym_from_upper(M.p, ld, (int)m, st))) return rc;
    if (n > 0)
        TAGS_HIP(hipMemcpy2DAsync(M.p + m, sizeof(c64) * ld, crm, sizeof(c64) * n,
                                  sizeof(c64) * n, m, hipMemcpyDeviceToDevice, st));
    {
        unsigned g = (unsigned)std::min<long>((m * m + 255) / 256, 8192);
        hipLaunchKernelGGL(k_identity, dim3(g), dim3(256), 0, st, M.p, ld, m + n, m);
        TAGS_HIP(hipGetLastError());
    }
    LUPlan plan;
    plan.phases.push_back(LUPhase{0, (int)m, (int)m, (int)m, (int)m, (int)ld});
    if ((rc = lu_factor<double>(w, M.p, ld, st, nullptr, &plan))) return rc;
    int flags[2];
    TAGS_HIP(hipMemcpyAsync(flags, w.flags, sizeof(flags), hipMemcpyDeviceToHost, st));
    TAGS_HIP(hipStreamSynchronize(st));
    if (flags[1]) return TAGS_ERR_BARRIER;
    if (flags[0]) return flags[0];
    if ((rc = lu_solve_cols<double>(w, M.p, ld, st, nullptr, (int)m, (int)(n + m)))) return rc;
    if ((rc = transpose(M.p + m, ld, w.perm, dy, m, m, n, 1.0 / s, false, st))) return rc;
    if ((rc = transpose(M.p + m + n, ld, w.perm, dz, m, m, m, 1.0, true, st))) return rc;
    TAGS_HIP(hipStreamSynchronize(st));
    return 0;
}
}  // namespace tags

extern "C" {

// Device copies of the host arrays of one admittance call.
struct AdmArrays {
    DevBuf<c64> zl, zt, arm, brm, yla, ytb, yn;
};

// Z^-1 (lower triangles, into zl and zt on the host), yla = Z_L^-1 A and
// ytb = Z_T^-1 B (on the device, column-major m x n) and the row-major A and
// -B (m x n) for the products.
static int adm_inverses(AdmArrays& d, TAGS_COMPLEX* zl, TAGS_COMPLEX* zt, const TAGS_COMPLEX* a,
                        const TAGS_COMPLEX* b, long m, long n, hipStream_t st)
{
    int rc;
    if (m == 0) return 0;
    if ((rc = d.zl.alloc((size_t)m * m)) || (rc = d.zt.alloc((size_t)m * m)) ||
        (rc = d.arm.alloc((size_t)m * n)) || (rc = d.brm.alloc((size_t)m * n)) ||
        (rc = d.yla.alloc((size_t)m * n)) || (rc = d.ytb.alloc((size_t)m * n)))
        return rc;
    {
        // a, b (column-major m x n) staged in yla / ytb, transposed to row-major
        TAGS_HIP(hipMemcpyAsync(d.yla.p, a, sizeof(c64) * m * n, hipMemcpyHostToDevice, st));
        TAGS_HIP(hipMemcpyAsync(d.ytb.p, b, sizeof(c64) * m * n, hipMemcpyHostToDevice, st));
        if ((rc = transpose(d.yla.p, m, nullptr, d.arm.p, n, n, m, 1.0, false, st)) ||
            (rc = transpose(d.ytb.p, m, nullptr, d.brm.p, n, n, m, -1.0, false, st)))
            return rc;
    }
    TAGS_HIP(hipMemcpyAsync(d.zl.p, zl, sizeof(c64) * m * m, hipMemcpyHostToDevice, st));
    TAGS_HIP(hipMemcpyAsync(d.zt.p, zt, sizeof(c64) * m * m, hipMemcpyHostToDevice, st));
    LUWork<double> w;
    // Z_L first, written back before Z_T is factored: as with zsytrf/zsytri in
    // the reference, a singular Z_T leaves zl already inverted
    if ((rc = adm_block(w, d.zl.p, d.arm.p, 1.0, m, n, d.yla.p, st))) return rc;
    TAGS_HIP(hipMemcpyAsync(zl, d.zl.p, sizeof(c64) * m * m, hipMemcpyDeviceToHost, st));
    TAGS_HIP(hipStreamSynchronize(st));
    if ((rc = adm_block(w, d.zt.p, d.brm.p, -1.0, m, n, d.ytb.p, st))) return rc;
    TAGS_HIP(hipMemcpyAsync(zt, d.zt.p, sizeof(c64) * m * m, hipMemcpyDeviceToHost, st));
    TAGS_HIP(hipStreamSynchronize(st));
    return 0;
}

// yn (n x n column-major) = A^T yla + B^T ytb: as row-major products,
// yn^T-view(v, u) = sum_i yla(i, v) A(i, u), i.e. C = P Q with P = yla read
// row-major (n x m) and Q = A row-major (m x n); the B term as C -= ytb (-B).
static int adm_products(AdmArrays& d, long m, long n, hipStream_t st)
{
    int rc;
    if ((rc = d.yn.alloc((size_t)n * n))) return rc;
    if (m == 0) {
        TAGS_HIP(hipMemsetAsync(d.yn.p, 0, sizeof(c64) * n * n, st));
        return 0;
    }
    GemmArgs<double> g{(int)n, (int)n, (int)m, d.yla.p, m, nullptr, d.arm.p, n, nullptr,
                       d.yn.p, n, nullptr, 0};
    if ((rc = launch_gemm<double>(g, st))) return rc;
    g.A = d.ytb.p;
    g.B = d.brm.p;
    g.mode = 1;
    return launch_gemm<double>(g, st);
}

int fill_impedance_adm(TAGS_COMPLEX* yn, TAGS_COMPLEX* zl, TAGS_COMPLEX* zt, TAGS_COMPLEX* a,
                       TAGS_COMPLEX* b, size_t num_electrodes, size_t num_nodes)
{
    const long m = (long)num_electrodes, n = (long)num_nodes;
    Stream S;
    AdmArrays d;
    int rc = adm_inverses(d, zl, zt, a, b, m, n, S.s);
    if (rc) return rc;
    if ((rc = adm_products(d, m, n, S.s))) return rc;
    TAGS_HIP(hipMemcpyAsync(yn, d.yn.p, sizeof(c64) * n * n, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    return 0;
}

int calculate_yla_ytb(TAGS_COMPLEX* yla, TAGS_COMPLEX* ytb, TAGS_COMPLEX* zl, TAGS_COMPLEX* zt,
                      TAGS_COMPLEX* a, TAGS_COMPLEX* b, size_t num_electrodes, size_t num_nodes)
{
    const long m = (long)num_electrodes, n = (long)num_nodes;
    if (m == 0) return 0;
    Stream S;
    AdmArrays d;
    int rc = adm_inverses(d, zl, zt, a, b, m, n, S.s);
    if (rc) return rc;
    TAGS_HIP(hipMemcpyAsync(yla, d.yla.p, sizeof(c64) * m * n, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipMemcpyAsync(ytb, d.ytb.p, sizeof(c64) * m * n, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    return 0;
}

int fill_impedance_adm2(TAGS_COMPLEX* yn, TAGS_COMPLEX* yla, TAGS_COMPLEX* ytb, TAGS_COMPLEX* a,
                        TAGS_COMPLEX* b, size_t num_electrodes, size_t num_nodes)
{
    const long m = (long)num_electrodes, n = (long)num_nodes;
    if (n == 0) return 0;
    Stream S;
    AdmArrays d;
    int rc;
    if (m > 0) {
        if ((rc = d.arm.alloc((size_t)m * n)) || (rc = d.brm.alloc((size_t)m * n)) ||
            (rc = d.yla.alloc((size_t)m * n)) || (rc = d.ytb.alloc((size_t)m * n)))
            return rc;
        // a and b pass through one staging buffer to their row-major copies
        DevBuf<c64> stage;
        if ((rc = stage.alloc((size_t)m * n))) return rc;
        TAGS_HIP(hipMemcpyAsync(stage.p, a, sizeof(c64) * m * n, hipMemcpyHostToDevice, S.s));
        if ((rc = transpose(stage.p, m, nullptr, d.arm.p, n, n, m, 1.0, false, S.s))) return rc;
        TAGS_HIP(hipStreamSynchronize(S.s));
        TAGS_HIP(hipMemcpyAsync(stage.p, b, sizeof(c64) * m * n, hipMemcpyHostToDevice, S.s));
        if ((rc = transpose(stage.p, m, nullptr, d.brm.p, n, n, m, -1.0, false, S.s))) return rc;
        TAGS_HIP(hipMemcpyAsync(d.yla.p, yla, sizeof(c64) * m * n, hipMemcpyHostToDevice, S.s));
        TAGS_HIP(hipMemcpyAsync(d.ytb.p, ytb, sizeof(c64) * m * n, hipMemcpyHostToDevice, S.s));
        TAGS_HIP(hipStreamSynchronize(S.s));
    }
    if ((rc = adm_products(d, m, n, S.s))) return rc;
    TAGS_HIP(hipMemcpyAsync(yn, d.yn.p, sizeof(c64) * n * n, hipMemcpyDeviceToHost, S.s));
    TAGS_HIP(hipStreamSynchronize(S.s));
    return 0;
}

int solve_admittance(TAGS_COMPLEX* yn, TAGS_COMPLEX* ic, size_t num_nodes)
{
    int rc = dense_solve_host((const c64*)yn, (c64*)ic, (long)num_nodes, TAGS_PREC_FP64, true,
                              nullptr);
    if (rc > 0) {
        std::printf("The diagonal element of the triangular factor of YN,\n");
        std::printf("U(%i,%i) is zero, so that YN is singular;\n", rc, rc);
        std::printf("the solution could not be computed.\n");
    }
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// device-resident sweep
// ---------------------------------------------------------------------------
// A stream owned for the lifetime of a slot: declared first in a slot, so that
// it is destroyed after everything that records work or events on it.
struct OwnedStream {
    hipStream_t s = nullptr;
    ~OwnedStream()
    {
        if (s) {
            hipStreamSynchronize(s);
            hipStreamDestroy(s);
        }
    }
};

// One frequency sample in flight: its own stream, W, right-hand side, solver
// (factor work space, complex64 copy) and kernel timers. The sweep queues a
// sample on every slot, then finishes them in turn from the calling thread,
// queuing the next sample on each slot as it frees up: one sample's trailing
// updates keep the GPU busy while another waits on its panel chain.
struct SweepSlot {
    OwnedStream stream;
    DevBuf<c64> W, rhs;        // cap items: W at stride N^2, rhs at stride N
    DevBuf<PairScalars> dsc;   // the group's per-sample scalars
    PairScalars* hsc = nullptr;  // pinned staging of dsc
    Profiler prof;
    Solver sv;
    int cap = 0;
    const tags_sweep* bound = nullptr;  // the sweep whose plans the solver holds
    int bound_kind = -1;                // and its plan kind then
    long f = -1;            // first sample of the group in flight (-1: idle)
    int cnt = 0;            // its samples f .. f + cnt - 1
    std::vector<tags_item_stats> stt;  // their statistics and infos, filled as they finish
    std::vector<int> info;
    std::vector<double> hz; // their frequencies |s| / 2 pi
    bool nodal = false;     // their complex64 factor is the nodal plan's
    size_t w_cap = 0, rhs_cap = 0;  // allocated elements of W and rhs
    // items [0, static_items) of W hold the frequency-independent part (zero
    // and incidence blocks) of static_sw's layout static_kind: the mixed form
    // factors a complex64 copy and leaves W as assembled, so the next group
    // of the same sweep writes only the impedance blocks
    const tags_sweep* static_sw = nullptr;
    int static_kind = -1;
    int static_items = 0;
    int init(long N, int batch)
    {
        if (!stream.s) TAGS_HIP(hipStreamCreate(&stream.s));
        int rc;
        const size_t nw = (size_t)N * N * batch, nr = (size_t)N * batch;
        if (nw > w_cap || nr > rhs_cap || batch > cap) {
            // grow to cover every earlier request as well (a lane of a
            // multi-case solve serves systems of different orders in turn)
            const size_t aw = std::max(nw, w_cap), ar = std::max(nr, rhs_cap);
            const int items = std::max(batch, cap);
            hipFree(W.p);
            hipFree(rhs.p);
            hipFree(dsc.p);
            if (hsc) hipHostFree(hsc);
            W.p = rhs.p = nullptr;
            dsc.p = nullptr;
            hsc = nullptr;
            cap = 0;
            w_cap = rhs_cap = 0;
            static_items = 0;
            if ((rc = W.alloc(aw)) || (rc = rhs.alloc(ar)) || (rc = dsc.alloc((size_t)items)))
                return rc;
            TAGS_HIP(hipHostMalloc(&hsc, sizeof(PairScalars) * items, 0));
            cap = items;
            w_cap = aw;
            rhs_cap = ar;
        }
        sv.st = stream.s;
        sv.pr = &prof;
        return 0;
    }
    ~SweepSlot()
    {
        if (stream.s) hipStreamSynchronize(stream.s);
        if (hsc) hipHostFree(hsc);
    }
};

// Device copies of one plan's row orders, owned by the sweep that owns the
// plan: uploaded (a blocking copy) only when the plan's content changes, at
// layout time, before any of the sweep's work is queued; factorisations then
// read them in place. Without them a work space that serves many plans in
// turn (a lane of a multi-case solve) would upload, with a blocking copy
// that waits for every stream, at each change of case.
struct PlanDev {
    int* init = nullptr;
    int* orders = nullptr;
    size_t init_cap = 0, orders_cap = 0;
    std::vector<int> h_init, h_orders;
    int attach(LUPlan& p)
    {
        p.dev_init = nullptr;
        p.dev_orders = nullptr;
        if (!p.init_perm.empty()) {
            if (p.init_perm != h_init) {
                if (p.init_perm.size() > init_cap) {
                    hipFree(init);
                    init = nullptr;
                    init_cap = 0;
                    TAGS_HIP(hipMalloc(&init, sizeof(int) * p.init_perm.size()));
                    init_cap = p.init_perm.size();
                }
                TAGS_HIP(hipMemcpy(init, p.init_perm.data(), sizeof(int) * p.init_perm.size(),
                                   hipMemcpyHostToDevice));
                h_init = p.init_perm;
            }
            p.dev_init = init;
        }
        if (!p.orders.empty()) {
            if (p.orders != h_orders) {
                if (p.orders.size() > orders_cap) {
                    hipFree(orders);
                    orders = nullptr;
                    orders_cap = 0;
                    TAGS_HIP(hipMalloc(&orders, sizeof(int) * p.orders.size()));
                    orders_cap = p.orders.size();
                }
                TAGS_HIP(hipMemcpy(orders, p.orders.data(), sizeof(int) * p.orders.size(),
                                   hipMemcpyHostToDevice));
                h_orders = p.orders;
            }
            p.dev_orders = orders;
        }
        return 0;
    }
    ~PlanDev()
    {
        hipFree(init);
        hipFree(orders);
    }
};

// refinement steps above which a nodal sample of a hybrid sweep counts as slow
// (a step costs a few per cent of a factorisation, a fallback to FP64 several
// factorisations)
constexpr int kHybridMaxIters = 10;

struct tags_sweep {
    long m = 0, n = 0, N = 0;
    int plan_kind = TAGS_PLAN_HYBRID;
    int layout_kind = -1;  // plan kind the current layout and grounding sums belong to
    int inflight = 4;      // groups of frequency samples in flight
    int batch_req = 0;     // samples per group as set (0: chosen from N and memory)
    int batch = 1;         // samples per group of the current call
    tags::LUPlan plan;
    tags::LUPlan plan32;
    // TAGS_PLAN_HYBRID: the mixed form's complex64 factor is the nodal plan's
    // at and above nodal_above Hz, ZL_FIRST's below; a nodal sample that falls
    // back to FP64 or needs more than kHybridMaxIters refinement steps raises
    // the switch to 1.5 times its frequency for the rest of the sweep
    tags::LUPlan plan_nodal;
    // device copies of plan, plan32 and plan_nodal's row orders
    PlanDev plan_dev[3];
    double nodal_above = 0.0;
    // nodal complex64 factors below gmres_below Hz refine by FGMRES
    // (gmres_steps Arnoldi steps per refinement step), above it classically
    double gmres_below = 2e4;
    int gmres_steps = 6;
    tags::BlockLayout layout{};
    hipStream_t st = nullptr;  // set-up and assembly stream
    tags::DevSegs e, img;
    int* sn = nullptr;
    int* en = nullptr;
    // each node's segment ends (node_ptr: n + 1, node_seg: 2 m at most) and the
    // structure of W in the current layout, for the refinement's products
    int* node_ptr = nullptr;
    int* node_seg = nullptr;
    tags::HemOp hem{};
    std::vector<std::unique_ptr<SweepSlot>> slots;
    bool profiling = false;
    // TAGS_FORM_MHEM: the gamma-free mHEM potentials (direct Z_L, Z_T, image
    // Z_L, Z_T; real m x m each), computed at the first sample after the switch
    int formulation = TAGS_FORM_HEM;
    double* pot = nullptr;
    // FP64 form: after each solve, assemble W again (the factor replaced it)
    // and report the backward error of x in the item statistics
    bool verify = false;
    // tuned tile shapes and schedule per precision (tags_sweep_set_tuning):
    // the kernel-level part is installed for the calling thread during a
    // solve, the sweep-level part (in flight, group size, switch) replaces the
    // sweep's settings for the call
    tags_tuning tune[2] = {};
    bool tuned[2] = {false, false};
    tags::Tuning ktune[2];
    // the plans a slot's solver takes for this sweep (set by sweep_layout)
    const tags::LUPlan* bind64 = nullptr;
    const tags::LUPlan* bind32 = nullptr;
    tags::Grounding bindgr{};
};

// Points a slot's solver at the sweep's plans (a slot of a multi-case solve
// serves many sweeps in turn); grounding sums are kept only while the slot
// stays with the same sweep and plan kind.
static void slot_bind(tags_sweep* sw, SweepSlot& sl)
{
    if (sl.bound != sw || sl.bound_kind != sw->layout_kind) sl.sv.ground_valid = false;
    sl.bound = sw;
    sl.bound_kind = sw->layout_kind;
    sl.sv.plan = sw->bind64;
    sl.sv.plan32 = sw->bind32;
    sl.sv.plan32_retry = nullptr;
    sl.sv.ground = sw->bindgr;
    sl.sv.gmres_k = 0;
    sl.sv.hem = &sw->hem;
}

// The first k slots exist and are allocated for order N (precision `mixed`).
// Samples per group when none is set: one from order kBatchOneAbove up (a
// sample's own launches fill the device), else as many as keep a slot's
// matrices (complex128 W and, mixed, its complex64 copy) within kBatchBytes,
// at most kBatchMax: 44 (mixed) or 64 for the 30 x 30 m transient grid
// (N = 3861), 64 below.
constexpr double kBatchBytes = 16.0 * (1 << 30);
constexpr int kBatchMax = 64;
constexpr long kBatchOneAbove = 8192;
static int sweep_batch(const tags_sweep* sw, bool mixed)
{
    if (sw->batch_req > 0) return sw->batch_req;
    if (sw->N >= kBatchOneAbove) return 1;
    const double N = (double)sw->N;
    const double per = N * N * (mixed ? 24.0 : 16.0) + N * 64.0 * 2 * (mixed ? 8 : 16) + N * 96.0;
    const double b = std::floor(kBatchBytes / per);
    return (int)std::max(1.0, std::min((double)kBatchMax, b));
}

// The first k slots exist and are allocated for groups of sw->batch samples of
// order N (precision `mixed`).
static int sweep_slots(tags_sweep* sw, int k, bool mixed)
{
    while ((int)sw->slots.size() < k) sw->slots.emplace_back(new SweepSlot());
    for (int i = 0; i < k; i++) {
        SweepSlot& sl = *sw->slots[i];
        int rc = sl.init(sw->N, sw->batch);
        if (rc) return rc;
        if ((rc = sl.sv.ensure((int)sw->N, mixed, sw->batch))) return rc;
        sl.prof.on = sw->profiling;
    }
    return 0;
}

static int sweep_layout(tags_sweep* sw)
{
    const long n = sw->n, m = sw->m;
    const bool changed = sw->layout_kind != sw->plan_kind;
    sw->layout_kind = sw->plan_kind;
    const LUPlan* p64 = &sw->plan;
    const LUPlan* p32 = nullptr;
    Grounding gr{};
    switch (sw->plan_kind) {
    case TAGS_PLAN_DENSE:
        sw->layout = BlockLayout{n, m, 0, n, n + m};
        p64 = nullptr;
        break;
    case TAGS_PLAN_STRUCTURED:
        sw->layout = BlockLayout{n, m, 0, n, n + m};
        sw->plan = LUPlan::nodes_first((int)m, (int)n);
        break;
    case TAGS_PLAN_GROUNDED:
    case TAGS_PLAN_GROUNDED_PIVOTED:
        sw->layout = BlockLayout{n, m, 2 * m, 0, m};
        sw->plan = LUPlan::zl_first((int)m, (int)n, false);
        sw->plan32 = sw->plan_kind == TAGS_PLAN_GROUNDED
                         ? LUPlan::zl_first_unpivoted((int)m, (int)n)
                         : LUPlan::zl_first((int)m, (int)n, false);
        p32 = &sw->plan32;
        gr = Grounding{(int)(2 * m), (int)(2 * m), (int)(2 * m + n)};
        break;
    case TAGS_PLAN_NODAL:
        sw->layout = BlockLayout{n, m, 2 * m, 0, m};
        sw->plan = LUPlan::zl_first((int)m, (int)n, false);
        sw->plan32 = LUPlan::nodal((int)m, (int)n);
        p32 = &sw->plan32;
        break;
    case TAGS_PLAN_HYBRID:  // complex64 plan chosen per sample (sweep_issue)
        sw->layout = BlockLayout{n, m, 2 * m, 0, m};
        sw->plan = LUPlan::zl_first((int)m, (int)n, false);
        sw->plan_nodal = LUPlan::nodal((int)m, (int)n, false);
        break;
    case TAGS_PLAN_BLOCK_PIVOT:
        // the FP64 form takes the fully pivoted factor: the Z_L block's weak
        // pivoting is for the complex64 factor, whose refinement answers for it
        sw->layout = BlockLayout{n, m, 2 * m, 0, m};
        sw->plan = LUPlan::zl_first((int)m, (int)n, false);
        sw->plan32 = LUPlan::zl_first((int)m, (int)n, true);
        p32 = &sw->plan32;
        break;
    default:
        sw->layout = BlockLayout{n, m, 2 * m, 0, m};
        sw->plan = LUPlan::zl_first((int)m, (int)n, false);
        break;
    }
    int rc;
    if ((rc = sw->plan_dev[0].attach(sw->plan)) || (rc = sw->plan_dev[1].attach(sw->plan32)) ||
        (rc = sw->plan_dev[2].attach(sw->plan_nodal)))
        return rc;
    sw->bind64 = p64;
    sw->bind32 = p32;
    sw->bindgr = gr;
    sw->hem.m = m;
    sw->hem.n = n;
    sw->hem.node_off = sw->layout.node_off;
    sw->hem.zl_off = sw->layout.zl_off;
    sw->hem.zt_off = sw->layout.zt_off;
    sw->hem.sn = sw->sn;
    sw->hem.en = sw->en;
    sw->hem.node_ptr = sw->node_ptr;
    sw->hem.node_seg = sw->node_seg;
    for (auto& sl : sw->slots) {
        if (changed) sl->sv.ground_valid = false;
        slot_bind(sw, *sl);
    }
    return 0;
}

// Per-frequency scalars (reference caller formulas) and the assembly of W in
// the sweep's block layout.
static PairScalars sample_scalars(TAGS_COMPLEX s_, TAGS_COMPLEX kappa_, double mur,
                                  const TAGS_COMPLEX* gamma_, const TAGS_COMPLEX* ref_l_,
                                  const TAGS_COMPLEX* ref_t_)
{
    const c64 sf = from_c(s_), kf = from_c(kappa_);
    c64 gf, rl, rt;
    if (gamma_) {
        gf = from_c(*gamma_);
    } else {
        // gamma = csqrt(s mu0 mur kappa), principal branch (examples/harmonic_impedance.c:118)
        const c64 z = cmul(c64{sf.x * kMu0 * mur, sf.y * kMu0 * mur}, kf);
        const std::complex<double> g = std::sqrt(std::complex<double>(z.x, z.y));
        gf = c64{g.real(), g.imag()};
    }
    rl = ref_l_ ? from_c(*ref_l_) : c64{1.0, 0.0};
    if (ref_t_) {
        rt = from_c(*ref_t_);
    } else {
        // (kappa - s eps0)/(kappa + s eps0), examples/harmonic_impedance.c:122
        const c64 se = c64{sf.x * kEps0, sf.y * kEps0};
        rt = cdiv(csub(kf, se), cadd(kf, se));
    }
    return make_scalars(gf, sf, mur, kf, rl, rt, INTG_DOUBLE);
}

// Assembly of nb samples' W (item z at W + z N^2) in the sweep's block layout:
// one sample from `sc`, several from their scalars `dsc` in device memory.
static int sweep_fill_items(tags_sweep* sw, c64* W, hipStream_t st, Profiler* prof, int nb,
                            const PairScalars& sc, const PairScalars* dsc,
                            bool fill_static = true)
{
    const long N = sw->N, sW = N * N;
    Timed tm(prof, FAM_FILL, st);
    int rc;
    const PairScalars* scs = nb > 1 ? dsc : nullptr;
    if (fill_static && (rc = launch_fill_static(W, N, sw->layout, sw->sn, sw->en, st, nb, sW)))
        return rc;
    if (sw->formulation == TAGS_FORM_MHEM) {
        if (!sw->pot) return TAGS_ERR_ARG;
        const size_t mm = (size_t)sw->m * sw->m;
        if ((rc = launch_fill_mhem(W, N, sw->layout, sw->e.soa(), sw->img.soa(), sw->pot,
                                   sw->pot + mm, sw->pot + 2 * mm, sw->pot + 3 * mm, sc, st, nb,
                                   sW, scs)))
            return rc;
    } else if ((rc = launch_fill_zblocks(W, N, sw->layout, sw->e.soa(), sw->img.soa(), sc, st, nb,
                                         sW, scs))) {
        return rc;
    }
    tm.end();
    return 0;
}

static int sweep_fill(tags_sweep* sw, c64* W, hipStream_t st, Profiler* prof, TAGS_COMPLEX s_,
                      TAGS_COMPLEX kappa_, double mur, const TAGS_COMPLEX* gamma_,
                      const TAGS_COMPLEX* ref_l_, const TAGS_COMPLEX* ref_t_)
{
    const PairScalars sc = sample_scalars(s_, kappa_, mur, gamma_, ref_l_, ref_t_);
    return sweep_fill_items(sw, W, st, prof, 1, sc, nullptr);
}

// The mHEM potentials, once per geometry (on the set-up stream, finished
// before any sample is queued).
static int sweep_potentials(tags_sweep* sw)
{
    if (sw->formulation != TAGS_FORM_MHEM || sw->pot) return 0;
    const size_t mm = (size_t)sw->m * sw->m;
    TAGS_HIP(hipMalloc(&sw->pot, sizeof(double) * 4 * mm));
    int rc = launch_mhem_potentials(sw->pot, sw->pot + mm, sw->pot + 2 * mm, sw->pot + 3 * mm,
                                    sw->m, sw->e.soa(), sw->img.soa(), sw->st);
    if (rc) return rc;
    TAGS_HIP(hipStreamSynchronize(sw->st));
    return 0;
}

extern "C" {

int tags_sweep_create(tags_sweep** out, const Electrode* electrodes, const Electrode* images,
                      size_t num_electrodes, const double* nodes, size_t num_nodes)
{
    if (!out || !electrodes || !nodes || num_electrodes == 0 || num_nodes == 0)
        return TAGS_ERR_ARG;
    std::unique_ptr<tags_sweep> sw(new tags_sweep());
    sw->m = (long)num_electrodes;
    sw->n = (long)num_nodes;
    sw->N = sw->n + 2 * sw->m;
    if (sw->N >= 65536) return TAGS_ERR_TOO_LARGE;
    TAGS_HIP(hipStreamCreate(&sw->st));
    int rc;
    std::vector<Electrode> im;
    if (!images) {
        im.assign(electrodes, electrodes + num_electrodes);
        for (auto& x : im) {
            x.start_point[2] = -x.start_point[2];
            x.end_point[2] = -x.end_point[2];
            x.middle_point[2] = -x.middle_point[2];
        }
        images = im.data();
    }
    if ((rc = sw->e.upload(electrodes, sw->m, sw->st))) return rc;
    if ((rc = sw->img.upload(images, sw->m, sw->st))) return rc;
    const long m = sw->m, n = sw->n;
    std::vector<double> sp = seg_points(electrodes, m);
    DevBuf<double> dsp, dnodes;
    DevBuf<int> dhits;
    if ((rc = dsp.alloc(6 * m)) || (rc = dnodes.alloc(3 * n)) || (rc = dhits.alloc(n + 2 * m)))
        return rc;
    TAGS_HIP(hipMalloc(&sw->sn, sizeof(int) * m));
    TAGS_HIP(hipMalloc(&sw->en, sizeof(int) * m));
    TAGS_HIP(hipMemsetAsync(sw->sn, 0xff, sizeof(int) * m, sw->st));
    TAGS_HIP(hipMemsetAsync(sw->en, 0xff, sizeof(int) * m, sw->st));
    TAGS_HIP(hipMemcpyAsync(dsp.p, sp.data(), sizeof(double) * 6 * m, hipMemcpyHostToDevice,
                            sw->st));
    TAGS_HIP(hipMemcpyAsync(dnodes.p, nodes, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                            sw->st));
    TAGS_HIP(hipMemsetAsync(dhits.p, 0, sizeof(int) * (n + 2 * m), sw->st));
    if ((rc = launch_incidence(dsp.p, m, dnodes.p, n, sw->sn, sw->en, dhits.p, sw->st))) return rc;
    std::vector<int> hits(n + 2 * m);
    TAGS_HIP(hipMemcpyAsync(hits.data(), dhits.p, sizeof(int) * (n + 2 * m),
                            hipMemcpyDeviceToHost, sw->st));
    TAGS_HIP(hipStreamSynchronize(sw->st));
    for (long i = 0; i < n; i++)
        if (hits[i] == 0) {
            std::printf("No electrode is connected to node[%i]\n", (int)i);
            return 1;
        }
    long cs = 0, ce = 0;
    for (long s = 0; s < m; s++) {
        cs += hits[n + 2 * s];
        ce += hits[n + 2 * s + 1];
        if (hits[n + 2 * s] != 1 || hits[n + 2 * s + 1] != 1) {
            // every segment end must match exactly one node for the sweep
            if (cs != ce) return (int)-(cs - ce);
            return TAGS_ERR_ARG;
        }
    }
    // each node's segment ends in segment order (a segment whose two ends
    // share a node has its start entry only, as W has one entry there)
    std::vector<int> hsn(m), hen(m);
    TAGS_HIP(hipMemcpy(hsn.data(), sw->sn, sizeof(int) * m, hipMemcpyDeviceToHost));
    TAGS_HIP(hipMemcpy(hen.data(), sw->en, sizeof(int) * m, hipMemcpyDeviceToHost));
    std::vector<int> ptr(n + 1, 0), seg;
    for (long s = 0; s < m; s++) {
        if (hsn[s] < 0 || hsn[s] >= n || hen[s] < 0 || hen[s] >= n) return TAGS_ERR_ARG;
        ptr[hsn[s] + 1]++;
        if (hen[s] != hsn[s]) ptr[hen[s] + 1]++;
    }
    for (long u = 0; u < n; u++) ptr[u + 1] += ptr[u];
    seg.assign(ptr[n], 0);
    {
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (long s = 0; s < m; s++) {
            seg[fill[hsn[s]]++] = (int)(2 * s);
            if (hen[s] != hsn[s]) seg[fill[hen[s]]++] = (int)(2 * s + 1);
        }
    }
    TAGS_HIP(hipMalloc(&sw->node_ptr, sizeof(int) * (n + 1)));
    TAGS_HIP(hipMalloc(&sw->node_seg, sizeof(int) * std::max<size_t>(1, seg.size())));
    TAGS_HIP(hipMemcpy(sw->node_ptr, ptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (!seg.empty())
        TAGS_HIP(hipMemcpy(sw->node_seg, seg.data(), sizeof(int) * seg.size(),
                           hipMemcpyHostToDevice));
    *out = sw.release();
    return 0;
}

int tags_sweep_assemble(tags_sweep* sw, TAGS_COMPLEX s, TAGS_COMPLEX kappa, double mur,
                        TAGS_COMPLEX* wg)
{
    if (!sw || !wg) return TAGS_ERR_ARG;
    const long N = sw->N;
    int lrc = sweep_layout(sw);
    if (lrc) return lrc;
    DevBuf<c64> W;
    int rc = sweep_potentials(sw);
    if (rc || (rc = W.alloc((size_t)N * N))) return rc;
    if ((rc = sweep_fill(sw, W.p, sw->st, nullptr, s, kappa, mur, nullptr, nullptr, nullptr)))
        return rc;
    std::vector<c64> h((size_t)N * N);
    TAGS_HIP(hipMemcpyAsync(h.data(), W.p, sizeof(c64) * N * N, hipMemcpyDeviceToHost, sw->st));
    TAGS_HIP(hipStreamSynchronize(sw->st));
    // unknown index in the reference order [U; I_L; I_T] -> storage index
    const BlockLayout& L = sw->layout;
    std::vector<long> pos(N);
    for (long i = 0; i < sw->n; i++) pos[i] = L.node_off + i;
    for (long i = 0; i < sw->m; i++) pos[sw->n + i] = L.zl_off + i;
    for (long i = 0; i < sw->m; i++) pos[sw->n + sw->m + i] = L.zt_off + i;
    c64* out = (c64*)wg;
    for (long c = 0; c < N; c++)
        for (long r = 0; r < N; r++) out[r + N * c] = h[(size_t)pos[r] * N + pos[c]];
    return 0;
}

}  // extern "C"

namespace tags {
// out[i] = W[idx[i]] (flat storage index)
__global__ void k_gather_entries(const c64* W, const long* idx, long count, c64* out)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) out[i] = W[idx[i]];
}
}  // namespace tags

extern "C" {

int tags_sweep_assemble_entries(tags_sweep* sw, TAGS_COMPLEX s, TAGS_COMPLEX kappa, double mur,
                                size_t count, const long* rows, const long* cols,
                                TAGS_COMPLEX* out)
{
    if (!sw || (count && (!rows || !cols || !out))) return TAGS_ERR_ARG;
    const long N = sw->N;
    const BlockLayout& L = sw->layout;
    int lrc = sweep_layout(sw);
    if (lrc) return lrc;
    // unknown index in the reference order [U; I_L; I_T] -> storage index
    auto pos = [&](long u) {
        if (u < sw->n) return L.node_off + u;
        if (u < sw->n + sw->m) return L.zl_off + (u - sw->n);
        return L.zt_off + (u - sw->n - sw->m);
    };
    std::vector<long> idx(count ? count : 1);
    for (size_t i = 0; i < count; i++) {
        if (rows[i] < 0 || rows[i] >= N || cols[i] < 0 || cols[i] >= N) return TAGS_ERR_ARG;
        idx[i] = pos(rows[i]) * N + pos(cols[i]);
    }
    DevBuf<c64> W, dout;
    DevBuf<long> didx;
    int rc = sweep_potentials(sw);
    if (rc || (rc = W.alloc((size_t)N * N)) || (rc = dout.alloc(count)) ||
        (rc = didx.alloc(count)))
        return rc;
    if ((rc = sweep_fill(sw, W.p, sw->st, nullptr, s, kappa, mur, nullptr, nullptr, nullptr)))
        return rc;
    if (count) {
        TAGS_HIP(hipMemcpyAsync(didx.p, idx.data(), sizeof(long) * count, hipMemcpyHostToDevice,
                                sw->st));
        hipLaunchKernelGGL(k_gather_entries, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                           sw->st, W.p, didx.p, (long)count, dout.p);
        TAGS_HIP(hipGetLastError());
        TAGS_HIP(hipMemcpyAsync(out, dout.p, sizeof(c64) * count, hipMemcpyDeviceToHost,
                                sw->st));
    }
    TAGS_HIP(hipStreamSynchronize(sw->st));
    return 0;
}

int tags_sweep_set_plan(tags_sweep* sw, int plan)
{
    if (!sw || plan < 0 || plan > TAGS_PLAN_HYBRID) return TAGS_ERR_ARG;
    sw->plan_kind = plan;
    return 0;
}

int tags_sweep_set_nodal_above(tags_sweep* sw, double hz)
{
    if (!sw || !(hz >= 0.0)) return TAGS_ERR_ARG;
    sw->nodal_above = hz;
    return 0;
}

int tags_sweep_set_inflight(tags_sweep* sw, int samples)
{
    if (!sw || samples < 1 || samples > 4) return TAGS_ERR_ARG;
    sw->inflight = samples;
    return 0;
}

int tags_sweep_set_profiling(tags_sweep* sw, int on)
{
    if (!sw) return TAGS_ERR_ARG;
    sw->profiling = on != 0;
    for (auto& sl : sw->slots) {
        TAGS_HIP(hipStreamSynchronize(sl->stream.s));
        sl->prof.on = sw->profiling;
        sl->prof.reset();
    }
    return 0;
}

int tags_sweep_get_timers(tags_sweep* sw, tags_timers* t)
{
    if (!sw || !t) return TAGS_ERR_ARG;
    std::memset(t, 0, sizeof(*t));
    for (auto& sl : sw->slots) {
        TAGS_HIP(hipStreamSynchronize(sl->stream.s));
        Profiler& p = sl->prof;
        p.resolve();
        t->ms_fill += p.ms[FAM_FILL];
        t->ms_panel += p.ms[FAM_PANEL];
        t->ms_gemm += p.ms[FAM_GEMM];
        t->ms_other += p.ms[FAM_OTHER];
        t->ms_solve += p.ms[FAM_SOLVE];
        t->ms_residual += p.ms[FAM_RESID];
        t->gemm_flops += p.gemm_flops;
        t->trailing_ms += p.trailing_ms;
        t->trailing_flops += p.trailing_flops;
        t->trailing_launches += p.trailing_launches;
    }
    return 0;
}

int tags_sweep_set_formulation(tags_sweep* sw, int formulation)
{
    if (!sw || (formulation != TAGS_FORM_HEM && formulation != TAGS_FORM_MHEM))
        return TAGS_ERR_ARG;
    sw->formulation = formulation;
    return 0;
}

void tags_sweep_destroy(tags_sweep* sw)
{
    if (!sw) return;
    hipStreamSynchronize(sw->st);
    sw->slots.clear();
    hipFree(sw->sn);
    hipFree(sw->en);
    hipFree(sw->node_ptr);
    hipFree(sw->node_seg);
    hipFree(sw->pot);
    hipStream_t st = sw->st;
    delete sw;
    hipStreamDestroy(st);
}

}  // extern "C"

namespace tags {
// b[z N + i] = (i == pos), i < N, for the nb items of a group
__global__ void k_unit_rhs(c64* b, long N, long pos)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < N) b[(long)blockIdx.z * N + i] = c64{i == pos ? 1.0 : 0.0, 0.0};
}
}  // namespace tags

extern "C" {

// Per-call arguments of tags_sweep_solve shared by the issue and finish steps.
struct SweepCall {
    size_t nfreq;
    const TAGS_COMPLEX *s, *kappa, *gamma, *ref_l, *ref_t;
    double mur;
    size_t inj_node;
    bool mixed;
    TAGS_COMPLEX *u_inj, *x_out;
    tags_item_stats* stats;
};

// Queues the group of samples starting at f on slot `sl`: up to sw->batch
// consecutive samples that take the same complex64 factor (the hybrid plan
// chooses per frequency); assembly, right-hand sides and the solve's first
// half, one set of launches for the group. Returns the samples taken (> 0) or
// an error.
static int sweep_issue(tags_sweep* sw, SweepSlot& sl, size_t f, const SweepCall& c)
{
    const long N = sw->N;
    const BlockLayout& L = sw->layout;
    int rc;
    auto hz_of = [&](size_t q) {
        const c64 sf = from_c(c.s[q]);
        return std::hypot(sf.x, sf.y) / (2.0 * M_PI);
    };
    const bool hybrid = sw->plan_kind == TAGS_PLAN_HYBRID && c.mixed;
    const bool nodal_plan = sw->plan_kind == TAGS_PLAN_NODAL && c.mixed;
    // the sample's class: 0 the plan's own complex64 factor with classic
    // refinement (the hybrid plan's Z_L-first one below its switch), 1 a nodal
    // factor refined by FGMRES, 2 a nodal factor refined classically
    auto cls = [&](size_t q) {
        if (!hybrid && !nodal_plan) return 0;
        const double h = hz_of(q);
        if (hybrid && h < sw->nodal_above) return 0;
        return h < sw->gmres_below ? 1 : 2;
    };
    const int c0 = cls(f);
    sl.nodal = hybrid && c0 > 0;
    int cnt = 0;
    sl.hz.clear();
    while (cnt < sw->batch && f + cnt < c.nfreq && cls(f + cnt) == c0) {
        sl.hz.push_back(hz_of(f + cnt));
        cnt++;
    }
    sl.sv.gmres_k = c0 == 1 ? sw->gmres_steps : 0;
    sl.f = (long)f;
    sl.cnt = cnt;
    sl.stt.assign(cnt, tags_item_stats{});
    for (auto& q : sl.stt)
        for (int i = 0; i < 8; i++) q.refine_hist[i] = -1.0;
    sl.info.assign(cnt, 0);
    if (sw->plan_kind == TAGS_PLAN_HYBRID) {
        sl.sv.plan32 = sl.nodal ? &sw->plan_nodal : nullptr;
        sl.sv.plan32_retry = sl.nodal ? &sw->plan : nullptr;
    }
    hipStream_t st = sl.stream.s;
    PairScalars sc0{};
    for (int z = 0; z < cnt; z++) {
        const size_t q = f + z;
        const PairScalars sc = sample_scalars(c.s[q], c.kappa[q], c.mur,
                                              c.gamma ? &c.gamma[q] : nullptr,
                                              c.ref_l ? &c.ref_l[q] : nullptr,
                                              c.ref_t ? &c.ref_t[q] : nullptr);
        if (z == 0) sc0 = sc;
        sl.hsc[z] = sc;
    }
    if (cnt > 1)
        TAGS_HIP(hipMemcpyAsync(sl.dsc.p, sl.hsc, sizeof(PairScalars) * cnt,
                                hipMemcpyHostToDevice, st));
    // the zero and incidence blocks stay from the last mixed group of this
    // sweep on this slot unless a factor in FP64 replaced W since
    const bool keep = c.mixed && sl.static_sw == sw && sl.static_kind == sw->layout_kind &&
                      sl.static_items >= cnt;
    if ((rc = sweep_fill_items(sw, sl.W.p, st, &sl.prof, cnt, sc0, sl.dsc.p, !keep))) return rc;
    if (!c.mixed) {
        sl.static_items = 0;  // the FP64 factor replaces W
    } else if (!keep) {
        sl.static_sw = sw;
        sl.static_kind = sw->layout_kind;
        sl.static_items = cnt;
    }
    // b = unit current at the injection node
    hipLaunchKernelGGL(k_unit_rhs, dim3((unsigned)((N + 255) / 256), 1, (unsigned)cnt), dim3(256),
                       0, st, sl.rhs.p, N, L.node_off + (long)c.inj_node);
    TAGS_HIP(hipGetLastError());
    rc = c.mixed ? sl.sv.issue_mixed(sl.W.p, sl.rhs.p, cnt, N * N)
                 : sl.sv.issue64(sl.W.p, sl.rhs.p, cnt, N * N);
    return rc ? rc : cnt;
}

// Backward error of the FP64 form's solution x of sample f (W assembled again
// at Wz, which the factor replaced; r = e_inj - W x in FP64).
static int sweep_verify64(tags_sweep* sw, SweepSlot& sl, c64* Wz, const c64* xz, size_t f,
                          const SweepCall& c, tags_item_stats* stt)
{
    const long N = sw->N;
    hipStream_t st = sl.stream.s;
    Solver& sv = sl.sv;
    int rc;
    if ((rc = sweep_fill(sw, Wz, st, nullptr, c.s[f], c.kappa[f], c.mur,
                         c.gamma ? &c.gamma[f] : nullptr, c.ref_l ? &c.ref_l[f] : nullptr,
                         c.ref_t ? &c.ref_t[f] : nullptr)))
        return rc;
    if ((rc = fill_c64(sv.b, N, c64{0.0, 0.0}, st)) ||
        (rc = fill_c64(sv.b + sw->layout.node_off + c.inj_node, 1, c64{1.0, 0.0}, st)) ||
        (rc = fill_c64((c64*)sv.norms, 2, c64{0.0, 0.0}, st)) ||
        (rc = residual64(Wz, N, (int)N, xz, sv.b, sv.r, sv.norms, true, st, nullptr, false)))
        return rc;
    TAGS_HIP(hipMemcpyAsync(sv.hnorms, sv.norms, sizeof(unsigned long long) * 3,
                            hipMemcpyDeviceToHost, st));
    TAGS_HIP(hipStreamSynchronize(st));
    const double rn = sv.norm_at(0, 0), xn = sv.norm_at(0, 1), wn = sv.norm_at(0, 2);
    stt->backward_err = rn == 0.0 ? 0.0 : (wn > 0 && xn > 0 ? rn / (wn * xn) : 1.0);
    return 0;
}

// Advances the group in flight on slot `sl`, whose stream has finished what
// was queued: the mixed form checks its refinement and may queue another step
// (*complete false); once complete the results are written. A singular sample
// (info > 0) gets NaN in u_inj and in its block of x_out.
static int sweep_advance(tags_sweep* sw, SweepSlot& sl, const SweepCall& c, bool* complete)
{
    const long N = sw->N, n = sw->n, m = sw->m;
    const BlockLayout& L = sw->layout;
    const size_t f0 = (size_t)sl.f;
    const int cnt = sl.cnt;
    std::vector<tags_item_stats>& stt = sl.stt;
    std::vector<int>& info = sl.info;
    *complete = false;
    int rc = c.mixed ? sl.sv.advance_mixed(sl.W.p, sl.rhs.p, stt.data(), info.data(), complete)
                     : sl.sv.finish64(stt.data(), info.data());
    if (rc < 0) return rc;
    if (!c.mixed) *complete = true;
    if (!*complete) return 0;
    sl.f = -1;
    hipStream_t st = sl.stream.s;
    if (!c.mixed && sw->verify)
        for (int z = 0; z < cnt; z++)
            if (info[z] == 0 &&
                (rc = sweep_verify64(sw, sl, sl.W.p + (long)z * N * N, sl.rhs.p + (long)z * N,
                                     f0 + z, c, &stt[z])))
                return rc;
    std::vector<c64> xs((size_t)N * cnt);
    TAGS_HIP(hipMemcpyAsync(xs.data(), sl.rhs.p, sizeof(c64) * N * cnt, hipMemcpyDeviceToHost,
                            st));
    TAGS_HIP(hipStreamSynchronize(st));
    for (int z = 0; z < cnt; z++) {
        const size_t f = f0 + z;
        if (c.stats) c.stats[f] = stt[z];
        // a fallback factored this sample's W in FP64
        if (c.mixed && stt[z].precision_used != TAGS_PREC_MIXED) sl.static_items = 0;
        if (sl.nodal &&
            (stt[z].precision_used != TAGS_PREC_MIXED || stt[z].refine_iters > kHybridMaxIters))
            sw->nodal_above = std::max(sw->nodal_above, 1.5 * sl.hz[z]);
        const c64* xz = xs.data() + (size_t)z * N;
        if (info[z] > 0) {
            const c64 nan = c64{NAN, NAN};
            if (c.u_inj) c.u_inj[f] = to_c(nan);
            if (c.x_out)
                for (long i = 0; i < N; i++) c.x_out[f * N + i] = to_c(nan);
            continue;
        }
        if (c.u_inj) c.u_inj[f] = to_c(xz[L.node_off + (long)c.inj_node]);
        if (c.x_out) {
            // reference order [U; I_L; I_T]
            TAGS_COMPLEX* xo = c.x_out + f * N;
            for (long i = 0; i < n; i++) xo[i] = to_c(xz[L.node_off + i]);
            for (long i = 0; i < m; i++) xo[n + i] = to_c(xz[L.zl_off + i]);
            for (long i = 0; i < m; i++) xo[n + m + i] = to_c(xz[L.zt_off + i]);
        }
    }
    return 0;
}

int tags_sweep_solve(tags_sweep* sw, size_t nfreq, const TAGS_COMPLEX* s,
                     const TAGS_COMPLEX* kappa, double mur, const TAGS_COMPLEX* gamma,
                     const TAGS_COMPLEX* ref_l, const TAGS_COMPLEX* ref_t, size_t inj_node,
                     int precision, TAGS_COMPLEX* u_inj, TAGS_COMPLEX* x_out,
                     tags_item_stats* stats)
{
    if (!sw || !s || !kappa) return TAGS_ERR_ARG;
    if ((long)inj_node >= sw->n) return TAGS_ERR_ARG;
    const bool mixed = precision == TAGS_PREC_MIXED;
    // the tuned values of this precision, if any, for this call
    const tags_tuning* tu = sw->tuned[mixed ? 1 : 0] ? &sw->tune[mixed ? 1 : 0] : nullptr;
    TuningScope scope(tu ? &sw->ktune[mixed ? 1 : 0] : tuning_current());
    const int batch_req0 = sw->batch_req;
    if (tu && tu->batch >= 0) sw->batch_req = tu->batch;
    // groups of samples share launches; grounded plans take one sample at a time
    sw->batch = (sw->plan_kind == TAGS_PLAN_GROUNDED ||
                 sw->plan_kind == TAGS_PLAN_GROUNDED_PIVOTED) && mixed
                    ? 1
                    : sweep_batch(sw, mixed);
    sw->batch_req = batch_req0;
    if (lu_panel_kind() == PANEL_PARTIAL) sw->batch = 1;
    const int k = tu && tu->inflight > 0 ? tu->inflight : sw->inflight;
    // every configured slot is allocated up front (a warm-up call sets all of
    // them up)
    int rc = sweep_slots(sw, k, mixed);
    if (rc || (rc = sweep_potentials(sw))) return rc;
    // after the slots exist: the layout sets every slot's plans
    if ((rc = sweep_layout(sw))) return rc;
    const SweepCall c{nfreq, s, kappa, gamma, ref_l, ref_t, mur, inj_node, mixed, u_inj, x_out,
                      stats};
    // the hybrid plan's switch may rise during this call; it is restored at
    // the end, so that every call starts from the configured switch
    const double nodal_above0 = sw->nodal_above;
    const int gmres_steps0 = sw->gmres_steps;
    if (tu && mixed && tu->nodal_above >= 0.0) sw->nodal_above = tu->nodal_above;
    if (tu && mixed && tu->gmres_steps > 0) sw->gmres_steps = tu->gmres_steps;
    size_t next = 0;
    for (int i = 0; i < k && next < nfreq; i++) {
        const int got = sweep_issue(sw, *sw->slots[i], next, c);
        if (got < 0) {
            rc = got;
            break;
        }
        next += (size_t)got;
    }
    // advance whichever slot's stream is idle; refill a slot whose group is done
    while (!rc) {
        int busy = 0, ready = -1;
        for (int i = 0; i < k; i++) {
            if (sw->slots[i]->f < 0) continue;
            busy++;
            const hipError_t q = hipStreamQuery(sw->slots[i]->stream.s);
            if (q == hipSuccess) {
                ready = i;
                break;
            }
            if (q != hipErrorNotReady) TAGS_HIP(q);
        }
        if (!busy) break;
        if (ready < 0) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            continue;
        }
        SweepSlot& sl = *sw->slots[ready];
        bool complete = false;
        if ((rc = sweep_advance(sw, sl, c, &complete))) break;
        if (complete && next < nfreq) {
            const int got = sweep_issue(sw, sl, next, c);
            if (got < 0) {
                rc = got;
                break;
            }
            next += (size_t)got;
        }
    }
    // after an error: let every queued sample end before returning
    sw->nodal_above = nodal_above0;
    sw->gmres_steps = gmres_steps0;
    for (int i = 0; i < k; i++) {
        TAGS_HIP(hipStreamSynchronize(sw->slots[i]->stream.s));
        sw->slots[i]->f = -1;
    }
    return rc;
}

int tags_sweep_set_verify(tags_sweep* sw, int on)
{
    if (!sw) return TAGS_ERR_ARG;
    sw->verify = on != 0;
    return 0;
}

int tags_sweep_set_gmres(tags_sweep* sw, double below_hz, int steps)
{
    if (!sw || !(below_hz >= 0.0) || steps < 1 || steps > 32) return TAGS_ERR_ARG;
    sw->gmres_below = below_hz;
    sw->gmres_steps = steps;
    return 0;
}

int tags_sweep_set_batch(tags_sweep* sw, int samples)
{
    if (!sw || samples < 0 || samples > 4096) return TAGS_ERR_ARG;
    sw->batch_req = samples;
    return 0;
}

}  // extern "C"


// ---------------------------------------------------------------------------
// Autotuner: tile shapes on the sweep's own update shapes, then the schedule by
// coordinate descent over timed probe solves, with an on-disk cache.
// ---------------------------------------------------------------------------
namespace {

constexpr double kTuneGain = 0.97;  // a candidate must be 3 % faster to replace

bool tuning_valid(const tags_tuning& t, int precision)
{
    const int nv = gemm_variant_count(precision == TAGS_PREC_MIXED ? 1 : 0);
    if (t.gemm_big >= nv || t.gemm_small >= nv) return false;
    if (t.step_width != 0 && t.step_width != 128 && t.step_width != 256) return false;
    if (t.update_chunks == 0 || t.update_chunks > 64) return false;
    if (t.solve_rows > 2) return false;
    if (t.inflight < 0 || t.inflight > 4) return false;
    if (t.batch > 4096) return false;
    if (t.gmres_steps < 0 || t.gmres_steps > 32) return false;
    return true;
}

void tuning_kernel_part(const tags_tuning& t, int precision, Tuning* k)
{
    const int p = precision == TAGS_PREC_MIXED ? 1 : 0;
    *k = Tuning();
    k->gemm_big[p] = t.gemm_big;
    k->gemm_small[p] = t.gemm_small;
    k->step_width[p] = t.step_width;
    k->update_chunks = t.update_chunks;
    k->solve_rows = t.solve_rows;
}

std::string tune_key(const tags_sweep* sw, int precision)
{
    int dev = 0;
    hipDeviceProp_t p;
    std::string arch = "gpu";
    int cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) {
        arch = p.gcnArchName;
        const size_t colon = arch.find(':');
        if (colon != std::string::npos) arch.resize(colon);
        cus = p.multiProcessorCount;
    }
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s cu=%d form=%s plan=%d prec=%s N=%ld m=%ld", arch.c_str(),
                  cus, sw->formulation == TAGS_FORM_MHEM ? "mhem" : "hem", sw->plan_kind,
                  precision == TAGS_PREC_MIXED ? "mixed" : "fp64", sw->N, sw->m);
    return buf;
}

std::string tune_value(const tags_tuning& t)
{
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "big=%d small=%d width=%d chunks=%d rows=%d inflight=%d batch=%d "
                  "nodal_above=%.6g gmres=%d ms=%.4f",
                  t.gemm_big, t.gemm_small, t.step_width, t.update_chunks, t.solve_rows,
                  t.inflight, t.batch, t.nodal_above, t.gmres_steps, t.ms_per_sample);
    return buf;
}

// "<key> | <value>" lines; '#' starts a comment line
bool tune_cache_read(const char* path, const std::string& key, tags_tuning* t)
{
    FILE* f = std::fopen(path, "r");
    if (!f) return false;
    char line[1024];
    bool found = false;
    while (std::fgets(line, sizeof(line), f)) {
        std::string l(line);
        while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
        if (l.empty() || l[0] == '#') continue;
        const size_t bar = l.find(" | ");
        if (bar == std::string::npos || l.compare(0, bar, key) != 0 || bar != key.size()) continue;
        tags_tuning v{};
        if (std::sscanf(l.c_str() + bar + 3,
                        "big=%d small=%d width=%d chunks=%d rows=%d inflight=%d batch=%d "
                        "nodal_above=%lf gmres=%d ms=%lf",
                        &v.gemm_big, &v.gemm_small, &v.step_width, &v.update_chunks,
                        &v.solve_rows, &v.inflight, &v.batch, &v.nodal_above,
                        &v.gmres_steps, &v.ms_per_sample) == 10) {
            *t = v;
            found = true;
        }
    }
    std::fclose(f);
    return found;
}

int tune_cache_write(const char* path, const std::string& key, const tags_tuning& t)
{
    std::vector<std::string> keep;
    if (FILE* f = std::fopen(path, "r")) {
        char line[1024];
        while (std::fgets(line, sizeof(line), f)) {
            std::string l(line);
            while (!l.empty() && (l.back() == '\n' || l.back() == '\r')) l.pop_back();
            const size_t bar = l.find(" | ");
            if (!l.empty() && l[0] != '#' && bar == key.size() && l.compare(0, bar, key) == 0)
                continue;
            keep.push_back(l);
        }
        std::fclose(f);
    }
    keep.push_back(key + " | " + tune_value(t));
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return TAGS_ERR_ARG;
    for (const std::string& l : keep) std::fprintf(f, "%s\n", l.c_str());
    if (std::fclose(f) != 0 || std::rename(tmp.c_str(), path) != 0) return TAGS_ERR_ARG;
    return 0;
}

// The fastest variant for C -= A B (M x N x K) in the given precision; *best =
// its index (-1 when none ran), *best_ms its mean ms per launch. With
// mid > 0 a variant is scored by its time at M x N plus mid times its time at
// (M / 2) x (N / 2): an LU's trailing updates shrink step by step, and a
// variant that wins only at the first one can lose over the rest.
int tune_gemm(int precision, int M, int N, int K, int iters, int* best, double* best_ms,
              double mid = 0.0, int* second = nullptr)
{
    *best = -1;
    *best_ms = 1e30;
    double best_score = 1e30, second_score = 1e30;
    if (second) *second = -1;
    const int nv = gemm_variant_count(precision == TAGS_PREC_MIXED ? 1 : 0);
    auto bench = [&](int m, int n, int v, double* ms) {
        return precision == TAGS_PREC_MIXED ? gemm_bench_t<float>(m, n, K, v, iters, ms)
                                            : gemm_bench_t<double>(m, n, K, v, iters, ms);
    };
    for (int v = 0; v < nv; v++) {
        double ms = 0, ms2 = 0;
        int rc = bench(M, N, v, &ms);
        if (rc == TAGS_ERR_ARG) continue;  // a variant that does not take this shape
        if (rc) return rc;
        double score = ms;
        if (mid > 0) {
            rc = bench(std::max(64, M / 2), std::max(64, N / 2), v, &ms2);
            if (rc == TAGS_ERR_ARG) continue;
            if (rc) return rc;
            score += mid * ms2;
        }
        if (score < best_score) {
            if (second) *second = *best;
            second_score = best_score;
            best_score = score;
            *best_ms = ms;
            *best = v;
        } else if (score < second_score) {
            second_score = score;
            if (second) *second = v;
        }
    }
    return 0;
}

}  // namespace

extern "C" {

int tags_sweep_set_tuning(tags_sweep* sw, int precision, const tags_tuning* t)
{
    if (!sw || (precision != TAGS_PREC_FP64 && precision != TAGS_PREC_MIXED)) return TAGS_ERR_ARG;
    const int p = precision == TAGS_PREC_MIXED ? 1 : 0;
    if (!t) {
        sw->tuned[p] = false;
        return 0;
    }
    if (!tuning_valid(*t, precision)) return TAGS_ERR_ARG;
    sw->tune[p] = *t;
    tuning_kernel_part(*t, precision, &sw->ktune[p]);
    sw->tuned[p] = true;
    return 0;
}

int tags_sweep_get_tuning(const tags_sweep* sw, int precision, tags_tuning* t)
{
    if (!sw || !t || (precision != TAGS_PREC_FP64 && precision != TAGS_PREC_MIXED))
        return TAGS_ERR_ARG;
    const int p = precision == TAGS_PREC_MIXED ? 1 : 0;
    if (!sw->tuned[p]) return TAGS_ERR_ARG;
    *t = sw->tune[p];
    return 0;
}

int tags_sweep_autotune(tags_sweep* sw, int precision, size_t nprobe, const TAGS_COMPLEX* s,
                        const TAGS_COMPLEX* kappa, double mur, size_t inj_node,
                        const char* cache_path, int force, tags_tuning* out)
{
    if (!sw || !s || !kappa || nprobe < 1 || (long)inj_node >= sw->n ||
        (precision != TAGS_PREC_FP64 && precision != TAGS_PREC_MIXED))
        return TAGS_ERR_ARG;
    const bool mixed = precision == TAGS_PREC_MIXED;
    const int p = mixed ? 1 : 0;
    const std::string key = tune_key(sw, precision);
    if (cache_path && !force) {
        tags_tuning t{};
        if (tune_cache_read(cache_path, key, &t) && tuning_valid(t, precision)) {
            t.from_cache = 1;
            t.evaluations = 0;
            int rc = tags_sweep_set_tuning(sw, precision, &t);
            if (rc) return rc;
            if (out) *out = t;
            return 0;
        }
    }
    // 1. tile variants on the plan's own update shapes, per step width
    int rc = sweep_layout(sw);
    if (rc) return rc;
    const LUPlan* plan = mixed ? (sw->plan_kind == TAGS_PLAN_HYBRID ? &sw->plan_nodal
                                  : sw->bind32                      ? sw->bind32
                                                                    : sw->bind64)
                               : sw->bind64;
    if (!plan || plan->phases.empty()) return TAGS_ERR_ARG;
    const LUPhase* wide = &plan->phases[0];
    auto cols = [](const LUPhase& ph) {
        return (ph.c_end - ph.c_begin) + std::max(0, ph.x_end - ph.x_begin);
    };
    for (const LUPhase& ph : plan->phases)
        if (cols(ph) > cols(*wide)) wide = &ph;
    const int widths[2] = {128, 256};
    int big[2] = {-1, -1}, small[2] = {-1, -1}, big2[2] = {-1, -1};
    for (int wi = 0; wi < 2; wi++) {
        const int w = widths[wi];
        const int M = std::min(12288, std::max(64, wide->row_end - wide->c_begin - w));
        const int N = std::min(12288, std::max(64, cols(*wide) - w));
        double ms = 0;
        if ((rc = tune_gemm(precision, M, N, w, 3, &big[wi], &ms, 2.0, &big2[wi]))) return rc;
        const int Ms = std::min(1024, M), Ns = std::min(1024, N);
        if ((rc = tune_gemm(precision, Ms, Ns, w, 10, &small[wi], &ms))) return rc;
        if (big[wi] < 0 || small[wi] < 0) return TAGS_ERR_ARG;
    }
    // 2. the schedule, by coordinate descent over timed probe solves
    const bool had = sw->tuned[p];
    const tags_tuning saved = sw->tune[p];
    std::vector<TAGS_COMPLEX> u(nprobe);
    int evals = 0;
    auto eval = [&](const tags_tuning& t, double* ms) -> int {
        int r = tags_sweep_set_tuning(sw, precision, &t);
        if (r) return r;
        // the warm-up sets up every slot (k groups of up to b samples)
        const int k = t.inflight > 0 ? t.inflight : sw->inflight;
        const size_t nw = std::min(nprobe, (size_t)k * (size_t)std::max(1, t.batch));
        r = tags_sweep_solve(sw, nw, s, kappa, mur, nullptr, nullptr, nullptr, inj_node,
                             precision, u.data(), nullptr, nullptr);
        if (r) return r;
        const auto t0 = std::chrono::steady_clock::now();
        r = tags_sweep_solve(sw, nprobe, s, kappa, mur, nullptr, nullptr, nullptr, inj_node,
                             precision, u.data(), nullptr, nullptr);
        const auto t1 = std::chrono::steady_clock::now();
        if (r) return r;
        evals++;
        *ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / (double)nprobe;
        for (size_t i = 0; i < nprobe; i++)
            if (!std::isfinite(__real__ u[i]) || !std::isfinite(__imag__ u[i])) *ms = 1e30;
        return 0;
    };
    const int w0 = lu_step_width(mixed) == 128 ? 0 : 1;
    tags_tuning cur{};
    cur.gemm_big = big[w0];
    cur.gemm_small = small[w0];
    cur.step_width = widths[w0];
    cur.update_chunks = -1;
    cur.solve_rows = -1;
    cur.inflight = sw->inflight;
    cur.batch = -1;
    cur.nodal_above = mixed ? sw->nodal_above : -1.0;
    cur.gmres_steps = 0;
    double best = 0;
    auto fail = [&](int r) {
        if (had) tags_sweep_set_tuning(sw, precision, &saved);
        else sw->tuned[p] = false;
        return r;
    };
    if ((rc = eval(cur, &best))) return fail(rc);
    auto consider = [&](tags_tuning cand) -> int {
        double ms = 0;
        const int r = eval(cand, &ms);
        if (r == TAGS_ERR_HIP) {
            // this candidate did not fit (memory): release what it took
            (void)hipGetLastError();
            return 0;
        }
        if (r) return r;
        if (ms < kTuneGain * best) {
            best = ms;
            cur = cand;
        }
        return 0;
    };
    {
        tags_tuning c = cur;
        c.step_width = widths[1 - w0];
        c.gemm_big = big[1 - w0];
        c.gemm_small = small[1 - w0];
        if ((rc = consider(c))) return fail(rc);
    }
    {
        // the runner-up of the product timings, on the probe solves (the
        // timings see two product sizes, the solves the whole factorisation)
        const int wi = cur.step_width == widths[0] ? 0 : 1;
        if (big2[wi] >= 0 && big2[wi] != cur.gemm_big) {
            tags_tuning c = cur;
            c.gemm_big = big2[wi];
            if ((rc = consider(c))) return fail(rc);
        }
    }
    for (int ch : {1, 2, 8}) {
        tags_tuning c = cur;
        c.update_chunks = ch;
        if ((rc = consider(c))) return fail(rc);
    }
    for (int rows : {0, 1}) {
        tags_tuning c = cur;
        c.solve_rows = rows;
        if ((rc = consider(c))) return fail(rc);
    }
    for (int k = 1; k <= 4; k++) {
        if (k == cur.inflight) continue;
        tags_tuning c = cur;
        c.inflight = k;
        if ((rc = consider(c))) return fail(rc);
    }
    if (sw->N < kBatchOneAbove) {
        const int b0 = sweep_batch(sw, mixed);
        for (int b : {b0 / 2, b0 / 4}) {
            if (b < 1 || b == b0) continue;
            tags_tuning c = cur;
            c.batch = b;
            if ((rc = consider(c))) return fail(rc);
        }
    } else {
        // large systems take one sample per group by default; groups of two
        // to four share each launch (and its place on the panel chain) among
        // their samples, as far as k groups' W (and W32) fit in 80 % of the
        // device's memory
        size_t mem_free = 0, mem_total = 0;
        if (hipMemGetInfo(&mem_free, &mem_total) != hipSuccess) mem_total = 0;
        const double per = (double)sw->N * (double)sw->N * (mixed ? 24.0 : 16.0);
        const int k = cur.inflight > 0 ? cur.inflight : sw->inflight;
        for (int b = 2; b <= 4; b++) {
            if ((double)b * k * per > 0.8 * (double)mem_total) break;
            tags_tuning c = cur;
            c.batch = b;
            if ((rc = consider(c))) return fail(rc);
        }
    }
    if (mixed && sw->plan_kind == TAGS_PLAN_HYBRID) {
        tags_tuning c = cur;
        c.nodal_above = cur.nodal_above > 0.0 ? 0.0 : 2e4;
        if ((rc = consider(c))) return fail(rc);
    }
    if (mixed && (sw->plan_kind == TAGS_PLAN_HYBRID || sw->plan_kind == TAGS_PLAN_NODAL)) {
        for (int k : {6, 8, 10}) {
            if (k == (cur.gmres_steps > 0 ? cur.gmres_steps : sw->gmres_steps)) continue;
            tags_tuning c = cur;
            c.gmres_steps = k;
            if ((rc = consider(c))) return fail(rc);
        }
    }
    cur.ms_per_sample = best;
    cur.from_cache = 0;
    cur.evaluations = evals;
    if ((rc = tags_sweep_set_tuning(sw, precision, &cur))) return rc;
    if (cache_path && (rc = tune_cache_write(cache_path, key, cur))) return rc;
    if (out) *out = cur;
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Many grounding cases: every (case, frequency) sample of every case, in
// groups of one case's consecutive samples, scheduled over a pool of lanes.
// The reference runs one case per driver run, its frequencies in an OpenMP
// loop with per-thread buffers (examples/harmonic_impedance.c:87-176,
// examples/grcev12pwrd01.c:88-97); here the groups of all cases form one
// queue, largest first (longest-processing-time order by the per-sample cost
// (8/3) N^3 + 6400 m^2: the factorisation and the pair quadrature), and each
// lane (own stream, buffers and solver) takes the next group as soon as its
// last one is finished.
// ---------------------------------------------------------------------------
struct tags_cases {
    std::vector<tags_sweep*> sw;
    std::vector<tags_case> cs;
    std::vector<std::unique_ptr<SweepSlot>> lanes;
    int nlanes = 3;
    int plan_kind = TAGS_PLAN_HYBRID;
    bool profiling = false;
    ~tags_cases()
    {
        lanes.clear();
        for (tags_sweep* p : sw) tags_sweep_destroy(p);
    }
};

extern "C" {

int tags_cases_create(tags_cases** out, const tags_case* cases, size_t ncases)
{
    if (!out || (ncases && !cases)) return TAGS_ERR_ARG;
    std::unique_ptr<tags_cases> cs(new tags_cases());
    for (size_t c = 0; c < ncases; c++) {
        const tags_case& d = cases[c];
        if (d.inj_node >= d.num_nodes || (d.nfreq && (!d.s || !d.kappa || !d.u_inj)))
            return TAGS_ERR_ARG;
        tags_sweep* p = nullptr;
        int rc = tags_sweep_create(&p, d.electrodes, d.images, d.num_electrodes, d.nodes,
                                   d.num_nodes);
        if (rc) return rc;
        cs->sw.push_back(p);
        cs->cs.push_back(d);
    }
    *out = cs.release();
    return 0;
}

int tags_cases_set_plan(tags_cases* cs, int plan)
{
    if (!cs || plan < 0 || plan > TAGS_PLAN_HYBRID) return TAGS_ERR_ARG;
    cs->plan_kind = plan;
    return 0;
}

int tags_cases_set_lanes(tags_cases* cs, int lanes)
{
    if (!cs || lanes < 1 || lanes > 4) return TAGS_ERR_ARG;
    cs->nlanes = lanes;
    return 0;
}

int tags_cases_autotune(tags_cases* cs, int precision, const char* cache_path, int force,
                        tags_tuning* classes, int* nclasses)
{
    if (!cs || (precision != TAGS_PREC_FP64 && precision != TAGS_PREC_MIXED)) return TAGS_ERR_ARG;
    const size_t nc = cs->cs.size();
    // class of order N: 0 below 256, then one per doubling, the last from 8192
    auto class_of = [](long N) {
        int k = 0;
        for (long lim = 256; N >= lim && k < 6; lim *= 2) k++;
        return k;
    };
    std::vector<long> rep(7, -1);
    for (size_t c = 0; c < nc; c++) {
        const int k = class_of(cs->sw[c]->N);
        if (cs->cs[c].nfreq == 0) continue;
        if (rep[k] < 0 || cs->sw[c]->N > cs->sw[rep[k]]->N) rep[k] = (long)c;
    }
    int nout = 0;
    for (int k = 0; k < 7; k++) {
        if (rep[k] < 0) continue;
        const size_t c = (size_t)rep[k];
        tags_sweep* sw = cs->sw[c];
        const tags_case& d = cs->cs[c];
        sw->plan_kind = cs->plan_kind;
        // probe: up to 64 of the case's samples, spread over them
        const size_t step = (d.nfreq + 63) / 64;
        std::vector<TAGS_COMPLEX> ps, pk;
        for (size_t f = 0; f < d.nfreq; f += step) {
            ps.push_back(d.s[f]);
            pk.push_back(d.kappa[f]);
        }
        tags_tuning t{};
        int rc = tags_sweep_autotune(sw, precision, ps.size(), ps.data(), pk.data(), d.mur,
                                     d.inj_node, cache_path, force, &t);
        // the case's own slots served the timing only: the lanes solve it
        sw->slots.clear();
        if (rc) return rc;
        tags_tuning inst = t;
        inst.inflight = 0;  // lanes, not slots
        for (size_t q = 0; q < nc; q++)
            if (class_of(cs->sw[q]->N) == k && (rc = tags_sweep_set_tuning(cs->sw[q], precision,
                                                                           &inst)))
                return rc;
        if (classes && nout < 16) classes[nout] = t;
        nout++;
    }
    if (nclasses) *nclasses = nout;
    return 0;
}

int tags_cases_set_profiling(tags_cases* cs, int on)
{
    if (!cs) return TAGS_ERR_ARG;
    cs->profiling = on != 0;
    for (auto& l : cs->lanes) {
        TAGS_HIP(hipStreamSynchronize(l->stream.s));
        l->prof.on = cs->profiling;
        l->prof.reset();
    }
    return 0;
}

int tags_cases_get_timers(tags_cases* cs, tags_timers* t)
{
    if (!cs || !t) return TAGS_ERR_ARG;
    std::memset(t, 0, sizeof(*t));
    for (auto& l : cs->lanes) {
        TAGS_HIP(hipStreamSynchronize(l->stream.s));
        Profiler& p = l->prof;
        p.resolve();
        t->ms_fill += p.ms[FAM_FILL];
        t->ms_panel += p.ms[FAM_PANEL];
        t->ms_gemm += p.ms[FAM_GEMM];
        t->ms_other += p.ms[FAM_OTHER];
        t->ms_solve += p.ms[FAM_SOLVE];
        t->ms_residual += p.ms[FAM_RESID];
        t->gemm_flops += p.gemm_flops;
        t->trailing_ms += p.trailing_ms;
        t->trailing_flops += p.trailing_flops;
        t->trailing_launches += p.trailing_launches;
    }
    return 0;
}

int tags_cases_solve(tags_cases* cs, int precision)
{
    if (!cs) return TAGS_ERR_ARG;
    const bool mixed = precision == TAGS_PREC_MIXED;
    const size_t nc = cs->cs.size();
    int rc;
    // per case: plans, layout and group size
    std::vector<SweepCall> call(nc);
    std::vector<double> cost(nc), nodal0(nc);
    std::vector<int> gmres0(nc);
    std::vector<size_t> next(nc, 0);
    for (size_t c = 0; c < nc; c++) {
        tags_sweep* sw = cs->sw[c];
        const tags_case& d = cs->cs[c];
        sw->plan_kind = cs->plan_kind;
        if ((rc = sweep_layout(sw))) return rc;
        const tags_tuning* tu = sw->tuned[mixed ? 1 : 0] ? &sw->tune[mixed ? 1 : 0] : nullptr;
        const int batch_req0 = sw->batch_req;
        if (tu && tu->batch >= 0) sw->batch_req = tu->batch;
        sw->batch = (sw->plan_kind == TAGS_PLAN_GROUNDED ||
                     sw->plan_kind == TAGS_PLAN_GROUNDED_PIVOTED) && mixed
                        ? 1
                        : sweep_batch(sw, mixed);
        sw->batch_req = batch_req0;
        if (lu_panel_kind() == PANEL_PARTIAL) sw->batch = 1;
        if ((rc = sweep_potentials(sw))) return rc;
        call[c] = SweepCall{d.nfreq, d.s, d.kappa, nullptr, nullptr, nullptr, d.mur, d.inj_node,
                            mixed, d.u_inj, nullptr, d.stats};
        const double N = (double)sw->N, m = (double)sw->m;
        cost[c] = 8.0 / 3.0 * N * N * N + 6400.0 * m * m;
        nodal0[c] = sw->nodal_above;
        gmres0[c] = sw->gmres_steps;
        if (tu && mixed && tu->nodal_above >= 0.0) sw->nodal_above = tu->nodal_above;
        if (tu && mixed && tu->gmres_steps > 0) sw->gmres_steps = tu->gmres_steps;
    }
    // the queue: cases by per-sample cost, largest first; a case's groups are
    // taken in frequency order
    std::vector<size_t> order(nc);
    for (size_t c = 0; c < nc; c++) order[c] = c;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    size_t qi = 0;  // first case in `order` with samples left
    while ((int)cs->lanes.size() < cs->nlanes) cs->lanes.emplace_back(new SweepSlot());
    // every lane's buffers sized for all cases before any work is queued: a
    // group size chosen by an N^2 budget makes N times the group grow as N
    // falls, and a reallocation between cases would wait for the whole
    // device. The cases with the largest matrix, vector and group needs come
    // first (buffers only grow), then the rest.
    {
        std::vector<size_t> prep(order);
        auto pick = [&](auto key) {
            size_t best = 0;
            for (size_t c = 1; c < nc; c++)
                if (key(c) > key(best)) best = c;
            return best;
        };
        if (nc) {
            auto nn = [&](size_t c) { return (double)cs->sw[c]->N; };
            auto bb = [&](size_t c) { return (double)cs->sw[c]->batch; };
            prep.insert(prep.begin(), {pick([&](size_t c) { return nn(c) * nn(c) * bb(c); }),
                                       pick([&](size_t c) { return nn(c) * bb(c); }),
                                       pick([&](size_t c) { return bb(c) * 1e6 + nn(c); })});
        }
        for (auto& lp : cs->lanes) {
            SweepSlot& L = *lp;
            for (size_t c : prep) {
                tags_sweep* sw = cs->sw[c];
                if ((rc = L.init(sw->N, sw->batch)) ||
                    (rc = L.sv.ensure((int)sw->N, mixed, sw->batch)))
                    return rc;
                if (mixed && (sw->plan_kind == TAGS_PLAN_HYBRID || sw->plan_kind == TAGS_PLAN_NODAL) &&
                    (rc = L.sv.gm.alloc((int)sw->N, sw->gmres_steps, sw->batch)))
                    return rc;
            }
        }
    }
    std::vector<long> lane_case(cs->nlanes, -1);
    auto issue = [&](int li) -> int {
        while (qi < nc && next[order[qi]] >= cs->cs[order[qi]].nfreq) qi++;
        if (qi >= nc) return 0;
        const size_t c = order[qi];
        tags_sweep* sw = cs->sw[c];
        SweepSlot& L = *cs->lanes[li];
        int r = L.init(sw->N, sw->batch);
        if (r || (r = L.sv.ensure((int)sw->N, mixed, sw->batch))) return r;
        L.prof.on = cs->profiling;
        slot_bind(sw, L);
        // the case's tuned tile shapes and schedule, if any
        TuningScope scope(sw->tuned[mixed ? 1 : 0] ? &sw->ktune[mixed ? 1 : 0] : tuning_current());
        const int got = sweep_issue(sw, L, next[c], call[c]);
        if (got < 0) return got;
        next[c] += (size_t)got;
        lane_case[li] = (long)c;
        return 0;
    };
    rc = 0;
    for (int li = 0; li < cs->nlanes && !rc; li++) rc = issue(li);
    // finish whichever lane is done first and refill it
    while (!rc) {
        int busy = 0, done = -1;
        for (int li = 0; li < cs->nlanes; li++) {
            if (cs->lanes[li]->f < 0) continue;
            busy++;
            const hipError_t q = hipStreamQuery(cs->lanes[li]->stream.s);
            if (q == hipSuccess) {
                done = li;
                break;
            }
            if (q != hipErrorNotReady) {
                TAGS_HIP(q);
            }
        }
        if (!busy) break;
        if (done < 0) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            continue;
        }
        SweepSlot& L = *cs->lanes[done];
        const size_t c = (size_t)lane_case[done];
        bool complete = false;
        {
            tags_sweep* sw = cs->sw[c];
            TuningScope scope(sw->tuned[mixed ? 1 : 0] ? &sw->ktune[mixed ? 1 : 0]
                                                       : tuning_current());
            if ((rc = sweep_advance(sw, L, call[c], &complete))) break;
        }
        if (!complete) continue;
        lane_case[done] = -1;
        rc = issue(done);
    }
    for (size_t c = 0; c < nc; c++) {
        cs->sw[c]->nodal_above = nodal0[c];
        cs->sw[c]->gmres_steps = gmres0[c];
    }
    for (auto& l : cs->lanes) {
        TAGS_HIP(hipStreamSynchronize(l->stream.s));
        l->f = -1;
    }
    return rc;
}

void tags_cases_destroy(tags_cases* cs)
{
    delete cs;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// calculate_impedances without allocation, for callers that cannot allocate
// device memory (work captured in HIP graphs, tight serving loops): device
// pointers, the caller's stream, and the segment geometry staged as structure
// of arrays in a fixed device array instead of a per-call buffer.
// ---------------------------------------------------------------------------
namespace tags {
// the staging array: 8 doubles per segment (start xyz, end xyz, length,
// radius), for at most kNoAllocMaxSegments segments (4 MiB per device)
constexpr long kNoAllocMaxSegments = 65536;
static __device__ __align__(256) double scratch[8 * kNoAllocMaxSegments];

// device address of `scratch` per device, taken once with that device current
constexpr int kNoAllocMaxDevices = 64;
static double* g_scratch_ptr[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_ptr(double** out)
{
    int dev = 0;
    TAGS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return TAGS_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        TAGS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return TAGS_ERR_HIP;
        // set-up of the first call on this device: the quadrature tables too
        int rc = gl_tables_init();
        if (rc) return rc;
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// soa[k m + i] = field k of segment i (start xyz, end xyz, length, radius)
__global__ void k_segments_soa(const Electrode* __restrict__ e, long m, double* soa)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    const Electrode s = e[i];
    soa[0 * m + i] = s.start_point[0];
    soa[1 * m + i] = s.start_point[1];
    soa[2 * m + i] = s.start_point[2];
    soa[3 * m + i] = s.end_point[0];
    soa[4 * m + i] = s.end_point[1];
    soa[5 * m + i] = s.end_point[2];
    soa[6 * m + i] = s.length;
    soa[7 * m + i] = s.radius;
}
}  // namespace tags

extern "C" {

int calculate_impedances_noalloc(TAGS_COMPLEX* zl, TAGS_COMPLEX* zt, const Electrode* electrodes,
                                 size_t num_electrodes, TAGS_COMPLEX gamma, TAGS_COMPLEX s,
                                 double mur, TAGS_COMPLEX kappa, size_t max_eval,
                                 double req_abs_error, double req_rel_error,
                                 int integration_type, void* stream)
{
    (void)max_eval;
    (void)req_abs_error;
    (void)req_rel_error;
    if (integration_type != INTG_DOUBLE && integration_type != INTG_NONE &&
        integration_type != INTG_MHEM && integration_type != INTG_SINGLE) {
        std::printf("integration failed with code: %i\n", -10);
        return -10;
    }
    const long m = (long)num_electrodes;
    if (m == 0) return 0;
    if (!zl || !zt || !electrodes) return TAGS_ERR_ARG;
    if (m > kNoAllocMaxSegments) {
        std::fprintf(stderr,
                     "calculate_impedances_noalloc: %ld segments exceed the staging array's "
                     "maximum of %ld\n",
                     m, kNoAllocMaxSegments);
        return TAGS_ERR_TOO_LARGE;
    }
    double* soa = nullptr;
    int rc = scratch_ptr(&soa);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_segments_soa, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st,
                       electrodes, m, soa);
    TAGS_HIP(hipGetLastError());
    const SegSoA g{soa, soa + m, soa + 2 * m, soa + 3 * m, soa + 4 * m, soa + 5 * m, soa + 6 * m,
                   soa + 7 * m};
    const PairScalars sc = make_scalars(from_c(gamma), from_c(s), mur, from_c(kappa), c64{1, 0},
                                        c64{1, 0}, integration_type);
    return launch_pairs_dropin((double*)zl, (double*)zt, (void*)zl == (void*)zt, g, g, m, sc,
                               false, st);
}

}  // extern "C"
