// Potentials, fields and voltages of the segment currents at arbitrary points
// (src/electrode.c:177-216, :433-668) and the examples' mHEM ground potential
// (examples/visacro57emc01.c:406-432), batched over points and right-hand
// sides on the GPU.
//
// Per (point, segment) pair two line integrals over the segment axis
// t in [0, 1] (field point P, axis point q(t), r = |P - q(t)|):
//   Ks = int e^{-gamma r} / r dt                      (INTG_SINGLE / L)
//   Kf = int (1 + gamma r) e^{-gamma r} / r^2 (P - q) / r dt
// The integrands are analytic in t except near the branch points
// t0 +- i d / L (t0: the closest axis parameter, d: the distance to the
// segment). The rule is composite 16-point Gauss-Legendre on panels graded
// geometrically away from t0 on either side: the first panel of width d / L,
// then each panel as wide as its distance from t0 (so that every panel sees
// the branch points at least one half-width outside its Bernstein ellipse of
// parameter ~4, about 1e-19 relative per panel) and at most 2 / (|gamma| L)
// wide (the oscillation of e^{-gamma r}).
//
// A work-group owns one field point and up to 4 right-hand sides; its threads
// stride over the segments and the sums are reduced in LDS.
#include <cmath>
#include <mutex>
#include <vector>

#include "tags_common.h"

namespace tags {

__constant__ double f_glx[16];
__constant__ double f_glw[16];
static bool f_gl_ready = false;
static std::mutex f_gl_mutex;

static int field_tables_init()
{
    std::lock_guard<std::mutex> lock(f_gl_mutex);
    if (f_gl_ready) return 0;
    double x[16], w[16];
    gauss_legendre(16, x, w);
    TAGS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(f_glx), x, sizeof(x)));
    TAGS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(f_glw), w, sizeof(w)));
    f_gl_ready = true;
    return 0;
}

constexpr int kFieldRhs = 4;  // right-hand sides per work-group

struct FieldArgs {
    const double* pts;  // np x 3
    long np;
    SegSoA e;
    long m;
    const c64* il;  // nrhs x m (right-hand side k at k m)
    const c64* it;
    int nrhs;
    c64 gamma;
    c64* ks;  // per point and rhs: sum_m it Ks            (np x nrhs)
    c64* ka;  // per point and rhs: sum_m il Ks (end - start) (np x nrhs x 3)
    c64* kf;  // per point and rhs: sum_m it Kf            (np x nrhs x 3)
};

// Ks and Kf of one segment at point P
__device__ void segment_kernels(const double P[3], const double a[3], const double b[3],
                                double len, c64 g, c64& ks, c64 kf[3])
{
    double d[3], w[3];
    double dd = 0.0, tw = 0.0;
    for (int i = 0; i < 3; i++) {
        d[i] = b[i] - a[i];
        w[i] = P[i] - a[i];
        dd += d[i] * d[i];
        tw += w[i] * d[i];
    }
    double t0 = dd > 0.0 ? tw / dd : 0.0;
    t0 = fmin(fmax(t0, 0.0), 1.0);
    double dist2 = 0.0;
    for (int i = 0; i < 3; i++) {
        const double q = a[i] + t0 * d[i] - P[i];
        dist2 += q * q;
    }
    const double L = sqrt(dd) > 0.0 ? sqrt(dd) : len;
    const double wmin = fmax(sqrt(dist2) / L, 1e-14);
    const double ga = sqrt(g.x * g.x + g.y * g.y);
    const double wmax = ga * L > 2.0 ? 2.0 / (ga * L) : 1.0;
    ks = c64{0.0, 0.0};
    kf[0] = kf[1] = kf[2] = c64{0.0, 0.0};
    for (int side = 0; side < 2; side++) {
        const double span = side ? 1.0 - t0 : t0;
        double u0 = 0.0;
        while (u0 < span) {
            const double step = u0 == 0.0 ? wmin : fmin(u0, wmax);
            const double u1 = fmin(u0 + step, span);
            const double h = u1 - u0;
            for (int q = 0; q < 16; q++) {
                const double u = u0 + h * f_glx[q];
                const double t = side ? t0 + u : t0 - u;
                double r2 = 0.0, dx[3];
                for (int i = 0; i < 3; i++) {
                    dx[i] = P[i] - (a[i] + t * d[i]);
                    r2 += dx[i] * dx[i];
                }
                double r = sqrt(r2);
                // src/electrode.c:211 (r >= DBL_EPSILON)
                const double rs = r < 2.220446049250313e-16 ? 2.220446049250313e-16 : r;
                double sn, cs;
                sincos(-g.y * rs, &sn, &cs);
                const double ex = exp(-g.x * rs);
                const c64 e = c64{ex * cs, ex * sn};
                const double wq = h * f_glw[q];
                ks.x += wq * e.x / rs;
                ks.y += wq * e.y / rs;
                // src/electrode.c:506-513: (1 + gamma r) e^{-gamma r} / r^2, the
                // r^2 floor at DBL_EPSILON
                const double r2s = r2 < 2.220446049250313e-16 ? 2.220446049250313e-16 : r2;
                const c64 f = cmul(c64{1.0 + g.x * r, g.y * r}, e);
                const double sc = wq / (r2s * r);
                for (int i = 0; i < 3; i++) {
                    kf[i].x += f.x * dx[i] * sc;
                    kf[i].y += f.y * dx[i] * sc;
                }
            }
            u0 = u1;
        }
    }
}

__global__ __launch_bounds__(256) void k_fields(FieldArgs A)
{
    const long p = blockIdx.x;
    const int r0 = blockIdx.y * kFieldRhs;
    const int nr = min(kFieldRhs, A.nrhs - r0);
    __shared__ double red[256];
    double P[3] = {A.pts[3 * p], A.pts[3 * p + 1], A.pts[3 * p + 2]};
    // accumulators: per rhs ks (2), ka (6), kf (6)
    double acc[kFieldRhs][14];
    for (int k = 0; k < kFieldRhs; k++)
        for (int c = 0; c < 14; c++) acc[k][c] = 0.0;
    for (long j = threadIdx.x; j < A.m; j += 256) {
        const double a[3] = {A.e.ax[j], A.e.ay[j], A.e.az[j]};
        const double b[3] = {A.e.bx[j], A.e.by[j], A.e.bz[j]};
        c64 ks, kf[3];
        segment_kernels(P, a, b, A.e.len[j], A.gamma, ks, kf);
        const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        for (int k = 0; k < nr; k++) {
            const c64 it = A.it ? A.it[(long)(r0 + k) * A.m + j] : c64{0.0, 0.0};
            const c64 il = A.il ? A.il[(long)(r0 + k) * A.m + j] : c64{0.0, 0.0};
            const c64 vs = cmul(it, ks);
            acc[k][0] += vs.x;
            acc[k][1] += vs.y;
            const c64 vl = cmul(il, ks);
            for (int i = 0; i < 3; i++) {
                acc[k][2 + 2 * i] += vl.x * d[i];
                acc[k][3 + 2 * i] += vl.y * d[i];
                const c64 vf = cmul(it, kf[i]);
                acc[k][8 + 2 * i] += vf.x;
                acc[k][9 + 2 * i] += vf.y;
            }
        }
    }
    for (int k = 0; k < nr; k++)
        for (int c = 0; c < 14; c++) {
            red[threadIdx.x] = acc[k][c];
            __syncthreads();
            for (int s = 128; s > 0; s >>= 1) {
                if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                const double v = red[0];
                const long o = p * A.nrhs + r0 + k;
                if (c < 2) {
                    if (A.ks) (&A.ks[o].x)[c] = v;
                } else if (c < 8) {
                    if (A.ka) (&A.ka[3 * o].x)[c - 2] = v;
                } else {
                    if (A.kf) (&A.kf[3 * o].x)[c - 8] = v;
                }
            }
            __syncthreads();
        }
}

// examples/visacro57emc01.c:406-432: sum_m it[m] one_4pik e^{-gamma rbar}
// log((r1 + r2 + L) / (r1 + r2 - L)) / L at each point, rbar the distance to
// the segment's middle point
__global__ __launch_bounds__(256) void k_mhem_gpd(const double* pts, long np, SegSoA e, long m,
                                                   const c64* it, int nrhs, c64 gamma,
                                                   c64 one_4pik, c64* out)
{
    const long p = blockIdx.x;
    const int k = blockIdx.y;
    __shared__ double red[2][256];
    const double P[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    double sx = 0.0, sy = 0.0;
    for (long j = threadIdx.x; j < m; j += 256) {
        const double a[3] = {e.ax[j], e.ay[j], e.az[j]};
        const double b[3] = {e.bx[j], e.by[j], e.bz[j]};
        double r1 = 0.0, r2 = 0.0, rb = 0.0;
        for (int i = 0; i < 3; i++) {
            const double mid = 0.5 * (a[i] + b[i]);
            r1 += (P[i] - a[i]) * (P[i] - a[i]);
            r2 += (P[i] - b[i]) * (P[i] - b[i]);
            rb += (P[i] - mid) * (P[i] - mid);
        }
        r1 = sqrt(r1);
        r2 = sqrt(r2);
        rb = sqrt(rb);
        const double L = e.len[j];
        const double lg = log((r1 + r2 + L) / (r1 + r2 - L)) / L;
        double sn, cs;
        sincos(-gamma.y * rb, &sn, &cs);
        const double ex = exp(-gamma.x * rb) * lg;
        const c64 v = cmul(it[(long)k * m + j], c64{ex * cs, ex * sn});
        sx += v.x;
        sy += v.y;
    }
    red[0][threadIdx.x] = sx;
    red[1][threadIdx.x] = sy;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[p * nrhs + k] = cmul(one_4pik, c64{red[0][0], red[1][0]});
}

}  // namespace tags

using namespace tags;

namespace {

struct FieldCall {
    hipStream_t st = nullptr;
    double* seg = nullptr;
    double* pts = nullptr;
    c64* cur = nullptr;  // il then it
    c64* out = nullptr;  // ks, ka, kf
    ~FieldCall()
    {
        if (st) hipStreamSynchronize(st);
        hipFree(seg);
        hipFree(pts);
        hipFree(cur);
        hipFree(out);
        if (st) hipStreamDestroy(st);
    }
    SegSoA soa(long m) const
    {
        return SegSoA{seg, seg + m, seg + 2 * m, seg + 3 * m, seg + 4 * m, seg + 5 * m,
                      seg + 6 * m, seg + 7 * m};
    }
};

int upload_segments(FieldCall& fc, const Electrode* e, long m)
{
    std::vector<double> h(8 * (size_t)m);
    for (long i = 0; i < m; i++) {
        for (int c = 0; c < 3; c++) {
            h[c * m + i] = e[i].start_point[c];
            h[(3 + c) * m + i] = e[i].end_point[c];
        }
        h[6 * m + i] = e[i].length;
        h[7 * m + i] = e[i].radius;
    }
    TAGS_HIP(hipMalloc(&fc.seg, sizeof(double) * 8 * (size_t)(m > 0 ? m : 1)));
    TAGS_HIP(hipMemcpyAsync(fc.seg, h.data(), sizeof(double) * 8 * m, hipMemcpyHostToDevice,
                            fc.st));
    return 0;
}

// ks (np x nrhs), ka and kf (np x nrhs x 3) on the host; il / it may be null
int fields_host(const double* points, long np, const Electrode* e, long m, const c64* il,
                const c64* it, int nrhs, c64 gamma, c64* ks, c64* ka, c64* kf)
{
    if (np <= 0 || nrhs <= 0) return 0;
    if (np > 2147483647L) return TAGS_ERR_TOO_LARGE;
    int rc = field_tables_init();
    if (rc) return rc;
    FieldCall fc;
    TAGS_HIP(hipStreamCreate(&fc.st));
    if ((rc = upload_segments(fc, e, m))) return rc;
    const size_t nc = (size_t)nrhs * m;
    const size_t no = (size_t)np * nrhs;
    TAGS_HIP(hipMalloc(&fc.pts, sizeof(double) * 3 * np));
    TAGS_HIP(hipMalloc(&fc.cur, sizeof(c64) * 2 * (nc ? nc : 1)));
    TAGS_HIP(hipMalloc(&fc.out, sizeof(c64) * 7 * no));
    TAGS_HIP(hipMemcpyAsync(fc.pts, points, sizeof(double) * 3 * np, hipMemcpyHostToDevice,
                            fc.st));
    if (il) TAGS_HIP(hipMemcpyAsync(fc.cur, il, sizeof(c64) * nc, hipMemcpyHostToDevice, fc.st));
    if (it)
        TAGS_HIP(hipMemcpyAsync(fc.cur + nc, it, sizeof(c64) * nc, hipMemcpyHostToDevice, fc.st));
    FieldArgs A{fc.pts, np, fc.soa(m), m, il ? fc.cur : nullptr, it ? fc.cur + nc : nullptr,
                nrhs, gamma, fc.out, fc.out + no, fc.out + 4 * no};
    dim3 grid((unsigned)np, (unsigned)((nrhs + kFieldRhs - 1) / kFieldRhs));
    hipLaunchKernelGGL(k_fields, grid, dim3(256), 0, fc.st, A);
    TAGS_HIP(hipGetLastError());
    if (ks) TAGS_HIP(hipMemcpyAsync(ks, fc.out, sizeof(c64) * no, hipMemcpyDeviceToHost, fc.st));
    if (ka)
        TAGS_HIP(hipMemcpyAsync(ka, fc.out + no, sizeof(c64) * 3 * no, hipMemcpyDeviceToHost,
                                fc.st));
    if (kf)
        TAGS_HIP(hipMemcpyAsync(kf, fc.out + 4 * no, sizeof(c64) * 3 * no, hipMemcpyDeviceToHost,
                                fc.st));
    TAGS_HIP(hipStreamSynchronize(fc.st));
    return 0;
}

inline c64 cdiv_h(c64 a, c64 b)
{
    const double den = b.x * b.x + b.y * b.y;
    return c64{(a.x * b.x + a.y * b.y) / den, (a.y * b.x - a.x * b.y) / den};
}

}  // namespace

extern "C" {

int tags_fields_batch(const double* points, size_t num_points, const Electrode* electrodes,
                      size_t num_electrodes, const TAGS_COMPLEX* il, const TAGS_COMPLEX* it,
                      size_t nrhs, TAGS_COMPLEX gamma, TAGS_COMPLEX s, double mur,
                      TAGS_COMPLEX kappa, TAGS_COMPLEX* pot, TAGS_COMPLEX* va, TAGS_COMPLEX* ve)
{
    const long np = (long)num_points, m = (long)num_electrodes;
    if (np == 0 || nrhs == 0) return 0;
    if (!points || !electrodes || nrhs > 65535 * (size_t)kFieldRhs) return TAGS_ERR_ARG;
    const size_t no = (size_t)np * nrhs;
    std::vector<c64> ks(no), ka(3 * no), kf(3 * no);
    int rc = fields_host(points, np, electrodes, m, (const c64*)il, (const c64*)it, (int)nrhs,
                         from_c(gamma), ks.data(), ka.data(), kf.data());
    if (rc) return rc;
    const c64 k4 = c64{kFourPi * from_c(kappa).x, kFourPi * from_c(kappa).y};
    const double u4 = mur * kMu0 / kFourPi;
    const c64 sv = from_c(s);
    const bool mag = std::sqrt(sv.x * sv.x + sv.y * sv.y) > 2.220446049250313e-16;
    for (size_t o = 0; o < no; o++) {
        if (pot) pot[o] = to_c(cdiv_h(ks[o], k4));
        for (int i = 0; i < 3; i++) {
            const c64 a = c64{ka[3 * o + i].x * u4, ka[3 * o + i].y * u4};
            if (va) va[3 * o + i] = to_c(a);
            if (ve) {
                c64 v = cdiv_h(kf[3 * o + i], k4);
                if (mag) v = csub(v, cmul(sv, a));
                ve[3 * o + i] = to_c(v);
            }
        }
    }
    return 0;
}

TAGS_COMPLEX electric_potential(const double* point, const Electrode* electrodes,
                                size_t num_electrodes, const TAGS_COMPLEX* it, TAGS_COMPLEX gamma,
                                TAGS_COMPLEX kappa, size_t max_eval, double req_abs_error,
                                double req_rel_error)
{
    (void)max_eval;
    (void)req_abs_error;
    (void)req_rel_error;
    TAGS_COMPLEX pot = to_c(c64{0.0, 0.0});
    TAGS_COMPLEX sz = to_c(c64{0.0, 0.0});
    int rc = tags_fields_batch(point, 1, electrodes, num_electrodes, nullptr, it, 1, gamma, sz,
                               1.0, kappa, &pot, nullptr, nullptr);
    return rc ? to_c(c64{(double)rc, 0.0}) : pot;
}

int magnetic_potential(const double* point, const Electrode* electrodes, size_t num_electrodes,
                       const TAGS_COMPLEX* il, TAGS_COMPLEX gamma, double mur, size_t max_eval,
                       double req_abs_error, double req_rel_error, TAGS_COMPLEX* va)
{
    (void)max_eval;
    (void)req_abs_error;
    (void)req_rel_error;
    TAGS_COMPLEX one = to_c(c64{1.0, 0.0}), zero = to_c(c64{0.0, 0.0});
    return tags_fields_batch(point, 1, electrodes, num_electrodes, il, nullptr, 1, gamma, zero,
                             mur, one, nullptr, va, nullptr);
}

int electric_field(const double* point, const Electrode* electrodes, size_t num_electrodes,
                   const TAGS_COMPLEX* il, const TAGS_COMPLEX* it, TAGS_COMPLEX gamma,
                   TAGS_COMPLEX s, double mur, TAGS_COMPLEX kappa, size_t max_eval,
                   double req_abs_error, double req_rel_error, TAGS_COMPLEX* ve)
{
    (void)max_eval;
    (void)req_abs_error;
    (void)req_rel_error;
    TAGS_COMPLEX e3[3];
    int rc = tags_fields_batch(point, 1, electrodes, num_electrodes, il, it, 1, gamma, s, mur,
                               kappa, nullptr, nullptr, e3);
    if (rc) return rc;
    // src/electrode.c:523-557: with |s| > DBL_EPSILON ve is overwritten by
    // -s A before the conservative part is added; otherwise that part is added
    // to what ve holds
    const c64 sv = from_c(s);
    const bool mag = std::sqrt(sv.x * sv.x + sv.y * sv.y) > 2.220446049250313e-16;
    for (int i = 0; i < 3; i++) {
        const c64 v = from_c(e3[i]);
        ve[i] = mag ? e3[i] : to_c(cadd(from_c(ve[i]), v));
    }
    return 0;
}

// src/electrode.c:626-660: the unit-interval integral of E(p(t)) . u along the
// straight path p(t) = p1 + t (p2 - p1), u its unit direction (the reference
// does not multiply by the path length). Composite Gauss-Legendre (16 points
// per panel) over 1, 2, 4, ... equal panels until two successive results agree
// to 1e-13 relative (at most 256 panels); each round is one batched field
// evaluation.
TAGS_COMPLEX voltage(const double* point1, const double* point2, const Electrode* electrodes,
                     size_t num_electrodes, const TAGS_COMPLEX* il, const TAGS_COMPLEX* it,
                     TAGS_COMPLEX gamma, TAGS_COMPLEX s, double mur, TAGS_COMPLEX kappa,
                     size_t max_eval, double req_abs_error, double req_rel_error)
{
    (void)max_eval;
    (void)req_abs_error;
    (void)req_rel_error;
    double dx[3], len = 0.0;
    for (int i = 0; i < 3; i++) {
        dx[i] = point2[i] - point1[i];
        len += dx[i] * dx[i];
    }
    len = std::sqrt(len);
    if (len < 2.220446049250313e-16) return to_c(c64{0.0, 0.0});
    double gx[16], gw[16];
    gauss_legendre(16, gx, gw);
    c64 prev = c64{0.0, 0.0};
    for (int panels = 1; panels <= 256; panels *= 2) {
        const long np = 16L * panels;
        std::vector<double> pts(3 * np), wts(np);
        for (int q = 0; q < panels; q++)
            for (int k = 0; k < 16; k++) {
                const double t = (q + gx[k]) / panels;
                for (int i = 0; i < 3; i++) pts[3 * (16 * q + k) + i] = point1[i] + t * dx[i];
                wts[16 * q + k] = gw[k] / panels;
            }
        std::vector<TAGS_COMPLEX> ve(3 * np);
        int rc = tags_fields_batch(pts.data(), np, electrodes, num_electrodes, il, it, 1, gamma, s,
                                   mur, kappa, nullptr, nullptr, ve.data());
        if (rc) return to_c(c64{(double)rc, 0.0});
        c64 v = c64{0.0, 0.0};
        for (long j = 0; j < np; j++)
            for (int i = 0; i < 3; i++) {
                const c64 e = from_c(ve[3 * j + i]);
                v.x += wts[j] * e.x * dx[i] / len;
                v.y += wts[j] * e.y * dx[i] / len;
            }
        const double dv = std::hypot(v.x - prev.x, v.y - prev.y);
        if (panels > 1 && dv <= 1e-13 * std::hypot(v.x, v.y)) return to_c(v);
        prev = v;
    }
    return to_c(prev);
}

int tags_mhem_gpd(const double* points, size_t num_points, const Electrode* electrodes,
                  size_t num_electrodes, const TAGS_COMPLEX* it, size_t nrhs, TAGS_COMPLEX gamma,
                  TAGS_COMPLEX kappa, TAGS_COMPLEX* out)
{
    const long np = (long)num_points, m = (long)num_electrodes;
    if (np == 0 || nrhs == 0) return 0;
    if (!points || !electrodes || !it || !out || nrhs > 65535) return TAGS_ERR_ARG;
    FieldCall fc;
    TAGS_HIP(hipStreamCreate(&fc.st));
    int rc = upload_segments(fc, electrodes, m);
    if (rc) return rc;
    const size_t nc = nrhs * (size_t)m, no = (size_t)np * nrhs;
    TAGS_HIP(hipMalloc(&fc.pts, sizeof(double) * 3 * np));
    TAGS_HIP(hipMalloc(&fc.cur, sizeof(c64) * (nc ? nc : 1)));
    TAGS_HIP(hipMalloc(&fc.out, sizeof(c64) * no));
    TAGS_HIP(hipMemcpyAsync(fc.pts, points, sizeof(double) * 3 * np, hipMemcpyHostToDevice,
                            fc.st));
    TAGS_HIP(hipMemcpyAsync(fc.cur, it, sizeof(c64) * nc, hipMemcpyHostToDevice, fc.st));
    const c64 kv = from_c(kappa);
    const c64 one_4pik = cdiv_h(c64{1.0, 0.0}, c64{kFourPi * kv.x, kFourPi * kv.y});
    hipLaunchKernelGGL(k_mhem_gpd, dim3((unsigned)np, (unsigned)nrhs), dim3(256), 0, fc.st, fc.pts,
                       np, fc.soa(m), m, fc.cur, (int)nrhs, from_c(gamma), one_4pik, fc.out);
    TAGS_HIP(hipGetLastError());
    TAGS_HIP(hipMemcpyAsync(out, fc.out, sizeof(c64) * no, hipMemcpyDeviceToHost, fc.st));
    TAGS_HIP(hipStreamSynchronize(fc.st));
    return 0;
}

}  // extern "C"
