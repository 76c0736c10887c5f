// Segment-pair impedance integrals and immittance-matrix assembly on gfx950.
//
// The reference fills Z_L / Z_T pair by pair with adaptive Cubature of
// exp(-gamma r)/r over the unit square of the two segment axes
// (src/electrode.c:176-197, 247-302, 304-431). Here every pair is one GPU lane:
//
//  * pairs that share an end point (adjacent segments of a conductor, or
//    segments meeting at a grid node) have a 1/r corner singularity. They are
//    split into two triangles with the Duffy map (u, v) = (rho, rho w); the
//    radial integral of exp(-gamma rho R)/R is done in closed form,
//    (1 - exp(-gamma R)) / (gamma R), and the remaining 1-D integral in w is
//    smooth and taken with Gauss-Legendre (graded panels when the two segments
//    meet at a sharp angle or have very different lengths);
//  * other pairs use a tensor Gauss-Legendre rule whose order follows the
//    ratio of the segment distance to the segment length and |gamma| L; pairs
//    closer than a quarter segment length are subdivided.
//
// The diagonal of the direct part is the reference's gamma-independent closed
// form (src/electrode.c:238-245, used at :346-348).
#include <cmath>
#include <mutex>
#include <vector>

#include "tags_common.h"

namespace tags {

// Gauss-Legendre rules on [0, 1] for the orders below, in constant memory.
constexpr int kNumOrders = 8;
constexpr int kOrders[kNumOrders] = {4, 5, 6, 8, 12, 16, 20, 32};
__constant__ double g_glx[128];
__constant__ double g_glw[128];
__constant__ int g_gloff[kNumOrders];

// set once per process under g_gl_mutex (callers may be concurrent threads)
static bool g_gl_ready = false;
static std::mutex g_gl_mutex;

// Legendre nodes by Newton iteration on P_n from the Chebyshev-like guess.
void gauss_legendre(int n, double* x, double* w)
{
    for (int i = 0; i < n; i++) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
        double dp = 0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; k++) {
                double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
            }
            dp = n * (z * p0 - p1) / (z * z - 1.0);
            double dz = p0 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        // map [-1, 1] -> [0, 1]
        x[i] = 0.5 * (1.0 - z);
        w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
}

int gl_tables_init()
{
    std::lock_guard<std::mutex> lock(g_gl_mutex);
    if (g_gl_ready) return 0;
    std::vector<double> x(128, 0.0), w(128, 0.0);
    int off[kNumOrders];
    int o = 0;
    for (int i = 0; i < kNumOrders; i++) {
        off[i] = o;
        gauss_legendre(kOrders[i], &x[o], &w[o]);
        o += kOrders[i];
    }
    TAGS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_glx), x.data(), 128 * sizeof(double)));
    TAGS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_glw), w.data(), 128 * sizeof(double)));
    TAGS_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_gloff), off, sizeof(off)));
    g_gl_ready = true;
    return 0;
}

// ---------------------------------------------------------------------------
// device geometry helpers
// ---------------------------------------------------------------------------
struct V3 {
    double x, y, z;
};
__device__ inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__device__ inline V3 vsub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 vadd(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 vmul(V3 a, double s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double vnorm(V3 a) { return sqrt(vdot(a, a)); }

struct DSeg {
    V3 a, b;     // start, end
    double len;  // length as stored by the caller
};

// Gauss-Legendre order of table entry i
__device__ inline int dev_order(int i)
{
    switch (i) {
    case 0: return 4;
    case 1: return 5;
    case 2: return 6;
    case 3: return 8;
    case 4: return 12;
    case 5: return 16;
    case 6: return 20;
    default: return 32;
    }
}

// table entry of the smallest order >= q
__device__ inline int order_index(int q)
{
    int i = 0;
    while (i < kNumOrders - 1 && dev_order(i) < q) i++;
    return i;
}

// exp(-g r) / r
__device__ inline c64 green(c64 g, double r)
{
    double mag = exp(-g.x * r) / r;
    double sn, cs;
    sincos(-g.y * r, &sn, &cs);
    return c64{mag * cs, mag * sn};
}

// (1 - exp(-z)) / z with a Taylor series near 0 (no cancellation)
__device__ inline c64 phi1(c64 z)
{
    double az = fabs(z.x) + fabs(z.y);
    if (az < 0.25) {
        // sum_{k>=0} (-z)^k / (k+1)!
        c64 term = c64{1.0, 0.0}, sum = c64{1.0, 0.0};
        c64 mz = c64{-z.x, -z.y};
        for (int k = 1; k < 16; k++) {
            term = cscale(cmul(term, mz), 1.0 / (k + 1));
            sum = cadd(sum, term);
        }
        return sum;
    }
    double e = exp(-z.x);
    double sn, cs;
    sincos(-z.y, &sn, &cs);
    c64 one_minus = c64{1.0 - e * cs, -e * sn};
    return cdiv(one_minus, z);
}

// Minimum distance between two segments (closest points, clamped).
__device__ double seg_seg_distance(const DSeg& S, const DSeg& R)
{
    V3 d1 = vsub(S.b, S.a), d2 = vsub(R.b, R.a), r = vsub(S.a, R.a);
    double a = vdot(d1, d1), e = vdot(d2, d2), f = vdot(d2, r);
    double c = vdot(d1, r), b = vdot(d1, d2);
    double den = a * e - b * b;
    double s = 0.0, t;
    if (den > 1e-14 * a * e) s = fmin(fmax((b * f - c * e) / den, 0.0), 1.0);
    t = (b * s + f) / e;
    if (t < 0.0) {
        t = 0.0;
        s = fmin(fmax(-c / a, 0.0), 1.0);
    } else if (t > 1.0) {
        t = 1.0;
        s = fmin(fmax((b - c) / a, 0.0), 1.0);
    }
    V3 p = vadd(S.a, vmul(d1, s)), q = vadd(R.a, vmul(d2, t));
    return vnorm(vsub(p, q));
}

// Tensor Gauss-Legendre rule of order index oi over the unit square.
__device__ c64 pair_gl(const DSeg& S, const DSeg& R, c64 g, int oi)
{
    const int q = dev_order(oi), off = g_gloff[oi];
    V3 ds = vsub(S.b, S.a), dr = vsub(R.b, R.a);
    double sx = 0.0, sy = 0.0;
    for (int p = 0; p < q; p++) {
        V3 pr = vadd(R.a, vmul(dr, g_glx[off + p]));
        double ix = 0.0, iy = 0.0;
        for (int k = 0; k < q; k++) {
            V3 ps = vadd(S.a, vmul(ds, g_glx[off + k]));
            c64 f = green(g, vnorm(vsub(pr, ps)));
            ix += g_glw[off + k] * f.x;
            iy += g_glw[off + k] * f.y;
        }
        sx += g_glw[off + p] * ix;
        sy += g_glw[off + p] * iy;
    }
    return c64{sx, sy};
}

__device__ inline int geometric_order(double rho, double glen)
{
    int qg = rho >= 16 ? 4 : rho >= 8 ? 5 : rho >= 4 ? 6 : rho >= 2 ? 8 : rho >= 1 ? 12
           : rho >= 0.5 ? 20 : 32;
    int qo = glen <= 0.5 ? 4 : glen <= 1 ? 6 : glen <= 2 ? 8 : glen <= 4 ? 12 : glen <= 8 ? 20
           : 32;
    return order_index(qg > qo ? qg : qo);
}

// Non-touching pair: choose the order from the geometry, subdivide very near
// pairs (distance < L/4) into ns x ns sub-pairs.
__device__ c64 pair_regular(const DSeg& S, const DSeg& R, c64 g, double dist)
{
    const double gabs = sqrt(g.x * g.x + g.y * g.y);
    const double lmax = fmax(S.len, R.len);
    const double rho = dist / lmax;
    if (rho >= 0.25) return pair_gl(S, R, g, geometric_order(rho, gabs * lmax));
    int ns = (int)ceil(1.0 / fmax(rho, 1.0 / 64.0));
    ns = ns > 64 ? 64 : ns;
    V3 ds = vmul(vsub(S.b, S.a), 1.0 / ns), dr = vmul(vsub(R.b, R.a), 1.0 / ns);
    double sx = 0.0, sy = 0.0;
    for (int i = 0; i < ns; i++) {
        DSeg Ss{vadd(S.a, vmul(ds, i)), vadd(S.a, vmul(ds, i + 1)), S.len / ns};
        for (int k = 0; k < ns; k++) {
            DSeg Rs{vadd(R.a, vmul(dr, k)), vadd(R.a, vmul(dr, k + 1)), R.len / ns};
            double d = seg_seg_distance(Ss, Rs);
            double rs = d / fmax(Ss.len, Rs.len);
            c64 v = pair_gl(Ss, Rs, g, geometric_order(rs < 0.25 ? 0.25 : rs,
                                                        gabs * fmax(Ss.len, Rs.len)));
            sx += v.x;
            sy += v.y;
        }
    }
    // each sub-square has area 1/ns^2 of the unit square
    double sc = 1.0 / ((double)ns * ns);
    return c64{sx * sc, sy * sc};
}

// One Duffy triangle: int_0^1 phi(g R(w)) / R(w) dw with R(w) = |p - w q|.
// Graded panels around the nearest approach of the complex root of R(w)^2.
__device__ c64 duffy_line(V3 p, V3 q, c64 g, bool w_on_p)
{
    // R(w) = |p - w q| (w_on_p == false) or |w p - q| (w_on_p == true)
    double pp = vdot(p, p), qq = vdot(q, q), pq = vdot(p, q);
    // roots of R^2 = 0 in w
    double den = w_on_p ? pp : qq;
    double re = pq / den;
    double cr2 = pp * qq - pq * pq;
    double im = sqrt(fmax(cr2, 0.0)) / den;
    double wc = fmin(fmax(re, 0.0), 1.0);
    double dist = sqrt((re - wc) * (re - wc) + im * im);
    const int oi = order_index(20);
    const int nq = dev_order(oi), off = g_gloff[oi];
    double sx = 0.0, sy = 0.0;
    // panel break points
    double bp[40];
    int nb = 0;
    if (dist >= 0.5) {
        bp[nb++] = 0.0;
        bp[nb++] = 1.0;
    } else {
        double h = fmax(dist, 1e-6);
        // left of wc
        double left[20], right[20];
        int nl = 0, nr = 0;
        double x = wc - h;
        while (x > 0.0 && nl < 19) {
            left[nl++] = x;
            h *= 2.0;
            x = wc - h;
        }
        h = fmax(dist, 1e-6);
        x = wc + h;
        while (x < 1.0 && nr < 19) {
            right[nr++] = x;
            h *= 2.0;
            x = wc + h;
        }
        bp[nb++] = 0.0;
        for (int i = nl - 1; i >= 0; i--) bp[nb++] = left[i];
        if (wc > 0.0 && wc < 1.0) bp[nb++] = wc;
        for (int i = 0; i < nr; i++) bp[nb++] = right[i];
        bp[nb++] = 1.0;
    }
    for (int s = 0; s + 1 < nb; s++) {
        double a = bp[s], b = bp[s + 1];
        if (b <= a) continue;
        for (int k = 0; k < nq; k++) {
            double w = a + (b - a) * g_glx[off + k];
            V3 d = w_on_p ? vsub(vmul(p, w), q) : vsub(p, vmul(q, w));
            double R = vnorm(d);
            c64 v = cscale(phi1(cscale(g, R)), 1.0 / R);
            sx += (b - a) * g_glw[off + k] * v.x;
            sy += (b - a) * g_glw[off + k] * v.y;
        }
    }
    return c64{sx, sy};
}

// Unit-square integral of exp(-g r)/r between segments S (sender) and R
// (receiver). Returns the integral over [0,1]^2 (the reference's `integral`
// multiplies this by L_r L_s).
__device__ c64 pair_integral_unit(const DSeg& S, const DSeg& R, c64 g)
{
    const double tol = 1e-9 * fmax(S.len, R.len);
    const double tol2 = tol * tol;
    V3 sp[2] = {S.a, S.b}, rp[2] = {R.a, R.b};
    for (int i = 0; i < 2; i++)
        for (int k = 0; k < 2; k++) {
            V3 d = vsub(sp[i], rp[k]);
            if (vdot(d, d) <= tol2) {
                V3 C = sp[i];
                V3 dsv = vsub(sp[1 - i], C);  // sender, from the shared point
                V3 drv = vsub(rp[1 - k], C);  // receiver, from the shared point
                // triangle v <= u: r = rho |d_r - w d_s|; triangle u <= v:
                // r = rho |w d_r - d_s|
                c64 t1 = duffy_line(drv, dsv, g, false);
                c64 t2 = duffy_line(drv, dsv, g, true);
                return cadd(t1, t2);
            }
        }
    return pair_regular(S, R, g, seg_seg_distance(S, R));
}

__device__ inline double self_integral_dev(double len, double rad)
{
    double k1 = rad / len;
    double k2 = sqrt(1.0 + k1 * k1);
    return 2.0 * len * (log((k2 + 1.0) / k1) - k2 + k1);
}

__device__ inline DSeg load_seg(const SegSoA& s, long i)
{
    return DSeg{V3{s.ax[i], s.ay[i], s.az[i]}, V3{s.bx[i], s.by[i], s.bz[i]}, s.len[i]};
}

// INTG_SINGLE (src/electrode.c:199-216, :279-284): the sender-axis integral of
// e^{-gamma r}/r to the receiver's middle point M over t in [0, 1] (r floored at
// DBL_EPSILON), by 16-point Gauss-Legendre panels graded away from the axis
// parameter closest to M: the first panel d/L wide (d: distance from M to the
// segment), each next one as wide as its distance from that parameter and at
// most 2/(|gamma| L) wide.
__device__ c64 single_unit(const DSeg& S, V3 M, c64 g)
{
    const int off = g_gloff[5];  // order 16
    const V3 d = vsub(S.b, S.a);
    const double dd = vdot(d, d);
    double t0 = dd > 0.0 ? vdot(vsub(M, S.a), d) / dd : 0.0;
    t0 = fmin(fmax(t0, 0.0), 1.0);
    const double L = dd > 0.0 ? sqrt(dd) : S.len;
    const double wmin = fmax(vnorm(vsub(vadd(S.a, vmul(d, t0)), M)) / L, 1e-14);
    const double ga = sqrt(g.x * g.x + g.y * g.y);
    const double wmax = ga * L > 2.0 ? 2.0 / (ga * L) : 1.0;
    double sx = 0.0, sy = 0.0;
    for (int side = 0; side < 2; side++) {
        const double span = side ? 1.0 - t0 : t0;
        double u0 = 0.0;
        while (u0 < span) {
            const double u1 = fmin(u0 + (u0 == 0.0 ? wmin : fmin(u0, wmax)), span);
            const double h = u1 - u0;
            for (int q = 0; q < 16; q++) {
                const double u = u0 + h * g_glx[off + q];
                const double t = side ? t0 + u : t0 - u;
                double r = vnorm(vsub(vadd(S.a, vmul(d, t)), M));
                if (r < 2.220446049250313e-16) r = 2.220446049250313e-16;
                const c64 f = green(g, r);
                sx += h * g_glw[off + q] * f.x;
                sy += h * g_glw[off + q] * f.y;
            }
            u0 = u1;
        }
    }
    return c64{sx, sy};
}

// Reference's `integral` result (L_r L_s times the unit-square integral, or the
// middle-point distance for INTG_NONE). INTG_MHEM (src/electrode.c:219-235,
// :288-293) is L_r times the receiver-axis integral of
// log((r1 + r2 + L_s) / (r1 + r2 - L_s)), which is the sender-axis integral of
// 1/r in closed form: the same quantity as the double integral of 1/r, i.e.
// the gamma = 0 value of the pair rule, with a zero imaginary part.
__device__ inline c64 pair_result(const DSeg& S, const DSeg& R, c64 g, int type)
{
    if (type == INTG_NONE) {
        V3 ms = vmul(vadd(S.a, S.b), 0.5), mr = vmul(vadd(R.a, R.b), 0.5);
        return c64{vnorm(vsub(ms, mr)), 0.0};
    }
    const double ll = R.len * S.len;
    if (type == INTG_SINGLE) {
        const c64 J = single_unit(S, vmul(vadd(R.a, R.b), 0.5), g);
        return c64{J.x * S.len, J.y * S.len};
    }
    if (type == INTG_MHEM) {
        const c64 J = pair_integral_unit(S, R, c64{0.0, 0.0});
        return c64{J.x * ll, 0.0};
    }
    c64 J = pair_integral_unit(S, R, g);
    return c64{J.x * ll, J.y * ll};
}

// |d_i . d_k| with un-normalised direction vectors (src/electrode.c:349-359)
__device__ inline double abs_dot(const DSeg& a, const DSeg& b)
{
    return fabs(vdot(vsub(a.b, a.a), vsub(b.b, b.a)));
}

// ---------------------------------------------------------------------------
// Drop-in calculate_impedances / impedances_images on a device copy of the
// caller's m*m arrays (lower-half convention, entry (i, k >= i) at i*m + k).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pairs_dropin(c64* zl, c64* zt, int alias, SegSoA e,
                                                       SegSoA img, long m, PairScalars sc,
                                                       int images)
{
    long k = blockIdx.x * 16 + threadIdx.x % 16;
    long i = blockIdx.y * 16 + threadIdx.x / 16;
    if (i >= m || k >= m || k < i) return;
    DSeg Si = load_seg(e, i);
    long idx = i * m + k;
    const double ls = Si.len;
    if (!images) {
        if (k == i) {
            // src/electrode.c:346-348
            double intg = self_integral_dev(ls, e.rad[i]);
            c64 vl = cscale(sc.zl_dir, intg);
            c64 vt = cscale(cscale(sc.zt_dir, 1.0 / (ls * ls)), intg);
            zl[idx] = vl;
            zt[idx] = vt;  // zt written last: wins when zl == zt
            return;
        }
        DSeg Rk = load_seg(e, k);
        const double lr = Rk.len;
        double cost = abs_dot(Si, Rk);
        c64 intg = pair_result(Si, Rk, sc.gamma, sc.intg_type);
        c64 vt = cmul(cscale(sc.zt_dir, 1.0 / (ls * lr)), intg);
        c64 vl = c64{0.0, 0.0};
        if (!(cost < 2.220446049250313e-16)) vl = cscale(cmul(sc.zl_dir, intg), cost / (ls * lr));
        if (alias) {
            zl[idx] = vl;  // zt then zl: zl wins off the diagonal
        } else {
            zt[idx] = vt;
            zl[idx] = vl;
        }
        return;
    }
    // images (src/electrode.c:403-429)
    DSeg Rk = load_seg(img, k);
    const double lr = Rk.len;
    double cost = abs_dot(Si, Rk);
    c64 intg = pair_result(Si, Rk, sc.gamma, sc.intg_type);
    c64 vt = cmul(cscale(sc.zt_img, 1.0 / (ls * lr)), intg);
    zt[idx] = cadd(zt[idx], vt);
    if (cost > 2.220446049250313e-16) {
        c64 vl = cscale(cmul(sc.zl_img, intg), cost / (ls * lr));
        zl[idx] = cadd(zl[idx], vl);  // same buffer twice when zl == zt
    }
}

int launch_pairs_dropin(double* zl, double* zt, bool alias, SegSoA e, SegSoA img, long m,
                        const PairScalars& sc, bool images, hipStream_t st)
{
    if (m <= 0) return 0;
    int rc = gl_tables_init();
    if (rc) return rc;
    dim3 grid((unsigned)((m + 15) / 16), (unsigned)((m + 15) / 16));
    hipLaunchKernelGGL(k_pairs_dropin, grid, dim3(256), 0, st, (c64*)zl, (c64*)zt,
                       alias ? 1 : 0, e, img, m, sc, images ? 1 : 0);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Sweep assembly: -Z_L and -Z_T (direct + images) straight into the row-major
// immittance matrix W (ld = ldw). Tiles of 32 x 32 segment pairs on and below
// the block diagonal; segment geometry of both tile edges staged in LDS; each
// tile is written once as (i, k) and once transposed as (k, i), both passes
// coalesced through an LDS staging tile.
// ---------------------------------------------------------------------------
constexpr int kFT = 32;  // tile edge of the mHEM fill

// Tiles are FT x FT pairs: 32 (four pairs per thread), or 16 (one pair per
// thread) when a launch would otherwise have too few work-groups to fill the
// device (few segments per sample: the 15 m conductor's 30 segments make one
// 32 x 32 tile, three 16 x 16 ones).
template <int FT>
__global__ __launch_bounds__(256) void k_fill_zblocks(c64* W, long ldw, long zl_off, long zt_off,
                                                       long m, SegSoA e, SegSoA img, PairScalars sc,
                                                       long sw, const PairScalars* scs)
{
    // item blockIdx.z of a batch: its own W and frequency-sample scalars
    if (scs) {
        sc = scs[blockIdx.z];
        W += (long)blockIdx.z * sw;
    }
    // triangular tile index -> (ti >= tk)
    const long b = blockIdx.x;
    long ti = (long)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= b) ti++;
    while (ti * (ti + 1) / 2 > b) ti--;
    const long tk = b - ti * (ti + 1) / 2;
    const long i0 = ti * FT, k0 = tk * FT;

    __shared__ double geo[2][8][FT];   // [row/col edge][ax ay az bx by bz len rad][seg]
    __shared__ c64 tl[FT][FT + 1];
    __shared__ c64 tt[FT][FT + 1];
    const int t = threadIdx.x;
    if (t < 2 * FT) {
        const int side = t / FT, s = t % FT;
        const long idx = (side == 0 ? i0 : k0) + s;
        if (idx < m) {
            geo[side][0][s] = e.ax[idx];
            geo[side][1][s] = e.ay[idx];
            geo[side][2][s] = e.az[idx];
            geo[side][3][s] = e.bx[idx];
            geo[side][4][s] = e.by[idx];
            geo[side][5][s] = e.bz[idx];
            geo[side][6][s] = e.len[idx];
            geo[side][7][s] = e.rad[idx];
        }
    }
    __syncthreads();
    for (int p = t; p < FT * FT; p += 256) {
        const int a = p / FT, c = p % FT;  // a: row in tile (i), c: column (k)
        const long i = i0 + a, k = k0 + c;
        c64 vl = c64{0, 0}, vt = c64{0, 0};
        if (i < m && k < m && k <= i) {
            DSeg Si{V3{geo[0][0][a], geo[0][1][a], geo[0][2][a]},
                    V3{geo[0][3][a], geo[0][4][a], geo[0][5][a]}, geo[0][6][a]};
            const double ls = Si.len;
            // direct part; the pair (k, i) with k < i is the reference's
            // (i' = k, k' = i): sender e[k], receiver e[i]
            if (k == i) {
                double intg = self_integral_dev(ls, geo[0][7][a]);
                vl = cscale(sc.zl_dir, intg);
                vt = cscale(cscale(sc.zt_dir, 1.0 / (ls * ls)), intg);
            } else {
                DSeg Sk{V3{geo[1][0][c], geo[1][1][c], geo[1][2][c]},
                        V3{geo[1][3][c], geo[1][4][c], geo[1][5][c]}, geo[1][6][c]};
                const double lk = Sk.len;
                c64 intg = pair_result(Sk, Si, sc.gamma, sc.intg_type);
                vt = cmul(cscale(sc.zt_dir, 1.0 / (lk * ls)), intg);
                double cost = abs_dot(Sk, Si);
                if (!(cost < 2.220446049250313e-16))
                    vl = cscale(cmul(sc.zl_dir, intg), cost / (lk * ls));
            }
            // image part: sender e[min], receiver image[max]
            {
                const long lo = k, hi = i;
                DSeg Sl{V3{geo[1][0][c], geo[1][1][c], geo[1][2][c]},
                        V3{geo[1][3][c], geo[1][4][c], geo[1][5][c]}, geo[1][6][c]};
                DSeg Ih = load_seg(img, hi);
                (void)lo;
                const double l1 = Sl.len, l2 = Ih.len;
                c64 intg = pair_result(Sl, Ih, sc.gamma, sc.intg_type);
                vt = cadd(vt, cmul(cscale(sc.zt_img, 1.0 / (l1 * l2)), intg));
                double cost = abs_dot(Sl, Ih);
                if (cost > 2.220446049250313e-16)
                    vl = cadd(vl, cscale(cmul(sc.zl_img, intg), cost / (l1 * l2)));
            }
        }
        tl[a][c] = c64{-vl.x, -vl.y};
        tt[a][c] = c64{-vt.x, -vt.y};
    }
    __syncthreads();
    // pass 1: entries (i, k), k <= i of the Z_L and Z_T blocks
    for (int p = t; p < FT * FT; p += 256) {
        const int a = p / FT, c = p % FT;
        const long i = i0 + a, k = k0 + c;
        if (i < m && k < m && k <= i) {
            W[(zl_off + i) * ldw + (zl_off + k)] = tl[a][c];
            W[(zt_off + i) * ldw + (zt_off + k)] = tt[a][c];
        }
    }
    // pass 2: transposed entries (k, i), k < i
    for (int p = t; p < FT * FT; p += 256) {
        const int c = p / FT, a = p % FT;  // consecutive threads walk i
        const long i = i0 + a, k = k0 + c;
        if (i < m && k < m && k < i) {
            W[(zl_off + k) * ldw + (zl_off + i)] = tl[a][c];
            W[(zt_off + k) * ldw + (zt_off + i)] = tt[a][c];
        }
    }
}

int launch_fill_zblocks(c64* W, long ldw, const BlockLayout& L, SegSoA e, SegSoA img,
                        const PairScalars& sc, hipStream_t st, int nb, long sw,
                        const PairScalars* scs)
{
    const long m = L.m;
    if (m <= 0 || nb <= 0) return 0;
    if (nb > 1 && !scs) return TAGS_ERR_ARG;
    int rc = gl_tables_init();
    if (rc) return rc;
    const long T32 = (m + 31) / 32;
    if (T32 * (T32 + 1) / 2 * nb >= 512) {
        hipLaunchKernelGGL(k_fill_zblocks<32>, dim3((unsigned)(T32 * (T32 + 1) / 2), 1,
                                                    (unsigned)nb),
                           dim3(256), 0, st, W, ldw, L.zl_off, L.zt_off, m, e, img, sc, sw, scs);
    } else {
        const long T16 = (m + 15) / 16;
        hipLaunchKernelGGL(k_fill_zblocks<16>, dim3((unsigned)(T16 * (T16 + 1) / 2), 1,
                                                    (unsigned)nb),
                           dim3(256), 0, st, W, ldw, L.zl_off, L.zt_off, m, e, img, sc, sw, scs);
    }
    TAGS_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// mHEM (examples/harmonic_impedance.c:78-85, :125-140): the gamma-free
// potentials, once per geometry, are the reference's calculate_impedances and
// impedances_images with INTG_MHEM (unit scalars, src/electrode.c:312-316,
// :391-395); every frequency sample scales them by exp(-gamma rbar) with rbar
// the middle-point distance of the pair (of segment and image for the image
// part) and by s mu / 4 pi, 1 / (4 pi kappa) and the reflection coefficients.
// Potentials: real row-major m x m arrays, entries (i, k <= i) only, for the
// reference pair (sender e[k], receiver e[i] or image[i]).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mhem_potentials(double* pl, double* pt, double* pli,
                                                          double* pti, long m, SegSoA e,
                                                          SegSoA img)
{
    const long k = blockIdx.x * 16 + threadIdx.x % 16;
    const long i = blockIdx.y * 16 + threadIdx.x / 16;
    if (i >= m || k > i) return;
    const long idx = i * m + k;
    const DSeg Si = load_seg(e, i), Sk = load_seg(e, k), Ii = load_seg(img, i);
    const double li = Si.len, lk = Sk.len;
    if (k == i) {
        // src/electrode.c:346-348
        const double intg = self_integral_dev(li, e.rad[i]);
        pl[idx] = intg;
        pt[idx] = intg / (li * li);
    } else {
        // src/electrode.c:352-372
        const double intg = pair_result(Sk, Si, c64{0.0, 0.0}, INTG_MHEM).x;
        pt[idx] = intg / (lk * li);
        const double cost = abs_dot(Sk, Si);
        pl[idx] = cost < 2.220446049250313e-16 ? 0.0 : intg * cost / (lk * li);
    }
    // src/electrode.c:403-428 (k <= i, the image of the receiver)
    const double intg = pair_result(Sk, Ii, c64{0.0, 0.0}, INTG_MHEM).x;
    pti[idx] = intg / (lk * Ii.len);
    const double cost = abs_dot(Sk, Ii);
    pli[idx] = cost > 2.220446049250313e-16 ? intg * cost / (lk * Ii.len) : 0.0;
}

int launch_mhem_potentials(double* pl, double* pt, double* pli, double* pti, long m, SegSoA e,
                           SegSoA img, hipStream_t st)
{
    if (m <= 0) return 0;
    if ((m + 15) / 16 > 65535) return TAGS_ERR_TOO_LARGE;
    int rc = gl_tables_init();
    if (rc) return rc;
    dim3 grid((unsigned)((m + 15) / 16), (unsigned)((m + 15) / 16));
    hipLaunchKernelGGL(k_mhem_potentials, grid, dim3(256), 0, st, pl, pt, pli, pti, m, e, img);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// One frequency sample of the mHEM impedance blocks, -Z_L and -Z_T straight
// into W (the tiling of k_fill_zblocks):
//   Z_L(i, k) = s mu/4pi (exp(-gamma rbar) pl + ref_l exp(-gamma rbar_img) pli),
//   Z_T(i, k) = 1/(4 pi kappa) (exp(-gamma rbar) pt + ref_t exp(-gamma rbar_img) pti),
// rbar = |mid(e[i]) - mid(e[k])|, rbar_img = |mid(e[i]) - mid(image[k])|.
__global__ __launch_bounds__(256) void k_fill_mhem(c64* W, long ldw, long zl_off, long zt_off,
                                                    long m, SegSoA e, SegSoA img,
                                                    const double* __restrict__ pl,
                                                    const double* __restrict__ pt,
                                                    const double* __restrict__ pli,
                                                    const double* __restrict__ pti,
                                                    PairScalars sc, long sw,
                                                    const PairScalars* scs)
{
    if (scs) {
        sc = scs[blockIdx.z];
        W += (long)blockIdx.z * sw;
    }
    const long b = blockIdx.x;
    long ti = (long)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= b) ti++;
    while (ti * (ti + 1) / 2 > b) ti--;
    const long tk = b - ti * (ti + 1) / 2;
    const long i0 = ti * kFT, k0 = tk * kFT;
    __shared__ double mid[3][3][kFT];  // [e rows, e cols, image cols][xyz][seg]
    __shared__ c64 tl[kFT][kFT + 1];
    __shared__ c64 tt[kFT][kFT + 1];
    const int t = threadIdx.x;
    if (t < 3 * kFT) {
        const int side = t / kFT, s = t % kFT;
        const long idx = (side == 0 ? i0 : k0) + s;
        const SegSoA& g = side == 2 ? img : e;
        if (idx < m) {
            mid[side][0][s] = 0.5 * (g.ax[idx] + g.bx[idx]);
            mid[side][1][s] = 0.5 * (g.ay[idx] + g.by[idx]);
            mid[side][2][s] = 0.5 * (g.az[idx] + g.bz[idx]);
        }
    }
    __syncthreads();
    for (int p = t; p < kFT * kFT; p += 256) {
        const int a = p / kFT, c = p % kFT;
        const long i = i0 + a, k = k0 + c;
        c64 vl = c64{0, 0}, vt = c64{0, 0};
        if (i < m && k < m && k <= i) {
            const long idx = i * m + k;
            const V3 mi = V3{mid[0][0][a], mid[0][1][a], mid[0][2][a]};
            const double rb = vnorm(vsub(mi, V3{mid[1][0][c], mid[1][1][c], mid[1][2][c]}));
            const double ri = vnorm(vsub(mi, V3{mid[2][0][c], mid[2][1][c], mid[2][2][c]}));
            // exp(-gamma rbar), exp(-gamma rbar_img)
            double sn, cs;
            sincos(-sc.gamma.y * rb, &sn, &cs);
            const double md = exp(-sc.gamma.x * rb);
            const c64 xd = c64{md * cs, md * sn};
            sincos(-sc.gamma.y * ri, &sn, &cs);
            const double mg = exp(-sc.gamma.x * ri);
            const c64 xi = c64{mg * cs, mg * sn};
            vl = cadd(cmul(sc.zl_dir, cscale(xd, pl[idx])), cmul(sc.zl_img, cscale(xi, pli[idx])));
            vt = cadd(cmul(sc.zt_dir, cscale(xd, pt[idx])), cmul(sc.zt_img, cscale(xi, pti[idx])));
        }
        tl[a][c] = c64{-vl.x, -vl.y};
        tt[a][c] = c64{-vt.x, -vt.y};
    }
    __syncthreads();
    for (int p = t; p < kFT * kFT; p += 256) {
        const int a = p / kFT, c = p % kFT;
        const long i = i0 + a, k = k0 + c;
        if (i < m && k < m && k <= i) {
            W[(zl_off + i) * ldw + (zl_off + k)] = tl[a][c];
            W[(zt_off + i) * ldw + (zt_off + k)] = tt[a][c];
        }
    }
    for (int p = t; p < kFT * kFT; p += 256) {
        const int c = p / kFT, a = p % kFT;
        const long i = i0 + a, k = k0 + c;
        if (i < m && k < m && k < i) {
            W[(zl_off + k) * ldw + (zl_off + i)] = tl[a][c];
            W[(zt_off + k) * ldw + (zt_off + i)] = tt[a][c];
        }
    }
}

int launch_fill_mhem(c64* W, long ldw, const BlockLayout& L, SegSoA e, SegSoA img,
                     const double* pl, const double* pt, const double* pli, const double* pti,
                     const PairScalars& sc, hipStream_t st, int nb, long sw,
                     const PairScalars* scs)
{
    const long m = L.m;
    if (m <= 0 || nb <= 0) return 0;
    if (nb > 1 && !scs) return TAGS_ERR_ARG;
    const long T = (m + kFT - 1) / kFT;
    const long nt = T * (T + 1) / 2;
    hipLaunchKernelGGL(k_fill_mhem, dim3((unsigned)nt, 1, (unsigned)nb), dim3(256), 0, st, W, ldw,
                       L.zl_off, L.zt_off, m, e, img, pl, pt, pli, pti, sc, sw, scs);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// Every entry of W outside the two impedance blocks: zero Y_E block, zero
// cross blocks, incidence A (+1 start node, -1 end node), B (0.5) and their
// transposes (src/linalg.c:41-66, 96-108), for any placement of the three
// unknown blocks (offsets of I_L, I_T and U).
__device__ inline int block_of(long i, const BlockLayout& L)
{
    if (i >= L.node_off && i < L.node_off + L.n) return 0;
    if (i >= L.zl_off && i < L.zl_off + L.m) return 1;
    return 2;
}

__global__ __launch_bounds__(256) void k_fill_static(c64* W, long ldw, BlockLayout L,
                                                      const int* __restrict__ sn,
                                                      const int* __restrict__ en, long sw)
{
    W += (long)blockIdx.z * sw;
    const long N = L.n + 2 * L.m;
    for (long r = blockIdx.y; r < N; r += gridDim.y) {
    const int rb = block_of(r, L);
    const long rl = rb == 0 ? r - L.node_off : (rb == 1 ? r - L.zl_off : r - L.zt_off);
    for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < N; c += (long)gridDim.x * 256) {
        const int cb = block_of(c, L);
        if (rb == cb && rb != 0) continue;  // impedance blocks
        const long cl = cb == 0 ? c - L.node_off : (cb == 1 ? c - L.zl_off : c - L.zt_off);
        double v = 0.0;
        if (rb == 0 && cb != 0) {
            const bool st = sn[cl] == rl, ed = en[cl] == rl;
            v = cb == 1 ? (st ? 1.0 : (ed ? -1.0 : 0.0)) : ((st || ed) ? 0.5 : 0.0);
        } else if (cb == 0 && rb != 0) {
            const bool st = sn[rl] == cl, ed = en[rl] == cl;
            v = rb == 1 ? (st ? 1.0 : (ed ? -1.0 : 0.0)) : ((st || ed) ? 0.5 : 0.0);
        }
        W[r * ldw + c] = c64{v, 0.0};
    }
    }
}

int launch_fill_static(c64* W, long ldw, const BlockLayout& L, const int* seg_start_node,
                       const int* seg_end_node, hipStream_t st, int nb, long sw)
{
    const long N = L.n + 2 * L.m;
    if (N <= 0 || nb <= 0) return 0;
    unsigned gx = (unsigned)((N + 255) / 256);
    if (gx > 16) gx = 16;
    const unsigned gy = (unsigned)(N < 65535 ? N : 65535);
    hipLaunchKernelGGL(k_fill_static, dim3(gx, gy, (unsigned)nb), dim3(256), 0, st, W, ldw, L,
                       seg_start_node, seg_end_node, sw);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// Single pair for the drop-in `integral`: out = {re, im, err_re, err_im}; the
// error is the difference to the same pair evaluated with a 3x finer split
// (INTG_SINGLE: the sender split in 3).
__global__ void k_single_pair(SegSoA e, SegSoA r, c64 g, int type, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DSeg S = load_seg(e, 0), R = load_seg(r, 0);
    if (type == INTG_SINGLE) {
        const V3 M = vmul(vadd(R.a, R.b), 0.5);
        const c64 J = single_unit(S, M, g);
        const V3 ds = vmul(vsub(S.b, S.a), 1.0 / 3);
        double sx = 0, sy = 0;
        for (int i = 0; i < 3; i++) {
            DSeg a{vadd(S.a, vmul(ds, i)), vadd(S.a, vmul(ds, i + 1)), S.len / 3};
            const c64 v = single_unit(a, M, g);
            sx += v.x / 3.0;
            sy += v.y / 3.0;
        }
        out[0] = J.x * S.len;
        out[1] = J.y * S.len;
        out[2] = fabs(J.x - sx) * S.len;
        out[3] = fabs(J.y - sy) * S.len;
        return;
    }
    c64 J = pair_integral_unit(S, R, g);
    // reference: split both segments in 3 and sum the 9 sub-pairs
    V3 ds = vmul(vsub(S.b, S.a), 1.0 / 3), dr = vmul(vsub(R.b, R.a), 1.0 / 3);
    double sx = 0, sy = 0;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) {
            DSeg a{vadd(S.a, vmul(ds, i)), vadd(S.a, vmul(ds, i + 1)), S.len / 3};
            DSeg b{vadd(R.a, vmul(dr, k)), vadd(R.a, vmul(dr, k + 1)), R.len / 3};
            c64 v = pair_integral_unit(a, b, g);
            sx += v.x / 9.0;
            sy += v.y / 9.0;
        }
    const double ll = S.len * R.len;
    out[0] = J.x * ll;
    out[1] = J.y * ll;
    out[2] = fabs(J.x - sx) * ll;
    out[3] = fabs(J.y - sy) * ll;
}

int launch_single_pair(SegSoA e, SegSoA r, c64 gamma, int type, double* out4, hipStream_t st)
{
    int rc = gl_tables_init();
    if (rc) return rc;
    hipLaunchKernelGGL(k_single_pair, dim3(1), dim3(64), 0, st, e, r, gamma, type, out4);
    TAGS_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Node/segment incidence (src/linalg.c:29-76): a node matches a segment start
// (or else its end) when every coordinate differs by at most 1e-5.
// seg_pts: m x 6 (start xyz, end xyz). Outputs per segment the matching node of
// its start and end (-1 none; the count of matches is accumulated in
// node_hits[n + 2 s] / [n + 2 s + 1]), per node the number of segments it
// touches in node_hits[node].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_incidence(const double* __restrict__ sp, long m,
                                                    const double* __restrict__ nodes, long n,
                                                    int* sn, int* en, int* hits)
{
    for (long s = blockIdx.y; s < m; s += gridDim.y) {
    const double sx = sp[6 * s], sy = sp[6 * s + 1], sz = sp[6 * s + 2];
    const double ex = sp[6 * s + 3], ey = sp[6 * s + 4], ez = sp[6 * s + 5];
    const double tol = 1e-5;
    for (long nd = (long)blockIdx.x * 256 + threadIdx.x; nd < n; nd += (long)gridDim.x * 256) {
        const double x = nodes[3 * nd], y = nodes[3 * nd + 1], z = nodes[3 * nd + 2];
        const bool st = fabs(sx - x) <= tol && fabs(sy - y) <= tol && fabs(sz - z) <= tol;
        const bool ed = !st && fabs(ex - x) <= tol && fabs(ey - y) <= tol && fabs(ez - z) <= tol;
        if (st) {
            sn[s] = (int)nd;
            atomicAdd(&hits[n + 2 * s], 1);
        }
        if (ed) {
            en[s] = (int)nd;
            atomicAdd(&hits[n + 2 * s + 1], 1);
        }
        if (st || ed) atomicAdd(&hits[nd], 1);
    }
    }
}

int launch_incidence(const double* seg_pts, long m, const double* nodes, long n,
                     int* seg_start_node, int* seg_end_node, int* node_hits, hipStream_t st)
{
    if (m <= 0 || n <= 0) return 0;
    unsigned gx = (unsigned)((n + 255) / 256);
    if (gx > 64) gx = 64;
    const unsigned gy = (unsigned)(m < 65535 ? m : 65535);
    hipLaunchKernelGGL(k_incidence, dim3(gx, gy), dim3(256), 0, st, seg_pts, m, nodes,
                       n, seg_start_node, seg_end_node, node_hits);
    TAGS_HIP(hipGetLastError());
    return 0;
}

}  // namespace tags
