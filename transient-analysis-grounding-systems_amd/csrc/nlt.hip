// Numerical Laplace transform and its inverse (src/auxiliary.c:240-329) and the
// frequency-dependent soil models (src/auxiliary.c:61-93).
//
// The reference calls FFTW's real-to-complex / complex-to-real plans on the
// host. Here the transforms run on the GPU as complex DFTs of length nt over a
// batch of signals (the examples transform one signal per node or right-hand
// side, src/auxiliary.c's layout: signal i at f[i nt ..], g[i ns ..]):
//  * nt a power of two: radix-2 Stockham passes (autosort, no bit reversal),
//    one launch per pass, ping-pong buffers, twiddles from a table computed on
//    the host (exp(-2 pi i t / nt), t < nt);
//  * any other nt: the direct sum with the exact table index (j k mod nt).
// The weights that the reference computes per sample (exp(-c k dt) dt, the
// data windows) are computed on the host with the same formulas, so that only
// the transform itself differs in rounding.
#include <cmath>
#include <complex>
#include <vector>

#include "tags_common.h"

namespace tags {

// y = one radix-2 Stockham pass of x for sub-transform size ns2 = 2 Ns:
// v0 = x[j], v1 = x[j + n/2] w^((j mod Ns) n / (2 Ns)); y[e] = v0 + v1,
// y[e + Ns] = v0 - v1 with e = (j / Ns) 2 Ns + j mod Ns. Batch b at b n.
__global__ __launch_bounds__(256) void k_stockham2(const c64* __restrict__ x, c64* __restrict__ y,
                                                    const c64* __restrict__ w, long n, long Ns,
                                                    long batch, int inverse)
{
    const long half = n / 2;
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < half * batch;
         idx += (long)gridDim.x * 256) {
        const long b = idx / half, j = idx % half;
        const c64* xb = x + b * n;
        c64* yb = y + b * n;
        const long k = j % Ns;
        c64 t = w[k * (n / (2 * Ns))];
        if (inverse) t.y = -t.y;
        const c64 v0 = xb[j];
        const c64 v1 = cmul(xb[j + half], t);
        const long e = (j / Ns) * 2 * Ns + k;
        yb[e] = cadd(v0, v1);
        yb[e + Ns] = csub(v0, v1);
    }
}

// y[k] = sum_j x[j] w^(j k mod n) (conjugate twiddles when inverse)
__global__ __launch_bounds__(256) void k_dft(const c64* __restrict__ x, c64* __restrict__ y,
                                              const c64* __restrict__ w, long n, long batch,
                                              int inverse)
{
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < n * batch;
         idx += (long)gridDim.x * 256) {
        const long b = idx / n, k = idx % n;
        const c64* xb = x + b * n;
        double re = 0.0, im = 0.0;
        long t = 0;  // j k mod n
        for (long j = 0; j < n; j++) {
            c64 tw = w[t];
            if (inverse) tw.y = -tw.y;
            const c64 v = xb[j];
            re += v.x * tw.x - v.y * tw.y;
            im += v.x * tw.y + v.y * tw.x;
            t += k;
            if (t >= n) t -= n;
        }
        y[idx] = c64{re, im};
    }
}

// a[b n + k] = (f[b n + k] w[k], 0): the reference's weighted samples
__global__ __launch_bounds__(256) void k_weight_real(const double* __restrict__ f,
                                                      const double* __restrict__ w, c64* a,
                                                      long n, long batch)
{
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < n * batch;
         idx += (long)gridDim.x * 256)
        a[idx] = c64{f[idx] * w[idx % n], 0.0};
}

// x[b n + k]: the windowed half spectrum g[b ns + k] win[k] for k < ns (the
// imaginary parts of the DC and, for even n, Nyquist terms dropped, as a real
// c2r transform does), its conjugate mirror above
__global__ __launch_bounds__(256) void k_hermitian(const c64* __restrict__ g,
                                                    const double* __restrict__ win, c64* x,
                                                    long n, long ns, long batch)
{
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < n * batch;
         idx += (long)gridDim.x * 256) {
        const long b = idx / n, k = idx % n;
        const long kk = k < ns ? k : n - k;
        const c64 v = g[b * ns + kk];
        c64 r = c64{win[kk] * v.x, win[kk] * v.y};
        if (kk == 0 || 2 * kk == n) r.y = 0.0;
        if (k >= ns) r.y = -r.y;
        x[idx] = r;
    }
}

// out[b n + k] = Re x[b n + k] sc[k]
__global__ __launch_bounds__(256) void k_real_scale(const c64* __restrict__ x,
                                                     const double* __restrict__ sc, double* out,
                                                     long n, long batch)
{
    for (long idx = (long)blockIdx.x * 256 + threadIdx.x; idx < n * batch;
         idx += (long)gridDim.x * 256)
        out[idx] = x[idx].x * sc[idx % n];
}

static unsigned grid_for(long work)
{
    return (unsigned)std::min<long>((work + 255) / 256, 16384);
}

// In-place complex DFT of `batch` signals of length n in buf (tmp: same size);
// the result ends in buf.
static int dft_batch(c64* buf, c64* tmp, const c64* w, long n, long batch, bool inverse,
                     hipStream_t st)
{
    if (n <= 1) return 0;
    if ((n & (n - 1)) == 0) {
        c64 *a = buf, *b = tmp;
        for (long Ns = 1; Ns < n; Ns *= 2) {
            hipLaunchKernelGGL(k_stockham2, dim3(grid_for(n / 2 * batch)), dim3(256), 0, st, a,
                               b, w, n, Ns, batch, inverse ? 1 : 0);
            TAGS_HIP(hipGetLastError());
            std::swap(a, b);
        }
        if (a != buf)
            TAGS_HIP(hipMemcpyAsync(buf, a, sizeof(c64) * n * batch, hipMemcpyDeviceToDevice, st));
        return 0;
    }
    hipLaunchKernelGGL(k_dft, dim3(grid_for(n * batch)), dim3(256), 0, st, buf, tmp, w, n, batch,
                       inverse ? 1 : 0);
    TAGS_HIP(hipGetLastError());
    TAGS_HIP(hipMemcpyAsync(buf, tmp, sizeof(c64) * n * batch, hipMemcpyDeviceToDevice, st));
    return 0;
}

// exp(-2 pi i t / n) for t < n, evaluated in extended precision and rounded
// once, so that every entry is within half an ulp or so
static std::vector<c64> twiddles(long n)
{
    std::vector<c64> w(n);
    for (long t = 0; t < n; t++) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)t /
                              (long double)n;
        w[t] = c64{(double)std::cos(a), (double)-std::sin(a)};
    }
    return w;
}

struct NltBuffers {
    c64* w = nullptr;       // twiddles
    c64* a = nullptr;       // transform buffers
    c64* b = nullptr;
    double* v = nullptr;    // per-sample weights (n)
    hipStream_t st = nullptr;
    ~NltBuffers()
    {
        if (st) hipStreamSynchronize(st);
        hipFree(w);
        hipFree(a);
        hipFree(b);
        hipFree(v);
        if (st) hipStreamDestroy(st);
    }
    int init(long n, long batch)
    {
        TAGS_HIP(hipStreamCreate(&st));
        TAGS_HIP(hipMalloc(&w, sizeof(c64) * n));
        TAGS_HIP(hipMalloc(&a, sizeof(c64) * n * batch));
        TAGS_HIP(hipMalloc(&b, sizeof(c64) * n * batch));
        TAGS_HIP(hipMalloc(&v, sizeof(double) * n));
        std::vector<c64> tw = twiddles(n);
        TAGS_HIP(hipMemcpyAsync(w, tw.data(), sizeof(c64) * n, hipMemcpyHostToDevice, st));
        return 0;
    }
};

}  // namespace tags

using namespace tags;

extern "C" {

int alipio_soil(TAGS_COMPLEX* sigma, TAGS_COMPLEX* epsr, double sigma0, TAGS_COMPLEX s, double h,
                double g, double eps_ratio)
{
    typedef std::complex<double> C;
    const c64 sv = from_c(s);
    const C f = C(sv.x, sv.y) / (kTwoPi * C(0.0, 1.0));
    const C sg = sigma0 + sigma0 * h * std::pow(f / 1e6, g);
    const double t = std::tan(kPi * g / 2) / (kTwoPi * kEps0 * std::pow(1e6, g));
    const C er = eps_ratio + t * sigma0 * h * std::pow(f, g - 1.0);
    *sigma = to_c(c64{sg.real(), sg.imag()});
    *epsr = to_c(c64{er.real(), er.imag()});
    return 0;
}

int smith_longmire_soil(TAGS_COMPLEX* sigma, TAGS_COMPLEX* epsr, double sigma0, TAGS_COMPLEX s,
                        double erinf)
{
    typedef std::complex<double> C;
    static const double a[13] = {3.4e6, 2.74e5, 2.58e4, 3.38e3, 5.26e2, 1.33e2, 2.72e1,
                                 1.25e1, 4.8e0, 2.17e0, 9.8e-1, 3.92e-1, 1.73e-1};
    const c64 sv = from_c(s);
    const C sc(sv.x, sv.y);
    const C fdc = std::pow(C(125.0 * sigma0, 0.0), 0.8312);
    C se = 0.0, ss = 0.0;
    for (int i = 0; i < 13; i++) {
        const C F = fdc * std::pow(C(10.0, 0.0), (double)i);
        const C r2 = std::pow(sc / (C(0.0, 1.0) * kTwoPi * F), 2.0);
        const C den = 1.0 + r2;
        se += a[i] / den;
        ss += a[i] * F * (r2 / den);
    }
    const C er = erinf + se;
    const C sg = sigma0 + kTwoPi * kEps0 * ss;
    *sigma = to_c(c64{sg.real(), sg.imag()});
    *epsr = to_c(c64{er.real(), er.imag()});
    return 0;
}

int tags_laplace_batch(const double* f, TAGS_COMPLEX* g, TAGS_COMPLEX* s, double tmax, size_t nt,
                       size_t howmany)
{
    if (nt < 2 || tmax <= 0) return TAGS_ERR_ARG;
    const long n = (long)nt, ns = n / 2 + 1, B = (long)howmany;
    const double c = std::log(std::pow((double)nt, 2.0)) / tmax;
    const double dt = tmax / (double)(nt - 1);
    const double dw = kTwoPi / tmax;
    c64* sg = (c64*)s;
    for (long k = 0; k < ns; k++) sg[k] = c64{c, dw * (double)k};
    if (B == 0) return 0;
    std::vector<double> wk(n);
    for (long k = 0; k < n; k++) wk[k] = dt * std::exp(-c * (double)k * dt);
    NltBuffers nb;
    int rc = nb.init(n, B);
    if (rc) return rc;
    // the real samples travel as doubles (staged in the second buffer)
    double* fd = (double*)nb.b;
    TAGS_HIP(hipMemcpyAsync(fd, f, sizeof(double) * n * B, hipMemcpyHostToDevice, nb.st));
    TAGS_HIP(hipMemcpyAsync(nb.v, wk.data(), sizeof(double) * n, hipMemcpyHostToDevice, nb.st));
    hipLaunchKernelGGL(k_weight_real, dim3(grid_for(n * B)), dim3(256), 0, nb.st, fd, nb.v, nb.a,
                       n, B);
    TAGS_HIP(hipGetLastError());
    if ((rc = dft_batch(nb.a, nb.b, nb.w, n, B, false, nb.st))) return rc;
    TAGS_HIP(hipMemcpy2DAsync(g, sizeof(c64) * ns, nb.a, sizeof(c64) * n, sizeof(c64) * ns, B,
                              hipMemcpyDeviceToHost, nb.st));
    TAGS_HIP(hipStreamSynchronize(nb.st));
    return 0;
}

// Data windows of src/auxiliary.c:271-305 (the Lanczos window is 0/0 at
// omega = 0, as in the reference).
static double inlt_window(int filter, double w, double omega)
{
    switch (filter) {
    case 1: {
        const double al = kPi * w / omega;
        return 0.42 + 0.5 * std::cos(al) + 0.08 * std::cos(2 * al);
    }
    case 2: return (1 + std::cos(kPi * w / omega)) / 2;
    case 3: {
        const double al = kPi * w / omega;
        return std::sin(al) / al;
    }
    case 4: return 1.0 - std::pow(std::fabs(w / omega), 2.0);
    default: return 1.0;
    }
}

int tags_inv_laplace_batch(double* f, const TAGS_COMPLEX* g, const TAGS_COMPLEX* s, double tmax,
                           size_t nt, int filter, size_t howmany)
{
    if (nt < 2 || tmax <= 0) return TAGS_ERR_ARG;
    const long n = (long)nt, ns = n / 2 + 1, B = (long)howmany;
    if (B == 0) return 0;
    const c64* gg = (const c64*)g;
    const c64* sv = (const c64*)s;
    const double dt = tmax / (double)(nt - 1);
    const double omega = sv[ns - 1].y;
    // window (ns) and output scale (n), by the reference's formulas on the host
    std::vector<double> win(ns), sc(n);
    for (long k = 0; k < ns; k++) win[k] = inlt_window(filter, sv[k].y, omega);
    const double c = sv[0].x;
    for (long k = 0; k < n; k++) sc[k] = std::exp(c * (double)k * dt) / ((double)nt * dt);
    NltBuffers nb;
    int rc = nb.init(n, B);
    if (rc) return rc;
    // the half spectra staged in the second buffer and the window after the
    // output scale in nb.v; the Hermitian extension built on the device
    TAGS_HIP(hipFree(nb.v));
    nb.v = nullptr;
    TAGS_HIP(hipMalloc(&nb.v, sizeof(double) * (n + ns)));
    double* dw = nb.v + n;
    TAGS_HIP(hipMemcpyAsync(nb.b, gg, sizeof(c64) * ns * B, hipMemcpyHostToDevice, nb.st));
    TAGS_HIP(hipMemcpyAsync(dw, win.data(), sizeof(double) * ns, hipMemcpyHostToDevice, nb.st));
    TAGS_HIP(hipMemcpyAsync(nb.v, sc.data(), sizeof(double) * n, hipMemcpyHostToDevice, nb.st));
    hipLaunchKernelGGL(k_hermitian, dim3(grid_for(n * B)), dim3(256), 0, nb.st, nb.b, dw, nb.a, n,
                       ns, B);
    TAGS_HIP(hipGetLastError());
    if ((rc = dft_batch(nb.a, nb.b, nb.w, n, B, true, nb.st))) return rc;
    double* fd = (double*)nb.b;
    hipLaunchKernelGGL(k_real_scale, dim3(grid_for(n * B)), dim3(256), 0, nb.st, nb.a, nb.v, fd,
                       n, B);
    TAGS_HIP(hipGetLastError());
    TAGS_HIP(hipMemcpyAsync(f, fd, sizeof(double) * n * B, hipMemcpyDeviceToHost, nb.st));
    TAGS_HIP(hipStreamSynchronize(nb.st));
    return 0;
}

int laplace_trans(double* f, TAGS_COMPLEX* g, TAGS_COMPLEX* s, double tmax, size_t nt)
{
    return tags_laplace_batch(f, g, s, tmax, nt, 1);
}

int inv_laplace_trans(double* f, TAGS_COMPLEX* g, TAGS_COMPLEX* s, double tmax, size_t nt,
                      int filter)
{
    return tags_inv_laplace_batch(f, g, s, tmax, nt, filter, 1);
}

}  // extern "C"
