// Shared definitions of the HIP implementation of the per-frequency HEM solve.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cmath>
#include <cstdio>

#include "tags_hip.h"

namespace tags {

// Interleaved complex value as the reference stores _Complex double.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};
using c64 = cx<double>;
using c32 = cx<float>;

template <typename T>
__host__ __device__ inline cx<T> cmk(T a, T b) { return cx<T>{a, b}; }
template <typename T>
__host__ __device__ inline cx<T> cadd(cx<T> a, cx<T> b) { return cx<T>{a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ inline cx<T> csub(cx<T> a, cx<T> b) { return cx<T>{a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
    return cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ inline cx<T> cscale(cx<T> a, T s) { return cx<T>{a.x * s, a.y * s}; }
// Reciprocal with the larger component as pivot (Smith's scaling) to avoid
// overflow for tiny or huge values.
template <typename T>
__host__ __device__ inline cx<T> crecip(cx<T> a) {
    T ax = a.x < 0 ? -a.x : a.x, ay = a.y < 0 ? -a.y : a.y;
    if (ax >= ay) {
        T r = a.y / a.x, d = a.x + a.y * r;
        return cx<T>{T(1) / d, -r / d};
    }
    T r = a.x / a.y, d = a.y + a.x * r;
    return cx<T>{r / d, T(-1) / d};
}
template <typename T>
__host__ __device__ inline cx<T> cdiv(cx<T> a, cx<T> b) { return cmul(a, crecip(b)); }
// a - l * u with four fused multiply-adds
template <typename T>
__host__ __device__ inline cx<T> cfnma(cx<T> a, cx<T> l, cx<T> u)
{
    T x = fma(-l.x, u.x, a.x);
    x = fma(l.y, u.y, x);
    T y = fma(-l.x, u.y, a.y);
    y = fma(-l.y, u.x, y);
    return cx<T>{x, y};
}

// c ? a : b, component-wise (a select of the struct itself would be lowered to
// a copy through a selected pointer, which keeps register arrays in memory)
template <typename T>
__host__ __device__ inline cx<T> csel(bool c, cx<T> a, cx<T> b)
{
    return cx<T>{c ? a.x : b.x, c ? a.y : b.y};
}
// |re| + |im| with the sign bit cleared (also for -0.0, which matters where the
// result is compared as an unsigned bit pattern)
__host__ __device__ inline double cabs1(cx<double> a) { return fabs(a.x) + fabs(a.y); }
__host__ __device__ inline float cabs1(cx<float> a) { return fabsf(a.x) + fabsf(a.y); }

// 1 / p for p != 0: p is scaled by a power of two so that |p|^2 lies in
// [1/4, 2), then the hardware reciprocal is refined by Newton steps (one for
// float, two for double) and the scale undone.
__device__ inline float rcp_nr(float d)
{
    float r = __builtin_amdgcn_rcpf(d);
    return r * fmaf(-d, r, 2.0f);
}
__device__ inline double rcp_nr(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    r = r * fma(-d, r, 2.0);
    return r * fma(-d, r, 2.0);
}
template <typename T>
__device__ inline cx<T> crecip_fast(cx<T> p)
{
    int e;
    (void)frexp(fmax(fabs(p.x), fabs(p.y)), &e);
    const T sx = ldexp(p.x, -e), sy = ldexp(p.y, -e);
    const T r = rcp_nr(fma(sx, sx, sy * sy));
    return cx<T>{ldexp(sx * r, -e), ldexp(-sy * r, -e)};
}

inline c64 from_c(double _Complex z) { return c64{__real__ z, __imag__ z}; }
inline double _Complex to_c(c64 z) {
    double _Complex r;
    __real__ r = z.x;
    __imag__ r = z.y;
    return r;
}

// Library error codes beyond the reference's own (all negative and below -10,
// which the reference already uses for an unknown integration type).
enum : int {
    TAGS_ERR_HIP = -100,          // a HIP runtime call failed
    TAGS_ERR_UNSUPPORTED = -11,   // integration type not implemented on the GPU
    TAGS_ERR_ARG = -12,           // bad size or pointer
    TAGS_ERR_TOO_LARGE = -13,     // problem exceeds a kernel's design limit
    TAGS_ERR_BARRIER = -14,       // a grid barrier timed out (kernel exited early)
};

// Reports and converts a HIP error; used by every host entry point.
#define TAGS_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "tags_hip: %s failed: %s (%s:%d)\n", #call,           \
                         hipGetErrorString(e_), __FILE__, __LINE__);                    \
            return TAGS_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

// Physical constants with the reference's literal values (src/auxiliary.c:17-21).
constexpr double kPi = 3.1415926535897932384626433832795029;
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kFourPi = 12.5663706143591729538505735331;
constexpr double kMu0 = 1.256637061435917e-6;
constexpr double kEps0 = 8.854187817620e-12;

// Segment geometry in structure-of-arrays form on the device.
struct SegSoA {
    const double* ax;  // start point
    const double* ay;
    const double* az;
    const double* bx;  // end point
    const double* by;
    const double* bz;
    const double* len;
    const double* rad;
};

// Scalars of one frequency sample that the pair kernels need.
struct PairScalars {
    c64 gamma;
    c64 zl_dir, zt_dir;   // s mu / 4 pi and 1 / (4 pi kappa)
    c64 zl_img, zt_img;   // ref_l s mu / 4 pi and ref_t / (4 pi kappa)
    int intg_type;        // INTG_DOUBLE, INTG_MHEM or INTG_NONE
};

// Placement of the three unknown blocks of the immittance system (row and
// column offsets): U (n nodes), I_L and I_T (m segments each). The reference
// order is node_off = 0, zl_off = n, zt_off = n + m.
struct BlockLayout {
    long n, m;
    long node_off, zl_off, zt_off;
};

// ---- fill.hip ----
int gl_tables_init();
int launch_pairs_dropin(double* zl, double* zt, bool alias, SegSoA e, SegSoA img,
                        const long m, const PairScalars& sc, bool images, hipStream_t st);
// Gauss-Legendre nodes and weights of order n on [0, 1] (host; Newton on P_n)
void gauss_legendre(int n, double* x, double* w);
int launch_mhem_potentials(double* pl, double* pt, double* pli, double* pti, long m, SegSoA e,
                           SegSoA img, hipStream_t st);
// Fills of nb frequency samples at once (nb > 1: W of item z at W + z sw, its
// scalars scs[z] in device memory; nb == 1 with scs null: `sc`).
int launch_fill_mhem(c64* W, long ldw, const BlockLayout& L, SegSoA e, SegSoA img,
                     const double* pl, const double* pt, const double* pli, const double* pti,
                     const PairScalars& sc, hipStream_t st, int nb = 1, long sw = 0,
                     const PairScalars* scs = nullptr);
int launch_fill_zblocks(c64* W, long ldw, const BlockLayout& L, SegSoA e, SegSoA img,
                        const PairScalars& sc, hipStream_t st, int nb = 1, long sw = 0,
                        const PairScalars* scs = nullptr);
int launch_fill_static(c64* W, long ldw, const BlockLayout& L, const int* seg_start_node,
                       const int* seg_end_node, hipStream_t st, int nb = 1, long sw = 0);
int launch_single_pair(SegSoA e, SegSoA r, c64 gamma, int type, double* out4, hipStream_t st);
int launch_incidence(const double* seg_pts, long m, const double* nodes, long n,
                     int* seg_start_node, int* seg_end_node, int* node_hits,
                     hipStream_t st);

}  // namespace tags
