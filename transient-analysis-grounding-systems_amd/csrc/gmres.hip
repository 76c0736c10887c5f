// Flexible GMRES in FP64, right-preconditioned by the complex64 LU factor: the
// correction solve of GMRES-based iterative refinement (Carson & Higham, "A
// new analysis of iterative refinement and its application to accurate
// solution of ill-conditioned sparse linear systems", SIAM J. Sci. Comput.
// 39 (2017); "Accelerating the solution of linear systems by iterative
// refinement in three precisions", SIAM J. Sci. Comput. 40 (2018)).
//
// One outer step solves W d = r approximately with k Arnoldi steps from d = 0:
//   v_1 = r / |r|_2;  for j = 1 .. k:
//     z_j = M^-1 v_j            (complex64 triangular solves of the factor)
//     w   = W z_j               (FP64 product with the resident W)
//     h   = V_j^H w, w -= V_j h; twice (classical Gram-Schmidt, two passes)
//     h_{j+1,j} = |w|_2, v_{j+1} = w / h_{j+1,j}
//     Givens rotations keep H upper triangular and g = Q^H |r|_2 e_1
//   y = R^-1 g,  d = Z y,  x += d.
// Everything runs on the device, for every item of a batch (blockIdx.z), with
// no host round trip inside a step. Dot products reduce in two stages in a
// fixed order, so results do not depend on timing.
#include "lu.h"

namespace tags {

constexpr int GM_CHUNK = 4096;  // vector entries per work-group of a dot product

// wave sum (lane shuffles, fixed order)
__device__ inline c64 gm_wave_sum(c64 v)
{
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    return v;
}

// part[(z * cnt + i) * nchunk + c] = sum over chunk c of conj(V_i[t]) w[t]
__global__ __launch_bounds__(256) void k_gm_dot_part(const c64* V, long sV, int cnt, const c64* w,
                                                      long sw, long n, c64* part, int nchunk)
{
    const long z = blockIdx.z;
    const int i = blockIdx.y, c = blockIdx.x;
    const c64* v = V + z * sV + (long)i * n;
    const c64* ww = w + z * sw;
    const long t0 = (long)c * GM_CHUNK, t1 = min(n, t0 + GM_CHUNK);
    c64 s = c64{0.0, 0.0};
    for (long t = t0 + threadIdx.x; t < t1; t += 256) {
        const c64 a = v[t], b = ww[t];
        s.x += a.x * b.x + a.y * b.y;
        s.y += a.x * b.y - a.y * b.x;
    }
    s = gm_wave_sum(s);
    __shared__ c64 red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        const c64 r = c64{(red[0].x + red[1].x) + (red[2].x + red[3].x),
                          (red[0].y + red[1].y) + (red[2].y + red[3].y)};
        part[(z * cnt + i) * nchunk + c] = r;
    }
}

// out[z * ldo + i] = sum_c part[(z * cnt + i) * nchunk + c], in chunk order
__global__ void k_gm_dot_final(const c64* part, int cnt, int nchunk, c64* out, long ldo)
{
    const long z = blockIdx.z;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        c64 s = c64{0.0, 0.0};
        for (int c = 0; c < nchunk; c++) s = cadd(s, part[(z * cnt + i) * nchunk + c]);
        out[z * ldo + i] = s;
    }
}

// w[t] -= sum_{i < cnt} h[i] V_i[t]
__global__ __launch_bounds__(256) void k_gm_axpy(c64* w, long sw, const c64* V, long sV,
                                                  const c64* h, long sh, int cnt, long n)
{
    const long z = blockIdx.z;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const c64* v = V + z * sV;
    const c64* hz = h + z * sh;
    c64 a = w[z * sw + t];
    for (int i = 0; i < cnt; i++) a = csub(a, cmul(hz[i], v[(long)i * n + t]));
    w[z * sw + t] = a;
}

// w = W z (row-major W, one work-group per row)
__global__ __launch_bounds__(256) void k_gm_matvec(const c64* W, long ld, long n, long sW,
                                                    const c64* x, long sx, c64* w, long sw)
{
    const long z = blockIdx.z;
    const long row = blockIdx.x;
    const c64* Wr = W + z * sW + row * ld;
    const c64* xz = x + z * sx;
    double sx_ = 0, sy_ = 0;
    for (long c = threadIdx.x; c < n; c += 256) {
        const c64 a = Wr[c], v = xz[c];
        sx_ += a.x * v.x - a.y * v.y;
        sy_ += a.x * v.y + a.y * v.x;
    }
    c64 s = gm_wave_sum(c64{sx_, sy_});
    __shared__ c64 red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        w[z * sw + row] = c64{(red[0].x + red[1].x) + (red[2].x + red[3].x),
                              (red[0].y + red[1].y) + (red[2].y + red[3].y)};
}

// t[i] = (float) (d[p] v[p]), p = perm[i] (the factor's logical order)
__global__ void k_gm_prep(const c64* v, long sv, const double* d, const int* perm, long n, c32* t)
{
    const long z = blockIdx.z;
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int p = perm[z * n + i];
    const double s = d[z * n + p];
    const c64 a = v[z * sv + p];
    t[z * n + i] = c32{(float)(a.x * s), (float)(a.y * s)};
}

// zj[j] = d[j] t[j] (the preconditioned vector, natural order, in FP64)
__global__ void k_gm_storez(c64* zj, long sz, const c32* t, const double* d, long n)
{
    const long z = blockIdx.z;
    const long j = (long)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const double s = d[z * n + j];
    const c32 a = t[z * n + j];
    zj[z * sz + j] = c64{s * (double)a.x, s * (double)a.y};
}

// beta = |r|_2 from nrm (the dot of r with itself): g = beta e_1, inv = 1 / beta
__global__ void k_gm_beta(const c64* nrm, c64* g, long sg, int kmax, double* inv)
{
    const long z = blockIdx.z;
    if (threadIdx.x != 0) return;
    const double beta = sqrt(fmax(nrm[z].x, 0.0));
    c64* gz = g + z * sg;
    gz[0] = c64{beta, 0.0};
    for (int i = 1; i <= kmax; i++) gz[i] = c64{0.0, 0.0};
    inv[z] = beta > 0.0 ? 1.0 / beta : 0.0;
}

// v = w inv[z]
__global__ void k_gm_scale(c64* v, long sv, const c64* w, long sw, const double* inv, long n)
{
    const long z = blockIdx.z;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    const double s = inv[z];
    const c64 a = w[z * sw + t];
    v[z * sv + t] = c64{a.x * s, a.y * s};
}

// Column j of H from the two Gram-Schmidt passes (h1 + h2) and |w|_2, the
// earlier rotations applied, a new rotation annihilating H[j+1][j], g rotated;
// inv = 1 / |w|_2 for v_{j+1}. One thread per item (j < kmax <= 32).
__global__ void k_gm_givens(const c64* h1, const c64* h2, long sh, const c64* nrm, c64* H,
                            long sH, int ldh, double* cs, c64* sn, long sr, c64* g, long sg,
                            int j, double* inv)
{
    const long z = blockIdx.z;
    if (threadIdx.x != 0) return;
    c64* Hz = H + z * sH;
    c64* col = Hz + (long)j * ldh;
    const c64* a1 = h1 + z * sh;
    const c64* a2 = h2 + z * sh;
    for (int i = 0; i <= j; i++) col[i] = cadd(a1[i], a2[i]);
    const double hn = sqrt(fmax(nrm[z].x, 0.0));
    col[j + 1] = c64{hn, 0.0};
    inv[z] = hn > 0.0 ? 1.0 / hn : 0.0;
    double* c = cs + z * sr;
    c64* s = sn + z * sr;
    // earlier rotations: [c, s; -conj(s), c]
    for (int i = 0; i < j; i++) {
        const c64 a = col[i], b = col[i + 1];
        const c64 sb = cmul(s[i], b);
        const c64 cs_ = c64{-s[i].x, s[i].y};  // -conj(s)
        col[i] = c64{c[i] * a.x + sb.x, c[i] * a.y + sb.y};
        const c64 t = cmul(cs_, a);
        col[i + 1] = c64{t.x + c[i] * b.x, t.y + c[i] * b.y};
    }
    // the new rotation: c = |a| / rho, s = (a / |a|) conj(b) / rho
    const c64 a = col[j], b = col[j + 1];
    const double na = sqrt(a.x * a.x + a.y * a.y), nb = sqrt(b.x * b.x + b.y * b.y);
    const double rho = sqrt(na * na + nb * nb);
    double cj;
    c64 sj, hj;
    if (rho == 0.0) {
        cj = 1.0;
        sj = c64{0.0, 0.0};
        hj = c64{0.0, 0.0};
    } else if (na == 0.0) {
        cj = 0.0;
        sj = c64{b.x / nb, -b.y / nb};
        hj = c64{nb, 0.0};
    } else {
        const c64 ph = c64{a.x / na, a.y / na};
        cj = na / rho;
        sj = cscale(cmul(ph, c64{b.x, -b.y}), 1.0 / rho);
        hj = cscale(ph, rho);
    }
    c[j] = cj;
    s[j] = sj;
    col[j] = hj;
    col[j + 1] = c64{0.0, 0.0};
    c64* gz = g + z * sg;
    const c64 gj = gz[j];
    gz[j] = cscale(gj, cj);
    const c64 t = cmul(c64{-sj.x, sj.y}, gj);
    gz[j + 1] = t;
}

// y = R^-1 g (R: the leading k x k of H, upper triangular); a zero pivot gives
// a zero component
__global__ void k_gm_solve_y(const c64* H, long sH, int ldh, const c64* g, long sg, c64* y,
                             long sy, int k)
{
    const long z = blockIdx.z;
    if (threadIdx.x != 0) return;
    const c64* Hz = H + z * sH;
    const c64* gz = g + z * sg;
    c64* yz = y + z * sy;
    for (int i = k - 1; i >= 0; i--) {
        c64 s = gz[i];
        for (int l = i + 1; l < k; l++) s = csub(s, cmul(Hz[(long)l * ldh + i], yz[l]));
        const c64 d = Hz[(long)i * ldh + i];
        yz[i] = (d.x == 0.0 && d.y == 0.0) ? c64{0.0, 0.0} : cdiv(s, d);
    }
}

// x += Z y (skipped where active[z] == 0); dxn: max |Z y| (zeroed beforehand)
__global__ __launch_bounds__(256) void k_gm_update(c64* x, const c64* Z, long sZ, const c64* y,
                                                    long sy, int k, long n,
                                                    unsigned long long* dxn, long sn,
                                                    const int* active)
{
    const long z = blockIdx.z;
    if (active && !active[z]) return;
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    double a = 0.0;
    if (t < n) {
        const c64* Zz = Z + z * sZ;
        const c64* yz = y + z * sy;
        c64 d = c64{0.0, 0.0};
        for (int j = 0; j < k; j++) d = cadd(d, cmul(yz[j], Zz[(long)j * n + t]));
        const c64 o = x[z * n + t];
        x[z * n + t] = cadd(o, d);
        a = sqrt(d.x * d.x + d.y * d.y);
    }
    for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off));
    if ((threadIdx.x & 63) == 0 && dxn)
        atomicMax(dxn + z * sn, (unsigned long long)__double_as_longlong(a));
}

int GmresWork::alloc(int n_, int kmax_, int cap_)
{
    if (n_ <= 0 || kmax_ < 1 || kmax_ > 32 || cap_ < 1) return TAGS_ERR_ARG;
    const long nch_ = (n_ + GM_CHUNK - 1) / GM_CHUNK;
    if (n_ * (long)cap_ <= n_cap && kmax_ <= kmax_cap && cap_ <= item_cap &&
        (long)(kmax_ + 1) * nch_ * cap_ <= part_cap) {
        n = n_;
        kmax = kmax_;
        cap = cap_;
        nchunk = (int)nch_;
        return 0;
    }
    // grow to cover every earlier request as well (a solver that serves
    // systems of different orders in turn reallocates, which waits for the
    // whole device, only while its largest request grows)
    const size_t NB = std::max((size_t)n_ * (size_t)cap_, (size_t)n_cap);
    const size_t K = (size_t)std::max(kmax_, kmax_cap);
    const size_t B = (size_t)std::max(cap_, item_cap);
    const size_t PB = std::max((size_t)(kmax_ + 1) * nch_ * cap_, (size_t)part_cap);
    release();
    n = n_;
    kmax = kmax_;
    cap = cap_;
    nchunk = (int)nch_;
    TAGS_HIP(hipMalloc(&V, sizeof(c64) * NB * (K + 1)));
    TAGS_HIP(hipMalloc(&Z, sizeof(c64) * NB * K));
    TAGS_HIP(hipMalloc(&w, sizeof(c64) * NB));
    TAGS_HIP(hipMalloc(&H, sizeof(c64) * (K + 1) * K * B));
    TAGS_HIP(hipMalloc(&g, sizeof(c64) * (K + 1) * B));
    TAGS_HIP(hipMalloc(&y, sizeof(c64) * K * B));
    TAGS_HIP(hipMalloc(&h1, sizeof(c64) * (K + 1) * B));
    TAGS_HIP(hipMalloc(&h2, sizeof(c64) * (K + 1) * B));
    TAGS_HIP(hipMalloc(&nrm, sizeof(c64) * B));
    TAGS_HIP(hipMalloc(&sn, sizeof(c64) * K * B));
    TAGS_HIP(hipMalloc(&cs, sizeof(double) * K * B));
    TAGS_HIP(hipMalloc(&inv, sizeof(double) * B));
    TAGS_HIP(hipMalloc(&part, sizeof(c64) * PB));
    n_cap = (long)NB;
    kmax_cap = (int)K;
    item_cap = (int)B;
    part_cap = (long)PB;
    return 0;
}

void GmresWork::release()
{
    hipFree(V);
    hipFree(Z);
    hipFree(w);
    hipFree(H);
    hipFree(g);
    hipFree(y);
    hipFree(h1);
    hipFree(h2);
    hipFree(nrm);
    hipFree(sn);
    hipFree(cs);
    hipFree(inv);
    hipFree(part);
    V = Z = w = H = g = y = h1 = h2 = nrm = sn = part = nullptr;
    cs = inv = nullptr;
    n_cap = part_cap = 0;
    kmax_cap = item_cap = 0;
}

// dots[z * ldo + i] = V_i^H w for i < cnt (V at stride sV per item, w at sw)
static int gm_dots(GmresWork& g, int nb, const c64* V, long sV, int cnt, const c64* w, long sw,
                   c64* out, long ldo, hipStream_t st)
{
    const int nch = (int)((g.n + GM_CHUNK - 1) / GM_CHUNK);
    hipLaunchKernelGGL(k_gm_dot_part, dim3(nch, cnt, nb), dim3(256), 0, st, V, sV, cnt, w, sw,
                       (long)g.n, g.part, nch);
    hipLaunchKernelGGL(k_gm_dot_final, dim3(1, 1, nb), dim3(64), 0, st, g.part, cnt, nch, out,
                       ldo);
    TAGS_HIP(hipGetLastError());
    return 0;
}

int gmres_step(GmresWork& g, int k, int nb, LUWork<float>& w32, const c32* W32, c32* t32,
               const double* d, const c64* W, long sW, const c64* r, c64* x,
               unsigned long long* dxn, long sn, const int* active, hipStream_t st, Profiler* pr,
               const HemOp* hem)
{
    const long n = g.n;
    if (k < 1 || k > g.kmax || nb < 1 || nb > g.cap) return TAGS_ERR_ARG;
    const long sV = n * (g.kmax + 1), sZ = n * g.kmax, sH = (long)(g.kmax + 1) * g.kmax;
    const int ldh = g.kmax + 1;
    const long sg = g.kmax + 1, sh = g.kmax + 1, sr = g.kmax;
    const dim3 vg((unsigned)((n + 255) / 256), 1, (unsigned)nb);
    int rc;
    Timed tm(pr, FAM_SOLVE, st);
    // v_1 = r / |r|_2, g = |r|_2 e_1
    if ((rc = gm_dots(g, nb, r, n, 1, r, n, g.nrm, 1, st))) return rc;
    hipLaunchKernelGGL(k_gm_beta, dim3(1, 1, nb), dim3(64), 0, st, g.nrm, g.g, sg, g.kmax, g.inv);
    hipLaunchKernelGGL(k_gm_scale, vg, dim3(256), 0, st, g.V, sV, r, n, g.inv, n);
    for (int j = 0; j < k; j++) {
        c64* vj = g.V + (long)j * n;
        c64* zj = g.Z + (long)j * n;
        // z_j = M^-1 v_j
        hipLaunchKernelGGL(k_gm_prep, vg, dim3(256), 0, st, vj, sV, d, w32.perm, n, t32);
        TAGS_HIP(hipGetLastError());
        if ((rc = lu_solve_inplace<float>(w32, W32, (long)n, t32, st, nullptr))) return rc;
        hipLaunchKernelGGL(k_gm_storez, vg, dim3(256), 0, st, zj, sZ, t32, d, n);
        // w = W z_j
        if (hem) {
            if ((rc = hem_apply(W, n, *hem, zj, sZ, nullptr, g.w, n, nullptr, 0, 0, nb, sW,
                                nullptr, st)))
                return rc;
        } else {
            hipLaunchKernelGGL(k_gm_matvec, dim3((unsigned)n, 1, (unsigned)nb), dim3(256), 0, st,
                               W, n, n, sW, zj, sZ, g.w, n);
        }
        // two Gram-Schmidt passes against v_1 .. v_j
        if ((rc = gm_dots(g, nb, g.V, sV, j + 1, g.w, n, g.h1, sh, st))) return rc;
        hipLaunchKernelGGL(k_gm_axpy, vg, dim3(256), 0, st, g.w, n, g.V, sV, g.h1, sh, j + 1, n);
        if ((rc = gm_dots(g, nb, g.V, sV, j + 1, g.w, n, g.h2, sh, st))) return rc;
        hipLaunchKernelGGL(k_gm_axpy, vg, dim3(256), 0, st, g.w, n, g.V, sV, g.h2, sh, j + 1, n);
        if ((rc = gm_dots(g, nb, g.w, n, 1, g.w, n, g.nrm, 1, st))) return rc;
        hipLaunchKernelGGL(k_gm_givens, dim3(1, 1, nb), dim3(64), 0, st, g.h1, g.h2, sh, g.nrm,
                           g.H, sH, ldh, g.cs, g.sn, sr, g.g, sg, j, g.inv);
        if (j + 1 < k)
            hipLaunchKernelGGL(k_gm_scale, vg, dim3(256), 0, st, g.V + (long)(j + 1) * n, sV,
                               g.w, n, g.inv, n);
        TAGS_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_gm_solve_y, dim3(1, 1, nb), dim3(64), 0, st, g.H, sH, ldh, g.g, sg, g.y,
                       (long)g.kmax, k);
    hipLaunchKernelGGL(k_gm_update, vg, dim3(256), 0, st, x, g.Z, sZ, g.y, (long)g.kmax, k, n,
                       dxn, sn, active);
    TAGS_HIP(hipGetLastError());
    tm.end();
    return 0;
}

}  // namespace tags
