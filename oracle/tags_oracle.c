/* CPU oracle for the per-frequency HEM solve -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates the reference algorithm so the HIP library can be checked against
 * it; it is never part of the product path. Each function cites the reference
 * file:line it follows (paths relative to the reference repository root).
 *
 * The reference integrates with the third-party h-adaptive `hcubature`
 * (stevengj/cubature, git submodule, absent from the reference tree; version
 * unknown). `or_hcubature2` below is an independent implementation of the
 * published Genz-Malik (1980) degree-7 rule with its embedded degree-5 error
 * estimate, bisection along the axis of largest fourth difference, a max-heap of
 * regions keyed on the paired (complex) error, and the (abs, rel, max_eval)
 * stopping test with paired error norm, which is the behaviour the reference
 * requests with ERROR_PAIRED (src/electrode.c:273-275).
 */
#include "tags_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants: src/auxiliary.c:17-21 (same literal values). */
const double OR_PI = 3.1415926535897932384626433832795029;
const double OR_TWO_PI = 6.283185307179586476925286766559;
const double OR_FOUR_PI = 12.5663706143591729538505735331;
const double OR_MU0 = 1.256637061435917e-6;
const double OR_EPS0 = 8.854187817620e-12;

/* src/auxiliary.c:110-125: Euclidean distance accumulated with pow(d, 2). */
double or_vector_length(const double a[3], const double b[3])
{
    double r = 0.0;
    for (int i = 0; i < 3; i++) r += pow(a[i] - b[i], 2.0);
    return sqrt(r);
}

/* src/auxiliary.c:24-36 */
int or_linspace(double a, double b, size_t n, double *u)
{
    if (n < 2 || !u) return 1;
    double c = (b - a) / (n - 1);
    for (size_t i = 0; i < n - 1; ++i) u[i] = a + i * c;
    u[n - 1] = b;
    return 0;
}

/* src/auxiliary.c:38-50 */
int or_logspace(double a, double b, size_t n, double *u)
{
    if (n < 2 || !u) return 1;
    double c = (b - a) / (n - 1);
    for (size_t i = 0; i < n - 1; ++i) u[i] = pow(10.0, a + i * c);
    u[n - 1] = pow(10.0, b);
    return 0;
}

/* src/auxiliary.c:101-108 */
static int or_equal_points_tol(const double *p, const double *q, double tol)
{
    for (int i = 0; i < 3; i++)
        if (fabs(p[i] - q[i]) > tol) return 0;
    return 1;
}

/* src/electrode.c:12-32 */
int or_populate_electrode(or_electrode *e, const double start[3], const double end[3],
                          double radius)
{
    if (or_equal_points_tol(start, end, DBL_EPSILON)) return 1;
    if (radius <= 0) return 2;
    e->radius = radius;
    e->length = or_vector_length(start, end);
    for (int i = 0; i < 3; i++) {
        e->start_point[i] = start[i];
        e->end_point[i] = end[i];
        e->middle_point[i] = (start[i] + end[i]) / 2.0;
    }
    return 0;
}

/* src/electrode.c:103-127 */
int or_segment_electrode(or_electrode *e, double *nodes, size_t num_segments,
                         const double *start, const double *end, double radius)
{
    if (num_segments < 1) return 1;
    size_t nn = num_segments + 1;
    double *x = malloc(3 * nn * sizeof(double));
    or_linspace(start[0], end[0], nn, x);
    or_linspace(start[1], end[1], nn, x + nn);
    or_linspace(start[2], end[2], nn, x + 2 * nn);
    for (size_t i = 0; i < nn; i++) {
        nodes[3 * i] = x[i];
        nodes[3 * i + 1] = x[nn + i];
        nodes[3 * i + 2] = x[2 * nn + i];
    }
    free(x);
    for (size_t i = 0; i < num_segments; i++)
        or_populate_electrode(&e[i], nodes + 3 * i, nodes + 3 * i + 3, radius);
    return 0;
}

/* src/grid.c:9-17 */
int or_number_segments(or_grid g)
{
    const unsigned N = g.edge_segments_x, vx = g.vertices_x;
    const unsigned M = g.edge_segments_y, vy = g.vertices_y;
    return (int)(N * vy * (vx - 1) + M * vx * (vy - 1));
}

/* src/grid.c:19-27 */
int or_number_nodes(or_grid g)
{
    const unsigned N = g.edge_segments_x, vx = g.vertices_x;
    const unsigned M = g.edge_segments_y, vy = g.vertices_y;
    return (int)(vx * vy + vx * (vy - 1) * (M - 1) + vy * (vx - 1) * (N - 1));
}

/* src/grid.c:29-91: horizontal edges row by row, then vertical edges column by
 * column; node 0 is the corner (0, 0, depth). */
int or_electrode_grid(or_grid g, or_electrode *e, double *nodes)
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
                or_populate_electrode(&e[ed++], sp, ep, g.radius);
                if (n == 0 && k == 0) {
                    memcpy(nodes + 3 * nd, sp, 3 * sizeof(double));
                    nd++;
                }
                memcpy(nodes + 3 * nd, ep, 3 * sizeof(double));
                nd++;
            }
    for (size_t gg = 0; gg < vx; gg++)
        for (size_t mm = 0; mm < vy - 1; mm++)
            for (size_t k = 0; k < M; k++) {
                sp[0] = lx * N * gg;
                sp[1] = ly * (mm * M + k);
                ep[0] = sp[0];
                ep[1] = sp[1] + ly;
                or_populate_electrode(&e[ed++], sp, ep, g.radius);
                if (k < M - 1) {
                    memcpy(nodes + 3 * nd, ep, 3 * sizeof(double));
                    nd++;
                }
            }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Adaptive Genz-Malik cubature on [0,1]^2, two components (re, im).         */
/* ------------------------------------------------------------------------- */
typedef struct {
    double c[2], h[2];
    double val[2], err[2];
    double key;   /* paired error, heap key */
    int split;
} gm_region;

static void gm_eval(gm_region *R, or_integrand2 f, void *data)
{
    /* Genz-Malik degree-7 rule for n = 2 with the embedded degree-5 rule. */
    const double l2 = sqrt(9.0 / 70.0), l4 = sqrt(9.0 / 10.0), l5 = sqrt(9.0 / 19.0);
    const double w1 = -3816.0 / 19683.0, w2 = 980.0 / 6561.0, w3 = 1020.0 / 19683.0;
    const double w4 = 200.0 / 19683.0, w5 = 6859.0 / 78732.0;
    const double v1 = -971.0 / 729.0, v2 = 245.0 / 486.0, v3 = 65.0 / 1458.0;
    const double v4 = 25.0 / 729.0;
    double f0[2], t[2], o[2], op[2];
    double s2[2] = {0, 0}, s3[2] = {0, 0}, s4[2] = {0, 0}, s5[2] = {0, 0};
    double diff[2];
    t[0] = R->c[0];
    t[1] = R->c[1];
    f(t, data, f0);
    for (int d = 0; d < 2; d++) {
        double a2[2] = {0, 0}, a3[2] = {0, 0};
        t[0] = R->c[0];
        t[1] = R->c[1];
        t[d] = R->c[d] + l2 * R->h[d];
        f(t, data, o);
        t[d] = R->c[d] - l2 * R->h[d];
        f(t, data, op);
        a2[0] = o[0] + op[0];
        a2[1] = o[1] + op[1];
        t[d] = R->c[d] + l4 * R->h[d];
        f(t, data, o);
        t[d] = R->c[d] - l4 * R->h[d];
        f(t, data, op);
        a3[0] = o[0] + op[0];
        a3[1] = o[1] + op[1];
        s2[0] += a2[0];
        s2[1] += a2[1];
        s3[0] += a3[0];
        s3[1] += a3[1];
        /* fourth difference along d (ratio (l2/l4)^2 = 1/7) */
        diff[d] = fabs(a2[0] - 2 * f0[0] - (a3[0] - 2 * f0[0]) / 7.0) +
                  fabs(a2[1] - 2 * f0[1] - (a3[1] - 2 * f0[1]) / 7.0);
    }
    for (int sx = -1; sx <= 1; sx += 2)
        for (int sy = -1; sy <= 1; sy += 2) {
            t[0] = R->c[0] + sx * l4 * R->h[0];
            t[1] = R->c[1] + sy * l4 * R->h[1];
            f(t, data, o);
            s4[0] += o[0];
            s4[1] += o[1];
            t[0] = R->c[0] + sx * l5 * R->h[0];
            t[1] = R->c[1] + sy * l5 * R->h[1];
            f(t, data, o);
            s5[0] += o[0];
            s5[1] += o[1];
        }
    const double vol = 4.0 * R->h[0] * R->h[1];
    for (int k = 0; k < 2; k++) {
        double i7 = vol * (w1 * f0[k] + w2 * s2[k] + w3 * s3[k] + w4 * s4[k] + w5 * s5[k]);
        double i5 = vol * (v1 * f0[k] + v2 * s2[k] + v3 * s3[k] + v4 * s4[k]);
        R->val[k] = i7;
        R->err[k] = fabs(i7 - i5);
    }
    R->key = sqrt(R->err[0] * R->err[0] + R->err[1] * R->err[1]);
    /* split along the larger fourth difference; near-ties go to the wider side */
    if (fabs(diff[0] - diff[1]) <= 1e-14 * fmax(diff[0], diff[1]))
        R->split = (R->h[1] > R->h[0]) ? 1 : 0;
    else
        R->split = (diff[1] > diff[0]) ? 1 : 0;
}

typedef struct {
    gm_region *a;
    size_t n, cap;
} gm_heap;

static void heap_push(gm_heap *H, gm_region r)
{
    if (H->n == H->cap) {
        H->cap = H->cap ? 2 * H->cap : 64;
        H->a = realloc(H->a, H->cap * sizeof(gm_region));
    }
    size_t i = H->n++;
    while (i > 0) {
        size_t p = (i - 1) / 2;
        if (H->a[p].key >= r.key) break;
        H->a[i] = H->a[p];
        i = p;
    }
    H->a[i] = r;
}

static gm_region heap_pop(gm_heap *H)
{
    gm_region top = H->a[0];
    gm_region last = H->a[--H->n];
    size_t i = 0;
    for (;;) {
        size_t l = 2 * i + 1, r = l + 1, m = i;
        double km = last.key;
        if (l < H->n && H->a[l].key > km) {
            m = l;
            km = H->a[l].key;
        }
        if (r < H->n && H->a[r].key > km) m = r;
        if (m == i) break;
        H->a[i] = H->a[m];
        i = m;
    }
    if (H->n) H->a[i] = last;
    return top;
}

int or_hcubature2(or_integrand2 f, void *data, size_t max_eval, double abs_tol,
                  double rel_tol, double val[2], double err[2], size_t *n_eval)
{
    const size_t pts = 17;
    gm_heap H = {0, 0, 0};
    gm_region R = {{0.5, 0.5}, {0.5, 0.5}, {0, 0}, {0, 0}, 0, 0};
    gm_eval(&R, f, data);
    size_t ne = pts;
    heap_push(&H, R);
    double tv[2] = {R.val[0], R.val[1]}, te[2] = {R.err[0], R.err[1]};
    size_t iter = 0;
    for (;;) {
        double e = sqrt(te[0] * te[0] + te[1] * te[1]);
        double v = sqrt(tv[0] * tv[0] + tv[1] * tv[1]);
        if (e <= abs_tol || e <= rel_tol * v) break;
        if (max_eval && ne + 2 * pts > max_eval) break;
        gm_region P = heap_pop(&H);
        gm_region A = P, B = P;
        int d = P.split;
        A.h[d] = B.h[d] = 0.5 * P.h[d];
        A.c[d] = P.c[d] - A.h[d];
        B.c[d] = P.c[d] + B.h[d];
        gm_eval(&A, f, data);
        gm_eval(&B, f, data);
        ne += 2 * pts;
        heap_push(&H, A);
        heap_push(&H, B);
        if (++iter % 64 == 0) {
            /* re-sum from the heap to stop drift of the running totals */
            tv[0] = tv[1] = te[0] = te[1] = 0;
            for (size_t i = 0; i < H.n; i++)
                for (int k = 0; k < 2; k++) {
                    tv[k] += H.a[i].val[k];
                    te[k] += H.a[i].err[k];
                }
        } else {
            for (int k = 0; k < 2; k++) {
                tv[k] += A.val[k] + B.val[k] - P.val[k];
                te[k] += A.err[k] + B.err[k] - P.err[k];
            }
        }
        if (H.n > (size_t)1 << 22) break; /* memory guard */
    }
    tv[0] = tv[1] = te[0] = te[1] = 0;
    for (size_t i = 0; i < H.n; i++)
        for (int k = 0; k < 2; k++) {
            tv[k] += H.a[i].val[k];
            te[k] += H.a[i].err[k];
        }
    val[0] = tv[0];
    val[1] = tv[1];
    err[0] = te[0];
    err[1] = te[1];
    if (n_eval) *n_eval = ne;
    free(H.a);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Integrals                                                                   */
/* ------------------------------------------------------------------------- */
typedef struct {
    const or_electrode *sender, *receiver;
    double g[2];
} or_idata;

/* src/electrode.c:176-197: e^{-gamma r}/r between points on the two segment
 * axes, t[0] along the receiver and t[1] along the sender; r below DBL_EPSILON
 * is replaced by the receiver radius. */
void or_integrand_double(const double t[2], const or_electrode *sender,
                         const or_electrode *receiver, const double gamma[2],
                         double out[2])
{
    double pr[3], ps[3];
    for (int i = 0; i < 3; i++) {
        pr[i] = t[0] * (receiver->end_point[i] - receiver->start_point[i]) +
                receiver->start_point[i];
        ps[i] = t[1] * (sender->end_point[i] - sender->start_point[i]) +
                sender->start_point[i];
    }
    double r = or_vector_length(ps, pr);
    if (r < DBL_EPSILON) r = receiver->radius;
    double mag = exp(-gamma[0] * r);
    double ph = -gamma[1] * r;
    out[0] = mag * cos(ph) / r;
    out[1] = mag * sin(ph) / r;
}

static void or_integrand_cb(const double t[2], void *data, double out[2])
{
    const or_idata *p = (const or_idata *)data;
    or_integrand_double(t, p->sender, p->receiver, p->g, out);
}

/* ------------------------------------------------------------------------- */
/* Adaptive 1-D Gauss-Kronrod (7, 15) on [0, 1], one component: the route    */
/* hcubature takes for a 1-D integrand (src/electrode.c:288-290, fdim 1,      */
/* ERROR_INDIVIDUAL). Cubature (stevengj/cubature, version not pinned) uses   */
/* QUADPACK's QK15 rule with QUADPACK's error heuristic; the region with the  */
/* largest error is bisected until err <= abs_tol or err <= rel_tol |val|.    */
/* ------------------------------------------------------------------------- */
/* QUADPACK dqk15 abscissae / weights (Piessens et al. 1983); the Gauss       */
/* 7-point rule uses the odd-indexed abscissae.                               */
static const double gk_x[8] = {
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.0};
static const double gk_wk[8] = {
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
static const double gk_wg[4] = {
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

void or_gk15_rule(double x[8], double wk[8], double wg[4])
{
    memcpy(x, gk_x, sizeof(gk_x));
    memcpy(wk, gk_wk, sizeof(gk_wk));
    memcpy(wg, gk_wg, sizeof(gk_wg));
}

typedef double (*or_integrand1)(double t, void *data);

static void gk15(or_integrand1 f, void *data, double a, double b, double *val, double *err)
{
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    double fv1[7], fv2[7];
    const double fc = f(c, data);
    double rg = fc * gk_wg[3], rk = fc * gk_wk[7], ra = fabs(rk);
    for (int j = 0; j < 3; j++) {
        const int jj = 2 * j + 1;
        const double x = h * gk_x[jj];
        const double f1 = f(c - x, data), f2 = f(c + x, data);
        fv1[jj] = f1;
        fv2[jj] = f2;
        rg += gk_wg[j] * (f1 + f2);
        rk += gk_wk[jj] * (f1 + f2);
        ra += gk_wk[jj] * (fabs(f1) + fabs(f2));
    }
    for (int j = 0; j < 4; j++) {
        const int jj = 2 * j;
        const double x = h * gk_x[jj];
        const double f1 = f(c - x, data), f2 = f(c + x, data);
        fv1[jj] = f1;
        fv2[jj] = f2;
        rk += gk_wk[jj] * (f1 + f2);
        ra += gk_wk[jj] * (fabs(f1) + fabs(f2));
    }
    const double mean = 0.5 * rk;
    double asc = gk_wk[7] * fabs(fc - mean);
    for (int j = 0; j < 7; j++) asc += gk_wk[j] * (fabs(fv1[j] - mean) + fabs(fv2[j] - mean));
    double e = fabs((rk - rg) * h);
    asc *= fabs(h);
    ra *= fabs(h);
    if (asc != 0.0 && e != 0.0) {
        const double q = pow(200.0 * e / asc, 1.5);
        e = asc * (q < 1.0 ? q : 1.0);
    }
    if (ra > DBL_MIN / (50.0 * DBL_EPSILON)) {
        const double emin = 50.0 * DBL_EPSILON * ra;
        if (emin > e) e = emin;
    }
    *val = rk * h;
    *err = e;
}

int or_quad1(or_integrand1 f, void *data, size_t max_eval, double abs_tol, double rel_tol,
             double *val, double *err)
{
    size_t cap = 64, n = 1, ne = 15;
    double *ra = malloc(cap * sizeof(double)), *rb = malloc(cap * sizeof(double));
    double *rv = malloc(cap * sizeof(double)), *re = malloc(cap * sizeof(double));
    ra[0] = 0.0;
    rb[0] = 1.0;
    gk15(f, data, 0.0, 1.0, &rv[0], &re[0]);
    for (;;) {
        double tv = 0, te = 0;
        size_t worst = 0;
        for (size_t i = 0; i < n; i++) {
            tv += rv[i];
            te += re[i];
            if (re[i] > re[worst]) worst = i;
        }
        *val = tv;
        *err = te;
        if (te <= abs_tol || te <= rel_tol * fabs(tv)) break;
        if (max_eval && ne + 30 > max_eval) break;
        if (n >= ((size_t)1 << 20)) break; /* memory guard */
        if (n == cap) {
            cap *= 2;
            ra = realloc(ra, cap * sizeof(double));
            rb = realloc(rb, cap * sizeof(double));
            rv = realloc(rv, cap * sizeof(double));
            re = realloc(re, cap * sizeof(double));
        }
        const double a = ra[worst], b = rb[worst], mid = 0.5 * (a + b);
        rb[worst] = mid;
        gk15(f, data, a, mid, &rv[worst], &re[worst]);
        ra[n] = mid;
        rb[n] = b;
        gk15(f, data, mid, b, &rv[n], &re[n]);
        n++;
        ne += 30;
    }
    free(ra);
    free(rb);
    free(rv);
    free(re);
    return 0;
}

/* src/electrode.c:219-235: at the receiver-axis point t, the sender-axis
 * integral of 1/r in closed form, log((r1 + r2 + L_s) / (r1 + r2 - L_s)). */
static double or_log_nf(double t, void *data)
{
    const or_idata *p = (const or_idata *)data;
    double pr[3];
    for (int i = 0; i < 3; i++)
        pr[i] = t * (p->receiver->end_point[i] - p->receiver->start_point[i]) +
                p->receiver->start_point[i];
    const double r1 = or_vector_length(pr, p->sender->start_point);
    const double r2 = or_vector_length(pr, p->sender->end_point);
    const double L = p->sender->length;
    return log((r1 + r2 + L) / (r1 + r2 - L));
}

/* src/electrode.c:238-245: closed form of the double integral of 1/r between two
 * parallel filaments of length L one radius apart; independent of gamma. */
double or_self_integral(const or_electrode *e)
{
    double ls = e->length;
    double k1 = e->radius / ls;
    double k2 = sqrt(1.0 + k1 * k1);
    return 2.0 * ls * (log((k2 + 1.0) / k1) - k2 + k1);
}

int or_single_pair(const or_electrode *sender, const or_electrode *receiver,
                   const double gamma[2], size_t max_eval, double abs_tol, double rel_tol,
                   double result[2], double error[2]);

/* src/electrode.c:247-302 (INTG_NONE, INTG_DOUBLE, INTG_SINGLE and INTG_MHEM
 * branches; -10 otherwise).
 * INTG_MHEM: result[0] = L_r times the 1-D integral, result[1] = 0; error[0]
 * is not scaled (as in the reference). */
int or_integral(const or_electrode *sender, const or_electrode *receiver,
                const double gamma[2], size_t max_eval, double abs_tol, double rel_tol,
                int integration_type, double result[2], double error[2])
{
    result[0] = result[1] = error[0] = error[1] = 0.0;
    switch (integration_type) {
    case OR_INTG_NONE:
        result[0] = or_vector_length(sender->middle_point, receiver->middle_point);
        return 0;
    case OR_INTG_DOUBLE: {
        or_idata d = {sender, receiver, {gamma[0], gamma[1]}};
        int fail = or_hcubature2(or_integrand_cb, &d, max_eval, abs_tol, rel_tol, result,
                                 error, NULL);
        result[0] *= receiver->length * sender->length;
        result[1] *= receiver->length * sender->length;
        return fail;
    }
    case OR_INTG_SINGLE:
        return or_single_pair(sender, receiver, gamma, max_eval, abs_tol, rel_tol, result,
                              error);
    case OR_INTG_MHEM: {
        or_idata d = {sender, receiver, {0.0, 0.0}};
        int fail = or_quad1(or_log_nf, &d, max_eval, abs_tol, rel_tol, &result[0], &error[0]);
        result[0] *= receiver->length;
        return fail;
    }
    default:
        return -10;
    }
}

static inline void cmul(const double a[2], const double b[2], double o[2])
{
    double re = a[0] * b[0] - a[1] * b[1];
    double im = a[0] * b[1] + a[1] * b[0];
    o[0] = re;
    o[1] = im;
}

static inline void cdiv(const double a[2], const double b[2], double o[2])
{
    double den = b[0] * b[0] + b[1] * b[1];
    double re = (a[0] * b[0] + a[1] * b[1]) / den;
    double im = (a[1] * b[0] - a[0] * b[1]) / den;
    o[0] = re;
    o[1] = im;
}

/* src/electrode.c:304-377. Diagonal from the closed-form self integral; for
 * k > i: Z_T = one_4pik/(Li Lk) * I, Z_L = iwu_4pi * I * |d_i . d_k|/(Li Lk), or 0
 * when the un-normalised dot product is below DBL_EPSILON. */
/* One off-diagonal entry (k > i) of src/electrode.c:350-372: I = integral of
 * the pair, Z_T = one_4pik I / (L_i L_k), Z_L = iwu_4pi I |d_i . d_k| / (L_i L_k)
 * or 0 when the un-normalised |d_i . d_k| < DBL_EPSILON. Returns the integral's
 * failure code (entries untouched then). */
static int z_direct_entry(const or_electrode *e, size_t i, size_t k, const double gamma[2],
                          const double iwu_4pi[2], const double one_4pik[2], size_t max_eval,
                          double abs_tol, double rel_tol, int integration_type, double zl[2],
                          double zt[2])
{
    double ls = e[i].length, lr = e[k].length, cost = 0.0, res[2], err[2], tmp[2];
    for (int c = 0; c < 3; c++)
        cost += (e[i].end_point[c] - e[i].start_point[c]) *
                (e[k].end_point[c] - e[k].start_point[c]);
    cost = fabs(cost);
    int f = or_integral(&e[i], &e[k], gamma, max_eval, abs_tol, rel_tol, integration_type, res,
                        err);
    if (f) return f;
    tmp[0] = one_4pik[0] / (ls * lr);
    tmp[1] = one_4pik[1] / (ls * lr);
    cmul(tmp, res, zt);
    if (cost < DBL_EPSILON) {
        zl[0] = 0.0;
        zl[1] = 0.0;
    } else {
        cost = cost / (ls * lr);
        cmul(iwu_4pi, res, tmp);
        zl[0] = tmp[0] * cost;
        zl[1] = tmp[1] * cost;
    }
    return 0;
}

/* One image entry (k >= i) of src/electrode.c:403-428, added to zl, zt: I = the
 * integral between segment i and image k, Z_T += one_4pik I / (L_i L_k) and, if
 * |d_i . d'_k| > DBL_EPSILON, Z_L += iwu_4pi I |d_i . d'_k| / (L_i L_k), the
 * scalars already carrying ref_t and ref_l. */
static int z_image_entry(const or_electrode *e, const or_electrode *images, size_t i, size_t k,
                         const double gamma[2], const double iwu_4pi[2],
                         const double one_4pik[2], size_t max_eval, double abs_tol,
                         double rel_tol, int integration_type, double zl[2], double zt[2])
{
    double ls = e[i].length, lr = images[k].length, cost = 0.0, res[2], err[2], tmp[2];
    for (int c = 0; c < 3; c++)
        cost += (e[i].end_point[c] - e[i].start_point[c]) *
                (images[k].end_point[c] - images[k].start_point[c]);
    cost = fabs(cost);
    int f = or_integral(&e[i], &images[k], gamma, max_eval, abs_tol, rel_tol, integration_type,
                        res, err);
    if (f) return f;
    tmp[0] = one_4pik[0] / (ls * lr);
    tmp[1] = one_4pik[1] / (ls * lr);
    double add[2];
    cmul(tmp, res, add);
    zt[0] += add[0];
    zt[1] += add[1];
    if (cost > DBL_EPSILON) {
        cost = cost / (ls * lr);
        cmul(iwu_4pi, res, add);
        zl[0] += add[0] * cost;
        zl[1] += add[1] * cost;
    }
    return 0;
}

/* Scalars of src/electrode.c:313-319 (direct) and :391-399 (images, with the
 * reflection coefficients); unit scalars for INTG_MHEM and INTG_NONE. */
static void z_scalars(const double s[2], double mur, const double kappa[2],
                      const double ref_l[2], const double ref_t[2], int integration_type,
                      double iwu_4pi[2], double one_4pik[2])
{
    if (integration_type == OR_INTG_MHEM || integration_type == OR_INTG_NONE) {
        iwu_4pi[0] = 1.0;
        iwu_4pi[1] = 0.0;
        one_4pik[0] = 1.0;
        one_4pik[1] = 0.0;
        return;
    }
    double t[2] = {s[0] * mur * OR_MU0 / OR_FOUR_PI, s[1] * mur * OR_MU0 / OR_FOUR_PI};
    double fk[2] = {OR_FOUR_PI * kappa[0], OR_FOUR_PI * kappa[1]};
    if (ref_l) {
        cmul(ref_l, t, iwu_4pi);
        cdiv(ref_t, fk, one_4pik);
    } else {
        double one[2] = {1.0, 0.0};
        iwu_4pi[0] = t[0];
        iwu_4pi[1] = t[1];
        cdiv(one, fk, one_4pik);
    }
}

int or_calculate_impedances(double *zl, double *zt, const or_electrode *e, size_t m,
                            const double gamma[2], const double s[2], double mur,
                            const double kappa[2], size_t max_eval, double abs_tol,
                            double rel_tol, int integration_type, int nthreads)
{
    double iwu_4pi[2], one_4pik[2];
    z_scalars(s, mur, kappa, NULL, NULL, integration_type, iwu_4pi, one_4pik);
    int failure = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (long ii = 0; ii < (long)m; ii++) {
        size_t i = (size_t)ii;
        double ls = e[i].length, tmp[2];
        double intg[2] = {or_self_integral(&e[i]), 0.0};
        cmul(iwu_4pi, intg, zl + 2 * (i * m + i));
        tmp[0] = one_4pik[0] / (ls * ls);
        tmp[1] = one_4pik[1] / (ls * ls);
        cmul(tmp, intg, zt + 2 * (i * m + i));
        for (size_t k = i + 1; k < m; k++) {
            int f = z_direct_entry(e, i, k, gamma, iwu_4pi, one_4pik, max_eval, abs_tol,
                                   rel_tol, integration_type, zl + 2 * (i * m + k),
                                   zt + 2 * (i * m + k));
            if (f) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
                failure = f;
            }
        }
    }
    return failure;
}

/* src/electrode.c:379-431: images accumulated (+=) for k >= i, including k == i
 * (numerically integrated), scaled by the reflection coefficients. */
int or_impedances_images(double *zl, double *zt, const or_electrode *e,
                         const or_electrode *images, size_t m, const double gamma[2],
                         const double s[2], double mur, const double kappa[2],
                         const double ref_l_in[2], const double ref_t_in[2],
                         size_t max_eval, double abs_tol, double rel_tol,
                         int integration_type, int nthreads)
{
    double iwu_4pi[2], one_4pik[2];
    z_scalars(s, mur, kappa, ref_l_in, ref_t_in, integration_type, iwu_4pi, one_4pik);
    int failure = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (long ii = 0; ii < (long)m; ii++) {
        size_t i = (size_t)ii;
        for (size_t k = i; k < m; k++) {
            int f = z_image_entry(e, images, i, k, gamma, iwu_4pi, one_4pik, max_eval, abs_tol,
                                  rel_tol, integration_type, zl + 2 * (i * m + k),
                                  zt + 2 * (i * m + k));
            if (f) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
                failure = f;
            }
        }
    }
    return failure;
}

/* Lower-half entries (i <= k; a pair given as i > k is swapped) of Z_L and Z_T
 * after calculate_impedances then impedances_images, for a list of pairs:
 * the same per-entry arithmetic as the two full-matrix routines above (the
 * diagonal from the closed form), so each entry equals theirs bit for bit. */
int or_z_entries(const or_electrode *e, const or_electrode *images, size_t m, const long *pairs,
                 size_t npairs, const double gamma[2], const double s[2], double mur,
                 const double kappa[2], const double ref_l[2], const double ref_t[2],
                 double rel_tol, double *zl, double *zt, int nthreads)
{
    double iwu_d[2], fk_d[2], iwu_i[2], fk_i[2];
    z_scalars(s, mur, kappa, NULL, NULL, OR_INTG_DOUBLE, iwu_d, fk_d);
    z_scalars(s, mur, kappa, ref_l, ref_t, OR_INTG_DOUBLE, iwu_i, fk_i);
    int failure = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (long p = 0; p < (long)npairs; p++) {
        size_t i = (size_t)pairs[2 * p], k = (size_t)pairs[2 * p + 1];
        if (i > k) {
            size_t t = i;
            i = k;
            k = t;
        }
        double *l = zl + 2 * p, *t = zt + 2 * p;
        int f = 0;
        if (i >= m || k >= m) {
            f = -1;
        } else if (i == k) {
            double ls = e[i].length, tmp[2];
            double intg[2] = {or_self_integral(&e[i]), 0.0};
            cmul(iwu_d, intg, l);
            tmp[0] = fk_d[0] / (ls * ls);
            tmp[1] = fk_d[1] / (ls * ls);
            cmul(tmp, intg, t);
        } else {
            f = z_direct_entry(e, i, k, gamma, iwu_d, fk_d, 0, 0.0, rel_tol, OR_INTG_DOUBLE, l,
                               t);
        }
        if (!f)
            f = z_image_entry(e, images, i, k, gamma, iwu_i, fk_i, 0, 0.0, rel_tol,
                              OR_INTG_DOUBLE, l, t);
        if (f) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            failure = f;
        }
    }
    return failure;
}

/* src/linalg.c:12-86: node/segment matching within 1e-5 per coordinate; A = +1
 * (node is the start point) or -1 (end point), B = 0.5, written to W and to its
 * transpose position; returns 1 for an unconnected node, -sum(A) or
 * sum(B) - m when the incidence sums are wrong. Indices use size_t here (the
 * reference uses int and overflows above N = 46,340). */
long or_fill_incidence_imm(double *wg, const or_electrode *e, size_t m,
                           const double *nodes, size_t n)
{
    const size_t ld = 2 * m + n;
    const double tol = 1e-5;
    long suma = 0;
    double sumb = 0.0;
    for (size_t nd = 0; nd < n; nd++) {
        int none = 1;
        for (size_t s = 0; s < m; s++) {
            int cond = 0;
            if (or_equal_points_tol(e[s].start_point, nodes + 3 * nd, tol))
                cond = 1;
            else if (or_equal_points_tol(e[s].end_point, nodes + 3 * nd, tol))
                cond = 2;
            size_t a = (s + n) + ld * nd, b = (s + n + m) + ld * nd;
            size_t c = nd + ld * (s + n), d = nd + ld * (s + n + m);
            double av = cond == 1 ? 1.0 : (cond == 2 ? -1.0 : 0.0);
            double bv = cond ? 0.5 : 0.0;
            wg[2 * a] = av;
            wg[2 * a + 1] = 0.0;
            wg[2 * b] = bv;
            wg[2 * b + 1] = 0.0;
            wg[2 * c] = av;
            wg[2 * c + 1] = 0.0;
            wg[2 * d] = bv;
            wg[2 * d + 1] = 0.0;
            suma += (long)av;
            sumb += bv;
            if (cond) none = 0;
        }
        if (none) return 1;
    }
    if (suma != 0) return -suma;
    if (sumb != (double)m) return (long)(sumb - m);
    return 0;
}

/* src/linalg.c:142-199: the same matching in the admittance layout, a and b
 * m x n column-major (entry (s, node) at m * node + s); b is written before a
 * and the sums read a and b after the writes, so with a == b the array ends as
 * A and sum(B) = sum(A). Returns as or_fill_incidence_imm. */
long or_fill_incidence_adm(double *a, double *b, const or_electrode *e, size_t m,
                           const double *nodes, size_t n)
{
    const double tol = 1e-5;
    long suma = 0;
    double sumb = 0.0;
    for (size_t nd = 0; nd < n; nd++) {
        int none = 1;
        for (size_t s = 0; s < m; s++) {
            int cond = 0;
            if (or_equal_points_tol(e[s].start_point, nodes + 3 * nd, tol))
                cond = 1;
            else if (or_equal_points_tol(e[s].end_point, nodes + 3 * nd, tol))
                cond = 2;
            const size_t pos = m * nd + s;
            b[2 * pos] = cond ? 0.5 : 0.0;
            b[2 * pos + 1] = 0.0;
            a[2 * pos] = cond == 1 ? 1.0 : (cond == 2 ? -1.0 : 0.0);
            a[2 * pos + 1] = 0.0;
            if (cond) none = 0;
            suma += (long)a[2 * pos];
            sumb += b[2 * pos];
        }
        if (none) return 1;
    }
    if (suma != 0) return -suma;
    if (sumb != (double)m) return (long)(sumb - m);
    return 0;
}

/* src/linalg.c:88-110: -Z_L and -Z_T copied into their blocks for all (i, k),
 * zero cross blocks and zero Y_E block. */
int or_fill_impedance_imm(double *wg, const double *zl, const double *zt, size_t m,
                          size_t n)
{
    const size_t ld = n + 2 * m;
    for (size_t k = 0; k < m; k++)
        for (size_t i = 0; i < m; i++) {
            size_t p1 = (i + n) + ld * (k + n), p2 = (i + n) + ld * (k + n + m);
            size_t p3 = (i + n + m) + ld * (k + n), p4 = (i + n + m) + ld * (k + n + m);
            wg[2 * p1] = -zl[2 * (i + m * k)];
            wg[2 * p1 + 1] = -zl[2 * (i + m * k) + 1];
            wg[2 * p2] = wg[2 * p2 + 1] = 0.0;
            wg[2 * p3] = wg[2 * p3 + 1] = 0.0;
            wg[2 * p4] = -zt[2 * (i + m * k)];
            wg[2 * p4 + 1] = -zt[2 * (i + m * k) + 1];
        }
    for (size_t k = 0; k < n; k++)
        for (size_t i = 0; i < n; i++) wg[2 * (i + ld * k)] = wg[2 * (i + ld * k) + 1] = 0.0;
    return 0;
}

int or_pair_sample(const or_electrode *e, const or_electrode *images, const long *pairs,
                   size_t npairs, const double gamma[2], double rel_tol, double *out,
                   int nthreads)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (long p = 0; p < (long)npairs; p++) {
        double res[2], err[2];
        long i = pairs[2 * p], k = pairs[2 * p + 1];
        if (i == k) {
            out[4 * p] = or_self_integral(&e[i]);
            out[4 * p + 1] = 0.0;
        } else {
            or_integral(&e[i], &e[k], gamma, 0, 0.0, rel_tol, OR_INTG_DOUBLE, res, err);
            out[4 * p] = res[0];
            out[4 * p + 1] = res[1];
        }
        or_integral(&e[i], &images[k], gamma, 0, 0.0, rel_tol, OR_INTG_DOUBLE, res, err);
        out[4 * p + 2] = res[0];
        out[4 * p + 3] = res[1];
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Fields and voltages (src/electrode.c:177-216 integrand_single,             */
/* :433-668). INTG_SINGLE follows the reference's call, hcubature over the    */
/* unit square of an integrand of t[0] only (src/electrode.c:279-284), here   */
/* or_hcubature2; the field integrand is a 6-component 1-D hcubature with     */
/* ERROR_PAIRED (:538-546), here the multi-component form of or_quad1.        */
/* ------------------------------------------------------------------------- */
typedef void (*or_integrandn)(double t, void *data, double *out);

/* GK15 on [a, b] for fdim components: values and QUADPACK error per component */
static void gk15n(or_integrandn f, void *data, unsigned fdim, double a, double b, double *val,
                  double *err)
{
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    double fv[15][8];
    double xs[15];
    xs[7] = c;
    for (int j = 0; j < 7; j++) {
        xs[j] = c - h * gk_x[j];
        xs[14 - j] = c + h * gk_x[j];
    }
    for (int q = 0; q < 15; q++) f(xs[q], data, fv[q]);
    for (unsigned d = 0; d < fdim; d++) {
        double f1[7], f2[7];
        for (int j = 0; j < 7; j++) {
            f1[j] = fv[j][d];
            f2[j] = fv[14 - j][d];
        }
        const double fc = fv[7][d];
        double rg = fc * gk_wg[3], rk = fc * gk_wk[7], ra = fabs(rk);
        for (int j = 0; j < 3; j++) {
            const int jj = 2 * j + 1;
            rg += gk_wg[j] * (f1[jj] + f2[jj]);
        }
        for (int j = 0; j < 7; j++) {
            rk += gk_wk[j] * (f1[j] + f2[j]);
            ra += gk_wk[j] * (fabs(f1[j]) + fabs(f2[j]));
        }
        const double mean = 0.5 * rk;
        double asc = gk_wk[7] * fabs(fc - mean);
        for (int j = 0; j < 7; j++) asc += gk_wk[j] * (fabs(f1[j] - mean) + fabs(f2[j] - mean));
        double e = fabs((rk - rg) * h);
        asc *= fabs(h);
        ra *= fabs(h);
        if (asc != 0.0 && e != 0.0) {
            const double q = pow(200.0 * e / asc, 1.5);
            e = asc * (q < 1.0 ? q : 1.0);
        }
        if (ra > DBL_MIN / (50.0 * DBL_EPSILON)) {
            const double emin = 50.0 * DBL_EPSILON * ra;
            if (emin > e) e = emin;
        }
        val[d] = rk * h;
        err[d] = e;
    }
}

/* ERROR_PAIRED convergence: every (re, im) pair's error norm within tolerance */
static int paired_ok(unsigned fdim, const double *v, const double *e, double abs_tol,
                     double rel_tol)
{
    for (unsigned j = 0; j + 1 < fdim; j += 2) {
        const double en = hypot(e[j], e[j + 1]), vn = hypot(v[j], v[j + 1]);
        if (en > abs_tol && en > rel_tol * vn) return 0;
    }
    if (fdim % 2) {
        const double en = e[fdim - 1], vn = fabs(v[fdim - 1]);
        if (en > abs_tol && en > rel_tol * vn) return 0;
    }
    return 1;
}

/* adaptive 1-D integration of fdim <= 8 components over [0, 1]: the region
 * with the largest component error is bisected until paired_ok */
int or_quadn(or_integrandn f, void *data, unsigned fdim, size_t max_eval, double abs_tol,
             double rel_tol, double *val, double *err)
{
    size_t cap = 64, n = 1, ne = 15;
    double *ra = malloc(cap * sizeof(double)), *rb = malloc(cap * sizeof(double));
    double *rv = malloc(cap * 8 * sizeof(double)), *re = malloc(cap * 8 * sizeof(double));
    ra[0] = 0.0;
    rb[0] = 1.0;
    gk15n(f, data, fdim, 0.0, 1.0, rv, re);
    for (;;) {
        double tv[8] = {0}, te[8] = {0}, worst_e = -1.0;
        size_t worst = 0;
        for (size_t i = 0; i < n; i++) {
            double emax = 0.0;
            for (unsigned d = 0; d < fdim; d++) {
                tv[d] += rv[8 * i + d];
                te[d] += re[8 * i + d];
                if (re[8 * i + d] > emax) emax = re[8 * i + d];
            }
            if (emax > worst_e) {
                worst_e = emax;
                worst = i;
            }
        }
        memcpy(val, tv, fdim * sizeof(double));
        memcpy(err, te, fdim * sizeof(double));
        if (paired_ok(fdim, tv, te, abs_tol, rel_tol)) break;
        if (max_eval && ne + 30 > max_eval) break;
        if (n >= ((size_t)1 << 14)) break; /* region guard */
        if (n == cap) {
            cap *= 2;
            ra = realloc(ra, cap * sizeof(double));
            rb = realloc(rb, cap * sizeof(double));
            rv = realloc(rv, cap * 8 * sizeof(double));
            re = realloc(re, cap * 8 * sizeof(double));
        }
        const double a = ra[worst], b = rb[worst], mid = 0.5 * (a + b);
        rb[worst] = mid;
        gk15n(f, data, fdim, a, mid, rv + 8 * worst, re + 8 * worst);
        ra[n] = mid;
        rb[n] = b;
        gk15n(f, data, fdim, mid, b, rv + 8 * n, re + 8 * n);
        n++;
        ne += 30;
    }
    free(ra);
    free(rb);
    free(rv);
    free(re);
    return 0;
}

typedef struct {
    const or_electrode *e;
    const double *point;
    double g[2];
} or_fdata;

/* src/electrode.c:199-216: e^{-gamma r}/r from the sender axis point at t[0]
 * to the receiver's middle point (here: the field point), r >= DBL_EPSILON */
static void or_single_cb(const double t[2], void *data, double out[2])
{
    const or_fdata *p = (const or_fdata *)data;
    double ps[3];
    for (int i = 0; i < 3; i++)
        ps[i] = t[0] * (p->e->end_point[i] - p->e->start_point[i]) + p->e->start_point[i];
    double r = or_vector_length(ps, p->point);
    if (r < DBL_EPSILON) r = DBL_EPSILON;
    const double mag = exp(-p->g[0] * r) / r, ph = -p->g[1] * r;
    out[0] = mag * cos(ph);
    out[1] = mag * sin(ph);
}

/* INTG_SINGLE (src/electrode.c:279-284): L_s times the unit-square integral */
static void or_single(const or_electrode *e, const double *point, const double g[2],
                      size_t max_eval, double abs_tol, double rel_tol, double res[2])
{
    or_fdata d = {e, point, {g[0], g[1]}};
    double err[2];
    or_hcubature2(or_single_cb, &d, max_eval, abs_tol, rel_tol, res, err, NULL);
    res[0] *= e->length;
    res[1] *= e->length;
}

/* INTG_SINGLE of `integral` (src/electrode.c:279-284): the receiver's middle
 * point as the field point */
int or_single_pair(const or_electrode *sender, const or_electrode *receiver,
                   const double gamma[2], size_t max_eval, double abs_tol, double rel_tol,
                   double result[2], double error[2])
{
    or_fdata d = {sender, receiver->middle_point, {gamma[0], gamma[1]}};
    or_hcubature2(or_single_cb, &d, max_eval, abs_tol, rel_tol, result, error, NULL);
    result[0] *= sender->length;
    result[1] *= sender->length;
    return 0;
}

/* src/electrode.c:435-456: sum_m I_single(m) it[m] / L_m / (4 pi kappa) */
void or_electric_potential(const double *point, const or_electrode *e, size_t m,
                           const double *it, const double g[2], const double kappa[2],
                           size_t max_eval, double abs_tol, double rel_tol, double out[2])
{
    double pr = 0.0, pi = 0.0, res[2];
    for (size_t k = 0; k < m; k++) {
        or_single(&e[k], point, g, max_eval, abs_tol, rel_tol, res);
        const double ar = res[0] / e[k].length, ai = res[1] / e[k].length;
        pr += ar * it[2 * k] - ai * it[2 * k + 1];
        pi += ar * it[2 * k + 1] + ai * it[2 * k];
    }
    const double dr = OR_FOUR_PI * kappa[0], di = OR_FOUR_PI * kappa[1], den = dr * dr + di * di;
    out[0] = (pr * dr + pi * di) / den;
    out[1] = (pi * dr - pr * di) / den;
}

/* src/electrode.c:458-490: va = mu/4pi sum_m il[m] I_single(m)/L_m (end - start),
 * electrodes with |il| < DBL_EPSILON skipped */
void or_magnetic_potential(const double *point, const or_electrode *e, size_t m,
                           const double *il, const double g[2], double mur, size_t max_eval,
                           double abs_tol, double rel_tol, double va[6])
{
    for (int i = 0; i < 6; i++) va[i] = 0.0;
    double res[2];
    for (size_t k = 0; k < m; k++) {
        if (hypot(il[2 * k], il[2 * k + 1]) < DBL_EPSILON) continue;
        or_single(&e[k], point, g, max_eval, abs_tol, rel_tol, res);
        const double ar = res[0] / e[k].length, ai = res[1] / e[k].length;
        const double pr = ar * il[2 * k] - ai * il[2 * k + 1];
        const double pi = ar * il[2 * k + 1] + ai * il[2 * k];
        for (int i = 0; i < 3; i++) {
            const double dx = e[k].end_point[i] - e[k].start_point[i];
            va[2 * i] += pr * dx;
            va[2 * i + 1] += pi * dx;
        }
    }
    const double u = mur * OR_MU0 / OR_FOUR_PI;
    for (int i = 0; i < 6; i++) va[i] *= u;
}

/* src/electrode.c:492-515 */
static void or_efield_cb(double t, void *data, double *out)
{
    const or_fdata *p = (const or_fdata *)data;
    double dx[3], r2 = 0.0;
    for (int i = 0; i < 3; i++) {
        const double q = t * (p->e->end_point[i] - p->e->start_point[i]) + p->e->start_point[i];
        dx[i] = p->point[i] - q;
        r2 += dx[i] * dx[i];
    }
    const double r1 = sqrt(r2);
    if (r2 < DBL_EPSILON) r2 = DBL_EPSILON;
    /* (1 + gamma r) e^{-gamma r} / r^2 */
    const double er = exp(-p->g[0] * r1), ph = -p->g[1] * r1;
    const double xr = er * cos(ph), xi = er * sin(ph);
    const double ar = 1.0 + p->g[0] * r1, ai = p->g[1] * r1;
    const double fr = (ar * xr - ai * xi) / r2, fi = (ar * xi + ai * xr) / r2;
    for (int i = 0; i < 3; i++) {
        out[2 * i] = fr * dx[i] / r1;
        out[2 * i + 1] = fi * dx[i] / r1;
    }
}

/* src/electrode.c:517-558: with |s| > DBL_EPSILON ve = -s A (overwritten);
 * then ve += sum_m it[m]/(4 pi kappa) times the 6-component field integral */
int or_electric_field(const double *point, const or_electrode *e, size_t m, const double *il,
                      const double *it, const double g[2], const double s[2], double mur,
                      const double kappa[2], size_t max_eval, double abs_tol, double rel_tol,
                      double ve[6])
{
    if (hypot(s[0], s[1]) > DBL_EPSILON) {
        or_magnetic_potential(point, e, m, il, g, mur, max_eval, abs_tol, rel_tol, ve);
        for (int i = 0; i < 3; i++) {
            const double vr = ve[2 * i], vi = ve[2 * i + 1];
            ve[2 * i] = -(s[0] * vr - s[1] * vi);
            ve[2 * i + 1] = -(s[0] * vi + s[1] * vr);
        }
    }
    const double dr = OR_FOUR_PI * kappa[0], di = OR_FOUR_PI * kappa[1], den = dr * dr + di * di;
    double res[6], err[6];
    for (size_t k = 0; k < m; k++) {
        if (hypot(it[2 * k], it[2 * k + 1]) < DBL_EPSILON) continue;
        or_fdata d = {&e[k], point, {g[0], g[1]}};
        or_quadn(or_efield_cb, &d, 6, max_eval, abs_tol, rel_tol, res, err);
        /* it / (4 pi kappa) */
        const double cr = (it[2 * k] * dr + it[2 * k + 1] * di) / den;
        const double ci = (it[2 * k + 1] * dr - it[2 * k] * di) / den;
        for (int i = 0; i < 3; i++) {
            ve[2 * i] += res[2 * i] * cr - res[2 * i + 1] * ci;
            ve[2 * i + 1] += res[2 * i] * ci + res[2 * i + 1] * cr;
        }
    }
    return 0;
}

typedef struct {
    const double *p1, *p2;
    const or_electrode *e;
    size_t m;
    const double *il, *it;
    double g[2], s[2], kappa[2], mur;
    size_t max_eval;
    double abs_tol, rel_tol;
} or_vdata;

/* src/electrode.c:594-624 (ve starts at zero here) */
static void or_voltage_cb(double t, void *data, double *out)
{
    const or_vdata *p = (const or_vdata *)data;
    double pt[3], dx[3], len = 0.0;
    for (int i = 0; i < 3; i++) {
        dx[i] = p->p2[i] - p->p1[i];
        pt[i] = t * dx[i] + p->p1[i];
        len += dx[i] * dx[i];
    }
    len = sqrt(len);
    double ve[6] = {0, 0, 0, 0, 0, 0};
    or_electric_field(pt, p->e, p->m, p->il, p->it, p->g, p->s, p->mur, p->kappa, p->max_eval,
                      p->abs_tol, p->rel_tol, ve);
    out[0] = (ve[0] * dx[0] + ve[2] * dx[1] + ve[4] * dx[2]) / len;
    out[1] = (ve[1] * dx[0] + ve[3] * dx[1] + ve[5] * dx[2]) / len;
}

/* src/electrode.c:626-660: the unit-interval integral of E . (p2 - p1)/|p2 - p1|
 * along the straight path (not multiplied by the path length, as in the
 * reference) */
void or_voltage(const double *p1, const double *p2, const or_electrode *e, size_t m,
                const double *il, const double *it, const double g[2], const double s[2],
                double mur, const double kappa[2], size_t max_eval, double abs_tol,
                double rel_tol, double out[2])
{
    or_vdata d = {p1, p2, e, m, il, it, {g[0], g[1]}, {s[0], s[1]}, {kappa[0], kappa[1]}, mur,
                  max_eval, abs_tol, rel_tol};
    double err[2];
    or_quadn(or_voltage_cb, &d, 2, max_eval, abs_tol, rel_tol, out, err);
}
