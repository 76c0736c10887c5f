/* Device-resident sweep of the 20 x 20 m grid (5 + 5 conductors, 10 segments
 * per edge, radius 6 mm, depth 0.5 m; soil rho = 1000 Ohm m, eps_r = 15) from
 * plain C through libtags_hip.so: geometry with the reference's electrode_grid
 * (src/grid.c:30-91), then tags_sweep_create, tags_sweep_autotune (tile shapes
 * and schedule, cached in the file given) and tags_sweep_solve over NFREQ
 * log-spaced frequencies from 100 Hz to 10 MHz in the mixed form and in FP64,
 * unit current at corner node 0.
 *
 * Usage: sweep_c2 CACHE_FILE [NFREQ]   prints the tuning of each precision
 * ("tuning <prec> big small width from_cache") and then "f Re(Z_mixed)
 * Im(Z_mixed) Re(Z_fp64) Im(Z_fp64) precision_used" per line. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "tags_hip.h"

int main(int argc, char **argv)
{
    if (argc < 2) return 1;
    const char *cache = argv[1];
    const size_t nf = argc > 2 ? strtoul(argv[2], NULL, 10) : 40;
    const Grid g = {5, 5, 20.0, 20.0, 10, 10, 6e-3, -0.5};
    const double sigma = 1e-3, epsr = 15.0, eps0 = 8.854187817620e-12;
    const double two_pi = 6.283185307179586476925286766559;
    const int m = number_segments(g), n = number_nodes(g);
    Electrode *e = malloc((size_t)m * sizeof *e);
    double *nodes = malloc(3 * (size_t)n * sizeof *nodes);
    _Complex double *s = malloc(nf * sizeof *s), *kap = malloc(nf * sizeof *kap);
    _Complex double *zm = malloc(nf * sizeof *zm), *z64 = malloc(nf * sizeof *z64);
    tags_item_stats *st = malloc(nf * sizeof *st);
    if (!e || !nodes || !s || !kap || !zm || !z64 || !st) return 2;
    if (electrode_grid(g, e, nodes)) return 3;
    for (size_t i = 0; i < nf; i++) {
        const double f = pow(10.0, 2.0 + 5.0 * (double)i / (double)(nf > 1 ? nf - 1 : 1));
        s[i] = I * two_pi * f;
        kap[i] = sigma + s[i] * epsr * eps0;
    }
    tags_sweep *sw = NULL;
    int rc = tags_sweep_create(&sw, e, NULL, (size_t)m, nodes, (size_t)n);
    if (rc) return 10 + rc;
    const int precs[2] = {TAGS_PREC_MIXED, TAGS_PREC_FP64};
    for (int p = 0; p < 2; p++) {
        tags_tuning t;
        /* the probe: every 4th sample */
        const size_t np = (nf + 3) / 4;
        _Complex double *ps = malloc(np * sizeof *ps), *pk = malloc(np * sizeof *pk);
        for (size_t i = 0; i < np; i++) {
            ps[i] = s[4 * i];
            pk[i] = kap[4 * i];
        }
        rc = tags_sweep_autotune(sw, precs[p], np, ps, pk, 1.0, 0, cache, 0, &t);
        free(ps);
        free(pk);
        if (rc) return 20 + p;
        printf("tuning %d %d %d %d %d\n", precs[p], t.gemm_big, t.gemm_small, t.step_width,
               t.from_cache);
        rc = tags_sweep_solve(sw, nf, s, kap, 1.0, NULL, NULL, NULL, 0, precs[p],
                              p == 0 ? zm : z64, NULL, p == 0 ? st : NULL);
        if (rc) return 30 + p;
    }
    for (size_t i = 0; i < nf; i++)
        printf("%.17g %.17g %.17g %.17g %.17g %d\n", cimag(s[i]) / two_pi, creal(zm[i]),
               cimag(zm[i]), creal(z64[i]), cimag(z64[i]), st[i].precision_used);
    tags_sweep_destroy(sw);
    free(e);
    free(nodes);
    free(s);
    free(kap);
    free(zm);
    free(z64);
    free(st);
    return 0;
}
