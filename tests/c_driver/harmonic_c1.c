/* Harmonic impedance of one horizontal conductor (15 m, radius 7 mm, depth
 * 0.5 m, 30 segments; soil rho = 100 Ohm m, eps_r = 10) through the reference's
 * C API, written as a reference driver would be and linked against
 * libtags_hip.so. Per frequency it makes the calls of the reference caller
 * (examples/harmonic_impedance.c:114-164): calculate_impedances,
 * impedances_images, fill_incidence_imm, fill_impedance_imm, solve_immittance;
 * Z_h = U[0] for a unit current injected at node 0.
 *
 * Usage: harmonic_c1 [NFREQ]   prints "f Re(Z_h) Im(Z_h)" per line for
 * NFREQ log-spaced frequencies from 100 Hz to 10 MHz. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tags_hip.h"

int main(int argc, char **argv)
{
    const size_t nf = argc > 1 ? strtoul(argv[1], NULL, 10) : 5;
    const size_t m = 30, n = m + 1, N = n + 2 * m;
    const double rho = 100.0, epsr = 10.0, mur = 1.0;
    const double mu0 = 1.256637061435917e-6, eps0 = 8.854187817620e-12;
    const double two_pi = 6.283185307179586476925286766559;
    const double start[3] = {0.0, 0.0, -0.5}, end[3] = {15.0, 0.0, -0.5};

    Electrode *e = malloc(m * sizeof *e);
    Electrode *img = malloc(m * sizeof *img);
    double *nodes = malloc(3 * n * sizeof *nodes);
    _Complex double *zl = malloc(m * m * sizeof *zl);
    _Complex double *zt = malloc(m * m * sizeof *zt);
    _Complex double *wg = malloc(N * N * sizeof *wg);
    _Complex double *ie = malloc(N * sizeof *ie);
    if (!e || !img || !nodes || !zl || !zt || !wg || !ie) return 1;
    if (segment_electrode(e, nodes, m, start, end, 7e-3)) return 2;
    /* single image for the soil-air interface: z -> -z */
    memcpy(img, e, m * sizeof *e);
    for (size_t i = 0; i < m; i++) {
        img[i].start_point[2] = -img[i].start_point[2];
        img[i].end_point[2] = -img[i].end_point[2];
        img[i].middle_point[2] = -img[i].middle_point[2];
    }
    for (size_t k = 0; k < nf; k++) {
        const double f = nf > 1 ? pow(10.0, 2.0 + 5.0 * (double)k / (double)(nf - 1)) : 100.0;
        const _Complex double s = I * two_pi * f;
        const _Complex double kappa = 1.0 / rho + s * epsr * eps0;
        const _Complex double gamma = csqrt(s * mur * mu0 * kappa);
        const _Complex double ref_l = 1.0;
        const _Complex double ref_t = (kappa - s * eps0) / (kappa + s * eps0);
        memset(zl, 0, m * m * sizeof *zl);
        memset(zt, 0, m * m * sizeof *zt);
        int rc = calculate_impedances(zl, zt, e, m, gamma, s, mur, kappa, 200, 1e-3, 1e-4,
                                      INTG_DOUBLE);
        if (rc) { fprintf(stderr, "calculate_impedances %d\n", rc); return 3; }
        rc = impedances_images(zl, zt, e, img, m, gamma, s, mur, kappa, ref_l, ref_t, 200,
                               1e-3, 1e-4, INTG_DOUBLE);
        if (rc) { fprintf(stderr, "impedances_images %d\n", rc); return 4; }
        memset(wg, 0, N * N * sizeof *wg);
        rc = fill_incidence_imm(wg, e, m, nodes, n);
        if (rc) { fprintf(stderr, "fill_incidence_imm %d\n", rc); return 5; }
        fill_impedance_imm(wg, zl, zt, m, n);
        memset(ie, 0, N * sizeof *ie);
        ie[0] = 1.0;
        rc = solve_immittance(wg, ie, m, n);
        if (rc) { fprintf(stderr, "solve_immittance %d\n", rc); return 6; }
        printf("%.12e %.17e %.17e\n", f, creal(ie[0]), cimag(ie[0]));
    }
    free(e); free(img); free(nodes); free(zl); free(zt); free(wg); free(ie);
    return 0;
}
