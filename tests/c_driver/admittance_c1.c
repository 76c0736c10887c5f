/* Harmonic impedance of the 15 m conductor of harmonic_c1.c by the route of the
 * reference's examples (examples/harmonic_impedance.c:66-164): mHEM potentials
 * once (calculate_impedances / impedances_images with INTG_MHEM), then per
 * frequency the e^{-gamma rbar} scaling of the potentials, the nodal
 * admittance matrix and its solve. Mode "adm" uses fill_impedance_adm, mode
 * "adm2" calculate_yla_ytb + fill_impedance_adm2 and also prints the sum of
 * the leakage currents I_T = ytb U (1 by Kirchhoff's current law).
 *
 * Usage: admittance_c1 [NFREQ] [adm|adm2]   prints "f Re(Z_h) Im(Z_h) [Re Im sum(I_T)]". */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tags_hip.h"

static double dist(const double *a, const double *b)
{
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                (a[2] - b[2]) * (a[2] - b[2]));
}

int main(int argc, char **argv)
{
    const size_t nf = argc > 1 ? strtoul(argv[1], NULL, 10) : 5;
    const int adm2 = argc > 2 && strcmp(argv[2], "adm2") == 0;
    const size_t ne = 30, nn = ne + 1;
    const double rho = 100.0, epsr = 10.0, mur = 1.0;
    const double mu0 = 1.256637061435917e-6, eps0 = 8.854187817620e-12;
    const double two_pi = 6.283185307179586476925286766559;
    const double four_pi = 12.5663706143591729538505735331;
    const double start[3] = {0.0, 0.0, -0.5}, end[3] = {15.0, 0.0, -0.5};

    Electrode *e = malloc(ne * sizeof *e);
    Electrode *img = malloc(ne * sizeof *img);
    double *nodes = malloc(3 * nn * sizeof *nodes);
    _Complex double *pl = calloc(ne * ne, sizeof *pl), *pt = calloc(ne * ne, sizeof *pt);
    _Complex double *pli = calloc(ne * ne, sizeof *pli), *pti = calloc(ne * ne, sizeof *pti);
    _Complex double *zl = malloc(ne * ne * sizeof *zl), *zt = malloc(ne * ne * sizeof *zt);
    _Complex double *a = malloc(ne * nn * sizeof *a), *b = malloc(ne * nn * sizeof *b);
    _Complex double *yla = malloc(ne * nn * sizeof *yla), *ytb = malloc(ne * nn * sizeof *ytb);
    _Complex double *yn = malloc(nn * nn * sizeof *yn), *ie = malloc(nn * sizeof *ie);
    if (!e || !img || !nodes || !pl || !pt || !pli || !pti || !zl || !zt || !a || !b || !yla ||
        !ytb || !yn || !ie)
        return 1;
    if (segment_electrode(e, nodes, ne, start, end, 7e-3)) return 2;
    memcpy(img, e, ne * sizeof *e);
    for (size_t i = 0; i < ne; i++) {
        img[i].start_point[2] = -img[i].start_point[2];
        img[i].end_point[2] = -img[i].end_point[2];
        img[i].middle_point[2] = -img[i].middle_point[2];
    }
    if (fill_incidence_adm(a, b, e, ne, nodes, nn)) return 3;
    if (calculate_impedances(pl, pt, e, ne, 0.0, 0.0, 0.0, 0.0, 200, 1e-3, 1e-4, INTG_MHEM))
        return 4;
    if (impedances_images(pli, pti, e, img, ne, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 200, 1e-3, 1e-4,
                          INTG_MHEM))
        return 5;
    for (size_t q = 0; q < nf; q++) {
        const double f = nf > 1 ? pow(10.0, 2.0 + 5.0 * (double)q / (double)(nf - 1)) : 100.0;
        const _Complex double s = I * two_pi * f;
        const _Complex double kappa = 1.0 / rho + s * epsr * eps0;
        const _Complex double gamma = csqrt(s * mur * mu0 * kappa);
        const _Complex double iwu_4pi = s * mur * mu0 / four_pi, one_4pik = 1.0 / (four_pi * kappa);
        const _Complex double ref_t = (kappa - s * eps0) / (kappa + s * eps0), ref_l = 1.0;
        for (size_t m = 0; m < ne; m++)
            for (size_t k = m; k < ne; k++) {
                const size_t i = m * ne + k;
                const _Complex double ed = cexp(-gamma * dist(e[k].middle_point, e[m].middle_point));
                const _Complex double ei =
                    cexp(-gamma * dist(e[k].middle_point, img[m].middle_point));
                zl[i] = (ed * pl[i] + ref_l * ei * pli[i]) * iwu_4pi;
                zt[i] = (ed * pt[i] + ref_t * ei * pti[i]) * one_4pik;
            }
        int rc;
        if (adm2) {
            rc = calculate_yla_ytb(yla, ytb, zl, zt, a, b, ne, nn);
            if (!rc) rc = fill_impedance_adm2(yn, yla, ytb, a, b, ne, nn);
        } else {
            rc = fill_impedance_adm(yn, zl, zt, a, b, ne, nn);
        }
        if (rc) { fprintf(stderr, "admittance %d\n", rc); return 6; }
        memset(ie, 0, nn * sizeof *ie);
        ie[0] = 1.0;
        if ((rc = solve_admittance(yn, ie, nn))) { fprintf(stderr, "solve %d\n", rc); return 7; }
        printf("%.12e %.17e %.17e", f, creal(ie[0]), cimag(ie[0]));
        if (adm2) {
            _Complex double sum = 0.0;
            for (size_t s_ = 0; s_ < ne; s_++)
                for (size_t k = 0; k < nn; k++) sum += ytb[k * ne + s_] * ie[k];
            printf(" %.17e %.17e", creal(sum), cimag(sum));
        }
        printf("\n");
    }
    free(e); free(img); free(nodes); free(pl); free(pt); free(pli); free(pti); free(zl);
    free(zt); free(a); free(b); free(yla); free(ytb); free(yn); free(ie);
    return 0;
}
