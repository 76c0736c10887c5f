/* CPU oracle for the per-frequency HEM solve -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm (pedrohnv/transient-analysis-
 * grounding-systems, files cited per function in tags_oracle.c). It is the
 * checker for the HIP library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. The product never links it.
 *
 * Complex numbers cross this interface as interleaved (re, im) double pairs so
 * that ctypes can call every function without by-value _Complex arguments.
 */
#ifndef TAGS_ORACLE_H_
#define TAGS_ORACLE_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and size (88 bytes) as the reference `Electrode`
 * (src/electrode.h:34-45). */
typedef struct {
    double start_point[3];
    double end_point[3];
    double middle_point[3];
    double length;
    double radius;
} or_electrode;

/* Same fields as the reference `Grid` (src/grid.h:35-52). */
typedef struct {
    unsigned int vertices_x;
    unsigned int vertices_y;
    double length_x;
    double length_y;
    unsigned int edge_segments_x;
    unsigned int edge_segments_y;
    double radius;
    double depth;
} or_grid;

enum { OR_INTG_NONE = 1, OR_INTG_DOUBLE = 2, OR_INTG_SINGLE = 3, OR_INTG_MHEM = 4 };

extern const double OR_PI, OR_TWO_PI, OR_FOUR_PI, OR_MU0, OR_EPS0;

double or_vector_length(const double a[3], const double b[3]);
int or_linspace(double a, double b, size_t n, double *u);
int or_logspace(double a, double b, size_t n, double *u);

int or_populate_electrode(or_electrode *e, const double start[3], const double end[3],
                          double radius);
int or_segment_electrode(or_electrode *e, double *nodes, size_t num_segments,
                         const double *start, const double *end, double radius);
int or_number_segments(or_grid g);
int or_number_nodes(or_grid g);
int or_electrode_grid(or_grid g, or_electrode *e, double *nodes);

/* Adaptive 2-D cubature (own Genz-Malik degree-7/5 h-adaptive rule with paired
 * (complex) error norm). Integrates f over [0,1]^2. Returns 0 on success. */
typedef void (*or_integrand2)(const double t[2], void *data, double out[2]);
int or_hcubature2(or_integrand2 f, void *data, size_t max_eval, double abs_tol,
                  double rel_tol, double val[2], double err[2], size_t *n_eval);

void or_integrand_double(const double t[2], const or_electrode *sender,
                         const or_electrode *receiver, const double gamma[2],
                         double out[2]);
double or_self_integral(const or_electrode *e);
int or_integral(const or_electrode *sender, const or_electrode *receiver,
                const double gamma[2], size_t max_eval, double abs_tol, double rel_tol,
                int integration_type, double result[2], double error[2]);

/* zl, zt: m*m complex (2*m*m doubles), reference "lower half" convention:
 * entry (i, k>=i) at flat index i*m + k. `nthreads` > 1 spreads rows of the
 * pair loop over OpenMP threads (results do not depend on it). */
int or_calculate_impedances(double *zl, double *zt, const or_electrode *e, size_t m,
                            const double gamma[2], const double s[2], double mur,
                            const double kappa[2], size_t max_eval, double abs_tol,
                            double rel_tol, int integration_type, int nthreads);
int or_impedances_images(double *zl, double *zt, const or_electrode *e,
                         const or_electrode *images, size_t m, const double gamma[2],
                         const double s[2], double mur, const double kappa[2],
                         const double ref_l[2], const double ref_t[2], size_t max_eval,
                         double abs_tol, double rel_tol, int integration_type,
                         int nthreads);

/* W: (2m+n)^2 complex, column-major, caller-zeroed. */
void or_gk15_rule(double x[8], double wk[8], double wg[4]);
long or_fill_incidence_imm(double *wg, const or_electrode *e, size_t m,
                           const double *nodes, size_t n);
long or_fill_incidence_adm(double *a, double *b, const or_electrode *e, size_t m,
                           const double *nodes, size_t n);
int or_fill_impedance_imm(double *wg, const double *zl, const double *zt, size_t m,
                          size_t n);

/* Direct + image integrals of a fixed list of pairs, for the CPU timing sample:
 * pairs[2*p] = i, pairs[2*p+1] = k; out[4*p..] = (direct re, im, image re, im). */
/* lower-half Z_L, Z_T entries (direct + image) of a list of pairs */
int or_z_entries(const or_electrode *e, const or_electrode *images, size_t m, const long *pairs,
                 size_t npairs, const double gamma[2], const double s[2], double mur,
                 const double kappa[2], const double ref_l[2], const double ref_t[2],
                 double rel_tol, double *zl, double *zt, int nthreads);
int or_pair_sample(const or_electrode *e, const or_electrode *images, const long *pairs,
                   size_t npairs, const double gamma[2], double rel_tol, double *out,
                   int nthreads);

#ifdef __cplusplus
}
#endif

void or_electric_potential(const double *point, const or_electrode *e, size_t m,
                           const double *it, const double g[2], const double kappa[2],
                           size_t max_eval, double abs_tol, double rel_tol, double out[2]);
void or_magnetic_potential(const double *point, const or_electrode *e, size_t m,
                           const double *il, const double g[2], double mur, size_t max_eval,
                           double abs_tol, double rel_tol, double va[6]);
int or_electric_field(const double *point, const or_electrode *e, size_t m, const double *il,
                      const double *it, const double g[2], const double s[2], double mur,
                      const double kappa[2], size_t max_eval, double abs_tol, double rel_tol,
                      double ve[6]);
void or_voltage(const double *p1, const double *p2, const or_electrode *e, size_t m,
                const double *il, const double *it, const double g[2], const double s[2],
                double mur, const double kappa[2], size_t max_eval, double abs_tol,
                double rel_tol, double out[2]);

#endif
