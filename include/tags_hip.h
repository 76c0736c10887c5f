/* C ABI of the MI355X (gfx950) implementation of the per-frequency HEM solve.
 *
 * Part 1 re-exports the reference's own C API for the solve path with unchanged
 * names, argument order, argument meaning, return codes and ownership (every
 * array is allocated by the caller; pointers are host pointers). A driver
 * written against the reference headers links against libtags_hip.so instead
 * of libhphem.so. Citations are reference file:line of the prototype each
 * declaration replaces.
 *
 * Part 2 adds entry points with no reference counterpart: a device-resident
 * sweep over frequency samples that never moves the N x N matrix over PCIe.
 */
#ifndef TAGS_HIP_H_
#define TAGS_HIP_H_

#include <stddef.h>

#ifdef __cplusplus
#define TAGS_COMPLEX double _Complex
extern "C" {
#else
#include <complex.h>
#define TAGS_COMPLEX _Complex double
#endif

/* ---------------------------------------------------------------------------
 * Part 1: the reference API
 * ------------------------------------------------------------------------- */

/* src/electrode.h:18-31 */
enum Integration_type { INTG_NONE = 1, INTG_DOUBLE = 2, INTG_SINGLE = 3, INTG_MHEM = 4 };

/* src/electrode.h:34-45 (88 bytes, same field order) */
typedef struct {
    double start_point[3];
    double end_point[3];
    double middle_point[3];
    double length;
    double radius;
} Electrode;

/* src/grid.h:35-52 */
typedef struct {
    unsigned int vertices_x;
    unsigned int vertices_y;
    double length_x;
    double length_y;
    unsigned int edge_segments_x;
    unsigned int edge_segments_y;
    double radius;
    double depth;
} Grid;

/* src/electrode.h:104-106. Host code. 0 ok, 1 equal points, 2 radius <= 0. */
int populate_electrode(Electrode *electrode, const double start_point[3],
                       const double end_point[3], double radius);

/* src/electrode.h:153-155. Host code. */
int segment_electrode(Electrode *electrodes, double *nodes, size_t num_segments,
                      const double *start_point, const double *end_point, double radius);

/* src/grid.h:55-71. Host code. */
int number_segments(const Grid grid);
int number_nodes(const Grid grid);
int electrode_grid(const Grid grid, Electrode *electrodes, double *nodes);

/* src/electrode.h:226-227. Closed form, independent of gamma. */
TAGS_COMPLEX self_integral(const Electrode *sender);

/* src/electrode.h:247-250. INTG_DOUBLE runs the GPU pair quadrature (fixed-
 * order Gauss-Legendre with analytic treatment of touching segments) and
 * returns L_r L_s times the unit-square integral of exp(-gamma r)/r, the same
 * quantity as the reference; `error` receives the difference to a rule of
 * higher order. INTG_MHEM returns L_r times the receiver-axis integral of
 * log((r1 + r2 + L_s)/(r1 + r2 - L_s)) (src/electrode.c:219-235), which is the
 * double integral of 1/r: the same pair rule at gamma = 0, imaginary part 0,
 * error per unit receiver length as in the reference. INTG_NONE returns the
 * distance between middle points. INTG_SINGLE returns L_s times the
 * sender-axis integral of exp(-gamma r)/r to the receiver's middle point
 * (src/electrode.c:199-216; 16-point Gauss-Legendre panels graded towards the
 * closest axis point). max_eval, req_abs_error and req_rel_error are accepted;
 * the rules meet a relative error far below any req_rel_error >= 1e-10.
 * Returns 0 or -10 for an unknown type. calculate_impedances and
 * impedances_images take every type, INTG_MHEM and INTG_NONE with the
 * reference's unit scalars. */
int integral(const Electrode *sender, const Electrode *receiver, TAGS_COMPLEX gamma,
             size_t max_eval, double req_abs_error, double req_rel_error,
             int integration_type, double result[2], double error[2]);

/* src/electrode.h:276-281. Writes the lower half (flat index i*m + k, k >= i)
 * of Z_L and Z_T; entries k < i are left untouched; aliasing zl == zt behaves as
 * in the reference (zt wins on the diagonal, zl off it). */
int calculate_impedances(TAGS_COMPLEX *zl, TAGS_COMPLEX *zt, const Electrode *electrodes,
                         size_t num_electrodes, TAGS_COMPLEX gamma, TAGS_COMPLEX s,
                         double mur, TAGS_COMPLEX kappa, size_t max_eval,
                         double req_abs_error, double req_rel_error, int integration_type);

/* src/electrode.h:310-317. Adds the image contributions (k >= i, k == i
 * included) to the lower half of Z_L and Z_T. */
int impedances_images(TAGS_COMPLEX *zl, TAGS_COMPLEX *zt, const Electrode *electrodes,
                      const Electrode *images, size_t num_electrodes, TAGS_COMPLEX gamma,
                      TAGS_COMPLEX s, double mur, TAGS_COMPLEX kappa, TAGS_COMPLEX ref_l,
                      TAGS_COMPLEX ref_t, size_t max_eval, double req_abs_error,
                      double req_rel_error, int integration_type);

/* src/linalg.h:38-41. W column-major, ld = 2m + n. Returns 0, 1 if a node has
 * no segment, -sum(A) or sum(B) - m if the incidence sums are wrong. */
int fill_incidence_imm(TAGS_COMPLEX *wg, const Electrode *electrodes, size_t num_electrodes,
                       const double *nodes, size_t num_nodes);

/* src/linalg.h:56-59. Copies -Z_L, -Z_T into W (all i, k), zeroes the cross
 * blocks and the Y_E block; the incidence blocks are not touched. */
int fill_impedance_imm(TAGS_COMPLEX *wg, const TAGS_COMPLEX *zl, const TAGS_COMPLEX *zt,
                       size_t num_electrodes, size_t num_nodes);

/* src/linalg.h:79-81. Solves W x = ie for the symmetric W whose lower triangle
 * is stored; x replaces ie ([U; I_L; I_T]). Factorisation: complex LU in FP64
 * on the GPU with tournament pivoting (or partial pivoting, tags_set_panel);
 * the reference uses LAPACK zsysv. Returns 0, or i > 0 when U(i,i) is exactly
 * zero. W itself is not modified. */
int solve_immittance(TAGS_COMPLEX *wg, TAGS_COMPLEX *ie, size_t num_electrodes,
                     size_t num_nodes);

/* Nodal admittance formulation, Y_N U = I_E with Y_N = A^T Z_L^-1 A + B^T Z_T^-1 B.
 * a, b, yla, ytb: m x n column-major (entry (segment s, node k) at m k + s);
 * zl, zt: m x m, lower triangle read (the calculate_impedances convention);
 * yn: n x n column-major. */

/* src/linalg.h:103-106. A = +1 / -1 where the node is a segment's start / end
 * point (1e-5 per coordinate), B = 0.5 there, 0 elsewhere, for the nodes up to
 * the first unconnected one. Same return codes as fill_incidence_imm; with
 * a == b the array ends as A and B is reported wrong, as in the reference. */
int fill_incidence_adm(TAGS_COMPLEX *a, TAGS_COMPLEX *b, const Electrode *electrodes,
                       size_t num_electrodes, const double *nodes, size_t num_nodes);

/* src/linalg.h:123-126. Writes all of yn. As in the reference (zsytrf/zsytri
 * in place) the lower triangles of zl and zt are replaced by those of Z_L^-1
 * and Z_T^-1; their strict upper triangles are left as they were. Z^-1 and
 * Z^-1 A come from one complex LU with row exchanges per block on the GPU
 * (FP64, the identity and A carried as extra columns, then a blocked backward
 * solve), the products on the MFMA cores. Returns 0, or the first exactly
 * zero pivot (1-based). On a singular Z_L neither zl nor zt changes; on a
 * singular Z_T zl holds its inverse (as after the reference's zsytri) and zt
 * is unchanged (the reference leaves zsytrf's partial factor there). */
int fill_impedance_adm(TAGS_COMPLEX *yn, TAGS_COMPLEX *zl, TAGS_COMPLEX *zt, TAGS_COMPLEX *a,
                       TAGS_COMPLEX *b, size_t num_electrodes, size_t num_nodes);

/* src/linalg.h:145-149. yla = Z_L^-1 A, ytb = Z_T^-1 B; zl and zt as in
 * fill_impedance_adm. */
int calculate_yla_ytb(TAGS_COMPLEX *yla, TAGS_COMPLEX *ytb, TAGS_COMPLEX *zl, TAGS_COMPLEX *zt,
                      TAGS_COMPLEX *a, TAGS_COMPLEX *b, size_t num_electrodes,
                      size_t num_nodes);

/* src/linalg.h:169-173. yn = A^T yla + B^T ytb (all of yn); the inputs are
 * not modified. */
int fill_impedance_adm2(TAGS_COMPLEX *yn, TAGS_COMPLEX *yla, TAGS_COMPLEX *ytb, TAGS_COMPLEX *a,
                        TAGS_COMPLEX *b, size_t num_electrodes, size_t num_nodes);

/* src/linalg.h:192-193. Solves Y_N x = ic for the symmetric yn whose lower
 * triangle is stored; x replaces ic. FP64 complex LU on the GPU as in
 * solve_immittance (the reference uses zsysv); yn is not modified. Returns 0,
 * or i > 0 when U(i,i) is exactly zero. */
int solve_admittance(TAGS_COMPLEX *yn, TAGS_COMPLEX *ic, size_t num_nodes);

/* Potentials, fields and voltages of given segment currents (src/electrode.c:
 * 433-668). The line integrals over each segment run on the GPU with a
 * composite 16-point Gauss-Legendre rule graded towards the point's closest
 * axis parameter (relative error ~1e-14 for points off the segment axes);
 * max_eval, req_abs_error and req_rel_error are accepted. il, it: the m
 * longitudinal / transversal segment currents. */

/* src/electrode.h:333-337. sum_m it[m] int e^{-gamma r}/r dl / L_m / (4 pi kappa)
 * (INTG_SINGLE from each segment axis to `point`). On an error the error code
 * is returned as the (real) value, as in the reference. */
TAGS_COMPLEX electric_potential(const double *point, const Electrode *electrodes,
                                size_t num_electrodes, const TAGS_COMPLEX *it, TAGS_COMPLEX gamma,
                                TAGS_COMPLEX kappa, size_t max_eval, double req_abs_error,
                                double req_rel_error);

/* src/electrode.h:355-360. va (3 complex) = mu_r mu0 / 4 pi sum_m il[m]
 * int e^{-gamma r}/r dl / L_m (end - start)_m. Returns 0. */
int magnetic_potential(const double *point, const Electrode *electrodes, size_t num_electrodes,
                       const TAGS_COMPLEX *il, TAGS_COMPLEX gamma, double mur, size_t max_eval,
                       double req_abs_error, double req_rel_error, TAGS_COMPLEX *va);

/* src/electrode.h:398-404. E = -grad u - s A at `point`: for |s| > DBL_EPSILON
 * ve is overwritten with -s A and the conservative part added; otherwise the
 * conservative part is added to what ve holds (as in the reference). */
int electric_field(const double *point, const Electrode *electrodes, size_t num_electrodes,
                   const TAGS_COMPLEX *il, const TAGS_COMPLEX *it, TAGS_COMPLEX gamma,
                   TAGS_COMPLEX s, double mur, TAGS_COMPLEX kappa, size_t max_eval,
                   double req_abs_error, double req_rel_error, TAGS_COMPLEX *ve);

/* src/electrode.h:440-445. The unit-interval integral of E(p(t)) . u along
 * p(t) = point1 + t (point2 - point1), u the unit direction: the voltage
 * divided by the path length, as the reference computes it. Composite
 * Gauss-Legendre along the path, panels doubled until two results agree to
 * 1e-13. */
TAGS_COMPLEX voltage(const double *point1, const double *point2, const Electrode *electrodes,
                     size_t num_electrodes, const TAGS_COMPLEX *il, const TAGS_COMPLEX *it,
                     TAGS_COMPLEX gamma, TAGS_COMPLEX s, double mur, TAGS_COMPLEX kappa,
                     size_t max_eval, double req_abs_error, double req_rel_error);

/* Case files (host code). */

/* src/electrode.h:127-128 (src/electrode.c:45-80). Reads num_electrodes lines
 * "x0, y0, z0, x1, y1, z1, radius" and populates the electrodes. Returns 0, -10
 * when the file cannot be opened, fscanf's count - 7 at the first short line, or
 * populate_electrode's code for a bad record; messages as the reference. */
int electrodes_file(const char file_name[], Electrode *electrodes, size_t num_electrodes);

/* src/electrode.h:140-141 (src/electrode.c:82-101). Reads num_nodes lines
 * "x, y, z". Returns 0, -10, or fscanf's count - 3 at the first short line. */
int nodes_file(const char file_name[], double *nodes, size_t num_nodes);

/* Transient wrap: frequency-dependent soils and the numerical Laplace
 * transform pair (Gomez & Uribe 2009), used on the s_k = c + j k 2 pi / tmax
 * grid of laplace_trans. */

/* src/auxiliary.h:101-103. Alipio-Visacro soil: sigma(s) and eps_r(s) with
 * f = s / (2 pi j) (complex), principal-branch complex powers. Host code. */
int alipio_soil(TAGS_COMPLEX *sigma, TAGS_COMPLEX *epsr, double sigma0, TAGS_COMPLEX s, double h,
                double g, double eps_ratio);

/* src/auxiliary.h:119-121. Smith-Longmire soil (13 terms). Host code. */
int smith_longmire_soil(TAGS_COMPLEX *sigma, TAGS_COMPLEX *epsr, double sigma0, TAGS_COMPLEX s,
                        double erinf);

/* src/auxiliary.h:320-321. g[k] = sum_j f[j] dt e^{-c j dt} e^{-2 pi i jk/nt},
 * k < nt/2 + 1, with c = log(nt^2)/tmax, dt = tmax/(nt - 1), and
 * s[k] = c + j k 2 pi / tmax. The transform runs on the GPU (radix-2 Stockham
 * passes for a power-of-two nt, the direct sum otherwise); the reference uses
 * FFTW's r2c plan. Returns 0, or TAGS_ERR_ARG for nt < 2 or tmax <= 0. */
int laplace_trans(double *f, TAGS_COMPLEX *g, TAGS_COMPLEX *s, double tmax, size_t nt);

/* src/auxiliary.h:329-346 */
enum INLT_Filter {
    FILTER_NONE = 0,
    FILTER_BLACKMAN = 1,
    FILTER_HANNING = 2,
    FILTER_LANCZOS = 3,
    FILTER_RIESZ = 4
};

/* src/auxiliary.h:369-370. f[k] = e^{c k dt} / (nt dt) times the real inverse
 * DFT of the Hermitian extension of sigma(w_k) g[k] (the imaginary parts of the
 * DC and, for even nt, Nyquist terms do not enter, as in a c2r transform);
 * sigma is the data window `filter` with Omega = Im s[nt/2]. As in the
 * reference, FILTER_LANCZOS is 0/0 at w = 0, so its output is NaN. GPU
 * transform as in laplace_trans. */
int inv_laplace_trans(double *f, TAGS_COMPLEX *g, TAGS_COMPLEX *s, double tmax, size_t nt,
                      int filter);

/* ---------------------------------------------------------------------------
 * Part 2: device-resident sweep (no reference counterpart)
 * ------------------------------------------------------------------------- */

enum tags_precision {
    TAGS_PREC_FP64 = 0,   /* complex128 LU throughout (the plain form) */
    TAGS_PREC_MIXED = 1,  /* complex64 LU on MFMA + FP64 iterative refinement,
                             automatic FP64 fallback when refinement stalls */
};

typedef struct {
    int precision_used;   /* tags_precision actually used for this sample */
    int refine_iters;     /* refinement steps (mixed) */
    int info;             /* 0, or first zero pivot (1-based) */
    double backward_err;  /* ||b - W x||_inf / (||W||_inf ||x||_inf) in FP64 (-1 when
                             not computed: the FP64 form of a sweep without
                             tags_sweep_set_verify) */
    double refine_hist[8];/* backward error after each refinement step (mixed) */
    double forward_est;   /* mixed: estimated error left in x relative to ||x||_inf,
                             rho / (1 - rho) ||dx||_inf / ||x||_inf with dx the last
                             correction and rho the observed contraction per step
                             (||dx|| itself where rho >= 1/2); accepted at <= 1e-10;
                             -1 for the FP64 form */
} tags_item_stats;

typedef struct {
    double ms_fill;       /* pair quadrature + assembly */
    double ms_panel;      /* cooperative panel factorisation */
    double ms_gemm;       /* MFMA trailing / block updates */
    double ms_other;      /* diagonal-block inverses, gathers, conversions */
    double ms_solve;      /* triangular solves */
    double ms_residual;   /* FP64 residuals */
    double gemm_flops;    /* real flops issued by the GEMM kernel (4 real products
                             per complex multiply-add, 8 flops) */
    double trailing_ms;   /* time of trailing-update GEMM launches only */
    double trailing_flops;
    long trailing_launches;
} tags_timers;

typedef struct tags_sweep tags_sweep;

/* Uploads a geometry. `images` may be NULL (z -> -z images of `electrodes`).
 * Node/segment incidence is matched on the GPU with the reference tolerance
 * (1e-5 per coordinate, src/linalg.c:26). Returns the fill_incidence_imm code. */
int tags_sweep_create(tags_sweep **out, const Electrode *electrodes,
                      const Electrode *images, size_t num_electrodes, const double *nodes,
                      size_t num_nodes);

/* Solves `nfreq` frequency samples for a unit current injected at `inj_node`.
 * s, kappa: per-sample complex frequency and soil complex conductivity.
 * gamma, ref_l, ref_t may be NULL: then gamma = csqrt(s mu0 mur kappa),
 * ref_l = 1, ref_t = (kappa - s eps0)/(kappa + s eps0) (the reference caller's
 * formulas). u_inj[f] receives U[inj_node] (= Z_h for unit current).
 * x_out (nullable): nfreq * N complex, the full [U; I_L; I_T] per sample.
 * stats (nullable): nfreq entries. A sample whose factor has an exactly zero
 * pivot gets stats info > 0 and NaN in u_inj[f] and in its block of x_out. */
int tags_sweep_solve(tags_sweep *sw, size_t nfreq, const TAGS_COMPLEX *s,
                     const TAGS_COMPLEX *kappa, double mur, const TAGS_COMPLEX *gamma,
                     const TAGS_COMPLEX *ref_l, const TAGS_COMPLEX *ref_t, size_t inj_node,
                     int precision, TAGS_COMPLEX *u_inj, TAGS_COMPLEX *x_out,
                     tags_item_stats *stats);

/* Factorisation plan of the sweep (default TAGS_PLAN_HYBRID, the fastest
 * measured: ZL_FIRST's FP64 factor, the nodal complex64 factor from the switch
 * frequency up). Every plan
 * but the mixed factor of GROUNDED pivots with the process's panel algorithm
 * (tags_set_panel); the structured
 * ones skip rows and columns that are exactly zero for the columns being
 * factored, which changes no arithmetic on the non-zero entries.
 * TAGS_PLAN_DENSE       unknowns in the reference order [U; I_L; I_T], all N
 *                       rows and columns;
 * TAGS_PLAN_STRUCTURED  reference order; the node rows are zero in the node
 *                       columns (Y_E = 0), so the first n columns pivot and
 *                       update over the 2m impedance rows only;
 * TAGS_PLAN_ZL_FIRST    unknowns stored as [I_L; I_T; U]; the Z_T rows are
 *                       zero in the Z_L columns and are skipped there (fewer
 *                       flops; its complex64 factor is less accurate at the
 *                       lowest frequencies);
 * TAGS_PLAN_BLOCK_PIVOT stored and factored as ZL_FIRST in FP64; the mixed
 *                       form's complex64 factor lets the Z_L columns pivot among
 *                       Z_L rows only, which also skips the Z_T columns (fewest
 *                       flops, weakest pivoting: refinement answers for it);
 * TAGS_PLAN_GROUNDED    stored as ZL_FIRST and factored as ZL_FIRST in FP64.
 *                       The mixed form's complex64 factor is of the grounded
 *                       matrix T^T W T (node voltages replaced by the first
 *                       node's voltage and the differences to it), factored
 *                       without row exchanges: Z_L columns, Z_T columns, then
 *                       the node block, each pivoting on its own diagonal.
 *                       FP64 refinement runs on W itself, and a complex64
 *                       breakdown or stalled refinement falls back to the
 *                       pivoted FP64 factor;
 * TAGS_PLAN_GROUNDED_PIVOTED as GROUNDED, but the complex64 factor of T^T W T
 *                       pivots like ZL_FIRST;
 * TAGS_PLAN_NODAL       stored and factored as ZL_FIRST in FP64. The mixed
 *                       form's complex64 factor is a block elimination of the
 *                       currents: Z_L and Z_T are factored with row exchanges
 *                       among their own rows, X = Z^-1 A for each, and the
 *                       node block is set to its Schur complement
 *                       -(A^T X_L + B^T X_T) from the sparse incidence rows,
 *                       then factored (about half the flops of ZL_FIRST: the
 *                       incidence blocks are not treated as dense). It needs
 *                       more refinement steps toward low frequencies, where a
 *                       stalled refinement falls back to the FP64 factor;
 * TAGS_PLAN_HYBRID      stored and factored as ZL_FIRST in FP64. In a sweep's
 *                       mixed form, each sample takes NODAL's complex64 factor,
 *                       without row exchanges (Z_L, Z_T and S are complex
 *                       symmetric with a definite real or imaginary part), at
 *                       and above the switch frequency
 *                       (tags_sweep_set_nodal_above) and ZL_FIRST's below it;
 *                       tags_immittance_solve, which has no frequency, treats
 *                       it as ZL_FIRST. */
enum tags_plan {
    TAGS_PLAN_DENSE = 0,
    TAGS_PLAN_STRUCTURED = 1,
    TAGS_PLAN_ZL_FIRST = 2,
    TAGS_PLAN_BLOCK_PIVOT = 3,
    TAGS_PLAN_GROUNDED = 4,
    TAGS_PLAN_GROUNDED_PIVOTED = 5,
    TAGS_PLAN_NODAL = 6,
    TAGS_PLAN_HYBRID = 7
};
int tags_sweep_set_plan(tags_sweep *sw, int plan);

/* Batched fields, no reference counterpart as a function: for every point p
 * and right-hand side k (currents il[k m ..], it[k m ..]; either may be NULL for
 * zero), pot[p nrhs + k] (electric_potential), va[3 (p nrhs + k) ..]
 * (magnetic_potential) and ve[3 (p nrhs + k) ..] = -grad u - s A (the
 * electric_field of a zeroed ve, -s A only when |s| > DBL_EPSILON); outputs may
 * be NULL. */
int tags_fields_batch(const double *points, size_t num_points, const Electrode *electrodes,
                      size_t num_electrodes, const TAGS_COMPLEX *il, const TAGS_COMPLEX *it,
                      size_t nrhs, TAGS_COMPLEX gamma, TAGS_COMPLEX s, double mur,
                      TAGS_COMPLEX kappa, TAGS_COMPLEX *pot, TAGS_COMPLEX *va, TAGS_COMPLEX *ve);

/* The mHEM ground potential of the examples (examples/visacro57emc01.c:406-432):
 * out[p nrhs + k] = sum_m it[k m + m'] / (4 pi kappa) e^{-gamma rbar}
 * log((r1 + r2 + L)/(r1 + r2 - L)) / L, rbar the distance to the segment's
 * middle point, r1, r2 to its ends (the caller applies the (1 + ref_t) image
 * factor to it, as the examples do). */
int tags_mhem_gpd(const double *points, size_t num_points, const Electrode *electrodes,
                  size_t num_electrodes, const TAGS_COMPLEX *it, size_t nrhs, TAGS_COMPLEX gamma,
                  TAGS_COMPLEX kappa, TAGS_COMPLEX *out);

/* The examples' case-file helpers (examples/harmonic_impedance.c:178-257), no
 * library counterpart in the reference: the line count the examples size
 * their arrays with (newlines + 1; -1 if the file cannot be opened), a list of
 * n frequencies (one "%lf" each; returns the number not read, or -10), and a
 * result file with one "%e%+eim" line per value (0, or -10). */
long tags_count_lines(const char file_name[]);
int tags_frequencies_file(const char file_name[], double *freq, size_t n);
int tags_write_complex_file(const char file_name[], const TAGS_COMPLEX *z, size_t n);

/* Batched laplace_trans / inv_laplace_trans over `howmany` signals (the
 * examples' fftw_plan_many layouts, e.g. examples/sunjerga173powsys.c:56-90):
 * signal i at f[i nt ..] and g[i (nt/2 + 1) ..]; s as in the single calls.
 * No reference counterpart as a function. */
int tags_laplace_batch(const double *f, TAGS_COMPLEX *g, TAGS_COMPLEX *s, double tmax, size_t nt,
                       size_t howmany);
int tags_inv_laplace_batch(double *f, const TAGS_COMPLEX *g, const TAGS_COMPLEX *s, double tmax,
                           size_t nt, int filter, size_t howmany);

/* Impedance formulation of a sweep's samples.
 * TAGS_FORM_HEM   (default) every sample integrates e^{-gamma r}/r over every
 *                 segment pair (calculate_impedances + impedances_images with
 *                 INTG_DOUBLE, src/electrode.c:305-431).
 * TAGS_FORM_MHEM  the reference examples' modified HEM
 *                 (examples/harmonic_impedance.c:78-85, :125-140): the
 *                 gamma-free INTG_MHEM potentials are integrated once per
 *                 geometry (4 m^2 doubles of device memory), and each sample
 *                 scales them by exp(-gamma rbar) (rbar: middle-point distance
 *                 of the pair, of segment and image for the image part), s mu /
 *                 4 pi, 1 / (4 pi kappa) and the reflection coefficients. The
 *                 examples then solve the nodal admittance system; the sweep
 *                 solves the immittance system with the same Z_L, Z_T, whose
 *                 node potentials are the same (Y_N is its Schur complement).
 * Returns 0, or TAGS_ERR_ARG for another value. */
enum tags_formulation { TAGS_FORM_HEM = 0, TAGS_FORM_MHEM = 1 };
int tags_sweep_set_formulation(tags_sweep *sw, int formulation);

/* Refinement of nodal complex64 factors (TAGS_PLAN_HYBRID and TAGS_PLAN_NODAL,
 * mixed form): a sample at |s| / 2 pi < below_hz (default 2e4 Hz) refines by
 * FGMRES in FP64 right-preconditioned by the complex64 factor, `steps`
 * (1 .. 32, default 6) Arnoldi steps per refinement step (GMRES-based
 * refinement), where the classic correction from the factor alone converges
 * slowly or not at all (the node block's Schur complement is dominated by
 * A^T Z_L^-1 A, whose complex64 rounding swamps the grounding terms at low
 * frequency); a higher one refines classically. No counterpart in the
 * reference. Returns 0, or TAGS_ERR_ARG. */
int tags_sweep_set_gmres(tags_sweep *sw, double below_hz, int steps);

/* Switch frequency of TAGS_PLAN_HYBRID (default 0 Hz: the nodal factor at
 * every frequency, refined as tags_sweep_set_gmres sets): in the mixed form a
 * sample at |s| / 2 pi >= hz takes the nodal plan's complex64 factor, a lower
 * one ZL_FIRST's. A nodal sample that falls back to FP64 or needs more than 10
 * refinement steps raises the switch to 1.5 times its frequency for the rest
 * of that tags_sweep_solve call (each call starts from the configured switch).
 * No counterpart in the reference, which factors every sample alike. Returns 0,
 * or TAGS_ERR_ARG for hz < 0. */
int tags_sweep_set_nodal_above(tags_sweep *sw, double hz);

/* Frequency samples per group (0 .. 4096; 0, the default, chooses from N: one
 * sample from N = 8192 up, else as many as keep a group's matrices within
 * 16 GiB, at most 64). A group's samples are assembled, factored, solved and
 * refined together: every kernel launch covers all of them (one work-group
 * grid per sample, side by side), so that mid-size systems fill the device.
 * Consecutive samples form a group; the hybrid plan starts a new group where
 * its choice of complex64 factor changes. Each sample's arithmetic is its own:
 * results agree with one-sample groups to rounding (the GEMM tile shape is
 * chosen by the size of the whole launch), and are independent of the
 * samples in flight. Grounded plans in the mixed form and the partial-
 * pivoting panel take one sample at a time. No counterpart in the reference. */
int tags_sweep_set_batch(tags_sweep *sw, int samples);

/* Tile shapes and schedule of one sweep and precision (no counterpart in the
 * reference). Fields < 0 (0 for step_width and inflight) leave the
 * process-wide settings (tags_gemm_select, tags_set_step_width,
 * tags_set_update_chunks, tags_set_solve_rows) or the sweep's own
 * (tags_sweep_set_inflight, tags_sweep_set_batch, tags_sweep_set_nodal_above)
 * as they are. */
typedef struct {
    int gemm_big;          /* MFMA tile variant of large products (tags_gemm_variants) */
    int gemm_small;        /* of small products */
    int step_width;        /* columns per factorisation step: 128 or 256 */
    int update_chunks;     /* trailing-update launches beside the look-ahead panel */
    int solve_rows;        /* triangular solves: 0, 1 or 2 (tags_set_solve_rows) */
    int inflight;          /* groups of samples in flight, 1..4 */
    int batch;             /* samples per group (0: chosen from N) */
    double nodal_above;    /* hybrid plan's switch (Hz), mixed form only */
    int gmres_steps;       /* Arnoldi steps per FGMRES refinement step, mixed form
                              only (1..32; 0 leaves the sweep's) */
    double ms_per_sample;  /* out: time per probe sample with these values */
    int from_cache;        /* out: 1 read from the cache file, 0 measured */
    int evaluations;       /* out: timed probe solves */
} tags_tuning;

/* Autotuner: picks the tile shapes and the schedule of `precision` for this
 * sweep and installs them (tags_sweep_set_tuning). First every MFMA tile
 * variant is timed on the sweep's own trailing-update shapes (the largest
 * update of the plan's widest phase, and a 1024 x 1024 tail product) at both
 * step widths; then, by coordinate descent from the current settings, step
 * width (with its fastest variants), update launches, solve rows, samples in
 * flight, samples per group (half and a quarter of the automatic size below
 * order 8192; 2, 3 and 4 from 8192 up, as far as the groups in flight fit in
 * 80 % of the device's memory) and, for the mixed form of
 * the hybrid plan, the nodal switch (0 or 20 kHz) and the Arnoldi steps per
 * FGMRES refinement step (6, 8, 10) are each kept when a
 * timed solve of the `nprobe` probe samples (s, kappa, mur, inj_node as in
 * tags_sweep_solve; one untimed solve first for the allocations; for a fair
 * comparison of group sizes nprobe should be a few times the samples in
 * flight times the largest group tried) runs more
 * than 3 % faster. With cache_path (nullable) an entry for the same device,
 * formulation, plan, precision, N and m is used instead when present (unless
 * force != 0), and the result is written there (one line per key). `out`
 * (nullable) receives the values. Returns 0 or an error code. */
int tags_sweep_autotune(tags_sweep *sw, int precision, size_t nprobe, const TAGS_COMPLEX *s,
                        const TAGS_COMPLEX *kappa, double mur, size_t inj_node,
                        const char *cache_path, int force, tags_tuning *out);
/* Installs (set) or reads (get) the tuning of `precision`; set with NULL
 * removes it. Returns 0, or TAGS_ERR_ARG for values out of range. */
int tags_sweep_set_tuning(tags_sweep *sw, int precision, const tags_tuning *t);
int tags_sweep_get_tuning(const tags_sweep *sw, int precision, tags_tuning *t);

/* FP64 form only: with on != 0, W is assembled again after each sample's
 * factorisation (which replaced it) and the backward error
 * ||e_inj - W x||_inf / (||W||_inf ||x||_inf) goes to the item statistics
 * (one more fill and one FP64 residual per sample; off by default, when
 * backward_err is -1). No counterpart in the reference. */
int tags_sweep_set_verify(tags_sweep *sw, int on);

/* Groups of frequency samples a sweep keeps in flight (1 .. 4, default 4),
 * each with its own HIP stream, W and factor work space (about 11 GB for a
 * one-sample group of the 60 x 60 m grid in the mixed form). The calling
 * thread queues a group on every slot, then finishes them in issue order,
 * queuing the next group on each slot as it frees up, so that one group's
 * trailing updates keep the GPU busy while another waits on its panel chain. Results do not depend on the count, except
 * where TAGS_PLAN_HYBRID raises its switch during a call: samples already in
 * flight keep the factor they were issued with (both answers meet the same
 * acceptance rule). */
int tags_sweep_set_inflight(tags_sweep *sw, int samples);

/* Assembles W for one frequency sample on the GPU exactly as tags_sweep_solve
 * does (pair integrals, images, incidence) and copies it to `wg` (N x N,
 * column-major, reference unknown order [U; I_L; I_T], both triangles). */
int tags_sweep_assemble(tags_sweep *sw, TAGS_COMPLEX s, TAGS_COMPLEX kappa, double mur,
                        TAGS_COMPLEX *wg);

/* As tags_sweep_assemble, but copies only the `count` entries W(rows[i],
 * cols[i]) (reference unknown order, 0-based) to out[i]: spot checks of large
 * systems without moving N^2 values. Returns TAGS_ERR_ARG for an index out of
 * range. */
int tags_sweep_assemble_entries(tags_sweep *sw, TAGS_COMPLEX s, TAGS_COMPLEX kappa, double mur,
                                size_t count, const long *rows, const long *cols,
                                TAGS_COMPLEX *out);

/* Accumulates HIP-event timings of every kernel family on the sweep's stream. */
int tags_sweep_set_profiling(tags_sweep *sw, int on);
int tags_sweep_get_timers(tags_sweep *sw, tags_timers *t);
void tags_sweep_destroy(tags_sweep *sw);

/* ---------------------------------------------------------------------------
 * Many grounding cases at once (no reference counterpart: the reference's
 * drivers run one case per process, its frequencies in an OpenMP loop,
 * examples/harmonic_impedance.c:87-176). Each case is a geometry with its own
 * frequency samples; every (case, frequency) sample of every case is solved
 * for a unit current at the case's injection node, with the case's
 * reference-caller scalars (gamma = csqrt(s mu0 mur kappa), ref_l = 1,
 * ref_t = (kappa - s eps0)/(kappa + s eps0)).
 * ------------------------------------------------------------------------- */
typedef struct {
    const Electrode *electrodes;  /* m segments */
    const Electrode *images;      /* m images, or NULL for z -> -z images */
    size_t num_electrodes;
    const double *nodes;          /* n nodes, 3 doubles each */
    size_t num_nodes;
    size_t inj_node;              /* unit current injected here */
    size_t nfreq;                 /* frequency samples of the case */
    const TAGS_COMPLEX *s;        /* nfreq complex frequencies */
    const TAGS_COMPLEX *kappa;    /* nfreq soil complex conductivities */
    double mur;
    TAGS_COMPLEX *u_inj;          /* out: nfreq values U[inj_node] (Z_h) */
    tags_item_stats *stats;       /* out, nullable: nfreq entries */
} tags_case;

typedef struct tags_cases tags_cases;

/* Uploads every case's geometry (one tags_sweep each; the incidence is matched
 * on the GPU). The descriptors' arrays are read again by tags_cases_solve and
 * must stay valid until tags_cases_destroy. Returns 0, TAGS_ERR_ARG, or the
 * first case's tags_sweep_create code. */
int tags_cases_create(tags_cases **out, const tags_case *cases, size_t ncases);

/* Solves every sample of every case in `precision` (tags_precision). The
 * samples are grouped per case (tags_sweep_set_batch's automatic size), the
 * groups of all cases queued largest first (per-sample cost
 * (8/3) N^3 + 6400 m^2) and taken by the lanes (tags_cases_set_lanes) as each
 * finishes its last group. Results as tags_sweep_solve's for each case. */
int tags_cases_solve(tags_cases *cs, int precision);

/* Factorisation plan of every case (tags_plan, default TAGS_PLAN_HYBRID). */
int tags_cases_set_plan(tags_cases *cs, int plan);

/* Lanes (1 .. 4, default 3): groups in flight, each with its own stream, W and
 * factor work space, sized for the largest group it has taken. */
int tags_cases_set_lanes(tags_cases *cs, int lanes);

/* Autotuner over the cases (no counterpart in the reference): the cases are
 * split into size classes by N (below 256, then doubling up to 8192, then
 * above), the largest case of each class is tuned as tags_sweep_autotune does
 * (its own samples as probe, up to 64 spread over them; cache_path and force
 * as there, the entry keyed by that case), and its tile shapes, group size,
 * nodal switch and FGMRES steps are installed on every case of the class
 * (samples in flight do not apply: the lanes take groups as they free up).
 * `classes` (nullable) receives up to 16 entries, one per class tuned, and
 * *nclasses their number. Returns 0 or an error code. */
int tags_cases_autotune(tags_cases *cs, int precision, const char *cache_path, int force,
                        tags_tuning *classes, int *nclasses);

/* HIP-event timings of every lane's kernels, as tags_sweep_get_timers. */
int tags_cases_set_profiling(tags_cases *cs, int on);
int tags_cases_get_timers(tags_cases *cs, tags_timers *t);
void tags_cases_destroy(tags_cases *cs);

/* calculate_impedances (src/electrode.h:276-281) for callers that cannot
 * allocate device memory (work captured in HIP graphs, tight serving loops); no
 * counterpart in the reference. Same arguments and results (the lower half of
 * zl and zt, aliasing rule included), but zl, zt and electrodes are DEVICE
 * pointers and the work is queued on `stream` (a hipStream_t; NULL: the null
 * stream) without waiting. The library allocates nothing: the segment
 * geometry is staged in a fixed device array of the library (one per device,
 * 4 MiB, set up at the first call on that device), so num_electrodes is at
 * most 65536 (TAGS_ERR_TOO_LARGE above, before any device work). Calls that
 * overlap on one device race on that array: queue them on one stream. */
int calculate_impedances_noalloc(TAGS_COMPLEX *zl, TAGS_COMPLEX *zt, const Electrode *electrodes,
                                 size_t num_electrodes, TAGS_COMPLEX gamma, TAGS_COMPLEX s,
                                 double mur, TAGS_COMPLEX kappa, size_t max_eval,
                                 double req_abs_error, double req_rel_error,
                                 int integration_type, void *stream);

/* Dense solve of a device-resident or host complex matrix, for tests and
 * library use. `a` is row-major n x n (for the symmetric W the same as column-
 * major), `b` is n; on return `b` holds x. precision: tags_precision.
 * Host pointers. Returns 0, info > 0, or a negative error code. */
int tags_dense_solve(const TAGS_COMPLEX *a, TAGS_COMPLEX *b, size_t n, int precision,
                     tags_item_stats *stats);

/* As tags_dense_solve for an immittance matrix stored row-major in the
 * storage order of the given tags_plan: [U; I_L; I_T] for DENSE and
 * STRUCTURED, [I_L; I_T; U] for ZL_FIRST, BLOCK_PIVOT and GROUNDED. */
int tags_immittance_solve(const TAGS_COMPLEX *a, TAGS_COMPLEX *b, size_t num_electrodes,
                          size_t num_nodes, int precision, int plan, tags_item_stats *stats);

/* Panel algorithm of every LU factorisation in the process:
 * TAGS_PANEL_TOURNAMENT (default): tournament pivoting (communication-avoiding
 *   LU): candidate rows are reduced in a tree of local partial-pivoting
 *   factorisations; an ordinary launch, any number of concurrent callers;
 * TAGS_PANEL_PARTIAL: partial pivoting over the whole panel (LAPACK's pivot
 *   order) with a cooperative launch and one grid barrier per column; needs the
 *   device's cooperative-launch support, which one process holds at a time.
 * Returns 0, or TAGS_ERR_ARG (-12) for another value. */
enum tags_panel { TAGS_PANEL_TOURNAMENT = 0, TAGS_PANEL_PARTIAL = 1 };
int tags_set_panel(int kind);

/* Look-ahead of every later LU factorisation in the process: 1 (default)
 * factors the next panel on a second stream while the trailing update of the
 * current step runs; 0 runs the steps strictly in order on one stream. */
int tags_set_lookahead(int on);

/* Number of launches (split by columns) of the part of each step's trailing
 * update that runs beside the look-ahead panel (default 4; at most 2 when
 * that panel has no row exchanges). Each launch boundary lets panel
 * work-groups waiting for a CU start. Returns 0, or TAGS_ERR_ARG for n < 1. */
int tags_set_update_chunks(int n);

/* Columns per factorisation step of every LU in the process, both precisions:
 * 256 (four sub-panels of 64; the trailing update is a K = 256 product) or 128;
 * 0 restores the defaults (128 for complex128 factors, 256 for complex64). */
int tags_set_step_width(int columns);

/* Triangular solves of every LU in the process (no reference counterpart;
 * part of the solve that replaces zsysv, src/linalg.c:113-139): rows = 1 or 2
 * runs each range of the solve forward and backward as one launch each,
 * work-groups of 64 rows (1) or 128 rows (2; complex128 factors take 64)
 * taking their blocks in dependency order and streaming the off-diagonal
 * blocks while earlier blocks resolve (default 2); 0 launches one
 * diagonal-block solve and one
 * product per factorisation step. Results agree to rounding. Returns 0, or
 * TAGS_ERR_ARG for other values. */
int tags_set_solve_rows(int rows);

/* Library version string. */
const char *tags_version(void);

/* Waits for all work on the current device (hipDeviceSynchronize). */
int tags_device_synchronize(void);

/* GEMM tuning. The trailing update C -= A B of the LU has several MFMA tile
 * variants per precision (precision: tags_precision). Complex64 variants
 * 0-21 and 33 run on the fp32 MFMA cores; 22-32 run on the bf16 MFMA cores,
 * every fp32 operand value split into three bf16 parts (six part products per
 * real product, fp32 accumulation: the accuracy of the fp32 tiles) after a
 * packing pass into a work space that the LU's work space provides (K <= 256;
 * a product without one runs variant 19). tags_gemm_bench times one variant
 * on an M x N x K problem (mean ms per launch); tags_autotune times every
 * variant on a large (8192 x 8192 x K) and a small (1024 x 1024 x K) problem,
 * K the LU's step (128 complex128, 256 complex64), and installs the fastest
 * of each for this process; tags_gemm_select installs variants explicitly. */
int tags_gemm_bench(int M, int N, int K, int precision, int variant, int iters, double *ms);
int tags_gemm_variants(int precision);
/* One product through one variant, for validation: C (M x N) = A B (mode 0) or
 * C -= A B (mode 1), host operands row-major (A M x K, B K x N, C M x N) in
 * complex128, rounded to the precision's type on the way in; optional row
 * maps (permutations of [0, M), [0, K), [0, M)) address A, B and C rows as the
 * LU's row map does. No counterpart in the reference. */
int tags_gemm_run(int M, int N, int K, int precision, int variant, const TAGS_COMPLEX *A,
                  const TAGS_COMPLEX *B, TAGS_COMPLEX *C, const int *arow, const int *brow,
                  const int *crow, int mode);
int tags_gemm_select(int precision, int big_variant, int small_variant);
int tags_autotune(int precision, int *big_variant, int *small_variant);

#ifdef __cplusplus
}
#endif
#endif /* TAGS_HIP_H_ */
