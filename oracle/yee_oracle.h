/*
 * yee_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference's Yee update path, used as the checker by
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  The product
 * (fdtd3d_amd, libyee.so) never links, loads or calls it.
 *
 * It follows the reference's semantics literally: two (three with Drude+PML) time
 * layers per grid, D/B computed over the whole grid when PML is on, coefficients
 * recomputed per cell from averaged materials, TF/SF predicates on floating-point
 * coordinates with the reference's tolerances.  Citations are per function in
 * yee_oracle.c.
 */
#ifndef YEE_ORACLE_H
#define YEE_ORACLE_H

#include <stdint.h>
#include "../include/yee.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

oracle *oracle_create (const yee_config *cfg);
void oracle_destroy (oracle *o);
/* Same semantics as yee_set_material (FP64 [x][y][z]). */
int oracle_set_material (oracle *o, int32_t which, const double *values);
/* Sigma profiles and other set-up; call after the materials. */
int oracle_prepare (oracle *o);
int oracle_step (oracle *o, int32_t n);
/* Newest layer of a component ([x][y][z]); returns -1 if the component is absent. */
int oracle_get_field (oracle *o, int32_t comp, double *out);
/* Material grid as stored (after prepare); which: yee_material, or 6/7/8 = SigmaX/Y/Z. */
int oracle_get_material (oracle *o, int32_t which, double *out);
/* TF/SF predicate bits per cell for component comp (same encoding as yee_get_mask kind 0). */
int oracle_tfsf_mask (oracle *o, int32_t comp, uint8_t *out);
int oracle_get_incident (oracle *o, double *einc, double *hinc, int64_t len);
/* Scattered field of a component (same semantics as yee_get_scattered_field); -1 without TF/SF. */
int oracle_get_scattered (oracle *o, int32_t comp, double *out);
int oracle_get_scalars (oracle *o, double out[3]);
int64_t oracle_incident_len (oracle *o);
/* Test hook: phase added to every source argument (pi/2 gives the cos-source run). */
int oracle_set_source_phase (oracle *o, double phase);
/* Sources of the complex build: 1 its real part, 2 its imaginary part, 0 the real build. */
int oracle_set_source_kind (oracle *o, int kind);
int oracle_set_source_gaussian (oracle *o, double t0, double sigma);
/* Analytic test problems (YEE_FN_* kinds for start values, border values, right side and the
 * exact solution of the diff norm) and the norm box (0 entries: [1, N-1)); before prepare. */
int oracle_set_test_functions (oracle *o, int start, int border, int rhs, int exact, const int64_t ns[3],
                               const int64_t ne[3]);
/* Near-to-far-field value of the complex build at (teta, phi): the real and imaginary parts are two
 * oracles stepped together; out[0] = Pointing_scat / Pointing_inc, out[1] = Pointing_scat. */
int oracle_ntff (const oracle *re, const oracle *im, double teta, double phi, const int64_t left[3],
                 const int64_t right[3], double *out);
/* Multiplies every E / H value of every time layer by s (a test of the far field conjugates the
 * complex fields by scaling the imaginary part's oracle by -1). */
void oracle_scale_fields (oracle *o, double s);
/* Exact solution exp1 / exp2 / exp3 (YEE_FN_EXP1 .. EXP3) of a 1D pair (YEE_SCHEME_EXHY .. EZHY),
 * border = eps sphere centre - radius along the line; before prepare. */
int oracle_set_exact_exp (oracle *o, int kind, int pair, double border, const int64_t ns[3], const int64_t ne[3]);
/* Module norm (out[0]) and its percentage (out[1]) of the complex build from the real and
 * imaginary part oracles at the same step. */
int oracle_exp_mod_norm (const oracle *re, const oracle *im, int32_t comp, double out[2]);
/* Diff norm of a component at the last step: norm, percent of max |exact|, sum of squares, max. */
int oracle_get_diff_norm (oracle *o, int32_t comp, double out[4]);

/* FP64 adjoint of the reference-form update (see yee_oracle.c): runs T steps from a freshly
 * prepared oracle, returns J = sum over the last `window` steps of sum of Ez^2 on the monitor
 * rectangle [mlo, mhi) x {mz}, and dJ/deps over the design box [dlo, dhi) (x-major). */
typedef struct
{
  int64_t dlo[3], dhi[3];
  int64_t mz, mlo[2], mhi[2];
  int32_t window;
} oracle_adj_spec;
int oracle_adjoint (oracle *o, const oracle_adj_spec *sp, int32_t T, double *J, double *grad);

/* Material generators, restated from Approximation.cpp (used to build parity inputs). */
double oracle_sphere_accurate_3d (double mx, double my, double mz, double cx, double cy, double cz,
                                  double radius, double eps, double outside, int accuracy);
/* 1D segments (Approximation.cpp:390-448, 678-699): cell [m - 1/2, m + 1/2], segment [c - r, c + r]. */
double oracle_sphere_accurate_1d (double m, double c, double radius, double eps, double outside);
double oracle_sphere_stair_1d (double m, double c, double radius, double eps, double outside);
double oracle_sphere_accurate_2d (double mx, double my, double cx, double cy,
                                  double radius, double eps, double outside, int accuracy);
double oracle_sphere_stair_3d (double mx, double my, double mz, double cx, double cy, double cz,
                               double radius, double eps, double outside);

/* Numerical wave number of the incident wave (Approximation::approximateWaveNumber). */
double oracle_wave_number (double delta, double lambda, double courant, double nlambda, double a1, double a2);
double oracle_wave_number_general (double delta, double lambda, double courant, double nlambda, double a1, double a2);

#ifdef __cplusplus
}
#endif

#endif /* YEE_ORACLE_H */
