/*
 * yee.h -- C ABI of the MI355X Yee time-stepping engine (libyee.so).
 *
 * The engine replaces the reference's per-cell field update path:
 *   Scheme::performNStepsForBlock / performFieldSteps / calculateFieldStep
 *     (Source/Scheme/Scheme.inc.h:28-473, 785-1344)
 *   InternalScheme::calculateFieldStepIteration{,PML,PMLMetamaterials}, calculateTFSF,
 *   performPlaneWave{E,H}Steps, performPointSourceCalc
 *     (Source/Scheme/InternalScheme.inc.h:1386-1835, 2351-2640, 2818-2895, 3276-3357)
 * together with the host-side set-up those updates read:
 *   Scheme::initGrids (materials), SchemeHelper::initSigma{X,Y,Z} (PML profiles),
 *   InternalScheme::initScheme (dt, frequency, incident phase velocity)
 *     (Source/Scheme/Scheme.inc.h:2962-3165, Source/Scheme/SchemeHelper.h:574-718,
 *      Source/Scheme/InternalScheme.inc.h:3359-3394).
 *
 * The reference has no FFI for this path: its seam is the C++ call from Scheme into
 * InternalScheme / InternalSchemeGPU.  Every entry point below names the reference
 * member function whose role it takes.
 *
 * Conventions
 *   - All functions return 0 on success and a negative YEE_E* code on failure; the
 *     message of the last failure on the calling thread is yee_last_error().
 *   - Host pointers are borrowed for the duration of the call and copied.
 *   - Grids are dense [x][y][z] arrays with z fastest (the reference's storage order,
 *     Source/Grid/Grid.h:585-730); 2D (TMz, TEz) grids are [x][y].
 *   - A handle owns its device memory and one HIP stream; a handle is not thread-safe,
 *     distinct handles may be used concurrently.
 *   - Fields start at zero, as in the reference (field loading is commented out in
 *     Source/Settings/Settings.inc.h:250-272).
 */
#ifndef YEE_H
#define YEE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YEE_ABI_VERSION 1

/* SchemeType values of the reference (Source/Settings/Settings.h:47-62). */
#define YEE_SCHEME_EXHY 1   /* 1D schemes: the grid runs along z (ExHy, EyHx), y (ExHz, EzHx) or x (EyHz, EzHy) */
#define YEE_SCHEME_EXHZ 2
#define YEE_SCHEME_EYHX 3
#define YEE_SCHEME_EYHZ 4
#define YEE_SCHEME_EZHX 5
#define YEE_SCHEME_EZHY 6
#define YEE_SCHEME_TEX 7
#define YEE_SCHEME_TEY 8
#define YEE_SCHEME_TEZ 9
#define YEE_SCHEME_TMX 10
#define YEE_SCHEME_TMY 11
#define YEE_SCHEME_TMZ 12
#define YEE_SCHEME_DIM3 13

/* yee_config.flags */
#define YEE_USE_PML 0x1u            /* --use-pml */
#define YEE_USE_TFSF 0x2u           /* --use-tfsf */
#define YEE_USE_METAMATERIALS 0x4u  /* --use-metamaterials (requires --use-pml, Scheme.inc.h:889-892) */
#define YEE_COMPLEX_FIELDS 0x8u     /* complex field values (the reference's -DCOMPLEX_FIELD_VALUES build,
                                       Kernels/FieldValue.h:68-72); see yee_get_imag_part */
#define YEE_DOUBLE_MATERIAL_PRECISION 0x10u
                                    /* --use-double-material-precision: material grids (eps, mu) of twice
                                       the field grid's extent along each axis of the scheme, each field
                                       point averaging the 8 material cells around it
                                       (YeeGridLayout::initMaterialCoordinates, Layout/YeeGridLayout.h:866-1037;
                                       InternalScheme::getMaterial, Scheme/InternalScheme.inc.h:2924-2950);
                                       PML sigma profiles on the same fine grid (SchemeHelper.h:598, 640, 682).
                                       Excludes metamaterials: the reference's 8-cell Drude average is
                                       UNREACHABLE (Layout/Approximation.cpp:196) */

/* Analytic test problems of the reference's real build (Source/Scheme/CallBack.cpp:28-143), the
   kinds of --use-<kind>-start-values / --use-<kind>-border-condition / --use-<kind>-right-side /
   --calc-<kind>-diff-norm (Settings.inc.h:134-153); see yee_set_test_functions. */
#define YEE_FN_NONE 0
#define YEE_FN_POLINOM1 1   /* Ez = c^2 x^2 t, Hy = eps0 Ez; right side Jz, My */
#define YEE_FN_POLINOM2 2   /* Ex = c^2 t (y^2 + z^2) and its cyclic twins, H = eps0 E; right side J, M */
#define YEE_FN_POLINOM3 3   /* Ez = c^2 (x t)^2, Hy = eps0 Ez; right side Jz, My */
#define YEE_FN_SIN1 4       /* Ez = c sin (t - x), Hy = -c sin (t - x); no right side */
/* Exact solutions of a 1D plane wave meeting a dielectric half-space (yee_set_exact_exp): */
#define YEE_FN_EXP1 5       /* incident i e^{i (k (z - z0) - w (t - t0))} */
#define YEE_FN_EXP2 6       /* reflected -i/3 e^{i (-k (z - (2 zb - z0)) - w (t - t0))} */
#define YEE_FN_EXP3 7       /* transmitted 2i/3 e^{i (2k (z - (zb + z0)/2) - w (t - t0))} */

/* Field components (GridType order of the reference). */
enum yee_component
{
  YEE_EX = 0, YEE_EY = 1, YEE_EZ = 2, YEE_HX = 3, YEE_HY = 4, YEE_HZ = 5,
  /* auxiliary PML / Drude state (valid inside the PML slabs; see yee_get_field) */
  YEE_DX = 6, YEE_DY = 7, YEE_DZ = 8, YEE_BX = 9, YEE_BY = 10, YEE_BZ = 11,
  YEE_D1X = 12, YEE_D1Y = 13, YEE_D1Z = 14, YEE_B1X = 15, YEE_B1Y = 16, YEE_B1Z = 17
};

/* Material grids (Source/Scheme/Grids.inc.h). */
enum yee_material
{
  YEE_MAT_EPS = 0, YEE_MAT_MU = 1, YEE_MAT_OMEGAPE = 2, YEE_MAT_OMEGAPM = 3, YEE_MAT_GAMMAE = 4, YEE_MAT_GAMMAM = 5
};

/* Error codes. */
#define YEE_OK 0
#define YEE_EINVAL (-1)      /* bad argument (null pointer, size, unsupported option) */
#define YEE_ENOMEM (-2)      /* device or host allocation failed */
#define YEE_EDEVICE (-3)     /* HIP runtime error */
#define YEE_ESTATE (-4)      /* call out of order (e.g. step before prepare) */
#define YEE_ENOTSUP (-5)     /* feature not built or not supported */

/*
 * Run configuration, filled from the reference's option table
 * (Source/Settings/Settings.inc.h) by the CLI front end or by a binding.
 */
typedef struct yee_config
{
  int32_t scheme;             /* YEE_SCHEME_DIM3 (--3d) or a 2D scheme YEE_SCHEME_TEX .. TMZ (--2d-tex .. --2d-tmz);
                                 2D grids are [a][b] over the scheme's two axes: TEx/TMx (y, z), TEy/TMy (x, z),
                                 TEz/TMz (x, y) (main.cpp:817-1000); the third size is ignored.  1D schemes
                                 YEE_SCHEME_EXHY .. EZHY (--1d-exhy .. --1d-ezhy) take the size of their one
                                 axis (main.cpp:629-800) and ignore the other two */
  int32_t layout;             /* --layout-type: 0 = E_CENTERED, 1 = H_CENTERED (E and H positions swapped,
                                 YeeGridLayout.cpp:1937-2043; runs without metamaterials and double
                                 material precision; NTFF and the adjoint take E_CENTERED runs only) */
  int32_t precision;          /* 4 = FP32 fields, 8 = FP64 fields (reference: -DFLOAT_VALUES / -DDOUBLE_VALUES) */
  int32_t time_steps;         /* --time-steps T; step t = 0 performs no field update (Scheme.inc.h:1032) */
  int64_t size[3];            /* --size x:,y:,z: */
  int64_t pml[3];             /* --pml-size */
  int64_t tfsf_left[3];       /* --tfsf-size-left */
  int64_t tfsf_right[3];      /* --tfsf-size-right */
  double dx;                  /* --dx (m) */
  double wavelength;          /* --wavelength (m) */
  double courant;             /* --courant-factor */
  double angle_deg[3];        /* --angle-teta, --angle-phi, --angle-psi (degrees) */
  uint32_t flags;             /* YEE_USE_* */
  uint32_t point_source;      /* bit (1 << yee_component) for each --point-source-<comp> */
  int64_t point_source_pos[3];/* --point-source x:,y:,z: */
  uint32_t current_source;    /* bit (1 << yee_component) for each --current-source-j<x|y|z> (E) / -m<x|y|z> (H) */
  int64_t current_source_pos[3]; /* --current-source x:,y:,z: (a cell the step updates) */
} yee_config;

typedef struct yee_handle yee_handle;

/* Thread-local message of the last failed call ("" if none). */
const char *yee_last_error (void);

/* Library / device information: fills up to `len` bytes of `buf` with a one-line description. */
int yee_device_info (char *buf, int64_t len);

/*
 * Creates an engine for `cfg` and allocates its device state.
 * Replaces InternalScheme::init + allocateGrids/allocateGridsInc
 * (Source/Scheme/InternalScheme.specific.h:37-107) and InternalScheme::initScheme
 * (Source/Scheme/InternalScheme.inc.h:3359-3394).
 */
int yee_create (const yee_config *cfg, yee_handle **out);

void yee_destroy (yee_handle *h);

/*
 * Sets one material grid (full [x][y][z] FP64 array of size[0]*size[1]*size[2] values,
 * or [x][y] in 2D; the fine grid of yee_material_size with YEE_DOUBLE_MATERIAL_PRECISION).  Defaults: eps = mu = 1, omega/gamma = 0.  Omega values are in
 * rad/s (the reference scales --omegape-sphere by 2*pi*f before storing, Scheme.inc.h:3091).
 * Replaces the material part of Scheme::initGrids (Source/Scheme/Scheme.inc.h:2962-3165).
 */
int yee_set_material (yee_handle *h, int32_t which, const double *values);

/*
 * Extent of the material grids along x, y, z: the field grid's size, or twice it along each axis
 * of the scheme with YEE_DOUBLE_MATERIAL_PRECISION (1 along an axis the scheme lacks).  Replaces
 * YeeGridLayout::getEpsSize (Source/Layout/YeeGridLayout.h:144, sizeEps of YeeGridLayout.cpp:1972).
 * yee_set_material* and yee_get_material take [x][y][z] arrays of this size.
 */
int yee_material_size (yee_handle *h, int64_t out[3]);

/* Same as yee_set_material from an FP32 array (the reference's -DFLOAT_VALUES grids). */
int yee_set_material_f32 (yee_handle *h, int32_t which, const float *values);

/*
 * Fills material `which` with a sphere the way Scheme::initGrids does
 * (Source/Scheme/Scheme.inc.h:2979-3012 for eps; mu, omegaPE, omegaPM likewise at :3020-3165):
 * every cell gets Approximation::approximateSphereAccurate (volume fraction over an
 * accuracy x accuracy raster; Source/Physics/Approximation.cpp:563-676) or, when stair != 0,
 * approximateSphereStair (:730-764) of `value` inside and `outside` elsewhere.  `center` is in
 * cell indices (the reference adds 0.5, the eps-point offset); `radius` in cells.  The value
 * replaces the whole grid, as in the reference (a sphere overrides a loaded file).  Omega values
 * are in rad/s: the reference passes --omegape-sphere * 2*pi*f here.  With
 * YEE_DOUBLE_MATERIAL_PRECISION the sphere is drawn on the fine grid with centre 2 center + 0.5
 * and radius 2 radius (Scheme.inc.h:2992-3009).
 */
int yee_set_material_sphere (yee_handle *h, int32_t which, double value, double outside, const int64_t center[3],
                             double radius, int32_t accuracy, int32_t stair);

/*
 * Host FP64 set-up of everything the kernels read: PML sigma profiles
 * (SchemeHelper::initSigma{X,Y,Z}, Source/Scheme/SchemeHelper.h:574-718), the material
 * encoding, Drude coefficients, TF/SF face tables and the incident 1D grid parameters.
 * Must be called after the last yee_set_material and before yee_step.
 */
int yee_prepare (yee_handle *h);

/*
 * Advances `n` time steps (t = current .. current + n - 1), enqueued on the handle's
 * stream; returns after the work is enqueued.  Replaces Scheme::performNStepsForBlock
 * (Source/Scheme/Scheme.inc.h:28-473).
 */
int yee_step (yee_handle *h, int32_t n);

/*
 * Gaussian-modulated hard point source: the value the reference's performPointSourceCalc
 * assigns (Source/Scheme/InternalScheme.inc.h:2818-2895, sin(2 pi f dt ts) with ts = t + 1/2
 * for E and t + 1 for H) is multiplied by exp(-((ts - t0) / sigma)^2), t0 and sigma in time
 * steps.  No counterpart in the reference (it only has the plain sinusoid, sigma <= 0 here);
 * it is the source of the batched TMz sweep.  May be called at any time; applies from the
 * next step on.
 */
int yee_set_point_source_gaussian (yee_handle *h, double t0, double sigma);

/*
 * Complex field values (YEE_COMPLEX_FIELDS): the reference's update multiplies every field by
 * real coefficients (computeCaCb and the PML / Drude coefficients are FIELDVALUE (c, 0)), so its
 * real and imaginary parts evolve independently and meet only in the sources, which the complex
 * build makes complex: the incident wave (sin, cos) of dt t 2 pi f (InternalScheme.inc.h:3310-3316),
 * the point source (sin, cos) (:2888-2893), the current source (cos, -sin) of dt (ts - 1) 2 pi f
 * (Scheme.inc.h:1230-1235).  A complex handle holds two real parts, each with its part of every
 * source; calls on the handle act on both.  yee_get_field & co. read the real part; this call
 * returns the handle of the imaginary part (owned by h: read from it, never destroy or step it
 * by itself).  YEE_EINVAL for a real-field handle.
 */
int yee_get_imag_part (yee_handle *h, yee_handle **imag);

/*
 * Analytic test problem (YEE_FN_* kinds, YEE_FN_NONE for none), replacing the callbacks
 * Scheme::initCallBacks installs for --use-<kind>-start-values / -border-condition / -right-side
 * and --calc-<kind>-diff-norm (Source/Scheme/Scheme.inc.h:2150-2430):
 *   start_values: every cell of the kind's components at E: dt / 2, H: dt
 *                 (Scheme::initGridWithInitialVals, Scheme.inc.h:2863-2958);
 *   border:       every step, after a component's update, its cells with an index 0 or N-1
 *                 (calculateFieldStepIterationBorder, InternalScheme.inc.h:2646-2705);
 *   right_side:   J (E) / M (H) subtracted in the curl update, calcField (InternalScheme.inc.h:765-772);
 *   exact:        every step the difference norm of the kind's components over the box
 *                 [norm_start, norm_end) (0 entries: [1, N-1); --norm-start / --norm-end), read
 *                 with yee_get_diff_norm (calculateFieldStepIterationExact, InternalScheme.inc.h:
 *                 2707-2816; Scheme.inc.h:1377-1550).
 * Before yee_prepare.  Runs on the half-step kernels.  YEE_ENOTSUP with PML, metamaterials or
 * complex fields (the reference's border and norm then act on its D grids), and for start values
 * without border values (its two time layers then alternate the boundary cells between the start
 * values and zero).
 */
int yee_set_test_functions (yee_handle *h, int32_t start_values, int32_t border, int32_t right_side, int32_t exact,
                            const int64_t norm_start[3], const int64_t norm_end[3]);

/*
 * Difference norm of component comp against the exact solution at the last step, as
 * Scheme::calculateFieldStep prints it ("DIFF NORM", Scheme.inc.h:1543-1549): out[0] =
 * sqrt (sum of (exact - field)^2 over the norm box / cells of the grid), out[1] = out[0] * 100 /
 * max |exact| (percent), out[2] = the sum, out[3] = max |exact|.  Taken after the component's
 * update and border values and before its point source, as the reference does.
 */
int yee_get_diff_norm (yee_handle *h, int32_t comp, double out[4]);

/*
 * Exact solution exp1 / exp2 / exp3 (YEE_FN_EXP1 .. EXP3) of the 1D pair `pair`
 * (YEE_SCHEME_EXHY .. EZHY: its E and H component, the line axis) for the reference's
 * --calc-exp<n>-<pair>-diff-norm (Scheme.inc.h:2437-2760; CallBack.cpp:280-760): E1 the
 * incident wave, E2 the wave reflected and E3 the wave transmitted by a dielectric of eps = 4
 * starting at `border` (the reference's eps sphere centre - radius along the line), each
 * H = E k / (mu0 w) (-k for E2, 2k for E3), z0 = the TF/SF left size along the line - 2.5,
 * k the numerical wave number, t0 = dt / 2, and the signs of the reference's per-pair wrappers.
 * Every step the diff norm of the pair's components over the norm box (as yee_set_test_functions),
 * read with yee_get_diff_norm (the real part; a complex handle's imaginary part handle gives
 * the imaginary part) or yee_get_diff_norm_complex.  Runs with PML (the norm reads E, as the
 * reference's normGrid = gridPML2) and complex fields; YEE_ENOTSUP with metamaterials and, for a
 * complex handle, point sources (its module norm is taken after the step).  Before yee_prepare.
 */
int yee_set_exact_exp (yee_handle *h, int32_t kind, int32_t pair, double border, const int64_t norm_start[3],
                       const int64_t norm_end[3]);

/*
 * The complex build's diff norm of component comp at the last step (Scheme.inc.h:1537-1542):
 * out[0..2] = the norms of the real part, the imaginary part and the module
 * (sqrt (sum / cells of the grid), the module's sum over (|exact| - |field|)^2),
 * out[3..5] their percentages of max |Re exact|, max |Im exact|, max |exact|, out[6..8] those
 * maxima.  A complex handle with an exp exact solution (yee_set_exact_exp).
 */
int yee_get_diff_norm_complex (yee_handle *h, int32_t comp, double out[9]);

/*
 * Near-to-far-field transform (--use-ntff) of a complex 3D handle at its newest step, as
 * Scheme::saveNTFF computes one diagram (Scheme.inc.h:4535-4640): for each of the nphi angles
 * phi[i] (radians) at the elevation teta (radians; the reference passes the incident wave's
 * teta), out[i] = Pointing_scat / Pointing_inc (Scheme.inc.h:4036-4110) over the box
 * [left, right] (the reference's leftNTFF / rightNTFF: size - left + 1 on its right side), whose
 * six faces sum the surface currents N (from H averaged over the newest two steps) and L (from E
 * averaged across the face), SchemeHelper::ntffN3D / ntffL3D (SchemeHelper.h:1030-2148).
 * yee_ntff_snapshot keeps the H of the step before: call it, step once, then call yee_ntff
 * (YEE_ESTATE otherwise).  YEE_ENOTSUP without YEE_COMPLEX_FIELDS, for 1D / 2D schemes (the
 * reference's are unreachable) and odd sizes; YEE_EINVAL unless 2 <= left < right <= size - 2.
 */
int yee_ntff_snapshot (yee_handle *h);
int yee_ntff (yee_handle *h, double teta, const double *phi, int32_t nphi, const int64_t left[3],
              const int64_t right[3], double *out);

/* Waits for all enqueued work of the handle. */
int yee_sync (yee_handle *h);

/* Number of time steps performed so far. */
int yee_time (yee_handle *h, int32_t *t);

/*
 * Copies the newest time layer of a component to host memory as FP64
 * ([x][y][z], or [x][y] in 2D).  Components absent from the scheme return YEE_EINVAL.
 * D/B (and D1/B1) hold the reference's values inside the PML slabs; in the interior the
 * engine updates E/H directly, which equals the reference's D->E stage with sigma = 0
 * up to rounding, and D/B read back as 0 there.
 */
int yee_get_field (yee_handle *h, int32_t comp, double *out);

/* Same as yee_get_field, in the handle's own precision (float* when precision == 4). */
int yee_get_field_raw (yee_handle *h, int32_t comp, void *out);

/* Material grid `which` (yee_material) as FP64 [x][y][z], as the engine holds it: the values given
 * to yee_set_material* (relative eps / mu, omega_p and gamma in rad/s) in the handle's precision;
 * a grid never set reads as its default (1 for eps / mu, 0 for the Drude terms).  Replaces the
 * material grids Scheme::initGrids dumps with --save-materials (Scheme.inc.h:3166-3200). */
int yee_get_material (yee_handle *h, int32_t which, double *out);

/* Scattered field of component comp (EX..HZ) as FP64 [x][y][z]: the newest field minus the
 * incident plane wave at the component's grid points inside the TF/SF box, the total field
 * outside it.  YEE_EINVAL without TF/SF.  Replaces Scheme::makeGridScattered (Scheme.inc.h:4113-4228),
 * used by --save-scattered-field-res. */
int yee_get_scattered_field (yee_handle *h, int32_t comp, double *out);

/*
 * Same as yee_get_scattered_field, allocating nothing: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that sample the field every few steps).
 * No counterpart in the reference.  The field is staged through a fixed 32 MiB device array
 * (4 Mi FP64 values, one per device, held until the process ends) in chunks, so any grid size
 * works; the call must be made on the device the handle was created on (YEE_EINVAL otherwise).
 * Calls that overlap on one GPU share that array and race.
 */
int yee_get_scattered_field_noalloc (yee_handle *h, int32_t comp, double *out);

/*
 * Mask export for bit-exact checks of the set-up tables.
 *   kind 0: TF/SF correction mask of component `comp`: bit d set when the reference's
 *           doNeedTFSFUpdate<comp>Border(pos, dir d) is true, d in LEFT=0, RIGHT=1, DOWN=2,
 *           UP=3, BACK=4, FRONT=5 (Source/Layout/YeeGridLayout.cpp:272-1740).
 *   kind 1: PML-slab mask (1 where the engine runs the D/B two-stage update).
 */
int yee_get_mask (yee_handle *h, int32_t kind, int32_t comp, uint8_t *out);

/* Copies the incident 1D grids (newest layer), `len` values each; either pointer may be null. */
int yee_get_incident (yee_handle *h, double *einc, double *hinc, int64_t len);

/* Scalars of InternalScheme::initScheme: out[0] = dt, out[1] = frequency, out[2] = relPhaseVelocity. */
int yee_get_scalars (yee_handle *h, double out[3]);

/*
 * Timing hook for benchmarks: average device time (ms) per step of the main stencil kernel
 * over the last yee_step call, measured with HIP events on the handle's stream.
 */
int yee_last_step_ms (yee_handle *h, double *ms);

/*
 * Runs `n` steps like yee_step with a HIP event pair around every group of kernel launches
 * and returns average device milliseconds per step of each group: out[0] incident 1D
 * update, out[1] E half-step (with the fused step: the E update of the PML slabs), out[2]
 * H half-step (fused: the H update of the PML slabs), out[3] fused E+H step kernel of the
 * PML-free box (0 when not used), out[4] whole step (yee_last_step_ms of the run).
 * Benchmark hook; the extra events cost a few microseconds per step.
 */
int yee_profile_steps (yee_handle *h, int32_t n, double out[5]);

/*
 * Forward/backward pair (shape of a torch.autograd.Function): J(eps) and dJ/d eps over a
 * design box, by the discrete adjoint (time-reversed transpose) of the engine's own 3D update.
 * No counterpart in the reference, which has no differentiation.
 *   J = sum over the last `window` steps of the sum of Ez^2 on the rectangle
 *       [monitor_lo[0], monitor_hi[0]) x [monitor_lo[1], monitor_hi[1]) of plane z = monitor_z.
 * checkpoint_every = 0 records E on the design box after every step during yee_forward;
 * K > 0 keeps a full-state snapshot every K steps instead and re-runs each segment during
 * yee_backward (memory: one snapshot per K steps, K + 1 design-box slots).
 */
typedef struct yee_adjoint_spec
{
  int64_t design_lo[3], design_hi[3];   /* eps voxels [lo, hi) (inside the PML-free interior) */
  int64_t monitor_z;
  int64_t monitor_lo[2], monitor_hi[2];
  int32_t window;                       /* 1 <= window <= time_steps - 1 */
  int32_t checkpoint_every;
} yee_adjoint_spec;

/* After yee_prepare: allocates the adjoint state (3D, no metamaterials). */
int yee_adjoint_setup (yee_handle *h, const yee_adjoint_spec *spec);
/* Sets the design-box eps (x-major, FP64), resets every field to zero and runs time_steps
 * steps; returns J.  The rest of the eps grid is the one set with yee_set_material. */
int yee_forward (yee_handle *h, const double *design_eps, double *J);
/* dJ/d eps of the last yee_forward, times the upstream gradient dJ, into grad_eps (design box,
 * x-major).  Leaves the fields in an unspecified state (re-run yee_forward to use them). */
int yee_backward (yee_handle *h, double dJ, double *grad_eps);
/* out[0] = device ms of the last yee_forward, out[1] = of the last yee_backward */
int yee_adjoint_timing (yee_handle *h, double out[2]);

/*
 * Batched TMz sweep: `nsim` independent TMz simulations (each its own yee_config: size, PML,
 * point source; one precision for the batch; no TF/SF, no metamaterials) advanced together,
 * two launches per time step for the whole batch, replayed from a captured HIP graph.  The
 * reference runs one simulation per process (Source/main.cpp); the batch has no counterpart
 * there and replaces running the same simulations one by one.  Every member computes exactly
 * what a yee_handle with its configuration computes.  Errors: yee_batch_last_error().
 */
typedef struct yee_batch yee_batch;

const char *yee_batch_last_error (void);
int yee_batch_create (const yee_config *cfgs, int32_t nsim, yee_batch **out);
void yee_batch_destroy (yee_batch *b);
/* eps (YEE_MAT_EPS) or mu (YEE_MAT_MU) of simulation `sim`, [x][y] */
int yee_batch_set_material (yee_batch *b, int32_t sim, int32_t which, const double *values);
int yee_batch_set_material_f32 (yee_batch *b, int32_t sim, int32_t which, const float *values);
/* Gaussian envelope of simulation sim's point source (see yee_set_point_source_gaussian) */
int yee_batch_set_point_source_gaussian (yee_batch *b, int32_t sim, double t0, double sigma);
int yee_batch_prepare (yee_batch *b);
/* Advances every simulation by n steps; the first call with enough steps also times the
 * tile-shape candidates on its first steps and keeps the fastest (out: yee_batch_tuning). */
int yee_batch_step (yee_batch *b, int32_t n);
int yee_batch_sync (yee_batch *b);
int yee_batch_time (yee_batch *b, int32_t *t);
/* average device ms per step of the last yee_batch_step call (HIP events on the batch stream) */
int yee_batch_last_step_ms (yee_batch *b, double *ms);
/* out[0] = chosen tile height (x rows per 64-cell tile), out[1] = workgroups per launch of
 * the first group, out[2] = groups (simulations advanced together through all their steps
 * before the next group, sized to stay cache resident); cand_ms[0..ncand) = ms per step the
 * tuner measured for each candidate */
int yee_batch_tuning (yee_batch *b, int32_t out[3], double *cand_ms, int32_t ncand);
/* newest layer of Ez, Hx, Hy (or Dz, Bx, By) of simulation sim as FP64 [x][y] */
int yee_batch_get_field (yee_batch *b, int32_t sim, int32_t comp, double *out);

#ifdef __cplusplus
}
#endif

#endif /* YEE_H */
