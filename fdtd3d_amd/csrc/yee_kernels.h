// Device-side argument blocks of the Yee kernels (shared by the kernels and the host driver).
#pragma once

#include <cstdint>

namespace yee {

constexpr int kMaxDim = 3;

// y rows per thread of the fused step kernel (FP64 holds half as many values per register)
#define FUSED_BY(T) (sizeof (T) == 4 ? 2 : 1)

// Incident-wave geometry of the TF/SF correction (FP64; InternalScheme::approximateIncidentWave).
template <typename T>
struct TfsfGeom
{
  double zinc[3];              // zeroIncCoordFP
  double S1, C1, S2, C2;       // sin / cos of the incidence angles
  double proj[6];              // get<comp>FromIncident factors
  const T *einc;               // EInc after this step's E update (read by H)
  const T *hinc[2];            // HInc by step parity (E reads the layer after the previous step)
  int64_t inc_len;
};

// Geometry and tables common to every half-step kernel.
template <typename T>
struct StepArgs
{
  // grid: [x][y][z], z fastest; TMz uses nz == 1
  int nx, ny, nz;
  int64_t sx;   // ny * nz
  int sy;       // nz
  int sz;       // 1 (0 along an axis the scheme lacks, as sx / sy then are)

  // cells that run the reference's two-stage D/B update: the PML slabs
  // (p < slab_lo[a] or p >= slab_hi[a]) and, with metamaterials, the Drude box
  int slab_lo[3], slab_hi[3];
  int box_lo[3], box_hi[3];
  // cells a launch updates: [run_lo, run_hi) on every axis (the full interior [1, N-1), a
  // PML slab, or the PML-free box of the fused step kernel)
  int run_lo[3], run_hi[3];

  // state read by this half-step / step
  T *f[6];      // Ex, Ey, Ez, Hx, Hy, Hz
  T *d[6];      // D/B (PML and Drude cells)
  T *dp[6];     // D/B one step older (Drude)
  T *d1[6];     // D1/B1 (Drude)
  T *d1p[6];    // D1/B1 one step older (Drude)
  // state written (the same arrays for the in-place half-step kernels, the other
  // buffer of each pair for the fused step kernel)
  T *f_n[6], *d_n[6], *dp_n[6], *d1_n[6], *d1p_n[6];

  // materials: per-cell arrays, or a uniform value when the pointer is null
  const T *eps, *mu, *ope, *opm, *ge, *gm;
  T eps_u, mu_u;
  // double material precision: the averaged eps (E) / mu (H) of each component per field cell,
  // read in place of the 2- / 4-cell averages of eps / mu when non-null
  const T *matc[6];

  // scalars in the field precision
  T dt, dx, eps0, mu0;
  // per-component 1D PML coefficient tables, [comp][role][tab_len]
  const T *pml;
  int tab_len;

  // TF/SF: the box bounds are tested for every cell; the incident-wave geometry is only read
  // by the O(N^2) corrected cells and lives in device memory (keeps scalar registers free)
  int tfsf;
  int tl[3], tr[3];            // total-field box [tl, tr] in integer coordinates
  const struct TfsfGeom<T> *tg;
  int hinc_sel;                // HInc buffer holding the layer after the previous step

  // fused step kernel: work decomposition and the point sources it applies in registers
  int lx;                      // x planes per chunk
  int nstrips;                 // y strips per chunk
  // half-step tile grid (e_half_3d / h_half_3d): z origin of the tiles, tiles along z and y,
  // planes, tiles per XCD band, total tiles
  int tz0, ntz, nty, ntx, per_xcd, ntile;
  uint32_t src_mask;           // bit c: hard source on component c at flat index src_q
  int64_t src_q;
  int src_pos[3];              // the same position as (x, y, z)
  uint32_t cur_mask;           // bit c: current source J (E) / M (H) on component c at cur_pos
  int cur_pos[3];
  double cur_val;              // the current of this half step: sin (dt ts 2 pi f) mu0 / dt / 2
  // TF/SF corrections of the cells on the z faces of the TF/SF box (planes k = L-1, R-1, R),
  // precomputed per step: [part (E, H)][face][x - (L-1)][y - (L-1)][comp (Ex/Ey or Hx/Hy)][slot lo/hi]
  const double *zcorr;
  int zc_n[2];                 // extents of the face tables along x and y
  // analytic test problem (YEE_FN_* kind, 0 = none): the right side J / M subtracted from the
  // curl of the half-step kernels, evaluated at time rhs_time (s) at (p + min coordinate) * fn_dx
  // (fn_dx = dx along the scheme's axes, 0 along an axis it lacks)
  int rhs_kind;
  double rhs_time;
  double fn_dx[3];
  // exact solutions exp1 .. exp3 of a 1D pair (yee_set_exact_exp): the pair (0 ExHy .. 5 EzHy),
  // the wave number of the phase (k, -k or 2k), the phase origin in cells, w, t0, the H factor
  // (k, -k or 2k) / (mu0 w), and the part this handle holds (0 real, 1 imaginary)
  int exp_pair, exp_part;
  double exp_k, exp_d, exp_w, exp_t0, exp_m;
  // Yee layout: 0 E_CENTERED, 1 H_CENTERED (component positions of the test problems)
  int layout;
};

// Near-to-far-field transform (yee_ntff): the complex fields' two parts, the snapshot of H before
// the last step, the box and the angles.
template <typename T>
struct NtffArgs
{
  const T *e_re[3], *e_im[3], *h_re[3], *h_im[3], *hp_re[3], *hp_im[3];
  int n[3];
  int64_t sx;
  int sy;
  int left[3], right[3];
  double dx, k, teta;
  const double *phis;
  double *part;   // [angle][12 face passes][4]: teta sum (re, im), phi sum (re, im)
};

}  // namespace yee
