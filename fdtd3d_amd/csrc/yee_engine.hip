// Host driver and C ABI of the Yee engine (include/yee.h).
//
// One handle = one simulation on one GPU: device state, the HIP stream its work is
// queued on, and the FP64 set-up tables.  yee_step enqueues per time step:
//   incident_step (TF/SF only)  ->  E half-step  ->  E point sources  ->  H half-step  ->  H point sources
// which is the ordering of Scheme::performNStepsForBlock (Source/Scheme/Scheme.inc.h:28-473).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/yee.h"
#include "yee_internal.h"
#include "yee_kernels.h"
#include "yee_setup.h"

namespace yee {

template <typename T, int MODE> __global__ void e_half_3d (StepArgs<T>);
template <typename T, int MODE> __global__ void e_half_3d_hc (StepArgs<T>);
template <typename T, int MODE> __global__ void h_half_3d_hc (StepArgs<T>);
template <typename T, int MODE> __global__ void h_half_3d (StepArgs<T>);
template <typename T, int MODE> __global__ void e_half_tmz (StepArgs<T>);
template <typename T, int MODE> __global__ void h_half_tmz (StepArgs<T>);
template <typename T, int MODE> __global__ void e_half_tez (StepArgs<T>);
template <typename T, int MODE> __global__ void h_half_tez (StepArgs<T>);
template <typename T> __global__ void incident_step (T *, const T *, T *, int64_t, int, double, double, double);
template <typename T> __global__ void point_source (T *, int64_t, double);
template <typename T> __global__ void tfsf_mask_kernel (StepArgs<T>, int, int, uint8_t *);
template <typename T, int COMP, int DIM> __global__ void scattered_kernel (StepArgs<T>, const T *, double *, int, int64_t, int64_t);
template <typename T, int BY, int MAT, int MAXW> __global__ void step_3d_fused (StepArgs<T>);
template <typename T> __global__ void tfsf_ztable (StepArgs<T>, double *);
template <typename T> __global__ void test_fill (StepArgs<T>, T *, int, int, double, int);
template <typename T> __global__ void ntff_faces (NtffArgs<T>);
template <typename T>
__global__ void test_norm (StepArgs<T>, const T *, const T *, int, int, double, int, int, int, int, int, int, double *);
template <typename T>
__global__ void fine_average (const T *, int64_t, int64_t, int64_t, int, int, int, int, int, T *, T *, T *);

static thread_local std::string g_error;

int fail (int code, const std::string &msg)
{
  g_error = msg;
  return code;
}



// Tile grid of a half-step launch over the run box of `a` (half_cell): bx z cells from a
// multiple of bx, by rows, one plane; one workgroup per tile on a 1-D grid.
template <typename T>
bool half_tiles (StepArgs<T> &a, unsigned bx, unsigned by, dim3 &grid)
{
  // tiles start at a z multiple of their width; a box no wider than one tile starts it at the
  // box (a 16-cell PML z slab row then fills the tile instead of half of an aligned one)
  a.tz0 = (a.run_hi[2] - a.run_lo[2] <= (int) bx) ? a.run_lo[2] : (a.run_lo[2] & ~(int) (bx - 1));
  a.ntz = (int) ((a.run_hi[2] - a.tz0 + (int) bx - 1) / (int) bx);
  a.nty = (int) ((a.run_hi[1] - a.run_lo[1] + (int) by - 1) / (int) by);
  a.ntx = a.run_hi[0] - a.run_lo[0];
  const int64_t ntile = (int64_t) a.ntz * a.nty * a.ntx;
  if (ntile <= 0 || ntile > (int64_t) 1 << 30) return false;
  a.ntile = (int) ntile;
  a.per_xcd = (int) ((ntile + 7) / 8);
  grid = dim3 ((unsigned) (8 * a.per_xcd), 1, 1);
  return true;
}
template bool half_tiles<float> (StepArgs<float> &, unsigned, unsigned, dim3 &);
template bool half_tiles<double> (StepArgs<double> &, unsigned, unsigned, dim3 &);
}  // namespace yee

using namespace yee;


namespace {

// Which buffers a launch reads and writes (double-buffered state only; in place otherwise).
enum ArgRole
{
  kRoleE,       // E update: old E, old H -> new E
  kRoleH,       // H update: new E, old H -> new H
  kRoleFused,   // whole step: old E, old H -> new E, new H
};

template <typename T>
StepArgs<T> make_args (const yee_handle *h, int parity, ArgRole role = kRoleE)
{
  StepArgs<T> a {};
  a.nx = (int) h->n[0];
  a.ny = (int) h->n[1];
  a.nz = (int) h->n[2];
  a.sx = h->n[1] * h->n[2];
  a.sy = (int) h->n[2];
  for (int ax = 0; ax < 3; ++ax)
  {
    // PML two-stage cells: p < P or p >= N - P - 1 (every averaged sigma is zero elsewhere)
    a.slab_lo[ax] = h->use_pml () ? (int) h->cfg.pml[ax] : -1;
    a.slab_hi[ax] = h->use_pml () ? (int) (h->n[ax] - h->cfg.pml[ax] - 1) : 1 << 30;
    a.box_lo[ax] = (int) h->box_lo[ax];
    a.box_hi[ax] = (int) h->box_hi[ax];
    a.tl[ax] = (int) h->cfg.tfsf_left[ax];
    a.tr[ax] = (int) (h->cfg.size[ax] - h->cfg.tfsf_right[ax]);
  }
  if (h->dim == 2)
  {
    a.slab_lo[2] = -1;
    a.slab_hi[2] = 1 << 30;
  }
  const int src = h->cur, dst = h->fused ? 1 - h->cur : h->cur;
  for (int c = 0; c < 6; ++c)
  {
    a.f[c] = (T *) h->f[(role == kRoleH && c < 3) ? dst : src][c];
    a.d[c] = (T *) h->d[src][c];
    a.dp[c] = (T *) h->dp[src][c];
    a.d1[c] = (T *) h->d1[src][c];
    a.d1p[c] = (T *) h->d1p[src][c];
    a.f_n[c] = (T *) h->f[dst][c];
    a.d_n[c] = (T *) h->d[dst][c];
    a.dp_n[c] = (T *) h->dp[dst][c];
    a.d1_n[c] = (T *) h->d1[dst][c];
    a.d1p_n[c] = (T *) h->d1p[dst][c];
  }
  for (int ax = 0; ax < 3; ++ax)
  {
    a.run_lo[ax] = 1;
    a.run_hi[ax] = (int) h->n[ax] - 1;
  }
  // current source J / M (Scheme.inc.h:1171-1260): the current of this half step, with the
  // reference's operations, at timestep t (E) or t + 1/2 (H); applied by the half-step kernels
  a.cur_mask = h->cfg.current_source & 0x3fu;
  for (int ax = 0; ax < 3; ++ax) a.cur_pos[ax] = (int) h->cfg.current_source_pos[ax];
  {
    const double ts = role == kRoleH ? (double) h->t + 0.5 : (double) h->t;
    double current;
    if (h->src_kind == 0) current = std::sin (h->sc.dt * ts * 2 * kPi * h->sc.freq);
    // the complex build's current is (cos, -sin) of dt (ts - 1) 2 pi f (Scheme.inc.h:1230-1235)
    else if (h->src_kind == 1) current = std::cos (h->sc.dt * (ts - 1) * 2 * kPi * h->sc.freq);
    else current = -std::sin (h->sc.dt * (ts - 1) * 2 * kPi * h->sc.freq);
    a.cur_val = current * mu0 () / h->sc.dt / 2.0;
  }
  a.sz = 1;
  for (int m = 0; m < 3; ++m)
  {
    if (!h->lack[m]) continue;
    // a 1D / 2D scheme on the 3D kernels: its one cell along a lacked axis is updated, a
    // neighbour along that axis is the cell itself (stride 0), and the PML and TF/SF tests
    // never fire along it
    a.run_lo[m] = 0;
    a.run_hi[m] = 1;
    if (m == 0) a.sx = 0;
    else if (m == 1) a.sy = 0;
    else a.sz = 0;
    a.slab_lo[m] = -1;
    a.slab_hi[m] = 1 << 30;
    a.tl[m] = -(1 << 28);
    a.tr[m] = 1 << 28;
  }
  a.eps = (const T *) h->mat[YEE_MAT_EPS];
  a.mu = (const T *) h->mat[YEE_MAT_MU];
  a.ope = (const T *) h->mat[YEE_MAT_OMEGAPE];
  a.opm = (const T *) h->mat[YEE_MAT_OMEGAPM];
  a.ge = (const T *) h->mat[YEE_MAT_GAMMAE];
  a.gm = (const T *) h->mat[YEE_MAT_GAMMAM];
  a.eps_u = (T) h->mat_u[YEE_MAT_EPS];
  a.mu_u = (T) h->mat_u[YEE_MAT_MU];
  if (h->fine ())
  {
    // eps / mu live on the fine grid: the kernels read the per-component averages instead
    a.eps = a.mu = nullptr;
    for (int c = 0; c < 6; ++c) a.matc[c] = (const T *) h->matc[c];
  }
  a.dt = (T) h->sc.dt;
  a.dx = (T) h->cfg.dx;
  a.eps0 = (T) eps0 ();
  a.mu0 = (T) mu0 ();
  a.pml = (const T *) h->pml;
  a.tab_len = h->tab_len;
  a.tfsf = h->use_tfsf () ? 1 : 0;
  a.tg = (const TfsfGeom<T> *) h->tg;
  a.zcorr = (const double *) h->zcorr;
  a.zc_n[0] = (int) h->zc_n[0];
  a.zc_n[1] = (int) h->zc_n[1];
  // the E half-step of step t reads HInc after step t-1, kept in buffer (t+1) & 1
  a.hinc_sel = parity;
  // analytic test problem: the right side of this half step, at time step t (E) or t + 1/2 (H)
  // (Scheme.inc.h:1062-1098)
  a.rhs_kind = h->fn_rhs;
  a.rhs_time = (role == kRoleH ? (double) h->t + 0.5 : (double) h->t) * h->sc.dt;
  for (int ax = 0; ax < 3; ++ax) a.fn_dx[ax] = axis_active (h->cfg, ax) ? h->cfg.dx : 0.0;
  a.exp_pair = h->exp_pair;
  a.exp_part = h->src_kind == 2 ? 1 : 0;
  a.exp_k = h->exp_k;
  a.exp_d = h->exp_d;
  a.exp_w = h->exp_w;
  a.exp_t0 = h->exp_t0;
  a.exp_m = h->exp_m;
  a.layout = h->cfg.layout;
  return a;
}


template <typename T>
int launch_half (yee_handle *h, bool e_half, int parity, const int64_t *lo = nullptr, const int64_t *hi = nullptr)
{
  StepArgs<T> a = make_args<T> (h, parity, e_half ? kRoleE : kRoleH);
  if (lo)
    for (int ax = 0; ax < 3; ++ax)
    {
      a.run_lo[ax] = (int) lo[ax];
      a.run_hi[ax] = (int) hi[ax];
    }
  if (h->dim == 3)
  {
    const int64_t ext[3] = { a.run_hi[0] - a.run_lo[0], a.run_hi[1] - a.run_lo[1], a.run_hi[2] - a.run_lo[2] };
    if (ext[0] <= 0 || ext[1] <= 0 || ext[2] <= 0) return YEE_OK;
    // 64 z cells x 4 rows; boxes thin in z (PML z slabs) take 16 or 32 z cells x 16 or 8 rows,
    // boxes one row high (TF/SF y faces) 64 x 1, so that few lanes idle.
    const unsigned bx = ext[2] <= 16 ? 16 : (ext[2] <= 32 ? 32 : 64);
    const unsigned by = (unsigned) std::min<int64_t> (256 / bx, std::max<int64_t> (1, ext[1]));
    dim3 block (bx, by, 1), grid;
    if (!half_tiles (a, bx, by, grid)) return fail (YEE_EINVAL, "half-step box too large for one launch");
    // a box that misses the Drude box takes the PML kernels, which every such cell of a Drude
    // run takes anyway (e_half_3d)
    int mode = h->mode;
    if (mode == 2)
    {
      bool meets = true;
      for (int ax = 0; ax < 3; ++ax) meets = meets && a.run_lo[ax] < a.box_hi[ax] && a.box_lo[ax] < a.run_hi[ax];
      if (!meets) mode = 1;
    }
    if (h->cfg.layout == 1)
    {
      // H_CENTERED: plain or PML runs only (validate refuses the rest)
      switch (mode * 2 + (e_half ? 0 : 1))
      {
        case 0: hipLaunchKernelGGL ((e_half_3d_hc<T, 0>), grid, block, 0, h->stream, a); break;
        case 1: hipLaunchKernelGGL ((h_half_3d_hc<T, 0>), grid, block, 0, h->stream, a); break;
        case 2: hipLaunchKernelGGL ((e_half_3d_hc<T, 1>), grid, block, 0, h->stream, a); break;
        default: hipLaunchKernelGGL ((h_half_3d_hc<T, 1>), grid, block, 0, h->stream, a); break;
      }
    }
    else switch (mode * 2 + (e_half ? 0 : 1))
    {
      case 0: hipLaunchKernelGGL ((e_half_3d<T, 0>), grid, block, 0, h->stream, a); break;
      case 1: hipLaunchKernelGGL ((h_half_3d<T, 0>), grid, block, 0, h->stream, a); break;
      case 2: hipLaunchKernelGGL ((e_half_3d<T, 1>), grid, block, 0, h->stream, a); break;
      case 3: hipLaunchKernelGGL ((h_half_3d<T, 1>), grid, block, 0, h->stream, a); break;
      case 4: hipLaunchKernelGGL ((e_half_3d<T, 2>), grid, block, 0, h->stream, a); break;
      default: hipLaunchKernelGGL ((h_half_3d<T, 2>), grid, block, 0, h->stream, a); break;
    }
  }
  else if (h->cfg.scheme == YEE_SCHEME_TEZ)
  {
    dim3 block (64, 4, 1);
    dim3 grid ((unsigned) ((h->n[1] + 63) / 64), (unsigned) ((h->n[0] + 3) / 4), 1);
    switch (h->mode * 2 + (e_half ? 0 : 1))
    {
      case 0: hipLaunchKernelGGL ((e_half_tez<T, 0>), grid, block, 0, h->stream, a); break;
      case 1: hipLaunchKernelGGL ((h_half_tez<T, 0>), grid, block, 0, h->stream, a); break;
      case 2: hipLaunchKernelGGL ((e_half_tez<T, 1>), grid, block, 0, h->stream, a); break;
      case 3: hipLaunchKernelGGL ((h_half_tez<T, 1>), grid, block, 0, h->stream, a); break;
      case 4: hipLaunchKernelGGL ((e_half_tez<T, 2>), grid, block, 0, h->stream, a); break;
      default: hipLaunchKernelGGL ((h_half_tez<T, 2>), grid, block, 0, h->stream, a); break;
    }
  }
  else
  {
    dim3 block (64, 4, 1);
    dim3 grid ((unsigned) ((h->n[1] + 63) / 64), (unsigned) ((h->n[0] + 3) / 4), 1);
    switch (h->mode * 2 + (e_half ? 0 : 1))
    {
      case 0: hipLaunchKernelGGL ((e_half_tmz<T, 0>), grid, block, 0, h->stream, a); break;
      case 1: hipLaunchKernelGGL ((h_half_tmz<T, 0>), grid, block, 0, h->stream, a); break;
      case 2: hipLaunchKernelGGL ((e_half_tmz<T, 1>), grid, block, 0, h->stream, a); break;
      case 3: hipLaunchKernelGGL ((h_half_tmz<T, 1>), grid, block, 0, h->stream, a); break;
      case 4: hipLaunchKernelGGL ((e_half_tmz<T, 2>), grid, block, 0, h->stream, a); break;
      default: hipLaunchKernelGGL ((h_half_tmz<T, 2>), grid, block, 0, h->stream, a); break;
    }
  }
  HIP_TRY (hipGetLastError ());
  return YEE_OK;
}

// Hard point-source value of the E (t + 1/2) or H (t + 1) update of step t
// (performPointSourceCalc, InternalScheme.inc.h:2818-2895), with the C library's sin.
double source_value (const yee_handle *h, bool e_half, int32_t t)
{
  const double ts = e_half ? t + 0.5 : t + 1.0;
  // the complex build's point source is (sin, cos) of the same argument (InternalScheme.inc.h:2888-2893)
  const double arg = h->sc.dt * ts * 2 * kPi * h->sc.freq;
  double v = h->src_kind == 2 ? std::cos (arg) : std::sin (arg);
  if (h->src_sigma > 0)
  {
    const double u = (ts - h->src_t0) / h->src_sigma;
    v *= std::exp (-u * u);
  }
  return v;
}

// A cell next to an x or y face of the TF/SF box (the kAxFRange / kAxFPlane bits of the fused
// kernel): inside [L-1, R] on every axis and on one of the planes L-1, R-1, R of x or y.
bool tfsf_near_host (const yee_handle *h, const int64_t p[3])
{
  bool face = false;
  for (int ax = 0; ax < 3; ++ax)
  {
    const int64_t L = h->cfg.tfsf_left[ax], R = h->n[ax] - h->cfg.tfsf_right[ax];
    if (p[ax] < L - 1 || p[ax] > R) return false;
    face = face || (ax < 2 && (p[ax] == L - 1 || p[ax] == R - 1 || p[ax] == R));
  }
  return face;
}

// Is cell p inside the fused kernel's box and not one of the cells it leaves to the half-step
// kernels for other reasons (TF/SF x/y faces, Drude box)?  (own_mask of the kernels, without
// the source cell.)
bool fused_owns_host (const yee_handle *h, const int64_t p[3])
{
  bool drude = h->mode == 2;
  for (int ax = 0; ax < 3; ++ax)
  {
    if (p[ax] < h->fbox_lo[ax] || p[ax] >= h->fbox_hi[ax]) return false;
    drude = drude && p[ax] >= h->box_lo[ax] && p[ax] < h->box_hi[ax];
  }
  return !(h->use_tfsf () && tfsf_near_host (h, p)) && !drude;
}

bool source_in_fused_box (const yee_handle *h)
{
  return (h->cfg.point_source & 0x3fu) && fused_owns_host (h, h->cfg.point_source_pos);
}

// Point sources as separate one-cell launches into buffer `buf`; with the fused step the
// sources inside its box are applied by that kernel instead (skip_box).
template <typename T>
int launch_sources (yee_handle *h, bool e_half, int32_t t, int buf, bool skip_box)
{
  const int c0 = e_half ? 0 : 3;
  for (int c = c0; c < c0 + 3; ++c)
  {
    if (!(h->cfg.point_source & (1u << c)) || !h->exposed[c]) continue;
    const int64_t *p = h->cfg.point_source_pos;
    const int64_t pz = h->dim == 3 ? p[2] : 0;
    // getFieldValueOrNullByAbsolutePos: positions outside the grid are ignored
    if (p[0] < 0 || p[0] >= h->n[0] || p[1] < 0 || p[1] >= h->n[1] || pz < 0 || pz >= h->n[2]) continue;
    if (skip_box && source_in_fused_box (h)) continue;
    const int64_t q = (p[0] * h->n[1] + p[1]) * h->n[2] + pz;
    hipLaunchKernelGGL ((point_source<T>), dim3 (1), dim3 (64), 0, h->stream, (T *) h->f[buf][c], q,
                       source_value (h, e_half, t));
    HIP_TRY (hipGetLastError ());
  }
  return YEE_OK;
}

// Components an analytic test problem kind defines (Scheme::initCallBacks, Scheme.inc.h:
// 2150-2430): polinom2 all six, polinom1 / polinom3 / sin1 Ez and Hy.
bool test_defines (int kind, int comp)
{
  if (kind >= YEE_FN_EXP1) return false;   // per pair: exp_has
  if (kind == YEE_FN_POLINOM2) return true;
  if (kind == YEE_FN_POLINOM1 || kind == YEE_FN_POLINOM3 || kind == YEE_FN_SIN1) return comp == YEE_EZ || comp == YEE_HY;
  return false;
}

// Components of an exp exact solution: the E and H component of the handle's pair
// (Scheme.inc.h:2437-2760).
bool exp_has (const yee_handle *h, int comp)
{
  static const int kE[6] = { YEE_EX, YEE_EX, YEE_EY, YEE_EY, YEE_EZ, YEE_EZ };
  static const int kH[6] = { YEE_HY, YEE_HZ, YEE_HX, YEE_HZ, YEE_HX, YEE_HY };
  return h->fn_exact >= YEE_FN_EXP1 && (comp == kE[h->exp_pair] || comp == kH[h->exp_pair]);
}

// Records an event pair around the next launch when kernel timing is on.
struct ProfScope
{
  yee_handle *h;
  int kind;
  bool on;
  ProfScope (yee_handle *hh, int k) : h (hh), kind (k), on (hh->prof_events != nullptr)
  {
    if (on) record ();
  }
  ~ProfScope ()
  {
    if (on) record ();
  }
  void record ()
  {
    hipEvent_t e;
    if (hipEventCreate (&e) == hipSuccess)
    {
      (void) hipEventRecord (e, h->stream);
      h->prof_events->push_back (e);
      h->prof_kinds->push_back (kind);
    }
  }
};

// The six PML slabs around the fused kernel's box: [lo, hi) boxes covering
// [1, N-1)^3 minus the box.
int slab_boxes (const yee_handle *h, int64_t lo[6][3], int64_t hi[6][3])
{
  const int64_t *bl = h->fbox_lo, *bh = h->fbox_hi;
  const int64_t n0 = h->n[0] - 1, n1 = h->n[1] - 1, n2 = h->n[2] - 1;
  const int64_t L[6][3] = { { 1, 1, 1 }, { bh[0], 1, 1 }, { bl[0], 1, 1 }, { bl[0], bh[1], 1 },
                            { bl[0], bl[1], 1 }, { bl[0], bl[1], bh[2] } };
  const int64_t H[6][3] = { { bl[0], n1, n2 }, { n0, n1, n2 }, { bh[0], bl[1], n2 }, { bh[0], n1, n2 },
                            { bh[0], bh[1], bl[2] }, { bh[0], bh[1], n2 } };
  int m = 0;
  for (int s = 0; s < 6; ++s)
  {
    if (H[s][0] <= L[s][0] || H[s][1] <= L[s][1] || H[s][2] <= L[s][2]) continue;
    for (int ax = 0; ax < 3; ++ax)
    {
      lo[m][ax] = L[s][ax];
      hi[m][ax] = H[s][ax];
    }
    ++m;
  }
  return m;
}

// The cells next to the x and y faces of the TF/SF box (tfsf_near_host): for x and y the planes
// L-1, R-1 and R, spanning [L-1, R] on the other axes, clipped to the fused kernel's box (the
// z faces are corrected inside the fused kernel; the face cells outside its box lie in the slab
// boxes).  The face boxes overlap one another only inside the fused box, where no cell has PML
// or Drude state: a cell updated twice reads the old E / H buffer both times and gets the same
// value, while the in-place D / B of the slab cells are updated exactly once.
int face_boxes (const yee_handle *h, int64_t lo[9][3], int64_t hi[9][3])
{
  if (!h->use_tfsf ()) return 0;
  int m = 0;
  for (int ax = 0; ax < 2; ++ax)
  {
    const int64_t L = h->cfg.tfsf_left[ax], R = h->n[ax] - h->cfg.tfsf_right[ax];
    const int64_t planes[3] = { L - 1, R - 1, R };
    for (int64_t pl : planes)
    {
      bool ok = true;
      for (int b = 0; b < 3; ++b)
      {
        const int64_t Lb = h->cfg.tfsf_left[b], Rb = h->n[b] - h->cfg.tfsf_right[b];
        const int64_t l = b == ax ? pl : Lb - 1, u = b == ax ? pl + 1 : Rb + 1;
        lo[m][b] = std::max<int64_t> (l, h->fbox_lo[b]);
        hi[m][b] = std::min<int64_t> (u, h->fbox_hi[b]);
        ok = ok && hi[m][b] > lo[m][b];
      }
      if (ok) ++m;
    }
  }
  return m;
}

// One time step (t > 0) with the fused kernel: E of the PML slabs and TF/SF faces, E sources
// the fused kernel does not own, fused E+H sweep, H of the slabs and faces, H sources; old
// state in buffer cur, new state in 1 - cur.
template <typename T>
int fused_step (yee_handle *h, int32_t t, int prev)
{
  int rc;
  const int dst = 1 - h->cur;
  int64_t lo[20][3], hi[20][3];
  int nslab = slab_boxes (h, lo, hi);
  nslab += face_boxes (h, lo + nslab, hi + nslab);
  if (h->mode == 2)
  {
    // the Drude box (its cells run the two-stage update with the ADE terms), clipped to the interior
    bool ok = true;
    for (int ax = 0; ax < 3; ++ax)
    {
      lo[nslab][ax] = std::max<int64_t> (1, h->box_lo[ax]);
      hi[nslab][ax] = std::min<int64_t> (h->n[ax] - 1, h->box_hi[ax]);
      ok = ok && hi[nslab][ax] > lo[nslab][ax];
    }
    if (ok) ++nslab;
  }
  const bool src_box = source_in_fused_box (h);
  if (src_box)
  {
    // the point-source cell: updated by the half-step kernels, then overwritten by the source
    // (the fused kernel leaves it out, as it does the TF/SF faces)
    for (int ax = 0; ax < 3; ++ax)
    {
      lo[nslab][ax] = h->cfg.point_source_pos[ax];
      hi[nslab][ax] = h->cfg.point_source_pos[ax] + 1;
    }
    ++nslab;
  }
  const bool cur_on_src = src_box && h->cfg.current_source_pos[0] == h->cfg.point_source_pos[0]
                          && h->cfg.current_source_pos[1] == h->cfg.point_source_pos[1]
                          && h->cfg.current_source_pos[2] == h->cfg.point_source_pos[2];
  if (h->cfg.current_source && !cur_on_src && fused_owns_host (h, h->cfg.current_source_pos))
  {
    // the current-source cell: the half-step kernels subtract the current after its curl
    // update (the fused kernel leaves it out, kAxCur)
    for (int ax = 0; ax < 3; ++ax)
    {
      lo[nslab][ax] = h->cfg.current_source_pos[ax];
      hi[nslab][ax] = h->cfg.current_source_pos[ax] + 1;
    }
    ++nslab;
  }
  {
    ProfScope ps (h, 1);
    for (int s = 0; s < nslab; ++s)
      if ((rc = launch_half<T> (h, true, prev, lo[s], hi[s]))) return rc;
  }
  if ((rc = launch_sources<T> (h, true, t, dst, false))) return rc;
  if (h->zcorr)
  {
    StepArgs<T> a = make_args<T> (h, prev);
    const int64_t n = 3 * h->zc_n[0] * h->zc_n[1];
    hipLaunchKernelGGL ((tfsf_ztable<T>), dim3 ((unsigned) ((n + 255) / 256)), dim3 (256), 0, h->stream, a,
                       (double *) h->zcorr);
    HIP_TRY (hipGetLastError ());
  }
  {
    StepArgs<T> a = make_args<T> (h, prev, kRoleFused);
    for (int ax = 0; ax < 3; ++ax)
    {
      a.run_lo[ax] = (int) h->fbox_lo[ax];
      a.run_hi[ax] = (int) h->fbox_hi[ax];
    }
    a.lx = h->fused_lx;
    a.nstrips = h->fused_nstrips;
    if (src_box)
    {
      // excluded from the kernel's cells (axis_bits, kAxSrc)
      a.src_mask = h->cfg.point_source & 0x3fu;
      for (int ax = 0; ax < 3; ++ax) a.src_pos[ax] = (int) h->cfg.point_source_pos[ax];
    }
    ProfScope ps (h, 3);
    const unsigned blocks = (unsigned) (h->fused_nstrips * h->fused_nchunks);
    const dim3 grid (blocks), block ((unsigned) h->fused_threads);
    // rows of 9-12 waves: the 12-wave instantiation (more registers, the same one workgroup per CU)
    const int w12 = (h->fused_threads > 512 && h->fused_threads <= 768) ? 4 : 0;
    switch (w12 | (a.eps ? 1 : 0) | (a.mu ? 2 : 0))
    {
      case 0: hipLaunchKernelGGL ((step_3d_fused<T, FUSED_BY (T), 0, 16>), grid, block, 0, h->stream, a); break;
      case 1: hipLaunchKernelGGL ((step_3d_fused<T, FUSED_BY (T), 1, 16>), grid, block, 0, h->stream, a); break;
      case 2: hipLaunchKernelGGL ((step_3d_fused<T, FUSED_BY (T), 2, 16>), grid, block, 0, h->stream, a); break;
      case 3: hipLaunchKernelGGL ((step_3d_fused<T, FUSED_BY (T), 3, 16>), grid, block, 0, h->stream, a); break;
      case 4: hipLaunchKernelGGL ((step_3d_fused<T, FUSED_BY (T), 0, 12>), grid, block, 0, h->stream, a); break;
      case 5: hipLaunchKernelGGL ((step_3d_fused<T, FUSED_BY (T), 1, 12>), grid, block, 0, h->stream, a); break;
      case 6: hipLaunchKernelGGL ((step_3d_fused<T, FUSED_BY (T), 2, 12>), grid, block, 0, h->stream, a); break;
      default: hipLaunchKernelGGL ((step_3d_fused<T, FUSED_BY (T), 3, 12>), grid, block, 0, h->stream, a); break;
    }
    HIP_TRY (hipGetLastError ());
  }
  {
    ProfScope ps (h, 2);
    for (int s = 0; s < nslab; ++s)
      if ((rc = launch_half<T> (h, false, prev, lo[s], hi[s]))) return rc;
  }
  return launch_sources<T> (h, false, t, dst, false);
}

// Analytic test problem, after the half step's update of every component and before its point
// sources (Scheme::calculateFieldStep, Scheme.inc.h:1346-1550, then performPointSourceCalc):
// border values, then the diff norm of each component with an exact solution, at time step
// t + 1/2 (E) or t + 1 (H), also at t == 0.
constexpr int kNormBlocks = 1024;

template <typename T>
int test_hooks (yee_handle *h, bool e_half)
{
  const double time = (e_half ? (double) h->t + 0.5 : (double) h->t + 1.0) * h->sc.dt;
  StepArgs<T> a = make_args<T> (h, 0);
  const int64_t blocks = std::max<int64_t> (1, std::min<int64_t> ((h->ncell + 255) / 256, 65536));
  for (int c = e_half ? 0 : 3; c < (e_half ? 3 : 6); ++c)
  {
    if (!h->present[c]) continue;
    T *f = (T *) h->f[h->cur][c];
    if (h->fn_border && test_defines (h->fn_border, c))
    {
      hipLaunchKernelGGL ((test_fill<T>), dim3 ((unsigned) blocks), dim3 (256), 0, h->stream, a, f, c, h->fn_border, time, 1);
      HIP_TRY (hipGetLastError ());
    }
    if (h->fn_exact && (test_defines (h->fn_exact, c) || exp_has (h, c)))
    {
      hipLaunchKernelGGL ((test_norm<T>), dim3 (kNormBlocks), dim3 (256), 0, h->stream, a, (const T *) f,
                         (const T *) nullptr, c, h->fn_exact,
                         time, (int) h->norm_lo[0], (int) h->norm_lo[1], (int) h->norm_lo[2], (int) h->norm_hi[0],
                         (int) h->norm_hi[1], (int) h->norm_hi[2], h->norm_part + (size_t) c * 2 * kNormBlocks);
      HIP_TRY (hipGetLastError ());
      h->norm_valid[c] = true;
    }
  }
  return YEE_OK;
}

// Start values of an analytic test problem (E at dt / 2, H at dt), at yee_prepare.
template <typename T>
int test_start (yee_handle *h)
{
  StepArgs<T> a = make_args<T> (h, 0);
  const int64_t blocks = std::max<int64_t> (1, std::min<int64_t> ((h->ncell + 255) / 256, 65536));
  for (int c = 0; c < 6; ++c)
  {
    if (!h->present[c] || !test_defines (h->fn_start, c)) continue;
    const double time = c >= 3 ? h->sc.dt : 0.5 * h->sc.dt;
    hipLaunchKernelGGL ((test_fill<T>), dim3 ((unsigned) blocks), dim3 (256), 0, h->stream, a, (T *) h->f[h->cur][c], c,
                       h->fn_start, time, 0);
    HIP_TRY (hipGetLastError ());
  }
  return YEE_OK;
}

template <typename T>
int step_once (yee_handle *h)
{
  const int32_t t = h->t;
  const int prev = (t + 1) & 1;   // HInc after step t-1
  const int cur = t & 1;          // HInc after step t
  if (h->use_tfsf ())
  {
    const double modE = h->sc.dt / (h->sc.rel * eps0 () * h->cfg.dx);
    const double modH = h->sc.dt / (h->sc.rel * mu0 () * h->cfg.dx);
    const double arg = h->sc.dt * t * 2 * kPi * h->sc.freq;
    ProfScope ps (h, 0);
    hipLaunchKernelGGL ((incident_step<T>), dim3 (1), dim3 (1024), 0, h->stream, (T *) h->einc,
                       (const T *) h->hinc[prev], (T *) h->hinc[cur], h->inc_len, (int) t, modE, modH,
                       h->src_kind == 2 ? std::cos (arg) : std::sin (arg));   // complex build: (sin, cos)
    HIP_TRY (hipGetLastError ());
  }
  int rc;
  if (h->fused && t > 0)
  {
    if ((rc = fused_step<T> (h, t, prev))) return rc;
    h->cur = 1 - h->cur;
    h->t++;
    return YEE_OK;
  }
  if (t > 0)
  {
    ProfScope ps (h, 1);
    if ((rc = launch_half<T> (h, true, prev))) return rc;
  }
  if (h->fn_any () && (rc = test_hooks<T> (h, true))) return rc;
  if ((rc = launch_sources<T> (h, true, t, h->cur, false))) return rc;
  if (t > 0)
  {
    ProfScope ps (h, 2);
    if ((rc = launch_half<T> (h, false, prev))) return rc;
  }
  if (h->fn_any () && (rc = test_hooks<T> (h, false))) return rc;
  if ((rc = launch_sources<T> (h, false, t, h->cur, false))) return rc;
  h->t++;
  return YEE_OK;
}

// Grid-sized arrays carry padding on both sides: the fused kernel reads the z neighbours of
// a row as immediate offsets (one element before the first row) and whole rows past the last
// one (up to FUSED_BY + 1 rows); those values only reach cells that are not updated.
constexpr int64_t kPadFront = 64;
int64_t pad_back (const yee_handle *h) { return 4 * h->n[2] + 64; }

int alloc_cells (yee_handle *h, void **view, int64_t n)
{
  const size_t elems = (size_t) (kPadFront + n + pad_back (h));
  void *raw = nullptr;
  HIP_TRY (hipMalloc (&raw, elems * h->wsize));
  h->grids.push_back (raw);
  HIP_TRY (hipMemsetAsync (raw, 0, elems * h->wsize, h->stream));
  *view = (char *) raw + kPadFront * h->wsize;
  return YEE_OK;
}

int alloc_grid (yee_handle *h, void **view)
{
  return alloc_cells (h, view, h->ncell);
}

int alloc_zero (yee_handle *h, void **p, size_t bytes)
{
  HIP_TRY (hipMalloc (p, bytes));
  HIP_TRY (hipMemsetAsync (*p, 0, bytes, h->stream));
  return YEE_OK;
}

template <typename T>
std::vector<T> convert (const double *src, size_t n)
{
  std::vector<T> v (n);
  for (size_t q = 0; q < n; ++q) v[q] = (T) src[q];
  return v;
}

int upload (yee_handle *h, void **dst, const double *src, size_t n)
{
  if (!*dst)
  {
    HIP_TRY (hipMalloc (dst, n * h->wsize));
  }
  if (h->fp32 ())
  {
    std::vector<float> v = convert<float> (src, n);
    HIP_TRY (hipMemcpy (*dst, v.data (), n * 4, hipMemcpyHostToDevice));
  }
  else
  {
    HIP_TRY (hipMemcpy (*dst, src, n * 8, hipMemcpyHostToDevice));
  }
  return YEE_OK;
}

int validate (const yee_config *c)
{
  if (!c) return fail (YEE_EINVAL, "null config");
  if (c->scheme < YEE_SCHEME_EXHY || c->scheme > YEE_SCHEME_DIM3)
    return fail (YEE_EINVAL, "scheme must be YEE_SCHEME_DIM3 (13), a 2D scheme TEx .. TMz (7 .. 12) or a 1D "
                             "scheme ExHy .. EzHy (1 .. 6)");
  if (c->layout != 0 && c->layout != 1) return fail (YEE_EINVAL, "layout must be 0 (E_CENTERED) or 1 (H_CENTERED)");
  if (c->layout == 1 && (c->flags & (YEE_USE_METAMATERIALS | YEE_DOUBLE_MATERIAL_PRECISION)))
    return fail (YEE_EINVAL, "the H_CENTERED layout runs without metamaterials and double material precision");
  if (c->precision != 4 && c->precision != 8) return fail (YEE_EINVAL, "precision must be 4 or 8");
  for (int a = 0; a < 3; ++a)
  {
    if (!axis_active (*c, a)) continue;
    if (c->size[a] < 3 || c->size[a] > (1 << 20)) return fail (YEE_EINVAL, "size out of range");
    // the PML sizes matter only with PML on (a 3-cell grid keeps the default size 2 otherwise)
    if ((c->flags & YEE_USE_PML) && (c->pml[a] < 0 || 2 * c->pml[a] > c->size[a]))
      return fail (YEE_EINVAL, "pml size out of range");
    // a 1D line may put its right TF/SF border on the grid's end (right size 0: the scattered
    // field only left of the left border, as Tests/suite t7.x do); its cells N - 1 and N are
    // never updated, so no right-border correction fires
    const int64_t min_right = line_axis (*c) >= 0 ? 0 : 1;
    if ((c->flags & YEE_USE_TFSF)
        && (c->tfsf_left[a] < 1 || c->tfsf_right[a] < min_right || c->tfsf_left[a] + c->tfsf_right[a] >= c->size[a]))
      return fail (YEE_EINVAL, "TF/SF box out of range");
  }
  int64_t cells = 1;
  for (int a = 0; a < 3; ++a) cells *= axis_active (*c, a) ? c->size[a] : 1;
  if (cells > (int64_t) 1 << 34) return fail (YEE_EINVAL, "grid too large");
  if (!(c->dx > 0) || !(c->wavelength > 0) || !(c->courant > 0))
    return fail (YEE_EINVAL, "dx, wavelength and courant factor must be positive");
  if ((c->flags & YEE_USE_METAMATERIALS) && !(c->flags & YEE_USE_PML))
    return fail (YEE_EINVAL, "metamaterials require PML (reference: Scheme.inc.h:889-892)");
  if ((c->flags & YEE_USE_METAMATERIALS) && (c->flags & YEE_DOUBLE_MATERIAL_PRECISION))
    return fail (YEE_EINVAL, "double material precision excludes metamaterials (the reference's 8-cell Drude "
                             "average is UNREACHABLE, Approximation.cpp:196)");
  if (c->time_steps < 0) return fail (YEE_EINVAL, "negative time steps");
  if (c->current_source & 0x3fu)
    for (int a = 0; a < 3; ++a)
      if (axis_active (*c, a) && (c->current_source_pos[a] < 1 || c->current_source_pos[a] > c->size[a] - 2))
        return fail (YEE_EINVAL, "the current source must lie on a cell the step updates ([1, N-1) on each axis)");
  return YEE_OK;
}

template <typename S>
int upload_as (yee_handle *h, void **dst, const S *src, size_t n)
{
  if (!*dst)
  {
    int rc = alloc_cells (h, dst, (int64_t) n);
    if (rc) return rc;
    HIP_TRY (hipStreamSynchronize (h->stream));
  }
  if (h->fp32 ())
  {
    std::vector<float> v (n);
    for (size_t q = 0; q < n; ++q) v[q] = (float) src[q];
    HIP_TRY (hipMemcpy (*dst, v.data (), n * 4, hipMemcpyHostToDevice));
  }
  else
  {
    std::vector<double> v (n);
    for (size_t q = 0; q < n; ++q) v[q] = (double) src[q];
    HIP_TRY (hipMemcpy (*dst, v.data (), n * 8, hipMemcpyHostToDevice));
  }
  return YEE_OK;
}

// Scans one material grid: uniform grids become a scalar, nonzero Drude grids extend the
// Drude box, everything else is converted to the field precision and uploaded.
template <typename S>
int set_material_impl (yee_handle *h, int32_t which, const S *values)
{
  if (!h || !values) return fail (YEE_EINVAL, "null argument");
  if (which < 0 || which > 5) return fail (YEE_EINVAL, "unknown material");
  if (h->prepared) return fail (YEE_ESTATE, "materials must be set before yee_prepare");
  const size_t n = (size_t) h->mcell;   // the fine grid with double material precision
  bool uniform = true;
  bool nonzero = false;
  int64_t lo[3] = { h->n[0], h->n[1], h->n[2] }, hi[3] = { -1, -1, -1 };
  for (size_t q = 0; q < n; ++q)
  {
    if (!std::isfinite ((double) values[q])) return fail (YEE_EINVAL, "material values must be finite");
    if (values[q] != values[0]) uniform = false;
    if (values[q] != 0)
    {
      nonzero = true;
      if (which >= 2)
      {
        const int64_t k = (int64_t) (q % h->mn[2]);
        const int64_t j = (int64_t) ((q / h->mn[2]) % h->mn[1]);
        const int64_t i = (int64_t) (q / ((size_t) h->mn[2] * h->mn[1]));
        const int64_t p[3] = { i, j, k };
        for (int a = 0; a < 3; ++a)
        {
          lo[a] = std::min (lo[a], p[a]);
          hi[a] = std::max (hi[a], p[a]);
        }
      }
    }
  }
  if ((which == YEE_MAT_EPS || which == YEE_MAT_MU))
  {
    for (size_t q = 0; q < n; ++q)
      if (values[q] == 0) return fail (YEE_EINVAL, "eps/mu must be nonzero everywhere");
  }
  h->mat_nonzero[which] = nonzero;
  if (which >= 2 && nonzero)
  {
    // field cells whose averages reach a nonzero Drude cell: p in [lo - 1, hi]
    for (int a = 0; a < 3; ++a)
    {
      const int64_t l = lo[a] - 1, u = hi[a] + 1;
      h->box_lo[a] = h->box_any ? std::min (h->box_lo[a], l) : l;
      h->box_hi[a] = h->box_any ? std::max (h->box_hi[a], u) : u;
    }
    h->box_any = true;
  }
  if ((uniform && which < 2) || !nonzero)
  {
    if (which < 2) h->mat_u[which] = (double) values[0];
    h->mat[which] = nullptr;   // an earlier grid allocation stays owned by h->grids
    return YEE_OK;
  }
  return upload_as (h, &h->mat[which], values, n);
}

// Double material precision: the 8-cell averages of eps (E components) and mu (H components)
// per field cell, computed on the device from the fine grids (a uniform material needs none:
// every average of a constant is the constant).
template <typename T>
int fine_averages (yee_handle *h)
{
  int act = 0;
  for (int a = 0; a < 3; ++a)
    if (axis_active (h->cfg, a)) act |= 1 << a;
  for (int g = 0; g < 2; ++g)
  {
    if (!h->mat[g]) continue;
    T *out[3] = { nullptr, nullptr, nullptr };
    for (int r = 0; r < 3; ++r)
    {
      const int c = 3 * g + r;
      if (!h->present[c]) continue;
      int rc = alloc_grid (h, &h->matc[c]);
      if (rc) return rc;
      out[r] = (T *) h->matc[c];
    }
    const int64_t blocks = std::min<int64_t> ((h->ncell + 255) / 256, 65536);
    hipLaunchKernelGGL ((fine_average<T>), dim3 ((unsigned) blocks), dim3 (256), 0, h->stream,
                        (const T *) h->mat[g], h->mn[0], h->mn[1], h->mn[2], (int) h->n[0], (int) h->n[1],
                        (int) h->n[2], act, 3 * g, out[0], out[1], out[2]);
    HIP_TRY (hipGetLastError ());
  }
  return YEE_OK;
}

template <typename T>
int upload_tfsf_geom (yee_handle *h)
{
  TfsfGeom<T> g {};
  for (int ax = 0; ax < 3; ++ax) g.zinc[ax] = h->sc.zero_inc[ax];
  g.S1 = h->sc.S1;
  g.C1 = h->sc.C1;
  g.S2 = h->sc.S2;
  g.C2 = h->sc.C2;
  for (int c = 0; c < 6; ++c) g.proj[c] = h->sc.proj[c];
  g.einc = (const T *) h->einc;
  g.hinc[0] = (const T *) h->hinc[0];
  g.hinc[1] = (const T *) h->hinc[1];
  g.inc_len = h->inc_len;
  if (!h->tg) HIP_TRY (hipMalloc (&h->tg, sizeof g));
  HIP_TRY (hipMemcpy (h->tg, &g, sizeof g, hipMemcpyHostToDevice));
  return YEE_OK;
}

int ntff_snapshot_part (yee_handle *h)
{
  for (int c = 0; c < 3; ++c)
  {
    if (!h->ntff_h[c])
    {
      int rc = alloc_grid (h, &h->ntff_h[c]);
      if (rc) return rc;
    }
    HIP_TRY (hipMemcpyAsync (h->ntff_h[c], h->f[h->cur][3 + c], (size_t) h->ncell * h->wsize, hipMemcpyDeviceToDevice,
                             h->stream));
  }
  h->ntff_t = h->t;
  return YEE_OK;
}

int ntff_check (const yee_handle *h)
{
  if (!h->imag) return fail (YEE_ENOTSUP, "the near-to-far-field transform needs complex field values (YEE_COMPLEX_FIELDS)");
  if (h->cfg.layout != 0) return fail (YEE_ENOTSUP, "the near-to-far-field transform runs on the E_CENTERED layout only");
  if (h->dim != 3 || h->miss >= 0)
    return fail (YEE_ENOTSUP, "the near-to-far-field transform runs on 3D schemes only (the reference's 1D / 2D "
                              "variants are unreachable)");
  for (int a = 0; a < 3; ++a)
    if (h->n[a] % 2) return fail (YEE_ENOTSUP, "the near-to-far-field transform needs even grid sizes");
  return YEE_OK;
}

template <typename T>
NtffArgs<T> ntff_args (const yee_handle *h, double teta, const int64_t left[3], const int64_t right[3])
{
  NtffArgs<T> a {};
  for (int c = 0; c < 3; ++c)
  {
    a.e_re[c] = (const T *) h->f[h->cur][c];
    a.e_im[c] = (const T *) h->imag->f[h->imag->cur][c];
    a.h_re[c] = (const T *) h->f[h->cur][3 + c];
    a.h_im[c] = (const T *) h->imag->f[h->imag->cur][3 + c];
    a.hp_re[c] = (const T *) h->ntff_h[c];
    a.hp_im[c] = (const T *) h->imag->ntff_h[c];
    a.n[c] = (int) h->n[c];
    a.left[c] = (int) left[c];
    a.right[c] = (int) right[c];
  }
  a.sx = h->n[1] * h->n[2];
  a.sy = (int) h->n[2];
  a.dx = h->cfg.dx;
  a.k = 2 * kPi / h->cfg.wavelength;
  a.teta = teta;
  return a;
}

// yee_set_exact_exp on one part's handle.
int set_exact_exp_part (yee_handle *h, int32_t kind, int32_t pair, double border, const int64_t ns[3], const int64_t ne[3])
{
  int rc = yee_set_test_functions (h, 0, 0, 0, 0, ns, ne);
  if (rc) return rc;
  if (!h->norm_part)
  {
    void *p = nullptr;
    HIP_TRY (hipMalloc (&p, (size_t) 12 * 2 * kNormBlocks * sizeof (double)));
    h->norm_part = (double *) p;
  }
  static const int kAxis[6] = { 2, 1, 2, 0, 1, 0 };
  const int pr = pair - YEE_SCHEME_EXHY;
  // CallBack::exp1_e / exp2_e / exp3_e and exp*_h (CallBack.cpp:535-760), constants with the
  // reference's operations
  const double delta = h->cfg.dx, lambda = h->cfg.wavelength, courant = h->cfg.courant;
  const double nl = lambda / delta;
  const double k = wave_number (delta, lambda, courant, nl, h->sc.a1, h->sc.a2);
  const double f = kC / lambda;
  const double w = 2 * kPi * f;
  const double step = delta * courant / kC;
  const double c0 = (double) h->cfg.tfsf_left[kAxis[pr]] - 2.5;
  double kk, d, hk;
  if (kind == YEE_FN_EXP1)
  {
    kk = k;
    d = c0;
    hk = k;
  }
  else if (kind == YEE_FN_EXP2)
  {
    kk = -k;
    d = 2 * border - c0;
    hk = -k;
  }
  else
  {
    kk = 2 * k;
    d = (border + c0) / 2.0;
    hk = kk;
  }
  h->fn_exact = kind;
  h->exp_pair = pr;
  h->exp_k = kk;
  h->exp_d = d;
  h->exp_w = w;
  h->exp_t0 = 0.5 * step;
  h->exp_m = hk / (mu0 () * w);
  h->fused = false;
  return YEE_OK;
}

}  // namespace

namespace yee {

int engine_step_once (yee_handle *h)
{
  return h->fp32 () ? step_once<float> (h) : step_once<double> (h);
}

int engine_alloc_grid (yee_handle *h, void **view)
{
  return alloc_grid (h, view);
}

template <typename T>
StepArgs<T> engine_args (const yee_handle *h, int parity)
{
  return make_args<T> (h, parity);
}
template StepArgs<float> engine_args<float> (const yee_handle *, int);
template StepArgs<double> engine_args<double> (const yee_handle *, int);

int engine_reset (yee_handle *h)
{
  const size_t bytes = (size_t) (kPadFront + h->ncell + pad_back (h)) * h->wsize;
  for (int bsel = 0; bsel < 2; ++bsel)
    for (int c = 0; c < 6; ++c)
    {
      void *ptrs[5] = { h->f[bsel][c], h->d[bsel][c], h->dp[bsel][c], h->d1[bsel][c], h->d1p[bsel][c] };
      for (void *p : ptrs)
        if (p) HIP_TRY (hipMemsetAsync ((char *) p - kPadFront * h->wsize, 0, bytes, h->stream));
    }
  if (h->einc) HIP_TRY (hipMemsetAsync (h->einc, 0, (size_t) h->inc_len * h->wsize, h->stream));
  for (void *p : h->hinc)
    if (p) HIP_TRY (hipMemsetAsync (p, 0, (size_t) h->inc_len * h->wsize, h->stream));
  h->t = 0;
  h->cur = 0;
  return YEE_OK;
}

}  // namespace yee

extern "C" {

const char *yee_last_error (void)
{
  return g_error.c_str ();
}

int yee_device_info (char *buf, int64_t len)
{
  if (!buf || len <= 0) return fail (YEE_EINVAL, "null buffer");
  int dev = 0;
  hipDeviceProp_t prop;
  HIP_TRY (hipGetDevice (&dev));
  HIP_TRY (hipGetDeviceProperties (&prop, dev));
  snprintf (buf, (size_t) len, "device %d: %s (%s), %d CUs, %.1f GiB", dev, prop.name, prop.gcnArchName,
            prop.multiProcessorCount, prop.totalGlobalMem / 1073741824.0);
  return YEE_OK;
}

int yee_create (const yee_config *cfg, yee_handle **out)
{
  if (!out) return fail (YEE_EINVAL, "null output handle");
  *out = nullptr;
  int rc = validate (cfg);
  if (rc) return rc;
  yee_handle *h = new yee_handle ();
  h->cfg = *cfg;
  // 2D schemes on the (y, z) and (x, z) planes run on the 3D kernels with the missing axis of
  // extent 1 (neighbours along it are the cell itself, the offsets the reference drops); TEz
  // and TMz have kernels of their own
  // 1D schemes run on the 3D kernels with both other axes of extent 1
  // (H_CENTERED runs TEz / TMz on the 3D kernels too: the 2D kernels are E_CENTERED only)
  const int miss = missing_axis (*cfg), line = line_axis (*cfg);
  const bool hc = cfg->layout == 1;
  h->miss = (miss == 2 && !hc) ? -1 : miss;
  h->dim = (cfg->scheme == YEE_SCHEME_DIM3 || h->miss >= 0 || line >= 0) ? 3 : 2;
  for (int a = 0; a < 3; ++a)
  {
    h->n[a] = axis_active (*cfg, a) ? cfg->size[a] : 1;
    if (axis_active (*cfg, a)) continue;
    h->lack[a] = a != 2 || line >= 0 || hc;   // TEz / TMz keep kernels of their own (E_CENTERED)
    h->cfg.pml[a] = 0;                     // no PML and no TF/SF border along a lacked axis
    h->cfg.point_source_pos[a] = 0;        // the source positions drop that coordinate
    h->cfg.current_source_pos[a] = 0;
  }
  h->ncell = h->n[0] * h->n[1] * h->n[2];
  for (int a = 0; a < 3; ++a) h->mn[a] = (h->fine () && axis_active (*cfg, a)) ? 2 * h->n[a] : h->n[a];
  h->mcell = h->mn[0] * h->mn[1] * h->mn[2];
  h->wsize = cfg->precision == 4 ? 4 : 8;
  h->mode = h->use_meta () ? 2 : (h->use_pml () ? 1 : 0);
  for (int c = 0; c < 6; ++c)
  {
    // InternalScheme component flags (InternalScheme.inc.h:3243-3272): TE<a> has H along the
    // missing axis a and E across it, TM<a> the reverse
    // (a 1D scheme: the two components it names, InternalScheme.inc.h:3243-3260)
    static const int k1d[7][2] = { { -1, -1 }, { YEE_EX, YEE_HY }, { YEE_EX, YEE_HZ }, { YEE_EY, YEE_HX },
                                   { YEE_EY, YEE_HZ }, { YEE_EZ, YEE_HX }, { YEE_EZ, YEE_HY } };
    const bool te = cfg->scheme <= YEE_SCHEME_TEZ, along = (c % 3) == miss;
    if (line >= 0) h->exposed[c] = c == k1d[cfg->scheme][0] || c == k1d[cfg->scheme][1];
    else h->exposed[c] = miss < 0 || ((c >= 3) == te ? along : !along);
    // the 3D kernels update all six components; the absent ones stay exactly 0 (their curl
    // operands cancel or are absent, their incident projections are 0)
    h->present[c] = h->dim == 3 || h->exposed[c];
  }
  for (int c = 0; c < 6; ++c)
    if (!h->exposed[c])
    {
      h->cfg.point_source &= ~(1u << c);
      h->cfg.current_source &= ~(1u << c);   // the reference updates no absent component
    }
  auto bail = [&] (int code) {
    yee_destroy (h);
    return code;
  };
  if (hipGetDevice (&h->device) != hipSuccess) return bail (fail (YEE_EDEVICE, "hipGetDevice failed"));
  hipError_t e = hipStreamCreateWithFlags (&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail (fail (YEE_EDEVICE, std::string ("hipStreamCreate: ") + hipGetErrorString (e)));
  if (hipEventCreate (&h->ev[0]) != hipSuccess || hipEventCreate (&h->ev[1]) != hipSuccess)
    return bail (fail (YEE_EDEVICE, "hipEventCreate failed"));
  // 3D runs use the fused step kernel on the PML-free box outside the Drude box (double-
  // buffered state) when the doubled state fits in HBM; TMz runs update in place with the
  // half-step kernels.  YEE_KERNEL=split in the environment selects the half-step kernels for 3D
  // as well (A/B checks).
  const char *kern = std::getenv ("YEE_KERNEL");
  const bool want_fused = !(kern && std::strcmp (kern, "split") == 0);
  h->fused = false;
  // (double material precision keeps the half-step kernels: they read per-component averages)
  // (the fused kernel is E_CENTERED: H_CENTERED runs take the half-step kernels)
  if (h->dim == 3 && h->miss < 0 && line < 0 && h->n[2] <= 1024 && want_fused && !h->fine () && cfg->layout == 0)
  {
    bool ok = true;
    for (int ax = 0; ax < 3; ++ax)
    {
      const int64_t P = h->use_pml () ? cfg->pml[ax] : 0;
      h->fbox_lo[ax] = std::max<int64_t> (1, P);
      h->fbox_hi[ax] = std::min<int64_t> (h->n[ax] - 1, h->n[ax] - P - 1);
      ok = ok && h->fbox_hi[ax] > h->fbox_lo[ax];
    }
    // two buffers of E/H (6), one of D/B (6 with PML), two of D/B and the Drude history with
    // metamaterials (24)
    const int arrays = 12 + (h->mode == 1 ? 6 : 0) + (h->mode == 2 ? 48 : 0);
    const double bytes = (double) arrays * (double) h->ncell * (double) h->wsize;
    h->fused = ok && bytes < 200e9;
  }
  if (h->fused)
  {
    const int by = h->fp32 () ? FUSED_BY (float) : FUSED_BY (double);
    int dev = 0, ncu = 256;
    if (hipGetDevice (&dev) == hipSuccess) (void) hipDeviceGetAttribute (&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t rows = h->fbox_hi[1] - h->fbox_lo[1], planes = h->fbox_hi[0] - h->fbox_lo[0];
    h->fused_nstrips = (int) ((rows + by - 1) / by);
    // x chunks: a CU holds 16 waves of this kernel (<= 128 VGPRs each), i.e. 16 / (waves per
    // row) workgroups.  The chunk count minimises rounds of resident workgroups x planes per
    // chunk (+1 for the chunk's first plane, read twice): enough workgroups to fill every CU,
    // and no last round left mostly empty.
    const int64_t waves = (h->n[2] + 63) / 64;
    const int64_t slots = (int64_t) ncu * std::max<int64_t> (1, 16 / waves);
    auto chunk_cost = [&] (int64_t c) {
      const int64_t rounds = (h->fused_nstrips * c + slots - 1) / slots;
      return (double) rounds * (double) ((planes + c - 1) / c + 1);
    };
    double best = 1e300;
    for (int64_t c = 1; c <= planes; ++c) best = std::min (best, chunk_cost (c));
    int64_t nch = 1;   // the fewest chunks within 1 % of the best
    while (nch < planes && chunk_cost (nch) > 1.01 * best) ++nch;
    if (const char *e = std::getenv ("YEE_FUSED_CHUNKS")) nch = std::max (1, std::atoi (e));
    nch = std::min (nch, planes);
    h->fused_lx = (int) ((planes + nch - 1) / nch);
    h->fused_nchunks = (int) ((planes + h->fused_lx - 1) / h->fused_lx);
    h->fused_threads = (int) ((h->n[2] + 63) / 64 * 64);
  }
  if (h->fused && h->use_tfsf ())
  {
    // z-face TF/SF correction table of the fused kernel: 2 parts x 3 faces x [L-1, R]^2 x 4
    for (int ax = 0; ax < 2; ++ax)
      h->zc_n[ax] = (h->n[ax] - cfg->tfsf_right[ax]) - cfg->tfsf_left[ax] + 2;
    if ((rc = alloc_zero (h, &h->zcorr, (size_t) (6 * h->zc_n[0] * h->zc_n[1] * 4) * sizeof (double)))) return bail (rc);
  }
  // E / H are double-buffered for the fused kernel (it recomputes neighbour cells from old
  // values); the PML D / B are touched only by the half-step kernels, which read every old value
  // of a cell before they store it and update each slab cell in exactly one launch (slab boxes
  // are disjoint, face boxes are clipped to the fused box), so they update in place: the second
  // buffer's views alias the first's.  The Drude box overlaps slab and face boxes, so Drude
  // runs keep two buffers.
  const int nbuf = h->fused ? 2 : 1;
  for (int bsel = 0; bsel < nbuf; ++bsel)
    for (int c = 0; c < 6; ++c)
    {
      if (!h->present[c]) continue;
      if ((rc = alloc_grid (h, &h->f[bsel][c]))) return bail (rc);
      if (bsel == 1 && h->mode == 1)
      {
        h->d[1][c] = h->d[0][c];
        h->dp[1][c] = h->dp[0][c];
        h->d1[1][c] = h->d1[0][c];
        h->d1p[1][c] = h->d1p[0][c];
        continue;
      }
      if (h->mode >= 1 && (rc = alloc_grid (h, &h->d[bsel][c]))) return bail (rc);
      if (h->mode == 2)
      {
        if ((rc = alloc_grid (h, &h->dp[bsel][c]))) return bail (rc);
        if ((rc = alloc_grid (h, &h->d1[bsel][c]))) return bail (rc);
        if ((rc = alloc_grid (h, &h->d1p[bsel][c]))) return bail (rc);
      }
    }
  if (h->use_tfsf ())
  {
    // allocateGridsInc: 500 * size.x cells (InternalScheme.specific.h:99-107)
    h->inc_len = 500 * h->n[first_active_axis (*cfg)];   // size.get1 () of the scheme's grid
    if ((rc = alloc_zero (h, &h->einc, h->inc_len * h->wsize))) return bail (rc);
    if ((rc = alloc_zero (h, &h->hinc[0], h->inc_len * h->wsize))) return bail (rc);
    if ((rc = alloc_zero (h, &h->hinc[1], h->inc_len * h->wsize))) return bail (rc);
  }
  if (hipStreamSynchronize (h->stream) != hipSuccess) return bail (fail (YEE_EDEVICE, "stream sync failed"));
  h->sc = compute_scalars (h->cfg);
  if (h->complex ())
  {
    // complex field values: this handle is the real part; the imaginary part is a second real
    // handle of the same configuration whose sources are the imaginary parts of the complex
    // build's sources
    h->src_kind = 1;
    yee_config ci = *cfg;
    ci.flags &= ~YEE_COMPLEX_FIELDS;
    yee_handle *im = nullptr;
    if ((rc = yee_create (&ci, &im))) return bail (rc);
    im->src_kind = 2;
    h->imag = im;
  }
  *out = h;
  return YEE_OK;
}

void yee_destroy (yee_handle *h)
{
  if (!h) return;
  if (h->imag) yee_destroy (h->imag);
  if (h->stream) (void) hipStreamSynchronize (h->stream);
  adjoint_free (h);
  for (void *p : h->grids) (void) hipFree (p);
  if (h->tg) (void) hipFree (h->tg);
  if (h->zcorr) (void) hipFree (h->zcorr);
  if (h->pml) (void) hipFree (h->pml);
  if (h->einc) (void) hipFree (h->einc);
  if (h->hinc[0]) (void) hipFree (h->hinc[0]);
  if (h->hinc[1]) (void) hipFree (h->hinc[1]);
  for (hipEvent_t e : h->ev)
    if (e) (void) hipEventDestroy (e);
  if (h->norm_part) (void) hipFree (h->norm_part);
  if (h->ntff_buf) (void) hipFree (h->ntff_buf);
  if (h->stream) (void) hipStreamDestroy (h->stream);
  delete h;
}

int yee_set_material_f32 (yee_handle *h, int32_t which, const float *values)
{
  int rc = set_material_impl (h, which, values);
  if (!rc && h->imag) rc = set_material_impl (h->imag, which, values);
  return rc;
}

int yee_set_material_sphere (yee_handle *h, int32_t which, double value, double outside, const int64_t center[3],
                             double radius, int32_t accuracy, int32_t stair)
{
  if (!h || !center) return fail (YEE_EINVAL, "null argument");
  if (which < 0 || which > 5) return fail (YEE_EINVAL, "unknown material");
  if (!stair && accuracy <= 0) return fail (YEE_EINVAL, "sphere accuracy must be positive");
  std::vector<double> v ((size_t) h->mcell);
  const int line = line_axis (h->cfg);
  // double material precision: the fine grid, centre 2 c + 0.5 and radius 2 r (Scheme.inc.h:2992-3009)
  const double mod = h->fine () ? 2.0 : 1.0;
  const double cx = (double) center[0] * mod + 0.5, cy = (double) center[1] * mod + 0.5,
               cz = (double) center[2] * mod + 0.5;
  radius *= mod;
  const int64_t nx = h->mn[0], ny = h->mn[1], nz = h->mn[2];
  // the accurate approximation integrates accuracy^2 columns per cell: split x planes over threads
  const unsigned nthr = std::max (1u, std::min (16u, std::thread::hardware_concurrency ()));
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nthr; ++t)
    pool.emplace_back ([&, t] () {
      for (int64_t i = t; i < nx; i += nthr)
        for (int64_t j = 0; j < ny; ++j)
          for (int64_t k = 0; k < nz; ++k)
          {
            // eps-point coordinates (i + 0.5, j + 0.5, k + 0.5)
            const double mx = i + 0.5, my = j + 0.5, mz = k + 0.5;
            double val;
            if (line >= 0)
            {
              // a 1D line: the segment on its one coordinate
              const double m = line == 0 ? mx : (line == 1 ? my : mz), c = line == 0 ? cx : (line == 1 ? cy : cz);
              val = stair ? sphere_stair_1d (m, c, radius, value, outside) : sphere_accurate_1d (m, c, radius, value, outside);
            }
            else if (h->miss >= 0)
            {
              // a 2D scheme on the 3D kernels ((y, z), (x, z), or (x, y) under H_CENTERED): the 2D
              // sphere on its two coordinates
              const double ma = h->miss == 0 ? my : mx, ca = h->miss == 0 ? cy : cx;
              const double mb = h->miss == 2 ? my : mz, cb = h->miss == 2 ? cy : cz;
              val = stair ? sphere_stair_2d (ma, mb, ca, cb, radius, value, outside)
                          : sphere_accurate_2d (ma, mb, ca, cb, radius, value, outside, accuracy);
            }
            else if (h->dim == 3)
              val = stair ? sphere_stair_3d (mx, my, mz, cx, cy, cz, radius, value, outside)
                          : sphere_accurate_3d (mx, my, mz, cx, cy, cz, radius, value, outside, accuracy);
            else
              val = stair ? sphere_stair_2d (mx, my, cx, cy, radius, value, outside)
                          : sphere_accurate_2d (mx, my, cx, cy, radius, value, outside, accuracy);
            v[(size_t) ((i * ny + j) * nz + k)] = val;
          }
    });
  for (std::thread &th : pool) th.join ();
  int rc = set_material_impl (h, which, v.data ());
  if (!rc && h->imag) rc = set_material_impl (h->imag, which, v.data ());
  return rc;
}

int yee_set_material (yee_handle *h, int32_t which, const double *values)
{
  int rc = set_material_impl (h, which, values);
  if (!rc && h->imag) rc = set_material_impl (h->imag, which, values);
  return rc;
}

int yee_prepare (yee_handle *h)
{
  if (!h) return fail (YEE_EINVAL, "null handle");
  if (h->imag)
  {
    int rc = yee_prepare (h->imag);
    if (rc) return rc;
  }
  if (h->prepared) return YEE_OK;
  // the Drude ADE divides by A = 4 M eps + ...: eps / mu must be nonzero (checked in set_material)
  if (h->mode == 2 && !h->box_any)
  {
    for (int a = 0; a < 3; ++a)
    {
      h->box_lo[a] = 0;
      h->box_hi[a] = 0;
    }
  }
  {
    // incident-wave geometry of the TF/SF correction, read by the corrected cells only
    int rc = h->fp32 () ? upload_tfsf_geom<float> (h) : upload_tfsf_geom<double> (h);
    if (rc) return rc;
  }
  if (h->fine ())
  {
    int rc = h->fp32 () ? fine_averages<float> (h) : fine_averages<double> (h);
    if (rc) return rc;
  }
  if (h->use_pml ())
  {
    h->tab_len = (int) std::max (h->n[0], std::max (h->n[1], h->n[2]));
    std::vector<double> tab = pml_tables (h->cfg, h->tab_len);
    int rc = upload (h, &h->pml, tab.data (), tab.size ());
    if (rc) return rc;
  }
  if (h->fn_start)
  {
    int rc = h->fp32 () ? test_start<float> (h) : test_start<double> (h);
    if (rc) return rc;
  }
  HIP_TRY (hipStreamSynchronize (h->stream));
  h->prepared = true;
  return YEE_OK;
}

int yee_step (yee_handle *h, int32_t n)
{
  if (!h) return fail (YEE_EINVAL, "null handle");
  if (!h->prepared) return fail (YEE_ESTATE, "yee_prepare has not been called");
  if (n < 0) return fail (YEE_EINVAL, "negative step count");
  HIP_TRY (hipEventRecord (h->ev[0], h->stream));
  if (h->imag) HIP_TRY (hipStreamWaitEvent (h->imag->stream, h->ev[0], 0));
  for (int32_t s = 0; s < n; ++s)
  {
    int rc = h->fp32 () ? step_once<float> (h) : step_once<double> (h);
    if (rc) return rc;
    // complex fields: the imaginary part on its own stream, alongside
    if (h->imag && (rc = h->fp32 () ? step_once<float> (h->imag) : step_once<double> (h->imag))) return rc;
  }
  if (h->imag)
  {
    // the step time covers both parts
    HIP_TRY (hipEventRecord (h->imag->ev[1], h->imag->stream));
    HIP_TRY (hipStreamWaitEvent (h->stream, h->imag->ev[1], 0));
  }
  HIP_TRY (hipEventRecord (h->ev[1], h->stream));
  HIP_TRY (hipEventSynchronize (h->ev[1]));
  float ms = 0;
  HIP_TRY (hipEventElapsedTime (&ms, h->ev[0], h->ev[1]));
  h->last_ms = n > 0 ? ms / n : 0;
  return YEE_OK;
}

int yee_profile_steps (yee_handle *h, int32_t n, double out[5])
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (!h->prepared) return fail (YEE_ESTATE, "yee_prepare has not been called");
  if (n <= 0) return fail (YEE_EINVAL, "step count must be positive");
  std::vector<hipEvent_t> events;
  std::vector<int> kinds;
  events.reserve ((size_t) n * 6);
  kinds.reserve ((size_t) n * 6);
  h->prof_events = &events;
  h->prof_kinds = &kinds;
  int rc = yee_step (h, n);
  h->prof_events = nullptr;
  h->prof_kinds = nullptr;
  double sum[4] = { 0, 0, 0, 0 };
  int cnt[4] = { 0, 0, 0, 0 };
  for (size_t e = 0; rc == YEE_OK && e + 1 < events.size (); e += 2)
  {
    float ms = 0;
    if (hipEventElapsedTime (&ms, events[e], events[e + 1]) == hipSuccess)
    {
      sum[kinds[e]] += ms;
      cnt[kinds[e]]++;
    }
  }
  for (hipEvent_t e : events) (void) hipEventDestroy (e);
  if (rc) return rc;
  for (int k = 0; k < 4; ++k) out[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
  out[4] = h->last_ms;
  return YEE_OK;
}

int yee_set_point_source_gaussian (yee_handle *h, double t0, double sigma)
{
  if (!h) return fail (YEE_EINVAL, "null handle");
  if (!std::isfinite (t0) || !std::isfinite (sigma)) return fail (YEE_EINVAL, "t0 and sigma must be finite");
  h->src_t0 = t0;
  h->src_sigma = sigma;
  if (h->imag)
  {
    h->imag->src_t0 = t0;
    h->imag->src_sigma = sigma;
  }
  return YEE_OK;
}

int yee_sync (yee_handle *h)
{
  if (!h) return fail (YEE_EINVAL, "null handle");
  HIP_TRY (hipStreamSynchronize (h->stream));
  if (h->imag) HIP_TRY (hipStreamSynchronize (h->imag->stream));
  return YEE_OK;
}

int yee_get_imag_part (yee_handle *h, yee_handle **imag)
{
  if (!h || !imag) return fail (YEE_EINVAL, "null argument");
  if (!h->imag) return fail (YEE_EINVAL, "the handle holds real field values (create it with YEE_COMPLEX_FIELDS)");
  *imag = h->imag;
  return YEE_OK;
}

int yee_time (yee_handle *h, int32_t *t)
{
  if (!h || !t) return fail (YEE_EINVAL, "null argument");
  *t = h->t;
  return YEE_OK;
}

int yee_last_step_ms (yee_handle *h, double *ms)
{
  if (!h || !ms) return fail (YEE_EINVAL, "null argument");
  *ms = h->last_ms;
  return YEE_OK;
}

int yee_get_field_raw (yee_handle *h, int32_t comp, void *out)
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (comp < 0 || comp > 17) return fail (YEE_EINVAL, "unknown component");
  const int base = comp % 6;
  if (!h->exposed[base]) return fail (YEE_EINVAL, "component not present in this scheme");
  void *src = comp < 6 ? h->f[h->cur][base] : (comp < 12 ? h->d[h->cur][base] : h->d1[h->cur][base]);
  if (!src) return fail (YEE_EINVAL, "component not allocated (enable PML / metamaterials)");
  HIP_TRY (hipStreamSynchronize (h->stream));
  HIP_TRY (hipMemcpy (out, src, (size_t) h->ncell * h->wsize, hipMemcpyDeviceToHost));
  return YEE_OK;
}

int yee_get_field (yee_handle *h, int32_t comp, double *out)
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (!h->fp32 ()) return yee_get_field_raw (h, comp, out);
  std::vector<float> tmp ((size_t) h->ncell);
  int rc = yee_get_field_raw (h, comp, tmp.data ());
  if (rc) return rc;
  for (size_t q = 0; q < tmp.size (); ++q) out[q] = tmp[q];
  return YEE_OK;
}

int yee_material_size (yee_handle *h, int64_t out[3])
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  for (int a = 0; a < 3; ++a) out[a] = h->mn[a];
  return YEE_OK;
}

int yee_set_test_functions (yee_handle *h, int32_t start_values, int32_t border, int32_t right_side, int32_t exact,
                            const int64_t norm_start[3], const int64_t norm_end[3])
{
  if (!h) return fail (YEE_EINVAL, "null handle");
  if (h->prepared) return fail (YEE_ESTATE, "test functions must be set before yee_prepare");
  const int32_t kinds[4] = { start_values, border, right_side, exact };
  for (int32_t k : kinds)
    if (k < YEE_FN_NONE || k > YEE_FN_SIN1) return fail (YEE_EINVAL, "unknown test function kind");
  if (right_side == YEE_FN_SIN1) return fail (YEE_EINVAL, "sin1 has no right side (CallBack.cpp)");
  const bool any = start_values || border || right_side || exact;
  if ((start_values || border || right_side) && (h->use_pml () || h->complex () || h->use_meta ()))
    return fail (YEE_ENOTSUP, "start values, border values and right sides run without PML, metamaterials and complex "
                              "fields (the reference's border values then go to its D grids)");
  if (exact && (h->complex () || h->use_meta ()))
    return fail (YEE_ENOTSUP, "the polinom / sin1 exact solutions are real-build callbacks (UNREACHABLE in the "
                              "complex build); metamaterials are not supported with them");
  if (start_values && !border)
    return fail (YEE_ENOTSUP, "start values need border values (the reference's two time layers otherwise alternate "
                              "the boundary cells between the start values and zero)");
  int64_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a)
  {
    const bool act = axis_active (h->cfg, a);
    lo[a] = act ? 1 : 0;
    hi[a] = act ? h->n[a] - 1 : 1;
    if (act && norm_start && norm_start[a] != 0) lo[a] = norm_start[a];
    if (act && norm_end && norm_end[a] != 0) hi[a] = norm_end[a];
    if (lo[a] < 0 || hi[a] > h->n[a] || lo[a] >= hi[a]) return fail (YEE_EINVAL, "norm box out of the grid");
  }
  if (exact && !h->norm_part)
  {
    void *p = nullptr;
    HIP_TRY (hipMalloc (&p, (size_t) 12 * 2 * kNormBlocks * sizeof (double)));
    h->norm_part = (double *) p;
  }
  h->fn_start = start_values;
  h->fn_border = border;
  h->fn_rhs = right_side;
  h->fn_exact = exact;
  for (int a = 0; a < 3; ++a)
  {
    h->norm_lo[a] = lo[a];
    h->norm_hi[a] = hi[a];
  }
  // the fused kernel has no right side: the half-step kernels update the whole grid in place
  if (any) h->fused = false;
  return YEE_OK;
}

int yee_get_diff_norm (yee_handle *h, int32_t comp, double out[4])
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (comp < 0 || comp > 5) return fail (YEE_EINVAL, "component must be E or H (0 .. 5)");
  if (!h->norm_valid[comp]) return fail (YEE_ESTATE, "no diff norm of this component yet (exact solution and a step)");
  std::vector<double> part ((size_t) 2 * kNormBlocks);
  HIP_TRY (hipStreamSynchronize (h->stream));
  HIP_TRY (hipMemcpy (part.data (), h->norm_part + (size_t) comp * 2 * kNormBlocks, part.size () * sizeof (double),
                      hipMemcpyDeviceToHost));
  double sum = 0.0, mx = 0.0;
  for (int b = 0; b < kNormBlocks; ++b)
  {
    sum += part[(size_t) 2 * b];
    mx = std::max (mx, part[(size_t) 2 * b + 1]);
  }
  const double norm = std::sqrt (sum / (double) h->ncell);
  out[0] = norm;
  out[1] = norm * 100.0 / mx;
  out[2] = sum;
  out[3] = mx;
  return YEE_OK;
}

int yee_ntff_snapshot (yee_handle *h)
{
  if (!h) return fail (YEE_EINVAL, "null handle");
  int rc = ntff_check (h);
  if (rc) return rc;
  if (!h->prepared) return fail (YEE_ESTATE, "yee_prepare has not been called");
  if ((rc = ntff_snapshot_part (h)) || (rc = ntff_snapshot_part (h->imag))) return rc;
  return YEE_OK;
}

int yee_ntff (yee_handle *h, double teta, const double *phi, int32_t nphi, const int64_t left[3], const int64_t right[3],
              double *out)
{
  if (!h || !phi || !left || !right || !out) return fail (YEE_EINVAL, "null argument");
  int rc = ntff_check (h);
  if (rc) return rc;
  if (!h->prepared) return fail (YEE_ESTATE, "yee_prepare has not been called");
  if (nphi <= 0 || nphi > 65536) return fail (YEE_EINVAL, "angle count must be in 1 .. 65536");
  for (int a = 0; a < 3; ++a)
    if (left[a] < 2 || right[a] > h->n[a] - 2 || right[a] <= left[a])
      return fail (YEE_EINVAL, "the box must satisfy 2 <= left < right <= size - 2 along every axis");
  for (int a = 0; a < 3; ++a)
  {
    // a face's cell count is a 32-bit loop bound of the face kernel
    const int64_t u = right[(a + 1) % 3] - left[(a + 1) % 3], v = right[(a + 2) % 3] - left[(a + 2) % 3];
    if (u * v >= ((int64_t) 1 << 31)) return fail (YEE_EINVAL, "box face too large");
  }
  if (h->t < 1 || h->ntff_t != h->t - 1 || h->imag->ntff_t != h->t - 1)
    return fail (YEE_ESTATE, "yee_ntff_snapshot must be taken one step before (the N sums average H over the "
                             "newest two steps)");
  const int64_t need = (int64_t) nphi * (1 + 12 * 4);
  if (need > h->ntff_cap)
  {
    if (h->ntff_buf) HIP_TRY (hipFree (h->ntff_buf));
    h->ntff_buf = nullptr;
    h->ntff_cap = 0;
    HIP_TRY (hipMalloc ((void **) &h->ntff_buf, (size_t) need * sizeof (double)));
    h->ntff_cap = need;
  }
  // both parts' state is final once both streams drained
  HIP_TRY (hipStreamSynchronize (h->imag->stream));
  HIP_TRY (hipMemcpyAsync (h->ntff_buf, phi, (size_t) nphi * sizeof (double), hipMemcpyHostToDevice, h->stream));
  double *part = h->ntff_buf + nphi;
  const dim3 grid ((unsigned) nphi * 12);
  if (h->fp32 ())
  {
    NtffArgs<float> a = ntff_args<float> (h, teta, left, right);
    a.phis = h->ntff_buf;
    a.part = part;
    hipLaunchKernelGGL ((ntff_faces<float>), grid, dim3 (256), 0, h->stream, a);
  }
  else
  {
    NtffArgs<double> a = ntff_args<double> (h, teta, left, right);
    a.phis = h->ntff_buf;
    a.part = part;
    hipLaunchKernelGGL ((ntff_faces<double>), grid, dim3 (256), 0, h->stream, a);
  }
  HIP_TRY (hipGetLastError ());
  std::vector<double> host ((size_t) nphi * 48);
  HIP_TRY (hipMemcpyAsync (host.data (), part, host.size () * sizeof (double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY (hipStreamSynchronize (h->stream));
  const double e0 = eps0 (), m0 = mu0 ();
  const double n0 = std::sqrt (m0 / e0);
  const double k = 2 * kPi / h->cfg.wavelength;
  for (int32_t i = 0; i < nphi; ++i)
  {
    // teta / phi sums of N and L, faces in the order ntffN / ntffL add them (x, y, z; left, right)
    double nt[2] = { 0, 0 }, np[2] = { 0, 0 }, lt[2] = { 0, 0 }, lp[2] = { 0, 0 };
    for (int ax = 0; ax < 3; ++ax)
      for (int kind = 0; kind < 2; ++kind)
        for (int side = 0; side < 2; ++side)
        {
          const double *q = host.data () + ((size_t) i * 12 + kind * 6 + ax * 2 + side) * 4;
          double *t = kind ? lt : nt, *f = kind ? lp : np;
          t[0] += q[0]; t[1] += q[1]; f[0] += q[2]; f[1] += q[3];
        }
    // Scheme::Pointing_scat (Scheme.inc.h:4085-4100) over Pointing_inc (:4103-4109)
    const double f1r = lp[0] + nt[0] * n0, f1i = lp[1] + nt[1] * n0;
    const double f2r = lt[0] - np[0] * n0, f2i = lt[1] - np[1] * n0;
    const double scat = (k * k) / (8 * kPi * n0) * ((f1r * f1r + f1i * f1i) + (f2r * f2r + f2i * f2i));
    out[i] = scat / (std::sqrt (e0 / m0) / 2);
  }
  return YEE_OK;
}

int yee_set_exact_exp (yee_handle *h, int32_t kind, int32_t pair, double border, const int64_t norm_start[3],
                       const int64_t norm_end[3])
{
  if (!h) return fail (YEE_EINVAL, "null handle");
  if (h->prepared) return fail (YEE_ESTATE, "test functions must be set before yee_prepare");
  if (kind < YEE_FN_EXP1 || kind > YEE_FN_EXP3) return fail (YEE_EINVAL, "kind must be YEE_FN_EXP1 .. YEE_FN_EXP3");
  if (pair < YEE_SCHEME_EXHY || pair > YEE_SCHEME_EZHY) return fail (YEE_EINVAL, "pair must be YEE_SCHEME_EXHY .. EZHY");
  if (!std::isfinite (border)) return fail (YEE_EINVAL, "border coordinate must be finite");
  if (h->use_meta ()) return fail (YEE_ENOTSUP, "exact solutions with metamaterials are not supported");
  if (h->imag && h->cfg.point_source)
    return fail (YEE_ENOTSUP, "the complex module norm is taken after the step: no point sources with it");
  int rc = set_exact_exp_part (h, kind, pair, border, norm_start, norm_end);
  if (!rc && h->imag) rc = set_exact_exp_part (h->imag, kind, pair, border, norm_start, norm_end);
  return rc;
}

int yee_get_diff_norm_complex (yee_handle *h, int32_t comp, double out[9])
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (!h->imag) return fail (YEE_EINVAL, "the handle holds real field values");
  if (comp < 0 || comp > 5) return fail (YEE_EINVAL, "component must be E or H (0 .. 5)");
  if (!exp_has (h, comp) || !h->norm_valid[comp] || !h->imag->norm_valid[comp])
    return fail (YEE_ESTATE, "no diff norm of this component yet (exp exact solution and a step)");
  if (h->t != h->imag->t) return fail (YEE_ESTATE, "the two parts are at different steps");
  double re[4], im[4];
  int rc = yee_get_diff_norm (h, comp, re);
  if (!rc) rc = yee_get_diff_norm (h->imag, comp, im);
  if (rc) return rc;
  // the module norm from both parts at the last step (calculateFieldStepIterationExact, complex)
  HIP_TRY (hipStreamSynchronize (h->imag->stream));
  const int64_t t = h->t - 1;
  const double time = (comp >= 3 ? (double) t + 1.0 : (double) t + 0.5) * h->sc.dt;
  double *mpart = h->norm_part + (size_t) (6 + comp) * 2 * kNormBlocks;
  if (h->fp32 ())
    hipLaunchKernelGGL ((test_norm<float>), dim3 (kNormBlocks), dim3 (256), 0, h->stream, make_args<float> (h, 0),
                       (const float *) h->f[h->cur][comp], (const float *) h->imag->f[h->imag->cur][comp], comp,
                       h->fn_exact, time, (int) h->norm_lo[0], (int) h->norm_lo[1], (int) h->norm_lo[2],
                       (int) h->norm_hi[0], (int) h->norm_hi[1], (int) h->norm_hi[2], mpart);
  else
    hipLaunchKernelGGL ((test_norm<double>), dim3 (kNormBlocks), dim3 (256), 0, h->stream, make_args<double> (h, 0),
                       (const double *) h->f[h->cur][comp], (const double *) h->imag->f[h->imag->cur][comp], comp,
                       h->fn_exact, time, (int) h->norm_lo[0], (int) h->norm_lo[1], (int) h->norm_lo[2],
                       (int) h->norm_hi[0], (int) h->norm_hi[1], (int) h->norm_hi[2], mpart);
  HIP_TRY (hipGetLastError ());
  std::vector<double> part ((size_t) 2 * kNormBlocks);
  HIP_TRY (hipMemcpyAsync (part.data (), mpart, part.size () * sizeof (double), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY (hipStreamSynchronize (h->stream));
  double sum = 0.0, mx = 0.0;
  for (int b = 0; b < kNormBlocks; ++b)
  {
    sum += part[(size_t) 2 * b];
    mx = std::max (mx, part[(size_t) 2 * b + 1]);
  }
  const double mod = std::sqrt (sum / (double) h->ncell);
  out[0] = re[0];
  out[1] = im[0];
  out[2] = mod;
  out[3] = re[1];
  out[4] = im[1];
  out[5] = mod * 100.0 / mx;
  out[6] = re[3];
  out[7] = im[3];
  out[8] = mx;
  return YEE_OK;
}

int yee_get_material (yee_handle *h, int32_t which, double *out)
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (which < 0 || which > 5) return fail (YEE_EINVAL, "unknown material");
  const size_t n = (size_t) h->mcell;
  HIP_TRY (hipStreamSynchronize (h->stream));
  if (!h->mat[which])
  {
    // a uniform grid is kept as a scalar (eps / mu) or not at all (Drude terms, all zero)
    const double u = which < 2 ? h->mat_u[which] : 0.0;
    const double v = h->fp32 () ? (double) (float) u : u;
    std::fill (out, out + n, v);
    return YEE_OK;
  }
  if (h->fp32 ())
  {
    std::vector<float> tmp (n);
    HIP_TRY (hipMemcpy (tmp.data (), h->mat[which], n * 4, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) out[q] = tmp[q];
  }
  else
  {
    HIP_TRY (hipMemcpy (out, h->mat[which], n * 8, hipMemcpyDeviceToHost));
  }
  return YEE_OK;
}

}  // extern "C"

namespace {

// The scattered field of cells [q0, q0 + count) into dev (FP64), enqueued on the handle's stream.
template <typename T>
int scattered (yee_handle *h, int comp, double *dev, int64_t q0, int64_t count)
{
  const StepArgs<T> a = make_args<T> (h, 0);
  const T *f = (const T *) h->f[h->cur][comp];
  const int hsel = (h->t + 1) & 1;   // buffer of the newest HInc (yee_get_incident)
  const dim3 grid ((unsigned) ((count + 255) / 256)), block (256);
  if (h->dim == 3)
    switch (comp)
    {
      case 0: hipLaunchKernelGGL ((scattered_kernel<T, 0, 3>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      case 1: hipLaunchKernelGGL ((scattered_kernel<T, 1, 3>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      case 2: hipLaunchKernelGGL ((scattered_kernel<T, 2, 3>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      case 3: hipLaunchKernelGGL ((scattered_kernel<T, 3, 3>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      case 4: hipLaunchKernelGGL ((scattered_kernel<T, 4, 3>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      default: hipLaunchKernelGGL ((scattered_kernel<T, 5, 3>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
    }
  else
    switch (comp)
    {
      case 0: hipLaunchKernelGGL ((scattered_kernel<T, 0, 2>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      case 1: hipLaunchKernelGGL ((scattered_kernel<T, 1, 2>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      case 2: hipLaunchKernelGGL ((scattered_kernel<T, 2, 2>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      case 3: hipLaunchKernelGGL ((scattered_kernel<T, 3, 2>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      case 4: hipLaunchKernelGGL ((scattered_kernel<T, 4, 2>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
      default: hipLaunchKernelGGL ((scattered_kernel<T, 5, 2>), grid, block, 0, h->stream, a, f, dev, hsel, q0, count); break;
    }
  HIP_TRY (hipGetLastError ());
  return YEE_OK;
}

int scattered_checks (yee_handle *h, int32_t comp, const double *out)
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (comp < 0 || comp > 5 || !h->exposed[comp]) return fail (YEE_EINVAL, "component not present in this scheme");
  // the reference interpolates EInc / HInc, which exist only with TF/SF
  if (!h->use_tfsf ()) return fail (YEE_EINVAL, "TF/SF is not enabled");
  if (!h->prepared) return fail (YEE_ESTATE, "yee_prepare has not been called");
  return YEE_OK;
}

// Staging array of yee_get_scattered_field_noalloc: a fixed device array, so that the entry
// point allocates nothing.  Each device that runs the library holds its own copy until the
// process ends; host code reaches it only through the address hipGetSymbolAddress returns on
// that device (the array's name is a host-side stand-in, not a device address).  Every call on
// a device shares its copy: calls that overlap on one GPU race.
constexpr int64_t kScratchValues = (int64_t) 4 << 20;   // 4 Mi FP64 values, 32 MiB
static __device__ __align__ (256) double scratch[kScratchValues];

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void *g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address (double **out)
{
  int dev = 0;
  HIP_TRY (hipGetDevice (&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail (YEE_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock (g_scratch_mu);
  if (!g_scratch_addr[dev])
  {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress (&p, HIP_SYMBOL (scratch));
    if (e != hipSuccess || !p)
      return fail (YEE_EDEVICE, std::string ("hipGetSymbolAddress (scratch): ") + hipGetErrorString (e));
    g_scratch_addr[dev] = p;
  }
  *out = (double *) g_scratch_addr[dev];
  return YEE_OK;
}

}  // namespace

extern "C" {

int yee_get_scattered_field (yee_handle *h, int32_t comp, double *out)
{
  int rc = scattered_checks (h, comp, out);
  if (rc) return rc;
  HIP_TRY (hipStreamSynchronize (h->stream));
  double *dev = nullptr;
  HIP_TRY (hipMalloc (&dev, (size_t) h->ncell * sizeof (double)));
  rc = h->fp32 () ? scattered<float> (h, comp, dev, 0, h->ncell) : scattered<double> (h, comp, dev, 0, h->ncell);
  hipError_t e = rc ? hipSuccess : hipStreamSynchronize (h->stream);
  if (!rc && e == hipSuccess) e = hipMemcpy (out, dev, (size_t) h->ncell * sizeof (double), hipMemcpyDeviceToHost);
  (void) hipFree (dev);
  if (rc) return rc;
  if (e != hipSuccess) return fail (YEE_EDEVICE, std::string ("scattered field: ") + hipGetErrorString (e));
  return YEE_OK;
}

int yee_get_scattered_field_noalloc (yee_handle *h, int32_t comp, double *out)
{
  int rc = scattered_checks (h, comp, out);
  if (rc) return rc;
  int dev = 0;
  HIP_TRY (hipGetDevice (&dev));
  if (dev != h->device) return fail (YEE_EINVAL, "call on the device the handle was created on");
  double *s = nullptr;
  if ((rc = scratch_address (&s))) return rc;
  HIP_TRY (hipStreamSynchronize (h->stream));
  // the grid in chunks of the staging array: each chunk is written whole by the kernel before
  // it is copied out, so nothing left from an earlier call is read
  for (int64_t q0 = 0; q0 < h->ncell; q0 += kScratchValues)
  {
    const int64_t cnt = std::min<int64_t> (kScratchValues, h->ncell - q0);
    rc = h->fp32 () ? scattered<float> (h, comp, s, q0, cnt) : scattered<double> (h, comp, s, q0, cnt);
    if (rc) return rc;
    HIP_TRY (hipMemcpyAsync (out + q0, s, (size_t) cnt * sizeof (double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY (hipStreamSynchronize (h->stream));
  }
  return YEE_OK;
}

int yee_get_mask (yee_handle *h, int32_t kind, int32_t comp, uint8_t *out)
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (kind == 0)
  {
    if (comp < 0 || comp > 5 || !h->exposed[comp]) return fail (YEE_EINVAL, "bad component");
    uint8_t *dev = nullptr;
    HIP_TRY (hipMalloc (&dev, (size_t) h->ncell));
    const int grid = (int) std::min<int64_t> (4096, (h->ncell + 255) / 256);
    if (h->fp32 ())
      hipLaunchKernelGGL ((tfsf_mask_kernel<float>), dim3 (grid), dim3 (256), 0, h->stream, make_args<float> (h, 0), comp, h->dim, dev);
    else
      hipLaunchKernelGGL ((tfsf_mask_kernel<double>), dim3 (grid), dim3 (256), 0, h->stream, make_args<double> (h, 0), comp, h->dim, dev);
    hipError_t e = hipGetLastError ();
    if (e == hipSuccess) e = hipStreamSynchronize (h->stream);
    if (e == hipSuccess) e = hipMemcpy (out, dev, (size_t) h->ncell, hipMemcpyDeviceToHost);
    (void) hipFree (dev);
    if (e != hipSuccess) return fail (YEE_EDEVICE, std::string ("tfsf mask: ") + hipGetErrorString (e));
    return YEE_OK;
  }
  if (kind == 1)
  {
    for (int64_t i = 0; i < h->n[0]; ++i)
      for (int64_t j = 0; j < h->n[1]; ++j)
        for (int64_t k = 0; k < h->n[2]; ++k)
        {
          const int64_t p[3] = { i, j, k };
          bool s = false;
          if (h->use_pml ())
            for (int a = 0; a < h->dim; ++a) s |= p[a] < h->cfg.pml[a] || p[a] >= h->n[a] - h->cfg.pml[a] - 1;
          if (h->mode == 2)
          {
            bool in = true;
            for (int a = 0; a < h->dim; ++a) in &= p[a] >= h->box_lo[a] && p[a] < h->box_hi[a];
            s |= in;
          }
          out[(i * h->n[1] + j) * h->n[2] + k] = s ? 1 : 0;
        }
    return YEE_OK;
  }
  return fail (YEE_EINVAL, "unknown mask kind");
}

int yee_get_incident (yee_handle *h, double *einc, double *hinc, int64_t len)
{
  if (!h) return fail (YEE_EINVAL, "null handle");
  if (!h->use_tfsf ()) return fail (YEE_EINVAL, "TF/SF is not enabled");
  if (len < 0 || len > h->inc_len) return fail (YEE_EINVAL, "length out of range");
  HIP_TRY (hipStreamSynchronize (h->stream));
  // newest HInc: written by the last completed step into buffer (t - 1) & 1
  const int newest = (h->t + 1) & 1;
  void *srcs[2] = { h->einc, h->hinc[newest] };
  double *dsts[2] = { einc, hinc };
  for (int s = 0; s < 2; ++s)
  {
    if (!dsts[s]) continue;
    if (h->fp32 ())
    {
      std::vector<float> tmp ((size_t) len);
      HIP_TRY (hipMemcpy (tmp.data (), srcs[s], (size_t) len * 4, hipMemcpyDeviceToHost));
      for (int64_t q = 0; q < len; ++q) dsts[s][q] = tmp[q];
    }
    else
    {
      HIP_TRY (hipMemcpy (dsts[s], srcs[s], (size_t) len * 8, hipMemcpyDeviceToHost));
    }
  }
  return YEE_OK;
}

int yee_get_scalars (yee_handle *h, double out[3])
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  out[0] = h->sc.dt;
  out[1] = h->sc.freq;
  out[2] = h->sc.rel;
  return YEE_OK;
}

}  // extern "C"
