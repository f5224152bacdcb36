// Yee E/H half-step kernels for gfx950 (3D and TMz, E_CENTERED layout).
//
// One thread owns a (y, z) column and marches along x over a chunk of planes, so the
// x-neighbours of the stencil stay in registers; z is the lane-contiguous axis, so every
// plane access of a wavefront is one coalesced 256-byte (FP32) / 512-byte (FP64) row.
//
// Per cell and half-step the kernel applies, in the reference's order of operations:
//   - the curl update                       calcField                 (InternalScheme.inc.h:765-772)
//   - the TF/SF correction of the operands  calculateTFSF             (InternalScheme.inc.h:1386-1835)
//   - in PML slabs: D/B curl update, optional Drude ADE, D->E stage
//                                           calculateFieldStepIteration{,PMLMetamaterials,PML}
//                                                                     (InternalScheme.inc.h:2351-2640)
// Outside the PML slabs and the Drude box the E/H update is the direct one
// (Ca = 1, Cb = dt / (m * m0 * dx), computeCaCb :2340-2344), which equals the
// reference's D->E stage with sigma = 0 up to rounding.
#include <hip/hip_runtime.h>

#include "yee_kernels.h"

namespace yee {

// E_CENTERED component tables (device copies of the host CompDesc plus the circuit offsets
// of YeeGridLayout.cpp:1937-1961).
__constant__ int kAxisD[6] = { 0, 1, 2, 0, 1, 2 };
// slot directions (LEFT=0, RIGHT=1, DOWN=2, UP=3, BACK=4, FRONT=5) of d11, d12, d21, d22
__constant__ int kDirD[6][4] = {
  { 2, 3, 4, 5 },   // Ex: DOWN UP BACK FRONT
  { 4, 5, 0, 1 },   // Ey: BACK FRONT LEFT RIGHT
  { 0, 1, 2, 3 },   // Ez: LEFT RIGHT DOWN UP
  { 4, 5, 2, 3 },   // Hx: BACK FRONT DOWN UP
  { 0, 1, 4, 5 },   // Hy: LEFT RIGHT BACK FRONT
  { 2, 3, 0, 1 },   // Hz: DOWN UP LEFT RIGHT
};
__constant__ int kOppD[6][2] = { { 5, 4 }, { 3, 5 }, { 4, 3 }, { 1, 2 }, { 2, 0 }, { 0, 1 } };
// offsets [slot][axis] of the four opposite-field operands
__constant__ int kOffD[6][4][3] = {
  { { 0, -1, 0 }, { 0, 0, 0 }, { 0, 0, -1 }, { 0, 0, 0 } },
  { { 0, 0, -1 }, { 0, 0, 0 }, { -1, 0, 0 }, { 0, 0, 0 } },
  { { -1, 0, 0 }, { 0, 0, 0 }, { 0, -1, 0 }, { 0, 0, 0 } },
  { { 0, 0, 0 }, { 0, 0, 1 }, { 0, 0, 0 }, { 0, 1, 0 } },
  { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 0, 0 }, { 0, 0, 1 } },
  { { 0, 0, 0 }, { 0, 1, 0 }, { 0, 0, 0 }, { 1, 0, 0 } },
};

// sigma axes (curl, sigma4, sigma5) per component
template <int COMP> struct SigAxes;
template <> struct SigAxes<0> { static constexpr int c = 1, s4 = 0, s5 = 2; };
template <> struct SigAxes<1> { static constexpr int c = 2, s4 = 1, s5 = 0; };
template <> struct SigAxes<2> { static constexpr int c = 0, s4 = 2, s5 = 1; };
template <> struct SigAxes<3> { static constexpr int c = 1, s4 = 0, s5 = 2; };
template <> struct SigAxes<4> { static constexpr int c = 2, s4 = 1, s5 = 0; };
template <> struct SigAxes<5> { static constexpr int c = 0, s4 = 2, s5 = 1; };

__device__ __forceinline__ int twice_coord_d (int comp, int axis, int p)
{
  const bool h = comp >= 3;
  const int own = kAxisD[comp];
  return 2 * p + (h ? (axis == own ? 1 : 2) : (axis == own ? 2 : 1));
}

// doNeedTFSFUpdate<comp>Border(pos, dir) on doubled integer coordinates (YeeGridLayout.cpp:272-1740;
// the FP tests c == b +- 0.5, c in (L - 0.1, R + 0.1), c in (L + 0.4, R - 0.4) are exact on half-integers).
template <typename T>
__device__ bool tfsf_pred (const StepArgs<T> &a, int comp, int dir, const int p[3], int dim)
{
  const int b = dir >> 1;
  const bool low = (dir & 1) == 0;
  const int ax = kAxisD[comp];
  const bool h = comp >= 3;
  if (b == ax || b >= dim) return false;
  const int t = 3 - b - ax;
  const int cb2 = twice_coord_d (comp, b, p[b]);
  bool ok = h ? (low ? cb2 == 2 * a.tl[b] : cb2 == 2 * a.tr[b])
              : (low ? cb2 == 2 * a.tl[b] - 1 : cb2 == 2 * a.tr[b] + 1);
  if (ok && ax < dim)
  {
    const int c2 = twice_coord_d (comp, ax, p[ax]);
    ok = h ? (c2 >= 2 * a.tl[ax] + 1 && c2 <= 2 * a.tr[ax] - 1) : (c2 >= 2 * a.tl[ax] && c2 <= 2 * a.tr[ax]);
  }
  if (ok && t < dim)
  {
    const int c2 = twice_coord_d (comp, t, p[t]);
    ok = h ? (c2 >= 2 * a.tl[t] && c2 <= 2 * a.tr[t]) : (c2 >= 2 * a.tl[t] + 1 && c2 <= 2 * a.tr[t] - 1);
  }
  return ok;
}

// A cell can carry a TF/SF correction only on the planes L-1, R-1, R of some axis while
// lying inside [L-1, R] on every axis.
template <typename T>
__device__ __forceinline__ bool tfsf_near (const StepArgs<T> &a, const int p[3], int dim)
{
  bool face = false;
  for (int ax = 0; ax < dim; ++ax)
  {
    if (p[ax] < a.tl[ax] - 1 || p[ax] > a.tr[ax]) return false;
    face |= (p[ax] == a.tl[ax] - 1) | (p[ax] == a.tr[ax] - 1) | (p[ax] == a.tr[ax]);
  }
  return face;
}

// Incident value of component `opp` at integer position pos: approximateIncidentWave +
// get<opp>FromIncident (InternalScheme.cpp:375-407, InternalScheme.inc.h:500-515).
template <typename T>
__device__ double incident_at (const StepArgs<T> &a, int opp, const int pos[3], int dim, const T *inc, double ddiff)
{
  double r[3];
  for (int ax = 0; ax < 3; ++ax)
  {
    r[ax] = ax < dim ? 0.5 * (double) twice_coord_d (opp, ax, pos[ax]) : 0.0;
  }
  const double x = r[0] - a.zinc[0];
  const double y = r[1] - a.zinc[1];
  double d;
  if (dim == 3)
  {
    const double z = r[2] - a.zinc[2];
    d = x * a.S1 * a.C2 + y * a.S1 * a.S2 + z * a.C1 - ddiff;
  }
  else
  {
    d = x * a.C2 + y * a.S2 - ddiff;
  }
  const double c1 = (double) (long long) d;
  long long i1 = (long long) c1;
  // the reference asserts d > 0; clamp so a degenerate box can never read out of bounds
  if (i1 < 0) i1 = 0;
  if (i1 > a.inc_len - 2) i1 = a.inc_len - 2;
  const double p2 = d - c1;
  const double p1 = 1 - p2;
  const double v = (double) inc[i1] * p1 + (double) inc[i1 + 1] * p2;
  return v * a.proj[opp];
}

// calculateTFSF for one component at cell p: corrects the operands v[0..3] = v11, v12, v21, v22.
template <typename T>
__device__ void tfsf_apply (const StepArgs<T> &a, int comp, const int p[3], int dim, T v[4])
{
  const bool h = comp >= 3;
  const bool u[4] = { tfsf_pred (a, comp, kDirD[comp][0], p, dim), tfsf_pred (a, comp, kDirD[comp][1], p, dim),
                      tfsf_pred (a, comp, kDirD[comp][2], p, dim), tfsf_pred (a, comp, kDirD[comp][3], p, dim) };
  const T *inc = h ? a.einc : a.hinc;
  const double ddiff = h ? 0.0 : 0.5;
  for (int pair = 0; pair < 2; ++pair)
  {
    const bool lo = u[2 * pair], hi = u[2 * pair + 1];
    if (!lo && !hi) continue;
    // E components (E_CENTERED) correct the operand opposite to the border side;
    // H components the operand on the border side, with the sign flipped.
    int slot;
    if (!h) slot = lo ? 2 * pair + 1 : 2 * pair;
    else slot = lo ? 2 * pair : 2 * pair + 1;
    const int opp = kOppD[comp][pair];
    int pos[3];
    for (int ax = 0; ax < 3; ++ax) pos[ax] = p[ax] + (ax < dim ? kOffD[comp][slot][ax] : 0);
    double diff = incident_at (a, opp, pos, dim, inc, ddiff);
    if (h) diff *= -1.0;
    v[slot] = (T) ((double) v[slot] - diff);
  }
}

template <typename T>
__device__ __forceinline__ bool is_special (const StepArgs<T> &a, const int p[3], int dim, int mode)
{
  if (mode == 0) return false;
  bool s = false;
  for (int ax = 0; ax < dim; ++ax) s |= (p[ax] < a.slab_lo[ax]) | (p[ax] >= a.slab_hi[ax]);
  if (mode == 2)
  {
    bool in = true;
    for (int ax = 0; ax < dim; ++ax) in &= (p[ax] >= a.box_lo[ax]) & (p[ax] < a.box_hi[ax]);
    s |= in;
  }
  return s;
}

// Field update of one component at flat index q.
//   tmp  : curl operand combination (o12 - o11 - o22 + o21)
//   mat  : averaged eps (E) or mu (H)
//   om/ga: averaged Drude omega / gamma (MODE 2 special cells)
template <typename T, int MODE, int COMP>
__device__ __forceinline__ void update_cell (const StepArgs<T> &a, int64_t q, const int p[3], T tmp, T mat,
                                             bool special, T om, T ga)
{
  T *F = a.f[COMP];
  const T modc = COMP >= 3 ? a.mu0 : a.eps0;
  const T f_old = F[q];
  if (MODE == 0 || !special)
  {
    // computeCaCb, non-PML branch: Ca = 1, Cb = dt / (material * modifier * dx)
    const T cb = a.dt / (mat * modc * a.dx);
    F[q] = f_old * (T) 1.0 + tmp * cb;
    return;
  }
  const int L = a.tab_len;
  const T *tab = a.pml + (size_t) COMP * 6 * L;
  const T cac = tab[0 * L + p[SigAxes<COMP>::c]];
  const T cbc = tab[1 * L + p[SigAxes<COMP>::c]];
  const T p4 = tab[2 * L + p[SigAxes<COMP>::s4]];
  const T m4 = tab[3 * L + p[SigAxes<COMP>::s4]];
  const T ca5 = tab[4 * L + p[SigAxes<COMP>::s5]];
  const T dd5 = tab[5 * L + p[SigAxes<COMP>::s5]];
  T *D = a.d[COMP];
  const T d_old = D[q];
  // curl update of D/B with the PML Ca/Cb (computeCaCb, PML branch)
  const T d_new = d_old * cac + tmp * cbc;
  if (MODE == 2)
  {
    const T d_pp = a.dp[COMP][q];
    const T m_old = a.d1[COMP][q];
    const T m_pp = a.d1p[COMP][q];
    // Drude ADE coefficients (InternalScheme.inc.h:2513-2518), M = eps0 or mu0
    const T M = modc, dt = a.dt;
    const T wdt2 = (dt * om) * (dt * om);
    const T A = (T) 4 * M * mat + (T) 2 * dt * M * mat * ga + M * wdt2;
    const T b0 = ((T) 4 + (T) 2 * dt * ga) / A;
    const T b1 = (T) -8 / A;
    const T b2 = ((T) 4 - (T) 2 * dt * ga) / A;
    const T a1 = ((T) 2 * M * wdt2 - (T) 8 * M * mat) / A;
    const T a2 = ((T) 4 * M * mat - (T) 2 * dt * M * mat * ga + M * wdt2) / A;
    const T m_new = d_new * b0 + d_old * b1 + d_pp * b2 - m_old * a1 - m_pp * a2;
    // D->E from D1 with modifier 1 (calculateFieldStepIterationPML, useMetamaterials)
    const T one = (T) 1;
    F[q] = f_old * ca5 + m_new * ((p4 / one) / dd5) - m_old * ((m4 / one) / dd5);
    a.dp[COMP][q] = d_old;
    a.d1[COMP][q] = m_new;
    a.d1p[COMP][q] = m_old;
  }
  else
  {
    const T mod = mat * modc;
    F[q] = f_old * ca5 + d_new * ((p4 / mod) / dd5) - d_old * ((m4 / mod) / dd5);
  }
  D[q] = d_new;
}

template <typename T>
__device__ __forceinline__ T curl (const T v[4])
{
  return v[1] - v[0] - v[3] + v[2];
}

template <typename T>
__device__ __forceinline__ T avg2 (T a, T b)
{
  return (a + b) / (T) 2.0;
}

template <typename T>
__device__ __forceinline__ T avg4 (T m11, T m12, T m21, T m22)
{
  return (avg2 (m11, m12) + avg2 (m21, m22)) / (T) 2.0;
}

template <typename T>
__device__ __forceinline__ T sum4 (T m11, T m12, T m21, T m22)
{
  return (m11 + m12 + m21 + m22) / (T) 4.0;
}

// ---------------------------------------------------------------------------------------
// 3D E half-step: Ex, Ey, Ez on [1, N-1)^3
// ---------------------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__ (256) void e_half_3d (StepArgs<T> a)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (j < 1 || j >= a.ny - 1 || k < 1 || k >= a.nz - 1) return;
  const int i0 = 1 + (int) blockIdx.z * a.chunk;
  const int i1 = min (a.nx - 1, i0 + a.chunk);
  if (i0 >= i1) return;
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const T *__restrict__ Hx = a.f[3];
  const T *__restrict__ Hy = a.f[4];
  const T *__restrict__ Hz = a.f[5];
  int64_t q = (int64_t) i0 * sx + (int64_t) j * sy + k;
  T hy_m = Hy[q - sx];
  T hz_m = Hz[q - sx];
  T e_here = a.eps ? a.eps[q] : a.eps_u;
  for (int i = i0; i < i1; ++i, q += sx)
  {
    const T hx = Hx[q], hy = Hy[q], hz = Hz[q];
    const int p[3] = { i, j, k };
    T vx[4] = { Hz[q - sy], hz, Hy[q - 1], hy };
    T vy[4] = { Hx[q - 1], hx, hz_m, hz };
    T vz[4] = { hy_m, hy, Hx[q - sy], hx };
    hy_m = hy;
    hz_m = hz;
    if (a.tfsf && tfsf_near (a, p, 3))
    {
      tfsf_apply (a, 0, p, 3, vx);
      tfsf_apply (a, 1, p, 3, vy);
      tfsf_apply (a, 2, p, 3, vz);
    }
    // materials: 2-cell averages along each component's axis (getMaterial, E_CENTERED)
    T ex_m, ey_m, ez_m, e_next;
    if (a.eps)
    {
      e_next = a.eps[q + sx];
      ex_m = avg2 (e_here, e_next);
      ey_m = avg2 (e_here, a.eps[q + sy]);
      ez_m = avg2 (e_here, a.eps[q + 1]);
    }
    else
    {
      e_next = a.eps_u;
      ex_m = ey_m = ez_m = a.eps_u;
    }
    const bool special = is_special (a, p, 3, MODE);
    T oxm = 0, oym = 0, ozm = 0, gxm = 0, gym = 0, gzm = 0;
    if (MODE == 2 && special)
    {
      // Drude 2-cell averages (getMetaMaterial / approximateDrudeModel); read only in
      // PML and Drude cells
      if (a.ope)
      {
        const T *o = a.ope;
        oxm = avg2 (o[q], o[q + sx]); oym = avg2 (o[q], o[q + sy]); ozm = avg2 (o[q], o[q + 1]);
      }
      if (a.ge)
      {
        const T *g = a.ge;
        gxm = avg2 (g[q], g[q + sx]); gym = avg2 (g[q], g[q + sy]); gzm = avg2 (g[q], g[q + 1]);
      }
    }
    update_cell<T, MODE, 0> (a, q, p, curl (vx), ex_m, special, oxm, gxm);
    update_cell<T, MODE, 1> (a, q, p, curl (vy), ey_m, special, oym, gym);
    update_cell<T, MODE, 2> (a, q, p, curl (vz), ez_m, special, ozm, gzm);
    e_here = e_next;
  }
}

// ---------------------------------------------------------------------------------------
// 3D H half-step: Hx, Hy, Hz on [1, N-1)^3
// ---------------------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__ (256) void h_half_3d (StepArgs<T> a)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (j < 1 || j >= a.ny - 1 || k < 1 || k >= a.nz - 1) return;
  const int i0 = 1 + (int) blockIdx.z * a.chunk;
  const int i1 = min (a.nx - 1, i0 + a.chunk);
  if (i0 >= i1) return;
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const T *__restrict__ Ex = a.f[0];
  const T *__restrict__ Ey = a.f[1];
  const T *__restrict__ Ez = a.f[2];
  int64_t q = (int64_t) i0 * sx + (int64_t) j * sy + k;
  T ey_c = Ey[q];
  T ez_c = Ez[q];
  for (int i = i0; i < i1; ++i, q += sx)
  {
    const T ex = Ex[q];
    const T ey_n = Ey[q + sx], ez_n = Ez[q + sx];
    const int p[3] = { i, j, k };
    T vx[4] = { ey_c, Ey[q + 1], ez_c, Ez[q + sy] };
    T vy[4] = { ez_c, ez_n, ex, Ex[q + 1] };
    T vz[4] = { ex, Ex[q + sy], ey_c, ey_n };
    if (a.tfsf && tfsf_near (a, p, 3))
    {
      tfsf_apply (a, 3, p, 3, vx);
      tfsf_apply (a, 4, p, 3, vy);
      tfsf_apply (a, 5, p, 3, vz);
    }
    ey_c = ey_n;
    ez_c = ez_n;
    // 4-cell transverse averages (getMaterial, E_CENTERED, H components)
    T hx_m, hy_m, hz_m;
    if (a.mu)
    {
      const T *m = a.mu;
      hx_m = avg4 (m[q], m[q + 1], m[q + sy], m[q + sy + 1]);
      hy_m = avg4 (m[q], m[q + 1], m[q + sx], m[q + sx + 1]);
      hz_m = avg4 (m[q], m[q + sy], m[q + sx], m[q + sx + sy]);
    }
    else
    {
      hx_m = hy_m = hz_m = a.mu_u;
    }
    const bool special = is_special (a, p, 3, MODE);
    T oxm = 0, oym = 0, ozm = 0, gxm = 0, gym = 0, gzm = 0;
    if (MODE == 2 && special)
    {
      const T *o = a.opm, *g = a.gm;
      if (o)
      {
        oxm = sum4 (o[q], o[q + 1], o[q + sy], o[q + sy + 1]);
        oym = sum4 (o[q], o[q + 1], o[q + sx], o[q + sx + 1]);
        ozm = sum4 (o[q], o[q + sy], o[q + sx], o[q + sx + sy]);
      }
      if (g)
      {
        gxm = sum4 (g[q], g[q + 1], g[q + sy], g[q + sy + 1]);
        gym = sum4 (g[q], g[q + 1], g[q + sx], g[q + sx + 1]);
        gzm = sum4 (g[q], g[q + sy], g[q + sx], g[q + sx + sy]);
      }
    }
    update_cell<T, MODE, 3> (a, q, p, curl (vx), hx_m, special, oxm, gxm);
    update_cell<T, MODE, 4> (a, q, p, curl (vy), hy_m, special, oym, gym);
    update_cell<T, MODE, 5> (a, q, p, curl (vz), hz_m, special, ozm, gzm);
  }
}

// ---------------------------------------------------------------------------------------
// TMz half-steps (Ez; Hx, Hy) on [1, N-1)^2; one thread per cell, y lane-contiguous
// ---------------------------------------------------------------------------------------
template <typename T, int MODE>
__global__ __launch_bounds__ (256) void e_half_tmz (StepArgs<T> a)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j < 1 || j >= a.ny - 1) return;
  const int64_t sx = a.sx;
  const int sy = a.sy;   // == 1
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy;
  const int p[3] = { i, j, 0 };
  const T *Hx = a.f[3], *Hy = a.f[4];
  T v[4] = { Hy[q - sx], Hy[q], Hx[q - sy], Hx[q] };
  if (a.tfsf && tfsf_near (a, p, 2)) tfsf_apply (a, 2, p, 2, v);
  // Ez sits on the eps point: the 2-cell average reads the same cell twice
  const T e = a.eps ? a.eps[q] : a.eps_u;
  const T mat = avg2 (e, e);
  const bool special = is_special (a, p, 2, MODE);
  T om = 0, ga = 0;
  if (MODE == 2 && special)
  {
    const T o = a.ope ? a.ope[q] : (T) 0, g = a.ge ? a.ge[q] : (T) 0;
    om = avg2 (o, o);
    ga = avg2 (g, g);
  }
  update_cell<T, MODE, 2> (a, q, p, curl (v), mat, special, om, ga);
}

template <typename T, int MODE>
__global__ __launch_bounds__ (256) void h_half_tmz (StepArgs<T> a)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j < 1 || j >= a.ny - 1) return;
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy;
  const int p[3] = { i, j, 0 };
  const T *Ez = a.f[2];
  const T ez = Ez[q];
  // Hx: Ey operands absent (0), Ez at (i, j) and (i, j+1); Hy: Ez at (i, j) and (i+1, j), Ex absent
  T vx[4] = { (T) 0, (T) 0, ez, Ez[q + sy] };
  T vy[4] = { ez, Ez[q + sx], (T) 0, (T) 0 };
  const bool near = a.tfsf && tfsf_near (a, p, 2);
  if (near)
  {
    tfsf_apply (a, 3, p, 2, vx);
    tfsf_apply (a, 4, p, 2, vy);
  }
  T mx, my;
  if (a.mu)
  {
    const T *m = a.mu;
    mx = avg4 (m[q], m[q], m[q + sy], m[q + sy]);
    my = avg4 (m[q], m[q], m[q + sx], m[q + sx]);
  }
  else
  {
    mx = my = a.mu_u;
  }
  const bool special = is_special (a, p, 2, MODE);
  T omx = 0, gmx = 0, omy = 0, gmy = 0;
  if (MODE == 2 && special)
  {
    const T *o = a.opm, *g = a.gm;
    if (o)
    {
      omx = sum4 (o[q], o[q], o[q + sy], o[q + sy]);
      omy = sum4 (o[q], o[q], o[q + sx], o[q + sx]);
    }
    if (g)
    {
      gmx = sum4 (g[q], g[q], g[q + sy], g[q + sy]);
      gmy = sum4 (g[q], g[q], g[q + sx], g[q + sx]);
    }
  }
  update_cell<T, MODE, 3> (a, q, p, curl (vx), mx, special, omx, gmx);
  update_cell<T, MODE, 4> (a, q, p, curl (vy), my, special, omy, gmy);
}

// ---------------------------------------------------------------------------------------
// Incident 1D grid: performPlaneWaveESteps then performPlaneWaveHSteps for step t
// (InternalScheme.inc.h:3276-3357).  One workgroup; only the cells the wave can have
// reached (i <= t + 2) are touched, the rest are exactly zero in the reference too.
// EInc is updated in place; HInc is written to the other buffer so the E half-step of
// this step can still read HInc of the previous step.  The source value sin(dt t 2 pi f)
// is computed on the host with the C library.
// ---------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__ (1024) void incident_step (T *einc, const T *hinc_prev, T *hinc_new, int64_t len,
                                                        int t, double modE, double modH, double src_value)
{
  const int64_t lim = min (len, (int64_t) t + 3);
  for (int64_t i = 1 + threadIdx.x; i < lim; i += blockDim.x)
  {
    einc[i] = einc[i] + (hinc_prev[i - 1] - hinc_prev[i]) * (T) modE;
  }
  if (threadIdx.x == 0)
  {
    einc[0] = (T) src_value;
  }
  __syncthreads ();
  const int64_t limh = min (len - 1, (int64_t) t + 3);
  for (int64_t i = threadIdx.x; i < limh; i += blockDim.x)
  {
    hinc_new[i] = hinc_prev[i] + (einc[i] - einc[i + 1]) * (T) modH;
  }
}

// Hard point source (performPointSourceCalc, InternalScheme.inc.h:2818-2895).  The sine is
// evaluated on the host with the C library, as in the reference, and passed in.
template <typename T>
__global__ void point_source (T *f, int64_t q, double value)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) f[q] = (T) value;
}

// TF/SF predicate mask of one component for every cell (set-up check export).
template <typename T>
__global__ void tfsf_mask_kernel (StepArgs<T> a, int comp, int dim, uint8_t *out)
{
  const int64_t n = (int64_t) a.nx * a.ny * a.nz;
  for (int64_t q = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; q < n; q += (int64_t) gridDim.x * blockDim.x)
  {
    const int k = (int) (q % a.nz);
    const int j = (int) ((q / a.nz) % a.ny);
    const int i = (int) (q / ((int64_t) a.nz * a.ny));
    const int p[3] = { i, j, k };
    uint8_t m = 0;
    for (int d = 0; d < 6; ++d)
      if (tfsf_pred (a, comp, d, p, dim)) m |= (uint8_t) (1u << d);
    out[q] = m;
  }
}

// Explicit instantiations used by the host driver.
#define YEE_INST(T)                                                                        \
  template __global__ void e_half_3d<T, 0> (StepArgs<T>);                                  \
  template __global__ void e_half_3d<T, 1> (StepArgs<T>);                                  \
  template __global__ void e_half_3d<T, 2> (StepArgs<T>);                                  \
  template __global__ void h_half_3d<T, 0> (StepArgs<T>);                                  \
  template __global__ void h_half_3d<T, 1> (StepArgs<T>);                                  \
  template __global__ void h_half_3d<T, 2> (StepArgs<T>);                                  \
  template __global__ void e_half_tmz<T, 0> (StepArgs<T>);                                 \
  template __global__ void e_half_tmz<T, 1> (StepArgs<T>);                                 \
  template __global__ void e_half_tmz<T, 2> (StepArgs<T>);                                 \
  template __global__ void h_half_tmz<T, 0> (StepArgs<T>);                                 \
  template __global__ void h_half_tmz<T, 1> (StepArgs<T>);                                 \
  template __global__ void h_half_tmz<T, 2> (StepArgs<T>);                                 \
  template __global__ void incident_step<T> (T *, const T *, T *, int64_t, int, double, double, double); \
  template __global__ void point_source<T> (T *, int64_t, double);                         \
  template __global__ void tfsf_mask_kernel<T> (StepArgs<T>, int, int, uint8_t *);

YEE_INST (float)
YEE_INST (double)

}  // namespace yee
