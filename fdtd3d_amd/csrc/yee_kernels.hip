// Yee E/H half-step kernels for gfx950 (3D and TMz, E_CENTERED layout).
//
// One thread per cell; z is the lane-contiguous axis, so every access of a wavefront is
// one coalesced 256-byte (FP32) / 512-byte (FP64) row.
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

#include "yee_device.h"

namespace yee {

// D/B state of a special cell before the update: D/B, D/B one step older, D1/B1 and D1/B1 one
// step older (the last three with Drude terms only).
template <typename T>
struct Old
{
  T d, dp, m, mp;
};

// PML coefficients of component COMP at cell p (computeCaCb, PML branch): Ca, Cb of the
// D/B update along the component's axis, the D/B -> field factors along the other two.
template <typename T>
struct PmlCoef
{
  T cac, cbc, p4, m4, ca5, dd5;
};

template <typename T, int COMP>
__device__ __forceinline__ PmlCoef<T> pml_coef (const StepArgs<T> &a, const int p[3])
{
  const int L = a.tab_len;
  const T *tab = a.pml + (size_t) COMP * 6 * L;
  PmlCoef<T> c;
  c.cac = gld (tab + 0 * L + p[SigAxes<COMP>::c]);
  c.cbc = gld (tab + 1 * L + p[SigAxes<COMP>::c]);
  c.p4 = gld (tab + 2 * L + p[SigAxes<COMP>::s4]);
  c.m4 = gld (tab + 3 * L + p[SigAxes<COMP>::s4]);
  c.ca5 = gld (tab + 4 * L + p[SigAxes<COMP>::s5]);
  c.dd5 = gld (tab + 5 * L + p[SigAxes<COMP>::s5]);
  return c;
}

// New value of one field component at flat index q (reference formulas, see file header),
// with the PML coefficients of the cell given (special cells of MODE 1 and 2).
//   f_old : the component's previous value
//   old   : its D/B history (special cells only; loaded by the caller)
//   tmp   : curl operand combination (o12 - o11 - o22 + o21)
//   mat   : averaged eps (E) or mu (H)
//   om/ga : averaged Drude omega / gamma (MODE 2 special cells)
//   store : write the auxiliary D/B (D1/B1) state of this cell
template <typename T, int MODE, int COMP>
__device__ __forceinline__ T cell_value_pc (const StepArgs<T> &a, int64_t q, T f_old, const Old<T> &old, T tmp, T mat,
                                            bool special, T om, T ga, bool store, const PmlCoef<T> &pc,
                                            T jdx = (T) 0)
{
  const T d_old = old.d;
  const T modc = COMP >= 3 ? a.mu0 : a.eps0;
  if (MODE == 0 || !special)
  {
    // computeCaCb, non-PML branch: Ca = 1, Cb = dt / (material * modifier * dx)
    const T cb = direct_cb (a.dt, mat, modc, a.dx);
    // a current source subtracts current * dx * Cb after the curl update (0 elsewhere: exact)
    const T f_new = f_old * (T) 1.0 + tmp * cb;
    return f_new - jdx * cb;
  }
  const T cac = pc.cac, cbc = pc.cbc, p4 = pc.p4, m4 = pc.m4, ca5 = pc.ca5, dd5 = pc.dd5;
  if constexpr (MODE == 2 && sizeof (T) == 4)
  {
    // FP32 Drude cells: the same recursion carried in increments.  The reference's form
    // b0 D + b1 D' + b2 D'' (b0 + b1 + b2 = 0) is a second difference of D, which FP32 values
    // of D (rounded to 6e-8 of D, while the difference is ~(w dt)^2 D) cannot resolve, and the
    // ADE's roots near 1 accumulate that error.  With dD = D - D', dD1 = D1 - D1' the algebra
    // gives (Drude ADE coefficients as below, c = a1 + a2 + 1 = 4 M (dt w_p)^2 / A):
    //   dD_new  = D (cac - 1) + tmp cbc,       D_new  = D + dD_new
    //   dD1_new = b0 dD_new - b2 dD - c D1 + a2 dD1,   D1_new = D1 + dD1_new
    //   f_new   = f ca5 + dD1_new p4 / dd5 + D1 (p4 - m4) / dd5
    // equal to the reference's expressions in real arithmetic.  dp / d1p hold dD / dD1 here
    // (values one step older in the FP64 state).  FP64 keeps the reference's operations.
    const T dn = d_old * (cac - (T) 1) + tmp * cbc - jdx * cbc;
    const T dd_old = old.dp, m_old = old.m, dm_old = old.mp;
    const T M = modc, dt = a.dt;
    const T wdt2 = (dt * om) * (dt * om);
    const T A = (T) 4 * M * mat + (T) 2 * dt * M * mat * ga + M * wdt2;
    const T b0 = ((T) 4 + (T) 2 * dt * ga) / A;
    const T b2 = ((T) 4 - (T) 2 * dt * ga) / A;
    const T a2 = ((T) 4 * M * mat - (T) 2 * dt * M * mat * ga + M * wdt2) / A;
    const T c = (T) 4 * M * wdt2 / A;
    const T dm = dn * b0 - dd_old * b2 - m_old * c + dm_old * a2;
    const T f_new = f_old * ca5 + dm * (p4 / dd5) + m_old * ((p4 - m4) / dd5);
    if (store)
    {
      gst (a.dp_n[COMP] + q, dn);
      gst (a.d1_n[COMP] + q, m_old + dm);
      gst (a.d1p_n[COMP] + q, dm);
      gst (a.d_n[COMP] + q, d_old + dn);
    }
    return f_new;
  }
  // curl update of D/B with the PML Ca/Cb (computeCaCb, PML branch), then the current source
  const T d_cur = d_old * cac + tmp * cbc;
  const T d_new = d_cur - jdx * cbc;
  T f_new;
  if (MODE == 2)
  {
    const T d_pp = old.dp, m_old = old.m, m_pp = old.mp;
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
    f_new = f_old * ca5 + m_new * ((p4 / one) / dd5) - m_old * ((m4 / one) / dd5);
    if (store)
    {
      gst (a.dp_n[COMP] + q, d_old);
      gst (a.d1_n[COMP] + q, m_new);
      gst (a.d1p_n[COMP] + q, m_old);
    }
  }
  else
  {
    const T mod = mat * modc;
    f_new = f_old * ca5 + d_new * ((p4 / mod) / dd5) - d_old * ((m4 / mod) / dd5);
  }
  if (store) gst (a.d_n[COMP] + q, d_new);
  return f_new;
}

// The same, loading the PML coefficients of cell p where they are needed.
template <typename T, int MODE, int COMP>
__device__ __forceinline__ T cell_value (const StepArgs<T> &a, int64_t q, const int p[3], T f_old, const Old<T> &old,
                                         T tmp, T mat, bool special, T om, T ga, bool store)
{
  PmlCoef<T> pc {};
  if (MODE != 0 && special) pc = pml_coef<T, COMP> (a, p);
  // current source J / M of this component at this cell (Scheme.inc.h:1171-1260): current * dx
  const bool cur = ((a.cur_mask >> COMP) & 1u) && p[0] == a.cur_pos[0] && p[1] == a.cur_pos[1] && p[2] == a.cur_pos[2];
  const T jdx = cur ? (T) a.cur_val * a.dx : (T) 0;
  return cell_value_pc<T, MODE, COMP> (a, q, f_old, old, tmp, mat, special, om, ga, store, pc, jdx);
}

// In-place update of one component (half-step kernels).  The old values are loaded by the
// caller before any store of the cell, so every load of a thread is in flight at once (the
// state pointers may alias as far as the compiler knows).
template <typename T, int MODE, int COMP>
__device__ __forceinline__ void update_cell (const StepArgs<T> &a, int64_t q, const int p[3], T f_old,
                                             const Old<T> &old, T tmp, T mat, bool special, T om, T ga)
{
  gst (a.f_n[COMP] + q, cell_value<T, MODE, COMP> (a, q, p, f_old, old, tmp, mat, special, om, ga, true));
}

// Previous value and (special cells) the D/B history of component COMP at q.
template <typename T, int MODE, int COMP>
__device__ __forceinline__ void load_old (const StepArgs<T> &a, int64_t q, bool special, T &f_old, Old<T> &old)
{
  f_old = gld (a.f[COMP] + q);
  old.d = old.dp = old.m = old.mp = (T) 0;
  if (MODE != 0 && special)
  {
    old.d = gld (a.d[COMP] + q);
    if (MODE == 2)
    {
      old.dp = gld (a.dp[COMP] + q);
      old.m = gld (a.d1[COMP] + q);
      old.mp = gld (a.d1p[COMP] + q);
    }
  }
}


// ---------------------------------------------------------------------------------------
// 3D E half-step: Ex, Ey, Ez on the run box (all of [1, N-1)^3, or one PML slab when the
// fused step kernel covers the rest).  One thread per cell, lane = z, so each wavefront
// access is a coalesced row (rows aligned to the tile width); the y/z/x neighbours of the
// stencil are served by L1/L2 (half_cell).  Reads f, writes f_n (the same arrays when
// updating in place).
// ---------------------------------------------------------------------------------------
template <typename T, int MODE, int LAY = 0>
__device__ __forceinline__ void e_half_cell (const StepArgs<T> &a, int i, int j, int k)
{
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const int sz = a.sz;   // 1, or 0 when the scheme lacks z (a 1D line along x or y)
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy + k;
  const int p[3] = { i, j, k };
  const bool special = is_special (a, p, 3, MODE);
  T fo0, fo1, fo2;
  Old<T> do0, do1, do2;
  load_old<T, MODE, 0> (a, q, special, fo0, do0);
  load_old<T, MODE, 1> (a, q, special, fo1, do1);
  load_old<T, MODE, 2> (a, q, special, fo2, do2);
  const T *Hx = a.f[3], *Hy = a.f[4], *Hz = a.f[5];
  const T hx = gld (Hx + q), hy = gld (Hy + q), hz = gld (Hz + q);
  // E_CENTERED: H at {-1, 0} along the derivative axis; H_CENTERED: at {0, +1}
  T vx[4], vy[4], vz[4];
  if constexpr (LAY == 0)
  {
    vx[0] = gld (Hz + q - sy); vx[1] = hz; vx[2] = gld (Hy + q - sz); vx[3] = hy;
    vy[0] = gld (Hx + q - sz); vy[1] = hx; vy[2] = gld (Hz + q - sx); vy[3] = hz;
    vz[0] = gld (Hy + q - sx); vz[1] = hy; vz[2] = gld (Hx + q - sy); vz[3] = hx;
  }
  else
  {
    vx[0] = hz; vx[1] = gld (Hz + q + sy); vx[2] = hy; vx[3] = gld (Hy + q + sz);
    vy[0] = hx; vy[1] = gld (Hx + q + sz); vy[2] = hz; vy[3] = gld (Hz + q + sx);
    vz[0] = hy; vz[1] = gld (Hy + q + sx); vz[2] = hx; vz[3] = gld (Hx + q + sy);
  }
  T ex_m = a.eps_u, ey_m = a.eps_u, ez_m = a.eps_u;
  if (a.matc[0])
  {
    // double material precision: the 8-cell averages, precomputed per component
    ex_m = gld (a.matc[0] + q);
    ey_m = gld (a.matc[1] + q);
    ez_m = gld (a.matc[2] + q);
  }
  else if (a.eps)
  {
    const T *m = a.eps;
    if constexpr (LAY == 0)
    {
      // 2-cell averages along each component's axis (getMaterial, E_CENTERED)
      const T e = gld (m + q);
      ex_m = avg2 (e, gld (m + q + sx));
      ey_m = avg2 (e, gld (m + q + sy));
      ez_m = avg2 (e, gld (m + q + sz));
    }
    else
    {
      // 4-cell transverse averages (getMaterial, H_CENTERED, E components)
      ex_m = avg4 (gld (m + q), gld (m + q + sz), gld (m + q + sy), gld (m + q + sy + sz));
      ey_m = avg4 (gld (m + q), gld (m + q + sz), gld (m + q + sx), gld (m + q + sx + sz));
      ez_m = avg4 (gld (m + q), gld (m + q + sy), gld (m + q + sx), gld (m + q + sx + sy));
    }
  }
  if (a.tfsf && tfsf_near<T, 3> (a, i, j, k))
  {
    tfsf_apply<T, 0, 3, LAY> (a, i, j, k, vx);
    tfsf_apply<T, 1, 3, LAY> (a, i, j, k, vy);
    tfsf_apply<T, 2, 3, LAY> (a, i, j, k, vz);
  }
  T oxm = 0, oym = 0, ozm = 0, gxm = 0, gym = 0, gzm = 0;
  if (MODE == 2 && special)
  {
    // Drude 2-cell averages (getMetaMaterial / approximateDrudeModel)
    if (a.ope)
    {
      const T *o = a.ope;
      oxm = avg2 (gld (o + q), gld (o + q + sx)); oym = avg2 (gld (o + q), gld (o + q + sy));
      ozm = avg2 (gld (o + q), gld (o + q + sz));
    }
    if (a.ge)
    {
      const T *g = a.ge;
      gxm = avg2 (gld (g + q), gld (g + q + sx)); gym = avg2 (gld (g + q), gld (g + q + sy));
      gzm = avg2 (gld (g + q), gld (g + q + sz));
    }
  }
  update_cell<T, MODE, 0> (a, q, p, fo0, do0, curl_rhs<T, 0> (a, vx, p), ex_m, special, oxm, gxm);
  update_cell<T, MODE, 1> (a, q, p, fo1, do1, curl_rhs<T, 1> (a, vy, p), ey_m, special, oym, gym);
  update_cell<T, MODE, 2> (a, q, p, fo2, do2, curl_rhs<T, 2> (a, vz, p), ez_m, special, ozm, gzm);
}

// ---------------------------------------------------------------------------------------
// 3D H half-step: Hx, Hy, Hz on the run box (one thread per cell).  E is read from f[0..2],
// which the host points at the new E when the state is double-buffered.
// ---------------------------------------------------------------------------------------
template <typename T, int MODE, int LAY = 0>
__device__ __forceinline__ void h_half_cell (const StepArgs<T> &a, int i, int j, int k)
{
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const int sz = a.sz;   // 1, or 0 when the scheme lacks z (a 1D line along x or y)
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy + k;
  const int p[3] = { i, j, k };
  const bool special = is_special (a, p, 3, MODE);
  T fo3, fo4, fo5;
  Old<T> do3, do4, do5;
  load_old<T, MODE, 3> (a, q, special, fo3, do3);
  load_old<T, MODE, 4> (a, q, special, fo4, do4);
  load_old<T, MODE, 5> (a, q, special, fo5, do5);
  const T *Ex = a.f[0], *Ey = a.f[1], *Ez = a.f[2];
  const T ex = gld (Ex + q), ey = gld (Ey + q), ez = gld (Ez + q);
  // E_CENTERED: E at {0, +1} along the derivative axis; H_CENTERED: at {-1, 0}
  T vx[4], vy[4], vz[4];
  if constexpr (LAY == 0)
  {
    vx[0] = ey; vx[1] = gld (Ey + q + sz); vx[2] = ez; vx[3] = gld (Ez + q + sy);
    vy[0] = ez; vy[1] = gld (Ez + q + sx); vy[2] = ex; vy[3] = gld (Ex + q + sz);
    vz[0] = ex; vz[1] = gld (Ex + q + sy); vz[2] = ey; vz[3] = gld (Ey + q + sx);
  }
  else
  {
    vx[0] = gld (Ey + q - sz); vx[1] = ey; vx[2] = gld (Ez + q - sy); vx[3] = ez;
    vy[0] = gld (Ez + q - sx); vy[1] = ez; vy[2] = gld (Ex + q - sz); vy[3] = ex;
    vz[0] = gld (Ex + q - sy); vz[1] = ex; vz[2] = gld (Ey + q - sx); vz[3] = ey;
  }
  if (a.tfsf && tfsf_near<T, 3> (a, i, j, k))
  {
    tfsf_apply<T, 3, 3, LAY> (a, i, j, k, vx);
    tfsf_apply<T, 4, 3, LAY> (a, i, j, k, vy);
    tfsf_apply<T, 5, 3, LAY> (a, i, j, k, vz);
  }
  T hx_m = a.mu_u, hy_m = a.mu_u, hz_m = a.mu_u;
  if (a.matc[3])
  {
    hx_m = gld (a.matc[3] + q);
    hy_m = gld (a.matc[4] + q);
    hz_m = gld (a.matc[5] + q);
  }
  else if (a.mu)
  {
    const T *m = a.mu;
    if constexpr (LAY == 0)
    {
      // 4-cell transverse averages (getMaterial, E_CENTERED, H components)
      hx_m = avg4 (gld (m + q), gld (m + q + sz), gld (m + q + sy), gld (m + q + sy + sz));
      hy_m = avg4 (gld (m + q), gld (m + q + sz), gld (m + q + sx), gld (m + q + sx + sz));
      hz_m = avg4 (gld (m + q), gld (m + q + sy), gld (m + q + sx), gld (m + q + sx + sy));
    }
    else
    {
      // 2-cell averages along each component's axis (getMaterial, H_CENTERED, H components)
      const T u = gld (m + q);
      hx_m = avg2 (u, gld (m + q + sx));
      hy_m = avg2 (u, gld (m + q + sy));
      hz_m = avg2 (u, gld (m + q + sz));
    }
  }
  T oxm = 0, oym = 0, ozm = 0, gxm = 0, gym = 0, gzm = 0;
  if (MODE == 2 && special)
  {
    const T *o = a.opm, *g = a.gm;
    if (o)
    {
      oxm = sum4 (gld (o + q), gld (o + q + sz), gld (o + q + sy), gld (o + q + sy + sz));
      oym = sum4 (gld (o + q), gld (o + q + sz), gld (o + q + sx), gld (o + q + sx + sz));
      ozm = sum4 (gld (o + q), gld (o + q + sy), gld (o + q + sx), gld (o + q + sx + sy));
    }
    if (g)
    {
      gxm = sum4 (gld (g + q), gld (g + q + sz), gld (g + q + sy), gld (g + q + sy + sz));
      gym = sum4 (gld (g + q), gld (g + q + sz), gld (g + q + sx), gld (g + q + sx + sz));
      gzm = sum4 (gld (g + q), gld (g + q + sy), gld (g + q + sx), gld (g + q + sx + sy));
    }
  }
  update_cell<T, MODE, 3> (a, q, p, fo3, do3, curl_rhs<T, 3> (a, vx, p), hx_m, special, oxm, gxm);
  update_cell<T, MODE, 4> (a, q, p, fo4, do4, curl_rhs<T, 4> (a, vy, p), hy_m, special, oym, gym);
  update_cell<T, MODE, 5> (a, q, p, fo5, do5, curl_rhs<T, 5> (a, vz, p), hz_m, special, ozm, gzm);
}

// Drude runs (MODE 2): a cell outside the Drude box has omega_p = gamma = 0 in every average,
// where the reference's ADE recursion (b0 + b1 + b2 = 0, a1 = -2, a2 = 1) keeps
// D1 = D / (M eps) exactly in real arithmetic, so its D1 -> field stage equals the PML D -> field
// stage: such a PML cell takes the PML update (MODE 1: D / B state only), as the cells outside
// the PML take the direct one.  Equal to the reference up to rounding, and without the ADE's
// cancellation (whose FP32 rounding the recursion amplifies).
template <typename T, int MODE>
__global__ __launch_bounds__ (256) void e_half_3d (StepArgs<T> a)
{
  int i, j, k;
  if (!half_cell (a, i, j, k)) return;
  const int p[3] = { i, j, k };
  if (MODE == 2 && !in_drude_box (a, p, 3)) e_half_cell<T, 1> (a, i, j, k);
  else e_half_cell<T, MODE> (a, i, j, k);
}

template <typename T, int MODE>
__global__ __launch_bounds__ (256) void h_half_3d (StepArgs<T> a)
{
  int i, j, k;
  if (!half_cell (a, i, j, k)) return;
  const int p[3] = { i, j, k };
  if (MODE == 2 && !in_drude_box (a, p, 3)) h_half_cell<T, 1> (a, i, j, k);
  else h_half_cell<T, MODE> (a, i, j, k);
}

// The H_CENTERED layout (--layout-type 1) on the same half-step cells: plain and PML runs
// (MODE 0 / 1; the host refuses Drude media and double material precision with it).
template <typename T, int MODE>
__global__ __launch_bounds__ (256) void e_half_3d_hc (StepArgs<T> a)
{
  int i, j, k;
  if (!half_cell (a, i, j, k)) return;
  e_half_cell<T, MODE, 1> (a, i, j, k);
}

template <typename T, int MODE>
__global__ __launch_bounds__ (256) void h_half_3d_hc (StepArgs<T> a)
{
  int i, j, k;
  if (!half_cell (a, i, j, k)) return;
  h_half_cell<T, MODE, 1> (a, i, j, k);
}

// ---------------------------------------------------------------------------------------
// Fused 3D time step: E of step t, then H of step t, in one sweep of the cells it owns.
//
// Owned cells: the PML-free box [run_lo, run_hi) minus the cells next to the TF/SF x and y
// faces, the Drude box and the point-source cell (own_mask).  The other cells are updated by
// e_half_3d / h_half_3d on slab and face boxes, the E updates before this kernel and the H
// updates after it; the z faces of the TF/SF box are corrected here from a per-step table.
// State is double-buffered: every old value comes from f, every new one goes to f_n, and the
// new E of cells this kernel does not own is read back from f_n.
//
// Each workgroup owns a full z row (one lane per z cell, up to 16 waves) of a strip of BY
// y rows and marches along x over a chunk of planes.  Per x plane it computes E one plane
// ahead of H: E at plane x needs H before the step at x and x-1 (x-1 carried in registers);
// H at plane x-1 needs E after the step at x-1 (carried) and x (just computed).  The loads of
// plane x+1 are issued before plane x is computed, so their latency overlaps the arithmetic,
// the stores and the barrier.  The y neighbour row j0+BY is recomputed (Ex, Ez only; not
// stored); the z+1 neighbours of a lane (new Ex, Ey; eps) come from the next lane by a DPP
// wave shift and, for the last lane of a wave, from lane 0 of the next wave through LDS.  Every
// state word crosses HBM once per step, except the recomputed row and the first plane of each
// chunk.
// ---------------------------------------------------------------------------------------

// Ownership bits of one coordinate of a cell (axis_bits): inside the run box, inside
// [L-1, R] of the TF/SF box (with TF/SF on), on a face plane L-1, R-1 or R of the TF/SF box
// (x and y only), inside the Drude box, on the point source (with a source on).
constexpr uint32_t kAxRun = 1, kAxFRange = 2, kAxFPlane = 4, kAxDrude = 8, kAxSrc = 16, kAxCur = 32;

// Thread geometry of the fused sweep.  A cell is updated by the fused kernel when it lies in
// the run box, is not next to an x or y face of the TF/SF box, not in the Drude box and not the
// point-source cell (the half-step kernels update those).  The test is split by axis: plane bits (uniform per x),
// row bits (uniform per strip row, packed 8 bits a row in rowb) and lane bits (zb, per thread),
// so that it costs a few scalar operations and one vector compare per row.
struct FusedPos
{
  int j0, k, kc;
  int zf;            // z face of the TF/SF box the lane lies on (zface_of), -1 if none
  uint32_t zb;       // axis_bits of k
  uint32_t rowb;     // axis_bits of rows j0 .. j0+BY, 8 bits each
  bool wave_box;     // every lane of the wave inside the run box along z
  bool wave_zface;   // a lane of the wave on a z face of the TF/SF box
};

// Direct E/H update (computeCaCb non-PML branch, calcField; the expression of cell_value) with
// the coefficient cb = direct_cb (dt, m, m0, dx) already computed.
template <typename T>
__device__ __forceinline__ T direct_value_cb (T f_old, T tmp, T cb)
{
  return f_old * (T) 1.0 + tmp * cb;
}

template <typename T>
__device__ __forceinline__ uint32_t axis_bits (const StepArgs<T> &a, int ax, int p)
{
  uint32_t b = (p >= a.run_lo[ax] && p < a.run_hi[ax]) ? kAxRun : 0u;
  b |= (a.tfsf && p >= a.tl[ax] - 1 && p <= a.tr[ax]) ? kAxFRange : 0u;
  b |= (ax < 2 && (p == a.tl[ax] - 1 || p == a.tr[ax] - 1 || p == a.tr[ax])) ? kAxFPlane : 0u;
  b |= (p >= a.box_lo[ax] && p < a.box_hi[ax]) ? kAxDrude : 0u;
  b |= (a.src_mask && p == a.src_pos[ax]) ? kAxSrc : 0u;
  b |= (a.cur_mask && p == a.cur_pos[ax]) ? kAxCur : 0u;
  return b;
}

// The z bits a lane must show, masked with the result, for the fused kernel to own cell
// (x, j, k) of plane bits xb and row bits yb: (zb & m) == kAxRun.  m = 0 owns no lane; m = kAxRun
// owns every lane inside the run box.  The point-source cell is left to the half-step kernels
// (the source overwrites it after their update).
__device__ __forceinline__ uint32_t own_mask (uint32_t xb, uint32_t yb)
{
  // branch-free: kAxFRange when inside the TF/SF range on x and y and on an x or y face plane
  // (kAxFPlane shifted onto kAxFRange), kAxDrude when inside the Drude box on x and y
  const uint32_t u = xb & yb;
  const uint32_t face = u & ((xb | yb) >> 1) & kAxFRange;
  return (0u - (u & kAxRun)) & (kAxRun | face | (u & (kAxDrude | kAxSrc | kAxCur)));
}

__device__ __forceinline__ bool lane_owns (uint32_t zb, uint32_t m)
{
  return (zb & m) == kAxRun;
}

// z face (0: k = L-1, 1: R-1, 2: R) of a lane, or -1.
template <typename T>
__device__ __forceinline__ int zface_of (const StepArgs<T> &a, int k)
{
  return k == a.tl[2] - 1 ? 0 : (k == a.tr[2] - 1 ? 1 : (k == a.tr[2] ? 2 : -1));
}

// One value of the z-face table for the LDS staging of the fused kernel: thread t of the
// first 24 (BY + 1) loads entry (part, face, row, slot) of plane xe (E part) or xe - 1 (H part)
// of the strip's rows; 0 outside the table.
template <typename T, int BY>
__device__ __forceinline__ double zstage_load (const StepArgs<T> &a, int t, int xe, int j0)
{
  constexpr int per_part = 3 * (BY + 1) * 4;
  if (!a.tfsf || t >= 2 * per_part) return 0.0;
  const int part = t / per_part, rem = t % per_part;
  const int f = rem / (4 * (BY + 1)), row = (rem / 4) % (BY + 1), slot = rem % 4;
  const int x = part == 0 ? xe : xe - 1, j = j0 + row;
  if (x < a.tl[0] - 1 || x > a.tr[0] || j < a.tl[1] - 1 || j > a.tr[1]) return 0.0;
  return a.zcorr[((((int64_t) part * 3 + f) * a.zc_n[0] + (x - (a.tl[0] - 1))) * a.zc_n[1] + (j - (a.tl[1] - 1))) * 4 + slot];
}

// Subtracts a tabulated correction pair (FP64) from two operands: (T) ((double) v - d), which is
// the identity for an uncorrected slot (d = 0).
template <typename T>
__device__ __forceinline__ void zface_apply (const double *c, T &v_lo, T &v_hi)
{
  v_lo = (T) ((double) v_lo - c[0]);
  v_hi = (T) ((double) v_hi - c[1]);
}

// Per-step table of the TF/SF corrections on the z faces of the TF/SF box: part 0 for the E
// update (Ex pair BACK/FRONT, Ey pair BACK/FRONT), part 1 for the H update (Hx, Hy likewise).
// The same predicates and incident interpolation as calculateTFSF (tfsf_pair_diff).
template <typename T>
__global__ void tfsf_ztable (StepArgs<T> a, double *out)
{
  const int64_t u = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t) a.zc_n[0] * a.zc_n[1];
  if (u >= 3 * per) return;
  const int f = (int) (u / per);
  const int ix = (int) ((u % per) / a.zc_n[1]), iy = (int) (u % a.zc_n[1]);
  const int x = a.tl[0] - 1 + ix, j = a.tl[1] - 1 + iy;
  const int k = f == 0 ? a.tl[2] - 1 : (f == 1 ? a.tr[2] - 1 : a.tr[2]);
  double *e = out + (((int64_t) 0 * 3 + f) * per + (int64_t) ix * a.zc_n[1] + iy) * 4;
  double *h = out + (((int64_t) 1 * 3 + f) * per + (int64_t) ix * a.zc_n[1] + iy) * 4;
  tfsf_pair_diff<T, 0, 1, 3> (a, x, j, k, e[0], e[1]);   // Ex: Hy operands (BACK, FRONT)
  tfsf_pair_diff<T, 1, 0, 3> (a, x, j, k, e[2], e[3]);   // Ey: Hx operands (BACK, FRONT)
  tfsf_pair_diff<T, 3, 0, 3> (a, x, j, k, h[0], h[1]);   // Hx: Ey operands (BACK, FRONT)
  tfsf_pair_diff<T, 4, 1, 3> (a, x, j, k, h[2], h[3]);   // Hy: Ex operands (BACK, FRONT)
}

// Lane byte offsets of the rows j0-1 .. j0+BY relative to row j0-1 (kept opaque, so that all
// arrays of a plane share them and each array needs one scalar base per plane).  Rows past
// the grid edge read the allocation padding; their values only reach cells that are not
// updated.
template <typename T, int BY>
__device__ __forceinline__ void row_offsets (const StepArgs<T> &a, const FusedPos &g, uint32_t (&ro)[BY + 2])
{
#pragma unroll
  for (int r = 0; r < BY + 2; ++r)
  {
    ro[r] = (uint32_t) g.kc * (uint32_t) sizeof (T) + (uint32_t) (r * a.sy) * (uint32_t) sizeof (T);
    asm volatile ("" : "+v"(ro[r]));
  }
}

// Base of plane x, row j0-1 of array `arr` (uniform).
template <typename T>
__device__ __forceinline__ const T *plane_row (const StepArgs<T> &a, const T *arr, int x, const FusedPos &g)
{
  const T *p = arr + (int64_t) x * a.sx + (int64_t) (g.j0 - 1) * a.sy;
  asm volatile ("" : "+s"(p));   // a scalar base: row accesses add only the 32-bit lane offset
  return p;
}

// Material arrays the fused kernel reads (template MAT): eps per cell, mu per cell; a uniform
// material (eps_u, mu_u) otherwise.
constexpr int kMatEps = 1, kMatMu = 2;

// A copy of the lane offsets that the compiler cannot relate to the originals, taken inside
// the block that uses it: each access then keeps zext (offset) beside its add and is selected
// as scalar base + 32-bit lane offset (the loop-invariant originals would otherwise be widened
// once, outside the loop, into 64-bit vector addresses).
template <int N>
__device__ __forceinline__ void fresh_offsets (const uint32_t (&ro)[N], uint32_t (&o)[N])
{
#pragma unroll
  for (int r = 0; r < N; ++r)
  {
    o[r] = ro[r];
    asm volatile ("" : "+v"(o[r]));
  }
}

// Everything the E update of plane x reads.
template <typename T, int BY>
struct PlaneLoads
{
  T hxr[BY + 2], hzr[BY + 2];          // H before the step, rows -1..BY
  T hy[BY + 1], hym[BY + 1], hxm[BY + 1];   // Hy rows 0..BY; z-1 neighbours of Hy, Hx
  T ex[BY + 1], ey[BY + 1], ez[BY + 1];     // owned cells: E before the step; others: E after it
  T ep1[BY + 1];                            // eps at plane x+1
  uint32_t ownb;                            // bit r: the kernel owns the lane's cell of row r
  uint32_t xb;                              // axis_bits of x
};

template <typename T, int BY, int MAT>
__device__ __forceinline__ void e_load (const StepArgs<T> &a, int x, const FusedPos &g, const uint32_t (&ro)[BY + 2],
                                        PlaneLoads<T, BY> &L)
{
  uint32_t o[BY + 2];
  fresh_offsets (ro, o);
  const T *Hx = plane_row (a, a.f[3], x, g), *Hy = plane_row (a, a.f[4], x, g), *Hz = plane_row (a, a.f[5], x, g);
  const T *Ex = plane_row (a, a.f[0], x, g), *Ey = plane_row (a, a.f[1], x, g), *Ez = plane_row (a, a.f[2], x, g);
#pragma unroll
  for (int r = -1; r <= BY; ++r)
  {
    L.hxr[r + 1] = rld (Hx, o[r + 1]);
    L.hzr[r + 1] = rld (Hz, o[r + 1]);
  }
  const uint32_t xb = axis_bits (a, 0, x);
  L.xb = xb;
  uint32_t ownb = 0;
  bool any_new = false;   // uniform: a lane of the wave reads E after the step for some row
#pragma unroll
  for (int r = 0; r <= BY; ++r)
  {
    const uint32_t off = o[r + 1];
    L.hy[r] = rld (Hy, off);
    L.hym[r] = rld (Hy, off, -1);
    if (r < BY) L.hxm[r] = rld (Hx, off, -1);
    L.ex[r] = rld (Ex, off);
    L.ez[r] = rld (Ez, off);
    if (r < BY) L.ey[r] = rld (Ey, off);
    const uint32_t m = own_mask (xb, g.rowb >> (8 * r));
    ownb |= lane_owns (g.zb, m) ? 1u << r : 0u;
    any_new |= (m != kAxRun) | !g.wave_box;
  }
  if (MAT & kMatEps)
  {
    const T *E1 = plane_row (a, a.eps, min (x + 1, a.nx - 1), g);
#pragma unroll
    for (int r = 0; r <= BY; ++r)
    {
      L.ep1[r] = rld (E1, o[r + 1]);
    }
  }
  else
  {
#pragma unroll
    for (int r = 0; r <= BY; ++r) L.ep1[r] = a.eps_u;
  }
  asm volatile ("" : "+v"(ownb));   // kept as lane bits, not as scalar lane masks
  L.ownb = ownb;
  if (any_new)
  {
    // cells the kernel does not own take E after the step, which the half-step kernels wrote
    // to f_n before this kernel
    const T *Nx = plane_row (a, (const T *) a.f_n[0], x, g), *Ny = plane_row (a, (const T *) a.f_n[1], x, g);
    const T *Nz = plane_row (a, (const T *) a.f_n[2], x, g);
#pragma unroll
    for (int r = 0; r <= BY; ++r)
    {
      if ((ownb >> r) & 1u) continue;
      uint32_t off = ro[r + 1];
      asm volatile ("" : "+v"(off));
      L.ex[r] = rld (Nx, off);
      L.ez[r] = rld (Nz, off);
      if (r < BY) L.ey[r] = rld (Ny, off);
    }
  }
}

// State carried from one x plane to the next: E after the step and H before it at plane i,
// eps at plane i+1, the plane bits of i.
template <typename T, int BY>
struct Carry
{
  T ex[BY + 1], ey[BY + 1], ez[BY + 1];   // E rows 0..BY (Ey: rows 0..BY-1)
  T hx[BY + 1], hy[BY + 1], hz[BY + 1];   // H rows 0..BY-1 (Hy: 0..BY)
  T ep1[BY + 1];                          // eps at plane i+1
  uint32_t ownb;                          // E ownership bits of plane i
  uint32_t xb;                            // axis_bits of i
};

// E update of plane x (rows 0..BY) from its loads L, the H / eps of plane x-1 carried in
// `prev` and eps of plane x at z+1 (epk); the result becomes the carried state of plane x.
template <typename T, int BY>
__device__ __forceinline__ void e_compute (const StepArgs<T> &a, int x, const FusedPos &g, const PlaneLoads<T, BY> &L,
                                           const Carry<T, BY> &prev, const T (&epk)[BY + 1],
                                           const double (*zl)[BY + 1][4], Carry<T, BY> &o)
{
  const uint32_t xb = L.xb;
  // z-face corrections: only waves holding a face lane, on planes inside the TF/SF box
  const bool zx = g.wave_zface && (xb & kAxFRange);
#pragma unroll
  for (int r = 0; r <= BY; ++r)
  {
    const uint32_t yb = g.rowb >> (8 * r);
    T vx[4] = { L.hzr[r], L.hzr[r + 1], L.hym[r], L.hy[r] };
    const T vz[4] = { prev.hy[r], L.hy[r], L.hxr[r], L.hxr[r + 1] };
    T vy[4] = { L.hxm[r], L.hxr[r + 1], prev.hz[r], L.hzr[r + 1] };
    const T e = prev.ep1[r];
    const bool own = (L.ownb >> r) & 1u;
    if (zx && (yb & kAxFRange) && own && g.zf >= 0)
    {
      // z-face TF/SF corrections, staged in LDS (Ex: Hy operands, Ey: Hx operands)
      zface_apply (&zl[g.zf][r][0], vx[2], vx[3]);
      zface_apply (&zl[g.zf][r][2], vy[0], vy[1]);
    }
    // cb = dt / (m eps0 dx), m the 2-cell average (a uniform eps averages to itself)
    const T cbx = direct_cb (a.dt, avg2 (e, L.ep1[r]), a.eps0, a.dx);
    const T cbz = direct_cb (a.dt, avg2 (e, epk[r]), a.eps0, a.dx);
    const T cby = r < BY ? direct_cb (a.dt, avg2 (e, prev.ep1[r + 1]), a.eps0, a.dx) : (T) 0;
    const T nx = direct_value_cb (L.ex[r], curl (vx), cbx);
    const T nz = direct_value_cb (L.ez[r], curl (vz), cbz);
    o.ex[r] = own ? nx : L.ex[r];
    o.ez[r] = own ? nz : L.ez[r];
    if (r < BY)
    {
      const T ny = direct_value_cb (L.ey[r], curl (vy), cby);
      o.ey[r] = own ? ny : L.ey[r];
      o.hx[r] = L.hxr[r + 1];
      o.hz[r] = L.hzr[r + 1];
    }
    o.hy[r] = L.hy[r];
    o.ep1[r] = L.ep1[r];
  }
  o.ownb = L.ownb;
  o.xb = xb;
}

// Stores the owned cells of rows 0..BY-1 of an E plane.
template <typename T, int BY>
__device__ __forceinline__ void e_store (const StepArgs<T> &a, int x, const FusedPos &g, const uint32_t (&ro)[BY + 2],
                                         const Carry<T, BY> &o)
{
  T *Ex = (T *) plane_row (a, (const T *) a.f_n[0], x, g);
  T *Ey = (T *) plane_row (a, (const T *) a.f_n[1], x, g);
  T *Ez = (T *) plane_row (a, (const T *) a.f_n[2], x, g);
#pragma unroll
  for (int r = 0; r < BY; ++r)
  {
    if (!((o.ownb >> r) & 1u)) continue;
    uint32_t off = ro[r + 1];
    asm volatile ("" : "+v"(off));
    rst (Ex, off, o.ex[r]);
    rst (Ey, off, o.ey[r]);
    rst (Ez, off, o.ez[r]);
  }
}

// LDS of the fused kernel: Ex, Ey after the step and eps of the next plane at lane 0 of every
// wave (the z+1 neighbours of lane 63 of the wave before; by plane parity), and the z-face
// TF/SF corrections of the strip ([parity][part][face][row][slot]).
template <typename T, int BY>
struct FusedLds
{
  T xch[2][16][BY][2];
  T epx[2][16][BY + 1];
  double zl[2][2][3][BY + 1][4];
};

// One plane of the sweep: E of plane x (from L and the state `c` of plane i = x-1), H of
// plane i, the stores, and the prefetch of plane x+1 into `nxt`.  PAR is the parity of the
// LDS slots of this plane.
template <typename T, int BY, int MAT>
__device__ __forceinline__ void fused_plane (const StepArgs<T> &a, const FusedPos &g, const uint32_t (&ro)[BY + 2],
                                             int x, int xend, int PAR, int lane, int wave, int nw, FusedLds<T, BY> &s,
                                             const Carry<T, BY> &c, const PlaneLoads<T, BY> &L, Carry<T, BY> &o,
                                             PlaneLoads<T, BY> &nxt)
{
  constexpr int kZStage = 2 * 3 * (BY + 1) * 4;   // z-face table entries per parity
  const int i = x - 1;                            // H plane
  e_load<T, BY, MAT> (a, min (x + 1, xend), g, ro, nxt);   // prefetch the next plane
  // z-face corrections of the next plane (written to LDS before the barrier)
  const double zpre = zstage_load<T, BY> (a, g.k, min (x + 1, xend), g.j0);
  // z+1 neighbours from the next lane (DPP shift); lane 63 takes lane 0 of the next wave,
  // which that wave left in LDS before the last barrier
  const bool edge = lane == 63 && wave + 1 < nw;
  T epk[BY + 1];
#pragma unroll
  for (int r = 0; r <= BY; ++r) epk[r] = lane_next (c.ep1[r], edge ? s.epx[PAR ^ 1][wave + 1][r] : c.ep1[r]);
  e_compute<T, BY> (a, x, g, L, c, epk, s.zl[PAR][0], o);
  if (lane == 0)
  {
#pragma unroll
    for (int r = 0; r < BY; ++r)
    {
      s.xch[PAR][wave][r][0] = o.ex[r];
      s.xch[PAR][wave][r][1] = o.ey[r];
    }
#pragma unroll
    for (int r = 0; r <= BY; ++r) s.epx[PAR][wave][r] = o.ep1[r];
  }
  // z+1 neighbours of Ex, Ey after the step at plane i
  T exk[BY + 1], eyk[BY + 1];
#pragma unroll
  for (int r = 0; r < BY; ++r)
  {
    exk[r] = lane_next (c.ex[r], edge ? s.xch[PAR ^ 1][wave + 1][r][0] : c.ex[r]);
    eyk[r] = lane_next (c.ey[r], edge ? s.xch[PAR ^ 1][wave + 1][r][1] : c.ey[r]);
  }
  // ---- H of plane i, rows 0..BY-1
  const uint32_t ib = c.xb;
  T hn[3][BY + 1];
  T mxm[BY + 1], mym[BY + 1], mzm[BY + 1];
  uint32_t hownb = 0;
  if (MAT & kMatMu)
  {
    uint32_t om[BY + 2];
    fresh_offsets (ro, om);
    const T *M0 = plane_row (a, a.mu, i, g), *M1 = plane_row (a, a.mu, i + 1, g);   // i + 1 <= nx - 2
#pragma unroll
    for (int r = 0; r < BY; ++r)
    {
      // 4-cell transverse averages (getMaterial, E_CENTERED, H components)
      const uint32_t o0 = om[r + 1], o1 = om[r + 2];
      mxm[r] = avg4 (rld (M0, o0), rld (M0, o0, 1), rld (M0, o1), rld (M0, o1, 1));
      mym[r] = avg4 (rld (M0, o0), rld (M0, o0, 1), rld (M1, o0), rld (M1, o0, 1));
      mzm[r] = avg4 (rld (M0, o0), rld (M0, o1), rld (M1, o0), rld (M1, o1));
    }
  }
  else
  {
#pragma unroll
    for (int r = 0; r < BY; ++r) mxm[r] = mym[r] = mzm[r] = a.mu_u;
  }
  const bool zx = g.wave_zface && (ib & kAxFRange);
#pragma unroll
  for (int r = 0; r < BY; ++r)
  {
    const uint32_t yb = g.rowb >> (8 * r);
    const bool own = lane_owns (g.zb, own_mask (ib, yb));
    hownb |= own ? 1u << r : 0u;
    T vx[4] = { c.ey[r], eyk[r], c.ez[r], c.ez[r + 1] };
    T vy[4] = { c.ez[r], o.ez[r], c.ex[r], exk[r] };
    const T vz[4] = { c.ex[r], c.ex[r + 1], c.ey[r], o.ey[r] };
    if (zx && (yb & kAxFRange) && own && g.zf >= 0)
    {
      // z-face TF/SF corrections (Hx: Ey operands, Hy: Ex operands)
      zface_apply (&s.zl[PAR][1][g.zf][r][0], vx[0], vx[1]);
      zface_apply (&s.zl[PAR][1][g.zf][r][2], vy[2], vy[3]);
    }
    // cb = dt / (m mu0 dx), m the 4-cell average (or the uniform mu)
    const T cbx = direct_cb (a.dt, mxm[r], a.mu0, a.dx);
    const T cby = direct_cb (a.dt, mym[r], a.mu0, a.dx);
    const T cbz = direct_cb (a.dt, mzm[r], a.mu0, a.dx);
    hn[0][r] = direct_value_cb (c.hx[r], curl (vx), cbx);
    hn[1][r] = direct_value_cb (c.hy[r], curl (vy), cby);
    hn[2][r] = direct_value_cb (c.hz[r], curl (vz), cbz);
  }
  // ---- stores: E of plane x when this chunk owns it, H of plane i (owned cells only)
  if (x < xend) e_store<T, BY> (a, x, g, ro, o);
  {
    T *Hx = (T *) plane_row (a, (const T *) a.f_n[3], i, g);
    T *Hy = (T *) plane_row (a, (const T *) a.f_n[4], i, g);
    T *Hz = (T *) plane_row (a, (const T *) a.f_n[5], i, g);
#pragma unroll
    for (int r = 0; r < BY; ++r)
    {
      if (!((hownb >> r) & 1u)) continue;
      uint32_t off = ro[r + 1];
      asm volatile ("" : "+v"(off));
      rst (Hx, off, hn[0][r]);
      rst (Hy, off, hn[1][r]);
      rst (Hz, off, hn[2][r]);
    }
  }
  if (a.tfsf)
  {
    if (g.k < kZStage) (&s.zl[PAR ^ 1][0][0][0][0])[g.k] = zpre;
    // a 64-thread block (short z rows) loads the entries past its width here
    for (int t = g.k + (int) blockDim.x; t < kZStage; t += (int) blockDim.x)
      (&s.zl[PAR ^ 1][0][0][0][0])[t] = zstage_load<T, BY> (a, t, min (x + 1, xend), g.j0);
  }
  __syncthreads ();
}

// MAXW: the most waves of a workgroup (rows of up to 64 MAXW cells); 12 lets the compiler use
// up to 170 VGPRs per lane, where 16 waves allow 128 (FP64 state would spill).
template <typename T, int BY, int MAT, int MAXW>
__global__ __launch_bounds__ (64 * MAXW) void step_3d_fused (StepArgs<T> a)
{
  __shared__ FusedLds<T, BY> s;
  constexpr int kZStage = 2 * 3 * (BY + 1) * 4;   // z-face table entries per parity

  FusedPos g;
  g.k = (int) threadIdx.x;
  const int lane = g.k & 63;
  const int wave = g.k >> 6;
  const int nw = (int) (blockDim.x >> 6);
  // block -> (chunk, strip): blocks are dispatched round-robin over the 8 XCDs; with
  // nstrips % 8 == 0 each XCD gets a contiguous band of strips, so the recomputed neighbour
  // row of a strip is mostly read from the same XCD's L2
  const int b = (int) blockIdx.x;
  const int chunk = b / a.nstrips;
  const int rr = b - chunk * a.nstrips;
  const int strip = (a.nstrips % 8 == 0) ? (rr % 8) * (a.nstrips / 8) + rr / 8 : rr;
  g.j0 = a.run_lo[1] + strip * BY;
  const int xa = a.run_lo[0] + chunk * a.lx;
  const int xend = min (xa + a.lx, a.run_hi[0]);   // H (and stored E) planes [xa, xend)
  g.kc = min (g.k, a.nz - 1);
  g.zb = axis_bits (a, 2, g.k);
  asm volatile ("" : "+v"(g.zb));
  g.zf = zface_of (a, g.k);
  g.rowb = 0;
#pragma unroll
  for (int r = 0; r <= BY; ++r) g.rowb |= axis_bits (a, 1, g.j0 + r) << (8 * r);
  {
    const int w0 = __builtin_amdgcn_readfirstlane (wave * 64);   // first k of the wave
    g.wave_box = w0 >= a.run_lo[2] && w0 + 63 < a.run_hi[2];
    const int f0 = a.tl[2] - 1, f1 = a.tr[2] - 1, f2 = a.tr[2];
    g.wave_zface = a.tfsf && ((f0 >= w0 && f0 < w0 + 64) || (f1 >= w0 && f1 < w0 + 64) || (f2 >= w0 && f2 < w0 + 64));
  }

  uint32_t ro[BY + 2];
  row_offsets<T, BY> (a, g, ro);

  // two copies of the carried state and the plane loads, alternating between the two planes
  // of one loop iteration (no register copies between planes)
  Carry<T, BY> c0, c1;
  PlaneLoads<T, BY> p0, p1;
  {
    // prologue: E at plane xa (H before the step at xa-1 and eps at xa are loaded here)
    Carry<T, BY> cm;
    T epk0[BY + 1];   // eps of plane xa at z+1
    {
      uint32_t o[BY + 2];
      fresh_offsets (ro, o);
      const T *Hy = plane_row (a, a.f[4], xa - 1, g), *Hz = plane_row (a, a.f[5], xa - 1, g);
#pragma unroll
      for (int r = 0; r <= BY; ++r)
      {
        cm.hy[r] = rld (Hy, o[r + 1]);
        cm.hz[r] = rld (Hz, o[r + 1]);
      }
      if (MAT & kMatEps)
      {
        const T *E0 = plane_row (a, a.eps, xa, g);
#pragma unroll
        for (int r = 0; r <= BY; ++r)
        {
          cm.ep1[r] = rld (E0, o[r + 1]);
          epk0[r] = rld (E0, o[r + 1], 1);
        }
      }
      else
      {
#pragma unroll
        for (int r = 0; r <= BY; ++r) cm.ep1[r] = epk0[r] = a.eps_u;
      }
    }
    PlaneLoads<T, BY> L;
    e_load<T, BY, MAT> (a, xa, g, ro, L);
    if (a.tfsf)
    {
      // corrections of plane xa (E part) into zl[0], then those of the first iteration; the
      // 24 (BY + 1) entries take two passes of a 64-thread block
      for (int t = g.k; t < kZStage; t += (int) blockDim.x) (&s.zl[0][0][0][0][0])[t] = zstage_load<T, BY> (a, t, xa, g.j0);
      __syncthreads ();
    }
    e_compute<T, BY> (a, xa, g, L, cm, epk0, s.zl[0][0], c0);
    const double zpre = zstage_load<T, BY> (a, g.k, min (xa + 1, xend), g.j0);
    e_store<T, BY> (a, xa, g, ro, c0);
    e_load<T, BY, MAT> (a, min (xa + 1, xend), g, ro, p0);
    if (lane == 0)
    {
#pragma unroll
      for (int r = 0; r < BY; ++r)
      {
        s.xch[0][wave][r][0] = c0.ex[r];
        s.xch[0][wave][r][1] = c0.ey[r];
      }
#pragma unroll
      for (int r = 0; r <= BY; ++r) s.epx[0][wave][r] = c0.ep1[r];
    }
    if (a.tfsf)
    {
      if (g.k < kZStage) (&s.zl[1][0][0][0][0])[g.k] = zpre;
      // a 64-thread block (short z rows) loads the entries past its width here
      for (int t = g.k + (int) blockDim.x; t < kZStage; t += (int) blockDim.x)
        (&s.zl[1][0][0][0][0])[t] = zstage_load<T, BY> (a, t, min (xa + 1, xend), g.j0);
    }
    __syncthreads ();
  }

  if (sizeof (T) == 4 || MAXW <= 12)
  {
    // unrolled by two planes: the two copies of the state swap roles without register copies
    for (int x = xa + 1; x <= xend; x += 2)
    {
      fused_plane<T, BY, MAT> (a, g, ro, x, xend, 1, lane, wave, nw, s, c0, p0, c1, p1);
      if (x + 1 > xend) break;
      fused_plane<T, BY, MAT> (a, g, ro, x + 1, xend, 0, lane, wave, nw, s, c1, p1, c0, p0);
    }
  }
  else
  {
    // FP64 in 16-wave workgroups (128 VGPRs; twice the registers per value): one plane per
    // iteration, fewer live values
    int par = 1;
    for (int x = xa + 1; x <= xend; ++x, par ^= 1)
    {
      fused_plane<T, BY, MAT> (a, g, ro, x, xend, par, lane, wave, nw, s, c0, p0, c1, p1);
      c0 = c1;
      p0 = p1;
    }
  }
}

// ---------------------------------------------------------------------------------------
// TMz half-steps (Ez; Hx, Hy) on [1, N-1)^2; one thread per cell, y lane-contiguous
// ---------------------------------------------------------------------------------------
template <typename T, int MODE>
__device__ __forceinline__ void tmz_e_cell (const StepArgs<T> &a, int i, int j)
{
  const int64_t sx = a.sx;
  const int sy = a.sy;   // == 1
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy;
  const int p[3] = { i, j, 0 };
  const bool special = is_special (a, p, 2, MODE);
  T fo;
  Old<T> dold;
  load_old<T, MODE, 2> (a, q, special, fo, dold);
  const T *Hx = a.f[3], *Hy = a.f[4];
  T v[4] = { Hy[q - sx], Hy[q], Hx[q - sy], Hx[q] };
  if (a.tfsf && tfsf_near<T, 2> (a, i, j, 0)) tfsf_apply<T, 2, 2> (a, i, j, 0, v);
  // Ez sits on the eps point: the 2-cell average reads the same cell twice
  const T e = a.eps ? a.eps[q] : a.eps_u;
  const T mat = a.matc[2] ? a.matc[2][q] : avg2 (e, e);
  T om = 0, ga = 0;
  if (MODE == 2 && special)
  {
    const T o = a.ope ? a.ope[q] : (T) 0, g = a.ge ? a.ge[q] : (T) 0;
    om = avg2 (o, o);
    ga = avg2 (g, g);
  }
  update_cell<T, MODE, 2> (a, q, p, fo, dold, curl_rhs<T, 2> (a, v, p), mat, special, om, ga);
}

template <typename T, int MODE>
__device__ __forceinline__ void tmz_h_cell (const StepArgs<T> &a, int i, int j)
{
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy;
  const int p[3] = { i, j, 0 };
  const bool special = is_special (a, p, 2, MODE);
  T fo3, fo4;
  Old<T> do3, do4;
  load_old<T, MODE, 3> (a, q, special, fo3, do3);
  load_old<T, MODE, 4> (a, q, special, fo4, do4);
  const T *Ez = a.f[2];
  const T ez = Ez[q];
  // Hx: Ey operands absent (0), Ez at (i, j) and (i, j+1); Hy: Ez at (i, j) and (i+1, j), Ex absent
  T vx[4] = { (T) 0, (T) 0, ez, Ez[q + sy] };
  T vy[4] = { ez, Ez[q + sx], (T) 0, (T) 0 };
  if (a.tfsf && tfsf_near<T, 2> (a, i, j, 0))
  {
    tfsf_apply<T, 3, 2> (a, i, j, 0, vx);
    tfsf_apply<T, 4, 2> (a, i, j, 0, vy);
  }
  T mx, my;
  if (a.matc[3])
  {
    mx = a.matc[3][q];
    my = a.matc[4][q];
  }
  else if (a.mu)
  {
    const T *m = a.mu;
    mx = avg4 (m[q], m[q], m[q + sy], m[q + sy]);
    my = avg4 (m[q], m[q], m[q + sx], m[q + sx]);
  }
  else
  {
    mx = my = a.mu_u;
  }
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
  update_cell<T, MODE, 3> (a, q, p, fo3, do3, curl_rhs<T, 3> (a, vx, p), mx, special, omx, gmx);
  update_cell<T, MODE, 4> (a, q, p, fo4, do4, curl_rhs<T, 4> (a, vy, p), my, special, omy, gmy);
}

// ---------------------------------------------------------------------------------------
// TEz half-steps (Ex, Ey; Hz) on [1, N-1)^2: the 3D expressions with the absent operands
// (Hy for Ex, Hx for Ey) read as 0, as for TMz; y lane-contiguous, updated in place
// ---------------------------------------------------------------------------------------
template <typename T, int MODE>
__device__ __forceinline__ void tez_e_cell (const StepArgs<T> &a, int i, int j)
{
  const int64_t sx = a.sx;
  const int sy = a.sy;   // == 1
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy;
  const int p[3] = { i, j, 0 };
  const bool special = is_special (a, p, 2, MODE);
  T fo0, fo1;
  Old<T> do0, do1;
  load_old<T, MODE, 0> (a, q, special, fo0, do0);
  load_old<T, MODE, 1> (a, q, special, fo1, do1);
  const T *Hz = a.f[5];
  const T hz = Hz[q];
  // Ex: Hz at (i, j-1) and (i, j), Hy absent; Ey: Hx absent, Hz at (i-1, j) and (i, j)
  T vx[4] = { Hz[q - sy], hz, (T) 0, (T) 0 };
  T vy[4] = { (T) 0, (T) 0, Hz[q - sx], hz };
  if (a.tfsf && tfsf_near<T, 2> (a, i, j, 0))
  {
    tfsf_apply<T, 0, 2> (a, i, j, 0, vx);
    tfsf_apply<T, 1, 2> (a, i, j, 0, vy);
  }
  // 2-cell averages along each component's axis
  T ex_m = a.eps_u, ey_m = a.eps_u;
  if (a.matc[0])
  {
    ex_m = a.matc[0][q];
    ey_m = a.matc[1][q];
  }
  else if (a.eps)
  {
    const T e = a.eps[q];
    ex_m = avg2 (e, a.eps[q + sx]);
    ey_m = avg2 (e, a.eps[q + sy]);
  }
  T oxm = 0, oym = 0, gxm = 0, gym = 0;
  if (MODE == 2 && special)
  {
    if (a.ope)
    {
      oxm = avg2 (a.ope[q], a.ope[q + sx]);
      oym = avg2 (a.ope[q], a.ope[q + sy]);
    }
    if (a.ge)
    {
      gxm = avg2 (a.ge[q], a.ge[q + sx]);
      gym = avg2 (a.ge[q], a.ge[q + sy]);
    }
  }
  update_cell<T, MODE, 0> (a, q, p, fo0, do0, curl_rhs<T, 0> (a, vx, p), ex_m, special, oxm, gxm);
  update_cell<T, MODE, 1> (a, q, p, fo1, do1, curl_rhs<T, 1> (a, vy, p), ey_m, special, oym, gym);
}

template <typename T, int MODE>
__device__ __forceinline__ void tez_h_cell (const StepArgs<T> &a, int i, int j)
{
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy;
  const int p[3] = { i, j, 0 };
  const bool special = is_special (a, p, 2, MODE);
  T fo5;
  Old<T> do5;
  load_old<T, MODE, 5> (a, q, special, fo5, do5);
  const T *Ex = a.f[0], *Ey = a.f[1];
  // Hz: Ex at (i, j) and (i, j+1), Ey at (i, j) and (i+1, j)
  T vz[4] = { Ex[q], Ex[q + sy], Ey[q], Ey[q + sx] };
  if (a.tfsf && tfsf_near<T, 2> (a, i, j, 0)) tfsf_apply<T, 5, 2> (a, i, j, 0, vz);
  // 4-cell transverse average
  T hz_m = a.mu_u;
  if (a.matc[5])
  {
    hz_m = a.matc[5][q];
  }
  else if (a.mu)
  {
    const T *m = a.mu;
    hz_m = avg4 (m[q], m[q + sy], m[q + sx], m[q + sx + sy]);
  }
  T ozm = 0, gzm = 0;
  if (MODE == 2 && special)
  {
    const T *o = a.opm, *g = a.gm;
    if (o) ozm = sum4 (o[q], o[q + sy], o[q + sx], o[q + sx + sy]);
    if (g) gzm = sum4 (g[q], g[q + sy], g[q + sx], g[q + sx + sy]);
  }
  update_cell<T, MODE, 5> (a, q, p, fo5, do5, curl_rhs<T, 5> (a, vz, p), hz_m, special, ozm, gzm);
}

template <typename T, int MODE>
__global__ __launch_bounds__ (256) void e_half_tez (StepArgs<T> a)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j < 1 || j >= a.ny - 1) return;
  const int p[3] = { i, j, 0 };
  // outside the Drude box: the PML update (see e_half_3d)
  if (MODE == 2 && !in_drude_box (a, p, 2)) tez_e_cell<T, 1> (a, i, j);
  else tez_e_cell<T, MODE> (a, i, j);
}

template <typename T, int MODE>
__global__ __launch_bounds__ (256) void h_half_tez (StepArgs<T> a)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j < 1 || j >= a.ny - 1) return;
  const int p[3] = { i, j, 0 };
  if (MODE == 2 && !in_drude_box (a, p, 2)) tez_h_cell<T, 1> (a, i, j);
  else tez_h_cell<T, MODE> (a, i, j);
}

template <typename T, int MODE>
__global__ __launch_bounds__ (256) void e_half_tmz (StepArgs<T> a)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j < 1 || j >= a.ny - 1) return;
  const int p[3] = { i, j, 0 };
  // outside the Drude box: the PML update (see e_half_3d)
  if (MODE == 2 && !in_drude_box (a, p, 2)) tmz_e_cell<T, 1> (a, i, j);
  else tmz_e_cell<T, MODE> (a, i, j);
}

template <typename T, int MODE>
__global__ __launch_bounds__ (256) void h_half_tmz (StepArgs<T> a)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i < 1 || i >= a.nx - 1 || j < 1 || j >= a.ny - 1) return;
  const int p[3] = { i, j, 0 };
  if (MODE == 2 && !in_drude_box (a, p, 2)) tmz_h_cell<T, 1> (a, i, j);
  else tmz_h_cell<T, MODE> (a, i, j);
}

// ---------------------------------------------------------------------------------------
// Batched TMz step: many independent simulations of different sizes, E then H of one time
// step fused in one launch.
//
// Each wave takes one tile of a tile table: (simulation, x rows [xa, xb), first column y0).
// Lane l holds column y0 + l; lanes 0..62 own their column, lane 63 recomputes E of the next
// tile's first column, which H of column y0 + 62 needs.  The wave marches along x with E one
// row ahead of H: E at row x needs H before the step at rows x and x-1 (x-1 carried in
// registers) and at column j-1 (the same row, one element back); H at row x-1 needs E after
// the step at rows x-1 (carried, and at column j+1 from the next lane) and x.  E of row xb is
// recomputed for H of row xb-1 and not stored.  The state is double-buffered by step parity
// (old values from f / d, new ones to f_n / d_n), so recomputed halo cells read the same old
// values as their owner.  Each cell's arithmetic is cell_value with the operands of
// tmz_e_cell / tmz_h_cell, so a batch member is bit-identical to the single-engine run.
//
// The per-simulation set-up (grid pointers of both parities, sizes, PML tables, coefficients)
// is a StepArgs record in device memory, read with scalar loads (uniform per tile).  The time
// step of each simulation comes from device memory (step[sim]), so a captured graph of several
// steps can be replayed and groups of simulations can run ahead of one another; hard point
// sources are applied to the source cell right after its update, from a per-simulation table.
// ---------------------------------------------------------------------------------------
// Old values of one row of a TMz tile (the lane's column), loaded one row ahead.
template <typename T>
struct TmzRow
{
  T ez, hx, hy, hx_l;   // Ez, Hx, Hy before the step; Hx at column j-1
  T eps;                // eps (a.eps_u without an eps grid)
  T dz, bx, by;         // D/B before the step (PML cells; 0 elsewhere)
  // PML coefficients that depend on x (uniform over the row): Ez Ca/Cb, Hx p4/m4, Hy ca5/dd5
  T ez_cac, ez_cbc, hx_p4, hx_m4, hy_ca5, hy_dd5;
  bool special;
};

// The PML coefficients of a TMz cell (x, j, 0) by the axis they depend on (SigAxes): x (the
// row, TmzRow), j (the lane's column, constant over a tile) or z = 0 (constant).
static_assert (SigAxes<2>::c == 0 && SigAxes<2>::s4 == 2 && SigAxes<2>::s5 == 1, "Ez PML axes");
static_assert (SigAxes<3>::c == 1 && SigAxes<3>::s4 == 0 && SigAxes<3>::s5 == 2, "Hx PML axes");
static_assert (SigAxes<4>::c == 2 && SigAxes<4>::s4 == 1 && SigAxes<4>::s5 == 0, "Hy PML axes");

template <typename T>
struct TmzColCoef
{
  T ez_ca5, ez_dd5, hx_cac, hx_cbc, hy_p4, hy_m4;   // column j
  T ez_p4, ez_m4, hx_ca5, hx_dd5, hy_cac, hy_cbc;   // z = 0
};

template <typename T>
__device__ __forceinline__ T pml_slot (const StepArgs<T> &a, int comp, int slot, int idx)
{
  return gld (a.pml + ((size_t) comp * 6 + slot) * a.tab_len + idx);
}

// The same as a vector load even for a uniform index (vector loads complete in order, so a
// value prefetched one row ahead is not waited for with the current row's; scalar loads are
// waited for all together).
template <typename T>
__device__ __forceinline__ T pml_slot_v (const StepArgs<T> &a, int comp, int slot, int idx)
{
  int z = 0;
  asm volatile ("" : "+v"(z));
  return gld (a.pml + ((size_t) comp * 6 + slot) * a.tab_len + idx + z);
}

template <typename T, int MODE>
__device__ __forceinline__ void tmz_row_load (const StepArgs<T> &a, int x, int j, int jc, TmzRow<T> &r)
{
  const int64_t q = (int64_t) x * a.sx + jc;
  r.ez = a.f[2][q];
  r.hx = a.f[3][q];
  r.hy = a.f[4][q];
  r.hx_l = a.f[3][q - (jc > 0 ? 1 : 0)];   // column 0 (a halo lane of the two-step tile) is never updated
  r.eps = a.eps ? a.eps[q] : a.eps_u;
  const int p[3] = { x, j, 0 };
  r.special = is_special (a, p, 2, MODE);
  r.dz = r.bx = r.by = (T) 0;
  if (MODE != 0)
  {
    r.ez_cac = pml_slot_v (a, 2, 0, x);
    r.ez_cbc = pml_slot_v (a, 2, 1, x);
    r.hx_p4 = pml_slot_v (a, 3, 2, x);
    r.hx_m4 = pml_slot_v (a, 3, 3, x);
    r.hy_ca5 = pml_slot_v (a, 4, 4, x);
    r.hy_dd5 = pml_slot_v (a, 4, 5, x);
    if (r.special)
    {
      r.dz = a.d[2][q];
      r.bx = a.d[3][q];
      r.by = a.d[4][q];
    }
  }
}

template <typename T, int MODE>
__device__ __forceinline__ void tmz_fused_tile (const StepArgs<T> &a, int xa, int xb, int j, int lane, T se, T sh)
{
  const int nx = a.nx, ny = a.ny;
  const int64_t sx = a.sx;   // == ny
  const int jc = min (j, ny - 1);   // lanes past the grid read its last column
  const bool jin = j < ny - 1;      // interior column (j >= 1 always)
  const bool own = jin && lane < 63;
  const bool src_col = j == a.src_pos[1];
  T *Ezn = a.f_n[2], *Hxn = a.f_n[3], *Hyn = a.f_n[4];
  TmzColCoef<T> cc {};
  if (MODE != 0)
  {
    cc.ez_ca5 = pml_slot (a, 2, 4, jc);
    cc.ez_dd5 = pml_slot (a, 2, 5, jc);
    cc.hx_cac = pml_slot (a, 3, 0, jc);
    cc.hx_cbc = pml_slot (a, 3, 1, jc);
    cc.hy_p4 = pml_slot (a, 4, 2, jc);
    cc.hy_m4 = pml_slot (a, 4, 3, jc);
    cc.ez_p4 = pml_slot (a, 2, 2, 0);
    cc.ez_m4 = pml_slot (a, 2, 3, 0);
    cc.hx_ca5 = pml_slot (a, 3, 4, 0);
    cc.hx_dd5 = pml_slot (a, 3, 5, 0);
    cc.hy_cac = pml_slot (a, 4, 0, 0);
    cc.hy_cbc = pml_slot (a, 4, 1, 0);
  }
  // carried: the old row x-1 and E after the step at row x-1; rows x+1 and x+2 are loaded
  // ahead (two rows of loads in flight per wave hide the memory latency of the march)
  TmzRow<T> prev, cur, nx1, nx2;
  prev.hy = a.f[4][(int64_t) (xa - 1) * sx + jc];
  prev.hx = prev.bx = prev.by = (T) 0;
  prev.special = false;
  T ez_p = (T) 0;
  tmz_row_load<T, MODE> (a, xa, j, jc, cur);
  if (xa < xb) tmz_row_load<T, MODE> (a, xa + 1, j, jc, nx1);
  for (int x = xa; x <= xb; ++x)
  {
    if (x + 2 <= xb) tmz_row_load<T, MODE> (a, x + 2, j, jc, nx2);   // prefetch
    const int64_t q = (int64_t) x * sx + jc;
    const bool xin = x < nx - 1;   // x >= 1 always
    T ez_n = cur.ez;
    if (jin && xin)
    {
      Old<T> dold;
      dold.d = cur.dz;
      dold.dp = dold.m = dold.mp = (T) 0;
      T v[4] = { prev.hy, cur.hy, cur.hx_l, cur.hx };
      const PmlCoef<T> pc = { cur.ez_cac, cur.ez_cbc, cc.ez_p4, cc.ez_m4, cc.ez_ca5, cc.ez_dd5 };
      ez_n = cell_value_pc<T, MODE, 2> (a, q, cur.ez, dold, curl (v), avg2 (cur.eps, cur.eps), cur.special, (T) 0,
                                        (T) 0, own && x < xb, pc);
      if ((a.src_mask & 4u) && src_col && x == a.src_pos[0]) ez_n = se;
      if (own && x < xb) Ezn[q] = ez_n;
    }
    const T ez_r = lane_next (ez_p, ez_p);   // E after the step at (x-1, j+1)
    const int i = x - 1;
    if (x > xa && own && i < nx - 1)
    {
      const int64_t qi = q - sx;
      Old<T> o3, o4;
      o3.dp = o3.m = o3.mp = (T) 0;
      o4 = o3;
      o3.d = prev.bx;
      o4.d = prev.by;
      // Hx: Ez at (i, j) and (i, j+1); Hy: Ez at (i, j) and (i+1, j)
      T vx[4] = { (T) 0, (T) 0, ez_p, ez_r };
      T vy[4] = { ez_p, ez_n, (T) 0, (T) 0 };
      T mx = a.mu_u, my = a.mu_u;
      if (a.mu)
      {
        const T *m = a.mu;
        mx = avg4 (m[qi], m[qi], m[qi + 1], m[qi + 1]);
        my = avg4 (m[qi], m[qi], m[q], m[q]);
      }
      const PmlCoef<T> px = { cc.hx_cac, cc.hx_cbc, prev.hx_p4, prev.hx_m4, cc.hx_ca5, cc.hx_dd5 };
      const PmlCoef<T> py = { cc.hy_cac, cc.hy_cbc, cc.hy_p4, cc.hy_m4, prev.hy_ca5, prev.hy_dd5 };
      T hxn = cell_value_pc<T, MODE, 3> (a, qi, prev.hx, o3, curl (vx), mx, prev.special, (T) 0, (T) 0, true, px);
      T hyn = cell_value_pc<T, MODE, 4> (a, qi, prev.hy, o4, curl (vy), my, prev.special, (T) 0, (T) 0, true, py);
      if (src_col && i == a.src_pos[0])
      {
        if (a.src_mask & 8u) hxn = sh;
        if (a.src_mask & 16u) hyn = sh;
      }
      Hxn[qi] = hxn;
      Hyn[qi] = hyn;
    }
    prev = cur;
    cur = nx1;
    nx1 = nx2;
    ez_p = ez_n;
  }
}

// One TMz component update for the two-step tile: cell_value_pc's arithmetic (direct, or the
// PML two-stage update of MODE 1), returning the new D/B in d_new (d_old for direct cells).
template <typename T, int MODE, int COMP>
__device__ __forceinline__ T tmz_update (const StepArgs<T> &a, T f_old, T d_old, T tmp, T mat, bool special,
                                         const PmlCoef<T> &pc, T &d_new)
{
  const T modc = COMP >= 3 ? a.mu0 : a.eps0;
  if (MODE == 0 || !special)
  {
    d_new = d_old;
    const T cb = direct_cb (a.dt, mat, modc, a.dx);
    return f_old * (T) 1.0 + tmp * cb;
  }
  d_new = d_old * pc.cac + tmp * pc.cbc;
  const T mod = mat * modc;
  return f_old * pc.ca5 + d_new * ((pc.p4 / mod) / pc.dd5) - d_old * ((pc.m4 / mod) / pc.dd5);
}

// Row-dependent H coefficients and PML flag of a row, kept for the second step's H update.
template <typename T>
struct TmzRowH
{
  T hx_p4, hx_m4, hy_ca5, hy_dd5;
  bool special;
};

// Two time steps (t+1 and t+2) of a TMz tile in one sweep.  Lane l holds column y0 - 1 + l:
// lanes 1..61 own their column; lane 0 (column y0 - 1) recomputes E and H of step t+1 for the
// E of step t+2 at column y0; lanes 62 and 63 recompute what H of step t+2 at column y0 + 60
// needs.  Along x the stages lag by one row: at row x, E(t+1) of row x, H(t+1) of x-1, E(t+2)
// of x-1 and H(t+2) of x-2; the sweep starts one row before the owned rows [xa, xb) and ends
// one row after, so that the owned rows of step t+2 see their neighbours of step t+1.  Only
// step t+2 is stored; step t+1 lives in registers.  Uniform mu only (no mu grid).
template <typename T, int MODE>
__device__ __forceinline__ void tmz_fused2_tile (const StepArgs<T> &a, int xa, int xb, int y0, int lane, T se0, T sh0,
                                                 T se1, T sh1)
{
  const int nx = a.nx, ny = a.ny;
  const int64_t sx = a.sx;   // == ny
  const int j = y0 - 1 + lane;
  const int jc = min (j, ny - 1);
  const bool jin = j >= 1 && j < ny - 1;
  const bool own = jin && lane >= 1 && lane <= 61;
  const bool src_col = j == a.src_pos[1];
  const int sxr = a.src_pos[0];
  const T mu = a.mu_u;
  T *Ezn = a.f_n[2], *Hxn = a.f_n[3], *Hyn = a.f_n[4];
  TmzColCoef<T> cc {};
  if (MODE != 0)
  {
    cc.ez_ca5 = pml_slot (a, 2, 4, jc);
    cc.ez_dd5 = pml_slot (a, 2, 5, jc);
    cc.hx_cac = pml_slot (a, 3, 0, jc);
    cc.hx_cbc = pml_slot (a, 3, 1, jc);
    cc.hy_p4 = pml_slot (a, 4, 2, jc);
    cc.hy_m4 = pml_slot (a, 4, 3, jc);
    cc.ez_p4 = pml_slot (a, 2, 2, 0);
    cc.ez_m4 = pml_slot (a, 2, 3, 0);
    cc.hx_ca5 = pml_slot (a, 3, 4, 0);
    cc.hx_dd5 = pml_slot (a, 3, 5, 0);
    cc.hy_cac = pml_slot (a, 4, 0, 0);
    cc.hy_cbc = pml_slot (a, 4, 1, 0);
  }
  // state t: rows x-1 (p), x (cur), x+1 and x+2 (loaded ahead)
  TmzRow<T> p, cur, nx1, nx2;
  p.hy = xa >= 2 ? a.f[4][(int64_t) (xa - 2) * sx + jc] : (T) 0;
  p.hx = p.bx = p.by = p.ez = p.dz = p.eps = (T) 0;
  p.special = false;
  tmz_row_load<T, MODE> (a, xa - 1, j, jc, cur);
  tmz_row_load<T, MODE> (a, xa, j, jc, nx1);
  // step t+1 at row x-1 (E, D) and x-2 (H, B); step t+2 E at row x-2; coefficients of row x-2
  T e1_p = (T) 0, dz1_p = (T) 0;
  T h1x_pp = (T) 0, h1y_pp = (T) 0, b1x_pp = (T) 0, b1y_pp = (T) 0, e2_pp = (T) 0;
  TmzRowH<T> rh_pp {};
  for (int x = xa - 1; x <= xb + 1; ++x)
  {
    if (x + 2 <= xb + 1) tmz_row_load<T, MODE> (a, min (x + 2, nx - 1), j, jc, nx2);   // prefetch
    const bool xin = x >= 1 && x < nx - 1;
    // S1: E(t+1) at row x
    T e1 = cur.ez, dz1 = cur.dz;
    if (jin && xin)
    {
      T v[4] = { p.hy, cur.hy, cur.hx_l, cur.hx };
      const PmlCoef<T> pc = { cur.ez_cac, cur.ez_cbc, cc.ez_p4, cc.ez_m4, cc.ez_ca5, cc.ez_dd5 };
      e1 = tmz_update<T, MODE, 2> (a, cur.ez, cur.dz, curl (v), avg2 (cur.eps, cur.eps), cur.special, pc, dz1);
      if ((a.src_mask & 4u) && src_col && x == sxr) e1 = se0;
    }
    // S2: H(t+1) at row x-1
    const int i1 = x - 1;
    const bool i1in = i1 >= 1 && i1 < nx - 1;
    const T e1r = lane_next (e1_p, e1_p);
    T h1x = p.hx, h1y = p.hy, b1x = p.bx, b1y = p.by;
    if (x >= xa && jin && i1in)
    {
      T vx[4] = { (T) 0, (T) 0, e1_p, e1r };
      T vy[4] = { e1_p, e1, (T) 0, (T) 0 };
      const PmlCoef<T> px = { cc.hx_cac, cc.hx_cbc, p.hx_p4, p.hx_m4, cc.hx_ca5, cc.hx_dd5 };
      const PmlCoef<T> py = { cc.hy_cac, cc.hy_cbc, cc.hy_p4, cc.hy_m4, p.hy_ca5, p.hy_dd5 };
      h1x = tmz_update<T, MODE, 3> (a, p.hx, p.bx, curl (vx), mu, p.special, px, b1x);
      h1y = tmz_update<T, MODE, 4> (a, p.hy, p.by, curl (vy), mu, p.special, py, b1y);
      if (src_col && i1 == sxr)
      {
        if (a.src_mask & 8u) h1x = sh0;
        if (a.src_mask & 16u) h1y = sh0;
      }
    }
    // S3: E(t+2) at row x-1
    const T h1xl = lane_prev (h1x, h1x);
    T e2 = e1_p;
    if (x >= xa + 1 && jin && i1in)
    {
      T v[4] = { h1y_pp, h1y, h1xl, h1x };
      const PmlCoef<T> pc = { p.ez_cac, p.ez_cbc, cc.ez_p4, cc.ez_m4, cc.ez_ca5, cc.ez_dd5 };
      T dz2;
      e2 = tmz_update<T, MODE, 2> (a, e1_p, dz1_p, curl (v), avg2 (p.eps, p.eps), p.special, pc, dz2);
      if ((a.src_mask & 4u) && src_col && i1 == sxr) e2 = se1;
      if (own && i1 < xb)
      {
        const int64_t q = (int64_t) i1 * sx + jc;
        Ezn[q] = e2;
        if (MODE != 0 && p.special) a.d_n[2][q] = dz2;
      }
    }
    // S4: H(t+2) at row x-2
    const int i2 = x - 2;
    const T e2r = lane_next (e2_pp, e2_pp);
    if (x >= xa + 2 && own && i2 >= 1 && i2 < nx - 1)
    {
      T vx[4] = { (T) 0, (T) 0, e2_pp, e2r };
      T vy[4] = { e2_pp, e2, (T) 0, (T) 0 };
      const PmlCoef<T> px = { cc.hx_cac, cc.hx_cbc, rh_pp.hx_p4, rh_pp.hx_m4, cc.hx_ca5, cc.hx_dd5 };
      const PmlCoef<T> py = { cc.hy_cac, cc.hy_cbc, cc.hy_p4, cc.hy_m4, rh_pp.hy_ca5, rh_pp.hy_dd5 };
      T b2x, b2y;
      T h2x = tmz_update<T, MODE, 3> (a, h1x_pp, b1x_pp, curl (vx), mu, rh_pp.special, px, b2x);
      T h2y = tmz_update<T, MODE, 4> (a, h1y_pp, b1y_pp, curl (vy), mu, rh_pp.special, py, b2y);
      if (src_col && i2 == sxr)
      {
        if (a.src_mask & 8u) h2x = sh1;
        if (a.src_mask & 16u) h2y = sh1;
      }
      const int64_t q = (int64_t) i2 * sx + jc;
      Hxn[q] = h2x;
      Hyn[q] = h2y;
      if (MODE != 0 && rh_pp.special)
      {
        a.d_n[3][q] = b2x;
        a.d_n[4][q] = b2y;
      }
    }
    // rotate
    h1x_pp = h1x;
    h1y_pp = h1y;
    b1x_pp = b1x;
    b1y_pp = b1y;
    e2_pp = e2;
    rh_pp.hx_p4 = p.hx_p4;
    rh_pp.hx_m4 = p.hx_m4;
    rh_pp.hy_ca5 = p.hy_ca5;
    rh_pp.hy_dd5 = p.hy_dd5;
    rh_pp.special = p.special;
    p = cur;
    e1_p = e1;
    dz1_p = dz1;
    cur = nx1;
    nx1 = nx2;
  }
}

template <typename T>
__global__ __launch_bounds__ (256) void batch_tmz_step2 (const StepArgs<T> *sims, const int4 *tiles, int ntiles,
                                                        const int *state, const T *src_tab, int src_len)
{
  const int lane = (int) (threadIdx.x & 63);
  const int tile = (int) blockIdx.x * 4 + __builtin_amdgcn_readfirstlane ((int) (threadIdx.x >> 6));
  if (tile >= ntiles) return;   // no barriers below
  const int4 td = tiles[tile];   // sim, xa, y0, xb
  const int t = state[2 * td.x];
  const StepArgs<T> a = sims[2 * td.x + state[2 * td.x + 1]];
  const T *st = src_tab + ((int64_t) td.x * src_len + t) * 2;
  if (a.pml) tmz_fused2_tile<T, 1> (a, td.y, td.w, td.z, lane, st[0], st[1], st[2], st[3]);
  else tmz_fused2_tile<T, 0> (a, td.y, td.w, td.z, lane, st[0], st[1], st[2], st[3]);
}

template <typename T>
__global__ __launch_bounds__ (256) void batch_tmz_step (const StepArgs<T> *sims, const int4 *tiles, int ntiles,
                                                       const int *state, const T *src_tab, int src_len)
{
  const int lane = (int) (threadIdx.x & 63);
  const int tile = (int) blockIdx.x * 4 + __builtin_amdgcn_readfirstlane ((int) (threadIdx.x >> 6));
  if (tile >= ntiles) return;   // no barriers below
  const int4 td = tiles[tile];   // sim, xa, y0, xb
  const int t = state[2 * td.x];
  // a private copy: loads through a reference could not be kept across the tile's stores
  // (the compiler cannot rule out that they alias the record)
  const StepArgs<T> a = sims[2 * td.x + state[2 * td.x + 1]];
  const T se = src_tab[((int64_t) td.x * src_len + t) * 2 + 0];
  const T sh = src_tab[((int64_t) td.x * src_len + t) * 2 + 1];
  if (a.pml) tmz_fused_tile<T, 1> (a, td.y, td.w, td.z + lane, lane, se, sh);
  else tmz_fused_tile<T, 0> (a, td.y, td.w, td.z + lane, lane, se, sh);
}

// Step 0 performs no field update (Scheme.inc.h:1032) but the hard sources still fire.
template <typename T>
__global__ void batch_tmz_sources (const StepArgs<T> *sims, int nsim, const int *state, const T *src_tab, int src_len)
{
  const int s = (int) (blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= nsim) return;
  const int t = state[2 * s];
  const StepArgs<T> &a = sims[2 * s + state[2 * s + 1]];
  if (!a.src_mask) return;
  const T ve = src_tab[((int64_t) s * src_len + t) * 2 + 0];
  const T vh = src_tab[((int64_t) s * src_len + t) * 2 + 1];
  if (a.src_mask & 4u) gst (a.f_n[2] + a.src_q, ve);
  if (a.src_mask & 8u) gst (a.f_n[3] + a.src_q, vh);
  if (a.src_mask & 16u) gst (a.f_n[4] + a.src_q, vh);
}

// After a launch that advanced simulations sims[0..n) by k steps (the last node of every
// captured step): their step counters move on by k and their newest fields are now in the
// other buffer.
__global__ void batch_advance (int *state, const int *sims, int n, int k)
{
  const int u = (int) (blockIdx.x * blockDim.x + threadIdx.x);
  if (u < n)
  {
    state[2 * sims[u]] += k;
    state[2 * sims[u] + 1] ^= 1;
  }
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

// Double material precision (--use-double-material-precision): the material grid has twice the
// field grid's extent along each axis of the scheme, and every field point averages the 8 fine
// cells around it (InternalScheme::getMaterial, InternalScheme.inc.h:2924-2950, with the cells
// of YeeGridLayout::initMaterialCoordinates, YeeGridLayout.h:866-1037, and the pairwise
// Approximation::approximateMaterial, Approximation.cpp:36-54).  Along an axis where the point
// sits at an integer coordinate (an E component's own axis, an H component's other axes) the
// fine cells are 2p+1, 2p+2, else 2p, 2p+1; an axis the scheme lacks contributes cell 0 twice.
// The pairs nest innermost-first along the component's (inner, mid, outer) axes: Ex (y, x, z),
// Ey (x, y, z), Ez (y, x, z), Hx (x, z, y), Hy (y, z, x), Hz (z, y, x), the order in which the
// reference lists the 8 cells.  One thread per field cell writes the averages of the three
// components cbase .. cbase + 2 (eps for E, mu for H) that have an output array; indices are
// clamped to the fine grid (the clamped values reach only cells the sweeps never update).
template <typename T>
__device__ __forceinline__ T fine_avg8 (const T *m, const int64_t mn[3], int act, int c, const int p[3])
{
  int64_t lo[3], hi[3];
  for (int ax = 0; ax < 3; ++ax)
  {
    const bool integer = c >= 3 ? ax != c - 3 : ax == c;
    if (!((act >> ax) & 1))
    {
      lo[ax] = hi[ax] = 0;
      continue;
    }
    const int64_t top = mn[ax] - 1, l = (int64_t) 2 * p[ax] + (integer ? 1 : 0);
    lo[ax] = l < top ? l : top;
    hi[ax] = l + 1 < top ? l + 1 : top;
  }
  int ai, am, ao;
  switch (c)
  {
    case 0: ai = 1; am = 0; ao = 2; break;
    case 1: ai = 0; am = 1; ao = 2; break;
    case 2: ai = 1; am = 0; ao = 2; break;
    case 3: ai = 0; am = 2; ao = 1; break;
    case 4: ai = 1; am = 2; ao = 0; break;
    default: ai = 2; am = 1; ao = 0; break;
  }
  const int64_t st[3] = { mn[1] * mn[2], mn[2], 1 };
  auto v = [&] (int si, int sm, int so) -> T {
    return m[(si ? hi[ai] : lo[ai]) * st[ai] + (sm ? hi[am] : lo[am]) * st[am] + (so ? hi[ao] : lo[ao]) * st[ao]];
  };
  const T q0 = avg2 (avg2 (v (0, 0, 0), v (1, 0, 0)), avg2 (v (0, 1, 0), v (1, 1, 0)));
  const T q1 = avg2 (avg2 (v (0, 0, 1), v (1, 0, 1)), avg2 (v (0, 1, 1), v (1, 1, 1)));
  return avg2 (q0, q1);
}

template <typename T>
__global__ void fine_average (const T *m, int64_t mn0, int64_t mn1, int64_t mn2, int nx, int ny, int nz, int act,
                              int cbase, T *o0, T *o1, T *o2)
{
  const int64_t n = (int64_t) nx * ny * nz;
  const int64_t mn[3] = { mn0, mn1, mn2 };
  for (int64_t q = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; q < n; q += (int64_t) gridDim.x * blockDim.x)
  {
    const int p[3] = { (int) (q / ((int64_t) ny * nz)), (int) ((q / nz) % ny), (int) (q % nz) };
    if (o0) gst (o0 + q, fine_avg8 (m, mn, act, cbase, p));
    if (o1) gst (o1 + q, fine_avg8 (m, mn, act, cbase + 1, p));
    if (o2) gst (o2 + q, fine_avg8 (m, mn, act, cbase + 2, p));
  }
}

// Analytic test problems: component `comp` of every cell (border 0: the start values of
// Scheme::initGridWithInitialVals, Scheme.inc.h:2863-2958) or of every cell with an index 0 or
// N-1 along an axis of the scheme (border 1: calculateFieldStepIterationBorder,
// InternalScheme.inc.h:2646-2705, with doSkipBorderFunc :539-555) set to the kind's value at
// time `time` (s).  Grid-stride over the cells; the border pass skips the interior.
template <typename T>
__global__ void test_fill (StepArgs<T> a, T *f, int comp, int kind, double time, int border)
{
  const int64_t n = (int64_t) a.nx * a.ny * a.nz;
  for (int64_t q = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; q < n; q += (int64_t) gridDim.x * blockDim.x)
  {
    const int p[3] = { (int) (q / ((int64_t) a.ny * a.nz)), (int) ((q / a.nz) % a.ny), (int) (q % a.nz) };
    if (border)
    {
      const int ext[3] = { a.nx, a.ny, a.nz };
      bool edge = false;
      for (int d = 0; d < 3; ++d) edge = edge || (a.fn_dx[d] != 0.0 && (p[d] == 0 || p[d] == ext[d] - 1));
      if (!edge) continue;
    }
    double r[3];
    test_coords (a, comp, p, r);
    gst (f + q, (T) test_field (kind, comp, r[0], r[1], r[2], time));
  }
}

// Difference norm against the exact solution (calculateFieldStepIterationExact,
// InternalScheme.inc.h:2707-2816) over the box [lo, hi): per workgroup the sum of
// (exact - field)^2 and the largest |exact| of its grid-stride share, written to part[2 b],
// part[2 b + 1] (the host adds the workgroups' sums in order).  The exact value is the part
// this handle holds (exp_part) of an exp solution, else the real analytic value.  With fim set
// (complex fields: f the real part, fim the imaginary part) the module norm instead:
// (|exact| - |field|)^2 and the largest |exact|.
template <typename T>
__global__ __launch_bounds__ (256) void test_norm (StepArgs<T> a, const T *f, const T *fim, int comp, int kind,
                                                   double time, int lo0, int lo1, int lo2, int hi0, int hi1, int hi2,
                                                   double *part)
{
  __shared__ double ss[256], sm[256];
  const int64_t e0 = hi0 - lo0, e1 = hi1 - lo1, e2 = hi2 - lo2;
  const int64_t n = e0 * e1 * e2;
  double sum = 0.0, mx = 0.0;
  for (int64_t u = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; u < n; u += (int64_t) gridDim.x * blockDim.x)
  {
    const int p[3] = { lo0 + (int) (u / (e1 * e2)), lo1 + (int) ((u / e2) % e1), lo2 + (int) (u % e2) };
    double r[3];
    test_coords (a, comp, p, r);
    double er, ei = 0.0;
    if (kind >= YEE_FN_EXP1)
      test_exp (a, kind, comp, r, time, er, ei);
    else
      er = test_field (kind, comp, r[0], r[1], r[2], time);
    const int64_t q = (int64_t) p[0] * a.sx + (int64_t) p[1] * a.sy + (int64_t) p[2] * a.sz;
    double exact, d;
    if (fim)
    {
      const double fr = (double) gld (f + q), fi = (double) gld (fim + q);
      exact = sqrt (er * er + ei * ei);
      d = exact - sqrt (fr * fr + fi * fi);
    }
    else
    {
      exact = a.exp_part ? ei : er;
      d = exact - (double) gld (f + q);
    }
    sum += d * d;
    mx = fmax (mx, fabs (exact));
  }
  ss[threadIdx.x] = sum;
  sm[threadIdx.x] = mx;
  __syncthreads ();
  for (int w = 128; w > 0; w >>= 1)
  {
    if ((int) threadIdx.x < w)
    {
      ss[threadIdx.x] += ss[threadIdx.x + w];
      sm[threadIdx.x] = fmax (sm[threadIdx.x], sm[threadIdx.x + w]);
    }
    __syncthreads ();
  }
  if (threadIdx.x == 0)
  {
    gst (part + 2 * blockIdx.x, ss[0]);
    gst (part + 2 * blockIdx.x + 1, sm[0]);
  }
}

// Near-to-far-field transform of the complex build (Scheme::Pointing_scat, Scheme.inc.h:4036-4101,
// over the face sums of SchemeHelper::ntffN3D_x/y/z and ntffL3D_x/y/z, SchemeHelper.h:1030-2148,
// E_CENTERED).  Workgroup (angle a, pass p): pass p / 6 = N (surface currents from H averaged
// over its last two time layers) or L (from E averaged across the face), axis (p % 6) / 2, side
// p % 2 (left / right plane of the box); the workgroup's threads stride over the face cells
// [left + 1/2, right - 1/2] of the two in-plane axes and reduce the complex teta and phi sums.
struct Cplx
{
  double re, im;
};

template <typename T>
__device__ __forceinline__ Cplx ntff_at (const T *re, const T *im, const NtffArgs<T> &a, int c, const double p[3])
{
  // minimum coordinates of the component (E_CENTERED, YeeGridLayout.cpp:1962-1969)
  const int axis = c % 3;
  const bool h = c >= 3;
  int64_t q[3];
  for (int d = 0; d < 3; ++d) q[d] = (int64_t) (p[d] - (((d == axis) != h) ? 1.0 : 0.5));
  const int64_t idx = q[0] * a.sx + q[1] * a.sy + q[2];
  return Cplx { (double) gld (re + idx), (double) gld (im + idx) };
}

template <typename T>
__global__ __launch_bounds__ (256) void ntff_faces (NtffArgs<T> a)
{
  __shared__ double red[4][256];
  const int angle = (int) blockIdx.x / 12, pass = (int) blockIdx.x % 12;
  const int kind = pass / 6, ax = (pass % 6) / 2, side = pass % 2;
  const int u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;
  const int p0 = side ? a.right[ax] : a.left[ax];
  const double sign = side ? 1.0 : -1.0;   // x0 == right ? 1 : -1 (left == right cannot occur)
  const double teta = a.teta, phi = a.phis[angle];
  const double st = sin (teta), ct = cos (teta), sp = sin (phi), cp = cos (phi);
  const double diff0[3] = { (double) (a.n[0] / 2), (double) (a.n[1] / 2), (double) (a.n[2] / 2) };
  // the two components of each face (N: H, L: E) and the extra -1 of the teta / phi sums
  const int c1 = kind == 0 ? (ax == 2 ? 4 : 5) : (ax == 0 ? 1 : 0);
  const int c2 = kind == 0 ? (ax == 0 ? 4 : 3) : (ax == 2 ? 1 : 2);
  const double ts = kind == 0 ? (ax == 0 ? -1.0 : 1.0) : (ax == 1 ? -1.0 : 1.0);
  const double fs = kind == 0 ? (ax == 2 ? 1.0 : -1.0) : (ax == 2 ? -1.0 : 1.0);
  const int nu = a.right[u] - a.left[u], nv = a.right[v] - a.left[v];
  double sum[4] = { 0, 0, 0, 0 };
  for (int w = (int) threadIdx.x; w < nu * nv; w += (int) blockDim.x)
  {
    double base[3];
    base[ax] = (double) p0;
    base[u] = a.left[u] + 0.5 + (double) (w / nv);
    base[v] = a.left[v] + 0.5 + (double) (w % nv);
    double ps[4][3];
    for (int q = 0; q < 4; ++q)
      for (int d = 0; d < 3; ++d) ps[q][d] = base[d];
    ps[0][u] -= 0.5; ps[1][u] += 0.5; ps[2][v] -= 0.5; ps[3][v] += 0.5;
    Cplx f[4];
    for (int q = 0; q < 4; ++q)
    {
      const int c = q < 2 ? c1 : c2;
      if (kind == 0)
      {
        // H averaged over the newest two time layers
        const Cplx x = ntff_at (a.h_re[c - 3], a.h_im[c - 3], a, c, ps[q]);
        const Cplx y = ntff_at (a.hp_re[c - 3], a.hp_im[c - 3], a, c, ps[q]);
        f[q] = Cplx { (x.re + y.re) / 2, (x.im + y.im) / 2 };
      }
      else
      {
        // E averaged across the face
        double lo[3] = { ps[q][0], ps[q][1], ps[q][2] }, hi[3] = { ps[q][0], ps[q][1], ps[q][2] };
        lo[ax] -= 0.5;
        hi[ax] += 0.5;
        const Cplx x = ntff_at (a.e_re[c], a.e_im[c], a, c, lo);
        const Cplx y = ntff_at (a.e_re[c], a.e_im[c], a, c, hi);
        f[q] = Cplx { (x.re + y.re) / 2, (x.im + y.im) / 2 };
      }
    }
    const Cplx A { (f[0].re + f[1].re) / 2, (f[0].im + f[1].im) / 2 };
    const Cplx B { (f[2].re + f[3].re) / 2, (f[2].im + f[3].im) / 2 };
    double arg = (base[0] - diff0[0]) * st * cp + (base[1] - diff0[1]) * st * sp + (base[2] - diff0[2]) * ct;
    arg *= a.dx;
    const double er = cos (a.k * arg), ei = sin (a.k * arg);
    double tr, ti, fr, fi;   // the teta and phi integrands before the phase
    if (ax == 0)
    {
      if (kind == 0)
      {
        tr = A.re * ct * sp + B.re * st; ti = A.im * ct * sp + B.im * st;
        fr = A.re * cp; fi = A.im * cp;
      }
      else
      {
        tr = B.re * ct * sp + A.re * st; ti = B.im * ct * sp + A.im * st;
        fr = B.re * cp; fi = B.im * cp;
      }
    }
    else if (ax == 1)
    {
      if (kind == 0)
      {
        tr = A.re * ct * cp + B.re * st; ti = A.im * ct * cp + B.im * st;
        fr = A.re * sp; fi = A.im * sp;
      }
      else
      {
        tr = B.re * ct * cp + A.re * st; ti = B.im * ct * cp + A.im * st;
        fr = B.re * sp; fi = B.im * sp;
      }
    }
    else
    {
      if (kind == 0)
      {
        tr = -A.re * ct * cp + B.re * ct * sp; ti = -A.im * ct * cp + B.im * ct * sp;
        fr = A.re * sp + B.re * cp; fi = A.im * sp + B.im * cp;
      }
      else
      {
        tr = B.re * ct * cp - A.re * ct * sp; ti = B.im * ct * cp - A.im * ct * sp;
        fr = B.re * sp + A.re * cp; fi = B.im * sp + A.im * cp;
      }
    }
    const double s2 = a.dx * a.dx;
    sum[0] += (tr * er - ti * ei) * s2 * ts * sign;
    sum[1] += (tr * ei + ti * er) * s2 * ts * sign;
    sum[2] += (fr * er - fi * ei) * s2 * fs * sign;
    sum[3] += (fr * ei + fi * er) * s2 * fs * sign;
  }
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = sum[q];
  __syncthreads ();
  for (int w = 128; w > 0; w >>= 1)
  {
    if ((int) threadIdx.x < w)
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads ();
  }
  if (threadIdx.x < 4) gst (a.part + ((size_t) angle * 12 + pass) * 4 + threadIdx.x, red[threadIdx.x][0]);
}

// TF/SF predicate mask of one component for every cell (set-up check export).
// Scattered field of one component (Scheme::makeGridScattered, Scheme.inc.h:4113-4228): the
// newest field minus the incident plane wave at the component's grid points inside the TF/SF
// box [L, R] (SchemeHelper::doSkipMakeScattered, SchemeHelper.h:905-911: points outside keep
// the total field).  E takes EInc, H the newest HInc buffer (hsel), interpolated as the TF/SF
// correction does.
// Cells [q0, q0 + count) of the flat grid; out[q - q0] receives the value in the field
// precision, widened to FP64.
// The incident value a cell of component COMP holds inside [L, R] of the TF/SF box (at the
// component's position in layout LAY), 0 outside.
template <typename T, int COMP, int DIM, int LAY>
__device__ double scattered_incident (const StepArgs<T> &a, int i, int j, int k, const T *inc)
{
  bool in = twice_coord<COMP, 0, LAY> (i) >= 2 * a.tl[0] && twice_coord<COMP, 0, LAY> (i) <= 2 * a.tr[0]
            && twice_coord<COMP, 1, LAY> (j) >= 2 * a.tl[1] && twice_coord<COMP, 1, LAY> (j) <= 2 * a.tr[1];
  if (DIM == 3) in = in && twice_coord<COMP, 2, LAY> (k) >= 2 * a.tl[2] && twice_coord<COMP, 2, LAY> (k) <= 2 * a.tr[2];
  return in ? incident_at<T, COMP, DIM, LAY> (a, i, j, k, inc, COMP >= 3 ? 0.5 : 0.0) : 0.0;
}

template <typename T, int COMP, int DIM>
__global__ void scattered_kernel (StepArgs<T> a, const T *f, double *out, int hsel, int64_t q0, int64_t count)
{
  const int nzz = DIM == 3 ? a.nz : 1;
  const int64_t u = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= count) return;
  const int64_t q = q0 + u;
  const int k = (int) (q % nzz);
  const int j = (int) ((q / nzz) % a.ny);
  const int i = (int) (q / ((int64_t) nzz * a.ny));
  double v = (double) f[q];
  const T *inc = COMP >= 3 ? a.tg->hinc[hsel] : a.tg->einc;
  if (a.layout == 0) v -= scattered_incident<T, COMP, DIM, 0> (a, i, j, k, inc);
  else v -= scattered_incident<T, COMP, DIM, 1> (a, i, j, k, inc);
  out[u] = (double) (T) v;
}

template <typename T>
__global__ void tfsf_mask_kernel (StepArgs<T> a, int comp, int dim, uint8_t *out)
{
  const int64_t n = (int64_t) a.nx * a.ny * a.nz;
  for (int64_t q = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; q < n; q += (int64_t) gridDim.x * blockDim.x)
  {
    const int k = (int) (q % a.nz);
    const int j = (int) ((q / a.nz) % a.ny);
    const int i = (int) (q / ((int64_t) a.nz * a.ny));
    uint8_t m = 0;
    if (dim == 3 && a.layout == 1)
    {
      switch (comp)
      {
        case 0: m = tfsf_bits<T, 0, 3, 1> (a, i, j, k); break;
        case 1: m = tfsf_bits<T, 1, 3, 1> (a, i, j, k); break;
        case 2: m = tfsf_bits<T, 2, 3, 1> (a, i, j, k); break;
        case 3: m = tfsf_bits<T, 3, 3, 1> (a, i, j, k); break;
        case 4: m = tfsf_bits<T, 4, 3, 1> (a, i, j, k); break;
        default: m = tfsf_bits<T, 5, 3, 1> (a, i, j, k); break;
      }
    }
    else if (dim == 3)
    {
      switch (comp)
      {
        case 0: m = tfsf_bits<T, 0, 3> (a, i, j, k); break;
        case 1: m = tfsf_bits<T, 1, 3> (a, i, j, k); break;
        case 2: m = tfsf_bits<T, 2, 3> (a, i, j, k); break;
        case 3: m = tfsf_bits<T, 3, 3> (a, i, j, k); break;
        case 4: m = tfsf_bits<T, 4, 3> (a, i, j, k); break;
        default: m = tfsf_bits<T, 5, 3> (a, i, j, k); break;
      }
    }
    else
    {
      switch (comp)
      {
        case 0: m = tfsf_bits<T, 0, 2> (a, i, j, k); break;
        case 1: m = tfsf_bits<T, 1, 2> (a, i, j, k); break;
        case 2: m = tfsf_bits<T, 2, 2> (a, i, j, k); break;
        case 3: m = tfsf_bits<T, 3, 2> (a, i, j, k); break;
        case 4: m = tfsf_bits<T, 4, 2> (a, i, j, k); break;
        default: m = tfsf_bits<T, 5, 2> (a, i, j, k); break;
      }
    }
    gst (out + q, m);
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
  template __global__ void e_half_3d_hc<T, 0> (StepArgs<T>);                               \
  template __global__ void e_half_3d_hc<T, 1> (StepArgs<T>);                               \
  template __global__ void h_half_3d_hc<T, 0> (StepArgs<T>);                               \
  template __global__ void h_half_3d_hc<T, 1> (StepArgs<T>);                               \
  template __global__ void e_half_tez<T, 0> (StepArgs<T>);                                 \
  template __global__ void e_half_tez<T, 1> (StepArgs<T>);                                 \
  template __global__ void e_half_tez<T, 2> (StepArgs<T>);                                 \
  template __global__ void h_half_tez<T, 0> (StepArgs<T>);                                 \
  template __global__ void h_half_tez<T, 1> (StepArgs<T>);                                 \
  template __global__ void h_half_tez<T, 2> (StepArgs<T>);                                 \
  template __global__ void e_half_tmz<T, 0> (StepArgs<T>);                                 \
  template __global__ void e_half_tmz<T, 1> (StepArgs<T>);                                 \
  template __global__ void e_half_tmz<T, 2> (StepArgs<T>);                                 \
  template __global__ void h_half_tmz<T, 0> (StepArgs<T>);                                 \
  template __global__ void h_half_tmz<T, 1> (StepArgs<T>);                                 \
  template __global__ void h_half_tmz<T, 2> (StepArgs<T>);                                 \
  template __global__ void batch_tmz_step<T> (const StepArgs<T> *, const int4 *, int, const int *, const T *, int); \
  template __global__ void batch_tmz_step2<T> (const StepArgs<T> *, const int4 *, int, const int *, const T *, int); \
  template __global__ void batch_tmz_sources<T> (const StepArgs<T> *, int, const int *, const T *, int); \
  template __global__ void incident_step<T> (T *, const T *, T *, int64_t, int, double, double, double); \
  template __global__ void step_3d_fused<T, FUSED_BY (T), 0, 16> (StepArgs<T>);            \
  template __global__ void step_3d_fused<T, FUSED_BY (T), 1, 16> (StepArgs<T>);            \
  template __global__ void step_3d_fused<T, FUSED_BY (T), 2, 16> (StepArgs<T>);            \
  template __global__ void step_3d_fused<T, FUSED_BY (T), 3, 16> (StepArgs<T>);            \
  template __global__ void step_3d_fused<T, FUSED_BY (T), 0, 12> (StepArgs<T>);            \
  template __global__ void step_3d_fused<T, FUSED_BY (T), 1, 12> (StepArgs<T>);            \
  template __global__ void step_3d_fused<T, FUSED_BY (T), 2, 12> (StepArgs<T>);            \
  template __global__ void step_3d_fused<T, FUSED_BY (T), 3, 12> (StepArgs<T>);            \
  template __global__ void tfsf_ztable<T> (StepArgs<T>, double *);                         \
  template __global__ void point_source<T> (T *, int64_t, double);                         \
  template __global__ void tfsf_mask_kernel<T> (StepArgs<T>, int, int, uint8_t *); \
  template __global__ void test_fill<T> (StepArgs<T>, T *, int, int, double, int); \
  template __global__ void ntff_faces<T> (NtffArgs<T>); \
  template __global__ void test_norm<T> (StepArgs<T>, const T *, const T *, int, int, double, int, int, int, int, int, int, double *); \
  template __global__ void fine_average<T> (const T *, int64_t, int64_t, int64_t, int, int, int, int, int, T *, T *, T *); \
  template __global__ void scattered_kernel<T, 0, 3> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 1, 3> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 2, 3> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 3, 3> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 4, 3> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 5, 3> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 0, 2> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 1, 2> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 5, 2> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 2, 2> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 3, 2> (StepArgs<T>, const T *, double *, int, int64_t, int64_t); \
  template __global__ void scattered_kernel<T, 4, 2> (StepArgs<T>, const T *, double *, int, int64_t, int64_t);

YEE_INST (float)
YEE_INST (double)

}  // namespace yee
