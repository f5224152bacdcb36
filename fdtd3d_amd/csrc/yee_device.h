// Device helpers shared by the Yee kernels: global-memory accessors, the E_CENTERED component
// tables, the TF/SF predicates and incident-wave interpolation, the PML-slab test and the
// material averages.  Included by yee_kernels.hip and yee_adjoint.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/yee.h"
#include "yee_kernels.h"

namespace yee {

// Global-memory accessors: the state pointers live in the kernel argument block, where the
// compiler cannot see their address space; casting to address space 1 keeps every access
// a global_load/global_store instead of a flat one.
template <typename T>
__device__ __forceinline__ T gld (const T *p)
{
  return *(const __attribute__ ((address_space (1))) T *) (p);
}

template <typename T>
__device__ __forceinline__ void gst (T *p, T v)
{
  *(__attribute__ ((address_space (1))) T *) (p) = v;
}

// Row-relative accessors for the fused kernel: `row` is a uniform base pinned to scalar
// registers (plane_row), `boff` the lane's byte offset and `delta` an element offset folded
// into the instruction.  The caller passes offsets taken with fresh_offsets (or an equivalent
// empty asm) in the same block, so that each access is selected as scalar base + 32-bit lane
// offset + immediate, with no 64-bit vector address.
template <typename T>
__device__ __forceinline__ T rld (const T *row, uint32_t boff, int delta = 0)
{
  typedef const __attribute__ ((address_space (1))) char gchar;
  typedef const __attribute__ ((address_space (1))) T gT;
  gchar *p = (gchar *) row + boff;
  return *((gT *) p + delta);
}

template <typename T>
__device__ __forceinline__ void rst (T *row, uint32_t boff, T v)
{
  typedef __attribute__ ((address_space (1))) char gchar;
  typedef __attribute__ ((address_space (1))) T gT;
  gchar *p = (gchar *) row + boff;
  *(gT *) p = v;
}

// The value of lane + 1 of the wavefront (lane 63: `edge`): a DPP wave shift (wave_shl:1),
// one vector op, with no LDS traffic or wait.
__device__ __forceinline__ float lane_next (float v, float edge)
{
  return __builtin_bit_cast (float, __builtin_amdgcn_update_dpp (__builtin_bit_cast (int, edge),
                                                                __builtin_bit_cast (int, v), 0x130, 0xf, 0xf, false));
}

__device__ __forceinline__ double lane_next (double v, double edge)
{
  const int2 vv = __builtin_bit_cast (int2, v), ee = __builtin_bit_cast (int2, edge);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp (ee.x, vv.x, 0x130, 0xf, 0xf, false);
  r.y = __builtin_amdgcn_update_dpp (ee.y, vv.y, 0x130, 0xf, 0xf, false);
  return __builtin_bit_cast (double, r);
}

// The value of lane - 1 (lane 0: `edge`): DPP wave_shr:1.
__device__ __forceinline__ float lane_prev (float v, float edge)
{
  return __builtin_bit_cast (float, __builtin_amdgcn_update_dpp (__builtin_bit_cast (int, edge),
                                                                __builtin_bit_cast (int, v), 0x138, 0xf, 0xf, false));
}

__device__ __forceinline__ double lane_prev (double v, double edge)
{
  const int2 vv = __builtin_bit_cast (int2, v), ee = __builtin_bit_cast (int2, edge);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp (ee.x, vv.x, 0x138, 0xf, 0xf, false);
  r.y = __builtin_amdgcn_update_dpp (ee.y, vv.y, 0x138, 0xf, 0xf, false);
  return __builtin_bit_cast (double, r);
}

// Cell of a half-step thread.  The tiles (blockDim.x z cells, from the aligned origin tz0, by
// blockDim.y rows, one x plane) are numbered plane-fastest, and the workgroups, which the
// hardware hands to the 8 XCDs round-robin, are mapped so that each XCD takes one contiguous
// band of tile numbers: the next plane of a tile then runs on the same XCD right after it,
// and its x-neighbour plane (and the y-neighbour rows of the next tile row) come from that
// XCD's L2 instead of HBM.  Returns false for threads outside the run box.
template <typename T>
__device__ __forceinline__ bool half_cell (const StepArgs<T> &a, int &i, int &j, int &k)
{
  const int b = (int) blockIdx.x;
  const int u = (b & 7) * a.per_xcd + (b >> 3);
  if (u >= a.ntile) return false;
  const int t = u / a.ntx;
  i = a.run_lo[0] + (u - t * a.ntx);
  const int ty = t / a.ntz;
  k = a.tz0 + (t - ty * a.ntz) * (int) blockDim.x + (int) threadIdx.x;
  j = a.run_lo[1] + ty * (int) blockDim.y + (int) threadIdx.y;
  return j < a.run_hi[1] && k >= a.run_lo[2] && k < a.run_hi[2];
}

// Direct-update coefficient cb = dt / (m m0 dx) of computeCaCb (InternalScheme.inc.h:2340-2344).
// FP64 divides exactly (the reference's operations); FP32 multiplies by the hardware reciprocal
// (1 ulp), several times cheaper than an IEEE division and far inside the FP32 tolerance.
// Every kernel that updates or transposes a direct cell uses this one function.
__device__ __forceinline__ double direct_cb (double dt, double m, double m0, double dx)
{
  return dt / (m * m0 * dx);
}

__device__ __forceinline__ float direct_cb (float dt, float m, float m0, float dx)
{
  return dt * __builtin_amdgcn_rcpf (m * m0 * dx);
}

// Component tables (the host CompDesc plus the circuit offsets of YeeGridLayout.cpp:1937-1961,
// E_CENTERED, and :1991-2015, H_CENTERED).  Every lookup below is resolved at compile time:
// runtime indexing of small arrays would push them to scratch memory.
//
// LAY selects the Yee layout: 0 = E_CENTERED, 1 = H_CENTERED.  H_CENTERED puts each E
// component where E_CENTERED puts the H component of the same axis and the reverse
// (YeeGridLayout.cpp:1962-1967, 2016-2021); every position-dependent rule (stencil offsets,
// TF/SF predicates and corrected slot, material averages, sigma positions) follows the
// position, i.e. the component's role: "H-like" for H under E_CENTERED and E under H_CENTERED.
template <int COMP, int LAY>
constexpr bool role_h ()
{
  return (COMP >= 3) != (LAY != 0);
}

struct Tab
{
  // slot directions (LEFT=0, RIGHT=1, DOWN=2, UP=3, BACK=4, FRONT=5) of d11, d12, d21, d22
  static constexpr int dir[6][4] = {
    { 2, 3, 4, 5 },   // Ex: DOWN UP BACK FRONT
    { 4, 5, 0, 1 },   // Ey: BACK FRONT LEFT RIGHT
    { 0, 1, 2, 3 },   // Ez: LEFT RIGHT DOWN UP
    { 4, 5, 2, 3 },   // Hx: BACK FRONT DOWN UP
    { 0, 1, 4, 5 },   // Hy: LEFT RIGHT BACK FRONT
    { 2, 3, 0, 1 },   // Hz: DOWN UP LEFT RIGHT
  };
  static constexpr int opp[6][2] = { { 5, 4 }, { 3, 5 }, { 4, 3 }, { 1, 2 }, { 2, 0 }, { 0, 1 } };
  // offsets [slot][axis] of the four opposite-field operands
  static constexpr int off[6][4][3] = {
    { { 0, -1, 0 }, { 0, 0, 0 }, { 0, 0, -1 }, { 0, 0, 0 } },
    { { 0, 0, -1 }, { 0, 0, 0 }, { -1, 0, 0 }, { 0, 0, 0 } },
    { { -1, 0, 0 }, { 0, 0, 0 }, { 0, -1, 0 }, { 0, 0, 0 } },
    { { 0, 0, 0 }, { 0, 0, 1 }, { 0, 0, 0 }, { 0, 1, 0 } },
    { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 0, 0 }, { 0, 0, 1 } },
    { { 0, 0, 0 }, { 0, 1, 0 }, { 0, 0, 0 }, { 1, 0, 0 } },
  };
  // the same under H_CENTERED: E reads H at {0, +1}, H reads E at {-1, 0}
  static constexpr int offh[6][4][3] = {
    { { 0, 0, 0 }, { 0, 1, 0 }, { 0, 0, 0 }, { 0, 0, 1 } },
    { { 0, 0, 0 }, { 0, 0, 1 }, { 0, 0, 0 }, { 1, 0, 0 } },
    { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 0, 0 }, { 0, 1, 0 } },
    { { 0, 0, -1 }, { 0, 0, 0 }, { 0, -1, 0 }, { 0, 0, 0 } },
    { { -1, 0, 0 }, { 0, 0, 0 }, { 0, 0, -1 }, { 0, 0, 0 } },
    { { 0, -1, 0 }, { 0, 0, 0 }, { -1, 0, 0 }, { 0, 0, 0 } },
  };
};

template <int COMP, int SLOT, int AX, int LAY>
constexpr int slot_off ()
{
  return LAY ? Tab::offh[COMP][SLOT][AX] : Tab::off[COMP][SLOT][AX];
}

// sigma axes (curl, sigma4, sigma5) per component
template <int COMP> struct SigAxes;
template <> struct SigAxes<0> { static constexpr int c = 1, s4 = 0, s5 = 2; };
template <> struct SigAxes<1> { static constexpr int c = 2, s4 = 1, s5 = 0; };
template <> struct SigAxes<2> { static constexpr int c = 0, s4 = 2, s5 = 1; };
template <> struct SigAxes<3> { static constexpr int c = 1, s4 = 0, s5 = 2; };
template <> struct SigAxes<4> { static constexpr int c = 2, s4 = 1, s5 = 0; };
template <> struct SigAxes<5> { static constexpr int c = 0, s4 = 2, s5 = 1; };

template <int AX>
__device__ __forceinline__ int pick (int i, int j, int k)
{
  return AX == 0 ? i : (AX == 1 ? j : k);
}

// Twice the real coordinate of component COMP along axis AX at integer index p.
template <int COMP, int AX, int LAY = 0>
__device__ __forceinline__ int twice_coord (int p)
{
  constexpr bool h = role_h<COMP, LAY> ();
  constexpr int own = COMP % 3;
  return 2 * p + (h ? (AX == own ? 1 : 2) : (AX == own ? 2 : 1));
}

// doNeedTFSFUpdate<comp>Border(pos, dir) on doubled integer coordinates (YeeGridLayout.cpp:272-1740;
// the FP tests c == b +- 0.5, c in (L - 0.1, R + 0.1), c in (L + 0.4, R - 0.4) are exact on half-integers).
// The H_CENTERED predicates (YeeGridLayout.cpp:44-1175) are those of the component role: an
// H_CENTERED E component tests what an E_CENTERED H component at the same point tests.
template <typename T, int COMP, int DIR, int DIM, int LAY = 0>
__device__ __forceinline__ bool tfsf_pred (const StepArgs<T> &a, int i, int j, int k)
{
  constexpr int b = DIR >> 1;
  constexpr bool low = (DIR & 1) == 0;
  constexpr int ax = COMP % 3;
  constexpr bool h = role_h<COMP, LAY> ();
  constexpr int t = 3 - b - ax;
  if (b == ax || b >= DIM) return false;
  const int cb2 = twice_coord<COMP, b, LAY> (pick<b> (i, j, k));
  bool ok = h ? (low ? cb2 == 2 * a.tl[b] : cb2 == 2 * a.tr[b])
              : (low ? cb2 == 2 * a.tl[b] - 1 : cb2 == 2 * a.tr[b] + 1);
  if (ax < DIM)
  {
    const int c2 = twice_coord<COMP, ax, LAY> (pick<ax> (i, j, k));
    ok = ok && (h ? (c2 >= 2 * a.tl[ax] + 1 && c2 <= 2 * a.tr[ax] - 1) : (c2 >= 2 * a.tl[ax] && c2 <= 2 * a.tr[ax]));
  }
  if (t < DIM)
  {
    const int c2 = twice_coord<COMP, t, LAY> (pick<t> (i, j, k));
    ok = ok && (h ? (c2 >= 2 * a.tl[t] && c2 <= 2 * a.tr[t]) : (c2 >= 2 * a.tl[t] + 1 && c2 <= 2 * a.tr[t] - 1));
  }
  return ok;
}

// Runtime-direction form (mask export only).
template <typename T, int COMP, int DIM, int LAY = 0>
__device__ uint8_t tfsf_bits (const StepArgs<T> &a, int i, int j, int k)
{
  return (uint8_t) ((tfsf_pred<T, COMP, 0, DIM, LAY> (a, i, j, k) ? 1 : 0) | (tfsf_pred<T, COMP, 1, DIM, LAY> (a, i, j, k) ? 2 : 0)
                    | (tfsf_pred<T, COMP, 2, DIM, LAY> (a, i, j, k) ? 4 : 0) | (tfsf_pred<T, COMP, 3, DIM, LAY> (a, i, j, k) ? 8 : 0)
                    | (tfsf_pred<T, COMP, 4, DIM, LAY> (a, i, j, k) ? 16 : 0) | (tfsf_pred<T, COMP, 5, DIM, LAY> (a, i, j, k) ? 32 : 0));
}

// A cell can carry a TF/SF correction only on the planes L-1, R-1, R of some axis while
// lying inside [L-1, R] on every axis.
template <typename T, int DIM>
__device__ __forceinline__ bool tfsf_near (const StepArgs<T> &a, int i, int j, int k)
{
  const int p[3] = { i, j, k };
  bool face = false;
#pragma unroll
  for (int ax = 0; ax < DIM; ++ax)
  {
    if (p[ax] < a.tl[ax] - 1 || p[ax] > a.tr[ax]) return false;
    face |= (p[ax] == a.tl[ax] - 1) | (p[ax] == a.tr[ax] - 1) | (p[ax] == a.tr[ax]);
  }
  return face;
}

// Incident value of component OPP at integer position (i, j, k): approximateIncidentWave +
// get<opp>FromIncident (InternalScheme.cpp:375-407, InternalScheme.inc.h:500-515).
template <typename T, int OPP, int DIM, int LAY = 0>
__device__ double incident_at (const StepArgs<T> &a, int i, int j, int k, const T *inc, double ddiff)
{
  const TfsfGeom<T> &tg = *a.tg;
  const double x = 0.5 * (double) twice_coord<OPP, 0, LAY> (i) - tg.zinc[0];
  const double y = 0.5 * (double) twice_coord<OPP, 1, LAY> (j) - tg.zinc[1];
  double d;
  if (DIM == 3)
  {
    const double z = 0.5 * (double) twice_coord<OPP, 2, LAY> (k) - tg.zinc[2];
    d = x * tg.S1 * tg.C2 + y * tg.S1 * tg.S2 + z * tg.C1 - ddiff;
  }
  else
  {
    d = x * tg.C2 + y * tg.S2 - ddiff;
  }
  const double c1 = (double) (long long) d;
  long long i1 = (long long) c1;
  // the reference asserts d > 0; clamp so a degenerate box can never read out of bounds
  if (i1 < 0) i1 = 0;
  if (i1 > tg.inc_len - 2) i1 = tg.inc_len - 2;
  const double p2 = d - c1;
  const double p1 = 1 - p2;
  const double v = (double) gld (inc + i1) * p1 + (double) gld (inc + i1 + 1) * p2;
  return v * tg.proj[OPP];
}

// One operand pair of calculateTFSF.  E components (E_CENTERED) correct the operand opposite
// to the border side; H components the operand on the border side, with the sign flipped.
// The correction of one operand pair as values to subtract (in FP64) from the lo / hi slot:
// 0 for a slot that is not corrected.  Returns false when neither slot is.
template <typename T, int COMP, int PAIR, int DIM>
__device__ __forceinline__ bool tfsf_pair_diff (const StepArgs<T> &a, int i, int j, int k, double &d_lo, double &d_hi)
{
  constexpr bool h = COMP >= 3;
  constexpr int OPP = Tab::opp[COMP][PAIR];
  d_lo = d_hi = 0.0;
  const bool lo = tfsf_pred<T, COMP, Tab::dir[COMP][2 * PAIR], DIM> (a, i, j, k);
  const bool hi = tfsf_pred<T, COMP, Tab::dir[COMP][2 * PAIR + 1], DIM> (a, i, j, k);
  if (!lo && !hi) return false;
  const T *inc = h ? a.tg->einc : a.tg->hinc[a.hinc_sel];
  const double ddiff = h ? 0.0 : 0.5;
  const bool use_hi_slot = h ? !lo : lo;
  constexpr int s_lo = 2 * PAIR, s_hi = 2 * PAIR + 1;
  double diff;
  if (use_hi_slot)
    diff = incident_at<T, OPP, DIM> (a, i + Tab::off[COMP][s_hi][0], j + Tab::off[COMP][s_hi][1],
                                     k + (DIM == 3 ? Tab::off[COMP][s_hi][2] : 0), inc, ddiff);
  else
    diff = incident_at<T, OPP, DIM> (a, i + Tab::off[COMP][s_lo][0], j + Tab::off[COMP][s_lo][1],
                                     k + (DIM == 3 ? Tab::off[COMP][s_lo][2] : 0), inc, ddiff);
  if (h) diff *= -1.0;
  if (use_hi_slot) d_hi = diff;
  else d_lo = diff;
  return true;
}

// The incident array and its half-step shift follow the physical type (E reads HInc, H reads
// EInc); the corrected slot and the sign follow the role (isRevertVals and the layout's sign
// flip, InternalScheme.inc.h:1405-1835): an H-like component corrects the slot on the border
// side with the sign flipped, an E-like one the slot opposite to it.
template <typename T, int COMP, int PAIR, int DIM, int LAY = 0>
__device__ __forceinline__ void tfsf_pair (const StepArgs<T> &a, int i, int j, int k, T &v_lo, T &v_hi)
{
  constexpr bool h = COMP >= 3;
  constexpr bool rh = role_h<COMP, LAY> ();
  constexpr int OPP = Tab::opp[COMP][PAIR];
  const bool lo = tfsf_pred<T, COMP, Tab::dir[COMP][2 * PAIR], DIM, LAY> (a, i, j, k);
  const bool hi = tfsf_pred<T, COMP, Tab::dir[COMP][2 * PAIR + 1], DIM, LAY> (a, i, j, k);
  if (!lo && !hi) return;
  const T *inc = h ? a.tg->einc : a.tg->hinc[a.hinc_sel];
  const double ddiff = h ? 0.0 : 0.5;
  // slot 2*PAIR is "lo" (d11 / d21), slot 2*PAIR+1 is "hi" (d12 / d22)
  const bool use_hi_slot = rh ? !lo : lo;
  constexpr int s_lo = 2 * PAIR, s_hi = 2 * PAIR + 1;
  double diff;
  if (use_hi_slot)
    diff = incident_at<T, OPP, DIM, LAY> (a, i + slot_off<COMP, s_hi, 0, LAY> (), j + slot_off<COMP, s_hi, 1, LAY> (),
                                          k + (DIM == 3 ? slot_off<COMP, s_hi, 2, LAY> () : 0), inc, ddiff);
  else
    diff = incident_at<T, OPP, DIM, LAY> (a, i + slot_off<COMP, s_lo, 0, LAY> (), j + slot_off<COMP, s_lo, 1, LAY> (),
                                          k + (DIM == 3 ? slot_off<COMP, s_lo, 2, LAY> () : 0), inc, ddiff);
  if (rh) diff *= -1.0;
  if (use_hi_slot) v_hi = (T) ((double) v_hi - diff);
  else v_lo = (T) ((double) v_lo - diff);
}

// calculateTFSF for one component at cell (i, j, k): corrects the operands v11, v12, v21, v22.
template <typename T, int COMP, int DIM, int LAY = 0>
__device__ __forceinline__ void tfsf_apply (const StepArgs<T> &a, int i, int j, int k, T v[4])
{
  tfsf_pair<T, COMP, 0, DIM, LAY> (a, i, j, k, v[0], v[1]);
  tfsf_pair<T, COMP, 1, DIM, LAY> (a, i, j, k, v[2], v[3]);
}

// Inside the Drude box: the field cells whose 2- or 4-cell averages reach a nonzero Drude cell.
template <typename T>
__device__ __forceinline__ bool in_drude_box (const StepArgs<T> &a, const int p[3], int dim)
{
  bool in = true;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
    if (ax < dim) in &= (p[ax] >= a.box_lo[ax]) & (p[ax] < a.box_hi[ax]);
  return in;
}

template <typename T>
__device__ __forceinline__ bool is_special (const StepArgs<T> &a, const int p[3], int dim, int mode)
{
  if (mode == 0) return false;
  bool s = false;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
    if (ax < dim) s |= (p[ax] < a.slab_lo[ax]) | (p[ax] >= a.slab_hi[ax]);
  if (mode == 2) s |= in_drude_box (a, p, dim);
  return s;
}


template <typename T>
__device__ __forceinline__ T curl (const T v[4])
{
  return v[1] - v[0] - v[3] + v[2];
}

// Analytic test problems of the reference's real build (Source/Scheme/CallBack.cpp:28-143;
// SQR (a) = a * a, the constants of PhysicsConst.h:30-32), evaluated in FP64 with the
// reference's operations: field values and the right-side currents J (E components) / M (H
// components).  Components a kind does not define read as 0 (Scheme::initCallBacks sets none).
__device__ __forceinline__ double test_field (int kind, int comp, double x, double y, double z, double t)
{
  const double c = 299792458.0, mu0 = 4 * 3.14159265358979323846 * 0.0000001, e0 = 1 / (mu0 * (c * c));
  if (kind == YEE_FN_POLINOM2)
  {
    const int a = comp % 3;
    const double u = a == 0 ? y : x, v = a == 2 ? y : z;
    const double e = (c * c) * t * ((u * u) + (v * v));
    return comp < 3 ? e : e0 * e;
  }
  if (comp != YEE_EZ && comp != YEE_HY) return 0.0;
  if (kind == YEE_FN_POLINOM1)
  {
    const double ez = (c * c) * (x * x) * t;
    return comp == YEE_EZ ? ez : e0 * ez;
  }
  if (kind == YEE_FN_POLINOM3)
  {
    const double ez = (c * c) * ((x * t) * (x * t));
    return comp == YEE_EZ ? ez : e0 * ez;
  }
  if (kind == YEE_FN_SIN1) return comp == YEE_EZ ? c * sin (t - x) : -c * sin (t - x);
  return 0.0;
}

__device__ __forceinline__ double test_rhs (int kind, int comp, double x, double y, double z, double t)
{
  const double c = 299792458.0, mu0 = 4 * 3.14159265358979323846 * 0.0000001, e0 = 1 / (mu0 * (c * c));
  if (kind == YEE_FN_POLINOM2)
  {
    const double p[3] = { x, y, z };
    const int a = comp % 3;
    const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
    if (comp < 3)
    {
      // Jx: 2 t (z - y) + y^2 + z^2, Jy: 2 t (x - z) + x^2 + z^2, Jz: 2 t (y - x) + x^2 + y^2
      const double d = a == 0 ? (p[2] - p[1]) : (a == 1 ? (p[0] - p[2]) : (p[1] - p[0]));
      return -(c * c) * e0 * (2 * t * d + (p[u] * p[u]) + (p[v] * p[v]));
    }
    // Mx: 2 t (y - z), My: 2 t (z - x), Mz: 2 t (x - y)
    const double d = a == 0 ? (p[1] - p[2]) : (a == 1 ? (p[2] - p[0]) : (p[0] - p[1]));
    return -(c * c) * (2 * t * d + e0 * mu0 * ((p[u] * p[u]) + (p[v] * p[v])));
  }
  if (comp != YEE_EZ && comp != YEE_HY) return 0.0;
  if (kind == YEE_FN_POLINOM1)
  {
    if (comp == YEE_EZ) return -(c * c) * e0 * (-2 * x * t + (x * x));
    return -(c * c) * (-2 * x * t + (x * x) * e0 * mu0);
  }
  if (kind == YEE_FN_POLINOM3)
  {
    if (comp == YEE_EZ) return -(c * c) * 2 * e0 * x * t * (x - t);
    return -(c * c) * 2 * x * t * (e0 * mu0 * x - t);
  }
  return 0.0;
}

// Exact solutions of the 1D plane wave meeting a dielectric half-space (Tests/suite t7.x;
// CallBack.cpp:280-760): kind YEE_FN_EXP1 i e^{i arg}, EXP2 -i/3 e^{i arg}, EXP3 2i/3 e^{i arg},
// arg = k (coord - d dx) - w (t - t0) with the host-computed k and phase origin d; H = E times
// (k / (mu0 w)).  The complex operations of the reference's CComplex (i * exponent (i arg) =
// (0 c - 1 s, 1 c + 0 s)), then the sign of the pair's wrapper.  Per pair (ExHy, ExHz, EyHx,
// EyHz, EzHx, EzHy): the line axis, the E and H components and their signs.
__device__ __forceinline__ int exp_axis (int pair)
{
  return pair == 0 || pair == 2 ? 2 : (pair == 1 || pair == 4 ? 1 : 0);
}

__device__ __forceinline__ bool exp_defines (int pair, int comp)
{
  const int e = pair < 2 ? YEE_EX : (pair < 4 ? YEE_EY : YEE_EZ);
  const int h = pair == 0 || pair == 5 ? YEE_HY : (pair == 1 || pair == 3 ? YEE_HZ : YEE_HX);
  return comp == e || comp == h;
}

template <typename T>
__device__ __forceinline__ void test_exp (const StepArgs<T> &a, int kind, int comp, const double r[3], double t,
                                          double &re, double &im)
{
  const int axis = exp_axis (a.exp_pair);
  const double arg = a.exp_k * (r[axis] - a.exp_d * a.fn_dx[axis]) - a.exp_w * (t - a.exp_t0);
  const double c = cos (arg), s = sin (arg);
  if (kind == YEE_FN_EXP1)
  {
    re = 0.0 * c - 1.0 * s;
    im = 1.0 * c + 0.0 * s;
  }
  else if (kind == YEE_FN_EXP2)
  {
    // (-i) (c, s) times 1, over 3
    re = ((-0.0 * c - (-1.0) * s) * 1.0) / 3.0;
    im = ((-1.0 * c + -0.0 * s) * 1.0) / 3.0;
  }
  else
  {
    re = ((0.0 * c - 1.0 * s) * 2.0) / 3.0;
    im = ((1.0 * c + 0.0 * s) * 2.0) / 3.0;
  }
  if (comp >= 3)
  {
    re = re * a.exp_m;
    im = im * a.exp_m;
  }
  // signs of the wrappers: ExHy (-, -), ExHz (+, -), EyHx (-, +), EyHz (-, -), EzHx (+, +), EzHy (+, -)
  const int p = a.exp_pair;
  const bool neg = comp >= 3 ? (p == 0 || p == 1 || p == 3 || p == 5) : (p == 0 || p == 2 || p == 3);
  if (neg)
  {
    re = -re;
    im = -im;
  }
}

// The real coordinate (m) of component COMP at cell p: (p + the component's minimum coordinate,
// E_CENTERED: 1 along an E component's own axis and 0.5 across it, the reverse for H; H_CENTERED
// the reverse of both) * fn_dx.
template <typename T>
__device__ __forceinline__ void test_coords (const StepArgs<T> &a, int comp, const int p[3], double r[3])
{
  const int axis = comp % 3;
  const bool h = (comp >= 3) != (a.layout != 0);
  for (int d = 0; d < 3; ++d) r[d] = ((double) p[d] + ((d == axis) != h ? 1.0 : 0.5)) * a.fn_dx[d];
}

// The curl of the half-step kernels with the right side of an analytic test problem
// subtracted last (calcField, InternalScheme.inc.h:765-772: o12 - o11 - o22 + o21 - J dx).
template <typename T, int COMP>
__device__ __forceinline__ T curl_rhs (const StepArgs<T> &a, const T v[4], const int p[3])
{
  const T cv = curl (v);
  if (!a.rhs_kind) return cv;
  double r[3];
  test_coords (a, COMP, p, r);
  return cv - (T) test_rhs (a.rhs_kind, COMP, r[0], r[1], r[2], a.rhs_time) * a.dx;
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

}  // namespace yee
