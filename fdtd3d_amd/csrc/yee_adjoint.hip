// Forward/backward pair of the 3D engine: J(eps) and dJ/d eps over a design box by the
// discrete adjoint (time-reversed transpose) of the engine's own update.
//
// No counterpart in the reference (it has no differentiation).  The forward map of one step
// is linear in the state (E, D, H, B) plus field-independent TF/SF and source terms, so the
// adjoint of a step is the transpose of its linear part, applied in reverse order (H, then E):
//   H cells:  plain   w = cb a;                      a unchanged
//             PML     mB = aB + kp a;  w = cbc mB;   aB <- cac mB - km a;  a <- ca5 a
//             with cb = dt / (m mu0 dx), kp = (p4 / mod) / dd5, km = (m4 / mod) / dd5 (the
//             coefficients of cell_value), a = adjoint of the cell's new value;
//   E cells:  a = aE - curl_E(w_H) first (the transpose of the H curl that read E), then the
//             same per-cell rule; aH -= curl_H(w_E) is applied at the start of the next
//             (earlier) reverse step, fused with its H transpose.
// A hard point source replaces the new value of its cell, so its adjoint is dropped.
// J = sum over the last `window` steps of the sum of Ez^2 on a monitor rectangle; its
// derivative 2 Ez is added to aEz before the reverse step of each of those steps.
// The update of an E cell depends on eps only through the averaged material m inside
// cb (1/m), so dE'/dm = -(E' - E) / m on the PML-free cells of the design box, and each of
// the two eps voxels of the average receives half of it.  E before and after every step on
// the design box comes from a trace recorded by the forward run (checkpoint_every = 0), or
// from re-running each segment of checkpoint_every steps from a snapshot of the full state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "yee_device.h"
#include "yee_internal.h"

namespace yee {

template <typename T>
struct AdjArgs
{
  StepArgs<T> a;              // forward set-up (sizes, materials, PML tables)
  T *af[6], *ad[6];           // adjoints of E/H and of D/B
  T *w[6];                    // curl weights of the current reverse step
  uint32_t src_mask;          // hard sources (adjoint dropped at src_q)
  int64_t src_q;
};

// The adjoint D/B and PML coefficients of a special cell, loaded before any store of the
// thread (the state pointers may alias as far as the compiler knows), so that every load of a
// cell is in flight at once.
template <typename T>
struct AdjOld
{
  T ad, cac, cbc, p4, m4, ca5, dd5;
};

template <typename T, int COMP>
__device__ __forceinline__ void adj_load (const AdjArgs<T> &g, int64_t q, const int p[3], bool special, AdjOld<T> &o)
{
  if (!special) return;
  const StepArgs<T> &a = g.a;
  const int L = a.tab_len;
  const T *tab = a.pml + (size_t) COMP * 6 * L;
  o.ad = g.ad[COMP][q];
  o.cac = tab[0 * L + p[SigAxes<COMP>::c]];
  o.cbc = tab[1 * L + p[SigAxes<COMP>::c]];
  o.p4 = tab[2 * L + p[SigAxes<COMP>::s4]];
  o.m4 = tab[3 * L + p[SigAxes<COMP>::s4]];
  o.ca5 = tab[4 * L + p[SigAxes<COMP>::s5]];
  o.dd5 = tab[5 * L + p[SigAxes<COMP>::s5]];
}

template <typename T, int COMP>
__device__ __forceinline__ T adj_cell (const AdjArgs<T> &g, int64_t q, T av, T mat, bool special, const AdjOld<T> &o)
{
  const StepArgs<T> &a = g.a;
  const T modc = COMP >= 3 ? a.mu0 : a.eps0;
  if ((g.src_mask & (1u << COMP)) && q == g.src_q) av = (T) 0;
  T w;
  if (!special)
  {
    const T cb = direct_cb (a.dt, mat, modc, a.dx);
    w = av * cb;
    g.af[COMP][q] = av;
  }
  else
  {
    const T mod = mat * modc;
    const T kp = (o.p4 / mod) / o.dd5, km = (o.m4 / mod) / o.dd5;
    const T mD = o.ad + kp * av;
    w = o.cbc * mD;
    g.ad[COMP][q] = o.cac * mD - km * av;
    g.af[COMP][q] = o.ca5 * av;
  }
  return w;
}

// aH -= curl_H(w_E) of the previous reverse step (the transpose of the E curl that read H;
// w_E is zero before the first one), then the transpose of this step's H update (H cells of
// [1, N-1)^3; tiles as half_cell).
template <typename T, int MODE>
__global__ __launch_bounds__ (256) void adj_h (AdjArgs<T> g)
{
  const StepArgs<T> &a = g.a;
  int i, j, k;
  if (!half_cell (a, i, j, k)) return;
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy + k;
  const int p[3] = { i, j, k };
  const bool special = MODE != 0 && is_special (a, p, 3, MODE);
  T mx = a.mu_u, my = a.mu_u, mz = a.mu_u;
  if (a.mu)
  {
    const T *m = a.mu;
    mx = avg4 (m[q], m[q + 1], m[q + sy], m[q + sy + 1]);
    my = avg4 (m[q], m[q + 1], m[q + sx], m[q + sx + 1]);
    mz = avg4 (m[q], m[q + sy], m[q + sx], m[q + sx + sy]);
  }
  const T *ex = g.w[0], *ey = g.w[1], *ez = g.w[2];
  // the H-update stencils applied to w_E (Hx: Ey BACK/FRONT, Ez DOWN/UP; cyclic)
  const T vx[4] = { ey[q], ey[q + 1], ez[q], ez[q + sy] };
  const T vy[4] = { ez[q], ez[q + sx], ex[q], ex[q + 1] };
  const T vz[4] = { ex[q], ex[q + sy], ey[q], ey[q + sx] };
  AdjOld<T> o3 {}, o4 {}, o5 {};
  adj_load<T, 3> (g, q, p, special, o3);
  adj_load<T, 4> (g, q, p, special, o4);
  adj_load<T, 5> (g, q, p, special, o5);
  const T a3 = g.af[3][q] - curl (vx), a4 = g.af[4][q] - curl (vy), a5 = g.af[5][q] - curl (vz);
  g.w[3][q] = adj_cell<T, 3> (g, q, a3, mx, special, o3);
  g.w[4][q] = adj_cell<T, 4> (g, q, a4, my, special, o4);
  g.w[5][q] = adj_cell<T, 5> (g, q, a5, mz, special, o5);
}

// aE -= curl_E(w_H), then the transpose of the E update (E cells of [1, N-1)^3).
template <typename T, int MODE>
__global__ __launch_bounds__ (256) void adj_e (AdjArgs<T> g)
{
  const StepArgs<T> &a = g.a;
  int i, j, k;
  if (!half_cell (a, i, j, k)) return;
  const int64_t sx = a.sx;
  const int sy = a.sy;
  const int64_t q = (int64_t) i * sx + (int64_t) j * sy + k;
  const int p[3] = { i, j, k };
  const bool special = MODE != 0 && is_special (a, p, 3, MODE);
  const T *wx = g.w[3], *wy = g.w[4], *wz = g.w[5];
  // the E-update stencils applied to w_H (Ex: Hz DOWN/UP, Hy BACK/FRONT; cyclic)
  const T vx[4] = { wz[q - sy], wz[q], wy[q - 1], wy[q] };
  const T vy[4] = { wx[q - 1], wx[q], wz[q - sx], wz[q] };
  const T vz[4] = { wy[q - sx], wy[q], wx[q - sy], wx[q] };
  T ex = a.eps_u, ey = a.eps_u, ez = a.eps_u;
  if (a.eps)
  {
    const T e = a.eps[q];
    ex = avg2 (e, a.eps[q + sx]);
    ey = avg2 (e, a.eps[q + sy]);
    ez = avg2 (e, a.eps[q + 1]);
  }
  AdjOld<T> o0 {}, o1 {}, o2 {};
  adj_load<T, 0> (g, q, p, special, o0);
  adj_load<T, 1> (g, q, p, special, o1);
  adj_load<T, 2> (g, q, p, special, o2);
  const T a0 = g.af[0][q] - curl (vx), a1 = g.af[1][q] - curl (vy), a2 = g.af[2][q] - curl (vz);
  g.w[0][q] = adj_cell<T, 0> (g, q, a0, ex, special, o0);
  g.w[1][q] = adj_cell<T, 1> (g, q, a1, ey, special, o1);
  g.w[2][q] = adj_cell<T, 2> (g, q, a2, ez, special, o2);
}

// aEz += 2 Ez on the monitor rectangle (mon: that step's recorded Ez, [x][y]).
template <typename T>
__global__ void adj_source (T *aez, const T *mon, int64_t sx, int sy, int64_t mz, int mx0, int my0, int mnx, int mny)
{
  const int u = (int) (blockIdx.x * blockDim.x + threadIdx.x);
  if (u >= mnx * mny) return;
  const int x = mx0 + u / mny, y = my0 + u % mny;
  const int64_t q = (int64_t) x * sx + (int64_t) y * sy + mz;
  aez[q] += (T) 2 * mon[u];
}

// Trace box geometry: E cells [lo, lo + n) of the three components, [comp][x][y][z].
struct TraceBox
{
  int lo[3], n[3];
  __host__ __device__ int64_t cells () const { return (int64_t) n[0] * n[1] * n[2]; }
  __host__ __device__ int64_t at (int c, int x, int y, int z) const
  {
    return ((c * (int64_t) n[0] + (x - lo[0])) * n[1] + (y - lo[1])) * n[2] + (z - lo[2]);
  }
};

// Copies E of the trace box (and Ez of the monitor rectangle when mon != null).
template <typename T>
__global__ void adj_record (const T *Ex, const T *Ey, const T *Ez, int64_t sx, int sy, TraceBox tb, T *slot)
{
  const int64_t u = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = tb.cells ();
  if (u >= nc) return;
  const int z = tb.lo[2] + (int) (u % tb.n[2]);
  const int y = tb.lo[1] + (int) ((u / tb.n[2]) % tb.n[1]);
  const int x = tb.lo[0] + (int) (u / ((int64_t) tb.n[2] * tb.n[1]));
  const int64_t q = (int64_t) x * sx + (int64_t) y * sy + z;
  slot[u] = Ex[q];
  slot[nc + u] = Ey[q];
  slot[2 * nc + u] = Ez[q];
}

template <typename T>
__global__ void adj_monitor (const T *Ez, int64_t sx, int sy, int64_t mz, int mx0, int my0, int mnx, int mny, T *out)
{
  const int u = (int) (blockIdx.x * blockDim.x + threadIdx.x);
  if (u >= mnx * mny) return;
  const int x = mx0 + u / mny, y = my0 + u % mny;
  out[u] = Ez[(int64_t) x * sx + (int64_t) y * sy + mz];
}

// grad(v) += sum over the six E cells whose material average contains voxel v of
// aE(c) * (-(E'(c) - E(c)) / m(c)) / 2, for the design voxels [dlo, dhi).
template <typename T>
__global__ void adj_grad (AdjArgs<T> g, TraceBox tb, const T *cur, const T *prev, int dlo0, int dlo1, int dlo2,
                          int dn0, int dn1, int dn2, double *grad)
{
  const StepArgs<T> &a = g.a;
  const int64_t u = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= (int64_t) dn0 * dn1 * dn2) return;
  const int vz = dlo2 + (int) (u % dn2);
  const int vy = dlo1 + (int) ((u / dn2) % dn1);
  const int vx = dlo0 + (int) (u / ((int64_t) dn2 * dn1));
  const int64_t sx = a.sx;
  const int sy = a.sy;
  double acc = 0;
  // (component, cell offset along that component's axis): cells c = v and c = v - e_axis
#pragma unroll
  for (int comp = 0; comp < 3; ++comp)
  {
#pragma unroll
    for (int back = 0; back < 2; ++back)
    {
      const int cx = vx - (comp == 0 ? back : 0), cy = vy - (comp == 1 ? back : 0), cz = vz - (comp == 2 ? back : 0);
      const int64_t qc = (int64_t) cx * sx + (int64_t) cy * sy + cz;
      const int64_t step = comp == 0 ? sx : (comp == 1 ? (int64_t) sy : 1);
      const T m = avg2 (a.eps[qc], a.eps[qc + step]);
      const int64_t tix = tb.at (comp, cx, cy, cz);
      const T de = cur[tix] - prev[tix];
      acc += (double) g.af[comp][qc] * (-(double) de / (double) m) * 0.5;
    }
  }
  grad[u] += acc;
}

// Writes the design-box values into the eps grid.
template <typename T>
__global__ void adj_set_design (T *eps, int64_t sx, int sy, int dlo0, int dlo1, int dlo2, int dn0, int dn1, int dn2,
                                const T *vals)
{
  const int64_t u = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= (int64_t) dn0 * dn1 * dn2) return;
  const int z = dlo2 + (int) (u % dn2);
  const int y = dlo1 + (int) ((u / dn2) % dn1);
  const int x = dlo0 + (int) (u / ((int64_t) dn2 * dn1));
  eps[(int64_t) x * sx + (int64_t) y * sy + z] = vals[u];
}

struct AdjointState
{
  yee_adjoint_spec sp {};
  int T = 0;
  TraceBox tb {};
  int64_t dn[3] = { 0, 0, 0 };
  int mnx = 0, mny = 0;
  std::vector<void *> allocs;
  void *af[6] = {}, *ad[6] = {}, *w[6] = {};
  void *trace = nullptr;
  int trace_slots = 0;
  void *mon = nullptr;
  double *grad = nullptr;
  void *stage = nullptr;          // design values staging (field precision)
  std::vector<void *> snap;       // checkpoint snapshots (full state + incident grids)
  std::vector<int32_t> snap_t;
  std::vector<int> snap_cur;
  size_t snap_bytes = 0;
  bool forward_done = false;
  bool last_traced = false;   // the forward recorded the trace of the last checkpoint segment
  double J = 0;
  double fwd_ms = 0, bwd_ms = 0;
};

namespace {

int adj_alloc (yee_handle *h, void **p, size_t bytes)
{
  HIP_TRY (hipMalloc (p, bytes));
  h->adj->allocs.push_back (*p);
  HIP_TRY (hipMemsetAsync (*p, 0, bytes, h->stream));
  return YEE_OK;
}

template <typename T>
AdjArgs<T> adj_args (const yee_handle *h)
{
  AdjArgs<T> g {};
  g.a = engine_args<T> (h, 0);
  for (int c = 0; c < 6; ++c)
  {
    g.af[c] = (T *) h->adj->af[c];
    g.ad[c] = (T *) h->adj->ad[c];
    g.w[c] = (T *) h->adj->w[c];
  }
  const int64_t *p = h->cfg.point_source_pos;
  g.src_mask = h->cfg.point_source & 0x3fu;
  g.src_q = (p[0] * h->n[1] + p[1]) * h->n[2] + p[2];
  return g;
}


template <typename T>
int record (yee_handle *h, int slot)
{
  AdjointState *s = h->adj;
  const int64_t nc = s->tb.cells ();
  T *dst = (T *) s->trace + (size_t) slot * 3 * nc;
  hipLaunchKernelGGL ((adj_record<T>), dim3 ((unsigned) ((nc + 255) / 256)), dim3 (256), 0, h->stream,
                     (const T *) h->f[h->cur][0], (const T *) h->f[h->cur][1], (const T *) h->f[h->cur][2],
                     h->n[1] * h->n[2], (int) h->n[2], s->tb, dst);
  HIP_TRY (hipGetLastError ());
  return YEE_OK;
}

template <typename T>
int monitor (yee_handle *h, int slot)
{
  AdjointState *s = h->adj;
  const int nm = s->mnx * s->mny;
  hipLaunchKernelGGL ((adj_monitor<T>), dim3 ((unsigned) ((nm + 255) / 256)), dim3 (256), 0, h->stream,
                     (const T *) h->f[h->cur][2], h->n[1] * h->n[2], (int) h->n[2], s->sp.monitor_z,
                     (int) s->sp.monitor_lo[0], (int) s->sp.monitor_lo[1], s->mnx, s->mny, (T *) s->mon + (size_t) slot * nm);
  HIP_TRY (hipGetLastError ());
  return YEE_OK;
}

// Snapshot k of the full state (fields, D/B, incident grids) and its step counter.
int snapshot (yee_handle *h, int k, bool restore)
{
  AdjointState *s = h->adj;
  const size_t grid = (size_t) h->ncell * h->wsize, inc = (size_t) h->inc_len * h->wsize;
  char *base = (char *) s->snap[k];
  if (restore)
  {
    h->t = s->snap_t[k];
    h->cur = s->snap_cur[k];
  }
  else
  {
    s->snap_t[k] = h->t;
    s->snap_cur[k] = h->cur;
  }
  size_t off = 0;
  auto move = [&] (void *state, size_t bytes) -> int {
    if (!state) return YEE_OK;
    if (restore) HIP_TRY (hipMemcpyAsync (state, base + off, bytes, hipMemcpyDeviceToDevice, h->stream));
    else HIP_TRY (hipMemcpyAsync (base + off, state, bytes, hipMemcpyDeviceToDevice, h->stream));
    off += bytes;
    return YEE_OK;
  };
  int rc;
  for (int c = 0; c < 6; ++c)
  {
    if ((rc = move (h->f[h->cur][c], grid))) return rc;
    if ((rc = move (h->d[h->cur][c], grid))) return rc;
  }
  if ((rc = move (h->einc, inc))) return rc;
  if ((rc = move (h->hinc[0], inc))) return rc;
  if ((rc = move (h->hinc[1], inc))) return rc;
  return YEE_OK;
}

template <typename T>
int reverse_step (yee_handle *h, int32_t t, int slot_cur, int slot_prev)
{
  AdjointState *s = h->adj;
  AdjArgs<T> g = adj_args<T> (h);
  // the interior [1, N-1)^3 in 64 x 4 tiles (half_cell)
  for (int ax = 0; ax < 3; ++ax)
  {
    g.a.run_lo[ax] = 1;
    g.a.run_hi[ax] = (int) h->n[ax] - 1;
  }
  dim3 grid;
  const dim3 block (64, 4, 1);
  if (!half_tiles (g.a, 64, 4, grid)) return fail (YEE_EINVAL, "adjoint grid too large for one launch");
  const int W = s->sp.window;
  if (t >= s->T - W)
  {
    const int nm = s->mnx * s->mny;
    hipLaunchKernelGGL ((adj_source<T>), dim3 ((unsigned) ((nm + 255) / 256)), dim3 (256), 0, h->stream, g.af[2],
                       (const T *) s->mon + (size_t) (t - (s->T - W)) * nm, h->n[1] * h->n[2], (int) h->n[2],
                       s->sp.monitor_z, (int) s->sp.monitor_lo[0], (int) s->sp.monitor_lo[1], s->mnx, s->mny);
  }
  if (h->mode == 1)
  {
    hipLaunchKernelGGL ((adj_h<T, 1>), grid, block, 0, h->stream, g);
    hipLaunchKernelGGL ((adj_e<T, 1>), grid, block, 0, h->stream, g);
  }
  else
  {
    hipLaunchKernelGGL ((adj_h<T, 0>), grid, block, 0, h->stream, g);
    hipLaunchKernelGGL ((adj_e<T, 0>), grid, block, 0, h->stream, g);
  }
  const int64_t nd = s->dn[0] * s->dn[1] * s->dn[2];
  const int64_t nc = s->tb.cells ();
  hipLaunchKernelGGL ((adj_grad<T>), dim3 ((unsigned) ((nd + 255) / 256)), dim3 (256), 0, h->stream, g, s->tb,
                     (const T *) s->trace + (size_t) slot_cur * 3 * nc, (const T *) s->trace + (size_t) slot_prev * 3 * nc,
                     (int) s->sp.design_lo[0], (int) s->sp.design_lo[1], (int) s->sp.design_lo[2], (int) s->dn[0],
                     (int) s->dn[1], (int) s->dn[2], s->grad);
  HIP_TRY (hipGetLastError ());
  return YEE_OK;
}

template <typename T>
int forward_impl (yee_handle *h, const double *design, double *J)
{
  AdjointState *s = h->adj;
  int rc;
  const int64_t nd = s->dn[0] * s->dn[1] * s->dn[2];
  {
    std::vector<T> v ((size_t) nd);
    for (int64_t u = 0; u < nd; ++u)
    {
      if (!(design[u] > 0) || !std::isfinite (design[u])) return fail (YEE_EINVAL, "design eps must be positive and finite");
      v[(size_t) u] = (T) design[u];
    }
    HIP_TRY (hipMemcpyAsync (s->stage, v.data (), (size_t) nd * sizeof (T), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL ((adj_set_design<T>), dim3 ((unsigned) ((nd + 255) / 256)), dim3 (256), 0, h->stream,
                       (T *) h->mat[YEE_MAT_EPS], h->n[1] * h->n[2], (int) h->n[2], (int) s->sp.design_lo[0],
                       (int) s->sp.design_lo[1], (int) s->sp.design_lo[2], (int) s->dn[0], (int) s->dn[1], (int) s->dn[2],
                       (const T *) s->stage);
    HIP_TRY (hipGetLastError ());
    HIP_TRY (hipStreamSynchronize (h->stream));
  }
  if ((rc = engine_reset (h))) return rc;
  const int K = s->sp.checkpoint_every;
  const int W = s->sp.window;
  // with checkpoints, the last segment's trace is recorded here (the backward starts with it
  // and need not re-run it)
  const int32_t t0_last = K > 0 ? ((s->T - 1 + K - 1) / K - 1) * K : 0;
  HIP_TRY (hipEventRecord (h->ev[0], h->stream));
  for (int32_t t = 0; t < s->T; ++t)
  {
    if ((rc = engine_step_once (h))) return rc;
    if (K == 0)
    {
      if ((rc = record<T> (h, t))) return rc;
    }
    else
    {
      if (t % K == 0 && (rc = snapshot (h, t / K, false))) return rc;
      if (t >= t0_last && (rc = record<T> (h, t - t0_last))) return rc;
    }
    if (t >= s->T - W && (rc = monitor<T> (h, t - (s->T - W)))) return rc;
  }
  HIP_TRY (hipEventRecord (h->ev[1], h->stream));
  HIP_TRY (hipEventSynchronize (h->ev[1]));
  float ms = 0;
  HIP_TRY (hipEventElapsedTime (&ms, h->ev[0], h->ev[1]));
  s->fwd_ms = ms;
  const size_t nm = (size_t) s->mnx * s->mny * W;
  std::vector<T> mon (nm);
  HIP_TRY (hipMemcpy (mon.data (), s->mon, nm * sizeof (T), hipMemcpyDeviceToHost));
  // compensated (Neumaier) sum: J is differenced by finite-difference checks, where the
  // rounding of a plain sum over millions of terms would dominate
  double acc = 0, comp = 0;
  for (size_t u = 0; u < nm; ++u)
  {
    const double v = (double) mon[u] * (double) mon[u];
    const double t = acc + v;
    comp += std::fabs (acc) >= std::fabs (v) ? (acc - t) + v : (v - t) + acc;
    acc = t;
  }
  acc += comp;
  s->J = acc;
  *J = acc;
  s->forward_done = true;
  s->last_traced = K > 0;
  return YEE_OK;
}

template <typename T>
int backward_impl (yee_handle *h, double dJ, double *grad)
{
  AdjointState *s = h->adj;
  int rc;
  const size_t grid_bytes = (size_t) h->ncell * h->wsize;
  for (int c = 0; c < 6; ++c)
  {
    if (s->af[c]) HIP_TRY (hipMemsetAsync (s->af[c], 0, grid_bytes, h->stream));
    if (s->ad[c]) HIP_TRY (hipMemsetAsync (s->ad[c], 0, grid_bytes, h->stream));
    if (s->w[c]) HIP_TRY (hipMemsetAsync (s->w[c], 0, grid_bytes, h->stream));
  }
  const int64_t nd = s->dn[0] * s->dn[1] * s->dn[2];
  HIP_TRY (hipMemsetAsync (s->grad, 0, (size_t) nd * sizeof (double), h->stream));
  HIP_TRY (hipEventRecord (h->ev[0], h->stream));
  const int K = s->sp.checkpoint_every;
  if (K == 0)
  {
    for (int32_t t = s->T - 1; t >= 1; --t)
      if ((rc = reverse_step<T> (h, t, t, t - 1))) return rc;
  }
  else
  {
    const int nseg = (s->T - 1 + K - 1) / K;   // segments of steps (kK, min((k+1)K, T-1)]
    for (int k = nseg - 1; k >= 0; --k)
    {
      const int32_t t0 = k * K, t1 = std::min ((k + 1) * K, s->T - 1);
      if (!(k == nseg - 1 && s->last_traced))
      {
        if ((rc = snapshot (h, k, true))) return rc;
        if ((rc = record<T> (h, 0))) return rc;   // state after step kK
        for (int32_t t = t0 + 1; t <= t1; ++t)
        {
          if ((rc = engine_step_once (h))) return rc;
          if ((rc = record<T> (h, t - t0))) return rc;
        }
      }
      for (int32_t t = t1; t >= t0 + 1; --t)
        if ((rc = reverse_step<T> (h, t, t - t0, t - t0 - 1))) return rc;
    }
    s->last_traced = false;   // the trace now holds segment 0
  }
  HIP_TRY (hipEventRecord (h->ev[1], h->stream));
  HIP_TRY (hipEventSynchronize (h->ev[1]));
  float ms = 0;
  HIP_TRY (hipEventElapsedTime (&ms, h->ev[0], h->ev[1]));
  s->bwd_ms = ms;
  HIP_TRY (hipMemcpy (grad, s->grad, (size_t) nd * sizeof (double), hipMemcpyDeviceToHost));
  for (int64_t u = 0; u < nd; ++u) grad[u] *= dJ;
  return YEE_OK;
}

template <typename T>
int fill_eps (yee_handle *h)
{
  // the design needs a per-cell eps grid: expand a uniform eps
  void *p = nullptr;
  int rc = engine_alloc_grid (h, &p);
  if (rc) return rc;
  std::vector<T> v ((size_t) h->ncell, (T) h->mat_u[YEE_MAT_EPS]);
  HIP_TRY (hipMemcpyAsync (p, v.data (), v.size () * sizeof (T), hipMemcpyHostToDevice, h->stream));
  HIP_TRY (hipStreamSynchronize (h->stream));
  h->mat[YEE_MAT_EPS] = p;
  return YEE_OK;
}

}  // namespace

void adjoint_free (yee_handle *h)
{
  if (!h || !h->adj) return;
  if (h->stream) (void) hipStreamSynchronize (h->stream);
  for (void *p : h->adj->allocs) (void) hipFree (p);
  for (void *p : h->adj->snap) (void) hipFree (p);
  delete h->adj;
  h->adj = nullptr;
}

}  // namespace yee

using namespace yee;

extern "C" {

int yee_adjoint_setup (yee_handle *h, const yee_adjoint_spec *sp)
{
  if (!h || !sp) return fail (YEE_EINVAL, "null argument");
  if (!h->prepared) return fail (YEE_ESTATE, "yee_prepare has not been called");
  if (h->cfg.scheme != YEE_SCHEME_DIM3 || h->mode == 2 || h->cfg.current_source || h->imag || h->fine () || h->fn_any ()
      || h->cfg.layout != 0)
    return fail (YEE_ENOTSUP, "the adjoint covers 3D E_CENTERED runs of real fields without metamaterials, current "
                              "sources, double material precision or test functions");
  const int T = h->cfg.time_steps;
  if (T < 2) return fail (YEE_EINVAL, "the adjoint needs --time-steps >= 2");
  if (sp->window < 1 || sp->window > T - 1) return fail (YEE_EINVAL, "window must be in [1, time_steps - 1]");
  if (sp->checkpoint_every < 0) return fail (YEE_EINVAL, "checkpoint_every must be >= 0");
  for (int a = 0; a < 3; ++a)
  {
    const int64_t P = h->use_pml () ? h->cfg.pml[a] : 0;
    // design voxels and their E cells (voxel - 1 .. voxel) must be direct-update cells
    if (sp->design_lo[a] < std::max<int64_t> (2, P + 1) || sp->design_hi[a] > h->n[a] - P - 2
        || sp->design_hi[a] <= sp->design_lo[a])
      return fail (YEE_EINVAL, "design box must lie inside the PML-free interior, one cell from its edge");
  }
  if (sp->monitor_z < 1 || sp->monitor_z >= h->n[2] - 1) return fail (YEE_EINVAL, "monitor plane outside the interior");
  for (int a = 0; a < 2; ++a)
    if (sp->monitor_lo[a] < 1 || sp->monitor_hi[a] > h->n[a] - 1 || sp->monitor_hi[a] <= sp->monitor_lo[a])
      return fail (YEE_EINVAL, "monitor rectangle outside the interior");
  if (h->cfg.point_source)
  {
    const int64_t *p = h->cfg.point_source_pos;
    for (int a = 0; a < 3; ++a)
      if (p[a] < 1 || p[a] >= h->n[a] - 1) return fail (YEE_EINVAL, "adjoint runs need interior point sources");
  }
  adjoint_free (h);
  h->adj = new AdjointState ();
  AdjointState *s = h->adj;
  s->sp = *sp;
  s->T = T;
  for (int a = 0; a < 3; ++a)
  {
    s->dn[a] = sp->design_hi[a] - sp->design_lo[a];
    s->tb.lo[a] = (int) sp->design_lo[a] - 1;
    s->tb.n[a] = (int) s->dn[a] + 1;
  }
  s->mnx = (int) (sp->monitor_hi[0] - sp->monitor_lo[0]);
  s->mny = (int) (sp->monitor_hi[1] - sp->monitor_lo[1]);
  int rc;
  if (!h->mat[YEE_MAT_EPS] && (rc = h->fp32 () ? fill_eps<float> (h) : fill_eps<double> (h))) return rc;
  const size_t grid = (size_t) h->ncell * h->wsize;
  for (int c = 0; c < 6; ++c)
  {
    if ((rc = adj_alloc (h, &s->af[c], grid))) return rc;
    if ((rc = adj_alloc (h, &s->w[c], grid))) return rc;
    if (h->use_pml () && (rc = adj_alloc (h, &s->ad[c], grid))) return rc;
  }
  const int K = sp->checkpoint_every;
  s->trace_slots = K == 0 ? T : K + 1;
  if ((rc = adj_alloc (h, &s->trace, (size_t) s->trace_slots * 3 * s->tb.cells () * h->wsize))) return rc;
  if ((rc = adj_alloc (h, &s->mon, (size_t) s->mnx * s->mny * sp->window * h->wsize))) return rc;
  void *g = nullptr;
  if ((rc = adj_alloc (h, &g, (size_t) (s->dn[0] * s->dn[1] * s->dn[2]) * sizeof (double)))) return rc;
  s->grad = (double *) g;
  if ((rc = adj_alloc (h, &s->stage, (size_t) (s->dn[0] * s->dn[1] * s->dn[2]) * h->wsize))) return rc;
  if (K > 0)
  {
    const int nsnap = (T - 1 + K - 1) / K + 1;
    size_t bytes = 0;
    for (int c = 0; c < 6; ++c)
    {
      if (h->f[0][c]) bytes += grid;
      if (h->d[0][c]) bytes += grid;
    }
    bytes += 3 * (size_t) h->inc_len * h->wsize;
    s->snap_bytes = bytes;
    s->snap.assign (nsnap, nullptr);
    s->snap_t.assign (nsnap, 0);
    s->snap_cur.assign (nsnap, 0);
    for (int k = 0; k < nsnap; ++k)
      if (hipMalloc (&s->snap[k], bytes) != hipSuccess) return fail (YEE_ENOMEM, "checkpoint snapshot allocation failed");
  }
  HIP_TRY (hipStreamSynchronize (h->stream));
  return YEE_OK;
}

int yee_forward (yee_handle *h, const double *design_eps, double *J)
{
  if (!h || !design_eps || !J) return fail (YEE_EINVAL, "null argument");
  if (!h->adj) return fail (YEE_ESTATE, "yee_adjoint_setup has not been called");
  return h->fp32 () ? forward_impl<float> (h, design_eps, J) : forward_impl<double> (h, design_eps, J);
}

int yee_backward (yee_handle *h, double dJ, double *grad_eps)
{
  if (!h || !grad_eps) return fail (YEE_EINVAL, "null argument");
  if (!h->adj || !h->adj->forward_done) return fail (YEE_ESTATE, "yee_forward has not been called");
  return h->fp32 () ? backward_impl<float> (h, dJ, grad_eps) : backward_impl<double> (h, dJ, grad_eps);
}

int yee_adjoint_timing (yee_handle *h, double out[2])
{
  if (!h || !out) return fail (YEE_EINVAL, "null argument");
  if (!h->adj) return fail (YEE_ESTATE, "yee_adjoint_setup has not been called");
  out[0] = h->adj->fwd_ms;
  out[1] = h->adj->bwd_ms;
  return YEE_OK;
}

}  // extern "C"
