// Batched TMz sweep: many independent TMz simulations (different grid sizes, materials and
// sources) advanced together, two launches per time step for the whole batch.
//
// The reference runs one simulation per process (Source/main.cpp); this has no counterpart
// there.  Every simulation is exactly what a single yee_handle with the same configuration
// computes (the same cell code, tmz_e_cell / tmz_h_cell), so a batch member is bit-identical
// to the run it replaces.
//
// Per step: batch_tmz_half (E) then batch_tmz_half (H) over a tile table of every
// simulation's interior, then batch_advance.  The step counter lives in device memory, so a
// graph captured over several steps is replayed as often as needed.  The tile height TX and
// the number of persistent workgroups are picked by a short timing of the first steps
// (every candidate computes the same values, so the tuning steps are ordinary steps).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/yee.h"
#include "yee_kernels.h"
#include "yee_setup.h"

namespace yee {

template <typename T, int TX>
__global__ void batch_tmz_half (const StepArgs<T> *, const int4 *, int, int, const int *, const T *, int);
template <typename T>
__global__ void batch_tmz_sources (const StepArgs<T> *, int, const int *, const T *, int);
__global__ void batch_advance (int *);

int batch_fail (int code, const std::string &msg);

}  // namespace yee

using namespace yee;

struct yee_batch
{
  int nsim = 0;
  size_t wsize = 4;
  std::vector<yee_config> cfg;
  std::vector<Scalars> sc;
  std::vector<double> t0, sigma;        // Gaussian envelopes of the point sources
  std::vector<double> mat_u;            // [sim][2] uniform eps, mu
  std::vector<void *> mat;              // [sim][2] eps, mu grids (nullptr: uniform)
  std::vector<void *> grids;            // every device allocation of the batch
  std::vector<void *> f, d;             // [sim][3] Ez, Hx, Hy and Dz, Bx, By
  std::vector<void *> pml;              // [sim] PML tables
  std::vector<int> tab_len;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = { nullptr, nullptr };
  void *sims = nullptr;                 // StepArgs<T>[nsim]
  int4 *tiles[4] = {};                  // tile tables for TX = 1, 2, 4, 8
  int ntiles[4] = {};
  int *step = nullptr;                  // device step counter
  void *src_tab = nullptr;              // [sim][src_len][2] source values
  int src_len = 0;
  int32_t t = 0;
  bool prepared = false;
  int tx_sel = -1, grid = 0;            // chosen tile height index and workgroup count
  int ncu = 256;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  double last_ms = 0;
  std::vector<double> tune_ms;          // [candidate] ms per step measured by the tuner
};

namespace {

static thread_local std::string g_berr;

#define BHIP(expr)                                                                             \
  do                                                                                           \
  {                                                                                            \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return yee::batch_fail (YEE_EDEVICE, std::string (#expr) + ": " + hipGetErrorString (e_)); \
  } while (0)

constexpr int kTx[4] = { 1, 2, 4, 8 };

int alloc_zero (yee_batch *b, void **p, size_t bytes)
{
  BHIP (hipMalloc (p, bytes));
  b->grids.push_back (*p);
  BHIP (hipMemsetAsync (*p, 0, bytes, b->stream));
  return YEE_OK;
}

template <typename S>
int upload_grid (yee_batch *b, int sim, int which, const S *src)
{
  const yee_config &c = b->cfg[sim];
  const size_t n = (size_t) (c.size[0] * c.size[1]);
  bool uniform = true;
  for (size_t q = 0; q < n; ++q)
  {
    if (!std::isfinite ((double) src[q]) || src[q] == 0) return batch_fail (YEE_EINVAL, "eps/mu must be finite and nonzero");
    if (src[q] != src[0]) uniform = false;
  }
  if (uniform)
  {
    b->mat_u[2 * sim + which] = (double) src[0];
    b->mat[2 * sim + which] = nullptr;
    return YEE_OK;
  }
  void *&dst = b->mat[2 * sim + which];
  if (!dst)
  {
    int rc = alloc_zero (b, &dst, n * b->wsize);
    if (rc) return rc;
  }
  if (b->wsize == 4)
  {
    std::vector<float> v (n);
    for (size_t q = 0; q < n; ++q) v[q] = (float) src[q];
    BHIP (hipMemcpyAsync (dst, v.data (), n * 4, hipMemcpyHostToDevice, b->stream));
    BHIP (hipStreamSynchronize (b->stream));
  }
  else
  {
    std::vector<double> v (n);
    for (size_t q = 0; q < n; ++q) v[q] = (double) src[q];
    BHIP (hipMemcpyAsync (dst, v.data (), n * 8, hipMemcpyHostToDevice, b->stream));
    BHIP (hipStreamSynchronize (b->stream));
  }
  return YEE_OK;
}

// Per-simulation StepArgs: the same set-up make_args performs for a single TMz handle.
template <typename T>
StepArgs<T> sim_args (const yee_batch *b, int s)
{
  const yee_config &c = b->cfg[s];
  const bool pml = (c.flags & YEE_USE_PML) != 0;
  StepArgs<T> a {};
  a.nx = (int) c.size[0];
  a.ny = (int) c.size[1];
  a.nz = 1;
  a.sx = c.size[1];
  a.sy = 1;
  for (int ax = 0; ax < 3; ++ax)
  {
    a.slab_lo[ax] = (ax < 2 && pml) ? (int) c.pml[ax] : -1;
    a.slab_hi[ax] = (ax < 2 && pml) ? (int) (c.size[ax] - c.pml[ax] - 1) : 1 << 30;
    a.run_lo[ax] = 1;
    a.run_hi[ax] = ax < 2 ? (int) c.size[ax] - 1 : 1;
  }
  for (int k = 0; k < 3; ++k)
  {
    const int comp = k == 0 ? YEE_EZ : (k == 1 ? YEE_HX : YEE_HY);
    a.f[comp] = a.f_n[comp] = (T *) b->f[3 * s + k];
    a.d[comp] = a.d_n[comp] = (T *) b->d[3 * s + k];
  }
  a.eps = (const T *) b->mat[2 * s];
  a.mu = (const T *) b->mat[2 * s + 1];
  a.eps_u = (T) b->mat_u[2 * s];
  a.mu_u = (T) b->mat_u[2 * s + 1];
  a.dt = (T) b->sc[s].dt;
  a.dx = (T) c.dx;
  a.eps0 = (T) eps0 ();
  a.mu0 = (T) mu0 ();
  a.pml = (const T *) b->pml[s];
  a.tab_len = b->tab_len[s];
  a.tfsf = 0;
  const int64_t *p = c.point_source_pos;
  a.src_mask = c.point_source & ((1u << YEE_EZ) | (1u << YEE_HX) | (1u << YEE_HY));
  a.src_q = p[0] * c.size[1] + p[1];
  return a;
}

// Source values of steps [0, len) for every simulation (host FP64 C library, as the single
// engine computes them), rounded to the field precision.
template <typename T>
int upload_sources (yee_batch *b, int len)
{
  std::vector<T> v ((size_t) b->nsim * len * 2);
  for (int s = 0; s < b->nsim; ++s)
    for (int t = 0; t < len; ++t)
      for (int h = 0; h < 2; ++h)
      {
        const double ts = h ? t + 1.0 : t + 0.5;
        double val = std::sin (b->sc[s].dt * ts * 2 * kPi * b->sc[s].freq);
        if (b->sigma[s] > 0)
        {
          const double u = (ts - b->t0[s]) / b->sigma[s];
          val *= std::exp (-u * u);
        }
        v[((size_t) s * len + t) * 2 + h] = (T) val;
      }
  BHIP (hipStreamSynchronize (b->stream));   // queued steps may still read the old table
  if (b->graph)
  {
    // the captured steps hold the old table's address
    BHIP (hipGraphExecDestroy (b->graph));
    b->graph = nullptr;
  }
  if (b->src_tab) BHIP (hipFree (b->src_tab));
  b->src_tab = nullptr;
  BHIP (hipMalloc (&b->src_tab, v.size () * sizeof (T)));
  BHIP (hipMemcpy (b->src_tab, v.data (), v.size () * sizeof (T), hipMemcpyHostToDevice));
  b->src_len = len;
  return YEE_OK;
}

template <typename T>
void launch_half (yee_batch *b, int sel, int grid, bool e_half)
{
  const int tx = kTx[sel];
  const dim3 block (64, (unsigned) tx, 1);
  const int nt = b->ntiles[sel];
  const StepArgs<T> *sims = (const StepArgs<T> *) b->sims;
  const T *tab = (const T *) b->src_tab;
  const int eh = e_half ? 1 : 0;
  switch (tx)
  {
    case 1: hipLaunchKernelGGL ((batch_tmz_half<T, 1>), dim3 (grid), block, 0, b->stream, sims, b->tiles[sel], nt, eh, b->step, tab, b->src_len); break;
    case 2: hipLaunchKernelGGL ((batch_tmz_half<T, 2>), dim3 (grid), block, 0, b->stream, sims, b->tiles[sel], nt, eh, b->step, tab, b->src_len); break;
    case 4: hipLaunchKernelGGL ((batch_tmz_half<T, 4>), dim3 (grid), block, 0, b->stream, sims, b->tiles[sel], nt, eh, b->step, tab, b->src_len); break;
    default: hipLaunchKernelGGL ((batch_tmz_half<T, 8>), dim3 (grid), block, 0, b->stream, sims, b->tiles[sel], nt, eh, b->step, tab, b->src_len); break;
  }
}

// One step t >= 1 (E, H, counter) with tile height kTx[sel] and `grid` workgroups.
template <typename T>
int enqueue_step (yee_batch *b, int sel, int grid)
{
  launch_half<T> (b, sel, grid, true);
  launch_half<T> (b, sel, grid, false);
  hipLaunchKernelGGL (batch_advance, dim3 (1), dim3 (64), 0, b->stream, b->step);
  BHIP (hipGetLastError ());
  return YEE_OK;
}

// Candidates of the tuner: (tile height index, workgroups per CU).
struct Cand
{
  int sel, per_cu;
};
// per_cu = 0: one tile per workgroup (grid = number of tiles, the hardware dispatcher is the
// queue); otherwise that many persistent workgroups per CU walk the tile table.
const Cand kCands[] = { { 2, 0 }, { 3, 0 }, { 1, 0 }, { 2, 8 }, { 3, 4 }, { 1, 16 }, { 2, 16 }, { 0, 0 } };
constexpr int kNumCands = (int) (sizeof kCands / sizeof kCands[0]);
constexpr int kTuneSteps = 3;
constexpr int kGraphSteps = 16;

int cand_grid (const yee_batch *b, const Cand &c)
{
  if (c.per_cu == 0) return std::max (1, b->ntiles[c.sel]);
  return std::max (1, std::min (b->ntiles[c.sel], c.per_cu * b->ncu));
}

template <typename T>
int build_graph (yee_batch *b)
{
  if (b->graph)
  {
    BHIP (hipGraphExecDestroy (b->graph));
    b->graph = nullptr;
  }
  hipGraph_t g = nullptr;
  BHIP (hipStreamBeginCapture (b->stream, hipStreamCaptureModeThreadLocal));
  int rc = YEE_OK;
  for (int s = 0; s < kGraphSteps && rc == YEE_OK; ++s) rc = enqueue_step<T> (b, b->tx_sel, b->grid);
  hipError_t e = hipStreamEndCapture (b->stream, &g);
  if (rc) return rc;
  if (e != hipSuccess) return batch_fail (YEE_EDEVICE, std::string ("hipStreamEndCapture: ") + hipGetErrorString (e));
  e = hipGraphInstantiate (&b->graph, g, nullptr, nullptr, 0);
  (void) hipGraphDestroy (g);
  if (e != hipSuccess) return batch_fail (YEE_EDEVICE, std::string ("hipGraphInstantiate: ") + hipGetErrorString (e));
  b->graph_steps = kGraphSteps;
  return YEE_OK;
}

template <typename T>
int step_impl (yee_batch *b, int32_t n)
{
  int rc;
  if (b->t + n > b->src_len && (rc = upload_sources<T> (b, std::max (b->t + n, 2 * b->src_len)))) return rc;
  // device pointers inside StepArgs do not change, but the source table may have moved
  BHIP (hipEventRecord (b->ev[0], b->stream));
  int32_t done = 0;
  if (b->t == 0 && n > 0)
  {
    hipLaunchKernelGGL ((batch_tmz_sources<T>), dim3 ((unsigned) ((b->nsim + 63) / 64)), dim3 (64), 0, b->stream,
                       (const StepArgs<T> *) b->sims, b->nsim, b->step, (const T *) b->src_tab, b->src_len);
    hipLaunchKernelGGL (batch_advance, dim3 (1), dim3 (64), 0, b->stream, b->step);
    BHIP (hipGetLastError ());
    ++done;
    ++b->t;
  }
  // tuner: kTuneSteps ordinary steps per candidate, timed, on the first call that has room
  if (b->tx_sel < 0 && n - done >= kNumCands * kTuneSteps)
  {
    b->tune_ms.assign (kNumCands, 0.0);
    hipEvent_t e0, e1;
    BHIP (hipEventCreate (&e0));
    BHIP (hipEventCreate (&e1));
    double best = 1e30;
    for (int c = 0; c < kNumCands; ++c)
    {
      const int grid = cand_grid (b, kCands[c]);
      BHIP (hipEventRecord (e0, b->stream));
      for (int s = 0; s < kTuneSteps; ++s)
        if ((rc = enqueue_step<T> (b, kCands[c].sel, grid))) return rc;
      BHIP (hipEventRecord (e1, b->stream));
      BHIP (hipEventSynchronize (e1));
      float ms = 0;
      BHIP (hipEventElapsedTime (&ms, e0, e1));
      b->tune_ms[c] = ms / kTuneSteps;
      if (ms < best)
      {
        best = ms;
        b->tx_sel = kCands[c].sel;
        b->grid = grid;
      }
      done += kTuneSteps;
      b->t += kTuneSteps;
    }
    (void) hipEventDestroy (e0);
    (void) hipEventDestroy (e1);
  }
  if (b->tx_sel < 0)
  {
    // too few steps to tune: a middle candidate
    b->tx_sel = kCands[1].sel;
    b->grid = cand_grid (b, kCands[1]);
  }
  const char *use_graph = std::getenv ("YEE_BATCH_GRAPH");
  const bool graphs = !(use_graph && std::strcmp (use_graph, "0") == 0);
  if (graphs && !b->graph && n - done >= kGraphSteps && (rc = build_graph<T> (b))) return rc;
  while (n - done > 0)
  {
    if (graphs && b->graph && n - done >= b->graph_steps)
    {
      BHIP (hipGraphLaunch (b->graph, b->stream));
      done += b->graph_steps;
      b->t += b->graph_steps;
    }
    else
    {
      if ((rc = enqueue_step<T> (b, b->tx_sel, b->grid))) return rc;
      ++done;
      ++b->t;
    }
  }
  BHIP (hipEventRecord (b->ev[1], b->stream));
  BHIP (hipEventSynchronize (b->ev[1]));
  float ms = 0;
  BHIP (hipEventElapsedTime (&ms, b->ev[0], b->ev[1]));
  b->last_ms = n > 0 ? ms / n : 0;
  return YEE_OK;
}

template <typename T>
int prepare_impl (yee_batch *b)
{
  int rc;
  std::vector<StepArgs<T>> args ((size_t) b->nsim);
  for (int s = 0; s < b->nsim; ++s)
  {
    const yee_config &c = b->cfg[s];
    if (c.flags & YEE_USE_PML)
    {
      b->tab_len[s] = (int) std::max (c.size[0], c.size[1]);
      std::vector<double> tab = pml_tables (c, b->tab_len[s]);
      std::vector<T> tv (tab.begin (), tab.end ());
      void *p = nullptr;
      if ((rc = alloc_zero (b, &p, tv.size () * sizeof (T)))) return rc;
      BHIP (hipMemcpyAsync (p, tv.data (), tv.size () * sizeof (T), hipMemcpyHostToDevice, b->stream));
      BHIP (hipStreamSynchronize (b->stream));
      b->pml[s] = p;
    }
    args[s] = sim_args<T> (b, s);
  }
  if ((rc = alloc_zero (b, &b->sims, args.size () * sizeof (StepArgs<T>)))) return rc;
  BHIP (hipMemcpyAsync (b->sims, args.data (), args.size () * sizeof (StepArgs<T>), hipMemcpyHostToDevice, b->stream));
  // tile tables, biggest simulations first so the tail of a sweep is made of small tiles
  std::vector<int> order ((size_t) b->nsim);
  for (int s = 0; s < b->nsim; ++s) order[s] = s;
  std::stable_sort (order.begin (), order.end (), [&] (int x, int y) {
    return b->cfg[x].size[0] * b->cfg[x].size[1] > b->cfg[y].size[0] * b->cfg[y].size[1];
  });
  for (int sel = 0; sel < 4; ++sel)
  {
    std::vector<int4> tv;
    for (int s : order)
    {
      const yee_config &c = b->cfg[s];
      const int mode = (c.flags & YEE_USE_PML) ? 1 : 0;
      for (int64_t x0 = 1; x0 < c.size[0] - 1; x0 += kTx[sel])
        for (int64_t y0 = 1; y0 < c.size[1] - 1; y0 += 64)
          tv.push_back (make_int4 (s, (int) x0, (int) y0, mode));
    }
    b->ntiles[sel] = (int) tv.size ();
    void *p = nullptr;
    if ((rc = alloc_zero (b, &p, std::max<size_t> (1, tv.size ()) * sizeof (int4)))) return rc;
    BHIP (hipMemcpyAsync (p, tv.data (), tv.size () * sizeof (int4), hipMemcpyHostToDevice, b->stream));
    b->tiles[sel] = (int4 *) p;
  }
  void *st = nullptr;
  if ((rc = alloc_zero (b, &st, 256))) return rc;
  b->step = (int *) st;
  BHIP (hipStreamSynchronize (b->stream));
  int maxT = 1;
  for (const yee_config &c : b->cfg) maxT = std::max (maxT, (int) c.time_steps);
  return upload_sources<T> (b, maxT + 1);
}

}  // namespace

namespace yee {
int batch_fail (int code, const std::string &msg)
{
  g_berr = msg;
  return code;
}
}  // namespace yee

extern "C" {

const char *yee_batch_last_error (void)
{
  return g_berr.c_str ();
}

int yee_batch_create (const yee_config *cfgs, int32_t nsim, yee_batch **out)
{
  if (!cfgs || !out || nsim <= 0) return batch_fail (YEE_EINVAL, "null argument or empty batch");
  *out = nullptr;
  const int32_t prec = cfgs[0].precision;
  for (int32_t s = 0; s < nsim; ++s)
  {
    const yee_config &c = cfgs[s];
    if (c.scheme != YEE_SCHEME_TMZ) return batch_fail (YEE_EINVAL, "batched runs are TMz (YEE_SCHEME_TMZ)");
    if (c.layout != 0) return batch_fail (YEE_EINVAL, "only the E_CENTERED layout is supported");
    if (c.precision != prec || (prec != 4 && prec != 8)) return batch_fail (YEE_EINVAL, "one precision (4 or 8) per batch");
    if (c.flags & (YEE_USE_TFSF | YEE_USE_METAMATERIALS))
      return batch_fail (YEE_ENOTSUP, "batched runs support PML and point sources (no TF/SF, no metamaterials)");
    for (int a = 0; a < 2; ++a)
    {
      if (c.size[a] < 3 || c.size[a] > (1 << 16)) return batch_fail (YEE_EINVAL, "size out of range");
      if (c.pml[a] < 0 || 2 * c.pml[a] > c.size[a]) return batch_fail (YEE_EINVAL, "pml size out of range");
    }
    if (!(c.dx > 0) || !(c.wavelength > 0) || !(c.courant > 0)) return batch_fail (YEE_EINVAL, "dx, wavelength, courant must be positive");
    if (c.point_source & ~((1u << YEE_EZ) | (1u << YEE_HX) | (1u << YEE_HY)))
      return batch_fail (YEE_EINVAL, "TMz point sources are Ez, Hx, Hy");
    const int64_t *p = c.point_source_pos;
    if (c.point_source && (p[0] < 1 || p[0] >= c.size[0] - 1 || p[1] < 1 || p[1] >= c.size[1] - 1))
      return batch_fail (YEE_EINVAL, "batched point sources must lie in the interior");
  }
  yee_batch *b = new yee_batch ();
  b->nsim = nsim;
  b->wsize = prec == 4 ? 4 : 8;
  b->cfg.assign (cfgs, cfgs + nsim);
  b->t0.assign (nsim, 0.0);
  b->sigma.assign (nsim, 0.0);
  b->mat_u.assign (2 * (size_t) nsim, 1.0);
  b->mat.assign (2 * (size_t) nsim, nullptr);
  b->f.assign (3 * (size_t) nsim, nullptr);
  b->d.assign (3 * (size_t) nsim, nullptr);
  b->pml.assign (nsim, nullptr);
  b->tab_len.assign (nsim, 0);
  auto bail = [&] (int code) {
    yee_batch_destroy (b);
    return code;
  };
  if (hipStreamCreateWithFlags (&b->stream, hipStreamNonBlocking) != hipSuccess) return bail (batch_fail (YEE_EDEVICE, "hipStreamCreate failed"));
  if (hipEventCreate (&b->ev[0]) != hipSuccess || hipEventCreate (&b->ev[1]) != hipSuccess)
    return bail (batch_fail (YEE_EDEVICE, "hipEventCreate failed"));
  int dev = 0;
  if (hipGetDevice (&dev) == hipSuccess) (void) hipDeviceGetAttribute (&b->ncu, hipDeviceAttributeMultiprocessorCount, dev);
  for (int32_t s = 0; s < nsim; ++s)
  {
    const size_t bytes = (size_t) (cfgs[s].size[0] * cfgs[s].size[1]) * b->wsize;
    for (int k = 0; k < 3; ++k)
    {
      int rc = alloc_zero (b, &b->f[3 * s + k], bytes);
      if (rc) return bail (rc);
      if ((cfgs[s].flags & YEE_USE_PML) && (rc = alloc_zero (b, &b->d[3 * s + k], bytes))) return bail (rc);
    }
    b->sc.push_back (compute_scalars (cfgs[s]));
  }
  if (hipStreamSynchronize (b->stream) != hipSuccess) return bail (batch_fail (YEE_EDEVICE, "stream sync failed"));
  *out = b;
  return YEE_OK;
}

void yee_batch_destroy (yee_batch *b)
{
  if (!b) return;
  if (b->stream) (void) hipStreamSynchronize (b->stream);
  if (b->graph) (void) hipGraphExecDestroy (b->graph);
  for (void *p : b->grids) (void) hipFree (p);
  if (b->src_tab) (void) hipFree (b->src_tab);
  for (hipEvent_t e : b->ev)
    if (e) (void) hipEventDestroy (e);
  if (b->stream) (void) hipStreamDestroy (b->stream);
  delete b;
}

int yee_batch_set_material (yee_batch *b, int32_t sim, int32_t which, const double *values)
{
  if (!b || !values) return batch_fail (YEE_EINVAL, "null argument");
  if (sim < 0 || sim >= b->nsim) return batch_fail (YEE_EINVAL, "simulation index out of range");
  if (which != YEE_MAT_EPS && which != YEE_MAT_MU) return batch_fail (YEE_EINVAL, "batched runs take eps and mu");
  if (b->prepared) return batch_fail (YEE_ESTATE, "materials must be set before yee_batch_prepare");
  return upload_grid (b, sim, which, values);
}

int yee_batch_set_material_f32 (yee_batch *b, int32_t sim, int32_t which, const float *values)
{
  if (!b || !values) return batch_fail (YEE_EINVAL, "null argument");
  if (sim < 0 || sim >= b->nsim) return batch_fail (YEE_EINVAL, "simulation index out of range");
  if (which != YEE_MAT_EPS && which != YEE_MAT_MU) return batch_fail (YEE_EINVAL, "batched runs take eps and mu");
  if (b->prepared) return batch_fail (YEE_ESTATE, "materials must be set before yee_batch_prepare");
  return upload_grid (b, sim, which, values);
}

int yee_batch_set_point_source_gaussian (yee_batch *b, int32_t sim, double t0, double sigma)
{
  if (!b) return batch_fail (YEE_EINVAL, "null handle");
  if (sim < 0 || sim >= b->nsim) return batch_fail (YEE_EINVAL, "simulation index out of range");
  if (b->prepared) return batch_fail (YEE_ESTATE, "sources must be set before yee_batch_prepare");
  if (!std::isfinite (t0) || !std::isfinite (sigma)) return batch_fail (YEE_EINVAL, "t0 and sigma must be finite");
  b->t0[sim] = t0;
  b->sigma[sim] = sigma;
  return YEE_OK;
}

int yee_batch_prepare (yee_batch *b)
{
  if (!b) return batch_fail (YEE_EINVAL, "null handle");
  if (b->prepared) return YEE_OK;
  int rc = b->wsize == 4 ? prepare_impl<float> (b) : prepare_impl<double> (b);
  if (rc) return rc;
  b->prepared = true;
  return YEE_OK;
}

int yee_batch_step (yee_batch *b, int32_t n)
{
  if (!b) return batch_fail (YEE_EINVAL, "null handle");
  if (!b->prepared) return batch_fail (YEE_ESTATE, "yee_batch_prepare has not been called");
  if (n < 0) return batch_fail (YEE_EINVAL, "negative step count");
  if (n == 0) return YEE_OK;
  return b->wsize == 4 ? step_impl<float> (b, n) : step_impl<double> (b, n);
}

int yee_batch_sync (yee_batch *b)
{
  if (!b) return batch_fail (YEE_EINVAL, "null handle");
  BHIP (hipStreamSynchronize (b->stream));
  return YEE_OK;
}

int yee_batch_time (yee_batch *b, int32_t *t)
{
  if (!b || !t) return batch_fail (YEE_EINVAL, "null argument");
  *t = b->t;
  return YEE_OK;
}

int yee_batch_last_step_ms (yee_batch *b, double *ms)
{
  if (!b || !ms) return batch_fail (YEE_EINVAL, "null argument");
  *ms = b->last_ms;
  return YEE_OK;
}

int yee_batch_tuning (yee_batch *b, int32_t out[2], double *cand_ms, int32_t ncand)
{
  if (!b || !out) return batch_fail (YEE_EINVAL, "null argument");
  out[0] = b->tx_sel >= 0 ? kTx[b->tx_sel] : 0;
  out[1] = b->grid;
  for (int32_t c = 0; cand_ms && c < ncand && c < (int32_t) b->tune_ms.size (); ++c) cand_ms[c] = b->tune_ms[c];
  return YEE_OK;
}

int yee_batch_get_field (yee_batch *b, int32_t sim, int32_t comp, double *out)
{
  if (!b || !out) return batch_fail (YEE_EINVAL, "null argument");
  if (sim < 0 || sim >= b->nsim) return batch_fail (YEE_EINVAL, "simulation index out of range");
  int k;
  void *src;
  switch (comp)
  {
    case YEE_EZ: k = 0; src = b->f[3 * sim]; break;
    case YEE_HX: k = 1; src = b->f[3 * sim + 1]; break;
    case YEE_HY: k = 2; src = b->f[3 * sim + 2]; break;
    case YEE_DZ: k = 0; src = b->d[3 * sim]; break;
    case YEE_BX: k = 1; src = b->d[3 * sim + 1]; break;
    case YEE_BY: k = 2; src = b->d[3 * sim + 2]; break;
    default: return batch_fail (YEE_EINVAL, "TMz components are Ez, Hx, Hy (and Dz, Bx, By with PML)");
  }
  (void) k;
  const size_t n = (size_t) (b->cfg[sim].size[0] * b->cfg[sim].size[1]);
  if (!src)
  {
    std::fill (out, out + n, 0.0);
    return YEE_OK;
  }
  BHIP (hipStreamSynchronize (b->stream));
  if (b->wsize == 4)
  {
    std::vector<float> v (n);
    BHIP (hipMemcpy (v.data (), src, n * 4, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) out[q] = v[q];
  }
  else
  {
    BHIP (hipMemcpy (out, src, n * 8, hipMemcpyDeviceToHost));
  }
  return YEE_OK;
}

}  // extern "C"
