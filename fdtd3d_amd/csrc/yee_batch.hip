// Batched TMz sweep: many independent TMz simulations (different grid sizes, materials and
// sources) advanced together, two launches per time step for the whole batch.
//
// The reference runs one simulation per process (Source/main.cpp); this has no counterpart
// there.  Every simulation is exactly what a single yee_handle with the same configuration
// computes (the same cell code, tmz_e_cell / tmz_h_cell), so a batch member is bit-identical
// to the run it replaces.
//
// The simulations are split into groups whose state fits a cache-sized budget (the 256 MB
// Infinity Cache of the MI355X sits in front of HBM): a call of yee_batch_step advances one
// group by all its steps before the next group starts, so after the first step a group's
// fields are read from the cache rather than from HBM.  Per step of a group: batch_tmz_step
// (E and H fused, double-buffered state) over the group's tile table, then batch_advance of
// its simulations' step counters.  The counters live in device memory (one per simulation),
// so a graph captured over several steps of a group is replayed as often as needed.  The
// grouping budget and the tile length (x rows per tile) are picked by a short timing of the
// first steps (every candidate computes the same values, so the tuning steps are ordinary
// steps).
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

template <typename T>
__global__ void batch_tmz_step (const StepArgs<T> *, const int4 *, int, const int *, const T *, int);
template <typename T>
__global__ void batch_tmz_step2 (const StepArgs<T> *, const int4 *, int, const int *, const T *, int);
template <typename T>
__global__ void batch_tmz_sources (const StepArgs<T> *, int, const int *, const T *, int);
__global__ void batch_advance (int *, const int *, int, int);

int batch_fail (int code, const std::string &msg);

}  // namespace yee

using namespace yee;

// Simulations advanced together, step after step, with their own tile tables and graph.
struct BatchGroup
{
  std::vector<int> sims;
  int *list = nullptr;                  // the same indices in device memory (batch_advance)
  int4 *tiles[2][4] = {};               // [two-step][length] tile tables (kLx[0..3] x rows)
  int ntiles[2][4] = {};
  hipGraphExec_t graph = nullptr;
};

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
  std::vector<void *> f[2], d[2];       // [parity][sim][3] Ez, Hx, Hy and Dz, Bx, By
  std::vector<void *> pml;              // [sim] PML tables
  std::vector<int> tab_len;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = { nullptr, nullptr };
  void *sims = nullptr;                 // StepArgs<T>[nsim][2] (by step parity)
  std::vector<std::vector<BatchGroup>> groupings;   // [grouping][group], one per cache budget
  int *state = nullptr;                 // device: [sim] steps done, buffer of the newest fields
  bool two_ok = true;                   // two-step launches allowed (no mu grid in the batch)
  int *all = nullptr;                   // 0 .. nsim-1 in device memory
  void *src_tab = nullptr;              // [sim][src_len][2] source values
  int src_len = 0;
  int32_t t = 0;
  bool prepared = false;
  int cand = -1;                        // chosen tuner candidate
  int ncu = 256;
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

constexpr int kLx[4] = { 13, 24, 40, 72 };   // x rows per tile

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

// Per-simulation StepArgs: the same set-up make_args performs for a single TMz handle, with
// the old state in buffer `parity` and the new one in the other.
template <typename T>
StepArgs<T> sim_args (const yee_batch *b, int s, int parity)
{
  const yee_config &c = b->cfg[s];
  const bool pml = (c.flags & YEE_USE_PML) != 0;
  StepArgs<T> a {};
  a.sz = 1;
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
    a.f[comp] = (T *) b->f[parity][3 * s + k];
    a.f_n[comp] = (T *) b->f[1 - parity][3 * s + k];
    a.d[comp] = (T *) b->d[parity][3 * s + k];
    a.d_n[comp] = (T *) b->d[1 - parity][3 * s + k];
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
  a.src_pos[0] = (int) p[0];
  a.src_pos[1] = (int) p[1];
  return a;
}

// Destroys every captured graph (they hold the source table's address).
int drop_graphs (yee_batch *b)
{
  for (auto &gs : b->groupings)
    for (BatchGroup &g : gs)
      if (g.graph)
      {
        BHIP (hipGraphExecDestroy (g.graph));
        g.graph = nullptr;
      }
  return YEE_OK;
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
  int rc = drop_graphs (b);                  // the captured steps hold the old table's address
  if (rc) return rc;
  if (b->src_tab) BHIP (hipFree (b->src_tab));
  b->src_tab = nullptr;
  BHIP (hipMalloc (&b->src_tab, v.size () * sizeof (T)));
  BHIP (hipMemcpy (b->src_tab, v.data (), v.size () * sizeof (T), hipMemcpyHostToDevice));
  b->src_len = len;
  return YEE_OK;
}

// One launch of group g: one step (batch_tmz_step) or two (batch_tmz_step2), tiles of
// length kLx[sel].
template <typename T>
void launch_step (yee_batch *b, const BatchGroup &g, int two, int sel)
{
  const int nt = g.ntiles[two][sel];
  const dim3 grid ((unsigned) std::max (1, (nt + 3) / 4)), block (256);
  if (two)
    hipLaunchKernelGGL ((batch_tmz_step2<T>), grid, block, 0, b->stream, (const StepArgs<T> *) b->sims, g.tiles[1][sel], nt,
                       b->state, (const T *) b->src_tab, b->src_len);
  else
    hipLaunchKernelGGL ((batch_tmz_step<T>), grid, block, 0, b->stream, (const StepArgs<T> *) b->sims, g.tiles[0][sel], nt,
                       b->state, (const T *) b->src_tab, b->src_len);
}

void launch_advance (yee_batch *b, const int *list, int n, int k)
{
  hipLaunchKernelGGL (batch_advance, dim3 ((unsigned) ((n + 63) / 64)), dim3 (64), 0, b->stream, b->state, list, n, k);
}

// Candidates of the tuner: (grouping, steps per launch, tile length index).
struct Cand
{
  int grouping, two, sel;
};
// Cache budgets of the groupings (MiB of state per group; 0: the whole batch in one group).
constexpr double kBudgetMiB[] = { 0, 160, 64 };
constexpr int kNumGroupings = (int) (sizeof kBudgetMiB / sizeof kBudgetMiB[0]);
const Cand kCands[] = { { 0, 0, 0 }, { 0, 0, 1 }, { 0, 0, 2 }, { 0, 1, 0 }, { 0, 1, 1 }, { 0, 1, 2 },
                        { 0, 1, 3 }, { 1, 1, 1 }, { 1, 1, 2 }, { 2, 1, 0 }, { 2, 1, 1 }, { 2, 1, 2 } };
constexpr int kNumCands = (int) (sizeof kCands / sizeof kCands[0]);
constexpr int kTuneSteps = 4;
constexpr int kGraphSteps = 16;

// `k` steps (1, or 2 with a two-step candidate) of group g with candidate c's tiles.
template <typename T>
int enqueue_steps (yee_batch *b, const BatchGroup &g, const Cand &c, int k)
{
  const int two = (c.two && k == 2) ? 1 : 0;
  launch_step<T> (b, g, two, c.sel);
  launch_advance (b, g.list, (int) g.sims.size (), two ? 2 : 1);
  BHIP (hipGetLastError ());
  return YEE_OK;
}

// Steps per launch of a candidate.
int cand_stride (const Cand &c)
{
  return c.two ? 2 : 1;
}

template <typename T>
int build_graph (yee_batch *b, BatchGroup &grp, const Cand &c)
{
  if (grp.graph)
  {
    BHIP (hipGraphExecDestroy (grp.graph));
    grp.graph = nullptr;
  }
  hipGraph_t g = nullptr;
  BHIP (hipStreamBeginCapture (b->stream, hipStreamCaptureModeThreadLocal));
  int rc = YEE_OK;
  for (int s = 0; s < kGraphSteps && rc == YEE_OK; s += cand_stride (c)) rc = enqueue_steps<T> (b, grp, c, cand_stride (c));
  hipError_t e = hipStreamEndCapture (b->stream, &g);
  if (rc) return rc;
  if (e != hipSuccess) return batch_fail (YEE_EDEVICE, std::string ("hipStreamEndCapture: ") + hipGetErrorString (e));
  e = hipGraphInstantiate (&grp.graph, g, nullptr, nullptr, 0);
  (void) hipGraphDestroy (g);
  if (e != hipSuccess) return batch_fail (YEE_EDEVICE, std::string ("hipGraphInstantiate: ") + hipGetErrorString (e));
  b->graph_steps = kGraphSteps;
  return YEE_OK;
}


template <typename T>
int step_impl (yee_batch *b, int32_t n)
{
  int rc = YEE_OK;
  if (b->t + n > b->src_len && (rc = upload_sources<T> (b, std::max (b->t + n, 2 * b->src_len)))) return rc;
  // device pointers inside StepArgs do not change, but the source table may have moved
  BHIP (hipEventRecord (b->ev[0], b->stream));
  int32_t done = 0;
  if (b->t == 0 && n > 0)
  {
    hipLaunchKernelGGL ((batch_tmz_sources<T>), dim3 ((unsigned) ((b->nsim + 63) / 64)), dim3 (64), 0, b->stream,
                       (const StepArgs<T> *) b->sims, b->nsim, b->state, (const T *) b->src_tab, b->src_len);
    launch_advance (b, b->all, b->nsim, 1);
    BHIP (hipGetLastError ());
    ++done;
    ++b->t;
  }
  // YEE_BATCH_CAND=<index> fixes the candidate (tests of every launch shape)
  if (const char *fc = std::getenv ("YEE_BATCH_CAND"))
  {
    const int ci = std::atoi (fc);
    if (ci < 0 || ci >= kNumCands || (kCands[ci].two && !b->two_ok)) return batch_fail (YEE_EINVAL, "YEE_BATCH_CAND out of range");
    if (b->cand != ci) rc = drop_graphs (b);
    if (rc) return rc;
    b->cand = ci;
  }
  // tuner: kTuneSteps ordinary steps of every simulation per candidate, timed, on the first
  // call that has room
  if (b->cand < 0 && n - done >= kNumCands * kTuneSteps)
  {
    b->tune_ms.assign (kNumCands, 0.0);
    hipEvent_t e0, e1;
    BHIP (hipEventCreate (&e0));
    BHIP (hipEventCreate (&e1));
    double best = 1e30;
    for (int c = 0; c < kNumCands; ++c)
    {
      if (kCands[c].two && !b->two_ok)
      {
        b->tune_ms[c] = 0;
        continue;
      }
      BHIP (hipEventRecord (e0, b->stream));
      const int stride = cand_stride (kCands[c]);
      for (const BatchGroup &g : b->groupings[kCands[c].grouping])
        for (int s = 0; s < kTuneSteps; s += stride)
          if ((rc = enqueue_steps<T> (b, g, kCands[c], stride))) return rc;
      BHIP (hipEventRecord (e1, b->stream));
      BHIP (hipEventSynchronize (e1));
      float ms = 0;
      BHIP (hipEventElapsedTime (&ms, e0, e1));
      b->tune_ms[c] = ms / kTuneSteps;
      if (ms < best)
      {
        best = ms;
        b->cand = c;
      }
      done += kTuneSteps;
      b->t += kTuneSteps;
    }
    (void) hipEventDestroy (e0);
    (void) hipEventDestroy (e1);
  }
  if (b->cand < 0) b->cand = 2;   // too few steps to tune: one group, 64-row tiles
  const Cand &c = kCands[b->cand];
  const char *use_graph = std::getenv ("YEE_BATCH_GRAPH");
  const bool graphs = !(use_graph && std::strcmp (use_graph, "0") == 0);
  // each group runs all the remaining steps before the next one starts
  const int32_t left = n - done;
  for (BatchGroup &g : b->groupings[c.grouping])
  {
    if (graphs && !g.graph && left >= kGraphSteps && (rc = build_graph<T> (b, g, c))) return rc;
    int32_t k = 0;
    while (k < left)
    {
      if (graphs && g.graph && left - k >= b->graph_steps)
      {
        BHIP (hipGraphLaunch (g.graph, b->stream));
        k += b->graph_steps;
      }
      else
      {
        const int kk = std::min (cand_stride (c), left - k);
        if ((rc = enqueue_steps<T> (b, g, c, kk))) return rc;
        k += kk;
      }
    }
  }
  b->t += left;
  BHIP (hipEventRecord (b->ev[1], b->stream));
  BHIP (hipEventSynchronize (b->ev[1]));
  float ms = 0;
  BHIP (hipEventElapsedTime (&ms, b->ev[0], b->ev[1]));
  b->last_ms = n > 0 ? ms / n : 0;
  return YEE_OK;
}

// Device bytes a simulation's steps touch: the field grids, eps/mu grids, and the D/B grids
// inside its PML frame.
double sim_bytes (const yee_batch *b, int s)
{
  const yee_config &c = b->cfg[s];
  const double cells = (double) c.size[0] * (double) c.size[1];
  double arrays = 3 + (b->mat[2 * s] ? 1 : 0) + (b->mat[2 * s + 1] ? 1 : 0);
  double bytes = cells * arrays * (double) b->wsize;
  if (c.flags & YEE_USE_PML)
  {
    const double inner = (double) std::max<int64_t> (0, c.size[0] - 2 * c.pml[0]) * (double) std::max<int64_t> (0, c.size[1] - 2 * c.pml[1]);
    bytes += (cells - inner) * 3 * (double) b->wsize;
  }
  return bytes;
}

template <typename T>
int prepare_impl (yee_batch *b)
{
  int rc;
  std::vector<StepArgs<T>> args (2 * (size_t) b->nsim);
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
    args[2 * s] = sim_args<T> (b, s, 0);
    args[2 * s + 1] = sim_args<T> (b, s, 1);
  }
  if ((rc = alloc_zero (b, &b->sims, args.size () * sizeof (StepArgs<T>)))) return rc;
  BHIP (hipMemcpyAsync (b->sims, args.data (), args.size () * sizeof (StepArgs<T>), hipMemcpyHostToDevice, b->stream));
  // groupings: simulations biggest first, packed greedily into groups of at most the budget
  // (a simulation above it gets a group of its own); tile tables per group, so the tail of
  // a sweep is made of small tiles
  std::vector<int> order ((size_t) b->nsim);
  for (int s = 0; s < b->nsim; ++s) order[s] = s;
  std::stable_sort (order.begin (), order.end (), [&] (int x, int y) {
    return b->cfg[x].size[0] * b->cfg[x].size[1] > b->cfg[y].size[0] * b->cfg[y].size[1];
  });
  b->groupings.assign (kNumGroupings, {});
  for (int gi = 0; gi < kNumGroupings; ++gi)
  {
    std::vector<BatchGroup> &gs = b->groupings[gi];
    const double budget = kBudgetMiB[gi] * 1048576.0;
    double used = 0;
    for (int s : order)
    {
      const double by = sim_bytes (b, s);
      if (gs.empty () || (budget > 0 && used + by > budget))
      {
        gs.emplace_back ();
        used = 0;
      }
      gs.back ().sims.push_back (s);
      used += by;
    }
    for (BatchGroup &g : gs)
    {
      void *p = nullptr;
      if ((rc = alloc_zero (b, &p, g.sims.size () * sizeof (int)))) return rc;
      BHIP (hipMemcpyAsync (p, g.sims.data (), g.sims.size () * sizeof (int), hipMemcpyHostToDevice, b->stream));
      g.list = (int *) p;
      for (int two = 0; two < 2; ++two)
        for (int sel = 0; sel < 4; ++sel)
        {
          // tiles (sim, xa, y0, xb): rows [xa, xb) of columns [y0, y0 + 63) (one step per
          // launch) or [y0, y0 + 61) (two steps)
          const int w = two ? 61 : 63;
          std::vector<int4> tv;
          for (int s : g.sims)
          {
            const yee_config &c = b->cfg[s];
            for (int64_t x0 = 1; x0 < c.size[0] - 1; x0 += kLx[sel])
              for (int64_t y0 = 1; y0 < c.size[1] - 1; y0 += w)
                tv.push_back (make_int4 (s, (int) x0, (int) y0, (int) std::min<int64_t> (x0 + kLx[sel], c.size[0] - 1)));
          }
          g.ntiles[two][sel] = (int) tv.size ();
          if ((rc = alloc_zero (b, &p, std::max<size_t> (1, tv.size ()) * sizeof (int4)))) return rc;
          BHIP (hipMemcpyAsync (p, tv.data (), tv.size () * sizeof (int4), hipMemcpyHostToDevice, b->stream));
          g.tiles[two][sel] = (int4 *) p;
        }
    }
  }
  {
    void *st = nullptr;
    if ((rc = alloc_zero (b, &st, std::max<size_t> (256, 2 * (size_t) b->nsim * sizeof (int))))) return rc;
    b->state = (int *) st;
    for (int s = 0; s < b->nsim; ++s) b->two_ok = b->two_ok && b->mat[2 * s + 1] == nullptr;
    std::vector<int> all ((size_t) b->nsim);
    for (int s = 0; s < b->nsim; ++s) all[s] = s;
    if ((rc = alloc_zero (b, &st, all.size () * sizeof (int)))) return rc;
    BHIP (hipMemcpyAsync (st, all.data (), all.size () * sizeof (int), hipMemcpyHostToDevice, b->stream));
    b->all = (int *) st;
  }
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
    if (c.flags & (YEE_USE_TFSF | YEE_USE_METAMATERIALS | YEE_COMPLEX_FIELDS | YEE_DOUBLE_MATERIAL_PRECISION))
      return batch_fail (YEE_ENOTSUP, "batched runs support PML and point sources (no TF/SF, metamaterials, complex "
                                      "fields or double material precision)");
    if (c.current_source)
      return batch_fail (YEE_ENOTSUP, "batched runs take point sources, not current sources");
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
  for (int par = 0; par < 2; ++par)
  {
    b->f[par].assign (3 * (size_t) nsim, nullptr);
    b->d[par].assign (3 * (size_t) nsim, nullptr);
  }
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
    for (int par = 0; par < 2; ++par)
      for (int k = 0; k < 3; ++k)
      {
        int rc = alloc_zero (b, &b->f[par][3 * s + k], bytes);
        if (rc) return bail (rc);
        if ((cfgs[s].flags & YEE_USE_PML) && (rc = alloc_zero (b, &b->d[par][3 * s + k], bytes))) return bail (rc);
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
  (void) drop_graphs (b);
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

int yee_batch_tuning (yee_batch *b, int32_t out[3], double *cand_ms, int32_t ncand)
{
  if (!b || !out) return batch_fail (YEE_EINVAL, "null argument");
  if (b->cand >= 0)
  {
    const Cand &c = kCands[b->cand];
    const std::vector<BatchGroup> &gs = b->groupings[c.grouping];
    out[0] = kLx[c.sel];
    out[1] = gs.empty () ? 0 : (gs[0].ntiles[c.two][c.sel] + 3) / 4;
    out[2] = (int32_t) gs.size ();
  }
  else
  {
    out[0] = out[1] = out[2] = 0;
  }
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
    case YEE_EZ: k = 0; break;
    case YEE_HX: k = 1; break;
    case YEE_HY: k = 2; break;
    case YEE_DZ: k = 3; break;
    case YEE_BX: k = 4; break;
    case YEE_BY: k = 5; break;
    default: return batch_fail (YEE_EINVAL, "TMz components are Ez, Hx, Hy (and Dz, Bx, By with PML)");
  }
  const size_t n = (size_t) (b->cfg[sim].size[0] * b->cfg[sim].size[1]);
  BHIP (hipStreamSynchronize (b->stream));
  // the buffer holding the newest fields of this simulation (flipped by every launch)
  int buf = 0;
  if (b->state) BHIP (hipMemcpy (&buf, b->state + 2 * sim + 1, sizeof (int), hipMemcpyDeviceToHost));
  src = k < 3 ? b->f[buf][3 * sim + k] : b->d[buf][3 * sim + k - 3];
  if (!src)
  {
    std::fill (out, out + n, 0.0);
    return YEE_OK;
  }
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
