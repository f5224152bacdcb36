// Internal declarations shared by the host files of libyee.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/yee.h"
#include "yee_kernels.h"
#include "yee_setup.h"

namespace yee {

// Records `msg` as the calling thread's last error (yee_last_error) and returns `code`.
int fail (int code, const std::string &msg);

#define HIP_TRY(expr)                                                                          \
  do                                                                                           \
  {                                                                                            \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return ::yee::fail (YEE_EDEVICE, std::string (#expr) + ": " + hipGetErrorString (e_));   \
  } while (0)

struct AdjointState;

}  // namespace yee

struct yee_handle
{
  yee_config cfg {};
  int dim = 3;
  int64_t n[3] = { 1, 1, 1 };
  int64_t ncell = 0;
  size_t wsize = 8;
  int mode = 0;   // 0 plain, 1 PML, 2 PML + Drude
  bool present[6] = {};   // components held and updated (all six on the 3D kernels)
  bool exposed[6] = {};   // components of the scheme (what the API reads and sources drive)
  bool lack[3] = { false, false, false };   // axes a 1D / 2D scheme run on the 3D kernels lacks (extent 1)
  int miss = -1;          // the single lacked axis of such a 2D scheme (TEx/TMx 0, TEy/TMy 1), else -1
  int device = 0;                // the device current at yee_create (its stream and memory live there)
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = { nullptr, nullptr };

  // state: two buffers per array when the fused 3D step is used (old -> new), else one
  void *f[2][6] = {}, *d[2][6] = {}, *dp[2][6] = {}, *d1[2][6] = {}, *d1p[2][6] = {};
  int cur = 0;          // buffer holding the newest state
  bool fused = false;   // 3D fused step kernel on the PML-free box + slab half-steps
  int64_t fbox_lo[3] = { 1, 1, 1 }, fbox_hi[3] = { 1, 1, 1 };   // PML-free box of the fused kernel
  int fused_lx = 0, fused_nstrips = 0, fused_nchunks = 0, fused_threads = 0;
  // padded grid allocations (every grid-sized array; the views above point inside them)
  std::vector<void *> grids;
  void *tg = nullptr;   // TfsfGeom<T> in device memory
  void *zcorr = nullptr;              // z-face TF/SF corrections of the fused kernel (FP64)
  int64_t zc_n[2] = { 0, 0 };

  void *mat[6] = {};
  double mat_u[6] = { 1, 1, 0, 0, 0, 0 };
  // material grid extent (twice the field grid along the scheme's axes with double material
  // precision) and, in that mode, the 8-cell averages of eps (E components) and mu (H
  // components) per field cell, computed at yee_prepare (null for a uniform material)
  int64_t mn[3] = { 1, 1, 1 };
  int64_t mcell = 0;
  void *matc[6] = {};
  bool mat_nonzero[6] = {};
  int64_t box_lo[3] = { 0, 0, 0 }, box_hi[3] = { 0, 0, 0 };   // Drude-material bounding box (eps cells)
  bool box_any = false;

  void *pml = nullptr;
  int tab_len = 0;
  void *einc = nullptr, *hinc[2] = { nullptr, nullptr };
  int64_t inc_len = 0;
  yee::Scalars sc {};
  double src_t0 = 0, src_sigma = 0;   // Gaussian envelope of the point source (sigma <= 0: none)
  int32_t t = 0;
  bool prepared = false;
  double last_ms = 0;
  // kernel timing (yee_profile_steps): events recorded around each launch when set
  std::vector<hipEvent_t> *prof_events = nullptr;
  std::vector<int> *prof_kinds = nullptr;

  bool fp32 () const { return wsize == 4; }
  bool use_pml () const { return (cfg.flags & YEE_USE_PML) != 0; }
  bool use_tfsf () const { return (cfg.flags & YEE_USE_TFSF) != 0; }
  bool use_meta () const { return (cfg.flags & YEE_USE_METAMATERIALS) != 0; }
  bool fine () const { return (cfg.flags & YEE_DOUBLE_MATERIAL_PRECISION) != 0; }

  yee::AdjointState *adj = nullptr;   // forward/backward pair state (yee_adjoint_setup)

  // complex field values: the part this handle's sources drive (0 the real build, 1 the real
  // part, 2 the imaginary part of the complex build), and the imaginary part's handle, which a
  // complex handle (the real part) owns
  int src_kind = 0;
  yee_handle *imag = nullptr;

  // analytic test problem (YEE_FN_* kinds; yee_set_test_functions): start values, border values,
  // right side, exact solution of the diff norm, its box, and per component the workgroup
  // partial sums of the last step's norm (test_norm)
  int fn_start = 0, fn_border = 0, fn_rhs = 0, fn_exact = 0;
  int64_t norm_lo[3] = { 0, 0, 0 }, norm_hi[3] = { 0, 0, 0 };
  double *norm_part = nullptr;
  bool norm_valid[6] = {};
  bool fn_any () const { return fn_start || fn_border || fn_rhs || fn_exact; }
  // exact solution exp1 .. exp3 of a 1D pair (yee_set_exact_exp): pair 0 (ExHy) .. 5 (EzHy) and
  // the host-computed constants of its phase and H factor (StepArgs exp_*)
  int exp_pair = 0;
  double exp_k = 0, exp_d = 0, exp_w = 0, exp_t0 = 0, exp_m = 0;

  // near-to-far-field transform: H of the step before the newest (yee_ntff_snapshot; each part's
  // handle holds its own), the step it was taken at, and the device buffer of the angles and
  // the per-face sums of yee_ntff
  void *ntff_h[3] = {};
  int32_t ntff_t = -1;
  double *ntff_buf = nullptr;
  int64_t ntff_cap = 0;
  bool complex () const { return (cfg.flags & YEE_COMPLEX_FIELDS) != 0; }
};

namespace yee {

// One time step of the handle (the body of yee_step), enqueued on its stream.
int engine_step_once (yee_handle *h);
// Padded, zero-filled grid allocation owned by the handle.
int engine_alloc_grid (yee_handle *h, void **view);
// Zeroes every field, auxiliary and incident array and resets the step counter.
int engine_reset (yee_handle *h);
// Kernel argument block of the handle's current state (in-place half-step view).
template <typename T>
StepArgs<T> engine_args (const yee_handle *h, int parity);
// Frees the adjoint state (called by yee_destroy).
void adjoint_free (yee_handle *h);
// Tile grid of a half-step style launch over a.run_lo .. a.run_hi (fills the tile fields of a;
// false when the box is empty or too large for one launch).
template <typename T>
bool half_tiles (StepArgs<T> &a, unsigned bx, unsigned by, dim3 &grid);

}  // namespace yee
