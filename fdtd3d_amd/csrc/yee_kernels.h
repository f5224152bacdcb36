// Device-side argument blocks of the Yee kernels (shared by the kernels and the host driver).
#pragma once

#include <cstdint>

namespace yee {

constexpr int kMaxDim = 3;

// Geometry and tables common to every half-step kernel.
template <typename T>
struct StepArgs
{
  // grid: [x][y][z], z fastest; TMz uses nz == 1
  int nx, ny, nz;
  int64_t sx;   // ny * nz
  int sy;       // nz

  // cells that run the reference's two-stage D/B update: the PML slabs
  // (p < slab_lo[a] or p >= slab_hi[a]) and, with metamaterials, the Drude box
  int slab_lo[3], slab_hi[3];
  int box_lo[3], box_hi[3];

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

  // scalars in the field precision
  T dt, dx, eps0, mu0;
  // per-component 1D PML coefficient tables, [comp][role][tab_len]
  const T *pml;
  int tab_len;

  // TF/SF (geometry kept in FP64; corrections touch O(N^2) cells only)
  int tfsf;
  int tl[3], tr[3];            // total-field box [tl, tr] in integer coordinates
  double zinc[3];
  double S1, C1, S2, C2;
  double proj[6];
  const T *einc;               // EInc after this step's E update (read by H)
  const T *hinc;               // HInc after the previous step (read by E)
  int64_t inc_len;
};

}  // namespace yee
