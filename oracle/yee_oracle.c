/*
 * yee_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker; never shipped or measured).
 *
 * FP64 CPU restatement of the reference's Yee update path (E_CENTERED layout,
 * --3d and --2d-tmz schemes, real fields).  Each function cites the reference code
 * it restates; expressions keep the reference's operand order so that, compiled
 * without FP contraction, the results are bit-identical to the reference's own
 * arithmetic (checked against oracle/_ref in tests/test_oracle_ref.py).
 *
 * Storage mirrors the reference: every state grid has `layers` time layers
 * (2, or 3 with metamaterials + PML: Source/Scheme/InternalScheme.specific.h:83-88);
 * layer 0 is written by the current update, layer 1 is the newest completed one;
 * shiftInTime rotates the layers (Source/Grid/Grid.h:1304-1319).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "yee_oracle.h"

/* Source/Physics/PhysicsConst.h:27-33 */
#define O_PI 3.14159265358979323846   /* M_PI of glibc, as PhysicsConst::Pi = M_PI */
#define O_C 299792458.0

static double o_mu0 (void) { return 4 * O_PI * 0.0000001; }
static double o_eps0 (void) { return 1 / (o_mu0 () * (O_C * O_C)); }

/*
 * Analytic test problems, the real build's callbacks (Source/Scheme/CallBack.cpp:28-143; SQR (a)
 * is a * a, PhysicsConst.h:30-32): field values at (x, y, z, t) and the right-side currents J
 * (E components) / M (H components).  Which components each kind defines follows
 * Scheme::initCallBacks (Scheme.inc.h:2150-2430): polinom1, polinom3, sin1 Ez and Hy,
 * polinom2 all six; right sides polinom1 / polinom3 Jz and My, polinom2 all six, sin1 none.
 */
static int o_fn_has (int kind, int comp)
{
  if (kind == YEE_FN_POLINOM2) return 1;
  if (kind == YEE_FN_POLINOM1 || kind == YEE_FN_POLINOM3 || kind == YEE_FN_SIN1) return comp == YEE_EZ || comp == YEE_HY;
  return 0;
}

static int o_rhs_has (int kind, int comp)
{
  if (kind == YEE_FN_POLINOM2) return 1;
  if (kind == YEE_FN_POLINOM1 || kind == YEE_FN_POLINOM3) return comp == YEE_EZ || comp == YEE_HY;
  return 0;
}

static double o_fn_value (int kind, int comp, double x, double y, double z, double t)
{
  const double c = O_C, e0 = o_eps0 ();
  switch (kind)
  {
    case YEE_FN_POLINOM1:
    {
      const double ez = (c * c) * (x * x) * t;
      return comp == YEE_EZ ? ez : e0 * ez;
    }
    case YEE_FN_POLINOM2:
    {
      const int a = comp % 3;
      const double u = a == 0 ? y : x, v = a == 2 ? y : z;   /* Ex: y, z; Ey: x, z; Ez: x, y */
      const double e = (c * c) * t * ((u * u) + (v * v));
      return comp < 3 ? e : e0 * e;
    }
    case YEE_FN_POLINOM3:
    {
      const double ez = (c * c) * ((x * t) * (x * t));
      return comp == YEE_EZ ? ez : e0 * ez;
    }
    case YEE_FN_SIN1:
      return comp == YEE_EZ ? c * sin (t - x) : -c * sin (t - x);
    default:
      return 0;
  }
}

static double o_rhs_value (int kind, int comp, double x, double y, double z, double t)
{
  const double c = O_C, e0 = o_eps0 (), m0 = o_mu0 ();
  switch (kind)
  {
    case YEE_FN_POLINOM1:
      if (comp == YEE_EZ) return -(c * c) * e0 * (-2 * x * t + (x * x));
      return -(c * c) * (-2 * x * t + (x * x) * e0 * m0);
    case YEE_FN_POLINOM2:
    {
      /* Jx: 2 t (z - y) + y^2 + z^2; Jy: 2 t (x - z) + x^2 + z^2; Jz: 2 t (y - x) + x^2 + y^2 */
      const double p[3] = { x, y, z };
      const int a = comp % 3;
      const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
      if (comp < 3)
      {
        const double d = a == 0 ? (p[2] - p[1]) : (a == 1 ? (p[0] - p[2]) : (p[1] - p[0]));
        return -(c * c) * e0 * (2 * t * d + (p[u] * p[u]) + (p[v] * p[v]));
      }
      /* Mx: 2 t (y - z); My: 2 t (z - x); Mz: 2 t (x - y) */
      const double d = a == 0 ? (p[1] - p[2]) : (a == 1 ? (p[2] - p[0]) : (p[0] - p[1]));
      return -(c * c) * (2 * t * d + e0 * m0 * ((p[u] * p[u]) + (p[v] * p[v])));
    }
    case YEE_FN_POLINOM3:
      if (comp == YEE_EZ) return -(c * c) * 2 * e0 * x * t * (x - t);
      return -(c * c) * 2 * x * t * (e0 * m0 * x - t);
    default:
      return 0;
  }
}

enum { LEFT = 0, RIGHT = 1, DOWN = 2, UP = 3, BACK = 4, FRONT = 5 };

/*
 * Per-component bindings of InternalScheme::calculateFieldStepInit
 * (Source/Scheme/InternalScheme.inc.h:1902-2295) and the E_CENTERED circuit offsets
 * (Source/Layout/YeeGridLayout.cpp:1937-1961) used by calculateFieldStepInitDiff (:1837-1900).
 * sig_curl/sig4/sig5: sigma axis of materialGrid / materialGrid4 / materialGrid5.
 */
typedef struct
{
  int axis, is_h, opp1, opp2;
  int d11[3], d12[3], d21[3], d22[3];
  int dir11, dir12, dir21, dir22;
  int sig_curl, sig4, sig5;
} comp_desc;

static const comp_desc CD[6] = {
  /* Ex: Hz (DOWN/UP), Hy (BACK/FRONT); SigmaY, SigmaX, SigmaZ */
  { 0, 0, YEE_HZ, YEE_HY, {0, -1, 0}, {0, 0, 0}, {0, 0, -1}, {0, 0, 0}, DOWN, UP, BACK, FRONT, 1, 0, 2 },
  /* Ey: Hx (BACK/FRONT), Hz (LEFT/RIGHT); SigmaZ, SigmaY, SigmaX */
  { 1, 0, YEE_HX, YEE_HZ, {0, 0, -1}, {0, 0, 0}, {-1, 0, 0}, {0, 0, 0}, BACK, FRONT, LEFT, RIGHT, 2, 1, 0 },
  /* Ez: Hy (LEFT/RIGHT), Hx (DOWN/UP); SigmaX, SigmaZ, SigmaY */
  { 2, 0, YEE_HY, YEE_HX, {-1, 0, 0}, {0, 0, 0}, {0, -1, 0}, {0, 0, 0}, LEFT, RIGHT, DOWN, UP, 0, 2, 1 },
  /* Hx: Ey (BACK/FRONT), Ez (DOWN/UP); SigmaY, SigmaX, SigmaZ */
  { 0, 1, YEE_EY, YEE_EZ, {0, 0, 0}, {0, 0, 1}, {0, 0, 0}, {0, 1, 0}, BACK, FRONT, DOWN, UP, 1, 0, 2 },
  /* Hy: Ez (LEFT/RIGHT), Ex (BACK/FRONT); SigmaZ, SigmaY, SigmaX */
  { 1, 1, YEE_EZ, YEE_EX, {0, 0, 0}, {1, 0, 0}, {0, 0, 0}, {0, 0, 1}, LEFT, RIGHT, BACK, FRONT, 2, 1, 0 },
  /* Hz: Ex (DOWN/UP), Ey (LEFT/RIGHT); SigmaX, SigmaZ, SigmaY */
  { 2, 1, YEE_EX, YEE_EY, {0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {1, 0, 0}, DOWN, UP, LEFT, RIGHT, 0, 2, 1 },
};

struct oracle
{
  yee_config cfg;
  int dim;                 /* 3 or 2 */
  int act[3];              /* axes of the scheme: a 2D scheme has one axis of extent 1, which it lacks */
  int64_t lo[3], hi[3];    /* updated range per axis: [1, n-1), or [0, 1) along the missing axis */
  int64_t n[3];            /* n[a] == 1 along the missing axis of a 2D scheme */
  int64_t ncell;
  int present[6];
  int use_pml, use_tfsf, use_meta;
  int layers;
  double dx, dt, freq, rel, courant, lambda;
  double a1, a2, a3;
  double zero_inc[3];
  double bl[3], br[3];     /* TF/SF borders as FP coordinates */
  double *mat[6];          /* eps, mu, omegaPE, omegaPM, gammaE, gammaM */
  double *sigma[3];
  double *f[6][3];         /* E/H layers */
  double *d[6][3];         /* D/B layers (PML) */
  double *d1[6][3];        /* D1/B1 layers (metamaterials) */
  double *einc[2], *hinc[2];
  int64_t inc_len;
  int32_t t;
  int prepared;
  double src_phase;        /* test hook: added to every source argument (0 = reference) */
  int src_kind;            /* 0 the real build; 1 / 2 the real / imaginary part of the complex build's sources */
  double src_t0, src_sigma; /* Gaussian envelope of the point source in time steps (sigma <= 0: none) */
  int fine;                /* --use-double-material-precision: material and sigma grids of extent mn */
  int64_t mn[3], mcell;
  /* analytic test problems (YEE_FN_*): start values, border values, right side, exact solution */
  int fn_start, fn_border, fn_rhs, fn_exact;
  /* exact solutions exp1 / exp2 / exp3 of a 1D scheme's pair (YEE_FN_EXP*): the pair, the
   * material border coordinate (eps sphere centre - radius along the line) */
  int exp_pair;
  double exp_border;
  int64_t norm_lo[3], norm_hi[3];
  double norm[6][4];       /* last diff norm per component: norm, percent, sum of squares, max |exact| */
  int norm_valid[6];
};

static int64_t o_idx (const oracle *o, int64_t i, int64_t j, int64_t k)
{
  return (i * o->n[1] + j) * o->n[2] + k;
}

/* Minimum real coordinate of a component, E_CENTERED (Source/Layout/YeeGridLayout.cpp:1962-1969). */
static double o_min_coord (int comp, int axis)
{
  const comp_desc *c = &CD[comp];
  if (!c->is_h)
  {
    return axis == c->axis ? 1.0 : 0.5;
  }
  return axis == c->axis ? 0.5 : 1.0;
}

static void o_real_coord (const oracle *o, int comp, int64_t i, int64_t j, int64_t k, double out[3])
{
  int64_t p[3] = { i, j, k };
  for (int a = 0; a < 3; ++a)
  {
    out[a] = o->act[a] ? (double) p[a] + o_min_coord (comp, a) : 0.0;
  }
}

/* getEpsCoord: integer coordinate of a real coordinate on the eps grid (minEpsCoordFP = 0.5). */
static void o_eps_index (const oracle *o, const double real[3], int64_t out[3])
{
  for (int a = 0; a < 3; ++a)
  {
    out[a] = o->act[a] ? (int64_t) (real[a] - 0.5) : 0;
  }
}

static double o_mval (const oracle *o, const double *grid, const int64_t p[3])
{
  return grid[(p[0] * o->mn[1] + p[1]) * o->mn[2] + p[2]];
}

/*
 * Neighbour eps-grid cells that InternalScheme::getMaterial averages for a field point
 * (YeeGridLayout::initMaterialCoordinates, Source/Layout/YeeGridLayout.h:806-1041,
 * non-double-precision branches).  E components: 2 cells along their own axis; H
 * components: 4 cells in the transverse plane.  Offsets along an axis the scheme does
 * not have are dropped, as TC::initAxesCoordinate does for 2D coordinates.
 */
static int o_material_cells (const oracle *o, int comp, int64_t i, int64_t j, int64_t k, int64_t cells[4][3])
{
  const comp_desc *c = &CD[comp];
  double real[3];
  o_real_coord (o, comp, i, j, k, real);
  if (!c->is_h)
  {
    double lo[3] = { real[0], real[1], real[2] };
    double hi[3] = { real[0], real[1], real[2] };
    if (o->act[c->axis])
    {
      lo[c->axis] -= 0.5;
      hi[c->axis] += 0.5;
    }
    o_eps_index (o, lo, cells[0]);
    o_eps_index (o, hi, cells[1]);
    return 2;
  }
  int b = c->axis == 0 ? 1 : 0;
  int cc = c->axis == 2 ? 1 : 2;
  double sb[4] = { -0.5, -0.5, 0.5, 0.5 };
  double sc[4] = { -0.5, 0.5, -0.5, 0.5 };
  for (int q = 0; q < 4; ++q)
  {
    double r[3] = { real[0], real[1], real[2] };
    if (o->act[b]) r[b] += sb[q];
    if (o->act[cc]) r[cc] += sc[q];
    o_eps_index (o, r, cells[q]);
  }
  return 4;
}

/*
 * The 8 fine material cells of a field point with double material precision
 * (YeeGridLayout::initMaterialCoordinates, Source/Layout/YeeGridLayout.h:866-1037, E_CENTERED,
 * isDoubleMaterialPrecision branches), in the reference's order 11, 12, 21, 22, 31, 32, 41, 42.
 * Each cell is getEpsCoord (realCoord + shift) * 2 + offset.  E components: the shift is -0.5
 * (lo) or +0.5 (hi) along their own axis; H components: one of four transverse shifts c1..c4
 * = (-,-), (-,+), (+,-), (+,+) on their two other axes in (x, y, z) order.  Shifts and
 * offsets along an axis the scheme lacks are dropped (TC::initAxesCoordinate), so the cell
 * index there is 0.
 */
static const int kFineBase[6][8] = {
  { 0, 0, 1, 1, 0, 0, 1, 1 },   /* Ex: lo, lo, hi, hi, lo, lo, hi, hi */
  { 0, 0, 1, 1, 0, 0, 1, 1 },   /* Ey */
  { 0, 0, 0, 0, 1, 1, 1, 1 },   /* Ez */
  { 0, 0, 1, 1, 2, 2, 3, 3 },   /* Hx: c1, c1, c2, c2, c3, c3, c4, c4 */
  { 0, 0, 1, 1, 2, 2, 3, 3 },   /* Hy */
  { 0, 0, 1, 1, 2, 2, 3, 3 },   /* Hz */
};
static const int kFineOff[6][8][3] = {
  { { 1, 0, 0 }, { 1, 1, 0 }, { 0, 0, 0 }, { 0, 1, 0 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 0, 1 }, { 0, 1, 1 } },
  { { 0, 1, 0 }, { 1, 1, 0 }, { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 1 }, { 1, 1, 1 }, { 0, 0, 1 }, { 1, 0, 1 } },
  { { 0, 0, 1 }, { 0, 1, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 0, 0 }, { 0, 1, 0 }, { 1, 0, 0 }, { 1, 1, 0 } },
  { { 0, 1, 1 }, { 1, 1, 1 }, { 0, 1, 0 }, { 1, 1, 0 }, { 0, 0, 1 }, { 1, 0, 1 }, { 0, 0, 0 }, { 1, 0, 0 } },
  { { 1, 0, 1 }, { 1, 1, 1 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 0, 1 }, { 0, 1, 1 }, { 0, 0, 0 }, { 0, 1, 0 } },
  { { 1, 1, 0 }, { 1, 1, 1 }, { 1, 0, 0 }, { 1, 0, 1 }, { 0, 1, 0 }, { 0, 1, 1 }, { 0, 0, 0 }, { 0, 0, 1 } },
};

static void o_material_cells_fine (const oracle *o, int comp, int64_t i, int64_t j, int64_t k, int64_t cells[8][3])
{
  const comp_desc *c = &CD[comp];
  double real[3];
  o_real_coord (o, comp, i, j, k, real);
  int64_t base[4][3];
  if (!c->is_h)
  {
    for (int s = 0; s < 2; ++s)
    {
      double r[3] = { real[0], real[1], real[2] };
      if (o->act[c->axis]) r[c->axis] += s ? 0.5 : -0.5;
      o_eps_index (o, r, base[s]);
    }
  }
  else
  {
    const int b = c->axis == 0 ? 1 : 0;
    const int cc = c->axis == 2 ? 1 : 2;
    for (int s = 0; s < 4; ++s)
    {
      double r[3] = { real[0], real[1], real[2] };
      if (o->act[b]) r[b] += (s & 2) ? 0.5 : -0.5;
      if (o->act[cc]) r[cc] += (s & 1) ? 0.5 : -0.5;
      o_eps_index (o, r, base[s]);
    }
  }
  for (int q = 0; q < 8; ++q)
    for (int a = 0; a < 3; ++a)
      cells[q][a] = o->act[a] ? base[kFineBase[comp][q]][a] * 2 + kFineOff[comp][q][a] : 0;
}

/* InternalScheme::getMaterial + Approximation::approximateMaterial (Approximation.cpp:36-54). */
static double o_get_material (const oracle *o, int comp, int64_t i, int64_t j, int64_t k, const double *grid)
{
  if (!grid)
  {
    return 0;
  }
  if (o->fine)
  {
    int64_t f[8][3];
    o_material_cells_fine (o, comp, i, j, k, f);
    double v[8];
    for (int q = 0; q < 8; ++q) v[q] = o_mval (o, grid, f[q]);
    double a = ((v[0] + v[1]) / 2.0 + (v[2] + v[3]) / 2.0) / 2.0;
    double b = ((v[4] + v[5]) / 2.0 + (v[6] + v[7]) / 2.0) / 2.0;
    return (a + b) / 2.0;
  }
  int64_t cells[4][3];
  int nc = o_material_cells (o, comp, i, j, k, cells);
  if (nc == 2)
  {
    return (o_mval (o, grid, cells[0]) + o_mval (o, grid, cells[1])) / 2.0;
  }
  double a = (o_mval (o, grid, cells[0]) + o_mval (o, grid, cells[1])) / 2.0;
  double b = (o_mval (o, grid, cells[2]) + o_mval (o, grid, cells[3])) / 2.0;
  return (a + b) / 2.0;
}

/* InternalScheme::getMetaMaterial + Approximation::approximateDrudeModel (Approximation.cpp:58-167). */
static double o_get_meta (const oracle *o, int comp, int64_t i, int64_t j, int64_t k, const double *eps,
                          const double *om, const double *ga, double *omega, double *gamma)
{
  int64_t cells[4][3];
  int nc = o_material_cells (o, comp, i, j, k, cells);
  if (nc == 2)
  {
    *omega = (o_mval (o, om, cells[0]) + o_mval (o, om, cells[1])) / 2.0;
    *gamma = (o_mval (o, ga, cells[0]) + o_mval (o, ga, cells[1])) / 2.0;
    return (o_mval (o, eps, cells[0]) + o_mval (o, eps, cells[1])) / 2.0;
  }
  *omega = (o_mval (o, om, cells[0]) + o_mval (o, om, cells[1]) + o_mval (o, om, cells[2]) + o_mval (o, om, cells[3])) / 4.0;
  *gamma = (o_mval (o, ga, cells[0]) + o_mval (o, ga, cells[1]) + o_mval (o, ga, cells[2]) + o_mval (o, ga, cells[3])) / 4.0;
  double a = (o_mval (o, eps, cells[0]) + o_mval (o, eps, cells[1])) / 2.0;
  double b = (o_mval (o, eps, cells[2]) + o_mval (o, eps, cells[3])) / 2.0;
  return (a + b) / 2.0;
}

/* --- TF/SF predicates: YeeGridLayoutHelper::tfsfBorder (Source/Layout/YeeGridLayout.h:498-547) --- */

static int o_fp_exact (double a, double b)
{
  return (a > b ? a - b : b - a) < 0.0000001;   /* IS_FP_EXACT, FPEXACT_ACCURACY (FieldValue.h:150-160) */
}

static int o_border_eq (double c, double border, int off)
{
  return o_fp_exact (c, border + off * 1.0 / 10.0);
}

static int o_border_range (double c, double b1, int o1, double b2, int o2)
{
  return c > b1 + o1 * 1.0 / 10.0 && c < b2 + o2 * 1.0 / 10.0;
}

/*
 * doNeedTFSFUpdate<comp>Border(pos, dir), E_CENTERED, Dim3
 * (Source/Layout/YeeGridLayout.cpp:272-306, 558-596, 844-882, 1130-1168, 1416-1454, 1702-1740)
 * and Dim2_TMz (:817-826, 1103-1112, 1389-1398).  All 24 3D predicates share one shape:
 * for an E component of axis a and a direction along axis b, the coordinate along b
 * equals left-0.5 / right+0.5, the coordinate along a lies within [left-0.1, right+0.1]
 * and the third coordinate within [left+0.4, right-0.4]; for H components the b
 * coordinate equals left / right, a lies within [left+0.4, right-0.4] and the third
 * within [left-0.1, right+0.1].  The TMz predicates are the same with the z test
 * removed (tfsfBorder2D*, YeeGridLayout.h:520-547).
 */
static int o_tfsf_pred (const oracle *o, int comp, int dir, const double real[3])
{
  const comp_desc *c = &CD[comp];
  int b = dir / 2;
  int low = (dir % 2) == 0;
  if (b == c->axis || !o->act[b])
  {
    return 0;
  }
  int third = 3 - b - c->axis;
  int cond_b, cond_a = 1, cond_t = 1;
  if (!c->is_h)
  {
    cond_b = low ? o_border_eq (real[b], o->bl[b], -5) : o_border_eq (real[b], o->br[b], 5);
    if (o->act[c->axis]) cond_a = o_border_range (real[c->axis], o->bl[c->axis], -1, o->br[c->axis], 1);
    if (o->act[third]) cond_t = o_border_range (real[third], o->bl[third], 4, o->br[third], -4);
  }
  else
  {
    cond_b = low ? o_border_eq (real[b], o->bl[b], 0) : o_border_eq (real[b], o->br[b], 0);
    if (o->act[c->axis]) cond_a = o_border_range (real[c->axis], o->bl[c->axis], 4, o->br[c->axis], -4);
    if (o->act[third]) cond_t = o_border_range (real[third], o->bl[third], -1, o->br[third], 1);
  }
  return cond_b && cond_a && cond_t;
}

/* Incident-field projections get<comp>FromIncident{E,H} (Source/Layout/YeeGridLayout.h:2194-2356;
 * Dim2_TEz :1939-2024). */
static double o_project (const oracle *o, int comp, double v)
{
  double a1 = o->a1, a2 = o->a2, a3 = o->a3;
  const double s1 = sin (a1), c1 = cos (a1);
  switch (o->cfg.scheme)
  {
    /* the 1D schemes (YeeGridLayout.h:1247-1760) */
    case YEE_SCHEME_EXHY: return comp == YEE_EX ? -v : comp == YEE_HY ? -v : 0;
    case YEE_SCHEME_EXHZ: return comp == YEE_EX ? v : comp == YEE_HZ ? -v : 0;
    case YEE_SCHEME_EYHX: return comp == YEE_EY ? -v : comp == YEE_HX ? v : 0;
    case YEE_SCHEME_EYHZ: return comp == YEE_EY ? -v : comp == YEE_HZ ? -v : 0;
    case YEE_SCHEME_EZHX: return comp == YEE_EZ ? v : comp == YEE_HX ? v : 0;
    case YEE_SCHEME_EZHY: return comp == YEE_EZ ? v : comp == YEE_HY ? -v : 0;
    case YEE_SCHEME_TEX:   /* YeeGridLayout.h:1761-1846 */
      return comp == YEE_EY ? v * (-c1) : comp == YEE_EZ ? v * s1 : comp == YEE_HX ? v : 0;
    case YEE_SCHEME_TEY:   /* :1847-1932 */
      return comp == YEE_EX ? v * (-c1) : comp == YEE_EZ ? v * s1 : comp == YEE_HY ? -v : 0;
    case YEE_SCHEME_TMX:   /* :2025-2104 */
      return comp == YEE_EX ? v : comp == YEE_HY ? v * c1 : comp == YEE_HZ ? -v * s1 : 0;
    case YEE_SCHEME_TMY:   /* :2105-2190 */
      return comp == YEE_EY ? -v : comp == YEE_HX ? v * c1 : comp == YEE_HZ ? -v * s1 : 0;
    default:
      break;
  }
  if (o->dim == 2 && o->cfg.scheme == YEE_SCHEME_TEZ)
  {
    switch (comp)
    {
      case YEE_EX: return v * (sin (a2));
      case YEE_EY: return v * (-cos (a2));
      case YEE_HZ: return -v;
      default: return 0;
    }
  }
  if (o->dim == 2)
  {
    switch (comp)
    {
      case YEE_EZ: return v;
      case YEE_HX: return v * (sin (a2));
      case YEE_HY: return v * (-cos (a2));
      default: return 0;
    }
  }
  switch (comp)
  {
    case YEE_EX: return v * (cos (a3) * sin (a2) - sin (a3) * cos (a1) * cos (a2));
    case YEE_EY: return v * (-cos (a3) * cos (a2) - sin (a3) * cos (a1) * sin (a2));
    case YEE_EZ: return v * (sin (a3) * sin (a1));
    case YEE_HX: return v * (sin (a3) * sin (a2) + cos (a3) * cos (a1) * cos (a2));
    case YEE_HY: return v * (-sin (a3) * cos (a2) + cos (a3) * cos (a1) * sin (a2));
    case YEE_HZ: return -v * (cos (a3) * sin (a1));
    default: return 0;
  }
}

/*
 * approximateIncidentWave (Source/Scheme/InternalScheme.cpp:375-407) and
 * approximateIncidentWaveHelper (Source/Scheme/InternalScheme.inc.h:500-515):
 * projection of the point onto the propagation direction, linear interpolation
 * of layer 1 of the 1D grid at the truncated index.
 */
static double o_incident (const oracle *o, const double real[3], double ddiff, const double *inc)
{
  double d;
  double x = real[0] - o->zero_inc[0];
  double y = real[1] - o->zero_inc[1];
  double z = real[2] - o->zero_inc[2];
  if (o->dim == 1)
  {
    /* the 1D schemes: the coordinate along the grid (InternalScheme.cpp:110-200) */
    d = (o->act[0] ? x : o->act[1] ? y : z) - ddiff;
  }
  else if (o->dim == 3)
  {
    d = x * sin (o->a1) * cos (o->a2) + y * sin (o->a1) * sin (o->a2) + z * cos (o->a1) - ddiff;
  }
  else if (!o->act[0])
  {
    d = y * sin (o->a1) + z * cos (o->a1) - ddiff;   /* Dim2_TEx / TMx (InternalScheme.cpp:210-222) */
  }
  else if (!o->act[1])
  {
    d = x * sin (o->a1) + z * cos (o->a1) - ddiff;   /* Dim2_TEy / TMy (:227-239) */
  }
  else
  {
    d = x * cos (o->a2) + y * sin (o->a2) - ddiff;   /* Dim2_TEz / TMz (:244-256) */
  }
  double c1 = (double) ((int64_t) d);
  double c2 = c1 + 1;
  double p2 = d - c1;
  double p1 = 1 - p2;
  double v1 = inc[(int64_t) c1];
  double v2 = inc[(int64_t) c2];
  return v1 * p1 + v2 * p2;
}

double oracle_wave_number_general (double delta, double lambda, double courant, double nl, double a1, double a2);

/* Approximation::approximateWaveNumber{,General} (Source/Layout/Approximation.cpp:259-339). */
double oracle_wave_number (double delta, double lambda, double courant, double nl, double a1, double a2)
{
  const double pi = O_PI;
  if ((a1 == 0 || a1 == pi / 2 || a1 == pi || a1 == 3 * pi / 2)
      && (a2 == 0 || a2 == pi / 2 || a2 == pi || a2 == 3 * pi / 2))
  {
    return 2.0 / delta * asin (sin (pi * courant / nl) / courant);
  }
  if (((a1 == 0 || a1 == pi / 2 || a1 == pi || a1 == 3 * pi / 2)
       && (a2 == pi / 4 || a2 == 3 * pi / 4 || a2 == 5 * pi / 4 || a2 == 7 * pi / 4))
      || ((a1 == pi / 4 || a1 == 3 * pi / 4 || a1 == 5 * pi / 4 || a1 == 7 * pi / 4)
          && (a2 == 0 || a2 == pi / 2 || a2 == pi || a2 == 3 * pi / 2)))
  {
    return 2.0 * sqrt (2.0) / delta * asin (sin (pi * courant / nl) / (courant * sqrt (2.0)));
  }
  return oracle_wave_number_general (delta, lambda, courant, nl, a1, a2);
}

/* Newton solution of the 3D Yee dispersion relation (Approximation.cpp:313-339). */
double oracle_wave_number_general (double delta, double lambda, double courant, double nl, double a1, double a2)
{
  const double pi = O_PI;
  const double acc = 0.0000001;
  double k = 2 * pi;
  double k_prev = k + 2 * sqrt (acc);
  double nd = delta / lambda;
  double A = nd * sin (a1) * cos (a2) / 2;
  double B = nd * sin (a1) * sin (a2) / 2;
  double C = nd * cos (a1) / 2;
  double s = sin (pi * courant / nl);
  double D = (s * s) / (courant * courant);
  while ((k_prev - k) * (k_prev - k) >= acc)
  {
    k_prev = k;
    double sa = sin (A * k), sb = sin (B * k), sc = sin (C * k);
    double diff1 = sa * sa + sb * sb + sc * sc - D;
    double diff2 = A * sin (2 * A * k) + B * sin (2 * B * k) + C * sin (2 * C * k);
    double diff = diff1 / diff2;
    k -= diff;
  }
  return k / lambda;
}

/* Approximation::phaseVelocityIncidentWave (Approximation.cpp:341-353). */
static double o_phase_velocity (double delta, double lambda, double courant, double nl, double a1, double a2)
{
  return O_C * 2 * O_PI / lambda / oracle_wave_number (delta, lambda, courant, nl, a1, a2);
}

/* SchemeHelper::initSigma (Source/Scheme/SchemeHelper.h:574-590). */
static double o_sigma_value (double dist, double boundary, double dx)
{
  double eps0 = o_eps0 ();
  double mu0 = o_mu0 ();
  unsigned exponent = 6;
  double r_err = 1e-16;
  double sigma_max_1 = -log (r_err) * (exponent + 1.0) / (2.0 * sqrt (mu0 / eps0) * boundary);
  double factor = sigma_max_1 / (dx * (pow (boundary, (double) exponent)) * (exponent + 1));
  double x1 = (dist + 1) * dx;
  double x2 = dist * dx;
  return factor * (pow (x1, (double) (exponent + 1)) - pow (x2, (double) (exponent + 1)));
}

/* Approximation::approximateSphereAccurate, 3D (Source/Layout/Approximation.cpp:563-676). */
double oracle_sphere_accurate_3d (double mx, double my, double mz, double cx, double cy, double cz,
                                  double radius, double eps, double outside, int accuracy)
{
  double sx = mx - 0.5, sy = my - 0.5, sz = mz - 0.5;
  double ez = mz + 0.5;
  int num = accuracy;
  double step = 1.0 / num;
  double elem = step * step;
  double volume = 0;
  for (int i = 0; i < num; ++i)
  {
    for (int j = 0; j < num; ++j)
    {
      double px = sx + i * step;
      double py = sy + j * step;
      double pz;
      double temp = radius * radius - (px - cx) * (px - cx) - (py - cy) * (py - cy);
      if (temp < 0)
      {
        pz = 0.0;
      }
      else
      {
        int above;
        if (mz > cz)
        {
          above = 1;
          pz = cz + sqrt (temp);
        }
        else
        {
          above = 0;
          pz = cz - sqrt (temp);
        }
        if (pz < sz) pz = sz;
        else if (pz > ez) pz = ez;
        pz = above ? pz - sz : ez - pz;
      }
      volume += pz * elem;
    }
  }
  if (volume > 1.0) volume = 1.0;
  return eps * volume + outside * (1 - volume);
}

/* Approximation::approximateSphereAccurate, 2D (Source/Layout/Approximation.cpp:471-561). */
/* approximateSphereAccurate, 1D (Approximation.cpp:390-448): its inside test is
 * "start >= left || end <= right", so a cell overlapping the segment is filled completely. */
double oracle_sphere_accurate_1d (double m, double c, double radius, double eps, double outside)
{
  const double start = m - 0.5, end = m + 0.5, left = c - radius, right = c + radius;
  double volume = 0;
  if (end <= left || start >= right)
    ;
  else if (start >= left || end <= right)
    volume = 1;
  else if (start < left)
    volume = end - left;
  else if (end > right)
    volume = right - start;
  if (volume > 1) volume = 1;
  return eps * volume + outside * (1 - volume);
}

/* approximateSphereStair, 1D (Approximation.cpp:678-699): both ends within the radius. */
double oracle_sphere_stair_1d (double m, double c, double radius, double eps, double outside)
{
  const double a = m - 0.5 - c, b = m + 0.5 - c;
  return (sqrt (a * a) <= radius && sqrt (b * b) <= radius) ? eps : outside;
}

double oracle_sphere_accurate_2d (double mx, double my, double cx, double cy,
                                  double radius, double eps, double outside, int accuracy)
{
  double sx = mx - 0.5, sy = my - 0.5, ey = my + 0.5;
  int num = accuracy;
  double step = 1.0 / num;
  double elem = step;
  double volume = 0;
  for (int i = 0; i < num; ++i)
  {
    double px = sx + i * step;
    double py;
    double temp = radius * radius - (px - cx) * (px - cx);
    if (temp < 0)
    {
      py = 0.0;
    }
    else
    {
      int above;
      if (my > cy)
      {
        above = 1;
        py = cy + sqrt (temp);
      }
      else
      {
        above = 0;
        py = cy - sqrt (temp);
      }
      if (py < sy) py = sy;
      else if (py > ey) py = ey;
      py = above ? py - sy : ey - py;
    }
    volume += py * elem;
  }
  if (volume > 1.0) volume = 1.0;
  return eps * volume + outside * (1 - volume);
}

/* Approximation::approximateSphereStair, 3D (Source/Layout/Approximation.cpp:730-764). */
double oracle_sphere_stair_3d (double mx, double my, double mz, double cx, double cy, double cz,
                               double radius, double eps, double outside)
{
  for (int q = 0; q < 8; ++q)
  {
    double x = mx + ((q & 1) ? 0.5 : -0.5);
    double y = my + ((q & 2) ? 0.5 : -0.5);
    double z = mz + ((q & 4) ? 0.5 : -0.5);
    double dist = sqrt ((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz));
    if (!(dist <= radius))
    {
      return outside;
    }
  }
  return eps;
}

oracle *oracle_create (const yee_config *cfg)
{
  if (!cfg || cfg->layout != 0 || cfg->scheme < YEE_SCHEME_EXHY || cfg->scheme > YEE_SCHEME_DIM3)
  {
    return NULL;
  }
  oracle *o = (oracle *) calloc (1, sizeof (oracle));
  o->cfg = *cfg;
  o->dim = cfg->scheme == YEE_SCHEME_DIM3 ? 3 : cfg->scheme >= YEE_SCHEME_TEX ? 2 : 1;
  /* 2D schemes lack one axis (main.cpp:817-1000: TEx/TMx on (y, z), TEy/TMy on (x, z), TEz/TMz on
   * (x, y)); the grids keep the order of the remaining two, which a 3D index with extent 1 along
   * the missing axis reproduces */
  const int missing = cfg->scheme == YEE_SCHEME_DIM3 ? -1
                      : (cfg->scheme == YEE_SCHEME_TEX || cfg->scheme == YEE_SCHEME_TMX) ? 0
                      : (cfg->scheme == YEE_SCHEME_TEY || cfg->scheme == YEE_SCHEME_TMY) ? 1 : 2;
  /* 1D schemes run along the axis across both of their components (main.cpp:629-800: ExHy and
   * EyHx along z, ExHz and EzHx along y, EyHz and EzHy along x) */
  static const int k1dAxis[7] = { -1, 2, 1, 2, 0, 1, 0 };
  for (int a = 0; a < 3; ++a)
  {
    o->act[a] = o->dim == 1 ? a == k1dAxis[cfg->scheme] : a != missing;
    o->n[a] = o->act[a] ? cfg->size[a] : 1;
    o->lo[a] = o->act[a] ? 1 : 0;
    o->hi[a] = o->act[a] ? o->n[a] - 1 : 1;
  }
  o->ncell = o->n[0] * o->n[1] * o->n[2];
  /* --use-double-material-precision: sizeEps = 2 size along the scheme's axes (YeeGridLayout.cpp:1972) */
  o->fine = (cfg->flags & YEE_DOUBLE_MATERIAL_PRECISION) != 0;
  for (int a = 0; a < 3; ++a) o->mn[a] = (o->fine && o->act[a]) ? 2 * o->n[a] : o->n[a];
  o->mcell = o->mn[0] * o->mn[1] * o->mn[2];
  o->use_pml = (cfg->flags & YEE_USE_PML) != 0;
  o->use_tfsf = (cfg->flags & YEE_USE_TFSF) != 0;
  o->use_meta = (cfg->flags & YEE_USE_METAMATERIALS) != 0;
  if (o->use_meta && !o->use_pml)
  {
    free (o);
    return NULL;   /* Scheme.inc.h:889-892: metamaterials without PML are unreachable */
  }
  if (o->use_meta && o->fine)
  {
    free (o);
    return NULL;   /* the 8-cell Drude average is UNREACHABLE (Approximation.cpp:196) */
  }
  o->layers = (o->use_meta && o->use_pml) ? 3 : 2;
  for (int c = 0; c < 6; ++c)
  {
    /* InternalScheme component flags (InternalScheme.inc.h:3243-3272): TMz Ez, Hx, Hy; TEz Ex, Ey, Hz */
    /* TE<a>: H along the missing axis a and the two E components across it; TM<a> the reverse */
    /* 1D: the two components the scheme names (InternalScheme.inc.h:3243-3260) */
    static const int k1dComps[7][2] = { { -1, -1 }, { YEE_EX, YEE_HY }, { YEE_EX, YEE_HZ }, { YEE_EY, YEE_HX },
                                        { YEE_EY, YEE_HZ }, { YEE_EZ, YEE_HX }, { YEE_EZ, YEE_HY } };
    if (o->dim == 3) o->present[c] = 1;
    else if (o->dim == 1) o->present[c] = c == k1dComps[cfg->scheme][0] || c == k1dComps[cfg->scheme][1];
    else
    {
      const int te = cfg->scheme <= YEE_SCHEME_TEZ;
      const int along = (c % 3) == missing;
      o->present[c] = (c >= 3) == te ? along : !along;
    }
  }
  for (int m = 0; m < 6; ++m)
  {
    o->mat[m] = (double *) malloc (sizeof (double) * o->mcell);
    double v = m < 2 ? 1.0 : 0.0;
    for (int64_t q = 0; q < o->mcell; ++q) o->mat[m][q] = v;
  }
  for (int c = 0; c < 6; ++c)
  {
    if (!o->present[c]) continue;
    for (int l = 0; l < o->layers; ++l)
    {
      o->f[c][l] = (double *) calloc (o->ncell, sizeof (double));
      if (o->use_pml) o->d[c][l] = (double *) calloc (o->ncell, sizeof (double));
      if (o->use_meta) o->d1[c][l] = (double *) calloc (o->ncell, sizeof (double));
    }
  }
  /* InternalScheme::initScheme (Source/Scheme/InternalScheme.inc.h:3359-3394) */
  o->dx = cfg->dx;
  o->lambda = cfg->wavelength;
  o->courant = cfg->courant;
  o->freq = O_C / o->lambda;
  o->dt = o->dx * o->courant / O_C;
  /* angles as main.cpp:1101-1110 converts them */
  o->a1 = cfg->angle_deg[0] * O_PI / 180.0;
  o->a2 = cfg->angle_deg[1] * O_PI / 180.0;
  o->a3 = cfg->angle_deg[2] * O_PI / 180.0;
  double nl = o->lambda / o->dx;
  double pv0 = o_phase_velocity (o->dx, o->lambda, o->courant, nl, O_PI / 2, 0);
  double pv = o_phase_velocity (o->dx, o->lambda, o->courant, nl, o->a1, o->a2);
  o->rel = pv0 / pv;
  /* TF/SF borders and zeroIncCoordFP (YeeGridLayout.cpp:1954-1958, 2583-2599, 2636-2656) */
  for (int a = 0; a < 3; ++a)
  {
    o->bl[a] = 0.0 + (double) cfg->tfsf_left[a];
    o->br[a] = 0.0 + (double) (cfg->size[a] - cfg->tfsf_right[a]);
  }
  if (o->dim == 1)
  {
    /* the 1D layouts: leftBorderTotalField - 2.5 along the scheme's axis (YeeGridLayout.cpp:2062-2263) */
    for (int a = 0; a < 3; ++a) o->zero_inc[a] = o->act[a] ? cfg->tfsf_left[a] - 2.5 : 0;
  }
  else if (o->dim == 3)
  {
    o->zero_inc[0] = cfg->tfsf_left[0] - 2.5 * sin (o->a1) * cos (o->a2);
    o->zero_inc[1] = cfg->tfsf_left[1] - 2.5 * sin (o->a1) * sin (o->a2);
    o->zero_inc[2] = cfg->tfsf_left[2] - 2.5 * cos (o->a1);
  }
  else if (!o->act[0] || !o->act[1])
  {
    /* TEx, TMx (y, z) and TEy, TMy (x, z): YeeGridLayout.cpp:2332-2333, 2385-2386, 2491-2492, 2544-2545 */
    const int first = o->act[0] ? 0 : 1;
    o->zero_inc[first] = cfg->tfsf_left[first] - 2.5 * sin (o->a1);
    o->zero_inc[2] = cfg->tfsf_left[2] - 2.5 * cos (o->a1);
    o->zero_inc[o->act[0] ? 1 : 0] = 0;
  }
  else
  {
    o->zero_inc[0] = cfg->tfsf_left[0] - 2.5 * cos (o->a2);
    o->zero_inc[1] = cfg->tfsf_left[1] - 2.5 * sin (o->a2);
    o->zero_inc[2] = 0;
  }
  if (o->use_tfsf)
  {
    /* allocateGridsInc: 500 * size.x cells, 2 layers (InternalScheme.specific.h:99-107) */
    o->inc_len = 500 * (o->act[0] ? o->n[0] : o->act[1] ? o->n[1] : o->n[2]);   /* size.get1 () of the scheme's grid */
    for (int l = 0; l < 2; ++l)
    {
      o->einc[l] = (double *) calloc (o->inc_len, sizeof (double));
      o->hinc[l] = (double *) calloc (o->inc_len, sizeof (double));
    }
  }
  return o;
}

void oracle_destroy (oracle *o)
{
  if (!o) return;
  for (int m = 0; m < 6; ++m) free (o->mat[m]);
  for (int a = 0; a < 3; ++a) free (o->sigma[a]);
  for (int c = 0; c < 6; ++c)
    for (int l = 0; l < 3; ++l)
    {
      free (o->f[c][l]);
      free (o->d[c][l]);
      free (o->d1[c][l]);
    }
  for (int l = 0; l < 2; ++l)
  {
    free (o->einc[l]);
    free (o->hinc[l]);
  }
  free (o);
}

int oracle_set_material (oracle *o, int32_t which, const double *values)
{
  if (!o || !values || which < 0 || which > 5) return -1;
  memcpy (o->mat[which], values, sizeof (double) * o->mcell);
  return 0;
}

/* SchemeHelper::initSigmaX/Y/Z (Source/Scheme/SchemeHelper.h:592-718); x and y/z differ on purpose. */
int oracle_prepare (oracle *o)
{
  if (!o) return -1;
  if (o->use_pml)
  {
    int first = 0;
    while (first < 2 && !o->act[first]) ++first;
    for (int a = 0; a < 3; ++a)
    {
      if (!o->act[a]) continue;
      /* double material precision: the fine grid, PML size 2 P (SchemeHelper.h:598, 640, 682) */
      o->sigma[a] = (double *) calloc (o->mcell, sizeof (double));
      double P = (double) (o->cfg.pml[a] * (o->fine ? 2 : 1));
      double boundary = P * o->dx;
      for (int64_t i = 0; i < o->mn[0]; ++i)
        for (int64_t j = 0; j < o->mn[1]; ++j)
          for (int64_t k = 0; k < o->mn[2]; ++k)
          {
            int64_t p[3] = { i, j, k };
            double pos = (double) p[a] + 0.5;   /* getEpsCoordFP */
            double val = 0;
            /* initSigmaX on the scheme's first coordinate (Scheme.cpp:141-420: y for TEx / TMx),
             * initSigmaY / Z on the others */
            if (a == first)
            {
              double size = (double) o->mn[a];
              if (pos < P) val = o_sigma_value (P - pos, boundary, o->dx);
              else if (pos >= size - P) val = o_sigma_value (pos - (size - P), boundary, o->dx);
            }
            else
            {
              double size = (double) o->mn[a] + 0.5;
              if (pos < P) val = o_sigma_value ((double) (int64_t) (P - pos), boundary, o->dx);
              else if (pos >= size - P) val = o_sigma_value ((double) (int64_t) (pos - (size - P)), boundary, o->dx);
            }
            o->sigma[a][(i * o->mn[1] + j) * o->mn[2] + k] = val;
          }
    }
  }
  /* start values of an analytic test problem on every cell, E at dt / 2 and H at dt
   * (Scheme::initGridWithInitialVals, Scheme.inc.h:2863-2958, called at :3255-3279) */
  if (o->fn_start)
    for (int c = 0; c < 6; ++c)
    {
      if (!o->present[c] || !o_fn_has (o->fn_start, c)) continue;
      const double time = CD[c].is_h ? o->dt : 0.5 * o->dt;
      for (int64_t i = 0; i < o->n[0]; ++i)
        for (int64_t j = 0; j < o->n[1]; ++j)
          for (int64_t k = 0; k < o->n[2]; ++k)
          {
            double r[3];
            o_real_coord (o, c, i, j, k, r);
            o->f[c][0][o_idx (o, i, j, k)] = o_fn_value (o->fn_start, c, r[0] * o->dx, r[1] * o->dx, r[2] * o->dx, time);
          }
    }
  o->prepared = 1;
  return 0;
}

/*
 * Exact solutions of the 1D plane wave meeting a dielectric half-space (Tests/suite t7.x;
 * CallBack.cpp:280-760): incident E1 = i e^{i (k (z - z0 dx) - w (t - t0))}, reflected
 * E2 = -i/3 e^{i (-k (z - (2 zb - z0) dx) - w (t - t0))}, transmitted
 * E3 = 2i/3 e^{i (2k (z - (zb + z0)/2 dx) - w (t - t0))}, H = E k / (mu0 w) (-k for E2, 2k for
 * E3); k the numerical wave number (approximateWaveNumber), w = 2 pi c / lambda,
 * t0 = dt / 2, z0 = the TF/SF left size along the line - 2.5.  Per pair (ExHy, ExHz, EyHx, EyHz,
 * EzHx, EzHy) the line axis and the signs of E and H of the reference's wrappers.  The complex
 * value with the reference's complex operations (i * exponent (i arg) = (0 c - 1 s, 0 s + 1 c));
 * the real build returns its real part with the same operations.
 */
static __attribute__ ((noinline)) double o_sin_only (double x)
{
  return sin (x);
}

static __attribute__ ((noinline)) void o_sin_cos (double x, double *s, double *c)
{
  *c = cos (x);
  *s = sin (x);
}

static const int kExpAxis[6] = { 2, 1, 2, 0, 1, 0 };
static const int kExpE[6] = { YEE_EX, YEE_EX, YEE_EY, YEE_EY, YEE_EZ, YEE_EZ };
static const int kExpH[6] = { YEE_HY, YEE_HZ, YEE_HX, YEE_HZ, YEE_HX, YEE_HY };
static const double kExpSignE[6] = { -1, 1, -1, -1, 1, 1 };
static const double kExpSignH[6] = { -1, -1, 1, -1, 1, -1 };

static int o_exp_has (const oracle *o, int comp)
{
  return o->fn_exact >= YEE_FN_EXP1 && (comp == kExpE[o->exp_pair] || comp == kExpH[o->exp_pair]);
}

static void o_exp_value (const oracle *o, int comp, double coord, double t, double *re, double *im)
{
  const int pair = o->exp_pair, n = o->fn_exact - YEE_FN_EXP1 + 1;
  const double lambda = o->lambda, courant = o->courant, delta = o->dx;
  const double nl = lambda / delta;
  double k = oracle_wave_number (delta, lambda, courant, nl, o->a1, o->a2);
  const double f = O_C / lambda;
  const double w = 2 * O_PI * f;
  const double step = delta * courant / O_C;
  const double t0 = 0.5 * step;
  const double c0 = (double) o->cfg.tfsf_left[kExpAxis[pair]] - 2.5;
  const double cb = o->exp_border;
  double arg, hk;
  if (n == 1)
  {
    arg = k * (coord - c0 * delta) - w * (t - t0);
    hk = k;
  }
  else if (n == 2)
  {
    const double diff = 2 * cb - c0;
    arg = -k * (coord - diff * delta) - w * (t - t0);
    hk = -k;
  }
  else
  {
    k = 2 * k;
    const double diff = (cb + c0) / 2.0;
    arg = k * (coord - diff * delta) - w * (t - t0);
    hk = k;
  }
  /* the real build evaluates sin (arg) alone; the complex build CComplex (cos, sin) (Complex.h:139),
   * whose two calls the compiler merges into one sincos, which differs from sin in the last bit
   * for rare arguments: each build's form is kept apart */
  double s, c = 0.0;
  if (o->src_kind)
    o_sin_cos (arg, &s, &c);
  else
    s = o_sin_only (arg);
  double r, i;
  if (n == 1)
  {
    r = o->src_kind ? 0.0 * c - 1.0 * s : -s;
    i = 1.0 * c + 0.0 * s;
  }
  else if (n == 2)
  {
    /* (-i) (c, s) = (-0 c - (-1) s, (-1) c + -0 s), times 1, over 3 */
    r = o->src_kind ? ((-0.0 * c - (-1.0) * s) * 1.0) / 3.0 : s * 1.0 / 3.0;
    i = ((-1.0 * c + -0.0 * s) * 1.0) / 3.0;
  }
  else
  {
    r = o->src_kind ? ((0.0 * c - 1.0 * s) * 2.0) / 3.0 : -s * 2.0 / 3.0;
    i = ((1.0 * c + 0.0 * s) * 2.0) / 3.0;
  }
  if (comp >= 3)
  {
    const double m = hk / (o_mu0 () * w);
    r = r * m;
    i = i * m;
  }
  const double sg = comp >= 3 ? kExpSignH[pair] : kExpSignE[pair];
  *re = sg < 0 ? -r : r;
  *im = sg < 0 ? -i : i;
}

/* Analytic test problems (YEE_FN_* kinds; 0 = none) and the norm box: [1, N-1) on each axis of the
 * scheme, an axis with a nonzero start / end taking it (--norm-start / --norm-end,
 * Scheme.inc.h:1394-1420).  Rejects PML (the border and norm then act on D / E grids this
 * restatement does not follow) and start values without border values (the reference then
 * alternates its boundary cells between the start values and zero with its two time layers). */
int oracle_set_test_functions (oracle *o, int start, int border, int rhs, int exact, const int64_t ns[3],
                               const int64_t ne[3])
{
  if (!o || o->prepared) return -1;
  if (start > YEE_FN_SIN1 || border > YEE_FN_SIN1 || rhs > YEE_FN_SIN1 || exact > YEE_FN_SIN1) return -1;
  /* with PML the border values go to D; the norm reads E (normGrid = gridPML2) */
  if ((start || border || rhs) && o->use_pml) return -1;
  if (start && !border) return -1;
  o->fn_start = start;
  o->fn_border = border;
  o->fn_rhs = rhs;
  o->fn_exact = exact;
  for (int a = 0; a < 3; ++a)
  {
    o->norm_lo[a] = o->act[a] ? 1 : 0;
    o->norm_hi[a] = o->act[a] ? o->n[a] - 1 : 1;
    if (o->act[a] && ns && ns[a] != 0) o->norm_lo[a] = ns[a];
    if (o->act[a] && ne && ne[a] != 0) o->norm_hi[a] = ne[a];
  }
  return 0;
}

/*
 * Near-to-far-field transform of the complex build (Scheme::Pointing_scat, Scheme.inc.h:4036-4101,
 * over the faces of SchemeHelper::ntffN3D / ntffL3D, SchemeHelper.h:1030-2148, E_CENTERED): the
 * real and imaginary parts come from two oracles (re, im) stepped in lock step.  N sums the
 * surface currents from H averaged over its newest two time layers, L from E averaged across the
 * face; both over the cells [left + 1/2, right - 1/2] of each face of the box, with the phase
 * exp (i k r.n) of the cell's offset from the grid centre.  Returns Pointing_scat / Pointing_inc.
 */
typedef struct { double re, im; } o_cplx;

static o_cplx o_cadd (o_cplx a, o_cplx b) { o_cplx r = { a.re + b.re, a.im + b.im }; return r; }
static o_cplx o_cscale (o_cplx a, double s) { o_cplx r = { a.re * s, a.im * s }; return r; }
static o_cplx o_cmul (o_cplx a, o_cplx b) { o_cplx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re }; return r; }

/* minimum coordinates, E_CENTERED (YeeGridLayout.cpp:1962-1969) */
static const double kMinCoord[6][3] = { { 1.0, 0.5, 0.5 }, { 0.5, 1.0, 0.5 }, { 0.5, 0.5, 1.0 },
                                        { 0.5, 1.0, 1.0 }, { 1.0, 0.5, 1.0 }, { 1.0, 1.0, 0.5 } };

/* value of component c at real position p (integer after subtracting the minimum coordinate) of
 * layer `layer` as a complex number */
static o_cplx o_at (const oracle *re, const oracle *im, int c, const double p[3], int layer)
{
  int64_t q[3];
  for (int a = 0; a < 3; ++a) q[a] = (int64_t) (p[a] - kMinCoord[c][a]);
  const int64_t idx = o_idx (re, q[0], q[1], q[2]);
  o_cplx v = { re->f[c][layer][idx], im->f[c][layer][idx] };
  return v;
}

/* N (kind 0) or L (kind 1) of face axis `ax` at plane p0 */
static void o_ntff_face (const oracle *re, const oracle *im, int kind, int ax, int64_t p0, double teta, double phi,
                         const int64_t left[3], const int64_t right[3], o_cplx *sum_teta, o_cplx *sum_phi)
{
  const double dx = re->dx;
  const double diff0[3] = { (double) (re->n[0] / 2), (double) (re->n[1] / 2), (double) (re->n[2] / 2) };
  const int u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;   /* the two in-plane axes, in order */
  const double sign = p0 == right[ax] ? 1.0 : -1.0;
  const double k = 2 * O_PI / re->lambda;
  /* components: N: H (Hz/Hy on x, Hz/Hx on y, Hy/Hx on z); L: E likewise */
  static const int kComp[2][3][2] = { { { YEE_HZ, YEE_HY }, { YEE_HZ, YEE_HX }, { YEE_HY, YEE_HX } },
                                      { { YEE_EY, YEE_EZ }, { YEE_EX, YEE_EZ }, { YEE_EX, YEE_EY } } };
  const int c1 = kComp[kind][ax][0], c2 = kComp[kind][ax][1];
  o_cplx st = { 0, 0 }, sp = { 0, 0 };
  for (double cu = left[u] + 0.5; cu <= right[u] - 0.5; ++cu)
    for (double cv = left[v] + 0.5; cv <= right[v] - 0.5; ++cv)
    {
      double base[3];
      base[ax] = (double) p0;
      base[u] = cu;
      base[v] = cv;
      /* pos1 / pos2: -/+ 1/2 along u (first component), pos3 / pos4: -/+ 1/2 along v (second) */
      double p1[3] = { base[0], base[1], base[2] }, p2[3] = { base[0], base[1], base[2] };
      double p3[3] = { base[0], base[1], base[2] }, p4[3] = { base[0], base[1], base[2] };
      p1[u] -= 0.5; p2[u] += 0.5; p3[v] -= 0.5; p4[v] += 0.5;
      o_cplx a1, a2, b1, b2;
      if (kind == 0)
      {
        /* H averaged over the newest two time layers (CurrentAfterShift: layer 1;
         * PreviousAfterShift: layer 2 of three, else layer 0; Grid.h:1184-1208) */
        const int pl = re->layers > 2 ? 2 : 0;
        a1 = o_cscale (o_cadd (o_at (re, im, c1, p1, 1), o_at (re, im, c1, p1, pl)), 1.0 / 2);
        a2 = o_cscale (o_cadd (o_at (re, im, c1, p2, 1), o_at (re, im, c1, p2, pl)), 1.0 / 2);
        b1 = o_cscale (o_cadd (o_at (re, im, c2, p3, 1), o_at (re, im, c2, p3, pl)), 1.0 / 2);
        b2 = o_cscale (o_cadd (o_at (re, im, c2, p4, 1), o_at (re, im, c2, p4, pl)), 1.0 / 2);
        /* the reference divides by 2 (FPValue (2)); / 2 and * 1/2 are the same in binary FP */
      }
      else
      {
        /* E averaged across the face (-/+ 1/2 along the face normal) */
        double q[4][2][3];
        const double *ps[4] = { p1, p2, p3, p4 };
        for (int w = 0; w < 4; ++w)
          for (int s = 0; s < 2; ++s)
          {
            for (int a = 0; a < 3; ++a) q[w][s][a] = ps[w][a];
            q[w][s][ax] += s ? 0.5 : -0.5;
          }
        a1 = o_cscale (o_cadd (o_at (re, im, c1, q[0][0], 1), o_at (re, im, c1, q[0][1], 1)), 1.0 / 2);
        a2 = o_cscale (o_cadd (o_at (re, im, c1, q[1][0], 1), o_at (re, im, c1, q[1][1], 1)), 1.0 / 2);
        b1 = o_cscale (o_cadd (o_at (re, im, c2, q[2][0], 1), o_at (re, im, c2, q[2][1], 1)), 1.0 / 2);
        b2 = o_cscale (o_cadd (o_at (re, im, c2, q[3][0], 1), o_at (re, im, c2, q[3][1], 1)), 1.0 / 2);
      }
      const double r[3] = { base[0] - diff0[0], base[1] - diff0[1], base[2] - diff0[2] };
      double arg = r[0] * sin (teta) * cos (phi) + r[1] * sin (teta) * sin (phi) + r[2] * cos (teta);
      arg *= dx;
      const o_cplx ex = { cos (k * arg), sin (k * arg) };
      const o_cplx A = o_cscale (o_cadd (a1, a2), 1.0 / 2), B = o_cscale (o_cadd (b1, b2), 1.0 / 2);
      o_cplx t, f;
      double ts, fs;   /* extra factor -1 of the teta / phi sums */
      if (ax == 0)
      {
        /* N_x: A = Hz, B = Hy; L_x: A = Ey, B = Ez */
        if (kind == 0)
        {
          t = o_cadd (o_cscale (o_cscale (A, cos (teta)), sin (phi)), o_cscale (B, sin (teta)));
          f = o_cscale (A, cos (phi));
          ts = -1; fs = -1;
        }
        else
        {
          t = o_cadd (o_cscale (o_cscale (B, cos (teta)), sin (phi)), o_cscale (A, sin (teta)));
          f = o_cscale (B, cos (phi));
          ts = 1; fs = 1;
        }
      }
      else if (ax == 1)
      {
        /* N_y: A = Hz, B = Hx; L_y: A = Ex, B = Ez */
        if (kind == 0)
        {
          t = o_cadd (o_cscale (o_cscale (A, cos (teta)), cos (phi)), o_cscale (B, sin (teta)));
          f = o_cscale (A, sin (phi));
          ts = 1; fs = -1;
        }
        else
        {
          t = o_cadd (o_cscale (o_cscale (B, cos (teta)), cos (phi)), o_cscale (A, sin (teta)));
          f = o_cscale (B, sin (phi));
          ts = -1; fs = 1;
        }
      }
      else
      {
        /* N_z: A = Hy, B = Hx; L_z: A = Ex, B = Ey */
        if (kind == 0)
        {
          t = o_cadd (o_cscale (o_cscale (o_cscale (A, -1.0), cos (teta)), cos (phi)), o_cscale (o_cscale (B, cos (teta)), sin (phi)));
          f = o_cadd (o_cscale (A, sin (phi)), o_cscale (B, cos (phi)));
          ts = 1; fs = 1;
        }
        else
        {
          t = o_cadd (o_cscale (o_cscale (B, cos (teta)), cos (phi)), o_cscale (o_cscale (o_cscale (A, cos (teta)), sin (phi)), -1.0));
          f = o_cadd (o_cscale (B, sin (phi)), o_cscale (A, cos (phi)));
          ts = 1; fs = -1;
        }
      }
      st = o_cadd (st, o_cscale (o_cscale (o_cscale (o_cmul (t, ex), dx * dx), ts), sign));
      sp = o_cadd (sp, o_cscale (o_cscale (o_cscale (o_cmul (f, ex), dx * dx), fs), sign));
    }
  *sum_teta = o_cadd (*sum_teta, st);
  *sum_phi = o_cadd (*sum_phi, sp);
}

int oracle_ntff (const oracle *re, const oracle *im, double teta, double phi, const int64_t left[3],
                 const int64_t right[3], double *out)
{
  if (!re || !im || !out || re->dim != 3 || re->layers < 2) return -1;
  for (int a = 0; a < 3; ++a)
    if (re->n[a] % 2 || left[a] < 2 || right[a] > re->n[a] - 2 || right[a] <= left[a]) return -1;
  o_cplx nt = { 0, 0 }, np = { 0, 0 }, lt = { 0, 0 }, lp = { 0, 0 };
  for (int ax = 0; ax < 3; ++ax)
  {
    o_ntff_face (re, im, 0, ax, left[ax], teta, phi, left, right, &nt, &np);
    o_ntff_face (re, im, 0, ax, right[ax], teta, phi, left, right, &nt, &np);
    o_ntff_face (re, im, 1, ax, left[ax], teta, phi, left, right, &lt, &lp);
    o_ntff_face (re, im, 1, ax, right[ax], teta, phi, left, right, &lt, &lp);
  }
  const double n0 = sqrt (o_mu0 () / o_eps0 ());
  const double k = 2 * O_PI / re->lambda;
  const o_cplx first = o_cadd (lp, o_cscale (nt, n0));
  const o_cplx second = o_cadd (lt, o_cscale (np, -n0));
  const double a1 = first.re * first.re + first.im * first.im, a2 = second.re * second.re + second.im * second.im;
  const double scat = (k * k) / (8 * O_PI * n0) * (a1 + a2);
  out[0] = scat / (sqrt (o_eps0 () / o_mu0 ()) / 2);   /* Pointing_inc (Scheme.inc.h:4103-4109) */
  out[1] = scat;
  return 0;
}

void oracle_scale_fields (oracle *o, double s)
{
  for (int c = 0; c < 6; ++c)
    for (int l = 0; l < o->layers; ++l)
      if (o->f[c][l])
        for (int64_t i = 0; i < o->ncell; ++i) o->f[c][l][i] *= s;
}

/* Exact solution exp<kind> of a 1D pair (kind YEE_FN_EXP1 .. EXP3, pair YEE_SCHEME_EXHY .. EZHY)
 * for --calc-exp<n>-<pair>-diff-norm (Scheme.inc.h:2437-2760), border = the material border
 * coordinate (eps sphere centre - radius along the line), and the norm box. */
int oracle_set_exact_exp (oracle *o, int kind, int pair, double border, const int64_t ns[3], const int64_t ne[3])
{
  if (!o || o->prepared || kind < YEE_FN_EXP1 || kind > YEE_FN_EXP3 || pair < YEE_SCHEME_EXHY || pair > YEE_SCHEME_EZHY)
    return -1;
  if (oracle_set_test_functions (o, 0, 0, 0, 0, ns, ne)) return -1;
  o->fn_exact = kind;
  o->exp_pair = pair - YEE_SCHEME_EXHY;
  o->exp_border = border;
  return 0;
}

/* Module norm of the complex build at the last step: sqrt (sum (|exact| - |field|)^2 / cells) and
 * its percentage of max |exact|, from the real and imaginary part oracles at the same step. */
int oracle_exp_mod_norm (const oracle *re, const oracle *im, int32_t comp, double out[2])
{
  if (!re || !im || !out || comp < 0 || comp > 5 || !o_exp_has (re, comp) || re->t != im->t || re->t < 1) return -1;
  const int64_t t = re->t - 1;
  const double time = (comp >= 3 ? (double) t + 1.0 : (double) t + 0.5) * re->dt;
  const int axis = kExpAxis[re->exp_pair];
  double sum = 0.0, mx = 0.0;
  for (int64_t i = re->norm_lo[0]; i < re->norm_hi[0]; ++i)
    for (int64_t j = re->norm_lo[1]; j < re->norm_hi[1]; ++j)
      for (int64_t k = re->norm_lo[2]; k < re->norm_hi[2]; ++k)
      {
        double r[3], er, ei;
        o_real_coord (re, comp, i, j, k, r);
        o_exp_value (re, comp, r[axis] * re->dx, time, &er, &ei);
        const int64_t q = o_idx (re, i, j, k);
        /* after the step's shift the newest values are layer 1 (CurrentAfterShift) */
        const double fr = re->f[comp][1][q], fi = im->f[comp][1][q];
        const double me = sqrt (er * er + ei * ei), mf = sqrt (fr * fr + fi * fi);
        sum += (me - mf) * (me - mf);
        if (mx < me) mx = me;
      }
  const double norm = sqrt (sum / (double) re->ncell);
  out[0] = norm;
  out[1] = norm * 100.0 / mx;
  return 0;
}

int oracle_get_diff_norm (oracle *o, int32_t comp, double out[4])
{
  if (!o || !out || comp < 0 || comp > 5 || !o->norm_valid[comp]) return -1;
  for (int q = 0; q < 4; ++q) out[q] = o->norm[comp][q];
  return 0;
}

static void o_rotate (double **layers, int n)
{
  double *oldest = layers[n - 1];
  for (int l = n - 1; l >= 1; --l) layers[l] = layers[l - 1];
  layers[0] = oldest;
}

/* performPlaneWaveESteps (Source/Scheme/InternalScheme.inc.h:3276-3318). */
static void o_inc_e (oracle *o, int32_t t)
{
  double mod = o->dt / (o->rel * o_eps0 () * o->dx);
  double *e0 = o->einc[0], *e1 = o->einc[1], *h1 = o->hinc[1];
  for (int64_t i = 1; i < o->inc_len; ++i)
  {
    e0[i] = e1[i] + (h1[i - 1] - h1[i]) * mod;
  }
  double arg = o->dt * t * 2 * O_PI * o->freq;
  /* the complex build sets EInc[0] = (sin (arg), cos (arg)) (InternalScheme.inc.h:3310-3316) */
  e0[0] = o->src_kind == 2 ? cos (arg + o->src_phase) : sin (arg + o->src_phase);
}

/* performPlaneWaveHSteps (Source/Scheme/InternalScheme.inc.h:3320-3357); the last cell is skipped. */
static void o_inc_h (oracle *o)
{
  double mod = o->dt / (o->rel * o_mu0 () * o->dx);
  double *h0 = o->hinc[0], *h1 = o->hinc[1], *e1 = o->einc[1];
  for (int64_t i = 0; i < o->inc_len - 1; ++i)
  {
    h0[i] = h1[i] + (e1[i] - e1[i + 1]) * mod;
  }
}

/* Reads layer 1 of a present grid at an offset position; absent grids read as 0. */
static double o_read1 (const oracle *o, int comp, int64_t i, int64_t j, int64_t k)
{
  if (!o->present[comp]) return 0;
  return o->f[comp][1][o_idx (o, i, j, k)];
}

/* calculateTFSF (Source/Scheme/InternalScheme.inc.h:1386-1835), E_CENTERED. */
static void o_tfsf (const oracle *o, int comp, int64_t i, int64_t j, int64_t k,
                    double *v11, double *v12, double *v21, double *v22)
{
  const comp_desc *c = &CD[comp];
  double real[3];
  o_real_coord (o, comp, i, j, k, real);
  int u11 = o_tfsf_pred (o, comp, c->dir11, real);
  int u12 = o_tfsf_pred (o, comp, c->dir12, real);
  int u21 = o_tfsf_pred (o, comp, c->dir21, real);
  int u22 = o_tfsf_pred (o, comp, c->dir22, real);
  int revert = !c->is_h;   /* E_CENTERED: E components revert, H do not */
  const double *inc = c->is_h ? o->einc[1] : o->hinc[1];
  double ddiff = c->is_h ? 0.0 : 0.5;
  double diff1 = 0, diff2 = 0;
  const int *p1 = NULL, *p2 = NULL;
  if (revert)
  {
    if (u11) p1 = c->d12; else if (u12) p1 = c->d11;
    if (u21) p2 = c->d22; else if (u22) p2 = c->d21;
  }
  else
  {
    if (u11) p1 = c->d11; else if (u12) p1 = c->d12;
    if (u21) p2 = c->d21; else if (u22) p2 = c->d22;
  }
  if (u11 || u12)
  {
    double r[3];
    o_real_coord (o, c->opp1, i + o->act[0] * p1[0], j + o->act[1] * p1[1], k + o->act[2] * p1[2], r);
    diff1 = o_project (o, c->opp1, o_incident (o, r, ddiff, inc));
    if (c->is_h) diff1 *= -1.0;
  }
  if (u21 || u22)
  {
    double r[3];
    o_real_coord (o, c->opp2, i + o->act[0] * p2[0], j + o->act[1] * p2[1], k + o->act[2] * p2[2], r);
    diff2 = o_project (o, c->opp2, o_incident (o, r, ddiff, inc));
    if (c->is_h) diff2 *= -1.0;
  }
  if (revert)
  {
    if (u11) *v12 -= diff1; else if (u12) *v11 -= diff1;
    if (u21) *v22 -= diff2; else if (u22) *v21 -= diff2;
  }
  else
  {
    if (u11) *v11 -= diff1; else if (u12) *v12 -= diff1;
    if (u21) *v21 -= diff2; else if (u22) *v22 -= diff2;
  }
}

static double o_sig (const oracle *o, int axis, int comp, int64_t i, int64_t j, int64_t k)
{
  if (!o->act[axis]) return 0;   /* the sigma of the missing axis of a 2D scheme reads as 0 */
  return o_get_material (o, comp, i, j, k, o->sigma[axis]);
}

/* One component's update for step t: Scheme::calculateFieldStep (Scheme.inc.h:876-1344). */
static void o_field_step (oracle *o, int comp, int32_t t)
{
  const comp_desc *c = &CD[comp];
  const double eps0 = o_eps0 ();
  const double modifier = c->is_h ? o_mu0 () : eps0;
  const double *mgrid = c->is_h ? o->mat[YEE_MAT_MU] : o->mat[YEE_MAT_EPS];
  const double *omg = c->is_h ? o->mat[YEE_MAT_OMEGAPM] : o->mat[YEE_MAT_OMEGAPE];
  const double *gam = c->is_h ? o->mat[YEE_MAT_GAMMAM] : o->mat[YEE_MAT_GAMMAE];
  const double dt = o->dt, dx = o->dx;
  int o1 = c->opp1, o2 = c->opp2;
  const int dx_ = o->act[0], dy_ = o->act[1], dz = o->act[2];   /* offsets along a missing axis drop */

  if (t > 0)
  {
    double *out = o->use_pml ? o->d[comp][0] : o->f[comp][0];
    const double *prev = o->use_pml ? o->d[comp][1] : o->f[comp][1];
    /* sweep 1: calculateFieldStepIteration (InternalScheme.inc.h:2351-2417) */
    for (int64_t i = o->lo[0]; i < o->hi[0]; ++i)
      for (int64_t j = o->lo[1]; j < o->hi[1]; ++j)
        for (int64_t k = o->lo[2]; k < o->hi[2]; ++k)
        {
          int64_t q = o_idx (o, i, j, k);
          double val = prev[q];
          double ca, cb;
          /* computeCaCb (InternalScheme.inc.h:2297-2349) */
          if (o->use_pml)
          {
            double material = o_sig (o, c->sig_curl, comp, i, j, k);
            double dd = (2 * eps0 * 1.0 + material * dt);
            ca = (2 * eps0 * 1.0 - material * dt) / dd;
            cb = (2 * eps0 * dt / dx) / dd;
          }
          else
          {
            double material = o_get_material (o, comp, i, j, k, mgrid);
            ca = 1.0;
            cb = dt / (material * modifier * dx);
          }
          double v11 = o_read1 (o, o1, i + dx_ * c->d11[0], j + dy_ * c->d11[1], k + dz * c->d11[2]);
          double v12 = o_read1 (o, o1, i + dx_ * c->d12[0], j + dy_ * c->d12[1], k + dz * c->d12[2]);
          double v21 = o_read1 (o, o2, i + dx_ * c->d21[0], j + dy_ * c->d21[1], k + dz * c->d21[2]);
          double v22 = o_read1 (o, o2, i + dx_ * c->d22[0], j + dy_ * c->d22[1], k + dz * c->d22[2]);
          if (o->use_tfsf)
          {
            o_tfsf (o, comp, i, j, k, &v11, &v12, &v21, &v22);
          }
          /* calcField (InternalScheme.inc.h:765-772); the right side J / M of an analytic test
           * problem at the component's point, at time step t (E) or t + 1/2 (H)
           * (Scheme.inc.h:1062-1098, InternalScheme.inc.h:2410-2413), else 0 */
          double rhs = 0.0;
          if (o->fn_rhs && o_rhs_has (o->fn_rhs, comp))
          {
            double r[3];
            o_real_coord (o, comp, i, j, k, r);
            const double timestep = c->is_h ? (double) t + 0.5 : (double) t;
            rhs = o_rhs_value (o->fn_rhs, comp, r[0] * dx, r[1] * dx, r[2] * dx, timestep * dt);
          }
          double tmp = v12 - v11 - v22 + v21 - rhs * dx;
          out[q] = val * ca + tmp * cb;
        }
    /*
     * Current source (Scheme::calculateFieldStep, Scheme.inc.h:1171-1260) after the curl sweep:
     * current = sin (dt * timestep * 2 pi f) * mu0 / dt / 2 with timestep t (E) or t + 1/2 (H),
     * subtracted as current * dx * Cb at the source cell, Cb from computeCaCb there
     * (calculateFieldStepIterationCurrent, InternalScheme.inc.h:2419-2450).
     */
    if (o->cfg.current_source & (1u << comp))
    {
      int64_t p[3];
      int inside = 1;
      for (int a = 0; a < 3; ++a)
      {
        p[a] = o->act[a] ? o->cfg.current_source_pos[a] : 0;
        inside = inside && p[a] >= 0 && p[a] < o->n[a];
      }
      if (inside)
      {
        const double timestep = c->is_h ? (double) t + 0.5 : (double) t;
        /* the complex build's current is (cos, -sin) of dt (timestep - 1) 2 pi f
         * (Scheme.inc.h:1230-1235) */
        double current = o->src_kind == 0 ? sin (dt * timestep * 2 * O_PI * o->freq)
                         : (o->src_kind == 1 ? cos (dt * (timestep - 1) * 2 * O_PI * o->freq)
                                             : -sin (dt * (timestep - 1) * 2 * O_PI * o->freq));
        current = current * o_mu0 () / dt / 2.0;
        double cb;
        if (o->use_pml)
        {
          double material = o_sig (o, c->sig_curl, comp, p[0], p[1], p[2]);
          double dd = (2 * eps0 * 1.0 + material * dt);
          cb = (2 * eps0 * dt / dx) / dd;
        }
        else
        {
          double material = o_get_material (o, comp, p[0], p[1], p[2], mgrid);
          cb = dt / (material * modifier * dx);
        }
        out[o_idx (o, p[0], p[1], p[2])] -= current * dx * cb;
      }
    }
    if (o->use_pml)
    {
      if (o->use_meta)
      {
        /* sweep 2: calculateFieldStepIterationPMLMetamaterials (InternalScheme.inc.h:2452-2549) */
        double *d0 = o->d[comp][0], *dp = o->d[comp][1], *dpp = o->d[comp][2];
        double *m0 = o->d1[comp][0], *mp = o->d1[comp][1], *mpp = o->d1[comp][2];
        for (int64_t i = o->lo[0]; i < o->hi[0]; ++i)
          for (int64_t j = o->lo[1]; j < o->hi[1]; ++j)
            for (int64_t k = o->lo[2]; k < o->hi[2]; ++k)
            {
              int64_t q = o_idx (o, i, j, k);
              double m1, m2;
              double material = o_get_meta (o, comp, i, j, k, mgrid, omg, gam, &m1, &m2);
              double M = modifier;
              double A = 4 * M * material + 2 * dt * M * material * m2 + M * ((dt * m1) * (dt * m1));
              double b0 = (4 + 2 * dt * m2) / A;
              double b1 = -8 / A;
              double b2 = (4 - 2 * dt * m2) / A;
              double a1 = (2 * M * ((dt * m1) * (dt * m1)) - 8 * M * material) / A;
              double a2 = (4 * M * material - 2 * dt * M * material * m2 + M * ((dt * m1) * (dt * m1))) / A;
              /* calcFieldDrude (InternalScheme.inc.h:774-780) */
              m0[q] = d0[q] * b0 + dp[q] * b1 + dpp[q] * b2 - mp[q] * a1 - mpp[q] * a2;
            }
      }
      /* sweep 3: calculateFieldStepIterationPML (InternalScheme.inc.h:2551-2640) */
      const double *cur = o->use_meta ? o->d1[comp][0] : o->d[comp][0];
      const double *prv = o->use_meta ? o->d1[comp][1] : o->d[comp][1];
      double *e0 = o->f[comp][0];
      const double *e1 = o->f[comp][1];
      for (int64_t i = o->lo[0]; i < o->hi[0]; ++i)
        for (int64_t j = o->lo[1]; j < o->hi[1]; ++j)
          for (int64_t k = o->lo[2]; k < o->hi[2]; ++k)
          {
            int64_t q = o_idx (o, i, j, k);
            double material1 = o_get_material (o, comp, i, j, k, mgrid);
            double material4 = o_sig (o, c->sig4, comp, i, j, k);
            double material5 = o_sig (o, c->sig5, comp, i, j, k);
            double mod = material1 * modifier;
            if (o->use_meta) mod = 1;
            double dd = (2 * eps0 * 1.0 + material5 * dt);
            double ca = (2 * eps0 * 1.0 - material5 * dt) / dd;
            double cb = ((2 * eps0 * 1.0 + material4 * dt) / (mod)) / dd;
            double cc = ((2 * eps0 * 1.0 - material4 * dt) / (mod)) / dd;
            /* calcFieldFromDOrB (InternalScheme.inc.h:782-787) */
            e0[q] = e1[q] * ca + cur[q] * cb - prv[q] * cc;
          }
    }
  }

  /* border values of an analytic test problem on every cell with an index 0 or N-1, also at
   * t == 0 (calculateFieldStepIterationBorder, InternalScheme.inc.h:2646-2705; doSkipBorderFunc
   * :539-555), at time step t + 1/2 (E) or t + 1 (H) */
  if (o->fn_border && o_fn_has (o->fn_border, comp))
  {
    const double time = (c->is_h ? (double) t + 1.0 : (double) t + 0.5) * dt;
    for (int64_t i = 0; i < o->n[0]; ++i)
      for (int64_t j = 0; j < o->n[1]; ++j)
        for (int64_t k = 0; k < o->n[2]; ++k)
        {
          const int64_t p[3] = { i, j, k };
          int edge = 0;
          for (int a = 0; a < 3; ++a) edge = edge || (o->act[a] && (p[a] == 0 || p[a] == o->n[a] - 1));
          if (!edge) continue;
          double r[3];
          o_real_coord (o, comp, i, j, k, r);
          o->f[comp][0][o_idx (o, i, j, k)] = o_fn_value (o->fn_border, comp, r[0] * dx, r[1] * dx, r[2] * dx, time);
        }
  }
  /* difference norm against the exact solution over the norm box, in the iteration order of the
   * reference (calculateFieldStepIterationExact, InternalScheme.inc.h:2707-2816; Scheme.inc.h:
   * 1377-1550): sqrt (sum of squares / cells of the grid) and its percentage of max |exact| */
  if (o->fn_exact && (o_fn_has (o->fn_exact, comp) || o_exp_has (o, comp)))
  {
    const double time = (c->is_h ? (double) t + 1.0 : (double) t + 0.5) * dt;
    double sum = 0.0, mx = 0.0;
    const double *f0 = o->f[comp][0];
    const int exp = o->fn_exact >= YEE_FN_EXP1;
    for (int64_t i = o->norm_lo[0]; i < o->norm_hi[0]; ++i)
      for (int64_t j = o->norm_lo[1]; j < o->norm_hi[1]; ++j)
        for (int64_t k = o->norm_lo[2]; k < o->norm_hi[2]; ++k)
        {
          double r[3];
          o_real_coord (o, comp, i, j, k, r);
          double exact;
          if (exp)
          {
            /* the part this oracle carries: the real part, or the imaginary part of the complex
             * build's imaginary-part run */
            double er, ei;
            o_exp_value (o, comp, r[kExpAxis[o->exp_pair]] * dx, time, &er, &ei);
            exact = o->src_kind == 2 ? ei : er;
          }
          else
            exact = o_fn_value (o->fn_exact, comp, r[0] * dx, r[1] * dx, r[2] * dx, time);
          const double diff = exact - f0[o_idx (o, i, j, k)];
          sum += diff * diff;
          const double ea = fabs (exact);
          if (mx < ea) mx = ea;
        }
    const double norm = sqrt (sum / (double) o->ncell);
    o->norm[comp][0] = norm;
    o->norm[comp][1] = norm * 100.0 / mx;
    o->norm[comp][2] = sum;
    o->norm[comp][3] = mx;
    o->norm_valid[comp] = 1;
  }

  /* performPointSourceCalc (InternalScheme.inc.h:2818-2895), also at t == 0 */
  if (o->cfg.point_source & (1u << comp))
  {
    int64_t p[3];
    for (int a = 0; a < 3; ++a) p[a] = o->act[a] ? o->cfg.point_source_pos[a] : 0;
    if (p[0] >= 0 && p[0] < o->n[0] && p[1] >= 0 && p[1] < o->n[1] && p[2] >= 0 && p[2] < o->n[2])
    {
      double ts = c->is_h ? t + 1.0 : t + 0.5;
      /* the complex build's point source is (sin, cos) of the same argument (:2888-2893) */
      double v = o->src_kind == 2 ? cos (o->dt * ts * 2 * O_PI * o->freq + o->src_phase)
                                  : sin (o->dt * ts * 2 * O_PI * o->freq + o->src_phase);
      /* Gaussian envelope (no reference counterpart; oracle_set_source_gaussian) */
      if (o->src_sigma > 0) v *= exp (-((ts - o->src_t0) / o->src_sigma) * ((ts - o->src_t0) / o->src_sigma));
      o->f[comp][0][o_idx (o, p[0], p[1], p[2])] = v;
    }
  }
}

static void o_shift (oracle *o, int comp)
{
  o_rotate (o->f[comp], o->layers);
  if (o->use_pml) o_rotate (o->d[comp], o->layers);
  if (o->use_meta) o_rotate (o->d1[comp], o->layers);
}

/* Scheme::performNStepsForBlock (Source/Scheme/Scheme.inc.h:28-473). */
int oracle_step (oracle *o, int32_t n)
{
  if (!o || !o->prepared || n < 0) return -1;
  for (int32_t s = 0; s < n; ++s)
  {
    int32_t t = o->t;
    if (o->use_tfsf)
    {
      o_inc_e (o, t);
      o_rotate (o->einc, 2);
    }
    for (int c = 0; c < 3; ++c) if (o->present[c]) o_field_step (o, c, t);
    for (int c = 0; c < 3; ++c) if (o->present[c]) o_shift (o, c);
    if (o->use_tfsf)
    {
      o_inc_h (o);
      o_rotate (o->hinc, 2);
    }
    for (int c = 3; c < 6; ++c) if (o->present[c]) o_field_step (o, c, t);
    for (int c = 3; c < 6; ++c) if (o->present[c]) o_shift (o, c);
    o->t++;
  }
  return 0;
}

int oracle_get_field (oracle *o, int32_t comp, double *out)
{
  if (!o || !out || comp < 0 || comp > 17) return -1;
  int base = comp % 6;
  if (!o->present[base]) return -1;
  double *src;
  if (comp < 6) src = o->f[base][1];
  else if (comp < 12) src = o->use_pml ? o->d[base][1] : NULL;
  else src = o->use_meta ? o->d1[base][1] : NULL;
  if (!src) return -1;
  memcpy (out, src, sizeof (double) * o->ncell);
  return 0;
}

int oracle_get_material (oracle *o, int32_t which, double *out)
{
  if (!o || !out) return -1;
  if (which >= 0 && which < 6)
  {
    memcpy (out, o->mat[which], sizeof (double) * o->mcell);
    return 0;
  }
  if (which >= 6 && which < 9 && o->sigma[which - 6])
  {
    memcpy (out, o->sigma[which - 6], sizeof (double) * o->mcell);
    return 0;
  }
  return -1;
}

int oracle_tfsf_mask (oracle *o, int32_t comp, uint8_t *out)
{
  if (!o || !out || comp < 0 || comp > 5 || !o->present[comp]) return -1;
  for (int64_t i = 0; i < o->n[0]; ++i)
    for (int64_t j = 0; j < o->n[1]; ++j)
      for (int64_t k = 0; k < o->n[2]; ++k)
      {
        double real[3];
        o_real_coord (o, comp, i, j, k, real);
        uint8_t m = 0;
        for (int d = 0; d < 6; ++d)
        {
          if (o_tfsf_pred (o, comp, d, real)) m |= (uint8_t) (1u << d);
        }
        out[o_idx (o, i, j, k)] = m;
      }
  return 0;
}

int oracle_get_incident (oracle *o, double *einc, double *hinc, int64_t len)
{
  if (!o || !o->use_tfsf || len > o->inc_len) return -1;
  if (einc) memcpy (einc, o->einc[1], sizeof (double) * len);
  if (hinc) memcpy (hinc, o->hinc[1], sizeof (double) * len);
  return 0;
}

/*
 * Scheme::makeGridScattered (Source/Scheme/Scheme.inc.h:4113-4228): layer 1 of the component
 * minus the projected incident wave at its real coordinate (E from EInc with diff 0, H from
 * HInc with diff 0.5, InternalScheme.inc.h:525-536), except where doSkipMakeScattered
 * (Source/Scheme/SchemeHelper.h:897-911) skips the point: a real coordinate below the left
 * or above the right TF/SF border on some axis keeps the total field.
 */
int oracle_get_scattered (oracle *o, int32_t comp, double *out)
{
  if (!o || !out || comp < 0 || comp > 5 || !o->present[comp] || !o->use_tfsf) return -1;
  const int is_h = CD[comp].is_h;
  const double *inc = is_h ? o->hinc[1] : o->einc[1];
  const double ddiff = is_h ? 0.5 : 0.0;
  for (int64_t i = 0; i < o->n[0]; ++i)
    for (int64_t j = 0; j < o->n[1]; ++j)
      for (int64_t k = 0; k < o->n[2]; ++k)
      {
        double real[3];
        o_real_coord (o, comp, i, j, k, real);
        int skip = 0;
        for (int a = 0; a < 3; ++a)
        {
          if (o->act[a] && (real[a] < o->bl[a] || real[a] > o->br[a])) skip = 1;
        }
        const int64_t q = o_idx (o, i, j, k);
        const double v = o->f[comp][1][q];
        out[q] = skip ? v : v - o_project (o, comp, o_incident (o, real, ddiff, inc));
      }
  return 0;
}

/*
 * Test hook: shifts the phase of every source.  With phase pi/2 the run is the imaginary
 * part of the reference's complex-field build, whose sources are sin + i cos
 * (InternalScheme.inc.h:2886-2888, 3311-3312); together with the phase-0 run it gives
 * the field magnitudes the reference's unit test checks (unit-test-internalscheme.cpp:941-988).
 */
int oracle_set_source_phase (oracle *o, double phase)
{
  if (!o) return -1;
  o->src_phase = phase;
  return 0;
}

/*
 * Complex field values: the reference's complex build (-DCOMPLEX_FIELD_VALUES) multiplies fields
 * by real coefficients only, so its real and imaginary parts are two real runs that differ in
 * their sources.  kind 1 gives the real part, kind 2 the imaginary part, kind 0 the real build.
 */
int oracle_set_source_kind (oracle *o, int kind)
{
  if (!o || kind < 0 || kind > 2) return -1;
  o->src_kind = kind;
  return 0;
}

/* Gaussian-modulated hard point source: the reference's sine times exp(-((ts - t0) / sigma)^2),
 * ts = t + 1/2 (E) or t + 1 (H) in time steps.  No reference counterpart (the batched TMz
 * sweep's source); pinned by the closed form in tests/test_oracle_known_answers.py. */
int oracle_set_source_gaussian (oracle *o, double t0, double sigma)
{
  if (!o) return -1;
  o->src_t0 = t0;
  o->src_sigma = sigma;
  return 0;
}

int64_t oracle_incident_len (oracle *o)
{
  return o ? o->inc_len : 0;
}

int oracle_get_scalars (oracle *o, double out[3])
{
  if (!o) return -1;
  out[0] = o->dt;
  out[1] = o->freq;
  out[2] = o->rel;
  return 0;
}

/* ------------------------------------------------------------------------------------------
 * FP64 CPU adjoint (dJ/d eps over a design box) of the reference-form update above.
 *
 * No reference counterpart: the reference has no differentiation.  The forward map of one
 * step is linear in the state (E, D, H, B) plus field-independent TF/SF and source terms, so
 * the adjoint step is its transpose, applied in reverse (H first, then E):
 *   H:  mB = aB + kp aH;  w = cb_curl mB;  aB <- cac mB - km aH;  aH <- ca5 aH;
 *       every E operand of the H curl receives +-w (the transpose of the curl gather);
 *   E:  the same with (E, D) and the H operands.
 * Without PML the cell update is V' = V + cb tmp (w = cb aV).  The update of an E cell
 * depends on eps only through the averaged material m = (eps(v1) + eps(v2)) / 2 inside kp, km
 * (or cb), each proportional to 1/m, so dE'/dm = -(E' - ca5 E) / m, and each of the two
 * eps voxels receives half of it.
 * A hard point source replaces the new value of its cell, so no adjoint flows back from it.
 * J = sum over the last `window` steps of sum over the monitor rectangle of Ez^2.
 * Pinned by central finite differences of the forward oracle (tests).
 * ------------------------------------------------------------------------------------------ */

typedef struct
{
  double cac, cbc, ca5, kp, km, cb;   /* two-stage (PML) or direct (cb) coefficients */
  double m;                           /* averaged material */
} o_coef;

static void o_cell_coef (const oracle *o, int comp, int64_t i, int64_t j, int64_t k, o_coef *r)
{
  const comp_desc *c = &CD[comp];
  const double eps0 = o_eps0 ();
  const double modifier = c->is_h ? o_mu0 () : eps0;
  const double *mgrid = c->is_h ? o->mat[YEE_MAT_MU] : o->mat[YEE_MAT_EPS];
  const double dt = o->dt, dx = o->dx;
  r->m = o_get_material (o, comp, i, j, k, mgrid);
  if (o->use_pml)
  {
    double s = o_sig (o, c->sig_curl, comp, i, j, k);
    double dd = (2 * eps0 * 1.0 + s * dt);
    r->cac = (2 * eps0 * 1.0 - s * dt) / dd;
    r->cbc = (2 * eps0 * dt / dx) / dd;
    double s4 = o_sig (o, c->sig4, comp, i, j, k);
    double s5 = o_sig (o, c->sig5, comp, i, j, k);
    double mod = r->m * modifier;
    double dd5 = (2 * eps0 * 1.0 + s5 * dt);
    r->ca5 = (2 * eps0 * 1.0 - s5 * dt) / dd5;
    r->kp = ((2 * eps0 * 1.0 + s4 * dt) / (mod)) / dd5;
    r->km = ((2 * eps0 * 1.0 - s4 * dt) / (mod)) / dd5;
    r->cb = 0;
  }
  else
  {
    r->cb = dt / (r->m * modifier * dx);
    r->cac = r->cbc = r->kp = r->km = 0;
    r->ca5 = 1.0;
  }
}

/* Adds s * w to the four curl operands of component comp at (i, j, k) in the adjoint
 * arrays a[] (tmp = v12 - v11 - v22 + v21). */
static void o_scatter (const oracle *o, int comp, int64_t i, int64_t j, int64_t k, double w, double *a[6])
{
  const comp_desc *c = &CD[comp];
  const int dz = o->dim == 3;
  if (o->present[c->opp1])
  {
    a[c->opp1][o_idx (o, i + c->d12[0], j + c->d12[1], k + dz * c->d12[2])] += w;
    a[c->opp1][o_idx (o, i + c->d11[0], j + c->d11[1], k + dz * c->d11[2])] -= w;
  }
  if (o->present[c->opp2])
  {
    a[c->opp2][o_idx (o, i + c->d22[0], j + c->d22[1], k + dz * c->d22[2])] -= w;
    a[c->opp2][o_idx (o, i + c->d21[0], j + c->d21[1], k + dz * c->d21[2])] += w;
  }
}

int oracle_adjoint (oracle *o, const oracle_adj_spec *sp, int32_t T, double *J_out, double *grad)
{
  if (!o || !sp || !J_out || !grad || !o->prepared || o->t != 0 || o->dim != 3 || o->use_meta || o->fine || T < 2)
    return -1;
  if (sp->window < 1 || sp->window > T - 1) return -1;
  const int64_t nc = o->ncell;
  const int nstate = o->use_pml ? 12 : 6;
  /* trajectory: state after steps 0 .. T-1 */
  double *traj = (double *) malloc (sizeof (double) * (size_t) nstate * nc * T);
  if (!traj) return -2;
  double J = 0;
  for (int32_t t = 0; t < T; ++t)
  {
    oracle_step (o, 1);
    for (int c = 0; c < 6; ++c)
    {
      memcpy (traj + ((size_t) t * nstate + c) * nc, o->f[c][1], sizeof (double) * nc);
      if (o->use_pml) memcpy (traj + ((size_t) t * nstate + 6 + c) * nc, o->d[c][1], sizeof (double) * nc);
    }
    if (t >= T - sp->window)
    {
      for (int64_t x = sp->mlo[0]; x < sp->mhi[0]; ++x)
        for (int64_t y = sp->mlo[1]; y < sp->mhi[1]; ++y)
        {
          double v = o->f[YEE_EZ][1][o_idx (o, x, y, sp->mz)];
          J += v * v;
        }
    }
  }
  *J_out = J;

  const int64_t dn[3] = { sp->dhi[0] - sp->dlo[0], sp->dhi[1] - sp->dlo[1], sp->dhi[2] - sp->dlo[2] };
  memset (grad, 0, sizeof (double) * (size_t) (dn[0] * dn[1] * dn[2]));
  double *aF[6], *aD[6], *w = (double *) calloc ((size_t) nc, sizeof (double));
  for (int c = 0; c < 6; ++c)
  {
    aF[c] = (double *) calloc ((size_t) nc, sizeof (double));
    aD[c] = (double *) calloc ((size_t) nc, sizeof (double));
  }
  for (int32_t t = T - 1; t >= 1; --t)
  {
    const double *Scur = traj + (size_t) t * nstate * nc, *Sprev = traj + (size_t) (t - 1) * nstate * nc;
    if (t >= T - sp->window)
      for (int64_t x = sp->mlo[0]; x < sp->mhi[0]; ++x)
        for (int64_t y = sp->mlo[1]; y < sp->mhi[1]; ++y)
        {
          int64_t q = o_idx (o, x, y, sp->mz);
          aF[YEE_EZ][q] += 2 * Scur[(size_t) YEE_EZ * nc + q];
        }
    for (int pass = 0; pass < 2; ++pass)
    {
      const int c0 = pass == 0 ? 3 : 0;   /* H transpose first, then E */
      for (int comp = c0; comp < c0 + 3; ++comp)
      {
        memset (w, 0, sizeof (double) * (size_t) nc);
        for (int64_t i = 1; i < o->n[0] - 1; ++i)
          for (int64_t j = 1; j < o->n[1] - 1; ++j)
            for (int64_t k = 1; k < o->n[2] - 1; ++k)
            {
              int64_t q = o_idx (o, i, j, k);
              o_coef cf;
              o_cell_coef (o, comp, i, j, k, &cf);
              double a = aF[comp][q];
              /* a hard source overwrites the new value: nothing flows back from it */
              if ((o->cfg.point_source & (1u << comp)) && i == o->cfg.point_source_pos[0]
                  && j == o->cfg.point_source_pos[1] && k == o->cfg.point_source_pos[2])
                a = 0;
              if (comp < 3)
              {
                /* dE'/dm = -(E' - ca5 E) / m, half to each eps voxel of the average */
                double dEdm = -(Scur[(size_t) comp * nc + q] - cf.ca5 * Sprev[(size_t) comp * nc + q]) / cf.m;
                int64_t cells[4][3];
                o_material_cells (o, comp, i, j, k, cells);
                for (int h = 0; h < 2; ++h)
                {
                  int64_t *v = cells[h];
                  if (v[0] < sp->dlo[0] || v[0] >= sp->dhi[0] || v[1] < sp->dlo[1] || v[1] >= sp->dhi[1]
                      || v[2] < sp->dlo[2] || v[2] >= sp->dhi[2])
                    continue;
                  grad[((v[0] - sp->dlo[0]) * dn[1] + (v[1] - sp->dlo[1])) * dn[2] + (v[2] - sp->dlo[2])] += a * dEdm * 0.5;
                }
              }
              if (o->use_pml)
              {
                double mD = aD[comp][q] + cf.kp * a;
                w[q] = cf.cbc * mD;
                aD[comp][q] = cf.cac * mD - cf.km * a;
                aF[comp][q] = cf.ca5 * a;
              }
              else
              {
                w[q] = cf.cb * a;
                aF[comp][q] = a;
              }
            }
        for (int64_t i = 1; i < o->n[0] - 1; ++i)
          for (int64_t j = 1; j < o->n[1] - 1; ++j)
            for (int64_t k = 1; k < o->n[2] - 1; ++k)
            {
              int64_t q = o_idx (o, i, j, k);
              if (w[q] != 0) o_scatter (o, comp, i, j, k, w[q], aF);
            }
      }
    }
  }
  for (int c = 0; c < 6; ++c)
  {
    free (aF[c]);
    free (aD[c]);
  }
  free (w);
  free (traj);
  return 0;
}
