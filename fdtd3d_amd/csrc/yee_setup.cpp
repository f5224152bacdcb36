// Host-side FP64 set-up of the Yee engine (see yee_setup.h).
#include "yee_setup.h"

#include <algorithm>
#include <cmath>

namespace yee {

// E_CENTERED bindings (InternalScheme.inc.h:1910-2295):
//   Ex: Hz, Hy, sigma Y / X / Z    Hx: Ey, Ez, sigma Y / X / Z
//   Ey: Hx, Hz, sigma Z / Y / X    Hy: Ez, Ex, sigma Z / Y / X
//   Ez: Hy, Hx, sigma X / Z / Y    Hz: Ex, Ey, sigma X / Z / Y
const CompDesc kComp[6] = {
  { 0, 0, YEE_HZ, YEE_HY, 1, 0, 2 },
  { 1, 0, YEE_HX, YEE_HZ, 2, 1, 0 },
  { 2, 0, YEE_HY, YEE_HX, 0, 2, 1 },
  { 0, 1, YEE_EY, YEE_EZ, 1, 0, 2 },
  { 1, 1, YEE_EZ, YEE_EX, 2, 1, 0 },
  { 2, 1, YEE_EX, YEE_EY, 0, 2, 1 },
};

// Approximation::approximateWaveNumber / approximateWaveNumberGeneral (Approximation.cpp:259-339):
// closed forms for propagation along an axis or at 45 degrees, Newton iteration otherwise.
double wave_number (double delta, double lambda, double courant, double nl, double a1, double a2)
{
  const double pi = kPi;
  auto axial = [pi] (double a) { return a == 0 || a == pi / 2 || a == pi || a == 3 * pi / 2; };
  auto diagonal = [pi] (double a) { return a == pi / 4 || a == 3 * pi / 4 || a == 5 * pi / 4 || a == 7 * pi / 4; };
  if (axial (a1) && axial (a2))
  {
    return 2.0 / delta * std::asin (std::sin (pi * courant / nl) / courant);
  }
  if ((axial (a1) && diagonal (a2)) || (diagonal (a1) && axial (a2)))
  {
    return 2.0 * std::sqrt (2.0) / delta * std::asin (std::sin (pi * courant / nl) / (courant * std::sqrt (2.0)));
  }
  const double tol = 0.0000001;
  double k = 2 * pi;
  double k_prev = k + 2 * std::sqrt (tol);
  const double nd = delta / lambda;
  const double A = nd * std::sin (a1) * std::cos (a2) / 2;
  const double B = nd * std::sin (a1) * std::sin (a2) / 2;
  const double C = nd * std::cos (a1) / 2;
  const double s = std::sin (pi * courant / nl);
  const double D = (s * s) / (courant * courant);
  while ((k_prev - k) * (k_prev - k) >= tol)
  {
    k_prev = k;
    const double sa = std::sin (A * k), sb = std::sin (B * k), sc = std::sin (C * k);
    const double f = sa * sa + sb * sb + sc * sc - D;
    const double df = A * std::sin (2 * A * k) + B * std::sin (2 * B * k) + C * std::sin (2 * C * k);
    k -= f / df;
  }
  return k / lambda;
}

Scalars compute_scalars (const yee_config &cfg)
{
  Scalars s {};
  const bool dim3 = cfg.scheme == YEE_SCHEME_DIM3;
  // InternalScheme::initScheme (InternalScheme.inc.h:3359-3394)
  s.freq = kC / cfg.wavelength;
  s.dt = cfg.dx * cfg.courant / kC;
  s.a1 = cfg.angle_deg[0] * kPi / 180.0;
  s.a2 = cfg.angle_deg[1] * kPi / 180.0;
  s.a3 = cfg.angle_deg[2] * kPi / 180.0;
  const double nl = cfg.wavelength / cfg.dx;
  auto phase_velocity = [&] (double a1, double a2) {
    return kC * 2 * kPi / cfg.wavelength / wave_number (cfg.dx, cfg.wavelength, cfg.courant, nl, a1, a2);
  };
  s.rel = phase_velocity (kPi / 2, 0) / phase_velocity (s.a1, s.a2);
  s.S1 = std::sin (s.a1);
  s.C1 = std::cos (s.a1);
  s.S2 = std::sin (s.a2);
  s.C2 = std::cos (s.a2);
  const double a1 = s.a1, a2 = s.a2, a3 = s.a3;
  if (dim3)
  {
    // YeeGridLayout.h:2279-2356
    s.proj[YEE_EX] = std::cos (a3) * std::sin (a2) - std::sin (a3) * std::cos (a1) * std::cos (a2);
    s.proj[YEE_EY] = -std::cos (a3) * std::cos (a2) - std::sin (a3) * std::cos (a1) * std::sin (a2);
    s.proj[YEE_EZ] = std::sin (a3) * std::sin (a1);
    s.proj[YEE_HX] = std::sin (a3) * std::sin (a2) + std::cos (a3) * std::cos (a1) * std::cos (a2);
    s.proj[YEE_HY] = -std::sin (a3) * std::cos (a2) + std::cos (a3) * std::cos (a1) * std::sin (a2);
    // -valH * f: negating the factor instead gives the same bits
    s.proj[YEE_HZ] = -(std::cos (a3) * std::sin (a1));
    // YeeGridLayout.cpp:2636-2656
    s.zero_inc[0] = cfg.tfsf_left[0] - 2.5 * std::sin (a1) * std::cos (a2);
    s.zero_inc[1] = cfg.tfsf_left[1] - 2.5 * std::sin (a1) * std::sin (a2);
    s.zero_inc[2] = cfg.tfsf_left[2] - 2.5 * std::cos (a1);
  }
  else if (line_axis (cfg) >= 0)
  {
    // the 1D schemes, run on the 3D kernels with the two other axes of extent 1: their incident
    // wave d is the coordinate along the line (InternalScheme.cpp:110-200), the 3D expression
    // x S1 C2 + y S1 S2 + z C1 with factor 1 on the line's axis and 0 on the others (exact)
    const int line = line_axis (cfg);
    s.S1 = line == 2 ? 0.0 : 1.0;
    s.C1 = line == 2 ? 1.0 : 0.0;
    s.C2 = line == 0 ? 1.0 : 0.0;
    s.S2 = line == 1 ? 1.0 : 0.0;
    // leftBorderTotalField - 2.5 along the line (YeeGridLayout.cpp:2062-2263)
    for (int a = 0; a < 3; ++a) s.zero_inc[a] = a == line ? cfg.tfsf_left[a] - 2.5 : 0.0;
    for (int c = 0; c < 6; ++c) s.proj[c] = 0.0;
    // YeeGridLayout.h:1247-1760: -valE / -valH written as val * -1 (the same bits)
    switch (cfg.scheme)
    {
      case YEE_SCHEME_EXHY: s.proj[YEE_EX] = -1.0; s.proj[YEE_HY] = -1.0; break;
      case YEE_SCHEME_EXHZ: s.proj[YEE_EX] = 1.0; s.proj[YEE_HZ] = -1.0; break;
      case YEE_SCHEME_EYHX: s.proj[YEE_EY] = -1.0; s.proj[YEE_HX] = 1.0; break;
      case YEE_SCHEME_EYHZ: s.proj[YEE_EY] = -1.0; s.proj[YEE_HZ] = -1.0; break;
      case YEE_SCHEME_EZHX: s.proj[YEE_EZ] = 1.0; s.proj[YEE_HX] = 1.0; break;
      default: s.proj[YEE_EZ] = 1.0; s.proj[YEE_HY] = -1.0; break;   // EzHy
    }
  }
  else if (missing_axis (cfg) != 2)
  {
    // TEx / TMx on (y, z) and TEy / TMy on (x, z), run on the 3D kernels with the missing axis
    // of extent 1.  Their incident wave d = a sin(theta) + z cos(theta) (InternalScheme.cpp:
    // 210-239, 261-288) is the 3D expression x S1 C2 + y S1 S2 + z C1 with the factors of the
    // missing axis set to 0 and the other to 1 (exact: x * S1 * 1.0 == x * S1, + 0.0 is exact).
    const int miss = missing_axis (cfg);
    s.S1 = std::sin (a1);
    s.C1 = std::cos (a1);
    s.C2 = miss == 0 ? 0.0 : 1.0;
    s.S2 = miss == 0 ? 1.0 : 0.0;
    const int first = miss == 0 ? 1 : 0;
    // YeeGridLayout.cpp:2332-2333, 2385-2386, 2491-2492, 2544-2545
    s.zero_inc[first] = cfg.tfsf_left[first] - 2.5 * std::sin (a1);
    s.zero_inc[2] = cfg.tfsf_left[2] - 2.5 * std::cos (a1);
    s.zero_inc[miss] = 0;
    for (int c = 0; c < 6; ++c) s.proj[c] = 0.0;
    // YeeGridLayout.h:1761-1932 (TEx, TEy) and 2025-2190 (TMx, TMy)
    switch (cfg.scheme)
    {
      case YEE_SCHEME_TEX:
        s.proj[YEE_EY] = -std::cos (a1);
        s.proj[YEE_EZ] = std::sin (a1);
        s.proj[YEE_HX] = 1.0;
        break;
      case YEE_SCHEME_TEY:
        s.proj[YEE_EX] = -std::cos (a1);
        s.proj[YEE_EZ] = std::sin (a1);
        s.proj[YEE_HY] = -1.0;
        break;
      case YEE_SCHEME_TMX:
        s.proj[YEE_EX] = 1.0;
        s.proj[YEE_HY] = std::cos (a1);
        s.proj[YEE_HZ] = -std::sin (a1);
        break;
      default:   // TMy
        s.proj[YEE_EY] = -1.0;
        s.proj[YEE_HX] = std::cos (a1);
        s.proj[YEE_HZ] = -std::sin (a1);
        break;
    }
  }
  else if (cfg.scheme == YEE_SCHEME_TEZ)
  {
    // YeeGridLayout.h:1939-2024 and YeeGridLayout.cpp:2424-2440 (TEz); Hz = -valH, i.e. valH * -1
    s.proj[YEE_EX] = std::sin (a2);
    s.proj[YEE_EY] = -std::cos (a2);
    s.proj[YEE_HZ] = -1.0;
    s.zero_inc[0] = cfg.tfsf_left[0] - 2.5 * std::cos (a2);
    s.zero_inc[1] = cfg.tfsf_left[1] - 2.5 * std::sin (a2);
    s.zero_inc[2] = 0;
  }
  else
  {
    // YeeGridLayout.h:2194-2277 and YeeGridLayout.cpp:2583-2599 (TMz)
    s.proj[YEE_EZ] = 1.0;
    s.proj[YEE_HX] = std::sin (a2);
    s.proj[YEE_HY] = -std::cos (a2);
    s.zero_inc[0] = cfg.tfsf_left[0] - 2.5 * std::cos (a2);
    s.zero_inc[1] = cfg.tfsf_left[1] - 2.5 * std::sin (a2);
    s.zero_inc[2] = 0;
  }
  if (missing_axis (cfg) == 2 && cfg.layout == 1)
  {
    // H_CENTERED TEz / TMz run on the 3D kernels: their d = x C2 + y S2 is the 3D expression
    // x S1 C2 + y S1 S2 + z C1 with S1 = 1 and C1 = 0 (exact)
    s.S1 = 1.0;
    s.C1 = 0.0;
  }
  return s;
}

// SchemeHelper::initSigma (SchemeHelper.h:574-590): polynomial grading, m = 6, R = 1e-16.
static double graded_sigma (double dist, double boundary, double dx)
{
  const unsigned m = 6;
  const double r = 1e-16;
  const double smax = -std::log (r) * (m + 1.0) / (2.0 * std::sqrt (mu0 () / eps0 ()) * boundary);
  const double factor = smax / (dx * (std::pow (boundary, (double) m)) * (m + 1));
  const double hi = (dist + 1) * dx;
  const double lo = dist * dx;
  return factor * (std::pow (hi, (double) (m + 1)) - std::pow (lo, (double) (m + 1)));
}

int line_axis (const yee_config &cfg)
{
  static const int kAxis[7] = { -1, 2, 1, 2, 0, 1, 0 };
  return cfg.scheme >= YEE_SCHEME_EXHY && cfg.scheme <= YEE_SCHEME_EZHY ? kAxis[cfg.scheme] : -1;
}

bool axis_active (const yee_config &cfg, int axis)
{
  const int line = line_axis (cfg);
  return line >= 0 ? axis == line : axis != missing_axis (cfg);
}

int first_active_axis (const yee_config &cfg)
{
  const int line = line_axis (cfg);
  return line >= 0 ? line : (missing_axis (cfg) == 0 ? 1 : 0);
}

int missing_axis (const yee_config &cfg)
{
  switch (cfg.scheme)
  {
    case YEE_SCHEME_TEX: case YEE_SCHEME_TMX: return 0;
    case YEE_SCHEME_TEY: case YEE_SCHEME_TMY: return 1;
    case YEE_SCHEME_TEZ: case YEE_SCHEME_TMZ: return 2;
    default: return -1;
  }
}

std::vector<double> sigma_profile (const yee_config &cfg, int axis)
{
  // double material precision: the sigma grid is the fine material grid, 2 N cells with a PML
  // of 2 P cells and boundary 2 P dx (SchemeHelper.h:598, 640, 682)
  const int64_t mod = (cfg.flags & YEE_DOUBLE_MATERIAL_PRECISION) ? 2 : 1;
  const int64_t n = cfg.size[axis] * mod;
  std::vector<double> prof (n, 0.0);
  const double P = (double) (cfg.pml[axis] * mod);
  const double boundary = P * cfg.dx;
  for (int64_t p = 0; p < n; ++p)
  {
    const double pos = (double) p + 0.5;   // eps-grid coordinate
    if (axis == first_active_axis (cfg))
    {
      // the scheme's first coordinate (initSigmaX, Scheme.cpp:141-420; y for TEx / TMx):
      // fractional distance, right edge at N - P
      const double edge = (double) n;
      if (pos < P) prof[p] = graded_sigma (P - pos, boundary, cfg.dx);
      else if (pos >= edge - P) prof[p] = graded_sigma (pos - (edge - P), boundary, cfg.dx);
    }
    else
    {
      // its other coordinates (initSigmaY / Z): distance truncated to an integer, right edge
      // at N + 0.5 - P
      const double edge = (double) n + 0.5;
      if (pos < P) prof[p] = graded_sigma ((double) (int64_t) (P - pos), boundary, cfg.dx);
      else if (pos >= edge - P) prof[p] = graded_sigma ((double) (int64_t) (pos - (edge - P)), boundary, cfg.dx);
    }
  }
  return prof;
}

// A component's field point lies at an integer coordinate along `axis` when it is an E
// component of that axis or an H component of another axis; getMaterial then averages the
// two eps cells p and p+1.  At a half-integer coordinate the cells coincide and the average
// reproduces the cell value exactly, as do the 4-cell averages of H components.
double sigma_at (const std::vector<double> &prof, int comp, int axis, int64_t p, int64_t n, int layout)
{
  const CompDesc &c = kComp[comp];
  // H_CENTERED swaps the E and H positions (YeeGridLayout.cpp:1962-1967, 2016-2021)
  const bool h_like = c.is_h != (layout == 1);
  const bool integer_coord = h_like ? (axis != c.axis) : (axis == c.axis);
  if ((int64_t) prof.size () == 2 * n)
  {
    // fine sigma grid (double material precision): the 8-cell average of a profile that varies
    // along one axis only reduces exactly to the mean of its two fine cells on that axis,
    // 2p+1 and 2p+2 at an integer coordinate, 2p and 2p+1 at a half-integer one
    const int64_t f = 2 * p + (integer_coord ? 1 : 0);
    const double hi = f + 1 < 2 * n ? prof[f + 1] : 0.0;
    return (prof[f] + hi) / 2.0;
  }
  if (!integer_coord)
  {
    return prof[p];
  }
  const double hi = p + 1 < n ? prof[p + 1] : 0.0;
  return (prof[p] + hi) / 2.0;
}

std::vector<double> pml_tables (const yee_config &cfg, int64_t tab_len)
{
  std::vector<double> tab ((size_t) 6 * kNumRoles * tab_len, 0.0);
  if (!(cfg.flags & YEE_USE_PML))
  {
    return tab;
  }
  std::vector<double> prof[3];
  for (int a = 0; a < 3; ++a)
    if (axis_active (cfg, a)) prof[a] = sigma_profile (cfg, a);
  const Scalars s = compute_scalars (cfg);
  const double e0 = eps0 (), dt = s.dt, dx = cfg.dx;
  auto at = [&] (int comp, int axis, int64_t p) -> double {
    if (!axis_active (cfg, axis)) return 0.0;   // no sigma along an axis the scheme lacks: reads as 0
    return sigma_at (prof[axis], comp, axis, p, cfg.size[axis], cfg.layout);
  };
  for (int comp = 0; comp < 6; ++comp)
  {
    const CompDesc &c = kComp[comp];
    double *T = &tab[(size_t) comp * kNumRoles * tab_len];
    for (int64_t p = 0; p < tab_len; ++p)
    {
      if (axis_active (cfg, c.sig_curl) ? p < cfg.size[c.sig_curl] : p == 0)
      {
        const double m = at (comp, c.sig_curl, p);
        const double dd = (2 * e0 * 1.0 + m * dt);
        T[kCAC * tab_len + p] = (2 * e0 * 1.0 - m * dt) / dd;
        T[kCBC * tab_len + p] = (2 * e0 * dt / dx) / dd;
      }
      if (axis_active (cfg, c.sig4) ? p < cfg.size[c.sig4] : p == 0)
      {
        const double m = at (comp, c.sig4, p);
        T[kP4 * tab_len + p] = 2 * e0 * 1.0 + m * dt;
        T[kM4 * tab_len + p] = 2 * e0 * 1.0 - m * dt;
      }
      if (axis_active (cfg, c.sig5) ? p < cfg.size[c.sig5] : p == 0)
      {
        const double m = at (comp, c.sig5, p);
        const double dd = (2 * e0 * 1.0 + m * dt);
        T[kCA5 * tab_len + p] = (2 * e0 * 1.0 - m * dt) / dd;
        T[kDD5 * tab_len + p] = dd;
      }
    }
  }
  return tab;
}

double sphere_accurate_3d (double mx, double my, double mz, double cx, double cy, double cz,
                           double radius, double eps, double outside, int accuracy)
{
  // Volume fraction of the cell [m-0.5, m+0.5]^3 inside the sphere, integrated column by
  // column over an accuracy x accuracy raster of the xy face.
  const double x0 = mx - 0.5, y0 = my - 0.5, z0 = mz - 0.5, z1 = mz + 0.5;
  const double h = 1.0 / accuracy;
  const double area = h * h;
  const bool upper = mz > cz;
  double vol = 0;
  for (int i = 0; i < accuracy; ++i)
  {
    for (int j = 0; j < accuracy; ++j)
    {
      const double px = x0 + i * h;
      const double py = y0 + j * h;
      const double r2 = radius * radius - (px - cx) * (px - cx) - (py - cy) * (py - cy);
      double height = 0.0;
      if (!(r2 < 0))
      {
        double zs = upper ? cz + std::sqrt (r2) : cz - std::sqrt (r2);
        zs = std::min (std::max (zs, z0), z1);
        height = upper ? zs - z0 : z1 - zs;
      }
      vol += height * area;
    }
  }
  vol = std::min (vol, 1.0);
  return eps * vol + outside * (1 - vol);
}

double sphere_accurate_2d (double mx, double my, double cx, double cy, double radius, double eps,
                           double outside, int accuracy)
{
  const double x0 = mx - 0.5, y0 = my - 0.5, y1 = my + 0.5;
  const double h = 1.0 / accuracy;
  const bool upper = my > cy;
  double vol = 0;
  for (int i = 0; i < accuracy; ++i)
  {
    const double px = x0 + i * h;
    const double r2 = radius * radius - (px - cx) * (px - cx);
    double height = 0.0;
    if (!(r2 < 0))
    {
      double ys = upper ? cy + std::sqrt (r2) : cy - std::sqrt (r2);
      ys = std::min (std::max (ys, y0), y1);
      height = upper ? ys - y0 : y1 - ys;
    }
    vol += height * h;
  }
  vol = std::min (vol, 1.0);
  return eps * vol + outside * (1 - vol);
}

double sphere_stair_3d (double mx, double my, double mz, double cx, double cy, double cz,
                        double radius, double eps, double outside)
{
  for (int q = 0; q < 8; ++q)
  {
    const double x = mx + ((q & 1) ? 0.5 : -0.5) - cx;
    const double y = my + ((q & 2) ? 0.5 : -0.5) - cy;
    const double z = mz + ((q & 4) ? 0.5 : -0.5) - cz;
    if (!(std::sqrt (x * x + y * y + z * z) <= radius)) return outside;
  }
  return eps;
}

double sphere_stair_2d (double mx, double my, double cx, double cy, double radius, double eps, double outside)
{
  for (int q = 0; q < 4; ++q)
  {
    const double x = mx + ((q & 1) ? 0.5 : -0.5) - cx;
    const double y = my + ((q & 2) ? 0.5 : -0.5) - cy;
    if (!(std::sqrt (x * x + y * y) <= radius)) return outside;
  }
  return eps;
}

// Approximation::approximateSphereAccurate on a 1D grid (Approximation.cpp:390-448): the cell
// [m - 1/2, m + 1/2] against [c - r, c + r].  The reference tests "start >= left || end <= right"
// for a cell inside, so every cell that overlaps the segment takes the full value (the partial
// branches after it are reached only by cells wider than the segment).
double sphere_accurate_1d (double m, double c, double radius, double eps, double outside)
{
  const double start = m - 0.5, end = m + 0.5;
  const double left = c - radius, right = c + radius;
  double vol = 0;
  if (end <= left || start >= right)
    vol = 0;
  else if (start >= left || end <= right)
    vol = 1;
  else if (start < left)
    vol = end - left;
  else if (end > right)
    vol = right - start;
  if (vol > 1) vol = 1;
  return eps * vol + outside * (1 - vol);
}

// Approximation::approximateSphereStair on a 1D grid (Approximation.cpp:678-699): both cell ends
// within the radius.
double sphere_stair_1d (double m, double c, double radius, double eps, double outside)
{
  const double d1 = m - 0.5 - c, d2 = m + 0.5 - c;
  return (std::sqrt (d1 * d1) <= radius && std::sqrt (d2 * d2) <= radius) ? eps : outside;
}

}  // namespace yee
