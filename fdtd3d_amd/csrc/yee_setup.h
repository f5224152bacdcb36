// Host-side FP64 set-up of the Yee engine: physical constants, time step and
// incident-wave phase velocity, PML sigma profiles and their per-component
// coefficient tables, TF/SF geometry, material generators.
//
// Everything here is computed once per run in double precision with the same
// expressions (and operand order) as the reference, so the FP64 engine sees the
// coefficients the reference computes per cell.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/yee.h"

namespace yee {

// Source/Physics/PhysicsConst.h:27-33
constexpr double kPi = 3.14159265358979323846;
constexpr double kC = 299792458.0;
inline double mu0 () { return 4 * kPi * 0.0000001; }
inline double eps0 () { return 1 / (mu0 () * (kC * kC)); }

// Field-point descriptors of the E_CENTERED Yee cell (Source/Layout/YeeGridLayout.cpp:1937-1969)
// and the grid bindings of InternalScheme::calculateFieldStepInit (InternalScheme.inc.h:1902-2295).
struct CompDesc
{
  int axis;        // field direction
  int is_h;
  int opp1, opp2;  // opposite components read by the curl
  int sig_curl, sig4, sig5;  // sigma axes of the curl, D->E numerator and denominator terms
};
extern const CompDesc kComp[6];

// Twice the real coordinate of component `comp` at integer index p along `axis`.
inline int64_t twice_coord (int comp, int axis, int64_t p)
{
  const CompDesc &c = kComp[comp];
  int64_t off = c.is_h ? (axis == c.axis ? 1 : 2) : (axis == c.axis ? 2 : 1);
  return 2 * p + off;
}

struct Scalars
{
  double dt, freq, rel;            // InternalScheme::initScheme
  double a1, a2, a3;               // angles in radians (main.cpp:1101-1110)
  double S1, C1, S2, C2;           // sin/cos of a1, a2
  double proj[6];                  // get<comp>FromIncident factors
  double zero_inc[3];              // zeroIncCoordFP
};

Scalars compute_scalars (const yee_config &cfg);

// The axis a 2D scheme lacks (main.cpp:817-1000: TEx/TMx on (y, z), TEy/TMy on (x, z),
// TEz/TMz on (x, y)); -1 for 3D.
int missing_axis (const yee_config &cfg);

// The axis of a 1D scheme's grid (main.cpp:629-800: ExHy / EyHx z, ExHz / EzHx y, EyHz / EzHy x);
// -1 for the other schemes.
int line_axis (const yee_config &cfg);

// Whether the scheme's grid extends along `axis` (3D: every axis; 2D: two; 1D: one).
bool axis_active (const yee_config &cfg, int axis);

// The axis of the scheme's first grid coordinate (y for TEx / TMx, the line's axis for 1D, x
// otherwise): the one whose PML sigma follows initSigmaX.
int first_active_axis (const yee_config &cfg);

double wave_number (double delta, double lambda, double courant, double nlambda, double a1, double a2);

// Sigma of eps-cell p along `axis` (SchemeHelper::initSigmaX/Y/Z, SchemeHelper.h:592-718).
std::vector<double> sigma_profile (const yee_config &cfg, int axis);

// Per-component 1D PML coefficient tables (length = size along the table's axis):
//   CAC, CBC : curl stage, along sig_curl   (computeCaCb, InternalScheme.inc.h:2327-2333)
//   P4, M4   : 2 eps0 +- sigma4 dt, along sig4  (calculateFieldStepIterationPML, :2620-2627)
//   CA5, DD5 : Ca and denominator, along sig5
enum PmlRole { kCAC = 0, kCBC = 1, kP4 = 2, kM4 = 3, kCA5 = 4, kDD5 = 5, kNumRoles = 6 };
// Returns [comp][role][tab_len] (tab_len = max size), zero-filled where unused.
std::vector<double> pml_tables (const yee_config &cfg, int64_t tab_len);

// Averaged sigma of `comp` along `axis` at index p (InternalScheme::getMaterial on a sigma grid).
double sigma_at (const std::vector<double> &prof, int comp, int axis, int64_t p, int64_t n, int layout = 0);

// Material generators restating Approximation::approximateSphere* (Approximation.cpp:563-764).
double sphere_accurate_3d (double mx, double my, double mz, double cx, double cy, double cz,
                           double radius, double eps, double outside, int accuracy);
double sphere_accurate_2d (double mx, double my, double cx, double cy, double radius, double eps,
                           double outside, int accuracy);
double sphere_stair_3d (double mx, double my, double mz, double cx, double cy, double cz,
                        double radius, double eps, double outside);
double sphere_stair_2d (double mx, double my, double cx, double cy, double radius, double eps, double outside);
// The 1D segment of a "sphere" on a 1D line (Approximation.cpp:390-448, 678-699).
double sphere_accurate_1d (double m, double c, double radius, double eps, double outside);
double sphere_stair_1d (double m, double c, double radius, double eps, double outside);

}  // namespace yee
