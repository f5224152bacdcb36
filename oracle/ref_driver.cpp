/*
 * ref_driver.cpp -- test infrastructure, NOT product code.
 *
 * A small driver that runs the reference's OWN per-cell arithmetic (compiled from
 * the reference sources by oracle/Makefile into oracle/_ref/) so that the C oracle
 * in oracle/yee_oracle.c can be checked against it, and golden fixtures can be
 * written for tests/golden/.
 *
 * What comes from the reference (linked, not copied):
 *   - option parsing                      Settings::SetupFromCmd   (Settings/Settings.cpp:558)
 *   - Yee layout, TF/SF predicates        YeeGridLayout            (Layout/YeeGridLayout.cpp)
 *   - per-cell updates                    InternalScheme::calculateFieldStepIteration / ...PML /
 *                                         ...PMLMetamaterials       (Scheme/InternalScheme.inc.h:2351-2640)
 *   - incident 1D grid, point source      performPlaneWave{E,H}Steps, performPointSourceCalc
 *                                         (Scheme/InternalScheme.inc.h:3276-3357, 2818-2895)
 *   - sphere approximation                Approximation::approximateSphere{Accurate,Stair}
 *                                         (Layout/Approximation.cpp:563-764)
 *
 * What this driver restates itself (because the reference's versions live in
 * Scheme/SchemeHelper, whose headers pull in the un-vendored EasyBMP):
 *   - the step ordering of Scheme::performNStepsForBlock (Scheme/Scheme.inc.h:28-473)
 *     and the three sweeps of Scheme::calculateFieldStep (Scheme/Scheme.inc.h:876-1344);
 *   - the material set-up of Scheme::initGrids (Scheme/Scheme.inc.h:2962-3165);
 *   - the PML sigma profiles of SchemeHelper::initSigma{,X,Y,Z} (Scheme/SchemeHelper.h:574-718);
 *   - with --use-ca-cb / --use-ca-cb-pml, the coefficient grids Scheme::initGrids precomputes
 *     (Scheme/Scheme.inc.h:3309-3548, 3549-3791), which the reference's own
 *     usePrecomputedGrids = true iterations then read.  --use-ca-cb-pml-metamaterials is
 *     refused: the reference's binding (InternalScheme.inc.h:1902-2295) never sets the a1 / a2
 *     grids (CA1 / CA2 stay null, Scheme.inc.h:932) and the precomputed Drude iteration
 *     dereferences them (InternalScheme.inc.h:2501-2502).
 *
 * With env YEE_REF_TXT set to a directory, every E / H component of a 3D run is also written
 * there through the reference's own TXTDumper (File-Management/Dumper/TXTDumper.h:55-92), whole
 * grid, newest layer, as <dir>/<name>.txt.
 *
 * Output: one binary file (path from env YEE_REF_OUT, default "ref_out.bin") with
 * records  [char name[16]][int64 count][double values[count]]  holding the latest
 * time layer of every state grid, the material grids and the incident 1D grids.
 * Stdout carries "Total time = ..." for the stepping loop only, as the reference CLI
 * prints it (Source/main.cpp:1211).
 */

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <sys/time.h>

#include <iomanip>
#include <limits>

#include "InternalScheme.h"
#include "Approximation.h"
#include "CallBack.h"
#include "TXTDumper.h"   /* expects <iomanip> and <limits> first, as its own TXTDumper.cpp includes them */

namespace
{

FILE *g_out = NULL;

void putRecord (const char *name, const std::vector<double> &v)
{
  char buf[16];
  memset (buf, 0, sizeof (buf));
  strncpy (buf, name, 15);
  int64_t n = (int64_t) v.size ();
  fwrite (buf, 1, 16, g_out);
  fwrite (&n, sizeof (n), 1, g_out);
  if (n)
  {
    fwrite (v.data (), sizeof (double), v.size (), g_out);
  }
}

/* Polynomial-graded sigma of one cell; restated from SchemeHelper::initSigma (SchemeHelper.h:574-590). */
FPValue sigmaValue (FPValue dist, FPValue boundary, FPValue gridStep)
{
  FPValue eps0 = PhysicsConst::Eps0;
  FPValue mu0 = PhysicsConst::Mu0;
  uint32_t exponent = 6;
  FPValue R_err = 1e-16;
  FPValue sigma_max_1 = -log (R_err) * (exponent + 1.0) / (2.0 * sqrt (mu0 / eps0) * boundary);
  FPValue boundaryFactor = sigma_max_1 / (gridStep * (pow (boundary, exponent)) * (exponent + 1));
  FPValue x1 = (dist + 1) * gridStep;
  FPValue x2 = dist * gridStep;
  return boundaryFactor * (pow (x1, (exponent + 1)) - pow (x2, (exponent + 1)));
}

/* Real / imaginary part of a field value (the complex build, -DCOMPLEX_FIELD_VALUES, dumps both). */
double partOf (const FieldValue &v, int part)
{
#ifdef COMPLEX_FIELD_VALUES
  return part ? (double) v.imag () : (double) v.real ();
#else
  return part ? 0.0 : (double) v;
#endif
}

/* Material files hold FPValue values (real materials in either build). */
bool readRaw (const std::string &name, std::vector<FPValue> &dst, size_t n)
{
  FILE *f = fopen (name.c_str (), "rb");
  if (!f)
  {
    fprintf (stderr, "cannot open %s\n", name.c_str ());
    return false;
  }
  dst.resize (n);
  size_t got = fread (dst.data (), sizeof (FPValue), n, f);
  fclose (f);
  if (got != n)
  {
    fprintf (stderr, "short file %s: %zu of %zu values\n", name.c_str (), got, n);
    return false;
  }
  return true;
}

/* Coordinate slot sl of a 1D, 2D or 3D coordinate. */
inline FPValue fpSlot (const GridCoordinateFP1D &c, int) { return c.get1 (); }
inline FPValue fpSlot (const GridCoordinateFP2D &c, int sl) { return sl == 0 ? c.get1 () : c.get2 (); }
inline FPValue fpSlot (const GridCoordinateFP3D &c, int sl) { return sl == 0 ? c.get1 () : (sl == 1 ? c.get2 () : c.get3 ()); }
inline grid_coord tcSlot (const GridCoordinate1D &c, int) { return c.get1 (); }
inline grid_coord tcSlot (const GridCoordinate2D &c, int sl) { return sl == 0 ? c.get1 () : c.get2 (); }
inline grid_coord tcSlot (const GridCoordinate3D &c, int sl) { return sl == 0 ? c.get1 () : (sl == 1 ? c.get2 () : c.get3 ()); }

/* One run of the reference's InternalScheme for scheme Type with Yee layout LT
 * (E_CENTERED or H_CENTERED, --layout-type 0 / 1). */
template <SchemeType_t Type, template <typename, bool> class TCoord, LayoutType LT>
class RefRun
{
  typedef TCoord<grid_coord, true> TC;
  typedef TCoord<grid_coord, false> TCS;
  typedef TCoord<FPValue, true> TCFP;
  typedef InternalScheme<Type, TCoord, LT> IS;
  typedef YeeGridLayout<Type, TCoord, LT> Layout;

  CoordinateType ct1, ct2, ct3;
  TC size;
  Layout *layout;
  IS *is;
  int dim;
  int miss;   /* the axis a 2D scheme lacks (-1 in 3D): TEx/TMx 0, TEy/TMy 1, TEz/TMz 2 */
  /* diff norm and percent of each step per component (index 0 .. 5: Ex .. Hz) */
  std::vector<double> norms[6];
  int axis1;  /* the axis of a 1D scheme (-1 otherwise) */

  /* an axis the scheme lacks */
  bool missing (int a) const
  {
    return dim == 1 ? a != axis1 : a == miss;
  }

  /* slot of physical axis a among the scheme's coordinates */
  int slot (int a) const
  {
    if (dim == 1) return 0;
    return (miss >= 0 && a > miss) ? a - 1 : a;
  }

  TC coordOf (grid_coord i, grid_coord j, grid_coord k) const
  {
    return TC::initAxesCoordinate (i, j, k, ct1, ct2, ct3);
  }

  /* extent along physical axis a (1 along the missing axis) */
  grid_coord sz (int axis) const
  {
    return missing (axis) ? 1 : tcSlot (size, slot (axis));
  }

  /* material grid extent along physical axis a: sizeEps, twice the field grid with
   * --use-double-material-precision (YeeGridLayout.cpp:1972) */
  grid_coord msz (int axis) const
  {
    return missing (axis) ? 1 : sz (axis) * (layout->getIsDoubleMaterialPrecision () ? 2 : 1);
  }

  /* every cell of a material grid */
  template <class F>
  void forAllMat (F f)
  {
    for (grid_coord i = 0; i < msz (0); ++i)
      for (grid_coord j = 0; j < msz (1); ++j)
        for (grid_coord k = 0; k < msz (2); ++k)
          f (coordOf (i, j, k));
  }

  std::vector<double> dumpMat (Grid<TC> *g)
  {
    std::vector<double> v;
    forAllMat ([&] (TC p) { v.push_back (partOf (*g->getFieldValue (p, 0), 0)); });
    return v;
  }

  template <class F>
  void forAll (grid_coord lo, grid_coord hiOff, F f)
  {
    grid_coord l[3], h[3];
    for (int a = 0; a < 3; ++a)
    {
      l[a] = missing (a) ? 0 : lo;
      h[a] = missing (a) ? 1 : sz (a) - hiOff;
    }
    for (grid_coord i = l[0]; i < h[0]; ++i)
      for (grid_coord j = l[1]; j < h[1]; ++j)
        for (grid_coord k = l[2]; k < h[2]; ++k)
          f (coordOf (i, j, k));
  }

  std::vector<double> dumpGrid (Grid<TC> *g, int layer, int part = 0)
  {
    std::vector<double> v;
    forAll (0, 0, [&] (TC p) { v.push_back (partOf (*g->getFieldValue (p, layer), part)); });
    return v;
  }

  void loadMaterial (Grid<TC> *g, const std::string &file)
  {
    if (file.empty ())
    {
      return;
    }
    size_t n = (size_t) msz (0) * msz (1) * msz (2);
    std::vector<FPValue> raw;
    if (!readRaw (file, raw, n))
    {
      exit (2);
    }
    size_t idx = 0;
    forAllMat ([&] (TC p) { g->setFieldValue (FIELDVALUE (raw[idx++], FPValue (0)), p, 0); });
  }

  /* Sphere of `value` (outside `outside`) on the material grid; follows Scheme::initGrids
   * (Scheme.inc.h:2983-3012), modifier 2 with double material precision. */
  void sphere (Grid<TC> *g, FPValue value, FPValue outside, grid_coord cx, grid_coord cy, grid_coord cz, int radius)
  {
    const FPValue modifier = layout->getIsDoubleMaterialPrecision () ? 2 : 1;
    TCFP center = TCFP::initAxesCoordinate ((FPValue) cx, (FPValue) cy, (FPValue) cz, ct1, ct2, ct3) * modifier
                  + TCFP::initAxesCoordinate (0.5, 0.5, 0.5, ct1, ct2, ct3);
    forAllMat ([&] (TC p) {
      TCFP posAbs = layout->getEpsCoordFP (p);
      FieldValue v;
      if (SOLVER_SETTINGS.getDoUseStairApproximation ())
      {
        v = Approximation::approximateSphereStair (posAbs, center, radius * modifier, value, outside);
      }
      else
      {
        v = Approximation::approximateSphereAccurate (posAbs, center, radius * modifier, value, outside);
      }
      g->setFieldValue (v, p, 0);
    });
  }

  void initSigmas ()
  {
    FPValue dx = SOLVER_SETTINGS.getGridStep ();
    TC pml = layout->getLeftBorderPML ();
    /* Scheme::initSigmas (Scheme.cpp:141-420) calls initSigmaX on the scheme's first coordinate
     * and initSigmaY / initSigmaZ on its second / third, whatever physical axis they are (TEx /
     * TMx: initSigmaX for y).  initSigmaX: fractional distance, right threshold N - P
     * (SchemeHelper.h:592-630); initSigmaY / Z: truncated integer distance, right threshold
     * N + 0.5 - P (SchemeHelper.h:632-718). */
    for (int axis = 0; axis < 3; ++axis)
    {
      if (missing (axis)) continue;
      Grid<TC> *s = axis == 0 ? is->getSigmaX () : (axis == 1 ? is->getSigmaY () : is->getSigmaZ ());
      /* the sigma grid is a material grid: fine, with PML size 2 P, under double material
       * precision (SchemeHelper.h:598, 640, 682) */
      FPValue P = FPValue (tcSlot (pml, slot (axis)) * (layout->getIsDoubleMaterialPrecision () ? 2 : 1));
      FPValue boundary = P * dx;
      const bool first = slot (axis) == 0;
      FPValue n = first ? FPValue (msz (axis)) : (FPValue) msz (axis) + FPValue (0.5);
      forAllMat ([&] (TC p) {
        FPValue y = fpSlot (layout->getEpsCoordFP (p), slot (axis));
        FPValue val = 0;
        if (y < P)
        {
          val = first ? sigmaValue (P - y, boundary, dx) : sigmaValue ((grid_coord) (P - y), boundary, dx);
        }
        else if (y >= n - P)
        {
          val = first ? sigmaValue (y - (n - P), boundary, dx) : sigmaValue ((grid_coord) (y - (n - P)), boundary, dx);
        }
        s->setFieldValue (val, p, 0);
      });
    }
  }

  /* The coefficient grids of --use-ca-cb / --use-ca-cb-pml for component grid_type, as
   * Scheme::initGrids fills them (Scheme.inc.h:3309-3548 curl Ca / Cb, 3549-3791 PML D -> E
   * Ca / Cb / Cc), over the cells the sweeps update; the material grids come from the
   * reference's own binding (calculateFieldStepInit). */
  template <uint8_t grid_type, bool usePML, bool useMeta>
  void fillCoefficients ()
  {
    Grid<TC> *grid = NULLPTR;
    GridType gridType = GridType::NONE;
    Grid<TC> *mg = NULLPTR, *mg1 = NULLPTR, *mg2 = NULLPTR, *mg3 = NULLPTR, *mg4 = NULLPTR, *mg5 = NULLPTR;
    GridType mgt = GridType::NONE, mgt1 = GridType::NONE, mgt2 = GridType::NONE, mgt3 = GridType::NONE,
             mgt4 = GridType::NONE, mgt5 = GridType::NONE;
    Grid<TC> *opp1 = NULLPTR, *opp2 = NULLPTR, *pml1 = NULLPTR, *pml2 = NULLPTR;
    GridType pmlType1 = GridType::NONE, pmlType2 = GridType::NONE;
    Grid<TC> *Ca = NULLPTR, *Cb = NULLPTR, *CB0 = NULLPTR, *CB1 = NULLPTR, *CB2 = NULLPTR, *CA1 = NULLPTR,
             *CA2 = NULLPTR, *CaPML = NULLPTR, *CbPML = NULLPTR, *CcPML = NULLPTR;
    SourceCallBack rhs = NULLPTR, border = NULLPTR, exact = NULLPTR;
    FPValue modifier = 0;
    is->template calculateFieldStepInit<grid_type, usePML, useMeta> (&grid, &gridType,
      &mg, &mgt, &mg1, &mgt1, &mg2, &mgt2, &mg3, &mgt3, &mg4, &mgt4, &mg5, &mgt5,
      &opp1, &opp2, &pml1, &pmlType1, &pml2, &pmlType2, &rhs, &border, &exact, &modifier,
      &Ca, &Cb, &CB0, &CB1, &CB2, &CA1, &CA2, &CaPML, &CbPML, &CcPML);
    const FPValue dt = is->getGridTimeStep (), dx = is->getGridStep ();
    const FPValue eps0 = PhysicsConst::Eps0, k_mod = FPValue (1);
    forAll (1, 1, [&] (TC p) {
      TC posAbs = grid->getTotalPosition (p);
      if (Ca != NULLPTR)
      {
        FPValue ca, cb;
        if (usePML)
        {
          FPValue material = mg ? is->getMaterial (posAbs, gridType, mg, mgt) : 0;
          FPValue dd = (2 * eps0 * k_mod + material * dt);
          ca = (2 * eps0 * k_mod - material * dt) / dd;
          cb = (2 * eps0 * dt / dx) / dd;
        }
        else
        {
          FPValue material = is->getMaterial (posAbs, gridType, mg, mgt);
          ca = FPValue (1);
          cb = dt / (material * modifier * dx);   /* modifier: PhysicsConst::Eps0 (E) or Mu0 (H) */
        }
        Ca->setFieldValue (FIELDVALUE (ca, FPValue (0)), p, 0);
        Cb->setFieldValue (FIELDVALUE (cb, FPValue (0)), p, 0);
      }
      if (usePML && CaPML != NULLPTR)
      {
        FPValue material1 = is->getMaterial (posAbs, pmlType1, mg1, mgt1);
        FPValue material4 = mg4 ? is->getMaterial (posAbs, pmlType1, mg4, mgt4) : 0;
        FPValue material5 = mg5 ? is->getMaterial (posAbs, pmlType1, mg5, mgt5) : 0;
        FPValue mod = useMeta ? FPValue (1) : material1 * modifier;
        FPValue dd = (2 * eps0 * k_mod + material5 * dt);
        CaPML->setFieldValue (FIELDVALUE ((2 * eps0 * k_mod - material5 * dt) / dd, FPValue (0)), p, 0);
        CbPML->setFieldValue (FIELDVALUE (((2 * eps0 * k_mod + material4 * dt) / (mod)) / dd, FPValue (0)), p, 0);
        CcPML->setFieldValue (FIELDVALUE (((2 * eps0 * k_mod - material4 * dt) / (mod)) / dd, FPValue (0)), p, 0);
      }
    });
  }

  template <uint8_t grid_type>
  void fillCoefficientsDispatch ()
  {
    bool pml = SOLVER_SETTINGS.getDoUsePML ();
    bool meta = SOLVER_SETTINGS.getDoUseMetamaterials ();
    if (pml && meta) fillCoefficients<grid_type, true, true> ();
    else if (pml) fillCoefficients<grid_type, true, false> ();
    else fillCoefficients<grid_type, false, false> ();
  }

  template <uint8_t grid_type, bool usePML, bool useMeta>
  void fieldStep (time_step t)
  {
    Grid<TC> *grid = NULLPTR;
    GridType gridType = GridType::NONE;
    Grid<TC> *mg = NULLPTR, *mg1 = NULLPTR, *mg2 = NULLPTR, *mg3 = NULLPTR, *mg4 = NULLPTR, *mg5 = NULLPTR;
    GridType mgt = GridType::NONE, mgt1 = GridType::NONE, mgt2 = GridType::NONE, mgt3 = GridType::NONE,
             mgt4 = GridType::NONE, mgt5 = GridType::NONE;
    Grid<TC> *opp1 = NULLPTR, *opp2 = NULLPTR, *pml1 = NULLPTR, *pml2 = NULLPTR;
    GridType pmlType1 = GridType::NONE, pmlType2 = GridType::NONE;
    Grid<TC> *Ca = NULLPTR, *Cb = NULLPTR, *CB0 = NULLPTR, *CB1 = NULLPTR, *CB2 = NULLPTR, *CA1 = NULLPTR,
             *CA2 = NULLPTR, *CaPML = NULLPTR, *CbPML = NULLPTR, *CcPML = NULLPTR;
    SourceCallBack rhs = NULLPTR, border = NULLPTR, exact = NULLPTR;
    FPValue modifier = 0;
    TCS d11, d12, d21, d22;

    is->template calculateFieldStepInit<grid_type, usePML, useMeta> (&grid, &gridType,
      &mg, &mgt, &mg1, &mgt1, &mg2, &mgt2, &mg3, &mgt3, &mg4, &mgt4, &mg5, &mgt5,
      &opp1, &opp2, &pml1, &pmlType1, &pml2, &pmlType2, &rhs, &border, &exact, &modifier,
      &Ca, &Cb, &CB0, &CB1, &CB2, &CA1, &CA2, &CaPML, &CbPML, &CcPML);
    is->template calculateFieldStepInitDiff<grid_type> (&d11, &d12, &d21, &d22);

    /* the reference skips every field update at t == 0 (Scheme.inc.h:1032) */
    if (t > 0)
    {
      bool isE = grid_type == static_cast<uint8_t> (GridType::EX) || grid_type == static_cast<uint8_t> (GridType::EY)
                 || grid_type == static_cast<uint8_t> (GridType::EZ);
      FPValue timestep = isE ? FPValue (t) : FPValue (t) + FPValue (0.5);
      forAll (1, 1, [&] (TC p) {
        /* the right side is read at the component's point (Scheme.inc.h:1108-1148) */
        TCFP coordFP;
        if (rhs != NULLPTR) coordFP = compCoordFP<grid_type> (grid->getTotalPosition (p));
        if (Ca != NULLPTR)
          is->template calculateFieldStepIteration<grid_type, true> (timestep, p, grid->getTotalPosition (p),
            d11, d12, d21, d22, grid, coordFP, opp1, opp2, rhs, Ca, Cb, usePML, gridType, mg, mgt, modifier);
        else
          is->template calculateFieldStepIteration<grid_type, false> (timestep, p, grid->getTotalPosition (p),
            d11, d12, d21, d22, grid, coordFP, opp1, opp2, rhs, Ca, Cb, usePML, gridType, mg, mgt, modifier);
      });
      /* current source after the curl sweep (Scheme.inc.h:1171-1260) */
      bool cur = false;
      switch (grid_type)
      {
        case (static_cast<uint8_t> (GridType::EX)): cur = SOLVER_SETTINGS.getDoUseCurrentSourceJx (); break;
        case (static_cast<uint8_t> (GridType::EY)): cur = SOLVER_SETTINGS.getDoUseCurrentSourceJy (); break;
        case (static_cast<uint8_t> (GridType::EZ)): cur = SOLVER_SETTINGS.getDoUseCurrentSourceJz (); break;
        case (static_cast<uint8_t> (GridType::HX)): cur = SOLVER_SETTINGS.getDoUseCurrentSourceMx (); break;
        case (static_cast<uint8_t> (GridType::HY)): cur = SOLVER_SETTINGS.getDoUseCurrentSourceMy (); break;
        case (static_cast<uint8_t> (GridType::HZ)): cur = SOLVER_SETTINGS.getDoUseCurrentSourceMz (); break;
        default: break;
      }
      if (cur)
      {
#ifdef COMPLEX_FIELD_VALUES
        /* the complex build's current (Scheme.inc.h:1230-1235) */
        FieldValue current = FieldValue (
          cos (is->getGridTimeStep () * (timestep - 1) * FPValue (2) * PhysicsConst::Pi * is->getSourceFrequency ()),
          -sin (is->getGridTimeStep () * (timestep - 1) * FPValue (2) * PhysicsConst::Pi * is->getSourceFrequency ()));
#else
        FieldValue current = sin (is->getGridTimeStep () * timestep * FPValue (2) * PhysicsConst::Pi * is->getSourceFrequency ());
#endif
        current = current * PhysicsConst::Mu0 / is->getGridTimeStep () / FPValue (2.0);
        if (Ca != NULLPTR)
          is->template calculateFieldStepIterationCurrent<grid_type, true> (current, grid, Ca, Cb, usePML, gridType, mg,
                                                                            mgt, modifier);
        else
          is->template calculateFieldStepIterationCurrent<grid_type, false> (current, grid, Ca, Cb, usePML, gridType, mg,
                                                                             mgt, modifier);
      }
      if (usePML)
      {
        if (useMeta)
        {
          forAll (1, 1, [&] (TC p) {
            is->template calculateFieldStepIterationPMLMetamaterials<false> (t, p, grid, pml1,
              CB0, CB1, CB2, CA1, CA2, gridType, mg1, mgt1, mg2, mgt2, mg3, mgt3, modifier);
          });
        }
        forAll (1, 1, [&] (TC p) {
          if (CaPML != NULLPTR)
            is->template calculateFieldStepIterationPML<useMeta, true> (t, p, grid, pml1, pml2, CaPML, CbPML, CcPML,
              pmlType1, mg1, mgt1, mg4, mgt4, mg5, mgt5, modifier);
          else
            is->template calculateFieldStepIterationPML<useMeta, false> (t, p, grid, pml1, pml2, CaPML, CbPML, CcPML,
              pmlType1, mg1, mgt1, mg4, mgt4, mg5, mgt5, modifier);
        });
      }
    }

    /* border values and the diff norm of an analytic test problem, also at t == 0
     * (Scheme::calculateFieldStep, Scheme.inc.h:1346-1550) */
    if (border != NULLPTR)
    {
      forAll (0, 0, [&] (TC p) { is->template calculateFieldStepIterationBorder<grid_type> (t, grid->getTotalPosition (p), grid, border); });
    }
    if (exact != NULLPTR)
    {
      FPValue normRe = 0, normIm = 0, normMod = 0, maxRe = 0, maxIm = 0, maxMod = 0;
      grid_coord lo[3], hi[3];
      const grid_coord ns[3] = { SOLVER_SETTINGS.getExactSolutionCompareStartX (), SOLVER_SETTINGS.getExactSolutionCompareStartY (),
                                 SOLVER_SETTINGS.getExactSolutionCompareStartZ () };
      const grid_coord ne[3] = { SOLVER_SETTINGS.getExactSolutionCompareEndX (), SOLVER_SETTINGS.getExactSolutionCompareEndY (),
                                 SOLVER_SETTINGS.getExactSolutionCompareEndZ () };
      for (int a = 0; a < 3; ++a)
      {
        lo[a] = missing (a) ? 0 : (ns[a] != 0 ? ns[a] : 1);
        hi[a] = missing (a) ? 1 : (ne[a] != 0 ? ne[a] : sz (a) - 1);
      }
      Grid<TC> *normGrid = usePML ? pml2 : grid;
      for (grid_coord i = lo[0]; i < hi[0]; ++i)
        for (grid_coord j = lo[1]; j < hi[1]; ++j)
          for (grid_coord k = lo[2]; k < hi[2]; ++k)
            is->template calculateFieldStepIterationExact<grid_type> (t, normGrid->getTotalPosition (coordOf (i, j, k)),
              normGrid, exact, normRe, normIm, normMod, maxRe, maxIm, maxMod);
      normRe = sqrt (normRe / normGrid->getTotalSize ().calculateTotalCoord ());
      const int ci = grid_type == static_cast<uint8_t> (GridType::EX) ? 0 : grid_type == static_cast<uint8_t> (GridType::EY) ? 1
                     : grid_type == static_cast<uint8_t> (GridType::EZ) ? 2 : grid_type == static_cast<uint8_t> (GridType::HX) ? 3
                     : grid_type == static_cast<uint8_t> (GridType::HY) ? 4 : 5;
      std::vector<double> &rec = norms[ci];
      rec.push_back ((double) normRe);
      rec.push_back ((double) (normRe * 100.0 / maxRe));
#ifdef COMPLEX_FIELD_VALUES
      /* the complex build's line (Scheme.inc.h:1537-1542): imaginary part and module too */
      normIm = sqrt (normIm / normGrid->getTotalSize ().calculateTotalCoord ());
      normMod = sqrt (normMod / normGrid->getTotalSize ().calculateTotalCoord ());
      rec.push_back ((double) normIm);
      rec.push_back ((double) (normIm * 100.0 / maxIm));
      rec.push_back ((double) normMod);
      rec.push_back ((double) (normMod * 100.0 / maxMod));
#endif
    }

    bool src = false;
    switch (grid_type)
    {
      case (static_cast<uint8_t> (GridType::EX)): src = SOLVER_SETTINGS.getDoUsePointSourceEx (); break;
      case (static_cast<uint8_t> (GridType::EY)): src = SOLVER_SETTINGS.getDoUsePointSourceEy (); break;
      case (static_cast<uint8_t> (GridType::EZ)): src = SOLVER_SETTINGS.getDoUsePointSourceEz (); break;
      case (static_cast<uint8_t> (GridType::HX)): src = SOLVER_SETTINGS.getDoUsePointSourceHx (); break;
      case (static_cast<uint8_t> (GridType::HY)): src = SOLVER_SETTINGS.getDoUsePointSourceHy (); break;
      case (static_cast<uint8_t> (GridType::HZ)): src = SOLVER_SETTINGS.getDoUsePointSourceHz (); break;
      default: break;
    }
    if (src)
    {
      is->template performPointSourceCalc<grid_type> (t);
    }
  }

  template <uint8_t grid_type>
  void fieldStepDispatch (time_step t)
  {
    bool pml = SOLVER_SETTINGS.getDoUsePML ();
    bool meta = SOLVER_SETTINGS.getDoUseMetamaterials ();
    if (pml && meta) fieldStep<grid_type, true, true> (t);
    else if (pml) fieldStep<grid_type, true, false> (t);
    else fieldStep<grid_type, false, false> (t);
  }

  void shiftGroup (Grid<TC> *f, Grid<TC> *d, Grid<TC> *d1)
  {
    f->shiftInTime ();
    if (SOLVER_SETTINGS.getDoUsePML ())
    {
      d->shiftInTime ();
    }
    if (SOLVER_SETTINGS.getDoUseMetamaterials ())
    {
      d1->shiftInTime ();
    }
  }

public:

  /* d = 3; d = 2 with the missing axis; d = 1 with the scheme's axis */
  RefRun (TC s, TC pml, TC tfL, TC tfR, int d, int axis = -1)
    : size (s), dim (d), miss (d == 2 ? axis : -1), axis1 (d == 1 ? axis : -1)
  {
    static const CoordinateType kCt[3] = { CoordinateType::X, CoordinateType::Y, CoordinateType::Z };
    ct1 = miss == 0 ? CoordinateType::Y : CoordinateType::X;
    ct2 = miss == 2 ? CoordinateType::Y : CoordinateType::Z;
    if (d == 3) ct2 = CoordinateType::Y;
    ct3 = d == 3 ? CoordinateType::Z : CoordinateType::NONE;
    if (d == 1)
    {
      ct1 = kCt[axis1];
      ct2 = CoordinateType::NONE;
    }
    layout = new Layout (s, pml, tfL, tfR,
                         SOLVER_SETTINGS.getIncidentWaveAngle1 () * PhysicsConst::Pi / 180.0,
                         SOLVER_SETTINGS.getIncidentWaveAngle2 () * PhysicsConst::Pi / 180.0,
                         SOLVER_SETTINGS.getIncidentWaveAngle3 () * PhysicsConst::Pi / 180.0,
                         SOLVER_SETTINGS.getDoUseDoubleMaterialPrecision ());
    is = new IS ();
    is->init (layout, false);
    is->initScheme (SOLVER_SETTINGS.getGridStep (), SOLVER_SETTINGS.getSourceWaveLength ());
  }

  /* The real build's CPU branch of Scheme::initCallBacks (Scheme.inc.h:2150-2430): the analytic
   * test problem's callbacks on the reference's own InternalScheme, in its priority order. */
  void initCallBacks ()
  {
#ifndef COMPLEX_FIELD_VALUES
    if (SOLVER_SETTINGS.getDoUsePolinom1BorderCondition ())
    {
      is->setCallbackEzBorder (CallBack::polinom1_ez);
      is->setCallbackHyBorder (CallBack::polinom1_hy);
    }
    else if (SOLVER_SETTINGS.getDoUsePolinom2BorderCondition ())
    {
      is->setCallbackExBorder (CallBack::polinom2_ex);
      is->setCallbackEyBorder (CallBack::polinom2_ey);
      is->setCallbackEzBorder (CallBack::polinom2_ez);
      is->setCallbackHxBorder (CallBack::polinom2_hx);
      is->setCallbackHyBorder (CallBack::polinom2_hy);
      is->setCallbackHzBorder (CallBack::polinom2_hz);
    }
    else if (SOLVER_SETTINGS.getDoUsePolinom3BorderCondition ())
    {
      is->setCallbackEzBorder (CallBack::polinom3_ez);
      is->setCallbackHyBorder (CallBack::polinom3_hy);
    }
    else if (SOLVER_SETTINGS.getDoUseSin1BorderCondition ())
    {
      is->setCallbackEzBorder (CallBack::sin1_ez);
      is->setCallbackHyBorder (CallBack::sin1_hy);
    }
    if (SOLVER_SETTINGS.getDoUsePolinom1StartValues ())
    {
      is->setCallbackEzInitial (CallBack::polinom1_ez);
      is->setCallbackHyInitial (CallBack::polinom1_hy);
    }
    else if (SOLVER_SETTINGS.getDoUsePolinom2StartValues ())
    {
      is->setCallbackExInitial (CallBack::polinom2_ex);
      is->setCallbackEyInitial (CallBack::polinom2_ey);
      is->setCallbackEzInitial (CallBack::polinom2_ez);
      is->setCallbackHxInitial (CallBack::polinom2_hx);
      is->setCallbackHyInitial (CallBack::polinom2_hy);
      is->setCallbackHzInitial (CallBack::polinom2_hz);
    }
    else if (SOLVER_SETTINGS.getDoUsePolinom3StartValues ())
    {
      is->setCallbackEzInitial (CallBack::polinom3_ez);
      is->setCallbackHyInitial (CallBack::polinom3_hy);
    }
    else if (SOLVER_SETTINGS.getDoUseSin1StartValues ())
    {
      is->setCallbackEzInitial (CallBack::sin1_ez);
      is->setCallbackHyInitial (CallBack::sin1_hy);
    }
    if (SOLVER_SETTINGS.getDoUsePolinom1RightSide ())
    {
      is->setCallbackJz (CallBack::polinom1_jz);
      is->setCallbackMy (CallBack::polinom1_my);
    }
    else if (SOLVER_SETTINGS.getDoUsePolinom2RightSide ())
    {
      is->setCallbackJx (CallBack::polinom2_jx);
      is->setCallbackJy (CallBack::polinom2_jy);
      is->setCallbackJz (CallBack::polinom2_jz);
      is->setCallbackMx (CallBack::polinom2_mx);
      is->setCallbackMy (CallBack::polinom2_my);
      is->setCallbackMz (CallBack::polinom2_mz);
    }
    else if (SOLVER_SETTINGS.getDoUsePolinom3RightSide ())
    {
      is->setCallbackJz (CallBack::polinom3_jz);
      is->setCallbackMy (CallBack::polinom3_my);
    }
    if (SOLVER_SETTINGS.getDoCalculatePolinom1DiffNorm ())
    {
      is->setCallbackEzExact (CallBack::polinom1_ez);
      is->setCallbackHyExact (CallBack::polinom1_hy);
    }
    else if (SOLVER_SETTINGS.getDoCalculatePolinom2DiffNorm ())
    {
      is->setCallbackExExact (CallBack::polinom2_ex);
      is->setCallbackEyExact (CallBack::polinom2_ey);
      is->setCallbackEzExact (CallBack::polinom2_ez);
      is->setCallbackHxExact (CallBack::polinom2_hx);
      is->setCallbackHyExact (CallBack::polinom2_hy);
      is->setCallbackHzExact (CallBack::polinom2_hz);
    }
    else if (SOLVER_SETTINGS.getDoCalculatePolinom3DiffNorm ())
    {
      is->setCallbackEzExact (CallBack::polinom3_ez);
      is->setCallbackHyExact (CallBack::polinom3_hy);
    }
    else if (SOLVER_SETTINGS.getDoCalculateSin1DiffNorm ())
    {
      is->setCallbackEzExact (CallBack::sin1_ez);
      is->setCallbackHyExact (CallBack::sin1_hy);
    }
#endif
    /* the exp exact solutions of the 1D pairs, both builds (Scheme.inc.h:2437-2760): one chain
     * per pair, exp1 before exp2 before exp3 */
    expChains ();
  }

  /* One chain per pair of Scheme::initCallBacks' exp part. */
  void expChains ()
  {
    if (SOLVER_SETTINGS.getDoCalculateExp1ExHyDiffNorm ())
    { is->setCallbackExExact (CallBack::exp1_ex_exhy); is->setCallbackHyExact (CallBack::exp1_hy_exhy); }
    else if (SOLVER_SETTINGS.getDoCalculateExp2ExHyDiffNorm ())
    { is->setCallbackExExact (CallBack::exp2_ex_exhy); is->setCallbackHyExact (CallBack::exp2_hy_exhy); }
    else if (SOLVER_SETTINGS.getDoCalculateExp3ExHyDiffNorm ())
    { is->setCallbackExExact (CallBack::exp3_ex_exhy); is->setCallbackHyExact (CallBack::exp3_hy_exhy); }
    if (SOLVER_SETTINGS.getDoCalculateExp1ExHzDiffNorm ())
    { is->setCallbackExExact (CallBack::exp1_ex_exhz); is->setCallbackHzExact (CallBack::exp1_hz_exhz); }
    else if (SOLVER_SETTINGS.getDoCalculateExp2ExHzDiffNorm ())
    { is->setCallbackExExact (CallBack::exp2_ex_exhz); is->setCallbackHzExact (CallBack::exp2_hz_exhz); }
    else if (SOLVER_SETTINGS.getDoCalculateExp3ExHzDiffNorm ())
    { is->setCallbackExExact (CallBack::exp3_ex_exhz); is->setCallbackHzExact (CallBack::exp3_hz_exhz); }
    if (SOLVER_SETTINGS.getDoCalculateExp1EyHxDiffNorm ())
    { is->setCallbackEyExact (CallBack::exp1_ey_eyhx); is->setCallbackHxExact (CallBack::exp1_hx_eyhx); }
    else if (SOLVER_SETTINGS.getDoCalculateExp2EyHxDiffNorm ())
    { is->setCallbackEyExact (CallBack::exp2_ey_eyhx); is->setCallbackHxExact (CallBack::exp2_hx_eyhx); }
    else if (SOLVER_SETTINGS.getDoCalculateExp3EyHxDiffNorm ())
    { is->setCallbackEyExact (CallBack::exp3_ey_eyhx); is->setCallbackHxExact (CallBack::exp3_hx_eyhx); }
    if (SOLVER_SETTINGS.getDoCalculateExp1EyHzDiffNorm ())
    { is->setCallbackEyExact (CallBack::exp1_ey_eyhz); is->setCallbackHzExact (CallBack::exp1_hz_eyhz); }
    else if (SOLVER_SETTINGS.getDoCalculateExp2EyHzDiffNorm ())
    { is->setCallbackEyExact (CallBack::exp2_ey_eyhz); is->setCallbackHzExact (CallBack::exp2_hz_eyhz); }
    else if (SOLVER_SETTINGS.getDoCalculateExp3EyHzDiffNorm ())
    { is->setCallbackEyExact (CallBack::exp3_ey_eyhz); is->setCallbackHzExact (CallBack::exp3_hz_eyhz); }
    if (SOLVER_SETTINGS.getDoCalculateExp1EzHxDiffNorm ())
    { is->setCallbackEzExact (CallBack::exp1_ez_ezhx); is->setCallbackHxExact (CallBack::exp1_hx_ezhx); }
    else if (SOLVER_SETTINGS.getDoCalculateExp2EzHxDiffNorm ())
    { is->setCallbackEzExact (CallBack::exp2_ez_ezhx); is->setCallbackHxExact (CallBack::exp2_hx_ezhx); }
    else if (SOLVER_SETTINGS.getDoCalculateExp3EzHxDiffNorm ())
    { is->setCallbackEzExact (CallBack::exp3_ez_ezhx); is->setCallbackHxExact (CallBack::exp3_hx_ezhx); }
    if (SOLVER_SETTINGS.getDoCalculateExp1EzHyDiffNorm ())
    { is->setCallbackEzExact (CallBack::exp1_ez_ezhy); is->setCallbackHyExact (CallBack::exp1_hy_ezhy); }
    else if (SOLVER_SETTINGS.getDoCalculateExp2EzHyDiffNorm ())
    { is->setCallbackEzExact (CallBack::exp2_ez_ezhy); is->setCallbackHyExact (CallBack::exp2_hy_ezhy); }
    else if (SOLVER_SETTINGS.getDoCalculateExp3EzHyDiffNorm ())
    { is->setCallbackEzExact (CallBack::exp3_ez_ezhy); is->setCallbackHyExact (CallBack::exp3_hy_ezhy); }
  }

  /* Real coordinate of a component's point (YeeGridLayout::get<C>CoordFP). */
  template <uint8_t grid_type>
  TCFP compCoordFP (TC posAbs)
  {
    switch (grid_type)
    {
      case (static_cast<uint8_t> (GridType::EX)): return layout->getExCoordFP (posAbs);
      case (static_cast<uint8_t> (GridType::EY)): return layout->getEyCoordFP (posAbs);
      case (static_cast<uint8_t> (GridType::EZ)): return layout->getEzCoordFP (posAbs);
      case (static_cast<uint8_t> (GridType::HX)): return layout->getHxCoordFP (posAbs);
      case (static_cast<uint8_t> (GridType::HY)): return layout->getHyCoordFP (posAbs);
      default: return layout->getHzCoordFP (posAbs);
    }
  }

  /* Scheme::initGridWithInitialVals (Scheme.inc.h:2863-2958) for one component: every cell of
   * its grid at `timestep` (s), called as Scheme::initGrids does (:3255-3279). */
  template <uint8_t grid_type>
  void initialValues (Grid<TC> *g, SourceCallBack cb, FPValue timestep)
  {
    if (!g || cb == NULLPTR) return;
    forAll (0, 0, [&] (TC p) {
      TCFP realCoord = compCoordFP<grid_type> (g->getTotalPosition (p));
      g->setFieldValue (cb (expandTo3D (realCoord * is->getGridStep (), ct1, ct2, ct3), timestep), p, 0);
    });
  }

  void initValues ()
  {
    const FPValue dt = is->getGridTimeStep ();
    if (is->getDoNeedEx ()) initialValues<static_cast<uint8_t> (GridType::EX)> (is->getEx (), is->getCallbackExInitial (), 0.5 * dt);
    if (is->getDoNeedEy ()) initialValues<static_cast<uint8_t> (GridType::EY)> (is->getEy (), is->getCallbackEyInitial (), 0.5 * dt);
    if (is->getDoNeedEz ()) initialValues<static_cast<uint8_t> (GridType::EZ)> (is->getEz (), is->getCallbackEzInitial (), 0.5 * dt);
    if (is->getDoNeedHx ()) initialValues<static_cast<uint8_t> (GridType::HX)> (is->getHx (), is->getCallbackHxInitial (), dt);
    if (is->getDoNeedHy ()) initialValues<static_cast<uint8_t> (GridType::HY)> (is->getHy (), is->getCallbackHyInitial (), dt);
    if (is->getDoNeedHz ()) initialValues<static_cast<uint8_t> (GridType::HZ)> (is->getHz (), is->getCallbackHzInitial (), dt);
  }

  void initMaterials ()
  {
    /* Scheme::initScheme: the callbacks, then Scheme::initGrids (materials, then the start values) */
    initCallBacks ();
    is->getEps ()->initialize (FieldValueHelpers::getFieldValueRealOnly (1.0));
    loadMaterial (is->getEps (), SOLVER_SETTINGS.getEpsFileName ());
    if (SOLVER_SETTINGS.getEpsSphere () != 1)
    {
      sphere (is->getEps (), SOLVER_SETTINGS.getEpsSphere (), 1.0, SOLVER_SETTINGS.getEpsSphereCenterX (),
              SOLVER_SETTINGS.getEpsSphereCenterY (), SOLVER_SETTINGS.getEpsSphereCenterZ (),
              SOLVER_SETTINGS.getEpsSphereRadius ());
    }
    if (SOLVER_SETTINGS.getUseEpsAllNorm ())
    {
      /* Scheme.inc.h:3016-3025 */
      forAllMat ([&] (TC p) { is->getEps ()->setFieldValue (FieldValueHelpers::getFieldValueRealOnly (FPValue (1.0) / PhysicsConst::Eps0), p, 0); });
    }
    is->getMu ()->initialize (FieldValueHelpers::getFieldValueRealOnly (1.0));
    loadMaterial (is->getMu (), SOLVER_SETTINGS.getMuFileName ());
    if (SOLVER_SETTINGS.getMuSphere () != 1)
    {
      sphere (is->getMu (), SOLVER_SETTINGS.getMuSphere (), 1.0, SOLVER_SETTINGS.getMuSphereCenterX (),
              SOLVER_SETTINGS.getMuSphereCenterY (), SOLVER_SETTINGS.getMuSphereCenterZ (),
              SOLVER_SETTINGS.getMuSphereRadius ());
    }
    if (SOLVER_SETTINGS.getUseMuAllNorm ())
    {
      /* Scheme.inc.h:3066-3075 */
      forAllMat ([&] (TC p) { is->getMu ()->setFieldValue (FieldValueHelpers::getFieldValueRealOnly (FPValue (1.0) / PhysicsConst::Mu0), p, 0); });
    }
    if (SOLVER_SETTINGS.getDoUseMetamaterials ())
    {
      FPValue w = 2 * PhysicsConst::Pi * is->getSourceFrequency ();
      is->getOmegaPE ()->initialize (FieldValueHelpers::getFieldValueRealOnly (0.0));
      loadMaterial (is->getOmegaPE (), SOLVER_SETTINGS.getOmegaPEFileName ());
      if (SOLVER_SETTINGS.getOmegaPESphere () != 0)
      {
        sphere (is->getOmegaPE (), SOLVER_SETTINGS.getOmegaPESphere () * w, 0.0,
                SOLVER_SETTINGS.getOmegaPESphereCenterX (), SOLVER_SETTINGS.getOmegaPESphereCenterY (),
                SOLVER_SETTINGS.getOmegaPESphereCenterZ (), SOLVER_SETTINGS.getOmegaPESphereRadius ());
      }
      is->getOmegaPM ()->initialize (FieldValueHelpers::getFieldValueRealOnly (0.0));
      loadMaterial (is->getOmegaPM (), SOLVER_SETTINGS.getOmegaPMFileName ());
      if (SOLVER_SETTINGS.getOmegaPMSphere () != 0)
      {
        sphere (is->getOmegaPM (), SOLVER_SETTINGS.getOmegaPMSphere () * w, 0.0,
                SOLVER_SETTINGS.getOmegaPMSphereCenterX (), SOLVER_SETTINGS.getOmegaPMSphereCenterY (),
                SOLVER_SETTINGS.getOmegaPMSphereCenterZ (), SOLVER_SETTINGS.getOmegaPMSphereRadius ());
      }
      is->getGammaE ()->initialize (FieldValueHelpers::getFieldValueRealOnly (0.0));
      loadMaterial (is->getGammaE (), SOLVER_SETTINGS.getGammaEFileName ());
      is->getGammaM ()->initialize (FieldValueHelpers::getFieldValueRealOnly (0.0));
      loadMaterial (is->getGammaM (), SOLVER_SETTINGS.getGammaMFileName ());
    }
    if (SOLVER_SETTINGS.getDoUsePML ())
    {
      initSigmas ();
    }
    if (SOLVER_SETTINGS.getDoUseCaCbGrids () || SOLVER_SETTINGS.getDoUseCaCbPMLGrids ()
        || SOLVER_SETTINGS.getDoUseCaCbPMLMetaGrids ())
    {
      if (is->getDoNeedEx ()) fillCoefficientsDispatch<static_cast<uint8_t> (GridType::EX)> ();
      if (is->getDoNeedEy ()) fillCoefficientsDispatch<static_cast<uint8_t> (GridType::EY)> ();
      if (is->getDoNeedEz ()) fillCoefficientsDispatch<static_cast<uint8_t> (GridType::EZ)> ();
      if (is->getDoNeedHx ()) fillCoefficientsDispatch<static_cast<uint8_t> (GridType::HX)> ();
      if (is->getDoNeedHy ()) fillCoefficientsDispatch<static_cast<uint8_t> (GridType::HY)> ();
      if (is->getDoNeedHz ()) fillCoefficientsDispatch<static_cast<uint8_t> (GridType::HZ)> ();
    }
    initValues ();
  }

  double run ()
  {
    struct timeval tv1, tv2;
    gettimeofday (&tv1, NULL);
    bool tfsf = SOLVER_SETTINGS.getDoUseTFSF ();
    bool pml = SOLVER_SETTINGS.getDoUsePML ();
    bool meta = SOLVER_SETTINGS.getDoUseMetamaterials ();
    GridCoordinate1D zero1D = GRID_COORDINATE_1D (0, CoordinateType::X);
    for (time_step t = 0; t < SOLVER_SETTINGS.getNumTimeSteps (); ++t)
    {
      if (tfsf)
      {
        is->performPlaneWaveESteps (t, zero1D, is->getEInc ()->getSize ());
        is->getEInc ()->shiftInTime ();
      }
      if (is->getDoNeedEx ()) fieldStepDispatch<static_cast<uint8_t> (GridType::EX)> (t);
      if (is->getDoNeedEy ()) fieldStepDispatch<static_cast<uint8_t> (GridType::EY)> (t);
      if (is->getDoNeedEz ()) fieldStepDispatch<static_cast<uint8_t> (GridType::EZ)> (t);
      if (is->getDoNeedEx ()) shiftGroup (is->getEx (), pml ? is->getDx () : NULLPTR, meta ? is->getD1x () : NULLPTR);
      if (is->getDoNeedEy ()) shiftGroup (is->getEy (), pml ? is->getDy () : NULLPTR, meta ? is->getD1y () : NULLPTR);
      if (is->getDoNeedEz ()) shiftGroup (is->getEz (), pml ? is->getDz () : NULLPTR, meta ? is->getD1z () : NULLPTR);
      if (tfsf)
      {
        is->performPlaneWaveHSteps (t, zero1D, is->getHInc ()->getSize ());
        is->getHInc ()->shiftInTime ();
      }
      if (is->getDoNeedHx ()) fieldStepDispatch<static_cast<uint8_t> (GridType::HX)> (t);
      if (is->getDoNeedHy ()) fieldStepDispatch<static_cast<uint8_t> (GridType::HY)> (t);
      if (is->getDoNeedHz ()) fieldStepDispatch<static_cast<uint8_t> (GridType::HZ)> (t);
      if (is->getDoNeedHx ()) shiftGroup (is->getHx (), pml ? is->getBx () : NULLPTR, meta ? is->getB1x () : NULLPTR);
      if (is->getDoNeedHy ()) shiftGroup (is->getHy (), pml ? is->getBy () : NULLPTR, meta ? is->getB1y () : NULLPTR);
      if (is->getDoNeedHz ()) shiftGroup (is->getHz (), pml ? is->getBz () : NULLPTR, meta ? is->getB1z () : NULLPTR);
    }
    gettimeofday (&tv2, NULL);
    return (double) (tv2.tv_usec - tv1.tv_usec) / 1000000 + (double) (tv2.tv_sec - tv1.tv_sec);
  }

  /* The E / H components through the reference's TXTDumper (3D runs). */
  void dumpText (const char *dir)
  {
    if (dim != 3)
    {
      return;
    }
    TC zero = TC::initAxesCoordinate (0, 0, 0, ct1, ct2, ct3);
    Grid<TC> *grids[6] = { is->getEx (), is->getEy (), is->getEz (), is->getHx (), is->getHy (), is->getHz () };
    const char *names[6] = { "Ex", "Ey", "Ez", "Hx", "Hy", "Hz" };
    for (int c = 0; c < 6; ++c)
    {
      if (!grids[c])
      {
        continue;
      }
      /* the dumper names its file after the layer it writes: entry 1 for the newest layer */
      std::vector<std::string> files (2, std::string (dir) + "/" + names[c] + ".txt");
      TXTDumper<TC> dumper;
      dumper.dumpGrid (grids[c], zero, grids[c]->getSize (), SOLVER_SETTINGS.getNumTimeSteps (), 1, files);
    }
  }

  void dump ()
  {
    bool pml = SOLVER_SETTINGS.getDoUsePML ();
    bool meta = SOLVER_SETTINGS.getDoUseMetamaterials ();
    /* after the final shift the newest values sit in time layer 1 (Grid.h:1304-1319) */
#ifdef COMPLEX_FIELD_VALUES
    /* the imaginary parts as "<name>.im" */
#define DUMP_IF(need, name, getter)                                                                         \
  if (need)                                                                                                  \
  {                                                                                                          \
    putRecord (name, dumpGrid (is->getter (), 1));                                                           \
    putRecord (std::string (name).append (".im").c_str (), dumpGrid (is->getter (), 1, 1));                    \
  }
#else
#define DUMP_IF(need, name, getter) if (need) putRecord (name, dumpGrid (is->getter (), 1));
#endif
    DUMP_IF (is->getDoNeedEx (), "Ex", getEx)
    DUMP_IF (is->getDoNeedEy (), "Ey", getEy)
    DUMP_IF (is->getDoNeedEz (), "Ez", getEz)
    DUMP_IF (is->getDoNeedHx (), "Hx", getHx)
    DUMP_IF (is->getDoNeedHy (), "Hy", getHy)
    DUMP_IF (is->getDoNeedHz (), "Hz", getHz)
    DUMP_IF (pml && is->getDoNeedEx (), "Dx", getDx)
    DUMP_IF (pml && is->getDoNeedEy (), "Dy", getDy)
    DUMP_IF (pml && is->getDoNeedEz (), "Dz", getDz)
    DUMP_IF (pml && is->getDoNeedHx (), "Bx", getBx)
    DUMP_IF (pml && is->getDoNeedHy (), "By", getBy)
    DUMP_IF (pml && is->getDoNeedHz (), "Bz", getBz)
    DUMP_IF (meta && is->getDoNeedEx (), "D1x", getD1x)
    DUMP_IF (meta && is->getDoNeedEy (), "D1y", getD1y)
    DUMP_IF (meta && is->getDoNeedEz (), "D1z", getD1z)
    DUMP_IF (meta && is->getDoNeedHx (), "B1x", getB1x)
    DUMP_IF (meta && is->getDoNeedHy (), "B1y", getB1y)
    DUMP_IF (meta && is->getDoNeedHz (), "B1z", getB1z)
#undef DUMP_IF
    putRecord ("Eps", dumpMat (is->getEps ()));
    putRecord ("Mu", dumpMat (is->getMu ()));
    {
      static const char *kNormName[6] = { "NormEx", "NormEy", "NormEz", "NormHx", "NormHy", "NormHz" };
      for (int c = 0; c < 6; ++c)
        if (!norms[c].empty ()) putRecord (kNormName[c], norms[c]);
    }
    if (pml)
    {
      if (!missing (0)) putRecord ("SigmaX", dumpMat (is->getSigmaX ()));
      if (!missing (1)) putRecord ("SigmaY", dumpMat (is->getSigmaY ()));
      if (!missing (2)) putRecord ("SigmaZ", dumpMat (is->getSigmaZ ()));
    }
    if (meta)
    {
      putRecord ("OmegaPE", dumpGrid (is->getOmegaPE (), 0));
      putRecord ("OmegaPM", dumpGrid (is->getOmegaPM (), 0));
      putRecord ("GammaE", dumpGrid (is->getGammaE (), 0));
      putRecord ("GammaM", dumpGrid (is->getGammaM (), 0));
    }
    if (SOLVER_SETTINGS.getDoUseTFSF ())
    {
      std::vector<double> e, h, ei, hi;
      for (grid_coord i = 0; i < is->getEInc ()->getSize ().get1 (); ++i)
      {
        GridCoordinate1D p = GRID_COORDINATE_1D (i, CoordinateType::X);
        e.push_back (partOf (*is->getEInc ()->getFieldValue (p, 1), 0));
        h.push_back (partOf (*is->getHInc ()->getFieldValue (p, 1), 0));
        ei.push_back (partOf (*is->getEInc ()->getFieldValue (p, 1), 1));
        hi.push_back (partOf (*is->getHInc ()->getFieldValue (p, 1), 1));
      }
      putRecord ("EInc", e);
      putRecord ("HInc", h);
#ifdef COMPLEX_FIELD_VALUES
      putRecord ("EInc.im", ei);
      putRecord ("HInc.im", hi);
#endif
    }
    std::vector<double> scal;
    scal.push_back ((double) is->getGridTimeStep ());
    scal.push_back ((double) is->getSourceFrequency ());
    putRecord ("scalars", scal);
  }
};

/* Builds, runs and dumps one RefRun with the layout --layout-type selects. */
template <SchemeType_t Type, template <typename, bool> class TCoord, LayoutType LT>
double driveLayout (TCoord<grid_coord, true> s, TCoord<grid_coord, true> p, TCoord<grid_coord, true> l,
                    TCoord<grid_coord, true> r, int dim, int axis, bool text)
{
  RefRun<Type, TCoord, LT> run (s, p, l, r, dim, axis);
  run.initMaterials ();
  double secs = run.run ();
  run.dump ();
  if (text)
  {
    if (const char *txt = getenv ("YEE_REF_TXT")) run.dumpText (txt);
  }
  return secs;
}

template <SchemeType_t Type, template <typename, bool> class TCoord>
double drive (TCoord<grid_coord, true> s, TCoord<grid_coord, true> p, TCoord<grid_coord, true> l,
              TCoord<grid_coord, true> r, int dim, int axis = -1, bool text = false)
{
  if (SOLVER_SETTINGS.getLayoutType () == H_CENTERED)
    return driveLayout<Type, TCoord, H_CENTERED> (s, p, l, r, dim, axis, text);
  return driveLayout<Type, TCoord, E_CENTERED> (s, p, l, r, dim, axis, text);
}

} /* namespace */

int main (int argc, char **argv)
{
  solverSettings.SetupFromCmd (argc, argv);
  solverSettings.Initialize ();
  if (SOLVER_SETTINGS.getDoUseCaCbPMLMetaGrids ())
  {
    fprintf (stderr, "--use-ca-cb-pml-metamaterials: the reference never binds its a1 / a2 grids\n");
    return 3;
  }
  if (SOLVER_SETTINGS.getDoUseDoubleMaterialPrecision () && SOLVER_SETTINGS.getDoUseMetamaterials ())
  {
    fprintf (stderr, "--use-double-material-precision with metamaterials: the 8-cell Drude average is UNREACHABLE\n");
    return 3;
  }
  const char *outName = getenv ("YEE_REF_OUT");
  g_out = fopen (outName ? outName : "ref_out.bin", "wb");
  if (!g_out)
  {
    return 4;
  }
  double secs = 0;
  SchemeType st = SOLVER_SETTINGS.getSchemeType ();
#ifdef MODE_DIM3
  if (st == SchemeType::Dim3)
  {
    GridCoordinate3D s = GRID_COORDINATE_3D (SOLVER_SETTINGS.getSizeX (), SOLVER_SETTINGS.getSizeY (), SOLVER_SETTINGS.getSizeZ (),
                                             CoordinateType::X, CoordinateType::Y, CoordinateType::Z);
    GridCoordinate3D p = GRID_COORDINATE_3D (SOLVER_SETTINGS.getPMLSizeX (), SOLVER_SETTINGS.getPMLSizeY (), SOLVER_SETTINGS.getPMLSizeZ (),
                                             CoordinateType::X, CoordinateType::Y, CoordinateType::Z);
    GridCoordinate3D l = GRID_COORDINATE_3D (SOLVER_SETTINGS.getTFSFSizeLeftX (), SOLVER_SETTINGS.getTFSFSizeLeftY (), SOLVER_SETTINGS.getTFSFSizeLeftZ (),
                                             CoordinateType::X, CoordinateType::Y, CoordinateType::Z);
    GridCoordinate3D r = GRID_COORDINATE_3D (SOLVER_SETTINGS.getTFSFSizeRightX (), SOLVER_SETTINGS.getTFSFSizeRightY (), SOLVER_SETTINGS.getTFSFSizeRightZ (),
                                             CoordinateType::X, CoordinateType::Y, CoordinateType::Z);
    secs = drive<static_cast<SchemeType_t> (SchemeType::Dim3), GridCoordinate3DTemplate> (s, p, l, r, 3, -1, true);
  }
  else
#endif
#define REF_RUN_2D(MODE, MISSING, GA, GB, CA, CB)                                                           \
  if (st == SchemeType::MODE)                                                                                \
  {                                                                                                          \
    GridCoordinate2D s = GRID_COORDINATE_2D (SOLVER_SETTINGS.getSize##GA (), SOLVER_SETTINGS.getSize##GB (), CA, CB); \
    GridCoordinate2D p = GRID_COORDINATE_2D (SOLVER_SETTINGS.getPMLSize##GA (), SOLVER_SETTINGS.getPMLSize##GB (), CA, CB); \
    GridCoordinate2D l = GRID_COORDINATE_2D (SOLVER_SETTINGS.getTFSFSizeLeft##GA (), SOLVER_SETTINGS.getTFSFSizeLeft##GB (), CA, CB); \
    GridCoordinate2D r = GRID_COORDINATE_2D (SOLVER_SETTINGS.getTFSFSizeRight##GA (), SOLVER_SETTINGS.getTFSFSizeRight##GB (), CA, CB); \
    secs = drive<static_cast<SchemeType_t> (SchemeType::MODE), GridCoordinate2DTemplate> (s, p, l, r, 2, MISSING); \
  }                                                                                                          \
  else
  /* the lines of main.cpp:629-800 */
#define REF_RUN_1D(MODE, AXIS, GA, CA)                                                                      \
  if (st == SchemeType::MODE)                                                                                \
  {                                                                                                          \
    GridCoordinate1D s = GRID_COORDINATE_1D (SOLVER_SETTINGS.getSize##GA (), CA);                            \
    GridCoordinate1D p = GRID_COORDINATE_1D (SOLVER_SETTINGS.getPMLSize##GA (), CA);                         \
    GridCoordinate1D l = GRID_COORDINATE_1D (SOLVER_SETTINGS.getTFSFSizeLeft##GA (), CA);                    \
    GridCoordinate1D r = GRID_COORDINATE_1D (SOLVER_SETTINGS.getTFSFSizeRight##GA (), CA);                   \
    secs = drive<static_cast<SchemeType_t> (SchemeType::MODE), GridCoordinate1DTemplate> (s, p, l, r, 1, AXIS); \
  }                                                                                                          \
  else
#ifdef MODE_EX_HY
  REF_RUN_1D (Dim1_ExHy, 2, Z, CoordinateType::Z)
#endif
#ifdef MODE_EX_HZ
  REF_RUN_1D (Dim1_ExHz, 1, Y, CoordinateType::Y)
#endif
#ifdef MODE_EY_HX
  REF_RUN_1D (Dim1_EyHx, 2, Z, CoordinateType::Z)
#endif
#ifdef MODE_EY_HZ
  REF_RUN_1D (Dim1_EyHz, 0, X, CoordinateType::X)
#endif
#ifdef MODE_EZ_HX
  REF_RUN_1D (Dim1_EzHx, 1, Y, CoordinateType::Y)
#endif
#ifdef MODE_EZ_HY
  REF_RUN_1D (Dim1_EzHy, 0, X, CoordinateType::X)
#endif
  /* the planes of main.cpp:817-1000 */
#ifdef MODE_TEX
  REF_RUN_2D (Dim2_TEx, 0, Y, Z, CoordinateType::Y, CoordinateType::Z)
#endif
#ifdef MODE_TEY
  REF_RUN_2D (Dim2_TEy, 1, X, Z, CoordinateType::X, CoordinateType::Z)
#endif
#ifdef MODE_TEZ
  REF_RUN_2D (Dim2_TEz, 2, X, Y, CoordinateType::X, CoordinateType::Y)
#endif
#ifdef MODE_TMX
  REF_RUN_2D (Dim2_TMx, 0, Y, Z, CoordinateType::Y, CoordinateType::Z)
#endif
#ifdef MODE_TMY
  REF_RUN_2D (Dim2_TMy, 1, X, Z, CoordinateType::X, CoordinateType::Z)
#endif
#ifdef MODE_TMZ
  REF_RUN_2D (Dim2_TMz, 2, X, Y, CoordinateType::X, CoordinateType::Y)
#endif
  {
    fprintf (stderr, "scheme not driven by this build\n");
    return 5;
  }
  fclose (g_out);
  printf ("Total time = %f seconds\n", secs);
  solverSettings.Uninitialize ();
  return 0;
}
