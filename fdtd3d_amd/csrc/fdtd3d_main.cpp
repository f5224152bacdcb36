// fdtd3d-compatible command line front end of the Yee engine (libyee.so).
//
// Accepts the reference's flags for the Yee time-stepping path (Source/Settings/Settings.inc.h,
// parsed as Source/Settings/Settings.cpp:133-369 does: "x:<int>,y:<int>,z:<int>" coordinates,
// --same-size* copy rules, std::stof for floating-point values, --cmd-from-file), sets the run
// up as Scheme::initScheme / initGrids do (materials from DAT files and spheres), performs
// --time-steps steps and prints the reference's "Total time = ... seconds" line
// (Source/main.cpp:1175-1211).  Result dumps follow Scheme::saveGrids
// (Source/Scheme/Scheme.inc.h:4367-4531) and the material dumps of Scheme::initGrids
// (Scheme.inc.h:3166-3200) in the text format (TXTDumper,
// Source/File-Management/Dumper/TXTDumper.h:50-130) and the binary format (DATDumper,
// Source/File-Management/Dumper/DATDumper.h:28-86) under the reference's file names
// (GridFileManager::setFileNames, Source/File-Management/Commons.h:64-122).  Options of the
// reference outside this path are either accepted without effect (logging, NTFF and other
// options that do not change the fields or the dumps) or rejected with an error (other
// schemes and physics, image dumps).
//
// Built three times: fdtd3d (FP64 fields, the reference's default value type), fdtd3d-f32 and
// fdtd3d-complex (FP64 complex fields, the reference's -DCOMPLEX_FIELD_VALUES=ON build: text
// dumps carry the real and imaginary parts, binary dumps both doubles, and --use-ntff writes the
// far-field diagrams of Scheme::saveNTFF).
#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <set>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/yee.h"

#ifndef YEE_CLI_PRECISION
#define YEE_CLI_PRECISION 8
#endif
#ifndef YEE_CLI_COMPLEX
#define YEE_CLI_COMPLEX 0
#endif

namespace {

struct Sphere
{
  double value;
  int64_t center[3] = { 0, 0, 0 };
  int64_t radius = 0;
};

// Result dump options (Settings.inc.h:199-217, 279-281).
struct SaveOptions
{
  bool res = false, interm = false, txt = false, dat = false, per_process = false, manual = false, no_pml = false;
  bool materials = false;   // --save-materials
  bool einc = false, hinc = false;
  bool scat_res = false, scat_interm = false;   // --save-scattered-field-res / --save-scattered--field-interm
  int64_t interm_step = 100;
  int64_t start[3] = { 0, 0, 0 }, end[3] = { 0, 0, 0 };
  int orth = 2;   // orthogonal axis of the default 3D dump range (--orth-axis-x/y/z)
  std::vector<std::string> refused;   // image dumps asked for
};

struct Options
{
  SaveOptions save;
  yee_config cfg {};
  std::string file[6];            // --load-<material>-from-file
  Sphere sphere[4] = { { 1.0 }, { 1.0 }, { 0.0 }, { 0.0 } };   // eps, mu, omegape, omegapm
  int accuracy = 100;
  bool stair = false;
  // analytic test problem: YEE_FN_* kinds of the start values, border values, right side and
  // exact solution (Settings.inc.h:134-153), the norm box (--norm-start / --norm-end), and
  // --eps-normed / --mu-normed
  int fn[4] = { 0, 0, 0, 0 };
  int64_t norm_start[3] = { 0, 0, 0 }, norm_end[3] = { 0, 0, 0 };
  bool eps_normed = false, mu_normed = false;
  // --calc-exp<n>-<pair>-diff-norm: YEE_FN_EXP1 .. EXP3 and the pair's YEE_SCHEME_EXHY .. EZHY
  int exp_kind = 0, exp_pair = 0;
  // near-to-far-field transform (Settings.inc.h:85-88, 124, 181-187)
  bool ntff = false, ntff_reverse = false, ntff_stdout = false;
  int64_t ntff_size[3] = { 3, 3, 3 };
  int64_t ntff_diff = 1, ntff_step = 100;
  double ntff_angle_step = 10.0;
  std::string ntff_file = "ntff-res";
};

// --use-<kind>-start-values / -border-condition / -right-side, --calc-<kind>-diff-norm: the slot
// (0 start, 1 border, 2 right side, 3 exact) and kind, or -1 for another option.  sin1 has no
// right side.
int test_flag (const std::string &a, int &kind)
{
  static const char *kinds[4] = { "polinom1", "polinom2", "polinom3", "sin1" };
  for (int k = 0; k < 4; ++k)
  {
    const std::string n = kinds[k];
    kind = k + 1;
    if (a == "--use-" + n + "-start-values") return 0;
    if (a == "--use-" + n + "-border-condition") return 1;
    if (k < 3 && a == "--use-" + n + "-right-side") return 2;
    if (a == "--calc-" + n + "-diff-norm") return 3;
  }
  return -1;
}

[[noreturn]] void die (const std::string &msg)
{
  std::fprintf (stderr, "fdtd3d: %s\n", msg.c_str ());
  std::exit (1);
}

void parse_coord (const std::string &text, int64_t def, int64_t out[3])
{
  out[0] = out[1] = out[2] = def;
  size_t start = 0;
  while (start <= text.size ())
  {
    size_t end = text.find (',', start);
    if (end == std::string::npos) end = text.size ();
    const std::string part = text.substr (start, end - start);
    const size_t colon = part.find (':');
    if (colon == std::string::npos || colon == 0) die ("bad coordinate '" + text + "'");
    const char axis = part[0];
    const int idx = axis == 'x' ? 0 : axis == 'y' ? 1 : axis == 'z' ? 2 : -1;
    if (idx < 0) die ("bad coordinate axis in '" + text + "'");
    out[idx] = std::stoll (part.substr (colon + 1));
    start = end + 1;
  }
}

// std::stof, then widened: the reference stores every FP option through stof.
double fp_option (const std::string &text)
{
  return (double) std::stof (text);
}

std::vector<std::string> read_cmd_file (const std::string &path)
{
  std::ifstream in (path);
  if (!in) die ("cannot open " + path);
  std::vector<std::string> out;
  std::string line;
  while (std::getline (in, line))
  {
    const size_t b = line.find_first_not_of (" \t\r");
    if (b == std::string::npos) continue;
    const size_t e = line.find_last_not_of (" \t\r");
    line = line.substr (b, e - b + 1);
    if (line[0] == '#' || line.rfind ("//", 0) == 0) continue;
    out.push_back (line);
  }
  return out;
}

const std::set<std::string> kIgnoredWithArg = {
  "--log-level", "--buffer-size", "--topology-size", "--cuda-buffer-size",
  "--cuda-gpus", "--num-cuda-threads", "--rebalance-step", "--stored-steps", "--save-cmd-to-file",
};
const std::set<std::string> kIgnoredFlags = {
  "--same-size-topology", "--manual-topology", "--use-cuda",
  "--palette-gray", "--palette-rgb", "--dynamic-grid", "--check-disabling-cond",
  "--check-enabling-cond",
  // precomputed coefficient grids (Scheme.inc.h:3309-3791): the same expressions as the per-cell
  // path the engine computes, so the results are the same with or without them
  "--use-ca-cb", "--use-ca-cb-pml",
};
const std::set<std::string> kUnsupported = {
  // the reference's precomputed Drude path reads a1 / a2 grids its binding never sets
  // (InternalScheme.inc.h:1902-2295, :2501-2502; Scheme.inc.h:932)
  "--use-ca-cb-pml-metamaterials",
};

Options parse (int argc, char **argv)
{
  Options o;
  yee_config &c = o.cfg;
  // reference defaults (Settings.inc.h)
  c.scheme = YEE_SCHEME_DIM3;
  c.layout = 0;
  c.precision = YEE_CLI_PRECISION;
  if (YEE_CLI_COMPLEX) c.flags |= YEE_COMPLEX_FIELDS;
  c.time_steps = 100;
  for (int a = 0; a < 3; ++a)
  {
    c.size[a] = 100;
    c.pml[a] = 2;
    c.tfsf_left[a] = 4;
    c.tfsf_right[a] = 4;
  }
  c.dx = 0.0005;
  c.wavelength = 0.02;
  c.courant = 0.5;
  c.angle_deg[0] = 90;
  c.angle_deg[1] = 0;
  c.angle_deg[2] = 90;

  std::vector<std::string> args (argv + 1, argv + argc);
  for (size_t i = 0; i < args.size (); ++i)
  {
    const std::string a = args[i];
    auto val = [&] () -> std::string {
      if (i + 1 >= args.size ()) die ("option " + a + " needs an argument");
      return args[++i];
    };
    auto copy3 = [] (int64_t v[3]) { v[1] = v[2] = v[0]; };
    static const char *comps[6] = { "ex", "ey", "ez", "hx", "hy", "hz" };
    static const char *mats[6] = { "eps", "mu", "omegape", "omegapm", "gammae", "gammam" };
    bool done = true;
    if (a == "--cmd-from-file")
    {
      std::vector<std::string> rest = read_cmd_file (val ());
      args.insert (args.begin () + (long) i + 1, rest.begin (), rest.end ());
    }
    else if (a == "--3d") c.scheme = YEE_SCHEME_DIM3;
    else if (a == "--current-source") parse_coord (val (), 0, c.current_source_pos);
    else if (a.rfind ("--current-source-", 0) == 0 && a.size () == 19 && (a[17] == 'j' || a[17] == 'm')
             && a[18] >= 'x' && a[18] <= 'z')
      c.current_source |= 1u << ((a[17] == 'j' ? 0 : 3) + (a[18] - 'x'));   // Jx..Jz on Ex..Ez, Mx..Mz on Hx..Hz
    else if (a == "--1d-exhy") c.scheme = YEE_SCHEME_EXHY;
    else if (a == "--1d-exhz") c.scheme = YEE_SCHEME_EXHZ;
    else if (a == "--1d-eyhx") c.scheme = YEE_SCHEME_EYHX;
    else if (a == "--1d-eyhz") c.scheme = YEE_SCHEME_EYHZ;
    else if (a == "--1d-ezhx") c.scheme = YEE_SCHEME_EZHX;
    else if (a == "--1d-ezhy") c.scheme = YEE_SCHEME_EZHY;
    else if (a == "--2d-tmz") c.scheme = YEE_SCHEME_TMZ;
    else if (a == "--2d-tez") c.scheme = YEE_SCHEME_TEZ;
    else if (a == "--2d-tmx") c.scheme = YEE_SCHEME_TMX;
    else if (a == "--2d-tmy") c.scheme = YEE_SCHEME_TMY;
    else if (a == "--2d-tex") c.scheme = YEE_SCHEME_TEX;
    else if (a == "--2d-tey") c.scheme = YEE_SCHEME_TEY;
    else if (a == "--size") parse_coord (val (), 100, c.size);
    else if (a == "--same-size") copy3 (c.size);
    else if (a == "--pml-size") parse_coord (val (), 2, c.pml);
    else if (a == "--same-size-pml") copy3 (c.pml);
    else if (a == "--tfsf-size-left") parse_coord (val (), 4, c.tfsf_left);
    else if (a == "--tfsf-size-right") parse_coord (val (), 4, c.tfsf_right);
    else if (a == "--same-size-tfsf")
    {
      copy3 (c.tfsf_left);
      copy3 (c.tfsf_right);
    }
    else if (a == "--time-steps") c.time_steps = std::stoi (val ());
    else if (a == "--angle-teta") c.angle_deg[0] = fp_option (val ());
    else if (a == "--angle-phi") c.angle_deg[1] = fp_option (val ());
    else if (a == "--angle-psi") c.angle_deg[2] = fp_option (val ());
    else if (a == "--dx") c.dx = fp_option (val ());
    else if (a == "--wavelength") c.wavelength = fp_option (val ());
    else if (a == "--courant-factor") c.courant = fp_option (val ());
    else if (a == "--use-tfsf") c.flags |= YEE_USE_TFSF;
    else if (a == "--use-pml") c.flags |= YEE_USE_PML;
    else if (a == "--use-metamaterials") c.flags |= YEE_USE_METAMATERIALS;
    else if (a == "--layout-type") c.layout = std::stoi (val ());
    else if (a == "--point-source") parse_coord (val (), 0, c.point_source_pos);
    else if (a == "--sphere-accuracy") o.accuracy = std::stoi (val ());
    else if (a == "--stair-sphere-approx") o.stair = true;
    else if (a == "--use-double-material-precision") o.cfg.flags |= YEE_DOUBLE_MATERIAL_PRECISION;
    else if (a == "--norm-start") parse_coord (val (), 0, o.norm_start);
    else if (a == "--norm-end") parse_coord (val (), 0, o.norm_end);
    else if (a == "--eps-normed") o.eps_normed = true;
    else if (a == "--mu-normed") o.mu_normed = true;
    else if (a.rfind ("--calc-exp", 0) == 0 && a.size () == 26 && a.compare (16, 10, "-diff-norm") == 0
             && a[10] >= '1' && a[10] <= '3' && a[11] == '-')
    {
      // --calc-exp<n>-<pair>-diff-norm (Settings.inc.h:153-175)
      static const char *pairs[6] = { "exhy", "exhz", "eyhx", "eyhz", "ezhx", "ezhy" };
      const std::string pr = a.substr (12, 4);
      int pair = -1;
      for (int q = 0; q < 6; ++q)
        if (pr == pairs[q]) pair = YEE_SCHEME_EXHY + q;
      if (pair < 0) die ("unknown option " + a);
      if (o.exp_kind && o.exp_pair != pair) die ("exact solutions of two different 1D pairs are not supported");
      const int kind = YEE_FN_EXP1 + (a[10] - '1');
      // Scheme::initCallBacks takes exp1 before exp2 before exp3 within a pair
      o.exp_kind = o.exp_kind ? std::min (o.exp_kind, kind) : kind;
      o.exp_pair = pair;
    }
    else if (int kind = 0, slot = test_flag (a, kind); slot >= 0)
    {
      // Scheme::initCallBacks takes the first of polinom1, polinom2, polinom3, sin1
      o.fn[slot] = o.fn[slot] == 0 ? kind : std::min (o.fn[slot], kind);
    }
    else if (a == "--save-res") o.save.res = true;
    else if (a == "--save-interm-res") o.save.interm = true;
    else if (a == "--interm-save-step") o.save.interm_step = std::stoll (val ());
    else if (a == "--save-as-txt") o.save.txt = true;
    else if (a == "--save-res-per-process") o.save.per_process = true;
    else if (a == "--manual-save-coords") o.save.manual = true;
    else if (a == "--save-start-coord") parse_coord (val (), 0, o.save.start);
    else if (a == "--save-end-coord") parse_coord (val (), 0, o.save.end);
    else if (a == "--save-no-pml") o.save.no_pml = true;
    else if (a == "--save-tfsf-e-incident") o.save.einc = true;
    else if (a == "--save-tfsf-h-incident") o.save.hinc = true;
    else if (a == "--orth-axis-x") o.save.orth = 0;
    else if (a == "--orth-axis-y") o.save.orth = 1;
    else if (a == "--orth-axis-z") o.save.orth = 2;
    else if (a == "--save-scattered-field-res") o.save.scat_res = true;
    else if (a == "--save-scattered--field-interm") o.save.scat_interm = true;
    else if (a == "--save-as-dat") o.save.dat = true;
    else if (a == "--save-materials") o.save.materials = true;
    else if (a == "--save-as-bmp") o.save.refused.push_back (a);
    else if (a == "--use-ntff") o.ntff = true;
    else if (a == "--ntff-reverse") o.ntff_reverse = true;
    else if (a == "--ntff-to-stdout") o.ntff_stdout = true;
    else if (a == "--ntff-size") parse_coord (val (), 3, o.ntff_size);
    else if (a == "--same-size-ntff") copy3 (o.ntff_size);
    else if (a == "--ntff-diff") o.ntff_diff = std::stoll (val ());
    else if (a == "--ntff-step-angle") o.ntff_angle_step = fp_option (val ());
    else if (a == "--ntff-filename") o.ntff_file = val ();
    else if (a == "--interm-ntff-step") o.ntff_step = std::stoll (val ());
    else done = false;
    if (done) continue;
    for (int k = 0; k < 6 && !done; ++k)
      if (a == std::string ("--point-source-") + comps[k])
      {
        c.point_source |= 1u << k;
        done = true;
      }
    for (int m = 0; m < 6 && !done; ++m)
    {
      const std::string base = std::string ("--") + mats[m];
      if (a == "--load-" + std::string (mats[m]) + "-from-file")
      {
        o.file[m] = val ();
        done = true;
      }
      else if (m < 4 && a == base + "-sphere")
      {
        o.sphere[m].value = fp_option (val ());
        done = true;
      }
      else if (m < 4 && a == base + "-sphere-center")
      {
        parse_coord (val (), 0, o.sphere[m].center);
        done = true;
      }
      else if (m < 4 && a == base + "-sphere-radius")
      {
        o.sphere[m].radius = std::stoll (val ());
        done = true;
      }
    }
    if (done) continue;
    if (kIgnoredWithArg.count (a)) (void) val ();
    else if (kIgnoredFlags.count (a)) continue;
    else if (kUnsupported.count (a)) die ("option " + a + " is outside the Yee engine (E_CENTERED, real fields)");
    else die ("unknown option " + a);
  }
  return o;
}

void check (int rc)
{
  if (rc != YEE_OK) die (yee_last_error ());
}

// Raw double array [x][y][z] of the grid's size (DATLoader, Source/Loader/DATLoader.h:62-86).
// The value type of the DAT files (FieldValue: double in the FP64 build, float in the FP32 one).
using FileValue = std::conditional<YEE_CLI_PRECISION == 4, float, double>::type;

// DATLoader::loadFromFile (Source/File-Management/Loader/DATLoader.h:55-90): N raw FieldValues in
// [x][y][z] order.
std::vector<double> load_dat (const std::string &path, int64_t n)
{
  std::ifstream in (path, std::ios::binary);
  if (!in) die ("cannot open " + path);
  std::vector<FileValue> raw ((size_t) n);
  in.read (reinterpret_cast<char *> (raw.data ()), (std::streamsize) (n * sizeof (FileValue)));
  if (in.gcount () != (std::streamsize) (n * sizeof (FileValue)))
    die (path + ": expected " + std::to_string (n) + " values of " + std::to_string (sizeof (FileValue)) + " bytes");
  return std::vector<double> (raw.begin (), raw.end ());
}

// TXTLoader::loadFromFile (Source/File-Management/Loader/TXTLoader.h:55-104, TXTLoader.cpp:25-78):
// one line per cell in [x][y][z] order (z fastest), the cell's dim indices then its value; the
// indices must name the cell the line is read for; values are read with std::stof (STOF,
// Helpers/PAssert.h:170), so they are float-rounded in either build; nothing may follow the last
// line.
std::vector<double> load_txt (const std::string &path, int dim, const int64_t size[3])
{
  std::ifstream in (path);
  if (!in) die ("cannot open " + path);
  const int64_t n = size[0] * size[1] * size[2];
  std::vector<double> v ((size_t) n);
  std::string line;
  int64_t q = 0;
  for (int64_t x = 0; x < size[0]; ++x)
    for (int64_t y = 0; y < size[1]; ++y)
      for (int64_t z = 0; z < size[2]; ++z, ++q)
      {
        if (!std::getline (in, line)) die (path + ": expected " + std::to_string (n) + " lines");
        std::istringstream ss (line);
        std::vector<std::string> tok;
        std::string w;
        while (ss >> w) tok.push_back (w);
        if ((int) tok.size () < dim + 1) die (path + ": line " + std::to_string (q + 1) + ": expected indices and a value");
        const int64_t want[3] = { x, y, z };
        for (int d = 0; d < dim; ++d)
          if (std::stoll (tok[d]) != want[d]) die (path + ": line " + std::to_string (q + 1) + ": indices do not match");
        v[(size_t) q] = (double) std::stof (tok[dim]);
      }
  in.peek ();
  if (!in.eof ()) die (path + ": more lines than cells");
  return v;
}

// Material file by extension, as GridFileManager::getFileType (File-Management/Commons.cpp:45-67)
// chooses the loader: .dat raw values, .txt text lines; .bmp images are not read.
std::vector<double> load_material (const std::string &path, int dim, const int64_t size[3])
{
  const size_t dot = path.find_last_of ('.');
  const std::string ext = dot == std::string::npos ? "" : path.substr (dot + 1);
  if (ext == "txt") return load_txt (path, dim, size);
  if (ext == "dat") return load_dat (path, size[0] * size[1] * size[2]);
  die (path + ": material files are read as .dat or .txt (the reference's BMP loader is not built)");
  return {};
}

// Minimum coordinate of a component, doubled (YeeGridLayout.cpp:1937-2043): E_CENTERED
// Ex (1, .5, .5), Ey (.5, 1, .5), Ez (.5, .5, 1), Hx (.5, 1, 1), Hy (1, .5, 1), Hz (1, 1, .5);
// H_CENTERED swaps the E and H rows.
int min_coord2 (int comp, int axis, int layout)
{
  static const int m[6][3] = { { 2, 1, 1 }, { 1, 2, 1 }, { 1, 1, 2 }, { 1, 2, 2 }, { 2, 1, 2 }, { 2, 2, 1 } };
  return m[layout == 1 ? (comp + 3) % 6 : comp][axis];
}

// Dimension of the scheme's grid.
int scheme_dim (const yee_config &c)
{
  return c.scheme == YEE_SCHEME_DIM3 ? 3 : (c.scheme >= YEE_SCHEME_TEX ? 2 : 1);
}

// The scheme's axes in the order of its coordinates: x, y, z in 3D, the two axes a 2D scheme has
// (main.cpp:817-1000: TEx/TMx y, z; TEy/TMy x, z; TEz/TMz x, y), the one of a 1D scheme
// (main.cpp:629-800: ExHy/EyHx z, ExHz/EzHx y, EyHz/EzHy x).
std::vector<int> scheme_axes (const yee_config &c)
{
  switch (c.scheme)
  {
    case YEE_SCHEME_EXHY: case YEE_SCHEME_EYHX: return { 2 };
    case YEE_SCHEME_EXHZ: case YEE_SCHEME_EZHX: return { 1 };
    case YEE_SCHEME_EYHZ: case YEE_SCHEME_EZHY: return { 0 };
    case YEE_SCHEME_TEX: case YEE_SCHEME_TMX: return { 1, 2 };
    case YEE_SCHEME_TEY: case YEE_SCHEME_TMY: return { 0, 2 };
    case YEE_SCHEME_TEZ: case YEE_SCHEME_TMZ: return { 0, 1 };
    default: return { 0, 1, 2 };
  }
}

// Whether a 2D scheme has component comp: TE<a> has H along the missing axis a and E across it,
// TM<a> the reverse (InternalScheme.inc.h:3243-3272).
bool scheme_has (const yee_config &c, int comp)
{
  if (c.scheme == YEE_SCHEME_DIM3) return true;
  if (c.scheme < YEE_SCHEME_TEX)
  {
    // a 1D scheme: the two components it names (InternalScheme.inc.h:3243-3260)
    static const int k1d[7][2] = { { -1, -1 }, { YEE_EX, YEE_HY }, { YEE_EX, YEE_HZ }, { YEE_EY, YEE_HX },
                                   { YEE_EY, YEE_HZ }, { YEE_EZ, YEE_HX }, { YEE_EZ, YEE_HY } };
    return comp == k1d[c.scheme][0] || comp == k1d[c.scheme][1];
  }
  const std::vector<int> ax = scheme_axes (c);
  const int miss = 3 - ax[0] - ax[1];
  const bool te = c.scheme <= YEE_SCHEME_TEZ, along = (comp % 3) == miss;
  return (comp >= 3) == te ? along : !along;
}

// Grid extent per scheme coordinate (1 past the scheme's dimension).
void scheme_size (const yee_config &c, int64_t size[3])
{
  const std::vector<int> ax = scheme_axes (c);
  for (int s = 0; s < 3; ++s) size[s] = s < (int) ax.size () ? c.size[ax[s]] : 1;
}

// Material grid extent per scheme coordinate: the field grid's, twice it with
// --use-double-material-precision (YeeGridLayout sizeEps, YeeGridLayout.cpp:1972).
void material_size (const yee_config &c, int64_t size[3])
{
  scheme_size (c, size);
  const int dim = c.scheme == YEE_SCHEME_DIM3 ? 3 : (c.scheme >= YEE_SCHEME_TEX ? 2 : 1);
  if (c.flags & YEE_DOUBLE_MATERIAL_PRECISION)
    for (int s = 0; s < dim; ++s) size[s] *= 2;
}

// Dump range [start, end) of a component, per scheme coordinate (Scheme::getStartCoord /
// getEndCoord, Scheme.inc.h:4647-4790; the middle plane along the orthogonal axis in 3D,
// SchemeHelper.h:860-990).  comp -1: a material grid, whose minimum coordinate is taken as 0
// there.
void dump_range (const Options &o, int comp, int dim, const int64_t size[3], int64_t start[3], int64_t end[3])
{
  const SaveOptions &sv = o.save;
  const std::vector<int> ax = scheme_axes (o.cfg);
  if (sv.manual)
  {
    // TC::initAxesCoordinate keeps the scheme's coordinates of the x/y/z options
    for (int s = 0; s < 3; ++s)
    {
      start[s] = s < dim ? sv.start[ax[s]] : 0;
      end[s] = s < dim ? sv.end[ax[s]] : 1;
    }
    return;
  }
  for (int a = 0; a < 3; ++a)
  {
    start[a] = 0;
    end[a] = size[a];
  }
  if (sv.no_pml && (o.cfg.flags & YEE_USE_PML))
  {
    // convertCoord (border - min) truncates the half-integer coordinate; start adds 1
    for (int s = 0; s < dim; ++s)
    {
      const int64_t L = o.cfg.pml[ax[s]], R = size[s] - o.cfg.pml[ax[s]];
      const int m2 = comp < 0 ? 0 : min_coord2 (comp, ax[s], o.cfg.layout);
      start[s] = (2 * L - m2) / 2 + 1;
      end[s] = (2 * R - m2) / 2;
    }
  }
  if (dim == 3)
  {
    start[sv.orth] = size[sv.orth] / 2;
    end[sv.orth] = size[sv.orth] / 2 + 1;
  }
}

// One dump of layer `layer` at time step t, file "previous-<layer>_[timestep=T]_[pid=0]_[name=N].<ext>"
// (GridFileManager::setFileNames; fields are dumped at layer 1, the layer after the step,
// materials at layer 0), of the cells [start, end) in [x][y][z] order:
//   text (TXTDumper::writeToFile): "x [y [z]] value" lines, values with 15 significant digits;
//   binary (DATDumper::writeToFile): the raw FieldValues.  The reference names its binary field
//   dumps from a list sized by the layer index and then reads the entry past its end
//   (DATDumper.h:62, 129), so it writes no usable file; here the binary dump takes the name its
//   text counterpart has, with the .dat extension.
void write_dump (bool binary, const std::string &name, int layer, int64_t t, int dim, const int64_t size[3],
                 const int64_t start[3], const int64_t end[3], const double *v, const double *im = nullptr)
{
  int64_t total = 1;
  for (int a = 0; a < dim; ++a) total *= size[a];
  std::printf ("Saving grid '%s' to %s. Time step: %d. Size: %lld (from startCoord to endCoord). \n", name.c_str (),
               binary ? "binary" : "text", layer, (long long) total);
  const std::string path = "previous-" + std::to_string (layer) + "_[timestep=" + std::to_string (t) + "]_[pid=0]_[name="
                           + name + "]" + (binary ? ".dat" : ".txt");
  FILE *f = std::fopen (path.c_str (), binary ? "wb" : "w");
  if (!f) die ("cannot write " + path);
  const int64_t nz = dim == 3 ? size[2] : 1;
  for (int64_t x = start[0]; x < end[0]; ++x)
    for (int64_t y = (dim >= 2 ? start[1] : 0); y < (dim >= 2 ? end[1] : 1); ++y)
      for (int64_t z = (dim == 3 ? start[2] : 0); z < (dim == 3 ? end[2] : 1); ++z)
      {
        const int64_t idx = (x * (dim >= 2 ? size[1] : 1) + y) * nz + z;
        const double val = v[idx];
        if (binary)
        {
          // a complex FieldValue is its real and its imaginary part
          const FileValue fv[2] = { (FileValue) val, (FileValue) (im ? im[idx] : 0.0) };
          if (std::fwrite (fv, sizeof (FileValue), im ? 2 : 1, f) != (im ? 2u : 1u)) die ("cannot write " + path);
          continue;
        }
        if (dim == 3) std::fprintf (f, "%lld %lld %lld ", (long long) x, (long long) y, (long long) z);
        else if (dim == 2) std::fprintf (f, "%lld %lld ", (long long) x, (long long) y);
        else std::fprintf (f, "%lld ", (long long) x);
        // TXTDumper::writeToFile: "real imag" in the complex build
        if (im) std::fprintf (f, "%.15g %.15g\n", val, im[idx]);
        else std::fprintf (f, "%.15g\n", val);
      }
  if (std::fclose (f) != 0) die ("cannot write " + path);
  std::printf ("Saved. \n");
}

// The dump formats asked for, in the reference's order (FILE_TYPE_DAT before FILE_TYPE_TXT).
std::vector<bool> dump_formats (const SaveOptions &sv)
{
  std::vector<bool> out;
  if (sv.dat) out.push_back (true);
  if (sv.txt) out.push_back (false);
  return out;
}

// Material dumps of Scheme::initGrids with --save-materials (Scheme.inc.h:3166-3200): Eps, Mu and,
// with metamaterials, OmegaPE, OmegaPM, GammaE, GammaM, at time step 0.
void save_materials (const Options &o, yee_handle *h)
{
  const yee_config &c = o.cfg;
  const int dim = scheme_dim (c);
  int64_t size[3];
  material_size (c, size);
  std::vector<double> v ((size_t) (size[0] * size[1] * size[2]));
  // the complex build holds materials as complex values with a zero imaginary part
  const std::vector<double> zero (YEE_CLI_COMPLEX ? v.size () : 0, 0.0);
  static const char *names[6] = { "Eps", "Mu", "OmegaPE", "OmegaPM", "GammaE", "GammaM" };
  const int count = (c.flags & YEE_USE_METAMATERIALS) ? 6 : 2;
  for (bool binary : dump_formats (o.save))
    for (int m = 0; m < count; ++m)
    {
      check (yee_get_material (h, m, v.data ()));
      int64_t start[3], end[3];
      dump_range (o, -1, dim, size, start, end);
      for (int a = 0; a < dim; ++a)
        if (start[a] < 0 || end[a] > size[a] || start[a] >= end[a]) die ("dump coordinates out of the grid");
      write_dump (binary, names[m], 0, 0, dim, size, start, end, v.data (), YEE_CLI_COMPLEX ? zero.data () : nullptr);
    }
}

// Scheme::saveGrids (time step t): per format, every component of the scheme, then EInc / HInc
// if asked.  With scattered set, the components are first made scattered
// (Scheme::gatherFieldsTotal, Scheme.inc.h:4335-4362); per-process dumps write the fields as
// they are.
void save_grids (const Options &o, yee_handle *h, int64_t t, bool scattered)
{
  const yee_config &c = o.cfg;
  const int dim = scheme_dim (c);
  int64_t size[3];
  scheme_size (c, size);
  const int64_t n = size[0] * size[1] * size[2];
  static const char *names[6] = { "Ex", "Ey", "Ez", "Hx", "Hy", "Hz" };
  // real part, and with complex fields the imaginary part (its own handle)
  yee_handle *parts[2] = { h, nullptr };
  if (YEE_CLI_COMPLEX) check (yee_get_imag_part (h, &parts[1]));
  const int nparts = YEE_CLI_COMPLEX ? 2 : 1;
  std::vector<std::vector<double>> fields[2] = { std::vector<std::vector<double>> (6), std::vector<std::vector<double>> (6) };
  for (int comp = 0; comp < 6; ++comp)
  {
    if (dim < 3 && !scheme_has (c, comp)) continue;
    for (int p = 0; p < nparts; ++p)
    {
      fields[p][comp].resize ((size_t) n);
      if (scattered && !o.save.per_process) check (yee_get_scattered_field (parts[p], comp, fields[p][comp].data ()));
      else check (yee_get_field (parts[p], comp, fields[p][comp].data ()));
    }
  }
  std::vector<double> e[2], hv[2];
  const int64_t len = 500 * size[0];   // allocateGridsInc: 500 * size.get1 ()
  const bool inc = (o.save.einc || o.save.hinc) && (c.flags & YEE_USE_TFSF);
  if (inc)
    for (int p = 0; p < nparts; ++p)
    {
      e[p].resize ((size_t) len);
      hv[p].resize ((size_t) len);
      check (yee_get_incident (parts[p], e[p].data (), hv[p].data (), len));
    }
  auto imag = [&] (const std::vector<double> &v) { return v.empty () ? nullptr : v.data (); };
  for (bool binary : dump_formats (o.save))
  {
    for (int comp = 0; comp < 6; ++comp)
    {
      if (fields[0][comp].empty ()) continue;
      int64_t start[3], end[3];
      if (o.save.per_process)
      {
        for (int a = 0; a < 3; ++a)
        {
          start[a] = 0;
          end[a] = size[a];
        }
      }
      else
      {
        dump_range (o, comp, dim, size, start, end);
      }
      for (int a = 0; a < dim; ++a)
        if (start[a] < 0 || end[a] > size[a] || start[a] >= end[a]) die ("dump coordinates out of the grid");
      write_dump (binary, names[comp], 1, t, dim, size, start, end, fields[0][comp].data (), imag (fields[1][comp]));
    }
    if (inc)
    {
      const int64_t s1[3] = { len, 1, 1 }, b[3] = { 0, 0, 0 };
      if (o.save.einc) write_dump (binary, "EInc", 1, t, 1, s1, b, s1, e[0].data (), imag (e[1]));
      if (o.save.hinc) write_dump (binary, "HInc", 1, t, 1, s1, b, s1, hv[0].data (), imag (hv[1]));
    }
  }
}

// Scheme::saveNTFF (Scheme.inc.h:4535-4640) at time step t: for each box step s of --ntff-diff,
// the box [ntff-size + s, size - ntff-size + 1 - s] and one diagram, forward (phi from 0 while
// phi <= 2 pi + pi / 180, by --ntff-step-angle degrees, added up in FPValue as the reference's loop
// does) or, with --ntff-reverse, the single angle phi of the incident wave; every angle at the
// incident wave's teta, in the file <--ntff-filename>_[timestep=t]_[step=s].txt or on stdout,
// values printed as std::ostream does by default (%g).
void save_ntff (const Options &o, yee_handle *h, int64_t t)
{
  const yee_config &c = o.cfg;
  const double pi = 3.14159265358979323846;
  const double ang[3] = { c.angle_deg[0] * pi / 180.0, c.angle_deg[1] * pi / 180.0, c.angle_deg[2] * pi / 180.0 };
  std::vector<double> phi;
  if (o.ntff_reverse)
    phi.push_back (ang[1]);
  else
  {
    const double end = 2 * pi + pi / 180, step = pi * o.ntff_angle_step / 180;
    if (!(step > 0)) die ("--ntff-step-angle must be positive");
    for (double a = 0.0; a <= end; a += step) phi.push_back (a);
  }
  std::vector<double> val (phi.size ());
  for (int64_t s = 0; s < o.ntff_diff; ++s)
  {
    int64_t left[3], right[3];
    for (int a = 0; a < 3; ++a)
    {
      left[a] = o.ntff_size[a] + s;
      right[a] = c.size[a] - o.ntff_size[a] + 1 - s;
    }
    check (yee_ntff (h, ang[0], phi.data (), (int32_t) phi.size (), left, right, val.data ()));
    FILE *f = stdout;
    std::string path;
    if (!o.ntff_stdout)
    {
      path = o.ntff_file + "_[timestep=" + std::to_string (t) + "]_[step=" + std::to_string (s) + "].txt";
      f = std::fopen (path.c_str (), "w");
      if (!f) die ("cannot write " + path);
    }
    std::fprintf (f, "==== NTFF (step=%lld) ====\n%s\n\n", (long long) s, o.ntff_reverse ? "Reverse diagram" : "Forward diagram");
    for (size_t i = 0; i < phi.size (); ++i)
      std::fprintf (f, "timestep = %lld, incident wave angle=(%g,%g,%g,), angle NTFF = %g, NTFF value = %g\n", (long long) t,
                    ang[0], ang[1], ang[2], phi[i], val[i]);
    if (f != stdout && std::fclose (f) != 0) die ("cannot write " + path);
  }
}

}  // namespace

int main (int argc, char **argv)
{
  Options o = parse (argc, argv);
  yee_config &c = o.cfg;
  int64_t ssize[3];
  scheme_size (c, ssize);
  const int64_t n = ssize[0] * ssize[1] * ssize[2];

  char info[256];
  check (yee_device_info (info, sizeof info));
  std::printf ("%s\n", info);

  yee_handle *h = nullptr;
  check (yee_create (&c, &h));
  if (o.fn[0] || o.fn[1] || o.fn[2] || o.fn[3])
    check (yee_set_test_functions (h, o.fn[0], o.fn[1], o.fn[2], o.fn[3], o.norm_start, o.norm_end));
  if (o.exp_kind)
  {
    // the dielectric border of the exp solutions: eps sphere centre - radius along the pair's line
    static const int kAxis[6] = { 2, 1, 2, 0, 1, 0 };
    const int ax = kAxis[o.exp_pair - YEE_SCHEME_EXHY];
    const double border = (double) (o.sphere[0].center[ax] - o.sphere[0].radius);
    check (yee_set_exact_exp (h, o.exp_kind, o.exp_pair, border, o.norm_start, o.norm_end));
  }
  double sc[3];
  check (yee_get_scalars (h, sc));
  const double two_pi_f = 2 * 3.14159265358979323846 * sc[1];

  // Scheme::initGrids order: file, then sphere (which replaces the grid); omega/gamma only
  // with metamaterials (Scheme.inc.h:2962-3165)
  const bool meta = (c.flags & YEE_USE_METAMATERIALS) != 0;
  for (int m = 0; m < 6; ++m)
  {
    if (m >= 2 && !meta) continue;
    if (!o.file[m].empty ())
    {
      int64_t msize[3];
      material_size (c, msize);
      std::vector<double> v = load_material (o.file[m], scheme_dim (c), msize);
      check (yee_set_material (h, m, v.data ()));
    }
    if (m < 4)
    {
      const Sphere &s = o.sphere[m];
      const double neutral = m < 2 ? 1.0 : 0.0;
      if (s.value != neutral)
      {
        const double value = m < 2 ? s.value : s.value * two_pi_f;
        check (yee_set_material_sphere (h, m, value, neutral, s.center, (double) s.radius, o.accuracy, o.stair ? 1 : 0));
      }
    }
  }
  // --eps-normed / --mu-normed: 1 / eps0, 1 / mu0 everywhere, after files and spheres
  // (Scheme.inc.h:3016-3025, 3066-3075)
  {
    const double cl = 299792458.0, mu0 = 4 * 3.14159265358979323846 * 0.0000001, eps0 = 1 / (mu0 * (cl * cl));
    int64_t msize[3];
    material_size (c, msize);
    const size_t mc = (size_t) (msize[0] * msize[1] * msize[2]);
    if (o.eps_normed) check (yee_set_material (h, YEE_MAT_EPS, std::vector<double> (mc, 1.0 / eps0).data ()));
    if (o.mu_normed) check (yee_set_material (h, YEE_MAT_MU, std::vector<double> (mc, 1.0 / mu0).data ()));
  }
  check (yee_prepare (h));

  const SaveOptions &sv = o.save;
  if (sv.materials && !sv.refused.empty ())
    die ("option " + sv.refused[0] + " is not supported (text and binary dumps: --save-as-txt, --save-as-dat)");
  if (sv.materials) save_materials (o, h);
  const bool dumping = (sv.res || sv.interm) && (sv.txt || sv.dat);
  if ((sv.res || sv.interm) && !sv.refused.empty ())
    die ("option " + sv.refused[0] + " is not supported (text and binary dumps: --save-as-txt, --save-as-dat)");
  if (sv.interm && sv.interm_step <= 0) die ("--interm-save-step must be positive");
  // Scheme.h:195-204 asserts that per-process result dumps are not scattered; the incident
  // wave that is subtracted exists only with TF/SF
  if (dumping && sv.res && sv.per_process && sv.scat_res)
    die ("--save-scattered-field-res cannot be combined with --save-res-per-process");
  if (dumping && ((sv.res && sv.scat_res) || (sv.interm && sv.scat_interm)) && !(c.flags & YEE_USE_TFSF))
    die ("scattered-field dumps need --use-tfsf");
  if (o.ntff)
  {
    // Scheme::Pointing_scat asserts the complex build (Scheme.inc.h:4096-4098)
    if (!YEE_CLI_COMPLEX)
      die ("--use-ntff: Solver is not compiled with support of complex values (use fdtd3d-complex, the "
           "-DCOMPLEX_FIELD_VALUES=ON build)");
    if (o.ntff_step <= 0) die ("--interm-ntff-step must be positive");
  }
  const auto t0 = std::chrono::steady_clock::now ();
  {
    // SchemeHelper::performNSteps3D (SchemeHelper.h:100-118): after step t, the NTFF and then the
    // intermediate grids are saved when t / step < (t + 1) / step, as time step t + 1; with an
    // exact solution every step prints the "DIFF NORM" line of every component that has one
    // (Scheme.inc.h:1543-1549), E before H as the reference computes them.  Steps between
    // these events run as one yee_step.
    static const char *names[6] = { "Ex", "Ey", "Ez", "Hx", "Hy", "Hz" };
    const bool interm = dumping && sv.interm;
    int64_t done = 0;
    while (done < c.time_steps)
    {
      const bool exact = o.fn[3] || o.exp_kind;
      int64_t next = exact ? done + 1 : c.time_steps;
      if (interm) next = std::min<int64_t> (next, (done / sv.interm_step + 1) * sv.interm_step);
      if (o.ntff) next = std::min<int64_t> (next, (done / o.ntff_step + 1) * o.ntff_step);
      const bool ntff_here = o.ntff && next % o.ntff_step == 0;
      if (ntff_here)
      {
        // the N sums average H over the last two steps: keep it one step before
        if (next - 1 > done) check (yee_step (h, (int32_t) (next - 1 - done)));
        check (yee_ntff_snapshot (h));
        check (yee_step (h, 1));
      }
      else
        check (yee_step (h, (int32_t) (next - done)));
      done = next;
      if (exact)
        for (int comp = 0; comp < 6; ++comp)
        {
          if (YEE_CLI_COMPLEX)
          {
            // the complex build's line (Scheme.inc.h:1537-1542)
            double nv[9];
            if (yee_get_diff_norm_complex (h, comp, nv) != YEE_OK) continue;
            std::printf ("-> DIFF NORM %s. Timestep %" PRId64 ". Value = ( %.20f , %.20f ) = ( %.20f %% , %.20f %% ), "
                         "module = %.20f = ( %.20f %% )\n",
                         names[comp], done - 1, nv[0], nv[1], nv[3], nv[4], nv[2], nv[5]);
            continue;
          }
          double nv[4];
          if (yee_get_diff_norm (h, comp, nv) != YEE_OK) continue;
          std::printf ("-> DIFF NORM %s. Timestep %" PRId64 ". Value = ( %.20f ) = ( %.20f %% )\n", names[comp], done - 1,
                       nv[0], nv[1]);
        }
      if (ntff_here) save_ntff (o, h, done);
      if (interm && done % sv.interm_step == 0) save_grids (o, h, done, sv.scat_interm);
    }
  }
  if (dumping && sv.res) save_grids (o, h, c.time_steps, sv.scat_res);
  check (yee_sync (h));
  const double secs = std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();
  std::printf ("Total time = %f seconds\n", secs);
  const double upd = (double) n * (double) (c.time_steps > 0 ? c.time_steps - 1 : 0);
  std::printf ("Cell updates = %.0f (%.2f Mcell-updates/s)\n", upd, secs > 0 ? upd / secs / 1e6 : 0.0);
  yee_destroy (h);
  return 0;
}
