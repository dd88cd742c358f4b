// genetIC-hip: paramfile front-end for libgic_hip.so.
//
// Keeps the reference's command surface for the field-generation path
// (genetIC/src/main.cpp:22-172) and its parser contract
// (tools/parser.hpp:40-211): case-insensitive command names, '#'/'%'
// comments, arguments parsed before the call, and
// 'Error "<what>" on line N ("<line>")' then exit 1.  Grid and flag
// bookkeeping, the zoom-window choice, the particle mapper and the writers
// are host integer work; every field operation is a gic_* call into the HIP
// library.
#include "../../include/gic.h"

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <strings.h>
#include <tuple>
#include <vector>

namespace fe {

// ---------------------------------------------------------------- logging --
// logging::entry() (tools/logging.hpp): local time stamp, then three spaces
std::ostream &log() {
  std::time_t t = std::time(nullptr);
  std::cerr << std::put_time(std::localtime(&t), " %F %T   ");
  return std::cerr;
}

// ---------------------------------------------------------------- geometry -
struct I3 {
  int x = 0, y = 0, z = 0;
  int &operator[](int i) { return i == 0 ? x : i == 1 ? y : z; }
  int operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
};

// One level's grid (simulation/grid/grid.hpp): index = (x*n + y)*n + z,
// centroid = c*dx + offset + dx/2, flags kept sorted and unique.
struct Grid {
  double periodic = 0, dx = 0, ox = 0, oy = 0, oz = 0;
  int n = 0;
  unsigned gadgetType = 1;
  std::vector<size_t> flags;
  size_t n3() const { return size_t(n) * n * n; }
  double size() const { return dx * n; }
  int simEquivalent() const { return int(std::llround(periodic / dx)); }
  bool coversFull() const { return periodic == size(); }
  I3 coord(size_t id) const {
    I3 c;
    c.x = int(id / (size_t(n) * n));
    c.y = int((id / n) % n);
    c.z = int(id % n);
    return c;
  }
  size_t index(int x, int y, int z) const { return (size_t(x) * n + y) * n + z; }
  std::array<double, 3> centroidOf(const I3 &c) const {
    return {c.x * dx + ox + dx / 2, c.y * dx + oy + dx / 2, c.z * dx + oz + dx / 2};
  }
  std::array<double, 3> centroid(size_t id) const { return centroidOf(coord(id)); }
  double wrap(double v) const {
    v = std::fmod(v, periodic);
    if (v < 0) v += periodic;
    return v;
  }
  double wrappedOffset(double a, double b) const {
    double r = a - b;
    if (r > periodic / 2) r -= periodic;
    if (r < -periodic / 2) r += periodic;
    return r;
  }
  void flagCells(const std::vector<size_t> &ids) {
    std::vector<size_t> u(ids);
    u.insert(u.end(), flags.begin(), flags.end());
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    flags.swap(u);
  }
  // getCentreWrapped (grid.hpp:466-494)
  std::array<double, 3> flaggedCentre() const {
    if (flags.empty()) throw std::runtime_error("Cannot calculate the center of an empty region");
    auto p0 = centroid(flags[0]);
    double s[3] = {0, 0, 0};
    for (size_t i = 1; i < flags.size(); ++i) {
      auto pi = centroid(flags[i]);
      for (int d = 0; d < 3; ++d) s[d] += wrappedOffset(pi[d], p0[d]);
    }
    std::array<double, 3> r;
    for (int d = 0; d < 3; ++d) r[d] = wrap(s[d] / double(flags.size()) + p0[d]);
    return r;
  }
  // insertCubeIdsIntoVector (grid.hpp:704-731): the cells whose centres lie
  // in the cube of side dxc centred on the point
  void cubeIds(const std::array<double, 3> &c, double dxc, std::vector<size_t> &out) const {
    double p[3] = {wrap(c[0] - ox), wrap(c[1] - oy), wrap(c[2] - oz)};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = int(std::floor((p[d] - dxc / 2 + dx / 2) / dx));
      hi[d] = int(std::floor((p[d] + dxc / 2 - dx / 2) / dx));
      if (lo[d] < 0 || hi[d] >= n) throw std::runtime_error("Requested zoom region falls outside high resolution grid");
    }
    for (int x = lo[0]; x <= hi[0]; ++x)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int z = lo[2]; z <= hi[2]; ++z) out.push_back(index(x, y, z));
  }
};

// Integer window with wrapping (simulation/window.hpp), kept with the
// reference's behaviour, including expandSymmetricallyToSize comparing the
// new y corner against the old z corner.
struct Window {
  int wrapLength = 0;
  I3 lo, hi;   // lower inclusive, upper exclusive
  Window() = default;
  Window(int wl, I3 p) : wrapLength(wl), lo(p) {
    hi = p;
    hi.x += 1;
    hi.y += 1;
    hi.z += 1;
  }
  int wrap(int v) const {
    if (v < 0) v += wrapLength;
    if (v >= wrapLength) v -= wrapLength;
    return v;
  }
  static bool within(int lo, int hi, int v) {
    if (hi >= lo) return v < hi && v >= lo;
    return v < hi || v >= lo;
  }
  void wrapCorners() {
    for (int d = 0; d < 3; ++d) {
      lo[d] = wrap(lo[d]);
      hi[d] = wrap(hi[d] - 1) + 1;
    }
  }
  int maxDim() const {
    int m = 0;
    for (int d = 0; d < 3; ++d) m = std::max(m, wrap(hi[d] - lo[d] - 1) + 1);
    return m;
  }
  I3 centre() const {
    I3 c;
    for (int d = 0; d < 3; ++d) c[d] = wrap(lo[d] + wrap(hi[d] - 1 - lo[d]) / 2);
    return c;
  }
  void expandToInclude(const I3 &p) {
    for (int d = 0; d < 3; ++d) {
      if (!within(lo[d], hi[d], p[d])) {
        int sizeLower = wrap(hi[d] - p[d] - 1) + 1;
        int sizeUpper = wrap(p[d] + 1 - lo[d] - 1) + 1;
        bool canLower = within(p[d], hi[d], lo[d]);
        bool canUpper = within(lo[d], p[d] + 1, hi[d] - 1);
        if (canLower && (sizeLower < sizeUpper || !canUpper)) lo[d] = p[d];
        else hi[d] = p[d] + 1;
      }
    }
  }
  void expandSymmetricallyToSize(int newSize) {
    I3 oldLo = lo;
    if (newSize == wrapLength) {
      lo = I3{};
      hi.x = hi.y = hi.z = newSize;
      return;
    }
    I3 c = centre();
    for (int d = 0; d < 3; ++d) {
      lo[d] = wrap(c[d] - newSize / 2);
      if (newSize % 2 == 0) hi[d] = wrap(c[d] + newSize / 2 - 1) + 1;
      else hi[d] = wrap(c[d] + newSize / 2) + 1;
    }
    if (lo.x > oldLo.x) {
      lo.x -= 1;
      hi.x -= 1;
    }
    if (lo.y > oldLo.z) {
      lo.y -= 1;
      hi.y -= 1;
    }
    if (lo.z > oldLo.z) {
      lo.z -= 1;
      hi.z -= 1;
    }
  }
  bool contains(int v, int d) const { return within(lo[d], hi[d], v); }
  int wrappedOffset(int a, int b) const {
    int dist = b - a;
    if (dist < -wrapLength / 2) dist += wrapLength;
    if (dist > wrapLength / 2) dist -= wrapLength;
    return dist;
  }
  void shift(const I3 &s, int sign) {
    for (int d = 0; d < 3; ++d) {
      lo[d] += sign * s[d];
      hi[d] += sign * s[d];
    }
    wrapCorners();
  }
  void moveInto(const Window &o) {
    I3 mv;
    for (int d = 0; d < 3; ++d) {
      if (!o.contains(lo[d], d)) mv[d] = wrappedOffset(lo[d], o.lo[d]);
      if (!o.contains(hi[d] - 1, d)) mv[d] = wrappedOffset(hi[d], o.hi[d]);
    }
    shift(mv, +1);
  }
};

int ratioInt(double a, double b) {
  double r = a / b;
  long ri = std::lround(r);
  if (std::fabs(r - ri) > 1e-6 || ri < 1) throw std::runtime_error("Ratio of grid spacings must be a positive integer");
  return int(ri);
}

std::vector<size_t> argsort(const std::vector<size_t> &v) {
  std::vector<size_t> i(v.size());
  std::iota(i.begin(), i.end(), 0);
  std::sort(i.begin(), i.end(), [&](size_t a, size_t b) { return v[a] < v[b]; });
  return i;
}

// ---------------------------------------------------------------- npy ------
void saveNpyTyped(const std::string &fn, int n, const std::vector<double> &d, const char *descr) {
  std::ofstream f(fn, std::ios::binary);
  if (!f) throw std::runtime_error("Can't open file " + fn);
  std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (" +
                    std::to_string(n) + ", " + std::to_string(n) + ", " + std::to_string(n) + "), }";
  size_t total = 10 + hdr.size() + 1;
  hdr += std::string((64 - total % 64) % 64, ' ') + "\n";
  unsigned short hl = (unsigned short)hdr.size();
  f.write("\x93NUMPY\x01\x00", 8);
  f.write((const char *)&hl, 2);
  f.write(hdr.data(), hdr.size());
  f.write((const char *)d.data(), d.size() * sizeof(double));
}
void saveNpy(const std::string &fn, int n, const std::vector<double> &d) { saveNpyTyped(fn, n, d, "<f8"); }
// interleaved (re, im) pairs as complex128
void saveNpyComplex(const std::string &fn, int n, const std::vector<double> &d) { saveNpyTyped(fn, n, d, "<c16"); }

// ---------------------------------------------------------------- generator
enum class Fmt { unknown = 1, gadget2 = 2, gadget3 = 3, tipsy = 4, grafic = 5, gadgethdf = 6, swift = 7 };

// One TwoLevelParticleMapper (particles/mapper/twolevelmapper.hpp) of the
// nested chain: coarse = mapper of levels [0, L-1], fine = grid L.
struct LevelMap {
  std::vector<size_t> cells;       // level1CellsToReplace: sorted cells of level L-1
  std::vector<size_t> particles;   // level1ParticlesToReplace: their ids in the coarse mapper
  std::vector<size_t> hires;       // zoomParticleArrayHiresUnsorted: level-L cells in output order
  size_t nhr = 0, size1 = 0, first2 = 0, total = 0;
};

// The particles of one level (or of the gas copy of the finest level) in
// mapper order
struct Block {
  int level = 0;
  int factor = 1;   // the grid it is laid on: > 1 subsampled, < -1 supersampled by -factor
  bool strays = false;   // on the level's box-covering grid (strays_on)
  int chain = -1;        // the mapper chain entry it is laid on (-1: the gas copy)
  bool gas = false;
  unsigned type = 1;
  std::vector<uint64_t> cells;
  size_t first = 0;   // mapper index of the first particle
  double mass = 0, eps = 0;
  std::vector<float> pos, vel;
  size_t count() const { return cells.size(); }
};

// A level as written out: a real level, or a sub/supersampled copy of one
// (the virtual grids of virtualgrid.hpp), possibly re-centred for grafic
struct VLevel {
  int src = 0;         // real level holding the data
  int factor = 1;      // > 1: subsampled by factor; < -1: supersampled by -factor
  Grid g;              // geometry of this level (after any offset)
  Grid gu;             // the same before the offset (where fields are evaluated)
  bool centred = false;
  I3 shift;            // CenteredGrid offset, in this level's cells
  // stray particles (ResolutionMatchingGrid): g covers the box at this
  // level's resolution; cells in the window [wlo, wlo + wn) (wrapped) are the
  // level's own, the others interpolated from the levels above
  bool strays = false;
  I3 wlo;
  int wn = 0;
};

struct Generator {
  gic_context *ctx = nullptr;
  bool fp32 = false;
  double Om = 0.3, Ol = 0, Ob = 0, hubble = 0.701, s8 = 0, ns = 0.96, z = 0, TCMB = 2.725;
  double epsNorm = 0.01075;
  double smoothNorm = 0;   // num_neighbours: SPH smoothing scale per cell size (read by no writer built here)
  double velOffset[3] = {0, 0, 0};
  bool haveCamb = false, drawn = false, generated = false;
  std::vector<Grid> grids;
  std::vector<std::vector<size_t>> zoomParticles;   // zoomParticleArray: per level >= 1, cells of level-1
  std::vector<LevelMap> maps;
  // the mapper's chain of particle grids, coarse to fine: each level's own
  // (or sub/supersampled, or stray-covering) grid, with autopad's padding
  // grids between levels; maps[k] (k >= 1) relates entry k to entries < k
  std::vector<VLevel> outLv;
  double x0 = 0, y0 = 0, z0 = 0;
  Fmt fmt = Fmt::unknown;
  std::string outdir = "./", outname;
  int nGadgetFiles = 1;
  unsigned mruGadgetType = 1;
  bool usedGadgetType[6] = {false, false, false, false, false, false};
  // gadget_flagged_particle_type: the finest level's flagged cells as a
  // separate block of this type, the cells taken when the mapper is rebuilt
  bool flaggedTypeOn = false;
  unsigned flaggedType = 0;
  std::vector<size_t> flaggedTypeCells;
  double varianceScale = -1;
  size_t nModifications = 0;
  int initialNumberSteps = 10;   // ic.hpp:155-161
  double precision = 0.001;

  explicit Generator(bool f32) : fp32(f32) {}
  ~Generator() {
    if (ctx) gic_destroy(ctx);
  }

  // The device context is opened by the first command that needs it, so a
  // paramfile's syntax errors are reported the same way with or without a GPU.
  gic_context *c() {
    if (!ctx) {
      if (gic_create(&ctx, 0, fp32 ? GIC_FLOAT32 : GIC_FLOAT64) != 0) {
        ctx = nullptr;
        throw std::runtime_error("No usable HIP device for libgic_hip");
      }
      pushCosmology();
    }
    return ctx;
  }
  void check(int rc) {
    if (rc) throw std::runtime_error(gic_last_error(ctx));
  }
  void pushCosmology() {
    if (ctx) check(gic_set_cosmology(c(), Om, Ol, Ob, hubble, s8, ns, z));
  }
  double a() const { return 1. / (z + 1.); }

  // io/input.hpp:33-84: whitespace-separated numbers; the column count is
  // that of the first line
  void setCamb(const std::string &fn) {
    std::ifstream f(fn);
    if (!f.is_open()) throw std::runtime_error("File " + fn + " not found");
    std::vector<double> v;
    std::string line;
    int ncols = 0;
    while (std::getline(f, line)) {
      std::stringstream ss(line);
      double t;
      int c = 0;
      while (ss >> t) {
        v.push_back(t);
        ++c;
      }
      if (ncols == 0) ncols = c;
    }
    if (ncols == 0) throw std::runtime_error("Error reading file " + fn + ". Could not determine number of columns.");
    check(gic_set_camb_table(c(), v.data(), v.size(), ncols));
    haveCamb = true;
  }
  // setPowerLawAmplitude (ic.hpp:731-734): P(k) = A k^ns instead of a CAMB table
  void setPowerLaw(double amplitude) {
    check(gic_set_power_law(c(), amplitude));
    haveCamb = true;
  }

  // --- grids ---
  void addLevel(double size, int n, double ox, double oy, double oz) {
    Grid g;
    g.periodic = grids.empty() ? size : grids[0].periodic;
    g.n = n;
    g.dx = size / n;
    g.ox = ox;
    g.oy = oy;
    g.oz = oz;
    g.gadgetType = mruGadgetType;
    if (!grids.empty()) usedGadgetType[mruGadgetType] = true;
    grids.push_back(g);
  }
  void initBaseGrid(double box, int n) {
    if (!grids.empty()) throw std::runtime_error("Cannot re-initialize the base grid");
    check(gic_init_base_grid(c(), box, n));
    addLevel(box, n, 0, 0, 0);
    log() << "Initialized the base grid:" << std::endl;
    log() << "  Box length            = " << box << " Mpc/h" << std::endl;
    log() << "  n                     = " << n << std::endl;
    log() << "  dx                    = " << box / n << std::endl;
  }

  // initZoomGridWithLowLeftCornerAt (ic.hpp:501-601)
  void initZoomGridWithLowLeftCornerAt(int x0i, int y0i, int z0i, int zoomfac, int n) {
    if (drawn) throw std::runtime_error("Trying to initialize a grid after the random field was already drawn");
    if (grids.empty()) throw std::runtime_error("Cannot initialise a zoom grid before initialising the base grid");
    if (zoomfac < 1) throw std::runtime_error("Zoom factor must be positive");
    if (allowStrays && grids.size() >= 2)
      throw std::runtime_error("strays_on with more than one zoom level is not built here");
    Grid above = grids.back();
    std::vector<size_t> untrimmed = above.flags;
    int nAbove = above.n;
    int nc = nAbove / zoomfac;
    if (x0i < 0 || y0i < 0 || z0i < 0)
      throw std::runtime_error("The newly initialised zoom level is not fully contained within the previous level");
    I3 lower{x0i, y0i, z0i};
    I3 upper{x0i + nc, y0i + nc, z0i + nc};
    if (above.coversFull()) {
      int sim = above.simEquivalent();
      for (int d = 0; d < 3; ++d) {
        int v = (upper[d] - 1) % sim;
        if (v < 0) v += sim;
        upper[d] = v + 1;
      }
    } else if (upper.x > nAbove || upper.y > nAbove || upper.z > nAbove) {
      throw std::runtime_error("The newly initialised zoom level is not fully contained within the previous level");
    }
    // keep the flagged cells whose centroid lies in the zoom window (+-eps);
    // Window<T>::wrapCorners wraps the point just inside the upper corner
    double eps = above.dx * 1e-6;
    I3 um1{upper.x - 1, upper.y - 1, upper.z - 1};
    auto wlo = above.centroidOf(lower), whi = above.centroidOf(um1);
    std::vector<size_t> trimmed;
    size_t missed = 0;
    for (size_t id : untrimmed) {
      auto p = above.centroid(id);
      bool in = true;
      for (int d = 0; d < 3; ++d) {
        double lo = wlo[d] - eps, hi = whi[d] + eps;
        lo = lo < 0 ? lo + above.periodic : (lo >= above.periodic ? lo - above.periodic : lo);
        double hin = std::nextafter(hi, -INFINITY);
        hin = hin < 0 ? hin + above.periodic : (hin >= above.periodic ? hin - above.periodic : hin);
        hi = std::nextafter(hin, INFINITY);
        in &= hi >= lo ? (p[d] < hi && p[d] >= lo) : (p[d] < hi || p[d] >= lo);
      }
      // strays_on keeps the particles outside the window too (ic.hpp:541-564)
      if (in || allowStrays) trimmed.push_back(id);
      if (!in) ++missed;
    }
    if (missed > 0) {
      log() << "WARNING: the requested zoom particles do not all fit in the requested zoom window" << std::endl;
      if (allowStrays)
        log() << "         of " << untrimmed.size() << " particles, " << missed
              << " will be interpolated from LR grid (stray particle mode)" << std::endl;
      else
        log() << "         of " << untrimmed.size() << " particles, " << missed << " have been omitted" << std::endl;
      log() << "         to make a new zoom flag list of " << trimmed.size() << std::endl;
    }
    check(gic_init_zoom_grid_with_origin_at(c(), x0i, y0i, z0i, zoomfac, n));
    zoomParticles.push_back(trimmed);
    addLevel(above.size() / zoomfac, n, above.ox + x0i * above.dx, above.oy + y0i * above.dx,
             above.oz + z0i * above.dx);
    Grid &ng = grids.back();
    log() << "Initialized a zoom region:" << std::endl;
    log() << "  Subbox length         = " << ng.size() << " Mpc/h" << std::endl;
    log() << "  n                     = " << ng.n << std::endl;
    log() << "  dx                    = " << ng.dx << std::endl;
    log() << "  Zoom factor           = " << zoomfac << std::endl;
    log() << "  Num particles         = " << untrimmed.size() << std::endl;
    log() << "  Low-left corner in parent grid = (" << x0i << ", " << y0i << ", " << z0i << ")" << std::endl;
    log() << "  Low-left corner (h**-1 Mpc)    = " << ng.ox << ", " << ng.oy << ", " << ng.oz << std::endl;
    if (fmt == Fmt::grafic) {
      // grafic output keeps no particle mapper across levels: its "particles"
      // are the cells of the grafic hierarchy (GraficMapper, graficmapper.hpp:43-72)
      size_t total = 0;
      for (auto &v : graficStack({x0, y0, z0})) total += v.g.n3();
      log() << "  Total particles       = " << total << std::endl;
    } else {
      buildMapper();
      log() << "  Total particles       = " << mapperSize(int(outLv.size()) - 1) << std::endl;
    }
    // flag the fine cells under the flagged coarse cells (makeProxyGridToMatch:
    // supersample, then the section at the fine grid's offset, wrapped)
    Grid &fine = grids.back();
    int r = ratioInt(above.dx, fine.dx);
    int sim = fine.simEquivalent();
    long off[3] = {std::lround((fine.ox - above.ox) / fine.dx), std::lround((fine.oy - above.oy) / fine.dx),
                   std::lround((fine.oz - above.oz) / fine.dx)};
    auto w = [sim](long v) {
      v %= sim;
      if (v < 0) v += sim;
      return v;
    };
    std::vector<size_t> fineFlags;
    for (size_t id : above.flags) {
      I3 c = above.coord(id);
      for (int i = 0; i < r; ++i)
        for (int j = 0; j < r; ++j)
          for (int k = 0; k < r; ++k) {
            long sx = w(long(c.x) * r + i - off[0]), sy = w(long(c.y) * r + j - off[1]),
                 sz = w(long(c.z) * r + k - off[2]);
            if (sx < fine.n && sy < fine.n && sz < fine.n) fineFlags.push_back(fine.index(int(sx), int(sy), int(sz)));
          }
    }
    fine.flagCells(fineFlags);
    if (flaggedTypeOn) {   // the mapper rebuild takes the new finest level's flags
      flaggedTypeCells = fine.flags;
      std::sort(flaggedTypeCells.begin(), flaggedTypeCells.end());
    }
  }

  // initZoomGrid (ic.hpp:398-478): smallest window around the flags,
  // expanded symmetrically to the requested size, moved inside the level
  void initZoomGrid(int zoomfac, int n) {
    if (grids.empty()) throw std::runtime_error("Cannot initialise a zoom grid before initialising the base grid");
    if (zoomfac < 1) throw std::runtime_error("Zoom factor must be positive");
    Grid &above = grids.back();
    if (above.flags.empty()) throw std::runtime_error("Cannot initialise zoom without marking particles to be replaced");
    int sim = above.simEquivalent();
    Window zw(sim, above.coord(above.flags[0]));
    for (size_t id : above.flags) zw.expandToInclude(above.coord(id));
    int nreq = zw.maxDim();
    int nuser = above.n / zoomfac;
    if (nreq > nuser && !allowStrays)
      throw std::runtime_error("Zoom particles do not fit in specified sub-box. Decrease zoom, or choose different particles");
    if (nreq < nuser) zw.expandSymmetricallyToSize(nuser);
    I3 aboveLo{int(std::llround(above.ox / above.dx)), int(std::llround(above.oy / above.dx)),
               int(std::llround(above.oz / above.dx))};
    zw.shift(aboveLo, +1);
    Window actual(sim, aboveLo);
    actual.hi = I3{aboveLo.x + above.n, aboveLo.y + above.n, aboveLo.z + above.n};
    actual.wrapCorners();
    zw.moveInto(actual);
    zw.shift(aboveLo, -1);
    // flagged cells within 3 cells of the window's faces (ic.hpp:462-473;
    // Window::containsWithBorderSafety, window.hpp:296-302)
    const int borderSafety = 3;
    for (size_t id : above.flags) {
      I3 c = above.coord(id);
      bool in = true;
      for (int d = 0; d < 3; ++d) in &= Window::within(zw.lo[d] + borderSafety, zw.hi[d] - borderSafety, c[d]);
      if (!in) {
        log() << "WARNING: Opening a zoom where flagged particles are within " << borderSafety
              << " pixels of the edge. This is prone to numerical errors." << std::endl;
        break;
      }
    }
    initZoomGridWithLowLeftCornerAt(zw.lo.x, zw.lo.y, zw.lo.z, zoomfac, n);
  }

  // --- particle mapper ---------------------------------------------------
  size_t mapperSize(int L) const { return L == 0 ? outLv[0].g.n3() : maps[L].total; }

  // getFlaggedParticles with an explicit per-level flag set, in the indexing
  // of each level's output grid (twolevelmapper.hpp:473-513)
  std::vector<size_t> flaggedParticles(int L, const std::vector<std::vector<size_t>> &f) const {
    std::vector<size_t> out;
    if (L == 0) {
      out = f[0];
      std::sort(out.begin(), out.end());
      return out;
    }
    const LevelMap &m = maps[L];
    auto g1 = flaggedParticles(L - 1, f);
    size_t zi = 0;
    for (size_t i_lr : g1) {
      while (zi < m.particles.size() && m.particles[zi] < i_lr) ++zi;
      if (zi == m.particles.size() || m.particles[zi] != i_lr) out.push_back(i_lr - zi);
    }
    std::vector<size_t> g2 = f[L];
    std::sort(g2.begin(), g2.end());
    auto si = argsort(m.hires);
    size_t k = 0;
    for (size_t i_hr : g2) {
      while (k < m.hires.size() && m.hires[si[k]] < i_hr) ++k;
      if (k == m.hires.size() || m.hires[si[k]] != i_hr) continue;
      out.push_back(si[k] + m.first2);
    }
    std::sort(out.begin(), out.end());
    return out;
  }
  std::vector<size_t> flaggedParticles() const {
    std::vector<std::vector<size_t>> f;
    for (auto &v : outLv) f.push_back(toOutputCells(v, grids[v.src].flags));
    return flaggedParticles(int(outLv.size()) - 1, f);
  }

  // flagParticles (twolevelmapper.hpp:357-466); ids ascending
  void flagParticles(int L, const std::vector<size_t> &ids) {
    if (L == 0) {
      for (size_t id : ids)
        if (id >= outLv[0].g.n3()) throw std::runtime_error("Particle ID out of range");
      grids[outLv[0].src].flagCells(toRealCells(outLv[0], ids));
      return;
    }
    const LevelMap &m = maps[L];
    size_t nz = m.particles.size();
    size_t skipped = 0, last = 0;
    std::vector<size_t> lr, hr;
    size_t firstHr = ids.size();
    for (size_t i = 0; i < ids.size(); ++i) {
      size_t p = ids[i];
      if (p < last) throw std::runtime_error("The particle list must be in ascending order");
      last = p;
      if (p < m.first2) {
        while (skipped < nz && m.particles[skipped] <= p + skipped) ++skipped;
        lr.push_back(p + skipped);
      } else {
        firstHr = i;
        break;
      }
    }
    const Grid &gc = outLv[L - 1].g;
    const Grid &gf = outLv[L].g;
    size_t lastLr = SIZE_MAX;
    std::vector<size_t> cube;
    for (size_t i = firstHr; i < ids.size(); ++i) {
      size_t rel = ids[i] - m.first2;
      size_t lri = rel / m.nhr;
      if (lri >= nz) throw std::runtime_error("Particle ID out of range");
      if (lastLr != m.particles[lri]) {
        lastLr = m.particles[lri];
        lr.push_back(lastLr);
        cube.clear();
        gf.cubeIds(gc.centroid(m.cells[lri]), gc.dx, cube);
      }
      hr.push_back(cube[rel % m.nhr]);
    }
    std::sort(lr.begin(), lr.end());
    std::sort(hr.begin(), hr.end());
    flagParticles(L - 1, lr);
    grids[outLv[L].src].flagCells(toRealCells(outLv[L], hr));
  }

  // TwoLevelParticleMapper constructor (twolevelmapper.hpp:270-320), per level
  // setSubsample / setSupersample (ic.hpp:296-330) rebuild the particle
  // mapper at once; with particle output that needs the resampled mappers
  // setSubsample / setSupersample / setSupersampleGas (ic.hpp:305-336):
  // the particle mapper is rebuilt at once, so a zoom that no longer fits
  // its grids after resampling is reported here
  void setSubsample(int f) {
    if (f <= 0) throw std::runtime_error("Subsampling factor must be greater then zero");
    subsample = f;
    refreshMapper();
  }
  void setSupersample(int f) {
    if (f <= 1) throw std::runtime_error("Supersampling factor must be greater then one");
    supersample = f;
    refreshMapper();
  }
  void setSupersampleGas(int f) {
    if (f <= 1) throw std::runtime_error("Supersampling factor must be greater then one");
    supersampleGas = f;
    refreshMapper();
  }
  void refreshMapper() {
    if (grids.empty() || fmt == Fmt::grafic) return;
    buildMapper();
    if (Ob > 0 && supersampleGas > 1) gasCells(resampledLevel(int(grids.size()) - 1, -supersampleGas));
  }

  // The grid the particles of level l are laid on (updateParticleMapper and
  // superOrSubSample, ic.hpp:1129-1236): the finest level supersampled by
  // `super` (SuperSampleGrid, virtualgrid.hpp:126-208: cells 1/super the
  // size, mass / super^3, the level's own softening), then the base level
  // subsampled by `sub` (SubSampleGrid, virtualgrid.hpp:526-647: n / sub
  // cells, mass * sub^3, softening of the larger cell).  On one level the two
  // combine: a supersampled grid subsampled by the same factor is the level
  // itself, by a larger one a subsampled level (virtualgrid.hpp:180-196).
  VLevel outputLevel(int l, int super, int sub) const {
    int sup = l == int(grids.size()) - 1 ? super : 1;
    int sb = l == 0 ? sub : 1;
    int f = 1;
    if (sup > 1 && sb > 1) {
      if ((sb > sup ? sb % sup : sup % sb) != 0) throw std::runtime_error("Ratio is not an integer");
      f = sb == sup ? 1 : sb > sup ? sb / sup : -(sup / sb);
    } else if (sb > 1) {
      f = sb;
    } else if (sup > 1) {
      f = -sup;
    }
    VLevel hi = resampledLevel(l, f);
    if (!allowStrays || l == 0) return hi;
    // strays_on: the zoom level's particles lie on a grid covering the box
    // at its resolution (ResolutionMatchingGrid, virtualgrid.hpp:217-400,
    // made by MultiLevelGrid::addLevel, multilevelgrid.hpp:130-134, over the
    // base level; sub/supersampling applies to both of its grids): the
    // level's own cells inside its window, interpolated cells outside
    const Grid &g0 = grids[0];
    VLevel r = hi;
    r.strays = true;
    r.g.n = int(std::llround(g0.periodic / hi.g.dx));
    r.g.ox = g0.ox;
    r.g.oy = g0.oy;
    r.g.oz = g0.oz;
    r.wlo = I3{int(std::llround((hi.g.ox - g0.ox) / hi.g.dx)), int(std::llround((hi.g.oy - g0.oy) / hi.g.dx)),
               int(std::llround((hi.g.oz - g0.oz) / hi.g.dx))};
    r.wn = hi.g.n;
    r.gu = r.g;
    return r;
  }
  // Cell i of a stray-covering level inside the level's window: its index on
  // the level's own (resampled) grid, or SIZE_MAX outside the window
  size_t strayInside(const VLevel &v, size_t i) const {
    I3 c = v.g.coord(i);
    int n = v.g.n;
    int lx = ((c.x - v.wlo.x) % n + n) % n, ly = ((c.y - v.wlo.y) % n + n) % n, lz = ((c.z - v.wlo.z) % n + n) % n;
    if (lx >= v.wn || ly >= v.wn || lz >= v.wn) return SIZE_MAX;
    return (size_t(lx) * v.wn + ly) * v.wn + lz;
  }
  // Field value at cell i of an output level, f(l) giving real level l's
  // field: the level's own value (through its sub/supersampling), or for a
  // stray cell the tricubic interpolation of the level above at the cell's
  // centroid (ResolutionMatchingEvaluator, evaluator.hpp:216-256, over the
  // SuperSampleEvaluator of the base level)
  template <class F> double evalOut(const VLevel &v, const VLevel &hi, F f, size_t i) const {
    if (!v.strays) return v.factor == 1 ? f(v.src)[i] : evalVirtual(v, f(v.src), i);
    size_t h = strayInside(v, i);
    if (h != SIZE_MAX) return hi.factor == 1 ? f(v.src)[h] : evalVirtual(hi, f(v.src), h);
    return interpolate(v.src - 1, v.g.centroid(i), f(v.src - 1));
  }
  // The output level a block of particles is laid on
  VLevel blockLevel(const Block &b) const {
    return b.chain >= 0 ? outLv[b.chain] : resampledLevel(b.level, b.factor);
  }
  // Real level l subsampled (f > 1) or supersampled (f < -1) by |f|
  VLevel resampledLevel(int l, int f) const {
    VLevel v;
    v.src = l;
    v.factor = f;
    v.g = grids[l];
    v.g.flags.clear();
    if (f > 1) {
      v.g.n = grids[l].n / f;
      v.g.dx = grids[l].dx * f;
    } else if (f < -1) {
      v.g.n = grids[l].n * -f;
      v.g.dx = grids[l].dx / -f;
    }
    v.gu = v.g;
    return v;
  }
  // Flags of real level v.src as cells of its output grid (the virtual
  // grids' getFlaggedCells, virtualgrid.hpp:155-160,556-561): a subsampled
  // cell is flagged when any real cell under it is, a supersampled level
  // flags every subcell of a flagged cell.
  std::vector<size_t> toOutputCells(const VLevel &v, const std::vector<size_t> &cells) const {
    std::vector<size_t> out;
    if (v.strays) {   // the level's cells, shifted to their place on the box-covering grid
      VLevel hi = resampledLevel(v.src, v.factor);
      int n = v.g.n;
      for (size_t c : toOutputCells(hi, cells)) {
        I3 q = hi.g.coord(c);
        out.push_back(v.g.index((q.x + v.wlo.x) % n, (q.y + v.wlo.y) % n, (q.z + v.wlo.z) % n));
      }
      std::sort(out.begin(), out.end());
      return out;
    }
    const Grid &rg = grids[v.src];
    if (v.factor == 1) {
      out = cells;
    } else if (v.factor > 1) {
      int f = v.factor;
      for (size_t c : cells) {
        I3 q = rg.coord(c);
        if (q.x / f < v.g.n && q.y / f < v.g.n && q.z / f < v.g.n) out.push_back(v.g.index(q.x / f, q.y / f, q.z / f));
      }
    } else {
      int f = -v.factor;
      for (size_t c : cells) {
        I3 q = rg.coord(c);
        for (int i = 0; i < f; ++i)
          for (int j = 0; j < f; ++j)
            for (int k = 0; k < f; ++k) out.push_back(v.g.index(q.x * f + i, q.y * f + j, q.z * f + k));
      }
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  }
  // ...and output cells as the real cells they flag (the virtual grids'
  // flagCells): every real cell under a subsampled cell, the real cell of a
  // supersampled one.
  std::vector<size_t> toRealCells(const VLevel &v, const std::vector<size_t> &cells) const {
    if (v.strays) {   // cells outside the window flag nothing on the level
      VLevel hi = resampledLevel(v.src, v.factor);
      std::vector<size_t> in;
      for (size_t c : cells) {
        size_t h = strayInside(v, c);
        if (h != SIZE_MAX) in.push_back(h);
      }
      std::sort(in.begin(), in.end());
      return toRealCells(hi, in);
    }
    if (v.factor == 1) return cells;
    std::vector<size_t> out;
    const Grid &rg = grids[v.src];
    if (v.factor > 1) {
      int f = v.factor;
      for (size_t c : cells) {
        I3 q = v.g.coord(c);
        for (int x = q.x * f; x < std::min(q.x * f + f, rg.n); ++x)
          for (int y = q.y * f; y < std::min(q.y * f + f, rg.n); ++y)
            for (int z = q.z * f; z < std::min(q.z * f + f, rg.n); ++z) out.push_back(rg.index(x, y, z));
      }
    } else {
      int f = -v.factor;
      for (size_t c : cells) {
        I3 q = v.g.coord(c);
        out.push_back(rg.index(q.x / f, q.y / f, q.z / f));
      }
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
  }

  // The cells of a zoom level's output grid that replace each zoomed coarse
  // cell, in mapper order: the cube under every replaced cell of the coarse
  // output grid (zoomParticleArrayHiresUnsorted, twolevelmapper.hpp:270-320);
  // after resampling the cubes must still fit the fine grid
  // (superOrSubSample, twolevelmapper.hpp:582-610).
  void zoomCubes(const Grid &gc, const std::vector<size_t> &coarse, const Grid &gf, std::vector<size_t> &out) const {
    try {
      for (size_t cid : coarse) gf.cubeIds(gc.centroid(cid), gc.dx, out);
    } catch (std::runtime_error &) {
      if (subsample > 1 || supersample > 1 || supersampleGas > 1)
        throw std::runtime_error(
            "Following sub/super-sampling, the zoom particles no longer fit fully on their grids. Try rerunning with "
            "strays_on, or (better) increase the size of the zoom region.");
      throw;
    }
  }

  void buildMapper() {
    // the chain: one entry per level, each level's zoomed cells read through
    // the output grid of the level above
    outLv.assign(1, outputLevel(0, supersample, subsample));
    std::vector<std::vector<size_t>> zc(1);   // zc[k]: cells of entry k-1 that entry k replaces
    for (size_t l = 1; l < grids.size(); ++l) {
      zc.push_back(toOutputCells(outLv.back(), zoomParticles[l - 1]));
      outLv.push_back(outputLevel(int(l), supersample, subsample));
    }
    if (autopad > 0 && grids.size() > 1) insertPadding(zc);
    size_t K = outLv.size();
    maps.assign(K, LevelMap());
    for (size_t L = 1; L < K; ++L) {
      LevelMap &m = maps[L];
      const Grid &gc = outLv[L - 1].g, &gf = outLv[L].g;
      m.cells = zc[L];
      std::vector<std::vector<size_t>> f(L);
      f[L - 1] = m.cells;
      m.particles = flaggedParticles(int(L) - 1, f);
      if (m.particles.size() != m.cells.size())
        throw std::runtime_error("Consistency error when relating cells to particles");
      int r = ratioInt(gc.dx, gf.dx);
      m.nhr = size_t(r) * r * r;
      m.size1 = mapperSize(int(L) - 1);
      m.total = m.size1 + (m.nhr - 1) * m.particles.size();
      m.first2 = m.size1 - m.particles.size();
      m.hires.clear();
      m.hires.reserve(m.cells.size() * m.nhr);
      zoomCubes(gc, m.cells, gf, m.hires);
    }
  }

  // The supersampled-by-2 version of an output level (makeSupersampled,
  // virtualgrid.hpp:176-200,610-630: a subsampled level is un-subsampled
  // first)
  VLevel superBy2(const VLevel &c) const {
    if (c.strays) throw std::runtime_error("autopad with strays_on is not built here");
    int f = c.factor;
    int nf = f == 1 ? -2 : f < -1 ? 2 * f : (f == 2 ? 1 : f / 2);
    if (f > 2 && f % 2) throw std::runtime_error("Ratio is not an integer");
    return resampledLevel(c.src, nf);
  }
  // The 8 subcells on p (c supersampled by 2) of cells of c
  static std::vector<size_t> subcells2(const Grid &c, const Grid &p, const std::vector<size_t> &cells) {
    std::vector<size_t> out;
    out.reserve(cells.size() * 8);
    for (size_t id : cells) {
      I3 q = c.coord(id);
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          for (int k = 0; k < 2; ++k) out.push_back(p.index(2 * q.x + i, 2 * q.y + j, 2 * q.z + k));
    }
    std::sort(out.begin(), out.end());
    return out;
  }
  // Grid::getFlaggedCellsSize (grid.hpp:318-329) of a cell list
  static int cellsSize(const Grid &g, const std::vector<size_t> &cells) {
    if (cells.empty()) return 0;
    Window w(g.simEquivalent(), g.coord(cells[0]));
    for (size_t id : cells) w.expandToInclude(g.coord(id));
    return w.maxDim();
  }
  // autopad (TwoLevelParticleMapper::insertIntermediateResolutionPadding,
  // twolevelmapper.hpp:620-735, with ratio 2): where a level refines the one
  // above by more than 2, the coarse cells it replaces, grown by autopad
  // cells (and kept to cells that are particles), are replaced by the coarse
  // level supersampled by 2; its cells under the zoom are replaced by the
  // fine level, padding again until the steps are 2.
  void insertPadding(std::vector<std::vector<size_t>> &zc) {
    std::vector<VLevel> ch{outLv[0]};
    std::vector<std::vector<size_t>> z(1);
    for (size_t k = 1; k < outLv.size(); ++k) {
      const VLevel &F = outLv[k];
      std::vector<size_t> Z = zc[k];
      // the coarse entry's cells that are particles: all of the base, the
      // zoom cubes of a finer entry
      bool all = ch.size() == 1;
      std::vector<size_t> access;
      if (!all) {
        zoomCubes(ch[ch.size() - 2].g, z.back(), ch.back().g, access);
        std::sort(access.begin(), access.end());
      }
      while (ratioInt(ch.back().g.dx, F.g.dx) > 2) {
        VLevel C = ch.back();
        VLevel P = superBy2(C);
        std::vector<size_t> S = subcells2(C.g, P.g, Z);
        std::vector<size_t> E = Z;
        dilate(E, C.g.n, int(autopad));
        if (!all) {
          std::vector<size_t> keep;
          std::set_intersection(E.begin(), E.end(), access.begin(), access.end(), std::back_inserter(keep));
          E.swap(keep);
        }
        log() << "Added a padding region of effective resolution " << P.g.simEquivalent() << " of physical size ~"
              << cellsSize(C.g, E) * C.g.dx << "Mpc/h (inner resolution " << F.g.simEquivalent() << ", size ~"
              << cellsSize(P.g, S) * P.g.dx << "Mpc/h)" << std::endl;
        z.push_back(E);
        ch.push_back(P);
        access = subcells2(C.g, P.g, E);
        all = false;
        Z = S;
      }
      z.push_back(Z);
      ch.push_back(F);
    }
    outLv = ch;
    zc = z;
  }

  // Cells of each level in output order: each two-level mapper emits the
  // coarse mapper's particles minus the replaced ones, then the zoom cubes
  // (twolevelmapper.hpp:252-319,869-892).
  std::vector<std::vector<uint64_t>> levelOrder() const {
    size_t nl = outLv.size();
    std::vector<std::vector<uint64_t>> lists(1);
    lists[0].resize(outLv[0].g.n3());
    std::iota(lists[0].begin(), lists[0].end(), uint64_t(0));
    for (size_t L = 1; L < nl; ++L) {
      const LevelMap &m = maps[L];
      size_t p = 0, k = 0;
      for (auto &lst : lists) {
        std::vector<uint64_t> kept;
        kept.reserve(lst.size());
        for (uint64_t c : lst) {
          while (k < m.particles.size() && m.particles[k] < p) ++k;
          if (!(k < m.particles.size() && m.particles[k] == p)) kept.push_back(c);
          ++p;
        }
        lst.swap(kept);
      }
      lists.emplace_back(m.hires.begin(), m.hires.end());
    }
    return lists;
  }

  // --- selection (ic.hpp:1442-1523) ---
  void select(const std::function<bool(double, double, double)> &inc) {
    if (grids.empty()) throw std::runtime_error("No grid has yet been initialised");
    for (auto &g : grids) g.flags.clear();
    for (auto &g : grids) {
      std::vector<size_t> ids;
      for (size_t i = 0; i < g.n3(); ++i) {
        auto p = g.centroid(i);
        double dx = grids[0].wrappedOffset(p[0], x0), dy = grids[0].wrappedOffset(p[1], y0),
               dz = grids[0].wrappedOffset(p[2], z0);
        if (inc(dx, dy, dz)) ids.push_back(i);
      }
      g.flagCells(ids);
    }
  }
  void selectSphere(float radius) {
    double r2 = radius * radius;   // float product, as in the reference
    select([r2](double a, double b, double c) { return a * a + b * b + c * c < r2; });
  }
  void selectCube(float side) {
    double h = side / 2;
    select([h](double a, double b, double c) { return std::fabs(a) < h && std::fabs(b) < h && std::fabs(c) < h; });
  }
  void selectEllipse(float a, float b, float c) {
    double a2 = a * a, b2 = b * b, c2 = c * c;
    select([=](double x, double y, double z) { return x * x / a2 + y * y / b2 + z * z / c2 < 1.; });
  }
  void selectNearest() {
    if (grids.empty()) throw std::runtime_error("No grid has yet been initialised");
    for (auto &g : grids) g.flags.clear();
    Grid &g = grids.back();
    double p[3] = {x0 - g.ox, y0 - g.oy, z0 - g.oz};
    int c[3];
    for (int d = 0; d < 3; ++d) {
      c[d] = int(std::floor(g.wrap(p[d]) / g.dx));
      if (c[d] < 0 || c[d] >= g.n) throw std::runtime_error("Point is outside the deepest grid");
    }
    g.flagCells({g.index(c[0], c[1], c[2])});
  }
  void centreOnFlagged() {
    int level = -1;
    for (int l = 0; l < int(grids.size()); ++l)
      if (!grids[l].flags.empty()) level = l;
    if (level < 0) throw std::runtime_error("Cannot calculate the center of an empty region");
    auto c = grids[level].flaggedCentre();
    x0 = c[0];
    y0 = c[1];
    z0 = c[2];
    log() << "Centre of region is " << std::setprecision(12) << x0 << " " << y0 << " " << z0 << std::setprecision(6)
          << std::endl;
  }
  // loadID / appendID (ic.hpp:1390-1428)
  void loadIDFile(const std::string &fn, bool append) {
    std::ifstream f(fn);
    if (!f.is_open()) throw std::runtime_error("File " + fn + " not found");
    std::vector<size_t> ids;
    size_t v;
    while (f >> v) ids.push_back(v);
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    if (!append)
      for (auto &g : grids) g.flags.clear();
    if (fmt == Fmt::grafic) {
      graficFlagParticles(ids);
    } else {
      buildMapper();
      flagParticles(int(outLv.size()) - 1, ids);
    }
    centreOnFlagged();
  }
  void dumpIDFile(const std::string &fn) {
    std::vector<size_t> ids;
    if (fmt == Fmt::grafic) {
      ids = graficFlaggedParticles();
    } else {
      buildMapper();
      ids = flaggedParticles();
    }
    std::ofstream f(fn);
    for (size_t id : ids) f << id << std::endl;
  }
  // setFlaggedGadgetParticleType (ic.hpp:706-714): the flagged particles of
  // the finest level are split off into their own gadget type, behind every
  // other particle (a two-level mapper whose fine grid is the finest grid
  // itself, ic.hpp:1180-1190)
  void setFlaggedGadgetParticleType(unsigned t) {
    if (t >= 6) throw std::runtime_error("Gadget particle type must be between 0 and 5");
    if (usedGadgetType[t])
      throw std::runtime_error(
          "Cannot reuse a gadget particle type; this type was already used in a different section of the ICs");
    if (grids.empty()) throw std::runtime_error("No grid has yet been initialised");
    flaggedTypeOn = true;
    flaggedType = t;
    usedGadgetType[t] = true;
    flaggedTypeCells = grids.back().flags;
    std::sort(flaggedTypeCells.begin(), flaggedTypeCells.end());
  }

  void setGadgetParticleType(unsigned t) {
    if (t >= 6) throw std::runtime_error("Gadget particle type must be between 0 and 5");
    if (t != mruGadgetType) {
      if (usedGadgetType[t])
        throw std::runtime_error(
            "Cannot reuse a gadget particle type; this type was already used in a different section of the ICs");
      mruGadgetType = t;
    }
    if (grids.empty())
      throw std::runtime_error("Can't set a gadget particle type until a level has been initialised to set it for");
    grids.back().gadgetType = t;
  }

  // --- generation ---
  void ensureDrawn() {
    if (grids.empty()) throw std::runtime_error("No grid has yet been initialised");
    if (!drawn) {
      check(gic_draw(c()));
      drawn = true;
    }
  }
  // --- white-noise adjustments (ic.hpp:669-671, 805-820, 1652-1706) ---
  void reverse() {
    ensureDrawn();
    check(gic_reverse(c()));
  }
  // splice (ic.hpp:1657-1686): the flagged cells take a new seed's field
  void splice(int newSeed) {
    ensureDrawn();
    syncFlags();
    if (grids.size() > 1)
      log() << "WARNING: Splicing operations on multiple levels work, but may have edge artefacts. Use "
               "experimentally/at your own risk!"
            << std::endl;
    log() << "Constructing new random field for exterior of splice" << std::endl;
    int iterations = 0;
    check(gic_splice(c(), (unsigned long)newSeed, &iterations));
    log() << "Finished constructing new random field. Beginning splice operation." << std::endl;
    log() << "Conjugate gradient ended after " << iterations << " iterations" << std::endl;
  }
  void reverseSmallK(double kmax) {
    ensureDrawn();
    check(gic_reverse_small_k(c(), kmax));
  }
  void importLevel(int level, const std::string &fn) {
    ensureDrawn();
    log() << "Importing random field on level " << level << " from " << fn << std::endl;
    if (level < 0 || level >= int(grids.size())) throw std::runtime_error("Specified level does not exist.");
    int n = grids[level].n;
    std::vector<double> d = loadNpy(fn, n);
    check(gic_import_level(c(), level, d.data(), d.size()));
    log() << "... success!" << std::endl;
  }
  // io::numpy::LoadArrayFromNumpy: a C-ordered n x n x n array of float64 or float32
  static std::vector<double> loadNpy(const std::string &fn, int n) {
    std::ifstream f(fn, std::ios::binary);
    if (!f) throw std::runtime_error("Unable to open file " + fn);
    char magic[6];
    f.read(magic, 6);
    if (!f || std::string(magic, 6) != "\x93NUMPY") throw std::runtime_error(fn + " is not a numpy file");
    unsigned char ver[2];
    f.read((char *)ver, 2);
    uint32_t hl = 0;
    if (ver[0] == 1) {
      uint16_t h16;
      f.read((char *)&h16, 2);
      hl = h16;
    } else {
      f.read((char *)&hl, 4);
    }
    std::string hdr(hl, ' ');
    f.read(&hdr[0], hl);
    bool f8 = hdr.find("'<f8'") != std::string::npos, f4 = hdr.find("'<f4'") != std::string::npos;
    if ((!f8 && !f4) || hdr.find("'fortran_order': True") != std::string::npos)
      throw std::runtime_error("Unsupported numpy array in " + fn);
    size_t p = hdr.find("'shape':");
    int dims[3] = {0, 0, 0}, nd = 0;
    if (p != std::string::npos) {
      size_t q = hdr.find('(', p), e = hdr.find(')', q);
      std::string sh = hdr.substr(q + 1, e - q - 1);
      std::stringstream ss(sh);
      std::string tok;
      while (std::getline(ss, tok, ',') && nd < 3)
        if (tok.find_first_of("0123456789") != std::string::npos) dims[nd++] = std::stoi(tok);
    }
    if (nd != 3 || dims[0] != n || dims[1] != n || dims[2] != n)
      throw std::runtime_error("Incorrect size for imported numpy array");
    size_t n3 = size_t(n) * n * n;
    std::vector<double> d(n3);
    if (f8) {
      f.read((char *)d.data(), n3 * 8);
    } else {
      std::vector<float> t(n3);
      f.read((char *)t.data(), n3 * 4);
      for (size_t i = 0; i < n3; ++i) d[i] = t[i];
    }
    if (!f) throw std::runtime_error("Truncated numpy file " + fn);
    return d;
  }
  // dumpGridFourier (ic.hpp:926-936): the full complex spectrum n x n x n
  // (kz over all n, the upper half from Hermitian symmetry) as complex128
  void dumpGridFourier(int level) {
    if (level < 0 || level >= int(grids.size())) throw std::runtime_error("Specified level does not exist.");
    ensureDrawn();
    int n = grids[level].n, nz = n / 2 + 1;
    std::vector<double> h(size_t(n) * n * nz * 2);
    check(gic_copy_field(c(), GIC_FIELD_OVERDENSITY, level, GIC_FOURIER_SPACE, h.data(), GIC_DTYPE_F64));
    std::vector<double> full(size_t(n) * n * n * 2);
    for (int x = 0; x < n; ++x)
      for (int y = 0; y < n; ++y)
        for (int z = 0; z < n; ++z) {
          double re, im;
          if (z < nz) {
            size_t q = ((size_t(x) * n + y) * nz + z) * 2;
            re = h[q];
            im = h[q + 1];
          } else {
            int xs = (n - x) % n, ys = (n - y) % n, zs = n - z;
            size_t q = ((size_t(xs) * n + ys) * nz + zs) * 2;
            re = h[q];
            im = -h[q + 1];
          }
          size_t o = ((size_t(x) * n + y) * n + z) * 2;
          full[o] = re;
          full[o + 1] = im;
        }
    saveNpyComplex(outdir + "/grid-" + std::to_string(level) + ".npy", n, full);
    std::ofstream inf(outdir + "/grid-info-" + std::to_string(level) + "-fourier.txt");
    const Grid &g = grids[level];
    inf << g.ox << " " << g.oy << " " << g.oz << " " << g.size() << std::endl;
    inf << "The line above contains information about grid level " << level << std::endl;
    inf << "It gives the x-offset, y-offset and z-offset of the low-left corner and also the box length" << std::endl;
  }

  // expandFlaggedRegion (ic.hpp:1527-1540, grid.hpp:282-300): n times, each
  // flagged cell's neighbours along z, then y, then x join the flags
  // (indices wrapped around the level's grid)
  // Grid::expandFlaggedRegion (grid.hpp:282-300): nCells times a step of
  // one cell both ways along z, then y, then x, wrapping within the grid
  static void dilate(std::vector<size_t> &flags, long n, int nCells) {
    for (int it = 0; it < nCells; ++it)
      for (int d = 2; d >= 0; --d) {
        size_t old = flags.size();
        std::vector<size_t> f(old * 3);
        for (size_t i = 0; i < old; ++i) {
          size_t id = flags[i];
          long c[3] = {long(id / size_t(n * n)), long((id / size_t(n)) % size_t(n)), long(id % size_t(n))};
          f[i] = id;
          for (int k = 0; k < 2; ++k) {
            long cc[3] = {c[0], c[1], c[2]};
            cc[d] = ((cc[d] + (k == 0 ? 1 : -1)) % n + n) % n;
            f[i + old * (k + 1)] = size_t((cc[0] * n + cc[1]) * n + cc[2]);
          }
        }
        std::sort(f.begin(), f.end());
        f.erase(std::unique(f.begin(), f.end()), f.end());
        flags = f;
      }
  }
  void expandFlaggedRegion(int nCells) {
    if (grids.empty()) throw std::runtime_error("No grid has yet been initialised");
    if (nCells < 0) throw std::runtime_error("expand_flagged_region takes a number of cells");
    log() << "Expand flagged region by " << nCells << " cells" << std::endl;
    for (size_t l = 0; l < grids.size(); ++l) {
      Grid &g = grids[l];
      size_t initial = g.flags.size();
      dilate(g.flags, g.n, nCells);
      log() << "  - level " << l << " increased number of flagged cells by " << g.flags.size() - initial << " (now "
            << g.flags.size() << ")" << std::endl;
    }
  }
  // adaptMask (ic.hpp:1548-1576): from the second-finest level up, the cells
  // each level replaces on the level above become the parents of the cells
  // it has zoomed, padded by nPad cells (the flags themselves are untouched)
  void adaptMask(int nPad) {
    if (grids.size() < 3) throw std::runtime_error("Adapting the mask requires there to be more than one zoom level");
    if (nPad < 0) throw std::runtime_error("adapt_mask takes a number of cells");
    for (size_t level = grids.size() - 2; level > 0; --level) {
      const Grid &g = grids[level], &ga = grids[level - 1];
      int r = ratioInt(ga.dx, g.dx);
      long sim = g.simEquivalent();
      long off[3] = {std::lround((g.ox - ga.ox) / g.dx), std::lround((g.oy - ga.oy) / g.dx),
                     std::lround((g.oz - ga.oz) / g.dx)};
      // makeProxyGridToMatch (grid.hpp:195-228): the section of this level
      // over the level above, then subsampled to its cells
      std::vector<size_t> parents;
      for (size_t id : zoomParticles[level]) {
        I3 q = g.coord(id);
        long s3[3] = {q.x + off[0], q.y + off[1], q.z + off[2]};
        bool in = true;
        for (long &v : s3) {
          v = (v % sim + sim) % sim;
          in &= v < long(ga.n) * r;
        }
        if (in) parents.push_back(ga.index(int(s3[0] / r), int(s3[1] / r), int(s3[2] / r)));
      }
      std::sort(parents.begin(), parents.end());
      parents.erase(std::unique(parents.begin(), parents.end()), parents.end());
      dilate(parents, ga.n, nPad);
      zoomParticles[level - 1] = parents;
    }
    refreshMapper();
  }
  void zeroLevel(int level) {
    log() << "*** Warning: your script calls zeroLevel(" << level << "). This is intended for testing purposes only!"
          << std::endl;
    if (level < 0 || level >= int(grids.size())) throw std::runtime_error("Specified level does not exist.");
    ensureDrawn();
    check(gic_zero_level(c(), level));
  }
  void generate() {
    if (generated) return;
    if (!haveCamb) throw std::runtime_error("No power spectrum has been set (use the camb command)");
    ensureDrawn();
    check(gic_apply_power_spectrum(c()));
    check(gic_zeldovich(c()));
    generated = true;
  }
  std::string outputPath() const {
    std::ostringstream s;
    if (outname.empty()) s << outdir << "/IC_" << (fp32 ? "float" : "doub") << "_z" << z << "_" << grids[0].n;
    else s << outdir << "/" << outname;
    return s.str();
  }

  // Mass and softening of a block's particles (ZeldovichParticleEvaluator::
  // getMass/getEps, zeldovich.hpp:110-119): a subsampled grid carries
  // factor^3 of the level's mass per cell and the larger cell's softening, a
  // supersampled one 1/factor^3 of it and the level's own softening
  // (virtualgrid.hpp:139-147,539-545).
  void setMassEps(Block &b) {
    check(gic_particle_mass_eps(c(), b.level, &b.mass, nullptr));
    double f = std::abs(b.factor);
    if (b.factor > 1) b.mass *= f * f * f;
    if (b.factor < -1) b.mass /= f * f * f;
    b.eps = grids[b.level].dx * epsNorm * (b.factor > 1 ? f : 1.0);
  }
  // The cells of the gas copy of the finest level, in mapper order, on the
  // finest grid supersampled by supersample_gas: the gas mapper keeps the
  // levels above without resampling and lists only the finest level's
  // particles (TwoLevelParticleMapper::splitMass with skipLevel1,
  // twolevelmapper.hpp:542-575; the gas resampled by itself, ic.hpp:1200-1201).
  std::vector<uint64_t> gasCells(const VLevel &gv) const {
    size_t nl = grids.size();
    std::vector<uint64_t> out;
    if (nl == 1) {
      out.resize(gv.g.n3());
      std::iota(out.begin(), out.end(), uint64_t(0));
      return out;
    }
    std::vector<size_t> coarse(zoomParticles[nl - 2]), hr;
    std::sort(coarse.begin(), coarse.end());
    zoomCubes(grids[nl - 2], coarse, gv.g, hr);
    out.assign(hr.begin(), hr.end());
    return out;
  }

  // The blocks of the particle list in mapper order, without their particles:
  // the DM levels on their output grids, then (gadget) or before them (tipsy)
  // the gas copy of the finest level when Ob > 0 (ic.hpp:1129-1222,
  // AddGasMapper in particles/mapper/gasmapper.hpp; the split masses follow
  // OneLevelParticleMapper::splitMass, onelevelmapper.hpp:155-173).
  std::vector<Block> layoutBlocks(bool gasFirst) {
    buildMapper();
    auto order = levelOrder();
    size_t nl = grids.size(), K = outLv.size();
    std::vector<Block> dm(K);
    double massRatio = Ob / Om;
    for (size_t k = 0; k < K; ++k) {
      // a padding entry takes the type of the level it supersamples
      // (getGadgetParticleTypeForFinestGrid, twolevelmapper.hpp:700-702)
      const VLevel &v = outLv[k];
      Block &b = dm[k];
      b.level = v.src;
      b.chain = int(k);
      b.factor = v.factor;
      b.strays = v.strays;
      b.type = grids[v.src].gadgetType;
      b.cells = std::move(order[k]);
      setMassEps(b);
      if (Ob > 0 && v.src == int(nl) - 1) b.mass *= 1.0 - massRatio;
    }
    if (flaggedTypeOn) {
      if (Ob > 0) throw std::runtime_error("gadget_flagged_particle_type with baryons is not built here");
      // the finest block keeps its unflagged cells in order; the flagged ones
      // follow as their own block, on the same (possibly resampled) grid
      auto flaggedOut = toOutputCells(outLv[K - 1], flaggedTypeCells);
      Block f = dm[K - 1];
      f.type = flaggedType;
      std::vector<uint64_t> keep, moved;
      for (uint64_t cell : dm[K - 1].cells) {
        if (std::binary_search(flaggedOut.begin(), flaggedOut.end(), size_t(cell))) moved.push_back(cell);
        else keep.push_back(cell);
      }
      std::sort(moved.begin(), moved.end());
      dm[K - 1].cells = std::move(keep);
      f.cells = std::move(moved);
      dm.push_back(std::move(f));
    }
    std::vector<Block> out;
    if (Ob > 0) {
      if (allowStrays && nl > 1) throw std::runtime_error("strays_on with baryons is not built here");
      VLevel gv = resampledLevel(int(nl - 1), supersampleGas > 1 ? -supersampleGas : 1);
      Block g;
      g.level = int(nl - 1);
      g.factor = gv.factor;
      g.gas = true;
      g.type = 0;
      g.cells = gasCells(gv);
      setMassEps(g);
      g.mass *= massRatio;
      if (gasFirst) out.push_back(std::move(g));
      for (auto &b : dm) out.push_back(std::move(b));
      if (!gasFirst) out.push_back(std::move(g));
    } else {
      out = std::move(dm);
    }
    size_t first = 0;
    for (auto &b : out) {
      b.first = first;
      first += b.count();
    }
    return out;
  }

  // The F64 real-space values of field k (GIC_FIELD_*) on real level l
  std::vector<double> levelField(int k, int l) {
    std::vector<double> f(grids[l].n3());
    check(gic_copy_field(c(), k, l, GIC_REAL_SPACE, f.data(), GIC_DTYPE_F64));
    return f;
  }

  // Particle blocks in mapper order with positions and velocities: real
  // levels from the device, resampled levels evaluated here from the
  // level's displacement fields through the virtual grid
  // (ZeldovichParticleEvaluator::getParticleNoWrap, zeldovich.hpp:87-108;
  // ParticleEvaluator::getParticle wraps, generator.hpp:49-53).
  std::vector<Block> particleBlocks(bool gasFirst) {
    auto out = layoutBlocks(gasFirst);
    double velRatio = 0;
    bool second = lptOrder == 2;
    for (auto &b : out) {
      size_t m = b.count();
      b.pos.resize(3 * m);
      b.vel.resize(3 * m);
      if (!m) continue;
      if (b.factor == 1 && !b.strays) {
        bool identity = m == grids[b.level].n3();
        for (size_t i = 0; identity && i < m; ++i) identity = b.cells[i] == i;
        check(gic_particles(c(), b.level, identity ? nullptr : b.cells.data(), 0, m, b.pos.data(), b.vel.data()));
      } else {
        if (velRatio == 0) check(gic_velocity_ratio(c(), &velRatio));
        VLevel v = blockLevel(b), hi = resampledLevel(b.level, b.factor);
        // the displacement fields of the level (and of the base level for
        // stray cells), fetched before the parallel loop
        std::map<int, std::vector<std::vector<double>>> psi;
        for (int l : {b.level, b.strays ? b.level - 1 : b.level})
          if (!psi.count(l))
            for (int k = 1; k <= (second ? 6 : 3); ++k) psi[l].push_back(levelField(k, l));
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < m; ++i) {
          size_t id = b.cells[i];
          auto cen = v.gu.centroid(id);
          for (int a3 = 0; a3 < 3; ++a3) {
            double p = evalOut(v, hi, [&](int l) -> const std::vector<double> & { return psi.at(l)[a3]; }, id);
            double vel = p * velRatio;
            if (second) {
              double p2 =
                  evalOut(v, hi, [&](int l) -> const std::vector<double> & { return psi.at(l)[3 + a3]; }, id);
              p += p2;
              vel += 2 * velRatio * p2;
            }
            b.pos[3 * i + a3] = float(v.g.wrap(p + cen[a3]));
            b.vel[3 * i + a3] = float(vel);
          }
        }
      }
      if (velOffset[0] != 0 || velOffset[1] != 0 || velOffset[2] != 0)
        for (size_t i = 0; i < m; ++i)
          for (int d = 0; d < 3; ++d) b.vel[3 * i + d] = float(b.vel[3 * i + d] + velOffset[d]);
    }
    return out;
  }

  // --- modifications (ic.hpp:1592-1631) ---
  void syncFlags() {
    for (size_t l = 0; l < grids.size(); ++l) {
      std::vector<uint64_t> f(grids[l].flags.begin(), grids[l].flags.end());
      check(gic_set_flags(c(), int(l), f.data(), f.size()));
    }
    check(gic_set_variance_filter_scale(c(), varianceScale));
  }
  // warnings of the Modification and FilteredVarianceModification constructors
  // (modification.hpp:36-49, quadraticmodification.hpp:131-150)
  void modificationWarnings(const std::string &name) {
    for (size_t l = 1; l < grids.size(); ++l)
      if (grids[l].flags.size() == grids[l].n3()) {
        log() << "WARNING: Region selected for modification is the entire zoom grid. This is likely because the cell "
                 "selection extends beyond the zoom boundaries."
              << std::endl;
        std::cerr << "By design, modifications are meant to be defined inside a zoom region. Increase the size of "
                     "your zoom grid or decrease your selection to avoid nasty surprises."
                  << std::endl;
      }
    if (strcasecmp(name.c_str(), "variance") == 0 && varianceScale >= 0 && varianceScale >= grids.back().dx) {
      const Grid &g = grids.back();
      double windowSize = 0;
      if (!g.flags.empty()) {
        Window w(g.simEquivalent(), g.coord(g.flags[0]));
        for (size_t id : g.flags) w.expandToInclude(g.coord(id));
        windowSize = double(w.maxDim()) * g.dx;
      }
      if (varianceScale > windowSize) {
        log() << "WARNING: High-pass filtering scale: " << varianceScale
              << " h**-1 Mpc is greater than the rough window size: " << windowSize
              << " h**-1 Mpc used for modifications." << std::endl;
        log() << "This is prone to numerical errors when modifying the field. Decrease filtering scale to avoid it."
              << std::endl;
      }
    }
  }
  void calculate(const std::string &name) {
    ensureDrawn();
    syncFlags();
    modificationWarnings(name);
    double v;
    check(gic_calculate(c(), name.c_str(), &v));
    log() << name << ": calculated value = " << v << std::endl;
  }
  void modify(const std::string &name, const std::string &type, float target) {
    ensureDrawn();
    bool relative = strcasecmp(type.c_str(), "relative") == 0;
    if (!relative && strcasecmp(type.c_str(), "absolute") != 0)
      throw std::runtime_error("Modification type must be either 'relative' or 'absolute'");
    syncFlags();
    modificationWarnings(name);
    check(gic_add_modification(c(), name.c_str(), relative ? 1 : 0, double(target), initialNumberSteps, precision));
    ++nModifications;
  }
  void applyModifications() {
    ensureDrawn();
    log() << std::endl << "Applying modifications" << std::endl;
    double pre, post;
    check(gic_apply_modifications(c(), &pre, &post));
    uint64_t dof = numDof();
    log() << "   Post-modification chi^2 = " << post << std::endl;
    log() << "  Modification Delta chi^2 = " << post - pre << std::endl;
    log() << "           d.o.f. in field = " << dof << std::endl;
    log() << std::endl;
  }
  void clearModifications() {
    log() << "Clearing modification list" << std::endl;
    check(gic_clear_modifications(c()));
    nModifications = 0;
  }
  uint64_t numDof() const {   // MultiLevelGrid::getNumDof (multilevelgrid.hpp:160-172)
    uint64_t dof = 0;
    for (size_t i = 0; i < grids.size(); ++i) {
      dof += grids[i].n3();
      if (i > 0) {
        uint64_t f = uint64_t(ratioInt(grids[i - 1].dx, grids[i].dx));
        dof -= grids[i].n3() / (f * f * f);
      }
    }
    return dof;
  }
  void chi2() {
    ensureDrawn();
    double v;
    uint64_t dof;
    check(gic_chi2(c(), &v, &dof));
    log() << "Calculated chi^2 = " << std::setprecision(10) << v << " (dof = " << dof << ")" << std::setprecision(6)
          << std::endl;
  }

  void done() {
    if (fmt == Fmt::unknown) throw std::runtime_error("No output format specified!");
    ensureDrawn();
    if (nModifications) applyModifications();
    generate();
    switch (fmt) {
      case Fmt::tipsy: {
        auto blocks = particleBlocks(true);
        logTotals(blocks);
        writeTipsy(outputPath() + ".tipsy", blocks);
        break;
      }
      case Fmt::gadget2:
      case Fmt::gadget3: {
        auto blocks = particleBlocks(false);
        logTotals(blocks);
        writeGadget(outputPath() + ".gadget", blocks, int(fmt));
        break;
      }
      case Fmt::grafic:
        writeGrafic(outputPath() + ".grafic");
        break;
      default:
        throw std::runtime_error("Output format " + std::to_string(int(fmt)) + " is not supported by this build");
    }
    log() << "Finished writing initial conditions" << std::endl;
  }
  void logTotals(const std::vector<Block> &blocks) {
    size_t ngas = 0, ndm = 0;
    for (auto &b : blocks) (b.gas ? ngas : ndm) += b.count();
    log() << "Writing output; number dm particles=" << ndm << ", number gas particles=" << ngas << std::endl;
  }

  // TipsyOutput::operator() (io/tipsy.hpp:207-266): gas first, masses scaled
  // to sum to Om, positions in box units centred on 0.
  void writeTipsy(const std::string &fn, const std::vector<Block> &blocks) {
    double boxLength = grids[0].periodic;
    double tot_mass = 0, min_mass = 1e300, max_mass = 0;
    for (auto &b : blocks) {
      for (size_t i = 0; i < b.count(); ++i) tot_mass += b.mass;
      if (b.count()) {
        min_mass = std::min(min_mass, b.mass);
        max_mass = std::max(max_mass, b.mass);
      }
    }
    double mass_factor = Om / tot_mass;
    double pos_factor = 1. / boxLength;
    double dKpcUnit = boxLength * 1000 / hubble;
    double dMsolUnit = 1e10 / hubble / mass_factor;
    double dKmsUnit = std::sqrt(4.3022682e-6 * dMsolUnit / dKpcUnit);
    double vel_factor = std::pow(a(), -0.5) / dKmsUnit;
    {
      std::ofstream pf("tipsy.param");
      pf << "dKpcUnit = " << dKpcUnit << std::endl;
      pf << "dMsolUnit = " << dMsolUnit << std::endl;
      pf << "dHubble0 = " << 0.1 * hubble * dKpcUnit / dKmsUnit << std::endl;
      pf << "bComove = 1 " << std::endl;
    }
    std::ofstream photo;
    if (min_mass != max_mass) photo.open("photogenic.txt");
    std::ofstream f(fn, std::ios::binary);
    if (!f) throw std::runtime_error("Can't open file " + fn);
    size_t total = 0, ngas = 0;
    for (auto &b : blocks) {
      total += b.count();
      if (b.gas) ngas += b.count();
    }
    // header {double a; int n, ndim, ngas, ndark, nstar;} padded to 32 bytes
    unsigned char hdr[32];
    memset(hdr, 0, sizeof hdr);
    double sf = a();
    int ints[5] = {int(total), 3, int(ngas), int(total - ngas), 0};
    memcpy(hdr, &sf, 8);
    memcpy(hdr + 8, ints, 20);
    f.write((const char *)hdr, 32);
    // gas temperature (cosmology/parameters.hpp:37-42), decoupling at a = 1/150
    double aDec = 1. / 150.;
    double temp = sf > aDec ? TCMB * aDec / (sf * sf) : TCMB / sf;
    std::vector<float> rec;
    size_t iord = 0;
    for (auto &b : blocks) {
      size_t m = b.count();
      int w = b.gas ? 12 : 9;   // gas: mass x y z vx vy vz rho temp eps metals phi
      rec.assign(size_t(w) * m, 0.f);
      float mass = float(b.mass * mass_factor), eps = float(b.eps * pos_factor);
      for (size_t i = 0; i < m; ++i) {
        float *r = &rec[size_t(w) * i];
        r[0] = mass;
        for (int d = 0; d < 3; ++d) {
          r[1 + d] = float(b.pos[3 * i + d] * pos_factor - 0.5);
          r[4 + d] = float(b.vel[3 * i + d] * vel_factor);
        }
        if (b.gas) {
          r[8] = float(temp);
          r[9] = eps;
        } else {
          r[7] = eps;
        }
      }
      f.write((const char *)rec.data(), rec.size() * 4);
      if (photo.is_open() && b.mass == min_mass)
        for (size_t i = 0; i < m; ++i) photo << iord + i << std::endl;
      iord += m;
    }
  }

  // GadgetOutput (io/gadget.hpp:26-232, 238-557): particles grouped by gadget
  // type, IDs = mapper index, masses in the header unless they vary.
  void writeGadget(const std::string &name, const std::vector<Block> &blocks, int version) {
    std::vector<const Block *> seq;   // file order: per type, its blocks in mapper order
    size_t npType[6] = {0, 0, 0, 0, 0, 0};
    double mmin[6], mmax[6];
    for (int t = 0; t < 6; ++t) {
      mmin[t] = 1e300;
      mmax[t] = 0;
      for (auto &b : blocks)
        if (int(b.type) == t && b.count()) {
          seq.push_back(&b);
          npType[t] += b.count();
          mmin[t] = std::min(mmin[t], b.mass);
          mmax[t] = std::max(mmax[t], b.mass);
        }
    }
    bool variable = false;
    double masses[6] = {0, 0, 0, 0, 0, 0};
    size_t nTotal = 0;
    log() << "Particles by gadget type:" << std::endl;
    for (int t = 0; t < 6; ++t)
      if (npType[t]) {
        if (mmin[t] != mmax[t]) variable = true;
        masses[t] = mmin[t];
        nTotal += npType[t];
        log() << "   Particle type " << t << ": " << npType[t] << " particles" << std::endl;
      }
    if (variable) {
      log() << "Using variable-mass gadget format" << std::endl;
      for (double &m : masses) m = 0;
    } else {
      log() << "Using fixed-mass gadget format" << std::endl;
    }
    int nFiles = nGadgetFiles;
    std::vector<size_t> perFile;
    for (int i = 0; i < nFiles; ++i)
      perFile.push_back(i == nFiles - 1 ? nTotal - (nTotal / nFiles) * (nFiles - 1) : nTotal / nFiles);
    // per-file per-type counts (gadget.hpp:415-447)
    std::vector<std::array<size_t, 6>> typePerFile;
    size_t remaining[6];
    std::copy(npType, npType + 6, remaining);
    for (int i = 0; i < nFiles; ++i) {
      size_t left = perFile[i];
      std::array<size_t, 6> c{};
      for (int t = 0; t < 6; ++t) {
        c[t] = std::min(remaining[t], left);
        left -= c[t];
        remaining[t] -= c[t];
      }
      typePerFile.push_back(c);
    }
    // walk the file-order sequence
    size_t bi = 0, bo = 0;
    auto next = [&](const Block *&b, size_t &i) {
      while (bo >= seq[bi]->count()) {
        ++bi;
        bo = 0;
      }
      b = seq[bi];
      i = bo++;
    };
    for (int fi = 0; fi < nFiles; ++fi) {
      std::string fn = name + std::to_string(version) + (nFiles == 1 ? "" : "." + std::to_string(fi));
      std::ofstream f(fn, std::ios::binary);
      if (!f) throw std::runtime_error("Can't open file " + fn);
      auto rec = [&](const void *p, uint32_t n) {
        f.write((const char *)&n, 4);
        f.write((const char *)p, n);
        f.write((const char *)&n, 4);
      };
      unsigned char hdr[256];
      memset(hdr, 0, sizeof hdr);
      size_t o = 0;
      auto put = [&](const void *p, size_t n) {
        memcpy(hdr + o, p, n);
        o += n;
      };
      int npart[6];
      unsigned int nall[6], nhigh[6];
      for (int t = 0; t < 6; ++t) {
        npart[t] = int(typePerFile[fi][t]);
        nall[t] = (unsigned)npType[t];
        nhigh[t] = (unsigned)(uint64_t(npType[t]) >> 32);
      }
      int baryonFlag = npType[0] > 0 ? 1 : 0, zero = 0;
      double time = a(), redshift = z, box = grids[0].periodic, om = Om, ol = Ol, hh = hubble;
      put(npart, 24);
      put(masses, 48);
      put(&time, 8);
      put(&redshift, 8);
      put(&baryonFlag, 4);   // flag_sfr
      put(&baryonFlag, 4);   // flag_feedback
      put(nall, 24);
      put(&baryonFlag, 4);   // flag_cooling
      put(&nFiles, 4);
      put(&box, 8);
      put(&om, 8);
      put(&ol, 8);
      put(&hh, 8);
      put(&zero, 4);   // flag_stellarage
      put(&zero, 4);   // flag_metals
      put(nhigh, 24);
      put(&zero, 4);   // flag_entropy_instead_u
      if (version == 3) {
        put(&zero, 4);   // flag_doubleprecision: float output
        int icInfo = 1;
        put(&icInfo, 4);
        float lpt = 0.f;
        put(&lpt, 4);
      }
      rec(hdr, 256);
      size_t nf = perFile[fi];
      std::vector<float> pos(3 * nf), vel(3 * nf), mass(nf);
      std::vector<uint64_t> ids(nf);
      for (size_t j = 0; j < nf; ++j) {
        const Block *b;
        size_t i;
        next(b, i);
        for (int d = 0; d < 3; ++d) {
          pos[3 * j + d] = b->pos[3 * i + d];
          vel[3 * j + d] = b->vel[3 * i + d];
        }
        ids[j] = b->first + i;
        mass[j] = float(b->mass);
      }
      rec(pos.data(), uint32_t(pos.size() * 4));
      rec(vel.data(), uint32_t(vel.size() * 4));
      rec(ids.data(), uint32_t(nf * 8));
      if (variable) rec(mass.data(), uint32_t(nf * 4));
      if (Ob > 0) {   // dummy internal energy of this file's gas particles
        std::vector<float> u(typePerFile[fi][0], 0.f);
        rec(u.data(), uint32_t(u.size() * 4));
      }
    }
  }

  // --- grafic ---
  // The grafic hierarchy (MultiLevelGrid::copyContextWithCenteredIntermediate,
  // multilevelgrid.hpp:338-473): extra subsampled copies of the base level
  // ('subsample'), subsampled intermediates so that successive levels differ
  // by a factor 2, supersampled copies of the finest level ('supersample');
  // then every level covering the box is re-centred on a point by an integer
  // cell shift (CenteredGrid, virtualgrid.hpp:688-800) and the others are
  // offset by the same distance (OffsetGrid, virtualgrid.hpp:806-840).
  int subsample = 1, supersample = 1, supersampleGas = 1, lptOrder = 1;
  bool allowStrays = false;   // strays_on (ic.hpp:257-260)
  size_t autopad = 0;         // autopad (ic.hpp:273-277): padding cells around each zoom
  double pvarValue = 1.0;
  bool centerOnTargetRegion = false;

  static int powerOf2(int v) {   // tools::findPowerOf(2, v)
    int p = 0;
    while (v > 1 && v % 2 == 0) {
      v /= 2;
      ++p;
    }
    if (v != 1)
      throw std::runtime_error(
          "Subsample and supersample cannot be converted into an integer number of grids with this base factor");
    return p;
  }
  VLevel subsampled(int l, int f) const {
    VLevel v;
    v.src = l;
    v.factor = f;
    v.g = grids[l];
    v.g.flags.clear();
    v.g.n = grids[l].n / f;
    v.g.dx = grids[l].dx * f;
    return v;
  }
  // Value at cell i of a virtual level (indexing of v.g) of a field f of its
  // real level (n^3 values, index order): the level itself; a subsampled
  // level averaging the real cells it covers (SubSampleGrid::forEachSubcell,
  // virtualgrid.hpp:576-607, which clips at the grid's far edge); a
  // supersampled level interpolating the real level at the cell's centroid
  // (SuperSampleEvaluator, field/evaluator.hpp:96-126: evaluateInterpolated,
  // field.hpp:483-508, with the tricubic of field.hpp:598-617), wrapping on a
  // level that covers the box and repeating edge values otherwise.
  double evalVirtual(const VLevel &v, const std::vector<double> &f, size_t i) const {
    const Grid &rg = grids[v.src];
    if (v.factor < -1) return interpolate(v.src, v.gu.centroid(i), f);
    if (v.factor <= 1) return f[i];
    I3 c0 = v.g.coord(i);
    double sum = 0;
    int cnt = 0;
    int fct = v.factor;
    for (int x = c0.x * fct; x < std::min(c0.x * fct + fct, rg.n); ++x)
      for (int y = c0.y * fct; y < std::min(c0.y * fct + fct, rg.n); ++y)
        for (int z = c0.z * fct; z < std::min(c0.z * fct + fct, rg.n); ++z) {
          sum += f[rg.index(x, y, z)];
          ++cnt;
        }
    return sum / cnt;
  }
  // Field f of real level l at point p: Field::evaluateInterpolated
  // (field.hpp:483-508) with the tricubic of field.hpp:598-617, wrapping on a
  // level that covers the box and repeating edge values otherwise.
  double interpolate(int l, const std::array<double, 3> &p, const std::vector<double> &f) const {
    const Grid &rg = grids[l];
    {
      double o[3] = {rg.ox, rg.oy, rg.oz};
      int p0[3];
      double w[3][4];
      for (int a3 = 0; a3 < 3; ++a3) {
        double u = rg.wrap(p[a3] - o[a3]) / rg.dx - 0.5;
        p0[a3] = int(std::floor(u));
        double t = u - p0[a3], t2 = t * t, t3 = t2 * t;
        w[a3][0] = -0.5 * t3 + t2 - 0.5 * t;
        w[a3][1] = 1.5 * t3 - 2.5 * t2 + 1.0;
        w[a3][2] = -1.5 * t3 + 2.0 * t2 + 0.5 * t;
        w[a3][3] = 0.5 * t3 - 0.5 * t2;
      }
      bool wrapc = rg.n == rg.simEquivalent();
      int sim = rg.simEquivalent();
      auto fix = [&](int c) {
        if (wrapc) {
          if (c > sim - 1) c -= sim;
          if (c < 0) c += sim;
        } else {
          c = std::min(std::max(c, 0), rg.n - 1);
        }
        return c;
      };
      double sum = 0;
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
          for (int cc = 0; cc < 4; ++cc)
            sum += w[0][a] * w[1][b] * w[2][cc] *
                   f[rg.index(fix(p0[0] + a - 1), fix(p0[1] + b - 1), fix(p0[2] + cc - 1))];
      return sum;
    }
  }

  std::vector<VLevel> graficStack(std::array<double, 3> centre) const {
    int extraLow = powerOf2(subsample), extraHigh = powerOf2(supersample);
    std::vector<VLevel> st;
    for (size_t l = 0; l < grids.size(); ++l) {
      int nint = extraLow;
      if (l > 0) nint = int(std::round(std::log10(grids[l - 1].dx / grids[l].dx) / std::log10(2.0))) - 1;
      for (int k = nint; k >= 1; --k) st.push_back(subsampled(int(l), 1 << k));
      VLevel r;
      r.src = int(l);
      r.g = grids[l];
      r.g.flags.clear();
      st.push_back(r);
    }
    for (int k = 1; k <= extraHigh; ++k) {
      VLevel v;
      v.src = int(grids.size()) - 1;
      v.factor = -(1 << k);
      v.g = grids.back();
      v.g.flags.clear();
      v.g.n = grids.back().n << k;
      v.g.dx = grids.back().dx / (1 << k);
      st.push_back(v);
    }
    // copyContextAndCenter (multilevelgrid.hpp:426-445)
    int deepest = -1;
    for (int i = int(st.size()) - 1; i >= 0; --i)
      if (st[i].g.coversFull()) {
        deepest = i;
        break;
      }
    if (deepest < 0) throw std::runtime_error("No coarse grid were found.");
    double off[3] = {0, 0, 0};
    for (int i = 0; i <= deepest; ++i) {
      VLevel &v = st[i];
      v.centred = true;
      for (int a = 0; a < 3; ++a) {
        v.shift[a] = int((0.5 * v.g.size() - centre[a]) / v.g.dx);   // Coordinate<int>: truncation
        off[a] = v.shift[a] * v.g.dx;
      }
    }
    for (auto &v : st) v.gu = v.g;
    for (size_t i = deepest + 1; i < st.size(); ++i) {
      Grid &g = st[i].g;
      g.ox = g.wrap(g.ox + off[0]);
      g.oy = g.wrap(g.oy + off[1]);
      g.oz = g.wrap(g.oz + off[2]);
    }
    return st;
  }
  // Grid::getIndexFromCoordinateNoWrap of a level of the stack: a centred
  // level steps back by its shift and wraps (getIndexFromIndexAndStep)
  static size_t vIndex(const VLevel &v, int x, int y, int z) {
    if (!v.centred) return v.g.index(x, y, z);
    int s = v.g.simEquivalent();
    int c[3] = {x - v.shift.x, y - v.shift.y, z - v.shift.z};
    for (int a = 0; a < 3; ++a) {
      if (c[a] > s - 1) c[a] -= s;
      if (c[a] < 0) c[a] += s;
    }
    return v.g.index(c[0], c[1], c[2]);
  }
  // getCentroidFromIndex / getIndexFromPoint of a stack level
  static std::array<double, 3> vCentroid(const VLevel &v, size_t id) {
    auto p = v.g.centroid(id);
    if (v.centred) {
      p[0] = v.g.wrap(p[0] + v.shift.x * v.g.dx);
      p[1] = v.g.wrap(p[1] + v.shift.y * v.g.dx);
      p[2] = v.g.wrap(p[2] + v.shift.z * v.g.dx);
    }
    return p;
  }
  static size_t vIndexFromPoint(const VLevel &v, std::array<double, 3> p) {
    if (v.centred)
      for (int a = 0; a < 3; ++a) p[a] -= v.shift[a] * v.g.dx;
    double o[3] = {v.g.ox, v.g.oy, v.g.oz};
    int c[3];
    for (int a = 0; a < 3; ++a) c[a] = int(std::floor(v.g.wrap(p[a] - o[a]) / v.g.dx));
    return v.g.index(c[0], c[1], c[2]);
  }
  // GraficMask (multilevelgrid/mask.hpp:46-231): per stack level, the
  // flagged cells; a level without flags counts as all in the mask
  std::vector<std::vector<size_t>> graficMask(const std::vector<VLevel> &st) const {
    size_t nl = st.size();
    std::vector<std::vector<size_t>> flagged(nl);
    const auto &input = zoomParticles;
    int deepest = -1;
    size_t il = 0;
    for (size_t l = 0; l + 1 < nl && il < input.size(); ++l) {
      // the first level where some cell index 0..size3-1 is in the next input list
      if (!input[il].empty() && input[il].front() < st[l].g.n3()) {
        flagged[l] = input[il];
        deepest = int(l);
        ++il;
      }
    }
    if (deepest < 0) {
      log() << "WARNING No zoom regions were ever opened. Grafic mask will not be generated in this case" << std::endl;
      return flagged;
    }
    for (int l = deepest; l >= 0; --l) {
      if (!flagged[l].empty()) continue;
      for (size_t i : flagged[l + 1]) flagged[l].push_back(vIndexFromPoint(st[l], vCentroid(st[l + 1], i)));
      std::sort(flagged[l].begin(), flagged[l].end());
      flagged[l].erase(std::unique(flagged[l].begin(), flagged[l].end()), flagged[l].end());
    }
    for (size_t l = deepest + 1; l + 1 < nl; ++l) {
      for (size_t i = 0; i < st[l].g.n3(); ++i) {
        size_t above = vIndexFromPoint(st[l - 1], vCentroid(st[l], i));
        if (std::binary_search(flagged[l - 1].begin(), flagged[l - 1].end(), above)) flagged[l].push_back(i);
      }
    }
    return flagged;
  }
  static float inMask(const std::vector<std::vector<size_t>> &flagged, size_t l, size_t i) {
    if (flagged[l].empty()) return 1.f;
    return std::binary_search(flagged[l].begin(), flagged[l].end(), i) ? 1.f : 0.f;
  }
  // Particle IDs of grafic output (GraficMapper, graficmapper.hpp:75-142):
  // the cells of the hierarchy's levels numbered one level after another.
  // Flags live on the real grids; a subsampled level sees them coarsened
  // (SubSampleGrid, virtualgrid.hpp:556-568), a supersampled one refined
  // (SuperSampleGrid, virtualgrid.hpp:161-175; Grid::upscaleCellFlagVector and
  // downscaleCellFlagVector, grid.hpp:384-436); centring and offsets move
  // centroids but keep cell ids.
  static std::vector<size_t> levelFlagsToReal(const VLevel &v, const Grid &real, const std::vector<size_t> &ids) {
    std::vector<size_t> out;
    if (v.factor > 1) {
      int f = v.factor;
      for (size_t id : ids) {
        I3 c = v.g.coord(id);
        for (int x = c.x * f; x < c.x * f + f && x < real.n; ++x)
          for (int y = c.y * f; y < c.y * f + f && y < real.n; ++y)
            for (int z = c.z * f; z < c.z * f + f && z < real.n; ++z) out.push_back(real.index(x, y, z));
      }
    } else if (v.factor < -1) {
      int f = -v.factor;
      for (size_t id : ids) {
        I3 c = v.g.coord(id);
        out.push_back(real.index(c.x / f, c.y / f, c.z / f));
      }
    } else {
      out = ids;
    }
    return out;
  }
  static std::vector<size_t> realFlagsToLevel(const VLevel &v, const Grid &real) {
    std::vector<size_t> out;
    if (v.factor > 1) {
      int f = v.factor;
      for (size_t id : real.flags) {
        I3 c = real.coord(id);
        out.push_back(v.g.index(c.x / f, c.y / f, c.z / f));
      }
      std::sort(out.begin(), out.end());
      out.erase(std::unique(out.begin(), out.end()), out.end());
    } else if (v.factor < -1) {
      // each coarse flag becomes its f^3 sub-cells, x slowest, in flag order
      int f = -v.factor;
      for (size_t id : real.flags) {
        I3 c = real.coord(id);
        for (int x = c.x * f; x < c.x * f + f; ++x)
          for (int y = c.y * f; y < c.y * f + f; ++y)
            for (int z = c.z * f; z < c.z * f + f; ++z) out.push_back(v.g.index(x, y, z));
      }
    } else {
      out = real.flags;
    }
    return out;
  }
  // GraficMapper::flagParticles + propagateFlagsThroughHierarchy
  // (graficmapper.hpp:75-108,163-184); ids ascending
  void graficFlagParticles(const std::vector<size_t> &ids) {
    auto st = graficStack({x0, y0, z0});
    size_t start = 0, i = 0;
    for (const VLevel &v : st) {
      size_t end = start + v.g.n3();
      std::vector<size_t> cells;
      while (i < ids.size() && ids[i] < end) cells.push_back(ids[i++] - start);
      if (v.factor != 1 && !cells.empty())
        log() << cells.size() << " input ids reference a GRAFIC intermediate grid - make sure you intended to do this !"
              << std::endl;
      grids[v.src].flagCells(levelFlagsToReal(v, grids[v.src], cells));
      start = end;
    }
    if (i != ids.size()) throw std::runtime_error("Ran out of grids when interpreting grafic cell IDs - check IDs?");
    // flagged cells of each full-resolution level flag the cells under their
    // centroids on the next coarser full-resolution level, finest first
    std::vector<size_t> real;
    for (size_t l = 0; l < st.size(); ++l)
      if (st[l].factor == 1) real.push_back(l);
    for (size_t k = real.size() - 1; k > 0; --k) {
      const VLevel &fine = st[real[k]], &coarse = st[real[k - 1]];
      std::vector<size_t> down;
      for (size_t id : grids[fine.src].flags) down.push_back(vIndexFromPoint(coarse, vCentroid(fine, id)));
      std::sort(down.begin(), down.end());
      down.erase(std::unique(down.begin(), down.end()), down.end());
      grids[coarse.src].flagCells(down);
    }
  }
  // GraficMapper::getFlaggedParticles (graficmapper.hpp:127-142)
  std::vector<size_t> graficFlaggedParticles() const {
    std::vector<size_t> out;
    size_t offset = 0;
    for (const VLevel &v : graficStack({x0, y0, z0})) {
      for (size_t id : realFlagsToLevel(v, grids[v.src])) out.push_back(id + offset);
      offset += v.g.n3();
    }
    return out;
  }
  std::array<double, 3> graficCentre() const {
    if (centerOnTargetRegion) return {x0, y0, z0};
    double h = grids[0].periodic / 2;
    return {h, h, h};
  }

  // dumpMask (ic.hpp:975-994): mask-<l>.npy of every level of the hierarchy
  void dumpMask() {
    if (grids.empty()) throw std::runtime_error("No grid has yet been initialised");
    auto st = graficStack(graficCentre());
    auto flagged = graficMask(st);
    for (size_t l = 0; l < st.size(); ++l) {
      int n = st[l].g.n;
      std::vector<double> d(size_t(n) * n * n);
      for (int x = 0; x < n; ++x)
        for (int y = 0; y < n; ++y)
          for (int z = 0; z < n; ++z) d[(size_t(x) * n + y) * n + z] = inMask(flagged, l, vIndex(st[l], x, y, z));
      saveNpy(outdir + "/mask-" + std::to_string(l) + ".npy", n, d);
      std::ofstream inf(outdir + "/mask-info-" + std::to_string(l) + ".txt");
      const Grid &g = st[l].g;
      inf << g.ox << " " << g.oy << " " << g.oz << " " << g.size() << std::endl;
      inf << "The line above contains information about grid level " << l << std::endl;
      inf << "It gives the x-offset, y-offset and z-offset of the low-left corner and also the box length" << std::endl;
    }
  }

  // GraficOutput::write (io/grafic.hpp:96-241): per level a directory
  // <name>_<effective size> with ten files, each a Fortran header record and
  // one record per z plane (x fastest).  Displacements in Mpc/h, velocities
  // in km/s (times a^1/2), lengths in the header in Mpc.
  void writeGrafic(const std::string &name) {
    auto st = graficStack(graficCentre());
    auto flagged = graficMask(st);
    double velRatio;
    check(gic_velocity_ratio(c(), &velRatio));
    double velFactor = std::pow(a(), 0.5);
    bool second = lptOrder == 2;
    size_t iordOffset = 0;
    const Grid &base = st[0].g;
    // the fields of each real level, fetched once
    std::vector<std::vector<std::vector<double>>> fields(grids.size());
    auto fieldsOf = [&](int l) -> std::vector<std::vector<double>> & {
      auto &f = fields[l];
      if (f.empty()) {
        int nf = second ? 7 : 4;
        f.resize(nf, std::vector<double>(grids[l].n3()));
        check(gic_copy_field(c(), GIC_FIELD_OVERDENSITY, l, GIC_REAL_SPACE, f[0].data(), GIC_DTYPE_F64));
        for (int k = 1; k < nf; ++k)
          check(gic_copy_field(c(), k, l, GIC_REAL_SPACE, f[k].data(), GIC_DTYPE_F64));
      }
      return f;
    };
    for (size_t l = 0; l < st.size(); ++l) {
      const VLevel &v = st[l];
      const Grid &g = v.g;
      int n = g.n;
      auto &f = fieldsOf(v.src);
      size_t effective = size_t(std::llround(base.dx * base.n / g.dx));
      std::string dir = name + "_" + std::to_string(effective);
      mkdir(dir.c_str(), 0777);
      static const char *names[10] = {"ic_velcx", "ic_velcy", "ic_velcz", "ic_poscx", "ic_poscy",
                                      "ic_poscz", "ic_deltab", "ic_refmap", "ic_pvar_00001", "ic_particle_ids"};
      std::vector<std::ofstream> out(10);
      struct {
        int32_t nx, ny, nz;
        float dx, xo, yo, zo, a, om, ol, h0;
      } hdr{n, n, n, float(g.dx / hubble), float(g.ox / hubble), float(g.oy / hubble), float(g.oz / hubble),
            float(a()), float(Om), float(Ol), float(hubble * 100)};
      static_assert(sizeof(hdr) == 44, "grafic header is 44 bytes");
      int32_t hl = 44;
      for (int k = 0; k < 10; ++k) {
        out[k].open(dir + "/" + names[k], std::ios::binary);
        if (!out[k]) throw std::runtime_error("Cannot open " + dir + "/" + names[k]);
        out[k].write((const char *)&hl, 4);
        out[k].write((const char *)&hdr, 44);
        out[k].write((const char *)&hl, 4);
      }
      size_t n2 = size_t(n) * n;
      std::vector<std::vector<float>> plane(9, std::vector<float>(n2));
      std::vector<uint64_t> ids(n2);
      // value of real-level field k at stack-level cell i: a subsampled level
      // averages the real cells it covers (SubSampleGrid::forEachSubcell,
      // virtualgrid.hpp:576-607, which clips at the grid's far edge)
      auto eval = [&](int k, size_t i) -> double { return evalVirtual(v, f[k], i); };
      for (int iz = 0; iz < n; ++iz) {
        for (int iy = 0; iy < n; ++iy)
          for (int ix = 0; ix < n; ++ix) {
            size_t i = vIndex(v, ix, iy, iz);
            size_t fi = size_t(iy) * n + ix;
            for (int a3 = 0; a3 < 3; ++a3) {
              double p = eval(1 + a3, i), vel = p * velRatio;
              if (second) {
                double p2 = eval(4 + a3, i);
                p += p2;
                vel += 2 * velRatio * p2;
              }
              plane[a3][fi] = float(vel * velFactor);
              plane[3 + a3][fi] = float(p);
            }
            plane[6][fi] = float(eval(0, i));
            float m = inMask(flagged, l, i);
            plane[7][fi] = m;
            plane[8][fi] = float(pvarValue * m);
            ids[fi] = i + iordOffset;
          }
        int32_t rl = int32_t(n2 * 4), il = int32_t(n2 * 8);
        for (int k = 0; k < 9; ++k) {
          out[k].write((const char *)&rl, 4);
          out[k].write((const char *)plane[k].data(), n2 * 4);
          out[k].write((const char *)&rl, 4);
        }
        out[9].write((const char *)&il, 4);
        out[9].write((const char *)ids.data(), n2 * 8);
        out[9].write((const char *)&il, 4);
      }
      iordOffset += g.n3();
    }
  }

  // --- diagnostics ---
  // dumpGridData (ic.hpp:852-883)
  void dumpGrid(int level) {
    if (level < 0 || level >= int(grids.size())) throw std::runtime_error("Specified level does not exist.");
    ensureDrawn();
    int n = grids[level].n;
    std::vector<double> d(size_t(n) * n * n);
    check(gic_copy_field(c(), GIC_FIELD_OVERDENSITY, level, GIC_REAL_SPACE, d.data(), GIC_DTYPE_F64));
    saveNpy(outdir + "/grid-" + std::to_string(level) + ".npy", n, d);
    std::ofstream inf(outdir + "/grid-info-" + std::to_string(level) + ".txt");
    const Grid &g = grids[level];
    inf << g.ox << " " << g.oy << " " << g.oz << " " << g.size() << std::endl;
    inf << "The line above contains information about grid level " << level << std::endl;
    inf << "It gives the x-offset, y-offset and z-offset of the low-left corner and also the box length" << std::endl;
  }

  // dumpVelocityX (ic.hpp:920-923): the level's first Zel'dovich offset field
  // (zeldovich.hpp:151-253) as vx-<level>.npy, generating the output first
  void dumpVx(int level) {
    if (level < 0 || level >= int(grids.size())) throw std::runtime_error("Specified level does not exist.");
    generate();
    int n = grids[level].n;
    std::vector<double> d(size_t(n) * n * n);
    check(gic_copy_field(c(), GIC_FIELD_PSI_X, level, GIC_REAL_SPACE, d.data(), GIC_DTYPE_F64));
    saveNpy(outdir + "/vx-" + std::to_string(level) + ".npy", n, d);
    std::ofstream inf(outdir + "/vx-info-" + std::to_string(level) + ".txt");
    const Grid &g = grids[level];
    inf << g.ox << " " << g.oy << " " << g.oz << " " << g.size() << std::endl;
    inf << "The line above contains information about grid level " << level << std::endl;
    inf << "It gives the x-offset, y-offset and z-offset of the low-left corner and also the box length" << std::endl;
  }

  // dumpPS / dumpPowerSpectrum (ic.hpp:943-963, cosmology/parameters.hpp:112-205):
  // 100 log bins between 2pi/L and pi N/L; the GPU returns per-|k|^2 sums.
  void dumpPS(int level) {
    if (level < 0 || level >= int(grids.size())) throw std::runtime_error("Specified level does not exist.");
    ensureDrawn();
    const Grid &g = grids[level];
    long h = g.n / 2;
    size_t nm = size_t(3 * h * h + 1);
    std::vector<double> sum(nm), cnt(nm), theory(nm);
    check(gic_mode_power(c(), level, sum.data(), cnt.data(), theory.data(), nm));
    const int nBins = 100;
    std::vector<double> inBin(nBins, 0), kbin(nBins, 0), Gx(nBins, 0), Px(nBins, 0);
    double boxLength = g.size();
    double kmax = M_PI / boxLength * double(g.n), kmin = 2.0f * M_PI / boxLength;
    double dklog = std::log10(kmax / kmin) / nBins, kw = 2.0f * M_PI / boxLength;
    for (size_t m = 0; m < nm; ++m) {
      if (cnt[m] == 0) continue;
      double k = std::sqrt(double(m)) * kw;
      int idx = (int)((1.0f / dklog * std::log10(k / kmin)));
      if (k >= kmin && k < kmax) {
        Gx[idx] += sum[m];
        Px[idx] += cnt[m] * theory[m];
        kbin[idx] += cnt[m] * k;
        inBin[idx] += cnt[m];
      }
    }
    double kwn = 2. * M_PI / boxLength;
    double psnorm = 1 / (kwn * kwn * kwn / std::pow(2.f * M_PI, 3.f) * double(g.n3()));
    std::string fn = outputPath() + "_" + char('0' + level) + ".ps";
    std::ofstream ofs(fn);
    for (int i = 0; i < nBins; ++i) {
      if (inBin[i] > 0) {
        ofs << std::setw(16) << std::pow(10., std::log10(kmin) + dklog * (i + 0.5)) << std::setw(16)
            << kbin[i] / inBin[i] << std::setw(16) << (Px[i] / inBin[i]) * psnorm << std::setw(16)
            << (Gx[i] / inBin[i]) * psnorm << std::setw(16) << inBin[i] << std::endl;
      }
    }
  }

  // saveFieldTipsyArray (io/tipsy.hpp:38-52): the overdensity at every
  // particle in mapper order, as float, after a 4-byte count
  void dumpTipsyArray(const std::string &fn) {
    generate();
    auto blocks = layoutBlocks(fmt == Fmt::tipsy);
    size_t total = 0;
    for (auto &b : blocks) total += b.count();
    std::ofstream f(fn, std::ios::binary);
    uint32_t len = uint32_t(total);
    f.write((const char *)&len, 4);
    std::vector<std::vector<double>> delta(grids.size());
    for (auto &b : blocks) {
      for (int l : {b.level, b.strays ? b.level - 1 : b.level})
        if (delta[l].empty()) delta[l] = levelField(GIC_FIELD_OVERDENSITY, l);
      VLevel v = blockLevel(b), hi = resampledLevel(b.level, b.factor);
      auto fl = [&](int l) -> const std::vector<double> & { return delta[l]; };
      std::vector<float> vals(b.count());
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < vals.size(); ++i) vals[i] = float(evalOut(v, hi, fl, b.cells[i]));
      f.write((const char *)vals.data(), vals.size() * 4);
    }
  }
};

// ---------------------------------------------------------------- dispatch -
// Argument parsing with the reference's rules (tools/parser.hpp:40-105):
// "Insufficient arguments" at the end of the line, "Error while parsing
// arguments" on a failed conversion, "Too many arguments" for trailing
// non-comment text -- all checked before the command runs.
template <typename T> T parseArg(std::istream &in) {
  T v{};
  if (in.eof()) throw std::runtime_error("Insufficient arguments");
  in >> v;
  if (in.fail()) throw std::runtime_error("Error while parsing arguments");
  return v;
}

void consumeComments(std::istream &in) {
  std::string s;
  in >> s;
  if (s[0] == '#' || s[0] == '%') {
    while (!in.eof()) in >> s;
  } else if (!s.empty()) {
    throw std::runtime_error("Too many arguments");
  }
}

struct Dispatcher {
  Generator &ic;
  std::ostream &echo;   // IC_output.params: the commands as parsed
  std::map<std::string, std::function<void(std::istream &, const std::string &)>> routes;

  template <typename... A> void add(const std::string &name, std::function<void(A...)> f) {
    std::string k = name;
    std::transform(k.begin(), k.end(), k.begin(), ::tolower);
    std::ostream *out = &echo;
    routes[k] = [f, out](std::istream &in, const std::string &cmd) {
      std::tuple<A...> args{parseArg<A>(in)...};   // braced init: left-to-right
      consumeComments(in);
      *out << cmd << " ";
      std::apply([out](const A &...v) { ((*out << v << " "), ...); }, args);
      *out << std::endl;
      std::apply(f, args);
    };
  }

  Dispatcher(Generator &g, std::ostream &e) : ic(g), echo(e) {
    using D = std::function<void(double)>;
    auto cosmo = [this](double Generator::*field) {
      return D([this, field](double v) {
        ic.*field = v;
        ic.pushCosmology();
      });
    };
    add("TCMB", cosmo(&Generator::TCMB));
    add("Om", cosmo(&Generator::Om));
    add("Ob", cosmo(&Generator::Ob));
    add("Ol", cosmo(&Generator::Ol));
    add("hubble", cosmo(&Generator::hubble));
    add("s8", cosmo(&Generator::s8));
    add("ns", cosmo(&Generator::ns));
    add("zin", cosmo(&Generator::z));
    add("z", cosmo(&Generator::z));
    auto seed = [this](int mode) {
      return std::function<void(int)>([this, mode](int s) { ic.check(gic_seed(ic.c(), (unsigned long)s, mode)); });
    };
    add("seed", seed(GIC_SEED_REAL_SPACE));
    add("random_seed_real_space", seed(GIC_SEED_REAL_SPACE));
    add("seedfourier", seed(GIC_SEED_FOURIER_SERIAL));
    add("random_seed_serial", seed(GIC_SEED_FOURIER_SERIAL));
    add("seedfourier_parallel", seed(GIC_SEED_FOURIER_PARALLEL));
    add("random_seed", seed(GIC_SEED_FOURIER_PARALLEL));
    add("seedfourier_reverse", seed(GIC_SEED_FOURIER_REVERSE));
    add("seed_field_fourier_reverse", seed(GIC_SEED_FOURIER_REVERSE));
    using S = std::function<void(std::string)>;
    add("camb", S([this](std::string f) { ic.setCamb(f); }));
    add("powerlaw_amplitude", D([this](double a) { ic.setPowerLaw(a); }));
    add("outdir", S([this](std::string s) { ic.outdir = s; }));
    add("outname", S([this](std::string s) { ic.outname = s; }));
    add("outformat", S([this](std::string s) {
          // io::OutputFormat is read as a number or as a name (io.hpp)
          static const std::map<std::string, Fmt> names{
              {"gadget2", Fmt::gadget2}, {"gadget3", Fmt::gadget3},     {"tipsy", Fmt::tipsy},
              {"grafic", Fmt::grafic},   {"gadgethdf", Fmt::gadgethdf}, {"swift", Fmt::swift}};
          char *end = nullptr;
          long v = std::strtol(s.c_str(), &end, 10);
          if (!s.empty() && end && *end == 0) {
            if (v < 1 || v > 7) throw std::runtime_error("Unknown output format");
            ic.fmt = Fmt(v);
          } else {
            std::string l = s;
            std::transform(l.begin(), l.end(), l.begin(), ::tolower);
            auto it = names.find(l);
            if (it == names.end()) throw std::runtime_error("Error while parsing arguments");
            ic.fmt = it->second;
          }
        }));
    std::function<void(double, int)> base = [this](double box, int n) { ic.initBaseGrid(box, n); };
    add("base_grid", base);
    add("basegrid", base);
    std::function<void(int, int)> zoom = [this](int f, int n) { ic.initZoomGrid(f, n); };
    add("zoom_grid", zoom);
    add("zoomgrid", zoom);
    std::function<void(int, int, int, int, int)> zoomAt = [this](int x, int y, int z, int f, int n) {
      ic.initZoomGridWithLowLeftCornerAt(x, y, z, f, n);
    };
    add("zoom_grid_with_origin_at", zoomAt);
    add("zoomgrid_with_origin_at", zoomAt);
    std::function<void(double, double, double)> centre = [this](double x, double y, double z) {
      ic.x0 = x;
      ic.y0 = y;
      ic.z0 = z;
    };
    add("centre", centre);
    add("center", centre);
    std::function<void(long)> centreOn = [this](long id) {
      if (ic.grids.empty()) throw std::runtime_error("No grid has yet been initialised");
      if (id < 0 || size_t(id) >= ic.grids[0].n3()) throw std::runtime_error("Particle ID out of range");
      auto p = ic.grids[0].centroid(size_t(id));
      ic.x0 = p[0];
      ic.y0 = p[1];
      ic.z0 = p[2];
    };
    add("centre_on", centreOn);
    add("center_on", centreOn);
    using F = std::function<void(float)>;
    add("select_sphere", F([this](float r) { ic.selectSphere(r); }));
    add("select_cube", F([this](float s) { ic.selectCube(s); }));
    add("select_ellipse",
        std::function<void(float, float, float)>([this](float a, float b, float c) { ic.selectEllipse(a, b, c); }));
    add("select_nearest", std::function<void()>([this] { ic.selectNearest(); }));
    S idfile = [this](std::string f) { ic.loadIDFile(f, false); };
    S mergeid = [this](std::string f) { ic.loadIDFile(f, true); };
    add("id_file", idfile);
    add("idfile", idfile);
    add("merge_id_file", mergeid);
    add("append_idfile", mergeid);
    S dumpId = [this](std::string f) { ic.dumpIDFile(f); };
    add("dump_id_file", dumpId);
    add("dump_idfile", dumpId);
    std::function<void(int)> zl = [this](int l) { ic.zeroLevel(l); };
    add("zero_level", zl);
    add("zerolevel", zl);
    // zero_level_field <level> <field>: one output field here (no baryon transfer function)
    add("zero_level_field", std::function<void(int, int)>([this](int l, int f) {
          if (f != 0) throw std::runtime_error("Field " + std::to_string(f) + " does not exist");
          ic.zeroLevel(l);
        }));
    add("expand_flagged_region", std::function<void(int)>([this](int nc) { ic.expandFlaggedRegion(nc); }));
    // setNumNeighbours (ic.hpp:1268-1271): the SPH smoothing scale, cell size x
    // (3 n / 4 pi)^(1/3); only the HDF5 writer (gadgethdf.hpp:233) reads it, so
    // it is kept but changes no output here
    std::function<void(int)> nn = [this](int v) { ic.smoothNorm = std::pow(3.0 * v / (4.0 * M_PI), 1.0 / 3.0); };
    add("num_neighbours", nn);
    add("num_neighbors", nn);
    add("eps_norm", D([this](double v) { ic.epsNorm = v; }));
    add("gadget_num_files", std::function<void(int)>([this](int n) {
          if (n < 1) throw std::runtime_error("Number of gadget files must be positive");
          ic.nGadgetFiles = n;
        }));
    add("gadget_particle_type", std::function<void(int)>([this](int t) { ic.setGadgetParticleType(unsigned(t)); }));
    add("gadget_flagged_particle_type",
        std::function<void(int)>([this](int t) { ic.setFlaggedGadgetParticleType(unsigned(t)); }));
    add("velocity_offset", std::function<void(double, double, double)>([this](double vx, double vy, double vz) {
          double c = std::pow(ic.a(), -0.5);   // km/s -> gadget units (ic.hpp:285-289)
          ic.velOffset[0] = vx * c;
          ic.velOffset[1] = vy * c;
          ic.velOffset[2] = vz * c;
        }));
    add("done", std::function<void()>([this] { ic.done(); }));
    add("calculate", S([this](std::string n) { ic.calculate(n); }));
    add("modify", std::function<void(std::string, std::string, float)>(
                      [this](std::string n, std::string t, float v) { ic.modify(n, t, v); }));
    add("apply_modifications", std::function<void()>([this] { ic.applyModifications(); }));
    add("clear_modifications", std::function<void()>([this] { ic.clearModifications(); }));
    add("chi2", std::function<void()>([this] { ic.chi2(); }));
    add("filtering_scale", D([this](double v) { ic.varianceScale = v; }));
    add("lpt_order", std::function<void(int)>([this](int o) {
          ic.check(gic_set_lpt_order(ic.c(), o));
          ic.lptOrder = o;
        }));
    // grafic hierarchy and mask (ic.hpp:301-350, main.cpp:59-74)
    add("subsample", std::function<void(int)>([this](int f) { ic.setSubsample(f); }));
    add("supersample", std::function<void(int)>([this](int f) { ic.setSupersample(f); }));
    add("supersample_gas", std::function<void(int)>([this](int f) { ic.setSupersampleGas(f); }));
    add("strays_on", std::function<void()>([this]() { ic.allowStrays = true; }));
    add("adapt_mask", std::function<void(int)>([this](int n) { ic.adaptMask(n); }));
    add("autopad", std::function<void(int)>([this](int n) {
      if (n < 0) throw std::runtime_error("autopad takes a number of cells");
      ic.autopad = size_t(n);
      ic.refreshMapper();
    }));
    add("pvar", std::function<void(double)>([this](double v) { ic.pvarValue = v; }));
    add("center_output", std::function<void()>([this]() { ic.centerOnTargetRegion = true; }));
    add("centre_output", std::function<void()>([this]() { ic.centerOnTargetRegion = true; }));
    add("center_grafic_output", std::function<void()>([this]() { ic.centerOnTargetRegion = true; }));
    add("dump_mask", std::function<void()>([this]() { ic.dumpMask(); }));
    add("dump_grid", std::function<void(int)>([this](int l) { ic.dumpGrid(l); }));
    add("dump_grid_fourier", std::function<void(int)>([this](int l) { ic.dumpGridFourier(l); }));
    add("dump_vx", std::function<void(int)>([this](int l) { ic.dumpVx(l); }));
    std::function<void()> fixp = [this]() { ic.check(gic_set_exact_power_spectrum(ic.c(), 1)); };
    add("fix_power", fixp);
    add("exact_power_spectrum_enforcement", fixp);
    add("reverse", std::function<void()>([this]() { ic.reverse(); }));
    add("splice", std::function<void(int)>([this](int s) { ic.splice(s); }));
    add("reverse_small_k", std::function<void(double)>([this](double k) { ic.reverseSmallK(k); }));
    add("import_level", std::function<void(int, std::string)>([this](int l, std::string f) { ic.importLevel(l, f); }));
    add("dump_ps", std::function<void(int)>([this](int l) { ic.dumpPS(l); }));
    add("dump_tipsy", S([this](std::string f) { ic.dumpTipsyArray(f); }));
  }

  // ClassDispatch::run (tools/parser.hpp:159-185)
  void run(std::istream &in) {
    std::string s;
    in >> s;
    if (s[0] == '#' || s[0] == '%') return;
    std::transform(s.begin(), s.end(), s.begin(), ::tolower);
    auto it = routes.find(s);
    if (it == routes.end()) throw std::runtime_error("Unknown command");
    it->second(in, s);
  }

  // run_loop (tools/parser.hpp:189-210)
  void runLoop(std::istream &in) {
    std::string str;
    int line = 0;
    while (std::getline(in, str)) {
      line += 1;
      if (str.empty()) continue;
      std::istringstream ss(str);
      try {
        run(ss);
      } catch (std::runtime_error &e) {
        log() << "Error \"" << e.what() << "\" on line " << line << " (\"" << str << "\")" << std::endl;
        exit(1);
      }
    }
  }
};

}  // namespace fe

int main(int argc, char **argv) {
  bool f32 = false;
  std::string fname;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-f") f32 = true;
    else if (a == "-c" && i + 1 < argc) ++i;   // transfer-function cache size: no cache to size here
    else fname = a;
  }
  if (fname.empty()) {
    std::cout << "Usage: genetIC-hip paramfile [-f]" << std::endl << std::endl
              << " The paramfile is a text file of genetIC commands." << std::endl
              << " With -f the fields are held in float (32-bit) instead of double (64-bit)." << std::endl;
    return 0;
  }
  std::ifstream in(fname);
  if (!in.is_open()) {
    fe::log() << "Error: could not open parameter file " << fname << std::endl;
    return 255;
  }
  // commands run relative to the paramfile's directory (main.cpp:205-213)
  std::string dir = fname.find('/') == std::string::npos ? "." : fname.substr(0, fname.rfind('/'));
  if (chdir(dir.c_str()) != 0) {
    fe::log() << "Error: cannot change directory to " << dir << std::endl;
    return 1;
  }
  std::ofstream echo("IC_output.params");
  if (!echo.is_open()) {
    fe::log() << "Error: could not open output file" << std::endl;
    return 1;
  }
  try {
    fe::Generator gen(f32);
    fe::Dispatcher d(gen, echo);
    d.runLoop(in);
  } catch (std::exception &e) {
    fe::log() << "Error \"" << e.what() << "\"" << std::endl;
    return 1;
  }
  return 0;
}
