// oeb_bench.cpp -- benchmark and verification driver for the ten stencils on
// the GPU, through the drop-in headers (util.h, kernels/<name>.h).
//
//   oeb_bench <kernel> [ALGO [domain_size domain_height]] [--count N]
//
// <kernel> is one of the ten stencils.  ALGO and the two extents follow the
// reference drivers' command line (src/<name>.cpp:16-29); every ALGO runs the
// same fused HIP kernel, since the reference's variants are CPU loop-fusion
// studies with one result.  The inputs are the reference drivers' own:
// fillMath with their per-field parameters, or initValue constants
// (src/<name>.cpp).  Unlike the reference's hdiffsa driver, which passes
// (in, mask, out) in the place of (out, in, mask) (src/hdiffsa.cpp:56), the
// table below passes hdiffsa's arguments in the order of its signature
// (kernels/hdiffsa.h:4).
//
// Output: the reference's header lines and timer line (timer.show_all), then
// one verification line per output and a JSON line with the mean kernel time
// from HIP events, Gpts/s, the fraction of the 8 TB/s HBM peak at the
// stencil's algorithmic bytes per point (in the build's precision), the output checksums (sum over each
// output's valid region) and the verification result.  The reference prints
// "-> starting verification" and verifies nothing (src/hdiffsa.cpp:50).  Here
// every output is checked after the timed loop: finite on its valid region,
// and every allocated point outside that region still holds the value it
// was initialised with (nothing written outside the region).  The JSON line
// carries each output's region sum and sum of magnitudes.  The exit status
// is 3 when a check fails.  The tests compare the checksums with the oracle.
//
// Built twice: with ElementType double (bin/oeb_bench) and, with
// -DOEB_BENCH_FLOAT, float (bin/oeb_bench_f32, the FP32 forms).
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>

int64_t domain_size = 64;
int64_t domain_height = 60;
int64_t halo_width = 3;
int64_t count = 1000;

#ifdef OEB_BENCH_FLOAT
typedef float ElementType;
#else
typedef double ElementType;
#endif

// uvbke's time step is a double in the reference driver whatever ElementType
// is (src/uvbke.cpp:11-12).
const double dt5 = 0.5 * 225.0;

#include "util.h"
#include "fastwavesuv.h"
#include "hadvuv.h"
#include "hadvuv5th.h"
#include "hdiffsa.h"
#include "hdiffsmag.h"
#include "laplace.h"
#include "nh_p_grad.h"
#include "p_grad_c.h"
#include "uvbke.h"
// fvtp2d's PPM constants are driver macros in the reference (src/fvtp2d.cpp:11-12).
#define p1 ElementType(0.583333)
#define p2 ElementType(0.083333)
#include "fvtp2d.h"
#include "oeb_driver.h"

namespace {

// How a field is initialised.
struct Fill {
  enum Kind { kMath, kConst, kNone } kind;
  double a, b, c, d, e, f;  // fillMath parameters; a is the constant of kConst
};
// The fillMath parameter sets of the reference drivers.
constexpr Fill kA{Fill::kMath, 1.1, 2.0, 1.5, 2.8, 2.0, 4.1};
constexpr Fill kB{Fill::kMath, 8.0, 9.4, 1.5, 1.7, 2.0, 3.5};
constexpr Fill kC{Fill::kMath, 5.0, 8.0, 1.5, 7.1, 2.0, 4.3};
constexpr Fill kD{Fill::kMath, 3.2, 7.0, 2.5, 6.1, 3.0, 2.3};
constexpr Fill kE{Fill::kMath, 4.5, 5.0, 2.5, 2.1, 3.0, 2.3};
constexpr Fill kF{Fill::kMath, 1.0, 3.3, 1.5, 1.5, 2.0, 4.0};
constexpr Fill kG{Fill::kMath, 4.0, 1.7, 1.5, 6.3, 2.0, 1.4};
constexpr Fill kScratch{Fill::kNone, 0, 0, 0, 0, 0, 0};
constexpr Fill value(double v) { return Fill{Fill::kConst, v, 0, 0, 0, 0, 0}; }

// A 3-D output and its valid region [n0i, n+n1i) x [n0j, n+n1j) x [0, nz).
struct Out {
  const char *name;
  int i0, i1, j0, j1;
};

struct Fields {
  std::map<std::string, Storage3D> f3;
  std::map<std::string, Storage1D> f1;
  Storage3D &operator[](const char *n) { return f3.at(n); }
  Storage1D &one(const char *n) { return f1.at(n); }
};

struct Kernel {
  const char *name;
  int bytes_f64, bytes_f32;  // algorithmic bytes per point (3-D inputs + outputs; FP64 islands count 8)
  std::vector<std::pair<const char *, Fill>> fields3;   // 3-D fields, outputs included
  std::vector<std::pair<const char *, Fill>> fields1;   // 1-D fields
  std::vector<Out> outputs;
  std::function<bool(Fields &, int)> run;               // one call of variant ALGO; false if it has none
};

// Calls variant ALGO of a stencil with the four reference variants.
#define OEB_FOUR_VARIANTS(fn, ...)                \
  switch (algo) {                                 \
    case DEFAULT: fn(__VA_ARGS__); return true;                  \
    case FULL_FUSION: fn##_fullfusion(__VA_ARGS__); return true; \
    case PARTIAL_FUSION: fn##_partialfusion(__VA_ARGS__); return true; \
    case OPENMP: fn##_openmp(__VA_ARGS__); return true;          \
  }                                               \
  return false

const ElementType kEddlat = ElementType(ldexpl(1.0, -11));  // src/hdiffsmag.cpp:48-52, src/hadvuv.cpp:53-54
const ElementType kEddlon = ElementType(ldexpl(1.5, -11));

std::vector<Kernel> kernels() {
  std::vector<Kernel> t;
  const Out interior_u{"uout", 0, 0, 0, 0}, interior_v{"vout", 0, 0, 0, 0};
  t.push_back({"laplace", 16, 8, {{"in", kA}, {"out", value(0.0)}}, {}, {{"out", 0, 0, 0, 0}},
               [](Fields &f, int algo) { OEB_FOUR_VARIANTS(laplace, f["in"], f["out"]); }});
  t.push_back({"hdiffsa", 24, 12,
               {{"in", kA}, {"mask", value(0.025)}, {"out", value(0.0)}, {"lap", kScratch}, {"flx", kScratch},
                {"fly", kScratch}},
               {{"crlato", kB}, {"crlatu", kB}},
               {{"out", 0, 0, 0, 0}},
               [](Fields &f, int algo) {
                 OEB_FOUR_VARIANTS(hdiffsa, f["out"], f["in"], f["mask"], f.one("crlato"), f.one("crlatu"),
                                   f["lap"], f["flx"], f["fly"]);
               }});
  t.push_back({"hdiffsmag", 40, 20,
               {{"uin", kA}, {"vin", kB}, {"mask", value(0.025)}, {"uout", value(0.0)}, {"vout", value(0.0)},
                {"T_sqr_s", value(0.0)}, {"S_sqr_uv", value(0.0)}},
               {{"crlavo", kB}, {"crlavu", kC}, {"crlato", kC}, {"crlatu", kD}, {"acrlat0", kE}},
               {interior_u, interior_v},
               [](Fields &f, int algo) {
                 OEB_FOUR_VARIANTS(hdiffsmag, f["uout"], f["vout"], f["uin"], f["vin"], f["mask"], f.one("crlavo"),
                                   f.one("crlavu"), f.one("crlato"), f.one("crlatu"), f.one("acrlat0"),
                                   f["T_sqr_s"], f["S_sqr_uv"], kEddlat, kEddlon, ElementType(0.025),
                                   ElementType(0.01));
               }});
  const std::vector<std::pair<const char *, Fill>> hadv_scratch = {
      {"uatupos", kScratch}, {"vatupos", kScratch}, {"uatvpos", kScratch}, {"vatvpos", kScratch},
      {"uavg", kScratch},    {"vavg", kScratch},    {"ures", kScratch},    {"vres", kScratch}};
  auto hadv_fields = [&](Fill u, Fill v) {
    std::vector<std::pair<const char *, Fill>> r = {{"uin", u}, {"vin", v}, {"uout", value(0.0)},
                                                    {"vout", value(0.0)}};
    r.insert(r.end(), hadv_scratch.begin(), hadv_scratch.end());
    return r;
  };
  t.push_back({"hadvuv", 32, 16, hadv_fields(kA, kB),
               {{"acrlat0", kB}, {"acrlat1", kC}, {"tgrlatda0", kC}, {"tgrlatda1", kD}},
               {interior_u, interior_v},
               [](Fields &f, int algo) {
                 OEB_FOUR_VARIANTS(hadvuv, f["uout"], f["vout"], f["uin"], f["vin"], f.one("acrlat0"),
                                   f.one("acrlat1"), f.one("tgrlatda0"), f.one("tgrlatda1"), f["uatupos"],
                                   f["vatupos"], f["uatvpos"], f["vatvpos"], f["uavg"], f["vavg"], f["ures"],
                                   f["vres"], kEddlat, kEddlon);
               }});
  t.push_back({"hadvuv5th", 32, 16, hadv_fields(kF, kA),
               {{"acrlat0", kA}, {"acrlat1", kB}, {"tgrlatda0", kB}, {"tgrlatda1", kC}},
               {interior_u, interior_v},
               [](Fields &f, int algo) {
                 OEB_FOUR_VARIANTS(hadvuv5th, f["uout"], f["vout"], f["uin"], f["vin"], f.one("acrlat0"),
                                   f.one("acrlat1"), f.one("tgrlatda0"), f.one("tgrlatda1"), f["uatupos"],
                                   f["vatupos"], f["uatvpos"], f["vatvpos"], f["uavg"], f["vavg"], f["ures"],
                                   f["vres"], kEddlat, kEddlon);
               }});
  t.push_back({"fastwavesuv", 80, 48,
               {{"uin", kF}, {"utens", kA}, {"vin", kA}, {"vtens", kB}, {"ppuv", kB}, {"wgtfac", kC}, {"hhl", kC},
                {"rho", kD}, {"uout", value(0.0)}, {"vout", value(0.0)}, {"ppgk", kScratch}, {"ppgc", kScratch},
                {"ppgu", kScratch}, {"ppgv", kScratch}},
               {{"fx", kE}},
               {interior_u, interior_v},
               [](Fields &f, int algo) {
                 // src/fastwavesuv.cpp:52-53
                 OEB_FOUR_VARIANTS(fastwavesuv, f["uout"], f["vout"], f["uin"], f["vin"], f["utens"], f["vtens"],
                                   f["wgtfac"], f["ppuv"], f["hhl"], f["rho"], f.one("fx"), f["ppgk"], f["ppgc"],
                                   f["ppgu"], f["ppgv"], ElementType(ldexpl(1.0, -11)), ElementType(10.0));
               }});
  t.push_back({"uvbke", 48, 24,
               {{"uc", kF}, {"vc", kA}, {"cosa", kG}, {"rsina", kB}, {"ub", value(-1.0)}, {"vb", value(-1.0)}},
               {},
               {{"ub", 0, 0, 0, 0}, {"vb", 0, 0, 0, 0}},
               [](Fields &f, int algo) {
                 OEB_FOUR_VARIANTS(uvbke, f["ub"], f["vb"], f["uc"], f["vc"], f["cosa"], f["rsina"]);
               }});
  t.push_back({"p_grad_c", 72, 44,
               {{"uin", kF}, {"vin", kA}, {"rdxc", kA}, {"rdyc", kB}, {"delpc", kB}, {"gz", kC}, {"pkc", kC},
                {"uout", value(0.0)}, {"vout", value(0.0)}, {"wk", kScratch}},
               {},
               {interior_u, interior_v},
               [](Fields &f, int algo) {
                 OEB_FOUR_VARIANTS(p_grad_c, f["uout"], f["vout"], f["uin"], f["vin"], f["rdxc"], f["rdyc"],
                                   f["delpc"], f["gz"], f["pkc"], f["wk"], ElementType(0.1));
               }});
  t.push_back({"nh_p_grad", 80, 52,
               {{"uin", kF}, {"vin", kA}, {"rdx", kA}, {"rdy", kB}, {"gz", kB}, {"pp", kC}, {"pk3", kC},
                {"wk1", kD}, {"uout", value(0.0)}, {"vout", value(0.0)}, {"wk", kScratch}, {"du", kScratch},
                {"dv", kScratch}},
               {},
               {interior_u, interior_v},
               [](Fields &f, int algo) {
                 OEB_FOUR_VARIANTS(nh_p_grad, f["uout"], f["vout"], f["uin"], f["vin"], f["rdx"], f["rdy"],
                                   f["gz"], f["pp"], f["pk3"], f["wk1"], f["wk"], f["du"], f["dv"],
                                   ElementType(0.1));
               }});
  t.push_back({"fvtp2d", 112, 56,
               {{"q", kF}, {"crx", kA}, {"cry", kA}, {"ra_x", kB}, {"ra_y", kB}, {"xfx", kC}, {"yfx", kC},
                {"area", kD}, {"q_i", value(-1.0)}, {"q_j", value(-1.0)}, {"fx1", value(-1.0)},
                {"fx2", value(-1.0)}, {"fy1", value(-1.0)}, {"fy2", value(-1.0)}, {"fxx", kScratch},
                {"fyy", kScratch}, {"al", kScratch}, {"almq", kScratch}, {"br", kScratch}, {"b0", kScratch},
                {"smt5", kScratch}},
               {},
               // Extended output regions (kernels/fvtp2d.h:8-113).
               {{"q_i", -3, 2, 0, 0}, {"q_j", 0, 0, -3, 2}, {"fx1", 0, 0, 0, 0}, {"fx2", 0, 1, -3, 2},
                {"fy1", 0, 0, 0, 0}, {"fy2", -3, 2, 0, 1}},
               [](Fields &f, int algo) {
                 // The reference has DEFAULT and fullfusion only (src/fvtp2d.cpp:78-98).
                 switch (algo) {
                   case DEFAULT:
                     fvtp2d(f["q_i"], f["q_j"], f["fx1"], f["fx2"], f["fy1"], f["fy2"], f["q"], f["crx"], f["cry"],
                            f["ra_x"], f["ra_y"], f["xfx"], f["yfx"], f["area"], f["fxx"], f["fyy"], f["al"],
                            f["almq"], f["br"], f["b0"], f["smt5"]);
                     return true;
                   case FULL_FUSION:
                     fvtp2d_fullfusion(f["q_i"], f["q_j"], f["fx1"], f["fx2"], f["fy1"], f["fy2"], f["q"], f["crx"],
                                       f["cry"], f["ra_x"], f["ra_y"], f["xfx"], f["yfx"], f["area"], f["fxx"],
                                       f["fyy"], f["al"], f["almq"], f["br"], f["b0"], f["smt5"]);
                     return true;
                 }
                 return false;
               }});
  return t;
}

void fill(Storage3D &s, const Fill &r) {
  if (r.kind == Fill::kMath) fillMath(r.a, r.b, r.c, r.d, r.e, r.f, s, domain_size, domain_height);
  if (r.kind == Fill::kConst) initValue(s, ElementType(r.a), domain_size, domain_height);
}

void fill(Storage1D &s, const Fill &r) {
  if (r.kind == Fill::kMath) fillMath(r.a, r.b, r.c, r.d, r.e, r.f, s, domain_size, domain_height);
}

// Checks one output after the timed loop: finite on its valid region, and
// every other allocated point still equal to its initial value.  Returns the
// region's sum and sum of magnitudes; on a failure sets *ok = false and
// reports the first bad point.
std::pair<double, double> verify(const char *kname, const Out &o, Storage3D &s, ElementType init, bool *ok) {
  toHost(s);
  const Storage3D &c = s;
  const int64_t h = halo_width, n = domain_size, nz = domain_height;
  double sum = 0.0, abs_sum = 0.0;
  int64_t bad = 0;
  for (int64_t k = -h; k < nz + h; ++k)
    for (int64_t j = -h; j < n + h; ++j)
      for (int64_t i = -h; i < n + h; ++i) {
        const ElementType v = c(i, j, k);
        const bool inside = k >= 0 && k < nz && i >= o.i0 && i < n + o.i1 && j >= o.j0 && j < n + o.j1;
        const bool good = inside ? std::isfinite(v) : std::memcmp(&v, &init, sizeof v) == 0;
        if (inside) {
          sum += double(v);
          abs_sum += std::fabs(double(v));
        }
        if (!good && bad++ == 0)
          std::printf("verification: %s.%s FAILED at (%lld, %lld, %lld): %.17g (%s)\n", kname, o.name,
                      (long long)i, (long long)j, (long long)k, double(v),
                      inside ? "not finite" : "written outside the output region");
      }
  if (bad == 0) std::printf("verification: %s.%s ok, region sum %.17g\n", kname, o.name, sum);
  *ok = *ok && bad == 0;
  return {sum, abs_sum};
}

int usage() {
  std::printf("Usage: ./oeb_bench <kernel> [ALGO [domain_size domain_height]] [--count N]\n  kernels:");
  for (const auto &k : kernels()) std::printf(" %s", k.name);
  std::printf("\n");
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<char *> rest;
  for (int a = 1; a < argc; ++a) {
    if (std::strcmp(argv[a], "--count") == 0 && a + 1 < argc) {
      count = std::atoll(argv[++a]);
    } else {
      rest.push_back(argv[a]);
    }
  }
  if (rest.empty()) return usage();
  const std::vector<Kernel> table = kernels();
  const Kernel *k = nullptr;
  for (const auto &e : table)
    if (e.name == std::string(rest[0])) k = &e;
  if (k == nullptr) return usage();
  // The reference's argument list after the kernel name.
  rest[0] = argv[0];
  const int algo = oeb_driver::parse(int(rest.size()), rest.data());

  const int64_t n1 = domain_size + 2 * halo_width;
  const std::array<int64_t, 3> sizes3 = {n1, n1, domain_height + 2 * halo_width};
  Fields f;
  std::map<std::string, ElementType> init;
  for (const auto &fs : k->fields3) {
    f.f3.emplace(fs.first, allocateStorage(sizes3));
    fill(f.f3.at(fs.first), fs.second);
    init[fs.first] = fs.second.kind == Fill::kConst ? ElementType(fs.second.a) : ElementType(0);
  }
  for (const auto &fs : k->fields1) {
    f.f1.emplace(fs.first, allocateStorage(n1));
    fill(f.f1.at(fs.first), fs.second);
  }

  std::cout << "-> starting verification" << std::endl;
  const std::string label = oeb_driver::label(k->name, algo);
  bool known = true;
  const double ms = oeb_driver::timed_loop(label.c_str(), [&] { known = k->run(f, algo); });
  if (!known) {
    std::cout << "Unknown ALGO" << std::endl;
    return 1;
  }
  timer.show_all();

  bool ok = true;
  std::vector<std::pair<const char *, std::pair<double, double>>> sums;
  for (const Out &o : k->outputs) sums.emplace_back(o.name, verify(k->name, o, f[o.name], init[o.name], &ok));
  oeb_driver::report(k->name, algo, ms, sizeof(ElementType) == 8 ? k->bytes_f64 : k->bytes_f32, sums, ok);
  for (auto &kv : f.f3) freeStorage(kv.second);
  for (auto &kv : f.f1) freeStorage(kv.second);
  return ok ? 0 : 3;
}
