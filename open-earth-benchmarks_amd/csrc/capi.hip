// capi.hip -- status strings, descriptor validation and error reporting
// shared by every entry point of liboeb.
#include <atomic>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <unordered_map>

#include "oeb_common.hpp"

namespace oeb {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int status, const std::string &msg) {
  set_error(msg);
  return status;
}

int check_domain(const char *kname, const oeb_domain *d) {
  if (d == nullptr) return fail(OEB_ERR_NULL, std::string(kname) + ": domain is NULL");
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) {
    std::ostringstream m;
    m << kname << ": domain extents must be >= 1 (nx=" << d->nx << " ny=" << d->ny << " nz=" << d->nz << ")";
    return fail(OEB_ERR_DOMAIN, m.str());
  }
  if (d->nx > (int64_t(1) << 24) || d->ny > (int64_t(1) << 24) || d->nz > (int64_t(1) << 24)) {
    return fail(OEB_ERR_DOMAIN, std::string(kname) + ": an extent exceeds 2^24");
  }
  return OEB_OK;
}

namespace {

template <class S, class T>
int check3_impl(const char *kname, const char *name, const S *s, const Box &b, Geo &g, bool first,
                const T **base) {
  std::ostringstream m;
  m << kname << ": array '" << name << "' ";
  if (s == nullptr || s->alignedPtr == nullptr) return fail(OEB_ERR_NULL, m.str() + "is NULL");
  const int64_t si = s->strides[2], sj = s->strides[1], sk = s->strides[0];
  if (si != 1) return fail(OEB_ERR_LAYOUT, m.str() + "must have i stride (strides[2]) 1");
  if (sj < 1 || sk < 1 || s->offset < 0 || s->sizes[0] < 1 || s->sizes[1] < 1 || s->sizes[2] < 1)
    return fail(OEB_ERR_LAYOUT, m.str() + "has a non-positive size or stride, or a negative offset");
  if (s->sizes[2] > sj || s->sizes[1] * sj > sk)
    return fail(OEB_ERR_LAYOUT, m.str() + "has strides smaller than its sizes (rows or planes overlap)");
  if (sk >= (int64_t(1) << 31))
    return fail(OEB_ERR_LAYOUT, m.str() + "has a k stride >= 2^31 elements (one plane must fit 32-bit indexing)");
  if (first) {
    g.sj = int(sj);
    g.sk = sk;
  } else if (sj != g.sj || sk != g.sk) {
    m << "has strides (k " << sk << ", j " << sj << ") that differ from the first array's (k " << g.sk << ", j "
      << g.sj << "); all 3-D arrays of one call must share a layout";
    return fail(OEB_ERR_LAYOUT, m.str());
  }
  // Position of element (0,0,0) inside the allocated box.
  const int64_t ok = s->offset / sk, rem = s->offset % sk;
  const int64_t oj = rem / sj, oi = rem % sj;
  const bool inside = b.ilo >= -oi && b.ihi <= s->sizes[2] - oi && b.jlo >= -oj && b.jhi <= s->sizes[1] - oj &&
                      b.klo >= -ok && b.khi <= s->sizes[0] - ok;
  if (!inside) {
    m << "does not cover the region the stencil touches: needs i [" << b.ilo << "," << b.ihi << ") j [" << b.jlo
      << "," << b.jhi << ") k [" << b.klo << "," << b.khi << "), allocation spans i [" << -oi << ","
      << s->sizes[2] - oi << ") j [" << -oj << "," << s->sizes[1] - oj << ") k [" << -ok << ","
      << s->sizes[0] - ok << ")";
    return fail(OEB_ERR_LAYOUT, m.str());
  }
  *base = s->alignedPtr + s->offset;
  return OEB_OK;
}

template <class S, class T>
int check1_impl(const char *kname, const char *name, const S *s, int64_t lo, int64_t hi, const T **base,
                int *stride) {
  std::ostringstream m;
  m << kname << ": 1-D array '" << name << "' ";
  if (s == nullptr || s->alignedPtr == nullptr) return fail(OEB_ERR_NULL, m.str() + "is NULL");
  const int64_t st = s->strides[0];
  if (st < 1 || st > (int64_t(1) << 20) || s->offset < 0 || s->sizes[0] < 1)
    return fail(OEB_ERR_LAYOUT, m.str() + "has a non-positive size or stride, or a negative offset");
  if (s->offset + lo * st < 0 || s->offset + (hi - 1) * st > (s->sizes[0] - 1) * st) {
    m << "does not cover indices [" << lo << "," << hi << ")";
    return fail(OEB_ERR_LAYOUT, m.str());
  }
  *base = s->alignedPtr + s->offset;
  *stride = int(st);
  return OEB_OK;
}

}  // namespace

int check3(const char *kname, const char *name, const oeb_s3 *s, const Box &b, Geo &g, bool first,
           const double **base) {
  return check3_impl(kname, name, s, b, g, first, base);
}
int check3(const char *kname, const char *name, const oeb_s3f *s, const Box &b, Geo &g, bool first,
           const float **base) {
  return check3_impl(kname, name, s, b, g, first, base);
}
int check1(const char *kname, const char *name, const oeb_s1 *s, int64_t lo, int64_t hi, const double **base,
           int *stride) {
  return check1_impl(kname, name, s, lo, hi, base, stride);
}
int check1(const char *kname, const char *name, const oeb_s1f *s, int64_t lo, int64_t hi, const float **base,
           int *stride) {
  return check1_impl(kname, name, s, lo, hi, base, stride);
}

void AliasSet::add_range(const char *nm, uintptr_t lo_b, uintptr_t hi_b, bool is_out) {
  if (n >= kMax) {
    overflow = true;
    return;
  }
  name[n] = nm;
  lo[n] = lo_b;
  hi[n] = hi_b;
  out[n] = is_out;
  ++n;
}

int AliasSet::check(const char *kname) const {
  if (overflow) return fail(OEB_ERR_ARG, std::string(kname) + ": too many arrays for the overlap check");
  for (int a = 0; a < n; ++a) {
    if (!out[a]) continue;
    for (int b = 0; b < n; ++b) {
      if (a == b) continue;
      if (lo[a] < hi[b] && lo[b] < hi[a]) {
        std::ostringstream m;
        m << kname << ": output '" << name[a] << "' overlaps " << (out[b] ? "output" : "input") << " '" << name[b]
          << "'";
        return fail(OEB_ERR_ALIAS, m.str());
      }
    }
  }
  // Every array must be memory a kernel may access: device memory, managed
  // memory or pinned host memory.  Pageable host memory would fault the GPU.
  for (int a = 0; a < n; ++a) {
    hipPointerAttribute_t attr;
    const hipError_t e = hipPointerGetAttributes(&attr, reinterpret_cast<const void *>(lo[a]));
    const bool ok = e == hipSuccess && (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged ||
                                        attr.type == hipMemoryTypeHost || attr.type == hipMemoryTypeUnified);
    if (!ok) {
      (void)hipGetLastError();  // a failed query leaves an error behind; clear it
      std::ostringstream m;
      m << kname << ": array '" << name[a] << "' is not in device memory (nor managed or pinned host memory)";
      if (e != hipSuccess) m << ": " << hipGetErrorString(e);
      return fail(OEB_ERR_MEMORY, m.str());
    }
  }
  return OEB_OK;
}

int check_tiles(const char *kname, const Tiles &t, int64_t ni, int64_t nj, int nz, int tile_i, int tile_j) {
  const int64_t n = ((ni + tile_i - 1) / tile_i) * ((nj + tile_j - 1) / tile_j) *
                    ((int64_t(nz) + t.kchunk - 1) / t.kchunk);
  if (n < 1 || n > (int64_t(1) << 30)) return fail(OEB_ERR_DOMAIN, std::string(kname) + ": domain too large");
  return OEB_OK;
}

int check_strips(const char *kname, const Strips &s) {
  if (s.n < 1 || s.lanes > 1024 || int64_t(s.ns) * s.nj > (int64_t(1) << 30))
    return fail(OEB_ERR_DOMAIN, std::string(kname) + ": domain too large");
  return OEB_OK;
}

namespace {
// The tuning table: (kernel, element size, nx, ny, nz) -> one value per knob
// (-1: not set).  Element size 0 stands for both precisions and a domain of
// (0, 0, 0) for every domain.
struct TuneKey {
  std::string kernel;
  size_t esize;
  int64_t nx, ny, nz;
  bool operator==(const TuneKey &o) const {
    return kernel == o.kernel && esize == o.esize && nx == o.nx && ny == o.ny && nz == o.nz;
  }
};
struct TuneHash {
  size_t operator()(const TuneKey &k) const {
    size_t h = std::hash<std::string>()(k.kernel);
    for (int64_t v : {int64_t(k.esize), k.nx, k.ny, k.nz}) h = h * 1000003u ^ std::hash<int64_t>()(v);
    return h;
  }
};
struct TuneVal {
  int v[KNOB_COUNT];
  TuneVal() {
    for (int &x : v) x = -1;
  }
  bool empty() const {
    for (int x : v)
      if (x >= 0) return false;
    return true;
  }
};
std::mutex g_tune_mu;
std::unordered_map<TuneKey, TuneVal, TuneHash> g_tune;

struct KnobInfo {
  const char *name;
  int lo, hi;
};
// Names and valid ranges of the knobs, in enum order.
constexpr KnobInfo kKnobs[KNOB_COUNT] = {
    {"kchunk", 1, 4096}, {"vector", 0, 1}, {"form", 0, 2},     {"rows", 1, 4096}, {"depth", 1, 8},
    {"order", 0, 3},     {"tile", 0, 17},  {"lanes", 64, 1024}, {"grid", 1, 1 << 30},
};

int knob_index(const char *name) {
  for (int k = 0; k < KNOB_COUNT; ++k)
    if (std::strcmp(kKnobs[k].name, name) == 0) return k;
  return -1;
}

// The table's value of knob k for this kernel, element size and domain, most
// specific entry first; -1 when no entry sets it.  Caller holds g_tune_mu.
int tune_lookup_locked(const char *kname, size_t esize, const oeb_domain *d, int k) {
  if (g_tune.empty()) return -1;
  const int64_t nx = d ? d->nx : 0, ny = d ? d->ny : 0, nz = d ? d->nz : 0;
  const TuneKey keys[4] = {{kname, esize, nx, ny, nz}, {kname, esize, 0, 0, 0}, {kname, 0, nx, ny, nz},
                           {kname, 0, 0, 0, 0}};
  for (const TuneKey &key : keys) {
    const auto it = g_tune.find(key);
    if (it != g_tune.end() && it->second.v[k] >= 0) return it->second.v[k];
  }
  return -1;
}
}  // namespace

int knob(const char *kname, size_t esize, const oeb_domain *d, Knob k, int dflt) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const int v = tune_lookup_locked(kname, esize, d, int(k));
  return v >= 0 ? v : dflt;
}

int launch_status(const char *kname) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(OEB_ERR_HIP, std::string(kname) + ": launch failed: " + hipGetErrorString(e));
  return OEB_OK;
}

int cu_count() {
  constexpr int kMaxDevices = 64;
  static std::atomic<int> cached[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return 256;
  }
  int n = cached[dev].load(std::memory_order_relaxed);
  if (n > 0) return n;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    n = 256;
  }
  cached[dev].store(n, std::memory_order_relaxed);
  return n;
}

int balanced_kchunk(int64_t tiles_per_level, int nz, int wg_per_cu, int lo, int hi, int dflt) {
  const int64_t cap = int64_t(wg_per_cu) * cu_count();
  int best = dflt;
  double best_eff = 0.0;
  for (int kc = hi; kc >= lo; --kc) {
    const int64_t n = tiles_per_level * ((nz + kc - 1) / kc);
    const double eff = double(n) / double(((n + cap - 1) / cap) * cap);
    if (eff > best_eff + 1e-9) best = kc, best_eff = eff;
  }
  return best_eff >= 0.9 ? best : dflt;
}

}  // namespace oeb

namespace oeb {
namespace {
int tune_args(const char *fn, const char *kernel, int element_bytes, const oeb_domain *d) {
  using namespace oeb;
  if (kernel == nullptr) return fail(OEB_ERR_NULL, std::string(fn) + ": kernel is NULL");
  if (element_bytes != 0 && element_bytes != 4 && element_bytes != 8)
    return fail(OEB_ERR_ARG, std::string(fn) + ": element_bytes must be 0, 4 or 8");
  if (d != nullptr && (d->nx < 0 || d->ny < 0 || d->nz < 0))
    return fail(OEB_ERR_DOMAIN, std::string(fn) + ": domain extents must be >= 0");
  return OEB_OK;
}

TuneKey tune_key(const char *kernel, int element_bytes, const oeb_domain *d) {
  return TuneKey{kernel, size_t(element_bytes), d ? d->nx : 0, d ? d->ny : 0, d ? d->nz : 0};
}
}  // namespace
}  // namespace oeb

extern "C" {

const char *oeb_version(void) { return "0.1.0"; }

const char *oeb_status_string(int status) {
  switch (status) {
    case OEB_OK: return "ok";
    case OEB_ERR_NULL: return "null descriptor or pointer";
    case OEB_ERR_DOMAIN: return "invalid domain extents";
    case OEB_ERR_LAYOUT: return "descriptor does not cover the stencil's region, or strides differ";
    case OEB_ERR_ALIAS: return "an output overlaps another array";
    case OEB_ERR_HIP: return "HIP runtime error";
    case OEB_ERR_ARG: return "an argument is out of its range";
    case OEB_ERR_MEMORY: return "an array is not in memory the GPU may access";
  }
  return "unknown status";
}

const char *oeb_last_error(void) { return oeb::g_last_error.c_str(); }


int oeb_tune_knob_set(const char *kernel, int element_bytes, const oeb_domain *d, const char *knob, int value) {
  using namespace oeb;
  OEB_TRY(tune_args("tune_knob_set", kernel, element_bytes, d));
  if (knob == nullptr) return fail(OEB_ERR_NULL, "tune_knob_set: knob is NULL");
  const int k = knob_index(knob);
  if (k < 0) return fail(OEB_ERR_ARG, std::string("tune_knob_set: unknown knob '") + knob + "'");
  if (value != -1 && (value < kKnobs[k].lo || value > kKnobs[k].hi)) {
    std::ostringstream m;
    m << "tune_knob_set: " << knob << " must lie in [" << kKnobs[k].lo << ", " << kKnobs[k].hi << "] or be -1";
    return fail(OEB_ERR_ARG, m.str());
  }
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const TuneKey key = tune_key(kernel, element_bytes, d);
  auto it = g_tune.find(key);
  if (it == g_tune.end()) {
    if (value == -1) return OEB_OK;
    it = g_tune.emplace(key, TuneVal()).first;
  }
  it->second.v[k] = value;
  if (it->second.empty()) g_tune.erase(it);
  return OEB_OK;
}

int oeb_tune_knob_get(const char *kernel, int element_bytes, const oeb_domain *d, const char *knob, int *value) {
  using namespace oeb;
  OEB_TRY(tune_args("tune_knob_get", kernel, element_bytes, d));
  if (knob == nullptr || value == nullptr) return fail(OEB_ERR_NULL, "tune_knob_get: knob or value is NULL");
  const int k = knob_index(knob);
  if (k < 0) return fail(OEB_ERR_ARG, std::string("tune_knob_get: unknown knob '") + knob + "'");
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const auto it = g_tune.find(tune_key(kernel, element_bytes, d));
  *value = it == g_tune.end() ? -1 : it->second.v[k];
  return OEB_OK;
}

int oeb_tune_set(const char *kernel, int element_bytes, const oeb_domain *d, int kchunk, int vector) {
  using namespace oeb;
  if (kernel == nullptr || d == nullptr) return fail(OEB_ERR_NULL, "tune_set: kernel or domain is NULL");
  if (element_bytes != 4 && element_bytes != 8) return fail(OEB_ERR_ARG, "tune_set: element_bytes must be 4 or 8");
  if (kchunk < 0 || kchunk > 4096 || vector < -1 || vector > 1)
    return fail(OEB_ERR_ARG, "tune_set: kchunk must lie in [0, 4096] and vector in [-1, 1]");
  OEB_TRY(oeb_tune_knob_set(kernel, element_bytes, d, "kchunk", kchunk == 0 ? -1 : kchunk));
  return oeb_tune_knob_set(kernel, element_bytes, d, "vector", vector);
}

int oeb_tune_get(const char *kernel, int element_bytes, const oeb_domain *d, int *kchunk, int *vector) {
  using namespace oeb;
  if (kernel == nullptr || d == nullptr || kchunk == nullptr || vector == nullptr)
    return fail(OEB_ERR_NULL, "tune_get: a pointer argument is NULL");
  int kc = -1;
  OEB_TRY(oeb_tune_knob_get(kernel, element_bytes, d, "kchunk", &kc));
  OEB_TRY(oeb_tune_knob_get(kernel, element_bytes, d, "vector", vector));
  *kchunk = kc < 0 ? 0 : kc;
  return OEB_OK;
}

int oeb_tune_clear(void) {
  std::lock_guard<std::mutex> lock(oeb::g_tune_mu);
  oeb::g_tune.clear();
  return OEB_OK;
}

}  // extern "C"
