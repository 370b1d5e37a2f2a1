// util.h -- drop-in replacement for the reference's include/util.h.
//
// Keeps the interface the reference drivers use (include/util.h of
// open-earth-benchmarks): the ALGO macros, Storage<ElementType, Rank> with its
// descriptor fields and operator(), allocateStorage, freeStorage, fillMath,
// initValue, mytimer_t and the global `timer`.  A driver that includes this
// file and kernels/<name>.h from this tree runs each stencil on the GPU
// through liboeb (include/oeb.h) without source changes.
//
// The driver defines domain_size, domain_height, halo_width and count, and
// `typedef double ElementType` (or float), before including this file, as in
// the reference (src/<name>.cpp:3-9).  With float, the stencils run liboeb's
// FP32 forms (oeb_<name>_f32); their FP64 island arrays (fastwavesuv ppuv,
// hhl; p_grad_c gz, pkc; nh_p_grad gz, pp, pk3) get an FP64 device copy of
// the float host data.
//
// Host and device copies.  Every Storage owns a host array (what operator()
// addresses) and a device array with the same element offsets.  A stencil
// call uploads each input whose host copy changed and marks its outputs as
// changed on the device.  Host access through operator() first downloads an
// array a kernel changed, and a non-const access marks the host copy as
// changed, so host reads see kernel results and host writes reach the next
// kernel that reads the array.  toHost / toDevice do the same explicitly.
//
// Other differences from the reference:
//  - allocateStorage pads the row pitch (strides[1]) so that element (0,j,k)
//    of every row lies on a 128-byte boundary.  Indexing goes through the
//    strides, so it is unchanged.
//  - fillMath and initValue also write the k-halo planes, which the
//    reference leaves uninitialised although fastwavesuv, p_grad_c and
//    nh_p_grad read them (kernels/fastwavesuv.h:10-11, p_grad_c.h:17-19,
//    nh_p_grad.h:10).
//  - freeStorage releases with the matching deallocator (the reference calls
//    delete on new[] memory, include/util.h:145).
//  - The stencil wrappers block until their kernel has finished, as the
//    reference's CPU functions return with the result computed, so mytimer_t
//    measures what it measures in the reference.
#ifndef UTIL_H
#define UTIL_H

#include <hip/hip_runtime.h>
#include <sys/time.h>

#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "oeb.h"

#define DEFAULT 0
#define FULL_FUSION 1
#define PARTIAL_FUSION 2
#define OPENMP 3

#define get_elapsed_time_ms(_s, _e) (1000.0 * (_e.tv_sec - _s.tv_sec) + (_e.tv_usec - _s.tv_usec) / 1000.0)

#define EARTH_RADIUS ((ElementType)6371.229e3)  // radius of the earth
#define EARTH_RADIUS_RECIP ((ElementType)1.0 / EARTH_RADIUS)

static_assert(std::is_same<ElementType, double>::value || std::is_same<ElementType, float>::value,
              "liboeb's stencils take ElementType double (oeb_<name>_f64) or float (oeb_<name>_f32)");

const ElementType pi(std::acos(-1.0));

#define OEB_HIP_CHECK(expr)                                                                        \
  do {                                                                                             \
    hipError_t oeb_e_ = (expr);                                                                    \
    if (oeb_e_ != hipSuccess) {                                                                    \
      std::fprintf(stderr, "%s:%d: %s: %s\n", __FILE__, __LINE__, #expr, hipGetErrorString(oeb_e_)); \
      std::exit(1);                                                                                \
    }                                                                                              \
  } while (0)

template <typename T, size_t Rank>
struct Storage {
  // The reference's descriptor (include/util.h:24-30), host side.
  T *allocatedPtr;
  T *alignedPtr;
  int64_t offset;
  std::array<int64_t, Rank> sizes;
  std::array<int64_t, Rank> strides;
  // Device array: the same element offsets from deviceAllocatedPtr.
  T *deviceAllocatedPtr;
  // FP64 device copy of a float array passed as an FP64 island (else NULL).
  double *deviceIsland;
  int64_t count;              // elements in the allocation
  mutable bool hostNewer;     // host copy changed since the last upload
  mutable bool islandNewer;   // host copy changed since the last island upload
  mutable bool deviceNewer;   // a kernel changed the device copy since the last download

  template <typename... I>
  T &operator()(I... idx) {
    static_assert(sizeof...(I) == Rank, "one index per dimension");
    sync_host();
    hostNewer = islandNewer = true;
    return alignedPtr[at(int64_t(idx)...)];
  }
  template <typename... I>
  const T &operator()(I... idx) const {
    static_assert(sizeof...(I) == Rank, "one index per dimension");
    sync_host();
    return alignedPtr[at(int64_t(idx)...)];
  }

  // Downloads the device copy if a kernel changed it since the last download.
  void sync_host() const {
    if (!deviceNewer) return;
    OEB_HIP_CHECK(hipMemcpy(allocatedPtr, deviceAllocatedPtr, size_t(count) * sizeof(T), hipMemcpyDeviceToHost));
    deviceNewer = false;
    islandNewer = true;
  }

 private:
  // Element (i[, j[, k]]): offset + strides[R-1]*i + ... + strides[0]*k, with
  // the strides stored in reverse axis order (include/util.h:56-94).
  int64_t at(int64_t i) const { return offset + strides[Rank - 1] * i; }
  int64_t at(int64_t i, int64_t j) const { return offset + strides[Rank - 1] * i + strides[Rank - 2] * j; }
  int64_t at(int64_t i, int64_t j, int64_t k) const {
    return offset + strides[Rank - 1] * i + strides[Rank - 2] * j + strides[Rank - 3] * k;
  }
};

typedef Storage<ElementType, 1> Storage1D;
typedef Storage<ElementType, 3> Storage3D;

// Uploads the host copy if it changed since the last upload.
template <size_t Rank>
inline void toDevice(const Storage<ElementType, Rank> &s) {
  s.sync_host();
  if (!s.hostNewer) return;
  OEB_HIP_CHECK(hipMemcpy(s.deviceAllocatedPtr, s.allocatedPtr, size_t(s.count) * sizeof(ElementType),
                          hipMemcpyHostToDevice));
  s.hostNewer = false;
}

// Downloads the device copy if a kernel changed it since the last download.
template <size_t Rank>
inline void toHost(const Storage<ElementType, Rank> &s) {
  s.sync_host();
}

namespace oeb_dropin {

constexpr int64_t kRowAlign = 16;  // elements of a row pitch multiple (128 B in FP64)

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

template <size_t Rank>
inline void alloc_both(Storage<ElementType, Rank> &s, int64_t n) {
  s.count = n;
  void *h = nullptr;
  if (posix_memalign(&h, 256, size_t(n) * sizeof(ElementType)) != 0) {
    std::fprintf(stderr, "allocateStorage: host allocation of %lld elements failed\n", (long long)n);
    std::exit(1);
  }
  std::memset(h, 0, size_t(n) * sizeof(ElementType));
  s.allocatedPtr = static_cast<ElementType *>(h);
  s.alignedPtr = s.allocatedPtr;
  OEB_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&s.deviceAllocatedPtr), size_t(n) * sizeof(ElementType)));
  OEB_HIP_CHECK(hipMemset(s.deviceAllocatedPtr, 0, size_t(n) * sizeof(ElementType)));
  s.deviceIsland = nullptr;
  s.hostNewer = s.islandNewer = true;
  s.deviceNewer = false;
}

// Marks every copy other than the host one as stale.
template <size_t Rank>
inline void host_written(Storage<ElementType, Rank> &s) {
  s.hostNewer = s.islandNewer = true;
  s.deviceNewer = false;
}

inline hipStream_t &stream() {
  static hipStream_t s = nullptr;  // the null stream unless a driver sets one
  return s;
}

}  // namespace oeb_dropin

// allocateStorage (include/util.h:99-141): the same arguments and index
// origin; the halo is halo_width on every axis, k included.
inline Storage1D allocateStorage(const int64_t size) {
  Storage1D r;
  r.sizes[0] = size;
  r.strides[0] = 1;
  r.offset = halo_width;
  oeb_dropin::alloc_both(r, size);
  return r;
}

inline Storage3D allocateStorage(const std::array<int64_t, 3> sizes) {
  Storage3D r;
  r.sizes[2] = sizes[0];
  r.sizes[1] = sizes[1];
  r.sizes[0] = sizes[2];
  // i = 0 of every row on a row-pitch boundary: `lead` elements precede it.
  const int64_t lead = oeb_dropin::round_up(halo_width, oeb_dropin::kRowAlign);
  r.strides[2] = 1;
  r.strides[1] = oeb_dropin::round_up(lead + sizes[0] - halo_width, oeb_dropin::kRowAlign);
  r.strides[0] = r.strides[1] * sizes[1];
  r.offset = lead + halo_width * r.strides[1] + halo_width * r.strides[0];
  oeb_dropin::alloc_both(r, r.strides[0] * sizes[2]);
  // sizes[2] counts from the row start, so it spans the leading pad too.
  r.sizes[2] = r.strides[1];
  return r;
}

template <typename S>
inline void freeStorage(S &ref) {
  std::free(ref.allocatedPtr);
  if (ref.deviceAllocatedPtr) OEB_HIP_CHECK(hipFree(ref.deviceAllocatedPtr));
  if (ref.deviceIsland) OEB_HIP_CHECK(hipFree(ref.deviceIsland));
  ref.allocatedPtr = nullptr;
  ref.alignedPtr = nullptr;
  ref.deviceAllocatedPtr = nullptr;
  ref.deviceIsland = nullptr;
}

// fillMath (include/util.h:150-188), over i, j in [-h, n+h) and, for 3-D
// fields, k in [-h, nz+h): the k-halo planes get the same formula.
inline void fillMath(ElementType a, ElementType b, ElementType c, ElementType d, ElementType e, ElementType f,
                     Storage3D &field, const int64_t domain_size, const int64_t domain_height) {
  const ElementType dx = ElementType(1.0) / (ElementType)(domain_size + 2 * halo_width);
  const ElementType dy = ElementType(1.0) / (ElementType)(domain_size + 2 * halo_width);
  field.sync_host();
  ElementType *p = field.alignedPtr;
  for (int64_t k = -halo_width; k < domain_height + halo_width; k++)
    for (int64_t j = -halo_width; j < domain_size + halo_width; j++)
      for (int64_t i = -halo_width; i < domain_size + halo_width; i++) {
        const ElementType x = dx * (ElementType)i;
        const ElementType y = dy * (ElementType)j;
        p[field.offset + i + j * field.strides[1] + k * field.strides[0]] =
            k * ElementType(10e-3) + a * (b + cos(pi * (x + c * y)) + sin(d * pi * (x + e * y))) / f;
      }
  oeb_dropin::host_written(field);
}

inline void fillMath(ElementType a, ElementType b, ElementType c, ElementType d, ElementType e, ElementType f,
                     Storage1D &field, const int64_t domain_size, const int64_t /*domain_height*/) {
  const ElementType dx = ElementType(1.0) / (ElementType)(domain_size + 2 * halo_width);
  field.sync_host();
  ElementType *p = field.alignedPtr;
  for (int64_t i = -halo_width; i < domain_size + halo_width; i++) {
    const ElementType x = dx * (ElementType)i;
    p[field.offset + i] = a * (b + cos(pi * (c * x)) + sin(d * pi * (e * x))) / f;
  }
  oeb_dropin::host_written(field);
}

// initValue (include/util.h:190-196): every allocated point.
inline void initValue(Storage3D &ref, const ElementType val, const int64_t /*domain_size*/,
                      const int64_t /*domain_height*/) {
  for (int64_t n = 0; n < ref.count; ++n) ref.allocatedPtr[n] = val;
  oeb_dropin::host_written(ref);
}

// The reference's timer interface (include/util.h:198-223): start / stop
// around each call of a named section, show_all prints one line per section,
// "<name>: <total> ms / <calls> count, avg <mean> ms".  Sections are keyed by
// their name's text and printed in the order they were first started.
struct mytimer_t {
  struct Section {
    std::string name;
    double total_ms = 0.0;
    int calls = 0;
    std::chrono::steady_clock::time_point began;
  };
  std::vector<Section> sections;

  Section &find(const char *func) {
    for (auto &s : sections)
      if (s.name == func) return s;
    sections.push_back(Section{func});
    return sections.back();
  }
  void start(const char *func) { find(func).began = std::chrono::steady_clock::now(); }
  void stop(const char *func) {
    const auto now = std::chrono::steady_clock::now();
    Section &s = find(func);
    s.total_ms += std::chrono::duration<double, std::milli>(now - s.began).count();
    s.calls += 1;
  }
  void show_all() const {
    for (const auto &s : sections)
      std::printf("%s: %lf ms / %d count, avg %lf ms\n", s.name.c_str(), s.total_ms, s.calls,
                  s.calls ? s.total_ms / s.calls : 0.0);
  }
};

inline mytimer_t timer;

// ------------------------------------------------- stencil wrapper plumbing
namespace oeb_dropin {

constexpr bool kFloat = std::is_same<ElementType, float>::value;
typedef std::conditional<kFloat, oeb_s3f, oeb_s3>::type DevStorage3D;
typedef std::conditional<kFloat, oeb_s1f, oeb_s1>::type DevStorage1D;

inline oeb_domain domain() { return oeb_domain{domain_size, domain_size, domain_height}; }

// The device view of a Storage: the reference descriptor with device pointers.
// Uploads the host copy first if it changed (outputs too, so that the points a
// stencil does not write keep the host's values on the device).
inline DevStorage3D dev(const Storage3D &s) {
  toDevice(s);
  DevStorage3D d;
  d.allocatedPtr = s.deviceAllocatedPtr;
  d.alignedPtr = s.deviceAllocatedPtr + (s.alignedPtr - s.allocatedPtr);
  d.offset = s.offset;
  for (int a = 0; a < 3; ++a) {
    d.sizes[a] = s.sizes[a];
    d.strides[a] = s.strides[a];
  }
  return d;
}

inline DevStorage1D dev(const Storage1D &s) {
  toDevice(s);
  DevStorage1D d;
  d.allocatedPtr = s.deviceAllocatedPtr;
  d.alignedPtr = s.deviceAllocatedPtr + (s.alignedPtr - s.allocatedPtr);
  d.offset = s.offset;
  d.sizes[0] = s.sizes[0];
  d.strides[0] = s.strides[0];
  return d;
}

// The FP64 device view of an island input.  FP64 storage: the array itself.
// Float storage: an FP64 device copy of the host data (same offsets),
// refreshed when the host copy changed since it was made.
template <class S = Storage3D>
inline oeb_s3 island(const S &s) {
  if constexpr (!kFloat) {
    return dev(s);
  } else {
    S &m = const_cast<S &>(s);
    m.sync_host();
    if (m.deviceIsland == nullptr) {
      OEB_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&m.deviceIsland), size_t(m.count) * sizeof(double)));
      m.islandNewer = true;
    }
    if (m.islandNewer) {
      std::vector<double> wide(m.allocatedPtr, m.allocatedPtr + m.count);
      OEB_HIP_CHECK(hipMemcpy(m.deviceIsland, wide.data(), wide.size() * sizeof(double), hipMemcpyHostToDevice));
      m.islandNewer = false;
    }
    oeb_s3 d;
    d.allocatedPtr = m.deviceIsland;
    d.alignedPtr = m.deviceIsland + (m.alignedPtr - m.allocatedPtr);
    d.offset = m.offset;
    for (int a = 0; a < 3; ++a) {
      d.sizes[a] = m.sizes[a];
      d.strides[a] = m.strides[a];
    }
    return d;
  }
}

// Calls the FP64 or the FP32 entry point of a stencil, by ElementType.
template <class F64, class F32, class... A>
inline int pick(F64 f64, F32 f32, A &&...args) {
  if constexpr (!kFloat) {
    (void)f32;
    return f64(std::forward<A>(args)...);
  } else {
    (void)f64;
    return f32(std::forward<A>(args)...);
  }
}
#define OEB_DROPIN_CALL(name, ...) oeb_dropin::pick(oeb_##name##_f64, oeb_##name##_f32, __VA_ARGS__)

// The reference's functions return void and have no error path; a rejected
// call here is a programming error, so it stops the program with the reason.
inline void check(int status, const char *name) {
  if (status != OEB_OK) {
    std::fprintf(stderr, "%s: %s (%s)\n", name, oeb_last_error(), oeb_status_string(status));
    std::exit(1);
  }
}

// Waits for the kernel and marks the outputs as changed on the device.
template <typename... S>
inline void finish(S &...outs) {
  OEB_HIP_CHECK(hipStreamSynchronize(stream()));
  (void)std::initializer_list<int>{((outs.deviceNewer = true), (outs.hostNewer = false), 0)...};
}

template <typename... S>
inline void upload(const S &...ins) {
  (void)std::initializer_list<int>{(toDevice(ins), 0)...};
}

}  // namespace oeb_dropin

#endif  // UTIL_H
