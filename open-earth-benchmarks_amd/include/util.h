// util.h -- drop-in replacement for the reference's include/util.h.
//
// Same names and call signatures as the reference harness (include/util.h of
// open-earth-benchmarks): the ALGO macros, Storage<ElementType, Rank> with its
// descriptor fields and operator(), allocateStorage, freeStorage, fillMath,
// initValue, mytimer_t and the global `timer`.  A driver that includes this
// file and kernels/<name>.h from this tree runs each stencil on the GPU
// through liboeb (include/oeb.h) without source changes.
//
// What differs, and why:
//  - Every Storage has a device mirror with the same element offsets.  Host
//    code fills it through operator() as before.  Before a stencil launches,
//    any array whose host copy changed since the last upload is uploaded
//    (toDevice).  Outputs become device-newer; call toHost(out) before
//    reading them on the host.  The reference drivers never read outputs.
//  - allocateStorage pads the row pitch (strides[1]) so that element (0,j,k)
//    of every row lies on a 128-byte boundary.  sizes[] keep the reference's
//    extents, and operator() uses the strides, so indexing is unchanged.
//  - fillMath and initValue also write the k-halo planes.  The reference
//    leaves them uninitialised, yet fastwavesuv, p_grad_c and nh_p_grad read
//    them (kernels/fastwavesuv.h:10-11, p_grad_c.h:17-19, nh_p_grad.h:10).
//  - freeStorage releases with the matching deallocator (the reference calls
//    delete on new[] memory, include/util.h:145).
//  - The stencil wrappers block until their kernel has finished, as the
//    reference's CPU functions return with the result computed; mytimer_t
//    then measures the same thing it measures in the reference.
//
// The driver defines domain_size, domain_height, halo_width and count, and
// `typedef double ElementType`, before including this file, as in the reference
// (src/<name>.cpp:3-9).
#ifndef UTIL_H
#define UTIL_H

#include <hip/hip_runtime.h>
#include <sys/time.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <type_traits>
#include <unordered_map>

#include "oeb.h"

#define DEFAULT 0
#define FULL_FUSION 1
#define PARTIAL_FUSION 2
#define OPENMP 3

#define get_elapsed_time_ms(_s, _e) (1000.0 * (_e.tv_sec - _s.tv_sec) + (_e.tv_usec - _s.tv_usec) / 1000.0)

#define EARTH_RADIUS ((ElementType)6371.229e3)  // radius of the earth
#define EARTH_RADIUS_RECIP ((ElementType)1.0 / EARTH_RADIUS)

static_assert(std::is_same<ElementType, double>::value, "the GPU stencils of liboeb are FP64 (ElementType double)");

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
  // Device mirror: same element offsets from deviceAllocatedPtr.
  T *deviceAllocatedPtr;
  int64_t count;             // elements in the allocation
  mutable bool hostNewer;    // host copy changed since the last upload
  mutable bool deviceNewer;  // a kernel wrote the device copy since the last download

  template <typename... I>
  T &operator()(I... idx) {
    static_assert(sizeof...(I) == Rank, "one index per dimension");
    return alignedPtr[at(int64_t(idx)...)];
  }
  template <typename... I>
  const T &operator()(I... idx) const {
    static_assert(sizeof...(I) == Rank, "one index per dimension");
    return alignedPtr[at(int64_t(idx)...)];
  }

 private:
  // Element (i[, j[, k]]): offset + strides[R-1]*i + ... + strides[0]*k (include/util.h:56-94).
  int64_t at(int64_t i) const { return offset + strides[Rank - 1] * i; }
  int64_t at(int64_t i, int64_t j) const { return offset + strides[Rank - 1] * i + strides[Rank - 2] * j; }
  int64_t at(int64_t i, int64_t j, int64_t k) const {
    return offset + strides[Rank - 1] * i + strides[Rank - 2] * j + strides[Rank - 3] * k;
  }
};

typedef Storage<ElementType, 1> Storage1D;
typedef Storage<ElementType, 3> Storage3D;

namespace oeb_dropin {

constexpr int64_t kRowAlign = 16;  // elements: 128 B

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
  s.hostNewer = true;
  s.deviceNewer = false;
}

inline hipStream_t &stream() {
  static hipStream_t s = nullptr;  // the null stream unless a driver sets one
  return s;
}

}  // namespace oeb_dropin

// Uploads the host copy if it changed since the last upload.
template <size_t Rank>
inline void toDevice(const Storage<ElementType, Rank> &s) {
  if (!s.hostNewer) return;
  OEB_HIP_CHECK(hipMemcpy(s.deviceAllocatedPtr, s.allocatedPtr, size_t(s.count) * sizeof(ElementType),
                          hipMemcpyHostToDevice));
  s.hostNewer = false;
}

// Downloads the device copy if a kernel wrote it since the last download.
template <size_t Rank>
inline void toHost(Storage<ElementType, Rank> &s) {
  if (!s.deviceNewer) return;
  OEB_HIP_CHECK(hipMemcpy(s.allocatedPtr, s.deviceAllocatedPtr, size_t(s.count) * sizeof(ElementType),
                          hipMemcpyDeviceToHost));
  s.deviceNewer = false;
}

// allocate Storages (include/util.h:99-141): the same arguments and index
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
  // i = 0 of every row on a 128 B boundary: `lead` elements precede it.
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
  ref.allocatedPtr = nullptr;
  ref.alignedPtr = nullptr;
  ref.deviceAllocatedPtr = nullptr;
}

// fillMath (include/util.h:150-188), over i, j in [-h, n+h) and, for 3-D
// fields, k in [-h, nz+h): the k-halo planes get the same formula.
inline void fillMath(ElementType a, ElementType b, ElementType c, ElementType d, ElementType e, ElementType f,
                     Storage3D &field, const int64_t domain_size, const int64_t domain_height) {
  const ElementType dx = ElementType(1.0) / (ElementType)(domain_size + 2 * halo_width);
  const ElementType dy = ElementType(1.0) / (ElementType)(domain_size + 2 * halo_width);
  for (int64_t k = -halo_width; k < domain_height + halo_width; k++)
    for (int64_t j = -halo_width; j < domain_size + halo_width; j++)
      for (int64_t i = -halo_width; i < domain_size + halo_width; i++) {
        const ElementType x = dx * (ElementType)i;
        const ElementType y = dy * (ElementType)j;
        field(i, j, k) = k * ElementType(10e-3) + a * (b + cos(pi * (x + c * y)) + sin(d * pi * (x + e * y))) / f;
      }
  field.hostNewer = true;
}

inline void fillMath(ElementType a, ElementType b, ElementType c, ElementType d, ElementType e, ElementType f,
                     Storage1D &field, const int64_t domain_size, const int64_t /*domain_height*/) {
  const ElementType dx = ElementType(1.0) / (ElementType)(domain_size + 2 * halo_width);
  for (int64_t i = -halo_width; i < domain_size + halo_width; i++) {
    const ElementType x = dx * (ElementType)i;
    field(i) = a * (b + cos(pi * (c * x)) + sin(d * pi * (e * x))) / f;
  }
  field.hostNewer = true;
}

// initValue (include/util.h:190-196): every allocated point.
inline void initValue(Storage3D &ref, const ElementType val, const int64_t /*domain_size*/,
                      const int64_t /*domain_height*/) {
  for (int64_t n = 0; n < ref.count; ++n) ref.allocatedPtr[n] = val;
  ref.hostNewer = true;
  ref.deviceNewer = false;
}

// Wall-clock timer with the reference's interface and output format
// (include/util.h:198-223).
struct mytimer_t {
  std::unordered_map<const char *, double> elapsed_time_ms;
  std::unordered_map<const char *, int> count;
  std::unordered_map<const char *, timeval> time_point;

  void start(const char *func) {
    timeval s;
    gettimeofday(&s, NULL);
    time_point[func] = s;
  }
  void stop(const char *func) {
    timeval e;
    gettimeofday(&e, NULL);
    count[func]++;
    timeval s = time_point[func];
    elapsed_time_ms[func] += get_elapsed_time_ms(s, e);
  }
  void show_all() {
    for (auto it = elapsed_time_ms.begin(); it != elapsed_time_ms.end(); ++it) {
      auto func = it->first;
      double t = it->second;
      int c = count[func];
      printf("%s: %lf ms / %d count, avg %lf ms\n", func, t, c, t / c);
    }
  }
};

inline mytimer_t timer;

// ------------------------------------------------- stencil wrapper plumbing
namespace oeb_dropin {

inline oeb_domain domain() { return oeb_domain{domain_size, domain_size, domain_height}; }

// The device view of a Storage: the reference descriptor with device pointers.
inline oeb_s3 dev(const Storage3D &s) {
  oeb_s3 d;
  d.allocatedPtr = s.deviceAllocatedPtr;
  d.alignedPtr = s.deviceAllocatedPtr + (s.alignedPtr - s.allocatedPtr);
  d.offset = s.offset;
  for (int a = 0; a < 3; ++a) {
    d.sizes[a] = s.sizes[a];
    d.strides[a] = s.strides[a];
  }
  return d;
}

inline oeb_s1 dev(const Storage1D &s) {
  oeb_s1 d;
  d.allocatedPtr = s.deviceAllocatedPtr;
  d.alignedPtr = s.deviceAllocatedPtr + (s.alignedPtr - s.allocatedPtr);
  d.offset = s.offset;
  d.sizes[0] = s.sizes[0];
  d.strides[0] = s.strides[0];
  return d;
}

// The reference's functions return void and have no error path; a rejected
// call here is a programming error, so it stops the program with the reason.
inline void check(int status, const char *name) {
  if (status != OEB_OK) {
    std::fprintf(stderr, "%s: %s (%s)\n", name, oeb_last_error(), oeb_status_string(status));
    std::exit(1);
  }
}

// Waits for the kernel and marks the outputs device-newer.
template <typename... S>
inline void finish(S &...outs) {
  OEB_HIP_CHECK(hipStreamSynchronize(stream()));
  (void)std::initializer_list<int>{((outs.deviceNewer = true), 0)...};
}

template <typename... S>
inline void upload(const S &...ins) {
  (void)std::initializer_list<int>{(toDevice(ins), 0)...};
}

}  // namespace oeb_dropin

#endif  // UTIL_H
