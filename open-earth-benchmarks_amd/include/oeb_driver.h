// oeb_driver.h -- pieces of the benchmark driver (src/oeb_bench.cpp): the
// reference's command line (src/<name>.cpp:16-29), the timed loop, and a
// result line with GPU event timing, bandwidth, output checksums and the
// verification result.
//
// Include after util.h and the kernel header.
#ifndef OEB_DRIVER_H
#define OEB_DRIVER_H

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

namespace oeb_driver {

// argv: [ALGO [domain_size domain_height]], as the reference drivers parse it.
inline int parse(int argc, char **argv) {
  int algo = DEFAULT;
  if (argc == 4) {
    algo = atoi(argv[1]);
    domain_size = atoi(argv[2]);
    domain_height = atoi(argv[3]);
  } else if (argc == 2) {
    algo = atoi(argv[1]);
  } else if (argc != 1) {
    std::cout << "Usage: ./kernel ${ALGO} ${domain_size} ${domain_height}" << std::endl;
    exit(1);
  }
  std::cout << "ALGO: " << algo << std::endl;
  std::cout << "domain_size: " << domain_size << std::endl;
  std::cout << "domain_height: " << domain_height << std::endl;
  return algo;
}

// Timer label of a variant, as the reference drivers name them.
inline std::string label(const char *name, int algo) {
  switch (algo) {
    case DEFAULT: return name;
    case FULL_FUSION: return std::string(name) + " fullfusion";
    case PARTIAL_FUSION: return std::string(name) + " partialfusion";
    case OPENMP: return std::string(name) + " openmp";
  }
  return std::string(name) + " unknown";
}

// Runs fn `count` times, each bracketed by the wall-clock timer (as the
// reference's loop, e.g. src/hdiffsa.cpp:54-58) and by HIP events.  Returns the
// mean event time per call in ms.
template <typename F>
double timed_loop(const char *label, F fn) {
  hipEvent_t e0, e1;
  OEB_HIP_CHECK(hipEventCreate(&e0));
  OEB_HIP_CHECK(hipEventCreate(&e1));
  double total = 0.0;
  int64_t n = 0;
  int64_t reps = count;
  while (reps--) {
    timer.start(label);
    OEB_HIP_CHECK(hipEventRecord(e0, oeb_dropin::stream()));
    fn();
    OEB_HIP_CHECK(hipEventRecord(e1, oeb_dropin::stream()));
    timer.stop(label);
    OEB_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    OEB_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    total += ms;
    ++n;
  }
  OEB_HIP_CHECK(hipEventDestroy(e0));
  OEB_HIP_CHECK(hipEventDestroy(e1));
  return n ? total / n : 0.0;
}

// One JSON result line: mean kernel time from HIP events, Gpts/s and the
// fraction of the 8 TB/s HBM peak at the stencil's algorithmic bytes per point
// (of the build's precision), the region
// sums of the outputs and whether every output passed verification.
inline void report(const char *name, int algo, double ms, int bytes_per_point,
                   const std::vector<std::pair<const char *, std::pair<double, double>>> &sums, bool verified) {
  const double pts = double(domain_size) * double(domain_size) * double(domain_height);
  const double gbs = bytes_per_point * pts / (ms * 1e-3) / 1e9;
  std::printf("{\"kernel\": \"%s\", \"algo\": %d, \"domain_size\": %lld, \"domain_height\": %lld, \"count\": %lld, "
              "\"gpu_ms\": %.6f, \"gpts_s\": %.3f, \"gbs\": %.1f, \"hbm_frac\": %.4f, \"checksums\": {",
              name, algo, (long long)domain_size, (long long)domain_height, (long long)count, ms,
              pts / (ms * 1e-3) / 1e9, gbs, gbs / 8000.0);
  bool first = true;
  for (const auto &kv : sums) {
    std::printf("%s\"%s\": %.17g", first ? "" : ", ", kv.first, kv.second.first);
    first = false;
  }
  std::printf("}, \"abs_checksums\": {");
  first = true;
  for (const auto &kv : sums) {
    std::printf("%s\"%s\": %.17g", first ? "" : ", ", kv.first, kv.second.second);
    first = false;
  }
  std::printf("}, \"verified\": %s}\n", verified ? "true" : "false");
}

}  // namespace oeb_driver

#endif  // OEB_DRIVER_H
