// hdiffsa.cpp -- benchmark driver for hdiffsa on the GPU (drop-in for the reference's src/hdiffsa.cpp).
// Same command line (ALGO [domain_size domain_height]), inputs (the reference's fillMath
// parameters), timed loop and timer output as the reference driver.  After the loop
// it prints one JSON line with the HIP-event time per call, Gpts/s, the fraction
// of HBM peak and the output checksums.  Every ALGO value runs the fused HIP kernel.
// Environment: OEB_COUNT overrides the repetition count.
// The reference driver passes (in, mask, out, ...) against the signature (out, in, mask, ...)
// (src/hdiffsa.cpp:56 vs kernels/hdiffsa.h:4); this driver passes the arguments in signature order.
#include <cmath>

// define the domain size and the halo width
int64_t domain_size = 64;
int64_t domain_height = 60;
int64_t halo_width = 3;
int64_t count = 1000;

typedef double ElementType;

#include "util.h"
#include "hdiffsa.h"
#include "oeb_driver.h"

int main(int argc, char **argv) {
  const int ALGO = oeb_driver::parse(argc, argv);
  const int64_t size1D = domain_size + 2 * halo_width;
  const std::array<int64_t, 3> sizes3D = {domain_size + 2 * halo_width, domain_size + 2 * halo_width,
                                          domain_height + 2 * halo_width};

  Storage3D in = allocateStorage(sizes3D);
  Storage3D mask = allocateStorage(sizes3D);
  Storage3D out = allocateStorage(sizes3D);
  Storage3D lap = allocateStorage(sizes3D);
  Storage3D flx = allocateStorage(sizes3D);
  Storage3D fly = allocateStorage(sizes3D);
  Storage1D crlato = allocateStorage(size1D);
  Storage1D crlatu = allocateStorage(size1D);

  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, in, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, crlato, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, crlatu, domain_size, domain_height);
  initValue(mask, 0.025, domain_size, domain_height);
  initValue(out, 0.0, domain_size, domain_height);

  const std::string label = oeb_driver::label("hdiffsa", ALGO);
  double ms = 0.0;
  switch (ALGO) {
    case DEFAULT:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hdiffsa(out, in, mask, crlato, crlatu, lap, flx, fly); });
      break;
    case FULL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hdiffsa_fullfusion(out, in, mask, crlato, crlatu, lap, flx, fly); });
      break;
    case PARTIAL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hdiffsa_partialfusion(out, in, mask, crlato, crlatu, lap, flx, fly); });
      break;
    case OPENMP:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hdiffsa_openmp(out, in, mask, crlato, crlatu, lap, flx, fly); });
      break;
    default:
      std::cout << "Unknown ALGO" << std::endl;
      return 1;
  }
  timer.show_all();
  oeb_driver::report("hdiffsa", ALGO, ms, 24, {{"out", oeb_driver::checksum(out)}});

  freeStorage(in);
  freeStorage(mask);
  freeStorage(out);
  freeStorage(lap);
  freeStorage(flx);
  freeStorage(fly);
  freeStorage(crlato);
  freeStorage(crlatu);
  return 0;
}
