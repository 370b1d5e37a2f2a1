// uvbke.cpp -- benchmark driver for uvbke on the GPU (drop-in for the reference's src/uvbke.cpp).
// Same command line (ALGO [domain_size domain_height]), inputs (the reference's fillMath
// parameters), timed loop and timer output as the reference driver.  After the loop
// it prints one JSON line with the HIP-event time per call, Gpts/s, the fraction
// of HBM peak and the output checksums.  Every ALGO value runs the fused HIP kernel.
// Environment: OEB_COUNT overrides the repetition count.
#include <cmath>

// define the domain size and the halo width
int64_t domain_size = 64;
int64_t domain_height = 60;
int64_t halo_width = 3;
int64_t count = 1000;

typedef double ElementType;

const double dt = 225.0;
const double dt5 = 0.5 * dt;

#include "util.h"
#include "uvbke.h"
#include "oeb_driver.h"

int main(int argc, char **argv) {
  const int ALGO = oeb_driver::parse(argc, argv);
  const std::array<int64_t, 3> sizes3D = {domain_size + 2 * halo_width, domain_size + 2 * halo_width,
                                          domain_height + 2 * halo_width};

  Storage3D uc = allocateStorage(sizes3D);
  Storage3D vc = allocateStorage(sizes3D);
  Storage3D cosa = allocateStorage(sizes3D);
  Storage3D rsina = allocateStorage(sizes3D);
  Storage3D ub = allocateStorage(sizes3D);
  Storage3D vb = allocateStorage(sizes3D);

  fillMath(1.0, 3.3, 1.5, 1.5, 2.0, 4.0, uc, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, vc, domain_size, domain_height);
  fillMath(4.0, 1.7, 1.5, 6.3, 2.0, 1.4, cosa, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, rsina, domain_size, domain_height);
  initValue(ub, -1.0, domain_size, domain_height);
  initValue(vb, -1.0, domain_size, domain_height);

  const std::string label = oeb_driver::label("uvbke", ALGO);
  double ms = 0.0;
  switch (ALGO) {
    case DEFAULT:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { uvbke(ub, vb, uc, vc, cosa, rsina); });
      break;
    case FULL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { uvbke_fullfusion(ub, vb, uc, vc, cosa, rsina); });
      break;
    case PARTIAL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { uvbke_partialfusion(ub, vb, uc, vc, cosa, rsina); });
      break;
    case OPENMP:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { uvbke_openmp(ub, vb, uc, vc, cosa, rsina); });
      break;
    default:
      std::cout << "Unknown ALGO" << std::endl;
      return 1;
  }
  timer.show_all();
  oeb_driver::report("uvbke", ALGO, ms, 48, {{"ub", oeb_driver::checksum(ub)}, {"vb", oeb_driver::checksum(vb)}});

  freeStorage(uc);
  freeStorage(vc);
  freeStorage(cosa);
  freeStorage(rsina);
  freeStorage(ub);
  freeStorage(vb);
  return 0;
}
