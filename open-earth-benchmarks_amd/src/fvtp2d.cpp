// fvtp2d.cpp -- benchmark driver for fvtp2d on the GPU (drop-in for the reference's src/fvtp2d.cpp).
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

#define p1 ElementType(0.583333)  // 7/12
#define p2 ElementType(0.083333)  // 1/12

#include "util.h"
#include "fvtp2d.h"
#include "oeb_driver.h"

int main(int argc, char **argv) {
  const int ALGO = oeb_driver::parse(argc, argv);
  const std::array<int64_t, 3> sizes3D = {domain_size + 2 * halo_width, domain_size + 2 * halo_width,
                                          domain_height + 2 * halo_width};

  Storage3D q = allocateStorage(sizes3D);
  Storage3D crx = allocateStorage(sizes3D);
  Storage3D cry = allocateStorage(sizes3D);
  Storage3D ra_x = allocateStorage(sizes3D);
  Storage3D ra_y = allocateStorage(sizes3D);
  Storage3D xfx = allocateStorage(sizes3D);
  Storage3D yfx = allocateStorage(sizes3D);
  Storage3D area = allocateStorage(sizes3D);
  Storage3D q_i = allocateStorage(sizes3D);
  Storage3D q_j = allocateStorage(sizes3D);
  Storage3D fx1 = allocateStorage(sizes3D);
  Storage3D fx2 = allocateStorage(sizes3D);
  Storage3D fy1 = allocateStorage(sizes3D);
  Storage3D fy2 = allocateStorage(sizes3D);
  Storage3D fxx = allocateStorage(sizes3D);
  Storage3D fyy = allocateStorage(sizes3D);
  Storage3D al = allocateStorage(sizes3D);
  Storage3D almq = allocateStorage(sizes3D);
  Storage3D br = allocateStorage(sizes3D);
  Storage3D b0 = allocateStorage(sizes3D);
  Storage3D smt5 = allocateStorage(sizes3D);

  fillMath(1.0, 3.3, 1.5, 1.5, 2.0, 4.0, q, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, crx, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, cry, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, ra_x, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, ra_y, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, xfx, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, yfx, domain_size, domain_height);
  fillMath(3.2, 7.0, 2.5, 6.1, 3.0, 2.3, area, domain_size, domain_height);
  initValue(q_i, -1.0, domain_size, domain_height);
  initValue(q_j, -1.0, domain_size, domain_height);
  initValue(fx1, -1.0, domain_size, domain_height);
  initValue(fx2, -1.0, domain_size, domain_height);
  initValue(fy1, -1.0, domain_size, domain_height);
  initValue(fy2, -1.0, domain_size, domain_height);

  const std::string label = oeb_driver::label("fvtp2d", ALGO);
  double ms = 0.0;
  switch (ALGO) {
    case DEFAULT:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { fvtp2d(q_i, q_j, fx1, fx2, fy1, fy2, q, crx, cry, ra_x, ra_y, xfx, yfx, area, fxx, fyy, al, almq, br, b0, smt5); });
      break;
    case FULL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { fvtp2d_fullfusion(q_i, q_j, fx1, fx2, fy1, fy2, q, crx, cry, ra_x, ra_y, xfx, yfx, area, fxx, fyy, al, almq, br, b0, smt5); });
      break;
    default:
      std::cout << "Unknown ALGO" << std::endl;
      return 1;
  }
  timer.show_all();
  oeb_driver::report("fvtp2d", ALGO, ms, 112, {{"q_i", oeb_driver::checksum(q_i, -3, 2, 0, 0)}, {"q_j", oeb_driver::checksum(q_j, 0, 0, -3, 2)}, {"fx1", oeb_driver::checksum(fx1)}, {"fx2", oeb_driver::checksum(fx2, 0, 1, -3, 2)}, {"fy1", oeb_driver::checksum(fy1)}, {"fy2", oeb_driver::checksum(fy2, -3, 2, 0, 1)}});

  freeStorage(q);
  freeStorage(crx);
  freeStorage(cry);
  freeStorage(ra_x);
  freeStorage(ra_y);
  freeStorage(xfx);
  freeStorage(yfx);
  freeStorage(area);
  freeStorage(q_i);
  freeStorage(q_j);
  freeStorage(fx1);
  freeStorage(fx2);
  freeStorage(fy1);
  freeStorage(fy2);
  freeStorage(fxx);
  freeStorage(fyy);
  freeStorage(al);
  freeStorage(almq);
  freeStorage(br);
  freeStorage(b0);
  freeStorage(smt5);
  return 0;
}
