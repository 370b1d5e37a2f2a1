// nh_p_grad.cpp -- benchmark driver for nh_p_grad on the GPU (drop-in for the reference's src/nh_p_grad.cpp).
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

#include "util.h"
#include "nh_p_grad.h"
#include "oeb_driver.h"

int main(int argc, char **argv) {
  const int ALGO = oeb_driver::parse(argc, argv);
  const std::array<int64_t, 3> sizes3D = {domain_size + 2 * halo_width, domain_size + 2 * halo_width,
                                          domain_height + 2 * halo_width};

  Storage3D uin = allocateStorage(sizes3D);
  Storage3D vin = allocateStorage(sizes3D);
  Storage3D rdx = allocateStorage(sizes3D);
  Storage3D rdy = allocateStorage(sizes3D);
  Storage3D gz = allocateStorage(sizes3D);
  Storage3D pp = allocateStorage(sizes3D);
  Storage3D pk3 = allocateStorage(sizes3D);
  Storage3D wk1 = allocateStorage(sizes3D);
  Storage3D uout = allocateStorage(sizes3D);
  Storage3D vout = allocateStorage(sizes3D);
  Storage3D wk = allocateStorage(sizes3D);
  Storage3D du = allocateStorage(sizes3D);
  Storage3D dv = allocateStorage(sizes3D);

  ElementType dt = 0.1;

  fillMath(1.0, 3.3, 1.5, 1.5, 2.0, 4.0, uin, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, vin, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, rdx, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, rdy, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, gz, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, pp, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, pk3, domain_size, domain_height);
  fillMath(3.2, 7.0, 2.5, 6.1, 3.0, 2.3, wk1, domain_size, domain_height);
  initValue(uout, 0.0, domain_size, domain_height);
  initValue(vout, 0.0, domain_size, domain_height);

  const std::string label = oeb_driver::label("nh_p_grad", ALGO);
  double ms = 0.0;
  switch (ALGO) {
    case DEFAULT:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { nh_p_grad(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv, dt); });
      break;
    case FULL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { nh_p_grad_fullfusion(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv, dt); });
      break;
    case PARTIAL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { nh_p_grad_partialfusion(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv, dt); });
      break;
    case OPENMP:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { nh_p_grad_openmp(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv, dt); });
      break;
    default:
      std::cout << "Unknown ALGO" << std::endl;
      return 1;
  }
  timer.show_all();
  oeb_driver::report("nh_p_grad", ALGO, ms, 80, {{"uout", oeb_driver::checksum(uout)}, {"vout", oeb_driver::checksum(vout)}});

  freeStorage(uin);
  freeStorage(vin);
  freeStorage(rdx);
  freeStorage(rdy);
  freeStorage(gz);
  freeStorage(pp);
  freeStorage(pk3);
  freeStorage(wk1);
  freeStorage(uout);
  freeStorage(vout);
  freeStorage(wk);
  freeStorage(du);
  freeStorage(dv);
  return 0;
}
