// p_grad_c.cpp -- benchmark driver for p_grad_c on the GPU (drop-in for the reference's src/p_grad_c.cpp).
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
#include "p_grad_c.h"
#include "oeb_driver.h"

int main(int argc, char **argv) {
  const int ALGO = oeb_driver::parse(argc, argv);
  const std::array<int64_t, 3> sizes3D = {domain_size + 2 * halo_width, domain_size + 2 * halo_width,
                                          domain_height + 2 * halo_width};

  Storage3D uin = allocateStorage(sizes3D);
  Storage3D vin = allocateStorage(sizes3D);
  Storage3D rdxc = allocateStorage(sizes3D);
  Storage3D rdyc = allocateStorage(sizes3D);
  Storage3D delpc = allocateStorage(sizes3D);
  Storage3D gz = allocateStorage(sizes3D);
  Storage3D pkc = allocateStorage(sizes3D);
  Storage3D uout = allocateStorage(sizes3D);
  Storage3D vout = allocateStorage(sizes3D);
  Storage3D wk = allocateStorage(sizes3D);

  ElementType dt2 = 0.1;

  fillMath(1.0, 3.3, 1.5, 1.5, 2.0, 4.0, uin, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, vin, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, rdxc, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, rdyc, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, delpc, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, gz, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, pkc, domain_size, domain_height);
  initValue(uout, 0.0, domain_size, domain_height);
  initValue(vout, 0.0, domain_size, domain_height);

  const std::string label = oeb_driver::label("p_grad_c", ALGO);
  double ms = 0.0;
  switch (ALGO) {
    case DEFAULT:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { p_grad_c(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2); });
      break;
    case FULL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { p_grad_c_fullfusion(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2); });
      break;
    case PARTIAL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { p_grad_c_partialfusion(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2); });
      break;
    case OPENMP:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { p_grad_c_openmp(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2); });
      break;
    default:
      std::cout << "Unknown ALGO" << std::endl;
      return 1;
  }
  timer.show_all();
  oeb_driver::report("p_grad_c", ALGO, ms, 72, {{"uout", oeb_driver::checksum(uout)}, {"vout", oeb_driver::checksum(vout)}});

  freeStorage(uin);
  freeStorage(vin);
  freeStorage(rdxc);
  freeStorage(rdyc);
  freeStorage(delpc);
  freeStorage(gz);
  freeStorage(pkc);
  freeStorage(uout);
  freeStorage(vout);
  freeStorage(wk);
  return 0;
}
