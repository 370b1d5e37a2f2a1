// fastwavesuv.cpp -- benchmark driver for fastwavesuv on the GPU (drop-in for the reference's src/fastwavesuv.cpp).
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
#include "fastwavesuv.h"
#include "oeb_driver.h"

int main(int argc, char **argv) {
  const int ALGO = oeb_driver::parse(argc, argv);
  const int64_t size1D = domain_size + 2 * halo_width;
  const std::array<int64_t, 3> sizes3D = {domain_size + 2 * halo_width, domain_size + 2 * halo_width,
                                          domain_height + 2 * halo_width};

  Storage3D uin = allocateStorage(sizes3D);
  Storage3D utens = allocateStorage(sizes3D);
  Storage3D vin = allocateStorage(sizes3D);
  Storage3D vtens = allocateStorage(sizes3D);
  Storage3D wgtfac = allocateStorage(sizes3D);
  Storage3D ppuv = allocateStorage(sizes3D);
  Storage3D hhl = allocateStorage(sizes3D);
  Storage3D rho = allocateStorage(sizes3D);
  Storage3D uout = allocateStorage(sizes3D);
  Storage3D vout = allocateStorage(sizes3D);
  Storage3D ppgk = allocateStorage(sizes3D);
  Storage3D ppgc = allocateStorage(sizes3D);
  Storage3D ppgu = allocateStorage(sizes3D);
  Storage3D ppgv = allocateStorage(sizes3D);
  Storage1D fx = allocateStorage(size1D);

  ElementType dt = 10.0;
  ElementType edadlat = ldexpl(1.0, -11);

  fillMath(1.0, 3.3, 1.5, 1.5, 2.0, 4.0, uin, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, utens, domain_size, domain_height);
  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, vin, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, vtens, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, ppuv, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, wgtfac, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, hhl, domain_size, domain_height);
  fillMath(3.2, 7.0, 2.5, 6.1, 3.0, 2.3, rho, domain_size, domain_height);
  fillMath(4.5, 5.0, 2.5, 2.1, 3.0, 2.3, fx, domain_size, domain_height);
  initValue(uout, 0.0, domain_size, domain_height);
  initValue(vout, 0.0, domain_size, domain_height);

  const std::string label = oeb_driver::label("fastwavesuv", ALGO);
  double ms = 0.0;
  switch (ALGO) {
    case DEFAULT:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { fastwavesuv(uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt); });
      break;
    case FULL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { fastwavesuv_fullfusion(uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt); });
      break;
    case PARTIAL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { fastwavesuv_partialfusion(uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt); });
      break;
    case OPENMP:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { fastwavesuv_openmp(uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt); });
      break;
    default:
      std::cout << "Unknown ALGO" << std::endl;
      return 1;
  }
  timer.show_all();
  oeb_driver::report("fastwavesuv", ALGO, ms, 80, {{"uout", oeb_driver::checksum(uout)}, {"vout", oeb_driver::checksum(vout)}});

  freeStorage(uin);
  freeStorage(utens);
  freeStorage(vin);
  freeStorage(vtens);
  freeStorage(wgtfac);
  freeStorage(ppuv);
  freeStorage(hhl);
  freeStorage(rho);
  freeStorage(uout);
  freeStorage(vout);
  freeStorage(ppgk);
  freeStorage(ppgc);
  freeStorage(ppgu);
  freeStorage(ppgv);
  freeStorage(fx);
  return 0;
}
