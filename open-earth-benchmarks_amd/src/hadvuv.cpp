// hadvuv.cpp -- benchmark driver for hadvuv on the GPU (drop-in for the reference's src/hadvuv.cpp).
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
#include "hadvuv.h"
#include "oeb_driver.h"

int main(int argc, char **argv) {
  const int ALGO = oeb_driver::parse(argc, argv);
  const int64_t size1D = domain_size + 2 * halo_width;
  const std::array<int64_t, 3> sizes3D = {domain_size + 2 * halo_width, domain_size + 2 * halo_width,
                                          domain_height + 2 * halo_width};

  Storage3D uin = allocateStorage(sizes3D);
  Storage3D vin = allocateStorage(sizes3D);
  Storage3D uout = allocateStorage(sizes3D);
  Storage3D vout = allocateStorage(sizes3D);
  Storage3D uatupos = allocateStorage(sizes3D);
  Storage3D vatupos = allocateStorage(sizes3D);
  Storage3D uatvpos = allocateStorage(sizes3D);
  Storage3D vatvpos = allocateStorage(sizes3D);
  Storage3D uavg = allocateStorage(sizes3D);
  Storage3D vavg = allocateStorage(sizes3D);
  Storage3D ures = allocateStorage(sizes3D);
  Storage3D vres = allocateStorage(sizes3D);
  Storage1D acrlat0 = allocateStorage(size1D);
  Storage1D acrlat1 = allocateStorage(size1D);
  Storage1D tgrlatda0 = allocateStorage(size1D);
  Storage1D tgrlatda1 = allocateStorage(size1D);

  ElementType eddlat = ldexpl(1.0, -11);
  ElementType eddlon = ldexpl(1.5, -11);

  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, uin, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, vin, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, acrlat0, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, acrlat1, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, tgrlatda0, domain_size, domain_height);
  fillMath(3.2, 7.0, 2.5, 6.1, 3.0, 2.3, tgrlatda1, domain_size, domain_height);
  initValue(uout, 0.0, domain_size, domain_height);
  initValue(vout, 0.0, domain_size, domain_height);

  const std::string label = oeb_driver::label("hadvuv", ALGO);
  double ms = 0.0;
  switch (ALGO) {
    case DEFAULT:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hadvuv(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres, eddlat, eddlon); });
      break;
    case FULL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hadvuv_fullfusion(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres, eddlat, eddlon); });
      break;
    case PARTIAL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hadvuv_partialfusion(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres, eddlat, eddlon); });
      break;
    case OPENMP:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hadvuv_openmp(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres, eddlat, eddlon); });
      break;
    default:
      std::cout << "Unknown ALGO" << std::endl;
      return 1;
  }
  timer.show_all();
  oeb_driver::report("hadvuv", ALGO, ms, 32, {{"uout", oeb_driver::checksum(uout)}, {"vout", oeb_driver::checksum(vout)}});

  freeStorage(uin);
  freeStorage(vin);
  freeStorage(uout);
  freeStorage(vout);
  freeStorage(uatupos);
  freeStorage(vatupos);
  freeStorage(uatvpos);
  freeStorage(vatvpos);
  freeStorage(uavg);
  freeStorage(vavg);
  freeStorage(ures);
  freeStorage(vres);
  freeStorage(acrlat0);
  freeStorage(acrlat1);
  freeStorage(tgrlatda0);
  freeStorage(tgrlatda1);
  return 0;
}
