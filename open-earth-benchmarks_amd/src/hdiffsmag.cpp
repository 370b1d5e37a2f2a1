// hdiffsmag.cpp -- benchmark driver for hdiffsmag on the GPU (drop-in for the reference's src/hdiffsmag.cpp).
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
#include "hdiffsmag.h"
#include "oeb_driver.h"

int main(int argc, char **argv) {
  const int ALGO = oeb_driver::parse(argc, argv);
  const int64_t size1D = domain_size + 2 * halo_width;
  const std::array<int64_t, 3> sizes3D = {domain_size + 2 * halo_width, domain_size + 2 * halo_width,
                                          domain_height + 2 * halo_width};

  Storage3D uin = allocateStorage(sizes3D);
  Storage3D vin = allocateStorage(sizes3D);
  Storage3D mask = allocateStorage(sizes3D);
  Storage3D uout = allocateStorage(sizes3D);
  Storage3D vout = allocateStorage(sizes3D);
  Storage3D T_sqr_s = allocateStorage(sizes3D);
  Storage3D S_sqr_uv = allocateStorage(sizes3D);
  Storage1D crlavo = allocateStorage(size1D);
  Storage1D crlavu = allocateStorage(size1D);
  Storage1D crlato = allocateStorage(size1D);
  Storage1D crlatu = allocateStorage(size1D);
  Storage1D acrlat0 = allocateStorage(size1D);

  ElementType eddlat = ldexpl(1.0, -11);
  ElementType eddlon = ldexpl(1.5, -11);
  ElementType tau_smag = 0.025;
  ElementType weight_smag = 0.01;

  fillMath(1.1, 2.0, 1.5, 2.8, 2.0, 4.1, uin, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, vin, domain_size, domain_height);
  fillMath(8.0, 9.4, 1.5, 1.7, 2.0, 3.5, crlavo, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, crlavu, domain_size, domain_height);
  fillMath(5.0, 8.0, 1.5, 7.1, 2.0, 4.3, crlato, domain_size, domain_height);
  fillMath(3.2, 7.0, 2.5, 6.1, 3.0, 2.3, crlatu, domain_size, domain_height);
  fillMath(4.5, 5.0, 2.5, 2.1, 3.0, 2.3, acrlat0, domain_size, domain_height);
  initValue(mask, 0.025, domain_size, domain_height);
  initValue(uout, 0.0, domain_size, domain_height);
  initValue(vout, 0.0, domain_size, domain_height);
  initValue(T_sqr_s, 0.0, domain_size, domain_height);
  initValue(S_sqr_uv, 0.0, domain_size, domain_height);

  const std::string label = oeb_driver::label("hdiffsmag", ALGO);
  double ms = 0.0;
  switch (ALGO) {
    case DEFAULT:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hdiffsmag(uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat, eddlon, tau_smag, weight_smag); });
      break;
    case FULL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hdiffsmag_fullfusion(uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat, eddlon, tau_smag, weight_smag); });
      break;
    case PARTIAL_FUSION:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hdiffsmag_partialfusion(uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat, eddlon, tau_smag, weight_smag); });
      break;
    case OPENMP:
      ms = oeb_driver::timed_loop(label.c_str(), [&] { hdiffsmag_openmp(uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat, eddlon, tau_smag, weight_smag); });
      break;
    default:
      std::cout << "Unknown ALGO" << std::endl;
      return 1;
  }
  timer.show_all();
  oeb_driver::report("hdiffsmag", ALGO, ms, 40, {{"uout", oeb_driver::checksum(uout)}, {"vout", oeb_driver::checksum(vout)}});

  freeStorage(uin);
  freeStorage(vin);
  freeStorage(mask);
  freeStorage(uout);
  freeStorage(vout);
  freeStorage(T_sqr_s);
  freeStorage(S_sqr_uv);
  freeStorage(crlavo);
  freeStorage(crlavu);
  freeStorage(crlato);
  freeStorage(crlatu);
  freeStorage(acrlat0);
  return 0;
}
