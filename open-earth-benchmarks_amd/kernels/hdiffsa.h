// hdiffsa.h -- drop-in for the reference's kernels/hdiffsa.h (kernels/hdiffsa.h:4-123).
// hdiffsa, hdiffsa_fullfusion, hdiffsa_partialfusion, hdiffsa_openmp keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_hdiffsa_f64, or oeb_hdiffsa_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef HDIFFSA_H
#define HDIFFSA_H

#include "util.h"

inline void hdiffsa(Storage3D& out, const Storage3D& in, const Storage3D& mask, const Storage1D& crlato,
                    const Storage1D& crlatu, Storage3D& lap, Storage3D& flx, Storage3D& fly) {
  (void)lap;
  (void)flx;
  (void)fly;
  oeb_dropin::upload(in, mask, crlato, crlatu);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_out = oeb_dropin::dev(out);
  const auto a_in = oeb_dropin::dev(in);
  const auto a_mask = oeb_dropin::dev(mask);
  const auto a_crlato = oeb_dropin::dev(crlato);
  const auto a_crlatu = oeb_dropin::dev(crlatu);
  oeb_dropin::check(OEB_DROPIN_CALL(hdiffsa, &d, &a_out, &a_in, &a_mask, &a_crlato, &a_crlatu, nullptr, nullptr, nullptr, oeb_dropin::stream()),
                    "hdiffsa");
  oeb_dropin::finish(out);
}

inline void hdiffsa_fullfusion(Storage3D& out, const Storage3D& in, const Storage3D& mask, const Storage1D& crlato,
                               const Storage1D& crlatu, Storage3D& lap, Storage3D& flx, Storage3D& fly) {
  hdiffsa(out, in, mask, crlato, crlatu, lap, flx, fly);
}

inline void hdiffsa_partialfusion(Storage3D& out, const Storage3D& in, const Storage3D& mask, const Storage1D& crlato,
                                  const Storage1D& crlatu, Storage3D& lap, Storage3D& flx, Storage3D& fly) {
  hdiffsa(out, in, mask, crlato, crlatu, lap, flx, fly);
}

inline void hdiffsa_openmp(Storage3D& out, const Storage3D& in, const Storage3D& mask, const Storage1D& crlato,
                           const Storage1D& crlatu, Storage3D& lap, Storage3D& flx, Storage3D& fly) {
  hdiffsa(out, in, mask, crlato, crlatu, lap, flx, fly);
}

#endif  // HDIFFSA_H
