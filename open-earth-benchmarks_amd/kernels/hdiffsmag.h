// hdiffsmag.h -- drop-in for the reference's kernels/hdiffsmag.h (kernels/hdiffsmag.h:4-263).
// hdiffsmag, hdiffsmag_fullfusion, hdiffsmag_partialfusion, hdiffsmag_openmp keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_hdiffsmag_f64, or oeb_hdiffsmag_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef HDIFFSMAG_H
#define HDIFFSMAG_H

#include "util.h"

inline void hdiffsmag(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                      const Storage3D& mask, const Storage1D& crlavo, const Storage1D& crlavu, const Storage1D& crlato,
                      const Storage1D& crlatu, const Storage1D& acrlat0, Storage3D& T_sqr_s, Storage3D& S_sqr_uv,
                      const ElementType eddlat, const ElementType eddlon, const ElementType tau_smag,
                      const ElementType weight_smag) {
  (void)T_sqr_s;
  (void)S_sqr_uv;
  oeb_dropin::upload(uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_uout = oeb_dropin::dev(uout);
  const auto a_vout = oeb_dropin::dev(vout);
  const auto a_uin = oeb_dropin::dev(uin);
  const auto a_vin = oeb_dropin::dev(vin);
  const auto a_mask = oeb_dropin::dev(mask);
  const auto a_crlavo = oeb_dropin::dev(crlavo);
  const auto a_crlavu = oeb_dropin::dev(crlavu);
  const auto a_crlato = oeb_dropin::dev(crlato);
  const auto a_crlatu = oeb_dropin::dev(crlatu);
  const auto a_acrlat0 = oeb_dropin::dev(acrlat0);
  oeb_dropin::check(OEB_DROPIN_CALL(hdiffsmag, &d, &a_uout, &a_vout, &a_uin, &a_vin, &a_mask, &a_crlavo, &a_crlavu, &a_crlato, &a_crlatu,
                          &a_acrlat0, nullptr, nullptr, eddlat, eddlon, tau_smag, weight_smag, oeb_dropin::stream()),
                    "hdiffsmag");
  oeb_dropin::finish(uout, vout);
}

inline void hdiffsmag_fullfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                 const Storage3D& mask, const Storage1D& crlavo, const Storage1D& crlavu, const Storage1D& crlato,
                                 const Storage1D& crlatu, const Storage1D& acrlat0, Storage3D& T_sqr_s, Storage3D& S_sqr_uv,
                                 const ElementType eddlat, const ElementType eddlon, const ElementType tau_smag,
                                 const ElementType weight_smag) {
  hdiffsmag(uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat, eddlon, tau_smag, weight_smag);
}

inline void hdiffsmag_partialfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                    const Storage3D& mask, const Storage1D& crlavo, const Storage1D& crlavu, const Storage1D& crlato,
                                    const Storage1D& crlatu, const Storage1D& acrlat0, Storage3D& T_sqr_s, Storage3D& S_sqr_uv,
                                    const ElementType eddlat, const ElementType eddlon, const ElementType tau_smag,
                                    const ElementType weight_smag) {
  hdiffsmag(uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat, eddlon, tau_smag, weight_smag);
}

inline void hdiffsmag_openmp(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                             const Storage3D& mask, const Storage1D& crlavo, const Storage1D& crlavu, const Storage1D& crlato,
                             const Storage1D& crlatu, const Storage1D& acrlat0, Storage3D& T_sqr_s, Storage3D& S_sqr_uv,
                             const ElementType eddlat, const ElementType eddlon, const ElementType tau_smag,
                             const ElementType weight_smag) {
  hdiffsmag(uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat, eddlon, tau_smag, weight_smag);
}

#endif  // HDIFFSMAG_H
