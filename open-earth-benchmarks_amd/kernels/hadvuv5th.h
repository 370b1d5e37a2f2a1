// hadvuv5th.h -- drop-in for the reference's kernels/hadvuv5th.h (kernels/hadvuv5th.h:41-201).
// hadvuv5th, hadvuv5th_fullfusion, hadvuv5th_partialfusion, hadvuv5th_openmp keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_hadvuv5th_f64, or oeb_hadvuv5th_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef HADVUV5TH_H
#define HADVUV5TH_H

#include "util.h"

inline void hadvuv5th(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                      const Storage1D& acrlat0, const Storage1D& acrlat1, const Storage1D& tgrlatda0,
                      const Storage1D& tgrlatda1, Storage3D& uatupos, Storage3D& vatupos, Storage3D& uatvpos,
                      Storage3D& vatvpos, Storage3D& uavg, Storage3D& vavg, Storage3D& ures, Storage3D& vres,
                      const ElementType eddlat, const ElementType eddlon) {
  (void)uatupos;
  (void)vatupos;
  (void)uatvpos;
  (void)vatvpos;
  (void)uavg;
  (void)vavg;
  (void)ures;
  (void)vres;
  oeb_dropin::upload(uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_uout = oeb_dropin::dev(uout);
  const auto a_vout = oeb_dropin::dev(vout);
  const auto a_uin = oeb_dropin::dev(uin);
  const auto a_vin = oeb_dropin::dev(vin);
  const auto a_acrlat0 = oeb_dropin::dev(acrlat0);
  const auto a_acrlat1 = oeb_dropin::dev(acrlat1);
  const auto a_tgrlatda0 = oeb_dropin::dev(tgrlatda0);
  const auto a_tgrlatda1 = oeb_dropin::dev(tgrlatda1);
  oeb_dropin::check(OEB_DROPIN_CALL(hadvuv5th, &d, &a_uout, &a_vout, &a_uin, &a_vin, &a_acrlat0, &a_acrlat1, &a_tgrlatda0, &a_tgrlatda1,
                           nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, eddlat, eddlon,
                           oeb_dropin::stream()),
                    "hadvuv5th");
  oeb_dropin::finish(uout, vout);
}

inline void hadvuv5th_fullfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                 const Storage1D& acrlat0, const Storage1D& acrlat1, const Storage1D& tgrlatda0,
                                 const Storage1D& tgrlatda1, Storage3D& uatupos, Storage3D& vatupos, Storage3D& uatvpos,
                                 Storage3D& vatvpos, Storage3D& uavg, Storage3D& vavg, Storage3D& ures, Storage3D& vres,
                                 const ElementType eddlat, const ElementType eddlon) {
  hadvuv5th(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres, eddlat, eddlon);
}

inline void hadvuv5th_partialfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                    const Storage1D& acrlat0, const Storage1D& acrlat1, const Storage1D& tgrlatda0,
                                    const Storage1D& tgrlatda1, Storage3D& uatupos, Storage3D& vatupos, Storage3D& uatvpos,
                                    Storage3D& vatvpos, Storage3D& uavg, Storage3D& vavg, Storage3D& ures, Storage3D& vres,
                                    const ElementType eddlat, const ElementType eddlon) {
  hadvuv5th(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres, eddlat, eddlon);
}

inline void hadvuv5th_openmp(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                             const Storage1D& acrlat0, const Storage1D& acrlat1, const Storage1D& tgrlatda0,
                             const Storage1D& tgrlatda1, Storage3D& uatupos, Storage3D& vatupos, Storage3D& uatvpos,
                             Storage3D& vatvpos, Storage3D& uavg, Storage3D& vavg, Storage3D& ures, Storage3D& vres,
                             const ElementType eddlat, const ElementType eddlon) {
  hadvuv5th(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres, eddlat, eddlon);
}

#endif  // HADVUV5TH_H
