// fastwavesuv.h -- drop-in for the reference's kernels/fastwavesuv.h (kernels/fastwavesuv.h:4-191).
// fastwavesuv, fastwavesuv_fullfusion, fastwavesuv_partialfusion, fastwavesuv_openmp keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_fastwavesuv_f64, or oeb_fastwavesuv_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef FASTWAVESUV_H
#define FASTWAVESUV_H

#include "util.h"

inline void fastwavesuv(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                        const Storage3D& utens, const Storage3D& vtens, const Storage3D& wgtfac,
                        const Storage3D& ppuv, const Storage3D& hhl, const Storage3D& rho, const Storage1D& fx,
                        Storage3D& ppgk, Storage3D& ppgc, Storage3D& ppgu, Storage3D& ppgv, const ElementType edadlat,
                        const ElementType dt) {
  (void)ppgk;
  (void)ppgc;
  (void)ppgu;
  (void)ppgv;
  oeb_dropin::upload(uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_uout = oeb_dropin::dev(uout);
  const auto a_vout = oeb_dropin::dev(vout);
  const auto a_uin = oeb_dropin::dev(uin);
  const auto a_vin = oeb_dropin::dev(vin);
  const auto a_utens = oeb_dropin::dev(utens);
  const auto a_vtens = oeb_dropin::dev(vtens);
  const auto a_wgtfac = oeb_dropin::dev(wgtfac);
  const auto a_ppuv = oeb_dropin::island(ppuv);
  const auto a_hhl = oeb_dropin::island(hhl);
  const auto a_rho = oeb_dropin::dev(rho);
  const auto a_fx = oeb_dropin::dev(fx);
  oeb_dropin::check(OEB_DROPIN_CALL(fastwavesuv, &d, &a_uout, &a_vout, &a_uin, &a_vin, &a_utens, &a_vtens, &a_wgtfac, &a_ppuv, &a_hhl,
                            &a_rho, &a_fx, nullptr, nullptr, nullptr, nullptr, edadlat, dt, oeb_dropin::stream()),
                    "fastwavesuv");
  oeb_dropin::finish(uout, vout);
}

inline void fastwavesuv_fullfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                   const Storage3D& utens, const Storage3D& vtens, const Storage3D& wgtfac,
                                   const Storage3D& ppuv, const Storage3D& hhl, const Storage3D& rho, const Storage1D& fx,
                                   Storage3D& ppgk, Storage3D& ppgc, Storage3D& ppgu, Storage3D& ppgv, const ElementType edadlat,
                                   const ElementType dt) {
  fastwavesuv(uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt);
}

inline void fastwavesuv_partialfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                      const Storage3D& utens, const Storage3D& vtens, const Storage3D& wgtfac,
                                      const Storage3D& ppuv, const Storage3D& hhl, const Storage3D& rho, const Storage1D& fx,
                                      Storage3D& ppgk, Storage3D& ppgc, Storage3D& ppgu, Storage3D& ppgv, const ElementType edadlat,
                                      const ElementType dt) {
  fastwavesuv(uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt);
}

inline void fastwavesuv_openmp(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                               const Storage3D& utens, const Storage3D& vtens, const Storage3D& wgtfac,
                               const Storage3D& ppuv, const Storage3D& hhl, const Storage3D& rho, const Storage1D& fx,
                               Storage3D& ppgk, Storage3D& ppgc, Storage3D& ppgu, Storage3D& ppgv, const ElementType edadlat,
                               const ElementType dt) {
  fastwavesuv(uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt);
}

#endif  // FASTWAVESUV_H
