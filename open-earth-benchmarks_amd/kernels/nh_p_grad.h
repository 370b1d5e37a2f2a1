// nh_p_grad.h -- drop-in for the reference's kernels/nh_p_grad.h (kernels/nh_p_grad.h:4-146).
// nh_p_grad, nh_p_grad_fullfusion, nh_p_grad_partialfusion, nh_p_grad_openmp keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_nh_p_grad_f64, or oeb_nh_p_grad_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef NH_P_GRAD_H
#define NH_P_GRAD_H

#include "util.h"

inline void nh_p_grad(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                      const Storage3D& rdx, const Storage3D& rdy, const Storage3D& gz, const Storage3D& pp,
                      const Storage3D& pk3, const Storage3D& wk1, Storage3D& wk, Storage3D& du, Storage3D& dv,
                      const ElementType dt) {
  (void)wk;
  (void)du;
  (void)dv;
  oeb_dropin::upload(uin, vin, rdx, rdy, gz, pp, pk3, wk1);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_uout = oeb_dropin::dev(uout);
  const auto a_vout = oeb_dropin::dev(vout);
  const auto a_uin = oeb_dropin::dev(uin);
  const auto a_vin = oeb_dropin::dev(vin);
  const auto a_rdx = oeb_dropin::dev(rdx);
  const auto a_rdy = oeb_dropin::dev(rdy);
  const auto a_gz = oeb_dropin::island(gz);
  const auto a_pp = oeb_dropin::island(pp);
  const auto a_pk3 = oeb_dropin::island(pk3);
  const auto a_wk1 = oeb_dropin::dev(wk1);
  oeb_dropin::check(OEB_DROPIN_CALL(nh_p_grad, &d, &a_uout, &a_vout, &a_uin, &a_vin, &a_rdx, &a_rdy, &a_gz, &a_pp, &a_pk3, &a_wk1, nullptr,
                          nullptr, nullptr, dt, oeb_dropin::stream()),
                    "nh_p_grad");
  oeb_dropin::finish(uout, vout);
}

inline void nh_p_grad_fullfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                 const Storage3D& rdx, const Storage3D& rdy, const Storage3D& gz, const Storage3D& pp,
                                 const Storage3D& pk3, const Storage3D& wk1, Storage3D& wk, Storage3D& du, Storage3D& dv,
                                 const ElementType dt) {
  nh_p_grad(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv, dt);
}

inline void nh_p_grad_partialfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                    const Storage3D& rdx, const Storage3D& rdy, const Storage3D& gz, const Storage3D& pp,
                                    const Storage3D& pk3, const Storage3D& wk1, Storage3D& wk, Storage3D& du, Storage3D& dv,
                                    const ElementType dt) {
  nh_p_grad(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv, dt);
}

inline void nh_p_grad_openmp(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                             const Storage3D& rdx, const Storage3D& rdy, const Storage3D& gz, const Storage3D& pp,
                             const Storage3D& pk3, const Storage3D& wk1, Storage3D& wk, Storage3D& du, Storage3D& dv,
                             const ElementType dt) {
  nh_p_grad(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk, du, dv, dt);
}

#endif  // NH_P_GRAD_H
