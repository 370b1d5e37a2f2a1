// p_grad_c.h -- drop-in for the reference's kernels/p_grad_c.h (kernels/p_grad_c.h:4-101).
// p_grad_c, p_grad_c_fullfusion, p_grad_c_partialfusion, p_grad_c_openmp keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_p_grad_c_f64, or oeb_p_grad_c_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef P_GRAD_C_H
#define P_GRAD_C_H

#include "util.h"

inline void p_grad_c(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                     const Storage3D& rdxc, const Storage3D& rdyc, const Storage3D& delpc, const Storage3D& gz,
                     const Storage3D& pkc, Storage3D& wk, const ElementType dt2) {
  (void)wk;
  oeb_dropin::upload(uin, vin, rdxc, rdyc, delpc, gz, pkc);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_uout = oeb_dropin::dev(uout);
  const auto a_vout = oeb_dropin::dev(vout);
  const auto a_uin = oeb_dropin::dev(uin);
  const auto a_vin = oeb_dropin::dev(vin);
  const auto a_rdxc = oeb_dropin::dev(rdxc);
  const auto a_rdyc = oeb_dropin::dev(rdyc);
  const auto a_delpc = oeb_dropin::dev(delpc);
  const auto a_gz = oeb_dropin::island(gz);
  const auto a_pkc = oeb_dropin::island(pkc);
  oeb_dropin::check(OEB_DROPIN_CALL(p_grad_c, &d, &a_uout, &a_vout, &a_uin, &a_vin, &a_rdxc, &a_rdyc, &a_delpc, &a_gz, &a_pkc, nullptr, dt2,
                         oeb_dropin::stream()),
                    "p_grad_c");
  oeb_dropin::finish(uout, vout);
}

inline void p_grad_c_fullfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                const Storage3D& rdxc, const Storage3D& rdyc, const Storage3D& delpc, const Storage3D& gz,
                                const Storage3D& pkc, Storage3D& wk, const ElementType dt2) {
  p_grad_c(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2);
}

inline void p_grad_c_partialfusion(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                                   const Storage3D& rdxc, const Storage3D& rdyc, const Storage3D& delpc, const Storage3D& gz,
                                   const Storage3D& pkc, Storage3D& wk, const ElementType dt2) {
  p_grad_c(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2);
}

inline void p_grad_c_openmp(Storage3D& uout, Storage3D& vout, const Storage3D& uin, const Storage3D& vin,
                            const Storage3D& rdxc, const Storage3D& rdyc, const Storage3D& delpc, const Storage3D& gz,
                            const Storage3D& pkc, Storage3D& wk, const ElementType dt2) {
  p_grad_c(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2);
}

#endif  // P_GRAD_C_H
