// uvbke.h -- drop-in for the reference's kernels/uvbke.h (kernels/uvbke.h:4-75; dt5 is the driver global, src/uvbke.cpp:11-12).
// uvbke, uvbke_fullfusion, uvbke_partialfusion, uvbke_openmp keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_uvbke_f64, or oeb_uvbke_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef UVBKE_H
#define UVBKE_H

#include "util.h"

inline void uvbke(Storage3D& ub, Storage3D& vb, const Storage3D& uc, const Storage3D& vc, const Storage3D& cosa,
                  const Storage3D& rsina) {
  oeb_dropin::upload(uc, vc, cosa, rsina);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_ub = oeb_dropin::dev(ub);
  const auto a_vb = oeb_dropin::dev(vb);
  const auto a_uc = oeb_dropin::dev(uc);
  const auto a_vc = oeb_dropin::dev(vc);
  const auto a_cosa = oeb_dropin::dev(cosa);
  const auto a_rsina = oeb_dropin::dev(rsina);
  oeb_dropin::check(OEB_DROPIN_CALL(uvbke, &d, &a_ub, &a_vb, &a_uc, &a_vc, &a_cosa, &a_rsina, dt5, oeb_dropin::stream()),
                    "uvbke");
  oeb_dropin::finish(ub, vb);
}

inline void uvbke_fullfusion(Storage3D& ub, Storage3D& vb, const Storage3D& uc, const Storage3D& vc, const Storage3D& cosa,
                             const Storage3D& rsina) {
  uvbke(ub, vb, uc, vc, cosa, rsina);
}

inline void uvbke_partialfusion(Storage3D& ub, Storage3D& vb, const Storage3D& uc, const Storage3D& vc, const Storage3D& cosa,
                                const Storage3D& rsina) {
  uvbke(ub, vb, uc, vc, cosa, rsina);
}

inline void uvbke_openmp(Storage3D& ub, Storage3D& vb, const Storage3D& uc, const Storage3D& vc, const Storage3D& cosa,
                         const Storage3D& rsina) {
  uvbke(ub, vb, uc, vc, cosa, rsina);
}

#endif  // UVBKE_H
