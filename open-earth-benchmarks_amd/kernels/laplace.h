// laplace.h -- drop-in for the reference's kernels/laplace.h (kernels/laplace.h:4-43).
// laplace, laplace_fullfusion, laplace_partialfusion, laplace_openmp keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_laplace_f64, or oeb_laplace_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef LAPLACE_H
#define LAPLACE_H

#include "util.h"

inline void laplace(const Storage3D& in, Storage3D& out) {
  oeb_dropin::upload(in);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_in = oeb_dropin::dev(in);
  const auto a_out = oeb_dropin::dev(out);
  oeb_dropin::check(OEB_DROPIN_CALL(laplace, &d, &a_in, &a_out, oeb_dropin::stream()),
                    "laplace");
  oeb_dropin::finish(out);
}

inline void laplace_fullfusion(const Storage3D& in, Storage3D& out) {
  laplace(in, out);
}

inline void laplace_partialfusion(const Storage3D& in, Storage3D& out) {
  laplace(in, out);
}

inline void laplace_openmp(const Storage3D& in, Storage3D& out) {
  laplace(in, out);
}

#endif  // LAPLACE_H
