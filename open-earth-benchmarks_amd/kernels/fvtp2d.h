// fvtp2d.h -- drop-in for the reference's kernels/fvtp2d.h (kernels/fvtp2d.h:4-336; p1/p2 are the driver macros, src/fvtp2d.cpp:11-12).
// fvtp2d, fvtp2d_fullfusion keep their names and parameter lists; each runs liboeb's fused
// HIP kernel (oeb_fvtp2d_f64, or oeb_fvtp2d_f32 when ElementType is float; include/oeb.h).  The reference variants are CPU
// loop-fusion studies with one result; on the GPU there is one implementation.
// Scratch arguments are accepted and left untouched.
#ifndef FVTP2D_H
#define FVTP2D_H

#include "util.h"

inline void fvtp2d(Storage3D& q_i, Storage3D& q_j, Storage3D& fx1, Storage3D& fx2, Storage3D& fy1, Storage3D& fy2,
                   const Storage3D& q, const Storage3D& crx, const Storage3D& cry, const Storage3D& ra_x,
                   const Storage3D& ra_y, const Storage3D& xfx, const Storage3D& yfx, const Storage3D& area,
                   Storage3D& fxx, Storage3D& fyy, Storage3D& al, Storage3D& almq, Storage3D& br, Storage3D& b0,
                   Storage3D& smt5) {
  (void)fxx;
  (void)fyy;
  (void)al;
  (void)almq;
  (void)br;
  (void)b0;
  (void)smt5;
  oeb_dropin::upload(q, crx, cry, ra_x, ra_y, xfx, yfx, area);
  const oeb_domain d = oeb_dropin::domain();
  const auto a_q_i = oeb_dropin::dev(q_i);
  const auto a_q_j = oeb_dropin::dev(q_j);
  const auto a_fx1 = oeb_dropin::dev(fx1);
  const auto a_fx2 = oeb_dropin::dev(fx2);
  const auto a_fy1 = oeb_dropin::dev(fy1);
  const auto a_fy2 = oeb_dropin::dev(fy2);
  const auto a_q = oeb_dropin::dev(q);
  const auto a_crx = oeb_dropin::dev(crx);
  const auto a_cry = oeb_dropin::dev(cry);
  const auto a_ra_x = oeb_dropin::dev(ra_x);
  const auto a_ra_y = oeb_dropin::dev(ra_y);
  const auto a_xfx = oeb_dropin::dev(xfx);
  const auto a_yfx = oeb_dropin::dev(yfx);
  const auto a_area = oeb_dropin::dev(area);
  oeb_dropin::check(OEB_DROPIN_CALL(fvtp2d, &d, &a_q_i, &a_q_j, &a_fx1, &a_fx2, &a_fy1, &a_fy2, &a_q, &a_crx, &a_cry, &a_ra_x, &a_ra_y,
                       &a_xfx, &a_yfx, &a_area, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, p1, p2,
                       oeb_dropin::stream()),
                    "fvtp2d");
  oeb_dropin::finish(q_i, q_j, fx1, fx2, fy1, fy2);
}

inline void fvtp2d_fullfusion(Storage3D& q_i, Storage3D& q_j, Storage3D& fx1, Storage3D& fx2, Storage3D& fy1, Storage3D& fy2,
                              const Storage3D& q, const Storage3D& crx, const Storage3D& cry, const Storage3D& ra_x,
                              const Storage3D& ra_y, const Storage3D& xfx, const Storage3D& yfx, const Storage3D& area,
                              Storage3D& fxx, Storage3D& fyy, Storage3D& al, Storage3D& almq, Storage3D& br, Storage3D& b0,
                              Storage3D& smt5) {
  fvtp2d(q_i, q_j, fx1, fx2, fy1, fy2, q, crx, cry, ra_x, ra_y, xfx, yfx, area, fxx, fyy, al, almq, br, b0, smt5);
}

#endif  // FVTP2D_H
