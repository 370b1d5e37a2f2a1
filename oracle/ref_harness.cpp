// ref_harness.cpp -- exposes the reference's own kernels (kernels/*.h and
// include/util.h of open-earth-benchmarks) as extern "C" functions.
//
// TEST INFRASTRUCTURE ONLY.  It pins the oracle and serves as the CPU
// baseline.  oracle/Makefile compiles it against the reference headers in
// place (-I<reference>/include -I<reference>/kernels) into
// oracle/_ref/libref.so (and libref_f32.so).  No reference source is copied into this
// repository: this file #includes the headers where they lie.
//
// The reference drivers (src/<name>.cpp:1-12) define the globals
// domain_size, domain_height, halo_width and count, the ElementType typedef,
// and the per-kernel constants (p1/p2 in src/fvtp2d.cpp:11-12, dt/dt5 in
// src/uvbke.cpp:11-12) before they include util.h and the kernel header.  This
// file plays the driver's part in the same way.  The driver values are kept.
// Each kernel header is included inside its own namespace, because hadvuv.h
// and hadvuv5th.h both define advectionDriver.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include <omp.h>

// Before the driver macros below: p1/p2 are also parameter names in oeb.h.
#include "../include/oeb.h"

int64_t domain_size = 64;
int64_t domain_height = 60;
int64_t halo_width = 3;
int64_t count = 1;

// Built twice: ElementType double (_ref/libref.so) and, with -DOEB_REF_FLOAT,
// ElementType float (_ref/libref_f32.so), the two element types the
// reference drivers can select (src/<name>.cpp:9).  RS3 / RS1 are the
// descriptors of include/oeb.h for that element type.
#ifdef OEB_REF_FLOAT
typedef float ElementType;
typedef oeb_s3f RS3;
typedef oeb_s1f RS1;
#else
typedef double ElementType;
typedef oeb_s3 RS3;
typedef oeb_s1 RS1;
#endif

#define p1 ElementType(0.583333)
#define p2 ElementType(0.083333)
const double dt = 225.0;
const double dt5 = 0.5 * dt;

#include "util.h"

namespace rk_laplace {
#include "laplace.h"
}
namespace rk_hdiffsa {
#include "hdiffsa.h"
}
namespace rk_hdiffsmag {
#include "hdiffsmag.h"
}
namespace rk_hadvuv {
#include "hadvuv.h"
}
namespace rk_hadvuv5th {
#include "hadvuv5th.h"
}
namespace rk_fastwavesuv {
#include "fastwavesuv.h"
}
namespace rk_uvbke {
#include "uvbke.h"
}
namespace rk_p_grad_c {
#include "p_grad_c.h"
}
namespace rk_nh_p_grad {
#include "nh_p_grad.h"
}
namespace rk_fvtp2d {
#include "fvtp2d.h"
}

namespace {

// Variant numbers, as the ALGO values of include/util.h:10-13.
enum { V_DEFAULT = 0, V_FULL_FUSION = 1, V_PARTIAL_FUSION = 2, V_OPENMP = 3 };

Storage3D to_ref(const RS3 *s) {
  Storage3D r;
  r.allocatedPtr = s->allocatedPtr;
  r.alignedPtr = s->alignedPtr;
  r.offset = s->offset;
  for (int a = 0; a < 3; ++a) {
    r.sizes[a] = s->sizes[a];
    r.strides[a] = s->strides[a];
  }
  return r;
}

Storage1D to_ref(const RS1 *s) {
  Storage1D r;
  r.allocatedPtr = s->allocatedPtr;
  r.alignedPtr = s->alignedPtr;
  r.offset = s->offset;
  r.sizes[0] = s->sizes[0];
  r.strides[0] = s->strides[0];
  return r;
}

// Zero-filled scratch with the layout of `like`; released by ~Scratch.
struct Scratch {
  Storage3D s;
  explicit Scratch(const RS3 *like) {
    s = to_ref(like);
    const size_t n = size_t(like->sizes[0] * like->strides[0]) + 64;
    s.allocatedPtr = static_cast<ElementType *>(calloc(n, sizeof(ElementType)));
    s.alignedPtr = s.allocatedPtr;
  }
  ~Scratch() { free(s.allocatedPtr); }
  Scratch(const Scratch &) = delete;
  Scratch &operator=(const Scratch &) = delete;
};

// The reference kernels take one square extent; returns 0 if d is not square.
int set_domain(const oeb_domain *d) {
  if (d->nx != d->ny) return 0;
  domain_size = d->nx;
  domain_height = d->nz;
  return 1;
}

}  // namespace

extern "C" {

void ref_set_threads(int n) { omp_set_num_threads(n); }
int ref_max_threads(void) { return omp_get_max_threads(); }

// Each ref_<name> returns 0 on success, -1 for a non-square domain and -2 for
// a variant the reference does not have.

int ref_laplace(int variant, const oeb_domain *d, const RS3 *in, const RS3 *out) {
  if (!set_domain(d)) return -1;
  Storage3D i = to_ref(in), o = to_ref(out);
  switch (variant) {
    case V_DEFAULT: rk_laplace::laplace(i, o); return 0;
    case V_FULL_FUSION: rk_laplace::laplace_fullfusion(i, o); return 0;
    case V_PARTIAL_FUSION: rk_laplace::laplace_partialfusion(i, o); return 0;
    case V_OPENMP: rk_laplace::laplace_openmp(i, o); return 0;
  }
  return -2;
}

int ref_hdiffsa(int variant, const oeb_domain *d, const RS3 *out, const RS3 *in, const RS3 *mask,
                const RS1 *crlato, const RS1 *crlatu) {
  if (!set_domain(d)) return -1;
  Scratch lap(in), flx(in), fly(in);
  Storage3D o = to_ref(out), i = to_ref(in), m = to_ref(mask);
  Storage1D lo = to_ref(crlato), lu = to_ref(crlatu);
  switch (variant) {
    case V_DEFAULT: rk_hdiffsa::hdiffsa(o, i, m, lo, lu, lap.s, flx.s, fly.s); return 0;
    case V_FULL_FUSION: rk_hdiffsa::hdiffsa_fullfusion(o, i, m, lo, lu, lap.s, flx.s, fly.s); return 0;
    case V_PARTIAL_FUSION: rk_hdiffsa::hdiffsa_partialfusion(o, i, m, lo, lu, lap.s, flx.s, fly.s); return 0;
    case V_OPENMP: rk_hdiffsa::hdiffsa_openmp(o, i, m, lo, lu, lap.s, flx.s, fly.s); return 0;
  }
  return -2;
}

int ref_hdiffsmag(int variant, const oeb_domain *d, const RS3 *uout, const RS3 *vout, const RS3 *uin,
                  const RS3 *vin, const RS3 *mask, const RS1 *crlavo, const RS1 *crlavu,
                  const RS1 *crlato, const RS1 *crlatu, const RS1 *acrlat0, double eddlat, double eddlon,
                  double tau_smag, double weight_smag) {
  if (!set_domain(d)) return -1;
  Scratch t(uin), s(uin);
  Storage3D uo = to_ref(uout), vo = to_ref(vout), ui = to_ref(uin), vi = to_ref(vin), m = to_ref(mask);
  Storage1D a = to_ref(crlavo), b = to_ref(crlavu), c = to_ref(crlato), e = to_ref(crlatu), f = to_ref(acrlat0);
  switch (variant) {
    case V_DEFAULT:
      rk_hdiffsmag::hdiffsmag(uo, vo, ui, vi, m, a, b, c, e, f, t.s, s.s, eddlat, eddlon, tau_smag, weight_smag);
      return 0;
    case V_FULL_FUSION:
      rk_hdiffsmag::hdiffsmag_fullfusion(uo, vo, ui, vi, m, a, b, c, e, f, t.s, s.s, eddlat, eddlon, tau_smag,
                                          weight_smag);
      return 0;
    case V_PARTIAL_FUSION:
      rk_hdiffsmag::hdiffsmag_partialfusion(uo, vo, ui, vi, m, a, b, c, e, f, t.s, s.s, eddlat, eddlon, tau_smag,
                                             weight_smag);
      return 0;
    case V_OPENMP:
      rk_hdiffsmag::hdiffsmag_openmp(uo, vo, ui, vi, m, a, b, c, e, f, t.s, s.s, eddlat, eddlon, tau_smag,
                                      weight_smag);
      return 0;
  }
  return -2;
}

#define HADV_SCRATCH(like) \
  Scratch s0(like), s1(like), s2(like), s3(like), s4(like), s5(like), s6(like), s7(like)
#define HADV_ARGS uo, vo, ui, vi, a0, a1, t0, t1, s0.s, s1.s, s2.s, s3.s, s4.s, s5.s, s6.s, s7.s, eddlat, eddlon

int ref_hadvuv(int variant, const oeb_domain *d, const RS3 *uout, const RS3 *vout, const RS3 *uin,
               const RS3 *vin, const RS1 *acrlat0, const RS1 *acrlat1, const RS1 *tgrlatda0,
               const RS1 *tgrlatda1, double eddlat, double eddlon) {
  if (!set_domain(d)) return -1;
  HADV_SCRATCH(uin);
  Storage3D uo = to_ref(uout), vo = to_ref(vout), ui = to_ref(uin), vi = to_ref(vin);
  Storage1D a0 = to_ref(acrlat0), a1 = to_ref(acrlat1), t0 = to_ref(tgrlatda0), t1 = to_ref(tgrlatda1);
  switch (variant) {
    case V_DEFAULT: rk_hadvuv::hadvuv(HADV_ARGS); return 0;
    case V_FULL_FUSION: rk_hadvuv::hadvuv_fullfusion(HADV_ARGS); return 0;
    case V_PARTIAL_FUSION: rk_hadvuv::hadvuv_partialfusion(HADV_ARGS); return 0;
    case V_OPENMP: rk_hadvuv::hadvuv_openmp(HADV_ARGS); return 0;
  }
  return -2;
}

int ref_hadvuv5th(int variant, const oeb_domain *d, const RS3 *uout, const RS3 *vout, const RS3 *uin,
                  const RS3 *vin, const RS1 *acrlat0, const RS1 *acrlat1, const RS1 *tgrlatda0,
                  const RS1 *tgrlatda1, double eddlat, double eddlon) {
  if (!set_domain(d)) return -1;
  HADV_SCRATCH(uin);
  Storage3D uo = to_ref(uout), vo = to_ref(vout), ui = to_ref(uin), vi = to_ref(vin);
  Storage1D a0 = to_ref(acrlat0), a1 = to_ref(acrlat1), t0 = to_ref(tgrlatda0), t1 = to_ref(tgrlatda1);
  switch (variant) {
    case V_DEFAULT: rk_hadvuv5th::hadvuv5th(HADV_ARGS); return 0;
    case V_FULL_FUSION: rk_hadvuv5th::hadvuv5th_fullfusion(HADV_ARGS); return 0;
    case V_PARTIAL_FUSION: rk_hadvuv5th::hadvuv5th_partialfusion(HADV_ARGS); return 0;
    case V_OPENMP: rk_hadvuv5th::hadvuv5th_openmp(HADV_ARGS); return 0;
  }
  return -2;
}

int ref_fastwavesuv(int variant, const oeb_domain *d, const RS3 *uout, const RS3 *vout, const RS3 *uin,
                    const RS3 *vin, const RS3 *utens, const RS3 *vtens, const RS3 *wgtfac,
                    const RS3 *ppuv, const RS3 *hhl, const RS3 *rho, const RS1 *fx, double edadlat,
                    double dtv) {
  if (!set_domain(d)) return -1;
  Scratch a(uin), b(uin), c(uin), e(uin);
  Storage3D uo = to_ref(uout), vo = to_ref(vout), ui = to_ref(uin), vi = to_ref(vin), ut = to_ref(utens),
            vt = to_ref(vtens), w = to_ref(wgtfac), p = to_ref(ppuv), h = to_ref(hhl), r = to_ref(rho);
  Storage1D f = to_ref(fx);
  switch (variant) {
    case V_DEFAULT:
      rk_fastwavesuv::fastwavesuv(uo, vo, ui, vi, ut, vt, w, p, h, r, f, a.s, b.s, c.s, e.s, edadlat, dtv);
      return 0;
    case V_FULL_FUSION:
      rk_fastwavesuv::fastwavesuv_fullfusion(uo, vo, ui, vi, ut, vt, w, p, h, r, f, a.s, b.s, c.s, e.s, edadlat,
                                              dtv);
      return 0;
    case V_PARTIAL_FUSION:
      rk_fastwavesuv::fastwavesuv_partialfusion(uo, vo, ui, vi, ut, vt, w, p, h, r, f, a.s, b.s, c.s, e.s, edadlat,
                                                 dtv);
      return 0;
    case V_OPENMP:
      rk_fastwavesuv::fastwavesuv_openmp(uo, vo, ui, vi, ut, vt, w, p, h, r, f, a.s, b.s, c.s, e.s, edadlat, dtv);
      return 0;
  }
  return -2;
}

// dt5 is the driver's constant (src/uvbke.cpp:11-12); the reference kernel
// reads the global, so this entry has no dt5 argument.
int ref_uvbke(int variant, const oeb_domain *d, const RS3 *ub, const RS3 *vb, const RS3 *uc,
              const RS3 *vc, const RS3 *cosa, const RS3 *rsina) {
  if (!set_domain(d)) return -1;
  Storage3D a = to_ref(ub), b = to_ref(vb), c = to_ref(uc), e = to_ref(vc), f = to_ref(cosa), g = to_ref(rsina);
  switch (variant) {
    case V_DEFAULT: rk_uvbke::uvbke(a, b, c, e, f, g); return 0;
    case V_FULL_FUSION: rk_uvbke::uvbke_fullfusion(a, b, c, e, f, g); return 0;
    case V_PARTIAL_FUSION: rk_uvbke::uvbke_partialfusion(a, b, c, e, f, g); return 0;
    case V_OPENMP: rk_uvbke::uvbke_openmp(a, b, c, e, f, g); return 0;
  }
  return -2;
}
double ref_uvbke_dt5(void) { return dt5; }

int ref_p_grad_c(int variant, const oeb_domain *d, const RS3 *uout, const RS3 *vout, const RS3 *uin,
                 const RS3 *vin, const RS3 *rdxc, const RS3 *rdyc, const RS3 *delpc, const RS3 *gz,
                 const RS3 *pkc, double dt2) {
  if (!set_domain(d)) return -1;
  Scratch wk(uin);
  Storage3D uo = to_ref(uout), vo = to_ref(vout), ui = to_ref(uin), vi = to_ref(vin), x = to_ref(rdxc),
            y = to_ref(rdyc), dp = to_ref(delpc), g = to_ref(gz), p = to_ref(pkc);
  switch (variant) {
    case V_DEFAULT: rk_p_grad_c::p_grad_c(uo, vo, ui, vi, x, y, dp, g, p, wk.s, dt2); return 0;
    case V_FULL_FUSION: rk_p_grad_c::p_grad_c_fullfusion(uo, vo, ui, vi, x, y, dp, g, p, wk.s, dt2); return 0;
    case V_PARTIAL_FUSION: rk_p_grad_c::p_grad_c_partialfusion(uo, vo, ui, vi, x, y, dp, g, p, wk.s, dt2); return 0;
    case V_OPENMP: rk_p_grad_c::p_grad_c_openmp(uo, vo, ui, vi, x, y, dp, g, p, wk.s, dt2); return 0;
  }
  return -2;
}

int ref_nh_p_grad(int variant, const oeb_domain *d, const RS3 *uout, const RS3 *vout, const RS3 *uin,
                  const RS3 *vin, const RS3 *rdx, const RS3 *rdy, const RS3 *gz, const RS3 *pp,
                  const RS3 *pk3, const RS3 *wk1, double dtv) {
  if (!set_domain(d)) return -1;
  Scratch wk(uin), du(uin), dv(uin);
  Storage3D uo = to_ref(uout), vo = to_ref(vout), ui = to_ref(uin), vi = to_ref(vin), x = to_ref(rdx),
            y = to_ref(rdy), g = to_ref(gz), p = to_ref(pp), p3 = to_ref(pk3), w1 = to_ref(wk1);
  switch (variant) {
    case V_DEFAULT: rk_nh_p_grad::nh_p_grad(uo, vo, ui, vi, x, y, g, p, p3, w1, wk.s, du.s, dv.s, dtv); return 0;
    case V_FULL_FUSION:
      rk_nh_p_grad::nh_p_grad_fullfusion(uo, vo, ui, vi, x, y, g, p, p3, w1, wk.s, du.s, dv.s, dtv);
      return 0;
    case V_PARTIAL_FUSION:
      rk_nh_p_grad::nh_p_grad_partialfusion(uo, vo, ui, vi, x, y, g, p, p3, w1, wk.s, du.s, dv.s, dtv);
      return 0;
    case V_OPENMP:
      rk_nh_p_grad::nh_p_grad_openmp(uo, vo, ui, vi, x, y, g, p, p3, w1, wk.s, du.s, dv.s, dtv);
      return 0;
  }
  return -2;
}

// Only DEFAULT (0).  fvtp2d_fullfusion reads the never-written br array
// (kernels/fvtp2d.h:195-196), so it is not exposed.
int ref_fvtp2d(int variant, const oeb_domain *d, const RS3 *q_i, const RS3 *q_j, const RS3 *fx1,
               const RS3 *fx2, const RS3 *fy1, const RS3 *fy2, const RS3 *q, const RS3 *crx,
               const RS3 *cry, const RS3 *ra_x, const RS3 *ra_y, const RS3 *xfx, const RS3 *yfx,
               const RS3 *area) {
  if (!set_domain(d)) return -1;
  if (variant != V_DEFAULT) return -2;
  Scratch fxx(q), fyy(q), al(q), almq(q), br(q), b0(q), smt5(q);
  Storage3D a = to_ref(q_i), b = to_ref(q_j), c = to_ref(fx1), e = to_ref(fx2), f = to_ref(fy1), g = to_ref(fy2),
            qq = to_ref(q), cx = to_ref(crx), cy = to_ref(cry), rx = to_ref(ra_x), ry = to_ref(ra_y), xf = to_ref(xfx),
            yf = to_ref(yfx), ar = to_ref(area);
  rk_fvtp2d::fvtp2d(a, b, c, e, f, g, qq, cx, cy, rx, ry, xf, yf, ar, fxx.s, fyy.s, al.s, almq.s, br.s, b0.s,
                     smt5.s);
  return 0;
}
double ref_fvtp2d_p1(void) { return p1; }
double ref_fvtp2d_p2(void) { return p2; }

}  // extern "C"
