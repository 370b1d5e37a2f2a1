// fvtp2d.hip -- FV3 2-D finite-volume tracer transport (PPM fluxes).
// Reference: kernels/fvtp2d.h:4-177 (the DEFAULT variant; there is no OpenMP
// variant).  The reference builds each PPM flux from scratch arrays
// al, almq, br, b0, smt5 (e.g. :9-37).  Here ppm_face() forms the same
// quantities at the three faces m-1, m, m+1 that one flux value needs, with
// the reference's expressions, and keeps them in registers.
//
// Three launches, in stream order:
//   A  fy2 = PPM_j(q, cry) on i [-3,nx+2) x j [0,ny+1)          (:9-37)
//      fx2 = PPM_i(q, crx) on i [0,nx+1)  x j [-3,ny+2)         (:84-113)
//   B  q_i = (q*area + yfx*fy2 - yfx*fy2[j+1]) / ra_y  on i [-3,nx+2) x j [0,ny)  (:39-51)
//      q_j = (q*area + xfx*fx2 - xfx*fx2[i+1]) / ra_x  on i [0,nx) x j [-3,ny+2)  (:115-127)
//   C  fx1 = 0.5*(PPM_i(q_i, crx) + fx2)*xfx,  fy1 = 0.5*(PPM_j(q_j, cry) + fy2)*yfx
//      on [0,nx) x [0,ny)                                       (:53-82, :129-176)
// fy2, fx2, q_i and q_j are outputs of the reference's signature, so each is
// written once in any case.  B and C read them back.
// Algorithmic traffic: 8 inputs + 6 outputs = 112 B per point (FP64).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 8;

// PPM flux at face m between cells m-1 and m of one row or column, given
// f(m-3) .. f(m+2) and the Courant number c at m (kernels/fvtp2d.h:12, 20-23, 28-34).
__device__ __forceinline__ double ppm_face(double fm3, double fm2, double fm1, double f0, double fp1, double fp2,
                                           double c, double p1, double p2) {
  const double al_m = ((p1 * (fm2 + fm1)) + (p2 * (fm3 + f0)));  // al(m-1)
  const double al_0 = ((p1 * (fm1 + f0)) + (p2 * (fm2 + fp1)));  // al(m)
  const double al_p = ((p1 * (f0 + fp1)) + (p2 * (fm1 + fp2)));  // al(m+1)
  const double almq_m = (al_m - fm1), br_m = (al_0 - fm1), b0_m = (almq_m + br_m);
  const double almq_0 = (al_0 - f0), br_0 = (al_p - f0), b0_0 = (almq_0 + br_0);
  const double smt_m = ((almq_m * br_m) < 0.0) ? 1.0 : 0.0;
  const double smt_0 = ((almq_0 * br_0) < 0.0) ? 1.0 : 0.0;
  const double smt = (smt_m + (smt_0 * (smt_m == 0.0 ? 1.0 : 0.0)));
  if (c > 0.0) {
    const double ctmp = ((1.0 - c) * (br_m - (c * b0_m)));
    return (fm1 + (ctmp * smt));
  }
  const double ctmp = ((1.0 + c) * (almq_0 + (c * b0_0)));
  return (f0 + (ctmp * smt));
}

// Flux of field f at (i,j) along the axis of stride s: face between (m-1) and m.
__device__ __forceinline__ double ppm_at(const double *__restrict__ f, int s, double c, double p1, double p2) {
  return ppm_face(f[-3 * s], f[-2 * s], f[-s], f[0], f[s], f[2 * s], c, p1, p2);
}

struct FvIn {
  const double *q, *crx, *cry, *ra_x, *ra_y, *xfx, *yfx, *area;
};
struct FvOut {
  double *q_i, *q_j, *fx1, *fx2, *fy1, *fy2;
};

// Launch A and B cover the box i [-3, nx+2) x j [-3, ny+2); C covers [0,nx) x [0,ny).
__global__ void __launch_bounds__(kTileI *kTileJ)
    fvtp2d_flux2_kernel(FvIn a, FvOut o, double p1, double p2, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x) - 3, j = c.j0 + int(threadIdx.y) - 3;
  if (i >= g.nx + 2 || j >= g.ny + 2) return;
  const bool do_fy2 = j >= 0 && j < g.ny + 1;
  const bool do_fx2 = i >= 0 && i < g.nx + 1;
  const int sj = g.sj;
  int64_t at = int64_t(c.k0) * g.sk + (i + int64_t(j) * sj);
  for (int k = c.k0; k < c.k1; ++k, at += g.sk) {
    if (do_fy2) o.fy2[at] = ppm_at(a.q + at, sj, a.cry[at], p1, p2);
    if (do_fx2) o.fx2[at] = ppm_at(a.q + at, 1, a.crx[at], p1, p2);
  }
}

__global__ void __launch_bounds__(kTileI *kTileJ)
    fvtp2d_qij_kernel(FvIn a, FvOut o, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x) - 3, j = c.j0 + int(threadIdx.y) - 3;
  if (i >= g.nx + 2 || j >= g.ny + 2) return;
  const bool do_qi = j >= 0 && j < g.ny;
  const bool do_qj = i >= 0 && i < g.nx;
  const int sj = g.sj;
  int64_t at = int64_t(c.k0) * g.sk + (i + int64_t(j) * sj);
  for (int k = c.k0; k < c.k1; ++k, at += g.sk) {
    const double qa = a.q[at] * a.area[at];
    if (do_qi) {
      const double fyy0 = (a.yfx[at] * o.fy2[at]);
      const double fyy1 = (a.yfx[at + sj] * o.fy2[at + sj]);
      o.q_i[at] = (((qa + fyy0) - fyy1) / a.ra_y[at]);
    }
    if (do_qj) {
      const double fxx0 = (a.xfx[at] * o.fx2[at]);
      const double fxx1 = (a.xfx[at + 1] * o.fx2[at + 1]);
      o.q_j[at] = (((qa + fxx0) - fxx1) / a.ra_x[at]);
    }
  }
}

__global__ void __launch_bounds__(kTileI *kTileJ)
    fvtp2d_flux1_kernel(FvIn a, FvOut o, double p1, double p2, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  int64_t at = int64_t(c.k0) * g.sk + (i + int64_t(j) * sj);
  for (int k = c.k0; k < c.k1; ++k, at += g.sk) {
    const double fx1 = ppm_at(o.q_i + at, 1, a.crx[at], p1, p2);
    const double fy1 = ppm_at(o.q_j + at, sj, a.cry[at], p1, p2);
    o.fx1[at] = ((0.5 * (fx1 + o.fx2[at])) * a.xfx[at]);
    o.fy1[at] = ((0.5 * (fy1 + o.fy2[at])) * a.yfx[at]);
  }
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_fvtp2d_f64(const oeb_domain *d, const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *fx1,
                              const oeb_s3 *fx2, const oeb_s3 *fy1, const oeb_s3 *fy2, const oeb_s3 *q,
                              const oeb_s3 *crx, const oeb_s3 *cry, const oeb_s3 *ra_x, const oeb_s3 *ra_y,
                              const oeb_s3 *xfx, const oeb_s3 *yfx, const oeb_s3 *area, const oeb_s3 *fxx,
                              const oeb_s3 *fyy, const oeb_s3 *al, const oeb_s3 *almq, const oeb_s3 *br,
                              const oeb_s3 *b0, const oeb_s3 *smt5, double p1, double p2, void *stream) {
  using namespace oeb;
  (void)fxx, (void)fyy, (void)al, (void)almq, (void)br, (void)b0, (void)smt5;
  const char *K = "fvtp2d";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_fy2{-3, nx + 2, 0, ny + 1, 0, nz}, r_fx2{0, nx + 1, -3, ny + 2, 0, nz};
  const Box r_qi{-3, nx + 2, 0, ny, 0, nz}, r_qj{0, nx, -3, ny + 2, 0, nz};
  const Box r_q{-3, nx + 3, -3, ny + 3, 0, nz}, r_area{-3, nx + 2, -3, ny + 2, 0, nz};
  FvIn a;
  FvOut o;
  const double *pq_i, *pq_j, *pfx1, *pfx2, *pfy1, *pfy2;
  OEB_TRY(check3(K, "q_i", q_i, r_qi, g, true, &pq_i));
  OEB_TRY(check3(K, "q_j", q_j, r_qj, g, false, &pq_j));
  OEB_TRY(check3(K, "fx1", fx1, r0, g, false, &pfx1));
  OEB_TRY(check3(K, "fx2", fx2, r_fx2, g, false, &pfx2));
  OEB_TRY(check3(K, "fy1", fy1, r0, g, false, &pfy1));
  OEB_TRY(check3(K, "fy2", fy2, r_fy2, g, false, &pfy2));
  OEB_TRY(check3(K, "q", q, r_q, g, false, &a.q));
  OEB_TRY(check3(K, "crx", crx, r_fx2, g, false, &a.crx));
  OEB_TRY(check3(K, "cry", cry, r_fy2, g, false, &a.cry));
  OEB_TRY(check3(K, "ra_x", ra_x, r_qj, g, false, &a.ra_x));
  OEB_TRY(check3(K, "ra_y", ra_y, r_qi, g, false, &a.ra_y));
  OEB_TRY(check3(K, "xfx", xfx, r_fx2, g, false, &a.xfx));
  OEB_TRY(check3(K, "yfx", yfx, r_fy2, g, false, &a.yfx));
  OEB_TRY(check3(K, "area", area, r_area, g, false, &a.area));
  AliasSet as;
  as.add("q_i", pq_i, r_qi, g, true);
  as.add("q_j", pq_j, r_qj, g, true);
  as.add("fx1", pfx1, r0, g, true);
  as.add("fx2", pfx2, r_fx2, g, true);
  as.add("fy1", pfy1, r0, g, true);
  as.add("fy2", pfy2, r_fy2, g, true);
  as.add("q", a.q, r_q, g, false);
  as.add("crx", a.crx, r_fx2, g, false);
  as.add("cry", a.cry, r_fy2, g, false);
  as.add("ra_x", a.ra_x, r_qj, g, false);
  as.add("ra_y", a.ra_y, r_qi, g, false);
  as.add("xfx", a.xfx, r_fx2, g, false);
  as.add("yfx", a.yfx, r_fy2, g, false);
  as.add("area", a.area, r_area, g, false);
  OEB_TRY(as.check(K));
  o.q_i = const_cast<double *>(pq_i);
  o.q_j = const_cast<double *>(pq_j);
  o.fx1 = const_cast<double *>(pfx1);
  o.fx2 = const_cast<double *>(pfx2);
  o.fy1 = const_cast<double *>(pfy1);
  o.fy2 = const_cast<double *>(pfy2);

  const hipStream_t s = hipStream_t(stream);
  const Tiles tb = make_tiles(nx + 5, ny + 5, int(nz), kTileI, kTileJ, kChunk);
  OEB_TRY(check_tiles(K, tb, nx + 5, ny + 5, int(nz), kTileI, kTileJ));
  const Tiles ti = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kChunk);
  hipLaunchKernelGGL(fvtp2d_flux2_kernel, dim3(tb.n), dim3(kTileI, kTileJ), 0, s, a, o, p1, p2, g, tb);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(fvtp2d_qij_kernel, dim3(tb.n), dim3(kTileI, kTileJ), 0, s, a, o, g, tb);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(fvtp2d_flux1_kernel, dim3(ti.n), dim3(kTileI, kTileJ), 0, s, a, o, p1, p2, g, ti);
  return launch_status(K);
}
