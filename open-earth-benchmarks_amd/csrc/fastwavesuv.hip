// fastwavesuv.hip -- COSMO fast-wave (u, v) update.
// Reference: kernels/fastwavesuv.h:154-191 (fastwavesuv_openmp), expressions
// :162-187.  ppgk (pressure at the half level below k) is formed at the
// centre C, east E = (i+1) and north N = (j+1) points.  It is carried from
// level k+1 to level k of the next step, and so are ppuv and hhl, so each plane
// of wgtfac, ppuv and hhl is loaded once per chunk.
// Reach: i, j in [0, +1]; k in [-1, +1] (the k-halo planes -1 and nz are read).
// Algorithmic traffic: 8 inputs + 2 outputs = 80 B per point (FP64).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 16;

__device__ __forceinline__ double ppgk_of(double wgt, double pp_k, double pp_km1) {
  return wgt * pp_k + (1.0 - wgt) * pp_km1;
}

__global__ void __launch_bounds__(kTileI *kTileJ)
    fastwavesuv_kernel(double *__restrict__ uout, double *__restrict__ vout, const double *__restrict__ uin,
                       const double *__restrict__ vin, const double *__restrict__ utens,
                       const double *__restrict__ vtens, const double *__restrict__ wgtfac,
                       const double *__restrict__ ppuv, const double *__restrict__ hhl,
                       const double *__restrict__ rho, const double *__restrict__ fx, int fx_stride, double edadlat,
                       double dt, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  const double fxj = fx[int64_t(j) * fx_stride];
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  const int64_t dn = at - sk;
  double ppC = ppuv[at], ppE = ppuv[at + 1], ppN = ppuv[at + sj];
  double gkC = ppgk_of(wgtfac[at], ppC, ppuv[dn]);
  double gkE = ppgk_of(wgtfac[at + 1], ppE, ppuv[dn + 1]);
  double gkN = ppgk_of(wgtfac[at + sj], ppN, ppuv[dn + sj]);
  double hC = hhl[at], hE = hhl[at + 1], hN = hhl[at + sj];
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const double ppC1 = ppuv[up], ppE1 = ppuv[up + 1], ppN1 = ppuv[up + sj];
    const double gkC1 = ppgk_of(wgtfac[up], ppC1, ppC);
    const double gkE1 = ppgk_of(wgtfac[up + 1], ppE1, ppE);
    const double gkN1 = ppgk_of(wgtfac[up + sj], ppN1, ppN);
    const double hC1 = hhl[up], hE1 = hhl[up + 1], hN1 = hhl[up + sj];
    const double rC = rho[at], rE = rho[at + 1], rN = rho[at + sj];

    const double gcC = gkC1 - gkC, gcE = gkE1 - gkE, gcN = gkN1 - gkN;
    const double ppgu = (ppE - ppC) + (gcE + gcC) * 0.5 * ((hC1 + hC) - (hE1 + hE)) / ((hC1 - hC) + (hE1 - hE));
    const double ppgv = (ppN - ppC) + (gcN + gcC) * 0.5 * ((hC1 + hC) - (hN1 + hN)) / ((hC1 - hC) + (hN1 - hN));
    st_nt(uout + at, ld_nt(uin + at) + dt * (ld_nt(utens + at) - ppgu * (2.0 * fxj / (rE + rC))));
    st_nt(vout + at, ld_nt(vin + at) + dt * (ld_nt(vtens + at) - ppgv * (2.0 * edadlat / (rN + rC))));

    ppC = ppC1;
    ppE = ppE1;
    ppN = ppN1;
    gkC = gkC1;
    gkE = gkE1;
    gkN = gkN1;
    hC = hC1;
    hE = hE1;
    hN = hN1;
  }
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_fastwavesuv_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                   const oeb_s3 *vin, const oeb_s3 *utens, const oeb_s3 *vtens, const oeb_s3 *wgtfac,
                                   const oeb_s3 *ppuv, const oeb_s3 *hhl, const oeb_s3 *rho, const oeb_s1 *fx,
                                   const oeb_s3 *ppgk, const oeb_s3 *ppgc, const oeb_s3 *ppgu, const oeb_s3 *ppgv,
                                   double edadlat, double dt, void *stream) {
  using namespace oeb;
  (void)ppgk;
  (void)ppgc;
  (void)ppgu;
  (void)ppgv;
  const char *K = "fastwavesuv";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_h{0, nx + 1, 0, ny + 1, 0, nz};
  const Box r_w{0, nx + 1, 0, ny + 1, 0, nz + 1}, r_p{0, nx + 1, 0, ny + 1, -1, nz + 1};
  const double *p_uo, *p_vo, *p_ui, *p_vi, *p_ut, *p_vt, *p_wg, *p_pp, *p_hh, *p_rh, *p_fx;
  int fx_stride = 1;
  OEB_TRY(check3(K, "uout", uout, r0, g, true, &p_uo));
  OEB_TRY(check3(K, "vout", vout, r0, g, false, &p_vo));
  OEB_TRY(check3(K, "uin", uin, r0, g, false, &p_ui));
  OEB_TRY(check3(K, "vin", vin, r0, g, false, &p_vi));
  OEB_TRY(check3(K, "utens", utens, r0, g, false, &p_ut));
  OEB_TRY(check3(K, "vtens", vtens, r0, g, false, &p_vt));
  OEB_TRY(check3(K, "wgtfac", wgtfac, r_w, g, false, &p_wg));
  OEB_TRY(check3(K, "ppuv", ppuv, r_p, g, false, &p_pp));
  OEB_TRY(check3(K, "hhl", hhl, r_w, g, false, &p_hh));
  OEB_TRY(check3(K, "rho", rho, r_h, g, false, &p_rh));
  OEB_TRY(check1(K, "fx", fx, 0, ny, &p_fx, &fx_stride));
  AliasSet as;
  as.add("uout", p_uo, r0, g, true);
  as.add("vout", p_vo, r0, g, true);
  as.add("uin", p_ui, r0, g, false);
  as.add("vin", p_vi, r0, g, false);
  as.add("utens", p_ut, r0, g, false);
  as.add("vtens", p_vt, r0, g, false);
  as.add("wgtfac", p_wg, r_w, g, false);
  as.add("ppuv", p_pp, r_p, g, false);
  as.add("hhl", p_hh, r_w, g, false);
  as.add("rho", p_rh, r_h, g, false);
  as.add1("fx", p_fx, 0, ny, fx_stride);
  OEB_TRY(as.check(K));
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kChunk);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  hipLaunchKernelGGL(fastwavesuv_kernel, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream),
                     const_cast<double *>(p_uo), const_cast<double *>(p_vo), p_ui, p_vi, p_ut, p_vt, p_wg, p_pp, p_hh,
                     p_rh, p_fx, fx_stride, edadlat, dt, g, t);
  return launch_status(K);
}
