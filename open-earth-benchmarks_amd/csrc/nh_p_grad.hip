// nh_p_grad.hip -- FV3 non-hydrostatic pressure-gradient update of (u, v).
// Reference: kernels/nh_p_grad.h:116-146 (nh_p_grad_openmp), expressions
// :123-142.  wk = pk3(k+1) - pk3(k) is formed on the fly at the centre, east
// and north points.  Reach: i, j in [0, +1], k in [0, +1].  Each lane marches
// up its column and carries the level k+1 values of gz, pk3 and pp over to the
// next step.
// Algorithmic traffic: 8 inputs + 2 outputs = 80 B per point (FP64).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 16;

__global__ void __launch_bounds__(kTileI *kTileJ)
    nh_p_grad_kernel(double *__restrict__ uout, double *__restrict__ vout, const double *__restrict__ uin,
                     const double *__restrict__ vin, const double *__restrict__ rdx, const double *__restrict__ rdy,
                     const double *__restrict__ gz, const double *__restrict__ pp, const double *__restrict__ pk3,
                     const double *__restrict__ wk1, double dt, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  // Level k at the centre C, east E = (i+1) and north N = (j+1).
  double gzC = gz[at], gzE = gz[at + 1], gzN = gz[at + sj];
  double pkC = pk3[at], pkE = pk3[at + 1], pkN = pk3[at + sj];
  double ppC = pp[at], ppE = pp[at + 1], ppN = pp[at + sj];
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const double gzC1 = gz[up], gzE1 = gz[up + 1], gzN1 = gz[up + sj];
    const double pkC1 = pk3[up], pkE1 = pk3[up + 1], pkN1 = pk3[up + sj];
    const double ppC1 = pp[up], ppE1 = pp[up + 1], ppN1 = pp[up + sj];
    const double w1C = wk1[at], w1E = wk1[at + 1], w1N = wk1[at + sj];
    const double wk0 = (pkC1 - pkC), wkE = (pkE1 - pkE), wkN = (pkN1 - pkN);

    const double du = ((dt / (wk0 + wkE)) * (((gzC1 - gzE) * (pkE1 - pkC)) + ((gzC - gzE1) * (pkC1 - pkE))));
    st_nt(uout + at, (((ld_nt(uin + at) + du) +
                       ((dt / (w1C + w1E)) * (((gzC1 - gzE) * (ppE1 - ppC)) + ((gzC - gzE1) * (ppC1 - ppE))))) *
                      ld_nt(rdx + at)));
    const double dv = ((dt / (wk0 + wkN)) * (((gzC1 - gzN) * (pkN1 - pkC)) + ((gzC - gzN1) * (pkC1 - pkN))));
    st_nt(vout + at, (((ld_nt(vin + at) + dv) +
                       ((dt / (w1C + w1N)) * (((gzC1 - gzN) * (ppN1 - ppC)) + ((gzC - gzN1) * (ppC1 - ppN))))) *
                      ld_nt(rdy + at)));

    gzC = gzC1;
    gzE = gzE1;
    gzN = gzN1;
    pkC = pkC1;
    pkE = pkE1;
    pkN = pkN1;
    ppC = ppC1;
    ppE = ppE1;
    ppN = ppN1;
  }
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_nh_p_grad_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                 const oeb_s3 *vin, const oeb_s3 *rdx, const oeb_s3 *rdy, const oeb_s3 *gz,
                                 const oeb_s3 *pp, const oeb_s3 *pk3, const oeb_s3 *wk1, const oeb_s3 *wk,
                                 const oeb_s3 *du, const oeb_s3 *dv, double dt, void *stream) {
  using namespace oeb;
  (void)wk;
  (void)du;
  (void)dv;
  const char *K = "nh_p_grad";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_h{0, nx + 1, 0, ny + 1, 0, nz}, r_k{0, nx + 1, 0, ny + 1, 0, nz + 1};
  const double *p_uo, *p_vo, *p_ui, *p_vi, *p_rx, *p_ry, *p_gz, *p_pp, *p_pk, *p_w1;
  OEB_TRY(check3(K, "uout", uout, r0, g, true, &p_uo));
  OEB_TRY(check3(K, "vout", vout, r0, g, false, &p_vo));
  OEB_TRY(check3(K, "uin", uin, r0, g, false, &p_ui));
  OEB_TRY(check3(K, "vin", vin, r0, g, false, &p_vi));
  OEB_TRY(check3(K, "rdx", rdx, r0, g, false, &p_rx));
  OEB_TRY(check3(K, "rdy", rdy, r0, g, false, &p_ry));
  OEB_TRY(check3(K, "gz", gz, r_k, g, false, &p_gz));
  OEB_TRY(check3(K, "pp", pp, r_k, g, false, &p_pp));
  OEB_TRY(check3(K, "pk3", pk3, r_k, g, false, &p_pk));
  OEB_TRY(check3(K, "wk1", wk1, r_h, g, false, &p_w1));
  AliasSet as;
  as.add("uout", p_uo, r0, g, true);
  as.add("vout", p_vo, r0, g, true);
  as.add("uin", p_ui, r0, g, false);
  as.add("vin", p_vi, r0, g, false);
  as.add("rdx", p_rx, r0, g, false);
  as.add("rdy", p_ry, r0, g, false);
  as.add("gz", p_gz, r_k, g, false);
  as.add("pp", p_pp, r_k, g, false);
  as.add("pk3", p_pk, r_k, g, false);
  as.add("wk1", p_w1, r_h, g, false);
  OEB_TRY(as.check(K));
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kChunk);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  hipLaunchKernelGGL(nh_p_grad_kernel, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream),
                     const_cast<double *>(p_uo), const_cast<double *>(p_vo), p_ui, p_vi, p_rx, p_ry, p_gz, p_pp, p_pk,
                     p_w1, dt, g, t);
  return launch_status(K);
}
