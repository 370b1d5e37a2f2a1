// p_grad_c.hip -- FV3 C-grid pressure-gradient update of (u, v).
// Reference: kernels/p_grad_c.h:80-101 (p_grad_c_openmp), expressions :87-99.
// delpc(i-1) and delpc(j-1) are read directly, as the OpenMP variant does.
// The DEFAULT variant instead reads an uninitialised copy at i = 0 and j = 0.
// Reach: i, j in [-1, 0], k in [0, +1].  Each lane marches up its column and
// carries the gz / pkc values of level k+1 over as level k of the next step,
// so every plane of gz and pkc is loaded once per chunk.
// Algorithmic traffic: 7 inputs + 2 outputs = 72 B per point (FP64).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 16;

__global__ void __launch_bounds__(kTileI *kTileJ)
    p_grad_c_kernel(double *__restrict__ uout, double *__restrict__ vout, const double *__restrict__ uin,
                    const double *__restrict__ vin, const double *__restrict__ rdxc, const double *__restrict__ rdyc,
                    const double *__restrict__ delpc, const double *__restrict__ gz, const double *__restrict__ pkc,
                    double dt2, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  // Level k at the centre C, west W = (i-1) and south S = (j-1).
  double gzC = gz[at], gzW = gz[at - 1], gzS = gz[at - sj];
  double pkC = pkc[at], pkW = pkc[at - 1], pkS = pkc[at - sj];
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const double gzC1 = gz[up], gzW1 = gz[up - 1], gzS1 = gz[up - sj];
    const double pkC1 = pkc[up], pkW1 = pkc[up - 1], pkS1 = pkc[up - sj];
    const double wk0 = delpc[at], wkW = delpc[at - 1], wkS = delpc[at - sj];
    st_nt(uout + at, (ld_nt(uin + at) + (((dt2 * ld_nt(rdxc + at)) / (wkW + wk0)) *
                                         (((gzW1 - gzC) * (pkC1 - pkW)) + ((gzW - gzC1) * (pkW1 - pkC))))));
    st_nt(vout + at, (ld_nt(vin + at) + (((dt2 * ld_nt(rdyc + at)) / (wkS + wk0)) *
                                         (((gzS1 - gzC) * (pkC1 - pkS)) + ((gzS - gzC1) * (pkS1 - pkC))))));
    gzC = gzC1;
    gzW = gzW1;
    gzS = gzS1;
    pkC = pkC1;
    pkW = pkW1;
    pkS = pkS1;
  }
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_p_grad_c_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                const oeb_s3 *vin, const oeb_s3 *rdxc, const oeb_s3 *rdyc, const oeb_s3 *delpc,
                                const oeb_s3 *gz, const oeb_s3 *pkc, const oeb_s3 *wk, double dt2, void *stream) {
  using namespace oeb;
  (void)wk;
  const char *K = "p_grad_c";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_dp{-1, nx, -1, ny, 0, nz}, r_k{-1, nx, -1, ny, 0, nz + 1};
  const double *p_uo, *p_vo, *p_ui, *p_vi, *p_rx, *p_ry, *p_dp, *p_gz, *p_pk;
  OEB_TRY(check3(K, "uout", uout, r0, g, true, &p_uo));
  OEB_TRY(check3(K, "vout", vout, r0, g, false, &p_vo));
  OEB_TRY(check3(K, "uin", uin, r0, g, false, &p_ui));
  OEB_TRY(check3(K, "vin", vin, r0, g, false, &p_vi));
  OEB_TRY(check3(K, "rdxc", rdxc, r0, g, false, &p_rx));
  OEB_TRY(check3(K, "rdyc", rdyc, r0, g, false, &p_ry));
  OEB_TRY(check3(K, "delpc", delpc, r_dp, g, false, &p_dp));
  OEB_TRY(check3(K, "gz", gz, r_k, g, false, &p_gz));
  OEB_TRY(check3(K, "pkc", pkc, r_k, g, false, &p_pk));
  AliasSet as;
  as.add("uout", p_uo, r0, g, true);
  as.add("vout", p_vo, r0, g, true);
  as.add("uin", p_ui, r0, g, false);
  as.add("vin", p_vi, r0, g, false);
  as.add("rdxc", p_rx, r0, g, false);
  as.add("rdyc", p_ry, r0, g, false);
  as.add("delpc", p_dp, r_dp, g, false);
  as.add("gz", p_gz, r_k, g, false);
  as.add("pkc", p_pk, r_k, g, false);
  OEB_TRY(as.check(K));
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kChunk);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  hipLaunchKernelGGL(p_grad_c_kernel, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream),
                     const_cast<double *>(p_uo), const_cast<double *>(p_vo), p_ui, p_vi, p_rx, p_ry, p_dp, p_gz, p_pk,
                     dt2, g, t);
  return launch_status(K);
}
