// hadvuv.hip -- horizontal upwind advection of (u, v) with metric terms, in
// two orders of accuracy that share one body:
//   hadvuv     kernels/hadvuv.h:178-210 (hadvuv_openmp), 3rd-order upwind
//              stencil advectionDriver kernels/hadvuv.h:4-27, reach (2,2,0);
//   hadvuv5th  kernels/hadvuv5th.h:177-201 (hadvuv5th_openmp), 5th-order
//              upwind stencil advectionDriver kernels/hadvuv5th.h:4-39, reach (3,3,0).
// The reference has both helpers under one name; here they are upwind3 and
// upwind5.  The branch on the sign of the averaged velocity (strict > 0, zero
// takes the second form) is kept.
// Algorithmic traffic: 2 inputs + 2 outputs = 32 B per point (FP64).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kChunk = 1;
constexpr double kEarthRadiusRecip = 1.0 / 6371.229e3;

// 3rd-order upwind along the axis of stride s (kernels/hadvuv.h:10-24).
__device__ __forceinline__ double upwind3_axis(const double *__restrict__ f, int s, double a) {
  if (a > 0) {
    return a * (-1.0 / 6.0 * f[-2 * s] + f[-s] + -0.5 * f[0] + -1.0 / 3.0 * f[s]);
  }
  return -a * (-1.0 / 3.0 * f[-s] + -0.5 * f[0] + f[s] + -1.0 / 6.0 * f[2 * s]);
}

// 5th-order upwind along the axis of stride s (kernels/hadvuv5th.h:10-36).
__device__ __forceinline__ double upwind5_axis(const double *__restrict__ f, int s, double a) {
  if (a > 0) {
    return a * (1.0 / 30.0 * f[-3 * s] + -1.0 / 4.0 * f[-2 * s] + f[-s] + -1.0 / 3.0 * f[0] + -1.0 / 2.0 * f[s] +
                1.0 / 20.0 * f[2 * s]);
  }
  return -a * (1.0 / 20.0 * f[-2 * s] + -1.0 / 2.0 * f[-s] + -1.0 / 3.0 * f[0] + f[s] + -1.0 / 4.0 * f[2 * s] +
               1.0 / 30.0 * f[3 * s]);
}

template <int ORDER>
__device__ __forceinline__ double upwind(const double *__restrict__ f, int sj, double ua, double va, double eddlat,
                                         double eddlon) {
  double rx, ry;
  if (ORDER == 3) {
    rx = upwind3_axis(f, 1, ua);
    ry = upwind3_axis(f, sj, va);
  } else {
    rx = upwind5_axis(f, 1, ua);
    ry = upwind5_axis(f, sj, va);
  }
  return eddlat * rx + eddlon * ry;
}

// One level of a 64 x 16 tile per workgroup.  u and v are staged in LDS with
// the stencil's halo (R = 2 or 3), so each value is read from memory once per
// tile.  The 14 neighbour reads of each field per point then come from LDS:
// read through L1 they made this stencil TA-bound, and the LDS version
// measured 1.67x faster at 256x256x80 and 512x512x80.  256 lanes; each lane
// computes 4 rows of one column.
constexpr int kTI = 64, kTJ = 16, kNT = 256;

template <int ORDER>
__global__ void __launch_bounds__(kNT)
    hadvuv_kernel(double *__restrict__ uout, double *__restrict__ vout, const double *__restrict__ uin,
                  const double *__restrict__ vin, const double *__restrict__ acrlat0, int s0,
                  const double *__restrict__ acrlat1, int s1, const double *__restrict__ tgrlatda0, int s2,
                  const double *__restrict__ tgrlatda1, int s3, double eddlat, double eddlon, Geo g, Tiles t) {
  constexpr int R = ORDER == 3 ? 2 : 3;
  constexpr int W = kTI + 2 * R;
  __shared__ double su[kTJ + 2 * R][W];
  __shared__ double sv[kTJ + 2 * R][W];
  const TileCoord c = tile_coord(t, kTI, kTJ, g.nz);
  const int tid = int(threadIdx.x) + kTI * int(threadIdx.y);
  const int sj = g.sj;
  for (int k = c.k0; k < c.k1; ++k) {
    const int64_t plane = int64_t(k) * g.sk;
    stage_tile<kTI, kTJ, R, kNT>(su, uin + plane, c.i0, c.j0, sj, g.nx + R, g.ny + R, tid);
    stage_tile<kTI, kTJ, R, kNT>(sv, vin + plane, c.i0, c.j0, sj, g.nx + R, g.ny + R, tid);
    __syncthreads();
    const int i = c.i0 + int(threadIdx.x);
#pragma unroll
    for (int t4 = 0; t4 < kTJ / 4; ++t4) {
      const int lj = int(threadIdx.y) * (kTJ / 4) + t4;
      const int j = c.j0 + lj;
      if (i < g.nx && j < g.ny) {
        const double *u = &su[lj + R][threadIdx.x + R];
        const double *v = &sv[lj + R][threadIdx.x + R];
        const double a0 = acrlat0[int64_t(j) * s0], a1 = acrlat1[int64_t(j) * s1];
        const double tg0 = tgrlatda0[int64_t(j) * s2], tg1 = tgrlatda1[int64_t(j) * s3];
        const int64_t at = plane + i + int64_t(j) * sj;
        // u component: kernels/hadvuv.h:186-195.
        const double uatupos = (1.0 / 3.0) * (u[-1] + u[0] + u[1]);
        const double vatupos = 0.25 * (v[1] + v[1 - W] + v[0] + v[-W]);
        const double ures = upwind<ORDER>(u, W, a0 * uatupos, kEarthRadiusRecip * vatupos, eddlat, eddlon);
        st_nt(uout + at, ures + tg0 * u[0] * vatupos);
        // v component: kernels/hadvuv.h:197-206.
        const double uatvpos = 0.25 * (u[-1] + u[0] + u[W] + u[W - 1]);
        const double vatvpos = 1.0 / 3.0 * (v[-W] + v[0] + v[W]);
        const double vres = upwind<ORDER>(v, W, a1 * uatvpos, kEarthRadiusRecip * vatvpos, eddlat, eddlon);
        st_nt(vout + at, vres - tg1 * uatvpos * uatvpos);
      }
    }
    if (k + 1 < c.k1) __syncthreads();
  }
}

template <int ORDER>
int hadv_entry(const char *K, const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
               const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0,
               const oeb_s1 *tgrlatda1, double eddlat, double eddlon, void *stream) {
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz, h = ORDER == 3 ? 2 : 3;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_in{-h, nx + h, -h, ny + h, 0, nz};
  const double *p_uo, *p_vo, *p_ui, *p_vi, *p_a0, *p_a1, *p_t0, *p_t1;
  int s0 = 1, s1 = 1, s2 = 1, s3 = 1;
  OEB_TRY(check3(K, "uout", uout, r0, g, true, &p_uo));
  OEB_TRY(check3(K, "vout", vout, r0, g, false, &p_vo));
  OEB_TRY(check3(K, "uin", uin, r_in, g, false, &p_ui));
  OEB_TRY(check3(K, "vin", vin, r_in, g, false, &p_vi));
  OEB_TRY(check1(K, "acrlat0", acrlat0, 0, ny, &p_a0, &s0));
  OEB_TRY(check1(K, "acrlat1", acrlat1, 0, ny, &p_a1, &s1));
  OEB_TRY(check1(K, "tgrlatda0", tgrlatda0, 0, ny, &p_t0, &s2));
  OEB_TRY(check1(K, "tgrlatda1", tgrlatda1, 0, ny, &p_t1, &s3));
  AliasSet as;
  as.add("uout", p_uo, r0, g, true);
  as.add("vout", p_vo, r0, g, true);
  as.add("uin", p_ui, r_in, g, false);
  as.add("vin", p_vi, r_in, g, false);
  as.add1("acrlat0", p_a0, 0, ny, s0);
  as.add1("acrlat1", p_a1, 0, ny, s1);
  as.add1("tgrlatda0", p_t0, 0, ny, s2);
  as.add1("tgrlatda1", p_t1, 0, ny, s3);
  OEB_TRY(as.check(K));
  const Tiles t = make_tiles(nx, ny, int(nz), kTI, kTJ, kchunk_for(K, kChunk));
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTI, kTJ));
  hipLaunchKernelGGL(hadvuv_kernel<ORDER>, dim3(t.n), dim3(kTI, kNT / kTI), 0, hipStream_t(stream),
                     const_cast<double *>(p_uo), const_cast<double *>(p_vo), p_ui, p_vi, p_a0, s0, p_a1, s1, p_t0, s2,
                     p_t1, s3, eddlat, eddlon, g, t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_hadvuv_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                              const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1,
                              const oeb_s1 *tgrlatda0, const oeb_s1 *tgrlatda1, const oeb_s3 *uatupos,
                              const oeb_s3 *vatupos, const oeb_s3 *uatvpos, const oeb_s3 *vatvpos, const oeb_s3 *uavg,
                              const oeb_s3 *vavg, const oeb_s3 *ures, const oeb_s3 *vres, double eddlat,
                              double eddlon, void *stream) {
  (void)uatupos, (void)vatupos, (void)uatvpos, (void)vatvpos, (void)uavg, (void)vavg, (void)ures, (void)vres;
  return oeb::hadv_entry<3>("hadvuv", d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat,
                            eddlon, stream);
}

extern "C" int oeb_hadvuv5th_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                 const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1,
                                 const oeb_s1 *tgrlatda0, const oeb_s1 *tgrlatda1, const oeb_s3 *uatupos,
                                 const oeb_s3 *vatupos, const oeb_s3 *uatvpos, const oeb_s3 *vatvpos,
                                 const oeb_s3 *uavg, const oeb_s3 *vavg, const oeb_s3 *ures, const oeb_s3 *vres,
                                 double eddlat, double eddlon, void *stream) {
  (void)uatupos, (void)vatupos, (void)uatvpos, (void)vatvpos, (void)uavg, (void)vavg, (void)ures, (void)vres;
  return oeb::hadv_entry<5>("hadvuv5th", d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat,
                            eddlon, stream);
}
