// hadvuv.hip -- horizontal upwind advection of (u, v) with metric terms, in
// two orders of accuracy that share one body:
//   hadvuv     kernels/hadvuv.h:178-210 (hadvuv_openmp), 3rd-order upwind
//              stencil advectionDriver kernels/hadvuv.h:4-27, reach (2,2,0);
//   hadvuv5th  kernels/hadvuv5th.h:177-201 (hadvuv5th_openmp), 5th-order
//              upwind stencil advectionDriver kernels/hadvuv5th.h:4-39, reach (3,3,0).
// The reference has both helpers under one name; here they are upwind3 and
// upwind5.  The branch on the sign of the averaged velocity (strict > 0, zero
// takes the second form) is kept.
// Algorithmic traffic: 2 inputs + 2 outputs = 32 B per point (FP64), 16 B (FP32).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kChunk = 4;

// EARTH_RADIUS_RECIP, include/util.h:17-18, formed in the element type.
template <class T>
constexpr T kEarthRadiusRecip = T(1.0) / T(6371.229e3);

// 3rd-order upwind along the axis of stride s (kernels/hadvuv.h:10-24).
template <class T>
__device__ __forceinline__ T upwind3_axis(const T *f, int s, T a) {
  if (a > 0) {
    return a * (T(-1.0) / T(6.0) * f[-2 * s] + f[-s] + T(-0.5) * f[0] + T(-1.0) / T(3.0) * f[s]);
  }
  return -a * (T(-1.0) / T(3.0) * f[-s] + T(-0.5) * f[0] + f[s] + T(-1.0) / T(6.0) * f[2 * s]);
}

// 5th-order upwind along the axis of stride s (kernels/hadvuv5th.h:10-36).
// FP32 forms both branches and selects one (each is the reference's
// expression, so the result has its bits).  On winds whose sign changes
// within a wave, against the branches on one box: 92.7-93.9 / 29.4-29.7 us
// against 95.7-96.1 / 30.6-30.9 at 512x512x80 / 256x256x80.  (The same for
// the 3rd order measured 78.9-79.1 / 25.3-25.5 against 80.1-80.8 /
// 24.5-25.1, and it keeps its branches.)  FP64 branches, which keeps its
// 1024-lane tile at 62 VGPRs (two workgroups per CU) instead of 75; selecting,
// FP64 hadvuv5th took 187.7-188.5 us with 1024 lanes and 153.7 with 512
// against 149.5 at 512x512x80.
template <class T>
__device__ __forceinline__ T upwind5_axis(const T *f, int s, T a) {
  if constexpr (sizeof(T) == 4) {
    const T up = a * (T(1.0) / T(30.0) * f[-3 * s] + T(-1.0) / T(4.0) * f[-2 * s] + f[-s] +
                      T(-1.0) / T(3.0) * f[0] + T(-1.0) / T(2.0) * f[s] + T(1.0) / T(20.0) * f[2 * s]);
    const T dn = -a * (T(1.0) / T(20.0) * f[-2 * s] + T(-1.0) / T(2.0) * f[-s] + T(-1.0) / T(3.0) * f[0] + f[s] +
                       T(-1.0) / T(4.0) * f[2 * s] + T(1.0) / T(30.0) * f[3 * s]);
    return a > 0 ? up : dn;
  }
  if (a > 0) {
    return a * (T(1.0) / T(30.0) * f[-3 * s] + T(-1.0) / T(4.0) * f[-2 * s] + f[-s] + T(-1.0) / T(3.0) * f[0] +
                T(-1.0) / T(2.0) * f[s] + T(1.0) / T(20.0) * f[2 * s]);
  }
  return -a * (T(1.0) / T(20.0) * f[-2 * s] + T(-1.0) / T(2.0) * f[-s] + T(-1.0) / T(3.0) * f[0] + f[s] +
               T(-1.0) / T(4.0) * f[2 * s] + T(1.0) / T(30.0) * f[3 * s]);
}

template <int ORDER, class T>
__device__ __forceinline__ T upwind(const T *f, int sj, T ua, T va, T eddlat, T eddlon) {
  T rx, ry;
  if (ORDER == 3) {
    rx = upwind3_axis(f, 1, ua);
    ry = upwind3_axis(f, sj, va);
  } else {
    rx = upwind5_axis(f, 1, ua);
    ry = upwind5_axis(f, sj, va);
  }
  return eddlat * rx + eddlon * ry;
}

// Tiles of 64 x 16 columns, 256 lanes; each lane computes 4 rows of one
// column.  u and v are staged in LDS with the stencil's halo (R = 2 or 3), so
// each value is read from memory once per tile and the 14 neighbour reads of
// each field per point come from LDS.  Read through L1, those reads made the
// stencil TA-bound; LDS measured 1.67x faster.  The workgroup walks a chunk
// of levels with two LDS buffers: while it computes level k from one, the
// loads of level k+1 are already in flight into registers, and are written
// to the other buffer after the compute.  The loads' latency then overlaps the
// arithmetic.  Against one buffer per level this measured 1.2-1.25x faster at
// 256x256x80 and 512x512x80 (5th order).
constexpr int kTI = 64, kTJ = 16, kNT = 256;

template <class T>
struct HadvArgs {
  T *uout, *vout;
  const T *uin, *vin, *acrlat0, *acrlat1, *tgrlatda0, *tgrlatda1;
  int s0, s1, s2, s3;
  T eddlat, eddlon;
};

// Logical block `id` of a launch.  Ends with a barrier, so a persistent
// workgroup may run it again at once (batched launches).
template <int ORDER, class T, int kNT = kNT>
__device__ __forceinline__ void hadvuv_body(const HadvArgs<T> &a, const Geo &g, const Tiles &t, int id) {
  constexpr int kRows = kTJ / (kNT / kTI);  // rows per lane
  constexpr int R = ORDER == 3 ? 2 : 3;
  constexpr int W = kTI + 2 * R, H = kTJ + 2 * R, kPer = (W * H + kNT - 1) / kNT;
  constexpr T rr = kEarthRadiusRecip<T>;
  __shared__ T su[2][H][W];
  __shared__ T sv[2][H][W];
  T *__restrict__ uout = a.uout;
  T *__restrict__ vout = a.vout;
  const T *__restrict__ uin = a.uin;
  const T *__restrict__ vin = a.vin;
  const T eddlat = a.eddlat, eddlon = a.eddlon;
  const TileCoord c = tile_coord_of(t, id, kTI, kTJ, g.nz);
  const int tid = int(threadIdx.x) + kTI * int(threadIdx.y);
  const int sj = g.sj;
  const int i = c.i0 + int(threadIdx.x);

  // Metric coefficients of this lane's rows (the clamp keeps rows past ny,
  // which compute nothing, on a valid index).
  T a0[kRows], a1[kRows], tg0[kRows], tg1[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int64_t j = min(c.j0 + int(threadIdx.y) * kRows + r, g.ny - 1);
    a0[r] = a.acrlat0[j * a.s0], a1[r] = a.acrlat1[j * a.s1];
    tg0[r] = a.tgrlatda0[j * a.s2], tg1[r] = a.tgrlatda1[j * a.s3];
  }
  // This lane's share of the staged tile: element e = tid + q*kNT of the
  // H x W tile, at in-plane offset off[q].  Elements past the field's valid
  // box (i >= nx + R or j >= ny + R) are set to 0 and never used.
  int off[kPer];
  bool ok[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int e = tid + q * kNT, r = e / W, col = e % W;
    const int gi = c.i0 - R + col, gj = c.j0 - R + r;
    ok[q] = e < W * H && gi < g.nx + R && gj < g.ny + R;
    off[q] = ok[q] ? gi + gj * sj : 0;
  }
  T ru[kPer], rv[kPer];
  auto load = [&](int64_t plane) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      ru[q] = ok[q] ? uin[plane + off[q]] : T(0);
      rv[q] = ok[q] ? vin[plane + off[q]] : T(0);
    }
  };
  auto store = [&](int b) {
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = tid + q * kNT;
      if (e < W * H) {
        su[b][e / W][e % W] = ru[q];
        sv[b][e / W][e % W] = rv[q];
      }
    }
  };

  load(int64_t(c.k0) * g.sk);
  store(0);
  lds_barrier();
  int cur = 0;
  for (int k = c.k0; k < c.k1; ++k) {
    const int64_t plane = int64_t(k) * g.sk;
    const bool more = k + 1 < c.k1;
    if (more) load(plane + g.sk);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int lj = int(threadIdx.y) * kRows + r;
      const int j = c.j0 + lj;
      if (i < g.nx && j < g.ny) {
        const T *u = &su[cur][lj + R][threadIdx.x + R];
        const T *v = &sv[cur][lj + R][threadIdx.x + R];
        const int64_t at = plane + i + int64_t(j) * sj;
        // u component: kernels/hadvuv.h:186-195.
        const T uatupos = (T(1.0) / T(3.0)) * (u[-1] + u[0] + u[1]);
        const T vatupos = T(0.25) * (v[1] + v[1 - W] + v[0] + v[-W]);
        const T ures = upwind<ORDER>(u, W, a0[r] * uatupos, rr * vatupos, eddlat, eddlon);
        st_nt(uout + at, ures + tg0[r] * u[0] * vatupos);
        // v component: kernels/hadvuv.h:197-206.
        const T uatvpos = T(0.25) * (u[-1] + u[0] + u[W] + u[W - 1]);
        const T vatvpos = T(1.0) / T(3.0) * (v[-W] + v[0] + v[W]);
        const T vres = upwind<ORDER>(v, W, a1[r] * uatvpos, rr * vatvpos, eddlat, eddlon);
        st_nt(vout + at, vres - tg1[r] * uatvpos * uatvpos);
      }
    }
    // Buffer cur^1 was last read before the previous barrier, so it can be
    // overwritten now; the barrier below publishes it for the next level.
    if (more) store(cur ^ 1);
    lds_barrier();
    cur ^= 1;
  }
}

// Lanes per 64x16 tile, measured in the ten-stencil step at 256x256x80 and
// 512x512x80 over 256, 512 and 1024: FP64 is fastest with 1024 (one row per
// lane; 1.06-1.18x faster than 256), FP32 with 512 (1.05-1.2x faster).  More
// waves per CU hide the latency of each level's staging.
template <class T>
constexpr int kHadvLanes = sizeof(T) == 8 ? 1024 : 512;

template <int ORDER, class T, int NT>
__global__ void __launch_bounds__(NT) hadvuv_kernel(HadvArgs<T> a, Geo g, Tiles t) {
  hadvuv_body<ORDER, T, NT>(a, g, t, int(blockIdx.x));
}

template <int ORDER, class T>
hipError_t hadvuv_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  constexpr int NT = kHadvLanes<T>;
  return launch_batch<HadvArgs<T>, hadvuv_body<ORDER, T, NT>, NT>(tab, n, total, s, dim3(kTI, NT / kTI));
}
template <int ORDER, class T>
const BatchKind kHadvuvBatch{ORDER == 3 ? "hadvuv" : "hadvuv5th", hadvuv_batch<ORDER, T>};

template <int ORDER, class T, class S3, class S1>
int hadv_run(const char *K, const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
             const S1 *acrlat0, const S1 *acrlat1, const S1 *tgrlatda0, const S1 *tgrlatda1, T eddlat, T eddlon,
             void *stream) {
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz, h = ORDER == 3 ? 2 : 3;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_in{-h, nx + h, -h, ny + h, 0, nz};
  const T *p_uo, *p_vo, *p_ui, *p_vi, *p_a0, *p_a1, *p_t0, *p_t1;
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
  // Levels per workgroup: 4, and 5 for FP32 hadvuv, which measured 1.12x
  // faster with 5 at 256x256x80 (the same at 512x512x80).  FP64 (1024-lane
  // workgroups, two per CU) takes the largest chunk of 4..10 levels whose
  // workgroups fill whole rounds of the CUs (balanced_kchunk): 10 at
  // 256x256x80 (64 tiles a level: 512 workgroups), where 4 left the third
  // round half empty (1280 workgroups) and 5 or 10 measured 1.04-1.05x
  // faster (hadvuv5th 45.7-45.9 us against 47.8-50.6, hadvuv 41.2-41.9 against
  // 42.9-44.5); 10 at 512x512x80 (157 / 136 us against 159 / 137 for 8).
  const int64_t level_tiles = ((nx + kTI - 1) / kTI) * ((ny + kTJ - 1) / kTJ);
  const int kc_dflt = sizeof(T) == 4 ? (ORDER == 3 ? 5 : kChunk)
                                     : balanced_kchunk(level_tiles, int(nz), 2, 4, 10,
                                                       level_tiles >= 256 ? 8 : kChunk);
  // The tiles of a row on one XCD (Tiles::irow): FP64 hadvuv 141 us against
  // 145 at 512x512x80, hadvuv5th 1% faster, the same or better at 256x256x80.
  const int order = knob(K, sizeof(T), d, KNOB_ORDER, 2);
  const Tiles t = make_tiles(nx, ny, int(nz), kTI, kTJ, knob(K, sizeof(T), d, KNOB_KCHUNK, kc_dflt), order);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTI, kTJ));
  const HadvArgs<T> a{const_cast<T *>(p_uo), const_cast<T *>(p_vo), p_ui, p_vi, p_a0, p_a1, p_t0, p_t1,
                      s0, s1, s2, s3, eddlat, eddlon};
  if (batch_record(kHadvuvBatch<ORDER, T>, a, g, t, t.n, as)) return OEB_OK;
  // Lanes per 64x16 tile: 256 (4 rows each), 512 (2) or 1024 (1).
  const int nt = knob(K, sizeof(T), d, KNOB_LANES, kHadvLanes<T>);
  if (nt >= 1024)
    hipLaunchKernelGGL((hadvuv_kernel<ORDER, T, 1024>), dim3(t.n), dim3(kTI, 16), 0, hipStream_t(stream), a, g, t);
  else if (nt >= 512)
    hipLaunchKernelGGL((hadvuv_kernel<ORDER, T, 512>), dim3(t.n), dim3(kTI, 8), 0, hipStream_t(stream), a, g, t);
  else
    hipLaunchKernelGGL((hadvuv_kernel<ORDER, T, 256>), dim3(t.n), dim3(kTI, 4), 0, hipStream_t(stream), a, g, t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

// uatupos .. vres are the reference's scratch arrays; nothing is stored in them.
extern "C" int oeb_hadvuv_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                              const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1,
                              const oeb_s1 *tgrlatda0, const oeb_s1 *tgrlatda1, const oeb_s3 *uatupos,
                              const oeb_s3 *vatupos, const oeb_s3 *uatvpos, const oeb_s3 *vatvpos, const oeb_s3 *uavg,
                              const oeb_s3 *vavg, const oeb_s3 *ures, const oeb_s3 *vres, double eddlat,
                              double eddlon, void *stream) {
  (void)uatupos, (void)vatupos, (void)uatvpos, (void)vatvpos, (void)uavg, (void)vavg, (void)ures, (void)vres;
  return oeb::hadv_run<3>("hadvuv", d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat, eddlon,
                          stream);
}

extern "C" int oeb_hadvuv5th_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                 const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1,
                                 const oeb_s1 *tgrlatda0, const oeb_s1 *tgrlatda1, const oeb_s3 *uatupos,
                                 const oeb_s3 *vatupos, const oeb_s3 *uatvpos, const oeb_s3 *vatvpos,
                                 const oeb_s3 *uavg, const oeb_s3 *vavg, const oeb_s3 *ures, const oeb_s3 *vres,
                                 double eddlat, double eddlon, void *stream) {
  (void)uatupos, (void)vatupos, (void)uatvpos, (void)vatvpos, (void)uavg, (void)vavg, (void)ures, (void)vres;
  return oeb::hadv_run<5>("hadvuv5th", d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat,
                          eddlon, stream);
}

extern "C" int oeb_hadvuv_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                              const oeb_s3f *vin, const oeb_s1f *acrlat0, const oeb_s1f *acrlat1,
                              const oeb_s1f *tgrlatda0, const oeb_s1f *tgrlatda1, const oeb_s3f *uatupos,
                              const oeb_s3f *vatupos, const oeb_s3f *uatvpos, const oeb_s3f *vatvpos,
                              const oeb_s3f *uavg, const oeb_s3f *vavg, const oeb_s3f *ures, const oeb_s3f *vres,
                              float eddlat, float eddlon, void *stream) {
  (void)uatupos, (void)vatupos, (void)uatvpos, (void)vatvpos, (void)uavg, (void)vavg, (void)ures, (void)vres;
  return oeb::hadv_run<3>("hadvuv", d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat, eddlon,
                          stream);
}

extern "C" int oeb_hadvuv5th_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout,
                                 const oeb_s3f *uin, const oeb_s3f *vin, const oeb_s1f *acrlat0,
                                 const oeb_s1f *acrlat1, const oeb_s1f *tgrlatda0, const oeb_s1f *tgrlatda1,
                                 const oeb_s3f *uatupos, const oeb_s3f *vatupos, const oeb_s3f *uatvpos,
                                 const oeb_s3f *vatvpos, const oeb_s3f *uavg, const oeb_s3f *vavg,
                                 const oeb_s3f *ures, const oeb_s3f *vres, float eddlat, float eddlon,
                                 void *stream) {
  (void)uatupos, (void)vatupos, (void)uatvpos, (void)vatvpos, (void)uavg, (void)vavg, (void)ures, (void)vres;
  return oeb::hadv_run<5>("hadvuv5th", d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat,
                          eddlon, stream);
}
