// p_grad_c.hip -- FV3 C-grid pressure-gradient update of (u, v).
// Reference: kernels/p_grad_c.h:80-101 (p_grad_c_openmp), expressions :87-99.
// delpc(i-1) and delpc(j-1) are read directly, as the OpenMP variant does.
// The DEFAULT variant instead reads an uninitialised copy at i = 0 and j = 0.
// Reach: i, j in [-1, 0], k in [0, +1].  Each lane marches up its column and
// carries the gz / pkc values of level k+1 over as level k of the next step,
// so every plane of gz and pkc is loaded once per chunk.
// FP32 forms: gz and pkc are FP64 values (FP64 arrays in oeb_p_grad_c_f32,
// FP32 deviations from FP64 per-level profiles in oeb_p_grad_c_f32d), the
// bracket (dgz*dpk + dgz'*dpk') is formed in FP64 and rounded once to float;
// the rest is float.  Algorithmic traffic: 7 inputs + 2 outputs = 72 B per
// point (FP64); 44 B in the FP32 form (gz and pkc 8 B, the rest 4 B), 36 B
// with deviations.
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 4;
constexpr int kPairI = 2 * kTileI;

// One point (kernels/p_grad_c.h:87-99).  C = (i,j), W = (i-1,j), S = (i,j-1);
// suffix 1 = level k+1.
template <class T>
__device__ __forceinline__ void pgc_point(double gzC, double gzW, double gzS, double gzC1, double gzW1, double gzS1,
                                          double pkC, double pkW, double pkS, double pkC1, double pkW1, double pkS1,
                                          T wk0, T wkW, T wkS, T uin, T vin, T rdxc, T rdyc, T dt2, T &uo, T &vo) {
  const double bu = ((gzW1 - gzC) * (pkC1 - pkW)) + ((gzW - gzC1) * (pkW1 - pkC));
  const double bv = ((gzS1 - gzC) * (pkC1 - pkS)) + ((gzS - gzC1) * (pkS1 - pkC));
  uo = (uin + (((dt2 * rdxc) / (wkW + wk0)) * T(bu)));
  vo = (vin + (((dt2 * rdyc) / (wkS + wk0)) * T(bv)));
}

template <class T, class IS>
struct PgcArgs {
  T *uout, *vout;
  const T *uin, *vin, *rdxc, *rdyc, *delpc;
  IS gz, pkc;
};

// (The west point's gz, pkc and delpc taken from the lane to the left with a
// DPP wave shift, loaded only by a row's first lane, took FP64 286.7-287.8 /
// 74.5 us against 256.9 / 67.5 at 512x512x80 / 256x256x80; those loads hit the
// centre's cache lines, so they cost little.)
template <class T, class IS>
__global__ void __launch_bounds__(kTileI *kTileJ) p_grad_c_kernel(PgcArgs<T, IS> a, T dt2, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  const auto z0 = a.gz.level(c.k0);
  const auto p0 = a.pkc.level(c.k0);
  double gzC = a.gz.at(at, z0), gzW = a.gz.at(at - 1, z0), gzS = a.gz.at(at - sj, z0);
  double pkC = a.pkc.at(at, p0), pkW = a.pkc.at(at - 1, p0), pkS = a.pkc.at(at - sj, p0);
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const auto z1 = a.gz.level(k + 1);
    const auto p1 = a.pkc.level(k + 1);
    const double gzC1 = a.gz.at(up, z1), gzW1 = a.gz.at(up - 1, z1), gzS1 = a.gz.at(up - sj, z1);
    const double pkC1 = a.pkc.at(up, p1), pkW1 = a.pkc.at(up - 1, p1), pkS1 = a.pkc.at(up - sj, p1);
    T uo, vo;
    pgc_point(gzC, gzW, gzS, gzC1, gzW1, gzS1, pkC, pkW, pkS, pkC1, pkW1, pkS1, a.delpc[at], a.delpc[at - 1],
              a.delpc[at - sj], ld_nt(a.uin + at), ld_nt(a.vin + at), ld_nt(a.rdxc + at), ld_nt(a.rdyc + at), dt2,
              uo, vo);
    st_nt(a.uout + at, uo);
    st_nt(a.vout + at, vo);
    gzC = gzC1, gzW = gzW1, gzS = gzS1;
    pkC = pkC1, pkW = pkW1, pkS = pkS1;
  }
}

// Two adjacent columns (i, i+1) per lane, i even: the element-type fields
// move as pairs (16 bytes in FP64, 8 in FP32), gz and pkc as pairs of their
// storage (16-byte FP64 or 8-byte FP32 deviation pairs).  For column i+1 the
// west neighbour is column i of the pair.  Column i-1 is the second column of
// the pair to the left: FP64 takes it from the lane to the left (from_left);
// the first lane of a row, and the FP32 forms, load it (own_w).
template <class T, class IS>
__global__ void __launch_bounds__(kTileI *kTileJ) p_grad_c_pair_kernel(PgcArgs<T, IS> a, T dt2, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kPairI, kTileJ, g.nz);
  const int i = c.i0 + 2 * int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  const auto z0 = a.gz.level(c.k0);
  const auto p0 = a.pkc.level(c.k0);
  constexpr bool kShift = sizeof(T) == 8;
  const bool own_w = !kShift || threadIdx.x == 0;
  d2 gzC = a.gz.pair(at, z0), gzS = a.gz.pair(at - sj, z0);
  d2 pkC = a.pkc.pair(at, p0), pkS = a.pkc.pair(at - sj, p0);
  double gzW = 0, pkW = 0;
  if constexpr (kShift) gzW = from_left(gzC.y), pkW = from_left(pkC.y);
  if (own_w) gzW = a.gz.at(at - 1, z0), pkW = a.pkc.at(at - 1, p0);
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const auto z1 = a.gz.level(k + 1);
    const auto p1 = a.pkc.level(k + 1);
    const d2 gzC1 = a.gz.pair(up, z1), gzS1 = a.gz.pair(up - sj, z1);
    const d2 pkC1 = a.pkc.pair(up, p1), pkS1 = a.pkc.pair(up - sj, p1);
    const typename Vec<T, 2>::type dpC = ldv<T, 2>(a.delpc + at), dpS = ldv<T, 2>(a.delpc + at - sj);
    double gzW1 = 0, pkW1 = 0;
    T dpW = 0;
    if constexpr (kShift) gzW1 = from_left(gzC1.y), pkW1 = from_left(pkC1.y), dpW = from_left(T(dpC[1]));
    if (own_w) gzW1 = a.gz.at(up - 1, z1), pkW1 = a.pkc.at(up - 1, p1), dpW = a.delpc[at - 1];
    const typename Vec<T, 2>::type ui = ldv_nt<T, 2>(a.uin + at), vi = ldv_nt<T, 2>(a.vin + at);
    const typename Vec<T, 2>::type rx = ldv_nt<T, 2>(a.rdxc + at), ry = ldv_nt<T, 2>(a.rdyc + at);
    typename Vec<T, 2>::type uo, vo;
    T u0, v0, u1, v1;
    pgc_point(gzC.x, gzW, gzS.x, gzC1.x, gzW1, gzS1.x, pkC.x, pkW, pkS.x, pkC1.x, pkW1, pkS1.x, dpC[0], dpW, dpS[0],
              ui[0], vi[0], rx[0], ry[0], dt2, u0, v0);
    pgc_point(gzC.y, gzC.x, gzS.y, gzC1.y, gzC1.x, gzS1.y, pkC.y, pkC.x, pkS.y, pkC1.y, pkC1.x, pkS1.y, dpC[1], dpC[0],
              dpS[1], ui[1], vi[1], rx[1], ry[1], dt2, u1, v1);
    uo[0] = u0, uo[1] = u1;
    vo[0] = v0, vo[1] = v1;
    stv_nt<T, 2>(a.uout + at, uo);
    stv_nt<T, 2>(a.vout + at, vo);
    gzC = gzC1, gzS = gzS1, gzW = gzW1;
    pkC = pkC1, pkS = pkS1, pkW = pkW1;
  }
}

// Validates and launches.  gz and pkc: FP64 arrays (gz0 == pkc0 == NULL), or
// FP32 deviation arrays with their FP64 base profiles gz0, pkc0 over the
// levels [0, nz + 1).
template <class T, class S3, class SI>
int p_grad_c_run(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin, const S3 *rdxc,
                 const S3 *rdyc, const S3 *delpc, const SI *gz, const SI *pkc, const oeb_s1 *gz0, const oeb_s1 *pkc0,
                 T dt2, void *stream) {
  constexpr bool kDev = std::is_same<SI, oeb_s3f>::value;
  using IS = typename std::conditional<kDev, IslandDev, IslandF64>::type;
  const char *K = "p_grad_c";  // one name for both storages: knobs and messages
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_dp{-1, nx, -1, ny, 0, nz}, r_k{-1, nx, -1, ny, 0, nz + 1};
  const T *p_uo, *p_vo;
  PgcArgs<T, IS> a;
  OEB_TRY(check3(K, "uout", uout, r0, g, true, &p_uo));
  OEB_TRY(check3(K, "vout", vout, r0, g, false, &p_vo));
  OEB_TRY(check3(K, "uin", uin, r0, g, false, &a.uin));
  OEB_TRY(check3(K, "vin", vin, r0, g, false, &a.vin));
  OEB_TRY(check3(K, "rdxc", rdxc, r0, g, false, &a.rdxc));
  OEB_TRY(check3(K, "rdyc", rdyc, r0, g, false, &a.rdyc));
  OEB_TRY(check3(K, "delpc", delpc, r_dp, g, false, &a.delpc));
  OEB_TRY(check3(K, "gz", gz, r_k, g, false, &a.gz.p));
  OEB_TRY(check3(K, "pkc", pkc, r_k, g, false, &a.pkc.p));
  AliasSet as;
  if constexpr (kDev) {
    OEB_TRY(check1(K, "gz0", gz0, 0, nz + 1, &a.gz.base, &a.gz.bs));
    OEB_TRY(check1(K, "pkc0", pkc0, 0, nz + 1, &a.pkc.base, &a.pkc.bs));
    as.add1("gz0", a.gz.base, 0, nz + 1, a.gz.bs);
    as.add1("pkc0", a.pkc.base, 0, nz + 1, a.pkc.bs);
  } else {
    (void)gz0, (void)pkc0;
  }
  as.add("uout", p_uo, r0, g, true);
  as.add("vout", p_vo, r0, g, true);
  as.add("uin", a.uin, r0, g, false);
  as.add("vin", a.vin, r0, g, false);
  as.add("rdxc", a.rdxc, r0, g, false);
  as.add("rdyc", a.rdyc, r0, g, false);
  as.add("delpc", a.delpc, r_dp, g, false);
  as.add("gz", a.gz.p, r_k, g, false);
  as.add("pkc", a.pkc.p, r_k, g, false);
  OEB_TRY(as.check(K));
  a.uout = const_cast<T *>(p_uo);
  a.vout = const_cast<T *>(p_vo);
  // Levels per workgroup: 4, and 3 in the FP64 form, which measured 69.6 / 262.0 us against
  // 70.9-72.0 / 266.9-267.2 with 4 at 256x256x80 / 512x512x80 (the ten-stencil step).
  // Taller tiles, which re-read fewer j+1 rows, measured slower at 512x512x80: FP64 64x8 tiles
  // of 3 / 8 / 16 levels 259.9 / 270.9 / 282.9 us, 64x16 tiles 263.7-268.7, against 256.9;
  // FP32 128x8 tiles of 4 / 12 levels 181.5-182.5 / 189.5 against 181.4 (deviation islands
  // 151.4 / 151.3 against 147.7).
  const int kc = knob(K, sizeof(T), d, KNOB_KCHUNK, sizeof(T) == 8 ? 3 : kChunk);
  // The tiles of a row on one XCD (Tiles::irow): FP32 184 / 52 us against 189 / 53
  // at 512x512x80 / 256x256x80, FP64 the same (order knob 0 for row order).
  const int xo = knob(K, sizeof(T), d, KNOB_ORDER, 2);
  // Column pairs by default.  Measured in the ten-stencil step at 256x256x80
  // and 512x512x80: FP32 pairs take 52 / 189 us against 55 / 203 us for
  // single columns.  FP64 pairs, with the west column taken from the lane to
  // the left (from_left), take 67.2-67.4 / 252.6-252.9 us against 67.4-67.7 /
  // 256.1-256.6 for single columns, and 70.3-70.5 / 253.1 loading it (all on
  // one box).
  // Four columns per lane (16-byte FP32 accesses, 120 VGPRs) measured no
  // faster with deviation islands (149 against 148 us at 512x512x80, 52.5
  // against 43 at 256x256x80), 3% faster with FP64 islands at 512x512x80.
  if (knob_on(K, sizeof(T), d, KNOB_VECTOR, true) && nx % 2 == 0 &&
      pairs_ok(g, {p_uo, p_vo, a.uin, a.vin, a.rdxc, a.rdyc, a.delpc, a.gz.ptr(), a.pkc.ptr()})) {
    const Tiles t = make_tiles(nx, ny, int(nz), kPairI, kTileJ, kc, xo);
    OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kPairI, kTileJ));
    hipLaunchKernelGGL((p_grad_c_pair_kernel<T, IS>), dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a,
                       dt2, g, t);
    return launch_status(K);
  }
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kc, xo);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  hipLaunchKernelGGL((p_grad_c_kernel<T, IS>), dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a, dt2, g,
                     t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

// wk is the reference's scratch array; nothing is stored in it.
extern "C" int oeb_p_grad_c_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                const oeb_s3 *vin, const oeb_s3 *rdxc, const oeb_s3 *rdyc, const oeb_s3 *delpc,
                                const oeb_s3 *gz, const oeb_s3 *pkc, const oeb_s3 *wk, double dt2, void *stream) {
  (void)wk;
  return oeb::p_grad_c_run<double>(d, uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, nullptr, nullptr, dt2,
                                   stream);
}

extern "C" int oeb_p_grad_c_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                                const oeb_s3f *vin, const oeb_s3f *rdxc, const oeb_s3f *rdyc, const oeb_s3f *delpc,
                                const oeb_s3 *gz, const oeb_s3 *pkc, const oeb_s3f *wk, float dt2, void *stream) {
  (void)wk;
  return oeb::p_grad_c_run<float>(d, uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, nullptr, nullptr, dt2,
                                  stream);
}

extern "C" int oeb_p_grad_c_f32d(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                                 const oeb_s3f *vin, const oeb_s3f *rdxc, const oeb_s3f *rdyc, const oeb_s3f *delpc,
                                 const oeb_s3f *gz, const oeb_s1 *gz0, const oeb_s3f *pkc, const oeb_s1 *pkc0,
                                 const oeb_s3f *wk, float dt2, void *stream) {
  (void)wk;
  return oeb::p_grad_c_run<float>(d, uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, gz0, pkc0, dt2, stream);
}
