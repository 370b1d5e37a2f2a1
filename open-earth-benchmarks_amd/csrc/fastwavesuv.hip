// fastwavesuv.hip -- COSMO fast-wave (u, v) update.
// Reference: kernels/fastwavesuv.h:154-191 (fastwavesuv_openmp), expressions
// :162-187.  ppgk (pressure at the half level below k) is formed at the
// centre C, east E = (i+1) and north N = (j+1) points.  It is carried from
// level k+1 to level k of the next step, and so are ppuv and hhl, so each plane
// of wgtfac, ppuv and hhl is loaded once per chunk.
// Reach: i, j in [0, +1]; k in [-1, +1] (the k-halo planes -1 and nz are read).
// FP32 forms: ppuv and hhl are FP64 values (FP64 arrays in
// oeb_fastwavesuv_f32, FP32 deviations from FP64 per-level profiles in
// oeb_fastwavesuv_f32d; pressure ~1e5 Pa and heights lose their small
// differences to cancellation in FP32), and ppgk, ppgc and ppgu/ppgv are
// formed in FP64; ppgu/ppgv are rounded once to float.
// Algorithmic traffic: 8 inputs + 2 outputs = 80 B per point (FP64); 48 B in
// the FP32 form (ppuv and hhl 8 B, the rest 4 B), 40 B with deviations.
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 8;
constexpr int kPairI = 2 * kTileI;

template <class T>
__device__ __forceinline__ double ppgk_of(T wgt, double pp_k, double pp_km1) {
  const double w = wgt;
  return w * pp_k + (1.0 - w) * pp_km1;
}

// Level-k and level-(k+1) values of one quantity at C = (i,j), E = (i+1,j),
// N = (i,j+1).
struct CEN {
  double C, E, N, C1, E1, N1;
};

// One point (kernels/fastwavesuv.h:312-337).  pp: ppuv at level k (C, E, N);
// gk: ppgk at k and k+1; h: hhl at k and k+1; r*: rho at level k.
template <class T>
__device__ __forceinline__ void fw_point(double ppC, double ppE, double ppN, const CEN &gk, const CEN &h, T rC, T rE,
                                         T rN, T uin, T vin, T utens, T vtens, T fxj, T edadlat, T dt, T &uo, T &vo) {
  const double gcC = gk.C1 - gk.C, gcE = gk.E1 - gk.E, gcN = gk.N1 - gk.N;
  const double ppgu =
      (ppE - ppC) + (gcE + gcC) * 0.5 * ((h.C1 + h.C) - (h.E1 + h.E)) / ((h.C1 - h.C) + (h.E1 - h.E));
  const double ppgv =
      (ppN - ppC) + (gcN + gcC) * 0.5 * ((h.C1 + h.C) - (h.N1 + h.N)) / ((h.C1 - h.C) + (h.N1 - h.N));
  uo = uin + dt * (utens - T(ppgu) * (T(2.0) * fxj / (rE + rC)));
  vo = vin + dt * (vtens - T(ppgv) * (T(2.0) * edadlat / (rN + rC)));
}

template <class T, class IS>
struct FwArgs {
  T *uout, *vout;
  const T *uin, *vin, *utens, *vtens, *wgtfac, *rho, *fx;
  IS ppuv, hhl;
  int fx_stride;
  T edadlat, dt;
};

// Logical block `id` of a launch: one 64 x 4 tile of columns over a k-chunk.
template <class T, class IS>
__device__ __forceinline__ void fastwavesuv_body(const FwArgs<T, IS> &a, const Geo &g, const Tiles &t, int id) {
  const T edadlat = a.edadlat, dt = a.dt;
  const TileCoord c = tile_coord_of(t, id, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  const T fxj = a.fx[int64_t(j) * a.fx_stride];
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  const int64_t dn = at - sk;
  const auto p0 = a.ppuv.level(c.k0), pm = a.ppuv.level(c.k0 - 1);
  const auto h0 = a.hhl.level(c.k0);
  double ppC = a.ppuv.at(at, p0), ppE = a.ppuv.at(at + 1, p0), ppN = a.ppuv.at(at + sj, p0);
  CEN gk, h;
  gk.C = ppgk_of(a.wgtfac[at], ppC, a.ppuv.at(dn, pm));
  gk.E = ppgk_of(a.wgtfac[at + 1], ppE, a.ppuv.at(dn + 1, pm));
  gk.N = ppgk_of(a.wgtfac[at + sj], ppN, a.ppuv.at(dn + sj, pm));
  h.C = a.hhl.at(at, h0), h.E = a.hhl.at(at + 1, h0), h.N = a.hhl.at(at + sj, h0);
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const auto p1 = a.ppuv.level(k + 1);
    const auto h1 = a.hhl.level(k + 1);
    const double ppC1 = a.ppuv.at(up, p1), ppE1 = a.ppuv.at(up + 1, p1), ppN1 = a.ppuv.at(up + sj, p1);
    gk.C1 = ppgk_of(a.wgtfac[up], ppC1, ppC);
    gk.E1 = ppgk_of(a.wgtfac[up + 1], ppE1, ppE);
    gk.N1 = ppgk_of(a.wgtfac[up + sj], ppN1, ppN);
    h.C1 = a.hhl.at(up, h1), h.E1 = a.hhl.at(up + 1, h1), h.N1 = a.hhl.at(up + sj, h1);
    T uo, vo;
    fw_point(ppC, ppE, ppN, gk, h, a.rho[at], a.rho[at + 1], a.rho[at + sj], ld_nt(a.uin + at), ld_nt(a.vin + at),
             ld_nt(a.utens + at), ld_nt(a.vtens + at), fxj, edadlat, dt, uo, vo);
    st_nt(a.uout + at, uo);
    st_nt(a.vout + at, vo);
    ppC = ppC1, ppE = ppE1, ppN = ppN1;
    gk.C = gk.C1, gk.E = gk.E1, gk.N = gk.N1;
    h.C = h.C1, h.E = h.E1, h.N = h.N1;
  }
}

template <class T, class IS>
__global__ void __launch_bounds__(kTileI *kTileJ) fastwavesuv_kernel(FwArgs<T, IS> a, Geo g, Tiles t) {
  fastwavesuv_body(a, g, t, int(blockIdx.x));
}

template <class T, class IS>
hipError_t fastwavesuv_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<FwArgs<T, IS>, fastwavesuv_body<T, IS>>(tab, n, total, s, dim3(kTileI, kTileJ));
}
template <class T, class IS>
const BatchKind kFastwavesuvBatch{"fastwavesuv", fastwavesuv_batch<T, IS>};

// Tiles of 128 columns x TJ rows (4 or 8).  Two adjacent columns (i, i+1)
// per lane, i even: the element-type fields move as pairs (16 bytes in FP64,
// 8 in FP32), ppuv and hhl as pairs of their storage (16-byte FP64 or 8-byte
// FP32 deviation pairs).  The east neighbour of column i is column i+1 of the
// pair; that of column i+1 is a single load at i+2.  Suffix P: the pair at (i, i+1);
// Q: the pair at row j+1; X: column i+2 at row j.
template <class T>
struct Pr {
  T x, y;
};
template <class T>
__device__ __forceinline__ Pr<T> ldp(const T *p) {
  const typename Vec<T, 2>::type v = ldv<T, 2>(p);
  return Pr<T>{v[0], v[1]};
}
template <class T>
__device__ __forceinline__ Pr<T> ldp_nt(const T *p) {
  const typename Vec<T, 2>::type v = ldv_nt<T, 2>(p);
  return Pr<T>{v[0], v[1]};
}
template <class T>
__device__ __forceinline__ void stp_nt(T *p, T x, T y) {
  typename Vec<T, 2>::type v;
  v[0] = x;
  v[1] = y;
  stv_nt<T, 2>(p, v);
}

// Column i+2 of a pair is the first column of the pair to its right, which
// the lane to the right holds: from_right moves lane n+1's value to lane n
// (a wave is one row of 64 pairs).  The last lane of a row, and the lane of
// the domain's last pair, load their column i+2 instead (own_x below).  FP64
// takes it so: four loads fewer per level and 134 VGPRs instead of 138; at
// 512x512x80 / 256x256x80 (ten-stencil step, both builds on one box)
// 287.4-288.0 / 79.9-80.2 us against 292.1-293.0 / 81.2-81.6.  The FP32 forms load it in every lane: shifted,
// they took 206.2-207.0 / 57.5-57.8 us against 197.7-197.9 / 54.0-54.3, and
// with own_x's loads issued one level ahead 205.5-206.5 (deviation islands
// 211-212 against 165) at 512x512x80; FP64 then took the same, 285.0 us.
template <class T, class IS, int TJ = kTileJ>
__device__ __forceinline__ void fastwavesuv_pair_body(const FwArgs<T, IS> &a, const Geo &g, const Tiles &t, int id) {
  const T edadlat = a.edadlat, dt = a.dt;
  const TileCoord c = tile_coord_of(t, id, kPairI, TJ, g.nz);
  const int i = c.i0 + 2 * int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  const T fxj = a.fx[int64_t(j) * a.fx_stride];
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  const int64_t dn = at - sk;
  const auto p0 = a.ppuv.level(c.k0), pm = a.ppuv.level(c.k0 - 1);
  const auto h0 = a.hhl.level(c.k0);
  constexpr bool kShift = sizeof(T) == 8;
  const bool own_x = !kShift || int(threadIdx.x) == kTileI - 1 || i + 2 >= g.nx;
  d2 ppP = a.ppuv.pair(at, p0), ppQ = a.ppuv.pair(at + sj, p0);
  d2 gkP, gkQ;
  {
    const Pr<T> wP = ldp(a.wgtfac + at), wQ = ldp(a.wgtfac + at + sj);
    const d2 mP = a.ppuv.pair(dn, pm), mQ = a.ppuv.pair(dn + sj, pm);
    gkP = d2{ppgk_of(wP.x, ppP.x, mP.x), ppgk_of(wP.y, ppP.y, mP.y)};
    gkQ = d2{ppgk_of(wQ.x, ppQ.x, mQ.x), ppgk_of(wQ.y, ppQ.y, mQ.y)};
  }
  d2 hP = a.hhl.pair(at, h0), hQ = a.hhl.pair(at + sj, h0);
  double ppX = 0, gkX = 0, hX = 0;
  if constexpr (kShift) ppX = from_right(ppP.x), gkX = from_right(gkP.x), hX = from_right(hP.x);
  if (own_x) {
    ppX = a.ppuv.at(at + 2, p0);
    gkX = ppgk_of(a.wgtfac[at + 2], ppX, a.ppuv.at(dn + 2, pm));
    hX = a.hhl.at(at + 2, h0);
  }
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const auto p1 = a.ppuv.level(k + 1);
    const auto h1 = a.hhl.level(k + 1);
    const d2 ppP1 = a.ppuv.pair(up, p1), ppQ1 = a.ppuv.pair(up + sj, p1);
    const Pr<T> wP = ldp(a.wgtfac + up), wQ = ldp(a.wgtfac + up + sj);
    const d2 gkP1{ppgk_of(wP.x, ppP1.x, ppP.x), ppgk_of(wP.y, ppP1.y, ppP.y)};
    const d2 gkQ1{ppgk_of(wQ.x, ppQ1.x, ppQ.x), ppgk_of(wQ.y, ppQ1.y, ppQ.y)};
    const d2 hP1 = a.hhl.pair(up, h1), hQ1 = a.hhl.pair(up + sj, h1);
    const Pr<T> rP = ldp(a.rho + at), rQ = ldp(a.rho + at + sj);
    double ppX1 = 0, gkX1 = 0, hX1 = 0;
    T rX = 0;
    if constexpr (kShift)
      ppX1 = from_right(ppP1.x), gkX1 = from_right(gkP1.x), hX1 = from_right(hP1.x), rX = from_right(rP.x);
    if (own_x) {
      ppX1 = a.ppuv.at(up + 2, p1);
      gkX1 = ppgk_of(a.wgtfac[up + 2], ppX1, ppX);
      hX1 = a.hhl.at(up + 2, h1);
      rX = a.rho[at + 2];
    }
    const Pr<T> ui = ldp_nt(a.uin + at), vi = ldp_nt(a.vin + at);
    const Pr<T> ut = ldp_nt(a.utens + at), vt = ldp_nt(a.vtens + at);
    T u0, v0, u1, v1;
    fw_point(ppP.x, ppP.y, ppQ.x, CEN{gkP.x, gkP.y, gkQ.x, gkP1.x, gkP1.y, gkQ1.x},
             CEN{hP.x, hP.y, hQ.x, hP1.x, hP1.y, hQ1.x}, rP.x, rP.y, rQ.x, ui.x, vi.x, ut.x, vt.x, fxj, edadlat, dt,
             u0, v0);
    fw_point(ppP.y, ppX, ppQ.y, CEN{gkP.y, gkX, gkQ.y, gkP1.y, gkX1, gkQ1.y}, CEN{hP.y, hX, hQ.y, hP1.y, hX1, hQ1.y},
             rP.y, rX, rQ.y, ui.y, vi.y, ut.y, vt.y, fxj, edadlat, dt, u1, v1);
    stp_nt(a.uout + at, u0, u1);
    stp_nt(a.vout + at, v0, v1);
    ppP = ppP1, ppQ = ppQ1, ppX = ppX1;
    gkP = gkP1, gkQ = gkQ1, gkX = gkX1;
    hP = hP1, hQ = hQ1, hX = hX1;
  }
}

template <class T, class IS, int TJ = kTileJ>
__global__ void __launch_bounds__(kTileI *TJ) fastwavesuv_pair_kernel(FwArgs<T, IS> a, Geo g, Tiles t) {
  fastwavesuv_pair_body<T, IS, TJ>(a, g, t, int(blockIdx.x));
}

template <class T, class IS>
hipError_t fastwavesuv_pair_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<FwArgs<T, IS>, fastwavesuv_pair_body<T, IS>>(tab, n, total, s, dim3(kTileI, kTileJ));
}
template <class T, class IS>
const BatchKind kFastwavesuvPairBatch{"fastwavesuv", fastwavesuv_pair_batch<T, IS>};

// Validates and launches.  ppuv and hhl: FP64 arrays (ppuv0, hhl0 NULL), or
// FP32 deviation arrays with FP64 base profiles over their levels, ppuv0
// over [-1, nz + 1) and hhl0 over [0, nz + 1).
template <class T, class S3, class S1, class SI>
int fastwavesuv_run(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
                    const S3 *utens, const S3 *vtens, const S3 *wgtfac, const SI *ppuv, const oeb_s1 *ppuv0,
                    const SI *hhl, const oeb_s1 *hhl0, const S3 *rho, const S1 *fx, T edadlat, T dt, void *stream) {
  constexpr bool kDev = std::is_same<SI, oeb_s3f>::value;
  using IS = typename std::conditional<kDev, IslandDev, IslandF64>::type;
  const char *K = "fastwavesuv";  // one name for both storages: knobs and messages
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_h{0, nx + 1, 0, ny + 1, 0, nz};
  const Box r_w{0, nx + 1, 0, ny + 1, 0, nz + 1}, r_p{0, nx + 1, 0, ny + 1, -1, nz + 1};
  const T *p_uo, *p_vo;
  FwArgs<T, IS> a;
  a.fx_stride = 1;
  OEB_TRY(check3(K, "uout", uout, r0, g, true, &p_uo));
  OEB_TRY(check3(K, "vout", vout, r0, g, false, &p_vo));
  OEB_TRY(check3(K, "uin", uin, r0, g, false, &a.uin));
  OEB_TRY(check3(K, "vin", vin, r0, g, false, &a.vin));
  OEB_TRY(check3(K, "utens", utens, r0, g, false, &a.utens));
  OEB_TRY(check3(K, "vtens", vtens, r0, g, false, &a.vtens));
  OEB_TRY(check3(K, "wgtfac", wgtfac, r_w, g, false, &a.wgtfac));
  OEB_TRY(check3(K, "ppuv", ppuv, r_p, g, false, &a.ppuv.p));
  OEB_TRY(check3(K, "hhl", hhl, r_w, g, false, &a.hhl.p));
  OEB_TRY(check3(K, "rho", rho, r_h, g, false, &a.rho));
  OEB_TRY(check1(K, "fx", fx, 0, ny, &a.fx, &a.fx_stride));
  AliasSet as;
  if constexpr (kDev) {
    OEB_TRY(check1(K, "ppuv0", ppuv0, -1, nz + 1, &a.ppuv.base, &a.ppuv.bs));
    OEB_TRY(check1(K, "hhl0", hhl0, 0, nz + 1, &a.hhl.base, &a.hhl.bs));
    as.add1("ppuv0", a.ppuv.base, -1, nz + 1, a.ppuv.bs);
    as.add1("hhl0", a.hhl.base, 0, nz + 1, a.hhl.bs);
  } else {
    (void)ppuv0, (void)hhl0;
  }
  as.add("uout", p_uo, r0, g, true);
  as.add("vout", p_vo, r0, g, true);
  as.add("uin", a.uin, r0, g, false);
  as.add("vin", a.vin, r0, g, false);
  as.add("utens", a.utens, r0, g, false);
  as.add("vtens", a.vtens, r0, g, false);
  as.add("wgtfac", a.wgtfac, r_w, g, false);
  as.add("ppuv", a.ppuv.p, r_p, g, false);
  as.add("hhl", a.hhl.p, r_w, g, false);
  as.add("rho", a.rho, r_h, g, false);
  as.add1("fx", a.fx, 0, ny, a.fx_stride);
  OEB_TRY(as.check(K));
  a.uout = const_cast<T *>(p_uo);
  a.vout = const_cast<T *>(p_vo);
  a.edadlat = edadlat;
  a.dt = dt;
  // Tile height: 8 rows on levels of more than 256x256 points in plain launches (tile knob 1),
  // else 4 (tile knob 0).  Batched launches take 4 rows, so that one launch covers the members of
  // a run whatever their sizes.  At 512x512x80 (the ten-stencil step), 8-row tiles of 16 levels
  // took FP64 290.5-290.8 us against 298.3 for 4 rows and 8 levels (12-40 levels: 290.6-292.9;
  // 4 rows with 12-40 levels: 298.1-311.3); FP32 8-row tiles of 12 levels 196.5-196.7 against
  // 198.6-198.8 (12-40 levels: 196.5-204.0).  The taller tile fetches the j+1 row of ppuv,
  // wgtfac, hhl and rho, which the tile above owns, for 8 rows instead of 4, and the longer chunk
  // fetches the k-halo planes once per 16 levels.  At 256x256x80 8-row tiles measured slower
  // (FP64 83.1-85.7 against 81.3 us, FP32 56.9-57.2 against 54.1-54.3).  16-row tiles of 1024
  // lanes spill registers.
  const bool big = int64_t(nx) * ny > 256 * 256;
  const bool pairs =
      knob(K, sizeof(T), d, KNOB_VECTOR, true) != 0 && nx % 2 == 0 &&
      pairs_ok(g, {p_uo, p_vo, a.uin, a.vin, a.utens, a.vtens, a.wgtfac, a.ppuv.ptr(), a.hhl.ptr(), a.rho});
  const bool tall = pairs && !batch_recording() && knob(K, sizeof(T), d, KNOB_TILE, big ? 1 : 0) == 1;
  const int tj = tall ? 8 : kTileJ;
  // Levels per workgroup with 4-row tiles: 8, and 4 in the FP32 form, which measured 56 / 203 us
  // against 59 / 204 with 8 at 256x256x80 / 512x512x80.  FP64 takes 3 on levels of at most
  // 256x256 points: 83.7 us against 85.1-85.9 with 8 at 256x256x80; at 512x512x80 the chunks
  // 2..10 measured 303-313 us, 8 among the fastest.
  const int kc_dflt = tall ? (sizeof(T) == 4 ? 12 : 16) : sizeof(T) == 4 ? 4 : (big ? kChunk : 3);
  const int kc = knob(K, sizeof(T), d, KNOB_KCHUNK, kc_dflt);
  // The tiles of a row on one XCD (Tiles::irow): FP32 204 / 58 us against 209 / 60
  // at 512x512x80 / 256x256x80, FP64 the same (order knob 0 for row order).
  const int xo = knob(K, sizeof(T), d, KNOB_ORDER, 2);
  if (pairs) {
    const Tiles t = make_tiles(nx, ny, int(nz), kPairI, tj, kc, xo);
    OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kPairI, tj));
    if (tall) {
      hipLaunchKernelGGL((fastwavesuv_pair_kernel<T, IS, 8>), dim3(t.n), dim3(kTileI, 8), 0, hipStream_t(stream), a,
                         g, t);
      return launch_status(K);
    }
    if (batch_record(kFastwavesuvPairBatch<T, IS>, a, g, t, t.n, as)) return OEB_OK;
    hipLaunchKernelGGL((fastwavesuv_pair_kernel<T, IS>), dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a,
                       g, t);
    return launch_status(K);
  }
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kc, xo);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  if (batch_record(kFastwavesuvBatch<T, IS>, a, g, t, t.n, as)) return OEB_OK;
  hipLaunchKernelGGL((fastwavesuv_kernel<T, IS>), dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a, g, t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

// ppgk, ppgc, ppgu and ppgv are the reference's scratch arrays; nothing is stored in them.
extern "C" int oeb_fastwavesuv_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                   const oeb_s3 *vin, const oeb_s3 *utens, const oeb_s3 *vtens, const oeb_s3 *wgtfac,
                                   const oeb_s3 *ppuv, const oeb_s3 *hhl, const oeb_s3 *rho, const oeb_s1 *fx,
                                   const oeb_s3 *ppgk, const oeb_s3 *ppgc, const oeb_s3 *ppgu, const oeb_s3 *ppgv,
                                   double edadlat, double dt, void *stream) {
  (void)ppgk, (void)ppgc, (void)ppgu, (void)ppgv;
  return oeb::fastwavesuv_run<double>(d, uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, nullptr, hhl, nullptr, rho,
                                      fx, edadlat, dt, stream);
}

extern "C" int oeb_fastwavesuv_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout,
                                   const oeb_s3f *uin, const oeb_s3f *vin, const oeb_s3f *utens,
                                   const oeb_s3f *vtens, const oeb_s3f *wgtfac, const oeb_s3 *ppuv, const oeb_s3 *hhl,
                                   const oeb_s3f *rho, const oeb_s1f *fx, const oeb_s3f *ppgk, const oeb_s3f *ppgc,
                                   const oeb_s3f *ppgu, const oeb_s3f *ppgv, float edadlat, float dt, void *stream) {
  (void)ppgk, (void)ppgc, (void)ppgu, (void)ppgv;
  return oeb::fastwavesuv_run<float>(d, uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, nullptr, hhl, nullptr, rho,
                                     fx, edadlat, dt, stream);
}

extern "C" int oeb_fastwavesuv_f32d(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout,
                                    const oeb_s3f *uin, const oeb_s3f *vin, const oeb_s3f *utens,
                                    const oeb_s3f *vtens, const oeb_s3f *wgtfac, const oeb_s3f *ppuv,
                                    const oeb_s1 *ppuv0, const oeb_s3f *hhl, const oeb_s1 *hhl0, const oeb_s3f *rho,
                                    const oeb_s1f *fx, const oeb_s3f *ppgk, const oeb_s3f *ppgc, const oeb_s3f *ppgu,
                                    const oeb_s3f *ppgv, float edadlat, float dt, void *stream) {
  (void)ppgk, (void)ppgc, (void)ppgu, (void)ppgv;
  return oeb::fastwavesuv_run<float>(d, uout, vout, uin, vin, utens, vtens, wgtfac, ppuv, ppuv0, hhl, hhl0, rho, fx,
                                     edadlat, dt, stream);
}
