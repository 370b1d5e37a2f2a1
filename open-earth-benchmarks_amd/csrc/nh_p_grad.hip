// nh_p_grad.hip -- FV3 non-hydrostatic pressure-gradient update of (u, v).
// Reference: kernels/nh_p_grad.h:116-146 (nh_p_grad_openmp), expressions
// :123-142.  wk = pk3(k+1) - pk3(k) is formed on the fly at the centre, east
// and north points.  Reach: i, j in [0, +1], k in [0, +1].  Each lane marches
// up its column and carries the level k+1 values of gz, pk3 and pp over to the
// next step.
// FP32 forms: gz, pp and pk3 are FP64 values (FP64 arrays in
// oeb_nh_p_grad_f32, FP32 deviations from FP64 per-level profiles in
// oeb_nh_p_grad_f32d).  wk, the pressure-thickness sums, du/dv and the pp
// bracket are formed in FP64; du/dv and the bracket are rounded once to
// float, and the rest is float.
// Algorithmic traffic: 8 inputs + 2 outputs = 80 B per point (FP64); 52 B in
// the FP32 form (gz, pp and pk3 8 B, the rest 4 B), 40 B with deviations.
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 8;
constexpr int kPairI = 2 * kTileI;

// Level-k and level-(k+1) values of one field at the centre C = (i,j), the
// east point E = (i+1,j) and the north point N = (i,j+1).
struct CEN {
  double C, E, N, C1, E1, N1;
};

// One point (kernels/nh_p_grad.h:123-142).
template <class T>
__device__ __forceinline__ void nh_point(const CEN &gz, const CEN &pk, const CEN &pp, T w1C, T w1E, T w1N, T uin,
                                         T vin, T rdx, T rdy, T dt, T &uo, T &vo) {
  const double dtd = dt;
  const double wk0 = (pk.C1 - pk.C), wkE = (pk.E1 - pk.E), wkN = (pk.N1 - pk.N);
  const double du = ((dtd / (wk0 + wkE)) * (((gz.C1 - gz.E) * (pk.E1 - pk.C)) + ((gz.C - gz.E1) * (pk.C1 - pk.E))));
  const double bu = ((gz.C1 - gz.E) * (pp.E1 - pp.C)) + ((gz.C - gz.E1) * (pp.C1 - pp.E));
  uo = (((uin + T(du)) + ((dt / (w1C + w1E)) * T(bu))) * rdx);
  const double dv = ((dtd / (wk0 + wkN)) * (((gz.C1 - gz.N) * (pk.N1 - pk.C)) + ((gz.C - gz.N1) * (pk.C1 - pk.N))));
  const double bv = ((gz.C1 - gz.N) * (pp.N1 - pp.C)) + ((gz.C - gz.N1) * (pp.C1 - pp.N));
  vo = (((vin + T(dv)) + ((dt / (w1C + w1N)) * T(bv))) * rdy);
}

template <class T, class IS>
struct NhArgs {
  T *uout, *vout;
  const T *uin, *vin, *rdx, *rdy, *wk1;
  IS gz, pp, pk3;
};

// (Capped at 5 waves per SIMD, amdgpu_waves_per_eu(5), this kernel fits 96
// VGPRs with 2 spilled and took 325.3-325.8 / 84.8-85.6 us against
// 311.1-311.7 / 77.7-77.8 at 512x512x80 / 256x256x80 on one box.)
// (The east point's gz, pk3, pp and wk1 taken from the lane to the right
// with a DPP wave shift, loaded only by a row's last lane, took FP64 331.7-332.6
// / 87.6 us against 310 / 78 at 512x512x80 / 256x256x80; those loads hit the
// centre's cache lines, so they cost little.)
template <class T, class IS>
__global__ void __launch_bounds__(kTileI *kTileJ) nh_p_grad_kernel(NhArgs<T, IS> a, T dt, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  CEN gz, pk, pp;
  {
    const auto z = a.gz.level(c.k0);
    const auto q = a.pk3.level(c.k0);
    const auto r = a.pp.level(c.k0);
    gz.C = a.gz.at(at, z), gz.E = a.gz.at(at + 1, z), gz.N = a.gz.at(at + sj, z);
    pk.C = a.pk3.at(at, q), pk.E = a.pk3.at(at + 1, q), pk.N = a.pk3.at(at + sj, q);
    pp.C = a.pp.at(at, r), pp.E = a.pp.at(at + 1, r), pp.N = a.pp.at(at + sj, r);
  }
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const auto z = a.gz.level(k + 1);
    const auto q = a.pk3.level(k + 1);
    const auto r = a.pp.level(k + 1);
    gz.C1 = a.gz.at(up, z), gz.E1 = a.gz.at(up + 1, z), gz.N1 = a.gz.at(up + sj, z);
    pk.C1 = a.pk3.at(up, q), pk.E1 = a.pk3.at(up + 1, q), pk.N1 = a.pk3.at(up + sj, q);
    pp.C1 = a.pp.at(up, r), pp.E1 = a.pp.at(up + 1, r), pp.N1 = a.pp.at(up + sj, r);
    T uo, vo;
    nh_point(gz, pk, pp, a.wk1[at], a.wk1[at + 1], a.wk1[at + sj], ld_nt(a.uin + at), ld_nt(a.vin + at),
             ld_nt(a.rdx + at), ld_nt(a.rdy + at), dt, uo, vo);
    st_nt(a.uout + at, uo);
    st_nt(a.vout + at, vo);
    gz.C = gz.C1, gz.E = gz.E1, gz.N = gz.N1;
    pk.C = pk.C1, pk.E = pk.E1, pk.N = pk.N1;
    pp.C = pp.C1, pp.E = pp.E1, pp.N = pp.N1;
  }
}

// Two adjacent columns (i, i+1) per lane, i even: the element-type fields
// move as pairs (16 bytes in FP64, 8 in FP32), gz, pp and pk3 as pairs of
// their storage (16-byte FP64 or 8-byte FP32 deviation pairs).  The east neighbour of column i is column i+1 of the pair;
// that of column i+1 is a single load at i+2.
struct PairCol {
  d2 C, N, C1, N1;   // pair at (i, i+1), rows j and j+1, levels k and k+1
  double E, E1;      // column i+2, row j, levels k and k+1
  __device__ CEN lo() const { return CEN{C.x, C.y, N.x, C1.x, C1.y, N1.x}; }
  __device__ CEN hi() const { return CEN{C.y, E, N.y, C1.y, E1, N1.y}; }
  template <class IS>
  __device__ void load(const IS &f, int64_t at, int sj, typename IS::Level lv) {
    C = f.pair(at, lv), N = f.pair(at + sj, lv), E = f.at(at + 2, lv);
  }
  template <class IS>
  __device__ void load_up(const IS &f, int64_t up, int sj, typename IS::Level lv) {
    C1 = f.pair(up, lv), N1 = f.pair(up + sj, lv), E1 = f.at(up + 2, lv);
  }
  __device__ void roll() { C = C1, N = N1, E = E1; }
};

template <class T, class IS>
__global__ void __launch_bounds__(kTileI *kTileJ) nh_p_grad_pair_kernel(NhArgs<T, IS> a, T dt, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kPairI, kTileJ, g.nz);
  const int i = c.i0 + 2 * int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t sk = g.sk;
  int64_t at = int64_t(c.k0) * sk + (i + j * sj);
  PairCol gz, pk, pp;
  gz.load(a.gz, at, sj, a.gz.level(c.k0));
  pk.load(a.pk3, at, sj, a.pk3.level(c.k0));
  pp.load(a.pp, at, sj, a.pp.level(c.k0));
  for (int k = c.k0; k < c.k1; ++k, at += sk) {
    const int64_t up = at + sk;
    const auto z = a.gz.level(k + 1);
    const auto q = a.pk3.level(k + 1);
    const auto r = a.pp.level(k + 1);
    gz.load_up(a.gz, up, sj, z);
    pk.load_up(a.pk3, up, sj, q);
    pp.load_up(a.pp, up, sj, r);
    typedef typename Vec<T, 2>::type P;
    const P w1C = ldv<T, 2>(a.wk1 + at), w1N = ldv<T, 2>(a.wk1 + at + sj);
    const T w1E = a.wk1[at + 2];
    const P ui = ldv_nt<T, 2>(a.uin + at), vi = ldv_nt<T, 2>(a.vin + at);
    const P rx = ldv_nt<T, 2>(a.rdx + at), ry = ldv_nt<T, 2>(a.rdy + at);
    T u0, v0, u1, v1;
    nh_point(gz.lo(), pk.lo(), pp.lo(), w1C[0], w1C[1], w1N[0], ui[0], vi[0], rx[0], ry[0], dt, u0, v0);
    nh_point(gz.hi(), pk.hi(), pp.hi(), w1C[1], w1E, w1N[1], ui[1], vi[1], rx[1], ry[1], dt, u1, v1);
    P uo, vo;
    uo[0] = u0, uo[1] = u1;
    vo[0] = v0, vo[1] = v1;
    stv_nt<T, 2>(a.uout + at, uo);
    stv_nt<T, 2>(a.vout + at, vo);
    gz.roll();
    pk.roll();
    pp.roll();
  }
}

// Validates and launches.  gz, pp and pk3: FP64 arrays (gz0, pp0, pk30 NULL),
// or FP32 deviation arrays with their FP64 base profiles over the levels
// [0, nz + 1).
template <class T, class S3, class SI>
int nh_p_grad_run(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin, const S3 *rdx,
                  const S3 *rdy, const SI *gz, const SI *pp, const SI *pk3, const oeb_s1 *gz0, const oeb_s1 *pp0,
                  const oeb_s1 *pk30, const S3 *wk1, T dt, void *stream) {
  constexpr bool kDev = std::is_same<SI, oeb_s3f>::value;
  using IS = typename std::conditional<kDev, IslandDev, IslandF64>::type;
  const char *K = "nh_p_grad";  // one name for both storages: knobs and messages
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_h{0, nx + 1, 0, ny + 1, 0, nz}, r_k{0, nx + 1, 0, ny + 1, 0, nz + 1};
  const T *p_uo, *p_vo;
  NhArgs<T, IS> a;
  OEB_TRY(check3(K, "uout", uout, r0, g, true, &p_uo));
  OEB_TRY(check3(K, "vout", vout, r0, g, false, &p_vo));
  OEB_TRY(check3(K, "uin", uin, r0, g, false, &a.uin));
  OEB_TRY(check3(K, "vin", vin, r0, g, false, &a.vin));
  OEB_TRY(check3(K, "rdx", rdx, r0, g, false, &a.rdx));
  OEB_TRY(check3(K, "rdy", rdy, r0, g, false, &a.rdy));
  OEB_TRY(check3(K, "gz", gz, r_k, g, false, &a.gz.p));
  OEB_TRY(check3(K, "pp", pp, r_k, g, false, &a.pp.p));
  OEB_TRY(check3(K, "pk3", pk3, r_k, g, false, &a.pk3.p));
  OEB_TRY(check3(K, "wk1", wk1, r_h, g, false, &a.wk1));
  AliasSet as;
  if constexpr (kDev) {
    OEB_TRY(check1(K, "gz0", gz0, 0, nz + 1, &a.gz.base, &a.gz.bs));
    OEB_TRY(check1(K, "pp0", pp0, 0, nz + 1, &a.pp.base, &a.pp.bs));
    OEB_TRY(check1(K, "pk30", pk30, 0, nz + 1, &a.pk3.base, &a.pk3.bs));
    as.add1("gz0", a.gz.base, 0, nz + 1, a.gz.bs);
    as.add1("pp0", a.pp.base, 0, nz + 1, a.pp.bs);
    as.add1("pk30", a.pk3.base, 0, nz + 1, a.pk3.bs);
  } else {
    (void)gz0, (void)pp0, (void)pk30;
  }
  as.add("uout", p_uo, r0, g, true);
  as.add("vout", p_vo, r0, g, true);
  as.add("uin", a.uin, r0, g, false);
  as.add("vin", a.vin, r0, g, false);
  as.add("rdx", a.rdx, r0, g, false);
  as.add("rdy", a.rdy, r0, g, false);
  as.add("gz", a.gz.p, r_k, g, false);
  as.add("pp", a.pp.p, r_k, g, false);
  as.add("pk3", a.pk3.p, r_k, g, false);
  as.add("wk1", a.wk1, r_h, g, false);
  OEB_TRY(as.check(K));
  a.uout = const_cast<T *>(p_uo);
  a.vout = const_cast<T *>(p_vo);
  // Levels per workgroup: 8, and 4 in the FP32 form, which measured 60 / 219 us against 62 / 219
  // with 8 at 256x256x80 / 512x512x80 (the ten-stencil step).  FP64 takes 3 on levels of at
  // most 256x256 points: 79.6 us against 82.3-83.1 with 8 at 256x256x80; at 512x512x80 the
  // chunks 2..10 measured 301-305 us, 8 among the fastest.  Taller tiles, which re-read fewer
  // j+1 rows, measured slower at 512x512x80: FP64 64x8 tiles of 8 / 16 levels 314.5 / 327.5 us,
  // 64x16 tiles 322-324, 64x4 tiles of 16 levels 312.7, against 309.9; FP32 128x8 tiles of 4 /
  // 12 levels 222.3 / 228.3 against 221.1 (deviation islands 186.7 / 173.2 against 172.4).
  const int kc_dflt = sizeof(T) == 4 ? 4 : (int64_t(nx) * ny <= 256 * 256 ? 3 : kChunk);
  const int kc = knob(K, sizeof(T), d, KNOB_KCHUNK, kc_dflt);
  // The tiles of a row on one XCD (Tiles::irow): FP32 219 / 62 us against 223 / 63
  // at 512x512x80 / 256x256x80, FP64 the same (order knob 0 for row order).
  const int xo = knob(K, sizeof(T), d, KNOB_ORDER, 2);
  // Column pairs by default in the FP32 form only.  Measured in the
  // ten-stencil step at 256x256x80 and 512x512x80: FP32 pairs take
  // 63 / 223 us against 68 / 241 us for single columns; FP64 pairs were slower.
  if (knob_on(K, sizeof(T), d, KNOB_VECTOR, sizeof(T) == 4) && nx % 2 == 0 &&
      pairs_ok(g, {p_uo, p_vo, a.uin, a.vin, a.rdx, a.rdy, a.gz.ptr(), a.pp.ptr(), a.pk3.ptr(), a.wk1})) {
    const Tiles t = make_tiles(nx, ny, int(nz), kPairI, kTileJ, kc, xo);
    OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kPairI, kTileJ));
    hipLaunchKernelGGL((nh_p_grad_pair_kernel<T, IS>), dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a, dt,
                       g, t);
    return launch_status(K);
  }
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kc, xo);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  hipLaunchKernelGGL((nh_p_grad_kernel<T, IS>), dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a, dt, g,
                     t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

// wk, du and dv are the reference's scratch arrays; nothing is stored in them.
extern "C" int oeb_nh_p_grad_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                 const oeb_s3 *vin, const oeb_s3 *rdx, const oeb_s3 *rdy, const oeb_s3 *gz,
                                 const oeb_s3 *pp, const oeb_s3 *pk3, const oeb_s3 *wk1, const oeb_s3 *wk,
                                 const oeb_s3 *du, const oeb_s3 *dv, double dt, void *stream) {
  (void)wk, (void)du, (void)dv;
  return oeb::nh_p_grad_run<double>(d, uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, nullptr, nullptr, nullptr, wk1,
                                    dt, stream);
}

extern "C" int oeb_nh_p_grad_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                                 const oeb_s3f *vin, const oeb_s3f *rdx, const oeb_s3f *rdy, const oeb_s3 *gz,
                                 const oeb_s3 *pp, const oeb_s3 *pk3, const oeb_s3f *wk1, const oeb_s3f *wk,
                                 const oeb_s3f *du, const oeb_s3f *dv, float dt, void *stream) {
  (void)wk, (void)du, (void)dv;
  return oeb::nh_p_grad_run<float>(d, uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, nullptr, nullptr, nullptr, wk1, dt,
                                   stream);
}

extern "C" int oeb_nh_p_grad_f32d(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                                  const oeb_s3f *vin, const oeb_s3f *rdx, const oeb_s3f *rdy, const oeb_s3f *gz,
                                  const oeb_s1 *gz0, const oeb_s3f *pp, const oeb_s1 *pp0, const oeb_s3f *pk3,
                                  const oeb_s1 *pk30, const oeb_s3f *wk1, const oeb_s3f *wk, const oeb_s3f *du,
                                  const oeb_s3f *dv, float dt, void *stream) {
  (void)wk, (void)du, (void)dv;
  return oeb::nh_p_grad_run<float>(d, uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, gz0, pp0, pk30, wk1, dt, stream);
}
