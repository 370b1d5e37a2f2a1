// hdiffsmag.hip -- Smagorinsky horizontal diffusion of (u, v), fused.
// Reference: kernels/hdiffsmag.h:205-263 (hdiffsmag_openmp).  Tension T and
// shear S (:213-226) are formed at the three points each output needs:
// T at (i,j), (i+1,j), (i,j+1) and S at (i,j), (i-1,j), (i,j-1).  Then the
// clamped coefficient (:233-236, :248-251) multiplies the Laplacian
// (:239-244, :254-259).  Both clamps are the C++ library's std::min and
// std::max, written as the same comparisons.
// Reach (1,1,0) plus the diagonals (i+1,j-1) and (i-1,j+1).
// Algorithmic traffic: 3 inputs + 2 outputs = 40 B per point (FP64), 20 B (FP32).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 32;
// EARTH_RADIUS_RECIP, include/util.h:17-18: both operands are ElementType,
// so the quotient is formed in the element type.
template <class T>
constexpr T kEarthRadiusRecip = T(1.0) / T(6371.229e3);

// std::max(a, b) / std::min(a, b) exactly as the C++ library defines them.
template <class T>
__device__ __forceinline__ T cxx_max(T a, T b) {
  return (a < b) ? b : a;
}
template <class T>
__device__ __forceinline__ T cxx_min(T a, T b) {
  return (b < a) ? b : a;
}

template <class T>
struct HsArgs {
  T *uout, *vout;
  const T *uin, *vin, *mask, *crlavo, *crlavu, *crlato, *crlatu, *acrlat0;
  int s_vo, s_vu, s_to, s_tu, s_ac;
  T eddlat, eddlon, tau_smag, weight_smag;
};

// Logical block `id` of a launch: one 64 x 4 tile of columns over a k-chunk.
template <class T>
__device__ __forceinline__ void hdiffsmag_body(const HsArgs<T> &a, const Geo &g, const Tiles &t, int id) {
  constexpr T rr = kEarthRadiusRecip<T>;
  const T eddlat = a.eddlat, eddlon = a.eddlon, tau_smag = a.tau_smag, weight_smag = a.weight_smag;
  const T *__restrict__ uin = a.uin;
  const T *__restrict__ vin = a.vin;
  const T *__restrict__ mask = a.mask;
  T *__restrict__ uout = a.uout;
  T *__restrict__ vout = a.vout;
  const TileCoord c = tile_coord_of(t, id, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const T ac_s = a.acrlat0[int64_t(j - 1) * a.s_ac], ac_c = a.acrlat0[int64_t(j) * a.s_ac],
          ac_n = a.acrlat0[int64_t(j + 1) * a.s_ac];
  const T lo = a.crlato[int64_t(j) * a.s_to], lu = a.crlatu[int64_t(j) * a.s_tu];
  const T vo = a.crlavo[int64_t(j) * a.s_vo], vu = a.crlavu[int64_t(j) * a.s_vu];
  int64_t at = int64_t(c.k0) * g.sk + (i + j * sj);
  for (int k = c.k0; k < c.k1; ++k, at += g.sk) {
    const T *u = uin + at;
    const T *v = vin + at;
    // Tension at (p,q) = (v(p,q-1) - v(p,q)) * eddlat * R^-1 - (u(p-1,q) - u(p,q)) * acrlat0(q) * eddlon.
    const T t0 = (v[-sj] - v[0]) * eddlat * rr - (u[-1] - u[0]) * ac_c * eddlon;
    const T t_ip = (v[1 - sj] - v[1]) * eddlat * rr - (u[0] - u[1]) * ac_c * eddlon;
    const T t_jp = (v[0] - v[sj]) * eddlat * rr - (u[sj - 1] - u[sj]) * ac_n * eddlon;
    // Shear at (p,q) = (u(p,q+1) - u(p,q)) * eddlat * R^-1 + (v(p+1,q) - v(p,q)) * acrlat0(q) * eddlon.
    const T s0 = (u[sj] - u[0]) * eddlat * rr + (v[1] - v[0]) * ac_c * eddlon;
    const T s_im = (u[sj - 1] - u[-1]) * eddlat * rr + (v[0] - v[-1]) * ac_c * eddlon;
    const T s_jm = (u[0] - u[-sj]) * eddlat * rr + (v[1 - sj] - v[-sj]) * ac_s * eddlon;
    const T tt0 = t0 * t0, tt_ip = t_ip * t_ip, tt_jp = t_jp * t_jp;
    const T ss0 = s0 * s0, ss_im = s_im * s_im, ss_jm = s_jm * s_jm;
    const T hdweight = weight_smag * ld_nt(mask + at);

    T smag_u = tau_smag * sqrt(T(0.5) * (tt_ip + tt0) + T(0.5) * (ss_jm + ss0)) - hdweight;
    smag_u = cxx_min(T(0.5), cxx_max(T(0.0), smag_u));
    const T u0 = u[0];
    const T lapu = u[1] + u[-1] - T(2.0) * u0 + lo * (u[sj] - u0) + lu * (u[-sj] - u0);
    st_nt(uout + at, u0 + smag_u * lapu);

    T smag_v = tau_smag * sqrt(T(0.5) * (tt_jp + tt0) + T(0.5) * (ss_im + ss0)) - hdweight;
    smag_v = cxx_min(T(0.5), cxx_max(T(0.0), smag_v));
    const T v0 = v[0];
    const T lapv = v[1] + v[-1] - T(2.0) * v0 + vo * (v[sj] - v0) + vu * (v[-sj] - v0);
    st_nt(vout + at, v0 + smag_v * lapv);
  }
}

template <class T>
__global__ void __launch_bounds__(kTileI *kTileJ) hdiffsmag_kernel(HsArgs<T> a, Geo g, Tiles t) {
  hdiffsmag_body(a, g, t, int(blockIdx.x));
}

template <class T>
hipError_t hdiffsmag_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<HsArgs<T>, hdiffsmag_body<T>>(tab, n, total, s, dim3(kTileI, kTileJ));
}
template <class T>
const BatchKind kHdiffsmagBatch{"hdiffsmag", hdiffsmag_batch<T>};

// Row-march form, cut into wave segments.  A wave owns the 64 V columns
// i = c0 + V ln .. of one level (lane ln holds V of them) and walks the rows
// [j0, j1) of its chunk.  Rows j - 1, j, j + 1 of u and v are in registers (V
// wide, with the single neighbours u(i - 1), u(i + V), v(i - 1), v(i + V) of
// each row), so a row of each input is fetched once per chunk; the neighbour
// loads hit the lines the adjacent lanes fetch.  Waves share nothing, so four
// of them make a 256-lane workgroup and the same body runs in the plain launch
// and in a batched (persistent) one.  Rows are loaded D rows ahead; addresses
// of lanes and rows outside the segment are clamped into it (never stored).
// The expressions are those of hdiffsmag_body, term for term.  (Carrying the
// squares of T and S between rows and columns, so that a lane forms 2V + 2 of
// them per row instead of 6V, gave the same bits and no gain in FP64: 182.8
// against 181-183 us at 512x512x80 with the same 3 waves per SIMD, 189.6 when
// the compiler fitted it in 114 VGPRs for 4 waves.  Taking the single
// neighbours u(i - 1), u(i + V), v(i - 1), v(i + V) from the adjacent lanes
// with DPP wave shifts, loaded only by a segment's end lanes, fitted 122
// VGPRs (4 waves) and took 185.1-185.5 / 59.7-59.9 us against 176.7-178 /
// 51.8-52.7 at 512x512x80 / 256x256x80.)
// Segment layout in a Tiles: ti = segments per row, tj = row chunks,
// kchunk = rows per chunk, nchunk = levels, n = workgroups.
constexpr int kSegLanes = 64, kSegWaves = 4;

inline Tiles make_segments(int64_t nx, int64_t ny, int nz, int v, int rows) {
  Tiles t;
  t.ti = int((nx + int64_t(kSegLanes) * v - 1) / (int64_t(kSegLanes) * v));
  t.kchunk = rows < 1 ? 1 : rows;
  t.tj = int((ny + t.kchunk - 1) / t.kchunk);
  t.nchunk = nz;
  const int64_t segs = int64_t(t.ti) * t.tj * nz;
  t.n = int((segs + kSegWaves - 1) / kSegWaves);
  return t;
}

template <class T, int V>
struct SmagRow {
  typename Vec<T, V>::type u, v;
  T uw, ue, vw, ve;  // u(i - 1), u(i + V), v(i - 1), v(i + V)
};

template <class T, int V, int D>
__device__ __forceinline__ void hdiffsmag_seg_body(const HsArgs<T> &a, const Geo &g, const Tiles &t, int id) {
  typedef typename Vec<T, V>::type VT;
  typedef SmagRow<T, V> Row;
  constexpr T rr = kEarthRadiusRecip<T>;
  const int tid = int(threadIdx.x) + int(blockDim.x) * int(threadIdx.y);
  const int ln = tid & (kSegLanes - 1), seg = id * kSegWaves + tid / kSegLanes;
  if (seg >= t.ti * t.tj * t.nchunk) return;  // a whole wave; no barriers follow
  const int si = seg % t.ti, rest = seg / t.ti, cj = rest % t.tj, k = rest / t.tj;
  const int c0 = si * kSegLanes * V, nl = min(kSegLanes, (g.nx - c0 + V - 1) / V);
  const int j0 = cj * t.kchunk, j1 = min(g.ny, j0 + t.kchunk), sj = g.sj;
  const T eddlat = a.eddlat, eddlon = a.eddlon, tau_smag = a.tau_smag, weight_smag = a.weight_smag;
  const bool out_lane = ln < nl;
  const int64_t col = int64_t(k) * g.sk + (c0 + V * min(ln, nl - 1));
  auto ldrow = [&](int r) -> Row {
    const int64_t at = col + int64_t(min(r, j1)) * sj;  // u, v rows [j0 - 1, j1]
    const T *pu = a.uin + at;
    const T *pv = a.vin + at;
    Row x;
    x.u = ldv<T, V>(pu);
    x.v = ldv<T, V>(pv);
    x.uw = pu[-1];
    x.ue = pu[V];
    x.vw = pv[-1];
    x.ve = pv[V];
    return x;
  };
  auto ldmask = [&](int r) -> VT { return ldv_nt<T, V>(a.mask + col + int64_t(min(r, j1 - 1)) * sj); };
  Row rs = ldrow(j0 - 1), r0 = ldrow(j0);
  Row cr[D];  // row j + 1 of iteration t0 + u
  VT cm[D];   // mask row j of iteration t0 + u
#pragma unroll
  for (int u = 0; u < D; ++u) {
    cr[u] = ldrow(j0 + 1 + u);
    cm[u] = ldmask(j0 + u);
  }
  const int nt = j1 - j0;
  for (int t0 = 0; t0 < nt; t0 += D) {
    Row nr[D];
    VT nm[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      nr[u] = ldrow(j0 + 1 + D + t0 + u);
      nm[u] = ldmask(j0 + D + t0 + u);
    }
#pragma unroll
    for (int u = 0; u < D; ++u) {
      if (t0 + u < nt) {
        const int j = j0 + t0 + u;
        const Row rn = cr[u];
        const VT mk = cm[u];
        const T ac_s = a.acrlat0[int64_t(j - 1) * a.s_ac], ac_c = a.acrlat0[int64_t(j) * a.s_ac],
                ac_n = a.acrlat0[int64_t(j + 1) * a.s_ac];
        const T lo = a.crlato[int64_t(j) * a.s_to], lu = a.crlatu[int64_t(j) * a.s_tu];
        const T vo = a.crlavo[int64_t(j) * a.s_vo], vu = a.crlavu[int64_t(j) * a.s_vu];
        VT ro, so;
#pragma unroll
        for (int x = 0; x < V; ++x) {
          const T u_c = r0.u[x], u_w = x == 0 ? r0.uw : r0.u[x - 1], u_e = x == V - 1 ? r0.ue : r0.u[x + 1];
          const T u_n = rn.u[x], u_s = rs.u[x], u_nw = x == 0 ? rn.uw : rn.u[x - 1];
          const T v_c = r0.v[x], v_w = x == 0 ? r0.vw : r0.v[x - 1], v_e = x == V - 1 ? r0.ve : r0.v[x + 1];
          const T v_n = rn.v[x], v_s = rs.v[x], v_se = x == V - 1 ? rs.ve : rs.v[x + 1];
          const T t0v = (v_s - v_c) * eddlat * rr - (u_w - u_c) * ac_c * eddlon;
          const T t_ip = (v_se - v_e) * eddlat * rr - (u_c - u_e) * ac_c * eddlon;
          const T t_jp = (v_c - v_n) * eddlat * rr - (u_nw - u_n) * ac_n * eddlon;
          const T s0 = (u_n - u_c) * eddlat * rr + (v_e - v_c) * ac_c * eddlon;
          const T s_im = (u_nw - u_w) * eddlat * rr + (v_c - v_w) * ac_c * eddlon;
          const T s_jm = (u_c - u_s) * eddlat * rr + (v_se - v_s) * ac_s * eddlon;
          const T tt0 = t0v * t0v, tt_ip = t_ip * t_ip, tt_jp = t_jp * t_jp;
          const T ss0 = s0 * s0, ss_im = s_im * s_im, ss_jm = s_jm * s_jm;
          const T hdweight = weight_smag * mk[x];
          T smag_u = tau_smag * sqrt(T(0.5) * (tt_ip + tt0) + T(0.5) * (ss_jm + ss0)) - hdweight;
          smag_u = cxx_min(T(0.5), cxx_max(T(0.0), smag_u));
          const T lapu = u_e + u_w - T(2.0) * u_c + lo * (u_n - u_c) + lu * (u_s - u_c);
          ro[x] = u_c + smag_u * lapu;
          T smag_v = tau_smag * sqrt(T(0.5) * (tt_jp + tt0) + T(0.5) * (ss_im + ss0)) - hdweight;
          smag_v = cxx_min(T(0.5), cxx_max(T(0.0), smag_v));
          const T lapv = v_e + v_w - T(2.0) * v_c + vo * (v_n - v_c) + vu * (v_s - v_c);
          so[x] = v_c + smag_v * lapv;
        }
        if (out_lane) {
          const int64_t at = col + int64_t(j) * sj;
          stv_nt<T, V>(a.uout + at, ro);
          stv_nt<T, V>(a.vout + at, so);
        }
        rs = r0;
        r0 = rn;
      }
    }
#pragma unroll
    for (int u = 0; u < D; ++u) {
      cr[u] = nr[u];
      cm[u] = nm[u];
    }
  }
}

template <class T, int V, int D>
__global__ void __launch_bounds__(kSegLanes *kSegWaves) hdiffsmag_seg_kernel(HsArgs<T> a, Geo g, Tiles t) {
  hdiffsmag_seg_body<T, V, D>(a, g, t, int(blockIdx.x));
}

template <class T, int V, int D>
hipError_t hdiffsmag_seg_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<HsArgs<T>, hdiffsmag_seg_body<T, V, D>>(tab, n, total, s, dim3(kSegLanes * kSegWaves));
}
template <class T, int V, int D>
const BatchKind kHdiffsmagSegBatch{"hdiffsmag", hdiffsmag_seg_batch<T, V, D>};

// Plain launch or batch record of the segment form with V columns per lane.
template <class T, int V>
int hdiffsmag_seg_run(const HsArgs<T> &a, const Geo &g, const oeb_domain *d, const AliasSet &as,
                      hipStream_t stream) {
  const char *K = "hdiffsmag";
  // Rows per chunk by the lanes that cover one row (default_strip_rows), so
  // that narrow rows get short chunks and enough waves; at most 8.  In the
  // ten-stencil step at 512x512x80 (where the formula gives 32): 16 rows took
  // 184 us against 194 for 32 in FP64, and 8 rows 178-179 against 183 for 16
  // (FP32 94.7 against 95.7); 6 rows 180, 10 and 12 183.  At 256x256x80 the
  // formula's chunks are already 8 rows or fewer.
  const int row_lanes = make_segments(d->nx, d->ny, int(d->nz), V, 1).ti * kSegLanes;
  const int rows = knob(K, sizeof(T), d, KNOB_ROWS, std::min(8, default_strip_rows(d->ny, row_lanes)));
  const Tiles t = make_segments(d->nx, d->ny, int(d->nz), V, rows);
  if (t.n < 1 || int64_t(t.ti) * t.tj * t.nchunk > (int64_t(1) << 30))
    return fail(OEB_ERR_DOMAIN, std::string(K) + ": domain too large");
  if (knob(K, sizeof(T), d, KNOB_DEPTH, 1) >= 2) {
    if (batch_record(kHdiffsmagSegBatch<T, V, 2>, a, g, t, t.n, as)) return OEB_OK;
    hipLaunchKernelGGL((hdiffsmag_seg_kernel<T, V, 2>), dim3(t.n), dim3(kSegLanes * kSegWaves), 0, stream, a, g, t);
  } else {
    if (batch_record(kHdiffsmagSegBatch<T, V, 1>, a, g, t, t.n, as)) return OEB_OK;
    hipLaunchKernelGGL((hdiffsmag_seg_kernel<T, V, 1>), dim3(t.n), dim3(kSegLanes * kSegWaves), 0, stream, a, g, t);
  }
  return launch_status(K);
}

template <class T, class S3, class S1>
int hdiffsmag_run(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin, const S3 *mask,
                  const S1 *crlavo, const S1 *crlavu, const S1 *crlato, const S1 *crlatu, const S1 *acrlat0,
                  T eddlat, T eddlon, T tau_smag, T weight_smag, void *stream) {
  const char *K = "hdiffsmag";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_uv{-1, nx + 1, -1, ny + 1, 0, nz};
  const T *p_uo, *p_vo, *p_ui, *p_vi, *p_m, *p_vo1, *p_vu1, *p_to1, *p_tu1, *p_ac;
  int s_vo = 1, s_vu = 1, s_to = 1, s_tu = 1, s_ac = 1;
  OEB_TRY(check3(K, "uout", uout, r0, g, true, &p_uo));
  OEB_TRY(check3(K, "vout", vout, r0, g, false, &p_vo));
  OEB_TRY(check3(K, "uin", uin, r_uv, g, false, &p_ui));
  OEB_TRY(check3(K, "vin", vin, r_uv, g, false, &p_vi));
  OEB_TRY(check3(K, "mask", mask, r0, g, false, &p_m));
  OEB_TRY(check1(K, "crlavo", crlavo, 0, ny, &p_vo1, &s_vo));
  OEB_TRY(check1(K, "crlavu", crlavu, 0, ny, &p_vu1, &s_vu));
  OEB_TRY(check1(K, "crlato", crlato, 0, ny, &p_to1, &s_to));
  OEB_TRY(check1(K, "crlatu", crlatu, 0, ny, &p_tu1, &s_tu));
  OEB_TRY(check1(K, "acrlat0", acrlat0, -1, ny + 1, &p_ac, &s_ac));
  AliasSet as;
  as.add("uout", p_uo, r0, g, true);
  as.add("vout", p_vo, r0, g, true);
  as.add("uin", p_ui, r_uv, g, false);
  as.add("vin", p_vi, r_uv, g, false);
  as.add("mask", p_m, r0, g, false);
  as.add1("crlavo", p_vo1, 0, ny, s_vo);
  as.add1("crlavu", p_vu1, 0, ny, s_vu);
  as.add1("crlato", p_to1, 0, ny, s_to);
  as.add1("crlatu", p_tu1, 0, ny, s_tu);
  as.add1("acrlat0", p_ac, -1, ny + 1, s_ac);
  OEB_TRY(as.check(K));
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, knob(K, sizeof(T), d, KNOB_KCHUNK, kChunk));
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  const HsArgs<T> a{const_cast<T *>(p_uo), const_cast<T *>(p_vo), p_ui, p_vi, p_m, p_vo1, p_vu1, p_to1, p_tu1, p_ac,
                    s_vo, s_vu, s_to, s_tu, s_ac, eddlat, eddlon, tau_smag, weight_smag};
  if (knob_on(K, sizeof(T), d, KNOB_FORM, true)) {
    constexpr int W = 16 / int(sizeof(T));
    if ((knob(K, sizeof(T), d, KNOB_VECTOR, true) != 0) && vec_ok(g, W, sizeof(T), {p_uo, p_vo, p_ui, p_vi, p_m}))
      return hdiffsmag_seg_run<T, W>(a, g, d, as, hipStream_t(stream));
    return hdiffsmag_seg_run<T, 1>(a, g, d, as, hipStream_t(stream));
  }
  if (batch_record(kHdiffsmagBatch<T>, a, g, t, t.n, as)) return OEB_OK;
  hipLaunchKernelGGL(hdiffsmag_kernel<T>, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a, g, t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

// T_sqr_s and S_sqr_uv are the reference's scratch arrays; nothing is stored in them.
extern "C" int oeb_hdiffsmag_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                                 const oeb_s3 *vin, const oeb_s3 *mask, const oeb_s1 *crlavo, const oeb_s1 *crlavu,
                                 const oeb_s1 *crlato, const oeb_s1 *crlatu, const oeb_s1 *acrlat0,
                                 const oeb_s3 *T_sqr_s, const oeb_s3 *S_sqr_uv, double eddlat, double eddlon,
                                 double tau_smag, double weight_smag, void *stream) {
  (void)T_sqr_s;
  (void)S_sqr_uv;
  return oeb::hdiffsmag_run<double>(d, uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, eddlat,
                                    eddlon, tau_smag, weight_smag, stream);
}

extern "C" int oeb_hdiffsmag_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout,
                                 const oeb_s3f *uin, const oeb_s3f *vin, const oeb_s3f *mask,
                                 const oeb_s1f *crlavo, const oeb_s1f *crlavu, const oeb_s1f *crlato,
                                 const oeb_s1f *crlatu, const oeb_s1f *acrlat0, const oeb_s3f *T_sqr_s,
                                 const oeb_s3f *S_sqr_uv, float eddlat, float eddlon, float tau_smag,
                                 float weight_smag, void *stream) {
  (void)T_sqr_s;
  (void)S_sqr_uv;
  return oeb::hdiffsmag_run<float>(d, uout, vout, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, eddlat,
                                   eddlon, tau_smag, weight_smag, stream);
}
