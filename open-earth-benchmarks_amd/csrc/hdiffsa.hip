// hdiffsa.hip -- COSMO horizontal diffusion with flux limiter, fused.
// Reference: kernels/hdiffsa.h:141-169 (hdiffsa_openmp).  Per level the
// reference makes three sweeps through scratch arrays:
//   lap (:147-149) -> flx, fly with limiters (:152-159) -> out (:164-165).
// Here a lane evaluates, for its point, the five Laplacians it depends on:
// centre and its four neighbours.  From them it forms the four limited fluxes
// across its cell faces, then writes out.  Nothing round-trips through
// memory.  Each Laplacian is evaluated with the reference expression, so a
// value shared between neighbouring points is the same number in both.
// Reach (2,2,0).  Algorithmic traffic: 2 inputs + 1 output = 24 B per point
// (FP64), 12 B (FP32).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 16;

// Laplacian at the point p points to, with that row's crlato / crlatu
// (kernels/hdiffsa.h:147).
template <class T>
__device__ __forceinline__ T lap_at(const T *__restrict__ p, int sj, T lo, T lu) {
  const T c = p[0];
  return p[-1] + p[1] - T(2.0) * c + lo * (p[sj] - c) + lu * (p[-sj] - c);
}

template <class T>
__global__ void __launch_bounds__(kTileI *kTileJ)
    hdiffsa_kernel(T *__restrict__ out, const T *__restrict__ in, const T *__restrict__ mask,
                   const T *__restrict__ crlato, int so, const T *__restrict__ crlatu, int su, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const T lo_s = crlato[int64_t(j - 1) * so], lo_c = crlato[int64_t(j) * so], lo_n = crlato[int64_t(j + 1) * so];
  const T lu_s = crlatu[int64_t(j - 1) * su], lu_c = crlatu[int64_t(j) * su], lu_n = crlatu[int64_t(j + 1) * su];
  int64_t at = int64_t(c.k0) * g.sk + (i + j * sj);
  for (int k = c.k0; k < c.k1; ++k, at += g.sk) {
    const T *p = in + at;
    const T lap_c = lap_at<T>(p, sj, lo_c, lu_c);
    const T lap_w = lap_at<T>(p - 1, sj, lo_c, lu_c);
    const T lap_e = lap_at<T>(p + 1, sj, lo_c, lu_c);
    const T lap_s = lap_at<T>(p - sj, sj, lo_s, lu_s);
    const T lap_n = lap_at<T>(p + sj, sj, lo_n, lu_n);
    const T q0 = p[0];
    // flx(i-1,j), flx(i,j), fly(i,j-1), fly(i,j) with the limiter of :153-158.
    T flx_w = lap_c - lap_w;
    if (flx_w * (q0 - p[-1]) > 0) flx_w = T(0);
    T flx_e = lap_e - lap_c;
    if (flx_e * (p[1] - q0) > 0) flx_e = T(0);
    T fly_s = lo_s * (lap_c - lap_s);
    if (fly_s * (q0 - p[-sj]) > 0) fly_s = T(0);
    T fly_n = lo_c * (lap_n - lap_c);
    if (fly_n * (p[sj] - q0) > 0) fly_n = T(0);
    st_nt(out + at, q0 + (flx_w - flx_e + fly_s - fly_n) * ld_nt(mask + at));
  }
}

// Strip form (Strips, h = 2): lane l owns column i = i0 - 2 + l.  Iteration t
// handles row r = j0 - 1 + t: it publishes in(r) to LDS, forms lap(r) on
// [i0 - 1, i0 + w + 1) from in(r) and its row neighbours (registers) and its
// i neighbours (LDS), publishes lap(r), and from t = 2 on writes out(r - 1),
// whose lap(r - 2), lap(r - 1), lap(r) and in(r - 2 .. r) are all at hand.
// LDS holds 3 rows of in and 2 of lap; each row slot is rewritten only after
// the barrier that follows its last read, so one barrier per row suffices.
// Input and mask rows are loaded kDepth iterations ahead into register rings
// (the loop is unrolled by kDepth so that ring slots are registers).  The
// expressions are those of hdiffsa_kernel, term for term.
template <class T, int kDepth>
__global__ void __launch_bounds__(1024)
    hdiffsa_strip_kernel(T *__restrict__ out, const T *__restrict__ in, const T *__restrict__ mask,
                         const T *__restrict__ crlato, int so, const T *__restrict__ crlatu, int su, Geo g,
                         Strips st) {
  extern __shared__ __align__(16) unsigned char oeb_smem[];
  T *s_in = reinterpret_cast<T *>(oeb_smem);  // 3 x lanes
  T *s_lap = s_in + 3 * st.lanes;              // 2 x lanes
  const StripCoord c = strip_coord(st, int(blockIdx.x), g.nx, g.ny);
  const int l = int(threadIdx.x), L = st.lanes, sj = g.sj;
  const bool lap_lane = l >= 1 && l < c.w + 3;  // lap on [i0 - 1, i0 + w + 1)
  const bool out_lane = l >= 2 && l < c.w + 2;  // out on [i0, i0 + w)
  const int64_t col = int64_t(c.k) * g.sk + (c.i0 - 2 + l);
  const int r0 = c.j0 - 1, last_in = c.j1 + 1;  // in rows [r0 - 1, last_in]
  // Loads are unconditional, with the address clamped into the array's valid
  // box: a masked load (cond ? load : 0) makes the select wait for the load,
  // which would drain the prefetch ring.  Values of lanes or rows outside the
  // strip are never used.
  const int64_t col_in = int64_t(c.k) * g.sk + (c.i0 - 2 + min(l, c.w + 3));
  const int64_t col_mk = int64_t(c.k) * g.sk + (c.i0 + min(max(l - 2, 0), c.w - 1));
  auto ld_in = [&](int row) -> T { return in[col_in + int64_t(min(row, last_in)) * sj]; };
  auto ld_mask = [&](int row) -> T { return ld_nt(mask + col_mk + int64_t(min(max(row, c.j0), c.j1 - 1)) * sj); };
  // Rows are prefetched a block of kDepth iterations ahead: while block t0
  // runs on cin/cmk (in(r + 1) and mask(r - 1) of iteration t0 + u), the
  // loads of block t0 + kDepth are in flight in nin/nmk.
  T cin[kDepth], cmk[kDepth];
#pragma unroll
  for (int u = 0; u < kDepth; ++u) {
    cin[u] = ld_in(r0 + 1 + u);
    cmk[u] = ld_mask(r0 - 1 + u);
  }
  T x_m2 = T(0), x_m1 = ld_in(r0 - 1), x_0 = ld_in(r0);
  T lap_s = T(0), lap_c = T(0);
  const int nt = c.j1 - c.j0 + 2;
  for (int t0 = 0; t0 < nt; t0 += kDepth) {
    T nin[kDepth], nmk[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      nin[u] = ld_in(r0 + 1 + kDepth + t0 + u);
      nmk[u] = ld_mask(r0 - 1 + kDepth + t0 + u);
    }
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int t = t0 + u;
      if (t < nt) {  // uniform across the workgroup
        const int r = r0 + t;
        const T x_p1 = cin[u], m_0 = cmk[u];
        T *si = s_in + (t % 3) * L + l;
        si[0] = x_0;
        lds_barrier();
        T lap_n = T(0);
        if (lap_lane) {
          const T lo = crlato[int64_t(r) * so], lu = crlatu[int64_t(r) * su];
          lap_n = si[-1] + si[1] - T(2.0) * x_0 + lo * (x_p1 - x_0) + lu * (x_m1 - x_0);
        }
        s_lap[(t & 1) * L + l] = lap_n;
        if (t >= 2 && out_lane) {
          const int j = r - 1;
          const T *ri = s_in + ((t - 1) % 3) * L + l;   // in(j)
          const T *rl = s_lap + ((t - 1) & 1) * L + l;  // lap(j)
          const T lo_s = crlato[int64_t(j - 1) * so], lo_c = crlato[int64_t(j) * so];
          const T q0 = x_m1;
          T flx_w = lap_c - rl[-1];
          if (flx_w * (q0 - ri[-1]) > 0) flx_w = T(0);
          T flx_e = rl[1] - lap_c;
          if (flx_e * (ri[1] - q0) > 0) flx_e = T(0);
          T fly_s = lo_s * (lap_c - lap_s);
          if (fly_s * (q0 - x_m2) > 0) fly_s = T(0);
          T fly_n = lo_c * (lap_n - lap_c);
          if (fly_n * (x_0 - q0) > 0) fly_n = T(0);
          st_nt(out + col + int64_t(j) * sj, q0 + (flx_w - flx_e + fly_s - fly_n) * m_0);
        }
        x_m2 = x_m1;
        x_m1 = x_0;
        x_0 = x_p1;
        lap_s = lap_c;
        lap_c = lap_n;
      }
    }
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      cin[u] = nin[u];
      cmk[u] = nmk[u];
    }
  }
}

template <class T, class S3, class S1>
int hdiffsa_run(const oeb_domain *d, const S3 *out, const S3 *in, const S3 *mask, const S1 *crlato,
                const S1 *crlatu, void *stream) {
  const char *K = "hdiffsa";
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_in{-2, nx + 2, -2, ny + 2, 0, nz};
  const T *p_out, *p_in, *p_mask, *p_lo, *p_lu;
  int s_lo = 1, s_lu = 1;
  OEB_TRY(check3(K, "out", out, r0, g, true, &p_out));
  OEB_TRY(check3(K, "in", in, r_in, g, false, &p_in));
  OEB_TRY(check3(K, "mask", mask, r0, g, false, &p_mask));
  OEB_TRY(check1(K, "crlato", crlato, -1, ny + 1, &p_lo, &s_lo));
  OEB_TRY(check1(K, "crlatu", crlatu, -1, ny + 1, &p_lu, &s_lu));
  AliasSet as;
  as.add("out", p_out, r0, g, true);
  as.add("in", p_in, r_in, g, false);
  as.add("mask", p_mask, r0, g, false);
  as.add1("crlato", p_lo, -1, ny + 1, s_lo);
  as.add1("crlatu", p_lu, -1, ny + 1, s_lu);
  OEB_TRY(as.check(K));
  if (knob_on(K, sizeof(T), d, KNOB_FORM, true)) {
    // At most 16 rows per chunk, except FP32 on levels of more than 256x256
    // points: measured in the ten-stencil step against the 20 (256x256x80)
    // and 32 (512x512x80) rows of default_strip_rows, FP64 34 / 106 us against
    // 37 / 109, FP32 23 / 64 against 26 / 62.  Re-measured for FP32: 32 rows
    // 61.2-61.4 against 62.5-62.6 at 512x512x80 (12, 20, 24 rows 65.1-66.6),
    // 26.8 against 20.8 at 256x256x80.
    Strips st = make_strips(nx, ny, int(nz), 2, 1);
    const int rows_cap = sizeof(T) == 4 && int64_t(nx) * ny > 256 * 256 ? 32 : 16;
    st = make_strips(nx, ny, int(nz), 2,
                     knob(K, sizeof(T), d, KNOB_ROWS, std::min(rows_cap, default_strip_rows(ny, st.lanes))));
    OEB_TRY(check_strips(K, st));
    // Measured at 256x256x80 and 512x512x80: prefetching 2 rows ahead is
    // fastest in FP64 (4 and 8 rows: 1.2x and 1.8x slower), 4 in FP32.
    const int depth = knob(K, sizeof(T), d, KNOB_DEPTH, sizeof(T) == 8 ? 2 : 4);
    auto kern = depth <= 2 ? hdiffsa_strip_kernel<T, 2> : depth <= 4 ? hdiffsa_strip_kernel<T, 4>
                                                                     : hdiffsa_strip_kernel<T, 8>;
    hipLaunchKernelGGL(kern, dim3(unsigned(st.n)), dim3(unsigned(st.lanes)), size_t(5) * st.lanes * sizeof(T),
                       hipStream_t(stream), const_cast<T *>(p_out), p_in, p_mask, p_lo, s_lo, p_lu, s_lu, g, st);
    return launch_status(K);
  }
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, knob(K, sizeof(T), d, KNOB_KCHUNK, kChunk));
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  hipLaunchKernelGGL(hdiffsa_kernel<T>, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream),
                     const_cast<T *>(p_out), p_in, p_mask, p_lo, s_lo, p_lu, s_lu, g, t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

// lap, flx and fly are the reference's scratch arrays; nothing is stored in them.
extern "C" int oeb_hdiffsa_f64(const oeb_domain *d, const oeb_s3 *out, const oeb_s3 *in, const oeb_s3 *mask,
                               const oeb_s1 *crlato, const oeb_s1 *crlatu, const oeb_s3 *lap, const oeb_s3 *flx,
                               const oeb_s3 *fly, void *stream) {
  (void)lap;
  (void)flx;
  (void)fly;
  return oeb::hdiffsa_run<double>(d, out, in, mask, crlato, crlatu, stream);
}

extern "C" int oeb_hdiffsa_f32(const oeb_domain *d, const oeb_s3f *out, const oeb_s3f *in, const oeb_s3f *mask,
                               const oeb_s1f *crlato, const oeb_s1f *crlatu, const oeb_s3f *lap,
                               const oeb_s3f *flx, const oeb_s3f *fly, void *stream) {
  (void)lap;
  (void)flx;
  (void)fly;
  return oeb::hdiffsa_run<float>(d, out, in, mask, crlato, crlatu, stream);
}
