// laplace.hip -- 5-point horizontal Laplacian.
// Reference: kernels/laplace.h:34-43 (laplace_openmp).  Expression :39,
// parenthesization kept: out = -4*in + ((in[i-1] + in[i+1]) + (in[j+1] + in[j-1])).
// -4.0 is a double literal there, so with float storage the product and the
// final sum are formed in double and rounded once on the store.
// Algorithmic traffic: 1 input + 1 output field = 16 B per point (FP64), 8 B (FP32).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 1;

template <class T>
__device__ __forceinline__ T lap_point(T c, T w, T e, T n, T s) {
  return T(-4.0 * c + ((w + e) + (n + s)));
}

template <class T>
__global__ void __launch_bounds__(kTileI *kTileJ)
    laplace_kernel(const T *__restrict__ in, T *__restrict__ out, Geo g, Tiles t) {
  const TileCoord c = tile_coord(t, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t at = int64_t(c.k0) * g.sk + (i + j * sj);
  const T *p = in + at;
  T *q = out + at;
  for (int k = c.k0; k < c.k1; ++k) {
    st_nt(q, lap_point<T>(p[0], p[-1], p[1], p[sj], p[-sj]));
    p += g.sk;
    q += g.sk;
  }
}

// W adjacent columns (i .. i+W-1) per lane, i a multiple of W: 16-byte loads
// and stores (W = 2 doubles or 4 floats).  Same arithmetic as laplace_kernel:
// in[i-1] and in[i+W] are single loads and the inner neighbours come from the
// vector itself.  At 256x256x80 and 512x512x80 the FP64 form measured 9-11%
// faster this way.  Taking in[i-1] and in[i+W] from the neighbouring lanes
// with DPP wave shifts (loaded only by a row's end lanes) took 67.5-68.0 /
// 20.1 us in FP64 and 35.6-35.8 / 11.9-12.0 in FP32 against 58.9 / 17.6 and
// 32 / 10.5 at 512x512x80 / 256x256x80 (the single loads hit the lines of
// the vector loads, so they cost little).
template <class T>
constexpr int kVecW = 16 / int(sizeof(T));

template <class T>
__global__ void __launch_bounds__(kTileI *kTileJ)
    laplace_vec_kernel(const T *__restrict__ in, T *__restrict__ out, Geo g, Tiles t) {
  constexpr int W = kVecW<T>;
  const TileCoord c = tile_coord(t, W * kTileI, kTileJ, g.nz);
  const int i = c.i0 + W * int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const int64_t at = int64_t(c.k0) * g.sk + (i + j * sj);
  const T *p = in + at;
  T *q = out + at;
  for (int k = c.k0; k < c.k1; ++k) {
    const auto cc = ldv<T, W>(p), n = ldv<T, W>(p + sj), s = ldv<T, W>(p - sj);
    const T w = p[-1], e = p[W];
    typename Vec<T, W>::type r;
#pragma unroll
    for (int x = 0; x < W; ++x) {
      r[x] = lap_point<T>(cc[x], x == 0 ? w : cc[x - 1], x == W - 1 ? e : cc[x + 1], n[x], s[x]);
    }
    stv_nt<T, W>(q, r);
    p += g.sk;
    q += g.sk;
  }
}

// Row-march form (Strips, no halo lanes): lane l owns the V columns
// i = i0 + V l .. i + V - 1 of one level and walks the rows [j0, j1) of its
// chunk.  Row j - 1 and j come from registers, row j + 1 is loaded (V wide),
// and the i neighbours in(i - 1, j), in(i + V, j) are single loads that hit
// the lines the adjacent lanes fetch.  A row of the input is thus fetched once
// per chunk (plus the two halo rows).  Loads run D rows ahead; addresses of
// lanes and rows outside the strip are clamped into it (never stored).  Same
// arithmetic as laplace_kernel.
template <class T, int V, int D>
__global__ void __launch_bounds__(1024)
    laplace_strip_kernel(const T *__restrict__ in, T *__restrict__ out, Geo g, Strips st) {
  typedef typename Vec<T, V>::type VT;
  const StripCoord c = strip_coord(st, int(blockIdx.x), g.nx, g.ny);
  const int l = int(threadIdx.x), nl = c.w / V, sj = g.sj;
  const bool out_lane = l < nl;
  const int64_t col = int64_t(c.k) * g.sk + (c.i0 + V * min(l, nl - 1));
  auto ldr = [&](int row) -> VT { return ldv<T, V>(in + col + int64_t(min(row, c.j1)) * sj); };
  auto ldw = [&](int row) -> T { return in[col + int64_t(min(row, c.j1 - 1)) * sj - 1]; };
  auto lde = [&](int row) -> T { return in[col + int64_t(min(row, c.j1 - 1)) * sj + V]; };
  VT x_m1 = ldr(c.j0 - 1), x_0 = ldr(c.j0);
  VT cin[D];
  T cw[D], ce[D];
#pragma unroll
  for (int u = 0; u < D; ++u) {
    cin[u] = ldr(c.j0 + 1 + u);
    cw[u] = ldw(c.j0 + u);
    ce[u] = lde(c.j0 + u);
  }
  const int nt = c.j1 - c.j0;
  for (int t0 = 0; t0 < nt; t0 += D) {
    VT nin[D];
    T nw[D], ne[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      nin[u] = ldr(c.j0 + 1 + D + t0 + u);
      nw[u] = ldw(c.j0 + D + t0 + u);
      ne[u] = lde(c.j0 + D + t0 + u);
    }
#pragma unroll
    for (int u = 0; u < D; ++u) {
      if (t0 + u < nt) {
        const VT x_p1 = cin[u];
        VT r;
#pragma unroll
        for (int x = 0; x < V; ++x)
          r[x] = lap_point<T>(x_0[x], x == 0 ? cw[u] : x_0[x - 1], x == V - 1 ? ce[u] : x_0[x + 1], x_p1[x],
                              x_m1[x]);
        if (out_lane) stv_nt<T, V>(out + col + int64_t(c.j0 + t0 + u) * sj, r);
        x_m1 = x_0;
        x_0 = x_p1;
      }
    }
#pragma unroll
    for (int u = 0; u < D; ++u) {
      cin[u] = nin[u];
      cw[u] = nw[u];
      ce[u] = ne[u];
    }
  }
}

template <class T, int V>
int laplace_strip_launch(const T *p_in, T *p_out, const Geo &g, const oeb_domain *d, hipStream_t stream) {
  const char *K = "laplace";
  Strips st = make_strips(d->nx, d->ny, int(d->nz), 0, 1, V);
  st = make_strips(d->nx, d->ny, int(d->nz), 0, knob(K, sizeof(T), d, KNOB_ROWS, default_strip_rows(d->ny, st.lanes)), V);
  OEB_TRY(check_strips(K, st));
  const int depth = knob(K, sizeof(T), d, KNOB_DEPTH, 2);
  auto kern = depth <= 1 ? laplace_strip_kernel<T, V, 1>
                         : (depth <= 2 ? laplace_strip_kernel<T, V, 2> : laplace_strip_kernel<T, V, 4>);
  hipLaunchKernelGGL(kern, dim3(unsigned(st.n)), dim3(unsigned(st.lanes)), 0, stream, p_in, p_out, g, st);
  return launch_status(K);
}

template <class T, class S>
int laplace_run(const oeb_domain *d, const S *in, const S *out, void *stream) {
  const char *K = "laplace";
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r_in{-1, nx + 1, -1, ny + 1, 0, nz}, r_out{0, nx, 0, ny, 0, nz};
  const T *p_in, *p_out;
  OEB_TRY(check3(K, "in", in, r_in, g, true, &p_in));
  OEB_TRY(check3(K, "out", out, r_out, g, false, &p_out));
  AliasSet as;
  as.add("in", p_in, r_in, g, false);
  as.add("out", p_out, r_out, g, true);
  OEB_TRY(as.check(K));
  constexpr int W = kVecW<T>;
  // The row-march form is off by default: in the ten-stencil step (its input
  // not in the MALL) the tile form measured 1.15x faster at 256x256x80 and
  // 512x512x80; run back to back on one field (input in the MALL) the strip
  // form measured 1.1-1.2x faster.  form knob 1 selects it.
  if (knob_on(K, sizeof(T), d, KNOB_FORM, false)) {
    if ((knob(K, sizeof(T), d, KNOB_VECTOR, true) != 0) && vec_ok(g, W, sizeof(T), {p_in, p_out}))
      return laplace_strip_launch<T, W>(p_in, const_cast<T *>(p_out), g, d, hipStream_t(stream));
    return laplace_strip_launch<T, 1>(p_in, const_cast<T *>(p_out), g, d, hipStream_t(stream));
  }
  const int kc = knob(K, sizeof(T), d, KNOB_KCHUNK, kChunk);
  // Level chunks interleaved so that a tile and the tile below it run on one
  // XCD (Tiles::xcd): measured in the ten-stencil step, 1.05-1.26x faster
  // at 256x256x80 (from run to run) and 1.11x at 512x512x80 in FP64,
  // 1.05-1.16x in FP32.  The other tiled stencils measured no gain from it.
  // order knob 0 turns it off.
  // Tiles::irow (order knob 2) measured 1.04x slower.
  const int xo = knob(K, sizeof(T), d, KNOB_ORDER, 1);
  if ((knob(K, sizeof(T), d, KNOB_VECTOR, true) != 0) && vec_ok(g, W, sizeof(T), {p_in, p_out})) {
    const Tiles t = make_tiles(nx, ny, int(nz), W * kTileI, kTileJ, kc, xo);
    OEB_TRY(check_tiles(K, t, nx, ny, int(nz), W * kTileI, kTileJ));
    hipLaunchKernelGGL(laplace_vec_kernel<T>, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), p_in,
                       const_cast<T *>(p_out), g, t);
    return launch_status(K);
  }
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kc, xo);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  hipLaunchKernelGGL(laplace_kernel<T>, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), p_in,
                     const_cast<T *>(p_out), g, t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_laplace_f64(const oeb_domain *d, const oeb_s3 *in, const oeb_s3 *out, void *stream) {
  return oeb::laplace_run<double>(d, in, out, stream);
}

extern "C" int oeb_laplace_f32(const oeb_domain *d, const oeb_s3f *in, const oeb_s3f *out, void *stream) {
  return oeb::laplace_run<float>(d, in, out, stream);
}
