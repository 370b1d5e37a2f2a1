// uvbke.hip -- FV3 C-grid to B-grid wind.
// Reference: kernels/uvbke.h:61-75 (uvbke_openmp), expressions :67-71:
//   ub = ((dt5 * ((uc[j-1] + uc) - ((vc[i-1] + vc) * cosa))) * rsina)
//   vb = ((dt5 * ((vc[i-1] + vc) - ((uc[j-1] + uc) * cosa))) * rsina)
// dt5 is the driver's constant 0.5 * 225 (src/uvbke.cpp:11-12), an argument
// here.  It is a double whatever the element type, so with float storage the
// two outer products are formed in double and rounded once on the store.
// Algorithmic traffic: 4 inputs + 2 outputs = 48 B per point (FP64), 24 B (FP32).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTileI = 64, kTileJ = 4, kChunk = 1;

template <class T>
struct UvArgs {
  T *ub, *vb;
  const T *uc, *vc, *cosa, *rsina;
  double dt5;
};

// Logical block `id` of a launch: one 64 x 4 tile of columns over a k-chunk.
template <class T>
__device__ __forceinline__ void uvbke_body(const UvArgs<T> &a, const Geo &g, const Tiles &t, int id) {
  const TileCoord c = tile_coord_of(t, id, kTileI, kTileJ, g.nz);
  const int i = c.i0 + int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const double dt5 = a.dt5;
  int64_t at = int64_t(c.k0) * g.sk + (i + j * sj);
  for (int k = c.k0; k < c.k1; ++k, at += g.sk) {
    const T usum = a.uc[at - sj] + a.uc[at];
    const T vsum = a.vc[at - 1] + a.vc[at];
    const T ca = ld_nt(a.cosa + at), rs = ld_nt(a.rsina + at);
    st_nt(a.ub + at, T((dt5 * (usum - (vsum * ca))) * rs));
    st_nt(a.vb + at, T((dt5 * (vsum - (usum * ca))) * rs));
  }
}

template <class T>
__global__ void __launch_bounds__(kTileI *kTileJ) uvbke_kernel(UvArgs<T> a, Geo g, Tiles t) {
  uvbke_body(a, g, t, int(blockIdx.x));
}

template <class T>
hipError_t uvbke_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<UvArgs<T>, uvbke_body<T>>(tab, n, total, s, dim3(kTileI, kTileJ));
}
template <class T>
const BatchKind kUvbkeBatch{"uvbke", uvbke_batch<T>};

// W adjacent columns (i .. i+W-1) per lane, i a multiple of W: 16-byte
// accesses (W = 2 doubles or 4 floats), the arithmetic of uvbke_kernel.  The
// west neighbour vc[i-1] of the first column is a single load; the others come
// from the vector.
template <class T>
constexpr int kVecW = 16 / int(sizeof(T));

template <class T>
__device__ __forceinline__ void uvbke_vec_body(const UvArgs<T> &a, const Geo &g, const Tiles &t, int id) {
  constexpr int W = kVecW<T>;
  const TileCoord c = tile_coord_of(t, id, W * kTileI, kTileJ, g.nz);
  const int i = c.i0 + W * int(threadIdx.x), j = c.j0 + int(threadIdx.y);
  if (i >= g.nx || j >= g.ny) return;
  const int sj = g.sj;
  const double dt5 = a.dt5;
  int64_t at = int64_t(c.k0) * g.sk + (i + j * sj);
  for (int k = c.k0; k < c.k1; ++k, at += g.sk) {
    const auto us = ldv<T, W>(a.uc + at - sj), u0 = ldv<T, W>(a.uc + at), v0 = ldv<T, W>(a.vc + at);
    const T vw = a.vc[at - 1];
    const auto ca = ldv_nt<T, W>(a.cosa + at), rs = ldv_nt<T, W>(a.rsina + at);
    typename Vec<T, W>::type ro, rv;
#pragma unroll
    for (int x = 0; x < W; ++x) {
      const T usum = us[x] + u0[x];
      const T vsum = (x == 0 ? vw : v0[x - 1]) + v0[x];
      ro[x] = T((dt5 * (usum - (vsum * ca[x]))) * rs[x]);
      rv[x] = T((dt5 * (vsum - (usum * ca[x]))) * rs[x]);
    }
    stv_nt<T, W>(a.ub + at, ro);
    stv_nt<T, W>(a.vb + at, rv);
  }
}

template <class T>
__global__ void __launch_bounds__(kTileI *kTileJ) uvbke_vec_kernel(UvArgs<T> a, Geo g, Tiles t) {
  uvbke_vec_body(a, g, t, int(blockIdx.x));
}

template <class T>
hipError_t uvbke_vec_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<UvArgs<T>, uvbke_vec_body<T>>(tab, n, total, s, dim3(kTileI, kTileJ));
}
template <class T>
const BatchKind kUvbkeVecBatch{"uvbke", uvbke_vec_batch<T>};

template <class T, class S>
int uvbke_run(const oeb_domain *d, const S *ub, const S *vb, const S *uc, const S *vc, const S *cosa,
              const S *rsina, double dt5, void *stream) {
  const char *K = "uvbke";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_uc{0, nx, -1, ny, 0, nz}, r_vc{-1, nx, 0, ny, 0, nz};
  const T *p_ub, *p_vb, *p_uc, *p_vc, *p_ca, *p_rs;
  OEB_TRY(check3(K, "ub", ub, r0, g, true, &p_ub));
  OEB_TRY(check3(K, "vb", vb, r0, g, false, &p_vb));
  OEB_TRY(check3(K, "uc", uc, r_uc, g, false, &p_uc));
  OEB_TRY(check3(K, "vc", vc, r_vc, g, false, &p_vc));
  OEB_TRY(check3(K, "cosa", cosa, r0, g, false, &p_ca));
  OEB_TRY(check3(K, "rsina", rsina, r0, g, false, &p_rs));
  AliasSet as;
  as.add("ub", p_ub, r0, g, true);
  as.add("vb", p_vb, r0, g, true);
  as.add("uc", p_uc, r_uc, g, false);
  as.add("vc", p_vc, r_vc, g, false);
  as.add("cosa", p_ca, r0, g, false);
  as.add("rsina", p_rs, r0, g, false);
  OEB_TRY(as.check(K));
  const int kc = knob(K, sizeof(T), d, KNOB_KCHUNK, kChunk);
  // Tile order (the order knob): rows of tiles in order; Tiles::irow measured the same.
  const int xo = knob(K, sizeof(T), d, KNOB_ORDER, 0);
  const UvArgs<T> a{const_cast<T *>(p_ub), const_cast<T *>(p_vb), p_uc, p_vc, p_ca, p_rs, dt5};
  constexpr int W = kVecW<T>;
  if ((knob(K, sizeof(T), d, KNOB_VECTOR, true) != 0) && vec_ok(g, W, sizeof(T), {p_ub, p_vb, p_uc, p_vc, p_ca, p_rs})) {
    const Tiles t = make_tiles(nx, ny, int(nz), W * kTileI, kTileJ, kc, xo);
    OEB_TRY(check_tiles(K, t, nx, ny, int(nz), W * kTileI, kTileJ));
    if (batch_record(kUvbkeVecBatch<T>, a, g, t, t.n, as)) return OEB_OK;
    hipLaunchKernelGGL(uvbke_vec_kernel<T>, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a, g, t);
    return launch_status(K);
  }
  const Tiles t = make_tiles(nx, ny, int(nz), kTileI, kTileJ, kc, xo);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTileI, kTileJ));
  if (batch_record(kUvbkeBatch<T>, a, g, t, t.n, as)) return OEB_OK;
  hipLaunchKernelGGL(uvbke_kernel<T>, dim3(t.n), dim3(kTileI, kTileJ), 0, hipStream_t(stream), a, g, t);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_uvbke_f64(const oeb_domain *d, const oeb_s3 *ub, const oeb_s3 *vb, const oeb_s3 *uc,
                             const oeb_s3 *vc, const oeb_s3 *cosa, const oeb_s3 *rsina, double dt5, void *stream) {
  return oeb::uvbke_run<double>(d, ub, vb, uc, vc, cosa, rsina, dt5, stream);
}

extern "C" int oeb_uvbke_f32(const oeb_domain *d, const oeb_s3f *ub, const oeb_s3f *vb, const oeb_s3f *uc,
                             const oeb_s3f *vc, const oeb_s3f *cosa, const oeb_s3f *rsina, double dt5,
                             void *stream) {
  return oeb::uvbke_run<float>(d, ub, vb, uc, vc, cosa, rsina, dt5, stream);
}
