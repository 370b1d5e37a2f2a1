// fvtp2d.hip -- FV3 2-D finite-volume tracer transport (PPM fluxes), fused.
// Reference: kernels/fvtp2d.h:4-177 (the DEFAULT variant; there is no OpenMP
// variant).  The reference makes ~14 sweeps per level through scratch arrays
// (al, almq, br, b0, smt5, fxx, fyy).  Here one workgroup takes a tile of
// TI x TJ output points of one level at a time.  All intermediates stay in
// LDS, and the four PPM stages and both inner updates run in one launch:
//
//   A  q on the tile plus a 3-point halo                       -> sq
//   B  fy2 = PPM_j(q, cry) on i [i0-3, i0+TI+2) x j [j0, j0+TJ] -> sfy2   (:9-37)
//      fx2 = PPM_i(q, crx) on i [i0, i0+TI] x j [j0-3, j0+TJ+2) -> sfx2   (:84-113)
//   C  q_i = (q area + yfx fy2 - yfx[j+1] fy2[j+1]) / ra_y      -> sqi    (:39-51)
//      q_j = (q area + xfx fx2 - xfx[i+1] fx2[i+1]) / ra_x      -> sqj    (:115-127)
//   D  fx1 = 0.5 (PPM_i(q_i, crx) + fx2) xfx, fy1 = 0.5 (PPM_j(q_j, cry) + fy2) yfx
//                                                                       (:53-82, :129-176)
//
// ppm_face() forms al, almq, br, b0 and smt5 at the three faces one flux needs,
// with the reference's expressions, so every value equals the reference's
// value.  fy2, fx2, q_i and q_j are outputs on extended regions.  A tile
// computes them on its tile plus halo, and writes only the points it owns:
// its own rows and columns, plus the region's extension beyond the domain when
// the tile touches that edge.  So every point is written exactly once.
// FP32 form: the PPM stages are those of the reference's float build; the flux
// divergences of q_i and q_j are compensated (flux_div below).
// Algorithmic traffic: 8 inputs + 6 outputs = 112 B per interior point (FP64),
// 56 B (FP32).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTI = 64;   // tile width along i (one wavefront)
constexpr int kNT = 256;  // threads per workgroup
// Tile height along j and levels per workgroup of the launch.
constexpr int kTJdefault = 8;
constexpr int kChunk = 2;

// PPM flux at face m between cells m-1 and m along one axis, given
// f(m-3) .. f(m+2) and the Courant number c at m (kernels/fvtp2d.h:12, 20-34).
template <class T>
__device__ __forceinline__ T ppm_face(T fm3, T fm2, T fm1, T f0, T fp1, T fp2, T c, T p1, T p2) {
  const T al_m = ((p1 * (fm2 + fm1)) + (p2 * (fm3 + f0)));  // al(m-1)
  const T al_0 = ((p1 * (fm1 + f0)) + (p2 * (fm2 + fp1)));  // al(m)
  const T al_p = ((p1 * (f0 + fp1)) + (p2 * (fm1 + fp2)));  // al(m+1)
  const T almq_m = (al_m - fm1), br_m = (al_0 - fm1), b0_m = (almq_m + br_m);
  const T almq_0 = (al_0 - f0), br_0 = (al_p - f0), b0_0 = (almq_0 + br_0);
  const T smt_m = ((almq_m * br_m) < T(0)) ? T(1) : T(0);
  const T smt_0 = ((almq_0 * br_0) < T(0)) ? T(1) : T(0);
  const T smt = (smt_m + (smt_0 * (smt_m == T(0) ? T(1) : T(0))));
  const T up = (fm1 + (((T(1) - c) * (br_m - (c * b0_m))) * smt));
  const T dn = (f0 + (((T(1) + c) * (almq_0 + (c * b0_0))) * smt));
  return (c > T(0)) ? up : dn;
}

// Flux divergence ((q*area + xa*ya) - xb*yb) / ra: q_i with fyy = yfx*fy2 at
// j and j+1 (kernels/fvtp2d.h:39-51), q_j with fxx = xfx*fx2 at i and i+1
// (:115-127).  FP64 keeps the reference's expression.
__device__ __forceinline__ double flux_div(double q, double area, double xa, double ya, double xb, double yb,
                                           double ra) {
  const double fa = (xa * ya), fb = (xb * yb);
  return ((((q * area) + fa) - fb) / ra);
}

// FP32: the numerator as a compensated dot product (Ogita, Rump and Oishi,
// "Accurate sum and dot product", SIAM J. Sci. Comput. 26(6), 2005, Dot2).
// Each product's rounding error comes exactly from an fma (TwoProduct), each
// sum's from Knuth's TwoSum.  The errors are added once to the rounded
// numerator before the division.  The numerator is then as accurate as if it
// were formed in twice the working precision and rounded once.
__device__ __forceinline__ float flux_div(float q, float area, float xa, float ya, float xb, float yb, float ra) {
  const float p0 = q * area, e0 = fmaf(q, area, -p0);
  const float p1 = xa * ya, e1 = fmaf(xa, ya, -p1);
  const float p2 = xb * yb, e2 = fmaf(xb, yb, -p2);
  const float s1 = p0 + p1, z1 = s1 - p0, t1 = (p0 - (s1 - z1)) + (p1 - z1);
  const float s2 = s1 - p2, z2 = s2 - s1, t2 = (s1 - (s2 - z2)) + (-p2 - z2);
  const float err = ((e0 + e1) - e2) + (t1 + t2);
  return (s2 + err) / ra;
}

template <class T>
struct FvIn {
  const T *q, *crx, *cry, *ra_x, *ra_y, *xfx, *yfx, *area;
};
template <class T>
struct FvOut {
  T *q_i, *q_j, *fx1, *fx2, *fy1, *fy2;
};

template <int kTI, int kTJ, int kNT, class T>
__global__ void __launch_bounds__(kNT) fvtp2d_kernel(FvIn<T> a, FvOut<T> o, T p1, T p2, Geo g, Tiles t) {
  // LDS tile extents (elements).
  constexpr int kQW = kTI + 6, kQH = kTJ + 6;  // q:   i [i0-3, i0+TI+3) x j [j0-3, j0+TJ+3)
  constexpr int kYW = kTI + 5, kYH = kTJ + 1;  // fy2: i [i0-3, i0+TI+2) x j [j0, j0+TJ+1)
  constexpr int kXW = kTI + 1, kXH = kTJ + 5;  // fx2: i [i0, i0+TI+1)   x j [j0-3, j0+TJ+2)
  constexpr int kIW = kTI + 5, kIH = kTJ;      // q_i: i [i0-3, i0+TI+2) x j [j0, j0+TJ)
  constexpr int kJW = kTI, kJH = kTJ + 5;      // q_j: i [i0, i0+TI)     x j [j0-3, j0+TJ+2)
  __shared__ T sq[kQH][kQW];
  __shared__ T sfy2[kYH][kYW];
  __shared__ T sfx2[kXH][kXW];
  __shared__ T sqi[kIH][kIW];
  __shared__ T sqj[kJH][kJW];

  const TileCoord tc = tile_coord(t, kTI, kTJ, g.nz);
  const int i0 = tc.i0, j0 = tc.j0;
  const int nx = g.nx, ny = g.ny, sj = g.sj;
  const int tid = int(threadIdx.x);
  // Ownership of extended-region points: a tile owns its own columns/rows,
  // plus everything left of 0 (first tile) or right of its end (last tile).
  const bool first_i = i0 == 0, last_i = i0 + kTI >= nx;
  const bool first_j = j0 == 0, last_j = j0 + kTJ >= ny;
  auto own_i = [&](int gi) { return (gi >= i0 || first_i) && (gi < i0 + kTI || last_i); };
  auto own_j = [&](int gj) { return (gj >= j0 || first_j) && (gj < j0 + kTJ || last_j); };

  for (int k = tc.k0; k < tc.k1; ++k) {
    const int64_t plane = int64_t(k) * g.sk;
    const T *q = a.q + plane;

    // A: q tile with halo.  Points beyond the allocated halo are never read by
    // a valid task below; they are filled with 0.
    for (int e = tid; e < kQH * kQW; e += kNT) {
      const int r = e / kQW, c = e % kQW;
      const int gi = i0 - 3 + c, gj = j0 - 3 + r;
      sq[r][c] = (gi < nx + 3 && gj < ny + 3) ? q[gi + gj * sj] : T(0);
    }
    lds_barrier();

    // B: fy2 on i [-3, nx+2) x j [0, ny+1) and fx2 on i [0, nx+1) x j [-3, ny+2).
    for (int e = tid; e < kYH * kYW; e += kNT) {
      const int r = e / kYW, c = e % kYW;
      const int gi = i0 - 3 + c, gj = j0 + r;
      if (gi < nx + 2 && gj < ny + 1) {
        const int at = gi + gj * sj;
        const T v = ppm_face(sq[r][c], sq[r + 1][c], sq[r + 2][c], sq[r + 3][c], sq[r + 4][c], sq[r + 5][c],
                             a.cry[plane + at], p1, p2);
        sfy2[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.fy2 + plane + at, v);
      }
    }
    for (int e = tid; e < kXH * kXW; e += kNT) {
      const int r = e / kXW, c = e % kXW;
      const int gi = i0 + c, gj = j0 - 3 + r;
      if (gi < nx + 1 && gj < ny + 2) {
        const int at = gi + gj * sj;
        const T *row = &sq[r][c];
        const T v = ppm_face(row[0], row[1], row[2], row[3], row[4], row[5], a.crx[plane + at], p1, p2);
        sfx2[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.fx2 + plane + at, v);
      }
    }
    lds_barrier();

    // C: q_i on i [-3, nx+2) x j [0, ny), q_j on i [0, nx) x j [-3, ny+2).
    for (int e = tid; e < kIH * kIW; e += kNT) {
      const int r = e / kIW, c = e % kIW;
      const int gi = i0 - 3 + c, gj = j0 + r;
      if (gi < nx + 2 && gj < ny) {
        const int at = gi + gj * sj;
        const T v = flux_div(sq[r + 3][c], a.area[plane + at], a.yfx[plane + at], sfy2[r][c], a.yfx[plane + at + sj],
                             sfy2[r + 1][c], a.ra_y[plane + at]);
        sqi[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.q_i + plane + at, v);
      }
    }
    for (int e = tid; e < kJH * kJW; e += kNT) {
      const int r = e / kJW, c = e % kJW;
      const int gi = i0 + c, gj = j0 - 3 + r;
      if (gi < nx && gj < ny + 2) {
        const int at = gi + gj * sj;
        const T v = flux_div(sq[r][c + 3], a.area[plane + at], a.xfx[plane + at], sfx2[r][c], a.xfx[plane + at + 1],
                             sfx2[r][c + 1], a.ra_x[plane + at]);
        sqj[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.q_j + plane + at, v);
      }
    }
    lds_barrier();

    // D: fx1, fy1 on the tile's interior points.
    for (int e = tid; e < kTJ * kTI; e += kNT) {
      const int r = e / kTI, c = e % kTI;
      const int gi = i0 + c, gj = j0 + r;
      if (gi < nx && gj < ny) {
        const int64_t at = plane + gi + gj * sj;
        const T *ri = &sqi[r][c];
        const T fx1 = ppm_face(ri[0], ri[1], ri[2], ri[3], ri[4], ri[5], a.crx[at], p1, p2);
        const T fy1 = ppm_face(sqj[r][c], sqj[r + 1][c], sqj[r + 2][c], sqj[r + 3][c], sqj[r + 4][c], sqj[r + 5][c],
                               a.cry[at], p1, p2);
        st_nt(o.fx1 + at, ((T(0.5) * (fx1 + sfx2[r + 3][c])) * a.xfx[at]));
        st_nt(o.fy1 + at, ((T(0.5) * (fy1 + sfy2[r][c + 3])) * a.yfx[at]));
      }
    }
    // The next level's phase A writes only sq, which phase D does not read;
    // the barrier after phase A orders this phase D before the next phase B.
  }
}

template <int TI, int TJ, int NT, class T>
int fvtp2d_launch(const char *K, const FvIn<T> &a, const FvOut<T> &o, T p1, T p2, const Geo &g,
                  const oeb_domain *d, void *stream) {
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  const Tiles t = make_tiles(nx, ny, int(nz), TI, TJ, kchunk_for(K, sizeof(T), d, kChunk));
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), TI, TJ));
  hipLaunchKernelGGL((fvtp2d_kernel<TI, TJ, NT, T>), dim3(t.n), dim3(NT), 0, hipStream_t(stream), a, o, p1, p2, g, t);
  return launch_status(K);
}

template <class T, class S>
int fvtp2d_run(const oeb_domain *d, const S *q_i, const S *q_j, const S *fx1, const S *fx2, const S *fy1,
               const S *fy2, const S *q, const S *crx, const S *cry, const S *ra_x, const S *ra_y, const S *xfx,
               const S *yfx, const S *area, T p1, T p2, void *stream) {
  const char *K = "fvtp2d";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_fy2{-3, nx + 2, 0, ny + 1, 0, nz}, r_fx2{0, nx + 1, -3, ny + 2, 0, nz};
  const Box r_qi{-3, nx + 2, 0, ny, 0, nz}, r_qj{0, nx, -3, ny + 2, 0, nz};
  const Box r_q{-3, nx + 3, -3, ny + 3, 0, nz}, r_area{-3, nx + 2, -3, ny + 2, 0, nz};
  // yfx is read at (i, j+1) for q_i rows, xfx at (i+1, j) for q_j columns: both
  // stay inside the fy2 / fx2 regions.
  FvIn<T> a;
  FvOut<T> o;
  const T *pq_i, *pq_j, *pfx1, *pfx2, *pfy1, *pfy2;
  OEB_TRY(check3(K, "q_i", q_i, r_qi, g, true, &pq_i));
  OEB_TRY(check3(K, "q_j", q_j, r_qj, g, false, &pq_j));
  OEB_TRY(check3(K, "fx1", fx1, r0, g, false, &pfx1));
  OEB_TRY(check3(K, "fx2", fx2, r_fx2, g, false, &pfx2));
  OEB_TRY(check3(K, "fy1", fy1, r0, g, false, &pfy1));
  OEB_TRY(check3(K, "fy2", fy2, r_fy2, g, false, &pfy2));
  OEB_TRY(check3(K, "q", q, r_q, g, false, &a.q));
  OEB_TRY(check3(K, "crx", crx, r_fx2, g, false, &a.crx));
  OEB_TRY(check3(K, "cry", cry, r_fy2, g, false, &a.cry));
  OEB_TRY(check3(K, "ra_x", ra_x, r_qj, g, false, &a.ra_x));
  OEB_TRY(check3(K, "ra_y", ra_y, r_qi, g, false, &a.ra_y));
  OEB_TRY(check3(K, "xfx", xfx, r_fx2, g, false, &a.xfx));
  OEB_TRY(check3(K, "yfx", yfx, r_fy2, g, false, &a.yfx));
  OEB_TRY(check3(K, "area", area, r_area, g, false, &a.area));
  AliasSet as;
  as.add("q_i", pq_i, r_qi, g, true);
  as.add("q_j", pq_j, r_qj, g, true);
  as.add("fx1", pfx1, r0, g, true);
  as.add("fx2", pfx2, r_fx2, g, true);
  as.add("fy1", pfy1, r0, g, true);
  as.add("fy2", pfy2, r_fy2, g, true);
  as.add("q", a.q, r_q, g, false);
  as.add("crx", a.crx, r_fx2, g, false);
  as.add("cry", a.cry, r_fy2, g, false);
  as.add("ra_x", a.ra_x, r_qj, g, false);
  as.add("ra_y", a.ra_y, r_qi, g, false);
  as.add("xfx", a.xfx, r_fx2, g, false);
  as.add("yfx", a.yfx, r_fy2, g, false);
  as.add("area", a.area, r_area, g, false);
  OEB_TRY(as.check(K));
  o.q_i = const_cast<T *>(pq_i);
  o.q_j = const_cast<T *>(pq_j);
  o.fx1 = const_cast<T *>(pfx1);
  o.fx2 = const_cast<T *>(pfx2);
  o.fy1 = const_cast<T *>(pfy1);
  o.fy2 = const_cast<T *>(pfy2);

  // Tile shape and lanes, measured in the ten-stencil step at 256x256x80 and
  // 512x512x80 over 64x4, 64x8, 64x16, 128x8, 128x16 and 256x8 tiles of 256,
  // 384, 512 and 1024 lanes.  FP64: 64x16 tiles of 512 lanes (52 KB of LDS,
  // 3 workgroups = 24 waves per CU) are fastest, 1.09x / 1.17x faster than
  // 64x8 x 256 lanes (20 waves per CU); the four phases wait on global loads
  // of their operands, and more waves hide that.  1024 lanes hold fewer
  // workgroups and ran 1.2-1.3x slower.  FP32: 128x16 tiles of 512 lanes
  // (half the bytes per tile, so less halo for the same LDS) are fastest.
  // OEB_TILE_FVTP2D = 1..12 forces one of the measured shapes.
  switch (env_choice("OEB_TILE_", K, sizeof(T) == 4 ? 3 : 8, 0, 12)) {
    case 10: return fvtp2d_launch<64, 16, 1024, T>(K, a, o, p1, p2, g, d, stream);
    case 11: return fvtp2d_launch<128, 16, 1024, T>(K, a, o, p1, p2, g, d, stream);
    case 12: return fvtp2d_launch<128, 8, 1024, T>(K, a, o, p1, p2, g, d, stream);
    case 7: return fvtp2d_launch<64, 8, 512, T>(K, a, o, p1, p2, g, d, stream);
    case 8: return fvtp2d_launch<64, 16, 512, T>(K, a, o, p1, p2, g, d, stream);
    case 9: return fvtp2d_launch<64, 8, 384, T>(K, a, o, p1, p2, g, d, stream);
    case 5: return fvtp2d_launch<64, 4, 256, T>(K, a, o, p1, p2, g, d, stream);
    case 6: return fvtp2d_launch<256, 8, 512, T>(K, a, o, p1, p2, g, d, stream);
    case 1: return fvtp2d_launch<128, 8, 512, T>(K, a, o, p1, p2, g, d, stream);
    case 2: return fvtp2d_launch<64, 16, 256, T>(K, a, o, p1, p2, g, d, stream);
    case 3: return fvtp2d_launch<128, 16, 512, T>(K, a, o, p1, p2, g, d, stream);
    case 4: return fvtp2d_launch<128, 8, 256, T>(K, a, o, p1, p2, g, d, stream);
    default: return fvtp2d_launch<kTI, kTJdefault, kNT, T>(K, a, o, p1, p2, g, d, stream);
  }
}

}  // namespace
}  // namespace oeb

// fxx, fyy, al, almq, br, b0 and smt5 are the reference's scratch arrays;
// nothing is stored in them.
extern "C" int oeb_fvtp2d_f64(const oeb_domain *d, const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *fx1,
                              const oeb_s3 *fx2, const oeb_s3 *fy1, const oeb_s3 *fy2, const oeb_s3 *q,
                              const oeb_s3 *crx, const oeb_s3 *cry, const oeb_s3 *ra_x, const oeb_s3 *ra_y,
                              const oeb_s3 *xfx, const oeb_s3 *yfx, const oeb_s3 *area, const oeb_s3 *fxx,
                              const oeb_s3 *fyy, const oeb_s3 *al, const oeb_s3 *almq, const oeb_s3 *br,
                              const oeb_s3 *b0, const oeb_s3 *smt5, double p1, double p2, void *stream) {
  (void)fxx, (void)fyy, (void)al, (void)almq, (void)br, (void)b0, (void)smt5;
  return oeb::fvtp2d_run<double>(d, q_i, q_j, fx1, fx2, fy1, fy2, q, crx, cry, ra_x, ra_y, xfx, yfx, area, p1, p2,
                                 stream);
}

extern "C" int oeb_fvtp2d_f32(const oeb_domain *d, const oeb_s3f *q_i, const oeb_s3f *q_j, const oeb_s3f *fx1,
                              const oeb_s3f *fx2, const oeb_s3f *fy1, const oeb_s3f *fy2, const oeb_s3f *q,
                              const oeb_s3f *crx, const oeb_s3f *cry, const oeb_s3f *ra_x, const oeb_s3f *ra_y,
                              const oeb_s3f *xfx, const oeb_s3f *yfx, const oeb_s3f *area, const oeb_s3f *fxx,
                              const oeb_s3f *fyy, const oeb_s3f *al, const oeb_s3f *almq, const oeb_s3f *br,
                              const oeb_s3f *b0, const oeb_s3f *smt5, float p1, float p2, void *stream) {
  (void)fxx, (void)fyy, (void)al, (void)almq, (void)br, (void)b0, (void)smt5;
  return oeb::fvtp2d_run<float>(d, q_i, q_j, fx1, fx2, fy1, fy2, q, crx, cry, ra_x, ra_y, xfx, yfx, area, p1, p2,
                                stream);
}
