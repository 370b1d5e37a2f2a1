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
// A second form, fvtp2d_strip_kernel, marches rows instead of taking tiles (the
// FP32 default); both give the same bits.
// Algorithmic traffic: 8 inputs + 6 outputs = 112 B per interior point (FP64),
// 56 B (FP32).
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTI = 64;   // tile width along i (one wavefront)
constexpr int kNT = 256;  // threads per workgroup
// Tile height along j and levels per workgroup of the launch.  One level per
// workgroup: with the tiles of a row on one XCD it measured 150 us against
// 158 for two levels in FP64 at 256x256x80 (550 against 550 at 512x512x80).
constexpr int kTJdefault = 8;
constexpr int kChunk = 1;

// PPM along one axis (kernels/fvtp2d.h:12, 20-34), split into its edge, cell
// and face parts so that a march along the axis can keep the parts it shares
// with the next face.
//   al(m)   edge value between cells m-1 and m, from f(m-2) .. f(m+1)
//   cell m  almq = al(m) - f(m), br = al(m+1) - f(m), b0 = almq + br, and
//           ext = almq * br < 0 (smt5 of the cell)
//   face m  from cells m-1 and m and the Courant number c at m
// Each value is the reference's expression, so it has the reference's bits.
template <class T>
__device__ __forceinline__ T ppm_al(T fm2, T fm1, T f0, T fp1, T p1, T p2) {
  return ((p1 * (fm1 + f0)) + (p2 * (fm2 + fp1)));
}

template <class T>
struct PpmCell {
  T almq, br, b0;
  bool ext;
};

template <class T>
__device__ __forceinline__ PpmCell<T> ppm_cell(T f, T al_lo, T al_hi) {
  PpmCell<T> c;
  c.almq = (al_lo - f);
  c.br = (al_hi - f);
  c.b0 = (c.almq + c.br);
  c.ext = (c.almq * c.br) < T(0);
  return c;
}

// Face flux from cell m-1 (lo, value flo) and cell m (hi, value fhi).  The
// reference's smt = smt5(m-1) + smt5(m) * (smt5(m-1) == 0 ? 1 : 0) of two
// values in {0, 1} is 1 exactly when either cell is an extremum.
template <class T>
__device__ __forceinline__ T ppm_flux(const PpmCell<T> &lo, T flo, const PpmCell<T> &hi, T fhi, T c) {
  const T smt = (lo.ext || hi.ext) ? T(1) : T(0);
  const T up = (flo + (((T(1) - c) * (lo.br - (c * lo.b0))) * smt));
  const T dn = (fhi + (((T(1) + c) * (hi.almq + (c * hi.b0))) * smt));
  return (c > T(0)) ? up : dn;
}

// PPM flux at face m between cells m-1 and m, given f(m-3) .. f(m+2) and the
// Courant number c at m.
template <class T>
__device__ __forceinline__ T ppm_face(T fm3, T fm2, T fm1, T f0, T fp1, T fp2, T c, T p1, T p2) {
  const T al_m = ppm_al(fm3, fm2, fm1, f0, p1, p2);  // al(m-1)
  const T al_0 = ppm_al(fm2, fm1, f0, fp1, p1, p2);  // al(m)
  const T al_p = ppm_al(fm1, f0, fp1, fp2, p1, p2);  // al(m+1)
  return ppm_flux(ppm_cell(fm1, al_m, al_0), fm1, ppm_cell(f0, al_0, al_p), f0, c);
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

// ------------------------------------------------------------- hoisted tiles
// fvtp2d_kernel with the global loads of each phase issued together ahead of
// its arithmetic.  In fvtp2d_kernel a lane's elements of a phase (three or
// four of each region) each wait for their own loads in turn.  Here a lane
// first loads the operands of all its elements of a region, at addresses
// clamped into the region the phase computes (a clamped value belongs to an
// element that is not computed, and is not used), and then forms them.  The
// regions, expressions and stores are those of fvtp2d_kernel, so the results
// are the same bits.  HV = 1: each region's loads at the start of its phase,
// phase B's with phase A's.  HV = 2: one region ahead: the loads of q_i's
// operands go out before phase B's arithmetic, q_j's before q_i's, and phase
// D's before q_j's.  lds_barrier() orders LDS only, and the compiler moves
// global loads across it: left alone it would issue every phase's loads at
// the level's start (198 VGPRs, 2 waves per SIMD).  A compiler-only memory
// fence before each later group keeps it where it is written.
__device__ __forceinline__ void hold_loads() { __asm__ __volatile__("" ::: "memory"); }

template <int kTI, int kTJ, int kNT, int HV, int WPE, class T>
__global__ void __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(WPE)))
    fvtp2d_hkernel(FvIn<T> a, FvOut<T> o, T p1, T p2, Geo g, Tiles t) {
  constexpr int kQW = kTI + 6, kQH = kTJ + 6;  // q:   i [i0-3, i0+TI+3) x j [j0-3, j0+TJ+3)
  constexpr int kYW = kTI + 5, kYH = kTJ + 1;  // fy2: i [i0-3, i0+TI+2) x j [j0, j0+TJ+1)
  constexpr int kXW = kTI + 1, kXH = kTJ + 5;  // fx2: i [i0, i0+TI+1)   x j [j0-3, j0+TJ+2)
  constexpr int kIW = kTI + 5, kIH = kTJ;      // q_i: i [i0-3, i0+TI+2) x j [j0, j0+TJ)
  constexpr int kJW = kTI, kJH = kTJ + 5;      // q_j: i [i0, i0+TI)     x j [j0-3, j0+TJ+2)
  constexpr int nA = (kQH * kQW + kNT - 1) / kNT, nY = (kYH * kYW + kNT - 1) / kNT;
  constexpr int nX = (kXH * kXW + kNT - 1) / kNT, nI = (kIH * kIW + kNT - 1) / kNT;
  constexpr int nJ = (kJH * kJW + kNT - 1) / kNT, nD = (kTJ * kTI + kNT - 1) / kNT;
  __shared__ T sq[kQH][kQW];
  __shared__ T sfy2[kYH][kYW];
  __shared__ T sfx2[kXH][kXW];
  __shared__ T sqi[kIH][kIW];
  __shared__ T sqj[kJH][kJW];

  const TileCoord tc = tile_coord(t, kTI, kTJ, g.nz);
  const int i0 = tc.i0, j0 = tc.j0;
  const int nx = g.nx, ny = g.ny, sj = g.sj;
  const int tid = int(threadIdx.x);
  const bool first_i = i0 == 0, last_i = i0 + kTI >= nx;
  const bool first_j = j0 == 0, last_j = j0 + kTJ >= ny;
  auto own_i = [&](int gi) { return (gi >= i0 || first_i) && (gi < i0 + kTI || last_i); };
  auto own_j = [&](int gj) { return (gj >= j0 || first_j) && (gj < j0 + kTJ || last_j); };
  auto cl = [](int x, int lo, int hi) { return min(max(x, lo), hi - 1); };
  // Element n of the lane in a region of W columns and N elements, clamped
  // to the region's last element for the loads.
  auto elem = [&](int n, int N) { return min(tid + n * kNT, N - 1); };

  {  // one level per workgroup (the launch takes level chunks of 1)
    const int64_t plane = int64_t(tc.k0) * g.sk;
    const T *q = a.q + plane;
    // Operands: q (A), cry (fy2), crx (fx2), area, yfx at j and j+1, ra_y
    // (q_i), area, xfx at i and i+1, ra_x (q_j), and crx, cry, xfx, yfx (D).
    T vq[nA], vcy[nY], vcx[nX];
    T ia[nI], iy0[nI], iy1[nI], ir[nI];
    T ja[nJ], jx0[nJ], jx1[nJ], jr[nJ];
    T dcx[nD], dcy[nD], dxf[nD], dyf[nD];
    auto load_qi = [&] {
      hold_loads();
#pragma unroll
      for (int n = 0; n < nI; ++n) {
        const int e = elem(n, kIH * kIW);
        const int64_t at = plane + cl(i0 - 3 + e % kIW, -3, nx + 2) + int64_t(cl(j0 + e / kIW, 0, ny)) * sj;
        ia[n] = a.area[at], iy0[n] = a.yfx[at], iy1[n] = a.yfx[at + sj], ir[n] = a.ra_y[at];
      }
    };
    auto load_qj = [&] {
      hold_loads();
#pragma unroll
      for (int n = 0; n < nJ; ++n) {
        const int e = elem(n, kJH * kJW);
        const int64_t at = plane + cl(i0 + e % kJW, 0, nx) + int64_t(cl(j0 - 3 + e / kJW, -3, ny + 2)) * sj;
        ja[n] = a.area[at], jx0[n] = a.xfx[at], jx1[n] = a.xfx[at + 1], jr[n] = a.ra_x[at];
      }
    };
    auto load_d = [&] {
      hold_loads();
#pragma unroll
      for (int n = 0; n < nD; ++n) {
        const int e = elem(n, kTJ * kTI);
        const int64_t at = plane + cl(i0 + e % kTI, 0, nx) + int64_t(cl(j0 + e / kTI, 0, ny)) * sj;
        dcx[n] = a.crx[at], dcy[n] = a.cry[at], dxf[n] = a.xfx[at], dyf[n] = a.yfx[at];
      }
    };

    // A: q tile with halo (clamped into q's box [-3, nx+3) x [-3, ny+3); a
    // point beyond it is never read by a computed element), with the
    // operands of phase B.
#pragma unroll
    for (int n = 0; n < nA; ++n) {
      const int e = elem(n, kQH * kQW);
      vq[n] = q[cl(i0 - 3 + e % kQW, -3, nx + 3) + int64_t(cl(j0 - 3 + e / kQW, -3, ny + 3)) * sj];
    }
#pragma unroll
    for (int n = 0; n < nY; ++n) {
      const int e = elem(n, kYH * kYW);
      vcy[n] = a.cry[plane + cl(i0 - 3 + e % kYW, -3, nx + 2) + int64_t(cl(j0 + e / kYW, 0, ny + 1)) * sj];
    }
#pragma unroll
    for (int n = 0; n < nX; ++n) {
      const int e = elem(n, kXH * kXW);
      vcx[n] = a.crx[plane + cl(i0 + e % kXW, 0, nx + 1) + int64_t(cl(j0 - 3 + e / kXW, -3, ny + 2)) * sj];
    }
#pragma unroll
    for (int n = 0; n < nA; ++n) {
      const int e = tid + n * kNT;
      if (e < kQH * kQW) sq[e / kQW][e % kQW] = vq[n];
    }
    lds_barrier();
    if (HV == 2) load_qi();

    // B: fy2 on i [-3, nx+2) x j [0, ny+1) and fx2 on i [0, nx+1) x j [-3, ny+2).
#pragma unroll
    for (int n = 0; n < nY; ++n) {
      const int e = tid + n * kNT;
      const int r = e / kYW, c = e % kYW;
      const int gi = i0 - 3 + c, gj = j0 + r;
      if (e < kYH * kYW && gi < nx + 2 && gj < ny + 1) {
        const T v = ppm_face(sq[r][c], sq[r + 1][c], sq[r + 2][c], sq[r + 3][c], sq[r + 4][c], sq[r + 5][c], vcy[n],
                             p1, p2);
        sfy2[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.fy2 + plane + gi + int64_t(gj) * sj, v);
      }
    }
#pragma unroll
    for (int n = 0; n < nX; ++n) {
      const int e = tid + n * kNT;
      const int r = e / kXW, c = e % kXW;
      const int gi = i0 + c, gj = j0 - 3 + r;
      if (e < kXH * kXW && gi < nx + 1 && gj < ny + 2) {
        const T *row = &sq[r][c];
        const T v = ppm_face(row[0], row[1], row[2], row[3], row[4], row[5], vcx[n], p1, p2);
        sfx2[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.fx2 + plane + gi + int64_t(gj) * sj, v);
      }
    }
    lds_barrier();

    // C: q_i on i [-3, nx+2) x j [0, ny), q_j on i [0, nx) x j [-3, ny+2).
    if (HV == 1) load_qi();
    if (HV == 2) load_qj();
#pragma unroll
    for (int n = 0; n < nI; ++n) {
      const int e = tid + n * kNT;
      const int r = e / kIW, c = e % kIW;
      const int gi = i0 - 3 + c, gj = j0 + r;
      if (e < kIH * kIW && gi < nx + 2 && gj < ny) {
        const T v = flux_div(sq[r + 3][c], ia[n], iy0[n], sfy2[r][c], iy1[n], sfy2[r + 1][c], ir[n]);
        sqi[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.q_i + plane + gi + int64_t(gj) * sj, v);
      }
    }
    if (HV == 1) load_qj();
    if (HV == 2) load_d();
#pragma unroll
    for (int n = 0; n < nJ; ++n) {
      const int e = tid + n * kNT;
      const int r = e / kJW, c = e % kJW;
      const int gi = i0 + c, gj = j0 - 3 + r;
      if (e < kJH * kJW && gi < nx && gj < ny + 2) {
        const T v = flux_div(sq[r][c + 3], ja[n], jx0[n], sfx2[r][c], jx1[n], sfx2[r][c + 1], jr[n]);
        sqj[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.q_j + plane + gi + int64_t(gj) * sj, v);
      }
    }
    lds_barrier();

    // D: fx1, fy1 on the tile's interior points.
    if (HV == 1) load_d();
#pragma unroll
    for (int n = 0; n < nD; ++n) {
      const int e = tid + n * kNT;
      const int r = e / kTI, c = e % kTI;
      const int gi = i0 + c, gj = j0 + r;
      if (e < kTJ * kTI && gi < nx && gj < ny) {
        const int64_t at = plane + gi + gj * sj;
        const T *ri = &sqi[r][c];
        const T fx1 = ppm_face(ri[0], ri[1], ri[2], ri[3], ri[4], ri[5], dcx[n], p1, p2);
        const T fy1 = ppm_face(sqj[r][c], sqj[r + 1][c], sqj[r + 2][c], sqj[r + 3][c], sqj[r + 4][c], sqj[r + 5][c],
                               dcy[n], p1, p2);
        st_nt(o.fx1 + at, ((T(0.5) * (fx1 + sfx2[r + 3][c])) * dxf[n]));
        st_nt(o.fy1 + at, ((T(0.5) * (fy1 + sfy2[r][c + 3])) * dyf[n]));
      }
    }
  }
}

// ------------------------------------------------------------ LDS-staged tiles
// The tile form with every input of a level brought into LDS up front: the
// eight input regions of a tile level are loaded straight into LDS
// (global_load_lds_dword: each active lane's dword lands at the instruction's
// LDS base + 4 x lane) before phase B, so the phases read no global memory
// and a level waits on its loads once instead of once per phase, with no
// registers spent on loads in flight.  Not the default: measured in the
// ten-stencil step at 256x256x80 / 512x512x80, FP64 64x8 tiles of 512 lanes
// (72 KB of LDS, two workgroups per CU) took 157.6-160.6 / 606.6-622.1 us
// against 144.6-144.9 / 544.8-545.1 for fvtp2d_kernel, 64x16 tiles of 1024
// lanes (124 KB, one per CU) 162.0-170.3 / 645.5-648.1; FP32 93.3-99.3 /
// 351.7-368.5 against the strip's 91.7-92.8 / 303.7-306.7.  A level's loads
// are one batch that the workgroup waits for: at 7.9 us per tile level it
// moved 3.2 TB/s, while the register-staged tile keeps more bytes in flight.

typedef __attribute__((address_space(3))) void lds_void_t;

// A lane's dwords in the rows of a W-element region: instruction q of a row
// moves dwords [64 q, 64 q + 64) of the row's W E dwords (E dwords per
// element).  Columns are clamped into the array's box [bi0, bi1); offsets are
// bytes from column bi0 of the row.
template <int W, class T>
struct RegionCols {
  static constexpr int E = int(sizeof(T)) / 4, D = W * E, NQ = (D + 63) / 64;
  unsigned off[NQ];
  bool on[NQ];
  __device__ __forceinline__ RegionCols(int i_lo, int bi0, int bi1, int lane) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int d = 64 * q + lane;
      on[q] = d < D;
      const int c = d / E, part = d - c * E;
      const int gi = min(max(i_lo + c, bi0), bi1 - 1);
      off[q] = unsigned(((gi - bi0) * E + part) * 4);
    }
  }
};

// Issues the loads of rows j_lo .. j_lo + H - 1 of one level of an array
// (rows clamped into [bj0, bj1)) into the row-major H x W LDS region dst.
// Wave w takes rows w, w + nwaves, ...; a row's address is a scalar row base
// plus the lane's 32-bit offset.
template <int W, int H, class T>
__device__ __forceinline__ void lds_load_region(T *dst, const T *level, const RegionCols<W, T> &cols, int j_lo,
                                                int bj0, int bj1, int bi0, int sj, int wave, int nwaves) {
  using C = RegionCols<W, T>;
  for (int r = wave; r < H; r += nwaves) {
    const int gj = min(max(j_lo + r, bj0), bj1 - 1);
    const char *row = reinterpret_cast<const char *>(level + (int64_t(gj) * sj + bi0));
    unsigned *d = reinterpret_cast<unsigned *>(dst + r * W);
#pragma unroll
    for (int q = 0; q < C::NQ; ++q)
      if (cols.on[q])
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const unsigned *>(row + cols.off[q]),
                                         (lds_void_t *)(d + 64 * q), 4, 0, 0);
  }
}

// Tiles of 64 x TJ columns, one level at a time, NT lanes.  The regions and
// phases are those of fvtp2d_kernel (same expressions, same bits); the inputs
// come from LDS:
//   q          [i0-3, i0+67) x [j0-3, j0+TJ+3)
//   cry, yfx   [i0-3, i0+66) x [j0, j0+TJ+1)      (fy2 region)
//   crx, xfx   [i0, i0+65)   x [j0-3, j0+TJ+2)    (fx2 region)
//   area       [i0-3, i0+66) x [j0-3, j0+TJ+2)    (q_i and q_j regions)
//   ra_y       [i0-3, i0+66) x [j0, j0+TJ)        (q_i region)
//   ra_x       [i0, i0+64)   x [j0-3, j0+TJ+2)    (q_j region)
// Each clamped into its array's box, so every load stays inside the array;
// a clamped value feeds only points that are not stored.  Four barriers per
// level: after the loads, after phase B, after phase C, and after phase D
// (the next level's loads overwrite what phase D reads).
template <int TJ, int NT, class T>
__global__ void __launch_bounds__(NT) fvtp2d_lds_kernel(FvIn<T> a, FvOut<T> o, T p1, T p2, Geo g, Tiles t) {
  constexpr int TI = 64;
  constexpr int kQW = TI + 6, kQH = TJ + 6;
  constexpr int kYW = TI + 5, kYH = TJ + 1;
  constexpr int kXW = TI + 1, kXH = TJ + 5;
  constexpr int kAH = TJ + 5;
  constexpr int kIH = TJ;
  constexpr int kJW = TI, kJH = TJ + 5;
  constexpr int kWaves = NT / 64;
  __shared__ T sq[kQH][kQW];
  __shared__ T scry[kYH][kYW], syfx[kYH][kYW], sfy2[kYH][kYW];
  __shared__ T scrx[kXH][kXW], sxfx[kXH][kXW], sfx2[kXH][kXW];
  __shared__ T sarea[kAH][kYW];
  __shared__ T sray[kIH][kYW], sqi[kIH][kYW];
  __shared__ T srax[kJH][kJW], sqj[kJH][kJW];

  const TileCoord tc = tile_coord(t, TI, TJ, g.nz);
  const int i0 = tc.i0, j0 = tc.j0;
  const int nx = g.nx, ny = g.ny, sj = g.sj;
  const int tid = int(threadIdx.x), lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool first_i = i0 == 0, last_i = i0 + TI >= nx;
  const bool first_j = j0 == 0, last_j = j0 + TJ >= ny;
  auto own_i = [&](int gi) { return (gi >= i0 || first_i) && (gi < i0 + TI || last_i); };
  auto own_j = [&](int gj) { return (gj >= j0 || first_j) && (gj < j0 + TJ || last_j); };
  const RegionCols<kQW, T> cq(i0 - 3, -3, nx + 3, lane);
  const RegionCols<kYW, T> cy(i0 - 3, -3, nx + 2, lane);   // cry, yfx, area, ra_y
  const RegionCols<kXW, T> cx(i0, 0, nx + 1, lane);        // crx, xfx
  const RegionCols<kJW, T> cr(i0, 0, nx, lane);            // ra_x

  for (int k = tc.k0; k < tc.k1; ++k) {
    const int64_t plane = int64_t(k) * g.sk;
    lds_load_region<kQW, kQH>(&sq[0][0], a.q + plane, cq, j0 - 3, -3, ny + 3, -3, sj, wave, kWaves);
    lds_load_region<kYW, kYH>(&scry[0][0], a.cry + plane, cy, j0, 0, ny + 1, -3, sj, wave, kWaves);
    lds_load_region<kYW, kYH>(&syfx[0][0], a.yfx + plane, cy, j0, 0, ny + 1, -3, sj, wave, kWaves);
    lds_load_region<kXW, kXH>(&scrx[0][0], a.crx + plane, cx, j0 - 3, -3, ny + 2, 0, sj, wave, kWaves);
    lds_load_region<kXW, kXH>(&sxfx[0][0], a.xfx + plane, cx, j0 - 3, -3, ny + 2, 0, sj, wave, kWaves);
    lds_load_region<kYW, kAH>(&sarea[0][0], a.area + plane, cy, j0 - 3, -3, ny + 2, -3, sj, wave, kWaves);
    lds_load_region<kYW, kIH>(&sray[0][0], a.ra_y + plane, cy, j0, 0, ny, -3, sj, wave, kWaves);
    lds_load_region<kJW, kJH>(&srax[0][0], a.ra_x + plane, cr, j0 - 3, -3, ny + 2, 0, sj, wave, kWaves);
    // The loads write LDS and count in vmcnt: wait for all of them, then
    // publish them to the other waves.
    __builtin_amdgcn_s_waitcnt(0);
    lds_barrier();

    // B: fy2 on i [-3, nx+2) x j [0, ny+1) and fx2 on i [0, nx+1) x j [-3, ny+2).
    for (int e = tid; e < kYH * kYW; e += NT) {
      const int r = e / kYW, c = e % kYW;
      const int gi = i0 - 3 + c, gj = j0 + r;
      if (gi < nx + 2 && gj < ny + 1) {
        const T v = ppm_face(sq[r][c], sq[r + 1][c], sq[r + 2][c], sq[r + 3][c], sq[r + 4][c], sq[r + 5][c],
                             scry[r][c], p1, p2);
        sfy2[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.fy2 + plane + gi + int64_t(gj) * sj, v);
      }
    }
    for (int e = tid; e < kXH * kXW; e += NT) {
      const int r = e / kXW, c = e % kXW;
      const int gi = i0 + c, gj = j0 - 3 + r;
      if (gi < nx + 1 && gj < ny + 2) {
        const T *row = &sq[r][c];
        const T v = ppm_face(row[0], row[1], row[2], row[3], row[4], row[5], scrx[r][c], p1, p2);
        sfx2[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.fx2 + plane + gi + int64_t(gj) * sj, v);
      }
    }
    lds_barrier();

    // C: q_i on i [-3, nx+2) x j [0, ny), q_j on i [0, nx) x j [-3, ny+2).
    for (int e = tid; e < kIH * kYW; e += NT) {
      const int r = e / kYW, c = e % kYW;
      const int gi = i0 - 3 + c, gj = j0 + r;
      if (gi < nx + 2 && gj < ny) {
        const T v = flux_div(sq[r + 3][c], sarea[r + 3][c], syfx[r][c], sfy2[r][c], syfx[r + 1][c], sfy2[r + 1][c],
                             sray[r][c]);
        sqi[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.q_i + plane + gi + int64_t(gj) * sj, v);
      }
    }
    for (int e = tid; e < kJH * kJW; e += NT) {
      const int r = e / kJW, c = e % kJW;
      const int gi = i0 + c, gj = j0 - 3 + r;
      if (gi < nx && gj < ny + 2) {
        const T v = flux_div(sq[r][c + 3], sarea[r][c + 3], sxfx[r][c], sfx2[r][c], sxfx[r][c + 1], sfx2[r][c + 1],
                             srax[r][c]);
        sqj[r][c] = v;
        if (own_i(gi) && own_j(gj)) st_nt(o.q_j + plane + gi + int64_t(gj) * sj, v);
      }
    }
    lds_barrier();

    // D: fx1, fy1 on the tile's interior points.
    for (int e = tid; e < TJ * TI; e += NT) {
      const int r = e / TI, c = e % TI;
      const int gi = i0 + c, gj = j0 + r;
      if (gi < nx && gj < ny) {
        const int64_t at = plane + gi + int64_t(gj) * sj;
        const T *ri = &sqi[r][c];
        const T fx1 = ppm_face(ri[0], ri[1], ri[2], ri[3], ri[4], ri[5], scrx[r + 3][c], p1, p2);
        const T fy1 = ppm_face(sqj[r][c], sqj[r + 1][c], sqj[r + 2][c], sqj[r + 3][c], sqj[r + 4][c], sqj[r + 5][c],
                               scry[r][c + 3], p1, p2);
        st_nt(o.fx1 + at, ((T(0.5) * (fx1 + sfx2[r + 3][c])) * sxfx[r + 3][c]));
        st_nt(o.fy1 + at, ((T(0.5) * (fy1 + sfy2[r][c + 3])) * syfx[r][c + 3]));
      }
    }
    lds_barrier();
  }
}

// One row's loads for the strip form, for lead row L of the march (rows
// clamped into the ranges the strip reads).
template <class T>
struct FvRow {
  T q;          // q(i, L)
  T cry, yfx;   // at (i, L-2): fy2 and fy1 of row L-2
  T ra_y;       // at (i, L-3): q_i of row L-3
  T crx, xfx;   // at (i, L)
  T area;       // at (i, L)
  T ra_x;       // at (i, L-1): q_j of row L-1
};

// Strip form (Strips, h = 3): a workgroup owns one level, the columns
// [i0, i0 + w) and the rows [j0, j1); lane l holds column i = i0 - 3 + l.
// Iteration t handles the lead row L = j0 - 3 + t, from j0 - 3 to j1 + 2:
//   before the barrier, from the lane's own column (registers):
//     fy2(L-2) = PPM_j(q rows L-5 .. L, cry)                   (:9-37)
//     q_i(L-3) from fy2(L-3), fy2(L-2)                          (:39-51)
//     and it publishes q(L), q_i(L-3), crx(L), xfx(L) and area(L) to LDS;
//   after it, with the row's i neighbours from LDS:
//     fx2(L) = PPM_i(q row L, crx), into LDS                   (:84-113)
//     q_j(L-1) from fx2(L-1) at faces i and i+1                 (:115-127)
//     fy1(L-3) = PPM_j(q_j rows L-6 .. L-1, cry), and its flux  (:129-162, :173-176)
//     fx1(L-3) = PPM_i(q_i row L-3, crx), and its flux          (:53-82, :166-170)
// The two PPM_j march with the rows: each row adds one edge value and one
// cell, and the face takes the cell before it from the previous row
// (ppm_al, ppm_cell, ppm_flux), where a face on its own forms three edges and
// two cells.  q and q_i rows take two LDS slots each; crx, xfx, area and fx2
// take rings of four, which also hold a lane's own values of rows L-3 .. L-1
// until fx1, q_i and q_j use them (so they need no registers).  q_j lags fx2
// by a row so that face i+1 of row L-1 is read from the ring after a barrier
// instead of being computed a second time.  Each slot is rewritten only after
// the barrier that follows its last read by another lane, so one barrier per
// row suffices.  The expressions are those of fvtp2d_kernel (ppm_face,
// flux_div), so the results are the same bits.  Rows are loaded kDepth
// iterations ahead, a block at a time, with addresses clamped into the box
// the strip reads (a masked load would make its select wait and drain the
// prefetch); a row's address is a scalar row offset plus the lane's 32-bit
// byte offset into the array's level (RowBuf).  A tile re-fetches the halo of
// every input for every tile it borders; a strip re-fetches only its chunk's
// five halo rows and one line at each row end.
template <class T, int kDepth, int kMaxLanes>
__global__ void __launch_bounds__(kMaxLanes)
    fvtp2d_strip_kernel(FvIn<T> a, FvOut<T> o, T p1, T p2, Geo g, Strips st) {
  extern __shared__ __align__(16) unsigned char oeb_smem[];
  const int NL = st.lanes;
  T *s_q = reinterpret_cast<T *>(oeb_smem);  // 2 slots: q(L)
  T *s_qi = s_q + 2 * NL;                     // 2 slots: q_i(L-3)
  T *s_cx = s_qi + 2 * NL;                    // 4 slots: crx(L)
  T *s_xf = s_cx + 4 * NL;                    // 4 slots: xfx(L)
  T *s_ar = s_xf + 4 * NL;                    // 4 slots: area(L)
  T *s_f2 = s_ar + 4 * NL;                    // 4 slots: fx2(L)
  const StripCoord c = strip_coord(st, int(blockIdx.x), g.nx, g.ny);
  const int l = int(threadIdx.x), sj = g.sj, w = c.w, j0 = c.j0, j1 = c.j1;
  const int gi = c.i0 - 3 + l;
  const bool first_i = c.i0 == 0, last_i = c.i0 + w >= g.nx;
  const bool first_j = j0 == 0, last_j = j1 >= g.ny;
  const bool y_lane = l < w + 5;            // fy2, q_i on [i0 - 3, i0 + w + 2)
  const bool x_lane = l >= 3 && l < w + 3;  // q_j, fx1, fy1 on [i0, i0 + w)
  const bool f_lane = l >= 3 && l < w + 4;  // fx2 on faces [i0, i0 + w]
  // Extended-region points: a strip owns its own columns and rows, plus what
  // lies beyond the domain on the side of a first or last strip / chunk.
  const bool own_col = (gi >= c.i0 || first_i) && (gi < c.i0 + w || last_i);
  auto own_row = [&](int r) { return (r >= j0 || first_j) && (r < j1 || last_j); };
  // Each array's level as a buffer based at (i0 - 3, -3): row r lies at
  // byte (r + 3) sj E from there, and a lane's column at its byte offset.
  constexpr int E = int(sizeof(T));
  const int64_t base = int64_t(c.k) * g.sk + (c.i0 - 3) - 3 * int64_t(sj);
  auto row = [&](int r) { return (r + 3) * sj * E; };
  const RowBuf bq = row_buf(a.q + base), bcry = row_buf(a.cry + base), byfx = row_buf(a.yfx + base),
               bray = row_buf(a.ra_y + base), bcrx = row_buf(a.crx + base), bxfx = row_buf(a.xfx + base),
               barea = row_buf(a.area + base), brax = row_buf(a.ra_x + base);
  const RowBuf bqi = row_buf(o.q_i + base), bqj = row_buf(o.q_j + base), bfx1 = row_buf(o.fx1 + base),
               bfx2 = row_buf(o.fx2 + base), bfy1 = row_buf(o.fy1 + base), bfy2 = row_buf(o.fy2 + base);
  const unsigned oq = unsigned(min(l, w + 5) * E);                   // q lanes [0, w + 6)
  const unsigned oy = unsigned(min(l, w + 4) * E);                   // y lanes [0, w + 5)
  const unsigned ox = unsigned((3 + min(max(l - 3, 0), w)) * E);     // faces [i0, i0 + w]
  const unsigned orx = unsigned((3 + min(max(l - 3, 0), w - 1)) * E);  // x lanes [3, w + 3)
  const unsigned ol = unsigned(l * E);                               // the lane's column
  auto ld = [&](int L) -> FvRow<T> {
    FvRow<T> r;
    const int rq = min(L, j1 + 2);                    // q rows [j0 - 3, j1 + 2]
    const int r2 = min(max(L - 2, j0), j1);           // rows [j0, j1]
    const int r3 = min(max(L - 3, j0), j1 - 1);       // rows [j0, j1)
    const int rx = min(L, j1 + 1);                    // rows [j0 - 3, j1 + 1]
    const int r1 = min(max(L - 1, j0 - 3), j1 + 1);
    r.q = buf_ld<T>(bq, oq, row(rq));
    r.cry = buf_ld<T>(bcry, oy, row(r2));
    r.yfx = buf_ld<T>(byfx, oy, row(r2));
    r.ra_y = buf_ld<T, kBufNT>(bray, oy, row(r3));
    r.crx = buf_ld<T>(bcrx, ox, row(rx));
    r.xfx = buf_ld<T>(bxfx, ox, row(rx));
    r.area = buf_ld<T>(barea, oy, row(rx));
    r.ra_x = buf_ld<T, kBufNT>(brax, orx, row(r1));
    return r;
  };
  FvRow<T> cur[kDepth];
#pragma unroll
  for (int u = 0; u < kDepth; ++u) cur[u] = ld(j0 - 3 + u);
  // Column windows: q of rows L-3 .. L-1, q_j of rows L-4 .. L-2; fy2, yfx
  // and cry of row L-3.  The column PPM state: q's edge value al(L-2) and
  // cell L-3, q_j's edge value al(L-3) and cell L-4.
  T q3 = T(0), q2 = T(0), q1 = T(0);
  T j4 = T(0), j3 = T(0), j2 = T(0);
  T fy2p = T(0), yfxp = T(0), cryp = T(0);
  T yal = T(0), jal = T(0);
  PpmCell<T> ycell{T(0), T(0), T(0), false}, jcell{T(0), T(0), T(0), false};
  const int nt = j1 - j0 + 6;
  for (int t0 = 0; t0 < nt; t0 += kDepth) {
    FvRow<T> nxt[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) nxt[u] = ld(j0 - 3 + t0 + kDepth + u);
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int t = t0 + u;
      if (t < nt) {  // uniform across the workgroup
        const int L = j0 - 3 + t;
        const FvRow<T> &r = cur[u];
        const T q0 = r.q;
        const int s2 = (t & 1) * NL + l;        // this row's slot of s_q, s_qi
        const int s4 = (t & 3) * NL + l;        // this row's slot of the rings
        const int s4m1 = ((t + 3) & 3) * NL + l;  // row L-1's slot of the rings
        const int s4m3 = ((t + 1) & 3) * NL + l;  // row L-3's slot of the rings
        // fy2(L-2) from edge L-1 and cell L-2 (new) and cell L-3; q_i(L-3).
        const T yal_n = ppm_al(q3, q2, q1, q0, p1, p2);
        const PpmCell<T> ycell_n = ppm_cell(q2, yal, yal_n);
        T fy2n = T(0), qi = T(0);
        if (L - 2 >= j0 && y_lane) {
          fy2n = ppm_flux(ycell, q3, ycell_n, q2, r.cry);
          if (own_col && own_row(L - 2)) buf_st_nt(bfy2, ol, row(L - 2), fy2n);
          if (L - 3 >= j0) {
            qi = flux_div(q3, s_ar[s4m3], yfxp, fy2p, r.yfx, fy2n, r.ra_y);
            if (own_col) buf_st_nt(bqi, ol, row(L - 3), qi);
          }
        }
        s_q[s2] = q0;
        s_qi[s2] = qi;
        s_cx[s4] = r.crx;
        s_xf[s4] = r.xfx;
        s_ar[s4] = r.area;
        lds_barrier();
        // fx2(L) from the q row; q_j(L-1) from fx2(L-1) at faces i and i+1.
        if (L <= j1 + 1 && f_lane) {
          const T *sq = s_q + s2;
          const T fx2 = ppm_face(sq[-3], sq[-2], sq[-1], sq[0], sq[1], sq[2], r.crx, p1, p2);
          s_f2[s4] = fx2;
          if (own_row(L) && (l < w + 3 || last_i)) buf_st_nt(bfx2, ol, row(L), fx2);
        }
        T qj = T(0);
        if (L - 1 >= j0 - 3 && L - 1 <= j1 + 1 && x_lane) {
          qj = flux_div(q1, s_ar[s4m1], s_xf[s4m1], s_f2[s4m1], s_xf[s4m1 + 1], s_f2[s4m1 + 1], r.ra_x);
          if (own_row(L - 1)) buf_st_nt(bqj, ol, row(L - 1), qj);
        }
        // fy1(L-3) from edge L-2 and cell L-3 (new) and cell L-4.
        const T jal_n = ppm_al(j4, j3, j2, qj, p1, p2);
        const PpmCell<T> jcell_n = ppm_cell(j3, jal, jal_n);
        if (L - 3 >= j0 && x_lane) {
          const T fy1 = ppm_flux(jcell, j4, jcell_n, j3, cryp);
          buf_st_nt(bfy1, ol, row(L - 3), ((T(0.5) * (fy1 + fy2p)) * yfxp));
        }
        if (L - 3 >= j0 && x_lane) {
          const T *si = s_qi + s2;
          const T fx1 = ppm_face(si[-3], si[-2], si[-1], si[0], si[1], si[2], s_cx[s4m3], p1, p2);
          buf_st_nt(bfx1, ol, row(L - 3), ((T(0.5) * (fx1 + s_f2[s4m3])) * s_xf[s4m3]));
        }
        q3 = q2, q2 = q1, q1 = q0;
        j4 = j3, j3 = j2, j2 = qj;
        yal = yal_n, ycell = ycell_n, jal = jal_n, jcell = jcell_n;
        fy2p = fy2n, yfxp = r.yfx, cryp = r.cry;
      }
    }
#pragma unroll
    for (int u = 0; u < kDepth; ++u) cur[u] = nxt[u];
  }
}

template <class T, int kDepth>
int fvtp2d_strip_launch(const char *K, const FvIn<T> &a, const FvOut<T> &o, T p1, T p2, const Geo &g,
                        const Strips &st, void *stream) {
  const size_t lds = size_t(20) * st.lanes * sizeof(T);
  if (st.lanes <= 512)
    hipLaunchKernelGGL((fvtp2d_strip_kernel<T, kDepth, 512>), dim3(unsigned(st.n)), dim3(unsigned(st.lanes)), lds,
                       hipStream_t(stream), a, o, p1, p2, g, st);
  else
    hipLaunchKernelGGL((fvtp2d_strip_kernel<T, kDepth, 1024>), dim3(unsigned(st.n)), dim3(unsigned(st.lanes)), lds,
                       hipStream_t(stream), a, o, p1, p2, g, st);
  return launch_status(K);
}

template <int TJ, int NT, class T>
int fvtp2d_lds_launch(const char *K, const FvIn<T> &a, const FvOut<T> &o, T p1, T p2, const Geo &g,
                      const oeb_domain *d, void *stream) {
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  const int order = knob(K, sizeof(T), d, KNOB_ORDER, 2);
  const Tiles t = make_tiles(nx, ny, int(nz), 64, TJ, knob(K, sizeof(T), d, KNOB_KCHUNK, kChunk), order);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), 64, TJ));
  hipLaunchKernelGGL((fvtp2d_lds_kernel<TJ, NT, T>), dim3(t.n), dim3(NT), 0, hipStream_t(stream), a, o, p1, p2, g, t);
  return launch_status(K);
}

template <int TI, int TJ, int NT, class T>
int fvtp2d_launch(const char *K, const FvIn<T> &a, const FvOut<T> &o, T p1, T p2, const Geo &g,
                  const oeb_domain *d, void *stream) {
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  // The tiles of a row on one XCD (Tiles::irow): 546 / 157 us against 554 /
  // 159 in FP64 at 512x512x80 / 256x256x80 (order knob 0 for row order).
  const int order = knob(K, sizeof(T), d, KNOB_ORDER, 2);
  const Tiles t = make_tiles(nx, ny, int(nz), TI, TJ, knob(K, sizeof(T), d, KNOB_KCHUNK, kChunk), order);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), TI, TJ));
  hipLaunchKernelGGL((fvtp2d_kernel<TI, TJ, NT, T>), dim3(t.n), dim3(NT), 0, hipStream_t(stream), a, o, p1, p2, g, t);
  return launch_status(K);
}

template <int TI, int TJ, int NT, int HV, int WPE, class T>
int fvtp2d_hlaunch(const char *K, const FvIn<T> &a, const FvOut<T> &o, T p1, T p2, const Geo &g,
                   const oeb_domain *d, void *stream) {
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  // The tiles of a row on one XCD (Tiles::irow).
  const int order = knob(K, sizeof(T), d, KNOB_ORDER, 2);
  const Tiles t = make_tiles(nx, ny, int(nz), TI, TJ, 1, order);
  OEB_TRY(check_tiles(K, t, nx, ny, int(nz), TI, TJ));
  hipLaunchKernelGGL((fvtp2d_hkernel<TI, TJ, NT, HV, WPE, T>), dim3(t.n), dim3(NT), 0, hipStream_t(stream), a, o, p1, p2,
                     g, t);
  return launch_status(K);
}

template <class T, class S>
int fvtp2d_run(const oeb_domain *d, const S *q_i, const S *q_j, const S *fx1, const S *fx2, const S *fy1,
               const S *fy2, const S *q, const S *crx, const S *cry, const S *ra_x, const S *ra_y, const S *xfx,
               const S *yfx, const S *area, T p1, T p2, void *stream) {
  const char *K = "fvtp2d";
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
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

  // Tiles by default (the hoisted tiles below); the form knob forces the
  // strips (1) or the LDS-staged tiles (2).  The strips were the FP32 default
  // until the hoisted tiles: in the ten-stencil step at 256x256x80 /
  // 512x512x80, 93-94 / 304 us against 80.5 / 295-296 for the hoisted tiles.
  // Measured in the ten-stencil step at 256x256x80 and 512x512x80: FP32
  // strips of 320 lanes take 93 / 318 us against 118 / 396 us for the best
  // tiles (1.25x).  FP64 strips read 1.5x fewer bytes than the tiles
  // (rocprofv3 FETCH_SIZE at 256x256x80: 445 MB against 671 MB), but at 109
  // VGPRs a SIMD holds only 4 of their waves, each row a chain of dependent
  // FP64 operations between two barriers: 163 / 596 us at best (256 lanes,
  // 32 rows) against 163 / 572 us for the tiles, which run at 5.6 TB/s of
  // fetched plus written bytes.  A third form, 64-column bands marching down
  // chunks of rows in 16-row blocks that hand their shared rows on in
  // registers (no j-halo re-fetched or recomputed), fetched 599 MB and took
  // 164 / 570 us: the tiles' extra bytes are mostly the i-halo (a 64-wide FP64
  // row with 3 halo points each side spans 6 cache lines instead of 4).
  // The strip form addresses a level as a buffer with byte offsets below 2^31
  // (RowBuf); a larger level takes the tiles.
  const bool strip_fits = (ny + 6) * int64_t(g.sj) * int64_t(sizeof(T)) + (nx + 6) * int64_t(sizeof(T)) <= INT32_MAX;
  // The form knob: 0 tiles, 1 strips, 2 LDS-staged tiles (fvtp2d_lds_kernel,
  // with the tile knob: 0 64x8 of 512 lanes, 1 64x16 of 1024, 2 64x8 of 256).
  const int form = knob(K, sizeof(T), d, KNOB_FORM, 0);
  if (form == 2) {
    switch (knob(K, sizeof(T), d, KNOB_TILE, 0)) {
      case 1: return fvtp2d_lds_launch<16, 1024, T>(K, a, o, p1, p2, g, d, stream);
      case 2: return fvtp2d_lds_launch<8, 256, T>(K, a, o, p1, p2, g, d, stream);
      default: return fvtp2d_lds_launch<8, 512, T>(K, a, o, p1, p2, g, d, stream);
    }
  }
  if (strip_fits && form == 1) {
    // Strips of at most 512 lanes (the lanes knob: 64 .. 1024).
    const int max_lanes = knob(K, sizeof(T), d, KNOB_LANES, 512);
    // At most 12 rows per chunk: FP32 308-310 us against 321-322 for the 32
    // rows of default_strip_rows at 512x512x80, 94-95 us against 95-96 for
    // its 20 at 256x256x80 (the ten-stencil step, repeated).  Rows 8-24,
    // 512 or 1024 lanes and prefetch depth 1 or 2 measured 304-419 us at
    // 512x512x80, this default among the fastest.  One-wave strips (64
    // lanes, which need no s_barrier) took 456-504 us and 128 lanes 391 us:
    // a 58-column row spans three cache lines per array where 256 columns
    // span little more than their own.
    Strips st = make_strips(nx, ny, int(nz), 3, 1, 1, max_lanes);
    st = make_strips(nx, ny, int(nz), 3, knob(K, sizeof(T), d, KNOB_ROWS, std::min(12, default_strip_rows(ny, st.lanes))), 1,
                     max_lanes);
    OEB_TRY(check_strips(K, st));
    if (knob(K, sizeof(T), d, KNOB_DEPTH, 1) >= 2) return fvtp2d_strip_launch<T, 2>(K, a, o, p1, p2, g, st, stream);
    return fvtp2d_strip_launch<T, 1>(K, a, o, p1, p2, g, st, stream);
  }

  // Tile shape and lanes, measured in the ten-stencil step at 256x256x80 and
  // 512x512x80 over 64x4, 64x8, 64x16, 128x8, 128x16 and 256x8 tiles of 256,
  // 384, 512 and 1024 lanes.  FP64: 64x16 tiles of 512 lanes (52 KB of LDS,
  // 3 workgroups = 24 waves per CU) are fastest, 1.09x / 1.17x faster than
  // 64x8 x 256 lanes (20 waves per CU); the four phases wait on global loads
  // of their operands, and more waves hide that.  1024 lanes hold fewer
  // workgroups and ran 1.2-1.3x slower.  FP32: 128x16 tiles of 512 lanes
  // (half the bytes per tile, so less halo for the same LDS) are fastest.
  // the tile knob = 1..12 forces one of the measured shapes, 0 the 64x8
  // tiles of 256 lanes.  13-17 are the hoisted tiles (fvtp2d_hkernel): 64x16
  // of 512 lanes with each region's loads at the start of its phase (13) or
  // one region ahead (14, the default), and one region ahead 64x8 of 256
  // lanes (15), 128x16 of 1024 (16) and 128x8 of 512 (17).  In the
  // ten-stencil step at 256x256x80 / 512x512x80, FP64: 14 took 133.1-134.1
  // / 516.2-516.9 us, 13 138.2-139.1 / 526.2-527.4, 8 144.1-144.6 /
  // 546.4-548.1; FP32: 14 80.5 / 295.2-295.7, 16 in the level-interleaved
  // order 80.4-81.1 / 305.1-305.9.  Launched alone: 15 took 142.5-143.3 /
  // 545.8-546.6 in FP64, 17 150 / 563; 256x8 tiles of 1024 lanes 75-76 /
  // 315-318 in FP32, the 128x16 tiles of 512 lanes (3) 81.6 / 389.5.  Taller
  // hoisted tiles, with less j-halo per point, measured no faster in the step:
  // 64x24 of 768 lanes (two workgroups per CU) the same as 14, 64x24 of 512
  // and 64x32 of 1024 lanes 0.5-2% slower.  Addressing every array of the
  // hoisted tile as a buffer based at the tile's corner (one 32-bit offset
  // per element for all its operands, only upper clamps) cut its vector
  // integer instructions from 453 to 324 and all its vector instructions by
  // 12%, and measured 2% faster in FP32 and 1-2% slower in FP64.  Global
  // loads and stores off a scalar base at the corner, with one 32-bit offset
  // per element (upper clamps only, a 24-bit multiply, no 64-bit address
  // arithmetic) cut the vector instructions by 18% in FP32 (1117 to 918) and
  // 21% in FP64 (1053 to 836), and took 291.6 / 79.6-79.9 us in FP32 against
  // 292.2-293.9 / 81.6-82.6, and 533-535 / 137.6 in FP64 against 519-522 /
  // 134 (512x512x80 / 256x256x80): the kernel is not bound by its vector
  // instruction issue, even in FP32.  Eight
  // waves per SIMD instead of six (64x12 tiles with q_i written over the q
  // tile after one more barrier: 58 VGPRs, 35 KB, four workgroups per CU)
  // took 540-543 / 143 us against 519 / 133 at 512x512x80 / 256x256x80.
  switch (knob(K, sizeof(T), d, KNOB_TILE, 14)) {
    case 13: return fvtp2d_hlaunch<64, 16, 512, 1, 6, T>(K, a, o, p1, p2, g, d, stream);
    case 14: return fvtp2d_hlaunch<64, 16, 512, 2, 6, T>(K, a, o, p1, p2, g, d, stream);
    case 15: return fvtp2d_hlaunch<64, 8, 256, 2, 5, T>(K, a, o, p1, p2, g, d, stream);
    case 16: return fvtp2d_hlaunch<128, 16, 1024, 2, 4, T>(K, a, o, p1, p2, g, d, stream);
    case 17: return fvtp2d_hlaunch<128, 8, 512, 2, 4, T>(K, a, o, p1, p2, g, d, stream);
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
