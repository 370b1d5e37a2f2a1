// adjoint.hip -- hand-written adjoints (vector-Jacobian products) of
// hdiffsa, hdiffsmag, hadvuv5th, fvtp2d and fastwavesuv.
//
// Each adjoint takes the forward inputs and the output cotangents (*_bar) and
// writes the cotangents of the differentiated inputs on the whole box the
// forward reads them on, halo included.  Coefficient fields (masks, metric
// terms, Courant numbers, densities, ...) are constants.  Every transposed
// stencil is a gather: a thread owns one input point and sums the
// contributions of the outputs that read it, so no atomics are needed.
// Branches (flux limiters, upwind direction, clamps, PPM monotonicity) are
// frozen at the forward pass's choice: each adjoint re-evaluates the forward
// condition from the same inputs with the forward's expression, and
// differentiates only the branch taken.
//
// Stages that a gather cannot express in one pass keep their intermediate
// cotangents in a caller-provided scratch array (the reference's own scratch
// arguments: lap for hdiffsa, T_sqr_s / S_sqr_uv for hdiffsmag, al / almq /
// br / b0 / smt5 for fvtp2d), on the same layout as the inputs.
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kTI = 64, kTJ = 4;

// EARTH_RADIUS_RECIP, include/util.h:17-18, formed in the element type.
template <class T>
constexpr T kEarthRadiusRecipAdj = T(1.0) / T(6371.229e3);

// A box [ilo, ihi) x [jlo, jhi) x [0, nz) walked by 64 x 4 tiles, one level
// per block row of the grid (grid.y = nz).
struct BoxGrid {
  int ilo, jlo, ni, nj;
};

__device__ __forceinline__ bool box_point(const BoxGrid &b, int &i, int &j) {
  const int ti = (b.ni + kTI - 1) / kTI;
  i = b.ilo + int(blockIdx.x % ti) * kTI + int(threadIdx.x);
  j = b.jlo + int(blockIdx.x / ti) * kTJ + int(threadIdx.y);
  return i < b.ilo + b.ni && j < b.jlo + b.nj;
}

inline dim3 box_blocks(const BoxGrid &b, int nz) {
  return dim3(unsigned(((b.ni + kTI - 1) / kTI) * ((b.nj + kTJ - 1) / kTJ)), unsigned(nz));
}

__device__ __forceinline__ bool in_range(int a, int lo, int hi) { return a >= lo && a < hi; }

// =================================================================== hdiffsa
// Forward (kernels/hdiffsa.h:141-169):
//   lap(p)    = in(p-1) + in(p+1) - 2 in(p) + lo(j) (in(j+1) - in(p)) + lu(j) (in(j-1) - in(p))
//   flx(i)    = lap(i+1) - lap(i),        0 if flx (in(i+1) - in(i)) > 0
//   fly(j)    = lo(j) (lap(j+1) - lap(j)), 0 if fly (in(j+1) - in(j)) > 0
//   out(p)    = in(p) + (flx(i-1) - flx(i) + fly(j-1) - fly(j)) mask(p)
// Adjoint stage 1, on [-1, nx+1) x [-1, ny+1): lap_bar(p), gathered from the
// cotangents of the (at most four) fluxes that read lap(p).  Stage 2, on
// [-2, nx+2) x [-2, ny+2): in_bar(q) = out_bar(q) [interior] + the transposed
// Laplacian of lap_bar.
template <class T>
struct HsaAdj {
  const T *in, *mask, *lo, *lu, *obar;
  int slo, slu;
  T *lapbar, *inbar;
};

template <class T>
__device__ __forceinline__ T hsa_lap(const T *__restrict__ in, int64_t at, int sj, T lo, T lu) {
  const T c = in[at];
  return in[at - 1] + in[at + 1] - T(2.0) * c + lo * (in[at + sj] - c) + lu * (in[at - sj] - c);
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) hdiffsa_adj_lap(HsaAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const int k = int(blockIdx.y), sj = g.sj, nx = g.nx, ny = g.ny;
  const int64_t at = int64_t(k) * g.sk + i + int64_t(j) * sj;
  auto lo = [&](int jj) { return a.lo[int64_t(jj) * a.slo]; };
  auto lu = [&](int jj) { return a.lu[int64_t(jj) * a.slu]; };
  auto lap = [&](int di, int dj) { return hsa_lap(a.in, at + di + int64_t(dj) * sj, sj, lo(j + dj), lu(j + dj)); };
  // Cotangent of out(i+di, j+dj) times its mask, 0 outside the interior.
  auto mo = [&](int di, int dj) {
    return (in_range(i + di, 0, nx) && in_range(j + dj, 0, ny))
               ? a.mask[at + di + int64_t(dj) * sj] * a.obar[at + di + int64_t(dj) * sj]
               : T(0);
  };
  T r = T(0);
  const T l0 = lap(0, 0);
  if (in_range(j, 0, ny)) {
    if (in_range(i - 1, -1, nx)) {  // flx(i-1) = lap(i) - lap(i-1): +1
      const T f = l0 - lap(-1, 0);
      if (!(f * (a.in[at] - a.in[at - 1]) > 0)) r += mo(0, 0) - mo(-1, 0);
    }
    if (in_range(i, -1, nx)) {  // flx(i) = lap(i+1) - lap(i): -1
      const T f = lap(1, 0) - l0;
      if (!(f * (a.in[at + 1] - a.in[at]) > 0)) r -= mo(1, 0) - mo(0, 0);
    }
  }
  if (in_range(i, 0, nx)) {
    if (in_range(j - 1, -1, ny)) {  // fly(j-1) = lo(j-1) (lap(j) - lap(j-1)): +lo(j-1)
      const T f = lo(j - 1) * (l0 - lap(0, -1));
      if (!(f * (a.in[at] - a.in[at - sj]) > 0)) r += lo(j - 1) * (mo(0, 0) - mo(0, -1));
    }
    if (in_range(j, -1, ny)) {  // fly(j) = lo(j) (lap(j+1) - lap(j)): -lo(j)
      const T f = lo(j) * (lap(0, 1) - l0);
      if (!(f * (a.in[at + sj] - a.in[at]) > 0)) r -= lo(j) * (mo(0, 1) - mo(0, 0));
    }
  }
  a.lapbar[at] = r;
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) hdiffsa_adj_in(HsaAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const int k = int(blockIdx.y), sj = g.sj, nx = g.nx, ny = g.ny;
  const int64_t at = int64_t(k) * g.sk + i + int64_t(j) * sj;
  // lap_bar lives on [-1, nx+1) x [-1, ny+1).
  auto lb = [&](int di, int dj) {
    return (in_range(i + di, -1, nx + 1) && in_range(j + dj, -1, ny + 1)) ? a.lapbar[at + di + int64_t(dj) * sj]
                                                                             : T(0);
  };
  T r = (in_range(i, 0, nx) && in_range(j, 0, ny)) ? a.obar[at] : T(0);
  if (in_range(i, -1, nx + 1) && in_range(j, -1, ny + 1))
    r += lb(0, 0) * (T(-2.0) - a.lo[int64_t(j) * a.slo] - a.lu[int64_t(j) * a.slu]);
  r += lb(-1, 0) + lb(1, 0);
  if (in_range(j - 1, -1, ny + 1)) r += lb(0, -1) * a.lo[int64_t(j - 1) * a.slo];
  if (in_range(j + 1, -1, ny + 1)) r += lb(0, 1) * a.lu[int64_t(j + 1) * a.slu];
  a.inbar[at] = r;
}

template <class T, class S3, class S1>
int hdiffsa_adj_run(const oeb_domain *d, const S3 *in, const S3 *mask, const S1 *crlato, const S1 *crlatu,
                    const S3 *out_bar, const S3 *in_bar, const S3 *lap, void *stream) {
  const char *K = "hdiffsa_adj";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz}, r_in{-2, nx + 2, -2, ny + 2, 0, nz}, r_l{-1, nx + 1, -1, ny + 1, 0, nz};
  HsaAdj<T> a;
  const T *p_ib, *p_lb;
  OEB_TRY(check3(K, "in", in, r_in, g, true, &a.in));
  OEB_TRY(check3(K, "mask", mask, r0, g, false, &a.mask));
  OEB_TRY(check3(K, "out_bar", out_bar, r0, g, false, &a.obar));
  OEB_TRY(check3(K, "in_bar", in_bar, r_in, g, false, &p_ib));
  OEB_TRY(check3(K, "lap", lap, r_l, g, false, &p_lb));
  OEB_TRY(check1(K, "crlato", crlato, -1, ny + 1, &a.lo, &a.slo));
  OEB_TRY(check1(K, "crlatu", crlatu, -1, ny + 1, &a.lu, &a.slu));
  AliasSet as;
  as.add("in", a.in, r_in, g, false);
  as.add("mask", a.mask, r0, g, false);
  as.add("out_bar", a.obar, r0, g, false);
  as.add("in_bar", p_ib, r_in, g, true);
  as.add("lap", p_lb, r_l, g, true);
  OEB_TRY(as.check(K));
  a.inbar = const_cast<T *>(p_ib);
  a.lapbar = const_cast<T *>(p_lb);
  const BoxGrid b1{-1, -1, int(nx + 2), int(ny + 2)}, b2{-2, -2, int(nx + 4), int(ny + 4)};
  hipLaunchKernelGGL(hdiffsa_adj_lap<T>, box_blocks(b1, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g, b1);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(hdiffsa_adj_in<T>, box_blocks(b2, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g, b2);
  return launch_status(K);
}

// =============================================================== fastwavesuv
// Forward (kernels/fastwavesuv.h:304-341), differentiated in uin, vin, ppuv:
//   gk(p,k)  = w(p,k) pp(p,k) + (1 - w(p,k)) pp(p,k-1)
//   gc(p,k)  = gk(p,k+1) - gk(p,k)
//   ppgu(p)  = (pp(i+1) - pp(i)) + (gc(i+1) + gc(i)) * 0.5 * Hu(p)
//   uout     = uin + dt (utens - ppgu * 2 fx(j) / (rho(i+1) + rho(i)))
// (Hu, Hv the hhl ratios; v and ppgv alike along j).  uin_bar = uout_bar and
// vin_bar = vout_bar on the interior.  ppuv_bar is gathered on
// [0, nx+1) x [0, ny+1) x [-1, nz+1) by one lane per column marching up k,
// carrying the gc cotangents of levels k-1 and k.  ppuv, hhl and the ppuv
// cotangent are FP64 in both forms (the FP64 islands of the FP32 form).
template <class T>
struct FwAdj {
  const T *wgtfac, *rho, *fx, *ubar, *vbar;
  const double *ppuv, *hhl;
  int fx_stride;
  T edadlat, dt;
  T *uinbar, *vinbar;
  double *ppbar;
};

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) fastwavesuv_adj_uv(FwAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * g.sj;
  a.uinbar[at] = a.ubar[at];
  a.vinbar[at] = a.vbar[at];
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) fastwavesuv_adj_pp(FwAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const int sj = g.sj, nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t sk = g.sk, col = i + int64_t(j) * sj;
  const double dt = a.dt;
  // Cotangents of ppgu(i+di, j+dj, k) and ppgv(...) (0 outside the interior).
  auto A = [&](int di, int dj, int k) -> double {
    if (!in_range(i + di, 0, nx) || !in_range(j + dj, 0, ny) || !in_range(k, 0, nz)) return 0.0;
    const int64_t at = int64_t(k) * sk + col + di + int64_t(dj) * sj;
    const T c = T(2.0) * a.fx[int64_t(j + dj) * a.fx_stride] / (a.rho[at + 1] + a.rho[at]);
    return -dt * double(c) * double(a.ubar[at]);
  };
  auto B = [&](int di, int dj, int k) -> double {
    if (!in_range(i + di, 0, nx) || !in_range(j + dj, 0, ny) || !in_range(k, 0, nz)) return 0.0;
    const int64_t at = int64_t(k) * sk + col + di + int64_t(dj) * sj;
    const T c = T(2.0) * a.edadlat / (a.rho[at + sj] + a.rho[at]);
    return -dt * double(c) * double(a.vbar[at]);
  };
  // The hhl ratio of ppgu at (i+di, j+dj, k) (hu) and of ppgv (hv).
  auto hu = [&](int di, int dj, int k) {
    const int64_t at = int64_t(k) * sk + col + di + int64_t(dj) * sj, up = at + sk;
    const double *h = a.hhl;
    return ((h[up] + h[at]) - (h[up + 1] + h[at + 1])) / ((h[up] - h[at]) + (h[up + 1] - h[at + 1]));
  };
  auto hv = [&](int di, int dj, int k) {
    const int64_t at = int64_t(k) * sk + col + di + int64_t(dj) * sj, up = at + sk;
    const double *h = a.hhl;
    return ((h[up] + h[at]) - (h[up + sj] + h[at + sj])) / ((h[up] - h[at]) + (h[up + sj] - h[at + sj]));
  };
  // gc_bar(i, j, k): ppgu(i) and ppgu(i-1) read gc(i); ppgv(j) and ppgv(j-1) read gc(j).
  auto gcbar = [&](int k) -> double {
    if (!in_range(k, 0, nz)) return 0.0;
    double r = 0.0;
    if (i < nx && j < ny) r += 0.5 * hu(0, 0, k) * A(0, 0, k) + 0.5 * hv(0, 0, k) * B(0, 0, k);
    if (i >= 1 && j < ny) r += 0.5 * hu(-1, 0, k) * A(-1, 0, k);
    if (j >= 1 && i < nx) r += 0.5 * hv(0, -1, k) * B(0, -1, k);
    return r;
  };
  // gk(k) is read by gc(k) (-1) and gc(k-1) (+1).
  double gc_km1 = 0.0, gc_k = gcbar(-1);  // gc cotangents of levels k-1 and k (k = -1 here)
  double gc_kp1 = gcbar(0);
  for (int k = -1; k <= nz; ++k) {
    // gk_bar at levels k and k+1 (gk lives on [0, nz]).
    const double gk_k = (k >= 0) ? (gc_km1 - gc_k) : 0.0;
    const double gk_kp1 = (k + 1 <= nz) ? (gc_k - gc_kp1) : 0.0;
    const int64_t at = int64_t(k) * sk + col;
    double r = (A(-1, 0, k) - A(0, 0, k)) + (B(0, -1, k) - B(0, 0, k));
    if (k >= 0) r += double(a.wgtfac[at]) * gk_k;
    if (k + 1 <= nz) r += (1.0 - double(a.wgtfac[at + sk])) * gk_kp1;
    a.ppbar[at] = r;
    gc_km1 = gc_k;
    gc_k = gc_kp1;
    gc_kp1 = gcbar(k + 2);
  }
}

template <class T, class S3, class S1>
int fastwavesuv_adj_run(const oeb_domain *d, const S3 *wgtfac, const oeb_s3 *ppuv, const oeb_s3 *hhl, const S3 *rho,
                        const S1 *fx, const S3 *uout_bar, const S3 *vout_bar, const S3 *uin_bar, const S3 *vin_bar,
                        const oeb_s3 *ppuv_bar, T edadlat, T dt, void *stream) {
  const char *K = "fastwavesuv_adj";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz}, r_h{0, nx + 1, 0, ny + 1, 0, nz};
  const Box r_w{0, nx + 1, 0, ny + 1, 0, nz + 1}, r_p{0, nx + 1, 0, ny + 1, -1, nz + 1};
  FwAdj<T> a;
  a.fx_stride = 1;
  a.edadlat = edadlat;
  a.dt = dt;
  const T *p_ub, *p_vb;
  const double *p_pb;
  OEB_TRY(check3(K, "wgtfac", wgtfac, r_w, g, true, &a.wgtfac));
  OEB_TRY(check3(K, "ppuv", ppuv, r_p, g, false, &a.ppuv));
  OEB_TRY(check3(K, "hhl", hhl, r_w, g, false, &a.hhl));
  OEB_TRY(check3(K, "rho", rho, r_h, g, false, &a.rho));
  OEB_TRY(check1(K, "fx", fx, 0, ny, &a.fx, &a.fx_stride));
  OEB_TRY(check3(K, "uout_bar", uout_bar, r0, g, false, &a.ubar));
  OEB_TRY(check3(K, "vout_bar", vout_bar, r0, g, false, &a.vbar));
  OEB_TRY(check3(K, "uin_bar", uin_bar, r0, g, false, &p_ub));
  OEB_TRY(check3(K, "vin_bar", vin_bar, r0, g, false, &p_vb));
  OEB_TRY(check3(K, "ppuv_bar", ppuv_bar, r_p, g, false, &p_pb));
  AliasSet as;
  as.add("wgtfac", a.wgtfac, r_w, g, false);
  as.add("ppuv", a.ppuv, r_p, g, false);
  as.add("hhl", a.hhl, r_w, g, false);
  as.add("rho", a.rho, r_h, g, false);
  as.add("uout_bar", a.ubar, r0, g, false);
  as.add("vout_bar", a.vbar, r0, g, false);
  as.add("uin_bar", p_ub, r0, g, true);
  as.add("vin_bar", p_vb, r0, g, true);
  as.add("ppuv_bar", p_pb, r_p, g, true);
  OEB_TRY(as.check(K));
  a.uinbar = const_cast<T *>(p_ub);
  a.vinbar = const_cast<T *>(p_vb);
  a.ppbar = const_cast<double *>(p_pb);
  const BoxGrid bi{0, 0, int(nx), int(ny)}, bp{0, 0, int(nx + 1), int(ny + 1)};
  hipLaunchKernelGGL(fastwavesuv_adj_uv<T>, box_blocks(bi, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g,
                     bi);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(fastwavesuv_adj_pp<T>, box_blocks(bp, 1), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g, bp);
  return launch_status(K);
}

// ================================================================= hadvuv5th
// Forward (kernels/hadvuv5th.h:177-201), differentiated in uin and vin.  For
// the u output at p:
//   U  = (u(i-1) + u + u(i+1)) / 3,  Vu = (v(i+1) + v(i+1,j-1) + v + v(j-1)) / 4
//   ua = a0(j) U,  va = R Vu
//   X  = |ua| Sx,  Sx = sum_k c_k u(i+k), c = the branch's coefficients
//        (c+ on offsets -3..2 when ua > 0, c- on -2..3 otherwise)
//   Y  = |va| Sy  along j,   uout = eddlat X + eddlon Y + tg0(j) u Vu
// and the v output alike (Uv, V, ua2 = a1(j) Uv, va2 = R V,
// vout = eddlat X2 + eddlon Y2 - tg1(j) Uv^2).  With the branch frozen, X is
// linear in u through Sx with weight |ua| and in ua through Sx:
//   d X / d u(i+k) = |ua| c_k,   d X / d ua = s Sx,  s = +1 if ua > 0 else -1.
// Stage 1 (interior output points) stores in the scratch arrays: the averaged
// velocities ua, va, ua2, va2 (signed: their sign is the frozen branch) and
// the cotangents that flow through them and through the metric terms:
//   gu  = eddlat ob s Sx a0 / 3        (to u(i-1), u, u(i+1))
//   gv  = eddlon ob s Sy R / 4 + ob tg0 u / 4              (to Vu's 4 v)
//   gu2 = eddlat ob2 s2 Sx2 a1 / 4 - ob2 tg1 Uv / 2        (to Uv's 4 u)
//   gv2 = eddlon ob2 s2 Sy2 R / 3                          (to v(j-1), v, v(j+1))
// Stage 2 (input points on [-3, nx+3) x [-3, ny+3)) gathers: the transposed
// upwind stencils (weights eddlat ob |ua| etc. with each source point's own
// branch), the metric paths above, and ob tg0 Vu at the point itself.
template <class T>
struct HadvAdj {
  const T *uin, *vin, *a0, *a1, *t0, *t1, *ubar, *vbar;
  int s0, s1, s2, s3;
  T eddlat, eddlon;
  T *ua, *va, *ua2, *va2, *gu, *gv, *gu2, *gv2;  // scratch, interior
  T *uinbar, *vinbar;
};

// The branch's coefficient of offset k (-3..3) of the 5th-order stencil
// (kernels/hadvuv5th.h:10-36); 0 where the branch has no term.
template <class T>
__device__ __forceinline__ T c5(bool plus, int k) {
  if (plus) {
    switch (k) {
      case -3: return T(1.0) / T(30.0);
      case -2: return T(-1.0) / T(4.0);
      case -1: return T(1.0);
      case 0: return T(-1.0) / T(3.0);
      case 1: return T(-1.0) / T(2.0);
      case 2: return T(1.0) / T(20.0);
      default: return T(0);
    }
  }
  switch (k) {
    case -2: return T(1.0) / T(20.0);
    case -1: return T(-1.0) / T(2.0);
    case 0: return T(-1.0) / T(3.0);
    case 1: return T(1.0);
    case 2: return T(-1.0) / T(4.0);
    case 3: return T(1.0) / T(30.0);
    default: return T(0);
  }
}

// Sum of the branch's stencil along stride s, as the forward forms it
// (same terms in the same order, kernels/hadvuv5th.h:10-36).
template <class T>
__device__ __forceinline__ T s5(const T *f, int s, bool plus) {
  if (plus) {
    return (T(1.0) / T(30.0) * f[-3 * s] + T(-1.0) / T(4.0) * f[-2 * s] + f[-s] + T(-1.0) / T(3.0) * f[0] +
            T(-1.0) / T(2.0) * f[s] + T(1.0) / T(20.0) * f[2 * s]);
  }
  return (T(1.0) / T(20.0) * f[-2 * s] + T(-1.0) / T(2.0) * f[-s] + T(-1.0) / T(3.0) * f[0] + f[s] +
          T(-1.0) / T(4.0) * f[2 * s] + T(1.0) / T(30.0) * f[3 * s]);
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) hadvuv5th_adj_stage1(HadvAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const int sj = g.sj;
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * sj;
  constexpr T rr = kEarthRadiusRecipAdj<T>;
  const T *u = a.uin + at;
  const T *v = a.vin + at;
  const T ob = a.ubar[at], ob2 = a.vbar[at];
  const T a0 = a.a0[int64_t(j) * a.s0], a1 = a.a1[int64_t(j) * a.s1];
  const T tg0 = a.t0[int64_t(j) * a.s2], tg1 = a.t1[int64_t(j) * a.s3];
  // u output (kernels/hadvuv5th.h:183-188).
  const T U = (T(1.0) / T(3.0)) * (u[-1] + u[0] + u[1]);
  const T Vu = T(0.25) * (v[1] + v[1 - sj] + v[0] + v[-sj]);
  const T ua = a0 * U, va = rr * Vu;
  const bool px = ua > 0, py = va > 0;
  const T sx = s5(u, 1, px), sy = s5(u, sj, py);
  a.ua[at] = ua;
  a.va[at] = va;
  a.gu[at] = a.eddlat * ob * (px ? sx : -sx) * a0 * (T(1.0) / T(3.0));
  a.gv[at] = a.eddlon * ob * (py ? sy : -sy) * rr * T(0.25) + ob * tg0 * u[0] * T(0.25);
  // v output (kernels/hadvuv5th.h:190-195).
  const T Uv = T(0.25) * (u[-1] + u[0] + u[sj] + u[sj - 1]);
  const T V = T(1.0) / T(3.0) * (v[-sj] + v[0] + v[sj]);
  const T ua2 = a1 * Uv, va2 = rr * V;
  const bool px2 = ua2 > 0, py2 = va2 > 0;
  const T sx2 = s5(v, 1, px2), sy2 = s5(v, sj, py2);
  a.ua2[at] = ua2;
  a.va2[at] = va2;
  a.gu2[at] = a.eddlat * ob2 * (px2 ? sx2 : -sx2) * a1 * T(0.25) - ob2 * tg1 * Uv * T(0.5);
  a.gv2[at] = a.eddlon * ob2 * (py2 ? sy2 : -sy2) * rr * (T(1.0) / T(3.0));
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) hadvuv5th_adj_stage2(HadvAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const int sj = g.sj, nx = g.nx, ny = g.ny;
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * sj;
  auto inside = [&](int di, int dj) { return in_range(i + di, 0, nx) && in_range(j + dj, 0, ny); };
  auto ix = [&](int di, int dj) { return at + di + int64_t(dj) * sj; };
  T ru = T(0), rv = T(0);
  // Transposed upwind stencils: the output at p = q - k reads q with offset k.
#pragma unroll
  for (int k = -3; k <= 3; ++k) {
    if (inside(-k, 0)) {
      const int64_t p = ix(-k, 0);
      const T w = a.ua[p], w2 = a.ua2[p];
      ru += a.eddlat * a.ubar[p] * (w > 0 ? w : -w) * c5<T>(w > 0, k);
      rv += a.eddlat * a.vbar[p] * (w2 > 0 ? w2 : -w2) * c5<T>(w2 > 0, k);
    }
    if (inside(0, -k)) {
      const int64_t p = ix(0, -k);
      const T w = a.va[p], w2 = a.va2[p];
      ru += a.eddlon * a.ubar[p] * (w > 0 ? w : -w) * c5<T>(w > 0, k);
      rv += a.eddlon * a.vbar[p] * (w2 > 0 ? w2 : -w2) * c5<T>(w2 > 0, k);
    }
  }
  // U of the u output at (i+m) reads u(i) for m = -1, 0, 1.
  for (int m = -1; m <= 1; ++m)
    if (inside(m, 0)) ru += a.gu[ix(m, 0)];
  // Uv of the v output at p reads u at p, p-1, p+sj, p+sj-1: p in {q, q+1, q-sj, q-sj+1}.
  if (inside(0, 0)) ru += a.gu2[ix(0, 0)];
  if (inside(1, 0)) ru += a.gu2[ix(1, 0)];
  if (inside(0, -1)) ru += a.gu2[ix(0, -1)];
  if (inside(1, -1)) ru += a.gu2[ix(1, -1)];
  // The metric term tg0 u Vu at q itself.
  if (inside(0, 0)) {
    const T *v = a.vin + at;
    const T Vu = T(0.25) * (v[1] + v[1 - sj] + v[0] + v[-sj]);
    ru += a.ubar[at] * a.t0[int64_t(j) * a.s2] * Vu;
  }
  // Vu of the u output at p reads v at p+1, p+1-sj, p, p-sj: p in {q-1, q-1+sj, q, q+sj}.
  if (inside(-1, 0)) rv += a.gv[ix(-1, 0)];
  if (inside(-1, 1)) rv += a.gv[ix(-1, 1)];
  if (inside(0, 0)) rv += a.gv[ix(0, 0)];
  if (inside(0, 1)) rv += a.gv[ix(0, 1)];
  // V of the v output at (j+m) reads v(j) for m = -1, 0, 1.
  for (int m = -1; m <= 1; ++m)
    if (inside(0, m)) rv += a.gv2[ix(0, m)];
  a.uinbar[at] = ru;
  a.vinbar[at] = rv;
}

template <class T, class S3, class S1>
int hadvuv5th_adj_run(const oeb_domain *d, const S3 *uin, const S3 *vin, const S1 *acrlat0, const S1 *acrlat1,
                      const S1 *tgrlatda0, const S1 *tgrlatda1, const S3 *uout_bar, const S3 *vout_bar,
                      const S3 *uin_bar, const S3 *vin_bar, const S3 *const scratch[8], T eddlat, T eddlon,
                      void *stream) {
  const char *K = "hadvuv5th_adj";
  OEB_TRY(check_domain(K, d));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz}, r_in{-3, nx + 3, -3, ny + 3, 0, nz};
  HadvAdj<T> a;
  a.eddlat = eddlat;
  a.eddlon = eddlon;
  const T *p_ub, *p_vb, *p_s[8];
  OEB_TRY(check3(K, "uin", uin, r_in, g, true, &a.uin));
  OEB_TRY(check3(K, "vin", vin, r_in, g, false, &a.vin));
  OEB_TRY(check3(K, "uout_bar", uout_bar, r0, g, false, &a.ubar));
  OEB_TRY(check3(K, "vout_bar", vout_bar, r0, g, false, &a.vbar));
  OEB_TRY(check3(K, "uin_bar", uin_bar, r_in, g, false, &p_ub));
  OEB_TRY(check3(K, "vin_bar", vin_bar, r_in, g, false, &p_vb));
  static const char *const kScratch[8] = {"uatupos", "vatupos", "uatvpos", "vatvpos", "uavg", "vavg", "ures", "vres"};
  for (int q = 0; q < 8; ++q) OEB_TRY(check3(K, kScratch[q], scratch[q], r0, g, false, &p_s[q]));
  OEB_TRY(check1(K, "acrlat0", acrlat0, 0, ny, &a.a0, &a.s0));
  OEB_TRY(check1(K, "acrlat1", acrlat1, 0, ny, &a.a1, &a.s1));
  OEB_TRY(check1(K, "tgrlatda0", tgrlatda0, 0, ny, &a.t0, &a.s2));
  OEB_TRY(check1(K, "tgrlatda1", tgrlatda1, 0, ny, &a.t1, &a.s3));
  AliasSet as;
  as.add("uin", a.uin, r_in, g, false);
  as.add("vin", a.vin, r_in, g, false);
  as.add("uout_bar", a.ubar, r0, g, false);
  as.add("vout_bar", a.vbar, r0, g, false);
  as.add("uin_bar", p_ub, r_in, g, true);
  as.add("vin_bar", p_vb, r_in, g, true);
  for (int q = 0; q < 8; ++q) as.add(kScratch[q], p_s[q], r0, g, true);
  OEB_TRY(as.check(K));
  T **dst[8] = {&a.ua, &a.va, &a.ua2, &a.va2, &a.gu, &a.gv, &a.gu2, &a.gv2};
  for (int q = 0; q < 8; ++q) *dst[q] = const_cast<T *>(p_s[q]);
  a.uinbar = const_cast<T *>(p_ub);
  a.vinbar = const_cast<T *>(p_vb);
  const BoxGrid b1{0, 0, int(nx), int(ny)}, b2{-3, -3, int(nx + 6), int(ny + 6)};
  hipLaunchKernelGGL(hadvuv5th_adj_stage1<T>, box_blocks(b1, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g,
                     b1);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(hadvuv5th_adj_stage2<T>, box_blocks(b2, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g,
                     b2);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_hdiffsa_adj_f64(const oeb_domain *d, const oeb_s3 *in, const oeb_s3 *mask, const oeb_s1 *crlato,
                                   const oeb_s1 *crlatu, const oeb_s3 *out_bar, const oeb_s3 *in_bar,
                                   const oeb_s3 *lap, void *stream) {
  return oeb::hdiffsa_adj_run<double>(d, in, mask, crlato, crlatu, out_bar, in_bar, lap, stream);
}

extern "C" int oeb_hdiffsa_adj_f32(const oeb_domain *d, const oeb_s3f *in, const oeb_s3f *mask,
                                   const oeb_s1f *crlato, const oeb_s1f *crlatu, const oeb_s3f *out_bar,
                                   const oeb_s3f *in_bar, const oeb_s3f *lap, void *stream) {
  return oeb::hdiffsa_adj_run<float>(d, in, mask, crlato, crlatu, out_bar, in_bar, lap, stream);
}

extern "C" int oeb_fastwavesuv_adj_f64(const oeb_domain *d, const oeb_s3 *wgtfac, const oeb_s3 *ppuv,
                                       const oeb_s3 *hhl, const oeb_s3 *rho, const oeb_s1 *fx,
                                       const oeb_s3 *uout_bar, const oeb_s3 *vout_bar, const oeb_s3 *uin_bar,
                                       const oeb_s3 *vin_bar, const oeb_s3 *ppuv_bar, double edadlat, double dt,
                                       void *stream) {
  return oeb::fastwavesuv_adj_run<double>(d, wgtfac, ppuv, hhl, rho, fx, uout_bar, vout_bar, uin_bar, vin_bar,
                                          ppuv_bar, edadlat, dt, stream);
}

extern "C" int oeb_fastwavesuv_adj_f32(const oeb_domain *d, const oeb_s3f *wgtfac, const oeb_s3 *ppuv,
                                       const oeb_s3 *hhl, const oeb_s3f *rho, const oeb_s1f *fx,
                                       const oeb_s3f *uout_bar, const oeb_s3f *vout_bar, const oeb_s3f *uin_bar,
                                       const oeb_s3f *vin_bar, const oeb_s3 *ppuv_bar, float edadlat, float dt,
                                       void *stream) {
  return oeb::fastwavesuv_adj_run<float>(d, wgtfac, ppuv, hhl, rho, fx, uout_bar, vout_bar, uin_bar, vin_bar,
                                         ppuv_bar, edadlat, dt, stream);
}

extern "C" int oeb_hadvuv5th_adj_f64(const oeb_domain *d, const oeb_s3 *uin, const oeb_s3 *vin, const oeb_s1 *acrlat0,
                                     const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0, const oeb_s1 *tgrlatda1,
                                     const oeb_s3 *uout_bar, const oeb_s3 *vout_bar, const oeb_s3 *uin_bar,
                                     const oeb_s3 *vin_bar, const oeb_s3 *uatupos, const oeb_s3 *vatupos,
                                     const oeb_s3 *uatvpos, const oeb_s3 *vatvpos, const oeb_s3 *uavg,
                                     const oeb_s3 *vavg, const oeb_s3 *ures, const oeb_s3 *vres, double eddlat,
                                     double eddlon, void *stream) {
  const oeb_s3 *const sc[8] = {uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres};
  return oeb::hadvuv5th_adj_run<double>(d, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uout_bar, vout_bar,
                                        uin_bar, vin_bar, sc, eddlat, eddlon, stream);
}

extern "C" int oeb_hadvuv5th_adj_f32(const oeb_domain *d, const oeb_s3f *uin, const oeb_s3f *vin,
                                     const oeb_s1f *acrlat0, const oeb_s1f *acrlat1, const oeb_s1f *tgrlatda0,
                                     const oeb_s1f *tgrlatda1, const oeb_s3f *uout_bar, const oeb_s3f *vout_bar,
                                     const oeb_s3f *uin_bar, const oeb_s3f *vin_bar, const oeb_s3f *uatupos,
                                     const oeb_s3f *vatupos, const oeb_s3f *uatvpos, const oeb_s3f *vatvpos,
                                     const oeb_s3f *uavg, const oeb_s3f *vavg, const oeb_s3f *ures,
                                     const oeb_s3f *vres, float eddlat, float eddlon, void *stream) {
  const oeb_s3f *const sc[8] = {uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres};
  return oeb::hadvuv5th_adj_run<float>(d, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, uout_bar, vout_bar,
                                       uin_bar, vin_bar, sc, eddlat, eddlon, stream);
}
