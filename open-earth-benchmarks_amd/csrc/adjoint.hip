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
#include <algorithm>
#include <mutex>
#include <sstream>
#include <vector>

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

// Fused row-march form, used when one strip spans nx (nx + 6 <= 1024): lane
// l owns column i = l - 3 of [-3, nx + 3) of one level, and the workgroup
// walks the in_bar rows [j0, j1) of its chunk of [-2, ny + 2).  With front row
// r it forms lap(r) (on [-1, nx+1) x [-1, ny+1)), lap_bar(r - 1) and
// in_bar(r - 2), so lap_bar never goes to memory: in, mask and out_bar are
// read once per chunk and in_bar is written once.  Each lane keeps rows of
// in, lap, lap_bar and mask x out_bar of its column in registers; the i
// neighbours come from LDS rows: in (3 slots, published before the barrier
// of iteration t and read in t and t + 1), mask x out_bar (2 slots, read in
// t), lap and lap_bar (2 slots each, written after the barrier of t and read
// in t + 1).  Values outside their boxes are exact zeros, as in the two-pass
// form; the arithmetic is that of hdiffsa_adj_lap and hdiffsa_adj_in, term
// for term.  Loads run two rows ahead, with addresses clamped into the
// arrays' boxes; a clamped value is never used.  crlato and crlatu of the
// chunk's rows are copied to LDS tables first: loaded where they are used,
// under lane conditions, they would be vector loads whose waits also wait for
// the rows loaded ahead.
template <class T>
__global__ void __launch_bounds__(1024) hdiffsa_adj_strip(HsaAdj<T> a, Geo g, int rows, int nchunk) {
  extern __shared__ __align__(16) unsigned char oeb_smem[];
  const int L = int(blockDim.x);
  T *s_in = reinterpret_cast<T *>(oeb_smem);  // 3 x L
  T *s_mo = s_in + 3 * L;                      // 2 x L
  T *s_lap = s_mo + 2 * L;                     // 2 x L
  T *s_lb = s_lap + 2 * L;                     // 2 x L
  T *s_lo = s_lb + 2 * L;                      // rows + 7: crlato of rows [j0 - 5, j1 + 2)
  T *s_lu = s_lo + rows + 7;                   // crlatu alike
  const int nx = g.nx, ny = g.ny, sj = g.sj;
  const int cj = int(blockIdx.x) % nchunk, k = int(blockIdx.x) / nchunk;
  const int j0 = -2 + cj * rows, j1 = min(ny + 2, j0 + rows);
  const int l = int(threadIdx.x), i = l - 3;
  const int tb = j0 - 5;  // first row of the tables; read after the first barrier
  for (int e = l; e < j1 - j0 + 7; e += L) {
    const int64_t row = min(max(tb + e, -1), ny);  // crlato, crlatu live on [-1, ny + 1)
    s_lo[e] = a.lo[row * a.slo];
    s_lu[e] = a.lu[row * a.slu];
  }
  const bool col_in = i >= -2 && i < nx + 2;   // in_bar columns
  const bool col_lap = i >= -1 && i < nx + 1;  // lap, lap_bar columns
  const bool col_int = i >= 0 && i < nx;       // interior columns
  const int64_t plane = int64_t(k) * g.sk;
  const int64_t c_in = plane + min(max(i, -2), nx + 1), c_mo = plane + min(max(i, 0), nx - 1);
  auto ld_in = [&](int r) -> T { return a.in[c_in + int64_t(min(max(r, -2), ny + 1)) * sj]; };
  auto ld_mk = [&](int r) -> T { return a.mask[c_mo + int64_t(min(max(r, 0), ny - 1)) * sj]; };
  auto ld_ob = [&](int r) -> T { return a.obar[c_mo + int64_t(min(max(r, 0), ny - 1)) * sj]; };
  auto lo = [&](int jj) { return s_lo[jj - tb]; };
  auto lu = [&](int jj) { return s_lu[jj - tb]; };
  // mask x out_bar at (i, r), 0 outside the interior (hdiffsa_adj_lap's mo).
  auto mo = [&](int r, T mk, T ob) { return (col_int && r >= 0 && r < ny) ? mk * ob : T(0); };
  constexpr int D = 2;
  const int r0 = j0 - 2, nt = j1 - j0 + 4;  // front rows r0 .. j1 + 1
  T x_m2 = ld_in(r0 - 2), x_m1 = ld_in(r0 - 1), x_0 = ld_in(r0);
  T mk_m2 = ld_mk(r0 - 2), ob_m2 = ld_ob(r0 - 2), mk_m1 = ld_mk(r0 - 1), ob_m1 = ld_ob(r0 - 1);
  T lap_m2 = T(0), lap_m1 = T(0), lb_m3 = T(0), lb_m2 = T(0);
  T cin[D], cmk[D], cob[D];  // in(r + 1), mask(r), out_bar(r) of iteration t0 + u
#pragma unroll
  for (int u = 0; u < D; ++u) {
    cin[u] = ld_in(r0 + 1 + u);
    cmk[u] = ld_mk(r0 + u);
    cob[u] = ld_ob(r0 + u);
  }
  for (int t0 = 0; t0 < nt; t0 += D) {
    T nin[D], nmk[D], nob[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      nin[u] = ld_in(r0 + 1 + D + t0 + u);
      nmk[u] = ld_mk(r0 + D + t0 + u);
      nob[u] = ld_ob(r0 + D + t0 + u);
    }
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int t = t0 + u;
      if (t < nt) {  // uniform across the workgroup
        const int r = r0 + t;
        const T x_p1 = cin[u], mk_0 = cmk[u], ob_0 = cob[u];
        T *si = s_in + (t % 3) * L + l;                // in row r
        const T *si1 = s_in + ((t + 2) % 3) * L + l;   // in row r - 1
        T *sm = s_mo + (t & 1) * L + l;                // mo row r - 1
        const T mo_m2 = mo(r - 2, mk_m2, ob_m2), mo_m1 = mo(r - 1, mk_m1, ob_m1), mo_0 = mo(r, mk_0, ob_0);
        si[0] = x_0;
        sm[0] = mo_m1;
        lds_barrier();
        // lap(r), kernels/hdiffsa.h:147 (hsa_lap).
        T lap_0 = T(0);
        if (col_lap && r >= -1 && r < ny + 1)
          lap_0 = si[-1] + si[1] - T(2.0) * x_0 + lo(r) * (x_p1 - x_0) + lu(r) * (x_m1 - x_0);
        s_lap[(t & 1) * L + l] = lap_0;
        // lap_bar(i, jj), jj = r - 1 (hdiffsa_adj_lap).
        const int jj = r - 1;
        T lb_m1 = T(0);
        if (col_lap && jj >= -1 && jj < ny + 1) {
          const T *sl = s_lap + ((t + 1) & 1) * L + l;  // lap row jj
          const T l0 = lap_m1;
          T rr = T(0);
          if (jj >= 0 && jj < ny) {
            if (i - 1 >= -1) {
              const T f = l0 - sl[-1];
              if (!(f * (x_m1 - si1[-1]) > 0)) rr += mo_m1 - sm[-1];
            }
            if (i < nx) {
              const T f = sl[1] - l0;
              if (!(f * (si1[1] - x_m1) > 0)) rr -= sm[1] - mo_m1;
            }
          }
          if (i >= 0 && i < nx) {
            if (jj - 1 >= -1) {
              const T f = lo(jj - 1) * (l0 - lap_m2);
              if (!(f * (x_m1 - x_m2) > 0)) rr += lo(jj - 1) * (mo_m1 - mo_m2);
            }
            if (jj < ny) {
              const T f = lo(jj) * (lap_0 - l0);
              if (!(f * (x_0 - x_m1) > 0)) rr -= lo(jj) * (mo_0 - mo_m1);
            }
          }
          lb_m1 = rr;
        }
        s_lb[(t & 1) * L + l] = lb_m1;
        // in_bar(i, q), q = r - 2 (hdiffsa_adj_in).
        const int q = r - 2;
        if (col_in && q >= j0 && q < j1) {
          const T *sb = s_lb + ((t + 1) & 1) * L + l;  // lap_bar row q
          T rr = (col_int && q >= 0 && q < ny) ? ob_m2 : T(0);
          if (col_lap && q >= -1 && q < ny + 1) rr += lb_m2 * (T(-2.0) - lo(q) - lu(q));
          rr += (i - 1 >= -1 ? sb[-1] : T(0)) + (i + 1 < nx + 1 ? sb[1] : T(0));
          if (q - 1 >= -1 && q - 1 < ny + 1) rr += lb_m3 * lo(q - 1);
          if (q + 1 >= -1 && q + 1 < ny + 1) rr += lb_m1 * lu(q + 1);
          a.inbar[plane + i + int64_t(q) * sj] = rr;
        }
        x_m2 = x_m1;
        x_m1 = x_0;
        x_0 = x_p1;
        mk_m2 = mk_m1;
        ob_m2 = ob_m1;
        mk_m1 = mk_0;
        ob_m1 = ob_0;
        lap_m2 = lap_m1;
        lap_m1 = lap_0;
        lb_m3 = lb_m2;
        lb_m2 = lb_m1;
      }
    }
#pragma unroll
    for (int u = 0; u < D; ++u) {
      cin[u] = nin[u];
      cmk[u] = nmk[u];
      cob[u] = nob[u];
    }
  }
}

template <class T, class S3, class S1>
int hdiffsa_adj_run(const oeb_domain *d, const S3 *in, const S3 *mask, const S1 *crlato, const S1 *crlatu,
                    const S3 *out_bar, const S3 *in_bar, const S3 *lap, void *stream) {
  const char *K = "hdiffsa_adj";
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
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
  if (nx + 6 <= 1024 && knob_on(K, sizeof(T), d, KNOB_FORM, true)) {
    const int lanes = int((nx + 6 + 63) / 64 * 64);
    // Row chunks measured at 256x256x80 over 16-64: 24 rows fastest in both
    // forms (FP64 58.5 us against 61 with 32; FP32 45.8 us against 48.5 with
    // 16).  At most 256 rows, which bounds the crlato / crlatu tables in LDS.
    const int rows = std::min(knob(K, sizeof(T), d, KNOB_ROWS, 24), 256);
    const int nchunk = int((ny + 4 + rows - 1) / rows);
    if (int64_t(nchunk) * nz > (int64_t(1) << 30)) return fail(OEB_ERR_DOMAIN, std::string(K) + ": domain too large");
    hipLaunchKernelGGL(hdiffsa_adj_strip<T>, dim3(unsigned(nchunk * nz)), dim3(unsigned(lanes)),
                       (size_t(9) * lanes + 2 * size_t(rows + 7)) * sizeof(T), hipStream_t(stream), a, g, rows,
                       nchunk);
    return launch_status(K);
  }
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

// One lane per column (i, j) of [0, nx+1) x [0, ny+1) and a chunk of levels
// [kb, ke) of the ppuv cotangent's levels [-1, nz+1) (grid.y = chunks).  Per
// level the lane forms A(i, j), A(i-1, j), B(i, j), B(i, j-1) (the ppgu /
// ppgv cotangents around it) and the two hhl ratios that weigh each, and from
// them gc_bar(k) and the horizontal part of ppuv_bar(k); it carries gc_bar of
// levels k-1, k, k+1 in registers.  A chunk starts one level early to have
// gc_bar(kb - 1).  Interior columns also write uin_bar = uout_bar and
// vin_bar = vout_bar on their interior levels.
template <class T>
__global__ void __launch_bounds__(kTI *kTJ) fastwavesuv_adj_kernel(FwAdj<T> a, Geo g, BoxGrid b, int kc) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const int sj = g.sj, nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t sk = g.sk, col = i + int64_t(j) * sj;
  const double dt = a.dt;
  const bool in_i = i < nx, in_j = j < ny, has_w = i >= 1 && in_j, has_s = j >= 1 && in_i;
  const bool interior = in_i && in_j;
  const T fx_c = in_j ? a.fx[int64_t(j) * a.fx_stride] : T(0);
  const T fx_w = fx_c;  // A(i - 1, j) has the same row
  // The cotangents and hhl ratios of one level (0 outside the interior).
  struct Lv {
    double gc, hs;  // gc_bar(k), (A(-1,0) - A(0,0)) + (B(0,-1) - B(0,0))
  };
  auto level = [&](int k) -> Lv {
    Lv r{0.0, 0.0};
    if (k < 0 || k >= nz) return r;
    const int64_t at = int64_t(k) * sk + col, up = at + sk;
    const double *h = a.hhl;
    double A0 = 0.0, Aw = 0.0, B0 = 0.0, Bs = 0.0;
    if (interior) {
      const T cu = T(2.0) * fx_c / (a.rho[at + 1] + a.rho[at]);
      A0 = -dt * double(cu) * double(a.ubar[at]);
      const T cv = T(2.0) * a.edadlat / (a.rho[at + sj] + a.rho[at]);
      B0 = -dt * double(cv) * double(a.vbar[at]);
      const double hu = ((h[up] + h[at]) - (h[up + 1] + h[at + 1])) / ((h[up] - h[at]) + (h[up + 1] - h[at + 1]));
      const double hv =
          ((h[up] + h[at]) - (h[up + sj] + h[at + sj])) / ((h[up] - h[at]) + (h[up + sj] - h[at + sj]));
      r.gc += 0.5 * hu * A0 + 0.5 * hv * B0;
    }
    if (has_w) {
      const int64_t aw = at - 1, uw = aw + sk;
      const T cu = T(2.0) * fx_w / (a.rho[aw + 1] + a.rho[aw]);
      Aw = -dt * double(cu) * double(a.ubar[aw]);
      const double hu = ((h[uw] + h[aw]) - (h[uw + 1] + h[aw + 1])) / ((h[uw] - h[aw]) + (h[uw + 1] - h[aw + 1]));
      r.gc += 0.5 * hu * Aw;
    }
    if (has_s) {
      const int64_t as = at - sj, us = as + sk;
      const T cv = T(2.0) * a.edadlat / (a.rho[as + sj] + a.rho[as]);
      Bs = -dt * double(cv) * double(a.vbar[as]);
      const double hv =
          ((h[us] + h[as]) - (h[us + sj] + h[as + sj])) / ((h[us] - h[as]) + (h[us + sj] - h[as + sj]));
      r.gc += 0.5 * hv * Bs;
    }
    r.hs = (Aw - A0) + (Bs - B0);
    return r;
  };
  const int kb = -1 + int(blockIdx.y) * kc, ke = min(kb + kc, nz + 1);
  Lv lm1 = level(kb - 1), l0 = level(kb);
  for (int k = kb; k < ke; ++k) {
    const Lv lp1 = level(k + 1);
    // gk(k) is read by gc(k) (-1) and gc(k-1) (+1); gk lives on [0, nz].
    const double gk_k = (k >= 0) ? (lm1.gc - l0.gc) : 0.0;
    const double gk_kp1 = (k + 1 <= nz) ? (l0.gc - lp1.gc) : 0.0;
    const int64_t at = int64_t(k) * sk + col;
    double r = l0.hs;
    if (k >= 0) r += double(a.wgtfac[at]) * gk_k;
    if (k + 1 <= nz) r += (1.0 - double(a.wgtfac[at + sk])) * gk_kp1;
    a.ppbar[at] = r;
    if (interior && k >= 0 && k < nz) {
      a.uinbar[at] = a.ubar[at];
      a.vinbar[at] = a.vbar[at];
    }
    lm1 = l0;
    l0 = lp1;
  }
}

template <class T, class S3, class S1>
int fastwavesuv_adj_run(const oeb_domain *d, const S3 *wgtfac, const oeb_s3 *ppuv, const oeb_s3 *hhl, const S3 *rho,
                        const S1 *fx, const S3 *uout_bar, const S3 *vout_bar, const S3 *uin_bar, const S3 *vin_bar,
                        const oeb_s3 *ppuv_bar, T edadlat, T dt, void *stream) {
  const char *K = "fastwavesuv_adj";
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
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
  const BoxGrid bp{0, 0, int(nx + 1), int(ny + 1)};
  const int kc = knob(K, sizeof(T), d, KNOB_KCHUNK, 8);
  const int nchunk = int((nz + 2 + kc - 1) / kc);
  hipLaunchKernelGGL(fastwavesuv_adj_kernel<T>, box_blocks(bp, nchunk), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g,
                     bp, kc);
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

// s5 with both branches formed from the seven values f(-3s) .. f(3s), read
// once, and one of them selected: lanes of a wave whose branches differ run
// one path instead of both in turn.  The same terms in the same order as s5.
template <class T>
__device__ __forceinline__ T s5sel(const T *f, int s, bool plus) {
  const T fm3 = f[-3 * s], fm2 = f[-2 * s], fm1 = f[-s], f0 = f[0], fp1 = f[s], fp2 = f[2 * s], fp3 = f[3 * s];
  const T up = (T(1.0) / T(30.0) * fm3 + T(-1.0) / T(4.0) * fm2 + fm1 + T(-1.0) / T(3.0) * f0 +
                T(-1.0) / T(2.0) * fp1 + T(1.0) / T(20.0) * fp2);
  const T dn = (T(1.0) / T(20.0) * fm2 + T(-1.0) / T(2.0) * fm1 + T(-1.0) / T(3.0) * f0 + fp1 +
                T(-1.0) / T(4.0) * fp2 + T(1.0) / T(30.0) * fp3);
  return plus ? up : dn;
}

// s5 on seven values f(-3) .. f(3) passed by value (the same terms in the same
// order), for register windows: taking a window's address would place it in
// scratch memory.  (In the FP32 row march, forming both branches and selecting,
// in s5v and in the x sums, took 88.1 / 377.0 us against 71.9 / 304.2 at
// 256x256x80 / 512x512x80, where the forward gained from it.)
template <class T>
__device__ __forceinline__ T s5v(T fm3, T fm2, T fm1, T f0, T fp1, T fp2, T fp3, bool plus) {
  if (plus)
    return (T(1.0) / T(30.0) * fm3 + T(-1.0) / T(4.0) * fm2 + fm1 + T(-1.0) / T(3.0) * f0 + T(-1.0) / T(2.0) * fp1 +
            T(1.0) / T(20.0) * fp2);
  return (T(1.0) / T(20.0) * fm2 + T(-1.0) / T(2.0) * fm1 + T(-1.0) / T(3.0) * f0 + fp1 + T(-1.0) / T(4.0) * fp2 +
          T(1.0) / T(30.0) * fp3);
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

// Stage 1 of hadvuv5th_adj_stage1 at one point (i, p) of the row march:
// the quantities the gathers need, 0 outside the interior.
template <class T>
struct HadvSt1 {
  T wux, wvx, wuy, wvy;  // eddlat ob |ua|, eddlat ob2 |ua2|, eddlon ob |va|, eddlon ob2 |va2|
  T gu, gv, gu2, gv2, tgt;
  int br;        // bit 0: ua > 0, bit 1: ua2 > 0
  bool py, py2;  // va > 0, va2 > 0
};

// Fused row-march form (the FP32 default; FP64 takes the tiled form below),
// used when one strip spans nx (nx + 12 <= 640, so that the FP64 form keeps
// its registers without spilling): lane l owns
// column i = l - 6 of [-6, nx + 6) of one level; the workgroup walks the
// output rows p of [j0 - 3, j1 + 3) for the uin_bar / vin_bar rows [j0, j1)
// of its chunk of [-3, ny + 3).  Per row p it gathers the contributions of
// the stage-1 quantities of hadvuv5th_adj_stage1 at (i, p) without memory:
//  * along i (the x upwind stencils and the averages U, Uv, Vu), from one
//    LDS row: lane i gathers the contributions of the outputs (i - 3 .. i + 3,
//    p) to the inputs (i, p), and the 2x2 contributions of gu2 and gv to rows
//    p + 1 and p - 1;
//  * along j (the y upwind stencils, V and the rows above and below), by
//    pushing them into a per-lane queue of the 7 input rows p - 3 .. p + 3 of
//    its own column.
// Row p - 3 of the queue is then complete and is written.  Stage 1 runs one
// row ahead: iteration t publishes, before its one barrier, the u row p + 2
// and v row p + 1 (3 LDS slots each) and the stage-1 row p formed in
// iteration t - 1 (2 slots); after the barrier it gathers row p and forms
// stage 1 of row p + 1, two independent chains.  acrlat0/1 and tgrlatda0/1 of
// the chunk's rows come from LDS tables.  Rows of u and v are loaded D rows
// ahead.  The terms are those of the two-pass form; they are summed in
// another order.  The FP64 form needs 133 registers, which would hold 3
// waves per SIMD; asked for 4 (amdgpu_waves_per_eu), it keeps 128 and spills
// 24 bytes per lane, and measured 1.1x faster at 256x256x80.
template <class T, int D>
__global__ void __launch_bounds__(640) __attribute__((amdgpu_waves_per_eu(4)))
    hadvuv5th_adj_strip(HadvAdj<T> a, Geo g, int rows, int nchunk) {
  extern __shared__ __align__(16) unsigned char oeb_smem[];
  constexpr T rr = kEarthRadiusRecipAdj<T>;
  const int L = int(blockDim.x);
  T *s_u = reinterpret_cast<T *>(oeb_smem);  // 3 x L: u rows
  T *s_v = s_u + 3 * L;                       // 3 x L: v rows
  T *s_wux = s_v + 3 * L;                     // 2 x L each: stage-1 rows
  T *s_wvx = s_wux + 2 * L;
  T *s_gu = s_wvx + 2 * L;
  T *s_gu2 = s_gu + 2 * L;
  T *s_gv = s_gu2 + 2 * L;
  T *s_a0 = s_gv + 2 * L;  // rows + 10 each: rows [j0 - 6, j1 + 4)
  T *s_a1 = s_a0 + rows + 10;
  T *s_t0 = s_a1 + rows + 10;
  T *s_t1 = s_t0 + rows + 10;
  int *s_br = reinterpret_cast<int *>(s_t1 + rows + 10);  // 2 x L
  const int nx = g.nx, ny = g.ny, sj = g.sj;
  const int cj = int(blockIdx.x) % nchunk, k = int(blockIdx.x) / nchunk;
  const int j0 = -3 + cj * rows, j1 = min(ny + 3, j0 + rows);
  const int l = int(threadIdx.x), i = l - 6;
  const bool col_io = i >= -3 && i < nx + 3;  // uin_bar / vin_bar columns
  const bool col_out = i >= 0 && i < nx;      // output (stage-1) columns
  const int tb = j0 - 6;                      // first row of the tables
  for (int e = l; e < j1 - j0 + 10; e += L) {
    const int64_t r = min(max(tb + e, 0), ny - 1);  // the 1-D arrays live on [0, ny)
    s_a0[e] = a.a0[r * a.s0];
    s_a1[e] = a.a1[r * a.s1];
    s_t0[e] = a.t0[r * a.s2];
    s_t1[e] = a.t1[r * a.s3];
  }
  const int64_t plane = int64_t(k) * g.sk;
  const int64_t c_uv = plane + min(max(i, -3), nx + 2), c_o = plane + min(max(i, 0), nx - 1);
  auto ldu = [&](int r) -> T { return a.uin[c_uv + int64_t(min(max(r, -3), ny + 2)) * sj]; };
  auto ldv_ = [&](int r) -> T { return a.vin[c_uv + int64_t(min(max(r, -3), ny + 2)) * sj]; };
  auto ldub = [&](int r) -> T { return a.ubar[c_o + int64_t(min(max(r, 0), ny - 1)) * sj]; };
  auto ldvb = [&](int r) -> T { return a.vbar[c_o + int64_t(min(max(r, 0), ny - 1)) * sj]; };
  // Stage 1 at (i, p) (kernels/hadvuv5th.h:183-195), given the LDS rows u(p),
  // u(p + 1), v(p), v(p - 1) and the column windows w (rows p - 3 .. p + 3).
  auto stage1 = [&](int p, const T *u, const T *u1, const T *v, const T *vm, const T (&ur)[7], const T (&vr)[7],
                    T ob_c, T ob2_c) {
    HadvSt1<T> S{};
    if (col_out && p >= 0 && p < ny) {
      const T a0 = s_a0[p - tb], a1 = s_a1[p - tb], tg0 = s_t0[p - tb], tg1 = s_t1[p - tb];
      const T U = (T(1.0) / T(3.0)) * (u[-1] + u[0] + u[1]);
      const T Vu = T(0.25) * (v[1] + vm[1] + v[0] + vm[0]);
      const T ua = a0 * U, va = rr * Vu;
      const bool px = ua > 0;
      S.py = va > 0;
      const T sx = s5(u, 1, px), sy = s5v(ur[0], ur[1], ur[2], ur[3], ur[4], ur[5], ur[6], S.py);
      S.wux = a.eddlat * ob_c * (px ? ua : -ua);
      S.wuy = a.eddlon * ob_c * (S.py ? va : -va);
      S.gu = a.eddlat * ob_c * (px ? sx : -sx) * a0 * (T(1.0) / T(3.0));
      S.gv = a.eddlon * ob_c * (S.py ? sy : -sy) * rr * T(0.25) + ob_c * tg0 * u[0] * T(0.25);
      S.tgt = ob_c * tg0 * Vu;
      const T Uv = T(0.25) * (u[-1] + u[0] + u1[0] + u1[-1]);
      const T V = T(1.0) / T(3.0) * (vr[2] + vr[3] + vr[4]);
      const T ua2 = a1 * Uv, va2 = rr * V;
      const bool px2 = ua2 > 0;
      S.py2 = va2 > 0;
      const T sx2 = s5(v, 1, px2), sy2 = s5v(vr[0], vr[1], vr[2], vr[3], vr[4], vr[5], vr[6], S.py2);
      S.wvx = a.eddlat * ob2_c * (px2 ? ua2 : -ua2);
      S.wvy = a.eddlon * ob2_c * (S.py2 ? va2 : -va2);
      S.gu2 = a.eddlat * ob2_c * (px2 ? sx2 : -sx2) * a1 * T(0.25) - ob2_c * tg1 * Uv * T(0.5);
      S.gv2 = a.eddlon * ob2_c * (S.py2 ? sy2 : -sy2) * rr * (T(1.0) / T(3.0));
      S.br = (px ? 1 : 0) | (px2 ? 2 : 0);
    }
    return S;
  };
  const int p0 = j0 - 3, nt = j1 - j0 + 6;  // output rows p0 .. j1 + 2
  T ur[7], vr[7];                            // rows p0 - 3 .. p0 + 3, then one row ahead
#pragma unroll
  for (int x = 0; x < 7; ++x) {
    ur[x] = ldu(p0 - 3 + x);
    vr[x] = ldv_(p0 - 3 + x);
  }
  T qu[7], qv[7];  // partial cotangents of input rows p - 3 .. p + 3
#pragma unroll
  for (int x = 0; x < 7; ++x) qu[x] = qv[x] = T(0);
  // Stage 1 of row p0 from slots 1 (u(p0), v(p0 - 1)) and 2 (u(p0 + 1), v(p0)).
  s_u[1 * L + l] = ur[3];
  s_u[2 * L + l] = ur[4];
  s_v[1 * L + l] = vr[2];
  s_v[2 * L + l] = vr[3];
  lds_barrier();
  HadvSt1<T> S = stage1(p0, s_u + L + l, s_u + 2 * L + l, s_v + 2 * L + l, s_v + L + l, ur, vr, ldub(p0), ldvb(p0));
#pragma unroll
  for (int x = 0; x < 6; ++x) {
    ur[x] = ur[x + 1];
    vr[x] = vr[x + 1];
  }
  ur[6] = ldu(p0 + 4);
  vr[6] = ldv_(p0 + 4);
  // One iteration, given the rows it loads: u, v at p + 5 and out_bar at p + 1.
  auto step = [&](int t, T u_next, T v_next, T ob_n, T ob2_n) {
    const int p = p0 + t;
    const int s3 = (t % 3) * L + l, s3m = ((t + 2) % 3) * L + l, s2 = (t & 1) * L + l;
    s_u[s3] = ur[4];  // u row p + 2 (ur holds rows p - 2 .. p + 4)
    s_v[s3] = vr[3];  // v row p + 1
    s_wux[s2] = S.wux;
    s_wvx[s2] = S.wvx;
    s_gu[s2] = S.gu;
    s_gu2[s2] = S.gu2;
    s_gv[s2] = S.gv;
    s_br[s2] = S.br;
    lds_barrier();
    // Along i: row p of the inputs gathers from outputs (i - 3 .. i + 3, p).
    if (col_io) {
      T ru = T(0), rv = T(0);
#pragma unroll
      for (int kk = -3; kk <= 3; ++kk) {
        const int br = s_br[s2 - kk];
        ru += s_wux[s2 - kk] * c5<T>((br & 1) != 0, kk);
        rv += s_wvx[s2 - kk] * c5<T>((br & 2) != 0, kk);
      }
      ru += s_gu[s2 - 1] + s_gu[s2] + s_gu[s2 + 1];
      const T g2 = s_gu2[s2] + s_gu2[s2 + 1];  // to (i, p) and (i, p + 1)
      const T hv = s_gv[s2 - 1] + s_gv[s2];    // to (i, p) and (i, p - 1)
      qu[3] += ru + g2 + S.tgt;
      qu[4] += g2;
      qv[3] += rv + hv;
      qv[2] += hv;
    }
    // Along j: output (i, p) to inputs (i, p + kk) of its own column.
#pragma unroll
    for (int kk = -3; kk <= 3; ++kk) {
      qu[3 + kk] += S.wuy * c5<T>(S.py, kk);
      qv[3 + kk] += S.wvy * c5<T>(S.py2, kk);
    }
    qv[2] += S.gv2;
    qv[3] += S.gv2;
    qv[4] += S.gv2;
    // Stage 1 of row p + 1 from u(p + 1) (slot of t - 1), u(p + 2), v(p + 1) and
    // v(p) (slot of t - 1).
    S = stage1(p + 1, s_u + s3m, s_u + s3, s_v + s3, s_v + s3m, ur, vr, ob_n, ob2_n);
    // Input row p - 3 has all its contributions.
    const int q = p - 3;
    if (col_io && q >= j0 && q < j1) {
      const int64_t at = plane + i + int64_t(q) * sj;
      a.uinbar[at] = qu[0];
      a.vinbar[at] = qv[0];
    }
#pragma unroll
    for (int x = 0; x < 6; ++x) {
      qu[x] = qu[x + 1];
      qv[x] = qv[x + 1];
      ur[x] = ur[x + 1];
      vr[x] = vr[x + 1];
    }
    qu[6] = qv[6] = T(0);
    ur[6] = u_next;
    vr[6] = v_next;
  };
  // Rows are loaded D iterations ahead, a block of D at a time.
  T cu[D], cv[D], cb[D], cc[D];
#pragma unroll
  for (int u = 0; u < D; ++u) {
    cu[u] = ldu(p0 + 5 + u);
    cv[u] = ldv_(p0 + 5 + u);
    cb[u] = ldub(p0 + 1 + u);
    cc[u] = ldvb(p0 + 1 + u);
  }
  for (int t0 = 0; t0 < nt; t0 += D) {
    T nu[D], nv[D], nb[D], nc[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      nu[u] = ldu(p0 + 5 + D + t0 + u);
      nv[u] = ldv_(p0 + 5 + D + t0 + u);
      nb[u] = ldub(p0 + 1 + D + t0 + u);
      nc[u] = ldvb(p0 + 1 + D + t0 + u);
    }
#pragma unroll
    for (int u = 0; u < D; ++u)
      if (t0 + u < nt) step(t0 + u, cu[u], cv[u], cb[u], cc[u]);  // uniform across the workgroup
#pragma unroll
    for (int u = 0; u < D; ++u) {
      cu[u] = nu[u];
      cv[u] = nv[u];
      cb[u] = nb[u];
      cc[u] = nc[u];
    }
  }
}

// Tiled form: a workgroup takes a 64 x TJ tile of one level at a time, for
// each level of its chunk.
//  1. u and v on the tile plus 6 points each side go to LDS (addresses
//     clamped into the box [-3, nx+3) x [-3, ny+3); a value outside it feeds
//     no interior stage-1 point).  The next level's are loaded into
//     registers during step 3 of this one.
//  2. Stage 1 of hadvuv5th_adj_stage1 (the same expressions) at every interior
//     point of the tile plus 3 each side, into LDS fields of their own: the
//     four upwind weights, gu, gv, gu2, gv2, the metric term and the four
//     branch bits.
//  3. Every uin_bar / vin_bar point the tile owns gathers exactly as
//     hadvuv5th_adj_stage2 does (same terms, same order, so both forms give
//     the same bits), with the stage-1 values read from LDS.  A tile owns its
//     own points, plus the halo beyond the domain where it touches an edge,
//     so every point of the box is written once.
// Two barriers per tile level: after step 1 (u, v written; the last level's
// gathers done with the fields) and after step 2 (fields written; stage 1
// done with u and v).  The stage-1 region is 1.5 (TJ = 16) to 1.9 (TJ = 8)
// times the tile, so stage 1 is formed that much more often than in the
// two-pass form, but nothing but uin_bar and vin_bar goes to memory.
template <class T, int TJ>
struct HadvTileLds {
  static constexpr int kW2 = kTI + 12, kH2 = TJ + 12;  // u, v: [i0-6, i0+70) x [j0-6, j0+TJ+6)
  static constexpr int kW1 = kTI + 6, kH1 = TJ + 6;    // stage 1: [i0-3, i0+67) x [j0-3, j0+TJ+3)
  static constexpr int kN2 = kW2 * kH2, kN1 = kW1 * kH1;
  // 1-D tables of acrlat0/1 and tgrlatda0/1 over the stage-1 rows.
  static constexpr size_t kBytes =
      (2 * size_t(kN2) + 9 * size_t(kN1) + 4 * size_t(kH1)) * sizeof(T) + size_t(kN1) * sizeof(int);
};

// At least 4 waves per SIMD: at most 128 registers.
template <class T, int TJ, int NT>
__global__ void __launch_bounds__(NT, 4) hadvuv5th_adj_tile(HadvAdj<T> a, Geo g, Tiles t) {
  using Lds = HadvTileLds<T, TJ>;
  constexpr int kW2 = Lds::kW2, kN2 = Lds::kN2, kW1 = Lds::kW1, kN1 = Lds::kN1;
  constexpr int kLoad = (kN2 + NT - 1) / NT;
  __shared__ __align__(16) unsigned char smem[Lds::kBytes];
  T *s_u = reinterpret_cast<T *>(smem), *s_v = s_u + kN2;
  T *s_wux = s_v + kN2, *s_wvx = s_wux + kN1, *s_wuy = s_wvx + kN1, *s_wvy = s_wuy + kN1;
  T *s_gu = s_wvy + kN1, *s_gv = s_gu + kN1, *s_gu2 = s_gv + kN1, *s_gv2 = s_gu2 + kN1;
  T *s_tgt = s_gv2 + kN1;
  T *s_a0 = s_tgt + kN1, *s_a1 = s_a0 + Lds::kH1, *s_t0 = s_a1 + Lds::kH1, *s_t1 = s_t0 + Lds::kH1;
  int *s_br = reinterpret_cast<int *>(s_t1 + Lds::kH1);
  constexpr T rr = kEarthRadiusRecipAdj<T>;
  const TileCoord tc = tile_coord(t, kTI, TJ, g.nz);
  const int i0 = tc.i0, j0 = tc.j0;
  const int nx = g.nx, ny = g.ny, sj = g.sj;
  const int tid = int(threadIdx.x);
  // Owned uin_bar / vin_bar points: the tile, plus the halo beyond an edge.
  const int oi0 = i0 == 0 ? -3 : i0, oi1 = i0 + kTI >= nx ? nx + 3 : i0 + kTI;
  const int oj0 = j0 == 0 ? -3 : j0, oj1 = j0 + TJ >= ny ? ny + 3 : j0 + TJ;
  const int ow = oi1 - oi0, on = ow * (oj1 - oj0);
  const bool far = i0 >= 3 && i0 + kTI + 3 <= nx && j0 >= 3 && j0 + TJ + 3 <= ny;
  auto inside = [&](int i, int j) { return in_range(i, 0, nx) && in_range(j, 0, ny); };
  // The 1-D tables: stage-1 row pj at pj - j0 + 3, clamped into [0, ny) (a
  // row outside it belongs to no interior point and is not read).
  if (tid < Lds::kH1) {
    const int64_t jr = min(max(j0 - 3 + tid, 0), ny - 1);
    s_a0[tid] = a.a0[jr * a.s0];
    s_a1[tid] = a.a1[jr * a.s1];
    s_t0[tid] = a.t0[jr * a.s2];
    s_t1[tid] = a.t1[jr * a.s3];
  }
  // u, v of level k with a 6-point halo, into registers.
  T lu[kLoad], lv[kLoad];
  auto load_uv = [&](int k) {
    const int64_t plane = int64_t(k) * g.sk;
#pragma unroll
    for (int n = 0; n < kLoad; ++n) {
      const int e = tid + n * NT;
      const int gi = min(max(i0 - 6 + e % kW2, -3), nx + 2), gj = min(max(j0 - 6 + e / kW2, -3), ny + 2);
      const int64_t at = plane + gi + int64_t(gj) * sj;
      lu[n] = a.uin[at];
      lv[n] = a.vin[at];
    }
  };
  load_uv(tc.k0);

  for (int k = tc.k0; k < tc.k1; ++k) {
    const int64_t plane = int64_t(k) * g.sk;
    // 1. u, v to LDS.  (Loading stage 1's uout_bar and vout_bar here, before
    // the barrier, spilled 48 bytes per lane and measured 208 us against
    // 170 at 256x256x80.)
#pragma unroll
    for (int n = 0; n < kLoad; ++n) {
      const int e = tid + n * NT;
      if (e < kN2) {
        s_u[e] = lu[n];
        s_v[e] = lv[n];
      }
    }
    lds_barrier();

    // 2. Stage 1 (kernels/hadvuv5th.h:183-195).
    // Stage 1 at element e of the stage-1 region, given uout_bar and
    // vout_bar there (read only for interior points).
    auto stage1 = [&](int e, T ob, T ob2) {
      const int c1 = e % kW1, r1 = e / kW1;
      const int pi = i0 - 3 + c1, pj = j0 - 3 + r1;
      T f0 = T(0), f1 = T(0), f2 = T(0), f3 = T(0), f4 = T(0), f5 = T(0), f6 = T(0), f7 = T(0), f8 = T(0);
      int br = 0;
      if (inside(pi, pj)) {
        const T a0 = s_a0[r1], a1 = s_a1[r1];
        const T tg0 = s_t0[r1], tg1 = s_t1[r1];
        const int e2 = (r1 + 3) * kW2 + c1 + 3;
        const T *u = s_u + e2, *v = s_v + e2;
        const T U = (T(1.0) / T(3.0)) * (u[-1] + u[0] + u[1]);
        const T Vu = T(0.25) * (v[1] + v[1 - kW2] + v[0] + v[-kW2]);
        const T ua = a0 * U, va = rr * Vu;
        const bool px = ua > 0, py = va > 0;
        const T sx = s5sel(u, 1, px), sy = s5sel(u, kW2, py);
        f0 = a.eddlat * ob * (px ? ua : -ua);
        f2 = a.eddlon * ob * (py ? va : -va);
        f4 = a.eddlat * ob * (px ? sx : -sx) * a0 * (T(1.0) / T(3.0));
        f5 = a.eddlon * ob * (py ? sy : -sy) * rr * T(0.25) + ob * tg0 * u[0] * T(0.25);
        f8 = ob * tg0 * Vu;
        const T Uv = T(0.25) * (u[-1] + u[0] + u[kW2] + u[kW2 - 1]);
        const T V = T(1.0) / T(3.0) * (v[-kW2] + v[0] + v[kW2]);
        const T ua2 = a1 * Uv, va2 = rr * V;
        const bool px2 = ua2 > 0, py2 = va2 > 0;
        const T sx2 = s5sel(v, 1, px2), sy2 = s5sel(v, kW2, py2);
        f1 = a.eddlat * ob2 * (px2 ? ua2 : -ua2);
        f3 = a.eddlon * ob2 * (py2 ? va2 : -va2);
        f6 = a.eddlat * ob2 * (px2 ? sx2 : -sx2) * a1 * T(0.25) - ob2 * tg1 * Uv * T(0.5);
        f7 = a.eddlon * ob2 * (py2 ? sy2 : -sy2) * rr * (T(1.0) / T(3.0));
        br = (px ? 1 : 0) | (px2 ? 2 : 0) | (py ? 4 : 0) | (py2 ? 8 : 0);
      }
      s_wux[e] = f0;
      s_wvx[e] = f1;
      s_wuy[e] = f2;
      s_wvy[e] = f3;
      s_gu[e] = f4;
      s_gv[e] = f5;
      s_gu2[e] = f6;
      s_gv2[e] = f7;
      s_tgt[e] = f8;
      s_br[e] = br;
    };
    // (Loading both stage-1 points' uout_bar and vout_bar first, at clamped
    // addresses, spilled 40 bytes per lane and took 186 us against 152 at
    // 256x256x80.  With the lane index passed through an opaque register at
    // each level, so that no address is held across levels, it fit in 108
    // VGPRs and took 160 us; loading them one level ahead, during the
    // gathers, 110 VGPRs and 153-158 us, against 151.)
    for (int e = tid; e < kN1; e += NT) {
      const int pi = i0 - 3 + e % kW1, pj = j0 - 3 + e / kW1;
      T ob = T(0), ob2 = T(0);
      if (inside(pi, pj)) {
        const int64_t at = plane + pi + int64_t(pj) * sj;
        ob = a.ubar[at];
        ob2 = a.vbar[at];
      }
      stage1(e, ob, ob2);
    }
    lds_barrier();

    // 3. The gathers of hadvuv5th_adj_stage2 at the owned points, with the
    // next level's u and v in flight.  A tile 3 or more points from every
    // edge owns only its own points, all of whose neighbours are interior:
    // the same terms in the same order without the tests.
    if (k + 1 < tc.k1) load_uv(k + 1);
    if (far) {
      for (int e = tid; e < kTI * TJ; e += NT) {
        const int i = i0 + e % kTI, j = j0 + e / kTI;
        const int c = (e % kTI + 3) + (e / kTI + 3) * kW1;
        T ru = T(0), rv = T(0);
#pragma unroll
        for (int kk = -3; kk <= 3; ++kk) {
          const int q = c - kk, b = s_br[q];
          ru += s_wux[q] * c5<T>((b & 1) != 0, kk);
          rv += s_wvx[q] * c5<T>((b & 2) != 0, kk);
          const int qy = c - kk * kW1, by = s_br[qy];
          ru += s_wuy[qy] * c5<T>((by & 4) != 0, kk);
          rv += s_wvy[qy] * c5<T>((by & 8) != 0, kk);
        }
        ru += s_gu[c - 1];
        ru += s_gu[c];
        ru += s_gu[c + 1];
        ru += s_gu2[c];
        ru += s_gu2[c + 1];
        ru += s_gu2[c - kW1];
        ru += s_gu2[c - kW1 + 1];
        ru += s_tgt[c];
        rv += s_gv[c - 1];
        rv += s_gv[c - 1 + kW1];
        rv += s_gv[c];
        rv += s_gv[c + kW1];
        rv += s_gv2[c - kW1];
        rv += s_gv2[c];
        rv += s_gv2[c + kW1];
        const int64_t at = plane + i + int64_t(j) * sj;
        a.uinbar[at] = ru;
        a.vinbar[at] = rv;
      }
      continue;
    }
    for (int e = tid; e < on; e += NT) {
      const int i = oi0 + e % ow, j = oj0 + e / ow;
      const int c = (i - i0 + 3) + (j - j0 + 3) * kW1;  // (i, j) in the stage-1 region
      T ru = T(0), rv = T(0);
#pragma unroll
      for (int kk = -3; kk <= 3; ++kk) {
        if (inside(i - kk, j)) {
          const int q = c - kk;
          const int b = s_br[q];
          ru += s_wux[q] * c5<T>((b & 1) != 0, kk);
          rv += s_wvx[q] * c5<T>((b & 2) != 0, kk);
        }
        if (inside(i, j - kk)) {
          const int q = c - kk * kW1;
          const int b = s_br[q];
          ru += s_wuy[q] * c5<T>((b & 4) != 0, kk);
          rv += s_wvy[q] * c5<T>((b & 8) != 0, kk);
        }
      }
      for (int m = -1; m <= 1; ++m)
        if (inside(i + m, j)) ru += s_gu[c + m];
      if (inside(i, j)) ru += s_gu2[c];
      if (inside(i + 1, j)) ru += s_gu2[c + 1];
      if (inside(i, j - 1)) ru += s_gu2[c - kW1];
      if (inside(i + 1, j - 1)) ru += s_gu2[c - kW1 + 1];
      if (inside(i, j)) ru += s_tgt[c];
      if (inside(i - 1, j)) rv += s_gv[c - 1];
      if (inside(i - 1, j + 1)) rv += s_gv[c - 1 + kW1];
      if (inside(i, j)) rv += s_gv[c];
      if (inside(i, j + 1)) rv += s_gv[c + kW1];
      for (int m = -1; m <= 1; ++m)
        if (inside(i, j + m)) rv += s_gv2[c + m * kW1];
      const int64_t at = plane + i + int64_t(j) * sj;
      a.uinbar[at] = ru;
      a.vinbar[at] = rv;
    }
    // The next level's step 1 writes only u and v, which this step does not
    // read; the barrier after it orders this step before the next step 2.
  }
}

template <class T, class S3, class S1>
int hadvuv5th_adj_run(const oeb_domain *d, const S3 *uin, const S3 *vin, const S1 *acrlat0, const S1 *acrlat1,
                      const S1 *tgrlatda0, const S1 *tgrlatda1, const S3 *uout_bar, const S3 *vout_bar,
                      const S3 *uin_bar, const S3 *vin_bar, const S3 *const scratch[8], T eddlat, T eddlon,
                      void *stream) {
  const char *K = "hadvuv5th_adj";
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
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
  // the tile knob: 1 = the tiled form, 0 = the row march (or the
  // two-pass form).  FP64 takes 64x16 tiles of 1024 lanes (151 KB of LDS, one
  // workgroup per CU); FP32 64x8 tiles of 512 lanes (51 KB), or 64x16 of 1024
  // with tile knob 2.  Measured at 256x256x80: FP64 tiles 173 us
  // (4 levels per workgroup; 180 at 2) against 211 for the row march; FP32
  // tiles 107 us against 86, so FP32 keeps the row march.  FP64 tiles take 5
  // levels per workgroup in the level-interleaved order (Tiles::klev):
  // 147.7 / 535.6 us against 159.7 / 592.8 for 4 levels in row order at
  // 256x256x80 / 512x512x80 (kernel time over HIP-graph replays; 4 and 6
  // levels and orders 0, 2 and 3 measured 148.8-174.3 / 528.9-592.8).
  const int tile = knob(K, sizeof(T), d, KNOB_TILE, sizeof(T) == 8 ? 1 : 0);
  if (tile != 0) {
    const int kc = knob(K, sizeof(T), d, KNOB_KCHUNK, sizeof(T) == 8 ? 5 : 4);
    if (sizeof(T) == 4 && tile == 1) {
      const Tiles t = make_tiles(nx, ny, int(nz), kTI, 8, kc);
      OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTI, 8));
      if constexpr (sizeof(T) == 4)
        hipLaunchKernelGGL((hadvuv5th_adj_tile<T, 8, 512>), dim3(t.n), dim3(512), 0, hipStream_t(stream), a, g, t);
    } else {
      const int order = knob(K, sizeof(T), d, KNOB_ORDER, sizeof(T) == 8 ? 3 : 0);
      const Tiles t = make_tiles(nx, ny, int(nz), kTI, 16, kc, order);
      OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTI, 16));
      hipLaunchKernelGGL((hadvuv5th_adj_tile<T, 16, 1024>), dim3(t.n), dim3(1024), 0, hipStream_t(stream), a, g, t);
    }
    return launch_status(K);
  }
  if (nx + 12 <= 640 && knob_on(K, sizeof(T), d, KNOB_FORM, true)) {
    const int lanes = int((nx + 12 + 63) / 64 * 64);
    // Row chunks measured at 256x256x80 over 8-64: 32 rows fastest in FP64.
    // FP32 takes 22: 71.9 / 304.2 us against 78.8 / 309.8 for 32 at
    // 256x256x80 / 512x512x80; 20-28 rows took 71.9-91.6 / 296.9-315.2.  At
    // 256x256x80, 22 rows is the shortest chunk whose 960 workgroups fit one
    // resident round of the CUs; 20 rows (1040) spill into a second.  At
    // most 256 rows, which bounds the 1-D tables in LDS.
    const int rows = std::min(knob(K, sizeof(T), d, KNOB_ROWS, sizeof(T) == 4 ? 22 : 32), 256);
    const int nchunk = int((ny + 6 + rows - 1) / rows);
    if (int64_t(nchunk) * nz > (int64_t(1) << 30)) return fail(OEB_ERR_DOMAIN, std::string(K) + ": domain too large");
    // Rows loaded one ahead: two ahead spills registers and measured slower.
    auto kern = knob(K, sizeof(T), d, KNOB_DEPTH, 1) >= 2 ? hadvuv5th_adj_strip<T, 2> : hadvuv5th_adj_strip<T, 1>;
    hipLaunchKernelGGL(kern, dim3(unsigned(nchunk * nz)), dim3(unsigned(lanes)),
                       (size_t(lanes) * 16 + 4 * size_t(rows + 10)) * sizeof(T) + size_t(lanes) * 2 * sizeof(int),
                       hipStream_t(stream), a, g, rows, nchunk);
    return launch_status(K);
  }
  const BoxGrid b1{0, 0, int(nx), int(ny)}, b2{-3, -3, int(nx + 6), int(ny + 6)};
  hipLaunchKernelGGL(hadvuv5th_adj_stage1<T>, box_blocks(b1, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g,
                     b1);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(hadvuv5th_adj_stage2<T>, box_blocks(b2, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g,
                     b2);
  return launch_status(K);
}

// ================================================================= hdiffsmag
// Forward (kernels/hdiffsmag.h:205-263), differentiated in uin and vin:
//   T(p)  = (v(j-1) - v) eddlat R - (u(i-1) - u) ac(j) eddlon
//   S(p)  = (u(j+1) - u) eddlat R + (v(i+1) - v) ac(j) eddlon
//   Qu    = 0.5 (T(i+1)^2 + T^2) + 0.5 (S(j-1)^2 + S^2)
//   smag_u = clamp(tau sqrt(Qu) - w mask, 0, 0.5),  uout = u + smag_u lapu(u)
//   Qv    = 0.5 (T(j+1)^2 + T^2) + 0.5 (S(i-1)^2 + S^2)
//   smag_v = clamp(tau sqrt(Qv) - w mask, 0, 0.5),  vout = v + smag_v lapv(v)
// Stage 1 (interior): gu = ubar lapu tau / (2 sqrt(Qu)) (0 when clamped), the
// cotangent of Qu; gv alike; wu = smag_u ubar and wv = smag_v vbar, the
// weights of the transposed Laplacians.  Stage 2 (on [-1, nx+1) x
// [-1, ny+1)): T_bar(x) = T(x) (gu(x) + gu(x-i) + gv(x) + gv(x-j)),
// S_bar(y) = S(y) (gu(y) + gu(y+j) + gv(y) + gv(y+i)), and the gathers of
// T_bar, S_bar and the Laplacian weights onto u and v, plus ubar, vbar.
template <class T>
struct HsmAdj {
  const T *uin, *vin, *mask, *vo, *vu, *lo, *lu, *ac, *ubar, *vbar;
  int svo, svu, slo, slu, sac;
  T eddlat, eddlon, tau, weight;
  T *gu, *gv, *wu, *wv;  // scratch, interior
  T *uinbar, *vinbar;
};

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) hdiffsmag_adj_stage1(HsmAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  constexpr T rr = kEarthRadiusRecipAdj<T>;
  const int sj = g.sj;
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * sj;
  const T *u = a.uin + at;
  const T *v = a.vin + at;
  const T eddlat = a.eddlat, eddlon = a.eddlon;
  const T ac_s = a.ac[int64_t(j - 1) * a.sac], ac_c = a.ac[int64_t(j) * a.sac], ac_n = a.ac[int64_t(j + 1) * a.sac];
  // The forward's expressions (kernels/hdiffsmag.h:213-259), as hdiffsmag.hip forms them.
  const T t0 = (v[-sj] - v[0]) * eddlat * rr - (u[-1] - u[0]) * ac_c * eddlon;
  const T t_ip = (v[1 - sj] - v[1]) * eddlat * rr - (u[0] - u[1]) * ac_c * eddlon;
  const T t_jp = (v[0] - v[sj]) * eddlat * rr - (u[sj - 1] - u[sj]) * ac_n * eddlon;
  const T s0 = (u[sj] - u[0]) * eddlat * rr + (v[1] - v[0]) * ac_c * eddlon;
  const T s_im = (u[sj - 1] - u[-1]) * eddlat * rr + (v[0] - v[-1]) * ac_c * eddlon;
  const T s_jm = (u[0] - u[-sj]) * eddlat * rr + (v[1 - sj] - v[-sj]) * ac_s * eddlon;
  const T tt0 = t0 * t0, tt_ip = t_ip * t_ip, tt_jp = t_jp * t_jp;
  const T ss0 = s0 * s0, ss_im = s_im * s_im, ss_jm = s_jm * s_jm;
  const T hdweight = a.weight * a.mask[at];
  const T qu = T(0.5) * (tt_ip + tt0) + T(0.5) * (ss_jm + ss0);
  const T qv = T(0.5) * (tt_jp + tt0) + T(0.5) * (ss_im + ss0);
  const T su = a.tau * sqrt(qu) - hdweight, sv = a.tau * sqrt(qv) - hdweight;
  // min(0.5, max(0, x)) with std::max(a, b) = a < b ? b : a and
  // std::min(a, b) = b < a ? b : a: the result is x itself when 0 < x < 0.5.
  const bool pass_u = (T(0) < su) && (su < T(0.5)), pass_v = (T(0) < sv) && (sv < T(0.5));
  const T smag_u = pass_u ? su : ((T(0) < su) ? T(0.5) : T(0));
  const T smag_v = pass_v ? sv : ((T(0) < sv) ? T(0.5) : T(0));
  const T lapu = u[1] + u[-1] - T(2.0) * u[0] + a.lo[int64_t(j) * a.slo] * (u[sj] - u[0]) +
                 a.lu[int64_t(j) * a.slu] * (u[-sj] - u[0]);
  const T lapv = v[1] + v[-1] - T(2.0) * v[0] + a.vo[int64_t(j) * a.svo] * (v[sj] - v[0]) +
                 a.vu[int64_t(j) * a.svu] * (v[-sj] - v[0]);
  const T ub = a.ubar[at], vb = a.vbar[at];
  a.gu[at] = pass_u ? ub * lapu * a.tau * T(0.5) / sqrt(qu) : T(0);
  a.gv[at] = pass_v ? vb * lapv * a.tau * T(0.5) / sqrt(qv) : T(0);
  a.wu[at] = smag_u * ub;
  a.wv[at] = smag_v * vb;
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) hdiffsmag_adj_stage2(HsmAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  constexpr T rr = kEarthRadiusRecipAdj<T>;
  const int sj = g.sj, nx = g.nx, ny = g.ny;
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * sj;
  const T eddlat = a.eddlat, eddlon = a.eddlon;
  auto inside = [&](int di, int dj) { return in_range(i + di, 0, nx) && in_range(j + dj, 0, ny); };
  auto get = [&](const T *f, int di, int dj) { return inside(di, dj) ? f[at + di + int64_t(dj) * sj] : T(0); };
  auto acr = [&](int jj) { return a.ac[int64_t(jj) * a.sac]; };
  // T(x), S(x) at x = q + (di, dj).
  auto Tx = [&](int di, int dj) {
    const int64_t x = at + di + int64_t(dj) * sj;
    return (a.vin[x - sj] - a.vin[x]) * eddlat * rr - (a.uin[x - 1] - a.uin[x]) * acr(j + dj) * eddlon;
  };
  auto Sx = [&](int di, int dj) {
    const int64_t x = at + di + int64_t(dj) * sj;
    return (a.uin[x + sj] - a.uin[x]) * eddlat * rr + (a.vin[x + 1] - a.vin[x]) * acr(j + dj) * eddlon;
  };
  // Cotangent weights of T(x) and S(x); zero where no output reads them.
  auto tw = [&](int di, int dj) {
    return get(a.gu, di, dj) + get(a.gu, di - 1, dj) + get(a.gv, di, dj) + get(a.gv, di, dj - 1);
  };
  auto sw = [&](int di, int dj) {
    return get(a.gu, di, dj) + get(a.gu, di, dj + 1) + get(a.gv, di, dj) + get(a.gv, di + 1, dj);
  };
  auto Tbar = [&](int di, int dj) {
    const T w = tw(di, dj);
    return w != T(0) ? Tx(di, dj) * w : T(0);
  };
  auto Sbar = [&](int di, int dj) {
    const T w = sw(di, dj);
    return w != T(0) ? Sx(di, dj) * w : T(0);
  };
  T ru = inside(0, 0) ? a.ubar[at] : T(0);
  T rv = inside(0, 0) ? a.vbar[at] : T(0);
  // d(T_sqr)/d: T(q) reads u(q) (+ac(j) eddlon), v(q) (-eddlat R); T(q+i)
  // reads u(q) (-ac(j) eddlon); T(q+j) reads v(q) (+eddlat R).
  // (d Q / d T(x) = T(x): the 0.5 of Q cancels the square's 2.)
  const T tq = Tbar(0, 0);
  ru += tq * acr(j) * eddlon - Tbar(1, 0) * acr(j) * eddlon;
  rv += Tbar(0, 1) * eddlat * rr - tq * eddlat * rr;
  // S(q) reads u(q) (-eddlat R), v(q) (-ac(j) eddlon); S(q-j) reads u(q)
  // (+eddlat R); S(q-i) reads v(q) (+ac(j) eddlon).
  const T sq = Sbar(0, 0);
  ru += Sbar(0, -1) * eddlat * rr - sq * eddlat * rr;
  rv += Sbar(-1, 0) * acr(j) * eddlon - sq * acr(j) * eddlon;
  // Transposed Laplacians: lapu(p) = u(p-1) + u(p+1) - 2u + lo(j)(u(j+1) - u) + lu(j)(u(j-1) - u).
  if (inside(0, 0)) {
    ru += get(a.wu, 0, 0) * (T(-2.0) - a.lo[int64_t(j) * a.slo] - a.lu[int64_t(j) * a.slu]);
    rv += get(a.wv, 0, 0) * (T(-2.0) - a.vo[int64_t(j) * a.svo] - a.vu[int64_t(j) * a.svu]);
  }
  ru += get(a.wu, -1, 0) + get(a.wu, 1, 0);
  rv += get(a.wv, -1, 0) + get(a.wv, 1, 0);
  if (inside(0, -1)) {
    ru += get(a.wu, 0, -1) * a.lo[int64_t(j - 1) * a.slo];
    rv += get(a.wv, 0, -1) * a.vo[int64_t(j - 1) * a.svo];
  }
  if (inside(0, 1)) {
    ru += get(a.wu, 0, 1) * a.lu[int64_t(j + 1) * a.slu];
    rv += get(a.wv, 0, 1) * a.vu[int64_t(j + 1) * a.svu];
  }
  a.uinbar[at] = ru;
  a.vinbar[at] = rv;
}

// Fused row-march form, used when one strip spans nx (nx + 4 <= 1024): lane
// l owns column i = l - 2 of [-2, nx + 2) of one level; the workgroup walks
// the uin_bar / vin_bar rows [j0, j1) of its chunk of [-1, ny + 1).  With
// front row r, one iteration
//  * publishes u(r) and v(r - 1) to LDS (3 slots each);
//  * forms stage 1 (hdiffsmag_adj_stage1) at (i, r - 1), and T(i, r),
//    S(i, r - 1) (the forward's tension and shear, on the points whose
//    cotangent weight can be non-zero);
//  * forms T_bar and S_bar at (i, r - 2) from the gu, gv of rows r - 3 ..
//    r - 1;
//  * gathers uin_bar, vin_bar at (i, r - 3) (hdiffsmag_adj_stage2).
// The i neighbours of gu, gv (2 slots), wu, wv (3 slots), T_bar, S_bar (2 slots) and
// u, v come from LDS rows, published after the barrier of one iteration and
// read in later ones; the j neighbours are rows of the lane's own column in
// registers.  One barrier per row.  The expressions are those of the
// two-pass form, term for term.  Loads run D rows ahead, with addresses
// clamped into the arrays' boxes; a clamped value is never used.  The
// chunk's rows of the five 1-D arrays are copied to LDS tables first (as in
// hdiffsa_adj_strip: loads under lane conditions would wait for the rows
// loaded ahead).
template <class T, int D>
__global__ void __launch_bounds__(1024) hdiffsmag_adj_strip(HsmAdj<T> a, Geo g, int rows, int nchunk) {
  extern __shared__ __align__(16) unsigned char oeb_smem[];
  constexpr T rr = kEarthRadiusRecipAdj<T>;
  const int L = int(blockDim.x);
  T *s_u = reinterpret_cast<T *>(oeb_smem);  // 3 x L: u rows
  T *s_v = s_u + 3 * L;                       // 3 x L: v rows
  T *s_gu = s_v + 3 * L;                      // 2 x L each: gu, gv rows (read one iteration later)
  T *s_gv = s_gu + 2 * L;
  T *s_wu = s_gv + 2 * L;                     // 3 x L each: wu, wv rows (read two iterations later)
  T *s_wv = s_wu + 3 * L;
  T *s_tb = s_wv + 3 * L;  // 2 x L each: T_bar, S_bar rows
  T *s_sb = s_tb + 2 * L;
  const int nt1 = rows + 9;  // table rows [j0 - 6, j1 + 3)
  T *s_ac = s_sb + 2 * L;     // acrlat0
  T *s_lo = s_ac + nt1;       // crlato
  T *s_lu = s_lo + nt1;       // crlatu
  T *s_vo = s_lu + nt1;       // crlavo
  T *s_vu = s_vo + nt1;       // crlavu
  const int nx = g.nx, ny = g.ny, sj = g.sj;
  const int cj = int(blockIdx.x) % nchunk, k = int(blockIdx.x) / nchunk;
  const int j0 = -1 + cj * rows, j1 = min(ny + 1, j0 + rows);
  const int l = int(threadIdx.x), i = l - 2;
  const int tb = j0 - 6;  // first row of the tables; read after the first barrier
  for (int e = l; e < j1 - j0 + 9; e += L) {
    const int64_t ra = min(max(tb + e, -1), ny);     // acrlat0 lives on [-1, ny + 1)
    const int64_t rl = min(max(tb + e, 0), ny - 1);  // the others on [0, ny)
    s_ac[e] = a.ac[ra * a.sac];
    s_lo[e] = a.lo[rl * a.slo];
    s_lu[e] = a.lu[rl * a.slu];
    s_vo[e] = a.vo[rl * a.svo];
    s_vu[e] = a.vu[rl * a.svu];
  }
  auto lo = [&](int jj) { return s_lo[jj - tb]; };
  auto lu = [&](int jj) { return s_lu[jj - tb]; };
  auto vo = [&](int jj) { return s_vo[jj - tb]; };
  auto vu = [&](int jj) { return s_vu[jj - tb]; };
  const bool col_uv = i >= -1 && i < nx + 1;  // uin_bar, vin_bar columns
  const bool col_int = i >= 0 && i < nx;
  const bool col_t = i >= 0 && i <= nx;    // T with a possibly non-zero weight
  const bool col_s = i >= -1 && i < nx;    // S alike
  const T eddlat = a.eddlat, eddlon = a.eddlon;
  const int64_t plane = int64_t(k) * g.sk;
  const int64_t c_uv = plane + min(max(i, -1), nx), c_o = plane + min(max(i, 0), nx - 1);
  auto ldu = [&](int r) -> T { return a.uin[c_uv + int64_t(min(max(r, -1), ny)) * sj]; };
  auto ldv_ = [&](int r) -> T { return a.vin[c_uv + int64_t(min(max(r, -1), ny)) * sj]; };
  auto ldo = [&](const T *f, int r) -> T { return f[c_o + int64_t(min(max(r, 0), ny - 1)) * sj]; };
  auto acr = [&](int jj) { return s_ac[jj - tb]; };
  const int r0 = j0 - 2, nt = j1 - j0 + 5;  // front rows r0 .. j1 + 2
  // Own-column rows (index: row relative to the front row r).
  T u_m2 = ldu(r0 - 2), u_m1 = ldu(r0 - 1), v_m2 = ldv_(r0 - 2), v_m1 = ldv_(r0 - 1);
  T t_m2 = T(0), t_m1 = T(0);               // T rows r - 2, r - 1
  T s_m2 = T(0);                            // S row r - 2
  T gu_m2 = T(0), gv_m3 = T(0), gv_m2 = T(0);  // gu row r - 2, gv rows r - 3, r - 2
  T wu_m4 = T(0), wu_m3 = T(0), wu_m2 = T(0), wv_m4 = T(0), wv_m3 = T(0), wv_m2 = T(0);
  T tb_m3 = T(0), sb_m4 = T(0), sb_m3 = T(0);  // T_bar row r - 3, S_bar rows r - 4, r - 3
  T ub_m2 = T(0), ub_m3 = T(0), vb_m2 = T(0), vb_m3 = T(0);  // out_bar rows r - 2, r - 3
  // One iteration, given the rows it loads: u(r), v(r) and mask, out_bar at r - 1.
  auto step = [&](int t, T u_0, T v_0, T mk, T ub_m1, T vb_m1) {
    const int r = r0 + t, p = r - 1, q = r - 3;
    // Slots written in this iteration, the previous one and the one before.
    const int s0 = t % 3, s1 = (t + 2) % 3, s2 = (t + 1) % 3;
    s_u[s0 * L + l] = u_0;    // u row r
    s_v[s0 * L + l] = v_m1;   // v row r - 1
    lds_barrier();
    const T *U0 = s_u + s0 * L + l;  // u row r
    const T *U1 = s_u + s1 * L + l;  // u row r - 1
    const T *V1 = s_v + s0 * L + l;  // v row r - 1
    const T *V2 = s_v + s1 * L + l;  // v row r - 2
    // Stage 1 at (i, p): hdiffsmag_adj_stage1.
    T gu_m1 = T(0), gv_m1 = T(0), wu_m1 = T(0), wv_m1 = T(0);
    if (col_int && p >= 0 && p < ny) {
      const T ac_s = acr(p - 1), ac_c = acr(p), ac_n = acr(p + 1);
      const T t0 = (v_m2 - v_m1) * eddlat * rr - (U1[-1] - u_m1) * ac_c * eddlon;
      const T t_ip = (V2[1] - V1[1]) * eddlat * rr - (u_m1 - U1[1]) * ac_c * eddlon;
      const T t_jp = (v_m1 - v_0) * eddlat * rr - (U0[-1] - u_0) * ac_n * eddlon;
      const T s0v = (u_0 - u_m1) * eddlat * rr + (V1[1] - v_m1) * ac_c * eddlon;
      const T s_im = (U0[-1] - U1[-1]) * eddlat * rr + (v_m1 - V1[-1]) * ac_c * eddlon;
      const T s_jm = (u_m1 - u_m2) * eddlat * rr + (V2[1] - v_m2) * ac_s * eddlon;
      const T tt0 = t0 * t0, tt_ip = t_ip * t_ip, tt_jp = t_jp * t_jp;
      const T ss0 = s0v * s0v, ss_im = s_im * s_im, ss_jm = s_jm * s_jm;
      const T hdweight = a.weight * mk;
      const T qu = T(0.5) * (tt_ip + tt0) + T(0.5) * (ss_jm + ss0);
      const T qv = T(0.5) * (tt_jp + tt0) + T(0.5) * (ss_im + ss0);
      const T su = a.tau * sqrt(qu) - hdweight, sv = a.tau * sqrt(qv) - hdweight;
      const bool pass_u = (T(0) < su) && (su < T(0.5)), pass_v = (T(0) < sv) && (sv < T(0.5));
      const T smag_u = pass_u ? su : ((T(0) < su) ? T(0.5) : T(0));
      const T smag_v = pass_v ? sv : ((T(0) < sv) ? T(0.5) : T(0));
      const T lapu = U1[1] + U1[-1] - T(2.0) * u_m1 + lo(p) * (u_0 - u_m1) + lu(p) * (u_m2 - u_m1);
      const T lapv = V1[1] + V1[-1] - T(2.0) * v_m1 + vo(p) * (v_0 - v_m1) + vu(p) * (v_m2 - v_m1);
      gu_m1 = pass_u ? ub_m1 * lapu * a.tau * T(0.5) / sqrt(qu) : T(0);
      gv_m1 = pass_v ? vb_m1 * lapv * a.tau * T(0.5) / sqrt(qv) : T(0);
      wu_m1 = smag_u * ub_m1;
      wv_m1 = smag_v * vb_m1;
    }
    s_gu[(t & 1) * L + l] = gu_m1;
    s_gv[(t & 1) * L + l] = gv_m1;
    s_wu[s0 * L + l] = wu_m1;
    s_wv[s0 * L + l] = wv_m1;
    // T(i, r) and S(i, r - 1) (the stage-2 lambdas Tx, Sx).
    T t_0 = T(0), s_m1 = T(0);
    if (col_t && r >= 0 && r <= ny) t_0 = (v_m1 - v_0) * eddlat * rr - (U0[-1] - u_0) * acr(r) * eddlon;
    if (col_s && p >= -1 && p < ny) s_m1 = (u_0 - u_m1) * eddlat * rr + (V1[1] - v_m1) * acr(p) * eddlon;
    // T_bar, S_bar at (i, r - 2): weights from gu, gv of rows r - 3 .. r - 1.
    T tb_m2 = T(0), sb_m2 = T(0);
    if (l >= 1 && l < L - 1) {
      const T *GU = s_gu + ((t + 1) & 1) * L + l;  // gu row r - 2 (published in the previous iteration)
      const T *GV = s_gv + ((t + 1) & 1) * L + l;  // gv row r - 2
      const T tw = GU[0] + GU[-1] + GV[0] + gv_m3;
      const T sw = GU[0] + gu_m1 + GV[0] + GV[1];
      tb_m2 = tw != T(0) ? t_m2 * tw : T(0);
      sb_m2 = sw != T(0) ? s_m2 * sw : T(0);
    }
    s_tb[(t & 1) * L + l] = tb_m2;
    s_sb[(t & 1) * L + l] = sb_m2;
    // uin_bar, vin_bar at (i, q): hdiffsmag_adj_stage2.
    if (col_uv && q >= j0 && q < j1) {
      const T *TB = s_tb + ((t + 1) & 1) * L + l;  // T_bar row q
      const T *SB = s_sb + ((t + 1) & 1) * L + l;  // S_bar row q
      const T *WU = s_wu + s2 * L + l;             // wu row q (published two iterations ago)
      const T *WV = s_wv + s2 * L + l;
      const bool in0 = col_int && q >= 0 && q < ny;
      T ru = in0 ? ub_m3 : T(0);
      T rv = in0 ? vb_m3 : T(0);
      const T ac = acr(q);
      const T tq = tb_m3;
      ru += tq * ac * eddlon - TB[1] * ac * eddlon;
      rv += tb_m2 * eddlat * rr - tq * eddlat * rr;
      const T sq = sb_m3;
      ru += sb_m4 * eddlat * rr - sq * eddlat * rr;
      rv += SB[-1] * ac * eddlon - sq * ac * eddlon;
      if (in0) {
        ru += wu_m3 * (T(-2.0) - lo(q) - lu(q));
        rv += wv_m3 * (T(-2.0) - vo(q) - vu(q));
      }
      ru += (i - 1 >= 0 && i - 1 < nx ? WU[-1] : T(0)) + (i + 1 >= 0 && i + 1 < nx ? WU[1] : T(0));
      rv += (i - 1 >= 0 && i - 1 < nx ? WV[-1] : T(0)) + (i + 1 >= 0 && i + 1 < nx ? WV[1] : T(0));
      if (col_int && q - 1 >= 0 && q - 1 < ny) {
        ru += wu_m4 * lo(q - 1);
        rv += wv_m4 * vo(q - 1);
      }
      if (col_int && q + 1 >= 0 && q + 1 < ny) {
        ru += wu_m2 * lu(q + 1);
        rv += wv_m2 * vu(q + 1);
      }
      const int64_t at = plane + i + int64_t(q) * sj;
      a.uinbar[at] = ru;
      a.vinbar[at] = rv;
    }
    // Shift the own-column rows by one.
    u_m2 = u_m1;
    u_m1 = u_0;
    v_m2 = v_m1;
    v_m1 = v_0;
    t_m2 = t_m1;
    t_m1 = t_0;
    s_m2 = s_m1;
    gu_m2 = gu_m1;
    gv_m3 = gv_m2;
    gv_m2 = gv_m1;
    wu_m4 = wu_m3;
    wu_m3 = wu_m2;
    wu_m2 = wu_m1;
    wv_m4 = wv_m3;
    wv_m3 = wv_m2;
    wv_m2 = wv_m1;
    tb_m3 = tb_m2;
    sb_m4 = sb_m3;
    sb_m3 = sb_m2;
    ub_m3 = ub_m2;
    ub_m2 = ub_m1;
    vb_m3 = vb_m2;
    vb_m2 = vb_m1;
  };
  // Rows are loaded D iterations ahead, a block of D at a time.
  T cu[D], cv[D], cm[D], cb[D], cc[D];
#pragma unroll
  for (int u = 0; u < D; ++u) {
    cu[u] = ldu(r0 + u);
    cv[u] = ldv_(r0 + u);
    cm[u] = ldo(a.mask, r0 - 1 + u);
    cb[u] = ldo(a.ubar, r0 - 1 + u);
    cc[u] = ldo(a.vbar, r0 - 1 + u);
  }
  for (int t0 = 0; t0 < nt; t0 += D) {
    T nu[D], nv[D], nm[D], nb[D], nc[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      nu[u] = ldu(r0 + D + t0 + u);
      nv[u] = ldv_(r0 + D + t0 + u);
      nm[u] = ldo(a.mask, r0 - 1 + D + t0 + u);
      nb[u] = ldo(a.ubar, r0 - 1 + D + t0 + u);
      nc[u] = ldo(a.vbar, r0 - 1 + D + t0 + u);
    }
#pragma unroll
    for (int u = 0; u < D; ++u)
      if (t0 + u < nt) step(t0 + u, cu[u], cv[u], cm[u], cb[u], cc[u]);  // uniform across the workgroup
#pragma unroll
    for (int u = 0; u < D; ++u) {
      cu[u] = nu[u];
      cv[u] = nv[u];
      cm[u] = nm[u];
      cb[u] = nb[u];
      cc[u] = nc[u];
    }
  }
}

// Tiled form: a workgroup takes a 64 x TJ tile of one level at a time, for
// each level of its chunk, in four steps with a barrier after each of the
// first three:
//  1. u and v on the tile plus 3 points each side to LDS (addresses clamped
//     into the box [-1, nx+1) x [-1, ny+1)); the next level's are loaded
//     into registers during step 4 of this one.  mask, uout_bar and vout_bar
//     of the lane's stage-1 points are loaded here too, before the barrier,
//     so that their latency overlaps the wait.
//  2. Stage 1 of hdiffsmag_adj_stage1 (gu, gv, wu, wv) at the interior points
//     of the tile plus 2 each side, 0 elsewhere.
//  3. T_bar and S_bar of hdiffsmag_adj_stage2 at the tile plus 1 each side
//     (the points whose weight can be non-zero for the points the tile
//     owns; outside that region they are 0).  uout_bar and vout_bar of the
//     lane's owned points are loaded before the barrier that follows.
//  4. The gathers of hdiffsmag_adj_stage2 at the points the tile owns: its
//     own, plus the halo beyond the domain where it touches an edge.
// The rows of the five 1-D arrays that the tile reads (the same for every
// level) are copied to LDS tables once per workgroup.  With one 1024-lane
// workgroup per CU, every global load issued after a barrier waited in full:
// stage 1 measured 37 of 142 us at 256x256x80 before these loads moved.
// The same expressions, terms and order as the two-pass form, so both forms
// give the same bits.  Only uin_bar and vin_bar go to memory.
template <class T, int TJ>
struct HsmTileLds {
  static constexpr int kW2 = kTI + 6, kH2 = TJ + 6;  // u, v: [i0-3, i0+67) x [j0-3, j0+TJ+3)
  static constexpr int kW1 = kTI + 4, kH1 = TJ + 4;  // stage 1: [i0-2, i0+66) x [j0-2, j0+TJ+2)
  static constexpr int kW0 = kTI + 2, kH0 = TJ + 2;  // T_bar, S_bar: [i0-1, i0+65) x [j0-1, j0+TJ+1)
  static constexpr int kRows = TJ + 6;                // 1-D tables: rows [j0-3, j0+TJ+3)
  static constexpr int kN2 = kW2 * kH2, kN1 = kW1 * kH1, kN0 = kW0 * kH0;
  static constexpr size_t kBytes =
      (2 * size_t(kN2) + 4 * size_t(kN1) + 2 * size_t(kN0) + 5 * size_t(kRows)) * sizeof(T);
};

template <class T, int TJ, int NT>
__global__ void __launch_bounds__(NT) hdiffsmag_adj_tile(HsmAdj<T> a, Geo g, Tiles t) {
  using Lds = HsmTileLds<T, TJ>;
  constexpr int kW2 = Lds::kW2, kN2 = Lds::kN2, kW1 = Lds::kW1, kN1 = Lds::kN1, kW0 = Lds::kW0, kH0 = Lds::kH0,
                kN0 = Lds::kN0, kRows = Lds::kRows;
  constexpr int kLoad = (kN2 + NT - 1) / NT, kSt1 = (kN1 + NT - 1) / NT, kOwn = (kTI * TJ + NT - 1) / NT;
  __shared__ __align__(16) unsigned char smem[Lds::kBytes];
  T *s_u = reinterpret_cast<T *>(smem), *s_v = s_u + kN2;
  T *s_gu = s_v + kN2, *s_gv = s_gu + kN1, *s_wu = s_gv + kN1, *s_wv = s_wu + kN1;
  T *s_tb = s_wv + kN1, *s_sb = s_tb + kN0;
  T *s_ac = s_sb + kN0, *s_lo = s_ac + kRows, *s_lu = s_lo + kRows, *s_vo = s_lu + kRows, *s_vu = s_vo + kRows;
  constexpr T rr = kEarthRadiusRecipAdj<T>;
  const T eddlat = a.eddlat, eddlon = a.eddlon;
  const TileCoord tc = tile_coord(t, kTI, TJ, g.nz);
  const int i0 = tc.i0, j0 = tc.j0;
  const int nx = g.nx, ny = g.ny, sj = g.sj;
  const int tid = int(threadIdx.x);
  const int oi0 = i0 == 0 ? -1 : i0, oi1 = i0 + kTI >= nx ? nx + 1 : i0 + kTI;
  const int oj0 = j0 == 0 ? -1 : j0, oj1 = j0 + TJ >= ny ? ny + 1 : j0 + TJ;
  const int ow = oi1 - oi0, on = ow * (oj1 - oj0);
  const bool far = i0 >= 1 && i0 + kTI + 1 <= nx && j0 >= 1 && j0 + TJ + 1 <= ny;
  auto inside = [&](int i, int j) { return in_range(i, 0, nx) && in_range(j, 0, ny); };
  // The 1-D tables: row jj at jj - j0 + 3.  Rows outside an array's range
  // (acrlat0: [-1, ny + 1), the others [0, ny)) hold a clamped copy that no
  // point uses: only interior points read them.
  if (tid < kRows) {
    const int jj = j0 - 3 + tid;
    const int64_t ja = min(max(jj, -1), ny), jr = min(max(jj, 0), ny - 1);
    s_ac[tid] = a.ac[ja * a.sac];
    s_lo[tid] = a.lo[jr * a.slo];
    s_lu[tid] = a.lu[jr * a.slu];
    s_vo[tid] = a.vo[jr * a.svo];
    s_vu[tid] = a.vu[jr * a.svu];
  }
  auto acr = [&](int jj) { return s_ac[jj - j0 + 3]; };
  auto lo = [&](int jj) { return s_lo[jj - j0 + 3]; };
  auto lu = [&](int jj) { return s_lu[jj - j0 + 3]; };
  auto vo = [&](int jj) { return s_vo[jj - j0 + 3]; };
  auto vu = [&](int jj) { return s_vu[jj - j0 + 3]; };
  T lu_[kLoad], lv_[kLoad];
  auto load_uv = [&](int k) {
    const int64_t plane = int64_t(k) * g.sk;
#pragma unroll
    for (int n = 0; n < kLoad; ++n) {
      const int e = tid + n * NT;
      const int gi = min(max(i0 - 3 + e % kW2, -1), nx), gj = min(max(j0 - 3 + e / kW2, -1), ny);
      const int64_t at = plane + gi + int64_t(gj) * sj;
      lu_[n] = a.uin[at];
      lv_[n] = a.vin[at];
    }
  };
  load_uv(tc.k0);

  for (int k = tc.k0; k < tc.k1; ++k) {
    const int64_t plane = int64_t(k) * g.sk;
    // 1. u, v to LDS; stage 1's loads of mask, uout_bar and vout_bar.
#pragma unroll
    for (int n = 0; n < kLoad; ++n) {
      const int e = tid + n * NT;
      if (e < kN2) {
        s_u[e] = lu_[n];
        s_v[e] = lv_[n];
      }
    }
    T pm[kSt1], pub[kSt1], pvb[kSt1];
#pragma unroll
    for (int n = 0; n < kSt1; ++n) {
      const int e = tid + n * NT;
      const int i = i0 - 2 + e % kW1, j = j0 - 2 + e / kW1;
      pm[n] = pub[n] = pvb[n] = T(0);
      if (e < kN1 && inside(i, j)) {
        const int64_t at = plane + i + int64_t(j) * sj;
        pm[n] = a.mask[at];
        pub[n] = a.ubar[at];
        pvb[n] = a.vbar[at];
      }
    }
    lds_barrier();

    // 2. Stage 1 (kernels/hdiffsmag.h:213-259), as hdiffsmag_adj_stage1.
#pragma unroll
    for (int n = 0; n < kSt1; ++n) {
      const int e = tid + n * NT;
      if (e >= kN1) break;
      const int c1 = e % kW1, r1 = e / kW1;
      const int i = i0 - 2 + c1, j = j0 - 2 + r1;
      T gu = T(0), gv = T(0), wu = T(0), wv = T(0);
      if (inside(i, j)) {
        const T *u = s_u + (r1 + 1) * kW2 + c1 + 1, *v = s_v + (r1 + 1) * kW2 + c1 + 1;
        const T ac_s = acr(j - 1), ac_c = acr(j), ac_n = acr(j + 1);
        const T t0 = (v[-kW2] - v[0]) * eddlat * rr - (u[-1] - u[0]) * ac_c * eddlon;
        const T t_ip = (v[1 - kW2] - v[1]) * eddlat * rr - (u[0] - u[1]) * ac_c * eddlon;
        const T t_jp = (v[0] - v[kW2]) * eddlat * rr - (u[kW2 - 1] - u[kW2]) * ac_n * eddlon;
        const T s0 = (u[kW2] - u[0]) * eddlat * rr + (v[1] - v[0]) * ac_c * eddlon;
        const T s_im = (u[kW2 - 1] - u[-1]) * eddlat * rr + (v[0] - v[-1]) * ac_c * eddlon;
        const T s_jm = (u[0] - u[-kW2]) * eddlat * rr + (v[1 - kW2] - v[-kW2]) * ac_s * eddlon;
        const T tt0 = t0 * t0, tt_ip = t_ip * t_ip, tt_jp = t_jp * t_jp;
        const T ss0 = s0 * s0, ss_im = s_im * s_im, ss_jm = s_jm * s_jm;
        const T hdweight = a.weight * pm[n];
        const T qu = T(0.5) * (tt_ip + tt0) + T(0.5) * (ss_jm + ss0);
        const T qv = T(0.5) * (tt_jp + tt0) + T(0.5) * (ss_im + ss0);
        const T su = a.tau * sqrt(qu) - hdweight, sv = a.tau * sqrt(qv) - hdweight;
        const bool pass_u = (T(0) < su) && (su < T(0.5)), pass_v = (T(0) < sv) && (sv < T(0.5));
        const T smag_u = pass_u ? su : ((T(0) < su) ? T(0.5) : T(0));
        const T smag_v = pass_v ? sv : ((T(0) < sv) ? T(0.5) : T(0));
        const T lapu = u[1] + u[-1] - T(2.0) * u[0] + lo(j) * (u[kW2] - u[0]) + lu(j) * (u[-kW2] - u[0]);
        const T lapv = v[1] + v[-1] - T(2.0) * v[0] + vo(j) * (v[kW2] - v[0]) + vu(j) * (v[-kW2] - v[0]);
        const T ub = pub[n], vb = pvb[n];
        gu = pass_u ? ub * lapu * a.tau * T(0.5) / sqrt(qu) : T(0);
        gv = pass_v ? vb * lapv * a.tau * T(0.5) / sqrt(qv) : T(0);
        wu = smag_u * ub;
        wv = smag_v * vb;
      }
      s_gu[e] = gu;
      s_gv[e] = gv;
      s_wu[e] = wu;
      s_wv[e] = wv;
    }
    lds_barrier();

    // 3. T_bar, S_bar (hdiffsmag_adj_stage2's Tbar and Sbar); the owned
    // points' uout_bar and vout_bar.
    for (int e = tid; e < kN0; e += NT) {
      const int c0 = e % kW0, r0 = e / kW0;
      const int x1 = (r0 + 1) * kW1 + c0 + 1;  // in the stage-1 region
      const int x2 = (r0 + 2) * kW2 + c0 + 2;  // in the u, v region
      const int xj = j0 - 1 + r0;
      const T tw = s_gu[x1] + s_gu[x1 - 1] + s_gv[x1] + s_gv[x1 - kW1];
      const T sw = s_gu[x1] + s_gu[x1 + kW1] + s_gv[x1] + s_gv[x1 + 1];
      T tb = T(0), sb = T(0);
      if (tw != T(0))
        tb = ((s_v[x2 - kW2] - s_v[x2]) * eddlat * rr - (s_u[x2 - 1] - s_u[x2]) * acr(xj) * eddlon) * tw;
      if (sw != T(0))
        sb = ((s_u[x2 + kW2] - s_u[x2]) * eddlat * rr + (s_v[x2 + 1] - s_v[x2]) * acr(xj) * eddlon) * sw;
      s_tb[e] = tb;
      s_sb[e] = sb;
    }
    T oub[kOwn], ovb[kOwn];
    if (far) {
#pragma unroll
      for (int n = 0; n < kOwn; ++n) {
        const int e = tid + n * NT;
        oub[n] = ovb[n] = T(0);
        if (e < kTI * TJ) {
          const int64_t at = plane + i0 + e % kTI + int64_t(j0 + e / kTI) * sj;
          oub[n] = a.ubar[at];
          ovb[n] = a.vbar[at];
        }
      }
    }
    lds_barrier();

    // 4. The gathers at the owned points, with the next level's u and v in
    // flight.  T_bar and S_bar are 0 outside their region (no weight there).
    if (k + 1 < tc.k1) load_uv(k + 1);
    if (far) {
      // A tile 1 or more points from every edge owns only its own points,
      // all of whose neighbours are interior: the same terms in the same
      // order without the tests.
#pragma unroll
      for (int n = 0; n < kOwn; ++n) {
        const int e = tid + n * NT;
        if (e >= kTI * TJ) break;
        const int ci = e % kTI, cj = e / kTI;
        const int i = i0 + ci, j = j0 + cj;
        const int x0 = (ci + 1) + (cj + 1) * kW0, c1 = (ci + 2) + (cj + 2) * kW1;
        const int64_t at = plane + i + int64_t(j) * sj;
        const T ac_j = acr(j);
        T ru = oub[n], rv = ovb[n];
        const T tq = s_tb[x0];
        ru += tq * ac_j * eddlon - s_tb[x0 + 1] * ac_j * eddlon;
        rv += s_tb[x0 + kW0] * eddlat * rr - tq * eddlat * rr;
        const T sq = s_sb[x0];
        ru += s_sb[x0 - kW0] * eddlat * rr - sq * eddlat * rr;
        rv += s_sb[x0 - 1] * ac_j * eddlon - sq * ac_j * eddlon;
        ru += s_wu[c1] * (T(-2.0) - lo(j) - lu(j));
        rv += s_wv[c1] * (T(-2.0) - vo(j) - vu(j));
        ru += s_wu[c1 - 1] + s_wu[c1 + 1];
        rv += s_wv[c1 - 1] + s_wv[c1 + 1];
        ru += s_wu[c1 - kW1] * lo(j - 1);
        rv += s_wv[c1 - kW1] * vo(j - 1);
        ru += s_wu[c1 + kW1] * lu(j + 1);
        rv += s_wv[c1 + kW1] * vu(j + 1);
        a.uinbar[at] = ru;
        a.vinbar[at] = rv;
      }
      continue;
    }
    for (int e = tid; e < on; e += NT) {
      const int i = oi0 + e % ow, j = oj0 + e / ow;
      const int ci = i - i0, cj = j - j0;
      auto bar = [&](const T *f, int di, int dj) {
        const int x = ci + di + 1, y = cj + dj + 1;
        return (x >= 0 && x < kW0 && y >= 0 && y < kH0) ? f[x + y * kW0] : T(0);
      };
      const int c1 = (ci + 2) + (cj + 2) * kW1;
      auto w1 = [&](const T *f, int di, int dj) { return inside(i + di, j + dj) ? f[c1 + di + dj * kW1] : T(0); };
      const int64_t at = plane + i + int64_t(j) * sj;
      T ru = inside(i, j) ? a.ubar[at] : T(0);
      T rv = inside(i, j) ? a.vbar[at] : T(0);
      const T tq = bar(s_tb, 0, 0);
      ru += tq * acr(j) * eddlon - bar(s_tb, 1, 0) * acr(j) * eddlon;
      rv += bar(s_tb, 0, 1) * eddlat * rr - tq * eddlat * rr;
      const T sq = bar(s_sb, 0, 0);
      ru += bar(s_sb, 0, -1) * eddlat * rr - sq * eddlat * rr;
      rv += bar(s_sb, -1, 0) * acr(j) * eddlon - sq * acr(j) * eddlon;
      if (inside(i, j)) {
        ru += w1(s_wu, 0, 0) * (T(-2.0) - lo(j) - lu(j));
        rv += w1(s_wv, 0, 0) * (T(-2.0) - vo(j) - vu(j));
      }
      ru += w1(s_wu, -1, 0) + w1(s_wu, 1, 0);
      rv += w1(s_wv, -1, 0) + w1(s_wv, 1, 0);
      if (inside(i, j - 1)) {
        ru += w1(s_wu, 0, -1) * lo(j - 1);
        rv += w1(s_wv, 0, -1) * vo(j - 1);
      }
      if (inside(i, j + 1)) {
        ru += w1(s_wu, 0, 1) * lu(j + 1);
        rv += w1(s_wv, 0, 1) * vu(j + 1);
      }
      a.uinbar[at] = ru;
      a.vinbar[at] = rv;
    }
    // The next level's step 1 writes only u and v, which this step does not
    // read; the barrier after it orders this step before the next step 2.
  }
}

template <class T, class S3, class S1>
int hdiffsmag_adj_run(const oeb_domain *d, const S3 *uin, const S3 *vin, const S3 *mask, const S1 *crlavo,
                      const S1 *crlavu, const S1 *crlato, const S1 *crlatu, const S1 *acrlat0, const S3 *uout_bar,
                      const S3 *vout_bar, const S3 *uin_bar, const S3 *vin_bar, const S3 *const work[4], T eddlat,
                      T eddlon, T tau_smag, T weight_smag, void *stream) {
  const char *K = "hdiffsmag_adj";
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz}, r_uv{-1, nx + 1, -1, ny + 1, 0, nz};
  HsmAdj<T> a;
  a.eddlat = eddlat, a.eddlon = eddlon, a.tau = tau_smag, a.weight = weight_smag;
  const T *p_ub, *p_vb, *p_w[4];
  OEB_TRY(check3(K, "uin", uin, r_uv, g, true, &a.uin));
  OEB_TRY(check3(K, "vin", vin, r_uv, g, false, &a.vin));
  OEB_TRY(check3(K, "mask", mask, r0, g, false, &a.mask));
  OEB_TRY(check3(K, "uout_bar", uout_bar, r0, g, false, &a.ubar));
  OEB_TRY(check3(K, "vout_bar", vout_bar, r0, g, false, &a.vbar));
  OEB_TRY(check3(K, "uin_bar", uin_bar, r_uv, g, false, &p_ub));
  OEB_TRY(check3(K, "vin_bar", vin_bar, r_uv, g, false, &p_vb));
  static const char *const kWork[4] = {"work0", "work1", "work2", "work3"};
  for (int q = 0; q < 4; ++q) OEB_TRY(check3(K, kWork[q], work[q], r0, g, false, &p_w[q]));
  OEB_TRY(check1(K, "crlavo", crlavo, 0, ny, &a.vo, &a.svo));
  OEB_TRY(check1(K, "crlavu", crlavu, 0, ny, &a.vu, &a.svu));
  OEB_TRY(check1(K, "crlato", crlato, 0, ny, &a.lo, &a.slo));
  OEB_TRY(check1(K, "crlatu", crlatu, 0, ny, &a.lu, &a.slu));
  OEB_TRY(check1(K, "acrlat0", acrlat0, -1, ny + 1, &a.ac, &a.sac));
  AliasSet as;
  as.add("uin", a.uin, r_uv, g, false);
  as.add("vin", a.vin, r_uv, g, false);
  as.add("mask", a.mask, r0, g, false);
  as.add("uout_bar", a.ubar, r0, g, false);
  as.add("vout_bar", a.vbar, r0, g, false);
  as.add("uin_bar", p_ub, r_uv, g, true);
  as.add("vin_bar", p_vb, r_uv, g, true);
  for (int q = 0; q < 4; ++q) as.add(kWork[q], p_w[q], r0, g, true);
  OEB_TRY(as.check(K));
  a.gu = const_cast<T *>(p_w[0]), a.gv = const_cast<T *>(p_w[1]);
  a.wu = const_cast<T *>(p_w[2]), a.wv = const_cast<T *>(p_w[3]);
  a.uinbar = const_cast<T *>(p_ub);
  a.vinbar = const_cast<T *>(p_vb);
  // the tile knob: 1 = 64x8 tiles of 512 lanes, 2 = 64x16 tiles of
  // 1024 lanes, 0 = the row march (or the two-pass form).  Measured at
  // 256x256x80 with 4 levels per workgroup: FP64 64x16 tiles 146-148 us, 64x8
  // 149-155, row march 158-160; FP32 tiles 85-95 us against 72 for the row
  // march.  Loading mask, ubar and vbar of the next level during the gathers
  // raised the registers to 117 and took 170 us.
  const int tile = knob(K, sizeof(T), d, KNOB_TILE, sizeof(T) == 8 ? 2 : 0);
  if (tile != 0) {
    const int kc = knob(K, sizeof(T), d, KNOB_KCHUNK, 4);
    if (tile == 1) {
      const Tiles t = make_tiles(nx, ny, int(nz), kTI, 8, kc);
      OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTI, 8));
      hipLaunchKernelGGL((hdiffsmag_adj_tile<T, 8, 512>), dim3(t.n), dim3(512), 0, hipStream_t(stream), a, g, t);
    } else {
      // Level-interleaved order (Tiles::klev): 120.6 / 449.2 us against
      // 125.5 / 458.3 for Tiles::irow at 256x256x80 / 512x512x80 (kernel time
      // over HIP-graph replays); 3-5 levels and orders 1-3 took 118.8-134.5 /
      // 449.2-477.5.
      const int order = knob(K, sizeof(T), d, KNOB_ORDER, 3);
      const Tiles t = make_tiles(nx, ny, int(nz), kTI, 16, kc, order);
      OEB_TRY(check_tiles(K, t, nx, ny, int(nz), kTI, 16));
      hipLaunchKernelGGL((hdiffsmag_adj_tile<T, 16, 1024>), dim3(t.n), dim3(1024), 0, hipStream_t(stream), a, g, t);
    }
    return launch_status(K);
  }
  if (nx + 4 <= 1024 && knob_on(K, sizeof(T), d, KNOB_FORM, true)) {
    const int lanes = int((nx + 4 + 63) / 64 * 64);
    // Measured at 256x256x80: FP64 fastest with 16-row chunks and rows
    // loaded one row ahead (two rows ahead spills registers).  FP32 takes
    // 22-row chunks loaded one row ahead: 62.6 / 269.7 us against 67.7-70.8 /
    // 376.1 for 32 rows loaded two ahead at 256x256x80 / 512x512x80; 20-32
    // rows took 62.6-79.0 / 265.2-290.5.  At 256x256x80, 22 rows is the
    // shortest chunk whose 960 workgroups fit one resident round of the CUs.
    // At most 256 rows, which bounds the 1-D tables in LDS.
    const int rows = std::min(knob(K, sizeof(T), d, KNOB_ROWS, sizeof(T) == 8 ? 16 : 22), 256);
    const int nchunk = int((ny + 2 + rows - 1) / rows);
    if (int64_t(nchunk) * nz > (int64_t(1) << 30)) return fail(OEB_ERR_DOMAIN, std::string(K) + ": domain too large");
    auto kern = knob(K, sizeof(T), d, KNOB_DEPTH, 1) >= 2 ? hdiffsmag_adj_strip<T, 2> : hdiffsmag_adj_strip<T, 1>;
    hipLaunchKernelGGL(kern, dim3(unsigned(nchunk * nz)), dim3(unsigned(lanes)),
                       (size_t(20) * lanes + 5 * size_t(rows + 9)) * sizeof(T), hipStream_t(stream), a, g, rows,
                       nchunk);
    return launch_status(K);
  }
  const BoxGrid b1{0, 0, int(nx), int(ny)}, b2{-1, -1, int(nx + 2), int(ny + 2)};
  hipLaunchKernelGGL(hdiffsmag_adj_stage1<T>, box_blocks(b1, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g,
                     b1);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(hdiffsmag_adj_stage2<T>, box_blocks(b2, int(nz)), dim3(kTI, kTJ), 0, hipStream_t(stream), a, g,
                     b2);
  return launch_status(K);
}

// ==================================================================== fvtp2d
// Forward (kernels/fvtp2d.h:4-177), differentiated in q:
//   fy2 = PPM_j(q, cry)                                 i [-3, nx+2) j [0, ny+1)
//   q_i = (q area + yfx fy2 - yfx(j+1) fy2(j+1)) / ra_y   i [-3, nx+2) j [0, ny)
//   fx1' = PPM_i(q_i, crx), fx1 = 0.5 (fx1' + fx2) xfx    [0, nx) x [0, ny)
//   fx2 = PPM_i(q, crx)                                 i [0, nx+1)  j [-3, ny+2)
//   q_j = (q area + xfx fx2 - xfx(i+1) fx2(i+1)) / ra_x   i [0, nx)    j [-3, ny+2)
//   fy1' = PPM_j(q_j, cry), fy1 = 0.5 (fy1' + fy2) yfx    [0, nx) x [0, ny)
// With S (the monotonicity switch) and the upwind side frozen, a PPM flux is
// linear in its field: flux(m) = sum_{k=-3..2} w_k(m) f(m+k) (ppm_weights).
// The adjoint runs the stages backwards, each a gather:
//   K1  W0 = q_j_bar + PPM_j^T(0.5 yfx fy1_bar; q_j)        on q_j's region
//   K2  W1 = q_i_bar + PPM_i^T(0.5 xfx fx1_bar; q_i)        on q_i's region
//   K3  W2 = fx2_bar + 0.5 xfx fx1_bar + xfx (W0/ra_x - W0(i-1)/ra_x(i-1))   fx2's region
//   K4  W3 = fy2_bar + 0.5 yfx fy1_bar + yfx (W1/ra_y - W1(j-1)/ra_y(j-1))   fy2's region
//   K5  q_bar = area (W0/ra_x + W1/ra_y) + PPM_i^T(W2; q) + PPM_j^T(W3; q)   on q's box
// W0 and W1 are used only through W0/ra_x and W1/ra_y, so K1 and K2 store
// those quotients (work0, work1): each is divided once, and every later use
// reads the same double.
// q_i and q_j are the forward's outputs.  The FP32 forward forms q_i / q_j
// with a compensated sum; the derivative is that of the same expression.

// Weights w[k+3] of f(m+k), k = -3..2, in the flux at face m, from
// f(m-3) .. f(m+2) and the Courant number c (kernels/fvtp2d.h:12, 20-34).
template <class T>
__device__ __forceinline__ void ppm_weights(const T f[6], T c, T p1, T p2, T w[6]) {
  const T fm3 = f[0], fm2 = f[1], fm1 = f[2], f0 = f[3], fp1 = f[4], fp2 = f[5];
  const T al_m = ((p1 * (fm2 + fm1)) + (p2 * (fm3 + f0)));
  const T al_0 = ((p1 * (fm1 + f0)) + (p2 * (fm2 + fp1)));
  const T al_p = ((p1 * (f0 + fp1)) + (p2 * (fm1 + fp2)));
  const T almq_m = (al_m - fm1), br_m = (al_0 - fm1);
  const T almq_0 = (al_0 - f0), br_0 = (al_p - f0);
  const T smt_m = ((almq_m * br_m) < T(0)) ? T(1) : T(0);
  const T smt_0 = ((almq_0 * br_0) < T(0)) ? T(1) : T(0);
  const T S = (smt_m + (smt_0 * (smt_m == T(0) ? T(1) : T(0))));
  if (c > T(0)) {
    // flux = D f(m-1) + A al(m) + B al(m-1)
    const T A = S * (T(1) - c) * (T(1) - c), B = -S * (T(1) - c) * c;
    const T D = T(1) - S * (T(1) - c) * (T(1) - T(2) * c);
    w[0] = B * p2;
    w[1] = A * p2 + B * p1;
    w[2] = D + A * p1 + B * p1;
    w[3] = A * p1 + B * p2;
    w[4] = A * p2;
    w[5] = T(0);
  } else {
    // flux = D f(m) + A al(m) + B al(m+1)
    const T A = S * (T(1) + c) * (T(1) + c), B = S * (T(1) + c) * c;
    const T D = T(1) - S * (T(1) + c) * (T(1) + T(2) * c);
    w[0] = T(0);
    w[1] = A * p2;
    w[2] = A * p1 + B * p2;
    w[3] = D + A * p1 + B * p1;
    w[4] = A * p2 + B * p1;
    w[5] = B * p2;
  }
}

template <class T>
struct FvAdj {
  const T *q, *crx, *cry, *ra_x, *ra_y, *xfx, *yfx, *area, *q_i, *q_j;
  const T *qib, *qjb, *fx1b, *fx2b, *fy1b, *fy2b;
  T p1, p2;
  T *w0, *w1, *w2, *w3, *qbar;
};

// Regions (half-open, domain-relative) of the forward's outputs.
struct FvRegions {
  int nx, ny;
  __device__ bool qi(int i, int j) const { return in_range(i, -3, nx + 2) && in_range(j, 0, ny); }
  __device__ bool qj(int i, int j) const { return in_range(i, 0, nx) && in_range(j, -3, ny + 2); }
  __device__ bool fx2(int i, int j) const { return in_range(i, 0, nx + 1) && in_range(j, -3, ny + 2); }
  __device__ bool fy2(int i, int j) const { return in_range(i, -3, nx + 2) && in_range(j, 0, ny + 1); }
  __device__ bool in(int i, int j) const { return in_range(i, 0, nx) && in_range(j, 0, ny); }
};

// sum over faces m = q-2 .. q+3 along stride s (m inside `face`) of
// w_{q-m}(m) gbar(m), the weights taken from field f and Courant numbers c.
// `g(m)` returns the flux cotangent at offset m (from q, along the axis).
template <class T, class Face, class Gbar>
__device__ __forceinline__ T ppm_transpose(const T *f, const T *c, int64_t at, int s, T p1, T p2, Face face,
                                          Gbar gbar) {
  T r = T(0);
#pragma unroll
  for (int o = -2; o <= 3; ++o) {  // face m = q + o; q is offset k = -o from m
    if (!face(o)) continue;
    const T gb = gbar(o);
    if (gb == T(0)) continue;
    const int64_t m = at + int64_t(o) * s;
    T fv[6], w[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) fv[t] = f[m + int64_t(t - 3) * s];
    ppm_weights(fv, c[m], p1, p2, w);
    r += w[3 - o] * gb;
  }
  return r;
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) fvtp2d_adj_k1(FvAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const FvRegions R{g.nx, g.ny};
  const int sj = g.sj;
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * sj;
  const T r = ppm_transpose(
      a.q_j, a.cry, at, sj, a.p1, a.p2, [&](int o) { return R.in(i, j + o); },
      [&](int o) { return T(0.5) * a.yfx[at + int64_t(o) * sj] * a.fy1b[at + int64_t(o) * sj]; });
  a.w0[at] = (a.qjb[at] + r) / a.ra_x[at];
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) fvtp2d_adj_k2(FvAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const FvRegions R{g.nx, g.ny};
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * g.sj;
  const T r = ppm_transpose(
      a.q_i, a.crx, at, 1, a.p1, a.p2, [&](int o) { return R.in(i + o, j); },
      [&](int o) { return T(0.5) * a.xfx[at + o] * a.fx1b[at + o]; });
  a.w1[at] = (a.qib[at] + r) / a.ra_y[at];
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) fvtp2d_adj_k3(FvAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const FvRegions R{g.nx, g.ny};
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * g.sj;
  T r = a.fx2b[at];
  if (R.in(i, j)) r += T(0.5) * a.xfx[at] * a.fx1b[at];
  T d = T(0);
  if (R.qj(i, j)) d += a.w0[at];
  if (R.qj(i - 1, j)) d -= a.w0[at - 1];
  a.w2[at] = r + a.xfx[at] * d;
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) fvtp2d_adj_k4(FvAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const FvRegions R{g.nx, g.ny};
  const int sj = g.sj;
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * sj;
  T r = a.fy2b[at];
  if (R.in(i, j)) r += T(0.5) * a.yfx[at] * a.fy1b[at];
  T d = T(0);
  if (R.qi(i, j)) d += a.w1[at];
  if (R.qi(i, j - 1)) d -= a.w1[at - sj];
  a.w3[at] = r + a.yfx[at] * d;
}

template <class T>
__global__ void __launch_bounds__(kTI *kTJ) fvtp2d_adj_k5(FvAdj<T> a, Geo g, BoxGrid b) {
  int i, j;
  if (!box_point(b, i, j)) return;
  const FvRegions R{g.nx, g.ny};
  const int sj = g.sj;
  const int64_t at = int64_t(blockIdx.y) * g.sk + i + int64_t(j) * sj;
  T r = T(0);
  if (R.qj(i, j)) r += a.area[at] * a.w0[at];
  if (R.qi(i, j)) r += a.area[at] * a.w1[at];
  r += ppm_transpose(
      a.q, a.crx, at, 1, a.p1, a.p2, [&](int o) { return R.fx2(i + o, j); }, [&](int o) { return a.w2[at + o]; });
  r += ppm_transpose(
      a.q, a.cry, at, sj, a.p1, a.p2, [&](int o) { return R.fy2(i, j + o); },
      [&](int o) { return a.w3[at + int64_t(o) * sj]; });
  a.qbar[at] = r;
}

// Tiled form of the transposed PPM gathers (K1, K2, K5).  A workgroup takes
// a 64 x 8 tile of one level of the output box.  Per axis it forms, once for
// every face m whose flux reaches the tile, the six products w_k(m) gbar(m)
// (ppm_weights) into LDS; each point then adds the six products that name
// it, in the order ppm_transpose adds them (faces q-2 .. q+3; a face outside
// its region gives 0, one with a zero cotangent zero products).  ppm_transpose recomputes
// the weights of six faces per point instead.  K5 runs the x axis, then the y
// axis, through the same LDS, and forms each face's cotangent W2 or W3 (K3,
// K4) where it forms the face's products, so the tiled form is three launches
// and W2, W3 never go to memory.
// 64 x 16 tiles of 512 lanes (64.5 KB of LDS in FP64, two workgroups per CU):
// 393 us for the FP64 adjoint at 256x256x80 against 408 with 64 x 8 tiles of
// 256 lanes, which form the products of 13 rows for 8 instead of 21 for 16.
constexpr int kPX = 64, kPY = 16, kPNT = 512;
constexpr int kPPts = kPX * kPY / kPNT;  // points per lane

// The field tile of the hoisted form: the field the products read (q, q_i
// or q_j) on [i0-5, i0+69) x [j0-5, j0+21), the points that the faces of both
// axes of a tile read.
constexpr int kFW = kPX + 10, kFH = kPY + 10;

template <class T, bool H>
struct PpmLds {
  T p[6][kPX * (kPY + 5)];
  T f[H ? kFH * kFW : 1];
};

// Rows [R0, R1) x columns [C0, C1) of the field tile, from the level at
// `plane` of field f, each point clamped into the field's box [bi0, bi1) x
// [bj0, bj1).  A clamped value is read only by faces outside their region,
// which form no products.  Every lane's loads are issued before its first
// LDS store.
template <int R0, int R1, int C0, int C1, class T>
__device__ __forceinline__ void field_tile(T *sf, const T *f, int64_t plane, int sj, int i0, int j0, int bi0,
                                           int bi1, int bj0, int bj1) {
  constexpr int W = C1 - C0, N = (R1 - R0) * W, NU = (N + kPNT - 1) / kPNT;
  T v[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int e = min(int(threadIdx.x) + u * kPNT, N - 1);
    const int gi = min(max(i0 - 5 + C0 + e % W, bi0), bi1 - 1);
    const int gj = min(max(j0 - 5 + R0 + e / W, bj0), bj1 - 1);
    v[u] = f[plane + gi + int64_t(gj) * sj];
  }
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    const int e = int(threadIdx.x) + u * kPNT;
    if (e < N) sf[(R0 + e / W) * kFW + C0 + e % W] = v[u];
  }
}

// Products of the faces of the tile (i0, j0) along AXIS (0: i, 1: j).
template <int AXIS, class T, class Face, class Gbar>
__device__ __forceinline__ void ppm_products(PpmLds<T, false> &L, const T *f, const T *c, int64_t plane, int sj,
                                             int i0, int j0, T p1, T p2, Face face, Gbar gbar) {
  constexpr int FW = AXIS == 0 ? kPX + 5 : kPX, FH = AXIS == 0 ? kPY : kPY + 5;
  const int s = AXIS == 0 ? 1 : sj;
  for (int e = int(threadIdx.x); e < FW * FH; e += kPNT) {
    const int r = e / FW, cc = e % FW;
    const int fi = (AXIS == 0 ? i0 - 2 : i0) + cc, fj = (AXIS == 0 ? j0 : j0 - 2) + r;
    T pk[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (face(fi, fj)) {
      // No early exit on a zero cotangent: the field loads then do not wait
      // for the cotangent's loads (a zero cotangent gives zero products).
      const int64_t m = plane + fi + int64_t(fj) * sj;
      const T gb = gbar(m, fi, fj);
      T fv[6], w[6];
#pragma unroll
      for (int t = 0; t < 6; ++t) fv[t] = f[m + int64_t(t - 3) * s];
      ppm_weights(fv, c[m], p1, p2, w);
#pragma unroll
      for (int t = 0; t < 6; ++t) pk[t] = w[t] * gb;
    }
#pragma unroll
    for (int t = 0; t < 6; ++t) L.p[t][e] = pk[t];
  }
}

// Sum over faces q-2 .. q+3 along AXIS of the products that name the tile
// point (lx, ly).
template <int AXIS, class T, bool H>
__device__ __forceinline__ T ppm_gather(const PpmLds<T, H> &L, int lx, int ly) {
  constexpr int FW = AXIS == 0 ? kPX + 5 : kPX;
  T r = T(0);
#pragma unroll
  for (int o = -2; o <= 3; ++o) {
    const int idx = AXIS == 0 ? ly * FW + (lx + o + 2) : (ly + o + 2) * FW + lx;
    r += L.p[3 - o][idx];
  }
  return r;
}

// The two-pass face loop of the hoisted form.  Pass 1 issues, for every face
// of the lane, the global loads of its Courant number and of the inputs of its
// cotangent (FaceLd), at addresses clamped into the face's region or into the
// box of the array read; pass 2 forms the products from them and from the
// field tile in LDS.  A value loaded at a clamped address is used only where
// the forward's region test holds, and there the address is the unclamped
// one, so both passes give the bits of ppm_products.  The loop in
// ppm_products waits for each face's loads in turn.
template <class T>
struct FaceLd {
  T c, g0, g1, g2, g3, g4;
};

template <int AXIS>
struct FaceGrid {
  static constexpr int FW = AXIS == 0 ? kPX + 5 : kPX, FH = AXIS == 0 ? kPY : kPY + 5;
  static constexpr int N = FW * FH, NU = (N + kPNT - 1) / kPNT;
};

// Pass 1: ld(fi, fj) for every face of the lane, (fi, fj) clamped into
// [fi_lo, fi_hi) x [fj_lo, fj_hi).
template <int AXIS, class T, class Load>
__device__ __forceinline__ void face_loads(FaceLd<T> (&v)[FaceGrid<AXIS>::NU], int i0, int j0, int fi_lo, int fi_hi,
                                           int fj_lo, int fj_hi, Load ld) {
  using G = FaceGrid<AXIS>;
#pragma unroll
  for (int u = 0; u < G::NU; ++u) {
    const int e = min(int(threadIdx.x) + u * kPNT, G::N - 1);
    const int fi = min(max((AXIS == 0 ? i0 - 2 : i0) + e % G::FW, fi_lo), fi_hi - 1);
    const int fj = min(max((AXIS == 0 ? j0 : j0 - 2) + e / G::FW, fj_lo), fj_hi - 1);
    v[u] = ld(fi, fj);
  }
}

// Pass 2: the products of ppm_products, with gbar(v, fi, fj) the face's
// cotangent from its loaded values.
template <int AXIS, class T, class Face, class Gbar>
__device__ __forceinline__ void face_products(PpmLds<T, true> &L, const FaceLd<T> (&v)[FaceGrid<AXIS>::NU], int i0,
                                              int j0, T p1, T p2, Face face, Gbar gbar) {
  using G = FaceGrid<AXIS>;
#pragma unroll
  for (int u = 0; u < G::NU; ++u) {
    const int e = int(threadIdx.x) + u * kPNT;
    if (e < G::N) {
      const int r = e / G::FW, cc = e % G::FW;
      const int fi = (AXIS == 0 ? i0 - 2 : i0) + cc, fj = (AXIS == 0 ? j0 : j0 - 2) + r;
      T pk[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
      if (face(fi, fj)) {
        const T gb = gbar(v[u], fi, fj);
        // Face (fi, fj) reads the field at fi - 3 .. fi + 2 (tile columns
        // cc .. cc + 5 of row r + 5) or fj - 3 .. fj + 2 (tile rows r .. r + 5
        // of column cc + 5).
        T fv[6], w[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) fv[t] = AXIS == 0 ? L.f[(r + 5) * kFW + cc + t] : L.f[(r + t) * kFW + cc + 5];
        ppm_weights(fv, v[u].c, p1, p2, w);
#pragma unroll
        for (int t = 0; t < 6; ++t) pk[t] = w[t] * gb;
      }
#pragma unroll
      for (int t = 0; t < 6; ++t) L.p[t][e] = pk[t];
    }
  }
}

// Tile `blk` of stage MODE in the hoisted form: the field tile, the faces'
// loads (face_loads) and the tile points' own loads are issued before the
// first barrier; K5 issues its y faces' loads after forming the x products.
// The same expressions in the same order as fvtp2d_adj_ptile_body.
template <class T, int MODE>
__device__ __forceinline__ void fvtp2d_adj_htile_body(PpmLds<T, true> &L, const FvAdj<T> &a, const Geo &g,
                                                      const BoxGrid &b, int blk, int k) {
  const int ntx = (b.ni + kPX - 1) / kPX;
  const int i0 = b.ilo + (blk % ntx) * kPX, j0 = b.jlo + (blk / ntx) * kPY;
  const int sj = g.sj, nx = g.nx, ny = g.ny;
  const int64_t plane = int64_t(k) * g.sk;
  const FvRegions R{nx, ny};
  const T p1 = a.p1, p2 = a.p2;
  auto at_of = [&](int i, int j) { return plane + i + int64_t(j) * sj; };
  auto clampi = [](int x, int lo, int hi) { return min(max(x, lo), hi - 1); };
  auto point = [&](int u, int &lx, int &ly) {
    const int e = int(threadIdx.x) + u * kPNT;
    lx = e % kPX;
    ly = e / kPX;
    return in_range(i0 + lx, b.ilo, b.ilo + b.ni) && in_range(j0 + ly, b.jlo, b.jlo + b.nj);
  };
  if (MODE == 1 || MODE == 2) {  // K1 (along j over q_j) or K2 (along i over q_i)
    constexpr int AX = MODE == 1 ? 1 : 0;
    FaceLd<T> v[FaceGrid<AX>::NU];
    T tb[kPPts], tr[kPPts];  // q_j_bar, ra_x (K1) or q_i_bar, ra_y (K2) at the tile points
    if (MODE == 1)
      field_tile<0, kFH, 5, 5 + kPX>(L.f, a.q_j, plane, sj, i0, j0, 0, nx, -3, ny + 2);
    else
      field_tile<5, 5 + kPY, 0, kFW>(L.f, a.q_i, plane, sj, i0, j0, -3, nx + 2, 0, ny);
    face_loads<AX, T>(v, i0, j0, 0, nx, 0, ny, [&](int fi, int fj) {
      const int64_t m = at_of(fi, fj);
      return MODE == 1 ? FaceLd<T>{a.cry[m], a.yfx[m], a.fy1b[m], T(0), T(0), T(0)}
                       : FaceLd<T>{a.crx[m], a.xfx[m], a.fx1b[m], T(0), T(0), T(0)};
    });
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      point(u, lx, ly);
      const int64_t at = MODE == 1 ? at_of(clampi(i0 + lx, 0, nx), clampi(j0 + ly, -3, ny + 2))
                                   : at_of(clampi(i0 + lx, -3, nx + 2), clampi(j0 + ly, 0, ny));
      tb[u] = MODE == 1 ? a.qjb[at] : a.qib[at];
      tr[u] = MODE == 1 ? a.ra_x[at] : a.ra_y[at];
    }
    lds_barrier();
    face_products<AX, T>(L, v, i0, j0, p1, p2, [&](int i, int j) { return R.in(i, j); },
                         [&](const FaceLd<T> &x, int, int) { return T(0.5) * x.g0 * x.g1; });
    lds_barrier();
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      if (point(u, lx, ly)) {
        const int64_t at = at_of(i0 + lx, j0 + ly);
        if (MODE == 1)
          a.w0[at] = (tb[u] + ppm_gather<1>(L, lx, ly)) / tr[u];
        else
          a.w1[at] = (tb[u] + ppm_gather<0>(L, lx, ly)) / tr[u];
      }
    }
  } else {  // K5, with K3 and K4 formed per face
    FaceLd<T> vx[FaceGrid<0>::NU];
    field_tile<0, kFH, 0, kFW>(L.f, a.q, plane, sj, i0, j0, -3, nx + 3, -3, ny + 3);
    // x faces on fx2's region [0, nx+1) x [-3, ny+2): W2 from fx2_bar, xfx,
    // fx1_bar (read on [0, nx) x [0, ny)) and W0 at faces i and i-1 (read on
    // q_j's region [0, nx) x [-3, ny+2)).
    face_loads<0, T>(vx, i0, j0, 0, nx + 1, -3, ny + 2, [&](int fi, int fj) {
      const int64_t m = at_of(fi, fj);
      return FaceLd<T>{a.crx[m], a.fx2b[m], a.xfx[m], a.fx1b[at_of(min(fi, nx - 1), clampi(fj, 0, ny))],
                       a.w0[at_of(min(fi, nx - 1), fj)], a.w0[at_of(max(fi - 1, 0), fj)]};
    });
    lds_barrier();
    // W2 at face (i, j): fvtp2d_adj_k3's expression.
    face_products<0, T>(L, vx, i0, j0, p1, p2, [&](int i, int j) { return R.fx2(i, j); },
                        [&](const FaceLd<T> &x, int i, int j) {
                          T r = x.g0;
                          if (R.in(i, j)) r += T(0.5) * x.g1 * x.g2;
                          T d = T(0);
                          if (R.qj(i, j)) d += x.g3;
                          if (R.qj(i - 1, j)) d -= x.g4;
                          return r + x.g1 * d;
                        });
    // y faces on fy2's region [-3, nx+2) x [0, ny+1): W3 from fy2_bar, yfx,
    // fy1_bar (read on [0, nx) x [0, ny)) and W1 at faces j and j-1 (read on
    // q_i's region [-3, nx+2) x [0, ny)); and the tile points' area, W0, W1.
    FaceLd<T> vy[FaceGrid<1>::NU];
    face_loads<1, T>(vy, i0, j0, -3, nx + 2, 0, ny + 1, [&](int fi, int fj) {
      const int64_t m = at_of(fi, fj);
      return FaceLd<T>{a.cry[m], a.fy2b[m], a.yfx[m], a.fy1b[at_of(clampi(fi, 0, nx), min(fj, ny - 1))],
                       a.w1[at_of(fi, min(fj, ny - 1))], a.w1[at_of(fi, max(fj - 1, 0))]};
    });
    T ta[kPPts], t0[kPPts], t1[kPPts];
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      point(u, lx, ly);
      const int i = i0 + lx, j = j0 + ly;
      ta[u] = a.area[at_of(clampi(i, -3, nx + 2), clampi(j, -3, ny + 2))];
      t0[u] = a.w0[at_of(clampi(i, 0, nx), clampi(j, -3, ny + 2))];
      t1[u] = a.w1[at_of(clampi(i, -3, nx + 2), clampi(j, 0, ny))];
    }
    lds_barrier();
    T rx[kPPts];
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      point(u, lx, ly);
      rx[u] = ppm_gather<0>(L, lx, ly);
    }
    lds_barrier();
    // W3 at face (i, j): fvtp2d_adj_k4's expression.
    face_products<1, T>(L, vy, i0, j0, p1, p2, [&](int i, int j) { return R.fy2(i, j); },
                        [&](const FaceLd<T> &x, int i, int j) {
                          T r = x.g0;
                          if (R.in(i, j)) r += T(0.5) * x.g1 * x.g2;
                          T d = T(0);
                          if (R.qi(i, j)) d += x.g3;
                          if (R.qi(i, j - 1)) d -= x.g4;
                          return r + x.g1 * d;
                        });
    lds_barrier();
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      if (point(u, lx, ly)) {
        const int i = i0 + lx, j = j0 + ly;
        T r = T(0);
        if (R.qj(i, j)) r += ta[u] * t0[u];
        if (R.qi(i, j)) r += ta[u] * t1[u];
        r += rx[u];
        r += ppm_gather<1>(L, lx, ly);
        a.qbar[at_of(i, j)] = r;
      }
    }
  }
}

// Tile `blk` (of the box b) of one level of stage MODE, with the LDS of its
// workgroup.
template <class T, int MODE>
__device__ __forceinline__ void fvtp2d_adj_ptile_body(PpmLds<T, false> &L, const FvAdj<T> &a, const Geo &g,
                                                      const BoxGrid &b, int blk, int k) {
  const int ntx = (b.ni + kPX - 1) / kPX;
  const int i0 = b.ilo + (blk % ntx) * kPX, j0 = b.jlo + (blk / ntx) * kPY;
  const int sj = g.sj;
  const int64_t plane = int64_t(k) * g.sk;
  const FvRegions R{g.nx, g.ny};
  const T p1 = a.p1, p2 = a.p2;
  auto point = [&](int u, int &lx, int &ly) {
    const int e = int(threadIdx.x) + u * kPNT;
    lx = e % kPX;
    ly = e / kPX;
    return in_range(i0 + lx, b.ilo, b.ilo + b.ni) && in_range(j0 + ly, b.jlo, b.jlo + b.nj);
  };
  if (MODE == 1) {  // K1
    ppm_products<1>(L, a.q_j, a.cry, plane, sj, i0, j0, p1, p2, [&](int i, int j) { return R.in(i, j); },
                    [&](int64_t m, int, int) { return T(0.5) * a.yfx[m] * a.fy1b[m]; });
    lds_barrier();
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      if (point(u, lx, ly)) {
        const int64_t at = plane + (i0 + lx) + int64_t(j0 + ly) * sj;
        a.w0[at] = (a.qjb[at] + ppm_gather<1>(L, lx, ly)) / a.ra_x[at];
      }
    }
  } else if (MODE == 2) {  // K2
    ppm_products<0>(L, a.q_i, a.crx, plane, sj, i0, j0, p1, p2, [&](int i, int j) { return R.in(i, j); },
                    [&](int64_t m, int, int) { return T(0.5) * a.xfx[m] * a.fx1b[m]; });
    lds_barrier();
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      if (point(u, lx, ly)) {
        const int64_t at = plane + (i0 + lx) + int64_t(j0 + ly) * sj;
        a.w1[at] = (a.qib[at] + ppm_gather<0>(L, lx, ly)) / a.ra_y[at];
      }
    }
  } else {  // K5, with K3 and K4 formed per face (w2 and w3 never go to memory)
    // W2 at face (i, j): fvtp2d_adj_k3's expression.
    auto w2_at = [&](int64_t m, int i, int j) -> T {
      T r = a.fx2b[m];
      if (R.in(i, j)) r += T(0.5) * a.xfx[m] * a.fx1b[m];
      T d = T(0);
      if (R.qj(i, j)) d += a.w0[m];
      if (R.qj(i - 1, j)) d -= a.w0[m - 1];
      return r + a.xfx[m] * d;
    };
    // W3 at face (i, j): fvtp2d_adj_k4's expression.
    auto w3_at = [&](int64_t m, int i, int j) -> T {
      T r = a.fy2b[m];
      if (R.in(i, j)) r += T(0.5) * a.yfx[m] * a.fy1b[m];
      T d = T(0);
      if (R.qi(i, j)) d += a.w1[m];
      if (R.qi(i, j - 1)) d -= a.w1[m - sj];
      return r + a.yfx[m] * d;
    };
    ppm_products<0>(L, a.q, a.crx, plane, sj, i0, j0, p1, p2, [&](int i, int j) { return R.fx2(i, j); }, w2_at);
    lds_barrier();
    T rx[kPPts];
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      point(u, lx, ly);
      rx[u] = ppm_gather<0>(L, lx, ly);
    }
    lds_barrier();
    ppm_products<1>(L, a.q, a.cry, plane, sj, i0, j0, p1, p2, [&](int i, int j) { return R.fy2(i, j); }, w3_at);
    lds_barrier();
#pragma unroll
    for (int u = 0; u < kPPts; ++u) {
      int lx, ly;
      if (point(u, lx, ly)) {
        const int i = i0 + lx, j = j0 + ly;
        const int64_t at = plane + i + int64_t(j) * sj;
        T r = T(0);
        if (R.qj(i, j)) r += a.area[at] * a.w0[at];
        if (R.qi(i, j)) r += a.area[at] * a.w1[at];
        r += rx[u];
        r += ppm_gather<1>(L, lx, ly);
        a.qbar[at] = r;
      }
    }
  }
}

// With irow set, the tiles of one row of one level go to one XCD, in
// consecutive slots of its dispatch order (as Tiles::irow): linear block
// x + 8 (ci + ntx q) takes row r = x + 8 q (r = tile row + nty level).  The
// host sets it only when nty nz is a multiple of 8.
template <class T, int MODE, bool H>
__global__ void __launch_bounds__(kPNT) fvtp2d_adj_ptile(FvAdj<T> a, Geo g, BoxGrid b, int irow) {
  __shared__ PpmLds<T, H> L;
  int blk = int(blockIdx.x), k = int(blockIdx.y);
  if (irow) {
    const int ntx = (b.ni + kPX - 1) / kPX, nty = (b.nj + kPY - 1) / kPY;
    const int lin = int(blockIdx.x) + int(gridDim.x) * int(blockIdx.y);
    const int q = lin >> 3, r = (lin & 7) + 8 * (q / ntx);
    blk = q % ntx + ntx * (r % nty);
    k = r / nty;
  }
  if constexpr (H)
    fvtp2d_adj_htile_body<T, MODE>(L, a, g, b, blk, k);
  else
    fvtp2d_adj_ptile_body<T, MODE>(L, a, g, b, blk, k);
}

inline int ptile_count(const BoxGrid &b) { return ((b.ni + kPX - 1) / kPX) * ((b.nj + kPY - 1) / kPY); }

// K1 and K2 in one launch: the first n1 tiles of each level run K1 on b1,
// the others K2 on b2.  They read different fields and write work0 and work1.
template <class T, bool H>
__global__ void __launch_bounds__(kPNT) fvtp2d_adj_ptile12(FvAdj<T> a, Geo g, BoxGrid b1, BoxGrid b2, int n1) {
  __shared__ PpmLds<T, H> L;
  const int blk = int(blockIdx.x);
  if constexpr (H) {
    if (blk < n1)
      fvtp2d_adj_htile_body<T, 1>(L, a, g, b1, blk, int(blockIdx.y));
    else
      fvtp2d_adj_htile_body<T, 2>(L, a, g, b2, blk - n1, int(blockIdx.y));
  } else {
    if (blk < n1)
      fvtp2d_adj_ptile_body<T, 1>(L, a, g, b1, blk, int(blockIdx.y));
    else
      fvtp2d_adj_ptile_body<T, 2>(L, a, g, b2, blk - n1, int(blockIdx.y));
  }
}

inline dim3 ptile_blocks(const BoxGrid &b, int nz) { return dim3(unsigned(ptile_count(b)), unsigned(nz)); }

// Validates the arrays of an fvtp2d adjoint call other than its work arrays:
// fills a (inputs and q_bar), g and the alias set.
template <class T, class S>
int fvtp2d_adj_args(const char *K, const oeb_domain *d, const S *q, const S *crx, const S *cry, const S *ra_x,
                    const S *ra_y, const S *xfx, const S *yfx, const S *area, const S *q_i, const S *q_j,
                    const S *q_i_bar, const S *q_j_bar, const S *fx1_bar, const S *fx2_bar, const S *fy1_bar,
                    const S *fy2_bar, const S *q_bar, T p1, T p2, FvAdj<T> &a, Geo &g, AliasSet &as) {
  OEB_TRY(check_domain(K, d));
  OEB_TRY(refuse_while_recording(K));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  g = Geo{int(nx), int(ny), int(nz), 0, 0};
  const Box r0{0, nx, 0, ny, 0, nz};
  const Box r_fy2{-3, nx + 2, 0, ny + 1, 0, nz}, r_fx2{0, nx + 1, -3, ny + 2, 0, nz};
  const Box r_qi{-3, nx + 2, 0, ny, 0, nz}, r_qj{0, nx, -3, ny + 2, 0, nz};
  const Box r_q{-3, nx + 3, -3, ny + 3, 0, nz}, r_area{-3, nx + 2, -3, ny + 2, 0, nz};
  a.p1 = p1, a.p2 = p2;
  const T *p_qb;
  OEB_TRY(check3(K, "q", q, r_q, g, true, &a.q));
  OEB_TRY(check3(K, "crx", crx, r_fx2, g, false, &a.crx));
  OEB_TRY(check3(K, "cry", cry, r_fy2, g, false, &a.cry));
  OEB_TRY(check3(K, "ra_x", ra_x, r_qj, g, false, &a.ra_x));
  OEB_TRY(check3(K, "ra_y", ra_y, r_qi, g, false, &a.ra_y));
  OEB_TRY(check3(K, "xfx", xfx, r_fx2, g, false, &a.xfx));
  OEB_TRY(check3(K, "yfx", yfx, r_fy2, g, false, &a.yfx));
  OEB_TRY(check3(K, "area", area, r_area, g, false, &a.area));
  OEB_TRY(check3(K, "q_i", q_i, r_qi, g, false, &a.q_i));
  OEB_TRY(check3(K, "q_j", q_j, r_qj, g, false, &a.q_j));
  OEB_TRY(check3(K, "q_i_bar", q_i_bar, r_qi, g, false, &a.qib));
  OEB_TRY(check3(K, "q_j_bar", q_j_bar, r_qj, g, false, &a.qjb));
  OEB_TRY(check3(K, "fx1_bar", fx1_bar, r0, g, false, &a.fx1b));
  OEB_TRY(check3(K, "fx2_bar", fx2_bar, r_fx2, g, false, &a.fx2b));
  OEB_TRY(check3(K, "fy1_bar", fy1_bar, r0, g, false, &a.fy1b));
  OEB_TRY(check3(K, "fy2_bar", fy2_bar, r_fy2, g, false, &a.fy2b));
  OEB_TRY(check3(K, "q_bar", q_bar, r_q, g, false, &p_qb));
  as.add("q_bar", p_qb, r_q, g, true);
  as.add("q", a.q, r_q, g, false);
  as.add("q_i", a.q_i, r_qi, g, false);
  as.add("q_j", a.q_j, r_qj, g, false);
  as.add("q_i_bar", a.qib, r_qi, g, false);
  as.add("q_j_bar", a.qjb, r_qj, g, false);
  as.add("fx1_bar", a.fx1b, r0, g, false);
  as.add("fx2_bar", a.fx2b, r_fx2, g, false);
  as.add("fy1_bar", a.fy1b, r0, g, false);
  as.add("fy2_bar", a.fy2b, r_fy2, g, false);
  a.qbar = const_cast<T *>(p_qb);
  return OEB_OK;
}

// The two launches of the tiled adjoint (K1 and K2, then K5) on levels
// [0, g.nz) of the arrays in a.
template <class T>
int fvtp2d_adj_tiled_launch(const char *K, const oeb_domain *d, const FvAdj<T> &a, const Geo &g,
                            hipStream_t st) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const BoxGrid bqj{0, -3, nx, ny + 5}, bqi{-3, 0, nx + 5, ny};
  const BoxGrid bq{-3, -3, nx + 6, ny + 6};
  const int n1 = ptile_count(bqj), n2 = ptile_count(bqi);
  // Tile knob 2 (the default): the hoisted form, bit for bit the same as
  // knob 1, the per-face loop of ppm_products.  At 256x256x80 FP64 / FP32:
  // 295.6 / 150.5 us against 393.8 / 193.8 (1411 / 1057 us FP64 at
  // 512x512x80); hoisted K1 and K2 alone gave 384 us, hoisted K5 alone 304.
  const bool hoist = knob(K, sizeof(T), d, KNOB_TILE, 2) == 2;
  const dim3 g12(unsigned(n1 + n2), unsigned(nz));
  if (hoist)
    hipLaunchKernelGGL((fvtp2d_adj_ptile12<T, true>), g12, dim3(kPNT), 0, st, a, g, bqj, bqi, n1);
  else
    hipLaunchKernelGGL((fvtp2d_adj_ptile12<T, false>), g12, dim3(kPNT), 0, st, a, g, bqj, bqi, n1);
  OEB_TRY(launch_status(K));
  // K5 with the tiles of a row on one XCD (order knob 2, the default;
  // 0 for row order): 408 us against 411 for the FP64 adjoint with 64 x 8 tiles.
  const int nty = (bq.nj + kPY - 1) / kPY;
  const int irow = (knob(K, sizeof(T), d, KNOB_ORDER, 2) == 2 && (int64_t(nty) * nz) % 8 == 0) ? 1 : 0;
  if (hoist)
    hipLaunchKernelGGL((fvtp2d_adj_ptile<T, 5, true>), ptile_blocks(bq, nz), dim3(kPNT), 0, st, a, g, bq, irow);
  else
    hipLaunchKernelGGL((fvtp2d_adj_ptile<T, 5, false>), ptile_blocks(bq, nz), dim3(kPNT), 0, st, a, g, bq, irow);
  return launch_status(K);
}

template <class T, class S>
int fvtp2d_adj_run(const oeb_domain *d, const S *q, const S *crx, const S *cry, const S *ra_x, const S *ra_y,
                   const S *xfx, const S *yfx, const S *area, const S *q_i, const S *q_j, const S *q_i_bar,
                   const S *q_j_bar, const S *fx1_bar, const S *fx2_bar, const S *fy1_bar, const S *fy2_bar,
                   const S *q_bar, const S *const work[4], T p1, T p2, void *stream) {
  const char *K = "fvtp2d_adj";
  FvAdj<T> a;
  Geo g;
  AliasSet as;
  OEB_TRY(fvtp2d_adj_args<T>(K, d, q, crx, cry, ra_x, ra_y, xfx, yfx, area, q_i, q_j, q_i_bar, q_j_bar, fx1_bar,
                             fx2_bar, fy1_bar, fy2_bar, q_bar, p1, p2, a, g, as));
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  const Box r_fy2{-3, nx + 2, 0, ny + 1, 0, nz}, r_fx2{0, nx + 1, -3, ny + 2, 0, nz};
  const Box r_qi{-3, nx + 2, 0, ny, 0, nz}, r_qj{0, nx, -3, ny + 2, 0, nz};
  const Box wr[4] = {r_qj, r_qi, r_fx2, r_fy2};
  static const char *const kWork[4] = {"work0", "work1", "work2", "work3"};
  const T *p_w[4];
  for (int t = 0; t < 4; ++t) OEB_TRY(check3(K, kWork[t], work[t], wr[t], g, false, &p_w[t]));
  for (int t = 0; t < 4; ++t) as.add(kWork[t], p_w[t], wr[t], g, true);
  OEB_TRY(as.check(K));
  a.w0 = const_cast<T *>(p_w[0]), a.w1 = const_cast<T *>(p_w[1]);
  a.w2 = const_cast<T *>(p_w[2]), a.w3 = const_cast<T *>(p_w[3]);
  const BoxGrid bqj{0, -3, int(nx), int(ny + 5)}, bqi{-3, 0, int(nx + 5), int(ny)};
  const BoxGrid bfx2{0, -3, int(nx + 1), int(ny + 5)}, bfy2{-3, 0, int(nx + 5), int(ny + 1)};
  const BoxGrid bq{-3, -3, int(nx + 6), int(ny + 6)};
  hipStream_t st = hipStream_t(stream);
  // The tiled form (fvtp2d_adj_ptile: K1 and K2 in one launch, then K5, which
  // forms K3 and K4 inside) by default; tile knob 0 selects the five per-point passes,
  // with W2 and W3 through work2 and work3.
  const bool tiled = knob(K, sizeof(T), d, KNOB_TILE, 2) != 0;
  if (tiled) return fvtp2d_adj_tiled_launch<T>(K, d, a, g, st);
  hipLaunchKernelGGL(fvtp2d_adj_k1<T>, box_blocks(bqj, int(nz)), dim3(kTI, kTJ), 0, st, a, g, bqj);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(fvtp2d_adj_k2<T>, box_blocks(bqi, int(nz)), dim3(kTI, kTJ), 0, st, a, g, bqi);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(fvtp2d_adj_k3<T>, box_blocks(bfx2, int(nz)), dim3(kTI, kTJ), 0, st, a, g, bfx2);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(fvtp2d_adj_k4<T>, box_blocks(bfy2, int(nz)), dim3(kTI, kTJ), 0, st, a, g, bfy2);
  OEB_TRY(launch_status(K));
  hipLaunchKernelGGL(fvtp2d_adj_k5<T>, box_blocks(bq, int(nz)), dim3(kTI, kTJ), 0, st, a, g, bq);
  return launch_status(K);
}

}  // namespace
}  // namespace oeb

// ------------------------------------------------- allocation-free fvtp2d adjoint
// oeb_fvtp2d_adj_f64_noalloc keeps the two work planes of the tiled fvtp2d
// adjoint (W0 / ra_x on the region of q_j and W1 / ra_y on the region of q_i,
// written by K1 / K2 and read by K5 in the same call) in this file-scope
// device array instead of caller-passed work arrays.  Each GPU the library
// runs on holds its own copy until the process ends.  Host code reaches it
// only through the device address hipGetSymbolAddress returns for the current
// device, never through its name.
static constexpr size_t kNoAllocScratchDoubles = size_t(4) << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace oeb {
namespace {

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // per device ordinal: address of `scratch` on it

// The device address of `scratch` on the current device, looked up on the
// first call on each device.
int scratch_on_current_device(const char *K, double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(OEB_ERR_HIP, std::string(K) + ": hipGetDevice: " + hipGetErrorString(e));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (size_t(dev) >= g_scratch_ptr.size()) g_scratch_ptr.resize(size_t(dev) + 1, nullptr);
  if (g_scratch_ptr[size_t(dev)] == nullptr) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(OEB_ERR_HIP, std::string(K) + ": hipGetSymbolAddress(scratch): " + hipGetErrorString(e));
    g_scratch_ptr[size_t(dev)] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[size_t(dev)];
  return OEB_OK;
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_fvtp2d_adj_f64_noalloc(const oeb_domain *d, const oeb_s3 *q, const oeb_s3 *crx,
                                          const oeb_s3 *cry, const oeb_s3 *ra_x, const oeb_s3 *ra_y,
                                          const oeb_s3 *xfx, const oeb_s3 *yfx, const oeb_s3 *area,
                                          const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *q_i_bar,
                                          const oeb_s3 *q_j_bar, const oeb_s3 *fx1_bar, const oeb_s3 *fx2_bar,
                                          const oeb_s3 *fy1_bar, const oeb_s3 *fy2_bar, const oeb_s3 *q_bar,
                                          double p1, double p2, void *stream) {
  using namespace oeb;
  const char *K = "fvtp2d_adj_noalloc";
  FvAdj<double> a;
  Geo g;
  AliasSet as;
  OEB_TRY(fvtp2d_adj_args<double>(K, d, q, crx, cry, ra_x, ra_y, xfx, yfx, area, q_i, q_j, q_i_bar, q_j_bar,
                                  fx1_bar, fx2_bar, fy1_bar, fy2_bar, q_bar, p1, p2, a, g, as));
  OEB_TRY(as.check(K));
  double *scr = nullptr;
  OEB_TRY(scratch_on_current_device(K, &scr));
  // Element offsets, from (0, 0) of a level, that the work planes span: W0 on
  // i [0, nx) x j [-3, ny+2), W1 on i [-3, nx+2) x j [0, ny).  A chunk of kc
  // levels needs (kc - 1) sk + (hi - lo) elements per plane, two planes.
  const int64_t sj = g.sj, sk = g.sk;
  const int64_t lo = -3 * sj - 3, hi = (int64_t(g.ny) + 1) * sj + g.nx + 2;
  const int64_t cap = int64_t(kNoAllocScratchDoubles) / 2;
  if (hi - lo > cap) {
    std::ostringstream m;
    m << K << ": one level's two work planes need " << 2 * (hi - lo) * 8 << " bytes (j stride " << sj
      << ", ny " << g.ny << "); the scratch holds " << kNoAllocScratchDoubles * 8
      << " bytes (at most about 2.1 million elements per plane); use oeb_fvtp2d_adj_f64 with work arrays";
    return fail(OEB_ERR_DOMAIN, m.str());
  }
  const int kc = int(std::min<int64_t>(g.nz, (cap - (hi - lo)) / sk + 1));
  const int64_t span = int64_t(kc - 1) * sk + (hi - lo);
  // Level chunks one after another on `stream`: stream order keeps a chunk's
  // K5 ahead of the next chunk's K1 / K2, which rewrite the planes.
  for (int k0 = 0; k0 < g.nz; k0 += kc) {
    FvAdj<double> c = a;
    const int64_t sh = int64_t(k0) * sk;
    for (const double **f : {&c.q, &c.crx, &c.cry, &c.ra_x, &c.ra_y, &c.xfx, &c.yfx, &c.area, &c.q_i, &c.q_j,
                             &c.qib, &c.qjb, &c.fx1b, &c.fx2b, &c.fy1b, &c.fy2b})
      *f += sh;
    c.qbar += sh;
    c.w0 = scr - lo;
    c.w1 = scr + span - lo;
    c.w2 = c.w3 = nullptr;
    Geo gc = g;
    gc.nz = std::min(kc, g.nz - k0);
    OEB_TRY(fvtp2d_adj_tiled_launch<double>(K, d, c, gc, hipStream_t(stream)));
  }
  return OEB_OK;
}

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

extern "C" int oeb_hdiffsmag_adj_f64(const oeb_domain *d, const oeb_s3 *uin, const oeb_s3 *vin, const oeb_s3 *mask,
                                     const oeb_s1 *crlavo, const oeb_s1 *crlavu, const oeb_s1 *crlato,
                                     const oeb_s1 *crlatu, const oeb_s1 *acrlat0, const oeb_s3 *uout_bar,
                                     const oeb_s3 *vout_bar, const oeb_s3 *uin_bar, const oeb_s3 *vin_bar,
                                     const oeb_s3 *work0, const oeb_s3 *work1, const oeb_s3 *work2,
                                     const oeb_s3 *work3, double eddlat, double eddlon, double tau_smag,
                                     double weight_smag, void *stream) {
  const oeb_s3 *const w[4] = {work0, work1, work2, work3};
  return oeb::hdiffsmag_adj_run<double>(d, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, uout_bar,
                                        vout_bar, uin_bar, vin_bar, w, eddlat, eddlon, tau_smag, weight_smag, stream);
}

extern "C" int oeb_hdiffsmag_adj_f32(const oeb_domain *d, const oeb_s3f *uin, const oeb_s3f *vin,
                                     const oeb_s3f *mask, const oeb_s1f *crlavo, const oeb_s1f *crlavu,
                                     const oeb_s1f *crlato, const oeb_s1f *crlatu, const oeb_s1f *acrlat0,
                                     const oeb_s3f *uout_bar, const oeb_s3f *vout_bar, const oeb_s3f *uin_bar,
                                     const oeb_s3f *vin_bar, const oeb_s3f *work0, const oeb_s3f *work1,
                                     const oeb_s3f *work2, const oeb_s3f *work3, float eddlat, float eddlon,
                                     float tau_smag, float weight_smag, void *stream) {
  const oeb_s3f *const w[4] = {work0, work1, work2, work3};
  return oeb::hdiffsmag_adj_run<float>(d, uin, vin, mask, crlavo, crlavu, crlato, crlatu, acrlat0, uout_bar,
                                       vout_bar, uin_bar, vin_bar, w, eddlat, eddlon, tau_smag, weight_smag, stream);
}

extern "C" int oeb_fvtp2d_adj_f64(const oeb_domain *d, const oeb_s3 *q, const oeb_s3 *crx, const oeb_s3 *cry,
                                  const oeb_s3 *ra_x, const oeb_s3 *ra_y, const oeb_s3 *xfx, const oeb_s3 *yfx,
                                  const oeb_s3 *area, const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *q_i_bar,
                                  const oeb_s3 *q_j_bar, const oeb_s3 *fx1_bar, const oeb_s3 *fx2_bar,
                                  const oeb_s3 *fy1_bar, const oeb_s3 *fy2_bar, const oeb_s3 *q_bar,
                                  const oeb_s3 *work0, const oeb_s3 *work1, const oeb_s3 *work2, const oeb_s3 *work3,
                                  double p1, double p2, void *stream) {
  const oeb_s3 *const w[4] = {work0, work1, work2, work3};
  return oeb::fvtp2d_adj_run<double>(d, q, crx, cry, ra_x, ra_y, xfx, yfx, area, q_i, q_j, q_i_bar, q_j_bar,
                                     fx1_bar, fx2_bar, fy1_bar, fy2_bar, q_bar, w, p1, p2, stream);
}

extern "C" int oeb_fvtp2d_adj_f32(const oeb_domain *d, const oeb_s3f *q, const oeb_s3f *crx, const oeb_s3f *cry,
                                  const oeb_s3f *ra_x, const oeb_s3f *ra_y, const oeb_s3f *xfx, const oeb_s3f *yfx,
                                  const oeb_s3f *area, const oeb_s3f *q_i, const oeb_s3f *q_j,
                                  const oeb_s3f *q_i_bar, const oeb_s3f *q_j_bar, const oeb_s3f *fx1_bar,
                                  const oeb_s3f *fx2_bar, const oeb_s3f *fy1_bar, const oeb_s3f *fy2_bar,
                                  const oeb_s3f *q_bar, const oeb_s3f *work0, const oeb_s3f *work1,
                                  const oeb_s3f *work2, const oeb_s3f *work3, float p1, float p2, void *stream) {
  const oeb_s3f *const w[4] = {work0, work1, work2, work3};
  return oeb::fvtp2d_adj_run<float>(d, q, crx, cry, ra_x, ra_y, xfx, yfx, area, q_i, q_j, q_i_bar, q_j_bar, fx1_bar,
                                    fx2_bar, fy1_bar, fy2_bar, q_bar, w, p1, p2, stream);
}
