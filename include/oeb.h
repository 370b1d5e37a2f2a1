/*
 * oeb.h -- C-ABI of liboeb, the gfx950 (MI355X) implementation of the ten
 * open-earth-benchmarks climate stencils.
 *
 * One entry point per stencil.  Each replaces the C++ free function of the same
 * name in the reference's kernels/<name>.h (cited per entry below) and takes the
 * same arguments in the same order.  Three things the reference passes
 * implicitly are explicit here:
 *   - the extents (the reference reads the globals domain_size and
 *     domain_height, src/<name>.cpp:4-5), now an oeb_domain;
 *   - the driver-level constants (uvbke's dt5, src/uvbke.cpp:11-12; fvtp2d's
 *     p1/p2, src/fvtp2d.cpp:11-12), now scalar arguments;
 *   - the stream, a hipStream_t passed as void* (NULL = the null stream).
 *
 * Arrays are described by oeb_storage3d_f64 / oeb_storage1d_f64.  They have
 * the field order of the reference's Storage<double,3> / Storage<double,1>
 * (include/util.h:24-30).  Element (i,j,k) lives at
 *   alignedPtr[offset + strides[2]*i + strides[1]*j + strides[0]*k]
 * (include/util.h:84-94).  sizes[] and strides[] are in reverse order:
 * index 2 is the i (fastest) axis and index 0 is k.  Negative indices address
 * the halo.  The data pointers are DEVICE pointers.  The descriptor structs
 * themselves live in host memory.
 *
 * Conventions:
 *   - Every call validates its descriptors before launching anything.  The
 *     region the stencil touches, halo reach included, must lie inside each
 *     array's allocated box.  All 3-D arrays of one call must share
 *     strides[1] and strides[0], with strides[2] == 1.  Outputs may not overlap
 *     inputs.  Every array must lie in device, managed or pinned host memory
 *     (checked with hipPointerGetAttributes).  A failed check returns a
 *     non-zero status and launches nothing.
 *     oeb_last_error() then gives a message naming the array and the reason.
 *   - Launches are asynchronous on `stream`.  A return of OEB_OK means the
 *     kernels were enqueued.  It says nothing about completion.
 *   - Scratch arguments (the reference's temporaries: lap, flx, fly, T_sqr_s,
 *     ...) are accepted for signature compatibility.  They may be NULL.  The
 *     library never reads or writes them, because every temporary stays on
 *     chip.  Scratch contents after a call are unspecified in the reference
 *     too: its variants disagree on them.
 *   - Outputs are written only on each stencil's valid region (interior
 *     [0,nx)x[0,ny)x[0,nz), except fvtp2d's extended regions).
 *   - Calls on distinct streams are thread-safe.
 */
#ifndef OEB_H
#define OEB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define OEB_OK 0
#define OEB_ERR_NULL 1   /* a required descriptor or data pointer is NULL */
#define OEB_ERR_DOMAIN 2 /* nx, ny or nz is < 1, or too large for 32-bit in-plane indexing */
#define OEB_ERR_LAYOUT 3 /* a descriptor does not cover the region the stencil touches, or strides differ */
#define OEB_ERR_ALIAS 4  /* an output array overlaps an input array */
#define OEB_ERR_HIP 5    /* a HIP runtime call or kernel launch failed */
#define OEB_ERR_ARG 6    /* a non-array argument is out of its range, or a call made while recording a batch */
#define OEB_ERR_MEMORY 7 /* an array lies in pageable host memory, which a kernel cannot access */

typedef struct oeb_domain {
  int64_t nx; /* interior extent along i (the reference's domain_size) */
  int64_t ny; /* interior extent along j (domain_size; the reference is square) */
  int64_t nz; /* number of levels along k (domain_height) */
} oeb_domain;

typedef struct oeb_storage3d_f64 {
  double *allocatedPtr;
  double *alignedPtr;
  int64_t offset;
  int64_t sizes[3];   /* [k extent, j extent, i extent], halos included */
  int64_t strides[3]; /* [k stride, j stride, i stride == 1], in elements */
} oeb_storage3d_f64;

typedef struct oeb_storage1d_f64 {
  double *allocatedPtr;
  double *alignedPtr;
  int64_t offset;
  int64_t sizes[1];
  int64_t strides[1];
} oeb_storage1d_f64;

/* FP32 descriptors: the same fields with float data pointers (the reference's
 * Storage<float,3> / Storage<float,1> when ElementType is float). */
typedef struct oeb_storage3d_f32 {
  float *allocatedPtr;
  float *alignedPtr;
  int64_t offset;
  int64_t sizes[3];
  int64_t strides[3];
} oeb_storage3d_f32;

typedef struct oeb_storage1d_f32 {
  float *allocatedPtr;
  float *alignedPtr;
  int64_t offset;
  int64_t sizes[1];
  int64_t strides[1];
} oeb_storage1d_f32;

typedef oeb_storage3d_f64 oeb_s3;
typedef oeb_storage1d_f64 oeb_s1;
typedef oeb_storage3d_f32 oeb_s3f;
typedef oeb_storage1d_f32 oeb_s1f;

/* Library version as "major.minor.patch". */
const char *oeb_version(void);
/* Short, static description of a status code. */
const char *oeb_status_string(int status);
/* Message of the last failing call on this host thread ("" if none). */
const char *oeb_last_error(void);

/* laplace -- kernels/laplace.h:4 `laplace(const Storage3D& in, Storage3D& out)`;
 * arithmetic of laplace_openmp, kernels/laplace.h:34-43.  Reach (1,1,0). */
int oeb_laplace_f64(const oeb_domain *d, const oeb_s3 *in, const oeb_s3 *out, void *stream);

/* hdiffsa -- kernels/hdiffsa.h:4-5 `hdiffsa(out, in, mask, crlato, crlatu, lap, flx, fly)`;
 * arithmetic of hdiffsa_openmp, kernels/hdiffsa.h:95-123.  Reach (2,2,0).
 * lap, flx, fly are scratch (may be NULL). */
int oeb_hdiffsa_f64(const oeb_domain *d, const oeb_s3 *out, const oeb_s3 *in, const oeb_s3 *mask,
                    const oeb_s1 *crlato, const oeb_s1 *crlatu, const oeb_s3 *lap, const oeb_s3 *flx,
                    const oeb_s3 *fly, void *stream);

/* hdiffsmag -- kernels/hdiffsmag.h:4-7 `hdiffsmag(uout, vout, uin, vin, mask, crlavo, crlavu, crlato,
 * crlatu, acrlat0, T_sqr_s, S_sqr_uv, eddlat, eddlon, tau_smag, weight_smag)`;
 * arithmetic of hdiffsmag_openmp, kernels/hdiffsmag.h:205-263.  Reach (1,1,0).
 * T_sqr_s, S_sqr_uv are scratch (may be NULL). */
int oeb_hdiffsmag_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s3 *mask, const oeb_s1 *crlavo, const oeb_s1 *crlavu,
                      const oeb_s1 *crlato, const oeb_s1 *crlatu, const oeb_s1 *acrlat0, const oeb_s3 *T_sqr_s,
                      const oeb_s3 *S_sqr_uv, double eddlat, double eddlon, double tau_smag, double weight_smag,
                      void *stream);

/* hadvuv -- kernels/hadvuv.h:29-32 `hadvuv(uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0,
 * tgrlatda1, uatupos, vatupos, uatvpos, vatvpos, uavg, vavg, ures, vres, eddlat, eddlon)`;
 * arithmetic of hadvuv_openmp, kernels/hadvuv.h:178-210 (3rd-order upwind,
 * advectionDriver kernels/hadvuv.h:4-27).  Reach (2,2,0).  The eight arrays
 * uatupos..vres are scratch (may be NULL). */
int oeb_hadvuv_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                   const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0,
                   const oeb_s1 *tgrlatda1, const oeb_s3 *uatupos, const oeb_s3 *vatupos, const oeb_s3 *uatvpos,
                   const oeb_s3 *vatvpos, const oeb_s3 *uavg, const oeb_s3 *vavg, const oeb_s3 *ures,
                   const oeb_s3 *vres, double eddlat, double eddlon, void *stream);

/* hadvuv5th -- kernels/hadvuv5th.h:41-44, same parameter list as hadvuv;
 * arithmetic of hadvuv5th_openmp, kernels/hadvuv5th.h:177-201 (5th-order
 * upwind, advectionDriver kernels/hadvuv5th.h:4-39).  Reach (3,3,0). */
int oeb_hadvuv5th_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0,
                      const oeb_s1 *tgrlatda1, const oeb_s3 *uatupos, const oeb_s3 *vatupos,
                      const oeb_s3 *uatvpos, const oeb_s3 *vatvpos, const oeb_s3 *uavg, const oeb_s3 *vavg,
                      const oeb_s3 *ures, const oeb_s3 *vres, double eddlat, double eddlon, void *stream);

/* fastwavesuv -- kernels/fastwavesuv.h:4-7 `fastwavesuv(uout, vout, uin, vin, utens, vtens, wgtfac,
 * ppuv, hhl, rho, fx, ppgk, ppgc, ppgu, ppgv, edadlat, dt)`;
 * arithmetic of fastwavesuv_openmp, kernels/fastwavesuv.h:154-191.
 * Reach i,j in [0,+1], k in [-1,+1]: the k-halo planes -1 and nz are read.
 * ppgk, ppgc, ppgu, ppgv are scratch (may be NULL). */
int oeb_fastwavesuv_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                        const oeb_s3 *vin, const oeb_s3 *utens, const oeb_s3 *vtens, const oeb_s3 *wgtfac,
                        const oeb_s3 *ppuv, const oeb_s3 *hhl, const oeb_s3 *rho, const oeb_s1 *fx,
                        const oeb_s3 *ppgk, const oeb_s3 *ppgc, const oeb_s3 *ppgu, const oeb_s3 *ppgv,
                        double edadlat, double dt, void *stream);

/* uvbke -- kernels/uvbke.h:4-5 `uvbke(ub, vb, uc, vc, cosa, rsina)`; arithmetic of
 * uvbke_openmp, kernels/uvbke.h:61-75.  dt5 is the driver constant
 * (src/uvbke.cpp:11-12: 0.5 * 225.0).  Reach i,j in [-1,0]. */
int oeb_uvbke_f64(const oeb_domain *d, const oeb_s3 *ub, const oeb_s3 *vb, const oeb_s3 *uc, const oeb_s3 *vc,
                  const oeb_s3 *cosa, const oeb_s3 *rsina, double dt5, void *stream);

/* p_grad_c -- kernels/p_grad_c.h:4-6 `p_grad_c(uout, vout, uin, vin, rdxc, rdyc, delpc, gz, pkc, wk, dt2)`;
 * arithmetic of p_grad_c_openmp, kernels/p_grad_c.h:80-101 (reads delpc(i-1)
 * directly, so there is no uninitialised-scratch ring).  Reach i,j in [-1,0],
 * k in [0,+1].  wk is scratch (may be NULL). */
int oeb_p_grad_c_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                     const oeb_s3 *vin, const oeb_s3 *rdxc, const oeb_s3 *rdyc, const oeb_s3 *delpc,
                     const oeb_s3 *gz, const oeb_s3 *pkc, const oeb_s3 *wk, double dt2, void *stream);

/* nh_p_grad -- kernels/nh_p_grad.h:4-6 `nh_p_grad(uout, vout, uin, vin, rdx, rdy, gz, pp, pk3, wk1, wk,
 * du, dv, dt)`; arithmetic of nh_p_grad_openmp, kernels/nh_p_grad.h:116-146.
 * Reach i,j in [0,+1], k in [0,+1].  wk, du, dv are scratch (may be NULL). */
int oeb_nh_p_grad_f64(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s3 *rdx, const oeb_s3 *rdy, const oeb_s3 *gz, const oeb_s3 *pp,
                      const oeb_s3 *pk3, const oeb_s3 *wk1, const oeb_s3 *wk, const oeb_s3 *du, const oeb_s3 *dv,
                      double dt, void *stream);

/* fvtp2d -- kernels/fvtp2d.h:4-7 `fvtp2d(q_i, q_j, fx1, fx2, fy1, fy2, q, crx, cry, ra_x, ra_y, xfx, yfx,
 * area, fxx, fyy, al, almq, br, b0, smt5)`; arithmetic of the DEFAULT variant,
 * kernels/fvtp2d.h:4-177 (the reference has no OpenMP variant, and its
 * fullfusion variant reads an unwritten array).  p1, p2 are the PPM
 * constants, which the reference driver defines as 0.583333 and 0.083333
 * (src/fvtp2d.cpp:11-12).  Reach (3,3,0).  Output regions (half-open):
 *   fy2  i in [-3,nx+2), j in [0,ny+1)     q_i  i in [-3,nx+2), j in [0,ny)
 *   fx2  i in [0,nx+1),  j in [-3,ny+2)    q_j  i in [0,nx),    j in [-3,ny+2)
 *   fx1, fy1  [0,nx) x [0,ny)
 * fxx, fyy, al, almq, br, b0, smt5 are scratch (may be NULL). */
/* The reference drivers define p1 and p2 as macros before including any
 * header (src/fvtp2d.cpp:11-12); keep them from rewriting this prototype. */
#pragma push_macro("p1")
#pragma push_macro("p2")
#undef p1
#undef p2
int oeb_fvtp2d_f64(const oeb_domain *d, const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *fx1,
                   const oeb_s3 *fx2, const oeb_s3 *fy1, const oeb_s3 *fy2, const oeb_s3 *q, const oeb_s3 *crx,
                   const oeb_s3 *cry, const oeb_s3 *ra_x, const oeb_s3 *ra_y, const oeb_s3 *xfx, const oeb_s3 *yfx,
                   const oeb_s3 *area, const oeb_s3 *fxx, const oeb_s3 *fyy, const oeb_s3 *al, const oeb_s3 *almq,
                   const oeb_s3 *br, const oeb_s3 *b0, const oeb_s3 *smt5, double p1, double p2, void *stream);

/* ------------------------------------------------------------------- FP32
 * oeb_<name>_f32 runs the same stencil with FP32 storage and arithmetic: the
 * reference compiled with `typedef float ElementType` (src/<name>.cpp:9).
 * Arguments come in the same order as in the FP64 entry point.  Arrays are
 * oeb_storage3d_f32 / oeb_storage1d_f32, and scalars are float, the
 * reference's `const ElementType`.  uvbke's dt5 stays double: the reference
 * declares it `const double` whatever ElementType is (src/uvbke.cpp:11-12).
 * As in the reference's float build, laplace's -4.0 literal and uvbke's dt5
 * are double, so those two products are formed in double and rounded once
 * on the store (kernels/laplace.h:39, kernels/uvbke.h:67-71).
 *
 * laplace, hdiffsa, hdiffsmag, hadvuv, hadvuv5th and uvbke compute exactly
 * what the reference's float build computes.  Three stencils keep FP64
 * islands, because FP32 storage of pressure (~1e5 Pa) and geopotential
 * (~2e5 m^2/s^2) loses their small differences to cancellation:
 *   p_grad_c     gz and pkc are FP64 arrays (oeb_storage3d_f64).  The
 *                bracket (dgz*dpk + dgz'*dpk') is formed in FP64 and
 *                rounded once to float.
 *   nh_p_grad    gz, pp and pk3 are FP64 arrays.  wk = pk3(k+1) - pk3(k),
 *                the pressure-thickness sums, du/dv and the pp bracket are
 *                formed in FP64.  du/dv and the bracket are rounded once to
 *                float.
 *   fastwavesuv  ppuv and hhl are FP64 arrays.  ppgk, ppgc and ppgu/ppgv
 *                are formed in FP64.  ppgu/ppgv are rounded once to float.
 * All 3-D arrays of one call, FP32 and FP64 alike, share strides counted
 * in elements.  fvtp2d forms its flux divergences q*area + fyy - fyy(j+1)
 * and q*area + fxx - fxx(i+1) with error-free transformations (the product
 * by fma, the sums by TwoSum) and adds the collected rounding errors once
 * before the division.  The numerator then carries about one float rounding
 * instead of three.  Its PPM stages are those of the float build. */
int oeb_laplace_f32(const oeb_domain *d, const oeb_s3f *in, const oeb_s3f *out, void *stream);
int oeb_hdiffsa_f32(const oeb_domain *d, const oeb_s3f *out, const oeb_s3f *in, const oeb_s3f *mask,
                    const oeb_s1f *crlato, const oeb_s1f *crlatu, const oeb_s3f *lap, const oeb_s3f *flx,
                    const oeb_s3f *fly, void *stream);
int oeb_hdiffsmag_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                      const oeb_s3f *vin, const oeb_s3f *mask, const oeb_s1f *crlavo, const oeb_s1f *crlavu,
                      const oeb_s1f *crlato, const oeb_s1f *crlatu, const oeb_s1f *acrlat0,
                      const oeb_s3f *T_sqr_s, const oeb_s3f *S_sqr_uv, float eddlat, float eddlon, float tau_smag,
                      float weight_smag, void *stream);
int oeb_hadvuv_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                   const oeb_s3f *vin, const oeb_s1f *acrlat0, const oeb_s1f *acrlat1, const oeb_s1f *tgrlatda0,
                   const oeb_s1f *tgrlatda1, const oeb_s3f *uatupos, const oeb_s3f *vatupos,
                   const oeb_s3f *uatvpos, const oeb_s3f *vatvpos, const oeb_s3f *uavg, const oeb_s3f *vavg,
                   const oeb_s3f *ures, const oeb_s3f *vres, float eddlat, float eddlon, void *stream);
int oeb_hadvuv5th_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                      const oeb_s3f *vin, const oeb_s1f *acrlat0, const oeb_s1f *acrlat1,
                      const oeb_s1f *tgrlatda0, const oeb_s1f *tgrlatda1, const oeb_s3f *uatupos,
                      const oeb_s3f *vatupos, const oeb_s3f *uatvpos, const oeb_s3f *vatvpos, const oeb_s3f *uavg,
                      const oeb_s3f *vavg, const oeb_s3f *ures, const oeb_s3f *vres, float eddlat, float eddlon,
                      void *stream);
int oeb_fastwavesuv_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                        const oeb_s3f *vin, const oeb_s3f *utens, const oeb_s3f *vtens, const oeb_s3f *wgtfac,
                        const oeb_s3 *ppuv, const oeb_s3 *hhl, const oeb_s3f *rho, const oeb_s1f *fx,
                        const oeb_s3f *ppgk, const oeb_s3f *ppgc, const oeb_s3f *ppgu, const oeb_s3f *ppgv,
                        float edadlat, float dt, void *stream);
int oeb_uvbke_f32(const oeb_domain *d, const oeb_s3f *ub, const oeb_s3f *vb, const oeb_s3f *uc,
                  const oeb_s3f *vc, const oeb_s3f *cosa, const oeb_s3f *rsina, double dt5, void *stream);
int oeb_p_grad_c_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                     const oeb_s3f *vin, const oeb_s3f *rdxc, const oeb_s3f *rdyc, const oeb_s3f *delpc,
                     const oeb_s3 *gz, const oeb_s3 *pkc, const oeb_s3f *wk, float dt2, void *stream);
int oeb_nh_p_grad_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                      const oeb_s3f *vin, const oeb_s3f *rdx, const oeb_s3f *rdy, const oeb_s3 *gz,
                      const oeb_s3 *pp, const oeb_s3 *pk3, const oeb_s3f *wk1, const oeb_s3f *wk,
                      const oeb_s3f *du, const oeb_s3f *dv, float dt, void *stream);
int oeb_fvtp2d_f32(const oeb_domain *d, const oeb_s3f *q_i, const oeb_s3f *q_j, const oeb_s3f *fx1,
                   const oeb_s3f *fx2, const oeb_s3f *fy1, const oeb_s3f *fy2, const oeb_s3f *q,
                   const oeb_s3f *crx, const oeb_s3f *cry, const oeb_s3f *ra_x, const oeb_s3f *ra_y,
                   const oeb_s3f *xfx, const oeb_s3f *yfx, const oeb_s3f *area, const oeb_s3f *fxx,
                   const oeb_s3f *fyy, const oeb_s3f *al, const oeb_s3f *almq, const oeb_s3f *br,
                   const oeb_s3f *b0, const oeb_s3f *smt5, float p1, float p2, void *stream);

/* ------------------------------------------- FP32 with deviation islands
 * No counterpart in the reference.  oeb_<name>_f32d is oeb_<name>_f32 with
 * each FP64 island array replaced by two arguments: an FP32 array of
 * deviations (oeb_storage3d_f32, same box as the island) followed by an FP64
 * per-level base profile (oeb_storage1d_f64 indexed by k, covering the
 * island's levels: [-1, nz + 1) for fastwavesuv's ppuv, [0, nz + 1) for the
 * others).  The kernel reads the island value
 * base(k) + dev(i,j,k), the sum formed in FP64, and continues exactly as
 * oeb_<name>_f32 does on an FP64 island holding those values.  An island
 * moves 4 bytes per point instead of 8: p_grad_c 36 B per point instead of
 * 44, nh_p_grad and fastwavesuv 40 instead of 52 and 48.  Storage rounding now applies to the
 * deviation only: for a base near the level's mean pressure or geopotential
 * the deviation is some 1e-2 of the value, so the stored field carries
 * about 1e-9 relative error instead of 6e-8.  The base profile may be any
 * per-level values (0 gives the FP32 array itself). */
int oeb_fastwavesuv_f32d(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                         const oeb_s3f *vin, const oeb_s3f *utens, const oeb_s3f *vtens, const oeb_s3f *wgtfac,
                         const oeb_s3f *ppuv, const oeb_s1 *ppuv0, const oeb_s3f *hhl, const oeb_s1 *hhl0,
                         const oeb_s3f *rho, const oeb_s1f *fx, const oeb_s3f *ppgk, const oeb_s3f *ppgc,
                         const oeb_s3f *ppgu, const oeb_s3f *ppgv, float edadlat, float dt, void *stream);
int oeb_p_grad_c_f32d(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                      const oeb_s3f *vin, const oeb_s3f *rdxc, const oeb_s3f *rdyc, const oeb_s3f *delpc,
                      const oeb_s3f *gz, const oeb_s1 *gz0, const oeb_s3f *pkc, const oeb_s1 *pkc0,
                      const oeb_s3f *wk, float dt2, void *stream);
int oeb_nh_p_grad_f32d(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                       const oeb_s3f *vin, const oeb_s3f *rdx, const oeb_s3f *rdy, const oeb_s3f *gz,
                       const oeb_s1 *gz0, const oeb_s3f *pp, const oeb_s1 *pp0, const oeb_s3f *pk3,
                       const oeb_s1 *pk30, const oeb_s3f *wk1, const oeb_s3f *wk, const oeb_s3f *du,
                       const oeb_s3f *dv, float dt, void *stream);
#pragma pop_macro("p2")
#pragma pop_macro("p1")

/* ------------------------------------------------------------ halo update
 * No counterpart in the reference: its kernels write only their output
 * regions and only the drivers' fillMath (include/util.h:155-163) writes
 * halos.  A chain of stencils, whose outputs feed the next stencil's inputs,
 * refills the halo the next stencil reads with these.  One launch on
 * `stream` writes
 *   - the i/j ring [-halo_ij, nx+halo_ij) x [-halo_ij, ny+halo_ij) minus the
 *     interior, on levels [0, nz), and
 *   - the k-halo planes [-halo_k, 0) and [nz, nz+halo_k) over the whole i/j
 *     box.
 * Each of these points takes the value of the interior point at the wrapped
 * (OEB_HALO_PERIODIC: i mod nx, j mod ny) or clamped (OEB_HALO_ZERO_GRADIENT)
 * i/j index on the clamped level (so the k-halo planes are copies of planes 0
 * and nz-1, rings included, in both modes).
 * halo_ij and halo_k lie in [0, 16]; the descriptor must cover the box.
 * oeb_halo_fill_n refills 1..8 fields of one layout in one launch (what a
 * chain does before a stencil that reads several refilled fields); the
 * fields' boxes must not overlap.  Same result as one oeb_halo_fill per
 * field. */
#define OEB_HALO_PERIODIC 0
#define OEB_HALO_ZERO_GRADIENT 1
int oeb_halo_fill_f64(const oeb_domain *d, const oeb_s3 *field, int halo_ij, int halo_k, int mode, void *stream);
int oeb_halo_fill_f32(const oeb_domain *d, const oeb_s3f *field, int halo_ij, int halo_k, int mode, void *stream);
int oeb_halo_fill_n_f64(const oeb_domain *d, const oeb_s3 *const *fields, int nfields, int halo_ij, int halo_k,
                        int mode, void *stream);
int oeb_halo_fill_n_f32(const oeb_domain *d, const oeb_s3f *const *fields, int nfields, int halo_ij, int halo_k,
                        int mode, void *stream);

/* ----------------------------------------------------------------- graphs
 * No counterpart in the reference.  Records every launch the oeb_* entry
 * points enqueue on `stream` between oeb_graph_begin and oeb_graph_end as a
 * HIP graph (stream capture, thread-local mode), and replays it with one
 * oeb_graph_launch.  The descriptors are validated when the calls are
 * recorded; a replay runs the same kernels on the same device addresses.
 * `stream` must be a created stream, not the null stream.  The entry points
 * enqueue kernels only (no memset, copy or allocation), so a recorded step
 * replays exactly. */
typedef struct oeb_graph oeb_graph;
int oeb_graph_begin(void *stream);
int oeb_graph_end(void *stream, oeb_graph **graph);
int oeb_graph_launch(const oeb_graph *graph, void *stream);
int oeb_graph_destroy(oeb_graph *graph);

/* --------------------------------------------------------------- adjoints
 * No counterpart in the reference.  oeb_<name>_adj_<f64|f32> computes the
 * vector-Jacobian product of a stencil: given the forward inputs and the
 * cotangents of its outputs (*_bar), it writes the cotangents of the
 * differentiated inputs on the whole box the forward reads them on (halo
 * included; other points are left untouched).  Coefficient fields are
 * constants.  Limiter, upwind, clamp and PPM branches are frozen at the
 * forward pass's choice.  Scratch arrays (same layout as the inputs) carry
 * intermediate cotangents between the kernels of one call.
 *   hdiffsa      d/d in on [-2, nx+2) x [-2, ny+2); scratch lap on
 *                [-1, nx+1) x [-1, ny+1)
 *   fastwavesuv  d/d uin, vin on the interior; d/d ppuv on
 *                [0, nx+1) x [0, ny+1) x [-1, nz+1) (FP64 in both forms)
 *   hadvuv5th    d/d uin, vin on [-3, nx+3) x [-3, ny+3); scratch: the eight
 *                reference temporaries, on the interior
 *   hdiffsmag    d/d uin, vin on [-1, nx+1) x [-1, ny+1); scratch work0..3 on
 *                the interior
 *   fvtp2d       d/d q on [-3, nx+3) x [-3, ny+3), from the cotangents of all
 *                six outputs; q_i and q_j are the forward's outputs; scratch
 *                work0..3 cover the regions of q_j, q_i, fx2 and fy2 */
int oeb_hdiffsa_adj_f64(const oeb_domain *d, const oeb_s3 *in, const oeb_s3 *mask, const oeb_s1 *crlato,
                        const oeb_s1 *crlatu, const oeb_s3 *out_bar, const oeb_s3 *in_bar, const oeb_s3 *lap,
                        void *stream);
int oeb_hdiffsa_adj_f32(const oeb_domain *d, const oeb_s3f *in, const oeb_s3f *mask, const oeb_s1f *crlato,
                        const oeb_s1f *crlatu, const oeb_s3f *out_bar, const oeb_s3f *in_bar, const oeb_s3f *lap,
                        void *stream);
int oeb_fastwavesuv_adj_f64(const oeb_domain *d, const oeb_s3 *wgtfac, const oeb_s3 *ppuv, const oeb_s3 *hhl,
                            const oeb_s3 *rho, const oeb_s1 *fx, const oeb_s3 *uout_bar, const oeb_s3 *vout_bar,
                            const oeb_s3 *uin_bar, const oeb_s3 *vin_bar, const oeb_s3 *ppuv_bar, double edadlat,
                            double dt, void *stream);
int oeb_fastwavesuv_adj_f32(const oeb_domain *d, const oeb_s3f *wgtfac, const oeb_s3 *ppuv, const oeb_s3 *hhl,
                            const oeb_s3f *rho, const oeb_s1f *fx, const oeb_s3f *uout_bar, const oeb_s3f *vout_bar,
                            const oeb_s3f *uin_bar, const oeb_s3f *vin_bar, const oeb_s3 *ppuv_bar, float edadlat,
                            float dt, void *stream);
int oeb_hadvuv5th_adj_f64(const oeb_domain *d, const oeb_s3 *uin, const oeb_s3 *vin, const oeb_s1 *acrlat0,
                          const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0, const oeb_s1 *tgrlatda1,
                          const oeb_s3 *uout_bar, const oeb_s3 *vout_bar, const oeb_s3 *uin_bar,
                          const oeb_s3 *vin_bar, const oeb_s3 *uatupos, const oeb_s3 *vatupos, const oeb_s3 *uatvpos,
                          const oeb_s3 *vatvpos, const oeb_s3 *uavg, const oeb_s3 *vavg, const oeb_s3 *ures,
                          const oeb_s3 *vres, double eddlat, double eddlon, void *stream);
int oeb_hadvuv5th_adj_f32(const oeb_domain *d, const oeb_s3f *uin, const oeb_s3f *vin, const oeb_s1f *acrlat0,
                          const oeb_s1f *acrlat1, const oeb_s1f *tgrlatda0, const oeb_s1f *tgrlatda1,
                          const oeb_s3f *uout_bar, const oeb_s3f *vout_bar, const oeb_s3f *uin_bar,
                          const oeb_s3f *vin_bar, const oeb_s3f *uatupos, const oeb_s3f *vatupos,
                          const oeb_s3f *uatvpos, const oeb_s3f *vatvpos, const oeb_s3f *uavg, const oeb_s3f *vavg,
                          const oeb_s3f *ures, const oeb_s3f *vres, float eddlat, float eddlon, void *stream);

int oeb_hdiffsmag_adj_f64(const oeb_domain *d, const oeb_s3 *uin, const oeb_s3 *vin, const oeb_s3 *mask,
                          const oeb_s1 *crlavo, const oeb_s1 *crlavu, const oeb_s1 *crlato, const oeb_s1 *crlatu,
                          const oeb_s1 *acrlat0, const oeb_s3 *uout_bar, const oeb_s3 *vout_bar,
                          const oeb_s3 *uin_bar, const oeb_s3 *vin_bar, const oeb_s3 *work0, const oeb_s3 *work1,
                          const oeb_s3 *work2, const oeb_s3 *work3, double eddlat, double eddlon, double tau_smag,
                          double weight_smag, void *stream);
int oeb_hdiffsmag_adj_f32(const oeb_domain *d, const oeb_s3f *uin, const oeb_s3f *vin, const oeb_s3f *mask,
                          const oeb_s1f *crlavo, const oeb_s1f *crlavu, const oeb_s1f *crlato, const oeb_s1f *crlatu,
                          const oeb_s1f *acrlat0, const oeb_s3f *uout_bar, const oeb_s3f *vout_bar,
                          const oeb_s3f *uin_bar, const oeb_s3f *vin_bar, const oeb_s3f *work0,
                          const oeb_s3f *work1, const oeb_s3f *work2, const oeb_s3f *work3, float eddlat,
                          float eddlon, float tau_smag, float weight_smag, void *stream);
#pragma push_macro("p1")
#pragma push_macro("p2")
#undef p1
#undef p2
int oeb_fvtp2d_adj_f64(const oeb_domain *d, const oeb_s3 *q, const oeb_s3 *crx, const oeb_s3 *cry,
                       const oeb_s3 *ra_x, const oeb_s3 *ra_y, const oeb_s3 *xfx, const oeb_s3 *yfx,
                       const oeb_s3 *area, const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *q_i_bar,
                       const oeb_s3 *q_j_bar, const oeb_s3 *fx1_bar, const oeb_s3 *fx2_bar, const oeb_s3 *fy1_bar,
                       const oeb_s3 *fy2_bar, const oeb_s3 *q_bar, const oeb_s3 *work0, const oeb_s3 *work1,
                       const oeb_s3 *work2, const oeb_s3 *work3, double p1, double p2, void *stream);
int oeb_fvtp2d_adj_f32(const oeb_domain *d, const oeb_s3f *q, const oeb_s3f *crx, const oeb_s3f *cry,
                       const oeb_s3f *ra_x, const oeb_s3f *ra_y, const oeb_s3f *xfx, const oeb_s3f *yfx,
                       const oeb_s3f *area, const oeb_s3f *q_i, const oeb_s3f *q_j, const oeb_s3f *q_i_bar,
                       const oeb_s3f *q_j_bar, const oeb_s3f *fx1_bar, const oeb_s3f *fx2_bar,
                       const oeb_s3f *fy1_bar, const oeb_s3f *fy2_bar, const oeb_s3f *q_bar, const oeb_s3f *work0,
                       const oeb_s3f *work1, const oeb_s3f *work2, const oeb_s3f *work3, float p1, float p2,
                       void *stream);
#pragma pop_macro("p2")
#pragma pop_macro("p1")

/* oeb_fvtp2d_adj_f64_noalloc -- no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): oeb_fvtp2d_adj_f64 without
 * the four work arrays, same arguments otherwise, same q_bar bit for bit.  It
 * calls no allocating routine.  Its two work planes (the quotients the tiled
 * adjoint forms on the regions of q_j and q_i) live in a 32 MiB device array
 * inside the library, one per GPU, held until the process ends; levels are
 * processed in chunks that fit it, one after another on `stream`.  A domain
 * whose single level needs more than 16 MiB per plane (about 2.1 million
 * elements of (ny + 5) rows at the j stride) fails with OEB_ERR_DOMAIN.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (on two
 * streams, or from two host threads) race and give wrong results. */
#pragma push_macro("p1")
#pragma push_macro("p2")
#undef p1
#undef p2
int oeb_fvtp2d_adj_f64_noalloc(const oeb_domain *d, const oeb_s3 *q, const oeb_s3 *crx, const oeb_s3 *cry,
                               const oeb_s3 *ra_x, const oeb_s3 *ra_y, const oeb_s3 *xfx, const oeb_s3 *yfx,
                               const oeb_s3 *area, const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *q_i_bar,
                               const oeb_s3 *q_j_bar, const oeb_s3 *fx1_bar, const oeb_s3 *fx2_bar,
                               const oeb_s3 *fy1_bar, const oeb_s3 *fy2_bar, const oeb_s3 *q_bar, double p1,
                               double p2, void *stream);
#pragma pop_macro("p2")
#pragma pop_macro("p1")

/* ---------------------------------------------------------------- tuning
 * No counterpart in the reference: its drivers fix every loop order at
 * compile time.  Each entry point has launch parameters ("knobs") with
 * built-in defaults, which a process-wide table may override per kernel name
 * ("hadvuv5th", "fvtp2d", "hdiffsmag_adj", ...), element size and domain.
 * The project's autotuner (oeb.tune) times the candidates and fills the
 * table.  Knobs (a kernel ignores those it does not have):
 *   "kchunk" [1, 4096]  levels per workgroup of a tile launch
 *   "vector" [0, 1]     1: the 16-byte vector / column-pair body
 *   "form"   [0, 2]     1: the row-march (strip) form, 0: the tile form,
 *                       2: LDS-staged tiles (fvtp2d)
 *   "rows"   [1, 4096]  rows per strip chunk
 *   "depth"  [1, 8]     rows a strip loads ahead
 *   "order"  [0, 3]     tile order: 0 rows in order, 1 level chunks
 *                       interleaved between tile rows, 2 the tiles of a row
 *                       on one XCD, 3 whole level chunks per XCD
 *   "tile"   [0, 12]    tile shape (kernel-specific index)
 *   "lanes"  [64, 1024] workgroup size
 *   "grid"   [1, 2^30]  workgroups of a batched launch (kernel "batch")
 * oeb_tune_knob_set stores `value` for (kernel, element_bytes, d); value -1
 * removes it.  element_bytes 0 stands for both precisions and d == NULL for
 * every domain.  A call uses the most specific entry that sets the knob: its
 * domain and element size, then every domain, then both precisions.  A value
 * a kernel's launch cannot use falls back to that kernel's default.
 * oeb_tune_knob_get reports the entry for exactly that key (-1 if unset).
 * oeb_tune_set / oeb_tune_get handle "kchunk" (0 = unset) and "vector" (-1 =
 * unset) together.  Thread-safe. */
int oeb_tune_knob_set(const char *kernel, int element_bytes, const oeb_domain *d, const char *knob, int value);
int oeb_tune_knob_get(const char *kernel, int element_bytes, const oeb_domain *d, const char *knob, int *value);
int oeb_tune_set(const char *kernel, int element_bytes, const oeb_domain *d, int kchunk, int vector);
int oeb_tune_get(const char *kernel, int element_bytes, const oeb_domain *d, int *kchunk, int *vector);
int oeb_tune_clear(void);

/* --------------------------------------------------------------- batches
 * No counterpart in the reference (its drivers run one field at a time).
 * Runs one stencil over many members of an ensemble in ONE launch.
 *   oeb_batch_record(b)  from now on, on this host thread, the batchable
 *                        entry points (hadvuv, hadvuv5th, hdiffsmag,
 *                        fastwavesuv, uvbke and oeb_halo_fill(_n), both
 *                        precisions) validate their arguments as usual and
 *                        record their launch into b instead of launching;
 *                        the stream argument is ignored.  The other stencil
 *                        and adjoint entry points fail with OEB_ERR_ARG and
 *                        launch nothing while a batch records.
 *   oeb_batch_stop(b)    ends recording and uploads the member tables (the
 *                        only allocation: device memory, freed by destroy).
 *   oeb_batch_launch     enqueues one launch per run of consecutive records
 *                        of one kernel, on `stream`.  Kernels only: it can
 *                        be recorded in an oeb_graph.
 * The records of a run execute concurrently.  A record that reads what an
 * earlier record of its run writes, or writes what one reads or writes,
 * starts a new run, so the launches keep the recorded order of dependent
 * calls.  Record member after member within each stencil (all members' halo
 * fills, then all members' first stencil, ...) for the fewest launches.
 * oeb_batch_info reports the number of records and of launches. */
typedef struct oeb_batch oeb_batch;
int oeb_batch_create(oeb_batch **batch);
int oeb_batch_record(oeb_batch *batch);
int oeb_batch_stop(oeb_batch *batch);
int oeb_batch_launch(const oeb_batch *batch, void *stream);
int oeb_batch_info(const oeb_batch *batch, int *records, int *launches);
int oeb_batch_destroy(oeb_batch *batch);

/* ---------------------------------------------------------- measurement
 * No counterpart in the reference.  oeb_stream_copy copies `bytes` from src
 * to dst (device memory, non-overlapping; pointers and size multiples of 16)
 * with 16-byte non-temporal loads and stores, asynchronously on `stream`:
 * the practical ceiling of HBM bandwidth that the benchmark reports beside
 * the stencils. */
#include <stddef.h>
int oeb_stream_copy(void *dst, const void *src, size_t bytes, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* OEB_H */
