/*
 * oracle.c -- CPU restatement of the ten open-earth-benchmarks stencils.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of it.
 * Built by oracle/Makefile with -O2 -ffp-contract=off.
 *
 * Every expression keeps the reference's operand order and parenthesization,
 * so that with contraction off the result is determined bit for bit.  C
 * evaluates a + b - c + d as ((a + b) - c) + d, as C++ does.  The
 * semantics are those of the *_openmp variants (fvtp2d: DEFAULT).  They differ
 * from DEFAULT only where DEFAULT reads unwritten scratch.  Constants written
 * as a quotient in the reference (for example ElementType(-1.0) /
 * ElementType(6.0)) are formed the same way here; both compilers fold them to
 * the correctly rounded double.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Element (i,j,k) of a 3-D descriptor, as Storage::operator() in
 * include/util.h:84-94 of the reference. */
#define AT(s, i, j, k)                                                                             \
  ((s)->alignedPtr[(s)->offset + (int64_t)(i) * (s)->strides[2] + (int64_t)(j) * (s)->strides[1] + \
                   (int64_t)(k) * (s)->strides[0]])
/* Element j of a 1-D descriptor (include/util.h:56-64). */
#define AT1(s, j) ((s)->alignedPtr[(s)->offset + (int64_t)(j) * (s)->strides[0]])

/* EARTH_RADIUS and EARTH_RADIUS_RECIP, include/util.h:17-18. */
#define EARTH_R 6371.229e3
#define EARTH_R_RECIP (1.0 / EARTH_R)

#define FOR_INTERIOR(d)                          \
  for (int64_t k = 0; k < (d)->nz; ++k)          \
    for (int64_t j = 0; j < (d)->ny; ++j)        \
      for (int64_t i = 0; i < (d)->nx; ++i)

/* ------------------------------------------------------------------ laplace */
/* kernels/laplace.h:34-43 (expression :39). */
void oracle_laplace(const oeb_domain *d, const oeb_s3 *in, const oeb_s3 *out) {
  FOR_INTERIOR(d) {
    AT(out, i, j, k) = -4.0 * AT(in, i, j, k) +
                       ((AT(in, i - 1, j, k) + AT(in, i + 1, j, k)) + (AT(in, i, j + 1, k) + AT(in, i, j - 1, k)));
  }
}

/* ------------------------------------------------------------------ hdiffsa */
/* The Laplacian at (p,q) with the row's crlato/crlatu: kernels/hdiffsa.h:147-149
 * (the three _lap_* lines are this one pattern at (i,j), (i+1,j), (i,j+1)). */
static double hd_lap(const oeb_s3 *in, const oeb_s1 *lo, const oeb_s1 *lu, int64_t p, int64_t q, int64_t k) {
  const double c = AT(in, p, q, k);
  return AT(in, p - 1, q, k) + AT(in, p + 1, q, k) - 2.0 * c + AT1(lo, q) * (AT(in, p, q + 1, k) - c) +
         AT1(lu, q) * (AT(in, p, q - 1, k) - c);
}

/* Limited x flux between (p,q) and (p+1,q): kernels/hdiffsa.h:152-155. */
static double hd_flx(const oeb_s3 *in, const oeb_s1 *lo, const oeb_s1 *lu, int64_t p, int64_t q, int64_t k) {
  double f = hd_lap(in, lo, lu, p + 1, q, k) - hd_lap(in, lo, lu, p, q, k);
  if (f * (AT(in, p + 1, q, k) - AT(in, p, q, k)) > 0) f = 0;
  return f;
}

/* Limited y flux between (p,q) and (p,q+1): kernels/hdiffsa.h:156-159. */
static double hd_fly(const oeb_s3 *in, const oeb_s1 *lo, const oeb_s1 *lu, int64_t p, int64_t q, int64_t k) {
  double f = AT1(lo, q) * (hd_lap(in, lo, lu, p, q + 1, k) - hd_lap(in, lo, lu, p, q, k));
  if (f * (AT(in, p, q + 1, k) - AT(in, p, q, k)) > 0) f = 0;
  return f;
}

/* kernels/hdiffsa.h:141-169 (output expression :164-165). */
void oracle_hdiffsa(const oeb_domain *d, const oeb_s3 *out, const oeb_s3 *in, const oeb_s3 *mask,
                    const oeb_s1 *crlato, const oeb_s1 *crlatu) {
  FOR_INTERIOR(d) {
    AT(out, i, j, k) = AT(in, i, j, k) + (hd_flx(in, crlato, crlatu, i - 1, j, k) - hd_flx(in, crlato, crlatu, i, j, k) +
                                          hd_fly(in, crlato, crlatu, i, j - 1, k) - hd_fly(in, crlato, crlatu, i, j, k)) *
                                             AT(mask, i, j, k);
  }
}

/* ---------------------------------------------------------------- hdiffsmag */
/* Tension at (p,q): kernels/hdiffsmag.h:213 (the T_s_* lines are this pattern
 * at (i,j), (i+1,j), (i,j+1)). */
static double sm_tension(const oeb_s3 *u, const oeb_s3 *v, const oeb_s1 *ac0, double eddlat, double eddlon,
                         int64_t p, int64_t q, int64_t k) {
  return (AT(v, p, q - 1, k) - AT(v, p, q, k)) * eddlat * EARTH_R_RECIP -
         (AT(u, p - 1, q, k) - AT(u, p, q, k)) * AT1(ac0, q) * eddlon;
}

/* Shear at (p,q): kernels/hdiffsmag.h:221 (S_uv_* at (i,j), (i-1,j), (i,j-1)). */
static double sm_shear(const oeb_s3 *u, const oeb_s3 *v, const oeb_s1 *ac0, double eddlat, double eddlon, int64_t p,
                       int64_t q, int64_t k) {
  return (AT(u, p, q + 1, k) - AT(u, p, q, k)) * eddlat * EARTH_R_RECIP +
         (AT(v, p + 1, q, k) - AT(v, p, q, k)) * AT1(ac0, q) * eddlon;
}

/* std::max(a, b) and std::min(a, b) as the C++ library defines them. */
static double cxx_max(double a, double b) { return (a < b) ? b : a; }
static double cxx_min(double a, double b) { return (b < a) ? b : a; }

/* kernels/hdiffsmag.h:205-263. */
void oracle_hdiffsmag(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s3 *mask, const oeb_s1 *crlavo, const oeb_s1 *crlavu,
                      const oeb_s1 *crlato, const oeb_s1 *crlatu, const oeb_s1 *acrlat0, double eddlat,
                      double eddlon, double tau_smag, double weight_smag) {
  FOR_INTERIOR(d) {
    const double t0 = sm_tension(uin, vin, acrlat0, eddlat, eddlon, i, j, k);
    const double t_ip = sm_tension(uin, vin, acrlat0, eddlat, eddlon, i + 1, j, k);
    const double t_jp = sm_tension(uin, vin, acrlat0, eddlat, eddlon, i, j + 1, k);
    const double s0 = sm_shear(uin, vin, acrlat0, eddlat, eddlon, i, j, k);
    const double s_im = sm_shear(uin, vin, acrlat0, eddlat, eddlon, i - 1, j, k);
    const double s_jm = sm_shear(uin, vin, acrlat0, eddlat, eddlon, i, j - 1, k);
    const double tt0 = t0 * t0, tt_ip = t_ip * t_ip, tt_jp = t_jp * t_jp;
    const double ss0 = s0 * s0, ss_im = s_im * s_im, ss_jm = s_jm * s_jm;
    const double hdweight = weight_smag * AT(mask, i, j, k);

    double smag_u = tau_smag * sqrt(0.5 * (tt_ip + tt0) + 0.5 * (ss_jm + ss0)) - hdweight;
    smag_u = cxx_min(0.5, cxx_max(0.0, smag_u));
    const double u0 = AT(uin, i, j, k);
    const double lapu = AT(uin, i + 1, j, k) + AT(uin, i - 1, j, k) - 2.0 * u0 +
                        AT1(crlato, j) * (AT(uin, i, j + 1, k) - u0) + AT1(crlatu, j) * (AT(uin, i, j - 1, k) - u0);
    AT(uout, i, j, k) = u0 + smag_u * lapu;

    double smag_v = tau_smag * sqrt(0.5 * (tt_jp + tt0) + 0.5 * (ss_im + ss0)) - hdweight;
    smag_v = cxx_min(0.5, cxx_max(0.0, smag_v));
    const double v0 = AT(vin, i, j, k);
    const double lapv = AT(vin, i + 1, j, k) + AT(vin, i - 1, j, k) - 2.0 * v0 +
                        AT1(crlavo, j) * (AT(vin, i, j + 1, k) - v0) + AT1(crlavu, j) * (AT(vin, i, j - 1, k) - v0);
    AT(vout, i, j, k) = v0 + smag_v * lapv;
  }
}

/* ------------------------------------------------------------ hadvuv(5th) */
/* 3rd-order upwind advection of f at (i,j,k): kernels/hadvuv.h:4-27. */
static double upwind3(const oeb_s3 *f, int64_t i, int64_t j, int64_t k, double ua, double va, double eddlat,
                      double eddlon) {
  double rx, ry;
  if (ua > 0) {
    rx = ua * (-1.0 / 6.0 * AT(f, i - 2, j, k) + AT(f, i - 1, j, k) + -0.5 * AT(f, i, j, k) +
               -1.0 / 3.0 * AT(f, i + 1, j, k));
  } else {
    rx = -ua * (-1.0 / 3.0 * AT(f, i - 1, j, k) + -0.5 * AT(f, i, j, k) + AT(f, i + 1, j, k) +
                -1.0 / 6.0 * AT(f, i + 2, j, k));
  }
  if (va > 0) {
    ry = va * (-1.0 / 6.0 * AT(f, i, j - 2, k) + AT(f, i, j - 1, k) + -0.5 * AT(f, i, j, k) +
               -1.0 / 3.0 * AT(f, i, j + 1, k));
  } else {
    ry = -va * (-1.0 / 3.0 * AT(f, i, j - 1, k) + -0.5 * AT(f, i, j, k) + AT(f, i, j + 1, k) +
                -1.0 / 6.0 * AT(f, i, j + 2, k));
  }
  return eddlat * rx + eddlon * ry;
}

/* 5th-order upwind advection: kernels/hadvuv5th.h:4-39. */
static double upwind5(const oeb_s3 *f, int64_t i, int64_t j, int64_t k, double ua, double va, double eddlat,
                      double eddlon) {
  double rx, ry;
  if (ua > 0) {
    rx = ua * (1.0 / 30.0 * AT(f, i - 3, j, k) + -1.0 / 4.0 * AT(f, i - 2, j, k) + AT(f, i - 1, j, k) +
               -1.0 / 3.0 * AT(f, i, j, k) + -1.0 / 2.0 * AT(f, i + 1, j, k) + 1.0 / 20.0 * AT(f, i + 2, j, k));
  } else {
    rx = -ua * (1.0 / 20.0 * AT(f, i - 2, j, k) + -1.0 / 2.0 * AT(f, i - 1, j, k) + -1.0 / 3.0 * AT(f, i, j, k) +
                AT(f, i + 1, j, k) + -1.0 / 4.0 * AT(f, i + 2, j, k) + 1.0 / 30.0 * AT(f, i + 3, j, k));
  }
  if (va > 0) {
    ry = va * (1.0 / 30.0 * AT(f, i, j - 3, k) + -1.0 / 4.0 * AT(f, i, j - 2, k) + AT(f, i, j - 1, k) +
               -1.0 / 3.0 * AT(f, i, j, k) + -1.0 / 2.0 * AT(f, i, j + 1, k) + 1.0 / 20.0 * AT(f, i, j + 2, k));
  } else {
    ry = -va * (1.0 / 20.0 * AT(f, i, j - 2, k) + -1.0 / 2.0 * AT(f, i, j - 1, k) + -1.0 / 3.0 * AT(f, i, j, k) +
                AT(f, i, j + 1, k) + -1.0 / 4.0 * AT(f, i, j + 2, k) + 1.0 / 30.0 * AT(f, i, j + 3, k));
  }
  return eddlat * rx + eddlon * ry;
}

typedef double (*upwind_fn)(const oeb_s3 *, int64_t, int64_t, int64_t, double, double, double, double);

/* The shared body of hadvuv_openmp (kernels/hadvuv.h:178-210) and
 * hadvuv5th_openmp (kernels/hadvuv5th.h:177-201), which differ only in the
 * upwind stencil. */
static void hadv_body(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0,
                      const oeb_s1 *tgrlatda1, double eddlat, double eddlon, upwind_fn adv) {
  FOR_INTERIOR(d) {
    const double uatupos = (1.0 / 3.0) * (AT(uin, i - 1, j, k) + AT(uin, i, j, k) + AT(uin, i + 1, j, k));
    const double vatupos =
        0.25 * (AT(vin, i + 1, j, k) + AT(vin, i + 1, j - 1, k) + AT(vin, i, j, k) + AT(vin, i, j - 1, k));
    double uavg = AT1(acrlat0, j) * uatupos;
    double vavg = EARTH_R_RECIP * vatupos;
    const double ures = adv(uin, i, j, k, uavg, vavg, eddlat, eddlon);
    AT(uout, i, j, k) = ures + AT1(tgrlatda0, j) * AT(uin, i, j, k) * vatupos;

    const double uatvpos =
        0.25 * (AT(uin, i - 1, j, k) + AT(uin, i, j, k) + AT(uin, i, j + 1, k) + AT(uin, i - 1, j + 1, k));
    const double vatvpos = 1.0 / 3.0 * (AT(vin, i, j - 1, k) + AT(vin, i, j, k) + AT(vin, i, j + 1, k));
    uavg = AT1(acrlat1, j) * uatvpos;
    vavg = EARTH_R_RECIP * vatvpos;
    const double vres = adv(vin, i, j, k, uavg, vavg, eddlat, eddlon);
    AT(vout, i, j, k) = vres - AT1(tgrlatda1, j) * uatvpos * uatvpos;
  }
}

void oracle_hadvuv(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                   const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0,
                   const oeb_s1 *tgrlatda1, double eddlat, double eddlon) {
  hadv_body(d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat, eddlon, upwind3);
}

void oracle_hadvuv5th(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0,
                      const oeb_s1 *tgrlatda1, double eddlat, double eddlon) {
  hadv_body(d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat, eddlon, upwind5);
}

/* -------------------------------------------------------------- fastwavesuv */
/* Pressure interpolated to the half level below k at (p,q):
 * kernels/fastwavesuv.h:312 (the _ppgk_* lines are this pattern). */
static double fw_ppgk(const oeb_s3 *wgtfac, const oeb_s3 *ppuv, int64_t p, int64_t q, int64_t k) {
  return AT(wgtfac, p, q, k) * AT(ppuv, p, q, k) + (1.0 - AT(wgtfac, p, q, k)) * AT(ppuv, p, q, k - 1);
}

/* kernels/fastwavesuv.h:304-341. */
void oracle_fastwavesuv(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                        const oeb_s3 *vin, const oeb_s3 *utens, const oeb_s3 *vtens, const oeb_s3 *wgtfac,
                        const oeb_s3 *ppuv, const oeb_s3 *hhl, const oeb_s3 *rho, const oeb_s1 *fx,
                        double edadlat, double dt) {
  FOR_INTERIOR(d) {
    const double ppgk0 = fw_ppgk(wgtfac, ppuv, i, j, k);
    const double ppgk_i = fw_ppgk(wgtfac, ppuv, i + 1, j, k);
    const double ppgk_j = fw_ppgk(wgtfac, ppuv, i, j + 1, k);
    const double ppgk0_up = fw_ppgk(wgtfac, ppuv, i, j, k + 1);
    const double ppgk_i_up = fw_ppgk(wgtfac, ppuv, i + 1, j, k + 1);
    const double ppgk_j_up = fw_ppgk(wgtfac, ppuv, i, j + 1, k + 1);
    const double ppgc0 = ppgk0_up - ppgk0;
    const double ppgc_i = ppgk_i_up - ppgk_i;
    const double ppgc_j = ppgk_j_up - ppgk_j;

    const double ppgu =
        (AT(ppuv, i + 1, j, k) - AT(ppuv, i, j, k)) +
        (ppgc_i + ppgc0) * 0.5 * ((AT(hhl, i, j, k + 1) + AT(hhl, i, j, k)) - (AT(hhl, i + 1, j, k + 1) + AT(hhl, i + 1, j, k))) /
            ((AT(hhl, i, j, k + 1) - AT(hhl, i, j, k)) + (AT(hhl, i + 1, j, k + 1) - AT(hhl, i + 1, j, k)));
    const double ppgv =
        (AT(ppuv, i, j + 1, k) - AT(ppuv, i, j, k)) +
        (ppgc_j + ppgc0) * 0.5 * ((AT(hhl, i, j, k + 1) + AT(hhl, i, j, k)) - (AT(hhl, i, j + 1, k + 1) + AT(hhl, i, j + 1, k))) /
            ((AT(hhl, i, j, k + 1) - AT(hhl, i, j, k)) + (AT(hhl, i, j + 1, k + 1) - AT(hhl, i, j + 1, k)));

    AT(uout, i, j, k) = AT(uin, i, j, k) +
                        dt * (AT(utens, i, j, k) - ppgu * (2.0 * AT1(fx, j) / (AT(rho, i + 1, j, k) + AT(rho, i, j, k))));
    AT(vout, i, j, k) = AT(vin, i, j, k) +
                        dt * (AT(vtens, i, j, k) - ppgv * (2.0 * edadlat / (AT(rho, i, j + 1, k) + AT(rho, i, j, k))));
  }
}

/* -------------------------------------------------------------------- uvbke */
/* kernels/uvbke.h:61-75. */
void oracle_uvbke(const oeb_domain *d, const oeb_s3 *ub, const oeb_s3 *vb, const oeb_s3 *uc, const oeb_s3 *vc,
                  const oeb_s3 *cosa, const oeb_s3 *rsina, double dt5) {
  FOR_INTERIOR(d) {
    AT(ub, i, j, k) = ((dt5 * ((AT(uc, i, j - 1, k) + AT(uc, i, j, k)) -
                               ((AT(vc, i - 1, j, k) + AT(vc, i, j, k)) * AT(cosa, i, j, k)))) *
                       AT(rsina, i, j, k));
    AT(vb, i, j, k) = ((dt5 * ((AT(vc, i - 1, j, k) + AT(vc, i, j, k)) -
                               ((AT(uc, i, j - 1, k) + AT(uc, i, j, k)) * AT(cosa, i, j, k)))) *
                       AT(rsina, i, j, k));
  }
}

/* ----------------------------------------------------------------- p_grad_c */
/* kernels/p_grad_c.h:80-101: delpc(i-1) and delpc(j-1) are read directly. */
void oracle_p_grad_c(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                     const oeb_s3 *vin, const oeb_s3 *rdxc, const oeb_s3 *rdyc, const oeb_s3 *delpc,
                     const oeb_s3 *gz, const oeb_s3 *pkc, double dt2) {
  FOR_INTERIOR(d) {
    const double wk0 = AT(delpc, i, j, k);
    const double wk_im = AT(delpc, i - 1, j, k);
    const double wk_jm = AT(delpc, i, j - 1, k);
    AT(uout, i, j, k) =
        (AT(uin, i, j, k) + (((dt2 * AT(rdxc, i, j, k)) / (wk_im + wk0)) *
                             (((AT(gz, i - 1, j, k + 1) - AT(gz, i, j, k)) * (AT(pkc, i, j, k + 1) - AT(pkc, i - 1, j, k))) +
                              ((AT(gz, i - 1, j, k) - AT(gz, i, j, k + 1)) * (AT(pkc, i - 1, j, k + 1) - AT(pkc, i, j, k))))));
    AT(vout, i, j, k) =
        (AT(vin, i, j, k) + (((dt2 * AT(rdyc, i, j, k)) / (wk_jm + wk0)) *
                             (((AT(gz, i, j - 1, k + 1) - AT(gz, i, j, k)) * (AT(pkc, i, j, k + 1) - AT(pkc, i, j - 1, k))) +
                              ((AT(gz, i, j - 1, k) - AT(gz, i, j, k + 1)) * (AT(pkc, i, j - 1, k + 1) - AT(pkc, i, j, k))))));
  }
}

/* ---------------------------------------------------------------- nh_p_grad */
/* kernels/nh_p_grad.h:116-146. */
void oracle_nh_p_grad(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s3 *rdx, const oeb_s3 *rdy, const oeb_s3 *gz, const oeb_s3 *pp,
                      const oeb_s3 *pk3, const oeb_s3 *wk1, double dt) {
  FOR_INTERIOR(d) {
    const double wk0 = (AT(pk3, i, j, k + 1) - AT(pk3, i, j, k));
    const double wk_i = (AT(pk3, i + 1, j, k + 1) - AT(pk3, i + 1, j, k));
    const double wk_j = (AT(pk3, i, j + 1, k + 1) - AT(pk3, i, j + 1, k));
    const double du =
        ((dt / (wk0 + wk_i)) *
         (((AT(gz, i, j, k + 1) - AT(gz, i + 1, j, k)) * (AT(pk3, i + 1, j, k + 1) - AT(pk3, i, j, k))) +
          ((AT(gz, i, j, k) - AT(gz, i + 1, j, k + 1)) * (AT(pk3, i, j, k + 1) - AT(pk3, i + 1, j, k)))));
    AT(uout, i, j, k) =
        (((AT(uin, i, j, k) + du) +
          ((dt / (AT(wk1, i, j, k) + AT(wk1, i + 1, j, k))) *
           (((AT(gz, i, j, k + 1) - AT(gz, i + 1, j, k)) * (AT(pp, i + 1, j, k + 1) - AT(pp, i, j, k))) +
            ((AT(gz, i, j, k) - AT(gz, i + 1, j, k + 1)) * (AT(pp, i, j, k + 1) - AT(pp, i + 1, j, k)))))) *
         AT(rdx, i, j, k));
    const double dv =
        ((dt / (wk0 + wk_j)) *
         (((AT(gz, i, j, k + 1) - AT(gz, i, j + 1, k)) * (AT(pk3, i, j + 1, k + 1) - AT(pk3, i, j, k))) +
          ((AT(gz, i, j, k) - AT(gz, i, j + 1, k + 1)) * (AT(pk3, i, j, k + 1) - AT(pk3, i, j + 1, k)))));
    AT(vout, i, j, k) =
        (((AT(vin, i, j, k) + dv) +
          ((dt / (AT(wk1, i, j, k) + AT(wk1, i, j + 1, k))) *
           (((AT(gz, i, j, k + 1) - AT(gz, i, j + 1, k)) * (AT(pp, i, j + 1, k + 1) - AT(pp, i, j, k))) +
            ((AT(gz, i, j, k) - AT(gz, i, j + 1, k + 1)) * (AT(pp, i, j, k + 1) - AT(pp, i, j + 1, k)))))) *
         AT(rdy, i, j, k));
  }
}

/* ------------------------------------------------------------------- fvtp2d */
/* A scratch array with the layout of `like`, owned by the caller of
 * scratch_new (free with scratch_free). */
static int scratch_new(oeb_s3 *s, const oeb_s3 *like) {
  *s = *like;
  const size_t n = (size_t)(like->sizes[0] * like->strides[0]) + 64;
  s->allocatedPtr = (double *)calloc(n, sizeof(double));
  s->alignedPtr = s->allocatedPtr;
  return s->allocatedPtr == NULL;
}
static void scratch_free(oeb_s3 *s) { free(s->allocatedPtr); }

/* One PPM flux stage.  Runs on each level k, in the stage-by-stage order of
 * kernels/fvtp2d.h.  The stage along j is :9-37 and :129-158, and the stage
 * along i is :53-82 and :84-113.  f is the advected field and c the Courant
 * number.  al is formed on [al_lo, al_hi) along the stage axis and the
 * limiter arrays on [lim_lo, lim_hi).  The flux is written on
 * [out_lo, out_hi).  The other axis spans [o_lo, o_hi). */
static void ppm_stage(int along_j, const oeb_domain *d, const oeb_s3 *f, const oeb_s3 *c, const oeb_s3 *flux,
                      const oeb_s3 *al, const oeb_s3 *almq, const oeb_s3 *br, const oeb_s3 *b0, const oeb_s3 *smt5,
                      int64_t o_lo, int64_t o_hi, int64_t al_lo, int64_t al_hi, int64_t lim_lo, int64_t lim_hi,
                      int64_t out_lo, int64_t out_hi, double p1, double p2) {
  /* (a, m) -> (i, j, k): a is the other axis, m the stage axis. */
#define IX(a, m) (along_j ? (a) : (m))
#define JX(a, m) (along_j ? (m) : (a))
#define F(s, a, m) AT(s, IX(a, m), JX(a, m), k)
  for (int64_t k = 0; k < d->nz; ++k) {
    for (int64_t a = o_lo; a < o_hi; ++a)
      for (int64_t m = al_lo; m < al_hi; ++m)
        F(al, a, m) = ((p1 * (F(f, a, m - 1) + F(f, a, m))) + (p2 * (F(f, a, m - 2) + F(f, a, m + 1))));
  }
  for (int64_t k = 0; k < d->nz; ++k) {
    for (int64_t a = o_lo; a < o_hi; ++a)
      for (int64_t m = lim_lo; m < lim_hi; ++m) {
        F(almq, a, m) = (F(al, a, m) - F(f, a, m));
        F(br, a, m) = (F(al, a, m + 1) - F(f, a, m));
        F(b0, a, m) = (F(almq, a, m) + F(br, a, m));
        F(smt5, a, m) = ((F(almq, a, m) * F(br, a, m)) < 0) ? 1.0 : 0.0;
      }
    for (int64_t a = o_lo; a < o_hi; ++a)
      for (int64_t m = out_lo; m < out_hi; ++m) {
        const double smt5tmp = (F(smt5, a, m - 1) + (F(smt5, a, m) * (F(smt5, a, m - 1) == 0 ? 1.0 : 0.0)));
        const double cm = F(c, a, m);
        const double ctmp = ((cm > 0.) ? ((1. - cm) * (F(br, a, m - 1) - (cm * F(b0, a, m - 1))))
                                       : ((1. + cm) * (F(almq, a, m) + (cm * F(b0, a, m)))));
        F(flux, a, m) = ((cm > 0.) ? (F(f, a, m - 1) + (ctmp * smt5tmp)) : (F(f, a, m) + (ctmp * smt5tmp)));
      }
  }
#undef F
#undef JX
#undef IX
}

/* kernels/fvtp2d.h:4-177 (DEFAULT variant), stage by stage. */
int oracle_fvtp2d(const oeb_domain *d, const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *fx1,
                  const oeb_s3 *fx2, const oeb_s3 *fy1, const oeb_s3 *fy2, const oeb_s3 *q, const oeb_s3 *crx,
                  const oeb_s3 *cry, const oeb_s3 *ra_x, const oeb_s3 *ra_y, const oeb_s3 *xfx, const oeb_s3 *yfx,
                  const oeb_s3 *area, double p1, double p2) {
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  oeb_s3 fxx, fyy, al, almq, br, b0, smt5;
  int fail = 0;
  fail |= scratch_new(&fxx, q);
  fail |= scratch_new(&fyy, q);
  fail |= scratch_new(&al, q);
  fail |= scratch_new(&almq, q);
  fail |= scratch_new(&br, q);
  fail |= scratch_new(&b0, q);
  fail |= scratch_new(&smt5, q);
  if (!fail) {
    /* fy2 = PPM_j(q, cry): :9-37. */
    ppm_stage(1, d, q, cry, fy2, &al, &almq, &br, &b0, &smt5, -3, nx + 2, -1, ny + 2, -1, ny + 1, 0, ny + 1, p1, p2);
    /* q_i: :39-51. */
    for (int64_t k = 0; k < nz; ++k) {
      for (int64_t j = 0; j < ny + 1; ++j)
        for (int64_t i = -3; i < nx + 2; ++i) AT(&fyy, i, j, k) = (AT(yfx, i, j, k) * AT(fy2, i, j, k));
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = -3; i < nx + 2; ++i)
          AT(q_i, i, j, k) =
              ((((AT(q, i, j, k) * AT(area, i, j, k)) + AT(&fyy, i, j, k)) - AT(&fyy, i, j + 1, k)) / AT(ra_y, i, j, k));
    }
    /* fx1 = PPM_i(q_i, crx): :53-82. */
    ppm_stage(0, d, q_i, crx, fx1, &al, &almq, &br, &b0, &smt5, 0, ny, -1, nx + 1, -1, nx, 0, nx, p1, p2);
    /* fx2 = PPM_i(q, crx): :84-113. */
    ppm_stage(0, d, q, crx, fx2, &al, &almq, &br, &b0, &smt5, -3, ny + 2, -1, nx + 2, -1, nx + 1, 0, nx + 1, p1, p2);
    /* q_j: :115-127. */
    for (int64_t k = 0; k < nz; ++k) {
      for (int64_t j = -3; j < ny + 2; ++j)
        for (int64_t i = 0; i < nx + 1; ++i) AT(&fxx, i, j, k) = (AT(xfx, i, j, k) * AT(fx2, i, j, k));
      for (int64_t j = -3; j < ny + 2; ++j)
        for (int64_t i = 0; i < nx; ++i)
          AT(q_j, i, j, k) =
              ((((AT(q, i, j, k) * AT(area, i, j, k)) + AT(&fxx, i, j, k)) - AT(&fxx, i + 1, j, k)) / AT(ra_x, i, j, k));
    }
    /* fy1 = PPM_j(q_j, cry): :129-158. */
    ppm_stage(1, d, q_j, cry, fy1, &al, &almq, &br, &b0, &smt5, 0, nx, -1, ny + 1, -1, ny, 0, ny, p1, p2);
    /* transport_flux: :160-176. */
    for (int64_t k = 0; k < nz; ++k)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
          AT(fx1, i, j, k) = ((0.5 * (AT(fx1, i, j, k) + AT(fx2, i, j, k))) * AT(xfx, i, j, k));
          AT(fy1, i, j, k) = ((0.5 * (AT(fy1, i, j, k) + AT(fy2, i, j, k))) * AT(yfx, i, j, k));
        }
  }
  scratch_free(&fxx);
  scratch_free(&fyy);
  scratch_free(&al);
  scratch_free(&almq);
  scratch_free(&br);
  scratch_free(&b0);
  scratch_free(&smt5);
  return fail;
}
