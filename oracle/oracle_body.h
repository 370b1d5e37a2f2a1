/*
 * oracle_body.h -- the ten stencils, written once for an element type R.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  oracle.c includes this file
 * twice: with R = double (the FP64 oracle, oracle_<name>) and with R = float
 * (the FP32 oracle, oracle_<name>_f32).  Before each inclusion it defines
 *   R        the element type (ElementType of the reference, src/<name>.cpp:9)
 *   S3, S1   the 3-D / 1-D descriptor types of that element type
 *   ORC(n)   the exported name of stencil n
 *   LOC(n)   the name of file-local helper n
 *
 * Every expression keeps the reference's operand order and parenthesization,
 * so that with contraction off the result is determined bit for bit.  C
 * evaluates a + b - c + d as ((a + b) - c) + d, as C++ does.  The semantics
 * are those of the *_openmp variants (fvtp2d: DEFAULT).  They differ from
 * DEFAULT only where DEFAULT reads unwritten scratch.  Every constant the
 * reference writes as ElementType(x), or as a quotient of two, is formed in R
 * here too.  The reference's two double-typed constants stay double:
 * laplace's -4.0 (kernels/laplace.h:39) and uvbke's dt5 (src/uvbke.cpp:11-12).
 * <tgmath.h> makes sqrt the float or double routine by argument type, as
 * std::sqrt is.
 *
 * FP32 form, where it is more than the reference's float build:
 *   - fastwavesuv, p_grad_c and nh_p_grad read their pressure and
 *     geopotential arrays (and fastwavesuv its hhl) as FP64 islands, always
 *     oeb_s3.  Every difference of those is formed in double and rounded once
 *     to R where it meets an R value.  With R = double those roundings vanish
 *     and the expressions are the reference's.
 *   - fvtp2d's flux divergences are compensated (LOC(flux_div)).
 */

/* ------------------------------------------------------------------ laplace */
/* kernels/laplace.h:34-43 (expression :39).  -4.0 is a double literal. */
void ORC(laplace)(const oeb_domain *d, const S3 *in, const S3 *out) {
  FOR_INTERIOR(d) {
    AT(out, i, j, k) = (R)(-4.0 * AT(in, i, j, k) +
                           ((AT(in, i - 1, j, k) + AT(in, i + 1, j, k)) + (AT(in, i, j + 1, k) + AT(in, i, j - 1, k))));
  }
}

/* ------------------------------------------------------------------ hdiffsa */
/* The Laplacian at (p,q) with the row's crlato/crlatu: kernels/hdiffsa.h:147-149
 * (the three _lap_* lines are this one pattern at (i,j), (i+1,j), (i,j+1)). */
static R LOC(hd_lap)(const S3 *in, const S1 *lo, const S1 *lu, int64_t p, int64_t q, int64_t k) {
  const R c = AT(in, p, q, k);
  return AT(in, p - 1, q, k) + AT(in, p + 1, q, k) - (R)2.0 * c + AT1(lo, q) * (AT(in, p, q + 1, k) - c) +
         AT1(lu, q) * (AT(in, p, q - 1, k) - c);
}

/* Limited x flux between (p,q) and (p+1,q): kernels/hdiffsa.h:152-155. */
static R LOC(hd_flx)(const S3 *in, const S1 *lo, const S1 *lu, int64_t p, int64_t q, int64_t k) {
  R f = LOC(hd_lap)(in, lo, lu, p + 1, q, k) - LOC(hd_lap)(in, lo, lu, p, q, k);
  if (f * (AT(in, p + 1, q, k) - AT(in, p, q, k)) > 0) f = 0;
  return f;
}

/* Limited y flux between (p,q) and (p,q+1): kernels/hdiffsa.h:156-159. */
static R LOC(hd_fly)(const S3 *in, const S1 *lo, const S1 *lu, int64_t p, int64_t q, int64_t k) {
  R f = AT1(lo, q) * (LOC(hd_lap)(in, lo, lu, p, q + 1, k) - LOC(hd_lap)(in, lo, lu, p, q, k));
  if (f * (AT(in, p, q + 1, k) - AT(in, p, q, k)) > 0) f = 0;
  return f;
}

/* kernels/hdiffsa.h:141-169 (output expression :164-165). */
void ORC(hdiffsa)(const oeb_domain *d, const S3 *out, const S3 *in, const S3 *mask, const S1 *crlato,
                  const S1 *crlatu) {
  FOR_INTERIOR(d) {
    AT(out, i, j, k) =
        AT(in, i, j, k) + (LOC(hd_flx)(in, crlato, crlatu, i - 1, j, k) - LOC(hd_flx)(in, crlato, crlatu, i, j, k) +
                           LOC(hd_fly)(in, crlato, crlatu, i, j - 1, k) - LOC(hd_fly)(in, crlato, crlatu, i, j, k)) *
                              AT(mask, i, j, k);
  }
}

/* ---------------------------------------------------------------- hdiffsmag */
/* EARTH_RADIUS_RECIP, include/util.h:17-18: (ElementType)1.0 / (ElementType)6371.229e3. */
#define EARTH_R_RECIP_R ((R)1.0 / (R)6371.229e3)

/* Tension at (p,q): kernels/hdiffsmag.h:213 (the T_s_* lines are this pattern
 * at (i,j), (i+1,j), (i,j+1)). */
static R LOC(sm_tension)(const S3 *u, const S3 *v, const S1 *ac0, R eddlat, R eddlon, int64_t p, int64_t q,
                         int64_t k) {
  return (AT(v, p, q - 1, k) - AT(v, p, q, k)) * eddlat * EARTH_R_RECIP_R -
         (AT(u, p - 1, q, k) - AT(u, p, q, k)) * AT1(ac0, q) * eddlon;
}

/* Shear at (p,q): kernels/hdiffsmag.h:221 (S_uv_* at (i,j), (i-1,j), (i,j-1)). */
static R LOC(sm_shear)(const S3 *u, const S3 *v, const S1 *ac0, R eddlat, R eddlon, int64_t p, int64_t q,
                       int64_t k) {
  return (AT(u, p, q + 1, k) - AT(u, p, q, k)) * eddlat * EARTH_R_RECIP_R +
         (AT(v, p + 1, q, k) - AT(v, p, q, k)) * AT1(ac0, q) * eddlon;
}

/* std::max(a, b) and std::min(a, b) as the C++ library defines them. */
static R LOC(cxx_max)(R a, R b) { return (a < b) ? b : a; }
static R LOC(cxx_min)(R a, R b) { return (b < a) ? b : a; }

/* kernels/hdiffsmag.h:205-263. */
void ORC(hdiffsmag)(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
                    const S3 *mask, const S1 *crlavo, const S1 *crlavu, const S1 *crlato, const S1 *crlatu,
                    const S1 *acrlat0, R eddlat, R eddlon, R tau_smag, R weight_smag) {
  FOR_INTERIOR(d) {
    const R t0 = LOC(sm_tension)(uin, vin, acrlat0, eddlat, eddlon, i, j, k);
    const R t_ip = LOC(sm_tension)(uin, vin, acrlat0, eddlat, eddlon, i + 1, j, k);
    const R t_jp = LOC(sm_tension)(uin, vin, acrlat0, eddlat, eddlon, i, j + 1, k);
    const R s0 = LOC(sm_shear)(uin, vin, acrlat0, eddlat, eddlon, i, j, k);
    const R s_im = LOC(sm_shear)(uin, vin, acrlat0, eddlat, eddlon, i - 1, j, k);
    const R s_jm = LOC(sm_shear)(uin, vin, acrlat0, eddlat, eddlon, i, j - 1, k);
    const R tt0 = t0 * t0, tt_ip = t_ip * t_ip, tt_jp = t_jp * t_jp;
    const R ss0 = s0 * s0, ss_im = s_im * s_im, ss_jm = s_jm * s_jm;
    const R hdweight = weight_smag * AT(mask, i, j, k);

    R smag_u = tau_smag * sqrt((R)0.5 * (tt_ip + tt0) + (R)0.5 * (ss_jm + ss0)) - hdweight;
    smag_u = LOC(cxx_min)((R)0.5, LOC(cxx_max)((R)0.0, smag_u));
    const R u0 = AT(uin, i, j, k);
    const R lapu = AT(uin, i + 1, j, k) + AT(uin, i - 1, j, k) - (R)2.0 * u0 +
                   AT1(crlato, j) * (AT(uin, i, j + 1, k) - u0) + AT1(crlatu, j) * (AT(uin, i, j - 1, k) - u0);
    AT(uout, i, j, k) = u0 + smag_u * lapu;

    R smag_v = tau_smag * sqrt((R)0.5 * (tt_jp + tt0) + (R)0.5 * (ss_im + ss0)) - hdweight;
    smag_v = LOC(cxx_min)((R)0.5, LOC(cxx_max)((R)0.0, smag_v));
    const R v0 = AT(vin, i, j, k);
    const R lapv = AT(vin, i + 1, j, k) + AT(vin, i - 1, j, k) - (R)2.0 * v0 +
                   AT1(crlavo, j) * (AT(vin, i, j + 1, k) - v0) + AT1(crlavu, j) * (AT(vin, i, j - 1, k) - v0);
    AT(vout, i, j, k) = v0 + smag_v * lapv;
  }
}

/* ------------------------------------------------------------ hadvuv(5th) */
/* 3rd-order upwind advection of f at (i,j,k): kernels/hadvuv.h:4-27. */
static R LOC(upwind3)(const S3 *f, int64_t i, int64_t j, int64_t k, R ua, R va, R eddlat, R eddlon) {
  R rx, ry;
  if (ua > 0) {
    rx = ua * ((R)-1.0 / (R)6.0 * AT(f, i - 2, j, k) + AT(f, i - 1, j, k) + (R)-0.5 * AT(f, i, j, k) +
               (R)-1.0 / (R)3.0 * AT(f, i + 1, j, k));
  } else {
    rx = -ua * ((R)-1.0 / (R)3.0 * AT(f, i - 1, j, k) + (R)-0.5 * AT(f, i, j, k) + AT(f, i + 1, j, k) +
                (R)-1.0 / (R)6.0 * AT(f, i + 2, j, k));
  }
  if (va > 0) {
    ry = va * ((R)-1.0 / (R)6.0 * AT(f, i, j - 2, k) + AT(f, i, j - 1, k) + (R)-0.5 * AT(f, i, j, k) +
               (R)-1.0 / (R)3.0 * AT(f, i, j + 1, k));
  } else {
    ry = -va * ((R)-1.0 / (R)3.0 * AT(f, i, j - 1, k) + (R)-0.5 * AT(f, i, j, k) + AT(f, i, j + 1, k) +
                (R)-1.0 / (R)6.0 * AT(f, i, j + 2, k));
  }
  return eddlat * rx + eddlon * ry;
}

/* 5th-order upwind advection: kernels/hadvuv5th.h:4-39. */
static R LOC(upwind5)(const S3 *f, int64_t i, int64_t j, int64_t k, R ua, R va, R eddlat, R eddlon) {
  R rx, ry;
  if (ua > 0) {
    rx = ua * ((R)1.0 / (R)30.0 * AT(f, i - 3, j, k) + (R)-1.0 / (R)4.0 * AT(f, i - 2, j, k) + AT(f, i - 1, j, k) +
               (R)-1.0 / (R)3.0 * AT(f, i, j, k) + (R)-1.0 / (R)2.0 * AT(f, i + 1, j, k) +
               (R)1.0 / (R)20.0 * AT(f, i + 2, j, k));
  } else {
    rx = -ua * ((R)1.0 / (R)20.0 * AT(f, i - 2, j, k) + (R)-1.0 / (R)2.0 * AT(f, i - 1, j, k) +
                (R)-1.0 / (R)3.0 * AT(f, i, j, k) + AT(f, i + 1, j, k) + (R)-1.0 / (R)4.0 * AT(f, i + 2, j, k) +
                (R)1.0 / (R)30.0 * AT(f, i + 3, j, k));
  }
  if (va > 0) {
    ry = va * ((R)1.0 / (R)30.0 * AT(f, i, j - 3, k) + (R)-1.0 / (R)4.0 * AT(f, i, j - 2, k) + AT(f, i, j - 1, k) +
               (R)-1.0 / (R)3.0 * AT(f, i, j, k) + (R)-1.0 / (R)2.0 * AT(f, i, j + 1, k) +
               (R)1.0 / (R)20.0 * AT(f, i, j + 2, k));
  } else {
    ry = -va * ((R)1.0 / (R)20.0 * AT(f, i, j - 2, k) + (R)-1.0 / (R)2.0 * AT(f, i, j - 1, k) +
                (R)-1.0 / (R)3.0 * AT(f, i, j, k) + AT(f, i, j + 1, k) + (R)-1.0 / (R)4.0 * AT(f, i, j + 2, k) +
                (R)1.0 / (R)30.0 * AT(f, i, j + 3, k));
  }
  return eddlat * rx + eddlon * ry;
}

/* The shared body of hadvuv_openmp (kernels/hadvuv.h:178-210) and
 * hadvuv5th_openmp (kernels/hadvuv5th.h:177-201), which differ only in the
 * upwind stencil. */
static void LOC(hadv_body)(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
                           const S1 *acrlat0, const S1 *acrlat1, const S1 *tgrlatda0, const S1 *tgrlatda1,
                           R eddlat, R eddlon, int order) {
  FOR_INTERIOR(d) {
    const R uatupos = ((R)1.0 / (R)3.0) * (AT(uin, i - 1, j, k) + AT(uin, i, j, k) + AT(uin, i + 1, j, k));
    const R vatupos =
        (R)0.25 * (AT(vin, i + 1, j, k) + AT(vin, i + 1, j - 1, k) + AT(vin, i, j, k) + AT(vin, i, j - 1, k));
    R uavg = AT1(acrlat0, j) * uatupos;
    R vavg = EARTH_R_RECIP_R * vatupos;
    const R ures = order == 3 ? LOC(upwind3)(uin, i, j, k, uavg, vavg, eddlat, eddlon)
                              : LOC(upwind5)(uin, i, j, k, uavg, vavg, eddlat, eddlon);
    AT(uout, i, j, k) = ures + AT1(tgrlatda0, j) * AT(uin, i, j, k) * vatupos;

    const R uatvpos =
        (R)0.25 * (AT(uin, i - 1, j, k) + AT(uin, i, j, k) + AT(uin, i, j + 1, k) + AT(uin, i - 1, j + 1, k));
    const R vatvpos = (R)1.0 / (R)3.0 * (AT(vin, i, j - 1, k) + AT(vin, i, j, k) + AT(vin, i, j + 1, k));
    uavg = AT1(acrlat1, j) * uatvpos;
    vavg = EARTH_R_RECIP_R * vatvpos;
    const R vres = order == 3 ? LOC(upwind3)(vin, i, j, k, uavg, vavg, eddlat, eddlon)
                              : LOC(upwind5)(vin, i, j, k, uavg, vavg, eddlat, eddlon);
    AT(vout, i, j, k) = vres - AT1(tgrlatda1, j) * uatvpos * uatvpos;
  }
}

void ORC(hadvuv)(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
                 const S1 *acrlat0, const S1 *acrlat1, const S1 *tgrlatda0, const S1 *tgrlatda1, R eddlat,
                 R eddlon) {
  LOC(hadv_body)(d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat, eddlon, 3);
}

void ORC(hadvuv5th)(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
                    const S1 *acrlat0, const S1 *acrlat1, const S1 *tgrlatda0, const S1 *tgrlatda1, R eddlat,
                    R eddlon) {
  LOC(hadv_body)(d, uout, vout, uin, vin, acrlat0, acrlat1, tgrlatda0, tgrlatda1, eddlat, eddlon, 5);
}

/* -------------------------------------------------------------- fastwavesuv */
/* Pressure interpolated to the half level below k at (p,q):
 * kernels/fastwavesuv.h:312 (the _ppgk_* lines are this pattern).  ppuv is an
 * FP64 island; wgtfac is read in R and widened. */
static double LOC(fw_ppgk)(const S3 *wgtfac, const oeb_s3 *ppuv, int64_t p, int64_t q, int64_t k) {
  const double w = AT(wgtfac, p, q, k);
  return w * AT(ppuv, p, q, k) + (1.0 - w) * AT(ppuv, p, q, k - 1);
}

/* kernels/fastwavesuv.h:304-341.  ppuv and hhl are FP64 islands: ppgk, ppgc
 * and ppgu/ppgv are double, rounded once to R in the final update. */
void ORC(fastwavesuv)(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
                      const S3 *utens, const S3 *vtens, const S3 *wgtfac, const oeb_s3 *ppuv, const oeb_s3 *hhl,
                      const S3 *rho, const S1 *fx, R edadlat, R dt) {
  FOR_INTERIOR(d) {
    const double ppgk0 = LOC(fw_ppgk)(wgtfac, ppuv, i, j, k);
    const double ppgk_i = LOC(fw_ppgk)(wgtfac, ppuv, i + 1, j, k);
    const double ppgk_j = LOC(fw_ppgk)(wgtfac, ppuv, i, j + 1, k);
    const double ppgk0_up = LOC(fw_ppgk)(wgtfac, ppuv, i, j, k + 1);
    const double ppgk_i_up = LOC(fw_ppgk)(wgtfac, ppuv, i + 1, j, k + 1);
    const double ppgk_j_up = LOC(fw_ppgk)(wgtfac, ppuv, i, j + 1, k + 1);
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
                        dt * (AT(utens, i, j, k) - (R)ppgu * ((R)2.0 * AT1(fx, j) / (AT(rho, i + 1, j, k) + AT(rho, i, j, k))));
    AT(vout, i, j, k) = AT(vin, i, j, k) +
                        dt * (AT(vtens, i, j, k) - (R)ppgv * ((R)2.0 * edadlat / (AT(rho, i, j + 1, k) + AT(rho, i, j, k))));
  }
}

/* -------------------------------------------------------------------- uvbke */
/* kernels/uvbke.h:61-75.  dt5 is a double whatever R is (src/uvbke.cpp:11-12). */
void ORC(uvbke)(const oeb_domain *d, const S3 *ub, const S3 *vb, const S3 *uc, const S3 *vc, const S3 *cosa,
                const S3 *rsina, double dt5) {
  FOR_INTERIOR(d) {
    AT(ub, i, j, k) = (R)((dt5 * ((AT(uc, i, j - 1, k) + AT(uc, i, j, k)) -
                                  ((AT(vc, i - 1, j, k) + AT(vc, i, j, k)) * AT(cosa, i, j, k)))) *
                          AT(rsina, i, j, k));
    AT(vb, i, j, k) = (R)((dt5 * ((AT(vc, i - 1, j, k) + AT(vc, i, j, k)) -
                                  ((AT(uc, i, j - 1, k) + AT(uc, i, j, k)) * AT(cosa, i, j, k)))) *
                          AT(rsina, i, j, k));
  }
}

/* ----------------------------------------------------------------- p_grad_c */
/* kernels/p_grad_c.h:80-101: delpc(i-1) and delpc(j-1) are read directly.
 * gz and pkc are FP64 islands: the bracket is double, rounded once to R. */
void ORC(p_grad_c)(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
                   const S3 *rdxc, const S3 *rdyc, const S3 *delpc, const oeb_s3 *gz, const oeb_s3 *pkc, R dt2) {
  FOR_INTERIOR(d) {
    const R wk0 = AT(delpc, i, j, k);
    const R wk_im = AT(delpc, i - 1, j, k);
    const R wk_jm = AT(delpc, i, j - 1, k);
    const double bu = ((AT(gz, i - 1, j, k + 1) - AT(gz, i, j, k)) * (AT(pkc, i, j, k + 1) - AT(pkc, i - 1, j, k))) +
                      ((AT(gz, i - 1, j, k) - AT(gz, i, j, k + 1)) * (AT(pkc, i - 1, j, k + 1) - AT(pkc, i, j, k)));
    const double bv = ((AT(gz, i, j - 1, k + 1) - AT(gz, i, j, k)) * (AT(pkc, i, j, k + 1) - AT(pkc, i, j - 1, k))) +
                      ((AT(gz, i, j - 1, k) - AT(gz, i, j, k + 1)) * (AT(pkc, i, j - 1, k + 1) - AT(pkc, i, j, k)));
    AT(uout, i, j, k) = (AT(uin, i, j, k) + (((dt2 * AT(rdxc, i, j, k)) / (wk_im + wk0)) * (R)bu));
    AT(vout, i, j, k) = (AT(vin, i, j, k) + (((dt2 * AT(rdyc, i, j, k)) / (wk_jm + wk0)) * (R)bv));
  }
}

/* ---------------------------------------------------------------- nh_p_grad */
/* kernels/nh_p_grad.h:116-146.  gz, pp and pk3 are FP64 islands: wk, the
 * thickness sums, du/dv and the pp brackets are double, rounded once to R. */
void ORC(nh_p_grad)(const oeb_domain *d, const S3 *uout, const S3 *vout, const S3 *uin, const S3 *vin,
                    const S3 *rdx, const S3 *rdy, const oeb_s3 *gz, const oeb_s3 *pp, const oeb_s3 *pk3,
                    const S3 *wk1, R dt) {
  const double dtd = dt;
  FOR_INTERIOR(d) {
    const double wk0 = (AT(pk3, i, j, k + 1) - AT(pk3, i, j, k));
    const double wk_i = (AT(pk3, i + 1, j, k + 1) - AT(pk3, i + 1, j, k));
    const double wk_j = (AT(pk3, i, j + 1, k + 1) - AT(pk3, i, j + 1, k));
    const double du =
        ((dtd / (wk0 + wk_i)) *
         (((AT(gz, i, j, k + 1) - AT(gz, i + 1, j, k)) * (AT(pk3, i + 1, j, k + 1) - AT(pk3, i, j, k))) +
          ((AT(gz, i, j, k) - AT(gz, i + 1, j, k + 1)) * (AT(pk3, i, j, k + 1) - AT(pk3, i + 1, j, k)))));
    const double bu = ((AT(gz, i, j, k + 1) - AT(gz, i + 1, j, k)) * (AT(pp, i + 1, j, k + 1) - AT(pp, i, j, k))) +
                      ((AT(gz, i, j, k) - AT(gz, i + 1, j, k + 1)) * (AT(pp, i, j, k + 1) - AT(pp, i + 1, j, k)));
    AT(uout, i, j, k) =
        (((AT(uin, i, j, k) + (R)du) + ((dt / (AT(wk1, i, j, k) + AT(wk1, i + 1, j, k))) * (R)bu)) * AT(rdx, i, j, k));
    const double dv =
        ((dtd / (wk0 + wk_j)) *
         (((AT(gz, i, j, k + 1) - AT(gz, i, j + 1, k)) * (AT(pk3, i, j + 1, k + 1) - AT(pk3, i, j, k))) +
          ((AT(gz, i, j, k) - AT(gz, i, j + 1, k + 1)) * (AT(pk3, i, j, k + 1) - AT(pk3, i, j + 1, k)))));
    const double bv = ((AT(gz, i, j, k + 1) - AT(gz, i, j + 1, k)) * (AT(pp, i, j + 1, k + 1) - AT(pp, i, j, k))) +
                      ((AT(gz, i, j, k) - AT(gz, i, j + 1, k + 1)) * (AT(pp, i, j, k + 1) - AT(pp, i, j + 1, k)));
    AT(vout, i, j, k) =
        (((AT(vin, i, j, k) + (R)dv) + ((dt / (AT(wk1, i, j, k) + AT(wk1, i, j + 1, k))) * (R)bv)) * AT(rdy, i, j, k));
  }
}

/* ------------------------------------------------------------------- fvtp2d */
/* A scratch array with the layout of `like`, owned by the caller of
 * scratch_new (free with scratch_free). */
static int LOC(scratch_new)(S3 *s, const S3 *like) {
  *s = *like;
  const size_t n = (size_t)(like->sizes[0] * like->strides[0]) + 64;
  s->allocatedPtr = (R *)calloc(n, sizeof(R));
  s->alignedPtr = s->allocatedPtr;
  return s->allocatedPtr == NULL;
}
static void LOC(scratch_free)(S3 *s) { free(s->allocatedPtr); }

/* One PPM flux stage.  Runs on each level k, in the stage-by-stage order of
 * kernels/fvtp2d.h.  The stage along j is :9-37 and :129-158, and the stage
 * along i is :53-82 and :84-113.  f is the advected field and c the Courant
 * number.  al is formed on [al_lo, al_hi) along the stage axis and the
 * limiter arrays on [lim_lo, lim_hi).  The flux is written on
 * [out_lo, out_hi).  The other axis spans [o_lo, o_hi). */
static void LOC(ppm_stage)(int along_j, const oeb_domain *d, const S3 *f, const S3 *c, const S3 *flux, const S3 *al,
                           const S3 *almq, const S3 *br, const S3 *b0, const S3 *smt5, int64_t o_lo, int64_t o_hi,
                           int64_t al_lo, int64_t al_hi, int64_t lim_lo, int64_t lim_hi, int64_t out_lo,
                           int64_t out_hi, R p1, R p2) {
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
        F(smt5, a, m) = ((F(almq, a, m) * F(br, a, m)) < 0) ? (R)1.0 : (R)0.0;
      }
    for (int64_t a = o_lo; a < o_hi; ++a)
      for (int64_t m = out_lo; m < out_hi; ++m) {
        const R smt5tmp = (F(smt5, a, m - 1) + (F(smt5, a, m) * (F(smt5, a, m - 1) == 0 ? (R)1.0 : (R)0.0)));
        const R cm = F(c, a, m);
        const R ctmp = ((cm > (R)0.) ? (((R)1. - cm) * (F(br, a, m - 1) - (cm * F(b0, a, m - 1))))
                                     : (((R)1. + cm) * (F(almq, a, m) + (cm * F(b0, a, m)))));
        F(flux, a, m) = ((cm > (R)0.) ? (F(f, a, m - 1) + (ctmp * smt5tmp)) : (F(f, a, m) + (ctmp * smt5tmp)));
      }
  }
#undef F
#undef JX
#undef IX
}

/* The flux divergence ((q*area + fa) - fb) / ra of kernels/fvtp2d.h:48 and
 * :124, with fa = xa*ya and fb = xb*yb (fyy = yfx*fy2, fxx = xfx*fx2, formed
 * at :43 and :119).  With R = double this is the reference's expression.  With
 * R = float the numerator is a compensated dot product (Ogita, Rump and
 * Oishi, "Accurate sum and dot product", SIAM J. Sci. Comput. 26(6), 2005,
 * algorithm Dot2).  fmaf gives each product's rounding error exactly
 * (TwoProduct), Knuth's TwoSum gives each sum's, and their total is added once
 * before the division. */
static R LOC(flux_div)(R q, R area, R xa, R ya, R xb, R yb, R ra) {
#if OEB_ORACLE_R_IS_FLOAT
  const float p0 = q * area, e0 = fmaf(q, area, -p0);
  const float p1 = xa * ya, e1 = fmaf(xa, ya, -p1);
  const float p2 = xb * yb, e2 = fmaf(xb, yb, -p2);
  const float s1 = p0 + p1, z1 = s1 - p0, t1 = (p0 - (s1 - z1)) + (p1 - z1);
  const float s2 = s1 - p2, z2 = s2 - s1, t2 = (s1 - (s2 - z2)) + (-p2 - z2);
  const float err = ((e0 + e1) - e2) + (t1 + t2);
  return (s2 + err) / ra;
#else
  const R fa = (xa * ya), fb = (xb * yb);
  return ((((q * area) + fa) - fb) / ra);
#endif
}

/* kernels/fvtp2d.h:4-177 (DEFAULT variant), stage by stage. */
int ORC(fvtp2d)(const oeb_domain *d, const S3 *q_i, const S3 *q_j, const S3 *fx1, const S3 *fx2, const S3 *fy1,
                const S3 *fy2, const S3 *q, const S3 *crx, const S3 *cry, const S3 *ra_x, const S3 *ra_y,
                const S3 *xfx, const S3 *yfx, const S3 *area, R p1, R p2) {
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
  S3 al, almq, br, b0, smt5;
  int fail = 0;
  fail |= LOC(scratch_new)(&al, q);
  fail |= LOC(scratch_new)(&almq, q);
  fail |= LOC(scratch_new)(&br, q);
  fail |= LOC(scratch_new)(&b0, q);
  fail |= LOC(scratch_new)(&smt5, q);
  if (!fail) {
    /* fy2 = PPM_j(q, cry): :9-37. */
    LOC(ppm_stage)(1, d, q, cry, fy2, &al, &almq, &br, &b0, &smt5, -3, nx + 2, -1, ny + 2, -1, ny + 1, 0, ny + 1, p1,
                   p2);
    /* q_i: :39-51. */
    for (int64_t k = 0; k < nz; ++k)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = -3; i < nx + 2; ++i)
          AT(q_i, i, j, k) = LOC(flux_div)(AT(q, i, j, k), AT(area, i, j, k), AT(yfx, i, j, k), AT(fy2, i, j, k),
                                           AT(yfx, i, j + 1, k), AT(fy2, i, j + 1, k), AT(ra_y, i, j, k));
    /* fx1 = PPM_i(q_i, crx): :53-82. */
    LOC(ppm_stage)(0, d, q_i, crx, fx1, &al, &almq, &br, &b0, &smt5, 0, ny, -1, nx + 1, -1, nx, 0, nx, p1, p2);
    /* fx2 = PPM_i(q, crx): :84-113. */
    LOC(ppm_stage)(0, d, q, crx, fx2, &al, &almq, &br, &b0, &smt5, -3, ny + 2, -1, nx + 2, -1, nx + 1, 0, nx + 1, p1,
                   p2);
    /* q_j: :115-127. */
    for (int64_t k = 0; k < nz; ++k)
      for (int64_t j = -3; j < ny + 2; ++j)
        for (int64_t i = 0; i < nx; ++i)
          AT(q_j, i, j, k) = LOC(flux_div)(AT(q, i, j, k), AT(area, i, j, k), AT(xfx, i, j, k), AT(fx2, i, j, k),
                                           AT(xfx, i + 1, j, k), AT(fx2, i + 1, j, k), AT(ra_x, i, j, k));
    /* fy1 = PPM_j(q_j, cry): :129-158. */
    LOC(ppm_stage)(1, d, q_j, cry, fy1, &al, &almq, &br, &b0, &smt5, 0, nx, -1, ny + 1, -1, ny, 0, ny, p1, p2);
    /* transport_flux: :160-176. */
    for (int64_t k = 0; k < nz; ++k)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
          AT(fx1, i, j, k) = (((R)0.5 * (AT(fx1, i, j, k) + AT(fx2, i, j, k))) * AT(xfx, i, j, k));
          AT(fy1, i, j, k) = (((R)0.5 * (AT(fy1, i, j, k) + AT(fy2, i, j, k))) * AT(yfx, i, j, k));
        }
  }
  LOC(scratch_free)(&al);
  LOC(scratch_free)(&almq);
  LOC(scratch_free)(&br);
  LOC(scratch_free)(&b0);
  LOC(scratch_free)(&smt5);
  return fail;
}

#undef EARTH_R_RECIP_R
