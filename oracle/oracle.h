/*
 * oracle.h -- CPU restatement of the ten open-earth-benchmarks stencils.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is
 * compared against.  It is never linked into liboeb, and the product path never
 * calls it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.
 *
 * Each function follows one reference kernel expression by expression,
 * including the parenthesization and the left-to-right order of every sum.
 * The C compiler is run with -ffp-contract=off, so no FMA is introduced.
 * The reference is cited per function in oracle.c.  The semantics are those of
 * the reference's *_openmp variant, except fvtp2d, which has only the DEFAULT
 * variant.  Arrays use the descriptors of include/oeb.h, with HOST data
 * pointers.
 */
#ifndef OEB_ORACLE_H
#define OEB_ORACLE_H

#include "../include/oeb.h"

#ifdef __cplusplus
extern "C" {
#endif

void oracle_laplace(const oeb_domain *d, const oeb_s3 *in, const oeb_s3 *out);
void oracle_hdiffsa(const oeb_domain *d, const oeb_s3 *out, const oeb_s3 *in, const oeb_s3 *mask,
                    const oeb_s1 *crlato, const oeb_s1 *crlatu);
void oracle_hdiffsmag(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s3 *mask, const oeb_s1 *crlavo, const oeb_s1 *crlavu,
                      const oeb_s1 *crlato, const oeb_s1 *crlatu, const oeb_s1 *acrlat0, double eddlat,
                      double eddlon, double tau_smag, double weight_smag);
void oracle_hadvuv(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                   const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0,
                   const oeb_s1 *tgrlatda1, double eddlat, double eddlon);
void oracle_hadvuv5th(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s1 *acrlat0, const oeb_s1 *acrlat1, const oeb_s1 *tgrlatda0,
                      const oeb_s1 *tgrlatda1, double eddlat, double eddlon);
void oracle_fastwavesuv(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                        const oeb_s3 *vin, const oeb_s3 *utens, const oeb_s3 *vtens, const oeb_s3 *wgtfac,
                        const oeb_s3 *ppuv, const oeb_s3 *hhl, const oeb_s3 *rho, const oeb_s1 *fx,
                        double edadlat, double dt);
void oracle_uvbke(const oeb_domain *d, const oeb_s3 *ub, const oeb_s3 *vb, const oeb_s3 *uc, const oeb_s3 *vc,
                  const oeb_s3 *cosa, const oeb_s3 *rsina, double dt5);
void oracle_p_grad_c(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                     const oeb_s3 *vin, const oeb_s3 *rdxc, const oeb_s3 *rdyc, const oeb_s3 *delpc,
                     const oeb_s3 *gz, const oeb_s3 *pkc, double dt2);
void oracle_nh_p_grad(const oeb_domain *d, const oeb_s3 *uout, const oeb_s3 *vout, const oeb_s3 *uin,
                      const oeb_s3 *vin, const oeb_s3 *rdx, const oeb_s3 *rdy, const oeb_s3 *gz, const oeb_s3 *pp,
                      const oeb_s3 *pk3, const oeb_s3 *wk1, double dt);
/* Returns 0 on success, 1 if scratch allocation failed. */
int oracle_fvtp2d(const oeb_domain *d, const oeb_s3 *q_i, const oeb_s3 *q_j, const oeb_s3 *fx1,
                  const oeb_s3 *fx2, const oeb_s3 *fy1, const oeb_s3 *fy2, const oeb_s3 *q, const oeb_s3 *crx,
                  const oeb_s3 *cry, const oeb_s3 *ra_x, const oeb_s3 *ra_y, const oeb_s3 *xfx, const oeb_s3 *yfx,
                  const oeb_s3 *area, double p1, double p2);

#ifdef __cplusplus
}
#endif

#endif /* OEB_ORACLE_H */
