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
 * The reference is cited per function in oracle_body.h.  The semantics are those of
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

/* FP32 oracle: the same stencils with float arrays and scalars (the
 * reference's ElementType = float), except the FP64 island arrays (oeb_s3)
 * and uvbke's double dt5; oracle_body.h states the arithmetic. */
void oracle_laplace_f32(const oeb_domain *d, const oeb_s3f *in, const oeb_s3f *out);
void oracle_hdiffsa_f32(const oeb_domain *d, const oeb_s3f *out, const oeb_s3f *in, const oeb_s3f *mask,
                        const oeb_s1f *crlato, const oeb_s1f *crlatu);
void oracle_hdiffsmag_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                          const oeb_s3f *vin, const oeb_s3f *mask, const oeb_s1f *crlavo, const oeb_s1f *crlavu,
                          const oeb_s1f *crlato, const oeb_s1f *crlatu, const oeb_s1f *acrlat0, float eddlat,
                          float eddlon, float tau_smag, float weight_smag);
void oracle_hadvuv_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                       const oeb_s3f *vin, const oeb_s1f *acrlat0, const oeb_s1f *acrlat1, const oeb_s1f *tgrlatda0,
                       const oeb_s1f *tgrlatda1, float eddlat, float eddlon);
void oracle_hadvuv5th_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                          const oeb_s3f *vin, const oeb_s1f *acrlat0, const oeb_s1f *acrlat1,
                          const oeb_s1f *tgrlatda0, const oeb_s1f *tgrlatda1, float eddlat, float eddlon);
void oracle_fastwavesuv_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                            const oeb_s3f *vin, const oeb_s3f *utens, const oeb_s3f *vtens, const oeb_s3f *wgtfac,
                            const oeb_s3 *ppuv, const oeb_s3 *hhl, const oeb_s3f *rho, const oeb_s1f *fx,
                            float edadlat, float dt);
void oracle_uvbke_f32(const oeb_domain *d, const oeb_s3f *ub, const oeb_s3f *vb, const oeb_s3f *uc,
                      const oeb_s3f *vc, const oeb_s3f *cosa, const oeb_s3f *rsina, double dt5);
void oracle_p_grad_c_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                         const oeb_s3f *vin, const oeb_s3f *rdxc, const oeb_s3f *rdyc, const oeb_s3f *delpc,
                         const oeb_s3 *gz, const oeb_s3 *pkc, float dt2);
void oracle_nh_p_grad_f32(const oeb_domain *d, const oeb_s3f *uout, const oeb_s3f *vout, const oeb_s3f *uin,
                          const oeb_s3f *vin, const oeb_s3f *rdx, const oeb_s3f *rdy, const oeb_s3 *gz,
                          const oeb_s3 *pp, const oeb_s3 *pk3, const oeb_s3f *wk1, float dt);
int oracle_fvtp2d_f32(const oeb_domain *d, const oeb_s3f *q_i, const oeb_s3f *q_j, const oeb_s3f *fx1,
                      const oeb_s3f *fx2, const oeb_s3f *fy1, const oeb_s3f *fy2, const oeb_s3f *q,
                      const oeb_s3f *crx, const oeb_s3f *cry, const oeb_s3f *ra_x, const oeb_s3f *ra_y,
                      const oeb_s3f *xfx, const oeb_s3f *yfx, const oeb_s3f *area, float p1, float p2);

#ifdef __cplusplus
}
#endif

#endif /* OEB_ORACLE_H */
