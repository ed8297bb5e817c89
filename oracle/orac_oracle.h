/*
 * orac_oracle.h -- CPU oracle for the ORAC nonbonded force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C FP64 restatement of the
 * reference routines, used by the tests, by smoke() and by bench.py's
 * cpu_baseline leg as the checker.  The product library never links,
 * loads or calls it.
 *
 * Conventions follow the reference exactly: arrays are Fortran
 * column-major (co[i + 3*j] == co(i+1, j+1)), index arrays hold 1-based
 * Fortran indices, grids are Q(K1,K2,K3) with K1 fastest and the
 * half-complex grid is FQ(K1/2+1,K2,K3) stored as interleaved (re,im).
 * `nthreads` > 1 runs the OpenMP variant (group ownership and private
 * reduction arrays, as the reference's OpenMP build does); 1 runs the
 * serial loop order of the reference's non-OpenMP build.
 */
#ifndef ORAC_ORACLE_H
#define ORAC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* pme_fftw3.f90:165-230 load_bsp_moduli + DFTMOD */
void orc_load_bsp_moduli(int order, int nfft, double *bsp_mod);
/* pme_fftw3.f90:537-576 fill_bspline */
void orc_fill_bspline(double w, int order, double *array, double *darray);
/* pme_fftw3.f90:132-163 get_scaled_fractionals */
void orc_scaled_fractionals(int n, const double *x, const double *y, const double *z, const double *recip,
                            int k1, int k2, int k3, double *fr1, double *fr2, double *fr3);
/* pme_fftw3.f90:487-512 get_bspline_coeffs; theta(order, n) */
void orc_bspline_coeffs(int n, const double *fr1, const double *fr2, const double *fr3, int order,
                        double *th1, double *th2, double *th3, double *dth1, double *dth2, double *dth3);
/* pme_fftw3.f90:234-298 fill_charge_grid (FFTW layout Q(K1,K2,K3)) */
void orc_fill_charge_grid(int n, const double *q, const double *th1, const double *th2, const double *th3,
                          const double *fr1, const double *fr2, const double *fr3, int order,
                          int k1, int k2, int k3, double *Q, int nthreads);
/* unnormalised forward r2c / backward c2r, FFTW plan order (K3,K2,K1) */
void orc_fft_r2c(int k1, int k2, int k3, const double *Q, double *FQ, int nthreads);
void orc_fft_c2r(int k1, int k2, int k3, const double *FQ, double *Q, int nthreads);
/* 1-D complex DFT helper (sign -1 forward), exposed for tests */
void orc_fft1d(int n, double *data, int sign);
/* scalar_sum.f:1-135 (FFTW variant) */
void orc_scalar_sum(double *FQ, double ewaldcof, double volume, const double *recip, const double *bsp1,
                    const double *bsp2, const double *bsp3, int k1, int k2, int k3, double rkcut,
                    double *eer, double *vir, int nthreads);
/* pme_fftw3.f90:311-372 grad_sum */
void orc_grad_sum(int n, const double *q, const double *recip, const double *th1, const double *th2,
                  const double *th3, const double *dth1, const double *dth2, const double *dth3,
                  double *fx, double *fy, double *fz, const double *fr1, const double *fr2, const double *fr3,
                  int order, int k1, int k2, int k3, const double *Q, int nthreads);
/* fft_pme.f:1-115 around do_pmesh_kspace (pme_fftw3.f90:56-130).
 * variant 0: FFTW branch (r2c, half-complex scalar_sum, c2r) -- the
 * production build; variant 1: fftpack branch (full complex grid,
 * scalar_sum.f:137-217).  bsp moduli come from orc_load_bsp_moduli. */
int orc_fft_pme(int ntap, const double *xp, const double *yp, const double *zp, const double *xpcm,
                const double *ypcm, const double *zpcm, const double *pmechg, const double *co,
                const double *oc, double volume, double alphal, int order, int nfft1, int nfft2, int nfft3,
                double *eer, double *fpx, double *fpy, double *fpz, double *stressc, const int32_t *atomp,
                int32_t pressure, double rkcut, const double *bsp1, const double *bsp2, const double *bsp3,
                int nthreads, int variant);
/* scalar_sum.f:137-217 (fftpack variant) on a full complex grid Q(2,K1,K2,K3) */
void orc_scalar_sum_full(double *Qc, double ewaldcof, double volume, const double *recip, const double *bsp1,
                         const double *bsp2, const double *bsp3, int nfft1, int nfft2, int nfft3, double rkcut,
                         double *eer, double *vir);
/* Stage timers of the last orc_fft_pme call (seconds): spread, fft fwd, scalar_sum, fft bwd, gather */
void orc_last_pme_times(double *t5);

/* dist_ijk.f (function) and lc_index.f:3-125. Returns 0, or 1 when the
 * reference would STOP (indmax overflow or a cell visited twice). */
double orc_dist_ijk(int ni, int nj, int nk, double dx, double dy, double dz, const double *co);
int orc_lc_index(int indmax, int ncx, int ncy, int ncz, int32_t *nind, int32_t *indxi, int32_t *indxj,
                 int32_t *indxk, double ctoff, const double *co);
/* lc_list.f:1-237, non-OpenMP layout nnlpp = per group [nv, l1..lnv].
 * Returns total entries written, or -1 if capacity would be exceeded. */
long orc_lc_list(int ncx, int ncy, int ncz, const int32_t *nind, const int32_t *indxi, const int32_t *indxj,
                 const int32_t *indxk, double rcut, const double *co, const double *xp, const double *yp,
                 const double *zp, int natp, int32_t *nnlpp, long capacity);

/* force.f:1-917, non-OpenMP layouts (nnlpp, mapnl, nnlpp1 single lists).
 * rtolinn is passed by reference because the reference zeroes it when
 * rinn < 1e-5.  Returns 0, or 1 on nnlpp1 capacity overflow. */
int orc_force(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
              const double *xpg, const double *ypg, const double *zpg, const double *charge,
              const int32_t *nbtype, const int32_t *type, int ma, int nato, const int32_t *atomg,
              const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
              const double *co, const double *ecc12, const double *ecc6, double alphal, const int32_t *mapnl,
              int ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv, double *uconf_ss,
              double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx, double *fpy, double *fpz,
              double *stress, const int32_t *nnlpp, int32_t *nnlpp1, long nnlpp1_cap, double rneigh,
              double rinn, double rout, double *rtolinn, double rtolout, const double *gmass, int iz,
              int32_t *ncount1, const double *mass, double massinfty, int32_t lskip_ewald, int32_t ewald,
              int nthreads);
/* force_REM.f:1-995 (serial): orc_force with the segment pairs of
 * seg_select(rem_kind, rematom) (rem.f90:631-645) scaled by rem_f3 in force
 * and their energy summed into seg_ene; lskip_force / ljintra as
 * force_REM.f:148-158.  The routine's locals coul and rsp carry over from
 * pair to pair as in the reference (start values zero). */
int orc_force_rem(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                  const double *xpg, const double *ypg, const double *zpg, const double *charge,
                  const int32_t *nbtype, const int32_t *type, int ma, int nato, const int32_t *atomg,
                  const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
                  const double *co, const double *ecc12, const double *ecc6, double alphal, const int32_t *mapnl,
                  int ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv, double *uconf_ss,
                  double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx, double *fpy, double *fpz,
                  double *stress, const int32_t *nnlpp, int32_t *nnlpp1, long nnlpp1_cap, double rneigh,
                  double rinn, double rout, double *rtolinn, double rtolout, const double *gmass, int iz,
                  int32_t *ncount1, const double *mass, double massinfty, double *seg_ene,
                  const int32_t *rematom, int rem_kind, double rem_f3, int32_t lskip_force, int32_t ljintra,
                  int32_t lskip_ewald, int32_t ewald);
/* mts_forces.f:212-241 decide_force */
int32_t orc_decide_force(double rinn, double alphal);
/* ferrf.f:1-244 without the ERF_CORR spline branch. Returns 0, 1 if erf_corr set. */
int orc_ferrf(const int32_t *ss_index, double alphal, const double *charge, double fudge, const double *x0,
              const double *y0, const double *z0, int iz, const int32_t *list, int nlist, double *fs_slt,
              double *fs_slv, double *fpx, double *fpy, double *fpz, int32_t erf_corr, int nbonds_added,
              const double *erf_arr_corr, double delew, double rlew);
/* cself.f:1-64 */
void orc_cself(const int32_t *ss_index, int nato, double alphal, double kcut, const double *charge,
               double *self_slt, double *self_slv);
/* fnbgrp.f:1-213 (serial) */
void orc_fnbgrp(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                const double *charge, const int32_t *type, const double *ecc12, const double *ecc6,
                int32_t ewald, double alphal, const int32_t *ingrp, int ingrpp, double *fpx, double *fpy,
                double *fpz, double *uconf_slt, double *uconf_slv, double *ucoul_slt, double *ucoul_slv);


/* alchemy (force_alchemy.f, force_alchemy_shifted.f:732, ferrf_alchemy.f, cself_alchemy.f) */
void orc_alch_set(double *fact, double ladd, double lrm, int intramol, int ljintra, int qq);
int orc_force_alchemy(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                      const double *xpg, const double *ypg, const double *zpg, const double *charge,
                      const int32_t *ss_alchemy, const double *rmin, const double *epsm, int nato,
                      const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *atomp,
                      const double *co, double alphal, const int32_t *mapnl, int ngrp, const int32_t *grppt,
                      double *e6, double *fpx, double *fpy, double *fpz, double *stress, const int32_t *nnlpp,
                      int32_t *nnlpp1, long nnlpp1_cap, double rneigh, double rinn, double rout, double *rtolinn,
                      double rtolout, const double *gmass, int iz, int32_t *ncount1, const double *mass,
                      double massinfty, const double *lam, int intramol, int ljintra, int opls, double *dwrk,
                      int32_t lskip_ewald, int32_t ewald, double qsoft, double ljsoft, int shifted);
int orc_ferrf_alchemy(const int32_t *ss_index, double alphal, const double *charge, const double *charge0,
                      double fudge, const double *x0, const double *y0, const double *z0, int iz,
                      const int32_t *list, int nlist, double *fs_slt, double *fs_slv, double *fpx, double *fpy,
                      double *fpz, int nbonds_added, double *dwrk);
void orc_cself_alchemy(const int32_t *ss_index, int nato, double alphal, double kcut, const double *charge,
                       const double *charge0, double *self_slt, double *dwrk);
void orc_fnb14(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
               const double *charge, int nato, const double *ecc12, const double *ecc6, int cut, int ewald,
               double alphal, const int32_t *int14, int int14p, const int32_t *type14, double fudge, double lj_fudge,
               double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv, double *ucoul_slt,
               double *ucoul_slv);
int orc_fptors(const int32_t *ss_index, const int32_t *ltor, int ltors, const double *xp0, const double *yp0,
               const double *zp0, const double *pota, const double *potb, double *utors_slt, double *utors_slv,
               double *fpx, double *fpy, double *fpz);
int orc_fitors(const int32_t *ss_index, const int32_t *litr, int litor, const double *xp0, const double *yp0,
               const double *zp0, const double *pota, const double *potb, const double *potc, double *utors_slt,
               double *utors_slv, double *fpx, double *fpy, double *fpz);
/* REM twins (fnb14_rem.f, fptors_rem.f, fitors_rem.f, serial): per-term flags
 * (1 = in the segment) whose energy is returned in *seg_ene and whose force
 * is scaled by f2 = rem_factor(2) */
void orc_fnb14_rem(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                   const double *charge, int nato, const double *ecc12, const double *ecc6, int cut, int ewald,
                   double alphal, const int32_t *int14, int int14p, const int32_t *type14, double fudge,
                   double lj_fudge, double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv,
                   double *ucoul_slt, double *ucoul_slv, const int32_t *remint14, double f2, double *seg_ene);
int orc_fptors_rem(const int32_t *ss_index, const int32_t *ltor, int ltors, const double *xp0, const double *yp0,
                   const double *zp0, const double *pota, const double *potb, double *utors_slt, double *utors_slv,
                   double *fpx, double *fpy, double *fpz, const int32_t *remltor, double f2, double *seg_ene);
int orc_fitors_rem(const int32_t *ss_index, const int32_t *litr, int litor, const double *xp0, const double *yp0,
                   const double *zp0, const double *pota, const double *potb, const double *potc, double *utors_slt,
                   double *utors_slv, double *fpx, double *fpy, double *fpz, const int32_t *remitor, double f2,
                   double *seg_ene);
void orc_fpbond(const int32_t *ss_index, const int32_t *lbnd, int lbond, const double *xp0, const double *yp0,
                const double *zp0, const double *pota, const double *potb, double *ubond_slt, double *ubond_slv,
                double *fpx, double *fpy, double *fpz);
void orc_fpbend(const int32_t *ss_index, const int32_t *lbndg, int lbend, const double *xp0, const double *yp0,
                const double *zp0, const double *pota, const double *potb, const double *potc, const double *potd,
                const double *pote, const double *potf, const double *potg, double *ubend_slt, double *ubend_slv,
                double *fpx, double *fpy, double *fpz);
/* ferrf_alchemy_add.f:1-204 (list(3, nlist): ia, ib, facQ) */
int orc_ferrf_alchemy_add(const int32_t *ss_index, double alphal, const double *charge, const double *lambdaq,
                          const double *lambdaq0, const double *x0, const double *y0, const double *z0,
                          const int32_t *list, int nlist, double *fs_slt, double *fs_slv, double *fpx, double *fpy,
                          double *fpz, int32_t erf_corr, const double *erf_arr_corr, double delew, double rlew,
                          double *dwrk, int intramol);
/* fnb14_alchemy.f:1-361 (serial) */
void orc_fnb14_alchemy(const int32_t *ss_index, const double *xp0, const double *yp0, const double *zp0,
                       const double *charge, int nato, const double *ecc12, const double *ecc6, int cut, int ewald,
                       double alphal, const int32_t *int14, int int14p, const int32_t *type14, double fudge,
                       double lj_fudge, double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv,
                       double *ucoul_slt, double *ucoul_slv, const double *lambda, const double *lambdaq,
                       const double *lambda0, const double *lambdaq0, double *dwrk);
/* rattle_verlet.f:1-598 (SHAKE after a drift: xp1 corrected towards the
   constraints with displacements along the bonds at xp0, velocities by the
   displacement / dt) and rattle_correc.f:1-414 (RATTLE: velocities made
   orthogonal to the bonds at xp0).  cnst(2, ncnstr) 1-based; cnst_protl =
   per molecule [count, 1-based constraint indices...].  Return 0, or 1
   (no convergence / mim_lim beyond 20), 2 (a MIM molecule of more than 3
   constraints: LAPACK in the reference), 3 (a MIM molecule with two
   constraints on disjoint atoms). */
int orc_rattle_verlet(double dt, double *xp1, double *yp1, double *zp1, const double *xp0, const double *yp0,
                      const double *zp0, double *vpx, double *vpy, double *vpz, int nato, const int32_t *cnst,
                      const double *dss, const double *coeff, int ncnstr, const double *mass, int nprot,
                      const int32_t *cnst_protl, int mim_lim, double *dnit);
int orc_rattle_correc(double dt, const double *xp0, const double *yp0, const double *zp0, double *vpx, double *vpy,
                      double *vpz, int nato, const int32_t *cnst, const double *dss, const double *coeff, int ncnstr,
                      const double *mass, int nprot, const int32_t *cnst_protl, int mim_lim, double *dnit);

#ifdef __cplusplus
}
#endif
#endif
