/*
 * orac_hip.h -- C ABI of the MI355X (gfx950) nonbonded force path for ORAC.
 *
 * Drop-in for ORAC's per-step nonbonded routines.  Each plain-mode entry
 * point takes exactly the arrays of the Fortran routine it replaces (cited
 * below as file:line of the reference source tree), with the same meaning,
 * layouts and in/out semantics:
 *   - arrays are Fortran column-major, index arrays hold 1-based indices;
 *   - forces and stress are accumulated (+=), energies overwritten;
 *   - neighbour lists and exclusion maps use the reference's serial
 *     (non-OpenMP) layouts: per group / per atom [count, entries...].
 * Scalar inputs are passed by value; scalar outputs by pointer.  LOGICAL
 * arguments are int32_t (0 = .false.).  Every function returns 0 on success
 * or a positive status code; orac_hip_last_error() gives the message.  No
 * C++ exception crosses this boundary.  A context is not re-entrant: one
 * host thread per context, one HIP stream per context.
 *
 * Status codes: 1 bad argument, 2 capacity exceeded (the reference's
 * xerror paths), 3 HIP runtime error, 4 unsupported input, 5 wrong state.
 */
#ifndef ORAC_HIP_H
#define ORAC_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orac_hip_ctx orac_hip_ctx;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart: holds the device, stream and cached topology. */
int orac_hip_create(orac_hip_ctx **ctx, int32_t device);
int orac_hip_destroy(orac_hip_ctx *ctx);
/* Copies the last error message of this host thread (NUL-terminated). */
int orac_hip_last_error(char *buf, int32_t len);
/* Drops every cached topology array so the next call uploads them again. */
int orac_hip_invalidate(orac_hip_ctx *ctx);
/* Name of the device the context runs on, e.g. "gfx950". */
int orac_hip_device_name(orac_hip_ctx *ctx, char *buf, int32_t len);
/* No reference counterpart (a measurement aid): one plain device copy kernel
 * (k_probe_copy) that reads nbytes and writes nbytes, so that profiler byte
 * counters (FETCH_SIZE, WRITE_SIZE) can be scaled against a known size.
 * Synchronous; allocates and frees its two buffers. */
int orac_hip_probe_copy(orac_hip_ctx *ctx, int64_t nbytes);

/* ---- reciprocal space ---------------------------------------------------- */
/* Replaces fft_pme_init (fftw3_pme_init.f90:2-62): B-spline moduli
 * (load_bsp_moduli/DFTMOD, pme_fftw3.f90:165-230, with the default-real pi
 * of DFTMOD) into bsp_mod1..3, and the device FFT plans for K1 x K2 x K3.
 * The FFTW plan handles of the reference have no counterpart. */
int orac_hip_fft_pme_init(orac_hip_ctx *ctx, int32_t numatoms, int32_t nfft1, int32_t nfft2, int32_t nfft3,
                          int32_t order, double *bsp_mod1, double *bsp_mod2, double *bsp_mod3);

/* Replaces fft_pme (fft_pme.f:1-115) with the FFTW do_pmesh_kspace
 * (pme_fftw3.f90:56-130) and scalar_sum (scalar_sum.f:1-135): eer is set,
 * fpx/fpy/fpz += reciprocal forces, and when pressure != 0
 * stressc += -vir + molecular correction.  grppt is accepted for signature
 * parity and not read (as in the reference). */
int orac_hip_fft_pme(orac_hip_ctx *ctx, int32_t ntap, const double *xp, const double *yp, const double *zp,
                     const double *xpcm, const double *ypcm, const double *zpcm, const double *pmechg,
                     const double *co, const double *oc, double volume, double alphal, int32_t order,
                     int32_t nfft1, int32_t nfft2, int32_t nfft3, double *eer, double *fpx, double *fpy, double *fpz,
                     double *stressc, const int32_t *atomp, const int32_t *grppt, int32_t pressure, double rkcut);
/* No counterpart in the reference: orac_hip_fft_pme with the same arguments
 * and results, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every per-call device buffer lives in a static 32 MiB region per GPU, so
 * the call allocates nothing; it needs fft_pme_init on the same context
 * first (set-up allocates the plan) and takes at most 65,536 atoms and
 * 64^3 grid points (status 2 beyond that).  Calls that overlap on one GPU
 * share the region and race; the region keeps the last call's contents. */
int orac_hip_fft_pme_noalloc(orac_hip_ctx *ctx, int32_t ntap, const double *xp, const double *yp, const double *zp,
                             const double *xpcm, const double *ypcm, const double *zpcm, const double *pmechg,
                             const double *co, const double *oc, double volume, double alphal, int32_t order,
                             int32_t nfft1, int32_t nfft2, int32_t nfft3, double *eer, double *fpx, double *fpy,
                             double *fpz, double *stressc, const int32_t *atomp, const int32_t *grppt,
                             int32_t pressure, double rkcut);

/* ---- linked cells --------------------------------------------------------- */
/* Replaces lc_index (lc_index.f:3-125) with dist_ijk (dist_ijk.f): the
 * cell-offset stencil, on the host.  indxi/j/k are (2, indmax).  Returns 2
 * where the reference STOPs (indmax overflow, or a cell visited twice). */
int orac_hip_lc_index(int32_t indmax, int32_t ncx, int32_t ncy, int32_t ncz, int32_t *nind, int32_t *indxi,
                      int32_t *indxj, int32_t *indxk, double ctoff, const double *co);

/* Replaces lc_list (lc_list.f:1-237): group-pair list of the group centres
 * xp/yp/zp (natp groups) within rcut, written as nnlpp = per group
 * [nv, l1..lnv] (l > n), the reference's serial layout.  *nentries gets the
 * total length.  Returns 2 if nnlpp_capacity would be exceeded (the
 * reference's xerror path, lc_list.f:191-197). */
int orac_hip_lc_list(orac_hip_ctx *ctx, int32_t ncx, int32_t ncy, int32_t ncz, const int32_t *nind,
                     const int32_t *indxi, const int32_t *indxj, const int32_t *indxk, double rcut, const double *co,
                     const double *xp, const double *yp, const double *zp, int32_t natp, int32_t *nnlpp,
                     int64_t nnlpp_capacity, int64_t *nentries);

/* ---- direct space ---------------------------------------------------------- */
/* Replaces force (force.f:1-917) as called from mts_forces.f:148-152:
 * group-switched Ewald-erfc (A&S 7.1.26) + Lennard-Jones forces, energies
 * per pair class (slt-slt, slv-slv, slt-slv) and molecular virial
 * (stress += st . co^T), plus the next-inner-shell list nnlpp1 (capacity
 * nnlpp1_capacity entries) and its length *ncount1.  rtolinn is in/out: the
 * reference sets it to 0 when rinn < 1e-5.  groupp is accepted for
 * signature parity and not read (as in the reference).  iz = 0 (or
 * rinn + rtolinn >= rout) is the list-only "phony" call. */
int orac_hip_force(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                   const double *zp0, const double *xpg, const double *ypg, const double *zpg, const double *charge,
                   const int32_t *nbtype, const int32_t *type, int32_t ma, int32_t nato, const int32_t *atomg,
                   const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *groupp,
                   const int32_t *atomp, const double *co, const double *ecc12, const double *ecc6, double alphal,
                   const int32_t *mapnl, int32_t ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv,
                   double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx,
                   double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1,
                   int64_t nnlpp1_capacity, double rneigh, double rinn, double rout, double *rtolinn,
                   double rtolout, const double *gmass, int32_t iz, int32_t *ncount1, const double *mass,
                   double massinfty, int32_t lskip_ewald, int32_t ewald);

/* Replaces force_rem (force_REM.f:1-995) as mts_forces.f:120-124 calls it
 * for a REM segment: arguments as orac_hip_force, plus seg_ene (out: the
 * energy of the selected pairs), rematom (per atom, nonzero in the segment)
 * and what force_REM reads from modules rem and cpropar: rem_segkind (0
 * either atom in the segment, 1 both, 2 exactly one; rem.f90:631-645),
 * rem_factor(3) (the selected pairs' force factor), lskip_force and ljintra
 * (force_REM.f:143-157: without ljintra nothing is computed, with it the
 * charges count as zero).  Only the Ewald branch (lskip_ewald 0, ewald 1)
 * is taken: the others read a distance or Coulomb energy that an earlier
 * pair left in the routine's locals (force_REM.f:491, :667), status 4. */
int orac_hip_force_rem(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                       const double *zp0, const double *xpg, const double *ypg, const double *zpg, const double *charge,
                       const int32_t *nbtype, const int32_t *type, int32_t ma, int32_t nato, const int32_t *atomg,
                       const double *xpcm, const double *ypcm, const double *zpcm, const int32_t *groupp,
                       const int32_t *atomp, const double *co, const double *ecc12, const double *ecc6, double alphal,
                       const int32_t *mapnl, int32_t ngrp, const int32_t *grppt, double *uconf_slt,
                       double *uconf_slv, double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss,
                       double *fpx, double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1,
                       int64_t nnlpp1_capacity, double rneigh, double rinn, double rout, double *rtolinn,
                       double rtolout, const double *gmass, int32_t iz, int32_t *ncount1, const double *mass,
                       double massinfty, double *seg_ene, const int32_t *rematom, int32_t lskip_ewald, int32_t ewald,
                       int32_t rem_segkind, double rem_factor3, int32_t lskip_force, int32_t ljintra);

/* Replaces force_alchemy (force_alchemy.f:1-759, the soft-core GROMOS form)
 * as mts_forces.f calls it for an alchemical run.  Arguments as
 * orac_hip_force, plus what force_alchemy reads from module cpropar and its
 * own argument list: ss_alchemy (per atom 1 normal, 2 added, -2 removed),
 * lam = {ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0}, intramol,
 * ljintra, opls (ffield_type 'opls'), per-atom rmin and epsm, qsoft and
 * ljsoft.  dwrk is in/out: += U(lambda) - U(lambda0) over the shell, as the
 * serial reference does.  The per-atom lambda arrays of the Fortran
 * signature are not read by force_alchemy and are not taken here. */
int orac_hip_force_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                           const double *zp0, const double *xpg, const double *ypg, const double *zpg,
                           const double *charge, const int32_t *nbtype, const int32_t *type, int32_t ma,
                           int32_t nato, const int32_t *atomg, const double *xpcm, const double *ypcm,
                           const double *zpcm, const int32_t *groupp, const int32_t *atomp, const double *co,
                           const double *ecc12, const double *ecc6, double alphal, const int32_t *mapnl,
                           int32_t ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv,
                           double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss, double *fpx,
                           double *fpy, double *fpz, double *stress, const int32_t *nnlpp, int32_t *nnlpp1,
                           int64_t nnlpp1_capacity, double rneigh, double rinn, double rout, double *rtolinn,
                           double rtolout, const double *gmass, int32_t iz, int32_t *ncount1, const double *mass,
                           double massinfty, const int32_t *ss_alchemy, const double *lam, int32_t intramol,
                           int32_t ljintra, int32_t opls, const double *rmin, const double *epsm, double *dwrk,
                           int32_t lskip_ewald, int32_t ewald, double qsoft, double ljsoft);

/* Replaces force_alchemy_shift (force_alchemy_shifted.f:1-730), the
 * shifted-potential form mts_forces.f:130 selects with SHIFT-POT: arguments
 * as orac_hip_force_alchemy with qshift and ljshift in place of qsoft and
 * ljsoft. */
int orac_hip_force_alchemy_shift(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                                 const double *zp0, const double *xpg, const double *ypg, const double *zpg,
                                 const double *charge, const int32_t *nbtype, const int32_t *type, int32_t ma,
                                 int32_t nato, const int32_t *atomg, const double *xpcm, const double *ypcm,
                                 const double *zpcm, const int32_t *groupp, const int32_t *atomp, const double *co,
                                 const double *ecc12, const double *ecc6, double alphal, const int32_t *mapnl,
                                 int32_t ngrp, const int32_t *grppt, double *uconf_slt, double *uconf_slv,
                                 double *uconf_ss, double *ucoul_slt, double *ucoul_slv, double *ucoul_ss,
                                 double *fpx, double *fpy, double *fpz, double *stress, const int32_t *nnlpp,
                                 int32_t *nnlpp1, int64_t nnlpp1_capacity, double rneigh, double rinn, double rout,
                                 double *rtolinn, double rtolout, const double *gmass, int32_t iz, int32_t *ncount1,
                                 const double *mass, double massinfty, const int32_t *ss_alchemy, const double *lam,
                                 int32_t intramol, int32_t ljintra, int32_t opls, const double *rmin,
                                 const double *epsm, double *dwrk, int32_t lskip_ewald, int32_t ewald,
                                 double qshift, double ljshift);

/* Replaces decide_force (mts_forces.f:212-241): lskip_ewald = erfc_AS(alpha*rinn) < 1e-3. */
int orac_hip_decide_force(double rinn, double alphal, int32_t *lskip_ewald);

/* ---- corrections ------------------------------------------------------------ */
/* Replaces ferrf (ferrf.f:1-244): reciprocal-space erf correction of the
 * excluded pairs list(2, nlist) on cartesian x0/y0/z0; iz = 1 also adds
 * forces (fudge-scaled), iz = 0 gives the energy only.  With erf_corr the
 * reciprocal-cutoff correction is added from the spline table
 * erf_arr_corr(4, *) on r = rlew + (i - 1) delew (ferrf.f:68-105,
 * 170-193), as int_corr_erf_spline builds it; a pair below rlew returns 1. */
int orac_hip_ferrf(orac_hip_ctx *ctx, const int32_t *ss_index, double alphal, const double *charge, double fudge,
                   const double *x0, const double *y0, const double *z0, int32_t iz, const int32_t *list,
                   int32_t nlist, double *fscnstr_slt, double *fscnstr_slv, double *fpx, double *fpy, double *fpz,
                   int32_t erf_corr, const double *erf_arr_corr, double delew, double rlew, int32_t nbonds_added,
                   int32_t mb);
/* ERF_CORR table builders, called once at set-up (mtsmd.f:1595-1606):
 * int_corr_erf_spline (int_corr_erf_spline.f:1-89, with cubspl of
 * erfc_spline.f:71-193) fills erf_arr_corr(4, nbin + 1) and tau(nbin + 1)
 * for r in [rmin, rmax]; erf_corr_cutoff (erf_corr_cutoff.f:1-54) sets
 * rkcut from the cell and the grid (delew is accepted and not read). */
int orac_hip_int_corr_erf_spline(double rmin, double rmax, int32_t nbin, double alpha, double kcut,
                                 double *erf_arr_corr, double *tau);
int orac_hip_erf_corr_cutoff(const double *oc, double *delew, double *rkcut, int32_t nfft1, int32_t nfft2,
                             int32_t nfft3);

/* Replaces cself (cself.f:1-64): Ewald self energy per class. */
int orac_hip_cself(orac_hip_ctx *ctx, const int32_t *ss_index, int32_t nato, double alphal, double kcut,
                   const double *charge, double *self_slt, double *self_slv);

/* Replace fft_pme_alchemy (fft_pme_alchemy.f:1-117, do_pmesh_alchemy.f90),
 * ferrf_alchemy (ferrf_alchemy.f:1-204, without ERF_CORR) and cself_alchemy
 * (cself_alchemy.f:1-66) as mts_furier_alchemy.f:120-185 calls them: the
 * routines of orac_hip_fft_pme / _ferrf / _cself on `charge`, plus the
 * energy on the lambda0 charges charg0 (a second, energy-only grid for the
 * reciprocal sum), returned as dwrk = U(charge) - U(charg0).  cself_alchemy
 * sums solute atoms only and returns no solvent term, as the reference. */
int orac_hip_fft_pme_alchemy(orac_hip_ctx *ctx, int32_t ntap, const double *xp, const double *yp, const double *zp,
                             const double *xpcm, const double *ypcm, const double *zpcm, const double *pmechg,
                             const double *charg0, const double *co, const double *oc, double volume, double alphal,
                             int32_t order, int32_t nfft1, int32_t nfft2, int32_t nfft3, double *eer, double *fpx,
                             double *fpy, double *fpz, double *stressc, const int32_t *atomp, const int32_t *grppt,
                             int32_t pressure, double rkcut, double *dwrk);
int orac_hip_ferrf_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, double alphal, const double *charge,
                           const double *charge0, double fudge, const double *x0, const double *y0, const double *z0,
                           int32_t iz, const int32_t *list, int32_t nlist, double *fscnstr_slt, double *fscnstr_slv,
                           double *fpx, double *fpy, double *fpz, int32_t erf_corr, int32_t nbonds_added,
                           double *dwrk, int32_t mb);
int orac_hip_cself_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, int32_t nato, double alphal, double kcut,
                           const double *charge, const double *charge0, double *self_slt, double *dwrk);

/* Replaces fptors (fptors.f:1-345) as mts_intra_n1.f:118-119 calls it:
 * proper torsions ltor(4, ltors) on the coordinates as given, pota the
 * signed multiplicity (1-8, 12, 14) and potb the barrier; mb atoms.  Forces
 * accumulate, energies (by the ss_index of each torsion's first atom) are
 * overwritten.  A torsion with collinear bond angles returns 5 (the
 * reference perturbs its atoms at random); a multiplicity outside the
 * reference's table returns 4 (the reference stops). */
int orac_hip_fptors(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *ltor, int32_t ltors, const double *xp0,
                    const double *yp0, const double *zp0, const double *pota, const double *potb, double *utors_slt,
                    double *utors_slv, double *fpx, double *fpy, double *fpz, int32_t mb);

/* Replaces fitors (fitors.f:1-286) as mts_intra_n0.f:97-99 calls it:
 * improper torsions litr(4, litor) on the coordinates as given; potc > 0
 * selects the harmonic form potb (phi - pota)^2 on the signed dihedral,
 * otherwise pota is the signed multiplicity (1, 2, 3, 4 or 6) and potb the
 * barrier.  Forces accumulate, energies are overwritten.  Another
 * multiplicity returns 4 (the reference stops). */
int orac_hip_fitors(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *litr, int32_t litor, const double *xp0,
                    const double *yp0, const double *zp0, const double *pota, const double *potb, const double *potc,
                    double *utors_slt, double *utors_slv, double *fpx, double *fpy, double *fpz, int32_t mb);

/* Replace the REM twins as mts_intra_n1.f:74-78, :112-116 and
 * mts_intra_n0.f:91-95 call them for a REM segment: fptors_rem
 * (fptors_rem.f:1-352), fitors_rem (fitors_rem.f:1-301) and fnb14_rem
 * (fnb14_rem.f:1-234).  Arguments as orac_hip_fptors, orac_hip_fitors and
 * orac_hip_fnb14, plus the per-term flags (remltor, remitor, remint14:
 * nonzero in the segment), seg_ene (out: the flagged terms' energy) and
 * rem_factor(2) (module rem), which scales the flagged terms' forces.
 * fptors_rem evaluates the 60-degree-phase multiplicities 12 and 14 as the
 * REM routine does (no sign factor; no sine-derivative term where the sine
 * is below 1e-8).  With no 1-4 pair, fnb14_rem leaves seg_ene as it was
 * (fnb14_rem.f:100). */
int orac_hip_fptors_rem(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *ltor, int32_t ltors,
                        const double *xp0, const double *yp0, const double *zp0, const double *pota, const double *potb,
                        double *utors_slt, double *utors_slv, double *fpx, double *fpy, double *fpz,
                        const int32_t *remltor, double *seg_ene, double rem_factor2, int32_t mb);
int orac_hip_fitors_rem(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *litr, int32_t litor,
                        const double *xp0, const double *yp0, const double *zp0, const double *pota, const double *potb,
                        const double *potc, double *utors_slt, double *utors_slv, double *fpx, double *fpy,
                        double *fpz, const int32_t *remitor, double *seg_ene, double rem_factor2, int32_t mb);
int orac_hip_fnb14_rem(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                       const double *zp0, const double *charge, int32_t nato, const double *ecc12, const double *ecc6,
                       int32_t cut, int32_t ewald, double alphal, const int32_t *int14, int32_t int14p,
                       const int32_t *type14, double fudge, double lj_fudge, double *fpx, double *fpy, double *fpz,
                       double *uconf_slt, double *uconf_slv, double *ucoul_slt, double *ucoul_slv,
                       const int32_t *remint14, double *seg_ene, double rem_factor2);

/* Replaces fpbond (fpbond.f:1-110) as mts_intra_n0.f:73-74 calls it:
 * harmonic bonds lbnd(2, lbond) on the coordinates as given, pota the rest
 * length (potbo(1,2)) and potb the constant (potbo(1,1)); mb atoms.  Forces
 * accumulate, energies (by the ss_index of each bond's first atom) are
 * overwritten.  The reference's icnstr and itask select OpenMP work and have
 * no counterpart. */
int orac_hip_fpbond(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *lbnd, int32_t lbond, const double *xp0,
                    const double *yp0, const double *zp0, const double *pota, const double *potb, double *ubond_slt,
                    double *ubond_slv, double *fpx, double *fpy, double *fpz, int32_t mb);

/* Replaces fpbend (fpbend.f:1-228) as mts_intra_n0.f:82-85 calls it: bends
 * lbndg(3, lbend) (vertex second), pota the rest angle in radians, potb its
 * constant, potc/potd the Urey-Bradley rest distance and constant, and
 * pote/potf/potg the stretch-stretch coupling, with the reference's rules for
 * which of these terms apply; a rest angle within 0.01 rad of pi uses the
 * 2 potb (cos + 1) form.  Forces accumulate, energies are overwritten. */
int orac_hip_fpbend(orac_hip_ctx *ctx, const int32_t *ss_index, const int32_t *lbndg, int32_t lbend, const double *xp0,
                    const double *yp0, const double *zp0, const double *pota, const double *potb, const double *potc,
                    const double *potd, const double *pote, const double *potf, const double *potg,
                    double *ubend_slt, double *ubend_slv, double *fpx, double *fpy, double *fpz, int32_t mb);

/* Replaces fnb14 (fnb14.f:1-219): the 1-4 pairs int14(2, int14p) on the
 * coordinates as given (no PBC), LJ with the packed index type14(i) scaled
 * by lj_fudge, Coulomb scaled by fudge, erfc (A&S 7.1.26) when ewald and
 * cut are both set and plain Coulomb otherwise.  Forces accumulate,
 * energies (by the ss_index of the pair's first atom) are overwritten. */
int orac_hip_fnb14(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                   const double *zp0, const double *charge, int32_t nato, const double *ecc12, const double *ecc6,
                   int32_t cut, int32_t ewald, double alphal, const int32_t *int14, int32_t int14p,
                   const int32_t *type14, double fudge, double lj_fudge, double *fpx, double *fpy, double *fpz,
                   double *uconf_slt, double *uconf_slv, double *ucoul_slt, double *ucoul_slv);

/* Replaces fnbgrp (fnbgrp.f:1-213), serial layout: intra-group
 * non-excluded pairs ingrp(2, ingrpp), LJ + Ewald (or plain Coulomb). */
int orac_hip_fnbgrp(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                    const double *zp0, const double *charge, const int32_t *type, const double *ecc12,
                    const double *ecc6, int32_t ewald, double alphal, const int32_t *ingrp, int32_t ingrpp,
                    double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv, double *ucoul_slt,
                    double *ucoul_slv, int32_t nato);

/* Replaces ferrf_alchemy_add (ferrf_alchemy_add.f:1-204) as mtsmd.f:3034-3040
 * calls it on every M step when Ewald and alchemy are both on: the
 * intra-alchemical pairs beyond 1-4, list(3, nlist) = (ia, ib, facQ) as
 * initialize_lambda.f builds it, on cartesian x0/y0/z0 and the created
 * charges.  Each pair's erf(alpha r) q_a q_b / r term is scaled by
 * factl = (1 - |lambdaq_a|)(1 - |lambdaq_b|) - facQ (and by the lambdaq0
 * factor for en0); with intramol the facQ = 1 pairs are skipped; with
 * erf_corr the spline correction of erf_arr_corr is included as in ferrf.
 * Forces accumulate; fscnstr_slt/slv (by the ss_index of each pair's first
 * atom) and dwrk = slt + slv - en0 are overwritten; nato atoms. */
int orac_hip_ferrf_alchemy_add(orac_hip_ctx *ctx, const int32_t *ss_index, double alphal, const double *charge,
                               const double *lambdaq, const double *lambdaq0, const double *x0, const double *y0,
                               const double *z0, const int32_t *list, int32_t nlist, double *fscnstr_slt,
                               double *fscnstr_slv, double *fpx, double *fpy, double *fpz, int32_t erf_corr,
                               const double *erf_arr_corr, double delew, double rlew, double *dwrk, int32_t nato,
                               int32_t intramol);

/* Replaces fnb14_alchemy (fnb14_alchemy.f:1-361) as mts_intra_n1.f:79-83
 * calls it: fnb14's 1-4 pairs with the alchemical factors of each pair's
 * per-atom lambda, lambdaq (and lambda0, lambdaq0 for dwrk), keeping the
 * reference's own arithmetic (without Ewald the forces are unscaled; with
 * Ewald only the repulsive LJ force term is scaled, and pairs of two
 * alchemical atoms add an erf term).  Forces accumulate; energies and
 * dwrk = U(lambda) - U(lambda0) are overwritten (the itask argument selects
 * OpenMP work in the reference and has no counterpart). */
int orac_hip_fnb14_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                           const double *zp0, const double *charge, int32_t nato, const double *ecc12,
                           const double *ecc6, int32_t cut, int32_t ewald, double alphal, const int32_t *int14,
                           int32_t int14p, const int32_t *type14, double fudge, double lj_fudge, double *fpx,
                           double *fpy, double *fpz, double *uconf_slt, double *uconf_slv, double *ucoul_slt,
                           double *ucoul_slv, const double *lambda, const double *lambdaq, const double *lambda0,
                           const double *lambdaq0, double *dwrk);

/* Replaces fnbgrp_alchemy (fnbgrp_alchemy.f:1-220) as mts_intra_n1.f:97-100
 * calls it: the arithmetic is fnbgrp's (the reference derives alphal0 from
 * lambdaq and never reads it), so lambdaq is accepted and not read; nato is
 * the number of atoms the arrays hold. */
int orac_hip_fnbgrp_alchemy(orac_hip_ctx *ctx, const int32_t *ss_index, const double *xp0, const double *yp0,
                            const double *zp0, const double *charge, const int32_t *type, const double *ecc12,
                            const double *ecc6, int32_t ewald, double alphal, const int32_t *ingrp, int32_t ingrpp,
                            double *fpx, double *fpy, double *fpz, double *uconf_slt, double *uconf_slv,
                            double *ucoul_slt, double *ucoul_slv, const double *lambdaq, int32_t nato);

/* ---- device-resident systems (no reference counterpart) -------------------- */
/* A system whose topology, coordinates, lists and grids stay in HBM.  The
 * caller describes it once with the same arrays the plain-mode routines
 * take, then evaluates the full nonbonded force (lists when due, direct
 * space, PME, exclusion corrections) on the device without host copies.
 * precision: 0 = FP64 throughout; 1 = mixed (FP32 pair/spline/FFT
 * arithmetic, FP64 influence function, corrections and reductions). */
typedef struct orac_hip_system_desc {
  int32_t natoms, ngrp, nmol, ntypes;
  const double *charge, *mass;        /* natoms */
  const int32_t *nbtype;              /* natoms, 1-based */
  const int32_t *type_table;          /* ntypes x ntypes, 1-based packed index */
  const double *ecc12, *ecc6;         /* ntypes*(ntypes+1)/2 */
  const int32_t *grppt;               /* (2, ngrp), 1-based atom ranges */
  const int32_t *atomp;               /* natoms, 1-based molecule; molecules are contiguous */
  const int32_t *ss_index;            /* natoms, 1 solute / 2 solvent */
  const int32_t *mapnl;               /* exclusion map, mappa.f layout */
  int32_t ncorr;                      /* per-step reciprocal erf corrections (ferrf, iz = 1) */
  const int32_t *corr_pairs;          /* (2, ncorr), 1-based */
  const double *corr_fudge;           /* ncorr */
  const double *co;                   /* 3x3 half cell matrix */
  double alpha;
  int32_t nfft1, nfft2, nfft3, order;
  double rcut, rtol, rnei;            /* direct shell: cutoff, healing length, list skin */
  int32_t ncx, ncy, ncz;              /* linked-cell grid */
  int32_t precision;
} orac_hip_system_desc;

int orac_hip_system_create(orac_hip_ctx *ctx, const orac_hip_system_desc *desc, int32_t *sys);
int orac_hip_system_destroy(orac_hip_ctx *ctx, int32_t sys);
/* Host -> device reduced coordinates. */
int orac_hip_system_set_coords(orac_hip_ctx *ctx, int32_t sys, const double *xp, const double *yp, const double *zp);
/* Device pointers of the resident coordinate and force arrays (for
 * callers that keep the state on the device). */
int orac_hip_system_device_ptrs(orac_hip_ctx *ctx, int32_t sys, double **xp, double **yp, double **zp, double **fx,
                                double **fy, double **fz);
/* One full force evaluation on the device.  rebuild_list != 0 rebuilds the
 * group lists first.  Asynchronous on the context stream. */
int orac_hip_system_eval(orac_hip_ctx *ctx, int32_t sys, int32_t rebuild_list);
/* Waits for the context stream, copies forces (may be null) and the
 * energy record: [0] eer, [1..3] uconf slt/slv/ss, [4..6] ucoul slt/slv/ss,
 * [7..8] correction slt/slv, [9..17] stress (3x3, direct + reciprocal +
 * molecular correction), [18] group-pair entries of the last list build
 * (full list, both directions), [19..20] self energy slt/slv.  Needs room
 * for 21 doubles.  Returns 2 if the list outgrew its capacity. */
int orac_hip_system_results(orac_hip_ctx *ctx, int32_t sys, double *fx, double *fy, double *fz, double *energies);
/* Stage timing: when enabled, eval records HIP events between its stages;
 * stage_ms returns, for the last eval, the milliseconds of
 * [0] group and molecule centres, [1] list build, [2] direct-space pairs,
 * [3] PME (spread, FFT, convolution, gather), [4] corrections,
 * [5] the pair kernel alone (within [2]), and within [3] [6] the spread
 * chain (binning, B-splines, tiles), [7] the FFT stage (both transforms and
 * the convolution) and [8] the gather. */
int orac_hip_system_profile(orac_hip_ctx *ctx, int32_t sys, int32_t enable);
int orac_hip_system_stage_ms(orac_hip_ctx *ctx, int32_t sys, float *ms, int32_t n);
/* Device bytes held by a system (for admission against an HBM budget). */
int orac_hip_system_bytes(orac_hip_ctx *ctx, int32_t sys, int64_t *bytes);
int orac_hip_sync(orac_hip_ctx *ctx);
/* Stream handle of the context, as a void* (a hipStream_t). */
int orac_hip_stream(orac_hip_ctx *ctx, void **stream);

/* ---- alchemy on a resident system ------------------------------------------ */
/* The alchemical force step of the reference (mts_furier_alchemy.f:80-190
 * with force_alchemy.f in the direct shell) on a resident system:
 * soft-core pairs on the created charges scaled by the charge factors;
 * reciprocal sum, ferrf corrections and self term on the lambda charges
 * (1 - |lambdaq|) q (mts_furier_alchemy.f:98-111); and the lambda0 charges
 * on a second, energy-only grid (do_pmesh_alchemy.f90), in the corrections
 * and in the self term, for dwrk = U(lambda) - U(lambda0).  ss_alchemy,
 * rmin, epsm, qsoft, ljsoft, intramol, ljintra and opls as in
 * orac_hip_force_alchemy.  listqq adds ferrf_alchemy_add's intra-alchemical
 * correction to every evaluation (mtsmd.f:3034-3040; its energy joins the
 * correction energies and its dwrk the corrections' dwrk).  Once set the system stays alchemical; its
 * energies record reports the self energy of the solute only
 * (cself_alchemy.f), slv 0. */
typedef struct orac_hip_alchemy_desc {
  const int32_t *ss_alchemy;          /* natoms: 1 normal, 2 added, -2 removed */
  const double *rmin, *epsm;          /* natoms */
  double qsoft, ljsoft;
  int32_t intramol, ljintra, opls;
  int32_t shifted;                    /* 1: shifted potential (force_alchemy_shifted.f) with qshift, ljshift */
  double qshift, ljshift;
  int32_t nqq;                        /* intra-alchemical pairs of ferrf_alchemy_add, 0 for none */
  const int32_t *listqq;              /* (3, nqq): ia, ib (1-based), facQ (initialize_lambda.f:66-140) */
} orac_hip_alchemy_desc;
int orac_hip_system_set_alchemy(orac_hip_ctx *ctx, int32_t sys, const orac_hip_alchemy_desc *desc);
/* lam = ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0 (initially 1, 0, 1, 0, 1, 0, 1, 0). */
int orac_hip_system_set_lambda(orac_hip_ctx *ctx, int32_t sys, const double *lam);
/* After an eval: out[0] dwrk = U(lambda) - U(lambda0) of the force step,
 * [1] direct pairs, [2] reciprocal eer - eer0, [3] corrections, [4] self,
 * [5] eer0.  Needs room for 6 doubles. */
int orac_hip_system_alchemy_results(orac_hip_ctx *ctx, int32_t sys, double *out);

/* ---- batches (no reference counterpart) ------------------------------------ */
/* An ordered set of systems (replicas, windows, requests; sizes may differ)
 * evaluated together: each stage is one kernel launch for the whole batch.
 * All systems of a batch have the same precision.  A system's forces are the
 * same alone and batched; energies may differ in their last bits when the
 * batch's FFT tiling differs from the system's own.  Destroying a system
 * destroys the batches that hold it. */
int orac_hip_batch_create(orac_hip_ctx *ctx, int32_t nsys, const int32_t *systems, int32_t *batch);
int orac_hip_batch_destroy(orac_hip_ctx *ctx, int32_t batch);
/* One force evaluation of every system.  use_graph != 0 replays a HIP graph
 * of the step, captured on first use (one graph with and one without the
 * list build).  Asynchronous; results per system via system_results. */
int orac_hip_batch_eval(orac_hip_ctx *ctx, int32_t batch, int32_t rebuild_list, int32_t use_graph);
int orac_hip_batch_profile(orac_hip_ctx *ctx, int32_t batch, int32_t enable);
int orac_hip_batch_stage_ms(orac_hip_ctx *ctx, int32_t batch, float *ms, int32_t n);
/* Tuning vector: [0] LDS elements per FFT buffer (power of two, 64..2048),
 * [1] pair-kernel occupancy target in waves/SIMD (0 default, 2, 3, 4),
 * [2] LDS-tiled spread with the largest tiles (1), the smallest tiles (2),
 * or global atomics (0), [3] XCD-grouped pair
 * blocks (1) or hardware order (0), [4] pair-kernel owners and neighbour
 * records in linked-cell order (1) or group order (0), [5] pair kernel as
 * one persistent grid walking all work items of the batch (1) or one
 * workgroup per item (0).  No choice changes forces; [0] and [4] change
 * energies only in their last bits. */
int orac_hip_batch_set_tuning(orac_hip_ctx *ctx, int32_t batch, const int32_t *tuning, int32_t n);
int orac_hip_batch_get_tuning(orac_hip_ctx *ctx, int32_t batch, int32_t *tuning, int32_t n);
/* Times the batch's step for the tuning candidates (reps evaluations each)
 * and keeps the fastest; best_ms (may be null) receives its time. */
int orac_hip_batch_autotune(orac_hip_ctx *ctx, int32_t batch, int32_t reps, int32_t rebuild_list, double *best_ms);
/* Every candidate the batch's last autotune timed, in order (the starting
 * tuning first): tunings[k*6 .. k*6+5] in the order of
 * orac_hip_batch_get_tuning and ms[k] its time per step, for k < min(cap, *n);
 * *n receives the number of candidates. */
int orac_hip_batch_tune_log(orac_hip_ctx *ctx, int32_t batch, int32_t *tunings, double *ms, int32_t cap, int32_t *n);

/* ---- dynamics (no reference counterpart: the caller side of the path) ------ */
/* On-device r-RESPA NVE integration in the scheme of ORAC's multiple-time-step
 * integrator (mtsmd.f): a fast level of harmonic distance terms
 * E = k (r - r0)^2 (terms (2, nterm), 1-based; k in energy / A^2) integrated
 * with ninner substeps, and the full nonbonded force of the system once per
 * outer step.  Rigid units are 3 consecutive atoms (rigid_first, 1-based)
 * whose distances (0,1), (0,2), (1,2) are kept at rigid_len (3, nrigid) by
 * SHAKE after every drift and RATTLE after every kick (tolerance 1e-12).
 * Units: positions A (Cartesian), time fs, velocities A/fs, masses amu,
 * energies in the system's internal units (amu A^2 / fs^2).  Setup takes the
 * positions from the system's reduced coordinates and zeroes velocities. */
int orac_hip_md_setup(orac_hip_ctx *ctx, int32_t sys, int32_t nrigid, const int32_t *rigid_first,
                      const double *rigid_len, int32_t nterm, const int32_t *terms, const double *term_k,
                      const double *term_r0);
/* Cartesian positions and/or velocities (either triple may be null);
 * velocities are projected onto the rigid-unit constraints. */
int orac_hip_md_set_state(orac_hip_ctx *ctx, int32_t sys, const double *x, const double *y, const double *z,
                          const double *vx, const double *vy, const double *vz);
int orac_hip_md_get_state(orac_hip_ctx *ctx, int32_t sys, double *x, double *y, double *z, double *vx, double *vy,
                          double *vz);
/* nsteps outer steps of dt fs for every system of the batch; the group
 * list is rebuilt every list_every steps; use_graph replays a captured HIP
 * graph of one outer step.  Returns once the steps are done: status 2 if a
 * system's list outgrew its slab at some step (that system's trajectory is
 * invalid from there on; setting its positions resizes the list). */
int orac_hip_batch_md_run(orac_hip_ctx *ctx, int32_t batch, int32_t nsteps, double dt, int32_t ninner,
                          int32_t list_every, int32_t use_graph);
/* Two nonbonded shells as ORAC's r-RESPA splits them (mts_forces.f:57-93,
 * the m shell and the outer one): the inner shell 0 .. rcutm switched out
 * over rtolm, the outer shell rcutm .. rcut switched in over rtolm with its
 * real-space Ewald term skipped where decide_force says so
 * (mts_forces.f:110-115), PME and the corrections with the outer shell.
 * rcutm <= 0 goes back to one shell.  Evaluations then give both shells'
 * forces and energies together; not for alchemical, REM or half-list
 * systems (status 4). */
int orac_hip_system_set_shells(orac_hip_ctx *ctx, int32_t sys, double rcutm, double rtolm);
/* nsteps outer steps of dt for a batch whose systems all have two shells:
 * the outer shell once per dt, the inner shell nmid times per dt (hm =
 * dt/nmid) and the fast terms ninner times per hm, in mtsmd.f's order of
 * kicks, drifts and constraints (its m and n0 levels inside h).  Other
 * arguments and results as orac_hip_batch_md_run. */
int orac_hip_batch_md_run_shells(orac_hip_ctx *ctx, int32_t batch, int32_t nsteps, double dt, int32_t nmid,
                                 int32_t ninner, int32_t list_every, int32_t use_graph);
/* niter steepest-descent steps: every unit moves along its total force by
 * hstep F/m, at most dmax A per atom, rigid units re-constrained; velocities
 * are zeroed. */
int orac_hip_batch_md_relax(orac_hip_ctx *ctx, int32_t batch, int32_t niter, double hstep, double dmax,
                            int32_t list_every);
/* v *= scale[k] for system k of the batch (velocity rescaling). */
int orac_hip_batch_md_scale_velocities(orac_hip_ctx *ctx, int32_t batch, const double *scale);
/* out[5]: kinetic energy, harmonic-term energy, total momentum (x, y, z) at
 * the current state; *fail (may be null): constraint solves that did not
 * converge since setup.  The nonbonded energies come from system_results. */
int orac_hip_system_md_energies(orac_hip_ctx *ctx, int32_t sys, double *out, int32_t *fail);

#ifdef __cplusplus
}
#endif
#endif
