! orac_hip_shim.f90 -- ISO_C_BINDING drop-in for ORAC's nonbonded routines.
!
! Link this object in place of fft_pme.o, fftw3_pme_init.o, force.o,
! lc_list.o, lc_index.o, ferrf.o, cself.o and fnbgrp.o: the external
! subroutines below keep the reference names and argument lists, so
! mts_furier.f, mts_forces.f and mtsmd.f compile and link unchanged.  Each
! forwards to the C ABI of liborac_hip.so (include/orac_hip.h); a nonzero
! status calls the caller's xerror(msg, 80, 1, 2), ORAC's fatal path.
!
! Layouts are the reference's serial (non-OpenMP) build: nnlpp, nnlpp1 and
! mapnl are single lists.  Compiled with ORAC's own cpp flags, an -D_OMP_
! build whose thread count nthr exceeds 1 (thread-split nnlpp(mpp8,nthr) and
! mapnl, omp_integr.f90:236-281) is refused through xerror by force and
! lc_list.  The routines ORAC calls inside its own OpenMP regions with a
! thread index itask (the 1-4 pairs, intra-group pairs and bonded terms,
! mts_intra_n0.f:65-106, mts_intra_n1.f:58-130) do every term on itask = 1
! and return zero energies on the other threads, whose contributions the
! caller's REDUCTION clauses add; in the serial build itask is always 1.
! List capacity defaults to mpp = _NAT_PP_*_TGROUP_ of the shipped config.H
! (524*50000), the bound the reference checks; orac_hip_set_list_capacity
! changes it.
module orac_hip_iface
  use, intrinsic :: iso_c_binding
  implicit none
  type(c_ptr), save :: orac_ctx = c_null_ptr
  integer(c_int64_t), save :: orac_list_capacity = 524_c_int64_t * 50000_c_int64_t

  interface
    integer(c_int) function c_create(ctx, device) bind(C, name="orac_hip_create")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr) :: ctx
      integer(c_int32_t), value :: device
    end function
    integer(c_int) function c_last_error(buf, n) bind(C, name="orac_hip_last_error")
      import :: c_char, c_int, c_int32_t
      character(kind=c_char) :: buf(*)
      integer(c_int32_t), value :: n
    end function
    integer(c_int) function c_fft_pme_init(ctx, numatoms, n1, n2, n3, order, b1, b2, b3) &
         bind(C, name="orac_hip_fft_pme_init")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: numatoms, n1, n2, n3, order
      real(c_double) :: b1(*), b2(*), b3(*)
    end function
    integer(c_int) function c_fft_pme(ctx, ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, co, oc, volume, alphal, &
         order, n1, n2, n3, eer, fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut) bind(C, name="orac_hip_fft_pme")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ntap, order, n1, n2, n3, pressure
      real(c_double) :: xp(*), yp(*), zp(*), xpcm(*), ypcm(*), zpcm(*), pmechg(*), co(*), oc(*)
      real(c_double), value :: volume, alphal, rkcut
      real(c_double) :: eer, fpx(*), fpy(*), fpz(*), stressc(*)
      integer(c_int32_t) :: atomp(*), grppt(*)
    end function
    integer(c_int) function c_lc_index(indmax, ncx, ncy, ncz, nind, indxi, indxj, indxk, ctoff, co) &
         bind(C, name="orac_hip_lc_index")
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: indmax, ncx, ncy, ncz
      integer(c_int32_t) :: nind(*), indxi(*), indxj(*), indxk(*)
      real(c_double), value :: ctoff
      real(c_double) :: co(*)
    end function
    integer(c_int) function c_lc_list(ctx, ncx, ncy, ncz, nind, indxi, indxj, indxk, rcut, co, xp, yp, zp, natp, &
         nnlpp, cap, nent) bind(C, name="orac_hip_lc_list")
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ncx, ncy, ncz, natp
      integer(c_int32_t) :: nind(*), indxi(*), indxj(*), indxk(*), nnlpp(*)
      real(c_double), value :: rcut
      real(c_double) :: co(*), xp(*), yp(*), zp(*)
      integer(c_int64_t), value :: cap
      integer(c_int64_t) :: nent
    end function
    integer(c_int) function c_force(ctx, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, &
         atomg, xpcm, ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, &
         uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, cap, rneigh, rinn, rout, &
         rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, lskip_ewald, ewald) bind(C, name="orac_hip_force")
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), nbtype(*), type(*), atomg(*), groupp(*), atomp(*), mapnl(*), grppt(*)
      integer(c_int32_t) :: nnlpp(*), nnlpp1(*), ncount1
      integer(c_int32_t), value :: ma, nato, ngrp, iz, lskip_ewald, ewald
      real(c_double) :: xp0(*), yp0(*), zp0(*), xpg(*), ypg(*), zpg(*), charge(*), xpcm(*), ypcm(*), zpcm(*)
      real(c_double) :: co(*), ecc12(*), ecc6(*), gmass(*), mass(*), fpx(*), fpy(*), fpz(*), stress(*)
      real(c_double) :: uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, rtolinn
      real(c_double), value :: alphal, rneigh, rinn, rout, rtolout, massinfty
      integer(c_int64_t), value :: cap
    end function
    integer(c_int) function c_force_rem(ctx, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, &
         atomg, xpcm, ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, &
         uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, cap, rneigh, rinn, rout, &
         rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, seg_ene, rematom, lskip_ewald, ewald, segkind, &
         factor3, lskip_force, ljintra) bind(C, name="orac_hip_force_rem")
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), nbtype(*), type(*), atomg(*), groupp(*), atomp(*), mapnl(*), grppt(*)
      integer(c_int32_t) :: nnlpp(*), nnlpp1(*), ncount1, rematom(*)
      integer(c_int32_t), value :: ma, nato, ngrp, iz, lskip_ewald, ewald, segkind, lskip_force, ljintra
      real(c_double) :: xp0(*), yp0(*), zp0(*), xpg(*), ypg(*), zpg(*), charge(*), xpcm(*), ypcm(*), zpcm(*)
      real(c_double) :: co(*), ecc12(*), ecc6(*), gmass(*), mass(*), fpx(*), fpy(*), fpz(*), stress(*)
      real(c_double) :: uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, rtolinn, seg_ene
      real(c_double), value :: alphal, rneigh, rinn, rout, rtolout, massinfty, factor3
      integer(c_int64_t), value :: cap
    end function
    integer(c_int) function c_force_alchemy(ctx, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, &
         nato, atomg, xpcm, ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, &
         uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, cap, rneigh, &
         rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, ss_alchemy, lam, intramol, ljintra, &
         opls, rmin, epsm, dwrk, lskip_ewald, ewald, qsoft, ljsoft) bind(C, name="orac_hip_force_alchemy")
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), nbtype(*), type(*), atomg(*), groupp(*), atomp(*), mapnl(*), grppt(*)
      integer(c_int32_t) :: nnlpp(*), nnlpp1(*), ncount1, ss_alchemy(*)
      integer(c_int32_t), value :: ma, nato, ngrp, iz, lskip_ewald, ewald, intramol, ljintra, opls
      real(c_double) :: xp0(*), yp0(*), zp0(*), xpg(*), ypg(*), zpg(*), charge(*), xpcm(*), ypcm(*), zpcm(*)
      real(c_double) :: co(*), ecc12(*), ecc6(*), gmass(*), mass(*), fpx(*), fpy(*), fpz(*), stress(*)
      real(c_double) :: uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, rtolinn, dwrk
      real(c_double) :: lam(8), rmin(*), epsm(*)
      real(c_double), value :: alphal, rneigh, rinn, rout, rtolout, massinfty, qsoft, ljsoft
      integer(c_int64_t), value :: cap
    end function
    integer(c_int) function c_force_alchemy_shift(ctx, ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, &
         nato, atomg, xpcm, ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, &
         uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, cap, rneigh, &
         rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, ss_alchemy, lam, intramol, ljintra, &
         opls, rmin, epsm, dwrk, lskip_ewald, ewald, qsoft, ljsoft) bind(C, name="orac_hip_force_alchemy_shift")
      import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), nbtype(*), type(*), atomg(*), groupp(*), atomp(*), mapnl(*), grppt(*)
      integer(c_int32_t) :: nnlpp(*), nnlpp1(*), ncount1, ss_alchemy(*)
      integer(c_int32_t), value :: ma, nato, ngrp, iz, lskip_ewald, ewald, intramol, ljintra, opls
      real(c_double) :: xp0(*), yp0(*), zp0(*), xpg(*), ypg(*), zpg(*), charge(*), xpcm(*), ypcm(*), zpcm(*)
      real(c_double) :: co(*), ecc12(*), ecc6(*), gmass(*), mass(*), fpx(*), fpy(*), fpz(*), stress(*)
      real(c_double) :: uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, rtolinn, dwrk
      real(c_double) :: lam(8), rmin(*), epsm(*)
      real(c_double), value :: alphal, rneigh, rinn, rout, rtolout, massinfty, qsoft, ljsoft
      integer(c_int64_t), value :: cap
    end function
    integer(c_int) function c_fft_pme_alchemy(ctx, ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, charg0, co, oc, &
         volume, alphal, order, n1, n2, n3, eer, fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut, dwrk) &
         bind(C, name="orac_hip_fft_pme_alchemy")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ntap, order, n1, n2, n3, pressure
      real(c_double) :: xp(*), yp(*), zp(*), xpcm(*), ypcm(*), zpcm(*), pmechg(*), charg0(*), co(*), oc(*)
      real(c_double), value :: volume, alphal, rkcut
      real(c_double) :: eer, fpx(*), fpy(*), fpz(*), stressc(*), dwrk
      integer(c_int32_t) :: atomp(*), grppt(*)
    end function
    integer(c_int) function c_ferrf_alchemy(ctx, ss_index, alphal, charge, charge0, fudge, x0, y0, z0, iz, list, &
         nlist, fs_slt, fs_slv, fpx, fpy, fpz, erf_corr, nbonds_added, dwrk, mb) bind(C, name="orac_hip_ferrf_alchemy")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), list(*)
      real(c_double), value :: alphal, fudge
      real(c_double) :: charge(*), charge0(*), x0(*), y0(*), z0(*), fs_slt, fs_slv, fpx(*), fpy(*), fpz(*), dwrk
      integer(c_int32_t), value :: iz, nlist, erf_corr, nbonds_added, mb
    end function
    integer(c_int) function c_cself_alchemy(ctx, ss_index, nato, alphal, kcut, charge, charge0, self_slt, dwrk) &
         bind(C, name="orac_hip_cself_alchemy")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*)
      integer(c_int32_t), value :: nato
      real(c_double), value :: alphal, kcut
      real(c_double) :: charge(*), charge0(*), self_slt, dwrk
    end function
    integer(c_int) function c_fitors(ctx, ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, ut_slt, ut_slv, &
         fpx, fpy, fpz, mb) bind(C, name="orac_hip_fitors")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), litr(*)
      integer(c_int32_t), value :: litor, mb
      real(c_double) :: xp0(*), yp0(*), zp0(*), pota(*), potb(*), potc(*), ut_slt, ut_slv, fpx(*), fpy(*), fpz(*)
    end function
    integer(c_int) function c_fpbond(ctx, ss_index, lbnd, lbond, xp0, yp0, zp0, pota, potb, ub_slt, ub_slv, fpx, &
         fpy, fpz, mb) bind(C, name="orac_hip_fpbond")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), lbnd(*)
      integer(c_int32_t), value :: lbond, mb
      real(c_double) :: xp0(*), yp0(*), zp0(*), pota(*), potb(*), ub_slt, ub_slv, fpx(*), fpy(*), fpz(*)
    end function
    integer(c_int) function c_fpbend(ctx, ss_index, lbndg, lbend, xp0, yp0, zp0, pota, potb, potc, potd, pote, &
         potf, potg, ub_slt, ub_slv, fpx, fpy, fpz, mb) bind(C, name="orac_hip_fpbend")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), lbndg(*)
      integer(c_int32_t), value :: lbend, mb
      real(c_double) :: xp0(*), yp0(*), zp0(*), pota(*), potb(*), potc(*), potd(*), pote(*), potf(*), potg(*)
      real(c_double) :: ub_slt, ub_slv, fpx(*), fpy(*), fpz(*)
    end function
    integer(c_int) function c_fptors_rem(ctx, ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, ut_slt, ut_slv, &
         fpx, fpy, fpz, remf, seg, f2, mb) bind(C, name="orac_hip_fptors_rem")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), ltor(*), remf(*)
      integer(c_int32_t), value :: ltors, mb
      real(c_double) :: xp0(*), yp0(*), zp0(*), pota(*), potb(*), ut_slt, ut_slv, fpx(*), fpy(*), fpz(*), seg
      real(c_double), value :: f2
    end function
    integer(c_int) function c_fitors_rem(ctx, ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, ut_slt, &
         ut_slv, fpx, fpy, fpz, remf, seg, f2, mb) bind(C, name="orac_hip_fitors_rem")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), litr(*), remf(*)
      integer(c_int32_t), value :: litor, mb
      real(c_double) :: xp0(*), yp0(*), zp0(*), pota(*), potb(*), potc(*), ut_slt, ut_slv, fpx(*), fpy(*), fpz(*)
      real(c_double) :: seg
      real(c_double), value :: f2
    end function
    integer(c_int) function c_fnb14_rem(ctx, ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, alphal, &
         int14, int14p, type14, fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, remf, &
         seg, f2) bind(C, name="orac_hip_fnb14_rem")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), int14(*), type14(*), remf(*)
      integer(c_int32_t), value :: nato, cut, ewald, int14p
      real(c_double), value :: alphal, fudge, lj_fudge, f2
      real(c_double) :: xp0(*), yp0(*), zp0(*), charge(*), ecc12(*), ecc6(*), fpx(*), fpy(*), fpz(*)
      real(c_double) :: uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, seg
    end function
    integer(c_int) function c_fptors(ctx, ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, ut_slt, ut_slv, fpx, &
         fpy, fpz, mb) bind(C, name="orac_hip_fptors")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), ltor(*)
      integer(c_int32_t), value :: ltors, mb
      real(c_double) :: xp0(*), yp0(*), zp0(*), pota(*), potb(*), ut_slt, ut_slv, fpx(*), fpy(*), fpz(*)
    end function
    integer(c_int) function c_fnb14(ctx, ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, alphal, &
         int14, int14p, type14, fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv) &
         bind(C, name="orac_hip_fnb14")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), int14(*), type14(*)
      integer(c_int32_t), value :: nato, cut, ewald, int14p
      real(c_double) :: xp0(*), yp0(*), zp0(*), charge(*), ecc12(*), ecc6(*), fpx(*), fpy(*), fpz(*)
      real(c_double), value :: alphal, fudge, lj_fudge
      real(c_double) :: uconf_slt, uconf_slv, ucoul_slt, ucoul_slv
    end function
    integer(c_int) function c_ferrf(ctx, ss_index, alphal, charge, fudge, x0, y0, z0, iz, list, nlist, fs_slt, &
         fs_slv, fpx, fpy, fpz, erf_corr, erf_arr_corr, delew, rlew, nbonds_added, mb) bind(C, name="orac_hip_ferrf")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), list(*)
      real(c_double), value :: alphal, fudge, delew, rlew
      real(c_double) :: charge(*), x0(*), y0(*), z0(*), fs_slt, fs_slv, fpx(*), fpy(*), fpz(*), erf_arr_corr(*)
      integer(c_int32_t), value :: iz, nlist, erf_corr, nbonds_added, mb
    end function
    integer(c_int) function c_int_corr_erf_spline(rmin, rmax, nbin, alpha, kcut, erf_arr_corr, tau) &
         bind(C, name="orac_hip_int_corr_erf_spline")
      import :: c_int, c_int32_t, c_double
      real(c_double), value :: rmin, rmax, alpha, kcut
      integer(c_int32_t), value :: nbin
      real(c_double) :: erf_arr_corr(*), tau(*)
    end function
    integer(c_int) function c_erf_corr_cutoff(oc, delew, rkcut, n1, n2, n3) bind(C, name="orac_hip_erf_corr_cutoff")
      import :: c_int, c_int32_t, c_double
      real(c_double) :: oc(*), delew, rkcut
      integer(c_int32_t), value :: n1, n2, n3
    end function
    integer(c_int) function c_cself(ctx, ss_index, nato, alphal, kcut, charge, s_slt, s_slv) &
         bind(C, name="orac_hip_cself")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*)
      integer(c_int32_t), value :: nato
      real(c_double), value :: alphal, kcut
      real(c_double) :: charge(*), s_slt, s_slv
    end function
    integer(c_int) function c_fnbgrp(ctx, ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, ewald, alphal, ingrp, &
         ingrpp, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, nato) bind(C, name="orac_hip_fnbgrp")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), type(*), ingrp(*)
      real(c_double) :: xp0(*), yp0(*), zp0(*), charge(*), ecc12(*), ecc6(*), fpx(*), fpy(*), fpz(*)
      real(c_double) :: uconf_slt, uconf_slv, ucoul_slt, ucoul_slv
      integer(c_int32_t), value :: ewald, ingrpp, nato
      real(c_double), value :: alphal
    end function
    integer(c_int) function c_ferrf_alchemy_add(ctx, ss_index, alphal, charge, lambdaq, lambdaq0, x0, y0, z0, list, &
         nlist, fs_slt, fs_slv, fpx, fpy, fpz, erf_corr, erf_arr_corr, delew, rlew, dwrk, nato, intramol) &
         bind(C, name="orac_hip_ferrf_alchemy_add")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), list(*)
      real(c_double), value :: alphal, delew, rlew
      real(c_double) :: charge(*), lambdaq(*), lambdaq0(*), x0(*), y0(*), z0(*), fs_slt, fs_slv
      real(c_double) :: fpx(*), fpy(*), fpz(*), erf_arr_corr(*), dwrk
      integer(c_int32_t), value :: nlist, erf_corr, nato, intramol
    end function
    integer(c_int) function c_fnb14_alchemy(ctx, ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, &
         alphal, int14, int14p, type14, fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, &
         lambda, lambdaq, lambda0, lambdaq0, dwrk) bind(C, name="orac_hip_fnb14_alchemy")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), int14(*), type14(*)
      integer(c_int32_t), value :: nato, cut, ewald, int14p
      real(c_double) :: xp0(*), yp0(*), zp0(*), charge(*), ecc12(*), ecc6(*), fpx(*), fpy(*), fpz(*)
      real(c_double) :: lambda(*), lambdaq(*), lambda0(*), lambdaq0(*)
      real(c_double), value :: alphal, fudge, lj_fudge
      real(c_double) :: uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, dwrk
    end function
    integer(c_int) function c_fnbgrp_alchemy(ctx, ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, ewald, alphal, &
         ingrp, ingrpp, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, lambdaq, nato) &
         bind(C, name="orac_hip_fnbgrp_alchemy")
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: ss_index(*), type(*), ingrp(*)
      real(c_double) :: xp0(*), yp0(*), zp0(*), charge(*), ecc12(*), ecc6(*), fpx(*), fpy(*), fpz(*), lambdaq(*)
      real(c_double) :: uconf_slt, uconf_slv, ucoul_slt, ucoul_slv
      integer(c_int32_t), value :: ewald, ingrpp, nato
      real(c_double), value :: alphal
    end function
  end interface

contains

  ! force and lc_list take ORAC's serial list layouts: an OpenMP build running
  ! more than one thread hands them thread-split lists, which is fatal here
  subroutine serial_layouts(name)
#ifdef _OMP_
    use omp_integr, only: nthr
#endif
    character(len=*), intent(in) :: name
#ifdef _OMP_
    character(len=80) :: msg
    if (nthr > 1) then
      msg = name // ': thread-split lists (OpenMP, nthr > 1) are not supported'
      call xerror(msg, 80, 1, 2)
    end if
#endif
  end subroutine

  subroutine orac_hip_set_list_capacity(n)
    integer, intent(in) :: n
    orac_list_capacity = int(n, c_int64_t)
  end subroutine

  function ctx() result(p)
    type(c_ptr) :: p
    if (.not. c_associated(orac_ctx)) call check(c_create(orac_ctx, 0_c_int32_t))
    p = orac_ctx
  end function

  integer(c_int32_t) function lg(l)
    logical, intent(in) :: l
    lg = 0
    if (l) lg = 1
  end function

  subroutine check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char) :: buf(80)
    character(len=80) :: msg
    integer :: i, junk
    if (rc == 0) return
    junk = c_last_error(buf, 80_c_int32_t)
    msg = ' '
    do i = 1, 80
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    call xerror(msg, 80, 1, 2)
  end subroutine
end module orac_hip_iface

! fftw3_pme_init.f90:2 (FFTW build signature; the plan handles are unused)
subroutine fft_pme_init(numatoms, nfft1, nfft2, nfft3, order, bsp_mod1, bsp_mod2, bsp_mod3, planf, planb)
  use orac_hip_iface
  implicit none
  integer :: numatoms, nfft1, nfft2, nfft3, order
  real(8) :: bsp_mod1(*), bsp_mod2(*), bsp_mod3(*)
  type(c_ptr) :: planf, planb
  call check(c_fft_pme_init(ctx(), numatoms, nfft1, nfft2, nfft3, order, bsp_mod1, bsp_mod2, bsp_mod3))
  planf = c_null_ptr
  planb = c_null_ptr
end subroutine

! fft_pme.f:1-3
subroutine fft_pme(ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, co, oc, volume, alphal, order, nfft1, nfft2, nfft3, &
     eer, fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut)
  use orac_hip_iface
  implicit none
  integer :: ntap, order, nfft1, nfft2, nfft3, atomp(*), grppt(2,*)
  real(8) :: xp(*), yp(*), zp(*), xpcm(*), ypcm(*), zpcm(*), pmechg(*), co(3,3), oc(3,3), volume, alphal, eer
  real(8) :: fpx(*), fpy(*), fpz(*), stressc(3,3), rkcut
  logical :: pressure
  call check(c_fft_pme(ctx(), ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, co, oc, volume, alphal, order, nfft1, &
       nfft2, nfft3, eer, fpx, fpy, fpz, stressc, atomp, grppt, lg(pressure), rkcut))
end subroutine

! lc_index.f:3-4
subroutine lc_index(indmax, ncx, ncy, ncz, nind, indxi, indxj, indxk, ctoff, co)
  use orac_hip_iface
  implicit none
  integer :: indmax, ncx, ncy, ncz, nind(2), indxi(2,*), indxj(2,*), indxk(2,*)
  real(8) :: ctoff, co(3,3)
  call check(c_lc_index(indmax, ncx, ncy, ncz, nind, indxi, indxj, indxk, ctoff, co))
end subroutine

! lc_list.f:1-2
subroutine lc_list(ncx, ncy, ncz, nind, indxi, indxj, indxk, rcut, co, xp, yp, zp, natp, nnlpp, kprint)
  use orac_hip_iface
  implicit none
  integer :: ncx, ncy, ncz, nind(2), indxi(2,*), indxj(2,*), indxk(2,*), natp, nnlpp(*), kprint
  real(8) :: rcut, co(3,3), xp(*), yp(*), zp(*)
  integer(c_int64_t) :: nent
  call serial_layouts('lc_list')
  call check(c_lc_list(ctx(), ncx, ncy, ncz, nind, indxi, indxj, indxk, rcut, co, xp, yp, zp, natp, nnlpp, &
       orac_list_capacity, nent))
  write(kprint, '(/5x,''Neighbor Lists Dimensions *neighbor('',i8,'')* '')') nent
end subroutine

! force.f:1-6 (mts_forces.f:148-152)
subroutine force(ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm, zpcm, &
     groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss, ucoul_slt, &
     ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, rneigh, rinn, rout, rtolinn, rtolout, gmass, iz, &
     ncount1, mass, massinfty, lskip_ewald, ewald)
  use orac_hip_iface
  implicit none
  integer :: nato, ngrp, ma, iz, ncount1
  integer :: nbtype(*), type(ma,*), grppt(2,*), ss_index(*), groupp(*), atomp(*), atomg(*), nnlpp(*), nnlpp1(*)
  integer :: mapnl(*)
  real(8) :: xp0(*), yp0(*), zp0(*), co(3,3), charge(*), ecc12(*), ecc6(*), xpg(*), ypg(*), zpg(*)
  real(8) :: xpcm(*), ypcm(*), zpcm(*), gmass(*), stress(3,3), mass(*), massinfty
  real(8) :: alphal, uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, rneigh, rinn, rout
  real(8) :: rtolinn, rtolout, fpx(*), fpy(*), fpz(*)
  logical :: ewald, lskip_ewald
  call serial_layouts('force')
  call check(c_force(ctx(), ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, &
       ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss, &
       ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, orac_list_capacity, rneigh, rinn, rout, &
       rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, lg(lskip_ewald), lg(ewald)))
end subroutine

! ferrf.f:1-3
subroutine ferrf(ss_index, alphal, charge, fudge, x0, y0, z0, iz, list, nlist, fscnstr_slt, fscnstr_slv, fpx, fpy, &
     fpz, erf_corr, erf_arr_corr, delew, rlew, nbonds_added, mb)
  use orac_hip_iface
  implicit none
  integer :: iz, nlist, list(2,*), ss_index(*), nbonds_added, mb
  real(8) :: charge(*), x0(*), y0(*), z0(*), fpx(mb), fpy(mb), fpz(mb), fscnstr_slt, fscnstr_slv
  real(8) :: erf_arr_corr(4,*), delew, rlew, alphal, fudge
  logical :: erf_corr
  call check(c_ferrf(ctx(), ss_index, alphal, charge, fudge, x0, y0, z0, iz, list, nlist, fscnstr_slt, fscnstr_slv, &
       fpx, fpy, fpz, lg(erf_corr), erf_arr_corr, delew, rlew, nbonds_added, mb))
end subroutine

! cself.f:1-2
subroutine cself(ss_index, nato, alphal, kcut, charge, self_slt, self_slv)
  use orac_hip_iface
  implicit none
  integer :: ss_index(*), nato
  real(8) :: charge(*), alphal, self_slt, self_slv, kcut
  call check(c_cself(ctx(), ss_index, nato, alphal, kcut, charge, self_slt, self_slv))
end subroutine

! fnbgrp.f:1-3
subroutine fnbgrp(ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, ewald, alphal, ingrp, ingrpp, fpx, fpy, fpz, &
     uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, nato, itask)
  use orac_hip_iface
  implicit none
  integer :: nato, ingrpp, type(*), ingrp(2,*), ss_index(*), itask
  real(8) :: xp0(*), yp0(*), zp0(*), fpx(nato), fpy(nato), fpz(nato), ecc6(*), ecc12(*), alphal, charge(*)
  real(8) :: uconf_slt, ucoul_slt, uconf_slv, ucoul_slv
  logical :: ewald
  if (itask /= 1) then  ! another OpenMP thread: thread 1 does every term
    uconf_slt = 0.0d0; uconf_slv = 0.0d0; ucoul_slt = 0.0d0; ucoul_slv = 0.0d0
    return
  end if
  call check(c_fnbgrp(ctx(), ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, lg(ewald), alphal, ingrp, ingrpp, &
       fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, nato))
end subroutine

! fitors.f:1-2 (mts_intra_n0.f:97-99)
subroutine fitors(ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, utors_slt, utors_slv, fpx, fpy, fpz, mb, &
     itask)
  use orac_hip_iface
  implicit none
  integer :: litor, mb, itask, litr(4,*), ss_index(*)
  real(8) :: utors_slt, utors_slv, pota(*), potb(*), potc(*), xp0(*), yp0(*), zp0(*), fpx(mb), fpy(mb), fpz(mb)
  if (itask /= 1) then  ! another OpenMP thread: thread 1 does every term
    utors_slt = 0.0d0; utors_slv = 0.0d0
    return
  end if
  call check(c_fitors(ctx(), ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, utors_slt, utors_slv, fpx, fpy, &
       fpz, mb))
end subroutine

! fpbond.f:1-2 (mts_intra_n0.f:73-74); icnstr and itask select OpenMP work in the reference
subroutine fpbond(ss_index, lbnd, lbond, xp0, yp0, zp0, pota, potb, ubond_slt, ubond_slv, fpx, fpy, fpz, icnstr, &
     mb, itask)
  use orac_hip_iface
  implicit none
  integer :: lbond, icnstr, mb, itask, lbnd(2,*), ss_index(*)
  real(8) :: ubond_slt, ubond_slv, pota(*), potb(*), xp0(*), yp0(*), zp0(*), fpx(mb), fpy(mb), fpz(mb)
  if (itask /= 1) then  ! another OpenMP thread: thread 1 does every term
    ubond_slt = 0.0d0; ubond_slv = 0.0d0
    return
  end if
  call check(c_fpbond(ctx(), ss_index, lbnd, lbond, xp0, yp0, zp0, pota, potb, ubond_slt, ubond_slv, fpx, fpy, fpz, &
       mb))
end subroutine

! fpbend.f:1-3 (mts_intra_n0.f:82-85)
subroutine fpbend(ss_index, lbndg, lbend, xp0, yp0, zp0, pota, potb, potc, potd, pote, potf, potg, ubend_slt, &
     ubend_slv, fpx, fpy, fpz, mb, itask)
  use orac_hip_iface
  implicit none
  integer :: lbend, mb, itask, lbndg(3,*), ss_index(*)
  real(8) :: ubend_slt, ubend_slv, xp0(*), yp0(*), zp0(*), fpx(mb), fpy(mb), fpz(mb)
  real(8) :: pota(*), potb(*), potc(*), potd(*), pote(*), potf(*), potg(*)
  if (itask /= 1) then  ! another OpenMP thread: thread 1 does every term
    ubend_slt = 0.0d0; ubend_slv = 0.0d0
    return
  end if
  call check(c_fpbend(ctx(), ss_index, lbndg, lbend, xp0, yp0, zp0, pota, potb, potc, potd, pote, potf, potg, &
       ubend_slt, ubend_slv, fpx, fpy, fpz, mb))
end subroutine

! fptors.f:1-2 (mts_intra_n1.f:118-119)
subroutine fptors(ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, utors_slt, utors_slv, fpx, fpy, fpz, mb, itask)
  use orac_hip_iface
  implicit none
  integer :: ltors, mb, itask, ltor(4,*), ss_index(*)
  real(8) :: utors_slt, utors_slv, pota(*), potb(*), xp0(*), yp0(*), zp0(*), fpx(mb), fpy(mb), fpz(mb)
  if (itask /= 1) then  ! another OpenMP thread: thread 1 does every term
    utors_slt = 0.0d0; utors_slv = 0.0d0
    return
  end if
  call check(c_fptors(ctx(), ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, utors_slt, utors_slv, fpx, fpy, fpz, &
       mb))
end subroutine

! fnb14.f:1-3
subroutine fnb14(ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, alphal, int14, int14p, type14, &
     fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, itask)
  use orac_hip_iface
  implicit none
  integer :: nato, int14p, int14(2,*), type14(*), ss_index(*), itask
  real(8) :: xp0(nato), yp0(nato), zp0(nato), fpx(nato), fpy(nato), fpz(nato), ecc6(*), ecc12(*), alphal, &
       charge(nato), fudge, lj_fudge, ucoul_slt, ucoul_slv, uconf_slt, uconf_slv
  logical :: ewald, cut
  if (itask /= 1) then  ! another OpenMP thread: thread 1 does every term
    uconf_slt = 0.0d0; uconf_slv = 0.0d0; ucoul_slt = 0.0d0; ucoul_slv = 0.0d0
    return
  end if
  call check(c_fnb14(ctx(), ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, lg(cut), lg(ewald), alphal, int14, &
       int14p, type14, fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv))
end subroutine

! fft_pme_alchemy.f:1-3 (mts_furier_alchemy.f:125-128)
subroutine fft_pme_alchemy(ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, charg0, co, oc, volume, alphal, order, nfft1, &
     nfft2, nfft3, eer, fpx, fpy, fpz, stressc, atomp, grppt, pressure, rkcut, dwrk)
  use orac_hip_iface
  implicit none
  integer :: ntap, order, nfft1, nfft2, nfft3, atomp(*), grppt(2,*)
  real(8) :: xp(*), yp(*), zp(*), xpcm(*), ypcm(*), zpcm(*), pmechg(*), charg0(*), co(3,3), oc(3,3), volume, alphal
  real(8) :: eer, fpx(*), fpy(*), fpz(*), stressc(3,3), rkcut, dwrk
  logical :: pressure
  call check(c_fft_pme_alchemy(ctx(), ntap, xp, yp, zp, xpcm, ypcm, zpcm, pmechg, charg0, co, oc, volume, alphal, &
       order, nfft1, nfft2, nfft3, eer, fpx, fpy, fpz, stressc, atomp, grppt, lg(pressure), rkcut, dwrk))
end subroutine

! ferrf_alchemy.f:1-3
subroutine ferrf_alchemy(ss_index, alphal, charge, charge0, fudge, x0, y0, z0, iz, list, nlist, fscnstr_slt, &
     fscnstr_slv, fpx, fpy, fpz, erf_corr, erf_arr_corr, delew, rlew, nbonds_added, dwrk, nato)
  use orac_hip_iface
  implicit none
  integer :: nato, iz, nlist, list(2,*), ss_index(*), nbonds_added
  real(8) :: charge(*), charge0(*), x0(*), y0(*), z0(*), fpx(nato), fpy(nato), fpz(nato), fscnstr_slt, fscnstr_slv
  real(8) :: erf_arr_corr(4,*), delew, rlew, dwrk, alphal, fudge
  logical :: erf_corr
  call check(c_ferrf_alchemy(ctx(), ss_index, alphal, charge, charge0, fudge, x0, y0, z0, iz, list, nlist, &
       fscnstr_slt, fscnstr_slv, fpx, fpy, fpz, lg(erf_corr), nbonds_added, dwrk, nato))
end subroutine

! cself_alchemy.f:1-2
subroutine cself_alchemy(ss_index, nato, alphal, kcut, charge, charge0, self_slt, dwrk)
  use orac_hip_iface
  implicit none
  integer :: ss_index(*), nato
  real(8) :: charge(*), charge0(*), alphal, kcut, self_slt, dwrk
  call check(c_cself_alchemy(ctx(), ss_index, nato, alphal, kcut, charge, charge0, self_slt, dwrk))
end subroutine

! int_corr_erf_spline.f:1-2 (mtsmd.f:1597-1598)
subroutine int_corr_erf_spline(rmin, rmax, nbin, alpha, kcut, erf_arr_corr, tau)
  use orac_hip_iface
  implicit none
  integer :: nbin
  real(8) :: rmin, rmax, alpha, kcut, erf_arr_corr(4,*), tau(*)
  call check(c_int_corr_erf_spline(rmin, rmax, nbin, alpha, kcut, erf_arr_corr, tau))
end subroutine

! erf_corr_cutoff.f:1 (mtsmd.f:1596)
subroutine erf_corr_cutoff(oc, delew, rkcut, nfft1, nfft2, nfft3)
  use orac_hip_iface
  implicit none
  integer :: nfft1, nfft2, nfft3
  real(8) :: oc(3,3), delew, rkcut
  call check(c_erf_corr_cutoff(oc, delew, rkcut, nfft1, nfft2, nfft3))
end subroutine

! ferrf_alchemy_add.f:1-3 (mtsmd.f:3036-3040)
subroutine ferrf_alchemy_add(ss_index, alphal, charge, lambdaq, lambdaq0, x0, y0, z0, list, nlist, fscnstr_slt, &
     fscnstr_slv, fpx, fpy, fpz, erf_corr, erf_arr_corr, delew, rlew, dwrk, nato, intramol)
  use orac_hip_iface
  implicit none
  integer :: nlist, list(3,*), ss_index(*), nato
  real(8) :: charge(*), lambdaq(*), lambdaq0(*), x0(*), y0(*), z0(*), fpx(nato), fpy(nato), fpz(nato)
  real(8) :: fscnstr_slt, fscnstr_slv, erf_arr_corr(4,*), delew, rlew, dwrk, alphal
  logical :: erf_corr, intramol
  call check(c_ferrf_alchemy_add(ctx(), ss_index, alphal, charge, lambdaq, lambdaq0, x0, y0, z0, list, nlist, &
       fscnstr_slt, fscnstr_slv, fpx, fpy, fpz, lg(erf_corr), erf_arr_corr, delew, rlew, dwrk, nato, lg(intramol)))
end subroutine

! fnb14_alchemy.f:1-4 (mts_intra_n1.f:79-83)
subroutine fnb14_alchemy(ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, alphal, int14, int14p, &
     type14, fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, lambda, lambdaq, lambda0, &
     lambdaq0, dwrk, itask)
  use orac_hip_iface
  implicit none
  integer :: nato, int14p, int14(2,*), type14(*), ss_index(*), itask
  real(8) :: xp0(nato), yp0(nato), zp0(nato), fpx(nato), fpy(nato), fpz(nato), ecc6(*), ecc12(*), alphal, &
       charge(nato), fudge, lj_fudge, ucoul_slt, ucoul_slv, uconf_slt, uconf_slv, lambda(*), lambdaq(*), &
       lambda0(*), lambdaq0(*), dwrk
  logical :: ewald, cut
  if (itask /= 1) then  ! another OpenMP thread: thread 1 does every term (dwrk is the caller's reduction)
    uconf_slt = 0.0d0; uconf_slv = 0.0d0; ucoul_slt = 0.0d0; ucoul_slv = 0.0d0
    return
  end if
  call check(c_fnb14_alchemy(ctx(), ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, lg(cut), lg(ewald), alphal, &
       int14, int14p, type14, fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, lambda, &
       lambdaq, lambda0, lambdaq0, dwrk))
end subroutine

! fnbgrp_alchemy.f:1-3 (mts_intra_n1.f:97-100); the arrays hold the atoms ingrp names
subroutine fnbgrp_alchemy(ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, ewald, alphal, ingrp, ingrpp, fpx, &
     fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, lambdaq, itask)
  use orac_hip_iface
  implicit none
  integer :: ingrpp, type(*), ingrp(2,*), ss_index(*), itask, nato
  real(8) :: xp0(*), yp0(*), zp0(*), fpx(*), fpy(*), fpz(*), ecc6(*), ecc12(*), lambdaq(*), alphal, charge(*)
  real(8) :: uconf_slt, ucoul_slt, uconf_slv, ucoul_slv
  logical :: ewald
  if (itask /= 1 .or. ingrpp == 0) then
    uconf_slt = 0.0d0; uconf_slv = 0.0d0; ucoul_slt = 0.0d0; ucoul_slv = 0.0d0
    return
  end if
  nato = maxval(ingrp(1:2, 1:ingrpp))
  call check(c_fnbgrp_alchemy(ctx(), ss_index, xp0, yp0, zp0, charge, type, ecc12, ecc6, lg(ewald), alphal, ingrp, &
       ingrpp, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, lambdaq, nato))
end subroutine
