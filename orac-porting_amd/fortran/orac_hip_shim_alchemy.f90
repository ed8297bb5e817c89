! orac_hip_shim_alchemy.f90 -- drop-in force_alchemy for ORAC's own build.
!
! Keeps force_alchemy's and force_alchemy_shift's names and argument lists
! (force_alchemy.f:2-7, force_alchemy_shifted.f:2-8) and
! reads, as the reference routine does, the alchemical state from ORAC's
! modules: cpropar (ss_alchemy, ladd/lrm and their charge and lambda0
! twins, intramol, ljintra) and unit (ffield_type).  It needs those modules,
! so it is compiled inside ORAC's build (src/Makefile), next to
! orac_hip_shim.f90, whose module orac_hip_iface declares the C entry point.
! The per-atom lambda arrays are not read (neither does force_alchemy).  As
! force does, both refuse an OpenMP build running more than one thread.
subroutine force_alchemy(ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, &
     ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss, &
     ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, rneigh, rinn, rout, rtolinn, rtolout, &
     gmass, iz, ncount1, mass, massinfty, lambda, lambda0, lambdaq, lambdaq0, rmin, epsm, dwrk, lskip_ewald, ewald, &
     qsoft, ljsoft)
  use orac_hip_iface
  use unit, only: ffield_type
  use cpropar, only: ss_alchemy, ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0, intramol, ljintra
  implicit none
  integer :: nato, ngrp, ma, iz, ncount1
  integer :: nbtype(*), type(ma,*), grppt(2,*), ss_index(*), groupp(*), atomp(*), atomg(*), nnlpp(*), nnlpp1(*)
  integer :: mapnl(*)
  real(8) :: xp0(*), yp0(*), zp0(*), co(3,3), charge(*), ecc12(*), ecc6(*), xpg(*), ypg(*), zpg(*)
  real(8) :: xpcm(*), ypcm(*), zpcm(*), gmass(*), stress(3,3), mass(*), massinfty
  real(8) :: alphal, uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, rneigh, rinn, rout
  real(8) :: rtolinn, rtolout, fpx(*), fpy(*), fpz(*)
  real(8) :: lambda(*), lambda0(*), lambdaq(*), lambdaq0(*), rmin(*), epsm(*), dwrk, qsoft, ljsoft
  logical :: ewald, lskip_ewald
  real(8) :: lam(8)
  integer(c_int32_t) :: opls
  lam = [ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0]
  call serial_layouts('force_alchemy')
  opls = 0
  if (ffield_type(1:4) == 'opls') opls = 1
  call check(c_force_alchemy(ctx(), ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, &
       xpcm, ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, &
       uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, orac_list_capacity, rneigh, &
       rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, ss_alchemy, lam, lg(intramol), &
       lg(ljintra), opls, rmin, epsm, dwrk, lg(lskip_ewald), lg(ewald), qsoft, ljsoft))
end subroutine

! force_alchemy_shifted.f:1-8: the shifted-potential form (SHIFT-POT)
subroutine force_alchemy_shift(ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, &
     ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss, &
     ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, rneigh, rinn, rout, rtolinn, rtolout, &
     gmass, iz, ncount1, mass, massinfty, lambda, lambda0, lambdaq, lambdaq0, rmin, epsm, dwrk, lskip_ewald, ewald, &
     qshift, ljshift)
  use orac_hip_iface
  use unit, only: ffield_type
  use cpropar, only: ss_alchemy, ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0, intramol, ljintra
  implicit none
  integer :: nato, ngrp, ma, iz, ncount1
  integer :: nbtype(*), type(ma,*), grppt(2,*), ss_index(*), groupp(*), atomp(*), atomg(*), nnlpp(*), nnlpp1(*)
  integer :: mapnl(*)
  real(8) :: xp0(*), yp0(*), zp0(*), co(3,3), charge(*), ecc12(*), ecc6(*), xpg(*), ypg(*), zpg(*)
  real(8) :: xpcm(*), ypcm(*), zpcm(*), gmass(*), stress(3,3), mass(*), massinfty
  real(8) :: alphal, uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, rneigh, rinn, rout
  real(8) :: rtolinn, rtolout, fpx(*), fpy(*), fpz(*)
  real(8) :: lambda(*), lambda0(*), lambdaq(*), lambdaq0(*), rmin(*), epsm(*), dwrk, qshift, ljshift
  logical :: ewald, lskip_ewald
  real(8) :: lam(8)
  integer(c_int32_t) :: opls
  lam = [ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0]
  call serial_layouts('force_alchemy')
  opls = 0
  if (ffield_type(1:4) == 'opls') opls = 1
  call check(c_force_alchemy_shift(ctx(), ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, &
       xpcm, ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, &
       uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, orac_list_capacity, rneigh, &
       rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, ss_alchemy, lam, lg(intramol), &
       lg(ljintra), opls, rmin, epsm, dwrk, lg(lskip_ewald), lg(ewald), qshift, ljshift))
end subroutine
