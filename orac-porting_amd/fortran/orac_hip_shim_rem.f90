! orac_hip_shim_rem.f90 -- drop-in force_rem for ORAC's own build.
!
! Keeps force_rem's name and argument list (force_REM.f:1-6) and reads, as
! the reference routine does, the REM state from ORAC's modules: rem
! (rem_segkind, rem_factor, lskip_force) and cpropar (ljintra).  It needs
! those modules, so it is compiled inside ORAC's build (src/Makefile), next
! to orac_hip_shim.f90, whose module orac_hip_iface declares the C entry
! point.  rematom is ORAC's LOGICAL array, passed on as 0/1 flags.  As force
! does, it refuses an OpenMP build running more than one thread.  Only the
! Ewald branch is taken (lskip_ewald false, ewald true): the C entry point
! stops through xerror on the others, whose Coulomb term reads locals an
! earlier pair left (force_REM.f:491, :667).
subroutine force_rem(ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, xpcm, ypcm, &
     zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, uconf_ss, ucoul_slt, &
     ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, rneigh, rinn, rout, rtolinn, rtolout, gmass, iz, &
     ncount1, mass, massinfty, seg_ene, rematom, lskip_ewald, ewald)
  use orac_hip_iface
  use rem, only: rem_segkind, rem_factor, lskip_force
  use cpropar, only: ljintra
  implicit none
  integer :: nato, ngrp, ma, iz, ncount1
  integer :: nbtype(*), type(ma,*), grppt(2,*), ss_index(*), groupp(*), atomp(*), atomg(*), nnlpp(*), nnlpp1(*)
  integer :: mapnl(*)
  real(8) :: xp0(*), yp0(*), zp0(*), co(3,3), charge(*), ecc12(*), ecc6(*), xpg(*), ypg(*), zpg(*)
  real(8) :: xpcm(*), ypcm(*), zpcm(*), gmass(*), stress(3,3), mass(*), massinfty
  real(8) :: alphal, uconf_slt, uconf_slv, uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, rneigh, rinn, rout
  real(8) :: rtolinn, rtolout, fpx(*), fpy(*), fpz(*), seg_ene
  logical :: rematom(*), ewald, lskip_ewald
  integer(c_int32_t), allocatable :: flags(:)
  integer :: i
  call serial_layouts('force_rem')
  allocate(flags(max(nato, 1)))
  flags = 0
  do i = 1, nato
    if (rematom(i)) flags(i) = 1
  end do
  call check(c_force_rem(ctx(), ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type, ma, nato, atomg, &
       xpcm, ypcm, zpcm, groupp, atomp, co, ecc12, ecc6, alphal, mapnl, ngrp, grppt, uconf_slt, uconf_slv, &
       uconf_ss, ucoul_slt, ucoul_slv, ucoul_ss, fpx, fpy, fpz, stress, nnlpp, nnlpp1, orac_list_capacity, rneigh, &
       rinn, rout, rtolinn, rtolout, gmass, iz, ncount1, mass, massinfty, seg_ene, flags, lg(lskip_ewald), &
       lg(ewald), int(rem_segkind, c_int32_t), rem_factor(3), lg(lskip_force), lg(ljintra)))
  deallocate(flags)
end subroutine

! The bonded REM twins (fptors_rem.f:1-2, fitors_rem.f:1-3, fnb14_rem.f:1-4),
! with rem_factor(2) from module rem.  Called inside ORAC's OpenMP regions
! with a thread index, they do every term on itask = 1 and return zero
! energies on the other threads (the caller's REDUCTION adds them).
! fptors_rem has no atom count argument: the largest atom index of ltor
! bounds the forces it touches.
subroutine fptors_rem(ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, utors_slt, utors_slv, fpx, fpy, fpz, &
     remltor, itask, seg_ene)
  use orac_hip_iface
  use rem, only: rem_factor
  implicit none
  integer :: ltors, itask, ltor(4,*), ss_index(*)
  real(8) :: utors_slt, utors_slv, pota(*), potb(*), xp0(*), yp0(*), zp0(*), fpx(*), fpy(*), fpz(*), seg_ene
  logical :: remltor(*)
  integer(c_int32_t), allocatable :: flags(:)
  integer :: i, mb
  utors_slt = 0.0d0; utors_slv = 0.0d0; seg_ene = 0.0d0
  if (itask /= 1 .or. ltors <= 0) return
  mb = maxval(ltor(1:4, 1:ltors))
  allocate(flags(ltors))
  do i = 1, ltors
    flags(i) = lg(remltor(i))
  end do
  call check(c_fptors_rem(ctx(), ss_index, ltor, ltors, xp0, yp0, zp0, pota, potb, utors_slt, utors_slv, fpx, fpy, &
       fpz, flags, seg_ene, rem_factor(2), mb))
  deallocate(flags)
end subroutine

subroutine fitors_rem(ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, utors_slt, utors_slv, fpx, fpy, fpz, &
     remitor, mb, itask, seg_ene)
  use orac_hip_iface
  use rem, only: rem_factor
  implicit none
  integer :: litor, mb, itask, litr(4,*), ss_index(*)
  real(8) :: utors_slt, utors_slv, pota(*), potb(*), potc(*), xp0(*), yp0(*), zp0(*), fpx(mb), fpy(mb), fpz(mb)
  real(8) :: seg_ene
  logical :: remitor(*)
  integer(c_int32_t), allocatable :: flags(:)
  integer :: i
  utors_slt = 0.0d0; utors_slv = 0.0d0; seg_ene = 0.0d0
  if (itask /= 1 .or. litor <= 0) return
  allocate(flags(litor))
  do i = 1, litor
    flags(i) = lg(remitor(i))
  end do
  call check(c_fitors_rem(ctx(), ss_index, litr, litor, xp0, yp0, zp0, pota, potb, potc, utors_slt, utors_slv, fpx, &
       fpy, fpz, flags, seg_ene, rem_factor(2), mb))
  deallocate(flags)
end subroutine

subroutine fnb14_rem(ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, cut, ewald, alphal, int14, int14p, type14, &
     fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, remint14, itask, seg_ene)
  use orac_hip_iface
  use rem, only: rem_factor
  implicit none
  integer :: nato, int14p, int14(2,*), type14(*), ss_index(*), itask
  real(8) :: xp0(nato), yp0(nato), zp0(nato), fpx(nato), fpy(nato), fpz(nato), ecc6(*), ecc12(*), alphal, &
       charge(nato), fudge, lj_fudge, ucoul_slt, ucoul_slv, uconf_slt, uconf_slv, seg_ene
  logical :: ewald, cut, remint14(*)
  integer(c_int32_t), allocatable :: flags(:)
  integer :: i
  if (itask /= 1) then  ! another OpenMP thread: thread 1 does every term
    uconf_slt = 0.0d0; uconf_slv = 0.0d0; ucoul_slt = 0.0d0; ucoul_slv = 0.0d0
    if (int14p > 0) seg_ene = 0.0d0
    return
  end if
  allocate(flags(max(int14p, 1)))
  flags = 0
  do i = 1, int14p
    flags(i) = lg(remint14(i))
  end do
  call check(c_fnb14_rem(ctx(), ss_index, xp0, yp0, zp0, charge, nato, ecc12, ecc6, lg(cut), lg(ewald), alphal, &
       int14, int14p, type14, fudge, lj_fudge, fpx, fpy, fpz, uconf_slt, uconf_slv, ucoul_slt, ucoul_slv, flags, &
       seg_ene, rem_factor(2)))
  deallocate(flags)
end subroutine
