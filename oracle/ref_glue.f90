! ref_glue.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Glue compiled together with the reference's own hot-path sources (built
! in place from the reference tree by oracle/Makefile, target `ref`) so the
! tests can call the reference routines through ctypes.  It only sets the
! module state those routines read and exposes a few module arrays.
!
!  * refglue_init: module `unit` pi (setuns.f:34) and kprint; module
!    `omp_integr` omp_timing (read by lc_list.f:88).
!  * refglue_pme_init: fft_pme_init, fftpack branch
!    (fftw3_pme_init.f90:9-11,58-60), on module `spme` storage.
!  * refglue_get_bsp: copies the module bsp_mod arrays out.
subroutine refglue_init() bind(C, name="refglue_init")
  use unit, only: pi, kprint
  use omp_integr, only: omp_timing
  implicit none
  pi = 4.0d0*datan(1.0d0)
  kprint = 6
  omp_timing = .false.
end subroutine refglue_init

subroutine refglue_pme_init(numatoms, nfft1, nfft2, nfft3, order) bind(C, name="refglue_pme_init")
  use spme
  implicit none
  integer, value :: numatoms, nfft1, nfft2, nfft3, order
  call fft_pme_init(numatoms, nfft1, nfft2, nfft3, order, sizfftab, sizffwrk, siztheta, siz_Q, sizheap, &
       sizstack, bsp_mod1, bsp_mod2, bsp_mod3, fftable, ffwork)
end subroutine refglue_pme_init

subroutine refglue_get_bsp(b1, b2, b3, nfft1, nfft2, nfft3) bind(C, name="refglue_get_bsp")
  use spme
  implicit none
  integer, value :: nfft1, nfft2, nfft3
  real(8) :: b1(nfft1), b2(nfft2), b3(nfft3)
  b1(1:nfft1) = bsp_mod1(1:nfft1)
  b2(1:nfft2) = bsp_mod2(1:nfft2)
  b3(1:nfft3) = bsp_mod3(1:nfft3)
end subroutine refglue_get_bsp

! refglue_set_alchemy: the module state force_alchemy.f reads (cpropar:
! ss_alchemy, ladd/lrm and their charge and lambda0 twins, intramol,
! ljintra; unit: ffield_type).  lam = ladd, lrm, ladd0, lrm0, laddq, lrmq,
! laddq0, lrmq0.
subroutine refglue_set_alchemy(n, ss, lam, intra, ljin, opls) bind(C, name="refglue_set_alchemy")
  use unit, only: ffield_type
  use cpropar, only: ss_alchemy, ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0, intramol, ljintra
  implicit none
  integer, value :: n, intra, ljin, opls
  integer :: ss(n)
  real(8) :: lam(8)
  ss_alchemy(1:n) = ss(1:n)
  ladd = lam(1); lrm = lam(2); ladd0 = lam(3); lrm0 = lam(4)
  laddq = lam(5); lrmq = lam(6); laddq0 = lam(7); lrmq0 = lam(8)
  intramol = intra /= 0
  ljintra = ljin /= 0
  if (opls /= 0) then
    ffield_type = 'opls'
  else
    ffield_type = 'amber'
  end if
end subroutine refglue_set_alchemy

! refglue_set_rem: the module state force_REM.f reads (rem: rem_segkind,
! rem_factor, lskip_force; cpropar: ljintra).  f3 goes to rem_factor(3).
subroutine refglue_set_rem(kind, f3, lskip, ljin) bind(C, name="refglue_set_rem")
  use rem, only: rem_segkind, rem_factor, lskip_force
  use cpropar, only: ljintra
  implicit none
  integer, value :: kind, lskip, ljin
  real(8), value :: f3
  rem_segkind = kind
  rem_factor(3) = f3
  lskip_force = lskip /= 0
  ljintra = ljin /= 0
end subroutine refglue_set_rem

! refglue_set_rem_factor: rem_factor(k) (the bonded _rem routines read k = 2)
subroutine refglue_set_rem_factor(k, f) bind(C, name="refglue_set_rem_factor")
  use rem, only: rem_factor
  implicit none
  integer, value :: k
  real(8), value :: f
  rem_factor(k) = f
end subroutine refglue_set_rem_factor
