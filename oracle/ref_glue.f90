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
