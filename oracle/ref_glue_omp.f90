! ref_glue_omp.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Glue for the reference's hot path compiled with -D_OMP_ -fopenmp
! (oracle/Makefile target `ref_omp`, _ref/libref_omp.so), the build ORAC's
! own OpenMP runs use; timed as the CPU baseline.  It sets the module state
! ORAC's start-up gives the OpenMP routines and splits the contact map per
! thread as ORAC does:
!
!  * refglue_omp_setup: the thread counts of module `omp_integr` (all three
!    levels nthr, nthr1, nthr2 as read_input.f:511-512 sets them without
!    dynamic threads), the ownership vectors array_omp/array1_omp/
!    array2_omp from cache_line over 4*m1 entries with ORAC's default
!    cache_line_size 8 (orac.f:136-139, blkdta.f90:491-493), m8t
!    (orac.f:141), mpp8 (mtsmd.f:629) and m1t (mtsmd.f:634); returns mpp8
!    and m8t, the leading dimensions of the per-thread lists.
!  * refglue_omp_split_map: the per-thread contact maps mapnlt(m8t, nthr)
!    from the full map, split_map for every thread (mtsmd.f's set-up).
subroutine refglue_omp_setup(n, mpp8_out, m8t_out) bind(C, name="refglue_omp_setup")
  use parst, only: m1, m8, mpp
  use unit, only: pi, kprint
  use omp_integr
  implicit none
  integer, value :: n
  integer :: mpp8_out, m8t_out
  integer, external :: omp_get_max_threads
  pi = 4.0d0*datan(1.0d0)
  kprint = 6
  omp_timing = .false.
  omp_dynamic = .false.
  nthr = n
  nthr1 = n
  nthr2 = n
  cache_line_size = 8
  cache_line_size1 = 8
  cache_line_size2 = 8
  call cache_line(4*m1, nthr, array_omp, cache_line_size)
  call cache_line(4*m1, nthr1, array1_omp, cache_line_size1)
  call cache_line(4*m1, nthr2, array2_omp, cache_line_size2)
  call omp_set_num_threads(nthr)
  m8t = 1 + m8/nthr
  mpp8 = 1 + mpp/nthr
  m1t = 1 + 2*m1/nthr2
  mpp8_out = mpp8
  m8t_out = m8t
end subroutine refglue_omp_setup

subroutine refglue_omp_split_map(nato, mapnl, grppt, ngrp, mapnlt) bind(C, name="refglue_omp_split_map")
  use unit, only: kprint
  use omp_integr
  implicit none
  integer, value :: nato, ngrp
  integer :: mapnl(*), grppt(2, *)
  integer :: mapnlt(m8t, nthr)
  integer :: itask
  do itask = 1, nthr
    call split_map(nato, mapnl, grppt, ngrp, kprint, mapnlt(1, itask), itask, 0)
  end do
end subroutine refglue_omp_split_map

! refglue_omp_pme_init: fft_pme_init, fftpack branch, on module `spme`
! storage of this build (as refglue_pme_init in ref_glue.f90).
subroutine refglue_omp_pme_init(numatoms, nfft1, nfft2, nfft3, order) bind(C, name="refglue_omp_pme_init")
  use spme
  implicit none
  integer, value :: numatoms, nfft1, nfft2, nfft3, order
  call fft_pme_init(numatoms, nfft1, nfft2, nfft3, order, sizfftab, sizffwrk, siztheta, siz_Q, sizheap, &
       sizstack, bsp_mod1, bsp_mod2, bsp_mod3, fftable, ffwork)
end subroutine refglue_omp_pme_init

! refglue_omp_threads: the team size for parallel regions started by the
! calling thread (an OpenMP setting of each thread: the thread that runs the
! routines sets it).
subroutine refglue_omp_threads(n) bind(C, name="refglue_omp_threads")
  implicit none
  integer, value :: n
  call omp_set_num_threads(n)
end subroutine refglue_omp_threads
