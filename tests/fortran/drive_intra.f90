! drive_intra.f90 -- calls the intramolecular and alchemical drop-in routines
! the way ORAC's callers do (implicit interfaces, reference argument lists):
!   mts_intra_n0.f:73-99   fpbond, fpbend, fitors
!   mts_intra_n1.f:79-119  fnb14 (also on itask = 2, as another OpenMP thread),
!                          fnbgrp, fptors, fnb14_alchemy, fnbgrp_alchemy
!   mts_furier_alchemy.f   ferrf_alchemy, cself_alchemy
!   mtsmd.f:3036-3040      ferrf_alchemy_add
! Each call starts from zero forces; every record of the output file is five
! energies (zero padded) and the forces fx, fy, fz.
! Usage: drive_intra <input file> <output file>   (stream binary, see test_gpu_fortran.py)
program drive_intra
  implicit none
  integer :: n, nb, nbe, nt, ni, n14, nqq, ngi, npk, nbc
  real(8) :: alpha, fudge, lj_fudge, e(5)
  integer, allocatable :: ss(:), nbtype(:), lbnd(:,:), lbndg(:,:), ltor(:,:), litr(:,:), int14(:,:), type14(:)
  integer, allocatable :: ingrp(:,:), listqq(:,:), lbc(:,:)
  real(8), allocatable :: x(:), y(:), z(:), q(:), ecc12(:), ecc6(:), bp(:,:), bep(:,:), tp(:,:), ip(:,:)
  real(8), allocatable :: lam(:), lamq(:), lam0(:), lamq0(:), qa(:), qa0(:), gx(:), gy(:), gz(:), dummy(:)
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) n, nb, nbe, nt, ni, n14, nqq, ngi, npk, nbc
  read(10) alpha, fudge, lj_fudge
  allocate(ss(n), nbtype(n), x(n), y(n), z(n), q(n), ecc12(npk), ecc6(npk))
  allocate(lbnd(2,nb), bp(2,nb), lbndg(3,nbe), bep(7,nbe), ltor(4,nt), tp(nt,2), litr(4,ni), ip(3,ni))
  allocate(int14(2,n14), type14(n14), ingrp(2,ngi), listqq(3,nqq), lbc(2,nbc))
  allocate(lam(n), lamq(n), lam0(n), lamq0(n), qa(n), qa0(n), gx(n), gy(n), gz(n), dummy(8))
  read(10) ss, x, y, z, q, nbtype, ecc12, ecc6
  read(10) lbnd, bp, lbndg, bep, ltor, tp, litr, ip, int14, type14, ingrp, listqq, lbc
  read(10) lam, lamq, lam0, lamq0, qa, qa0
  close(10)
  dummy = 0.0d0
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')

  call zero()
  call fpbond(ss, lbnd, nb, x, y, z, bp(1,:), bp(2,:), e(1), e(2), gx, gy, gz, 0, n, 1)
  call put()
  call zero()
  call fpbend(ss, lbndg, nbe, x, y, z, bep(1,:), bep(2,:), bep(3,:), bep(4,:), bep(5,:), bep(6,:), bep(7,:), &
       e(1), e(2), gx, gy, gz, n, 1)
  call put()
  call zero()
  call fptors(ss, ltor, nt, x, y, z, tp(:,2), tp(:,1), e(1), e(2), gx, gy, gz, n, 1)
  call put()
  call zero()
  call fitors(ss, litr, ni, x, y, z, ip(1,:), ip(2,:), ip(3,:), e(1), e(2), gx, gy, gz, n, 1)
  call put()
  call zero()
  call fnb14(ss, x, y, z, q, n, ecc12, ecc6, .true., .true., alpha, int14, n14, type14, fudge, lj_fudge, gx, gy, gz, &
       e(1), e(2), e(3), e(4), 1)
  call put()
  call zero()
  call fnbgrp(ss, x, y, z, q, nbtype, ecc12, ecc6, .true., alpha, ingrp, ngi, gx, gy, gz, e(1), e(2), e(3), e(4), &
       n, 1)
  call put()
  call zero()
  e = 7.0d0
  call fnb14(ss, x, y, z, q, n, ecc12, ecc6, .true., .true., alpha, int14, n14, type14, fudge, lj_fudge, gx, gy, gz, &
       e(1), e(2), e(3), e(4), 2)
  e(5) = 0.0d0
  call put()
  call zero()
  call ferrf_alchemy_add(ss, alpha, q, lamq, lamq0, x, y, z, listqq, nqq, e(1), e(2), gx, gy, gz, .false., dummy, &
       0.0d0, 0.0d0, e(3), n, .false.)
  call put()
  call zero()
  call fnb14_alchemy(ss, x, y, z, q, n, ecc12, ecc6, .true., .true., alpha, int14, n14, type14, fudge, lj_fudge, &
       gx, gy, gz, e(1), e(2), e(3), e(4), lam, lamq, lam0, lamq0, e(5), 1)
  call put()
  call zero()
  call fnbgrp_alchemy(ss, x, y, z, q, nbtype, ecc12, ecc6, .true., alpha, ingrp, ngi, gx, gy, gz, e(1), e(2), &
       e(3), e(4), lamq, 1)
  call put()
  call zero()
  call ferrf_alchemy(ss, alpha, qa, qa0, 0.8d0, x, y, z, 1, lbc, nbc, e(1), e(2), gx, gy, gz, .false., dummy, &
       0.0d0, 0.0d0, 0, e(3), n)
  call put()
  call zero()
  call cself_alchemy(ss, n, alpha, 1.0d9, qa, qa0, e(1), e(2))
  call put()
  close(11)

contains

  subroutine zero()
    e = 0.0d0
    gx = 0.0d0
    gy = 0.0d0
    gz = 0.0d0
  end subroutine

  subroutine put()
    write(11) e, gx, gy, gz
  end subroutine
end program drive_intra

! ORAC's fatal error path (xerror.f), as the caller provides it
subroutine xerror(msg, n, i, j)
  implicit none
  character(*) :: msg
  integer :: n, i, j
  write(0, '(a,a)') 'xerror: ', msg(1:min(n, len(msg)))
  stop 2
end subroutine
