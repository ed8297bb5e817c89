! drive_shim.f90 -- calls the drop-in routines the way ORAC's callers do
! (implicit interfaces, reference argument lists):
!   mtsmd.f:1582          fft_pme_init
!   mts_furier.f:84-86    fft_pme
!   mtsmd.f:1891-1904     lc_index + lc_list
!   mts_forces.f:148-152  force (m shell: rinn = 0, rout = rcut, healing rtol)
!   mts_furier.f:108-130  ferrf (bonds, bends)
!   mtsmd.f:1641-1645     cself, ferrf(iz = 0) on the constraints
! Usage: drive_shim <input file> <output file>   (stream binary, see test_gpu_fortran.py)
program drive_shim
  implicit none
  integer :: n, ng, nprot, ma, k1, k2, k3, order, lmap, ncorr, ncons, ncx, npk, i
  real(8) :: alpha, volume, rcut, rtol, rnei, eer, massinfty, rtolinn, rc_list
  real(8) :: co(3,3), oc(3,3), stressc(3,3), stress(3,3)
  real(8) :: ucnp, ucns, ucnsp, ucop, ucos, ucosp, fbs, fbv, fcs, fcv, ss1, ss2
  real(8), allocatable :: xp0(:), yp0(:), zp0(:), charge(:), mass(:), pmass(:), xpg(:), ypg(:), zpg(:)
  real(8), allocatable :: xpcm(:), ypcm(:), zpcm(:), ecc12(:), ecc6(:), x0(:), y0(:), z0(:)
  real(8), allocatable :: fpx(:), fpy(:), fpz(:), gx(:), gy(:), gz(:), bsp1(:), bsp2(:), bsp3(:)
  integer, allocatable :: nbtype(:), atomg(:), atomp(:), ss_index(:), type_table(:,:), grppt(:,:), mapnl(:)
  integer, allocatable :: corr(:,:), cons(:,:), groupp(:), nnlpp(:), nnlpp1(:)
  integer :: nind(2), indxi(2,2500), indxj(2,2500), indxk(2,2500), iz1, npp
  logical :: lskip_ewald
  character(len=512) :: fin, fout
  integer(8) :: planf, planb

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) n, ng, nprot, ma, k1, k2, k3, order, lmap, ncorr, ncons, ncx
  read(10) alpha, volume, rcut, rtol, rnei
  npk = ma*(ma+1)/2
  allocate(xp0(n), yp0(n), zp0(n), charge(n), mass(n), pmass(n), xpg(ng), ypg(ng), zpg(ng))
  allocate(xpcm(nprot), ypcm(nprot), zpcm(nprot), ecc12(npk), ecc6(npk), x0(n), y0(n), z0(n))
  allocate(nbtype(n), atomg(n), atomp(n), ss_index(n), type_table(ma,ma), grppt(2,ng), mapnl(lmap))
  allocate(corr(2,max(ncorr,1)), cons(2,max(ncons,1)), groupp(ng), nnlpp(200*ng+16), nnlpp1(200*ng+16))
  allocate(fpx(n), fpy(n), fpz(n), gx(n), gy(n), gz(n), bsp1(k1), bsp2(k2), bsp3(k3))
  read(10) co, oc, xp0, yp0, zp0, charge, mass, pmass, xpg, ypg, zpg, xpcm, ypcm, zpcm, ecc12, ecc6, x0, y0, z0
  read(10) nbtype, atomg, atomp, ss_index, type_table, grppt, mapnl
  if (ncorr > 0) read(10) corr(:, 1:ncorr)
  if (ncons > 0) read(10) cons(:, 1:ncons)
  close(10)
  groupp = 0
  massinfty = 1.0d20

  ! reciprocal space
  call fft_pme_init(n, k1, k2, k3, order, bsp1, bsp2, bsp3, planf, planb)
  fpx = 0.0d0; fpy = 0.0d0; fpz = 0.0d0; stressc = 0.0d0
  call fft_pme(n, xp0, yp0, zp0, xpcm, ypcm, zpcm, charge, co, oc, volume, alpha, order, k1, k2, k3, eer, &
       fpx, fpy, fpz, stressc, atomp, grppt, .true., 1.0d9)
  ! lists
  rc_list = rcut + rtol + rnei
  call lc_index(2500, ncx, ncx, ncx, nind, indxi, indxj, indxk, rc_list, co)
  call lc_list(ncx, ncx, ncx, nind, indxi, indxj, indxk, rc_list, co, xpg, ypg, zpg, ng, nnlpp, 6)
  ! direct space, m shell
  call decide_force(0.0d0, alpha, lskip_ewald)
  gx = 0.0d0; gy = 0.0d0; gz = 0.0d0; stress = 0.0d0
  rtolinn = 0.0d0
  iz1 = 1
  call force(ss_index, xp0, yp0, zp0, xpg, ypg, zpg, charge, nbtype, type_table, ma, n, atomg, xpcm, ypcm, zpcm, &
       groupp, atomp, co, ecc12, ecc6, alpha, mapnl, ng, grppt, ucnp, ucns, ucnsp, ucop, ucos, ucosp, gx, gy, gz, &
       stress, nnlpp, nnlpp1, 0.0d0, 0.0d0, rcut, rtolinn, rtol, pmass, iz1, npp, mass, massinfty, lskip_ewald, &
       .true.)
  ! corrections
  fbs = 0.0d0; fbv = 0.0d0
  if (ncorr > 0) call ferrf(ss_index, alpha, charge, 1.0d0, x0, y0, z0, 1, corr, ncorr, fbs, fbv, gx, gy, gz, &
       .false., x0, 0.0d0, 0.0d0, 0, n)
  call ferrf(ss_index, alpha, charge, 1.0d0, x0, y0, z0, 0, cons, ncons, fcs, fcv, gx, gy, gz, .false., x0, &
       0.0d0, 0.0d0, 0, n)
  call cself(ss_index, n, alpha, 1.0d9, charge, ss1, ss2)

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) eer, stressc, fpx, fpy, fpz
  write(11) ucnp, ucns, ucnsp, ucop, ucos, ucosp, stress, gx, gy, gz
  write(11) fbs, fbv, fcs, fcv, ss1, ss2
  write(11) npp
  close(11)
end program drive_shim

! decide_force is mts_forces.f:212 (kept here as the caller owns it)
subroutine decide_force(rinn, alphal, lskip_ewald)
  implicit none
  real(8) :: rinn, alphal, qt, expcst, erfcst
  logical :: lskip_ewald
  real(8), parameter :: a1 = 0.2548296d0, a2 = -0.28449674d0, a3 = 1.4214137d0, a4 = -1.453152d0, &
       a5 = 1.0614054d0, qp = 0.3275911d0
  qt = 1.0d0/(1.0d0 + qp*alphal*rinn)
  expcst = dexp(-alphal*rinn*alphal*rinn)
  erfcst = ((((a5*qt + a4)*qt + a3)*qt + a2)*qt + a1)*qt*expcst
  lskip_ewald = erfcst .lt. 10.d-4
end subroutine

! ORAC's fatal error path (xerror.f), as the caller provides it
subroutine xerror(msg, n, i, j)
  implicit none
  character(*) :: msg
  integer :: n, i, j
  write(0, '(a,a)') 'xerror: ', msg(1:min(n, len(msg)))
  stop 2
end subroutine
