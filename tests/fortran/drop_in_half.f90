! A reference-style half-space caller: the iter protocol of
! elhfmm3dtriatargiter (src/elhfmm3dtria_iter.f:46-51; icomp_type 4 for the
! size of wsave, 2 to compute and keep the image near field, 3 to replay
! it, 1 to compute) and the direct sum elh3dtriadirecttarg
! (src/elh3dtria.f:53-58) at a subset of the targets, both through implicit
! interfaces, linked against libelfmm_gfx950.so instead of the reference's
! objects.
!
! usage: drop_in_half <input file> <output file>
! input (stream, native endian): int32 n, nt, iprec, nsub; real*8 rlam, rmu;
!   triangle(3,3,n), trinorm(3,n), source(3,n), sigma_sl(3,n),
!   sigma_dl(3,n), target(3,nt), targets of the direct sum (3,nsub)
! output: int32 ier4, lused, ier2, ier3, ier1; real*8 for icomp 2, 3, 1:
!   ptfrctarg(3,nt), straintarg(3,3,nt); then the direct sum's
!   ptfrctarg(3,nsub), straintarg(3,3,nsub)
program drop_in_half
  implicit none
  integer :: n, nt, iprec, nsub, u, lused, lwsave, icomp
  integer :: ier4, ier2, ier3, ier1
  real*8 :: rlam, rmu
  real*8, allocatable :: triangle(:,:,:), trinorm(:,:), source(:,:)
  real*8, allocatable :: sigma_sl(:,:), sigma_dl(:,:), target(:,:), tsub(:,:)
  real*8, allocatable :: pt(:,:,:), st(:,:,:,:), ptd(:,:), std(:,:,:)
  real*8, allocatable :: wsave(:)
  real*8 :: dum(3,3,1)
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, nt, iprec, nsub
  read(u) rlam, rmu
  allocate(triangle(3,3,n), trinorm(3,n), source(3,n), sigma_sl(3,n), sigma_dl(3,n))
  allocate(target(3,nt), tsub(3,nsub))
  read(u) triangle, trinorm, source, sigma_sl, sigma_dl, target, tsub
  close(u)
  allocate(pt(3,nt,3), st(3,3,nt,3), ptd(3,nsub), std(3,3,nsub))

  ! size of wsave
  allocate(wsave(1))
  lwsave = 0
  icomp = 4
  call elhfmm3dtriatargiter(ier4, iprec, rlam, rmu, triangle, trinorm, n, source, &
       1, sigma_sl, 1, sigma_dl, nt, target, 1, pt(1,1,1), 1, st(1,1,1,1), &
       icomp, wsave, lwsave, lused)
  deallocate(wsave)
  lwsave = max(lused, 1)
  allocate(wsave(lwsave))
  ! compute and keep, then replay
  icomp = 2
  call elhfmm3dtriatargiter(ier2, iprec, rlam, rmu, triangle, trinorm, n, source, &
       1, sigma_sl, 1, sigma_dl, nt, target, 1, pt(1,1,1), 1, st(1,1,1,1), &
       icomp, wsave, lwsave, lused)
  icomp = 3
  call elhfmm3dtriatargiter(ier3, iprec, rlam, rmu, triangle, trinorm, n, source, &
       1, sigma_sl, 1, sigma_dl, nt, target, 1, pt(1,1,2), 1, st(1,1,1,2), &
       icomp, wsave, lwsave, lused)
  ! plain compute
  icomp = 1
  call elhfmm3dtriatargiter(ier1, iprec, rlam, rmu, triangle, trinorm, n, source, &
       1, sigma_sl, 1, sigma_dl, nt, target, 1, pt(1,1,3), 1, st(1,1,1,3), &
       icomp, wsave, lwsave, lused)
  ! direct sum at the subset, targets only
  call elh3dtriadirecttarg(rlam, rmu, triangle, trinorm, n, source, &
       1, sigma_sl, 1, sigma_dl, 0, dum, 0, dum, &
       nsub, tsub, 1, ptd, 1, std)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) ier4, lused, ier2, ier3, ier1
  write(u) pt(:,:,1), st(:,:,:,1), pt(:,:,2), st(:,:,:,2), pt(:,:,3), st(:,:,:,3)
  write(u) ptd, std
  close(u)
end program drop_in_half
