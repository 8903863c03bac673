! A reference-style caller: calls elfmm3dtriatarg through an implicit
! interface with the argument list of src/elfmm3dtria.f:162-167, linked
! against libelfmm_gfx950.so instead of the reference's objects.
!
! usage: drop_in <input file> <output file>
! input (stream, native endian): int32 n, nt, iprec; real*8 rlam, rmu;
!   triangle(3,3,n), trinorm(3,n), source(3,n), sigma_sl(3,n),
!   sigma_dl(3,n), target(3,nt)
! output: int32 ier; real*8 ptfrc(3,n), strain(3,3,n), ptfrctarg(3,nt),
!   straintarg(3,3,nt)
program drop_in
  implicit none
  integer :: n, nt, iprec, ier, u
  real*8 :: rlam, rmu
  real*8, allocatable :: triangle(:,:,:), trinorm(:,:), source(:,:)
  real*8, allocatable :: sigma_sl(:,:), sigma_dl(:,:), target(:,:)
  real*8, allocatable :: ptfrc(:,:), strain(:,:,:), ptfrctarg(:,:), straintarg(:,:,:)
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n, nt, iprec
  read(u) rlam, rmu
  allocate(triangle(3,3,n), trinorm(3,n), source(3,n), sigma_sl(3,n), sigma_dl(3,n))
  allocate(target(3,max(nt,1)))
  read(u) triangle, trinorm, source, sigma_sl, sigma_dl
  if (nt > 0) read(u) target(:,1:nt)
  close(u)
  allocate(ptfrc(3,n), strain(3,3,n), ptfrctarg(3,max(nt,1)), straintarg(3,3,max(nt,1)))

  call elfmm3dtriatarg(ier, iprec, rlam, rmu, triangle, trinorm, n, source, &
       1, sigma_sl, 1, sigma_dl, 1, ptfrc, 1, strain, &
       nt, target, 1, ptfrctarg, 1, straintarg)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) ier
  write(u) ptfrc, strain
  if (nt > 0) write(u) ptfrctarg(:,1:nt), straintarg(:,:,1:nt)
  close(u)
end program drop_in
