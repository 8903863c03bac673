! Fortran interfaces of libelfmm_gfx950.so.
!
! The drop-in entry points elfmm3dtria, elfmm3dtriatarg, el3dtriadirecttarg,
! elfmm3dtriaiter and elfmm3dtriatargiter are exported under the reference's own linker names (lowercase, trailing
! underscore, every argument by reference), so a program that called the
! reference's subroutines through implicit interfaces links this library
! unchanged. This module adds explicit interfaces for those (argument
! checking at compile time) and bind(C) interfaces for the device-resident
! entry point and the diagnostics.
module elfmm_gfx950
  use iso_c_binding
  implicit none

  interface
     ! src/elfmm3dtria.f:71-74
     subroutine elfmm3dtria(ier, iprec, rlam, rmu, triangle, trinorm, nparts, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, ifptfrc, ptfrc, ifstrain, strain) &
          bind(C, name="elfmm3dtria_")
       import :: c_int, c_double
       integer(c_int) :: ier
       integer(c_int) :: iprec, nparts, ifsingle, ifdouble, ifptfrc, ifstrain
       real(c_double) :: rlam, rmu
       real(c_double) :: triangle(3, 3, *), trinorm(3, *), source(3, *)
       real(c_double) :: sigma_sl(3, *), sigma_dl(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
     end subroutine elfmm3dtria

     ! src/elfmm3dtria.f:162-167
     subroutine elfmm3dtriatarg(ier, iprec, rlam, rmu, triangle, trinorm, nparts, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, ifptfrc, ptfrc, ifstrain, strain, &
          ntargs, target, ifptfrctarg, ptfrctarg, ifstraintarg, straintarg) &
          bind(C, name="elfmm3dtriatarg_")
       import :: c_int, c_double
       integer(c_int) :: ier
       integer(c_int) :: iprec, nparts, ifsingle, ifdouble, ifptfrc, ifstrain
       integer(c_int) :: ntargs, ifptfrctarg, ifstraintarg
       real(c_double) :: rlam, rmu
       real(c_double) :: triangle(3, 3, *), trinorm(3, *), source(3, *)
       real(c_double) :: sigma_sl(3, *), sigma_dl(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
       real(c_double) :: target(3, *), ptfrctarg(3, *), straintarg(3, 3, *)
     end subroutine elfmm3dtriatarg

     ! src/elfmm3dtria.f:1453-1458
     subroutine el3dtriadirecttarg(rlam, rmu, triangle, trinorm, nsource, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, ifptfrc, ptfrc, ifstrain, strain, &
          ntarget, target, ifptfrctarg, ptfrctarg, ifstraintarg, straintarg) &
          bind(C, name="el3dtriadirecttarg_")
       import :: c_int, c_double
       integer(c_int) :: nsource, ifsingle, ifdouble, ifptfrc, ifstrain
       integer(c_int) :: ntarget, ifptfrctarg, ifstraintarg
       real(c_double) :: rlam, rmu
       real(c_double) :: triangle(3, 3, *), trinorm(3, *), source(3, *)
       real(c_double) :: sigma_sl(3, *), sigma_dl(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
       real(c_double) :: target(3, *), ptfrctarg(3, *), straintarg(3, 3, *)
     end subroutine el3dtriadirecttarg

     ! src/elfmm3dpart.f:104-109: point sources
     subroutine elfmm3dparttarg(ier, iprec, rlam, rmu, nparts, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, sigma_dv, ifptfrc, ptfrc, ifstrain, strain, &
          ntargs, target, ifptfrctarg, ptfrctarg, ifstraintarg, straintarg) &
          bind(C, name="elfmm3dparttarg_")
       import :: c_int, c_double
       integer(c_int) :: ier, iprec, nparts, ifsingle, ifdouble, ifptfrc, ifstrain
       integer(c_int) :: ntargs, ifptfrctarg, ifstraintarg
       real(c_double) :: rlam, rmu
       real(c_double) :: source(3, *), sigma_sl(3, *), sigma_dl(3, *), sigma_dv(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
       real(c_double) :: target(3, *), ptfrctarg(3, *), straintarg(3, 3, *)
     end subroutine elfmm3dparttarg

     ! src/elfmm3dpart.f:52-55
     subroutine elfmm3dpart(ier, iprec, rlam, rmu, nparts, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, sigma_dv, ifptfrc, ptfrc, ifstrain, strain) &
          bind(C, name="elfmm3dpart_")
       import :: c_int, c_double
       integer(c_int) :: ier, iprec, nparts, ifsingle, ifdouble, ifptfrc, ifstrain
       real(c_double) :: rlam, rmu
       real(c_double) :: source(3, *), sigma_sl(3, *), sigma_dl(3, *), sigma_dv(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
     end subroutine elfmm3dpart

     ! src/elfmm3dpart.f:506-511
     subroutine el3dpartdirecttarg(rlam, rmu, nsource, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, sigma_dv, ifptfrc, ptfrc, ifstrain, strain, &
          ntarget, target, ifptfrctarg, ptfrctarg, ifstraintarg, straintarg) &
          bind(C, name="el3dpartdirecttarg_")
       import :: c_int, c_double
       integer(c_int) :: nsource, ifsingle, ifdouble, ifptfrc, ifstrain
       integer(c_int) :: ntarget, ifptfrctarg, ifstraintarg
       real(c_double) :: rlam, rmu
       real(c_double) :: source(3, *), sigma_sl(3, *), sigma_dl(3, *), sigma_dv(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
       real(c_double) :: target(3, *), ptfrctarg(3, *), straintarg(3, 3, *)
     end subroutine el3dpartdirecttarg

     ! src/elhfmm3dpart.f:961-966 (points in z < 0)
     subroutine elh3dpartdirecttarg(rlam, rmu, nsource, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, sigma_dv, ifptfrc, ptfrc, ifstrain, strain, &
          ntarget, target, ifptfrctarg, ptfrctarg, ifstraintarg, straintarg) &
          bind(C, name="elh3dpartdirecttarg_")
       import :: c_int, c_double
       integer(c_int) :: nsource, ifsingle, ifdouble, ifptfrc, ifstrain
       integer(c_int) :: ntarget, ifptfrctarg, ifstraintarg
       real(c_double) :: rlam, rmu
       real(c_double) :: source(3, *), sigma_sl(3, *), sigma_dl(3, *), sigma_dv(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
       real(c_double) :: target(3, *), ptfrctarg(3, *), straintarg(3, 3, *)
     end subroutine elh3dpartdirecttarg

     ! src/elfmm3dtria_iter_1.f:160-166 (wsave carries a handle to the
     ! device-resident near-field interactions; lused = 4 for icomp 1-5)
     subroutine elfmm3dtriatargiter(ier, iprec, rlam, rmu, triangle, trinorm, nparts, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, ifptfrc, ptfrc, ifstrain, strain, &
          ntargs, target, ifptfrctarg, ptfrctarg, ifstraintarg, straintarg, &
          icomp_type, wsave, lwsave, lused) bind(C, name="elfmm3dtriatargiter_")
       import :: c_int, c_double
       integer(c_int) :: ier, lused
       integer(c_int) :: iprec, nparts, ifsingle, ifdouble, ifptfrc, ifstrain
       integer(c_int) :: ntargs, ifptfrctarg, ifstraintarg, icomp_type, lwsave
       real(c_double) :: rlam, rmu
       real(c_double) :: triangle(3, 3, *), trinorm(3, *), source(3, *)
       real(c_double) :: sigma_sl(3, *), sigma_dl(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
       real(c_double) :: target(3, *), ptfrctarg(3, *), straintarg(3, 3, *)
       real(c_double) :: wsave(*)
     end subroutine elfmm3dtriatargiter

     ! src/elfmm3dtria_iter_1.f:43-47
     subroutine elfmm3dtriaiter(ier, iprec, rlam, rmu, triangle, trinorm, nparts, source, &
          ifsingle, sigma_sl, ifdouble, sigma_dl, ifptfrc, ptfrc, ifstrain, strain, &
          icomp_type, wsave, lwsave, lused) bind(C, name="elfmm3dtriaiter_")
       import :: c_int, c_double
       integer(c_int) :: ier, lused
       integer(c_int) :: iprec, nparts, ifsingle, ifdouble, ifptfrc, ifstrain, icomp_type, lwsave
       real(c_double) :: rlam, rmu
       real(c_double) :: triangle(3, 3, *), trinorm(3, *), source(3, *)
       real(c_double) :: sigma_sl(3, *), sigma_dl(3, *)
       real(c_double) :: ptfrc(3, *), strain(3, 3, *)
       real(c_double) :: wsave(*)
     end subroutine elfmm3dtriaiter

     ! Frees the device-resident interactions named by wsave.
     integer(c_int) function elfmm_gfx950_iter_release(wsave, lwsave) &
          bind(C, name="elfmm_gfx950_iter_release")
       import :: c_int, c_double
       real(c_double) :: wsave(*)
       integer(c_int), value :: lwsave
     end function elfmm_gfx950_iter_release

     ! Device-resident elfmm3dtriatarg: every array argument is a device
     ! address (type(c_ptr)), stream is a hipStream_t (c_null_ptr = default).
     integer(c_int) function elfmm_gfx950_triatarg_dev(iprec, rlam, rmu, nparts, &
          d_triangle, d_trinorm, d_source, ifsingle, d_sigma_sl, ifdouble, d_sigma_dl, &
          ifptfrc, d_ptfrc, ifstrain, d_strain, ntargs, d_target, ifptfrctarg, d_ptfrctarg, &
          ifstraintarg, d_straintarg, stream) bind(C, name="elfmm_gfx950_triatarg_dev")
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: iprec, nparts, ifsingle, ifdouble, ifptfrc, ifstrain
       integer(c_int), value :: ntargs, ifptfrctarg, ifstraintarg
       real(c_double), value :: rlam, rmu
       type(c_ptr), value :: d_triangle, d_trinorm, d_source, d_sigma_sl, d_sigma_dl
       type(c_ptr), value :: d_ptfrc, d_strain, d_target, d_ptfrctarg, d_straintarg
       type(c_ptr), value :: stream
     end function elfmm_gfx950_triatarg_dev

     ! Phase times (ms) of the last call on this thread, 16 entries.
     subroutine elfmm_gfx950_last_timings(ms16) bind(C, name="elfmm_gfx950_last_timings")
       import :: c_double
       real(c_double) :: ms16(16)
     end subroutine elfmm_gfx950_last_timings

     ! Work counts of the last call on this thread, 16 entries.
     subroutine elfmm_gfx950_last_stats(st16) bind(C, name="elfmm_gfx950_last_stats")
       import :: c_double
       real(c_double) :: st16(16)
     end subroutine elfmm_gfx950_last_stats
  end interface
end module elfmm_gfx950
