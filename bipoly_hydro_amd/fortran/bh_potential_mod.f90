! bh_potential_mod.f90 -- ISO_C_BINDING interface to libbh_potential.so
! (the C ABI declared in include/bh_potential.h).
module bh_potential
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: BH_OK = 0
  integer(c_int), parameter :: BH_PREC_FP64 = 0
  integer(c_int), parameter :: BH_PREC_MIXED = 1
  integer(c_int), parameter :: BH_PREC_MIXED_TABLE32 = 2

  ! mirrors struct bh_grid_desc
  type, bind(C) :: bh_grid_desc
     integer(c_int) :: numr, numz, numphi, isym
     real(c_double) :: dr
     integer(c_int) :: mmax, nsteps_cold, nsteps_warm, refine_passes
  end type bh_grid_desc

  interface
     function bh_plan_create(desc, precision, hbm_budget, plan) bind(C, name="bh_plan_create") result(rc)
       import :: c_int, c_size_t, c_ptr, bh_grid_desc
       type(bh_grid_desc), intent(in) :: desc
       integer(c_int), value :: precision
       integer(c_size_t), value :: hbm_budget
       type(c_ptr), intent(out) :: plan
       integer(c_int) :: rc
     end function bh_plan_create

     function bh_plan_destroy(plan) bind(C, name="bh_plan_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: plan
       integer(c_int) :: rc
     end function bh_plan_destroy

     function bh_solve(plan, rho, pot, tstep) bind(C, name="bh_solve") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: plan
       real(c_double), intent(in) :: rho(*)
       real(c_double), intent(out) :: pot(*)
       integer(c_int), value :: tstep
       integer(c_int) :: rc
     end function bh_solve

     function bh_set_potential(plan, pot, on_device) bind(C, name="bh_set_potential") result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: plan
       real(c_double), intent(in) :: pot(*)
       integer(c_int), value :: on_device
       integer(c_int) :: rc
     end function bh_set_potential

     function bh_last_error_c() bind(C, name="bh_last_error") result(msg)
       import :: c_ptr
       type(c_ptr) :: msg
     end function bh_last_error_c
  end interface

contains

  ! message of the last failure, as a Fortran string
  function bh_last_error() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    msg = ' '
    cp = bh_last_error_c()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [256])
    do i = 1, 256
       if (p(i) == c_null_char) exit
       msg(i:i) = p(i)
    enddo
  end function bh_last_error

end module bh_potential
