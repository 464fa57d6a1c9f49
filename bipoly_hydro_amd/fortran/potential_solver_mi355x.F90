!  potential_solver_mi355x.F90 -- drop-in replacements for bipoly_hydro's
!  potsetup (potsetup.F90:6) and potential_solver (potential_solver.F90:6)
!  that run the solve on one MI355X through libbh_potential.so.
!
!  Build them in place of potential_solver.F90, potsetup.F90, bessel.F90,
!  helmadi.F90, realft.F90, tridagr.F90, tridagz.F90, tm.F90, sm.F90, bm.F90,
!  elle/ellf/rf/rd/gammln.F90, swap_*.F90 and comm_r/z_sweep.F90, together with
!  bh_potential_mod.f90, linking -lbh_potential.  Both routines keep the
!  reference's argument-less interface and COMMON blocks.
!
!  The hydro code keeps its 2-D (r, z) MPI decomposition; the solve itself
!  runs on the root rank's GPU: rho blocks are gathered to the root, solved,
!  and pot blocks (ghost layers included) are scattered back, so every rank
!  ends with what comm(potp) leaves in the reference (helmadi.F90:414).
!  Precision: BH_PRECISION=mixed in the environment selects the mixed
!  form, anything else the plain FP64 path.
module bh_solver_state
  use bh_potential
  implicit none
  type(c_ptr), save :: plan = c_null_ptr
end module bh_solver_state

subroutine potsetup
  use bh_solver_state
  implicit none
  include 'runhydro.h'
  include 'mpif.h'
  real :: dr, dz, dphi, drinv, dzinv, dphiinv
  common /coord_differentials/ dr, dz, dphi, drinv, dzinv, dphiinv
  integer :: isym
  integer, dimension(3) :: boundary_condition
  common /boundary_conditions/ isym, boundary_condition
  logical :: iam_on_top, iam_on_bottom, iam_on_axis, iam_on_edge, iam_root
  integer :: column_num, row_num
  integer :: iam, down_neighbor, up_neighbor, in_neighbor, out_neighbor, root, &
             REAL_SIZE, INT_SIZE, numprocs
  integer, dimension(numr_procs,numz_procs) :: pe_grid
  common /processor_grid/ iam, numprocs, iam_on_top, iam_on_bottom, iam_on_axis, &
                          iam_on_edge, down_neighbor, up_neighbor, in_neighbor,  &
                          out_neighbor, root, column_num, row_num, pe_grid,      &
                          iam_root, REAL_SIZE, INT_SIZE
  type(bh_grid_desc) :: desc
  integer(c_int) :: rc, prec
  character(len=16) :: env
  integer :: ierror

  if (iam == root) then
     desc%numr = numr
     desc%numz = numz
     desc%numphi = numphi
     desc%isym = isym
     desc%dr = dr
     desc%mmax = 30
     desc%nsteps_cold = 20
     desc%nsteps_warm = 5
     desc%refine_passes = 0
     env = ' '
     call get_environment_variable('BH_PRECISION', env)
     prec = BH_PREC_FP64
     if (trim(env) == 'mixed') prec = BH_PREC_MIXED
     if (trim(env) == 'mixed_t32') prec = BH_PREC_MIXED_TABLE32
     env = ' '
     call get_environment_variable('BH_REFINE_PASSES', env)
     if (len_trim(env) > 0) read(env, *) desc%refine_passes
     rc = bh_plan_create(desc, prec, 0_c_size_t, plan)
     if (rc /= BH_OK) then
        write(6,*) 'potsetup: bh_plan_create failed: ', trim(bh_last_error())
        call mpi_abort(MPI_COMM_WORLD, 1, ierror)
     endif
  endif
end subroutine potsetup

subroutine potential_solver
  use bh_solver_state
  implicit none
  include 'runhydro.h'
  include 'mpif.h'
  real, dimension(numr_dd,numz_dd,numphi) :: pot, rho
  common /poisson/ pot, rho
  real, dimension(numr_dd,numz_dd,numphi) :: potp, rhop
  common /potarrays/ potp, rhop
  real :: dt, time, dt_visc
  integer :: tstep
  common /timestep/ dt, time, dt_visc, tstep
  logical :: iam_on_top, iam_on_bottom, iam_on_axis, iam_on_edge, iam_root
  integer :: column_num, row_num
  integer :: iam, down_neighbor, up_neighbor, in_neighbor, out_neighbor, root, &
             REAL_SIZE, INT_SIZE, numprocs
  integer, dimension(numr_procs,numz_procs) :: pe_grid
  common /processor_grid/ iam, numprocs, iam_on_top, iam_on_bottom, iam_on_axis, &
                          iam_on_edge, down_neighbor, up_neighbor, in_neighbor,  &
                          out_neighbor, root, column_num, row_num, pe_grid,      &
                          iam_root, REAL_SIZE, INT_SIZE
  real(c_double), allocatable :: blocks(:,:,:,:), grho(:,:,:), gpot(:,:,:)
  integer :: ierror, nblk, src, pr, pz, j, k, jlo, jhi, klo, khi
  integer(c_int) :: rc

  nblk = numr_dd * numz_dd * numphi
  if (iam == root) then
     allocate(blocks(numr_dd, numz_dd, numphi, numprocs))
  else
     allocate(blocks(1, 1, 1, 1))
  endif
  call mpi_gather(rho, nblk, REAL_SIZE, blocks, nblk, REAL_SIZE, root, MPI_COMM_WORLD, ierror)

  if (iam == root) then
     allocate(grho(numr, numz, numphi), gpot(numr, numz, numphi))
     grho = 0.0
     do src = 0, numprocs - 1
        pr = mod(src, numr_procs)
        pz = src / numr_procs
        ! owned range: interior, plus the global boundary layers on edge ranks
        jlo = 2
        jhi = numr_dd - 1
        klo = 2
        khi = numz_dd - 1
        if (pr == 0) jlo = 1
        if (pr == numr_procs - 1) jhi = numr_dd
        if (pz == 0) klo = 1
        if (pz == numz_procs - 1) khi = numz_dd
        do k = klo, khi
           do j = jlo, jhi
              grho(j + pr*(numr_dd-2), k + pz*(numz_dd-2), :) = blocks(j, k, :, src + 1)
           enddo
        enddo
     enddo
     rc = bh_solve(plan, grho, gpot, int(tstep, c_int))
     if (rc /= BH_OK) then
        write(6,*) 'potential_solver: bh_solve failed: ', trim(bh_last_error())
        call mpi_abort(MPI_COMM_WORLD, 1, ierror)
     endif
     ! every rank's block, ghost layers included, from the global potential
     do src = 0, numprocs - 1
        pr = mod(src, numr_procs)
        pz = src / numr_procs
        do k = 1, numz_dd
           do j = 1, numr_dd
              blocks(j, k, :, src + 1) = gpot(j + pr*(numr_dd-2), k + pz*(numz_dd-2), :)
           enddo
        enddo
     enddo
     deallocate(grho, gpot)
  endif
  call mpi_scatter(blocks, nblk, REAL_SIZE, pot, nblk, REAL_SIZE, root, MPI_COMM_WORLD, ierror)
  deallocate(blocks)
  ! keep the reference's COMMON relation pot = potp (potential_solver.F90:106)
  potp = pot
end subroutine potential_solver
