!  bh_ref_harness.F90 -- driver for the reference's own potential_solver,
!  used ONLY to produce golden vectors in the build container.
!
!  It replaces main.F90 + setup_frac.F90 for the self-gravity path: it sets
!  up the 2-D process grid (the rules of main.F90:148-244), the coordinate
!  COMMONs (the formulas of setup_frac.F90:356-463), reads a global density,
!  calls the unmodified potsetup and potential_solver compiled from the
!  reference sources, gathers pot to rank 0 and writes it.
!
!  usage: [BH_REF_ISYM=1|2|3] bh_ref <rho1.bin> <pot1.bin> [<rho2.bin> <pot2.bin> [<tabprefix>]]
!    BH_REF_ISYM: the run-time symmetry flag isym (fort.7, setup_frac.F90:313),
!    1 when unset.
!    solve 1 is cold (tstep = 0), solve 2 warm (tstep = 2) from solve 1.
!    <tabprefix>: if given, each rank writes its interior blocks of tmr, bmr,
!    smz to <tabprefix>.<table>.<rank>.
!  Arrays are raw float64, Fortran order (numr, numz, numphi).
program bh_ref_harness
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
real, dimension(numr_dd) :: rhf, r, rhfinv, rinv
real, dimension(numz_dd) :: zhf
real, dimension(numphi) :: phi
common /grid/ rhf, r, rhfinv, rinv, zhf, phi
real, dimension(numr) :: rhf_g, r_g, rhfinv_g, rinv_g
real, dimension(numz) :: zhf_g
common /global_grid/ rhf_g, r_g, rhfinv_g, rinv_g, zhf_g
real :: dr, dz, dphi, drinv, dzinv, dphiinv
common /coord_differentials/ dr, dz, dphi, drinv, dzinv, dphiinv
real :: pi, grav, cfl_factor, den_cutoff, vlim_factor, viscosity
common /constants/ pi, grav, cfl_factor, den_cutoff, vlim_factor, viscosity
integer, parameter :: mmax_h = 30
real, dimension(numr_dd,numz_dd,numr,mmax_h) :: tmr, bmr
real, dimension(numr_dd,numz_dd,numz,mmax_h) :: smz
common /green_functions/ tmr, bmr, smz

real, allocatable :: gbuf(:,:,:), lbuf(:,:,:)
integer :: ierror, nargs, i, j, k, jg, kg, src, pr, pz
integer :: rank_of(0:numr_procs*numz_procs-1, 2)
character(len=512) :: frho1, fpot1, frho2, fpot2, ftab
character(len=16) :: rk, sym
real :: t0, t1, x, offset
integer :: istat(MPI_STATUS_SIZE)

call mpi_init(ierror)
REAL_SIZE = MPI_REAL8
INT_SIZE = MPI_INTEGER
call mpi_comm_rank(MPI_COMM_WORLD, iam, ierror)
numprocs = numr_procs * numz_procs
root = 0
iam_root = (iam == 0)

nargs = command_argument_count()
call get_command_argument(1, frho1)
call get_command_argument(2, fpot1)
frho2 = ''
fpot2 = ''
ftab = ''
if (nargs >= 4) then
   call get_command_argument(3, frho2)
   call get_command_argument(4, fpot2)
endif
if (nargs >= 5) call get_command_argument(5, ftab)

! process grid: rank = (column) + numr_procs*(row), column along r
do i = 0, numprocs - 1
   rank_of(i,1) = mod(i, numr_procs)
   rank_of(i,2) = i / numr_procs
enddo
do j = 1, numz_procs
   do i = 1, numr_procs
      pe_grid(i,j) = (i-1) + numr_procs*(j-1)
   enddo
enddo
column_num = rank_of(iam,1)
row_num = rank_of(iam,2)
iam_on_bottom = (row_num == 0)
iam_on_top = (row_num == numz_procs-1)
iam_on_axis = (column_num == 0)
iam_on_edge = (column_num == numr_procs-1)
up_neighbor = -1
down_neighbor = -1
in_neighbor = -1
out_neighbor = -1
if (.not. iam_on_top) up_neighbor = iam + numr_procs
if (.not. iam_on_bottom) down_neighbor = iam - numr_procs
if (.not. iam_on_axis) in_neighbor = iam - 1
if (.not. iam_on_edge) out_neighbor = iam + 1

! constants and coordinates
isym = 1
call get_environment_variable('BH_REF_ISYM', sym)
if (len_trim(sym) > 0) read(sym, *) isym
boundary_condition = 0
if (isym == 2 .or. isym == 3) boundary_condition(1) = 1
pi = acos(-1.0)
grav = 1.0
dr = 1.0 / (oldnumr - deltar_parameter)
dz = dr
dphi = 2.0 * pi * numphiinv
drinv = 1.0 / dr
dzinv = 1.0 / dz
dphiinv = 1.0 / dphi
x = 1.0
offset = column_num*(numr_dd-2)*dr
do j = 1, numr_dd
   r(j) = offset + (x - 2.0)*dr
   x = x + 1.0
enddo
rhf = r + 0.5*dr
where (r /= 0.0)
   rinv = 1.0/r
elsewhere
   rinv = 0.0
endwhere
rhfinv = 1.0/rhf
x = 1.0
if (isym /= 1) then
   offset = row_num*(numz_dd-2)*dz
   do k = 1, numz_dd
      zhf(k) = offset + (x - 1.5)*dz
      x = x + 1.0
   enddo
else
   offset = (row_num*(numz_dd-2) - numz/2)*dz
   do k = 1, numz_dd
      zhf(k) = offset + (x - 0.5)*dz
      x = x + 1.0
   enddo
endif
x = 0.0
do i = 1, numphi
   phi(i) = x*dphi
   x = x + 1.0
enddo
x = 1.0
do j = 1, numr
   r_g(j) = (x - 2.0)*dr
   x = x + 1.0
enddo
rhf_g = r_g + 0.5*dr
where (r_g /= 0.0)
   rinv_g = 1.0/r_g
elsewhere
   rinv_g = 0.0
endwhere
rhfinv_g = 1.0/rhf_g
x = 1.0
if (isym /= 1) then
   do k = 1, numz
      zhf_g(k) = (x - 1.5)*dz
      x = x + 1.0
   enddo
else
   offset = -(numz/2)*dz
   do k = 1, numz
      zhf_g(k) = offset + (x - 0.5)*dz
      x = x + 1.0
   enddo
endif

pot = 0.0
potp = 0.0
rhop = 0.0
tstep = 0

t0 = mpi_wtime()
call potsetup
t1 = mpi_wtime()
if (iam_root) write(*,'(a,f12.4)') 'potsetup_s ', t1 - t0

if (len_trim(ftab) > 0) call dump_tables()

call load_rho(frho1)
tstep = 0
call mpi_barrier(MPI_COMM_WORLD, ierror)
t0 = mpi_wtime()
call potential_solver
call mpi_barrier(MPI_COMM_WORLD, ierror)
t1 = mpi_wtime()
if (iam_root) write(*,'(a,f12.6)') 'solve_cold_s ', t1 - t0
call store_pot(fpot1)

if (len_trim(frho2) > 0) then
   call load_rho(frho2)
   tstep = 2
   call mpi_barrier(MPI_COMM_WORLD, ierror)
   t0 = mpi_wtime()
   call potential_solver
   call mpi_barrier(MPI_COMM_WORLD, ierror)
   t1 = mpi_wtime()
   if (iam_root) write(*,'(a,f12.6)') 'solve_warm_s ', t1 - t0
   call store_pot(fpot2)
endif

call mpi_finalize(ierror)

contains

subroutine load_rho(fname)
character(len=*) :: fname
integer :: u
allocate(gbuf(numr,numz,numphi))
open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old')
read(u) gbuf
close(u)
do k = 1, numz_dd
   kg = k + row_num*(numz_dd-2)
   do j = 1, numr_dd
      jg = j + column_num*(numr_dd-2)
      rho(j,k,:) = gbuf(jg,kg,:)
   enddo
enddo
deallocate(gbuf)
end subroutine load_rho

subroutine store_pot(fname)
character(len=*) :: fname
integer :: u, jlo, jhi, klo, khi
allocate(lbuf(numr_dd,numz_dd,numphi))
if (iam_root) then
   allocate(gbuf(numr,numz,numphi))
   gbuf = 0.0
   do src = 0, numprocs - 1
      if (src == 0) then
         lbuf = pot
      else
         call mpi_recv(lbuf, numr_dd*numz_dd*numphi, REAL_SIZE, src, 77, &
                       MPI_COMM_WORLD, istat, ierror)
      endif
      pr = rank_of(src,1)
      pz = rank_of(src,2)
      ! owned range: interior, plus the global boundary layers on edge ranks
      jlo = 2
      jhi = numr_dd - 1
      klo = 2
      khi = numz_dd - 1
      if (pr == 0) jlo = 1
      if (pr == numr_procs-1) jhi = numr_dd
      if (pz == 0) klo = 1
      if (pz == numz_procs-1) khi = numz_dd
      do k = klo, khi
         do j = jlo, jhi
            gbuf(j + pr*(numr_dd-2), k + pz*(numz_dd-2), :) = lbuf(j,k,:)
         enddo
      enddo
   enddo
   open(newunit=u, file=trim(fname), access='stream', form='unformatted', &
        status='replace')
   write(u) gbuf
   close(u)
   deallocate(gbuf)
else
   lbuf = pot
   call mpi_send(lbuf, numr_dd*numz_dd*numphi, REAL_SIZE, 0, 77, &
                 MPI_COMM_WORLD, ierror)
endif
deallocate(lbuf)
end subroutine store_pot

subroutine dump_tables()
integer :: u
write(rk,'(i0)') iam
open(newunit=u, file=trim(ftab)//'.tmr.'//trim(rk), access='stream', &
     form='unformatted', status='replace')
write(u) tmr(2:numr_dd-1, 2:numz_dd-1, :, :)
close(u)
open(newunit=u, file=trim(ftab)//'.bmr.'//trim(rk), access='stream', &
     form='unformatted', status='replace')
write(u) bmr(2:numr_dd-1, 2:numz_dd-1, :, :)
close(u)
open(newunit=u, file=trim(ftab)//'.smz.'//trim(rk), access='stream', &
     form='unformatted', status='replace')
write(u) smz(2:numr_dd-1, 2:numz_dd-1, :, :)
close(u)
end subroutine dump_tables

end program bh_ref_harness
