!  bh_hydro_harness.F90 -- driver for the reference's own hydro routines that
!  sit beside the self-gravity solve, used ONLY to produce golden vectors in
!  the build container:
!
!    the momentum source step that consumes the potential (source.F90 with
!    make_source_temp.F90, source_bc.F90, make_source_temp_bc.F90), and
!    the continuation-file writer (ritecont_frac_new.F90).
!
!  The process grid (rank = column + numr_procs*row) and the coordinate
!  COMMONs follow the same rules as bh_ref_harness.F90 (main.F90:148-244,
!  setup_frac.F90:374-463, isym = 1).  Fields are read as global raw float64
!  arrays, Fortran order (numr, numz, numphi), into every rank's block
!  including its ghost layers, as after a halo exchange.
!
!  usage:
!    bh_hydro cont <in> <tstep>
!        reads <in>.<f>.bin for f = s t a rho tau pot tempa frac1 frac2 and
!        calls ritecont_frac_new, which writes output/conts/fort.12.<tstep>.
!        <rank> and output/conts/fort.13.<tstep> under the current folder.
!    bh_hydro source <in> <out> <source_time> <tstep> <bc1> <bc2> <bc3>
!        reads <in>.<f>.bin for f = rho p pot s t a temps tempa and
!        <in>.com.bin (dt, omega_frame, R_com, cylin_a_com(3,numphi),
!        cylin_v_com(3,numphi)), calls source(source_time) and writes the
!        gathered s t a temps tempa to <out>.<f>.bin.
program bh_hydro_harness
implicit none
include 'runhydro.h'
include 'mpif.h'

real, dimension(numr_dd,numz_dd,numphi) :: pot, rho
common /poisson/ pot, rho
real, dimension(numr_dd,numz_dd,numphi) :: s, t, a
common /kinematic/ s, t, a
real :: phi_com, R_com, R_com_inv
real, dimension(3,3) :: com
real, dimension(3) :: v_com, a_com, delt
real, dimension(3,numphi) :: cylin_v_com, cylin_a_com
common /centerofmass/ phi_com, R_com, R_com_inv, com, v_com, a_com, delt, cylin_v_com, cylin_a_com
real, dimension(numr_dd,numz_dd,numphi) :: temps, tempa
common /source_temp/ temps, tempa
real, dimension(numr_dd,numz_dd,numphi) :: p, tau
common /thermo/ p, tau
real, dimension(numr_dd,numz_dd,numphi) :: frac1, frac2
common /fluid_frac/ frac1, frac2
real, dimension(numr_dd) :: rhf, r, rhfinv, rinv
real, dimension(numz_dd) :: zhf
real, dimension(numphi) :: phi
common /grid/ rhf, r, rhfinv, rinv, zhf, phi
real :: dr, dz, dphi, drinv, dzinv, dphiinv
common /coord_differentials/ dr, dz, dphi, drinv, dzinv, dphiinv
real :: pin, gamma, kappa1, kappa2, gammainv
common /polytrope/ pin, gamma, kappa1, kappa2, gammainv
real :: dt, time, dt_visc
integer :: tstep
common /timestep/ dt, time, dt_visc, tstep
real :: omega_frame, cirp, scf_omega
common /rotframe/ omega_frame, cirp, scf_omega
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

real, allocatable :: gbuf(:,:,:), lbuf(:,:,:), cbuf(:)
integer :: ierror, i, j, k, src, pr, pz, frnum, source_time
integer :: rank_of(0:numr_procs*numz_procs-1, 2)
character(len=512) :: mode, fin, fout, arg
real :: x, offset
integer :: istat(MPI_STATUS_SIZE)

call mpi_init(ierror)
REAL_SIZE = MPI_REAL8
INT_SIZE = MPI_INTEGER
call mpi_comm_rank(MPI_COMM_WORLD, iam, ierror)
numprocs = numr_procs * numz_procs
root = 0
iam_root = (iam == 0)

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

isym = 1
boundary_condition = 0
dr = 1.0 / (oldnumr - deltar_parameter)
dz = dr
dphi = 2.0 * acos(-1.0) * numphiinv
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
offset = (row_num*(numz_dd-2) - numz/2)*dz
do k = 1, numz_dd
   zhf(k) = offset + (x - 0.5)*dz
   x = x + 1.0
enddo
x = 0.0
do i = 1, numphi
   phi(i) = x*dphi
   x = x + 1.0
enddo
pin = 1.5
gamma = 5.0/3.0
kappa1 = 1.0
kappa2 = 1.0
gammainv = 1.0/gamma
time = 0.0
dt_visc = 0.0
cirp = 0.0
scf_omega = 0.0

call get_command_argument(1, mode)
call get_command_argument(2, fin)

if (trim(mode) == 'cont') then
   call get_command_argument(3, arg)
   read(arg, *) tstep
   call load(trim(fin)//'.s.bin', s)
   call load(trim(fin)//'.t.bin', t)
   call load(trim(fin)//'.a.bin', a)
   call load(trim(fin)//'.rho.bin', rho)
   call load(trim(fin)//'.tau.bin', tau)
   call load(trim(fin)//'.pot.bin', pot)
   call load(trim(fin)//'.tempa.bin', tempa)
   call load(trim(fin)//'.frac1.bin', frac1)
   call load(trim(fin)//'.frac2.bin', frac2)
   ! the scalars of fort.13: arbitrary, distinct values
   time = 1.25
   frnum = 1017
   phi_com = 0.5
   R_com = 0.125
   do i = 1, 3
      do j = 1, 3
         com(i,j) = 10.0*i + j
      enddo
      v_com(i) = 0.1*i
      a_com(i) = 0.01*i
      delt(i) = 0.001*i
   enddo
   do i = 1, numphi
      do j = 1, 3
         cylin_a_com(j,i) = 1.0e-3*j + 1.0e-6*i
         cylin_v_com(j,i) = 2.0e-3*j - 1.0e-6*i
      enddo
   enddo
   call ritecont_frac_new(frnum)
else if (trim(mode) == 'source') then
   call get_command_argument(3, fout)
   call get_command_argument(4, arg)
   read(arg, *) source_time
   call get_command_argument(5, arg)
   read(arg, *) tstep
   do i = 1, 3
      call get_command_argument(5 + i, arg)
      read(arg, *) boundary_condition(i)
   enddo
   call load(trim(fin)//'.rho.bin', rho)
   call load(trim(fin)//'.p.bin', p)
   call load(trim(fin)//'.pot.bin', pot)
   call load(trim(fin)//'.s.bin', s)
   call load(trim(fin)//'.t.bin', t)
   call load(trim(fin)//'.a.bin', a)
   call load(trim(fin)//'.temps.bin', temps)
   call load(trim(fin)//'.tempa.bin', tempa)
   allocate(cbuf(3 + 6*numphi))
   call load_vec(trim(fin)//'.com.bin', cbuf)
   dt = cbuf(1)
   omega_frame = cbuf(2)
   R_com = cbuf(3)
   do i = 1, numphi
      do j = 1, 3
         cylin_a_com(j,i) = cbuf(3 + 3*(i-1) + j)
         cylin_v_com(j,i) = cbuf(3 + 3*numphi + 3*(i-1) + j)
      enddo
   enddo
   call source(source_time)
   call store(trim(fout)//'.s.bin', s)
   call store(trim(fout)//'.t.bin', t)
   call store(trim(fout)//'.a.bin', a)
   call store(trim(fout)//'.temps.bin', temps)
   call store(trim(fout)//'.tempa.bin', tempa)
endif

call mpi_finalize(ierror)

contains

! every rank's block, ghost layers included, from the global array
subroutine load(fname, q)
character(len=*) :: fname
real, dimension(numr_dd,numz_dd,numphi) :: q
integer :: u, jg, kg
allocate(gbuf(numr,numz,numphi))
open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old')
read(u) gbuf
close(u)
do k = 1, numz_dd
   kg = k + row_num*(numz_dd-2)
   do j = 1, numr_dd
      jg = j + column_num*(numr_dd-2)
      q(j,k,:) = gbuf(jg,kg,:)
   enddo
enddo
deallocate(gbuf)
end subroutine load

subroutine load_vec(fname, v)
character(len=*) :: fname
real, dimension(:) :: v
integer :: u
open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old')
read(u) v
close(u)
end subroutine load_vec

! gather to rank 0: each rank's interior, plus the global boundary layers
! (ghosts included) on the edge ranks
subroutine store(fname, q)
character(len=*) :: fname
real, dimension(numr_dd,numz_dd,numphi) :: q
integer :: u, jlo, jhi, klo, khi
allocate(lbuf(numr_dd,numz_dd,numphi))
if (iam_root) then
   allocate(gbuf(numr,numz,numphi))
   gbuf = 0.0
   do src = 0, numprocs - 1
      if (src == 0) then
         lbuf = q
      else
         call mpi_recv(lbuf, numr_dd*numz_dd*numphi, REAL_SIZE, src, 78, &
                       MPI_COMM_WORLD, istat, ierror)
      endif
      pr = rank_of(src,1)
      pz = rank_of(src,2)
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
   open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
   write(u) gbuf
   close(u)
   deallocate(gbuf)
else
   lbuf = q
   call mpi_send(lbuf, numr_dd*numz_dd*numphi, REAL_SIZE, 0, 78, MPI_COMM_WORLD, ierror)
endif
deallocate(lbuf)
end subroutine store

end program bh_hydro_harness
