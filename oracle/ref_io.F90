! REF_IO.F90
! Test infrastructure only: a harness around the SEREN reference snapshot
! readers and writers (io/read_data_*.F90, io/write_data_*.F90), compiled
! from the reference's own sources by oracle/Makefile.ref (target io).  Tests
! use it to check seren_amd.snapshot against the reference byte for byte.
!
!   * ref_io_set   ~ the module state SEREN holds before write_data: units.F90
!                    (DIMENSIONLESS scales), allocate_memory.F90, then the
!                    particle arrays and time_module counters
!   * ref_io_write ~ write_data_seren_unform / _form / write_data_ascii
!   * ref_io_read  ~ read_data_seren_unform / _form / read_data_ascii
!   * ref_io_get   ~ copies the particle arrays and counters out
! The ASCII routines read the column list from ./asciicolumns.dat, as the
! reference does; callers run them from a directory holding that file.
! ============================================================================

#include "macros.h"

MODULE ref_io
  use iso_c_binding
  use definitions
  use particle_module
  use hydro_module
  use time_module
  use type_module
  use sink_module
  use scaling_module
  use neighbour_module
  implicit none

  logical, save :: io_allocated = .false.

CONTAINS

  function c_to_f_string(cs, n) result(fs)
    integer(c_int), value :: n
    character(kind=c_char), intent(in) :: cs(n)
    character(len=n) :: fs
    integer :: i
    do i=1,n
       fs(i:i) = cs(i)
    end do
  end function c_to_f_string

  subroutine io_free()
    if (io_allocated) call clean_up
    io_allocated = .false.
  end subroutine io_free

  subroutine ref_io_set(n, pos, mass, h, vel, rho_in, temp_in, porig_in, time_in, &
       & nsteps_in, snapshot_in, hfac_in, pp_gather_in) bind(C, name="ref_io_set")
    integer(c_int), value :: n, pp_gather_in
    real(c_double), intent(in) :: pos(3,n), mass(n), h(n), vel(3,n), rho_in(n), temp_in(n)
    integer(c_int), intent(in) :: porig_in(n)
    real(c_double), value :: time_in, hfac_in
    integer(c_long_long), value :: nsteps_in, snapshot_in
    integer :: p
    call io_free()
    call units
    ptot = n
    pgas = n ; pboundary = 0 ; picm = 0 ; pcdm = 0 ; pdust = 0 ; pion = 0
    stot = 0
    pgas_orig = n
    pp_gather = pp_gather_in
    pp_limit = 2*pp_gather_in
    call allocate_memory
    io_allocated = .true.
    do p=1,n
       porig(p) = porig_in(p)
       parray(1:NDIM,p) = pos(1:NDIM,p)
       parray(MASS,p) = mass(p)
       parray(SMOO,p) = h(p)
       v(1:VDIM,p) = vel(1:VDIM,p)
       rho(p) = rho_in(p)
       temp(p) = temp_in(p)
    end do
    mgas_orig = sum(mass(1:n))
    h_fac = hfac_in
    gamma = 1.4_PR
    mu_bar = 2.35_PR
    time = time_in
    lastsnap = time_in
    nsteps = nsteps_in
    snapshot = snapshot_in
    ntempnext = nsteps_in + 1
    ndiagnext = nsteps_in + 2
    nsnapnext = nsteps_in + 3
    nsinknext = nsteps_in + 4
  end subroutine ref_io_set

  subroutine ref_io_write(kind, fname, lname) bind(C, name="ref_io_write")
    integer(c_int), value :: kind, lname
    character(kind=c_char), intent(in) :: fname(lname)
    character(len=lname) :: f
    f = c_to_f_string(fname, lname)
    select case (kind)
    case (0)
       call write_data_seren_unform(f)
    case (1)
       call write_data_seren_form(f)
    case (2)
       call write_data_ascii(f)
    end select
    close(1)
  end subroutine ref_io_write

  subroutine ref_io_read(kind, fname, lname, pp_gather_in) bind(C, name="ref_io_read")
    integer(c_int), value :: kind, lname, pp_gather_in
    character(kind=c_char), intent(in) :: fname(lname)
    character(len=lname) :: f
    call io_free()
    call units
    pp_gather = pp_gather_in
    pp_limit = 2*pp_gather_in
    ptot = 0
    f = c_to_f_string(fname, lname)
    select case (kind)
    case (0)
       call read_data_seren_unform(f)
    case (1)
       call read_data_seren_form(f)
    case (2)
       call read_data_ascii(f)
    end select
    close(1)
    io_allocated = .true.
  end subroutine ref_io_read

  subroutine ref_io_get(n, pos, mass, h, vel, rho_out, temp_out, porig_out, &
       & counts, time_out, steps) bind(C, name="ref_io_get")
    integer(c_int), value :: n
    real(c_double), intent(out) :: pos(3,n), mass(n), h(n), vel(3,n), rho_out(n), temp_out(n)
    integer(c_int), intent(out) :: porig_out(n), counts(8)
    real(c_double), intent(out) :: time_out(3)
    integer(c_long_long), intent(out) :: steps(6)
    integer :: p
    do p=1,min(n,ptot)
       porig_out(p) = porig(p)
       pos(1:NDIM,p) = parray(1:NDIM,p)
       mass(p) = parray(MASS,p)
       h(p) = parray(SMOO,p)
       vel(1:VDIM,p) = v(1:VDIM,p)
       rho_out(p) = rho(p)
       temp_out(p) = temp(p)
    end do
    counts = (/ptot, stot, pboundary, picm, pgas, pcdm, pdust, pion/)
    time_out = (/time, lastsnap, mgas_orig/)
    steps = (/snapshot, nsteps, ntempnext, ndiagnext, nsnapnext, nsinknext/)
  end subroutine ref_io_get

  subroutine ref_io_ptot(n) bind(C, name="ref_io_ptot")
    integer(c_int), intent(out) :: n
    n = ptot
  end subroutine ref_io_ptot

END MODULE ref_io
