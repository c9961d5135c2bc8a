! SERENX_F.F90
! ISO_C_BINDING interface to libserenx.so (include/serenx.h) for a Fortran
! host.  Plain C types only; the context is an opaque c_ptr.
! ============================================================================
MODULE serenx_f
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: SERENX_OK = 0
  integer(c_int32_t), parameter :: SERENX_QUADRUPOLE = 2, SERENX_OCTUPOLE = 3, SERENX_OCTUPOLE_EXACT = 4
  integer(c_int32_t), parameter :: SERENX_MAC_GEOMETRIC = 0, SERENX_MAC_GADGET = 1, &
       & SERENX_MAC_GADGET2 = 2, SERENX_MAC_EIGEN = 3
  integer(c_int32_t), parameter :: SERENX_KERNEL_M4 = 0, SERENX_KERNEL_M4_TC = 1, &
       & SERENX_KERNEL_QUINTIC = 2, SERENX_KERNEL_QUINTIC_TC = 3
  integer(c_int32_t), parameter :: SERENX_FP64 = 0, SERENX_MIXED = 1
  integer(c_int32_t), parameter :: SERENX_THERMAL_ISOTHERMAL = 0, SERENX_THERMAL_ENERGY_EQN = 1
  integer(c_int32_t), parameter :: SERENX_SPH_GRADH = 0, SERENX_SPH_STANDARD = 1
  integer(c_int32_t), parameter :: SERENX_HFIND_NUMBER = 0, SERENX_HFIND_CONSTANT = 1

  type, bind(C) :: serenx_options
     integer(c_int32_t) :: multipole, mac, precision, kernel
     real(c_double) :: h_fac, thetamaxsqd, abserror, alpha, sound_speed
     integer(c_int32_t) :: thermal, sph
     real(c_double) :: gamma
     integer(c_int32_t) :: pp_gather, serial, hfind
     real(c_double) :: hmin
  end type serenx_options

  type, bind(C) :: serenx_host_state
     integer(c_int32_t) :: ptot, pp_limit
     type(c_ptr) :: parray, v, a, gpot, agravmag, rho, omega, press, sound, &
          & div_v, drhodt, pptot, pplist, accdo, rextent, du_dt
  end type serenx_host_state

  interface
     integer(c_int) function serenx_default_options(opt) bind(C, name="serenx_default_options")
       import :: c_int, serenx_options
       type(serenx_options), intent(out) :: opt
     end function
     integer(c_int) function serenx_create(opt, device, ctx) bind(C, name="serenx_create")
       import :: c_int, c_ptr, serenx_options
       type(serenx_options), intent(in) :: opt
       integer(c_int), value :: device
       type(c_ptr), intent(out) :: ctx
     end function
     integer(c_int) function serenx_destroy(ctx) bind(C, name="serenx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     type(c_ptr) function serenx_last_error() bind(C, name="serenx_last_error")
       import :: c_ptr
     end function
     integer(c_int) function serenx_bind_host_state(ctx, st) bind(C, name="serenx_bind_host_state")
       import :: c_int, c_ptr, serenx_host_state
       type(c_ptr), value :: ctx
       type(serenx_host_state), intent(in) :: st
     end function
     integer(c_int) function serenx_tree_update(ctx, dummylist) bind(C, name="serenx_tree_update")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: dummylist(*)
     end function
     integer(c_int) function serenx_tree_stock(ctx) bind(C, name="serenx_tree_stock")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function serenx_hguess(ctx) bind(C, name="serenx_hguess")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function serenx_sph_update(ctx) bind(C, name="serenx_sph_update")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function serenx_hydro_forces(ctx) bind(C, name="serenx_hydro_forces")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function serenx_grav_forces(ctx) bind(C, name="serenx_grav_forces")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
  end interface

  type(c_ptr), save :: serenx_ctx = c_null_ptr

CONTAINS

  subroutine serenx_f_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc == SERENX_OK) return
    call c_f_pointer(serenx_last_error(), msg, [512])
    write(6,*) "serenx error in ", what, " :"
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       write(6,'(a)',advance='no') msg(i)
    end do
    write(6,*)
    stop 'serenx failure'
  end subroutine serenx_f_check

END MODULE serenx_f
