! SERENX_DRIVERS.F90
! Drop-in replacements of SEREN's force-path driver subroutines.  A SEREN
! build that compiles this file instead of src/main/tree_update.F90,
! src/sph_sim/sph_update.F90, src/sph_sim/sph_hydro_forces.F90 and
! src/sph_sim/sph_grav_forces.F90 (and adds BHhydro_hguess below in place of
! src/BHtree/BHhydro_hguess.F90) runs its force evaluation on the GPU through
! libserenx.so.  The module arrays are bound by address before every call,
! so reallocation between steps is harmless.  Option set: GRAD_H_SPH (or
! SPH=STANDARD with HFIND=NUMBER), M4 or QUINTIC (TC or not),
! ISOTHERMAL or ENERGY_EQN (INTERNAL_ENERGY: du_dt comes back from
! sph_hydro_forces; thermal and the advance stay on the host), AB viscosity,
! KS gravity, BH tree, REORDER_PARTICLES, NEIGHBOUR_LISTS, DOUBLE_PRECISION,
! NDIM=3.
! ============================================================================

#include "macros.h"

MODULE serenx_seren_glue
  use serenx_f
  implicit none
CONTAINS

  ! Create the context on first use from SEREN's run-time parameters.
  subroutine serenx_seren_bind()
    use particle_module
    use hydro_module
    use neighbour_module
    use time_module, only : accdo
    use tree_module, only : thetamaxsqd, abserror
    type(serenx_options) :: opt
    type(serenx_host_state) :: st
    if (.not. c_associated(serenx_ctx)) then
       call serenx_f_check(serenx_default_options(opt), "default_options")
#if defined(QUINTIC_KERNEL) && defined(TC_KERNEL)
       opt%kernel = SERENX_KERNEL_QUINTIC_TC
#elif defined(QUINTIC_KERNEL)
       opt%kernel = SERENX_KERNEL_QUINTIC
#elif defined(TC_KERNEL)
       opt%kernel = SERENX_KERNEL_M4_TC
#endif
#if defined(OCTUPOLE)
       opt%multipole = SERENX_OCTUPOLE
#else
       opt%multipole = SERENX_QUADRUPOLE
#endif
#if defined(GADGET2_MAC)
       opt%mac = SERENX_MAC_GADGET2
#elif defined(GADGET_MAC)
       opt%mac = SERENX_MAC_GADGET
#elif defined(EIGEN_MAC)
       opt%mac = SERENX_MAC_EIGEN
#else
       opt%mac = SERENX_MAC_GEOMETRIC
#endif
       opt%h_fac = h_fac
       opt%thetamaxsqd = thetamaxsqd
       opt%abserror = abserror
       opt%alpha = alpha
#if defined(INTERNAL_ENERGY)
       opt%thermal = SERENX_THERMAL_ENERGY_EQN
       opt%gamma = gamma
#endif
#if !defined(GRAD_H_SPH)
       ! SPH=STANDARD, HFIND=NUMBER: h from the pp_gather-th neighbour
       opt%sph = SERENX_SPH_STANDARD
       opt%pp_gather = pp_gather
#endif
#if defined(CONSTANT_H)
       ! HFIND=CONSTANT: every h is hmin (sph_update.F90:107-109)
       opt%hfind = SERENX_HFIND_CONSTANT
       opt%hmin = hmin
#endif
       call serenx_f_check(serenx_create(opt, 0, serenx_ctx), "create")
    end if
    st%ptot = ptot
    st%pp_limit = pp_limit
    ! the module arrays are not TARGETs: take their addresses through
    ! assumed-size TARGET dummies (contiguous allocatables, passed by address)
#if defined(GRAD_H_SPH)
    call addresses(parray, v, a, gpot, rho, omega, press, sound, div_v, drhodt, &
         & pptot, pplist, accdo)
#else
    ! no omega without grad-h: rho stands in for the unused pointer, then NULL
    call addresses(parray, v, a, gpot, rho, rho, press, sound, div_v, drhodt, &
         & pptot, pplist, accdo)
    st%omega = c_null_ptr
#endif
#if !defined(GEOMETRIC_MAC)
    call address_of(agravmag, st%agravmag)
#else
    st%agravmag = c_null_ptr
#endif
#if defined(GRAD_H_SPH)
    call address_of_scalar(rextent, st%rextent)
#else
    st%rextent = c_null_ptr
#endif
#if defined(INTERNAL_ENERGY)
    call address_of(du_dt, st%du_dt)
#else
    st%du_dt = c_null_ptr
#endif
    call serenx_f_check(serenx_bind_host_state(serenx_ctx, st), "bind_host_state")
  contains
    subroutine addresses(x1, x2, x3, x4, x5, x6, x7, x8, x9, x10, i1, i2, l1)
      real(c_double), target, intent(inout) :: x1(*), x2(*), x3(*), x4(*), x5(*), &
           & x6(*), x7(*), x8(*), x9(*), x10(*)
      integer(c_int32_t), target, intent(inout) :: i1(*), i2(*)
      logical, target, intent(inout) :: l1(*)
      st%parray = c_loc(x1)
      st%v = c_loc(x2)
      st%a = c_loc(x3)
      st%gpot = c_loc(x4)
      st%rho = c_loc(x5)
      st%omega = c_loc(x6)
      st%press = c_loc(x7)
      st%sound = c_loc(x8)
      st%div_v = c_loc(x9)
      st%drhodt = c_loc(x10)
      st%pptot = c_loc(i1)
      st%pplist = c_loc(i2)
      st%accdo = c_loc(l1)
    end subroutine addresses
  end subroutine serenx_seren_bind

  subroutine address_of(x, p)
    real(c_double), target, intent(inout) :: x(*)
    type(c_ptr), intent(out) :: p
    p = c_loc(x)
  end subroutine address_of

  subroutine address_of_scalar(x, p)
    real(c_double), target, intent(inout) :: x
    type(c_ptr), intent(out) :: p
    p = c_loc(x)
  end subroutine address_of_scalar

END MODULE serenx_seren_glue


! tree_update.F90:9-92
SUBROUTINE tree_update
  use serenx_seren_glue
  use interface_module, only : reorder_particle_arrays
  use particle_module, only : ptot
  use time_module, only : nsteps, nbuild, nbuildstep, nstock
  implicit none
  integer(c_int32_t), allocatable :: dummylist(:)
  integer :: i

  call serenx_seren_bind()
  if (nbuild == nsteps) then
     allocate(dummylist(1:ptot))
     call serenx_f_check(serenx_tree_update(serenx_ctx, dummylist), "tree_update")
     ! BH_reorder_particles.F90:123: permute every particle array
     call reorder_particle_arrays(1, ptot, dummylist(1:ptot))
     deallocate(dummylist)
     nstock = nsteps
  else if (nstock == nsteps) then
     call serenx_f_check(serenx_tree_stock(serenx_ctx), "tree_stock")
  end if
  if (nbuild == nsteps) nbuild = nbuild + nbuildstep
#if defined(EULER) || defined(RUNGE_KUTTA)
  if (nstock == nsteps) nstock = nstock + 1
#else
  if (nstock == nsteps) nstock = nstock + 2
#endif
  i = 0
END SUBROUTINE tree_update


! BHhydro_hguess.F90:11-163
SUBROUTINE BHhydro_hguess
  use serenx_seren_glue
  implicit none
  call serenx_seren_bind()
  call serenx_f_check(serenx_hguess(serenx_ctx), "hguess")
END SUBROUTINE BHhydro_hguess


! sph_update.F90:10-158
SUBROUTINE sph_update
  use serenx_seren_glue
  implicit none
  call serenx_seren_bind()
  call serenx_f_check(serenx_sph_update(serenx_ctx), "sph_update")
END SUBROUTINE sph_update


! sph_hydro_forces.F90:9-120
SUBROUTINE sph_hydro_forces
  use serenx_seren_glue
  implicit none
  call serenx_seren_bind()
  call serenx_f_check(serenx_hydro_forces(serenx_ctx), "hydro_forces")
END SUBROUTINE sph_hydro_forces


! sph_grav_forces.F90:10-199
SUBROUTINE sph_grav_forces
  use serenx_seren_glue
  implicit none
  call serenx_seren_bind()
  call serenx_f_check(serenx_grav_forces(serenx_ctx), "grav_forces")
END SUBROUTINE sph_grav_forces
