! REF_DRIVER.F90
! Test infrastructure only: a dump harness around the SEREN reference
! routines, compiled from the reference's own sources by oracle/Makefile.ref.
! It is never part of the product path; tests use it to write golden vectors
! and bench.py times it as the CPU baseline.
!
! The harness plays the role of SEREN's setup code:
!   * ref_setup         ~ allocate_memory.F90:43-271 (only the arrays the
!                         force path touches), types.F90:23-47,
!                         initialize_seren_variables_1.F90:30-39 (pp_gather),
!                         initialize_sph_variables_1.F90:33-110, kernel.F90
!   * ref_tree_update   ~ tree_update.F90:9-92 with nbuild == nstock == nsteps
!   * ref_hguess        ~ sph_setup.F90:46
!   * ref_sph_update    ~ sph_update.F90:10-158
!   * ref_thermal_iso   ~ thermal.F90:52-60 with the sound speed injected
!                         (P = cs^2 rho, c = cs), because DIMENSIONLESS
!                         isothermal runs otherwise force c = 1
!   * ref_hydro_forces  ~ sph_hydro_forces.F90:9-120
!   * ref_grav_forces   ~ sph_grav_forces.F90:10-199
!   * ref_direct_gravity~ direct_sph_gravity.F90:10-107 for a target list
!   * ref_step_setup    ~ initialize_sph_variables_2.F90:56-70,157-175 and
!                         initialize_seren_variables_2.F90:60-61 (step state
!                         after the setup force evaluation)
!   * ref_integrate     ~ sph_integrate.F90:41-113 (tree_update, sph_update,
!                         thermal, forces, sph_timesteps, sph_advance with
!                         RUNGE_KUTTA) without sinks, output or extra physics
! plus getters that copy module state out through plain C pointers.
! ============================================================================

#include "macros.h"

MODULE ref_driver
  use iso_c_binding
  use definitions
  use particle_module
  use hydro_module
  use kernel_module
  use neighbour_module
  use time_module
  use type_module
  use tree_module
  implicit none

  logical, save :: ref_allocated = .false.

CONTAINS

! ----------------------------------------------------------------------------
  subroutine ref_free_all()
    if (.not. ref_allocated) return
    deallocate(porig, parray, v, a, r_old, v_old, gpot)
#if !defined(GEOMETRIC_MAC)
    deallocate(agravmag)
#endif
    deallocate(pptot, pplist)
    deallocate(rho, temp, press, sound, div_v, rho_old, drhodt)
#if defined(GRAD_H_SPH)
    deallocate(omega)
#endif
    deallocate(accdo, nlast, nlevel, laststep)
    deallocate(v_half)
#if defined(CHECK_NEIGHBOUR_TIMESTEPS)
    deallocate(nminneib)
#endif
#if defined(INTERNAL_ENERGY)
    deallocate(u, u_old, du_dt)
#endif
    deallocate(first_cell_grav, last_cell_grav, BHgrav)
    deallocate(first_cell_hydro, last_cell_hydro, BHhydro)
    deallocate(cellof, BHnextptcl, whichchild, BHtemp, BHstock)
    deallocate(w0, w1, w2, w3, w4, w5, w6)
#if defined(GRAD_H_SPH)
    deallocate(wh, wg)
#endif
    ref_allocated = .false.
  end subroutine ref_free_all

! ----------------------------------------------------------------------------
  subroutine ref_setup(n, pos, mass, h, vel, hfac_in, thetamaxsqd_in, &
       & abserror_in, alpha_in) bind(C, name="ref_setup")
    integer(c_int), value :: n
    real(c_double), intent(in) :: pos(3,n), mass(n), h(n), vel(3,n)
    real(c_double), value :: hfac_in, thetamaxsqd_in, abserror_in, alpha_in
    integer :: p

    call ref_free_all()

    ptot = n
    pmax = n
    pboundary = 0 ; picm = 0 ; pgas = n ; pcdm = 0 ; pdust = 0 ; pion = 0
    call types

    h_fac = hfac_in
    pp_gather = int((4.0_PR*PI/3.0_PR)*(KERNRANGE*h_fac)**3)
    pp_limit = pp_gather*2
    thetamaxsqd = thetamaxsqd_in
    abserror = abserror_in
    alpha = alpha_in
    beta = 2.0_PR*alpha_in
    hmin = 0.0_PR

    allocate(porig(1:pmax), parray(1:DATATOT,1:pmax), v(1:VDIM,1:pmax))
    allocate(a(1:VDIM,1:pmax), r_old(1:NDIM,1:pmax), v_old(1:VDIM,1:pmax))
    allocate(gpot(1:pmax))
#if !defined(GEOMETRIC_MAC)
    allocate(agravmag(1:pmax))
#endif
    allocate(pptot(1:pmax), pplist(1:pp_limit,1:pmax))
    allocate(rho(1:pmax), temp(1:pmax), press(1:pmax), sound(1:pmax))
    allocate(div_v(1:pmax), rho_old(1:pmax), drhodt(1:pmax))
#if defined(GRAD_H_SPH)
    allocate(omega(1:pmax))
#endif
    allocate(accdo(1:pmax), nlast(1:pmax), nlevel(1:pmax), laststep(1:pmax))
    allocate(v_half(1:VDIM,1:pmax))
#if defined(CHECK_NEIGHBOUR_TIMESTEPS)
    allocate(nminneib(1:pmax))
    nminneib(1:pmax) = 0
#endif
#if defined(INTERNAL_ENERGY)
    ! energy equation: ideal gas with gamma (params default 1.4), DIMENSIONLESS
    ! units (units.F90:58 Pconst = 1); u is set by ref_set_energy
    allocate(u(1:pmax), u_old(1:pmax), du_dt(1:pmax))
    u(1:pmax) = 0.0_PR
    u_old(1:pmax) = 0.0_PR
    du_dt(1:pmax) = 0.0_PR
    gamma = 1.4_PR
    Pconst = 1.0_PR
#endif

    cmax_grav  = 2*(pgas + pcdm)
    cmax_hydro = 2*pmax
    allocate(first_cell_grav(0:LMAX), last_cell_grav(0:LMAX))
    allocate(BHgrav(0:cmax_grav))
    allocate(first_cell_hydro(0:LMAX), last_cell_hydro(0:LMAX))
    allocate(BHhydro(0:cmax_hydro))
    allocate(cellof(1:pmax), BHnextptcl(1:pmax), whichchild(1:pmax))
    allocate(BHtemp(0:cmax_hydro), BHstock(0:cmax_hydro))
    allocate(w0(0:KERNTOT), w1(0:KERNTOT), w2(0:KERNTOT), w3(0:KERNTOT))
    allocate(w4(0:KERNTOT), w5(0:KERNTOT), w6(0:KERNTOT))
#if defined(GRAD_H_SPH)
    allocate(wh(0:KERNTOT), wg(0:KERNTOT))
#endif
    ref_allocated = .true.

    ! Tree node arrays start from a defined state so that unused plist slots
    ! are reproducible in dumps (only valid slots are ever compared).
    do p=0,cmax_grav
       BHgrav(p)%leaf = 0
       BHgrav(p)%plist(1:LEAFMAX) = 0
       BHgrav(p)%nextcell = 0
       BHgrav(p)%ifopen = 0
    end do
    do p=0,cmax_hydro
       BHtemp(p)%plist(1:LEAFMAX) = 0
       BHtemp(p)%ifopen = 0
       BHtemp(p)%nextcell = 0
    end do

    do p=1,n
       porig(p) = p
       parray(1:NDIM,p) = pos(1:NDIM,p)
       parray(MASS,p) = mass(p)
       parray(SMOO,p) = h(p)
       v(1:VDIM,p) = vel(1:VDIM,p)
       r_old(1:NDIM,p) = pos(1:NDIM,p)
       v_old(1:VDIM,p) = vel(1:VDIM,p)
       rho(p) = 0.0_PR
       rho_old(p) = 0.0_PR
       temp(p) = 1.0_PR
       press(p) = 0.0_PR
       sound(p) = 0.0_PR
       drhodt(p) = 0.0_PR
       pptot(p) = 0
    end do
    pplist(:,:) = 0

    ! initialize_sph_variables_1.F90:33-110
    do p=1,ptot
       accdo(p) = .true.
       nlevel(p) = 0
       nlast(p) = 0
       laststep(p) = 0.0_DP
       if (rho(p) == 0.0_PR) rho(p) = 1.0_PR
#if defined(GRAD_H_SPH)
       omega(p) = 1.0_PR
#endif
       parray(ZETA,p) = 0.0_PR   ! DATATOT = ZETA with GRAVITY; unused without grad-h
    end do
    a(1:VDIM,1:ptot) = 0.0_PR
#if !defined(GEOMETRIC_MAC)
    agravmag(1:ptot) = 0.0_PR
#endif
    div_v(1:ptot) = 0.0_PR
    gpot(1:ptot) = 0.0_PR

    call kernel

    nsteps = 0
    nbuild = 0
    nstock = 0
    nbuildstep = 1
  end subroutine ref_setup

  ! params hmin: the smoothing length HFIND=CONSTANT sets (sph_update.F90:107-109)
  subroutine ref_set_hmin(h) bind(C, name="ref_set_hmin")
    real(c_double), value :: h
    hmin = h
  end subroutine ref_set_hmin

! ----------------------------------------------------------------------------
  subroutine ref_tree_update() bind(C, name="ref_tree_update")
    nbuild = nsteps
    nstock = nsteps
    call tree_update
  end subroutine ref_tree_update

  subroutine ref_hguess() bind(C, name="ref_hguess")
    call BHhydro_hguess
  end subroutine ref_hguess

  subroutine ref_sph_update() bind(C, name="ref_sph_update")
    call sph_update
  end subroutine ref_sph_update

  subroutine ref_thermal_iso(cs) bind(C, name="ref_thermal_iso")
    real(c_double), value :: cs
    integer :: p
#if defined(INTERNAL_ENERGY)
    ! the energy equation's ideal-gas branch (thermal.F90:108-113) through
    ! the reference's own driver; cs is not used
    call update_thermal_properties
    return
#endif
    do p=1,ptot
       press(p) = cs*cs*rho(p)
       sound(p) = cs
       temp(p) = 1.0_PR
    end do
  end subroutine ref_thermal_iso

  subroutine ref_hydro_forces() bind(C, name="ref_hydro_forces")
    call sph_hydro_forces
  end subroutine ref_hydro_forces

  subroutine ref_grav_forces() bind(C, name="ref_grav_forces")
    call sph_grav_forces
  end subroutine ref_grav_forces

  subroutine ref_direct_gravity(nt, targets, agrav, pot) &
       & bind(C, name="ref_direct_gravity")
    integer(c_int), value :: nt
    integer(c_int), intent(in) :: targets(nt)
    real(c_double), intent(out) :: agrav(3,nt), pot(nt)
    integer :: i, p
    real(kind=DP) :: agravp(1:NDIM), potp
    !$OMP PARALLEL DO SCHEDULE(DYNAMIC,16) DEFAULT(SHARED) PRIVATE(p,agravp,potp)
    do i=1,nt
       p = targets(i)
       call direct_sph_gravity(p,1.0_PR/parray(SMOO,p),parray(1:NDIM,p),&
            &agravp(1:NDIM),potp)
       agrav(1:NDIM,i) = agravp(1:NDIM)
       pot(i) = potp
    end do
    !$OMP END PARALLEL DO
  end subroutine ref_direct_gravity

#if defined(DEBUG_DIAGNOSTICS)
! ----------------------------------------------------------------------------
! diagnostics.F90 (DEBUG_DIAGNOSTICS): appends its sums to <dir>ref.diag.
! The scales are unity under DIMENSIONLESS (units.F90:37-49).
  subroutine ref_diagnostics(nc, dir) bind(C, name="ref_diagnostics")
    use filename_module, only : run_dir, run_id
    use scaling_module
    integer(c_int), value :: nc
    character(kind=c_char), intent(in) :: dir(nc)
    integer :: i
    run_dir = ""
    do i=1,nc
       run_dir(i:i) = dir(i)
    end do
    run_id = "ref"
    rscale = 1.0_DP ; mscale = 1.0_DP ; vscale = 1.0_DP ; tscale = 1.0_DP
    momscale = 1.0_DP ; angmomscale = 1.0_DP ; Escale = 1.0_DP
    call diagnostics
  end subroutine ref_diagnostics
#endif

! ----------------------------------------------------------------------------
! Setters (state between calls, e.g. |a| from a previous evaluation)
! ----------------------------------------------------------------------------
  subroutine ref_set_agravmag(n, am) bind(C, name="ref_set_agravmag")
    integer(c_int), value :: n
    real(c_double), intent(in) :: am(n)
#if !defined(GEOMETRIC_MAC)
    agravmag(1:n) = am(1:n)
#endif
  end subroutine ref_set_agravmag

  subroutine ref_set_energy(n, uin, gam) bind(C, name="ref_set_energy")
    integer(c_int), value :: n
    real(c_double), intent(in) :: uin(n)
    real(c_double), value :: gam
#if defined(INTERNAL_ENERGY)
    u(1:n) = uin(1:n)
    gamma = gam
#endif
  end subroutine ref_set_energy

  subroutine ref_get_energy(n, uout, dudt) bind(C, name="ref_get_energy")
    integer(c_int), value :: n
    real(c_double), intent(out) :: uout(n), dudt(n)
#if defined(INTERNAL_ENERGY)
    uout(1:n) = u(1:n)
    dudt(1:n) = du_dt(1:n)
#else
    uout(1:n) = 0.0_DP
    dudt(1:n) = 0.0_DP
#endif
  end subroutine ref_get_energy

  subroutine ref_set_h(n, h) bind(C, name="ref_set_h")
    integer(c_int), value :: n
    real(c_double), intent(in) :: h(n)
    parray(SMOO,1:n) = h(1:n)
  end subroutine ref_set_h

! ----------------------------------------------------------------------------
! Getters
! ----------------------------------------------------------------------------
  subroutine ref_get_params(ipar, rpar) bind(C, name="ref_get_params")
    integer(c_int), intent(out) :: ipar(8)
    real(c_double), intent(out) :: rpar(4)
    ipar(1) = pp_gather
    ipar(2) = pp_limit
    ipar(3) = ctot_grav
    ipar(4) = ltot_grav
    ipar(5) = ctot_hydro
    ipar(6) = ltot_hydro
    ipar(7) = cmax_grav
    ipar(8) = cmax_hydro
#if defined(GRAD_H_SPH)
    rpar(1) = rextent
#else
    rpar(1) = 0.0_DP
#endif
    rpar(2) = thetamaxsqd
    rpar(3) = abserror
    rpar(4) = beta
  end subroutine ref_get_params

  subroutine ref_get_levels(first, last) bind(C, name="ref_get_levels")
    integer(c_int), intent(out) :: first(0:LMAX), last(0:LMAX)
    first(0:LMAX) = first_cell_grav(0:LMAX)
    last(0:LMAX) = last_cell_grav(0:LMAX)
  end subroutine ref_get_levels

  subroutine ref_get_particles(n, porig_o, parray_o, v_o, a_o, gpot_o, &
       & rho_o, omega_o, divv_o, drhodt_o, press_o, sound_o, agravmag_o) &
       & bind(C, name="ref_get_particles")
    integer(c_int), value :: n
    integer(c_int), intent(out) :: porig_o(n)
    real(c_double), intent(out) :: parray_o(DATATOT,n), v_o(3,n), a_o(3,n)
    real(c_double), intent(out) :: gpot_o(n), rho_o(n), omega_o(n)
    real(c_double), intent(out) :: divv_o(n), drhodt_o(n), press_o(n)
    real(c_double), intent(out) :: sound_o(n), agravmag_o(n)
    porig_o(1:n) = porig(1:n)
    parray_o(1:DATATOT,1:n) = parray(1:DATATOT,1:n)
    v_o(1:3,1:n) = v(1:3,1:n)
    a_o(1:3,1:n) = a(1:3,1:n)
    gpot_o(1:n) = gpot(1:n)
    rho_o(1:n) = rho(1:n)
#if defined(GRAD_H_SPH)
    omega_o(1:n) = omega(1:n)
#else
    omega_o(1:n) = 1.0_DP
#endif
    divv_o(1:n) = div_v(1:n)
    drhodt_o(1:n) = drhodt(1:n)
    press_o(1:n) = press(1:n)
    sound_o(1:n) = sound(1:n)
#if !defined(GEOMETRIC_MAC)
    agravmag_o(1:n) = agravmag(1:n)
#else
    agravmag_o(1:n) = 0.0_DP
#endif
  end subroutine ref_get_particles

  subroutine ref_get_neib(n, lim, pptot_o, pplist_o) bind(C, name="ref_get_neib")
    integer(c_int), value :: n, lim
    integer(c_int), intent(out) :: pptot_o(n), pplist_o(lim,n)
    pptot_o(1:n) = pptot(1:n)
    pplist_o(1:lim,1:n) = pplist(1:lim,1:n)
  end subroutine ref_get_neib

  subroutine ref_get_grav_tree(nc, leaf, plist, nextcell, ifopen, r, m, &
       & dminsqd, mac, q, s, hmax) bind(C, name="ref_get_grav_tree")
    integer(c_int), value :: nc
    integer(c_int), intent(out) :: leaf(nc), plist(LEAFMAX,nc)
    integer(c_int), intent(out) :: nextcell(nc), ifopen(nc)
    real(c_double), intent(out) :: r(3,nc), m(nc), dminsqd(nc), mac(nc)
    real(c_double), intent(out) :: q(5,nc), s(10,nc), hmax(nc)
    integer :: c
    do c=0,nc-1
       leaf(c+1) = BHgrav(c)%leaf
       plist(1:LEAFMAX,c+1) = BHgrav(c)%plist(1:LEAFMAX)
       nextcell(c+1) = BHgrav(c)%nextcell
       ifopen(c+1) = BHgrav(c)%ifopen
       r(1:3,c+1) = BHgrav(c)%r(1:3)
       m(c+1) = BHgrav(c)%m
       dminsqd(c+1) = BHgrav(c)%dminsqd
#if !defined(GEOMETRIC_MAC)
       mac(c+1) = BHgrav(c)%mac
#else
       mac(c+1) = 0.0_DP
#endif
#if defined(QUADRUPOLE)
       q(1:5,c+1) = BHgrav(c)%q(1:5)
#else
       q(1:5,c+1) = 0.0_DP
#endif
#if defined(OCTUPOLE)
       s(1:10,c+1) = BHgrav(c)%s(1:10)
#else
       s(1:10,c+1) = 0.0_DP
#endif
       hmax(c+1) = BHstock(c)%hmax
    end do
  end subroutine ref_get_grav_tree

  subroutine ref_get_hydro_tree(nc, r, rmax, hrangemax) &
       & bind(C, name="ref_get_hydro_tree")
    integer(c_int), value :: nc
    real(c_double), intent(out) :: r(3,nc), rmax(nc), hrangemax(nc)
    integer :: c
    do c=0,nc-1
       r(1:3,c+1) = BHhydro(c)%r(1:3)
       rmax(c+1) = BHhydro(c)%rmax
       hrangemax(c+1) = BHhydro(c)%hrangemax
    end do
  end subroutine ref_get_hydro_tree

  subroutine ref_get_kernel(tab) bind(C, name="ref_get_kernel")
    real(c_double), intent(out) :: tab(0:KERNTOT,9)
    tab(:,1) = w0(:) ; tab(:,2) = w1(:) ; tab(:,3) = w2(:)
    tab(:,4) = w3(:) ; tab(:,5) = w4(:) ; tab(:,6) = w5(:)
#if defined(GRAD_H_SPH)
    tab(:,7) = w6(:) ; tab(:,8) = wh(:) ; tab(:,9) = wg(:)
#else
    tab(:,7) = w6(:) ; tab(:,8) = 0.0_DP ; tab(:,9) = 0.0_DP
#endif
  end subroutine ref_get_kernel

! ----------------------------------------------------------------------------
  subroutine ref_step_setup(nlevels_in, nbuildstep_in, accel_in, courant_in, dtfixed_in) &
       & bind(C, name="ref_step_setup")
    integer(c_int), value :: nlevels_in, nbuildstep_in
    real(c_double), value :: accel_in, courant_in, dtfixed_in
    r_old(1:NDIM,1:ptot) = parray(1:NDIM,1:ptot)
    v_old(1:VDIM,1:ptot) = v(1:VDIM,1:ptot)
    v_half(1:VDIM,1:ptot) = v(1:VDIM,1:ptot)
    rho_old(1:ptot) = rho(1:ptot)
#if defined(INTERNAL_ENERGY)
    u_old(1:ptot) = u(1:ptot)            ! initialize_sph_variables_2.F90:75-77
#endif
    accdo(1:ptot) = .false.
    nlevels = nlevels_in
    nbuildstep = nbuildstep_in
    accel_mult = accel_in
    courant_mult = courant_in
    dt_fixed = dtfixed_in
    n = 0
    nresync = n
    time = 0.0_DP
    nbuild = nsteps + 1
    nstock = nsteps + 1
  end subroutine ref_step_setup

  subroutine ref_integrate(cs) bind(C, name="ref_integrate")
    real(c_double), value :: cs
    integer :: p
    call tree_update
    call sph_update
    call ref_thermal_iso(cs)
    do p=1,ptot
       if (accdo(p)) a(1:VDIM,p) = 0.0_PR
    end do
    call sph_hydro_forces
    call sph_grav_forces
    call sph_timesteps
#if defined(CHECK_NEIGHBOUR_TIMESTEPS)
    call check_neighbour_timesteps
#endif
    call sph_advance
#if defined(CHECK_NEIGHBOUR_TIMESTEPS) && defined(IMMEDIATE_TIMESTEP_REDUCTION)
    call reduce_timesteps
#endif
  end subroutine ref_integrate

  subroutine ref_get_step(istate, rstate) bind(C, name="ref_get_step")
    integer(c_long_long), intent(out) :: istate(8)
    real(c_double), intent(out) :: rstate(4)
    istate(1) = n ; istate(2) = nsteps ; istate(3) = nresync
    istate(4) = level_max ; istate(5) = level_step
    istate(6) = nbuild ; istate(7) = nstock ; istate(8) = 0
    rstate(1) = time ; rstate(2) = timestep ; rstate(3) = dt_max ; rstate(4) = 0.0_DP
  end subroutine ref_get_step

  subroutine ref_get_step_particles(np, nlev, nl, lstep, acc, vh, ro, vo, nmn) &
       & bind(C, name="ref_get_step_particles")
    integer(c_int), value :: np
    integer(c_long_long), intent(out) :: nlev(np), nl(np), nmn(np)
    real(c_double), intent(out) :: lstep(np), vh(3,np), ro(3,np), vo(3,np)
    integer(c_int), intent(out) :: acc(np)
    integer :: p
    do p=1,np
#if defined(CHECK_NEIGHBOUR_TIMESTEPS)
       nmn(p) = nminneib(p)
#else
       nmn(p) = 0
#endif
       nlev(p) = nlevel(p)
       nl(p) = nlast(p)
       lstep(p) = laststep(p)
       acc(p) = merge(1, 0, accdo(p))
       vh(1:3,p) = v_half(1:3,p)
       ro(1:3,p) = r_old(1:3,p)
       vo(1:3,p) = v_old(1:3,p)
    end do
  end subroutine ref_get_step_particles

END MODULE ref_driver
