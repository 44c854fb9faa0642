! The iteration of src/main.f90:124-227 (reference) driven through the
! ISO_C_BINDING module: a pure-H Stromgren sphere on a uniform car grid at
! fixed T, transport pass -> fixed-T cell update until the volume criterion
! holds, then the final consistency pass.
!   drive_stromgren layout        prints the byte size and field offsets of
!                                 mc_params (no GPU needed)
!   drive_stromgren [n] [npk]     runs the model and prints the iteration log,
!                                 the energy ledger and the sum of Gamma_HI
!                                 over the leaves
!   drive_stromgren dust [n] [npk] [nir]
!                                 the same sphere with grains (power-law
!                                 extinction, rhokap = n_H C_ext(V) d2cm):
!                                 one ionizing pass, then the block of
!                                 src/main.f90:242-260 (dust_temp_setup /
!                                 _compute, dust_emis_transport_setup,
!                                 transport_dust_emission with nir packets
!                                 per pass); prints the report and the sums of
!                                 Heat_dust and T_dust
program drive_stromgren
  use, intrinsic :: iso_c_binding
  use mochii_gpu
  implicit none
  type(mc_params), target :: p
  type(mc_element), target :: el
  type(c_ptr) :: ctx
  real(c_double), allocatable :: nH(:), rates(:)
  real(c_double) :: max_dx, dx_vol, led(3)
  integer(c_int32_t) :: nleaf, nnu
  real(c_double) :: Ltot
  character(len=32) :: arg
  integer :: it, n, rc, iarg
  integer(c_int64_t) :: npk
  logical :: dust
  real(c_double) :: nir

  call get_command_argument(1, arg)
  if (trim(arg) == 'layout') then
     call print_layout()
     call element_layout()
     stop
  end if
  dust = trim(arg) == 'dust'
  if (dust) then
     call run_dust()
     stop
  end if
  n = 16;  npk = 20000_c_int64_t
  if (len_trim(arg) > 0) read(arg, *) n
  call get_command_argument(2, arg)
  if (len_trim(arg) > 0) read(arg, *) npk

  p = mc_params_default()
  p%nphotons = npk
  p%tstar(1) = 4.0d4;  p%src_lum(1) = 3.177837d38
  p%nx = n;  p%ny = n;  p%nz = n
  p%xmax = 4d0;  p%ymax = 4d0;  p%zmax = 4d0;  p%distance2cm = 3.0856776d18
  p%He_abund = 0d0;  p%gas_tol = 2d-3;  p%conv_vol = 1;  p%gas_niter = 100

  call check(mc_init(p, ctx), 'mc_init')
  call check(mc_info(ctx, nleaf, nnu, Ltot, c_null_ptr), 'mc_info')
  allocate(nH(nleaf), rates(6 * nleaf))
  nH = 100d0
  call check(mc_gas_state_setup(ctx, nH, c_null_ptr), 'mc_gas_state_setup')
  do it = 1, p%gas_niter
     call check(mc_ionizing_field_and_rates(ctx, 0), 'mc_ionizing_field_and_rates')
     call check(mc_gas_equilibrium_update(ctx, max_dx, dx_vol), 'mc_gas_equilibrium_update')
     write(*, '(a,i4,2es14.6)') 'iteration', it, max_dx, dx_vol
     if (dx_vol < p%gas_tol) exit
  end do
  call check(mc_ionizing_field_and_rates(ctx, 0), 'final pass')
  call check(mc_energy_ledger(ctx, led), 'mc_energy_ledger')
  call check(mc_rates_get(ctx, rates), 'mc_rates_get')
  write(*, '(a,3es24.16)') 'ledger', led
  write(*, '(a,es24.16)') 'sum_gamma_HI', sum(rates(1:nleaf))
  call check(mc_finalize(ctx), 'mc_finalize')

contains

  ! The dusty sphere and the dust infrared band (src/main.f90:242-260).
  subroutine run_dust()
    integer, parameter :: nt = 600
    real(c_double) :: lam(nt), alb(nt), gc(nt), cext(nt), report(MC_DUSTRT_NREPORT)
    real(c_double), allocatable :: rk(:), heat(:), td(:)
    integer :: k
    n = 16;  npk = 20000_c_int64_t;  nir = 1d5
    call get_command_argument(2, arg)
    if (len_trim(arg) > 0) read(arg, *) n
    call get_command_argument(3, arg)
    if (len_trim(arg) > 0) read(arg, *) npk
    call get_command_argument(4, arg)
    if (len_trim(arg) > 0) read(arg, *) nir
    p = mc_params_default()
    p%nphotons = npk
    p%tstar(1) = 4.0d4;  p%src_lum(1) = 3.177837d38
    p%nx = n;  p%ny = n;  p%nz = n
    p%xmax = 4d0;  p%ymax = 4d0;  p%zmax = 4d0;  p%distance2cm = 3.0856776d18
    p%He_abund = 0d0;  p%ion_add_dust = 1;  p%ion_dust_scatter = 0;  p%lambda_ref = 0.55d0
    p%dust_no_photons = nir;  p%dust_niter = 2;  p%dust_tol = 1d-12
    do k = 1, nt                       ! C_ext ~ lambda^-1.6 from 0.01 to 1e4 um
       lam(k) = 1d-2 * (1d6)**(real(k - 1, c_double) / real(nt - 1, c_double))
       alb(k) = 0.35d0;  gc(k) = 0.55d0
       cext(k) = 4.868d-22 * (lam(k) / 0.55d0)**(-1.6d0)
    end do
    call check(mc_init(p, ctx), 'mc_init')
    call check(mc_info(ctx, nleaf, nnu, Ltot, c_null_ptr), 'mc_info')
    call check(mc_ion_dust_table(ctx, nt, lam, alb, gc, cext), 'mc_ion_dust_table')
    allocate(nH(nleaf), rk(nleaf), heat(nleaf), td(nleaf))
    nH = 100d0
    rk = nH * 4.868d-22 * p%distance2cm * 3d0
    call check(mc_gas_state_setup(ctx, nH, c_null_ptr), 'mc_gas_state_setup')
    call check(mc_state_set(ctx, MC_RHOKAP, rk), 'mc_state_set')
    call check(mc_ionizing_field_and_rates(ctx, 0), 'mc_ionizing_field_and_rates')
    call check(mc_dust_temp_setup(ctx), 'dust_temp_setup')
    call check(mc_dust_temp_compute(ctx), 'dust_temp_compute')
    call check(mc_dust_emis_transport_setup(ctx), 'dust_emis_transport_setup')
    call check(mc_transport_dust_emission(ctx, report), 'transport_dust_emission')
    call check(mc_dust_get(ctx, MC_DUST_HEAT, heat), 'heat_dust')
    call check(mc_dust_get(ctx, MC_DUST_TEMP, td), 't_dust')
    write(*, '(a,10es24.16)') 'report', report
    write(*, '(a,es24.16)') 'sum_heat_dust', sum(heat)
    write(*, '(a,es24.16)') 'sum_t_dust', sum(td)
    call check(mc_finalize(ctx), 'mc_finalize')
  end subroutine run_dust

  subroutine check(rc_, what)
    integer(c_int), intent(in) :: rc_
    character(len=*), intent(in) :: what
    if (rc_ /= 0) then
       write(*, '(3a)') trim(what), ' failed: ', mc_error_message()
       error stop 1
    end if
  end subroutine check

  subroutine print_layout()
    integer(c_intptr_t) :: b
    b = transfer(c_loc(p), b)
    write(*, '(a,i0)') 'sizeof ', c_sizeof(p)
    call off('nphotons', c_loc(p%nphotons));      call off('precision', c_loc(p%precision))
    call off('eion_min', c_loc(p%eion_min));      call off('nnu_ion', c_loc(p%nnu_ion))
    call off('tstar', c_loc(p%tstar));            call off('src_z', c_loc(p%src_z))
    call off('ext_geometry', c_loc(p%ext_geometry)); call off('ext_tstar', c_loc(p%ext_tstar))
    call off('grid_type', c_loc(p%grid_type));    call off('xmax', c_loc(p%xmax))
    call off('distance2cm', c_loc(p%distance2cm)); call off('He_abund', c_loc(p%He_abund))
    call off('te_fixed', c_loc(p%te_fixed));      call off('gas_tol_te', c_loc(p%gas_tol_te))
    call off('case_A', c_loc(p%case_A));          call off('gas_niter', c_loc(p%gas_niter))
    call off('xHI_init', c_loc(p%xHI_init));      call off('ion_add_dust', c_loc(p%ion_add_dust))
    call off('lambda_ref', c_loc(p%lambda_ref));  call off('diffuse_field', c_loc(p%diffuse_field))
    call off('diffuse_model', c_loc(p%diffuse_model))
    call off('nmetal', c_loc(p%nmetal));          call off('metal_eth', c_loc(p%metal_eth))
    call off('ion_metal_abs', c_loc(p%ion_metal_abs))
    call off('cooling_model', c_loc(p%cooling_model))
    call off('dust_niter', c_loc(p%dust_niter));  call off('dust_tol', c_loc(p%dust_tol))
    call off('hei_diffuse', c_loc(p%hei_diffuse))
  contains
    subroutine off(name, q)
      character(len=*), intent(in) :: name
      type(c_ptr), intent(in) :: q
      integer(c_intptr_t) :: a
      a = transfer(q, a)
      write(*, '(a,1x,i0)') name, a - b
    end subroutine off
  end subroutine print_layout

  ! the element record: size and field offsets, prefixed "element."
  subroutine element_layout()
    integer(c_intptr_t) :: eb
    eb = transfer(c_loc(el), eb)
    write(*, '(a,i0)') 'element.sizeof ', c_sizeof(el)
    write(*, '(a,1x,i0)') 'element.abund', transfer(c_loc(el%abund), eb) - eb
    write(*, '(a,1x,i0)') 'element.photo', transfer(c_loc(el%photo), eb) - eb
    write(*, '(a,1x,i0)') 'element.ndr', transfer(c_loc(el%ndr), eb) - eb
    write(*, '(a,1x,i0)') 'element.dr2', transfer(c_loc(el%dr2), eb) - eb
    write(*, '(a,1x,i0)') 'element.rr2', transfer(c_loc(el%rr2), eb) - eb
    write(*, '(a,1x,i0)') 'element.cxr', transfer(c_loc(el%cxr), eb) - eb
    write(*, '(a,1x,i0)') 'element.ncool', transfer(c_loc(el%ncool), eb) - eb
    write(*, '(a,1x,i0)') 'element.cool_T', transfer(c_loc(el%cool_T), eb) - eb
  end subroutine element_layout

end program drive_stromgren
