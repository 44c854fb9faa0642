! A stand-in for the reference's src/main.f90 over the drop-in modules: the
! same namelist input (&parameters par%... /, read as read_input reads it),
! a namelist-built car grid of uniform density par%nH_const in place of
! grid_create_amr, host gas-state and rate arrays in place of gas_state_mod /
! gas_rates_mod (initialised as gas_state_setup does), then the set-up and
! iteration calls of src/main.f90:112-227 under the reference's own module
! and procedure names.  Prints the iteration log, the hydrogen-balance
! residuals and sums of the final rates and state.
!   drive_main <input namelist>
program drive_main
  use define
  use ionizing_field_mod, only : ionizing_field_setup, ionizing_field_and_rates
  use ion_balance_mod,    only : gas_equilibrium_update, hydrogen_balance_residual_grid
  use thermal_mod,        only : gas_thermal_update
  use gas_opacity_mod,    only : gas_opacity_setup, gas_opacity_fill
  use mochii_dropin,      only : mochii_bind_state, mochii_bind_rates, mochii_bind_cooling, &
                                 mochii_bind_band_tally, mochii_ctx
  use mochii_gpu,         only : mc_info
  use, intrinsic :: iso_c_binding, only : c_int32_t, c_double, c_null_ptr
  implicit none

  real(kind=wp), allocatable, target :: nH(:), xHI(:), xHeI(:), xHeII(:), ne(:), Te(:)
  real(kind=wp), allocatable, target :: gHI(:), gHeI(:), gHeII(:), hHI(:), hHeI(:), hHeII(:)
  real(kind=wp), allocatable :: fitA(:), fitT(:)
  real(kind=wp), allocatable, target :: jt_ion(:, :)
  integer(c_int32_t) :: nl32, nnu_band
  real(c_double) :: ltot
  real(kind=wp) :: max_dx, max_dte, dx_vol, dte_vol, eps_hmax, eps_hvol
  logical :: converged, use_vol
  integer :: iter, niter, nleaf, il
  character(len=32) :: pass_label

  call read_namelist()
  nleaf = par%nx * par%ny * par%nz
  allocate(nH(nleaf), xHI(nleaf), xHeI(nleaf), xHeII(nleaf), ne(nleaf), Te(nleaf))
  allocate(gHI(nleaf), gHeI(nleaf), gHeII(nleaf), hHI(nleaf), hHeI(nleaf), hHeII(nleaf))
  ! gas_state_setup (src/gas_state_mod.f90:41-84) without a file column
  nH = par%nH_const
  xHI = par%xHI_init;  xHeI = par%xHeI_init;  xHeII = par%xHeII_init
  do il = 1, nleaf
     ne(il) = nH(il) * ((1.0_wp - xHI(il)) + par%He_abund * (xHeII(il) &
              + 2.0_wp * max(0.0_wp, 1.0_wp - xHeI(il) - xHeII(il))))
  end do
  Te = par%te_fixed
  call mochii_bind_state(nH, xHI, xHeI, xHeII, ne, Te)
  call mochii_bind_rates(gHI, gHeI, gHeII, hHI, hHeI, hHeII)
  if (par%solve_te) then
     call load_fit(trim(par%atomic_dir)//'/cooling_h_1.txt')
     call mochii_bind_cooling(fitA, fitT)
  end if

  call gas_opacity_setup()
  niter = max(par%gas_niter, 1)
  call ionizing_field_setup()
  ! jtally_mod's band tally, sized by the band ion_setup built
  if (mc_info(mochii_ctx, nl32, nnu_band, ltot, c_null_ptr) /= 0) stop 3
  allocate(jt_ion(nnu_band, nleaf))
  call mochii_bind_band_tally(jt_ion)
  converged = .false.
  max_dx = 0.0_wp;  max_dte = 0.0_wp;  dx_vol = 0.0_wp;  dte_vol = 0.0_wp
  do iter = 1, niter
     write(pass_label, '(a,i4)') 'iteration ', iter
     call ionizing_field_and_rates(pass_label)
     if (par%gas_niter < 1) exit
     use_vol = trim(par%conv_crit) == 'vol'
     if (par%solve_te) then
        call gas_thermal_update(max_dx, max_dte, dx_vol, dte_vol)
        if (use_vol) then
           converged = dx_vol < par%gas_tol .and. dte_vol < par%gas_tol_te
        else
           converged = max_dx < par%gas_tol .and. max_dte < par%gas_tol_te
        end if
        write(*, '(a,i4,4es24.16)') 'iteration', iter, max_dx, dx_vol, max_dte, dte_vol
     else
        call gas_equilibrium_update(max_dx, dx_vol)
        if (use_vol) then
           converged = dx_vol < par%gas_tol
        else
           converged = max_dx < par%gas_tol
        end if
        write(*, '(a,i4,2es24.16)') 'iteration', iter, max_dx, dx_vol
     end if
     call hydrogen_balance_residual_grid(eps_hmax, eps_hvol)
     write(*, '(a,i4,2es24.16)') 'residual', iter, eps_hmax, eps_hvol
     call gas_opacity_fill()
     if (converged) exit
  end do
  if (par%gas_niter >= 1) call ionizing_field_and_rates('final consistency pass')
  write(*, '(a,l2,i5)') 'converged', converged, min(iter, niter)
  write(*, '(a,6es24.16)') 'sum_rates', sum(gHI), sum(gHeI), sum(gHeII), sum(hHI), &
     sum(hHeI), sum(hHeII)
  write(*, '(a,5es24.16)') 'sum_state', sum(xHI), sum(xHeI), sum(xHeII), sum(ne), sum(Te)
  write(*, '(a,es24.16)') 'sum_jt', sum(jt_ion)

contains

  ! read_input (src/setup.f90:11-69, 578-600): the namelist, the photon
  ! count, the grid alias, case A with the diffuse field, the length unit
  subroutine read_namelist()
    character(len=256) :: fname
    integer :: unit
    namelist /parameters/ par
    if (command_argument_count() < 1) then
       write(*, '(a)') 'usage: drive_main <input namelist>'
       stop 2
    end if
    call get_command_argument(1, fname)
    open(newunit=unit, file=trim(fname), status='old')
    read(unit, parameters)
    close(unit)
    par%nphotons = int(par%no_photons, kind(par%nphotons))
    if (trim(par%grid_type) == 'uniform') par%grid_type = 'car'
    if (trim(par%grid_type) /= 'car') then
       write(*, '(a)') 'drive_main: only the namelist-built car grid'
       stop 2
    end if
    if (par%diffuse_field .and. trim(par%case_ab) /= 'A') par%case_ab = 'A'
    select case (trim(par%distance_unit))
    case ('kpc');  par%distance2cm = kpc2cm
    case ('pc');   par%distance2cm = pc2cm
    case ('');     par%distance2cm = 1.0_wp
    case default
       write(*, '(a)') 'drive_main: unknown distance_unit '//trim(par%distance_unit)
       stop 2
    end select
  end subroutine read_namelist

  ! cooling_load (src/cooling_mod.f90:98-123): '#' comment lines, the term
  ! count, then one (A_i, T_i) pair per line
  subroutine load_fit(fname)
    character(len=*), intent(in) :: fname
    character(len=256) :: line
    integer :: unit, n, i
    open(newunit=unit, file=fname, status='old')
    do
       read(unit, '(a)') line
       line = adjustl(line)
       if (len_trim(line) > 0 .and. line(1:1) /= '#') exit
    end do
    read(line, *) n
    allocate(fitA(n), fitT(n))
    do i = 1, n
       read(unit, *) fitA(i), fitT(i)
    end do
    close(unit)
  end subroutine load_fit

end program drive_main
