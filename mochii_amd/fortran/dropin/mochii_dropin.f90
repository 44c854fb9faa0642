! Shared part of the drop-in modules (ionizing_field_mod, ion_balance_mod,
! thermal_mod, gas_opacity_mod, dust_temp_mod, dust_emis_transport_mod in this
! folder): the library context, the reference's par mapped onto mc_params,
! and the host arrays in which the reference keeps the grid, the gas state
! and the rates.  Compiled inside the reference's tree, against its module
! define (par, src/define.f90:189-861); the modules replace the reference's
! files of the same names, and src/main.f90 keeps every call it makes.
!
! The reference's state lives in gas_state_mod / octree_mod / gas_rates_mod
! arrays; the driver binds them once, after grid_create_amr and
! gas_state_setup and before ionizing_field_setup:
!     call mochii_bind_state(gas_nH, gas_xHI, gas_xHeI, gas_xHeII, gas_ne, gas_Te)
!     call mochii_bind_rates(gamma_HI, gamma_HeI, gamma_HeII, heat_HI, heat_HeI, heat_HeII)
! (plus mochii_bind_octree, mochii_bind_dust, mochii_bind_registry,
! mochii_bind_cooling and mochii_bind_band_tally where
! they apply).  Every replaced procedure leaves the bound arrays as the
! reference's would: the rates after a pass, the gas state after an update.
module mochii_dropin
  use, intrinsic :: iso_c_binding
  use define, only : par, wp
  use mochii_gpu
  implicit none
  private

  ! the library context of the run; c_null_ptr until ionizing_field_setup
  type(c_ptr), save, public :: mochii_ctx = c_null_ptr
  ! arithmetic of the pass: MC_FP64 (the reference's), or MC_MIXED; -1 takes
  ! the environment variable MOCHII_PRECISION ('fp64' / 'mixed'), else MC_FP64
  integer(c_int), save, public :: mochii_precision = -1
  ! start of the thermal cell solve: MC_WARM_NONE (the reference's bisection),
  ! MC_WARM_NEWTON; -1 takes MC_WARM_NONE with MC_FP64, MC_WARM_NEWTON with
  ! MC_MIXED
  integer(c_int), save, public :: mochii_warm_start = -1

  real(wp), pointer, save :: s_nH(:) => null(), s_xHI(:) => null(), s_xHeI(:) => null()
  real(wp), pointer, save :: s_xHeII(:) => null(), s_ne(:) => null(), s_Te(:) => null()
  real(wp), pointer, save :: s_rhokap(:) => null()
  type rate_ptr
     real(wp), pointer :: a(:) => null()
  end type rate_ptr
  type(rate_ptr), save :: rate(6)
  real(wp), pointer, save :: s_heat_dust(:) => null(), s_t_dust(:) => null()
  real(wp), pointer, save :: s_jt(:, :) => null()
  ! octree leaf list (box-centred leaf centres, levels, root half-width)
  real(c_double), allocatable, save :: o_x(:), o_y(:), o_z(:)
  integer(c_int32_t), allocatable, save :: o_level(:)
  real(c_double), save :: o_half = 0.0_c_double
  logical, save :: have_octree = .false.
  ! grain optics table (ion_dust_kext rows, ascending wavelength)
  real(c_double), allocatable, save :: d_lam(:), d_alb(:), d_g(:), d_cext(:)
  ! trace-metal registry and H I line-cooling fit, as the reference loaded them
  type(mc_element), allocatable, target, save :: m_el(:)
  real(c_double), allocatable, save :: c_A(:), c_T(:)

  public :: mochii_bind_state, mochii_bind_rates, mochii_bind_octree, mochii_bind_dust
  public :: mochii_bind_registry, mochii_bind_cooling, mochii_bind_band_tally
  public :: mochii_params_from_par, mochii_open, mochii_check
  public :: mochii_pull_state, mochii_pull_rates, mochii_pull_dust, mochii_push_state

contains

  ! gas_state_mod's arrays (src/gas_state_mod.f90:23-28) and, with grains,
  ! octree_mod's rhokap (src/octree_mod.f90:50)
  subroutine mochii_bind_state(nH, xHI, xHeI, xHeII, ne, Te, rhokap)
    real(wp), target, intent(inout) :: nH(:), xHI(:), xHeI(:), xHeII(:), ne(:), Te(:)
    real(wp), target, intent(inout), optional :: rhokap(:)
    s_nH => nH;  s_xHI => xHI;  s_xHeI => xHeI;  s_xHeII => xHeII;  s_ne => ne;  s_Te => Te
    if (present(rhokap)) s_rhokap => rhokap
  end subroutine mochii_bind_state

  ! gas_rates_mod's Gamma and heating arrays (src/gas_rates_mod.f90:36-37)
  subroutine mochii_bind_rates(gHI, gHeI, gHeII, hHI, hHeI, hHeII)
    real(wp), target, intent(inout) :: gHI(:), gHeI(:), gHeII(:), hHI(:), hHeI(:), hHeII(:)
    rate(1)%a => gHI;  rate(2)%a => gHeI;  rate(3)%a => gHeII
    rate(4)%a => hHI;  rate(5)%a => hHeI;  rate(6)%a => hHeII
  end subroutine mochii_bind_rates

  ! jtally_mod's band tally jt_ion(nnu_band, nleaf) (src/jtally_mod.f90:
  ! 330-369), which the reference's gas_rates_write turns into J_nu
  subroutine mochii_bind_band_tally(jt_ion)
    real(wp), target, intent(inout) :: jt_ion(:, :)
    s_jt => jt_ion
  end subroutine mochii_bind_band_tally

  ! The leaf list of amr_grid (octree_mod): leaf il's centre relative to the
  ! box centre, its level and the root cell's half-width (the reader's
  ! input, src/read_generic_amr.f90:34-147)
  subroutine mochii_bind_octree(x, y, z, level, half)
    real(wp), intent(in) :: x(:), y(:), z(:), half
    integer, intent(in) :: level(:)
    o_x = real(x, c_double);  o_y = real(y, c_double);  o_z = real(z, c_double)
    o_level = int(level, c_int32_t);  o_half = real(half, c_double)
    have_octree = .true.
  end subroutine mochii_bind_octree

  ! The grain optics table of par%ion_dust_kext (wavelength [um], albedo,
  ! <cos>, C_ext per H) and gas_rates_mod's heat_dust / dust_temp_mod's
  ! T_dust, which the dust modules fill
  subroutine mochii_bind_dust(lam, albedo, gcos, cext, heat_dust, t_dust)
    real(wp), intent(in) :: lam(:), albedo(:), gcos(:), cext(:)
    real(wp), target, intent(inout) :: heat_dust(:)
    real(wp), target, intent(inout), optional :: t_dust(:)
    d_lam = real(lam, c_double);  d_alb = real(albedo, c_double)
    d_g = real(gcos, c_double);  d_cext = real(cext, c_double)
    s_heat_dust => heat_dust
    if (present(t_dust)) s_t_dust => t_dust
  end subroutine mochii_bind_dust

  ! The registry species_setup read (one record per element, the fields of
  ! read_element + cooling_load, src/species_mod.f90:54-99)
  subroutine mochii_bind_registry(elements)
    type(mc_element), intent(in) :: elements(:)
    m_el = elements
  end subroutine mochii_bind_registry

  ! H I collisional-excitation fit of cooling_setup (src/cooling_mod.f90:67-97)
  subroutine mochii_bind_cooling(A, Ti)
    real(wp), intent(in) :: A(:), Ti(:)
    c_A = real(A, c_double);  c_T = real(Ti, c_double)
  end subroutine mochii_bind_cooling

  ! par (src/define.f90:189-861) as mc_params.  A lone point source without
  ! an external field takes the legacy scalars (par%tstar, par%luminosity,
  ! par%xs_point..., src/ion_band_mod.f90:212-226); otherwise the src_*
  ! arrays.  Options the library does not carry stop the run.
  function mochii_params_from_par() result(p)
    type(mc_params) :: p
    integer :: k, n
    p = mc_params_default()
    p%nphotons = int(par%nphotons, c_int64_t)
    p%qmc_seed = int(par%qmc_seed, c_int32_t)
    p%precision = precision_choice()
    p%eion_min = par%eion_min;  p%eion_max = par%eion_max
    p%source_cdf_tol = par%source_cdf_tol
    p%nnu_ion = par%nnu_ion;  p%ion_align_edges = lg(par%ion_align_edges)
    if (trim(par%source_geometry) /= 'point') &
       call refuse('source_geometry '//trim(par%source_geometry))
    if (len_trim(par%ion_spectrum) > 0 .or. len_trim(par%src_spectrum_file) > 0) &
       call refuse('tabulated source spectra (ion_spectrum / src_spectrum_file)')
    p%nsource = par%nsource
    if (par%nsource > MC_MAXSRC) call refuse('nsource above 16')
    p%ext_on = lg(par%ext_intensity > 0.0_wp)
    if (par%nsource == 1 .and. p%ext_on == 0) then
       p%tstar(1) = par%tstar;  p%src_lum(1) = par%luminosity
       p%src_x(1) = par%xs_point;  p%src_y(1) = par%ys_point;  p%src_z(1) = par%zs_point
    else
       n = par%nsource
       do k = 1, n
          p%tstar(k) = par%src_tstar(k);  p%src_lum(k) = par%src_lum(k)
          p%src_x(k) = par%src_x(k);  p%src_y(k) = par%src_y(k);  p%src_z(k) = par%src_z(k)
       end do
    end if
    if (trim(par%ext_geometry) == 'sph') then
       p%ext_geometry = MC_EXT_SPH
    else
       p%ext_geometry = MC_EXT_REC
    end if
    p%ext_tstar = par%ext_tstar;  p%ext_intensity = par%ext_intensity;  p%rmax = par%rmax
    p%car_walk_amr = lg(trim(par%car_walk) == 'amr')
    if (par%use_amr_grid .or. trim(par%grid_type) == 'amr') then
       p%grid_type = MC_GRID_AMR
    else
       p%grid_type = MC_GRID_CAR
    end if
    p%nx = par%nx;  p%ny = par%ny;  p%nz = par%nz
    p%xmax = par%xmax;  p%ymax = par%ymax;  p%zmax = par%zmax
    p%distance2cm = par%distance2cm
    p%He_abund = par%He_abund
    p%te_fixed = par%te_fixed;  p%ion_relax = par%ion_relax
    p%te_min = par%te_min;  p%te_max = par%te_max
    p%gas_tol = par%gas_tol;  p%gas_tol_te = par%gas_tol_te
    p%case_A = lg(par%case_ab == 'A' .or. par%case_ab == 'a')
    p%solve_te = lg(par%solve_te)
    p%conv_vol = lg(trim(par%conv_crit) == 'vol')
    p%gas_niter = par%gas_niter
    p%xHI_init = par%xHI_init;  p%xHeI_init = par%xHeI_init;  p%xHeII_init = par%xHeII_init
    p%ion_add_dust = lg(par%ion_add_dust);  p%ion_dust_scatter = lg(par%ion_dust_scatter)
    p%lambda_ref = par%lambda_ref
    p%diffuse_field = lg(par%diffuse_field)
    select case (trim(par%diffuse_energy_model))
    case ('threshold');   p%diffuse_model = MC_DIFFUSE_THRESHOLD
    case ('exponential'); p%diffuse_model = MC_DIFFUSE_EXPONENTIAL
    case default;         p%diffuse_model = MC_DIFFUSE_MILNE
    end select
    p%ion_metal_abs = lg(par%ion_metal_abs);  p%metal_ne = lg(par%metal_ne)
    p%metal_heat = lg(par%metal_heat);  p%metal_freefree = lg(par%metal_freefree)
    if (trim(par%charge_exchange) == 'metal_only') then
       p%charge_exchange = MC_CX_METAL_ONLY
    else
       p%charge_exchange = MC_CX_TWO_WAY
    end if
    if (trim(par%ci_model) == 'dere_hybrid') then
       p%ci_model = MC_CI_DERE_HYBRID
    else
       p%ci_model = MC_CI_VORONOV
    end if
    if (trim(par%cooling_model) == 'local_ne') then
       p%cooling_model = MC_COOL_LOCAL_NE
    else
       p%cooling_model = MC_COOL_LOW_DENSITY
    end if
    p%dust_niter = par%dust_niter;  p%dust_no_photons = par%dust_no_photons
    p%dust_tol = par%dust_tol
    p%hei_diffuse = lg(par%diffuse_field .and. par%hei_diffuse)
    ! metal thresholds of the active registry, in registry order (src/main.f90:58-82)
    p%nmetal = 0
    if (par%use_metals .and. allocated(m_el)) then
       do k = 1, size(m_el)
          if (m_el(k)%abund <= 0.0_c_double) cycle
          do n = 1, m_el(k)%nstage - 1
             if (p%nmetal >= MC_MAX_METAL_TR) call refuse('more than 64 metal transitions')
             p%nmetal = p%nmetal + 1
             p%metal_eth(p%nmetal) = m_el(k)%eth(n)
          end do
       end do
    end if
  end function mochii_params_from_par

  ! The context for par and the bound arrays: mc_init, the grid, the gas
  ! state, grains, registry and cooling.  Called by ionizing_field_setup.
  subroutine mochii_open()
    type(mc_params) :: p
    integer(c_int32_t) :: nleaf, nnu
    real(c_double) :: Ltot
    if (c_associated(mochii_ctx)) return
    if (.not. associated(s_nH)) &
       call fatal('the gas state is not bound (mochii_bind_state before ionizing_field_setup)')
    if (trim(par%launch_sequence) /= 'sobol') &
       write(*, '(a)') ' mochii: launch_sequence '''//trim(par%launch_sequence)// &
          ''' runs as the scrambled Sobol launch (the only launch of the GPU pass)'
    p = mochii_params_from_par()
    call mochii_check(mc_init(p, mochii_ctx), 'mc_init')
    if (p%grid_type == MC_GRID_AMR) then
       if (.not. have_octree) call fatal('grid_type amr needs mochii_bind_octree')
       call mochii_check(mc_grid_octree(mochii_ctx, int(size(o_x), c_int), o_x, o_y, o_z, &
                                        o_level, o_half), 'mc_grid_octree')
    end if
    call mochii_check(mc_info(mochii_ctx, nleaf, nnu, Ltot, c_null_ptr), 'mc_info')
    if (size(s_nH) /= nleaf) call fatal('bound gas arrays do not match the grid''s leaves')
    if (par%ion_add_dust) then
       if (.not. allocated(d_lam)) call fatal('ion_add_dust needs mochii_bind_dust')
       call mochii_check(mc_ion_dust_table(mochii_ctx, int(size(d_lam), c_int), d_lam, d_alb, &
                                           d_g, d_cext), 'mc_ion_dust_table')
    end if
    if (par%use_metals) then
       if (.not. allocated(m_el)) call fatal('use_metals needs mochii_bind_registry')
       call mochii_check(mc_species_setup(mochii_ctx, int(size(m_el), c_int), m_el), &
                         'mc_species_setup')
    end if
    if (p%hei_diffuse /= 0) call hei_two_photon_load()
    call mochii_check(mc_gas_state_setup(mochii_ctx, real(s_nH, c_double), c_null_ptr), &
                      'mc_gas_state_setup')
    call mochii_push_state()
    if (par%solve_te) then
       if (.not. allocated(c_A)) call fatal('solve_te needs mochii_bind_cooling')
       call mochii_check(mc_cooling_setup(mochii_ctx, int(size(c_A), c_int), c_A, c_T), &
                         'mc_cooling_setup')
    end if
    call mochii_check(mc_set_warm_start(mochii_ctx, warm_choice(p%precision)), &
                      'mc_set_warm_start')
  end subroutine mochii_open

  ! host state -> device (the reference's arrays are the truth at set-up)
  subroutine mochii_push_state()
    call set(MC_XHI, s_xHI);  call set(MC_XHEI, s_xHeI);  call set(MC_XHEII, s_xHeII)
    call set(MC_NE, s_ne);  call set(MC_TE, s_Te)
    if (associated(s_rhokap)) call set(MC_RHOKAP, s_rhokap)
  contains
    subroutine set(f, a)
      integer(c_int), intent(in) :: f
      real(wp), intent(in) :: a(:)
      call mochii_check(mc_state_set(mochii_ctx, f, real(a, c_double)), 'mc_state_set')
    end subroutine set
  end subroutine mochii_push_state

  ! device state -> the bound gas_state_mod arrays
  subroutine mochii_pull_state()
    call get(MC_XHI, s_xHI);  call get(MC_XHEI, s_xHeI);  call get(MC_XHEII, s_xHeII)
    call get(MC_NE, s_ne);  call get(MC_TE, s_Te)
  contains
    subroutine get(f, a)
      integer(c_int), intent(in) :: f
      real(wp), intent(inout) :: a(:)
      real(c_double), allocatable :: t(:)
      allocate(t(size(a)))
      call mochii_check(mc_state_get(mochii_ctx, f, t), 'mc_state_get')
      a = real(t, wp)
    end subroutine get
  end subroutine mochii_pull_state

  ! the pass's rates -> the bound gas_rates_mod arrays (and the band tally
  ! -> the bound jt_ion)
  subroutine mochii_pull_rates()
    real(c_double), allocatable, target :: r(:), jt(:, :)
    integer :: n, k
    if (associated(s_jt)) then
       allocate(jt(size(s_jt, 1), size(s_jt, 2)))
       call mochii_check(mc_tallies_get(mochii_ctx, c_null_ptr, c_loc(jt), c_null_ptr, &
                                        c_null_ptr), 'mc_tallies_get')
       s_jt = real(jt, wp)
    end if
    if (.not. associated(rate(1)%a)) return
    n = size(s_nH)
    allocate(r(6 * n))
    call mochii_check(mc_rates_get(mochii_ctx, r), 'mc_rates_get')
    do k = 1, 6
       rate(k)%a = real(r((k - 1) * n + 1:k * n), wp)
    end do
  end subroutine mochii_pull_rates

  ! Heat_dust (and T_dust when bound) -> the bound arrays
  subroutine mochii_pull_dust()
    real(c_double), allocatable :: t(:)
    if (.not. associated(s_heat_dust)) return
    allocate(t(size(s_heat_dust)))
    call mochii_check(mc_dust_get(mochii_ctx, MC_DUST_HEAT, t), 'mc_dust_get')
    s_heat_dust = real(t, wp)
    if (associated(s_t_dust)) then
       call mochii_check(mc_dust_get(mochii_ctx, MC_DUST_TEMP, t), 'mc_dust_get')
       s_t_dust = real(t, wp)
    end if
  end subroutine mochii_pull_dust

  ! hei_2ph_setup's read (src/hei_twophoton_mod.f90:105-121): the rows of
  ! par%data_dir/HeI2phot.dat up to the first that is not two numbers
  subroutine hei_two_photon_load()
    integer, parameter :: maxpts = 128
    real(c_double) :: ty(maxpts), ta(maxpts), ya, fa
    integer :: unit, ios, n
    open(newunit=unit, file=trim(par%data_dir)//'/HeI2phot.dat', status='old', iostat=ios)
    if (ios /= 0) call fatal('cannot open '//trim(par%data_dir)//'/HeI2phot.dat (par%data_dir)')
    n = 0
    do
       read(unit, *, iostat=ios) ya, fa
       if (ios /= 0) exit
       n = n + 1
       if (n > maxpts) call fatal('HeI2phot.dat has more than 128 rows')
       ty(n) = ya;  ta(n) = fa
    end do
    close(unit)
    call mochii_check(mc_diffuse_hei_setup(mochii_ctx, int(n, c_int), ty, ta), &
                      'mc_diffuse_hei_setup')
  end subroutine hei_two_photon_load

  subroutine mochii_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) call fatal(what//' failed: '//mc_error_message())
  end subroutine mochii_check

  ! ---- helpers --------------------------------------------------------------
  integer(c_int32_t) function lg(b)
    logical, intent(in) :: b
    lg = merge(1_c_int32_t, 0_c_int32_t, b)
  end function lg

  integer(c_int32_t) function precision_choice() result(pr)
    character(len=16) :: env
    integer :: n, st
    if (mochii_precision >= 0) then
       pr = mochii_precision
       return
    end if
    pr = MC_FP64
    call get_environment_variable('MOCHII_PRECISION', env, n, st)
    if (st == 0 .and. trim(env) == 'mixed') pr = MC_MIXED
  end function precision_choice

  integer(c_int) function warm_choice(pr) result(w)
    integer(c_int32_t), intent(in) :: pr
    if (mochii_warm_start >= 0) then
       w = mochii_warm_start
    else if (pr == MC_MIXED) then
       w = MC_WARM_NEWTON
    else
       w = MC_WARM_NONE
    end if
  end function warm_choice

  subroutine refuse(what)
    character(len=*), intent(in) :: what
    call fatal('the GPU pass does not carry '//what)
  end subroutine refuse

  subroutine fatal(msg)
    character(len=*), intent(in) :: msg
    write(*, '(a)') ' ERROR (mochii): '//msg
    error stop 1
  end subroutine fatal

end module mochii_dropin
