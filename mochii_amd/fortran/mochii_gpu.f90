! ISO_C_BINDING interface to libmochii_gpu.so (include/mochii_gpu.h) for the
! reference's Fortran driver.  One interface per C entry point; mc_params is
! the bind(C) mirror of the C struct, field for field.  A main program (or a
! replacement body of ionizing_field_mod / ion_balance_mod / thermal_mod)
! uses this module and calls the entry points at pass granularity; see
! drive_stromgren.f90 for a complete iteration.
module mochii_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter, public :: MC_MAXSRC = 16
  integer(c_int), parameter, public :: MC_FP64 = 0, MC_MIXED = 1
  integer(c_int), parameter, public :: MC_GRID_CAR = 0, MC_GRID_AMR = 1
  integer(c_int), parameter, public :: MC_EXT_SPH = 1, MC_EXT_REC = 2
  integer(c_int), parameter, public :: MC_NH = 0, MC_XHI = 1, MC_XHEI = 2, MC_XHEII = 3, &
                                       MC_NE = 4, MC_TE = 5, MC_RHOKAP = 6
  integer(c_int), parameter, public :: MC_WARM_NONE = 0, MC_WARM_NEWTON = 1, MC_WARM_MLP = 2
  integer(c_int), parameter, public :: MC_DIFFUSE_THRESHOLD = 0, MC_DIFFUSE_EXPONENTIAL = 1, &
                                       MC_DIFFUSE_MILNE = 2
  integer, parameter, public :: MC_MAX_EL = 11, MC_MAX_ST = 6, MC_MAX_METAL_TR = 64, &
                                MC_MAX_COOLFIT = 8
  integer(c_int), parameter, public :: MC_CX_TWO_WAY = 0, MC_CX_METAL_ONLY = 1
  integer(c_int), parameter, public :: MC_CI_VORONOV = 0, MC_CI_DERE_HYBRID = 1
  integer(c_int), parameter, public :: MC_COOL_LOW_DENSITY = 0, MC_COOL_LOCAL_NE = 1
  ! dust thermal-infrared band: per-leaf fields and the report length
  integer(c_int), parameter, public :: MC_DUST_HEAT = 0, MC_DUST_TEMP = 1
  integer(c_int), parameter, public :: MC_DUST_HEAT_IR = 2, MC_DUST_IR_ABS = 3
  integer(c_int), parameter, public :: MC_DUSTRT_NREPORT = 10

  type, bind(C), public :: mc_params
     integer(c_int64_t) :: nphotons
     integer(c_int32_t) :: qmc_seed, precision
     real(c_double)     :: eion_min, eion_max, source_cdf_tol
     integer(c_int32_t) :: nnu_ion, ion_align_edges, nsource, ext_on
     real(c_double)     :: tstar(MC_MAXSRC), src_lum(MC_MAXSRC)
     real(c_double)     :: src_x(MC_MAXSRC), src_y(MC_MAXSRC), src_z(MC_MAXSRC)
     integer(c_int32_t) :: ext_geometry, car_walk_amr
     real(c_double)     :: ext_tstar, ext_intensity, rmax
     integer(c_int32_t) :: grid_type, nx, ny, nz
     real(c_double)     :: xmax, ymax, zmax
     real(c_double)     :: distance2cm
     real(c_double)     :: He_abund
     real(c_double)     :: te_fixed, ion_relax, te_min, te_max, gas_tol, gas_tol_te
     integer(c_int32_t) :: case_A, solve_te, conv_vol, gas_niter
     real(c_double)     :: xHI_init, xHeI_init, xHeII_init
     integer(c_int32_t) :: ion_add_dust, ion_dust_scatter
     real(c_double)     :: lambda_ref
     integer(c_int32_t) :: diffuse_field, diffuse_model
     integer(c_int32_t) :: nmetal
     real(c_double)     :: metal_eth(MC_MAX_METAL_TR)
     integer(c_int32_t) :: ion_metal_abs, metal_ne, metal_heat, metal_freefree
     integer(c_int32_t) :: charge_exchange, ci_model, cooling_model
     integer(c_int32_t) :: dust_niter
     real(c_double)     :: dust_no_photons, dust_tol
     integer(c_int32_t) :: hei_diffuse
  end type mc_params

  ! one registry element (read_element + cooling_load); C arrays [stage][k]
  ! are Fortran arrays (k, stage)
  type, bind(C), public :: mc_element
     character(kind=c_char) :: name(4)
     integer(c_int32_t) :: Z, nstage
     real(c_double)     :: abund
     real(c_double)     :: eth(MC_MAX_ST)
     integer(c_int32_t) :: photo_form(MC_MAX_ST)
     real(c_double)     :: photo(8, MC_MAX_ST)
     real(c_double)     :: ci(5, MC_MAX_ST)
     real(c_double)     :: rr(6, MC_MAX_ST)
     integer(c_int32_t) :: ndr(MC_MAX_ST)
     real(c_double)     :: dr_c(9, MC_MAX_ST), dr_e(9, MC_MAX_ST)
     integer(c_int32_t) :: has_dr2(MC_MAX_ST)
     real(c_double)     :: dr2(4, MC_MAX_ST)
     integer(c_int32_t) :: has_rr2(MC_MAX_ST)
     real(c_double)     :: rr2(2, MC_MAX_ST)
     integer(c_int32_t) :: cxi_form(MC_MAX_ST)
     real(c_double)     :: cxi(6, MC_MAX_ST)
     integer(c_int32_t) :: cxr_form(MC_MAX_ST)
     real(c_double)     :: cxr(6, MC_MAX_ST)
     integer(c_int32_t) :: ncool(MC_MAX_ST)
     real(c_double)     :: cool_A(MC_MAX_COOLFIT, MC_MAX_ST), cool_T(MC_MAX_COOLFIT, MC_MAX_ST)
  end type mc_element

  public :: mc_params_default, mc_error_message

  interface
     function mc_last_error() bind(C, name='mc_last_error') result(msg)
       import :: c_ptr
       type(c_ptr) :: msg
     end function mc_last_error

     function mc_init(p, ctx) bind(C, name='mc_init') result(rc)
       import :: c_int, c_ptr, mc_params
       type(mc_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
       integer(c_int) :: rc
     end function mc_init

     function mc_finalize(ctx) bind(C, name='mc_finalize') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mc_finalize

     function mc_grid_octree(ctx, nleaf, xleaf, yleaf, zleaf, level, half) &
          bind(C, name='mc_grid_octree') result(rc)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nleaf
       real(c_double), intent(in) :: xleaf(*), yleaf(*), zleaf(*)
       integer(c_int32_t), intent(in) :: level(*)
       real(c_double), value :: half
       integer(c_int) :: rc
     end function mc_grid_octree

     function mc_gas_state_setup(ctx, nH, xHI) bind(C, name='mc_gas_state_setup') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: nH(*)
       type(c_ptr), value :: xHI              ! c_null_ptr: par%xHI_init
       integer(c_int) :: rc
     end function mc_gas_state_setup

     function mc_state_set(ctx, field, src) bind(C, name='mc_state_set') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       real(c_double), intent(in) :: src(*)
       integer(c_int) :: rc
     end function mc_state_set

     function mc_state_get(ctx, field, dst) bind(C, name='mc_state_get') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       real(c_double), intent(out) :: dst(*)
       integer(c_int) :: rc
     end function mc_state_get

     function mc_ion_dust_table(ctx, n, lam, albedo, gcos, cext) &
          bind(C, name='mc_ion_dust_table') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       real(c_double), intent(in) :: lam(*), albedo(*), gcos(*), cext(*)
       integer(c_int) :: rc
     end function mc_ion_dust_table

     function mc_gas_opacity_fill(ctx) bind(C, name='mc_gas_opacity_fill') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mc_gas_opacity_fill

     function mc_ionizing_field_and_rates(ctx, with_peel) &
          bind(C, name='mc_ionizing_field_and_rates') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: with_peel
       integer(c_int) :: rc
     end function mc_ionizing_field_and_rates

     function mc_diffuse_info(ctx, lum, nphot) bind(C, name='mc_diffuse_info') result(rc)
       import :: c_int, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: lum
       integer(c_int64_t), intent(out) :: nphot
       integer(c_int) :: rc
     end function mc_diffuse_info

     function mc_diffuse_hei_setup(ctx, n, y, A) bind(C, name='mc_diffuse_hei_setup') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: n
       real(c_double), intent(in) :: y(*), A(*)
       integer(c_int) :: rc
     end function mc_diffuse_hei_setup

     function mc_gas_equilibrium_update(ctx, max_dx, dx_vol) &
          bind(C, name='mc_gas_equilibrium_update') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: max_dx, dx_vol
       integer(c_int) :: rc
     end function mc_gas_equilibrium_update

     function mc_cooling_setup(ctx, nfit, A, Ti) bind(C, name='mc_cooling_setup') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nfit
       real(c_double), intent(in) :: A(*), Ti(*)
       integer(c_int) :: rc
     end function mc_cooling_setup

     function mc_gas_thermal_update(ctx, max_dx, max_dte, dx_vol, dte_vol) &
          bind(C, name='mc_gas_thermal_update') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: max_dx, max_dte, dx_vol, dte_vol
       integer(c_int) :: rc
     end function mc_gas_thermal_update

     function mc_set_warm_start(ctx, mode) bind(C, name='mc_set_warm_start') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: mode
       integer(c_int) :: rc
     end function mc_set_warm_start

     function mc_hydrogen_balance_residual_grid(ctx, eps_max, eps_vol) &
          bind(C, name='mc_hydrogen_balance_residual_grid') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: eps_max, eps_vol
       integer(c_int) :: rc
     end function mc_hydrogen_balance_residual_grid

     function mc_tallies_get(ctx, hhe, jt_ion, dust_abs, segments) &
          bind(C, name='mc_tallies_get') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: hhe, jt_ion, dust_abs, segments   ! any may be c_null_ptr
       integer(c_int) :: rc
     end function mc_tallies_get

     function mc_rates_get(ctx, rates6) bind(C, name='mc_rates_get') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: rates6(*)
       integer(c_int) :: rc
     end function mc_rates_get

     function mc_energy_ledger(ctx, out3) bind(C, name='mc_energy_ledger') result(rc)
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out3(3)
       integer(c_int) :: rc
     end function mc_energy_ledger

     function mc_info(ctx, nleaf, nnu_band, Ltot, eedge) bind(C, name='mc_info') result(rc)
       import :: c_int, c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: nleaf, nnu_band
       real(c_double), intent(out) :: Ltot
       type(c_ptr), value :: eedge
       integer(c_int) :: rc
     end function mc_info

     function mc_set_precision(ctx, precision) bind(C, name='mc_set_precision') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: precision
       integer(c_int) :: rc
     end function mc_set_precision

     function mc_species_setup(ctx, n_elements, elements) bind(C, name='mc_species_setup') &
          result(rc)
       import :: c_int, c_ptr, mc_element
       type(c_ptr), value :: ctx
       integer(c_int), value :: n_elements
       type(mc_element), intent(in) :: elements(*)
       integer(c_int) :: rc
     end function mc_species_setup

     function mc_species_fractions_get(ctx, ie, frac) bind(C, name='mc_species_fractions_get') &
          result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ie
       real(c_double), intent(out) :: frac(*)
       integer(c_int) :: rc
     end function mc_species_fractions_get

     function mc_species_rates_get(ctx, ie, gam, heat) bind(C, name='mc_species_rates_get') &
          result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, gam, heat
       integer(c_int), value :: ie
       integer(c_int) :: rc
     end function mc_species_rates_get

     function mc_species_tallies_get(ctx, ie, raw) bind(C, name='mc_species_tallies_get') &
          result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ie
       real(c_double), intent(out) :: raw(*)
       integer(c_int) :: rc
     end function mc_species_tallies_get
     ! dust_temp_setup / dust_temp_compute / dust_emis_transport_setup /
     ! transport_dust_emission (src/dust_temp_mod.f90,
     ! src/dust_emis_transport_mod.f90) on the device-resident heat_dust
     function mc_dust_temp_setup(ctx) bind(C, name='mc_dust_temp_setup') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mc_dust_temp_setup
     function mc_dust_temp_compute(ctx) bind(C, name='mc_dust_temp_compute') result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mc_dust_temp_compute
     function mc_dust_emis_transport_setup(ctx) bind(C, name='mc_dust_emis_transport_setup') &
          result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: rc
     end function mc_dust_emis_transport_setup
     function mc_transport_dust_emission(ctx, report) bind(C, name='mc_transport_dust_emission') &
          result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: report(*)
       integer(c_int) :: rc
     end function mc_transport_dust_emission
     function mc_dust_get(ctx, field, dst) bind(C, name='mc_dust_get') result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       real(c_double), intent(out) :: dst(*)
       integer(c_int) :: rc
     end function mc_dust_get
     function mc_dust_set(ctx, field, src) bind(C, name='mc_dust_set') result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       real(c_double), intent(in) :: src(*)
       integer(c_int) :: rc
     end function mc_dust_set
     function mc_dust_ir_energy_budget(ctx, out3) bind(C, name='mc_dust_ir_energy_budget') &
          result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out3(3)
       integer(c_int) :: rc
     end function mc_dust_ir_energy_budget
     function mc_dust_ir_escaped_spectrum(ctx, lam, Llam, dlam) &
          bind(C, name='mc_dust_ir_escaped_spectrum') result(rc)
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: lam(*), Llam(*), dlam(*)
       integer(c_int) :: rc
     end function mc_dust_ir_escaped_spectrum
  end interface

  public :: mc_last_error, mc_init, mc_finalize, mc_grid_octree, mc_gas_state_setup
  public :: mc_state_set, mc_state_get, mc_ion_dust_table, mc_gas_opacity_fill
  public :: mc_ionizing_field_and_rates, mc_diffuse_info, mc_gas_equilibrium_update
  public :: mc_diffuse_hei_setup
  public :: mc_cooling_setup, mc_gas_thermal_update, mc_set_warm_start
  public :: mc_hydrogen_balance_residual_grid, mc_tallies_get, mc_rates_get
  public :: mc_energy_ledger, mc_info, mc_set_precision
  public :: mc_species_setup, mc_species_fractions_get, mc_species_rates_get
  public :: mc_species_tallies_get
  public :: mc_dust_temp_setup, mc_dust_temp_compute, mc_dust_emis_transport_setup
  public :: mc_transport_dust_emission, mc_dust_get, mc_dust_set
  public :: mc_dust_ir_energy_budget, mc_dust_ir_escaped_spectrum

contains

  ! The reference's defaults (src/define.f90) for the fields of mc_params.
  function mc_params_default() result(p)
    type(mc_params) :: p
    p%nphotons = 100000_c_int64_t;  p%qmc_seed = 12345;  p%precision = MC_FP64
    p%eion_min = 13.598d0;  p%eion_max = 100d0;  p%source_cdf_tol = 1d-8
    p%nnu_ion = 16;  p%ion_align_edges = 1;  p%nsource = 1;  p%ext_on = 0
    p%tstar = 0d0;  p%src_lum = 0d0;  p%src_x = 0d0;  p%src_y = 0d0;  p%src_z = 0d0
    p%ext_geometry = MC_EXT_REC;  p%car_walk_amr = 0
    p%ext_tstar = -999d0;  p%ext_intensity = -1d0;  p%rmax = -999d0
    p%grid_type = MC_GRID_CAR;  p%nx = 1;  p%ny = 1;  p%nz = 11
    p%xmax = 1d0;  p%ymax = 1d0;  p%zmax = 1d0;  p%distance2cm = 3.0856776d21
    p%He_abund = 0.1d0
    p%te_fixed = 1d4;  p%ion_relax = 1d0;  p%te_min = 1d3;  p%te_max = 5d4
    p%gas_tol = 1d-3;  p%gas_tol_te = 1d-3
    p%case_A = 0;  p%solve_te = 0;  p%conv_vol = 0;  p%gas_niter = 0
    p%xHI_init = 1d0;  p%xHeI_init = 1d0;  p%xHeII_init = 0d0
    p%ion_add_dust = 0;  p%ion_dust_scatter = 0;  p%lambda_ref = 0.55d0
    p%diffuse_field = 0;  p%diffuse_model = MC_DIFFUSE_MILNE
    p%nmetal = 0;  p%metal_eth = 0d0
    p%ion_metal_abs = 1;  p%metal_ne = 1;  p%metal_heat = 1;  p%metal_freefree = 1
    p%charge_exchange = MC_CX_TWO_WAY;  p%ci_model = MC_CI_VORONOV
    p%cooling_model = MC_COOL_LOW_DENSITY
    p%dust_niter = 2;  p%dust_no_photons = 1d6;  p%dust_tol = 1d-3
    p%hei_diffuse = 0
  end function mc_params_default

  ! mc_last_error() as a Fortran string (for fatal_error).
  function mc_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: cp
    integer :: n
    interface
       function c_strlen(str) bind(C, name='strlen') result(len)
         import :: c_ptr, c_size_t
         type(c_ptr), value :: str
         integer(c_size_t) :: len
       end function c_strlen
    end interface
    cp = mc_last_error()
    n = int(c_strlen(cp))
    call c_f_pointer(cp, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function mc_error_message

end module mochii_gpu
