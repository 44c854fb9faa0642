! Drop-in for src/ion_balance_mod.f90's grid-level procedures: the fixed-T
! equilibrium update and the hydrogen-balance residual, run by
! libmochii_gpu.so over every leaf (src/ion_balance_mod.f90:313-434).  The
! cell-level solve_ion_cell / charge_neutrality_ne / hydrogen_balance_residual
! are used only by the reference's thermal_mod, which thermal_mod.f90 here
! replaces.
module ion_balance_mod
  use, intrinsic :: iso_c_binding
  use define, only : par, wp
  use mochii_gpu
  use mochii_dropin
  implicit none
  private

  public :: gas_equilibrium_update, hydrogen_balance_residual_grid

contains

  ! gas_equilibrium_update (src/ion_balance_mod.f90:352-434): max and
  ! volume-weighted |delta x_HII|; the new state into the bound arrays
  subroutine gas_equilibrium_update(max_dx, dx_vol)
    real(kind=wp), intent(out) :: max_dx, dx_vol
    real(c_double) :: a, b
    call mochii_check(mc_gas_equilibrium_update(mochii_ctx, a, b), 'gas_equilibrium_update')
    max_dx = real(a, wp);  dx_vol = real(b, wp)
    call mochii_pull_state()
  end subroutine gas_equilibrium_update

  ! hydrogen_balance_residual_grid (src/ion_balance_mod.f90:313-350)
  subroutine hydrogen_balance_residual_grid(eps_max, eps_vol)
    real(kind=wp), intent(out) :: eps_max, eps_vol
    real(c_double) :: a, b
    call mochii_check(mc_hydrogen_balance_residual_grid(mochii_ctx, a, b), &
                      'hydrogen_balance_residual_grid')
    eps_max = real(a, wp);  eps_vol = real(b, wp)
  end subroutine hydrogen_balance_residual_grid

end module ion_balance_mod
