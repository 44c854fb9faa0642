! Drop-in for src/thermal_mod.f90: the per-leaf thermal and ionisation
! update, run by libmochii_gpu.so (src/thermal_mod.f90:124-249).
module thermal_mod
  use, intrinsic :: iso_c_binding
  use define, only : par, wp
  use mochii_gpu
  use mochii_dropin
  implicit none
  private

  public :: gas_thermal_update

contains

  ! gas_thermal_update (src/thermal_mod.f90:124-249): max and
  ! volume-weighted |delta x_HII| and |delta T_e|/T_e; the new state into the
  ! bound arrays
  subroutine gas_thermal_update(max_dx, max_dte, dx_vol, dte_vol)
    real(kind=wp), intent(out) :: max_dx, max_dte, dx_vol, dte_vol
    real(c_double) :: a, b, c, d
    call mochii_check(mc_gas_thermal_update(mochii_ctx, a, b, c, d), 'gas_thermal_update')
    max_dx = real(a, wp);  max_dte = real(b, wp);  dx_vol = real(c, wp);  dte_vol = real(d, wp)
    call mochii_pull_state()
  end subroutine gas_thermal_update

end module thermal_mod
