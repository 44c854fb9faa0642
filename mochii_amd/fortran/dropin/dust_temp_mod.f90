! Drop-in for src/dust_temp_mod.f90: grain temperatures from the absorbed
! power, run by libmochii_gpu.so (src/dust_temp_mod.f90:39-230).
module dust_temp_mod
  use, intrinsic :: iso_c_binding
  use define, only : par, wp
  use mochii_gpu
  use mochii_dropin
  implicit none
  private

  public :: dust_temp_setup, dust_temp_compute

contains

  subroutine dust_temp_setup()
    call mochii_check(mc_dust_temp_setup(mochii_ctx), 'dust_temp_setup')
  end subroutine dust_temp_setup

  ! heat_dust is the library's own Heat_dust of the last pass (the bound
  ! gas_rates_mod array receives it with T_dust)
  subroutine dust_temp_compute(heat_dust)
    real(kind=wp), intent(inout) :: heat_dust(:)
    call mochii_check(mc_dust_temp_compute(mochii_ctx), 'dust_temp_compute')
    call mochii_pull_dust()
  end subroutine dust_temp_compute

end module dust_temp_mod
