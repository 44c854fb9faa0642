! Drop-in for src/gas_opacity_mod.f90's set-up and fill, the two procedures
! src/main.f90 calls.  The grain optics lookups (ion_dust_*_at) are used only
! by the reference's packet transport, which the library replaces.
module gas_opacity_mod
  use, intrinsic :: iso_c_binding
  use define, only : par, wp
  use mochii_gpu
  use mochii_dropin
  implicit none
  private

  public :: gas_opacity_setup, gas_opacity_fill

contains

  ! gas_opacity_setup (src/gas_opacity_mod.f90:57-215): the library builds
  ! the band opacity when the context opens (ionizing_field_setup)
  subroutine gas_opacity_setup()
  end subroutine gas_opacity_setup

  ! gas_opacity_fill (src/gas_opacity_mod.f90:299-332): the per-leaf band
  ! opacity from the current state, on the device
  subroutine gas_opacity_fill()
    call mochii_check(mc_gas_opacity_fill(mochii_ctx), 'gas_opacity_fill')
  end subroutine gas_opacity_fill

end module gas_opacity_mod
