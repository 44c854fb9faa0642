! Drop-in for src/ionizing_field_mod.f90: the set-up and the transport pass
! with its rate integrals, run by libmochii_gpu.so.  Same module name, same
! public procedures and arguments (src/ionizing_field_mod.f90:54-72).
module ionizing_field_mod
  use, intrinsic :: iso_c_binding
  use define, only : par, wp
  use mochii_gpu
  use mochii_dropin
  implicit none
  private

  public :: ionizing_field_setup, ionizing_field_and_rates

contains

  ! The library context: par, grid, gas state, grains, registry, cooling
  ! (ionizing_field_setup, src/ionizing_field_mod.f90:54-64, after ion_setup /
  ! jtally_ion_setup / ion_score_setup, whose work mc_init does).
  subroutine ionizing_field_setup()
    call mochii_open()
  end subroutine ionizing_field_setup

  ! One pass: launch, walk, tallies, the rate integrals (gas_rates_compute)
  ! and the energy-closure check, then the rates into the bound
  ! gas_rates_mod arrays (src/ionizing_field_mod.f90:72-330).  Observer
  ! peel-off (with_peel = .true.) is not on this path and stops the run.
  subroutine ionizing_field_and_rates(label, with_peel)
    character(len=*), intent(in)  :: label
    logical, intent(in), optional :: with_peel
    integer(c_int) :: peel
    real(c_double) :: led(3)
    peel = 0
    if (present(with_peel)) then
       if (with_peel) peel = 1
    end if
    call mochii_check(mc_ionizing_field_and_rates(mochii_ctx, peel), &
                      'ionizing_field_and_rates ('//trim(label)//')')
    call mochii_check(mc_energy_ledger(mochii_ctx, led), 'mc_energy_ledger')
    write(*, '(2a,3(a,es12.4))') '     ', trim(label), ':  L_emit = ', led(1), &
       ',  L_abs = ', led(2), ',  L_esc = ', led(3)
    call mochii_pull_rates()
  end subroutine ionizing_field_and_rates

end module ionizing_field_mod
