! Drop-in for src/dust_emis_transport_mod.f90: transport of the grains'
! thermal infrared and the self-heating loop, run by libmochii_gpu.so
! (src/dust_emis_transport_mod.f90:100-370).
module dust_emis_transport_mod
  use, intrinsic :: iso_c_binding
  use define, only : par, wp
  use mochii_gpu
  use mochii_dropin
  implicit none
  private

  public :: dust_emis_transport_setup, transport_dust_emission

contains

  subroutine dust_emis_transport_setup()
    call mochii_check(mc_dust_emis_transport_setup(mochii_ctx), 'dust_emis_transport_setup')
  end subroutine dust_emis_transport_setup

  ! the passes until the emitted luminosity changes by less than
  ! par%dust_tol (at most par%dust_niter); Heat_dust and T_dust into the
  ! bound arrays, the report line as the reference prints it
  subroutine transport_dust_emission(heat_dust)
    real(kind=wp), intent(inout) :: heat_dust(:)
    real(c_double) :: rep(MC_DUSTRT_NREPORT)
    call mochii_check(mc_transport_dust_emission(mochii_ctx, rep), 'transport_dust_emission')
    write(*, '(a,i0,a,es12.4,a,es12.4,a,es10.2)') '     dust IR: ', nint(rep(1)), &
       ' passes, L_emit = ', rep(3), ', escaped = ', rep(6), ', closure = ', rep(7)
    call mochii_pull_dust()
  end subroutine transport_dust_emission

end module dust_emis_transport_mod
