! C-callable entry points into the reference's own Fortran modules
! (qmc_mod, energy_sampler_mod, photo_xsec, recomb_mod, gaunt,
! milne_recomb_spectrum_mod, nlevel_mod, nlevel_cooling_mod, hei_cascade_mod),
! compiled from
! their sources under the reference checkout by oracle/Makefile.  Test
! infrastructure only: the tests use it to check the C oracle against the
! reference code itself.
module ref_bind
  use, intrinsic :: iso_c_binding
  use define, only : wp, par
  use qmc_mod, only : qmc_setup, qmc_uniforms_stream, qmc_uniforms_raw
  use energy_sampler_mod, only : energy_sampler_type, build_planck_sampler, &
       build_tabulated_sampler, sample_energy_cdf, energy_cdf_value
  use photo_xsec, only : sigma_HI, sigma_HeI, sigma_HeII, sigma_vfky96, sigma_vy95
  use recomb_mod
  use milne_recomb_spectrum_mod, only : milne_setup, milne_sample_energy, &
       milne_energy_per_recomb
  use nlevel_mod, only : nlevel_atom_type, nlevel_load, nlevel_emissivities, &
       nlevel_cooling_at, nlevel_nlines, nlevel_line_ident
  use nlevel_cooling_mod, only : nlevel_cooling_add, nlevel_cooling_apply, &
       nlevel_cooling_lam_se
  use hei_cascade_mod, only : hei_branch_fractions, hei_584_ionizes_H
  implicit none
  type(energy_sampler_type), save :: smp
  type(nlevel_atom_type), save :: nla

  interface
     subroutine getGauntFF(z, log10Te, xlf, g, iflag)
        implicit none
        integer, intent(out) :: iflag
        real, intent(in)  :: log10Te, z
        real, dimension(:), intent(in)  :: xlf
        real, dimension(size(xlf)), intent(out) :: g
     end subroutine getGauntFF
  end interface

contains

  subroutine ref_qmc_setup(seed) bind(C, name='ref_qmc_setup')
    integer(c_int), value :: seed
    call qmc_setup(int(seed))
  end subroutine ref_qmc_setup

  subroutine ref_qmc_uniforms(idx, n, stream, u) bind(C, name='ref_qmc_uniforms')
    integer(c_int64_t), value :: idx
    integer(c_int), value :: n, stream
    real(c_double), intent(out) :: u(n)
    call qmc_uniforms_stream(int(idx, 8), u, int(stream))
  end subroutine ref_qmc_uniforms

  subroutine ref_qmc_raw(idx, n, u) bind(C, name='ref_qmc_raw')
    integer(c_int64_t), value :: idx
    integer(c_int), value :: n
    real(c_double), intent(out) :: u(n)
    call qmc_uniforms_raw(int(idx, 8), u)
  end subroutine ref_qmc_raw

  function ref_build_planck(T, emin, emax, rtol, thr, nthr) result(n) &
       bind(C, name='ref_build_planck')
    real(c_double), value :: T, emin, emax, rtol
    integer(c_int), value :: nthr
    real(c_double), intent(in) :: thr(nthr)
    integer(c_int) :: n
    call build_planck_sampler(smp, T, emin, emax, rtol, thr)
    n = size(smp%energy)
  end function ref_build_planck

  function ref_build_tabulated(e, f, n) result(m) bind(C, name='ref_build_tabulated')
    integer(c_int), value :: n
    real(c_double), intent(in) :: e(n), f(n)
    integer(c_int) :: m
    call build_tabulated_sampler(smp, e, f)
    m = size(smp%energy)
  end function ref_build_tabulated

  subroutine ref_sampler_get(e, d, c, total) bind(C, name='ref_sampler_get')
    real(c_double), intent(out) :: e(*), d(*), c(*), total
    integer :: n
    n = size(smp%energy)
    e(1:n) = smp%energy;  d(1:n) = smp%density;  c(1:n) = smp%cdf
    total = smp%total_integral
  end subroutine ref_sampler_get

  real(c_double) function ref_sample_energy(u) bind(C, name='ref_sample_energy')
    real(c_double), value :: u
    ref_sample_energy = sample_energy_cdf(smp, u)
  end function ref_sample_energy

  real(c_double) function ref_cdf_value(e) bind(C, name='ref_cdf_value')
    real(c_double), value :: e
    ref_cdf_value = energy_cdf_value(smp, e)
  end function ref_cdf_value

  real(c_double) function ref_sigma(which, e) bind(C, name='ref_sigma')
    integer(c_int), value :: which
    real(c_double), value :: e
    select case (which)
    case (0); ref_sigma = sigma_HI(e)
    case (1); ref_sigma = sigma_HeI(e)
    case default; ref_sigma = sigma_HeII(e)
    end select
  end function ref_sigma

  real(c_double) function ref_rate(which, T) bind(C, name='ref_rate')
    integer(c_int), value :: which
    real(c_double), value :: T
    par%recomb_model = 'badnell_milne'
    par%ci_model = 'voronov'
    select case (which)
    case (0); ref_rate = alphaA_HII(T)
    case (1); ref_rate = alphaB_HII(T)
    case (2); ref_rate = alphaA_HeII(T)
    case (3); ref_rate = alphaB_HeII(T)
    case (4); ref_rate = alphaA_HeIII(T)
    case (5); ref_rate = alphaB_HeIII(T)
    case (6); ref_rate = ci_HI(T)
    case (7); ref_rate = ci_HeI(T)
    case default; ref_rate = ci_HeII(T)
    end select
  end function ref_rate

  ! ci_HI / ci_HeI / ci_HeII (which 6, 7, 8) with par%ci_model = 'dere_hybrid'
  real(c_double) function ref_rate_dere(which, T) bind(C, name='ref_rate_dere')
    integer(c_int), value :: which
    real(c_double), value :: T
    par%ci_model = 'dere_hybrid'
    select case (which)
    case (6); ref_rate_dere = ci_HI(T)
    case (7); ref_rate_dere = ci_HeI(T)
    case default; ref_rate_dere = ci_HeII(T)
    end select
    par%ci_model = 'voronov'
  end function ref_rate_dere

  real(c_double) function ref_vfky96(e, p) bind(C, name='ref_vfky96')
    real(c_double), value :: e
    real(c_double), intent(in) :: p(8)
    ref_vfky96 = sigma_vfky96(e, p(1), p(2), p(3), p(4), p(5), p(6), p(7), p(8))
  end function ref_vfky96

  real(c_double) function ref_vy95(e, p) bind(C, name='ref_vy95')
    real(c_double), value :: e
    real(c_double), intent(in) :: p(7)
    ref_vy95 = sigma_vy95(e, p(1), p(2), p(3), p(4), p(5), p(6), p(7))
  end function ref_vy95

  real(c_double) function ref_ci_dere_ratio(Z, stage) bind(C, name='ref_ci_dere_ratio')
    integer(c_int), value :: Z, stage
    par%ci_model = 'dere_hybrid'
    ref_ci_dere_ratio = ci_dere_ratio(Z, stage)
    par%ci_model = 'voronov'
  end function ref_ci_dere_ratio

  subroutine ref_gaunt_ff(z, log10te, xlf, g, n) bind(C, name='ref_gaunt_ff')
    real(c_float), value :: z, log10te
    integer(c_int), value :: n
    real(c_float), intent(in) :: xlf(n)
    real(c_float), intent(out) :: g(n)
    integer :: iflag
    call getGauntFF(z, log10te, xlf, g, iflag)
  end subroutine ref_gaunt_ff

  subroutine ref_milne_setup(eion_max, te_min, te_max) bind(C, name='ref_milne_setup')
    real(c_double), value :: eion_max, te_min, te_max
    par%eion_max = eion_max
    par%te_min = te_min
    par%te_max = te_max
    par%diffuse_energy_model = 'milne'
    call milne_setup()
  end subroutine ref_milne_setup

  real(c_double) function ref_milne_sample(ich, T, u) bind(C, name='ref_milne_sample')
    integer(c_int), value :: ich
    real(c_double), value :: T, u
    ref_milne_sample = milne_sample_energy(int(ich), T, u)
  end function ref_milne_sample

  real(c_double) function ref_milne_eper(ich, T) bind(C, name='ref_milne_eper')
    integer(c_int), value :: ich
    real(c_double), value :: T
    ref_milne_eper = milne_energy_per_recomb(int(ich), T)
  end function ref_milne_eper

  ! ---- n-level atoms and the local-n_e cooling tables ----------------------
  function c_to_f(s, n) result(f)
    character(kind=c_char), intent(in) :: s(*)
    integer, intent(in) :: n
    character(len=n) :: f
    integer :: i
    do i = 1, n
       f(i:i) = s(i)
    end do
  end function c_to_f

  integer(c_int) function ref_nlevel_load(path, n) bind(C, name='ref_nlevel_load')
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: n
    logical :: ok
    call nlevel_load(c_to_f(path, int(n)), nla, ok)
    ref_nlevel_load = merge(nlevel_nlines(nla), -1, ok)
  end function ref_nlevel_load

  real(c_double) function ref_nlevel_cooling_at(T, ne) bind(C, name='ref_nlevel_cooling_at')
    real(c_double), value :: T, ne
    ref_nlevel_cooling_at = nlevel_cooling_at(nla, T, ne)
  end function ref_nlevel_cooling_at

  subroutine ref_nlevel_emissivities(T, ne, emis) bind(C, name='ref_nlevel_emissivities')
    real(c_double), value :: T, ne
    real(c_double), intent(out) :: emis(600)
    call nlevel_emissivities(nla, T, ne, emis)
  end subroutine ref_nlevel_emissivities

  real(c_double) function ref_nlevel_line_ident(k) bind(C, name='ref_nlevel_line_ident')
    integer(c_int), value :: k
    ref_nlevel_line_ident = nlevel_line_ident(nla, int(k))
  end function ref_nlevel_line_ident

  ! nlevel_cooling_add(name, stage) reading par%atomic_dir
  integer(c_int) function ref_nlevel_cooling_add(dir, nd, name, nn, stage) &
       bind(C, name='ref_nlevel_cooling_add')
    character(kind=c_char), intent(in) :: dir(*), name(*)
    integer(c_int), value :: nd, nn, stage
    par%atomic_dir = c_to_f(dir, int(nd))
    ref_nlevel_cooling_add = nlevel_cooling_add(c_to_f(name, int(nn)), int(stage))
  end function ref_nlevel_cooling_add

  real(c_double) function ref_nlevel_cooling_apply(idx, lam_fit, T, ne) &
       bind(C, name='ref_nlevel_cooling_apply')
    integer(c_int), value :: idx
    real(c_double), value :: lam_fit, T, ne
    ref_nlevel_cooling_apply = nlevel_cooling_apply(int(idx), lam_fit, T, ne)
  end function ref_nlevel_cooling_apply

  real(c_double) function ref_nlevel_cooling_lam_se(idx, T, ne) &
       bind(C, name='ref_nlevel_cooling_lam_se')
    integer(c_int), value :: idx
    real(c_double), value :: T, ne
    ref_nlevel_cooling_lam_se = nlevel_cooling_lam_se(int(idx), T, ne)
  end function ref_nlevel_cooling_lam_se

  subroutine ref_hei_branch_fractions(T, f3s, f1s, f1p) bind(C, name='ref_hei_branch_fractions')
    real(c_double), value :: T
    real(c_double), intent(out) :: f3s, f1s, f1p
    call hei_branch_fractions(T, f3s, f1s, f1p)
  end subroutine ref_hei_branch_fractions

  real(c_double) function ref_hei_584_ionizes_H(T, xHI, xHeI) bind(C, name='ref_hei_584_ionizes_H')
    real(c_double), value :: T, xHI, xHeI
    ref_hei_584_ionizes_H = hei_584_ionizes_H(T, xHI, xHeI)
  end function ref_hei_584_ionizes_H

end module ref_bind
