!=========================================================================
! m_mgw: Fortran 2008 iso_c_binding interface of libmgw (include/mgw.h).
!
! MOLGW keeps its own subroutines (calculate_eri_3center_mo, polarizability,
! polarizability_grid_scalapack, gw_selfenergy, gw_selfenergy_imag_scalapack,
! se%pade_fit, find_qp_energy_graphical/linearization); their bodies call
! the wrappers of this module, which pass the module state those routines
! read implicitly (eri_3center, index_basis, nsemin, ncore_G, ...) to the
! C-ABI explicitly. A non-zero return code ends the run with the library's
! message, as MOLGW's die() does.
!
! Arrays are passed as the caller holds them (column-major, real(dp) and
! complex(dp)); state ranges are 1-based.
!=========================================================================
module m_mgw
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter :: dp = kind(1.0d0)

  integer(c_int), parameter, public :: MGW_MEM_HOST = 0, MGW_MEM_DEVICE = 1
  integer(c_int), parameter, public :: MGW_FP64 = 0, MGW_FP32 = 1, MGW_BF16X3 = 2
  integer(c_int), parameter, public :: MGW_TECH_ONE_SHOT = 0, MGW_TECH_IMAGINARY_PADE = 1, MGW_TECH_CONTOUR = 2

  type, bind(C), public :: mgw_ranges
    integer(c_int32_t) :: ncore_G = 0
    integer(c_int32_t) :: nvirtual_G = 0
    integer(c_int32_t) :: nsemin = 0
    integer(c_int32_t) :: nsemax = 0
    integer(c_int32_t) :: nhomo_G = 0
    integer(c_int32_t) :: ncore_W = 0
    integer(c_int32_t) :: nvirtual_W = 0
    integer(c_int32_t) :: nlumo_G = 0        ! 0: nhomo_G + 1
  end type mgw_ranges

  type, bind(C), public :: mgw_g0w0_params
    integer(c_int32_t) :: technique = 0
    integer(c_int32_t) :: nomega_sigma = 51
    real(c_double)     :: step_sigma = 0.01_dp
    real(c_double)     :: eta = 0.001_dp
    integer(c_int32_t) :: nomega_chi_imag = 0
    integer(c_int32_t) :: nomega_sigma_calc = 1
    real(c_double)     :: step_sigma_calc = 0.03_dp
    real(c_double)     :: spin_fact = 2.0_dp
    integer(c_int32_t) :: precision = 0
    integer(c_int32_t) :: nomega_chi_real = 2
  end type mgw_g0w0_params

  public :: mgw_init, mgw_finalize, mgw_check
  public :: mgw_upload_eri_3center, mgw_upload_mo, mgw_upload_mo_spin, mgw_select_spin, mgw_set_w
  public :: mgw_inverse_sqrt_eri_2center, mgw_build_eri_3center
  public :: mgw_eri_3center_mo, mgw_destroy_mo_integrals
  public :: mgw_polarizability_rpa, mgw_polarizability_imag
  public :: mgw_selfenergy_poles, mgw_selfenergy_imag, mgw_selfenergy_contour, mgw_pade
  public :: mgw_qp_energy, mgw_one_shot
  public :: mgw_selfenergy_qs, mgw_gw_ingredients, mgw_exchange_m_vxc

  type(c_ptr), save :: ctx = c_null_ptr

  interface
    function c_last_error() bind(C, name='mgw_last_error') result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function c_last_error

    function c_ctx_create(device, handle) bind(C, name='mgw_ctx_create') result(ierr)
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: handle
      integer(c_int) :: ierr
    end function c_ctx_create

    function c_ctx_destroy(handle) bind(C, name='mgw_ctx_destroy') result(ierr)
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int) :: ierr
    end function c_ctx_destroy

    function c_set_eri_3center(handle, eri, npair, nauxil, index_basis, nbf, where) &
        bind(C, name='mgw_set_eri_3center') result(ierr)
      import :: c_int, c_int64_t, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: handle
      real(c_double), intent(in) :: eri(*)
      integer(c_int64_t), value :: npair
      integer(c_int), value :: nauxil
      integer(c_int32_t), intent(in) :: index_basis(*)
      integer(c_int), value :: nbf, where
      integer(c_int) :: ierr
    end function c_set_eri_3center

    function c_inverse_sqrt_eri_2center(handle, eri_2center, nauxil_bas, where, nauxil_kept, lowest, sqrtinv) &
        bind(C, name='mgw_inverse_sqrt_eri_2center') result(ierr)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(in) :: eri_2center(*)
      integer(c_int), value :: nauxil_bas, where
      integer(c_int) :: nauxil_kept
      real(c_double) :: lowest
      real(c_double) :: sqrtinv(*)
      integer(c_int) :: ierr
    end function c_inverse_sqrt_eri_2center

    function c_build_eri_3center(handle, tmp, ipair_first, mpair, npair, index_basis, nbf, where, eri) &
        bind(C, name='mgw_build_eri_3center') result(ierr)
      import :: c_int, c_int64_t, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: handle
      real(c_double), intent(in) :: tmp(*)
      integer(c_int64_t), value :: ipair_first, mpair, npair
      integer(c_int32_t), intent(in) :: index_basis(*)
      integer(c_int), value :: nbf, where
      real(c_double) :: eri(*)
      integer(c_int) :: ierr
    end function c_build_eri_3center

    function c_set_mo_spin(handle, c_matrix, nbf, nstate, nspin, energy, occupation, where) &
        bind(C, name='mgw_set_mo_spin') result(ierr)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(in) :: c_matrix(*), energy(*), occupation(*)
      integer(c_int), value :: nbf, nstate, nspin, where
      integer(c_int) :: ierr
    end function c_set_mo_spin

    function c_set_spin(handle, ispin) bind(C, name='mgw_set_spin') result(ierr)
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: ispin
      integer(c_int) :: ierr
    end function c_set_spin

    function c_set_spectral_function(handle, npole, nauxil, pole, w_s, where) &
        bind(C, name='mgw_set_spectral_function') result(ierr)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: npole, nauxil, where
      real(c_double), intent(in) :: pole(*), w_s(*)
      integer(c_int) :: ierr
    end function c_set_spectral_function

    function c_set_mo(handle, c_matrix, nbf, nstate, energy, occupation, where) &
        bind(C, name='mgw_set_mo') result(ierr)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(in) :: c_matrix(*), energy(*), occupation(*)
      integer(c_int), value :: nbf, nstate, where
      integer(c_int) :: ierr
    end function c_set_mo

    function c_eri_3center_mo(handle, mmin, mmax, nmin, nmax, out) &
        bind(C, name='mgw_calculate_eri_3center_mo') result(ierr)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: mmin, mmax, nmin, nmax
      real(c_double) :: out(*)
      integer(c_int) :: ierr
    end function c_eri_3center_mo

    function c_destroy_eri_3center_mo(handle) bind(C, name='mgw_destroy_eri_3center_mo') result(ierr)
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int) :: ierr
    end function c_destroy_eri_3center_mo

    function c_npole(handle, ranges, npole) bind(C, name='mgw_npole') result(ierr)
      import :: c_int, c_ptr, mgw_ranges
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      integer(c_int) :: npole
      integer(c_int) :: ierr
    end function c_npole

    function c_polarizability(handle, ranges, spin_fact, prec, en_rpa, en_gw, pole, w_s) &
        bind(C, name='mgw_polarizability') result(ierr)
      import :: c_int, c_ptr, c_double, mgw_ranges
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      real(c_double), value :: spin_fact
      integer(c_int), value :: prec
      real(c_double) :: en_rpa, en_gw
      type(c_ptr), value :: pole, w_s
      integer(c_int) :: ierr
    end function c_polarizability

    function c_polarizability_grid(handle, ranges, spin_fact, prec, nomega, omega_im, weight, &
                                   erpa, egw, chi) bind(C, name='mgw_polarizability_grid') result(ierr)
      import :: c_int, c_ptr, c_double, mgw_ranges
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      real(c_double), value :: spin_fact
      integer(c_int), value :: prec, nomega
      real(c_double), intent(in) :: omega_im(*), weight(*)
      real(c_double) :: erpa, egw
      type(c_ptr), value :: chi
      integer(c_int) :: ierr
    end function c_polarizability_grid

    function c_gw_selfenergy(handle, ranges, spin_fact, eta, nomega, omega, se_energy0, sigma, egw) &
        bind(C, name='mgw_gw_selfenergy') result(ierr)
      import :: c_int, c_ptr, c_double, mgw_ranges, c_double_complex
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      real(c_double), value :: spin_fact, eta
      integer(c_int), value :: nomega
      real(c_double), intent(in) :: omega(*), se_energy0(*)
      complex(c_double_complex) :: sigma(*)
      real(c_double) :: egw
      integer(c_int) :: ierr
    end function c_gw_selfenergy

    function c_gw_selfenergy_imag(handle, ranges, ncalc, omega_calc, sigma_calc) &
        bind(C, name='mgw_gw_selfenergy_imag') result(ierr)
      import :: c_int, c_ptr, mgw_ranges, c_double_complex
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      integer(c_int), value :: ncalc
      complex(c_double_complex), intent(in) :: omega_calc(*)
      complex(c_double_complex) :: sigma_calc(*)
      integer(c_int) :: ierr
    end function c_gw_selfenergy_imag

    function c_gw_selfenergy_contour(handle, ranges, spin_fact, nimag, nreal, eta, nomega, omega, se_energy0, &
                                     sigma, de_max) bind(C, name='mgw_gw_selfenergy_contour') result(ierr)
      import :: c_int, c_ptr, c_double, mgw_ranges, c_double_complex
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      real(c_double), value :: spin_fact, eta
      integer(c_int), value :: nimag, nreal, nomega
      real(c_double), intent(in) :: omega(*), se_energy0(*)
      complex(c_double_complex) :: sigma(*)
      real(c_double) :: de_max
      integer(c_int) :: ierr
    end function c_gw_selfenergy_contour

    function c_pade_fit(handle, nse, ncalc, omega_calc, sigma_calc, nomega, omega, eta, sigma) &
        bind(C, name='mgw_pade_fit') result(ierr)
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: handle
      integer(c_int), value :: nse, ncalc, nomega
      complex(c_double_complex), intent(in) :: omega_calc(*), sigma_calc(*)
      real(c_double), intent(in) :: omega(*)
      real(c_double), value :: eta
      complex(c_double_complex) :: sigma(*)
      integer(c_int) :: ierr
    end function c_pade_fit

    function c_find_qp_energy(handle, ranges, nomega, omega, sigma, se_energy0, exchange_m_vxc, &
                              energy0, nstate, energy_qp, zz, energy_qp_lin) &
        bind(C, name='mgw_find_qp_energy') result(ierr)
      import :: c_int, c_ptr, c_double, mgw_ranges, c_double_complex
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      integer(c_int), value :: nomega, nstate
      real(c_double), intent(in) :: omega(*), se_energy0(*), exchange_m_vxc(*), energy0(*)
      complex(c_double_complex), intent(in) :: sigma(*)
      real(c_double) :: energy_qp(*), zz(*), energy_qp_lin(*)
      integer(c_int) :: ierr
    end function c_find_qp_energy

    function c_g0w0(handle, params, ranges, exchange_m_vxc, energy_qp, zz, energy_qp_lin, sigma, energies) &
        bind(C, name='mgw_g0w0') result(ierr)
      import :: c_int, c_ptr, c_double, mgw_ranges, mgw_g0w0_params, c_double_complex
      type(c_ptr), value :: handle
      type(mgw_g0w0_params), intent(in) :: params
      type(mgw_ranges), intent(in) :: ranges
      real(c_double), intent(in) :: exchange_m_vxc(*)
      real(c_double) :: energy_qp(*), zz(*), energy_qp_lin(*)
      complex(c_double_complex) :: sigma(*)
      real(c_double) :: energies(2)
      integer(c_int) :: ierr
    end function c_g0w0

    function c_gw_selfenergy_qs(handle, ranges, spin_fact, eta, s_matrix, selfenergy_mo, selfenergy_ao) &
        bind(C, name='mgw_gw_selfenergy_qs') result(ierr)
      import :: c_int, c_ptr, c_double, mgw_ranges
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      real(c_double), value :: spin_fact, eta
      real(c_double), intent(in) :: s_matrix(*)
      real(c_double) :: selfenergy_mo(*), selfenergy_ao(*)
      integer(c_int) :: ierr
    end function c_gw_selfenergy_qs

    function c_gw_ingredients(handle, ranges, w, v) bind(C, name='mgw_gw_ingredients') result(ierr)
      import :: c_int, c_ptr, c_double, mgw_ranges
      type(c_ptr), value :: handle
      type(mgw_ranges), intent(in) :: ranges
      real(c_double) :: w(*), v(*)
      integer(c_int) :: ierr
    end function c_gw_ingredients

    function c_exchange_m_vxc(handle, hamiltonian_fock, exchange_m_vxc, diag) &
        bind(C, name='mgw_exchange_m_vxc') result(ierr)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), intent(in) :: hamiltonian_fock(*)
      real(c_double) :: exchange_m_vxc(*), diag(*)
      integer(c_int) :: ierr
    end function c_exchange_m_vxc

    function c_strlen(s) bind(C, name='strlen') result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function c_strlen
  end interface

contains

!=========================================================================
! Error handling: the reference ends with die() (src/m_warning.f90:78-89).
subroutine mgw_check(ierr, where)
  integer(c_int), intent(in)   :: ierr
  character(len=*), intent(in) :: where
  !=====
  type(c_ptr) :: p
  character(kind=c_char), pointer :: msg(:)
  character(len=:), allocatable :: text
  integer :: i, n
  !=====
  if( ierr == 0 ) return
  p = c_last_error()
  n = int(c_strlen(p))
  call c_f_pointer(p, msg, [n])
  allocate(character(len=n) :: text)
  do i = 1, n
    text(i:i) = msg(i)
  end do
  write(*, '(a,a,a,i0,a,a)') ' ', where, ': libmgw error ', ierr, ': ', text
  error stop 'libmgw'
end subroutine mgw_check

subroutine mgw_init(device)
  integer, intent(in) :: device
  if( c_associated(ctx) ) return
  call mgw_check(c_ctx_create(int(device, c_int), ctx), 'mgw_init')
end subroutine mgw_init

subroutine mgw_finalize()
  if( .NOT. c_associated(ctx) ) return
  call mgw_check(c_ctx_destroy(ctx), 'mgw_finalize')
  ctx = c_null_ptr
end subroutine mgw_finalize

!=========================================================================
! Module state of m_eri (eri_3center(npair, nauxil), index_basis(2, npair))
subroutine mgw_upload_eri_3center(eri_3center, index_basis, nbf)
  real(dp), intent(in) :: eri_3center(:, :)
  integer, intent(in)  :: index_basis(:, :)
  integer, intent(in)  :: nbf
  !=====
  integer(c_int32_t), allocatable :: ib(:, :)
  !=====
  ib = int(index_basis, c_int32_t)
  call mgw_check(c_set_eri_3center(ctx, eri_3center, int(SIZE(eri_3center, DIM=1), c_int64_t), &
                                   int(SIZE(eri_3center, DIM=2), c_int), ib, int(nbf, c_int), MGW_MEM_HOST), &
                 'mgw_upload_eri_3center')
end subroutine mgw_upload_eri_3center

!=========================================================================
! calculate_inverse_sqrt_eri_2center_scalapack (src/m_eri_calculate.f90:1049-1262),
! full-range branch: columns 1:nauxil_kept of eri_2center_sqrtinv are set.
subroutine mgw_inverse_sqrt_eri_2center(eri_2center, nauxil_kept, eri_2center_sqrtinv)
  real(dp), intent(in)    :: eri_2center(:, :)
  integer, intent(out)    :: nauxil_kept
  real(dp), intent(inout) :: eri_2center_sqrtinv(:, :)
  !=====
  integer(c_int) :: kept
  real(c_double) :: lowest
  !=====
  if( SIZE(eri_2center_sqrtinv) < SIZE(eri_2center) ) then
    write(*, '(a)') ' mgw_inverse_sqrt_eri_2center: eri_2center_sqrtinv is smaller than eri_2center'
    error stop 'libmgw'
  end if
  call mgw_check(c_inverse_sqrt_eri_2center(ctx, eri_2center, int(SIZE(eri_2center, DIM=1), c_int), MGW_MEM_HOST, &
                                            kept, lowest, eri_2center_sqrtinv), 'mgw_inverse_sqrt_eri_2center')
  nauxil_kept = kept
end subroutine mgw_inverse_sqrt_eri_2center

!=========================================================================
! Second part of calculate_eri_3center_scalapack (src/m_eri_calculate.f90:1877-1892)
! for the pairs ipair_first:ipair_first+SIZE(eri_3center_tmp,1)-1; the RI
! tensor is resident on the GPU once the last batch is in.
subroutine mgw_build_eri_3center(eri_3center_tmp, ipair_first, index_basis, nbf, eri_3center)
  real(dp), intent(in)    :: eri_3center_tmp(:, :)
  integer, intent(in)     :: ipair_first, nbf
  integer, intent(in)     :: index_basis(:, :)
  real(dp), intent(inout) :: eri_3center(:, :)
  !=====
  integer(c_int32_t), allocatable :: ib(:, :)
  !=====
  ib = int(index_basis, c_int32_t)
  call mgw_check(c_build_eri_3center(ctx, eri_3center_tmp, int(ipair_first, c_int64_t), &
                                     int(SIZE(eri_3center_tmp, DIM=1), c_int64_t), &
                                     int(SIZE(eri_3center, DIM=1), c_int64_t), ib, int(nbf, c_int), &
                                     MGW_MEM_HOST, eri_3center), 'mgw_build_eri_3center')
end subroutine mgw_build_eri_3center

!=========================================================================
! All spin channels at once: c_matrix(:,:,nspin), energy(:,nspin), occupation(:,nspin)
! as held by the reference; W sums the channels, the self-energy calls work
! on the channel chosen by mgw_select_spin.
subroutine mgw_upload_mo_spin(c_matrix, energy, occupation)
  real(dp), intent(in) :: c_matrix(:, :, :), energy(:, :), occupation(:, :)
  call mgw_check(c_set_mo_spin(ctx, c_matrix, int(SIZE(c_matrix, DIM=1), c_int), int(SIZE(c_matrix, DIM=2), c_int), &
                               int(SIZE(c_matrix, DIM=3), c_int), energy, occupation, MGW_MEM_HOST), &
                 'mgw_upload_mo_spin')
end subroutine mgw_upload_mo_spin

subroutine mgw_select_spin(ispin)
  integer, intent(in) :: ispin
  call mgw_check(c_set_spin(ctx, int(ispin, c_int)), 'mgw_select_spin')
end subroutine mgw_select_spin

!=========================================================================
! After read_spectral_function (src/m_spectral_function.f90:508-620): W from
! the poles and residues read from SCREENED_COULOMB.
subroutine mgw_set_w(pole, w_s)
  real(dp), intent(in) :: pole(:), w_s(:, :)
  call mgw_check(c_set_spectral_function(ctx, int(SIZE(pole), c_int), int(SIZE(w_s, DIM=1), c_int), pole, w_s, &
                                         MGW_MEM_HOST), 'mgw_set_w')
end subroutine mgw_set_w

subroutine mgw_upload_mo(c_matrix, energy, occupation)
  real(dp), intent(in) :: c_matrix(:, :), energy(:), occupation(:)
  call mgw_check(c_set_mo(ctx, c_matrix, int(SIZE(c_matrix, DIM=1), c_int), int(SIZE(c_matrix, DIM=2), c_int), &
                          energy, occupation, MGW_MEM_HOST), 'mgw_upload_mo')
end subroutine mgw_upload_mo

!=========================================================================
! calculate_eri_3center_mo (src/m_eri_ao_mo.f90:416-582)
subroutine mgw_eri_3center_mo(mstate_min, mstate_max, nstate_min, nstate_max, eri_3center_mo)
  integer, intent(in)   :: mstate_min, mstate_max, nstate_min, nstate_max
  real(dp), intent(out) :: eri_3center_mo(:, :, :)
  call mgw_check(c_eri_3center_mo(ctx, int(mstate_min, c_int), int(mstate_max, c_int), int(nstate_min, c_int), &
                                  int(nstate_max, c_int), eri_3center_mo), 'mgw_eri_3center_mo')
end subroutine mgw_eri_3center_mo

subroutine mgw_destroy_mo_integrals()
  call mgw_check(c_destroy_eri_3center_mo(ctx), 'mgw_destroy_mo_integrals')
end subroutine mgw_destroy_mo_integrals

!=========================================================================
! polarizability, RPA branch (src/m_linear_response.f90:33-498): fills
! wpol%pole(npole); wpol%w_s(nauxil, npole) stays resident on the GPU for
! the self-energy (copied out only when w_s is present); en_rpa, en_gw
subroutine mgw_polarizability_rpa(ranges, en_rpa, en_gw, pole, nauxil, w_s)
  type(mgw_ranges), intent(in) :: ranges
  real(dp), intent(out) :: en_rpa, en_gw
  real(dp), allocatable, target, intent(out) :: pole(:)
  integer, intent(in), optional :: nauxil
  real(dp), allocatable, target, intent(out), optional :: w_s(:, :)
  !=====
  integer(c_int) :: npole
  type(c_ptr) :: ws_ptr
  !=====
  call mgw_check(c_npole(ctx, ranges, npole), 'mgw_polarizability_rpa')
  allocate(pole(npole))
  ws_ptr = c_null_ptr
  if( PRESENT(w_s) ) then
    if( .NOT. PRESENT(nauxil) ) error stop 'mgw_polarizability_rpa: w_s needs nauxil'
    allocate(w_s(nauxil, npole))
    ws_ptr = c_loc(w_s)
  end if
  call mgw_check(c_polarizability(ctx, ranges, 2.0_dp, MGW_FP64, en_rpa, en_gw, c_loc(pole), ws_ptr), &
                 'mgw_polarizability_rpa')
end subroutine mgw_polarizability_rpa

! polarizability_grid_scalapack (src/m_gw_selfenergy_grid.f90:33-209); chi stays resident
subroutine mgw_polarizability_imag(ranges, omega_im, weight_quad, erpa, egw)
  type(mgw_ranges), intent(in) :: ranges
  real(dp), intent(in)  :: omega_im(:), weight_quad(:)
  real(dp), intent(out) :: erpa, egw
  call mgw_check(c_polarizability_grid(ctx, ranges, 2.0_dp, MGW_FP64, int(SIZE(omega_im), c_int), omega_im, &
                                       weight_quad, erpa, egw, c_null_ptr), 'mgw_polarizability_imag')
end subroutine mgw_polarizability_imag

!=========================================================================
! gw_selfenergy (src/m_gw_selfenergy_analytic.f90:29-196): se%sigma(-n:n, nsemin:nsemax)
subroutine mgw_selfenergy_poles(ranges, eta, omega, se_energy0, sigma, energy_gw)
  type(mgw_ranges), intent(in) :: ranges
  real(dp), intent(in)    :: eta, omega(:), se_energy0(:)
  complex(dp), intent(out) :: sigma(:, :)
  real(dp), intent(out)   :: energy_gw
  call mgw_check(c_gw_selfenergy(ctx, ranges, 2.0_dp, eta, int(SIZE(omega), c_int), omega, se_energy0, sigma, &
                                 energy_gw), 'mgw_selfenergy_poles')
end subroutine mgw_selfenergy_poles

! gw_selfenergy_imag_scalapack (src/m_gw_selfenergy_grid.f90:213-338): se%sigma_calc
subroutine mgw_selfenergy_imag(ranges, omega_calc, sigma_calc)
  type(mgw_ranges), intent(in) :: ranges
  complex(dp), intent(in)  :: omega_calc(:)
  complex(dp), intent(out) :: sigma_calc(:, :)
  call mgw_check(c_gw_selfenergy_imag(ctx, ranges, int(SIZE(omega_calc), c_int), omega_calc, sigma_calc), &
                 'mgw_selfenergy_imag')
end subroutine mgw_selfenergy_imag

! gw_selfenergy_contour (src/m_gw_selfenergy_grid.f90:342-551): se%sigma of
! every spin channel, sigma(-nomega:nomega, nsemin:nsemax, nspin) as in se_init
subroutine mgw_selfenergy_contour(ranges, spin_fact, nomega_chi_imag, nomega_chi_real, eta, omega, se_energy0, &
                                  sigma)
  type(mgw_ranges), intent(in) :: ranges
  real(dp), intent(in)     :: spin_fact, eta, omega(:), se_energy0(:, :)
  integer, intent(in)      :: nomega_chi_imag, nomega_chi_real
  complex(dp), intent(out) :: sigma(:, :, :)
  real(c_double) :: de_max
  call mgw_check(c_gw_selfenergy_contour(ctx, ranges, spin_fact, int(nomega_chi_imag, c_int), &
                                         int(nomega_chi_real, c_int), eta, int(SIZE(omega), c_int), omega, &
                                         se_energy0, sigma, de_max), 'mgw_selfenergy_contour')
end subroutine mgw_selfenergy_contour

! se_pade_fit (src/m_selfenergy_tools.f90:965-999)
subroutine mgw_pade(omega_calc, sigma_calc, omega, eta, sigma)
  complex(dp), intent(in)  :: omega_calc(:), sigma_calc(:, :)
  real(dp), intent(in)     :: omega(:), eta
  complex(dp), intent(out) :: sigma(:, :)
  call mgw_check(c_pade_fit(ctx, int(SIZE(sigma_calc, DIM=2), c_int), int(SIZE(omega_calc), c_int), omega_calc, &
                            sigma_calc, int(SIZE(omega), c_int), omega, eta, sigma), 'mgw_pade')
end subroutine mgw_pade

!=========================================================================
! find_qp_energy_graphical + find_qp_energy_linearization
! (src/m_selfenergy_tools.f90:202-420)
subroutine mgw_qp_energy(ranges, omega, sigma, se_energy0, exchange_m_vxc, energy0, energy_qp_g, zz, energy_qp_z)
  type(mgw_ranges), intent(in) :: ranges
  real(dp), intent(in)    :: omega(:), se_energy0(:), exchange_m_vxc(:), energy0(:)
  complex(dp), intent(in) :: sigma(:, :)
  real(dp), intent(out)   :: energy_qp_g(:), zz(:), energy_qp_z(:)
  call mgw_check(c_find_qp_energy(ctx, ranges, int(SIZE(omega), c_int), omega, sigma, se_energy0, exchange_m_vxc, &
                                  energy0, int(SIZE(energy0), c_int), energy_qp_g, zz, energy_qp_z), 'mgw_qp_energy')
end subroutine mgw_qp_energy

!=========================================================================
! One whole one-shot solve on the resident tensors (selfenergy_evaluation flow)
subroutine mgw_one_shot(params, ranges, exchange_m_vxc, energy_qp, zz, energy_qp_lin, sigma, energies)
  type(mgw_g0w0_params), intent(in) :: params
  type(mgw_ranges), intent(in)      :: ranges
  real(dp), intent(in)     :: exchange_m_vxc(:)
  real(dp), intent(out)    :: energy_qp(:), zz(:), energy_qp_lin(:), energies(2)
  complex(dp), intent(out) :: sigma(:, :)
  call mgw_check(c_g0w0(ctx, params, ranges, exchange_m_vxc, energy_qp, zz, energy_qp_lin, sigma, energies), &
                 'mgw_one_shot')
end subroutine mgw_one_shot

!=========================================================================
! gw_selfenergy_qs, GW branch (src/m_gw_selfenergy_analytic.f90:653-797) with
! the symmetrisation of apply_qs_approximation (src/m_selfenergy_tools.f90:809-828)
! on the resident Casida W, active spin channel: selfenergy_mo(nstate, nstate)
! (zero outside nsemin..nsemax) and selfenergy_ao = S C Sigma C^T S (nbf, nbf)
subroutine mgw_selfenergy_qs(ranges, spin_fact, eta, s_matrix, selfenergy_mo, selfenergy_ao)
  type(mgw_ranges), intent(in) :: ranges
  real(dp), intent(in)  :: spin_fact, eta, s_matrix(:, :)
  real(dp), intent(out) :: selfenergy_mo(:, :), selfenergy_ao(:, :)
  call mgw_check(c_gw_selfenergy_qs(ctx, ranges, spin_fact, eta, s_matrix, selfenergy_mo, selfenergy_ao), &
                 'mgw_selfenergy_qs')
end subroutine mgw_selfenergy_qs

! dump_gw_ingredients (src/m_gw_selfenergy_analytic.f90:801-881): the arrays of
! the w.bin and v.bin records for G states ncore_G+1 .. nvirtual_G-1 (ng):
! w(npole, ng, ng) on the resident W and v(nauxil, ng, ng)
subroutine mgw_gw_ingredients(ranges, w, v)
  type(mgw_ranges), intent(in) :: ranges
  real(dp), intent(out) :: w(:, :, :), v(:, :, :)
  call mgw_check(c_gw_ingredients(ctx, ranges, w, v), 'mgw_gw_ingredients')
end subroutine mgw_gw_ingredients

!=========================================================================
! setup_exchange_m_vxc (src/m_selfenergy_tools.f90:720-802), the branch without
! a DFT core-valence split: exchange_m_vxc(:,:,ispin) = C^T F C - diag(energy)
! for the uploaded MO data (mgw_upload_mo / mgw_upload_mo_spin)
subroutine mgw_exchange_m_vxc(hamiltonian_fock, exchange_m_vxc)
  real(dp), intent(in)  :: hamiltonian_fock(:, :, :)
  real(dp), intent(out) :: exchange_m_vxc(:, :, :)
  real(dp), allocatable :: diag(:, :)
  allocate(diag(SIZE(exchange_m_vxc, DIM=1), SIZE(exchange_m_vxc, DIM=3)))
  call mgw_check(c_exchange_m_vxc(ctx, hamiltonian_fock, exchange_m_vxc, diag), 'mgw_exchange_m_vxc')
  deallocate(diag)
end subroutine mgw_exchange_m_vxc

end module m_mgw
