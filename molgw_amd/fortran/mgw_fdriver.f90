!=========================================================================
! mgw_fdriver: runs the one-shot G0W0 flow of MOLGW's selfenergy_evaluation
! (src/m_selfenergy_evaluation.f90:135-261, 589-590) through the Fortran
! interface module m_mgw, on inputs read from a binary stream file, and
! writes the results to a second file.
!
! usage: mgw_fdriver input.bin output.bin
! input  (stream): int32 nbf, nauxil, nstate; int64 npair;
!                  real64 eri_3center(npair, nauxil); int32 index_basis(2, npair);
!                  real64 c_matrix(nbf, nstate), energy(nstate), occupation(nstate)
! output (stream): real64 energy_qp(nstate), zz(nstate), energy_qp_lin(nstate),
!                  en_rpa, en_gw, energy_qp_one_shot(nstate),
!                  energy_qp_contour(nstate) (gw_selfenergy_contour, 12 imaginary
!                  and 6 real W frequencies, then the graphical QP solve),
!                  selfenergy_mo(nstate, nstate), selfenergy_ao(nbf, nbf) (the
!                  QSGW matrix on the Casida W, S = identity),
!                  w(npole, nstate, nstate), v(nauxil, nstate, nstate)
!                  (dump_gw_ingredients arrays, all states), w_s(nauxil, npole),
!                  exchange_m_vxc(nstate, nstate) of the Fock matrix
!                  C diag(energy(p) + 0.001 p) C^T
!=========================================================================
program mgw_fdriver
  use m_mgw
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  integer :: nbf, nauxil, nstate, iu, istate, nhomo, k
  integer(8) :: npair
  real(dp), allocatable :: eri(:, :), c_matrix(:, :), energy(:), occupation(:)
  integer, allocatable :: index_basis(:, :)
  real(dp), allocatable :: pole(:), omega(:), se_energy0(:), exc(:)
  real(dp), allocatable :: energy_qp(:), zz(:), energy_qp_lin(:), eqp2(:), zz2(:), elin2(:)
  real(dp), allocatable :: eqp3(:), zz3(:), elin3(:), se_energy0_spin(:, :)
  complex(dp), allocatable :: sigma(:, :), sigma2(:, :), sigma3(:, :, :)
  real(dp) :: en_rpa, en_gw, energies(2)
  real(dp), allocatable :: s_matrix(:, :), sigma_qs_mo(:, :), sigma_qs_ao(:, :), w_ing(:, :, :), v_ing(:, :, :)
  real(dp), allocatable :: w_s(:, :), fock(:, :, :), x_m_vxc(:, :, :)
  type(mgw_ranges) :: ranges
  type(mgw_g0w0_params) :: params
  character(len=512) :: fin, fout
  integer, parameter :: nomega_sigma = 51
  real(dp), parameter :: step_sigma = 0.01_dp, eta = 0.001_dp

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=iu, file=TRIM(fin), access='stream', form='unformatted', status='old', action='read')
  read(iu) nbf, nauxil, nstate
  read(iu) npair
  allocate(eri(npair, nauxil), index_basis(2, npair), c_matrix(nbf, nstate), energy(nstate), occupation(nstate))
  read(iu) eri
  read(iu) index_basis
  read(iu) c_matrix
  read(iu) energy
  read(iu) occupation
  close(iu)

  ! selfenergy_set_state_range defaults: all states, no frozen core
  nhomo = 1
  do istate = 1, nstate
    if( occupation(istate) > 1.0e-5_dp ) nhomo = istate
  end do
  ranges%ncore_G = 0
  ranges%nvirtual_G = nstate + 1
  ranges%nsemin = 1
  ranges%nsemax = nstate
  ranges%nhomo_G = nhomo
  ranges%ncore_W = 0
  ranges%nvirtual_W = nstate + 1

  call mgw_init(0)
  call mgw_upload_eri_3center(eri, index_basis, nbf)
  call mgw_upload_mo(c_matrix, energy, occupation)

  ! W by Casida, then Sigma_c on se%omega = step_sigma * k centred on energy0
  call mgw_polarizability_rpa(ranges, en_rpa, en_gw, pole, nauxil, w_s)
  allocate(omega(nomega_sigma), se_energy0(nstate), exc(nstate), sigma(nomega_sigma, nstate))
  do k = 1, nomega_sigma
    omega(k) = step_sigma * (k - 1 - nomega_sigma / 2)
  end do
  se_energy0(:) = energy(:)
  exc(:) = 0.0_dp
  call mgw_selfenergy_poles(ranges, eta, omega, se_energy0, sigma, en_gw)
  allocate(energy_qp(nstate), zz(nstate), energy_qp_lin(nstate))
  call mgw_qp_energy(ranges, omega, sigma, se_energy0, exc, energy, energy_qp, zz, energy_qp_lin)

  ! the QSGW self-energy matrix and the GW ingredients on the same resident W
  allocate(s_matrix(nbf, nbf), sigma_qs_mo(nstate, nstate), sigma_qs_ao(nbf, nbf))
  s_matrix(:, :) = 0.0_dp
  do k = 1, nbf
    s_matrix(k, k) = 1.0_dp
  end do
  call mgw_selfenergy_qs(ranges, 2.0_dp, eta, s_matrix, sigma_qs_mo, sigma_qs_ao)
  allocate(w_ing(SIZE(pole), nstate, nstate), v_ing(nauxil, nstate, nstate))
  call mgw_gw_ingredients(ranges, w_ing, v_ing)

  ! Sigma_x - Vxc of a Fock matrix C diag(energy + 0.001 p) C^T (orthonormal C, S = 1)
  allocate(fock(nbf, nbf, 1), x_m_vxc(nstate, nstate, 1))
  fock(:, :, 1) = 0.0_dp
  do istate = 1, nstate
    do k = 1, nbf
      fock(:, k, 1) = fock(:, k, 1) + c_matrix(:, istate) * (energy(istate) + 0.001_dp * istate) * c_matrix(k, istate)
    end do
  end do
  call mgw_exchange_m_vxc(fock, x_m_vxc)

  ! the same solve in one call
  allocate(eqp2(nstate), zz2(nstate), elin2(nstate), sigma2(nomega_sigma, nstate))
  call mgw_one_shot(params, ranges, exc, eqp2, zz2, elin2, sigma2, energies)

  ! contour deformation on the same grid, then the QP equation
  allocate(eqp3(nstate), zz3(nstate), elin3(nstate), sigma3(nomega_sigma, nstate, 1), se_energy0_spin(nstate, 1))
  se_energy0_spin(:, 1) = energy(:)
  call mgw_selfenergy_contour(ranges, 2.0_dp, 12, 6, eta, omega, se_energy0_spin, sigma3)
  call mgw_qp_energy(ranges, omega, sigma3(:, :, 1), se_energy0, exc, energy, eqp3, zz3, elin3)
  call mgw_finalize()

  open(newunit=iu, file=TRIM(fout), access='stream', form='unformatted', status='replace', action='write')
  write(iu) energy_qp, zz, energy_qp_lin, en_rpa, en_gw, eqp2, eqp3
  write(iu) sigma_qs_mo, sigma_qs_ao, w_ing, v_ing, w_s, x_m_vxc
  close(iu)
  write(*, '(a,i0,a,f16.10)') ' mgw_fdriver: nstate=', nstate, ' E_RPA=', en_rpa
end program mgw_fdriver
