! ref_dms.f90 -- test harness around the reference's own DMS charge routine.
!
! Test infrastructure only (see oracle/Makefile, target ref): compiled with amdflang
! together with the reference's module.f90, multibead.f90, error.f, invert.f, images.f
! and qdetermine.f90, read in place from the reference tree, into oracle/_ref/.  It lets
! tests/golden/make_dms_golden.py record what qdetermine (qdetermine.f90:15-321) returns
! for given water geometries, so that the oracle's restatement (elec_oracle.c dms_water)
! is pinned by the reference's arithmetic rather than by invariants alone.
!
! ref_dms(nw, imcon, cell, gamma, xyz, chge, dp, moldip, chkdip):
!   xyz(3, 4*nw): O, H1, H2 of water i at columns 3i-2..3i (1-based), its M site at 3nw+i;
!   chge(4*nw): the charges qdetermine assigns (new assignment, MB-pol's lttmfold = F);
!   dp(6, nw): dp1/dr1, dp1/dr2, dp1/dcos, dp2/dr1, dp2/dr2, dp2/dcos per water;
!   moldip(3, nw), chkdip(3, nw): the reference's molecular and PS check dipoles.
subroutine ref_dms(nw, imcon, cell, gamma, xyz, chge, dp, moldip, chkdip) bind(c, name="ref_dms")
  use iso_c_binding, only: c_int, c_double
  use global_variables, only: mxatms
  use ps_type_pes, only: init_ps_type_pes
  use ps_type_dms, only: ldms, lttmfold, init_ps_type_dms, init_intra_grad_dms, end_intra_grad_dms, &
                         dp1dr1, dp1dr2, dp1dcos, dp2dr1, dp2dr2, dp2dcos, vesp
  use dipole_moments, only: init_dipole_moments, end_dipole_moments, moldipx, moldipy, moldipz, &
                            chkdipx, chkdipy, chkdipz
  implicit none
  integer(c_int), value :: nw, imcon
  real(c_double), intent(in) :: cell(9), xyz(3, 4 * nw)
  real(c_double), value :: gamma
  real(c_double), intent(out) :: chge(4 * nw), dp(6, nw), moldip(3, nw), chkdip(3, nw)

  logical, save :: tables_ready = .false.
  integer, allocatable :: lst(:)
  real(8), allocatable :: x(:), y(:), z(:), q(:)
  integer :: i, n

  n = 4 * nw
  mxatms = n
  if (.not. tables_ready) then
    call init_ps_type_pes
    call init_ps_type_dms
    tables_ready = .true.
  end if
  ldms = .true.
  lttmfold = .false.
  call init_intra_grad_dms(nw)
  call init_dipole_moments(nw)
  allocate (lst(n), x(n), y(n), z(n), q(n))
  do i = 1, n
    lst(i) = i
    x(i) = xyz(1, i)
    y(i) = xyz(2, i)
    z(i) = xyz(3, i)
  end do
  q(:) = 0.0d0
  call qdetermine(lst, nw, 3 * nw, .false., imcon, cell, q, gamma, x, y, z)
  chge(:) = q(:)
  do i = 1, nw
    dp(:, i) = (/ dp1dr1(i), dp1dr2(i), dp1dcos(i), dp2dr1(i), dp2dr2(i), dp2dcos(i) /)
    moldip(:, i) = (/ moldipx(i), moldipy(i), moldipz(i) /)
    chkdip(:, i) = (/ chkdipx(i), chkdipy(i), chkdipz(i) /)
  end do
  deallocate (lst, x, y, z, q)
  call end_dipole_moments
  call end_intra_grad_dms
  deallocate (vesp)     ! the one array end_intra_grad_dms leaves allocated
end subroutine ref_dms

! ref_qd(nw, imcon, cell, gamma, xyz, vesp_in, f, stress, virqdf, stress_qd_atoms, fm):
! qdetermine then qdforce (qdforce.f90:13-569, built with HEAT_CURRENT and without
! HEAT_STRESS as the reference's Makefile sets it) for the site potentials vesp_in(4*nw):
!   f(3, 4*nw):       the charge-derivative forces at the sites (M sites not redistributed);
!   stress(9), virqdf: its stress and virial;
!   stress_qd_atoms(9): the sum over atoms of atomic_stress_qd (heatcurrent.f90);
!   fm(4*nw, 4*nw, 3): the force-matrix entries force_matrix(i, j, :) qdforce records.
subroutine ref_qd(nw, imcon, cell, gamma, xyz, vesp_in, f, stress, virqdf, stress_qd_atoms, fm) &
    bind(c, name="ref_qd")
  use iso_c_binding, only: c_int, c_double
  use global_variables, only: mxatms
  use multibead, only: bead_rank, bead_size
  use ps_type_pes, only: init_ps_type_pes
  use ps_type_dms, only: ldms, lttmfold, init_ps_type_dms, init_intra_grad_dms, end_intra_grad_dms, vesp
  use dipole_moments, only: init_dipole_moments, end_dipole_moments
  use heatcurrent, only: init_heat, zero_heat, deallocate_heat, atomic_stress_qd, force_matrix
  implicit none
  integer(c_int), value :: nw, imcon
  real(c_double), intent(in) :: cell(9), xyz(3, 4 * nw), vesp_in(4 * nw)
  real(c_double), value :: gamma
  real(c_double), intent(out) :: f(3, 4 * nw), stress(9), virqdf, stress_qd_atoms(9), fm(4 * nw, 4 * nw, 3)

  logical, save :: tables_ready = .false.
  integer, allocatable :: lst(:)
  real(8), allocatable :: x(:), y(:), z(:), q(:), fx(:), fy(:), fz(:)
  integer :: i, a, b, n

  n = 4 * nw
  mxatms = n
  bead_rank = 0
  bead_size = 1
  if (.not. tables_ready) then
    call init_ps_type_pes
    call init_ps_type_dms
    tables_ready = .true.
  end if
  ldms = .true.
  lttmfold = .false.
  call init_intra_grad_dms(nw)
  call init_dipole_moments(nw)
  allocate (lst(n), x(n), y(n), z(n), q(n), fx(n), fy(n), fz(n))
  do i = 1, n
    lst(i) = i
    x(i) = xyz(1, i)
    y(i) = xyz(2, i)
    z(i) = xyz(3, i)
  end do
  q(:) = 0.0d0
  call qdetermine(lst, nw, 3 * nw, .false., imcon, cell, q, gamma, x, y, z)
  vesp(1:n) = vesp_in(1:n)
  call init_heat()
  call zero_heat()
  fx(:) = 0.0d0
  fy(:) = 0.0d0
  fz(:) = 0.0d0
  stress(:) = 0.0d0
  call qdforce(lst, nw, 3 * nw, n, imcon, cell, gamma, x, y, z, fx, fy, fz, stress, virqdf)
  do i = 1, n
    f(:, i) = (/ fx(i), fy(i), fz(i) /)
  end do
  do a = 1, 3
    do b = 1, 3
      stress_qd_atoms(3 * (a - 1) + b) = sum(atomic_stress_qd(:, a, b))
    end do
  end do
  fm(:, :, :) = force_matrix(1:n, 1:n, 1:3)
  call deallocate_heat()
  deallocate (lst, x, y, z, q, fx, fy, fz)
  call end_dipole_moments
  call end_intra_grad_dms
  deallocate (vesp)
end subroutine ref_qd
