! mbgpu.f90 -- Fortran 2003 (iso_c_binding) interface module of libmbgpu.so, the gfx950
! MB-pol force / heat-current path declared in mbgpu.h.
!
! The reference's host is Fortran (DL_POLY 2); a maintainer adds this module to its
! sources and replaces three call sites:
!   mbpol_init / mbpol_fini   (mbpol.f90:52-94)      -> mbgpu_create / mbgpu_destroy
!   call mbpol_forces(...)    (dlpoly_cmd.f:2504)    -> mbgpu_mbpol_forces
!   compute_heat_flux2        (heatcurrent.f90:703-728, called at dlpoly_cmd.f:3237, 3456)
!                                                    -> mbgpu_heat_flux
! and, to move the rest of the force evaluation of an MB-pol water box to the GPU, the
! polarizable part of forces (forces.f:268-404, 595, 734, 888, with qdetermine) and the
! water terms (bndfrc, angfrc/ps_type_couple, srfrce2, lrcorrect) -> mbgpu_polar_electrostatics
! and mbgpu_water_terms.  The leaves mbpol_2b_poly_ / mbpol_3b_poly_ / mbpol_*b_cutoff_ keep
! the reference's C signatures (mbpol-2b-poly.cpp:1878-1904, mbpol-3b-poly.cpp:1518-1542),
! so the reference's implicit-interface calls (mbpol.f90:64-65, 517, 894) link to them
! unchanged; explicit interfaces are given here as well.
!
! Every function returns 0 on success; mbgpu_error_text() gives the message of the calling
! thread's last failure.  Arrays are passed by reference (the reference's xxx(mxatms) etc.),
! scalars the C side takes by value carry the VALUE attribute, optional arrays the C side
! accepts as NULL are TYPE(C_PTR) (pass c_null_ptr or c_loc(array)).
module mbgpu
  use iso_c_binding, only: c_int, c_long, c_double, c_char, c_ptr, c_null_ptr, c_null_char, &
                           c_associated, c_f_pointer
  implicit none
  private :: c_int, c_long, c_double, c_char, c_ptr, c_null_ptr, c_null_char, c_associated, c_f_pointer
  private :: c_strlen

  ! mbgpu_elec_params (mbgpu.h): CONTROL, FIELD and vmbpol-preset inputs of the
  ! polarizable Ewald electrostatics
  type, bind(c) :: mbgpu_elec_params
    real(c_double) :: rcut, alpha
    integer(c_int) :: kmax(3)
    real(c_double) :: toler, omega
    integer(c_int) :: max_iter
    real(c_double) :: gamma, q_o
    real(c_double) :: polr(3), polr2(3)
    real(c_double) :: athole, athole12, athole13, ascc, ascd
    integer(c_int) :: compat_qd_column, distribute_msite, solver
  end type mbgpu_elec_params

  ! mbgpu_water_params (mbgpu.h): FIELD inputs of the monomer, dispersion and long-range
  ! terms; bond_prm(q, t) is the C array's bond_prm[t][q]
  type, bind(c) :: mbgpu_water_params
    real(c_double) :: rcut, rvdw
    real(c_double) :: c6(3), a6(3), c8(3), a8(3)
    integer(c_int) :: bond_key(2)
    real(c_double) :: bond_prm(3, 2)
    integer(c_int) :: ps_monomer, dispersion, long_range
  end type mbgpu_water_params

  integer, parameter :: mbgpu_ring_max_beads = 256
  integer, parameter :: mbgpu_spme_max_order = 12, mbgpu_spme_max_grid = 256
  integer, parameter :: mbgpu_noalloc_scratch_doubles = 32768

  interface
    ! ---- leaves (reference C ABI) ----
    subroutine mbpol_2b_poly(w1, w2, e, g1, g2, force_matrix) bind(c, name="mbpol_2b_poly_")
      import :: c_double
      real(c_double), intent(in) :: w1(9), w2(9)
      real(c_double) :: e, g1(9), g2(9), force_matrix(108)
    end subroutine
    subroutine mbpol_3b_poly(w1, w2, w3, e, g1, g2, g3, force_matrix) bind(c, name="mbpol_3b_poly_")
      import :: c_double
      real(c_double), intent(in) :: w1(9), w2(9), w3(9)
      real(c_double) :: e, g1(9), g2(9), g3(9), force_matrix(243)
    end subroutine
    subroutine mbpol_2b_cutoff(r) bind(c, name="mbpol_2b_cutoff_")
      import :: c_double
      real(c_double) :: r
    end subroutine
    subroutine mbpol_3b_cutoff(r) bind(c, name="mbpol_3b_cutoff_")
      import :: c_double
      real(c_double) :: r
    end subroutine

    ! ---- context ----
    function mbgpu_create(nw, natms) bind(c, name="mbgpu_create")
      import :: c_int, c_ptr
      integer(c_int), value :: nw, natms
      type(c_ptr) :: mbgpu_create
    end function
    function mbgpu_create_queues(nw, natms, dedicated) bind(c, name="mbgpu_create_queues")
      import :: c_int, c_ptr
      integer(c_int), value :: nw, natms, dedicated
      type(c_ptr) :: mbgpu_create_queues
    end function
    subroutine mbgpu_destroy(ctx) bind(c, name="mbgpu_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    function mbgpu_last_error() bind(c, name="mbgpu_last_error")
      import :: c_ptr
      type(c_ptr) :: mbgpu_last_error
    end function

    ! ---- MB-pol 2B + 3B (mbpol_forces) ----
    function mbgpu_mbpol_forces(ctx, imcon, cell, nttm2, listttm2, xxx, yyy, zzz, fxx, fyy, fzz, &
                                engsrp, virsrp, stress) bind(c, name="mbgpu_mbpol_forces")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: imcon, nttm2
      real(c_double), intent(in) :: cell(9), xxx(*), yyy(*), zzz(*)
      integer(c_int), intent(in) :: listttm2(*)
      real(c_double) :: fxx(*), fyy(*), fzz(*), engsrp, virsrp, stress(9)
      integer(c_int) :: mbgpu_mbpol_forces
    end function
    function mbgpu_mbpol_forces_dev(ctx, imcon, cell, nttm2, listttm2_dev, x_dev, y_dev, z_dev, &
                                    fx_dev, fy_dev, fz_dev, engsrp, virsrp, stress) &
        bind(c, name="mbgpu_mbpol_forces_dev")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, listttm2_dev, x_dev, y_dev, z_dev, fx_dev, fy_dev, fz_dev
      integer(c_int), value :: imcon, nttm2
      real(c_double), intent(in) :: cell(9)
      real(c_double) :: engsrp, virsrp, stress(9)
      integer(c_int) :: mbgpu_mbpol_forces_dev
    end function
    function mbgpu_check_listttm2(nttm2, listttm2, natms) bind(c, name="mbgpu_check_listttm2")
      import :: c_int
      integer(c_int), value :: nttm2, natms
      integer(c_int), intent(in) :: listttm2(*)
      integer(c_int) :: mbgpu_check_listttm2
    end function
    function mbgpu_get_decomposition(ctx, atomic_energy, heat_w, forces) bind(c, name="mbgpu_get_decomposition")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, atomic_energy, heat_w, forces
      integer(c_int) :: mbgpu_get_decomposition
    end function
    function mbgpu_set_precision(ctx, mode) bind(c, name="mbgpu_set_precision")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      integer(c_int) :: mbgpu_set_precision
    end function
    function mbgpu_set_elec_reciprocal(ctx, mode, order, grid) bind(c, name="mbgpu_set_elec_reciprocal")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode, order
      integer(c_int), intent(in) :: grid(3)
      integer(c_int) :: mbgpu_set_elec_reciprocal
    end function
    function mbgpu_set_graphs(ctx, on) bind(c, name="mbgpu_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
      integer(c_int) :: mbgpu_set_graphs
    end function
    function mbgpu_get_graph_counts(ctx, counts) bind(c, name="mbgpu_get_graph_counts")
      import :: c_int, c_long, c_ptr
      type(c_ptr), value :: ctx
      integer(c_long), intent(out) :: counts(2)
      integer(c_int) :: mbgpu_get_graph_counts
    end function
    function mbgpu_get_lists(ctx, dimers, dcap, nd, trimers, tcap, nt) bind(c, name="mbgpu_get_lists")
      import :: c_int, c_long, c_ptr
      type(c_ptr), value :: ctx, dimers, trimers
      integer(c_long), value :: dcap, tcap
      integer(c_long) :: nd, nt
      integer(c_int) :: mbgpu_get_lists
    end function

    ! ---- heat flux (compute_heat_flux2) ----
    function mbgpu_heat_flux(ctx, vxx, vyy, vzz, e_kin, e_pot_extra, w_extra, stress_ew1p, &
                             jkin, jpot, jtot) bind(c, name="mbgpu_heat_flux")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, e_pot_extra, w_extra, stress_ew1p
      real(c_double), intent(in) :: vxx(*), vyy(*), vzz(*), e_kin(*)
      real(c_double) :: jkin(3), jpot(3), jtot(3)
      integer(c_int) :: mbgpu_heat_flux
    end function
    function mbgpu_heat_flux_dev(ctx, vx_dev, vy_dev, vz_dev, e_kin_dev, e_pot_extra_dev, w_extra_dev, &
                                 stress_ew1p_dev, jkin, jpot, jtot) bind(c, name="mbgpu_heat_flux_dev")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, vx_dev, vy_dev, vz_dev, e_kin_dev, e_pot_extra_dev, w_extra_dev, stress_ew1p_dev
      real(c_double) :: jkin(3), jpot(3), jtot(3)
      integer(c_int) :: mbgpu_heat_flux_dev
    end function

    ! ---- polarizable electrostatics ----
    function mbgpu_polar_electrostatics(ctx, imcon, cell, p, xxx, yyy, zzz, dipx, dipy, dipz, chge, &
                                        fxx, fyy, fzz, engcpe, vircpe, stress, niter) &
        bind(c, name="mbgpu_polar_electrostatics")
      import :: c_int, c_double, c_ptr, mbgpu_elec_params
      type(c_ptr), value :: ctx, chge
      integer(c_int), value :: imcon
      real(c_double), intent(in) :: cell(9)
      type(mbgpu_elec_params), intent(in) :: p
      real(c_double) :: xxx(*), yyy(*), zzz(*), dipx(*), dipy(*), dipz(*), fxx(*), fyy(*), fzz(*)
      real(c_double) :: engcpe, vircpe, stress(9)
      integer(c_int) :: niter, mbgpu_polar_electrostatics
    end function
    function mbgpu_polar_electrostatics_dev(ctx, imcon, cell, p, x_dev, y_dev, z_dev, dip_dev, f_dev, &
                                            engcpe, vircpe, stress, niter) &
        bind(c, name="mbgpu_polar_electrostatics_dev")
      import :: c_int, c_double, c_ptr, mbgpu_elec_params
      type(c_ptr), value :: ctx, x_dev, y_dev, z_dev, dip_dev, f_dev
      integer(c_int), value :: imcon
      real(c_double), intent(in) :: cell(9)
      type(mbgpu_elec_params), intent(in) :: p
      real(c_double) :: engcpe, vircpe, stress(9)
      integer(c_int) :: niter, mbgpu_polar_electrostatics_dev
    end function
    function mbgpu_get_elec_details(ctx, scal, chge, vesp, e_real, w_real, s_real, f_site) &
        bind(c, name="mbgpu_get_elec_details")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, scal, chge, vesp, e_real, w_real, s_real, f_site
      integer(c_int) :: mbgpu_get_elec_details
    end function
    function mbgpu_get_dipoles(ctx, dip) bind(c, name="mbgpu_get_dipoles")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: dip(*)
      integer(c_int) :: mbgpu_get_dipoles
    end function

    ! ---- monomer, dispersion and long-range terms ----
    function mbgpu_water_terms(ctx, imcon, cell, p, xxx, yyy, zzz, fxx, fyy, fzz, eng, vir, stress) &
        bind(c, name="mbgpu_water_terms")
      import :: c_int, c_double, c_ptr, mbgpu_water_params
      type(c_ptr), value :: ctx, fxx, fyy, fzz
      integer(c_int), value :: imcon
      real(c_double), intent(in) :: cell(9), xxx(*), yyy(*), zzz(*)
      type(mbgpu_water_params), intent(in) :: p
      real(c_double) :: eng(4), vir(4), stress(9)
      integer(c_int) :: mbgpu_water_terms
    end function
    function mbgpu_water_terms_dev(ctx, imcon, cell, p, x_dev, y_dev, z_dev, fx_dev, fy_dev, fz_dev, &
                                   eng, vir, stress) bind(c, name="mbgpu_water_terms_dev")
      import :: c_int, c_double, c_ptr, mbgpu_water_params
      type(c_ptr), value :: ctx, x_dev, y_dev, z_dev, fx_dev, fy_dev, fz_dev
      integer(c_int), value :: imcon
      real(c_double), intent(in) :: cell(9)
      type(mbgpu_water_params), intent(in) :: p
      real(c_double) :: eng(4), vir(4), stress(9)
      integer(c_int) :: mbgpu_water_terms_dev
    end function
    function mbgpu_evaluate_dev(ctx, imcon, cell, nttm2, listttm2_dev, ep, wp, x_dev, y_dev, z_dev, dip_dev, &
                                vx_dev, vy_dev, vz_dev, e_kin_dev, out) bind(c, name="mbgpu_evaluate_dev")
      import :: c_int, c_double, c_ptr, mbgpu_elec_params, mbgpu_water_params
      type(c_ptr), value :: ctx, listttm2_dev, x_dev, y_dev, z_dev, dip_dev, vx_dev, vy_dev, vz_dev, e_kin_dev
      integer(c_int), value :: imcon, nttm2
      real(c_double), intent(in) :: cell(9)
      type(mbgpu_elec_params), intent(in) :: ep
      type(mbgpu_water_params), intent(in) :: wp
      real(c_double), intent(out) :: out(64)
      integer(c_int) :: mbgpu_evaluate_dev
    end function
    function mbgpu_evaluate_forces_dev(ctx, imcon, cell, nttm2, listttm2_dev, ep, wp, x_dev, y_dev, z_dev, &
                                       dip_dev, fx_dev, fy_dev, fz_dev, fsite_dev, vx_dev, vy_dev, vz_dev, &
                                       e_kin_dev, out) bind(c, name="mbgpu_evaluate_forces_dev")
      import :: c_int, c_double, c_ptr, mbgpu_elec_params, mbgpu_water_params
      type(c_ptr), value :: ctx, listttm2_dev, x_dev, y_dev, z_dev, dip_dev, fx_dev, fy_dev, fz_dev, fsite_dev
      type(c_ptr), value :: vx_dev, vy_dev, vz_dev, e_kin_dev
      integer(c_int), value :: imcon, nttm2
      real(c_double), intent(in) :: cell(9)
      type(mbgpu_elec_params), intent(in) :: ep
      type(mbgpu_water_params), intent(in) :: wp
      real(c_double), intent(out) :: out(64)
      integer(c_int) :: mbgpu_evaluate_forces_dev
    end function
    function mbgpu_get_forces(ctx, f) bind(c, name="mbgpu_get_forces")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: f(*)
      integer(c_int) :: mbgpu_get_forces
    end function
    function mbgpu_get_evaluate_graph_counts(ctx, counts) bind(c, name="mbgpu_get_evaluate_graph_counts")
      import :: c_int, c_long, c_ptr
      type(c_ptr), value :: ctx
      integer(c_long), intent(out) :: counts(2)
      integer(c_int) :: mbgpu_get_evaluate_graph_counts
    end function
    function mbgpu_get_water_details(ctx, e_real, w_real, f_real) bind(c, name="mbgpu_get_water_details")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, e_real, w_real, f_real
      integer(c_int) :: mbgpu_get_water_details
    end function

    function mbgpu_heat_check(ctx, e_kin_dev, out) bind(c, name="mbgpu_heat_check")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, e_kin_dev
      real(c_double) :: out(12)
      integer(c_int) :: mbgpu_heat_check
    end function

    ! ---- integrator, Green-Kubo, ring polymer, particle-mesh Ewald ----
    function mbgpu_verlet_stage(ctx, stage, dt, mass_dev, x_dev, y_dev, z_dev, vx_dev, vy_dev, vz_dev, &
                                fx_dev, fy_dev, fz_dev, fsite_dev, ekin_dev, ekin_total) &
        bind(c, name="mbgpu_verlet_stage")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, mass_dev, x_dev, y_dev, z_dev, vx_dev, vy_dev, vz_dev
      type(c_ptr), value :: fx_dev, fy_dev, fz_dev, fsite_dev, ekin_dev
      integer(c_int), value :: stage
      real(c_double), value :: dt
      real(c_double) :: ekin_total
      integer(c_int) :: mbgpu_verlet_stage
    end function
    function mbgpu_heat_acf(ctx, j, nt, ntau, acf) bind(c, name="mbgpu_heat_acf")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: j(*)
      integer(c_long), value :: nt
      integer(c_int), value :: ntau
      real(c_double) :: acf(*)
      integer(c_int) :: mbgpu_heat_acf
    end function
    function mbgpu_heat_acf_dev(ctx, j_dev, nt, ntau, acf_dev) bind(c, name="mbgpu_heat_acf_dev")
      import :: c_int, c_long, c_ptr
      type(c_ptr), value :: ctx, j_dev, acf_dev
      integer(c_long), value :: nt
      integer(c_int), value :: ntau
      integer(c_int) :: mbgpu_heat_acf_dev
    end function
    function mbgpu_ring_matrices(nbead, langevin, nmode_to_cart, cart_to_nmode, tcart_to_tnmode, lambda) &
        bind(c, name="mbgpu_ring_matrices")
      import :: c_int, c_double
      integer(c_int), value :: nbead, langevin
      real(c_double) :: nmode_to_cart(*), cart_to_nmode(*), tcart_to_tnmode(*), lambda(*)
      integer(c_int) :: mbgpu_ring_matrices
    end function
    function mbgpu_ring_transform_dev(ctx, nbead, ncoord, m, in_dev, out_dev) bind(c, name="mbgpu_ring_transform_dev")
      import :: c_int, c_long, c_double, c_ptr
      type(c_ptr), value :: ctx, in_dev, out_dev
      integer(c_int), value :: nbead
      integer(c_long), value :: ncoord
      real(c_double), intent(in) :: m(*)
      integer(c_int) :: mbgpu_ring_transform_dev
    end function
    function mbgpu_ring_spring_dev(ctx, nbead, natom, coeff, fac3, force_scale, mass_dev, x_dev, xc_dev, f_dev, &
                                   epot_spring, epot_deriv) bind(c, name="mbgpu_ring_spring_dev")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, mass_dev, x_dev, xc_dev, f_dev
      integer(c_int), value :: nbead, natom
      real(c_double), value :: coeff, fac3, force_scale
      real(c_double) :: epot_spring, epot_deriv
      integer(c_int) :: mbgpu_ring_spring_dev
    end function
    function mbgpu_ring_spring_dev_noalloc(ctx, nbead, natom, coeff, fac3, force_scale, mass_dev, x_dev, xc_dev, &
                                           f_dev, epot_spring, epot_deriv) &
        bind(c, name="mbgpu_ring_spring_dev_noalloc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, mass_dev, x_dev, xc_dev, f_dev
      integer(c_int), value :: nbead, natom
      real(c_double), value :: coeff, fac3, force_scale
      real(c_double) :: epot_spring, epot_deriv
      integer(c_int) :: mbgpu_ring_spring_dev_noalloc
    end function
    function mbgpu_spme_recip_dev(ctx, cell, order, grid, alpha, n, x_dev, y_dev, z_dev, q_dev, mu_dev, &
                                  phi_dev, efield_dev, f_dev, energy) bind(c, name="mbgpu_spme_recip_dev")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, x_dev, y_dev, z_dev, q_dev, mu_dev, phi_dev, efield_dev, f_dev
      real(c_double), intent(in) :: cell(9)
      integer(c_int), value :: order, n
      integer(c_int), intent(in) :: grid(3)
      real(c_double), value :: alpha
      real(c_double) :: energy
      integer(c_int) :: mbgpu_spme_recip_dev
    end function

    ! ---- timing helpers ----
    function mbgpu_time_forces(ctx, reps, ms_per_eval, ms_pip3b_kernel) bind(c, name="mbgpu_time_forces")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      real(c_double) :: ms_per_eval, ms_pip3b_kernel
      integer(c_int) :: mbgpu_time_forces
    end function
    function mbgpu_time_elec(ctx, reps, ms_sfac, flop_sfac) bind(c, name="mbgpu_time_elec")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      real(c_double) :: ms_sfac, flop_sfac
      integer(c_int) :: mbgpu_time_elec
    end function

    function c_strlen(s) bind(c, name="strlen")
      import :: c_ptr, c_long
      type(c_ptr), value :: s
      integer(c_long) :: c_strlen
    end function
  end interface

contains

  ! The calling thread's last error message as a Fortran string (for mb_abort messages).
  function mbgpu_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    p = mbgpu_last_error()
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate (character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function mbgpu_error_text

end module mbgpu
