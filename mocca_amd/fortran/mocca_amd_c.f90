!-------------------------------------------------------------------------------
! Fortran binding of the mocca_amd C ABI (include/mocca_amd.h).
!
! Every entry point of the library is declared here with bind(C), so that
! Fortran shims that keep MOCCa's module interfaces (ImaginaryTime,
! Densities, PotentialMixing, Coulomb, Energy, Pairing) can call the HIP
! library directly through iso_c_binding.  Array arguments are passed as
! assumed-size arrays of real(c_double) / integer(c_int) (contiguous host
! memory in Fortran order, which is the library's layout), structures as
! bind(C) derived types.  mocca_check turns a non-zero status into a stop
! with the library's message, as the reference's stp does.
!-------------------------------------------------------------------------------
module mocca_amd_c
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: MOCCA_OK = 0, MOCCA_ERR_ARG = -1,                &
  &                            MOCCA_ERR_UNSUPPORTED = -2, MOCCA_ERR_HIP = -3,  &
  &                            MOCCA_ERR_STATE = -4, MOCCA_ERR_NUMERIC = -5
  integer(c_int), parameter :: MOCCA_FP64 = 0, MOCCA_MIXED = 1

  type, bind(C) :: mocca_mesh
    integer(c_int) :: nx, ny, nz
    real(c_double) :: dx, dv
    integer(c_int) :: time_reversal, lagrange
  end type

  type, bind(C) :: mocca_fields
    type(c_ptr) :: U, B, NablaB, W, DivW, S, A, DivA
    real(c_double) :: kin(2)
  end type

  type, bind(C) :: mocca_density
    type(c_ptr) :: rho, tau, nablaJ, derrho, laprho, vecj, vecs, rotvecj, rots, ders
  end type

  type, bind(C) :: mocca_functional
    real(c_double) :: B(28)
    real(c_double) :: hbm(2), mass(2)
    real(c_double) :: e2, neutrons, protons
    integer(c_int) :: com1body, cexchange, coulomb, coulomb_maxit, max_moment
    real(c_double) :: damping
    integer(c_int) :: mixing_scheme, pad_
    real(c_double) :: preconu
  end type

  type, bind(C) :: mocca_constraint
    integer(c_int) :: l, m, type, isoswitch, iteration, pad_
    real(c_double) :: value(2)
    real(c_double) :: intensity
  end type

  type, bind(C) :: mocca_pairing
    integer(c_int) :: type, guess_kappa, iterations, solver
    real(c_double) :: strength(2), alpha(2), cut(2), mu(2)
    real(c_double) :: rhosat, hfb_mix, fermi_prec
    real(c_double) :: fermi(2)
  end type

  interface
    integer(c_int) function mocca_solve_create(mesh, nwt, precision, h) bind(C)
      import :: c_int, c_ptr, mocca_mesh
      type(mocca_mesh), intent(in) :: mesh
      integer(c_int), value :: nwt, precision
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function mocca_solve_destroy(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function mocca_last_error(h) bind(C)
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_synchronize(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_set_states_meta(h, qn, occ) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), intent(in) :: qn(*)
      real(c_double), intent(in) :: occ(*)
    end function
    integer(c_int) function mocca_set_occupations(h, occ) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: occ(*)
    end function
    integer(c_int) function mocca_set_energies(h, energy) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: energy(*)
    end function
    integer(c_int) function mocca_get_energies(h, energy, disp) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: energy(*), disp(*)
    end function
    integer(c_int) function mocca_upload_states(h, first, count, grid) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: first, count
      real(c_double), intent(in) :: grid(*)
    end function
    integer(c_int) function mocca_download_states(h, first, count, grid) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: first, count
      real(c_double), intent(out) :: grid(*)
    end function
    integer(c_int) function mocca_reset_updates(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_set_potentials(h, f) bind(C)
      import :: c_int, c_ptr, mocca_fields
      type(c_ptr), value :: h
      type(mocca_fields), intent(in) :: f
    end function
    integer(c_int) function mocca_derive_all(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_download_derivatives(h, state, der, lap) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: state
      real(c_double), intent(out) :: der(*), lap(*)
    end function
    integer(c_int) function mocca_hpsi(h, state, out) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: state
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function mocca_grad_desc(h, dt_over_hbar, momentum) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt_over_hbar, momentum
    end function
    integer(c_int) function mocca_gram_schmidt(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_compute_density(h, out) bind(C)
      import :: c_int, c_ptr, mocca_density
      type(c_ptr), value :: h
      type(mocca_density), intent(in) :: out
    end function
    integer(c_int) function mocca_scf_setup(h, f) bind(C)
      import :: c_int, c_ptr, mocca_functional
      type(c_ptr), value :: h
      type(mocca_functional), intent(in) :: f
    end function
    integer(c_int) function mocca_hf_fill(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_get_occupations(h, occ) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: occ(*)
    end function
    integer(c_int) function mocca_update_densities(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_mix_densities(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_construct_potentials(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_comp_energy(h, e) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: e(*)
    end function
    integer(c_int) function mocca_scf_start(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_convergence_check(h, prec, out) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: prec(*)
      integer(c_int), intent(out) :: out(*)
    end function
    integer(c_int) function mocca_set_lagrange_precision(h, mode) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mode
    end function
    integer(c_int) function mocca_angular_momentum(h, states, canonical, totals) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h, states, canonical, totals
    end function
    ! Checkpoint / restart in the MOCCa v1 format (WriteMOCCa_v1 /
    ! ReadMOCCa_v1, InOut.f90:905-1416); path: a NUL-terminated name
    ! (trim(name)//c_null_char); extra / info: c_null_ptr or the address of
    ! a mocca_v1_info record
    integer(c_int) function mocca_v1_save(h, path, extra) bind(C)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h, extra
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function mocca_v1_load(h, path, recalc, info) bind(C)
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h, info
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: recalc
    end function
    integer(c_int) function mocca_final_iteration(h, row) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: row(*)
    end function
    integer(c_int) function mocca_iterate(h, niter, dt_over_hbar, momentum, energies) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: niter
      real(c_double), value :: dt_over_hbar, momentum
      real(c_double), intent(out) :: energies(*)
    end function
    integer(c_int) function mocca_estimate_parameters(h, iteration, dt, momentum, out) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: iteration
      real(c_double), value :: dt, momentum
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function mocca_estimation_seed(h, grid) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: grid(*)
    end function
    integer(c_int) function mocca_iterate_estimated(h, first_iteration, niter, dt, momentum,   &
    &                                               energies, params) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: first_iteration, niter
      real(c_double), value :: dt, momentum
      real(c_double), intent(out) :: energies(*), params(*)
    end function
    ! solves(nsolves): handles; energies(nsolves): c_loc of each solve's
    ! rows(47, niter + 1), or c_null_ptr
    integer(c_int) function mocca_batch_run(solves, nsolves, niter, dt_over_hbar, momentum, &
    &                                       energies) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), intent(in) :: solves(*)
      integer(c_int), value :: nsolves, niter
      real(c_double), intent(in) :: dt_over_hbar(*), momentum(*)
      type(c_ptr), intent(in) :: energies(*)
    end function
    integer(c_int) function mocca_set_constraints(h, c, count, radd, acut, damping,        &
    &                                             readjust_slowdown) bind(C)
      import :: c_int, c_ptr, c_double, mocca_constraint
      type(c_ptr), value :: h
      type(mocca_constraint), intent(in) :: c(*)
      integer(c_int), value :: count
      real(c_double), value :: radd, acut, damping, readjust_slowdown
    end function
    integer(c_int) function mocca_readjust_constraints(h, iteration, momentum) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: iteration
      real(c_double), value :: momentum
    end function
    integer(c_int) function mocca_alternate_step(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_get_constraints(h, out) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function mocca_set_pairing(h, p) bind(C)
      import :: c_int, c_ptr, mocca_pairing
      type(c_ptr), value :: h
      type(mocca_pairing), intent(in) :: p
    end function
    integer(c_int) function mocca_solve_pairing(h) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mocca_get_pairing(h, out, canonical_occ) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: out(*), canonical_occ(*)
    end function
    integer(c_int) function mocca_get_moments(h, lm, q, count) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: lm(*)
      real(c_double), intent(out) :: q(*)
      integer(c_int), intent(out) :: count
    end function
    integer(c_int) function mocca_get_coulomb(h, vcoul, vcoulx, cpot, iters) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: vcoul(*), vcoulx(*), cpot(*)
      integer(c_int), intent(out) :: iters
    end function
    integer(c_int) function mocca_get_potentials(h, U, B, NablaB, W, DivW, S, A, DivA) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: U(*), B(*), NablaB(*), W(*), DivW(*), S(*), A(*), DivA(*)
    end function
    integer(c_int) function mocca_profile_enable(h, enable) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: enable
    end function
    integer(c_int) function mocca_profile_get(h, name, ms, count) bind(C)
      import :: c_int, c_ptr, c_double, c_char, c_long
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: name(*)
      real(c_double), intent(out) :: ms
      integer(c_long), intent(out) :: count
    end function
  end interface

contains

  function mocca_error_message(h) result(msg)
    ! The library's message for the last failed call on h (or on any handle
    ! when h is c_null_ptr), as a Fortran string.
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: n
    p = mocca_last_error(h)
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [4096])
    n = 0
    do while (n < 4096)
      if (cs(n + 1) == c_null_char) exit
      n = n + 1
    enddo
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function

  subroutine mocca_check(status, h, where)
    ! Stop with the library's message on a non-zero status (the shims'
    ! counterpart of stp, GenInfo.f90:91-177).
    integer(c_int), intent(in) :: status
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: where
    if (status .eq. MOCCA_OK) return
    print '(a, a, i0, a, a)', where, ': mocca_amd status ', status, ': ',            &
    &     mocca_error_message(h)
    stop 1
  end subroutine
end module mocca_amd_c
