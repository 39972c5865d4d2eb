! One GradDesc + ComputeDensity through the Fortran binding, the way a
! Fortran shim of ImaginaryTime/Densities drives the library.
!   graddesc_driver <in.bin> <out.bin>
! in.bin (stream): int32 nx ny nz trc lag nwt; real8 dx dv dt_over_hbar mu
!                  kin(2); int32 qn(4,nwt); real8 occ(nwt); real8 grid(N,4,nwt);
!                  real8 U(2N) B(2N) NablaB(6N) W(18N) DivW(6N) S(6N) A(6N) DivA(2N)
! out.bin: real8 grid(N,4,nwt) after the step, E(nwt), disp(nwt), rho(2N)
program graddesc_driver
  use mocca_amd_c
  implicit none
  character(len=512) :: fin, fout
  integer(c_int) :: nx, ny, nz, trc, lag, nwt, n, rc
  real(c_double) :: dx, dv, dth, mu, kin(2)
  integer(c_int), allocatable :: qn(:)
  real(c_double), allocatable, target :: occ(:), grid(:), U(:), B(:), NB(:), W(:), DW(:),  &
  &                                      S(:), A(:), DA(:), E(:), D(:), rho(:)
  type(mocca_mesh) :: mesh
  type(mocca_fields) :: f
  type(mocca_density) :: dens
  type(c_ptr) :: h
  integer :: iu

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old')
  read(iu) nx, ny, nz, trc, lag, nwt
  read(iu) dx, dv, dth, mu, kin
  n = nx * ny * nz
  allocate(qn(4*nwt), occ(nwt), grid(4*n*nwt), U(2*n), B(2*n), NB(6*n), W(18*n), DW(6*n), &
  &        S(6*n), A(6*n), DA(2*n), E(nwt), D(nwt), rho(2*n))
  read(iu) qn, occ, grid, U, B, NB, W, DW, S, A, DA
  close(iu)

  mesh = mocca_mesh(nx, ny, nz, dx, dv, trc, lag)
  call mocca_check(mocca_solve_create(mesh, nwt, MOCCA_FP64, h), c_null_ptr, 'create')
  call mocca_check(mocca_set_states_meta(h, qn, occ), h, 'set_states_meta')
  call mocca_check(mocca_upload_states(h, 0, nwt, grid), h, 'upload_states')
  f%U = c_loc(U); f%B = c_loc(B); f%NablaB = c_loc(NB); f%W = c_loc(W); f%DivW = c_loc(DW)
  f%S = c_loc(S); f%A = c_loc(A); f%DivA = c_loc(DA); f%kin = kin
  call mocca_check(mocca_set_potentials(h, f), h, 'set_potentials')
  call mocca_check(mocca_grad_desc(h, dth, mu), h, 'grad_desc')
  call mocca_check(mocca_download_states(h, 0, nwt, grid), h, 'download_states')
  call mocca_check(mocca_get_energies(h, E, D), h, 'get_energies')
  dens%rho = c_loc(rho)
  dens%tau = c_null_ptr; dens%nablaJ = c_null_ptr; dens%derrho = c_null_ptr
  dens%laprho = c_null_ptr; dens%vecj = c_null_ptr; dens%vecs = c_null_ptr
  dens%rotvecj = c_null_ptr; dens%rots = c_null_ptr; dens%ders = c_null_ptr
  call mocca_check(mocca_compute_density(h, dens), h, 'compute_density')
  rc = mocca_solve_destroy(h)

  open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(iu) grid, E, D, rho
  close(iu)
  print '(a)', 'FORTRAN_GRADDESC_OK'
end program graddesc_driver
