! Error path of the C ABI through the Fortran binding: no device needed.
program abi_errors
  use mocca_amd_c
  implicit none
  type(mocca_mesh) :: mesh
  type(c_ptr) :: h
  integer(c_int) :: rc
  mesh = mocca_mesh(4, 4, 4, 0.8d0, 4.096d0, 1, 0)   ! FD needs >= 6 points
  rc = mocca_solve_create(mesh, 8, MOCCA_FP64, h)
  if (rc .ne. MOCCA_ERR_ARG) stop 2
  print '(a)', 'status ok; message: '//mocca_error_message(c_null_ptr)
  if (index(mocca_error_message(c_null_ptr), 'at least 6 points') .eq. 0) stop 3
  print '(a)', 'FORTRAN_ABI_OK'
end program abi_errors
