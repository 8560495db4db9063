! Calls the engine through the Fortran binding without touching a device:
! version string, and the argument check of pnfam_gpu_create.
program check_binding
   use iso_c_binding
   use pnfam_gpu
   implicit none
   type(pnfam_gpu_basis) :: b
   type(c_ptr) :: h
   integer(c_int) :: rc
   print '(A,A)', 'version: ', pnfam_gpu_string(pnfam_gpu_version())
   b%nb = 0; b%dqp = 0; b%nghl = 0
   rc = pnfam_gpu_create(b, h)
   print '(A,I0)', 'create rc: ', rc
   print '(A,A)', 'error: ', pnfam_gpu_string(pnfam_gpu_last_error())
end program
