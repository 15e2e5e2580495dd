! C-callable entry point for the reference EQWM module (eqwm.f90:90-97).
! Test infrastructure: lets the parity tests call the reference's own
! solve_u_eqwm, compiled from its source, without changing it.
!
! solve_u_eqwm takes Maxval over y_cv(1:ncv+1) but never assigns
! y_cv(ncv+1) (eqwm.f90:42,116), so its result depends on whatever the heap
! held there.  Before each call this shim allocates, zeroes and frees two
! blocks of the same size, so the allocator hands them back to y_cv/y_fa
! with a zero in the unassigned slot: the maximum is then the maximum over
! the ncv assigned centroids, the evident intent.
subroutine cf_ref_solve_u_eqwm(umag, nu, rho, h, ncv, stretch, tau_in, tau, iter) &
     bind(C, name="cf_ref_solve_u_eqwm")
  use iso_c_binding
  use EQWM
  real(c_double), value :: umag, nu, rho, h, stretch, tau_in
  integer(c_int), value :: ncv
  real(c_double), intent(out) :: tau
  integer(c_int), intent(out) :: iter
  real(8) :: tau_f
  integer(4) :: iter_f
  real(8), allocatable :: a(:), b(:)
  allocate(a(ncv+1), b(ncv+1))
  a = 0.0d0
  b = 0.0d0
  deallocate(b)
  deallocate(a)
  call solve_u_eqwm(umag, nu, rho, h, ncv, stretch, tau_f, iter_f, tau_in, 0.0d0)
  tau = tau_f
  iter = iter_f
end subroutine cf_ref_solve_u_eqwm
