! C-callable entry point for the reference EQWM module (eqwm.f90:90-97).
! Test infrastructure: lets the parity tests call the reference's own
! solve_u_eqwm, compiled from its source, without changing it.
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
  call solve_u_eqwm(umag, nu, rho, h, ncv, stretch, tau_f, iter_f, tau_in, 0.0d0)
  tau = tau_f
  iter = iter_f
end subroutine cf_ref_solve_u_eqwm
