!-------------------------------------------------------------!
! Module state of the time step (the part of global.f90 the   !
! hot path reads).  The arrays are host mirrors of the        !
! device-resident state owned by the HIP context `ctx`;       !
! sync_to_host / sync_to_device move it when host code needs  !
! it (statistics, monitor, snapshots).                        !
!-------------------------------------------------------------!
Module global

  Use iso_fortran_env, Only : error_unit, Int32, Int64
  Use, Intrinsic :: iso_c_binding
  Use cf_hip
  Implicit None

  Integer(Int32) :: istep, rk_step
  Real   (Int64) :: pi = 4d0*datan(1d0)
  Character(200) :: filein, fileout
  Integer(Int32) :: nsave, nmonitor, nstats, nsteps, nstep_init, random_init
  Real   (Int64) :: Lx, Ly, Lz, dt, t, nu, CFL
  Integer(Int32) :: nx_global, ny_global, nz_global
  Integer(Int32) :: nx, ny, nz, nxg, nyg, nzg
  Real   (Int64), Allocatable, Dimension(:) :: x, y, z, xm, ym, zm, xg, yg, zg
  Real   (Int64), Allocatable, Dimension(:,:,:) :: U, V, W, P, nu_t
  Real   (Int64) :: dPdx, dPdy, dPdz, dPdx_ref, Qflow_x_0, Qflow_y_0, Omega_z
  Integer(Int32) :: x_mass_cte, y_mass_cte, Retau_int
  Integer(Int32) :: LES_model, iwall_model, Dirichlet_nu_t, jref_eqwm, est_Tij
  Real   (Int64) :: alpha_mean_x, alpha_mean_y, alpha_mean_z, alpha_std, freq_mult, fil_size_2
  Integer(Int32) :: gwm_cu, gwm_c_test, gwm_diag, gwm_ai(6)
  Real   (Int64) :: Retau, utau, Qflow_x, Qflow_y, Ucenterline_x
  Real   (Int64), Allocatable, Dimension(:) :: Umean, Vmean, Wmean, U2mean, V2mean, W2mean, UVmean, P2mean
  ! wall-model state saved for restarts: tau_eqwm_initial(nx,2,nzg)
  Real   (Int64), Allocatable, Dimension(:,:,:) :: tau_eqwm_initial
  ! set by compute_statistics when it ran this step; output_data then writes
  ! the pressure saved by the projection instead of computing it
  Logical :: pressure_computed = .False.
  ! monitor clock (system_clock counts)
  Integer(Int64) :: time1 = 0, time2 = 0
  ! settings the reference hard-codes, read from an optional trailing block
  Real   (Int64) :: stretch_y = 2.0d0, amp_ic = 0.1d0, Cs = 0.1d0
  Real   (Int64) :: eqwm_stretch = 1.10d0, eqwm_tol = 1d-7
  Integer(Int32) :: eqwm_ncv = 100, eqwm_max_iter = 100, seed = 2018, precision = 0
  Integer(Int32) :: use_graph = 1

  Type(c_ptr) :: ctx = c_null_ptr

Contains

  Subroutine sync_to_host
    Call cf_check(cf_get_state(ctx, U, V, W), 'cf_get_state')
  End Subroutine sync_to_host

  Subroutine sync_to_device
    Call cf_check(cf_set_state(ctx, U, V, W), 'cf_set_state')
  End Subroutine sync_to_device

  Subroutine sync_scalars_from_device
    Real(c_double) :: s(7)
    Call cf_check(cf_get_scalars(ctx, s), 'cf_get_scalars')
    t = s(1); dt = s(2); dPdx = s(3); dPdz = s(4); Qflow_x_0 = s(5); dPdy = s(6); Qflow_y_0 = s(7)
  End Subroutine sync_scalars_from_device

  Subroutine sync_scalars_to_device
    Real(c_double) :: s(7)
    s = (/ t, dt, dPdx, dPdz, Qflow_x_0, dPdy, Qflow_y_0 /)
    Call cf_check(cf_set_scalars(ctx, s), 'cf_set_scalars')
  End Subroutine sync_scalars_to_device

End Module global
