! dfn_hip -- ISO_C_BINDING layer between a Fortran host (the DemystiFicatioN
! solver) and the gfx950 library libdfn.so (include/dfn.h).
!
! 1. bind(C) mirrors of dfn_problem / dfn_run_stats and interfaces to the
!    extern "C" entry points a Fortran host calls: one problem (dfn_*), the
!    multi-RHS stress evaluation, ensembles of networks (dfn_ensemble_*) and the
!    step-size proposal network (dfn_mlp_*); dfn_hip_problem fills a dfn_problem
!    from the reference's own arrays.
! 2. Procedures with the reference's own interfaces, so that the reference's
!    pointers and calls can be bound to the device without changing their callers:
!      hip_kernel             -> kernel          (src/variables.f90:17-26)
!      ode_RateStateAgingHIP  -> friction_law    (src/variables.f90:9-14)
!      compute_P_hip          -> compute_P       (src/fluid_diffusion.f90:24)
!      rkqs_hip               -> rkqs            (src/ode_solver/ode_bs.f90:125)
!    They act on the handle created by dfn_hip_setup (one problem per process,
!    as the reference's module state is).
module dfn_hip
use iso_c_binding
implicit none
private

integer(c_int), parameter, public :: DFN_OK = 0, DFN_MODE_II = 2, DFN_MODE_III = 3
integer(c_int), parameter, public :: DFN_KERNEL_DIRECT = 0, DFN_KERNEL_HMATRIX = 1
integer(c_int), parameter, public :: DFN_ERR_CG_TIMEOUT = 9
integer(c_int), parameter, public :: DFN_PREC_FP64 = 0, DFN_PREC_FP32 = 1, DFN_PREC_MIXED = 2, DFN_PREC_FP32_STORE = 3
integer(c_int), parameter, public :: DFN_PREC_MIXED_SPLIT = 4
integer(c_int), parameter, public :: DFN_STEP_RKQS = 0, DFN_STEP_PI = 1, DFN_STEP_LEARNED = 2
integer(c_int), parameter, public :: DFN_MLP_INPUTS = 8, DFN_MLP_PARAMS = 4801

type, bind(C), public :: dfn_problem
  integer(c_int32_t) :: fracture_mode, static_kernel, nb_fault, nb_element
  type(c_ptr) :: fault_nb_element, node
  type(c_ptr) :: a, b, Dc, f0, V0, normal_loading_dot, shear_loading_dot
  type(c_ptr) :: V, theta, sigmaN, P
  type(c_ptr) :: permeability, kmin, kmax, Snk, Lk, Tk
  real(c_double) :: mu, cs, cp
  real(c_double) :: Sigma11_dot, Sigma12_dot, Sigma22_dot, Sigma31_dot, Sigma32_dot
  real(c_double) :: porosity, fluid_comp, rock_comp, dyn_viscosity, fluid_density
  real(c_double) :: tol_solver
  integer(c_int32_t) :: permeability_coupling, nb_source
  type(c_ptr) :: source_fault, source_index, source_Q, source_t_beg, source_t_end
  real(c_double) :: eta, eps_aca
  integer(c_int32_t) :: nleaf, max_rank
end type

type, bind(C), public :: dfn_run_stats
  integer(c_int64_t) :: accepted_steps, rejected_attempts, rhs_evaluations, cg_solves, cg_iterations, cg_halvings
  real(c_double) :: time, t_try, max_V
end type

public :: dfn_create, dfn_destroy, dfn_build_hmat, dfn_set_state, dfn_get_state, dfn_kernel, dfn_friction_law
public :: dfn_compute_P, dfn_rkqs, dfn_solve, dfn_get_stats, dfn_set_output, dfn_flush_output
public :: dfn_set_precision, dfn_kernel_multi
public :: dfn_ensemble_create, dfn_ensemble_destroy, dfn_ensemble_build_hmat, dfn_ensemble_set_precision
public :: dfn_ensemble_set_graph, dfn_ensemble_solve, dfn_ensemble_solve_until, dfn_ensemble_get_state
public :: dfn_ensemble_get_stats, dfn_ensemble_set_controller, dfn_ensemble_set_recording
public :: dfn_ensemble_take_records, dfn_ensemble_info, dfn_ensemble_kernel_timing, dfn_ensemble_kernel_timing_read
public :: dfn_mlp_create, dfn_mlp_destroy, dfn_mlp_set_params, dfn_mlp_get_params, dfn_mlp_forward
public :: dfn_mlp_loss_grad, dfn_mlp_train
public :: dfn_hip_problem, dfn_hip_setup, dfn_hip_close, hip_kernel, ode_RateStateAgingHIP, compute_P_hip, rkqs_hip
public :: dfn_hip_handle

interface
  integer(c_int) function dfn_create(p, h) bind(C, name='dfn_create')
    import :: dfn_problem, c_ptr, c_int
    type(dfn_problem), intent(in) :: p
    type(c_ptr), intent(out) :: h
  end function
  integer(c_int) function dfn_destroy(h) bind(C, name='dfn_destroy')
    import :: c_ptr, c_int
    type(c_ptr), value :: h
  end function
  integer(c_int) function dfn_build_hmat(h) bind(C, name='dfn_build_hmat')
    import :: c_ptr, c_int
    type(c_ptr), value :: h
  end function
  integer(c_int) function dfn_set_state(h, Y, P, time) bind(C, name='dfn_set_state')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    real(c_double), intent(in) :: Y(*), P(*)
    real(c_double), value :: time
  end function
  integer(c_int) function dfn_get_state(h, Y, P, P_dot, time) bind(C, name='dfn_get_state')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    real(c_double), intent(out) :: Y(*), P(*), P_dot(*), time
  end function
  integer(c_int) function dfn_get_stats(h, s) bind(C, name='dfn_get_stats')
    import :: c_ptr, c_int, dfn_run_stats
    type(c_ptr), value :: h
    type(dfn_run_stats), intent(out) :: s
  end function
  integer(c_int) function dfn_kernel(h, slip, stressT, stressN) bind(C, name='dfn_kernel')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    real(c_double), intent(in) :: slip(*)
    real(c_double), intent(out) :: stressT(*), stressN(*)
  end function
  integer(c_int) function dfn_friction_law(h, x, Y, Y_dot) bind(C, name='dfn_friction_law')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    real(c_double), value :: x
    real(c_double), intent(in) :: Y(*)
    real(c_double), intent(out) :: Y_dot(*)
  end function
  integer(c_int) function dfn_compute_P(h, dt, P_temp, P_dot, ier) bind(C, name='dfn_compute_P')
    import :: c_ptr, c_int, c_double, c_int32_t
    type(c_ptr), value :: h
    real(c_double), value :: dt
    real(c_double), intent(out) :: P_temp(*), P_dot(*)
    integer(c_int32_t), intent(out) :: ier
  end function
  integer(c_int) function dfn_rkqs(h, y, dydx, x, htry, eps, yscal, hdid, hnext) bind(C, name='dfn_rkqs')
    import :: c_ptr, c_int, c_double
    type(c_ptr), value :: h
    real(c_double), intent(inout) :: y(*), x
    real(c_double), intent(in) :: dydx(*), yscal(*)
    real(c_double), value :: htry, eps
    real(c_double), intent(out) :: hdid, hnext
  end function
  integer(c_int) function dfn_solve(h, nsteps, t_final, s) bind(C, name='dfn_solve')
    import :: c_ptr, c_int, c_double, c_int64_t, dfn_run_stats
    type(c_ptr), value :: h
    integer(c_int64_t), value :: nsteps
    real(c_double), value :: t_final
    type(dfn_run_stats), intent(out) :: s
  end function
  ! save_data / write_data of src/solver.f90:180-202, 211-317 inside dfn_solve;
  ! dir is a NUL-terminated directory name (c_null_char)
  integer(c_int) function dfn_set_output(h, dir, stride_time, freq_writing_file) bind(C, name='dfn_set_output')
    import :: c_ptr, c_int, c_char, c_int64_t
    type(c_ptr), value :: h
    character(kind=c_char), intent(in) :: dir(*)
    integer(c_int64_t), value :: stride_time, freq_writing_file
  end function
  integer(c_int) function dfn_flush_output(h) bind(C, name='dfn_flush_output')
    import :: c_ptr, c_int
    type(c_ptr), value :: h
  end function
  integer(c_int) function dfn_set_precision(h, precision) bind(C, name='dfn_set_precision')
    import :: c_ptr, c_int, c_int32_t
    type(c_ptr), value :: h
    integer(c_int32_t), value :: precision
  end function
  ! nrhs slip-rate vectors of N values one after another (slip(N, nrhs) in Fortran order)
  integer(c_int) function dfn_kernel_multi(h, nrhs, slip, stressT, stressN) bind(C, name='dfn_kernel_multi')
    import :: c_ptr, c_int, c_int32_t, c_double
    type(c_ptr), value :: h
    integer(c_int32_t), value :: nrhs
    real(c_double), intent(in) :: slip(*)
    real(c_double), intent(out) :: stressT(*), stressN(*)
  end function

  ! -- ensembles of independent networks (one shared device layout) ---------
  integer(c_int) function dfn_ensemble_create(problems, nnet, e) bind(C, name='dfn_ensemble_create')
    import :: dfn_problem, c_ptr, c_int, c_int32_t
    type(dfn_problem), intent(in) :: problems(*)
    integer(c_int32_t), value :: nnet
    type(c_ptr), intent(out) :: e
  end function
  integer(c_int) function dfn_ensemble_destroy(e) bind(C, name='dfn_ensemble_destroy')
    import :: c_ptr, c_int
    type(c_ptr), value :: e
  end function
  integer(c_int) function dfn_ensemble_build_hmat(e) bind(C, name='dfn_ensemble_build_hmat')
    import :: c_ptr, c_int
    type(c_ptr), value :: e
  end function
  integer(c_int) function dfn_ensemble_set_precision(e, precision) bind(C, name='dfn_ensemble_set_precision')
    import :: c_ptr, c_int, c_int32_t
    type(c_ptr), value :: e
    integer(c_int32_t), value :: precision
  end function
  integer(c_int) function dfn_ensemble_set_graph(e, use_graph) bind(C, name='dfn_ensemble_set_graph')
    import :: c_ptr, c_int, c_int32_t
    type(c_ptr), value :: e
    integer(c_int32_t), value :: use_graph
  end function
  integer(c_int) function dfn_ensemble_solve(e, nsteps) bind(C, name='dfn_ensemble_solve')
    import :: c_ptr, c_int, c_int64_t
    type(c_ptr), value :: e
    integer(c_int64_t), value :: nsteps
  end function
  integer(c_int) function dfn_ensemble_solve_until(e, t_end, max_steps) bind(C, name='dfn_ensemble_solve_until')
    import :: c_ptr, c_int, c_double, c_int64_t
    type(c_ptr), value :: e
    real(c_double), value :: t_end
    integer(c_int64_t), value :: max_steps
  end function
  ! net is 0-based; Y = [V | theta | sigmaN] (3N), P and P_dot (N)
  integer(c_int) function dfn_ensemble_get_state(e, net, Y, P, P_dot, time) bind(C, name='dfn_ensemble_get_state')
    import :: c_ptr, c_int, c_int32_t, c_double
    type(c_ptr), value :: e
    integer(c_int32_t), value :: net
    real(c_double), intent(out) :: Y(*), P(*), P_dot(*), time
  end function
  integer(c_int) function dfn_ensemble_get_stats(e, net, s) bind(C, name='dfn_ensemble_get_stats')
    import :: c_ptr, c_int, c_int32_t, dfn_run_stats
    type(c_ptr), value :: e
    integer(c_int32_t), value :: net
    type(dfn_run_stats), intent(out) :: s
  end function
  ! mode DFN_STEP_*; m: a dfn_mlp handle for DFN_STEP_LEARNED (c_null_ptr otherwise)
  integer(c_int) function dfn_ensemble_set_controller(e, mode, m) bind(C, name='dfn_ensemble_set_controller')
    import :: c_ptr, c_int, c_int32_t
    type(c_ptr), value :: e
    integer(c_int32_t), value :: mode
    type(c_ptr), value :: m
  end function
  integer(c_int) function dfn_ensemble_set_recording(e, on) bind(C, name='dfn_ensemble_set_recording')
    import :: c_ptr, c_int, c_int32_t
    type(c_ptr), value :: e
    integer(c_int32_t), value :: on
  end function
  ! X(DFN_MLP_INPUTS, cap) and T(cap); c_null_ptr for X or T only counts the pairs
  integer(c_int) function dfn_ensemble_take_records(e, X, T, cap, count) bind(C, name='dfn_ensemble_take_records')
    import :: c_ptr, c_int, c_int64_t
    type(c_ptr), value :: e, X, T
    integer(c_int64_t), value :: cap
    integer(c_int64_t), intent(out) :: count
  end function
  integer(c_int) function dfn_ensemble_info(e, info) bind(C, name='dfn_ensemble_info')
    import :: c_ptr, c_int, c_int64_t
    type(c_ptr), value :: e
    integer(c_int64_t), intent(out) :: info(9)
  end function
  integer(c_int) function dfn_ensemble_kernel_timing(e, on) bind(C, name='dfn_ensemble_kernel_timing')
    import :: c_ptr, c_int, c_int32_t
    type(c_ptr), value :: e
    integer(c_int32_t), value :: on
  end function
  integer(c_int) function dfn_ensemble_kernel_timing_read(e, ms_mean, count, bytes) &
      bind(C, name='dfn_ensemble_kernel_timing_read')
    import :: c_ptr, c_int, c_double, c_int64_t
    type(c_ptr), value :: e
    real(c_double), intent(out) :: ms_mean, bytes
    integer(c_int64_t), intent(out) :: count
  end function

  ! -- step-size proposal network ------------------------------------------
  integer(c_int) function dfn_mlp_create(m) bind(C, name='dfn_mlp_create')
    import :: c_ptr, c_int
    type(c_ptr), intent(out) :: m
  end function
  integer(c_int) function dfn_mlp_destroy(m) bind(C, name='dfn_mlp_destroy')
    import :: c_ptr, c_int
    type(c_ptr), value :: m
  end function
  ! params(DFN_MLP_PARAMS), norm(2 DFN_MLP_INPUTS + 2) (c_null_ptr keeps the normalisation)
  integer(c_int) function dfn_mlp_set_params(m, params, norm) bind(C, name='dfn_mlp_set_params')
    import :: c_ptr, c_int, c_float
    type(c_ptr), value :: m
    real(c_float), intent(in) :: params(*)
    type(c_ptr), value :: norm
  end function
  integer(c_int) function dfn_mlp_get_params(m, params, norm) bind(C, name='dfn_mlp_get_params')
    import :: c_ptr, c_int, c_float
    type(c_ptr), value :: m
    real(c_float), intent(out) :: params(*), norm(*)
  end function
  integer(c_int) function dfn_mlp_forward(m, X, n, Y) bind(C, name='dfn_mlp_forward')
    import :: c_ptr, c_int, c_float, c_int64_t
    type(c_ptr), value :: m
    real(c_float), intent(in) :: X(*)
    integer(c_int64_t), value :: n
    real(c_float), intent(out) :: Y(*)
  end function
  integer(c_int) function dfn_mlp_loss_grad(m, X, T, n, loss, grad) bind(C, name='dfn_mlp_loss_grad')
    import :: c_ptr, c_int, c_float, c_double, c_int64_t
    type(c_ptr), value :: m
    real(c_float), intent(in) :: X(*), T(*)
    integer(c_int64_t), value :: n
    real(c_double), intent(out) :: loss
    real(c_float), intent(out) :: grad(*)
  end function
  integer(c_int) function dfn_mlp_train(m, X, T, n, epochs, batch, lr, seed, final_loss) bind(C, name='dfn_mlp_train')
    import :: c_ptr, c_int, c_float, c_double, c_int64_t, c_int32_t
    type(c_ptr), value :: m
    real(c_float), intent(in) :: X(*), T(*)
    integer(c_int64_t), value :: n, seed
    integer(c_int32_t), value :: epochs, batch
    real(c_float), value :: lr
    real(c_double), intent(out) :: final_loss
  end function
end interface

type(c_ptr), save :: dfn_hip_handle = c_null_ptr
integer, save :: n_el = 0

contains

! Fills a dfn_problem with the addresses of the reference's own arrays (the
! content of src/variables.f90 after config()/ini_value()) and its scalars.  The
! arrays must stay allocated until the problem has been handed to dfn_create or
! dfn_ensemble_create.
subroutine dfn_hip_problem(prob, mode, kernel, nb_fault, nb_element, fnel, node, a, b, Dc, f0, V0, nld, sld, V, &
                           theta, sigmaN, P, permeability, kmin, kmax, Snk, Lk, Tk, mu, cs, cp, sdot, porosity, &
                           fluid_comp, rock_comp, dyn_viscosity, tol_solver, coupling, nb_source, src_fault, &
                           src_index, src_Q, src_tbeg, src_tend, eta, eps_aca, nleaf, max_rank)
  type(dfn_problem), intent(out) :: prob
  integer, intent(in) :: mode, kernel, nb_fault, nb_element, coupling, nb_source, nleaf, max_rank
  integer(c_int64_t), intent(in), target :: fnel(nb_fault), src_fault(*), src_index(*)
  real(c_double), intent(in), target :: node(2, nb_element+1)
  real(c_double), intent(in), target, dimension(nb_element) :: a, b, Dc, f0, V0, nld, sld, V, theta, sigmaN, P
  real(c_double), intent(in), target, dimension(nb_element) :: permeability, kmin, kmax, Snk, Lk, Tk
  real(c_double), intent(in), target :: src_Q(*), src_tbeg(*), src_tend(*)
  real(c_double), intent(in) :: mu, cs, cp, sdot(5), porosity, fluid_comp, rock_comp, dyn_viscosity, tol_solver
  real(c_double), intent(in) :: eta, eps_aca
  prob%fracture_mode = mode; prob%static_kernel = kernel; prob%nb_fault = nb_fault; prob%nb_element = nb_element
  prob%fault_nb_element = c_loc(fnel); prob%node = c_loc(node)
  prob%a = c_loc(a); prob%b = c_loc(b); prob%Dc = c_loc(Dc); prob%f0 = c_loc(f0); prob%V0 = c_loc(V0)
  prob%normal_loading_dot = c_loc(nld); prob%shear_loading_dot = c_loc(sld)
  prob%V = c_loc(V); prob%theta = c_loc(theta); prob%sigmaN = c_loc(sigmaN); prob%P = c_loc(P)
  prob%permeability = c_loc(permeability); prob%kmin = c_loc(kmin); prob%kmax = c_loc(kmax)
  prob%Snk = c_loc(Snk); prob%Lk = c_loc(Lk); prob%Tk = c_loc(Tk)
  prob%mu = mu; prob%cs = cs; prob%cp = cp
  prob%Sigma11_dot = sdot(1); prob%Sigma12_dot = sdot(2); prob%Sigma22_dot = sdot(3)
  prob%Sigma31_dot = sdot(4); prob%Sigma32_dot = sdot(5)
  prob%porosity = porosity; prob%fluid_comp = fluid_comp; prob%rock_comp = rock_comp
  prob%dyn_viscosity = dyn_viscosity; prob%fluid_density = 1000._c_double; prob%tol_solver = tol_solver
  prob%permeability_coupling = coupling; prob%nb_source = nb_source
  prob%source_fault = c_loc(src_fault); prob%source_index = c_loc(src_index)
  prob%source_Q = c_loc(src_Q); prob%source_t_beg = c_loc(src_tbeg); prob%source_t_end = c_loc(src_tend)
  prob%eta = eta; prob%eps_aca = eps_aca; prob%nleaf = nleaf; prob%max_rank = max_rank
end subroutine

! Creates the device problem from the reference's own arrays, and builds the
! H-matrix when static_kernel is 'hmatrix'.  Every array argument stays owned by
! the caller.
subroutine dfn_hip_setup(mode, kernel, nb_fault, nb_element, fnel, node, a, b, Dc, f0, V0, nld, sld, V, theta, &
                         sigmaN, P, permeability, kmin, kmax, Snk, Lk, Tk, mu, cs, cp, sdot, porosity, fluid_comp, &
                         rock_comp, dyn_viscosity, tol_solver, coupling, nb_source, src_fault, src_index, src_Q, &
                         src_tbeg, src_tend, eta, eps_aca, nleaf, ier)
  integer, intent(in) :: mode, kernel, nb_fault, nb_element, coupling, nb_source, nleaf
  integer(c_int64_t), intent(in), target :: fnel(nb_fault), src_fault(*), src_index(*)
  real(c_double), intent(in), target :: node(2, nb_element+1)
  real(c_double), intent(in), target, dimension(nb_element) :: a, b, Dc, f0, V0, nld, sld, V, theta, sigmaN, P
  real(c_double), intent(in), target, dimension(nb_element) :: permeability, kmin, kmax, Snk, Lk, Tk
  real(c_double), intent(in), target :: src_Q(*), src_tbeg(*), src_tend(*)
  real(c_double), intent(in) :: mu, cs, cp, sdot(5), porosity, fluid_comp, rock_comp, dyn_viscosity, tol_solver
  real(c_double), intent(in) :: eta, eps_aca
  integer, intent(out) :: ier
  type(dfn_problem) :: prob
  call dfn_hip_problem(prob, mode, kernel, nb_fault, nb_element, fnel, node, a, b, Dc, f0, V0, nld, sld, V, theta, &
                       sigmaN, P, permeability, kmin, kmax, Snk, Lk, Tk, mu, cs, cp, sdot, porosity, fluid_comp, &
                       rock_comp, dyn_viscosity, tol_solver, coupling, nb_source, src_fault, src_index, src_Q, &
                       src_tbeg, src_tend, eta, eps_aca, nleaf, 20)
  call dfn_hip_close()
  ier = dfn_create(prob, dfn_hip_handle)
  if (ier /= DFN_OK) return
  n_el = nb_element
  ier = dfn_build_hmat(dfn_hip_handle)
end subroutine

subroutine dfn_hip_close()
  integer(c_int) :: s
  if (c_associated(dfn_hip_handle)) s = dfn_destroy(dfn_hip_handle)
  dfn_hip_handle = c_null_ptr
end subroutine

! kernel interface (src/variables.f90:17-26)
subroutine hip_kernel(nelement, node, slip, iprec, stressT, stressN, ier, mu)
  integer :: nelement, iprec, ier
  real(kind=8), dimension(2, nelement) :: node
  real(kind=8), dimension(nelement) :: slip, stressT, stressN
  real(kind=8) :: mu
  ier = dfn_kernel(dfn_hip_handle, slip, stressT, stressN)
end subroutine

! friction_law interface (src/variables.f90:9-14)
subroutine ode_RateStateAgingHIP(x, nv, Y, Y_dot)
  real(kind=8), intent(in) :: x
  integer :: nv
  real(kind=8), dimension(:), intent(in) :: Y
  real(kind=8), dimension(:), intent(out) :: Y_dot
  integer(c_int) :: s
  s = dfn_friction_law(dfn_hip_handle, x, Y, Y_dot)
  if (s /= DFN_OK) then
    print *, 'dfn_friction_law failed, status ', s
    stop
  end if
end subroutine

! compute_P (src/fluid_diffusion.f90:24): ier = 1 on CG failure, outputs then untouched.
! A solve whose device workgroups lost contact is a fault of the device, not a
! reason to halve the step: it stops.
subroutine compute_P_hip(dt, P_temp, P_dot_temp, ier)
  real(kind=8) :: dt
  real(kind=8), dimension(n_el) :: P_temp, P_dot_temp
  integer :: ier
  integer(c_int32_t) :: e
  integer(c_int) :: s
  s = dfn_compute_P(dfn_hip_handle, dt, P_temp, P_dot_temp, e)
  if (s == DFN_ERR_CG_TIMEOUT) then
    print *, 'dfn_compute_P: device workgroups lost contact, status ', s
    stop
  end if
  ier = int(e)
  if (s /= DFN_OK) ier = 1
end subroutine

! rkqs (src/ode_solver/ode_bs.f90:125); derivs is the handle's friction law
subroutine rkqs_hip(y, dydx, nv, x, htry, eps, yscal, hdid, hnext, derivs)
  real(kind=8), dimension(:), intent(inout) :: y
  real(kind=8), dimension(:), intent(in) :: dydx, yscal
  integer :: nv
  real(kind=8), intent(inout) :: x
  real(kind=8), intent(in) :: htry, eps
  real(kind=8), intent(out) :: hdid, hnext
  external :: derivs
  integer(c_int) :: s
  s = dfn_rkqs(dfn_hip_handle, y, dydx, x, htry, eps, yscal, hdid, hnext)
  if (s == 4) then
    print *, 'stepsize underflow in rkqs'
    stop
  else if (s /= DFN_OK) then
    print *, 'dfn_rkqs failed, status ', s
    stop
  end if
end subroutine

end module dfn_hip
