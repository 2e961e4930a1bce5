! Fortran host of an ensemble of networks through the dfn_hip ISO_C_BINDING layer:
! reads problem files written by demystification_dfn_amd.fortran_io.write_problem,
! fills one dfn_problem per network (dfn_hip_problem), creates the ensemble,
! builds its operators, sets the operator precision and the step-size controller
! (a trained proposal network read from a raw FP32 weights file for the learned
! one), steps every network nsteps accepted steps while recording (features,
! target) pairs, and writes each network's state and counts, the pairs and, with
! a proposal network, its outputs on the pairs and the result of one training
! epoch on them.
!
!   drive_ensemble OUT NSTEPS PRECISION CONTROLLER WEIGHTS|- PROBLEM...
program drive_ensemble
use iso_c_binding
use dfn_hip
implicit none
type :: net_data
  integer(c_int64_t), allocatable :: fnel(:), sfault(:), sidx(:)
  real(c_double), allocatable :: node(:,:), arr(:,:), sQ(:), stb(:), ste(:)
  real(c_double) :: scal(14), hyp(4)
  integer :: N, nf, mode, kernel, nsrc
end type
type(net_data), allocatable, target :: nd(:)
type(dfn_problem), allocatable :: probs(:)
type(dfn_run_stats) :: rs
type(c_ptr) :: ens, mlp
character(len=512) :: fout, arg, wfile
integer :: nnet, i, u, prec, ctrl
integer(c_int64_t) :: nsteps, hdr(6), cnt, got
integer(c_int) :: st
real(c_double) :: time, loss
real(c_double), allocatable :: Y(:), P(:), Pd(:)
real(c_float), allocatable, target :: X(:,:), T(:), Yf(:), params(:), norm(:)

if (command_argument_count() < 6) then
  print *, 'usage: drive_ensemble OUT NSTEPS PRECISION CONTROLLER WEIGHTS|- PROBLEM...'
  stop 1
end if
call get_command_argument(1, fout)
call get_command_argument(2, arg); read(arg, *) nsteps
call get_command_argument(3, arg); read(arg, *) prec
call get_command_argument(4, arg); read(arg, *) ctrl
call get_command_argument(5, wfile)
nnet = command_argument_count() - 5
allocate(nd(nnet), probs(nnet))
do i = 1, nnet
  call get_command_argument(5 + i, arg)
  call read_problem(trim(arg), nd(i))
  call dfn_hip_problem(probs(i), nd(i)%mode, nd(i)%kernel, nd(i)%nf, nd(i)%N, nd(i)%fnel, nd(i)%node, &
                       nd(i)%arr(:,1), nd(i)%arr(:,2), nd(i)%arr(:,3), nd(i)%arr(:,4), nd(i)%arr(:,5), &
                       nd(i)%arr(:,6), nd(i)%arr(:,7), nd(i)%arr(:,8), nd(i)%arr(:,9), nd(i)%arr(:,10), &
                       nd(i)%arr(:,11), nd(i)%arr(:,12), nd(i)%arr(:,13), nd(i)%arr(:,14), nd(i)%arr(:,15), &
                       nd(i)%arr(:,16), nd(i)%arr(:,17), nd(i)%scal(1), nd(i)%scal(2), nd(i)%scal(3), &
                       nd(i)%scal(4:8), nd(i)%scal(9), nd(i)%scal(10), nd(i)%scal(11), nd(i)%scal(12), &
                       nd(i)%scal(13), int(nd(i)%hyp(4)), nd(i)%nsrc, nd(i)%sfault, nd(i)%sidx, nd(i)%sQ, &
                       nd(i)%stb, nd(i)%ste, nd(i)%scal(14), nd(i)%hyp(1), int(nd(i)%hyp(2)), int(nd(i)%hyp(3)))
end do
call check(dfn_ensemble_create(probs, int(nnet, c_int32_t), ens), 'dfn_ensemble_create')
call check(dfn_ensemble_build_hmat(ens), 'dfn_ensemble_build_hmat')
call check(dfn_ensemble_set_precision(ens, int(prec, c_int32_t)), 'dfn_ensemble_set_precision')
mlp = c_null_ptr
if (trim(wfile) /= '-') then
  allocate(params(DFN_MLP_PARAMS), norm(2*DFN_MLP_INPUTS + 2))
  open(newunit=u, file=trim(wfile), access='stream', form='unformatted', status='old')
  read(u) params, norm
  close(u)
  call check(dfn_mlp_create(mlp), 'dfn_mlp_create')
  call check(dfn_mlp_set_params(mlp, params, c_loc(norm)), 'dfn_mlp_set_params')
end if
call check(dfn_ensemble_set_controller(ens, int(ctrl, c_int32_t), mlp), 'dfn_ensemble_set_controller')
call check(dfn_ensemble_set_recording(ens, 1_c_int32_t), 'dfn_ensemble_set_recording')
call check(dfn_ensemble_solve(ens, nsteps), 'dfn_ensemble_solve')

open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
write(u) real(nnet, c_double)
do i = 1, nnet
  allocate(Y(3*nd(i)%N), P(nd(i)%N), Pd(nd(i)%N))
  call check(dfn_ensemble_get_state(ens, int(i - 1, c_int32_t), Y, P, Pd, time), 'dfn_ensemble_get_state')
  call check(dfn_ensemble_get_stats(ens, int(i - 1, c_int32_t), rs), 'dfn_ensemble_get_stats')
  write(u) time, Y, P, real(rs%accepted_steps, c_double), real(rs%rejected_attempts, c_double)
  deallocate(Y, P, Pd)
end do
call check(dfn_ensemble_take_records(ens, c_null_ptr, c_null_ptr, 0_c_int64_t, cnt), 'dfn_ensemble_take_records')
allocate(X(DFN_MLP_INPUTS, max(cnt, 1_c_int64_t)), T(max(cnt, 1_c_int64_t)))
got = 0
if (cnt > 0) call check(dfn_ensemble_take_records(ens, c_loc(X), c_loc(T), cnt, got), 'dfn_ensemble_take_records')
write(u) real(got, c_double), X(:, 1:got), T(1:got)
if (c_associated(mlp) .and. got > 0) then
  allocate(Yf(got))
  call check(dfn_mlp_forward(mlp, X, got, Yf), 'dfn_mlp_forward')
  call check(dfn_mlp_train(mlp, X, T, got, 1_c_int32_t, 64_c_int32_t, 1.e-3_c_float, 5_c_int64_t, loss), &
             'dfn_mlp_train')
  call check(dfn_mlp_get_params(mlp, params, norm), 'dfn_mlp_get_params')
  write(u) Yf, loss, params
end if
close(u)
if (c_associated(mlp)) st = dfn_mlp_destroy(mlp)
st = dfn_ensemble_destroy(ens)

contains

subroutine check(status, what)
  integer(c_int), intent(in) :: status
  character(len=*), intent(in) :: what
  if (status /= DFN_OK) then
    print *, what, ' failed, status ', status
    stop 2
  end if
end subroutine

subroutine read_problem(path, d)
  character(len=*), intent(in) :: path
  type(net_data), intent(inout) :: d
  integer :: v
  open(newunit=v, file=path, access='stream', form='unformatted', status='old')
  read(v) hdr
  d%N = int(hdr(1)); d%nf = int(hdr(2)); d%mode = int(hdr(3)); d%kernel = int(hdr(4)); d%nsrc = int(hdr(5))
  allocate(d%fnel(d%nf), d%node(2, d%N + 1), d%arr(d%N, 17))
  allocate(d%sfault(max(d%nsrc, 1)), d%sidx(max(d%nsrc, 1)), d%sQ(max(d%nsrc, 1)), d%stb(max(d%nsrc, 1)), &
           d%ste(max(d%nsrc, 1)))
  read(v) d%fnel
  read(v) d%node
  read(v) d%arr
  read(v) d%scal
  if (d%nsrc > 0) read(v) d%sfault(1:d%nsrc), d%sidx(1:d%nsrc), d%sQ(1:d%nsrc), d%stb(1:d%nsrc), d%ste(1:d%nsrc)
  read(v) d%hyp
  close(v)
end subroutine

end program
