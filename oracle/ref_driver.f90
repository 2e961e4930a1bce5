! ORACLE -- test infrastructure only.
!
! C-callable driver around the reference's own hot-path Fortran modules
! (compiled from /root/reference/src by oracle/Makefile.ref into
! oracle/_ref/libdfn_ref.so).  It fills the reference's module state
! (src/variables.f90) from caller arrays, binds the kernel and friction-law
! procedure pointers as src/initialisation.f90:94-110,116-117 does, and calls:
! directIII/directII, eval_kernel_ModeIII/II_T/II_N, ode_RateStateAgingModeIII/II,
! compute_P, rkqs, build_cluster_tree and build_block_cluster_tree.  The solver
! step (ref_solve) is this driver's own restatement of src/solver.f90:59-198,
! since solver.f90 itself needs HDF5, which the image lacks.
module ref_driver
use iso_c_binding
use variables
use fluid_diffusion, only: compute_P
use ode_bs, only: rkqs
use RateStateAgingModeIII, only: ode_RateStateAgingModeIII
use RateStateAgingModeII, only: ode_RateStateAgingModeII
use eval_kernel_hmatrix
use cluster_tree, only: build_cluster_tree
use block_cluster_tree, only: build_block_cluster_tree
implicit none
external :: directII, directIII
real(kind=dp) :: drv_t_try = 100._dp
contains

subroutine ref_init(n, nf, fnel, nodes, elem, dsv, tang, nrmv, av, bv, Dcv, f0v, V0v, nld, sld, perm, kminv, &
                    Snkv, Yv, Pv, scal, ints, nsrc, sfault, sidx, sQ, stb, ste, ncr, cr) bind(C, name='ref_init')
  integer(c_int), value :: n, nf, nsrc, ncr
  integer(c_int64_t), intent(in) :: fnel(nf), sfault(*), sidx(*), cr(6, *)
  real(c_double), intent(in) :: nodes(2, n+1), elem(2, n), dsv(n), tang(2, n), nrmv(2, n)
  real(c_double), intent(in) :: av(n), bv(n), Dcv(n), f0v(n), V0v(n), nld(n), sld(n), perm(n), kminv(n), Snkv(n)
  real(c_double), intent(in) :: Yv(3*n), Pv(n), sQ(*), stb(*), ste(*)
  ! scal = mu, cs, cp, porosity, fluid_comp, rock_comp, dyn_viscosity, tol_solver
  real(c_double), intent(in) :: scal(8)
  ! ints = fracture mode (2/3), kernel (0 direct), permeability_coupling
  integer(c_int), intent(in) :: ints(3)
  integer :: f, s, cnt, k
  call ref_free()
  nb_element = n
  nb_node = n + 1
  nb_fault = nf
  length_Y = 3*n
  allocate(faults(nf), mapping_fault(nf, 2))
  allocate(node(2, n+1), element(2, n), ds(n), tangent(2, n), normal(2, n))
  allocate(normal_left(2, n), normal_right(2, n))
  allocate(Y(3*n), Y_dot(3*n), a(n), b(n), Dc(n), f0(n), V0(n), normal_loading_dot(n), shear_loading_dot(n))
  allocate(SigmaN_dot(n), SigmaT_dot(n), permeability(n), permeability_star(n), permeability_x(1:n+1))
  allocate(kmin(n), Snk(n), P(n), P_dot(n))
  node = nodes; element = elem; ds = dsv; tangent = tang; normal = nrmv
  normal_left = nrmv; normal_right = nrmv
  a = av; b = bv; Dc = Dcv; f0 = f0v; V0 = V0v
  normal_loading_dot = nld; shear_loading_dot = sld
  Y = Yv; Y_dot = 0._dp; P = Pv; P_dot = 0._dp; SigmaN_dot = 0._dp; SigmaT_dot = 0._dp
  V => Y(1:n); theta => Y(n+1:2*n); SigmaN => Y(2*n+1:3*n)
  mu = scal(1); cs = scal(2); cp = scal(3); porosity = scal(4); fluid_comp = scal(5); rock_comp = scal(6)
  dyn_viscosity = scal(7); tol_solver = scal(8)
  permeability_coupling = ints(3)
  permeability = perm; kmin = kminv; Snk = Snkv
  cnt = 0
  do f = 1, nf
    faults(f)%nb_element = int(fnel(f))
    mapping_fault(f, 1) = cnt + 1
    mapping_fault(f, 2) = cnt + fnel(f)
    cnt = cnt + int(fnel(f)) + 1
    faults(f)%nb_source = 0
  end do
  do f = 1, nf
    k = 0
    do s = 1, nsrc
      if (sfault(s) == f) k = k + 1
    end do
    faults(f)%nb_source = k
    allocate(faults(f)%Q(max(k,1)), faults(f)%t_injection_beg(max(k,1)), faults(f)%t_injection_end(max(k,1)))
    allocate(faults(f)%index_injection(max(k,1)))
    k = 0
    do s = 1, nsrc
      if (sfault(s) == f) then
        k = k + 1
        faults(f)%Q(k) = sQ(s); faults(f)%t_injection_beg(k) = stb(s); faults(f)%t_injection_end(k) = ste(s)
        faults(f)%index_injection(k) = sidx(s)
      end if
    end do
  end do
  ! src/initialisation.f90:252-266
  permeability_x = 0._dp
  if (permeability_coupling == 1) then
    permeability_star = kmin + (permeability-kmin) * exp(abs(sigmaN+P)/abs(Snk))
  else
    permeability_star = 0._dp
  end if
  permeability_x(2:n) = 2._dp*permeability(1:n-1)*permeability(2:n)/(permeability(1:n-1)+permeability(2:n))
  nb_crossing = ncr
  allocate(crossing(max(ncr,1), 6))
  do k = 1, ncr
    crossing(k, :) = cr(:, k)
  end do
  time = 0._dp
  drv_t_try = 100._dp
  ! src/initialisation.f90:94-128 (direct kernels only: the H-matrix modules need LAPACK)
  if (ints(1) == 3) then
    fric_law_ptr => ode_RateStateAgingModeIII
    kernel_ptr => directIII
  else
    fric_law_ptr => ode_RateStateAgingModeII
    kernel_ptr => directII
  end if
end subroutine

subroutine ref_free() bind(C, name='ref_free')
  integer :: f
  if (allocated(faults)) then
    do f = 1, size(faults)
      if (allocated(faults(f)%Q)) deallocate(faults(f)%Q, faults(f)%t_injection_beg, faults(f)%t_injection_end, &
                                             faults(f)%index_injection)
    end do
    deallocate(faults)
  end if
  if (allocated(mapping_fault)) deallocate(mapping_fault)
  if (allocated(node)) deallocate(node, element, ds, tangent, normal, normal_left, normal_right)
  if (allocated(Y)) deallocate(Y, Y_dot, a, b, Dc, f0, V0, normal_loading_dot, shear_loading_dot)
  if (allocated(SigmaN_dot)) deallocate(SigmaN_dot, SigmaT_dot, permeability, permeability_x, kmin, Snk, P, P_dot)
  if (associated(permeability_star)) deallocate(permeability_star)
  if (allocated(crossing)) deallocate(crossing)
  if (allocated(node_left)) deallocate(node_left, node_right)
end subroutine

subroutine ref_set_state(Yv, Pv, t) bind(C, name='ref_set_state')
  real(c_double), intent(in) :: Yv(*), Pv(*)
  real(c_double), value :: t
  Y(1:3*nb_element) = Yv(1:3*nb_element)
  P(1:nb_element) = Pv(1:nb_element)
  time = t
end subroutine

subroutine ref_get_state(Yv, Pv, Pdv, t) bind(C, name='ref_get_state')
  real(c_double), intent(out) :: Yv(*), Pv(*), Pdv(*), t
  Yv(1:3*nb_element) = Y(1:3*nb_element)
  Pv(1:nb_element) = P(1:nb_element)
  Pdv(1:nb_element) = P_dot(1:nb_element)
  t = time
end subroutine

subroutine ref_kernel(slip, sT, sN) bind(C, name='ref_kernel')
  real(c_double), intent(inout) :: slip(*)
  real(c_double), intent(out) :: sT(*), sN(*)
  real(kind=dp), allocatable :: s1(:), t1(:), n1(:)
  integer :: ierr
  allocate(s1(nb_element), t1(nb_element), n1(nb_element))
  s1 = slip(1:nb_element); t1 = 0._dp; n1 = 0._dp
  call kernel_ptr(nb_element, node, s1, iprec, t1, n1, ierr, mu)
  sT(1:nb_element) = t1
  sN(1:nb_element) = n1
end subroutine

! M(k,l) = eval_kernel(k,l) with node_left/right = node(:,l)/node(:,l+1) in the
! original element order (identity permutation); op 0 = III or II_T, 1 = II_N.
subroutine ref_eval_matrix(op, mode3, M) bind(C, name='ref_eval_matrix')
  integer(c_int), value :: op, mode3
  real(c_double), intent(out) :: M(nb_element, nb_element)
  integer :: k, l
  real(kind=8) :: yn(2), cn(2,1)
  if (.not. allocated(node_left)) allocate(node_left(2, nb_element), node_right(2, nb_element))
  node_left(:, 1:nb_element) = node(:, 1:nb_element)
  node_right(:, 1:nb_element) = node(:, 2:nb_element+1)
  yn = 0._dp; cn = 0._dp
  do l = 1, nb_element
    do k = 1, nb_element
      if (mode3 == 1) then
        M(k, l) = real(eval_kernel_ModeIII(k, l, yn, cn, 1, 1, 1), 8)
      else if (op == 0) then
        M(k, l) = real(eval_kernel_ModeII_T(k, l, yn, cn, 1, 1, 1), 8)
      else
        M(k, l) = real(eval_kernel_ModeII_N(k, l, yn, cn, 1, 1, 1), 8)
      end if
    end do
  end do
end subroutine

subroutine ref_friction_law(x, Yv, Ydv) bind(C, name='ref_friction_law')
  real(c_double), value :: x
  real(c_double), intent(in) :: Yv(*)
  real(c_double), intent(out) :: Ydv(*)
  real(kind=dp), allocatable :: y1(:), yd1(:)
  allocate(y1(3*nb_element), yd1(3*nb_element))
  y1 = Yv(1:3*nb_element)
  call fric_law_ptr(x, length_Y, y1, yd1)
  Ydv(1:3*nb_element) = yd1
end subroutine

subroutine ref_compute_P(dt, Pt, Pd, ierr) bind(C, name='ref_compute_P')
  real(c_double), value :: dt
  real(c_double), intent(out) :: Pt(*), Pd(*)
  integer(c_int), intent(out) :: ierr
  real(kind=dp), allocatable :: ptmp(:)
  integer :: e
  allocate(ptmp(nb_element))
  ptmp = 0._dp
  call compute_P(dt, ptmp, P_dot, e)
  ierr = e
  Pt(1:nb_element) = ptmp
  Pd(1:nb_element) = P_dot
end subroutine

subroutine ref_rkqs(yv, dydxv, x, htry, eps, yscalv, hdid, hnext) bind(C, name='ref_rkqs')
  real(c_double), intent(inout) :: yv(*), x
  real(c_double), intent(in) :: dydxv(*), yscalv(*)
  real(c_double), value :: htry, eps
  real(c_double), intent(out) :: hdid, hnext
  real(kind=dp), allocatable :: y1(:), d1(:), s1(:)
  allocate(y1(3*nb_element), d1(3*nb_element), s1(3*nb_element))
  y1 = yv(1:3*nb_element); d1 = dydxv(1:3*nb_element); s1 = yscalv(1:3*nb_element)
  call rkqs(y1, d1, length_Y, x, htry, eps, s1, hdid, hnext, fric_law_ptr)
  yv(1:3*nb_element) = y1
end subroutine

subroutine ref_set_t_try(t) bind(C, name='ref_set_t_try')
  real(c_double), value :: t
  drv_t_try = t
end subroutine

! nsteps iterations of the step of src/solver.f90:59-198 (no output files).
subroutine ref_solve(nsteps, t_final, accepted) bind(C, name='ref_solve')
  integer(c_int64_t), value :: nsteps
  real(c_double), value :: t_final
  integer(c_int64_t), intent(out) :: accepted
  real(kind=dp), allocatable :: ptmp(:), yscal(:)
  real(kind=dp) :: t_did, t_next, time_temp
  integer(kind=dp) :: it, f, s
  integer :: e
  allocate(ptmp(nb_element), yscal(3*nb_element))
  accepted = 0
  do it = 1, nsteps
    if (t_final > 0._dp .and. time >= t_final) exit
    if (drv_t_try > 10._dp) then
      do f = 1, nb_fault
        do s = 1, faults(f)%nb_source
          if ((time < faults(f)%t_injection_beg(s)) .and. (time + drv_t_try > faults(f)%t_injection_beg(s))) then
            drv_t_try = 10._dp
          end if
        end do
      end do
    end if
    call compute_P(drv_t_try, ptmp, P_dot, e)
    do while (e == 1)
      drv_t_try = drv_t_try/2._dp
      call compute_P(drv_t_try, ptmp, P_dot, e)
    end do
    time_temp = 0._dp
    call fric_law_ptr(drv_t_try, length_Y, Y, Y_dot)
    yscal = abs(Y) + abs(Y_dot*drv_t_try) + tol_solver
    call rkqs(Y, Y_dot, length_Y, time_temp, drv_t_try, tol_solver, yscal, t_did, t_next, fric_law_ptr)
    drv_t_try = t_next
    if (drv_t_try > 24*60*60) drv_t_try = 24*60*60
    call compute_P(t_did, P, P_dot, e)
    time = time + t_did
    accepted = accepted + 1
    if (any(sigmaN > 0)) exit
  end do
end subroutine

! build_cluster_tree (src/hmatrix/cluster_tree.f90:9-65) on element midpoints cnod (2,n),
! then build_block_cluster_tree (block_cluster_tree.f90:11-53).  Outputs are copied into
! caller buffers of capacity cap_cl clusters / cap_bl blocks.
subroutine ref_trees(n, cnod, nleaf_in, eta_in, cap_cl, cap_bl, mc, nbcl, indc, nbbl, bltree, blinfo) &
           bind(C, name='ref_trees')
  integer(c_int), value :: n, nleaf_in, cap_cl, cap_bl
  real(c_double), value :: eta_in
  real(c_double), intent(inout) :: cnod(2, n)
  integer(c_int), intent(out) :: mc(n), nbcl, indc(*), nbbl, bltree(*)
  integer(c_int64_t), intent(out) :: blinfo(*)
  integer(kind=4), allocatable :: mcl(:), ind(:), sons(:), blt(:,:)
  integer(kind=8), allocatable :: info(:,:)
  integer(kind=4) :: nn, nl, nbc, nbb
  real(kind=8) :: cn(2, n), et
  integer :: i, j
  nn = n; nl = nleaf_in; et = eta_in
  cn = cnod
  call build_cluster_tree(nn, nbc, ind, sons, cn, nl, mcl)
  nbcl = nbc
  mc(1:n) = mcl(1:n)
  if (nbc <= cap_cl) indc(1:2*nbc) = ind(1:2*nbc)
  cnod = cn
  call build_block_cluster_tree(nbc, ind, sons, nn, cn, nbb, blt, info, et)
  nbbl = nbb
  if (nbb <= cap_bl) then
    do i = 1, nbb
      do j = 1, 3
        bltree(3*(i-1)+j) = blt(i, j)
      end do
      do j = 1, 6
        blinfo(6*(i-1)+j) = info(i, j)
      end do
    end do
  end if
end subroutine
end module ref_driver
