! ref_driver.f90 -- test harness over the reference's own lib/ modules
! (TEST INFRASTRUCTURE ONLY; linked into _ref/libdnnf90_ref.so by `make ref`).
!
! The reference's C ABI (lib/dnnf90.h) has no gradient read-out and no
! multi-output forward, so these bind(C) entries call the module routines
! directly: tabset_init (multi_index_bell_module.f90:692), net_init
! (net_module.f90:92), net_forward_point_multi / net_backward_point_multi
! (train_module.f90:226, :459), the filter (kalman_module.f90:107, :240, :280,
! :309, :512, :532) and the committee (committee_module.f90:79).  Weight and
! gradient vectors concatenate, for l = 2..nlayer, ndim(l) rows of
! (ndim(l-1)+1) doubles, bias first.
MODULE ref_driver
  use iso_c_binding
  use multi_index_bell_module, only: tabset_t, tabset_init, tabset_free
  use net_module, only: net_t, net_init, net_free
  use train_module, only: twork_t, grad_t, twork_init, twork_free, grad_init, &
       grad_free, net_forward_point_multi, net_backward_point_multi, opt_adam_step
  use kalman_module, only: kalman_t, kf_init, kf_free, kf_update_obs, kf_update_resid, &
       kf_update_resid_multi, kf_iekf_begin, kf_iekf_iter
  use committee_module, only: committee_t, comm_init, comm_free, comm_eval
  use net_module, only: work_t, work_init, work_free
  implicit none
CONTAINS

  SUBROUTINE build_net( d0, k, nseed, seeds, iact, nlayer, dims, w, ts, nt )
    integer(c_int),intent(IN) :: d0, k, nseed, iact, nlayer
    integer(c_int),intent(IN) :: seeds(*), dims(*)
    real(c_double),intent(IN) :: w(*)
    type(tabset_t),intent(OUT) :: ts
    type(net_t),intent(OUT) :: nt
    integer :: sa(max(d0,1),max(nseed,1)), s, l, j, p
    do s = 1, nseed
       sa(1:d0,s) = int( seeds(1+(s-1)*d0 : s*d0) )
    end do
    call tabset_init( ts, int(d0), int(k), int(nseed), sa )
    ts%iact = int(iact)
    call net_init( nt, int(nlayer), int(dims(1:nlayer)), ts )
    p = 0
    do l = 2, nlayer
       do j = 1, dims(l)
          nt%w(l,j,0:dims(l-1)) = w(p+1:p+dims(l-1)+1)
          p = p + dims(l-1) + 1
       end do
    end do
  END SUBROUTINE build_net

  !> tout(npt, nout, na) in C order: every carried derivative of every output.
  INTEGER(c_int) FUNCTION refx_eval( d0, k, nseed, seeds, iact, nlayer, dims, w, &
                                     npt, x, tout ) bind(C)
    integer(c_int),value :: d0, k, nseed, iact, nlayer, npt
    integer(c_int),intent(IN) :: seeds(*), dims(*)
    real(c_double),intent(IN) :: w(*), x(*)
    real(c_double),intent(OUT) :: tout(*)
    type(tabset_t) :: ts
    type(net_t) :: nt
    type(twork_t) :: tw
    real(8),allocatable :: tm(:,:)
    integer :: n, i, ia, nout, na, p
    call build_net( d0, k, nseed, seeds, iact, nlayer, dims, w, ts, nt )
    na = nt%tab%na;  nout = dims(nlayer)
    call twork_init( tw, nt )
    allocate( tm(nout,na) )
    p = 0
    do n = 1, npt
       call net_forward_point_multi( nt, tw, x((n-1)*d0+1:n*d0), tm )
       do i = 1, nout
          do ia = 1, na
             p = p + 1
             tout(p) = tm(i,ia)
          end do
       end do
    end do
    call twork_free( tw );  call net_free( nt );  call tabset_free( ts )
    refx_eval = 0
  END FUNCTION refx_eval

  !> grad(nw) += sum over points of the adjoint of seed(npt, nout, na).
  INTEGER(c_int) FUNCTION refx_grad( d0, k, nseed, seeds, iact, nlayer, dims, w, &
                                     npt, x, seed, grad ) bind(C)
    integer(c_int),value :: d0, k, nseed, iact, nlayer, npt
    integer(c_int),intent(IN) :: seeds(*), dims(*)
    real(c_double),intent(IN) :: w(*), x(*), seed(*)
    real(c_double),intent(INOUT) :: grad(*)
    type(tabset_t) :: ts
    type(net_t) :: nt
    type(twork_t) :: tw
    type(grad_t) :: g
    real(8),allocatable :: tm(:,:), sm(:,:)
    integer :: n, i, ia, nout, na, p, l, j
    call build_net( d0, k, nseed, seeds, iact, nlayer, dims, w, ts, nt )
    na = nt%tab%na;  nout = dims(nlayer)
    call twork_init( tw, nt )
    call grad_init( g, nt )
    allocate( tm(nout,na), sm(nout,na) )
    p = 0
    do n = 1, npt
       call net_forward_point_multi( nt, tw, x((n-1)*d0+1:n*d0), tm )
       do i = 1, nout
          do ia = 1, na
             p = p + 1
             sm(i,ia) = seed(p)
          end do
       end do
       call net_backward_point_multi( nt, tw, sm, g )
    end do
    p = 0
    do l = 2, nlayer
       do j = 1, dims(l)
          grad(p+1:p+dims(l-1)+1) = grad(p+1:p+dims(l-1)+1) + g%nabla(l,j,0:dims(l-1))
          p = p + dims(l-1) + 1
       end do
    end do
    call grad_free( g );  call twork_free( tw );  call net_free( nt );  call tabset_free( ts )
    refx_grad = 0
  END FUNCTION refx_grad

  !> A sequence of nobs filter updates on the reference's kalman_module.
  !! Observation n is the point x(:,n) with seed(n, nout, na) (C order) and
  !! the value val(n):
  !!   mode 0  kf_update_obs( x, seed(output 1), target = val(n) )
  !!   mode 1  kf_update_resid_multi( x, seedm, resid = val(n), rnoise(n) )
  !!   mode 2  kf_iekf_begin, then niter kf_iekf_iter on R = seed.T - val(n)
  !!           re-evaluated at the running weights, rnoise(n)
  !!   mode 3  the collocation observation of app/sgd_batch_module.f90:547-550:
  !!           R = val(n) + sum_k c_k [T_1 *] T_slot(k) on output 1, seed dR/dT
  !!           (set_pinn_seed with R = fac = 1), kf_update_resid( x, seed, R )
  !! slot(:) are 0-based slots.  Returns the weights, the covariance as the
  !! filter holds it (P(nw,nw), or Pb(maxb,maxb,nb)), lambda and the gated count.
  INTEGER(c_int) FUNCTION refx_kf_seq( d0, k, nseed, seeds, iact, nlayer, dims, w, &
       p0, lam_ini, lam0, decoupled, gate, q, mode, niter, nobs, x, seed, val, rnoise, &
       nterm, coeff, slot, nonlin, wout, pout, lam_out, ngated ) bind(C)
    integer(c_int),value :: d0, k, nseed, iact, nlayer, decoupled, mode, niter, nobs, nterm
    real(c_double),value :: p0, lam_ini, lam0, gate, q
    integer(c_int),intent(IN) :: seeds(*), dims(*), slot(*), nonlin(*)
    real(c_double),intent(IN) :: w(*), x(*), seed(*), val(*), rnoise(*), coeff(*)
    real(c_double),intent(OUT) :: wout(*), pout(*), lam_out
    integer(c_int),intent(OUT) :: ngated
    type(tabset_t) :: ts
    type(net_t) :: nt
    type(twork_t) :: tw
    type(grad_t) :: g
    type(kalman_t) :: kf
    real(8),allocatable :: sm(:,:), tm(:,:), sv(:), xo(:)
    real(8) :: R
    integer :: n, i, ia, nout, na, p, l, j, it, kk, s
    call build_net( d0, k, nseed, seeds, iact, nlayer, dims, w, ts, nt )
    na = nt%tab%na;  nout = dims(nlayer)
    call twork_init( tw, nt );  call grad_init( g, nt )
    call kf_init( kf, nt, p0, lam_ini, lam0, decoupled /= 0 )
    kf%gate = gate;  kf%q = q
    allocate( sm(nout,na), tm(nout,na), sv(na), xo(d0) )
    do n = 1, nobs
       p = (n-1)*nout*na
       do i = 1, nout
          do ia = 1, na
             sm(i,ia) = seed(p + (i-1)*na + ia)
          end do
       end do
       xo = x((n-1)*d0+1:n*d0)
       select case ( mode )
       case ( 0 )
          sv = sm(1,:)
          call kf_update_obs( nt, tw, g, kf, xo, sv, val(n) )
       case ( 1 )
          call kf_update_resid_multi( nt, tw, g, kf, xo, sm, val(n), rnoise(n) )
       case ( 2 )
          call kf_iekf_begin( nt, kf )
          do it = 1, niter
             call net_forward_point_multi( nt, tw, xo, tm )
             R = 0.d0
             do i = 1, nout
                do ia = 1, na
                   R = R + sm(i,ia)*tm(i,ia)
                end do
             end do
             R = R - val(n)
             call kf_iekf_iter( nt, tw, g, kf, xo, sm, R, rnoise(n), it == niter )
          end do
       case ( 3 )
          call net_forward_point_multi( nt, tw, xo, tm )
          R = val(n)
          sv = 0.d0
          do kk = 1, nterm
             s = slot(kk) + 1
             if ( nonlin(kk) /= 0 ) then
                R = R + coeff(kk)*tm(1,1)*tm(1,s)
                sv(s) = sv(s) + coeff(kk)*tm(1,1)
                sv(1) = sv(1) + coeff(kk)*tm(1,s)
             else
                R = R + coeff(kk)*tm(1,s)
                sv(s) = sv(s) + coeff(kk)
             end if
          end do
          call kf_update_resid( nt, tw, g, kf, xo, sv, R )
       case default
          refx_kf_seq = -1
          return
       end select
    end do
    p = 0
    do l = 2, nlayer
       do j = 1, dims(l)
          wout(p+1:p+dims(l-1)+1) = nt%w(l,j,0:dims(l-1))
          p = p + dims(l-1) + 1
       end do
    end do
    if ( kf%decoupled ) then
       pout(1:size(kf%Pb)) = reshape( kf%Pb, (/ size(kf%Pb) /) )
    else
       pout(1:size(kf%P)) = reshape( kf%P, (/ size(kf%P) /) )
    end if
    lam_out = kf%lam
    ngated = kf%ngated
    call kf_free( kf );  call grad_free( g );  call twork_free( tw )
    call net_free( nt );  call tabset_free( ts )
    refx_kf_seq = 0
  END FUNCTION refx_kf_seq

  !> comm_eval over nmem members of one shape, w(nmem, nw) in C order;
  !! tmean / tstd (npt, na) of output 1.
  INTEGER(c_int) FUNCTION refx_comm_eval( d0, k, nseed, seeds, iact, nlayer, dims, &
       nmem, w, npt, x, tmean, tstd ) bind(C)
    integer(c_int),value :: d0, k, nseed, iact, nlayer, nmem, npt
    integer(c_int),intent(IN) :: seeds(*), dims(*)
    real(c_double),intent(IN) :: w(*), x(*)
    real(c_double),intent(OUT) :: tmean(*), tstd(*)
    type(tabset_t) :: ts
    type(net_t) :: nt
    type(committee_t) :: cm
    type(work_t) :: wk
    real(8),allocatable :: tm(:), tsd(:)
    integer :: m, n, l, j, p, na, nw
    call build_net( d0, k, nseed, seeds, iact, nlayer, dims, w, ts, nt )
    na = nt%tab%na
    nw = 0
    do l = 2, nlayer
       nw = nw + dims(l)*(dims(l-1)+1)
    end do
    call comm_init( cm, int(nmem), int(nlayer), int(dims(1:nlayer)), ts )
    do m = 1, nmem
       p = (m-1)*nw
       do l = 2, nlayer
          do j = 1, dims(l)
             cm%mem(m)%w(l,j,0:dims(l-1)) = w(p+1:p+dims(l-1)+1)
             p = p + dims(l-1) + 1
          end do
       end do
    end do
    call work_init( wk, cm%mem(1) )
    allocate( tm(na), tsd(na) )
    do n = 1, npt
       call comm_eval( cm, wk, x((n-1)*d0+1:n*d0), tm, tsd )
       tmean((n-1)*na+1:n*na) = tm
       tstd((n-1)*na+1:n*na) = tsd
    end do
    call work_free( wk );  call comm_free( cm )
    call net_free( nt );  call tabset_free( ts )
    refx_comm_eval = 0
  END FUNCTION refx_comm_eval

  !> The collocation step on output 1, one point at a time: forward, the
  !! residual R = sum_k coeff(k) [T_1 *] T_slot(k) (slot 1-based; nonlin(k)
  !! multiplies by the value), the seed fac R dR/dT, the adjoint into one
  !! accumulator (app/pinn_module.f90:32-65, app/sgd_batch_module.f90:733-737
  !! with the trainer's lnet_* calls replaced by the module routines).
  !! grad(nw) += the batch gradient; rsq = sum_n R_n^2; secs = the seconds of
  !! the point loop alone (set-up excluded).
  INTEGER(c_int) FUNCTION refx_pinn_step( d0, k, nseed, seeds, iact, nlayer, dims, w, &
                                          npt, x, nterm, coeff, slot, nonlin, fac, &
                                          grad, rsq, secs ) bind(C)
    integer(c_int),value :: d0, k, nseed, iact, nlayer, npt, nterm
    integer(c_int),intent(IN) :: seeds(*), dims(*), slot(*), nonlin(*)
    real(c_double),value :: fac
    real(c_double),intent(IN) :: w(*), x(*), coeff(*)
    real(c_double),intent(INOUT) :: grad(*)
    real(c_double),intent(OUT) :: rsq, secs
    type(tabset_t) :: ts
    type(net_t) :: nt
    type(twork_t) :: tw
    type(grad_t) :: g
    real(8),allocatable :: tm(:,:), sm(:,:)
    real(8) :: R
    integer :: n, nout, na, p, l, j, it
    integer(8) :: c0, c1, crate
    call build_net( d0, k, nseed, seeds, iact, nlayer, dims, w, ts, nt )
    na = nt%tab%na;  nout = dims(nlayer)
    call twork_init( tw, nt )
    call grad_init( g, nt )
    allocate( tm(nout,na), sm(nout,na) )
    rsq = 0.d0
    call system_clock( c0, crate )
    do n = 1, npt
       call net_forward_point_multi( nt, tw, x((n-1)*d0+1:n*d0), tm )
       R = 0.d0
       do it = 1, nterm
          if ( nonlin(it) /= 0 ) then
             R = R + coeff(it)*tm(1,1)*tm(1,slot(it))
          else
             R = R + coeff(it)*tm(1,slot(it))
          end if
       end do
       rsq = rsq + R*R
       sm = 0.d0
       do it = 1, nterm
          if ( nonlin(it) /= 0 ) then
             sm(1,slot(it)) = sm(1,slot(it)) + fac*R*coeff(it)*tm(1,1)
             sm(1,1) = sm(1,1) + fac*R*coeff(it)*tm(1,slot(it))
          else
             sm(1,slot(it)) = sm(1,slot(it)) + fac*R*coeff(it)
          end if
       end do
       call net_backward_point_multi( nt, tw, sm, g )
    end do
    call system_clock( c1 )
    secs = dble(c1 - c0) / dble(crate)
    p = 0
    do l = 2, nlayer
       do j = 1, dims(l)
          grad(p+1:p+dims(l-1)+1) = grad(p+1:p+dims(l-1)+1) + g%nabla(l,j,0:dims(l-1))
          p = p + dims(l-1) + 1
       end do
    end do
    call grad_free( g );  call twork_free( tw );  call net_free( nt );  call tabset_free( ts )
    refx_pinn_step = 0
  END FUNCTION refx_pinn_step

  !> nsteps Adam steps of the reference's opt_adam_step
  !! (train_module.f90:656-683) with the fixed gradient g(nw), steps istep0,
  !! istep0+1, ...; w(nw) in and out (flat, layer blocks, bias first).
  INTEGER(c_int) FUNCTION refx_adam( nlayer, dims, w, g, eta, beta1, beta2, eps, &
                                     nbatch, istep0, nsteps ) bind(C)
    integer(c_int),value :: nlayer, nbatch, istep0, nsteps
    integer(c_int),intent(IN) :: dims(*)
    real(c_double),intent(INOUT) :: w(*)
    real(c_double),intent(IN) :: g(*)
    real(c_double),value :: eta, beta1, beta2, eps
    type(tabset_t) :: ts
    type(net_t) :: nt
    type(grad_t) :: gr
    integer :: s, l, j, p
    integer(c_int) :: dummy(1)
    dummy = 0
    call build_net( dims(1), 1_c_int, 0_c_int, dummy, 0_c_int, nlayer, dims, w, ts, nt )
    call grad_init( gr, nt )
    gr%nabla = 0.d0
    p = 0
    do l = 2, nlayer
       do j = 1, dims(l)
          gr%nabla(l,j,0:dims(l-1)) = g(p+1:p+dims(l-1)+1)
          p = p + dims(l-1) + 1
       end do
    end do
    do s = 0, nsteps - 1
       call opt_adam_step( nt, gr, eta, beta1, beta2, eps, int(nbatch), int(istep0 + s) )
    end do
    p = 0
    do l = 2, nlayer
       do j = 1, dims(l)
          w(p+1:p+dims(l-1)+1) = nt%w(l,j,0:dims(l-1))
          p = p + dims(l-1) + 1
       end do
    end do
    call grad_free( gr );  call net_free( nt );  call tabset_free( ts )
    refx_adam = 0
  END FUNCTION refx_adam

END MODULE ref_driver
