! ref_driver.f90 -- test harness over the reference's own lib/ modules
! (TEST INFRASTRUCTURE ONLY; linked into _ref/libdnnf90_ref.so by `make ref`).
!
! The reference's C ABI (lib/dnnf90.h) has no gradient read-out and no
! multi-output forward, so these bind(C) entries call the module routines
! directly: tabset_init (multi_index_bell_module.f90:692), net_init
! (net_module.f90:92), net_forward_point_multi / net_backward_point_multi
! (train_module.f90:226, :459).  Weight and gradient vectors concatenate, for
! l = 2..nlayer, ndim(l) rows of (ndim(l-1)+1) doubles, bias first.
MODULE ref_driver
  use iso_c_binding
  use multi_index_bell_module, only: tabset_t, tabset_init, tabset_free
  use net_module, only: net_t, net_init, net_free
  use train_module, only: twork_t, grad_t, twork_init, twork_free, grad_init, &
       grad_free, net_forward_point_multi, net_backward_point_multi
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

END MODULE ref_driver
