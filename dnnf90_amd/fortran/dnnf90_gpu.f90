! dnnf90_gpu.f90 -- Fortran binding of the MI355X library (iso_c_binding).
!
! Part 1 declares the C ABI of include/dnnf90.h and include/dnnf90_amd.h
! (the reference's lib/dnnf90.h entries keep their names).  Part 2 mirrors
! the reference's Fortran module API (lib/net_module.f90:48-51,
! lib/train_module.f90:49-53) with the same routine names and argument
! order -- net_init, net_eval_hod, net_forward_point, net_backward_point,
! net_grad_point, grad_init, grad_zero, opt_adam_step -- over handles, so a
! host written against those modules switches by `use dnnf90_gpu`.  A
! contract violation the reference answers with `stop` stops here too, with
! the library's message.
MODULE dnnf90_gpu
  use iso_c_binding
  implicit none

  PRIVATE
  PUBLIC :: dnnf90_tables_init_dense, dnnf90_tables_init_closure, dnnf90_nderiv, dnnf90_alpha
  PUBLIC :: dnnf90_tabset_dense, dnnf90_tabset_closure, dnnf90_tabset_nderiv, dnnf90_tabset_activation
  PUBLIC :: dnnf90_net_create, dnnf90_net_create_ts, dnnf90_net_load, dnnf90_net_load_ts
  PUBLIC :: dnnf90_net_set_layer, dnnf90_net_free, dnnf90_work_create, dnnf90_work_free, dnnf90_eval
  PUBLIC :: dnnf90_twork_create, dnnf90_twork_free, dnnf90_grad_create, dnnf90_grad_free
  PUBLIC :: dnnf90_grad_zero, dnnf90_grad_point, dnnf90_adam_step
  PUBLIC :: dnnf90_bwork_create, dnnf90_bwork_free, dnnf90_forward_batch, dnnf90_backward_batch
  PUBLIC :: dnnf90_grad_get, dnnf90_net_nweights
  PUBLIC :: net_t, work_t, twork_t, grad_t
  PUBLIC :: tables_dense, tables_closure, net_init, net_set_layer, net_free
  PUBLIC :: work_init, work_free, net_eval_hod, twork_init, twork_free
  PUBLIC :: grad_init, grad_free, grad_zero, grad_get
  PUBLIC :: net_forward_point, net_backward_point, net_grad_point, opt_adam_step

  interface
     integer(c_int) function dnnf90_tables_init_dense(d0, k) bind(C)
       import :: c_int
       integer(c_int), value :: d0, k
     end function
     integer(c_int) function dnnf90_tables_init_closure(d0, k, nseed, seeds) bind(C)
       import :: c_int
       integer(c_int), value :: d0, k, nseed
       integer(c_int), intent(in) :: seeds(*)
     end function
     integer(c_int) function dnnf90_nderiv() bind(C)
       import :: c_int
     end function
     integer(c_int) function dnnf90_alpha(ia, a) bind(C)
       import :: c_int
       integer(c_int), value :: ia
       integer(c_int), intent(out) :: a(*)
     end function
     integer(c_int) function dnnf90_tabset_dense(d0, k) bind(C)
       import :: c_int
       integer(c_int), value :: d0, k
     end function
     integer(c_int) function dnnf90_tabset_closure(d0, k, nseed, seeds) bind(C)
       import :: c_int
       integer(c_int), value :: d0, k, nseed
       integer(c_int), intent(in) :: seeds(*)
     end function
     integer(c_int) function dnnf90_tabset_nderiv(tid) bind(C)
       import :: c_int
       integer(c_int), value :: tid
     end function
     integer(c_int) function dnnf90_tabset_activation(tid, iact) bind(C)
       import :: c_int
       integer(c_int), value :: tid, iact
     end function
     integer(c_int) function dnnf90_net_create_ts(tid, nlayer, dims) bind(C)
       import :: c_int
       integer(c_int), value :: tid, nlayer
       integer(c_int), intent(in) :: dims(*)
     end function
     integer(c_int) function dnnf90_net_load(filename) bind(C)
       import :: c_char, c_int
       character(kind=c_char), intent(in) :: filename(*)
     end function
     integer(c_int) function dnnf90_net_load_ts(tid, filename) bind(C)
       import :: c_char, c_int
       integer(c_int), value :: tid
       character(kind=c_char), intent(in) :: filename(*)
     end function
     integer(c_int) function dnnf90_net_create(nlayer, dims) bind(C)
       import :: c_int
       integer(c_int), value :: nlayer
       integer(c_int), intent(in) :: dims(*)
     end function
     integer(c_int) function dnnf90_net_set_layer(nid, l, wrow) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, l
       real(c_double), intent(in) :: wrow(*)
     end function
     integer(c_int) function dnnf90_net_free(nid) bind(C)
       import :: c_int
       integer(c_int), value :: nid
     end function
     integer(c_int) function dnnf90_work_create(nid) bind(C)
       import :: c_int
       integer(c_int), value :: nid
     end function
     integer(c_int) function dnnf90_work_free(wid) bind(C)
       import :: c_int
       integer(c_int), value :: wid
     end function
     integer(c_int) function dnnf90_eval(nid, wid, x, t) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, wid
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(out) :: t(*)
     end function
     integer(c_int) function dnnf90_twork_create(nid) bind(C)
       import :: c_int
       integer(c_int), value :: nid
     end function
     integer(c_int) function dnnf90_twork_free(tid) bind(C)
       import :: c_int
       integer(c_int), value :: tid
     end function
     integer(c_int) function dnnf90_grad_create(nid) bind(C)
       import :: c_int
       integer(c_int), value :: nid
     end function
     integer(c_int) function dnnf90_grad_free(gid) bind(C)
       import :: c_int
       integer(c_int), value :: gid
     end function
     integer(c_int) function dnnf90_grad_zero(gid) bind(C)
       import :: c_int
       integer(c_int), value :: gid
     end function
     integer(c_int) function dnnf90_grad_point(nid, tid, x, seed, gid) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, tid, gid
       real(c_double), intent(in) :: x(*), seed(*)
     end function
     integer(c_int) function dnnf90_adam_step(nid, gid, eta, beta1, beta2, eps, nbatch, istep) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, gid, nbatch, istep
       real(c_double), value :: eta, beta1, beta2, eps
     end function
     integer(c_int) function dnnf90_bwork_create(nid, maxpt, prec) bind(C)
       import :: c_int
       integer(c_int), value :: nid, maxpt, prec
     end function
     integer(c_int) function dnnf90_bwork_free(bid) bind(C)
       import :: c_int
       integer(c_int), value :: bid
     end function
     integer(c_int) function dnnf90_forward_batch(nid, bid, npt, x, tout, flags) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, bid, npt, flags
       real(c_double), intent(in) :: x(*)
       real(c_double), intent(out) :: tout(*)
     end function
     integer(c_int) function dnnf90_backward_batch(nid, bid, seed, gid, flags) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, bid, gid, flags
       real(c_double), intent(in) :: seed(*)
     end function
     integer(c_int) function dnnf90_grad_get(gid, g) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: gid
       real(c_double), intent(out) :: g(*)
     end function
     integer(c_int) function dnnf90_net_nweights(nid) bind(C)
       import :: c_int
       integer(c_int), value :: nid
     end function
     type(c_ptr) function dnnf90_last_error() bind(C)
       import :: c_ptr
     end function
  end interface

  !> A network on the GPU (cf. net_t, lib/net_module.f90:54-60).
  TYPE :: net_t
     integer :: nid = 0, nlayer = 0, na = 0, nw = 0
     integer,allocatable :: ndim(:)
  END TYPE net_t
  !> Evaluation work space (cf. work_t, lib/net_module.f90:73-87).
  TYPE :: work_t
     integer :: wid = 0
  END TYPE work_t
  !> Training work space (cf. twork_t, lib/train_module.f90:74-90): a
  !! batched work space of one point, so forward and adjoint can be split.
  TYPE :: twork_t
     integer :: bid = 0
  END TYPE twork_t
  !> Gradient accumulator (cf. grad_t, lib/train_module.f90:110-114).
  TYPE :: grad_t
     integer :: gid = 0
  END TYPE grad_t

CONTAINS

  SUBROUTINE check( rc, where )
    integer(c_int),intent(IN) :: rc
    character(*),intent(IN) :: where
    character(kind=c_char),pointer :: msg(:)
    integer :: n
    if ( rc >= 0 ) return
    call c_f_pointer( dnnf90_last_error(), msg, [1024] )
    n = 0
    do while ( n < 1024 )
       if ( msg(n+1) == c_null_char ) exit
       n = n + 1
    end do
    write(*,*) where, ": return code", rc, " ", msg(1:n)
    stop
  END SUBROUTINE check

  !> Build the current table family (init_hod_tables, dense or closure).
  SUBROUTINE tables_dense( d0, k )
    integer,intent(IN) :: d0, k
    call check( dnnf90_tables_init_dense( int(d0,c_int), int(k,c_int) ), "tables_dense" )
  END SUBROUTINE tables_dense

  SUBROUTINE tables_closure( d0, k, nseed, seed_alpha )
    integer,intent(IN) :: d0, k, nseed
    integer,intent(IN) :: seed_alpha(d0,*)
    integer(c_int) :: s(d0*max(nseed,1))
    integer :: i
    do i = 1, nseed
       s((i-1)*d0+1:i*d0) = int( seed_alpha(1:d0,i), c_int )
    end do
    call check( dnnf90_tables_init_closure( int(d0,c_int), int(k,c_int), int(nseed,c_int), s ), &
                "tables_closure" )
  END SUBROUTINE tables_closure

  !> net_init on the current family (lib/net_module.f90:92-143).
  SUBROUTINE net_init( nt, nlayer, ndim )
    type(net_t),intent(OUT) :: nt
    integer,intent(IN) :: nlayer, ndim(nlayer)
    integer(c_int) :: rc
    rc = dnnf90_net_create( int(nlayer,c_int), int(ndim,c_int) )
    call check( rc, "net_init" )
    nt%nid = rc;  nt%nlayer = nlayer
    allocate( nt%ndim(nlayer) );  nt%ndim = ndim
    nt%na = dnnf90_nderiv()
    nt%nw = dnnf90_net_nweights( rc )
  END SUBROUTINE net_init

  !> Layer l (2..nlayer): w(j, 0:ndim(l-1)), bias at i = 0, as nn_weight.dat.
  SUBROUTINE net_set_layer( nt, l, w )
    type(net_t),intent(IN) :: nt
    integer,intent(IN) :: l
    real(8),intent(IN) :: w(:,0:)
    real(c_double) :: row(size(w,1)*size(w,2))
    integer :: j, n
    n = size(w,2)
    do j = 1, size(w,1)
       row((j-1)*n+1:j*n) = w(j,0:n-1)
    end do
    call check( dnnf90_net_set_layer( int(nt%nid,c_int), int(l,c_int), row ), "net_set_layer" )
  END SUBROUTINE net_set_layer

  SUBROUTINE net_free( nt )
    type(net_t),intent(INOUT) :: nt
    call check( dnnf90_net_free( int(nt%nid,c_int) ), "net_free" )
    nt%nid = 0
  END SUBROUTINE net_free

  SUBROUTINE work_init( wk, nt )
    type(work_t),intent(OUT) :: wk
    type(net_t),intent(IN) :: nt
    integer(c_int) :: rc
    rc = dnnf90_work_create( int(nt%nid,c_int) )
    call check( rc, "work_init" )
    wk%wid = rc
  END SUBROUTINE work_init

  SUBROUTINE work_free( wk )
    type(work_t),intent(INOUT) :: wk
    call check( dnnf90_work_free( int(wk%wid,c_int) ), "work_free" )
  END SUBROUTINE work_free

  !> Value and every carried derivative of output 1 (lib/net_module.f90:295-375).
  SUBROUTINE net_eval_hod( nt, wk, x, t )
    type(net_t),intent(IN) :: nt
    type(work_t),intent(INOUT) :: wk
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: t(:)
    call check( dnnf90_eval( int(nt%nid,c_int), int(wk%wid,c_int), x, t ), "net_eval_hod" )
  END SUBROUTINE net_eval_hod

  SUBROUTINE twork_init( tw, nt )
    type(twork_t),intent(OUT) :: tw
    type(net_t),intent(IN) :: nt
    integer(c_int) :: rc
    rc = dnnf90_bwork_create( int(nt%nid,c_int), 1_c_int, 0_c_int )
    call check( rc, "twork_init" )
    tw%bid = rc
  END SUBROUTINE twork_init

  SUBROUTINE twork_free( tw )
    type(twork_t),intent(INOUT) :: tw
    call check( dnnf90_bwork_free( int(tw%bid,c_int) ), "twork_free" )
  END SUBROUTINE twork_free

  SUBROUTINE grad_init( g, nt )
    type(grad_t),intent(OUT) :: g
    type(net_t),intent(IN) :: nt
    integer(c_int) :: rc
    rc = dnnf90_grad_create( int(nt%nid,c_int) )
    call check( rc, "grad_init" )
    g%gid = rc
  END SUBROUTINE grad_init

  SUBROUTINE grad_free( g )
    type(grad_t),intent(INOUT) :: g
    call check( dnnf90_grad_free( int(g%gid,c_int) ), "grad_free" )
  END SUBROUTINE grad_free

  SUBROUTINE grad_zero( g )
    type(grad_t),intent(INOUT) :: g
    call check( dnnf90_grad_zero( int(g%gid,c_int) ), "grad_zero" )
  END SUBROUTINE grad_zero

  !> The accumulated gradient as the flat ABI vector (layer blocks, bias first).
  SUBROUTINE grad_get( g, nt, v )
    type(grad_t),intent(IN) :: g
    type(net_t),intent(IN) :: nt
    real(8),intent(OUT) :: v(nt%nw)
    call check( dnnf90_grad_get( int(g%gid,c_int), v ), "grad_get" )
  END SUBROUTINE grad_get

  !> Forward half (lib/train_module.f90:210-218): tout(1:na) of output 1.
  SUBROUTINE net_forward_point( nt, tw, x, tout )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: tout(:)
    real(c_double) :: tm(nt%ndim(nt%nlayer)*nt%na)
    call check( dnnf90_forward_batch( int(nt%nid,c_int), int(tw%bid,c_int), 1_c_int, x, tm, 0_c_int ), &
                "net_forward_point" )
    tout(1:nt%na) = tm(1:nt%na)
  END SUBROUTINE net_forward_point

  !> Adjoint half (lib/train_module.f90:341-355): seed on output 1.
  SUBROUTINE net_backward_point( nt, tw, seed, g )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(8),intent(IN) :: seed(:)
    type(grad_t),intent(INOUT) :: g
    real(c_double) :: sm(nt%ndim(nt%nlayer)*nt%na)
    sm = 0.d0
    sm(1:nt%na) = seed(1:nt%na)
    call check( dnnf90_backward_batch( int(nt%nid,c_int), int(tw%bid,c_int), sm, int(g%gid,c_int), 0_c_int ), &
                "net_backward_point" )
  END SUBROUTINE net_backward_point

  !> Forward, adjoint and accumulation for one point (lib/train_module.f90:478-489).
  SUBROUTINE net_grad_point( nt, tw, x, seed, g )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(8),intent(IN) :: x(:), seed(:)
    type(grad_t),intent(INOUT) :: g
    real(8) :: tout(nt%na)
    call net_forward_point( nt, tw, x, tout )
    call net_backward_point( nt, tw, seed, g )
  END SUBROUTINE net_grad_point

  !> Adam (lib/train_module.f90:656-683).
  SUBROUTINE opt_adam_step( nt, g, eta, beta1, beta2, eps, nbatch, istep )
    type(net_t),intent(IN) :: nt
    type(grad_t),intent(IN) :: g
    real(8),intent(IN) :: eta, beta1, beta2, eps
    integer,intent(IN) :: nbatch, istep
    call check( dnnf90_adam_step( int(nt%nid,c_int), int(g%gid,c_int), eta, beta1, beta2, eps, &
                                  int(nbatch,c_int), int(istep,c_int) ), "opt_adam_step" )
  END SUBROUTINE opt_adam_step

END MODULE dnnf90_gpu
