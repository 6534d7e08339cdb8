! train_module.f90 -- the reference's training module
! (lib/train_module.f90:49-53, 74-114, 163-683) over the MI355X library's C
! ABI.
!
! grad_t keeps the reference's public gradient cube nabla(nlayer, ndmax,
! 0:ndmax); every accumulation adds into it: the cube's live slice is
! copied to the GPU accumulator, the adjoint adds there, and the result is
! copied back, so a host may clear, read or combine nabla directly between
! calls (grad_add sums accumulators in the host's fixed order, as the
! reference's threaded reduction does, tools/example_train.f90:128-170).
! twork_t and bwork_t are GPU work spaces; one twork_t and one grad_t per
! thread keep the calls re-entrant, as in the reference.
MODULE train_module
  use iso_c_binding
  use dnnf90_capi
  use net_module, only: net_t, net_sync, net_pull, net_nweights, cube_to_flat, flat_to_cube

  implicit none

  PRIVATE
  PUBLIC :: twork_t, grad_t, bwork_t
  PUBLIC :: twork_init, twork_free, grad_init, grad_free, grad_zero, grad_add
  PUBLIC :: net_grad_point, net_forward_point, net_forward_point_multi, net_backward_point, net_backward_point_multi
  PUBLIC :: bwork_init, bwork_free, net_grad_batch
  PUBLIC :: opt_sgd_step, opt_adam_step

  !> Training work space: a one-point batched work space on the GPU, which
  !! keeps every layer's state between the forward and the adjoint.
  TYPE :: twork_t
     integer :: bid = 0
  END TYPE twork_t

  !> Value-only batch work space (net_grad_batch).
  TYPE :: bwork_t
     integer :: bid = 0
     integer :: nb = 0
  END TYPE bwork_t

  TYPE :: grad_t
     real(8),allocatable :: nabla(:,:,:)   ! (nlayer, ndmax, 0:ndmax)
     !> The Adam moments stay with the GPU accumulator (opt_adam_step);
     !! these reference components are left unallocated.
     real(8),allocatable :: m(:,:,:)
     real(8),allocatable :: v(:,:,:)
     integer :: gid = 0
     integer :: nlayer = 0
     integer,allocatable :: ndim(:)
  END TYPE grad_t

CONTAINS

  SUBROUTINE twork_init( tw, nt )
    type(twork_t),intent(OUT) :: tw
    type(net_t),intent(IN) :: nt
    integer(c_int) :: rc
    rc = dnnf90_bwork_create( int(nt%nid,c_int), 1_c_int, 0_c_int )
    call capi_check( rc, "twork_init" )
    tw%bid = rc
  END SUBROUTINE twork_init

  SUBROUTINE twork_free( tw )
    type(twork_t),intent(INOUT) :: tw
    if ( tw%bid > 0 ) call capi_check( dnnf90_bwork_free( int(tw%bid,c_int) ), "twork_free" )
    tw%bid = 0
  END SUBROUTINE twork_free

  SUBROUTINE grad_init( g, nt )
    type(grad_t),intent(OUT) :: g
    type(net_t),intent(IN) :: nt
    integer(c_int) :: rc
    rc = dnnf90_grad_create( int(nt%nid,c_int) )
    call capi_check( rc, "grad_init" )
    g%gid = rc
    g%nlayer = nt%nlayer
    allocate( g%ndim(nt%nlayer) )
    g%ndim = nt%ndim
    allocate( g%nabla(nt%nlayer,nt%ndmax,0:nt%ndmax) )
    g%nabla = 0.d0
  END SUBROUTINE grad_init

  SUBROUTINE grad_free( g )
    type(grad_t),intent(INOUT) :: g
    if ( g%gid > 0 ) call capi_check( dnnf90_grad_free( int(g%gid,c_int) ), "grad_free" )
    g%gid = 0
    if ( allocated(g%nabla) ) deallocate( g%nabla )
    if ( allocated(g%m) ) deallocate( g%m )
    if ( allocated(g%v) ) deallocate( g%v )
    if ( allocated(g%ndim) ) deallocate( g%ndim )
  END SUBROUTINE grad_free

  SUBROUTINE grad_zero( g )
    type(grad_t),intent(INOUT) :: g
    g%nabla = 0.d0
  END SUBROUTINE grad_zero

  !> gsum <- gsum + gadd (lib/train_module.f90:191-196).
  SUBROUTINE grad_add( gsum, gadd )
    type(grad_t),intent(INOUT) :: gsum
    type(grad_t),intent(IN) :: gadd
    gsum%nabla = gsum%nabla + gadd%nabla
  END SUBROUTINE grad_add

  ! nabla's live slice to the GPU accumulator, and back after an adjoint
  SUBROUTINE grad_push( g )
    type(grad_t),intent(IN) :: g
    real(8) :: v(sum( g%ndim(2:g%nlayer)*(g%ndim(1:g%nlayer-1)+1) ))
    call cube_to_flat( g%nlayer, g%ndim, g%nabla, v )
    call capi_check( dnnf90_grad_set( int(g%gid,c_int), v ), "grad" )
  END SUBROUTINE grad_push

  SUBROUTINE grad_pull( g )
    type(grad_t),intent(INOUT) :: g
    real(8) :: v(sum( g%ndim(2:g%nlayer)*(g%ndim(1:g%nlayer-1)+1) ))
    call capi_check( dnnf90_grad_get( int(g%gid,c_int), v ), "grad" )
    call flat_to_cube( g%nlayer, g%ndim, v, g%nabla )
  END SUBROUTINE grad_pull

  !> Forward half (lib/train_module.f90:210-218): tout(1:na) of output 1;
  !! the work space keeps every layer's state for net_backward_point.
  SUBROUTINE net_forward_point( nt, tw, x, tout )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: tout(:)
    real(c_double) :: flat(nt%ndim(nt%nlayer)*nt%tab%na)
    call net_sync( nt )
    call capi_check( dnnf90_forward_batch( int(nt%nid,c_int), int(tw%bid,c_int), 1_c_int, x, flat, 0_c_int ), &
                     "net_forward_point" )
    tout(1:nt%tab%na) = flat(1:nt%tab%na)
  END SUBROUTINE net_forward_point

  !> Every output's carried derivatives, tm(1:nout, 1:na)
  !! (lib/train_module.f90:226-237).
  SUBROUTINE net_forward_point_multi( nt, tw, x, tm )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: tm(:,:)
    integer :: nout, na, k
    real(c_double) :: flat(nt%ndim(nt%nlayer)*nt%tab%na)
    nout = nt%ndim(nt%nlayer)
    na = nt%tab%na
    call net_sync( nt )
    call capi_check( dnnf90_forward_batch( int(nt%nid,c_int), int(tw%bid,c_int), 1_c_int, x, flat, 0_c_int ), &
                     "net_forward_point_multi" )
    do k = 1, nout
       tm(k,1:na) = flat((k-1)*na+1:k*na)
    end do
  END SUBROUTINE net_forward_point_multi

  ! the adjoint of the seed plane sm [nout][na] into g%nabla
  SUBROUTINE adjoint( nt, tw, sm, g, where )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(c_double),intent(IN) :: sm(:)
    type(grad_t),intent(INOUT) :: g
    character(*),intent(IN) :: where
    ! a weight change since the forward is refused by the library (the
    ! work space's planes belong to the old weights)
    call net_sync( nt )
    call grad_push( g )
    call capi_check( dnnf90_backward_batch( int(nt%nid,c_int), int(tw%bid,c_int), sm, int(g%gid,c_int), 0_c_int ), &
                     where )
    call grad_pull( g )
  END SUBROUTINE adjoint

  !> Adjoint half (lib/train_module.f90:341-355): seed(1:na) = dL/dT of
  !! output 1, the gradient added to g%nabla.
  SUBROUTINE net_backward_point( nt, tw, seed, g )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(8),intent(IN) :: seed(:)
    type(grad_t),intent(INOUT) :: g
    real(c_double) :: sm(nt%ndim(nt%nlayer)*nt%tab%na)
    sm = 0.d0
    sm(1:nt%tab%na) = seed(1:nt%tab%na)
    call adjoint( nt, tw, sm, g, "net_backward_point" )
  END SUBROUTINE net_backward_point

  !> Adjoint of every output's seed, seedm(1:nout, 1:na)
  !! (lib/train_module.f90:459-475).
  SUBROUTINE net_backward_point_multi( nt, tw, seedm, g )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(8),intent(IN) :: seedm(:,:)
    type(grad_t),intent(INOUT) :: g
    integer :: nout, na, k
    real(c_double) :: sm(nt%ndim(nt%nlayer)*nt%tab%na)
    nout = nt%ndim(nt%nlayer)
    na = nt%tab%na
    if ( size(seedm,1) < nout .or. size(seedm,2) < na ) then
       write(*,*) "net_backward_point_multi: seedm must be at least", nout, " by", na
       stop 1
    end if
    do k = 1, nout
       sm((k-1)*na+1:k*na) = seedm(k,1:na)
    end do
    call adjoint( nt, tw, sm, g, "net_backward_point_multi" )
  END SUBROUTINE net_backward_point_multi

  !> Forward, adjoint and accumulation for one point (lib/train_module.f90:478-489).
  SUBROUTINE net_grad_point( nt, tw, x, seed, g )
    type(net_t),intent(IN) :: nt
    type(twork_t),intent(INOUT) :: tw
    real(8),intent(IN) :: x(:)
    real(8),intent(IN) :: seed(:)
    type(grad_t),intent(INOUT) :: g
    real(8) :: tdum(nt%tab%na)
    call net_forward_point( nt, tw, x, tdum )
    call net_backward_point( nt, tw, seed, g )
  END SUBROUTINE net_grad_point

  SUBROUTINE bwork_init( bw, nt, nb )
    type(bwork_t),intent(OUT) :: bw
    type(net_t),intent(IN) :: nt
    integer,intent(IN) :: nb
    integer(c_int) :: rc
    rc = dnnf90_bwork_create( int(nt%nid,c_int), int(max(nb,1),c_int), 0_c_int )
    call capi_check( rc, "bwork_init" )
    bw%bid = rc
    bw%nb = nb
  END SUBROUTINE bwork_init

  SUBROUTINE bwork_free( bw )
    type(bwork_t),intent(INOUT) :: bw
    if ( bw%bid > 0 ) call capi_check( dnnf90_bwork_free( int(bw%bid,c_int) ), "bwork_free" )
    bw%bid = 0
    bw%nb = 0
  END SUBROUTINE bwork_free

  !> Gradient of L = (coef/2) sum_n ( N(x_n) - y_n )^2 over the batch X(D0, nb),
  !! added to g%nabla (lib/train_module.f90:525-631): the value-only path.
  SUBROUTINE net_grad_batch( nt, bw, X, Y, coef, g )
    type(net_t),intent(IN) :: nt
    type(bwork_t),intent(INOUT) :: bw
    real(8),intent(IN) :: X(:,:), Y(:)
    real(8),intent(IN) :: coef
    type(grad_t),intent(INOUT) :: g
    integer :: nb
    nb = size(X,2)
    if ( nb > bw%nb ) then
       write(*,*) "net_grad_batch: batch of", nb, " points in a work space of", bw%nb
       stop 1
    end if
    call net_sync( nt )
    call grad_push( g )
    call capi_check( dnnf90_grad_batch( int(nt%nid,c_int), int(bw%bid,c_int), int(nb,c_int), X, Y, coef, &
                                        int(g%gid,c_int), 0_c_int ), "net_grad_batch" )
    call grad_pull( g )
  END SUBROUTINE net_grad_batch

  !> w <- w - eta nabla / nbatch (lib/train_module.f90:635-652).
  SUBROUTINE opt_sgd_step( nt, g, eta, nbatch )
    type(net_t),intent(INOUT) :: nt
    type(grad_t),intent(IN) :: g
    real(8),intent(IN) :: eta
    integer,intent(IN) :: nbatch
    integer :: l, j, i
    do i = 0, nt%ndmax
       do j = 1, nt%ndmax
          do l = 2, nt%nlayer
             nt%w(l,j,i) = nt%w(l,j,i) - eta*( g%nabla(l,j,i)/dble(nbatch) )
          end do
       end do
    end do
  END SUBROUTINE opt_sgd_step

  !> Adam (lib/train_module.f90:656-683) on the GPU: the moments stay with
  !! the accumulator, the updated weights come back into nt%w.
  SUBROUTINE opt_adam_step( nt, g, eta, beta1, beta2, eps, nbatch, istep )
    type(net_t),intent(INOUT) :: nt
    type(grad_t),intent(INOUT) :: g
    real(8),intent(IN) :: eta, beta1, beta2, eps
    integer,intent(IN) :: nbatch, istep
    call net_sync( nt )
    call grad_push( g )
    call capi_check( dnnf90_adam_step( int(nt%nid,c_int), int(g%gid,c_int), eta, beta1, beta2, eps, &
                                       int(nbatch,c_int), int(istep,c_int) ), "opt_adam_step" )
    call net_pull( nt )
  END SUBROUTINE opt_adam_step

END MODULE train_module
