! lib_net_module.f90 -- the stock trainer's bridge to the network library
! (app/lib_net_module.f90:26-618), routed to the MI355X library's C ABI.
!
! The trainer (app/sgd_batch_module.f90, validation_module, io_module,
! hod_check_module, api_module) calls these lnet_* routines per point: a
! forward pass that publishes every carried derivative of the point, then the
! adjoint of a seed the trainer forms from them, accumulated into the
! trainer's own gradient cube, or written as a Gauss-Newton row; the value-only
! batch path; and the extended Kalman filter.  This module keeps that
! interface (the same names, arguments and stops) and runs every evaluation,
! adjoint and filter update on the GPU: a one-point FP64 work space for the
! per-point calls, a batched one for the value path, the device-resident
! filter for the Kalman entries.  The trainer's master weights stay its own;
! lnet_sync_weights copies them to the GPU network, lnet_export_weights
! brings back what the filter moved.
!
! The network's table family is the trainer's: when the trainer has built
! its derivative tables (a high-order or collocation run) the library builds
! the same downward-closed set from its carried multi-indices (same set,
! same slot order, checked); otherwise the one-index (value) set, or the
! dense first-order set when derivatives are written out.
MODULE lib_net_module
  use iso_c_binding
  use dnnf90_capi
  use global_variables, only: kalman_gate, kalman_q, kalman_decoupled, sys_wcomp, Nlayer, ndim, ndim_max, &
       Activation_type, iswitch_out_deriv, weight_gen, weight_best_gen
  use multi_index_bell_module, only: hod_tables_ready, HOD_D0, HOD_KMAX, NUM_alpha, alpha_list
  implicit none
  private
  public :: lnet_forward_hod_multi, lnet_seed_grad_multi
  public :: lnet_value_grad_multi, lnet_seed_row_multi
  public :: lnet_kalman_resid_multi
  public :: lnet_kalman_iekf_begin, lnet_kalman_iekf_iter
  public :: lnet_sync_weights, lnet_sync_best, lnet_free
  public :: lnet_forward_value, lnet_forward_hod, lnet_nalpha
  public :: lnet_seed_grad, lnet_seed_row, lnet_value_row, lnet_value_grad
  public :: lnet_batch_grad
  public :: lnet_kalman_init, lnet_kalman_slot, lnet_kalman_resid
  public :: lnet_kalman_active, lnet_export_weights, lnet_nweights
  public :: lnet_kalman_ngated

  integer,save :: tid_l = 0, nid_l = 0, bid_l = 0, gid_l = 0   ! table set, network, one-point work space, accumulator
  integer,save :: bwb_l = 0, nb_l = 0                        ! value-path work space and its point count
  integer,save :: kid_l = 0                                  ! filter
  integer,save :: na_l = 0, nout_l = 0, nw_l = 0
  logical,save :: net_ready = .false., kf_ready = .false.
  integer,save :: synced_gen = -1
  integer,save :: synced_src = 0   ! 1 = weight, 2 = weight_best
  real(8),allocatable,save :: flat_l(:), seedp_l(:), tbuf_l(:)

CONTAINS

  SUBROUTINE lnet_init
    integer(c_int) :: rc, a(64)
    integer(c_int),allocatable :: seeds(:)
    integer :: d0, ia, iact
    d0 = ndim(1)
    if ( hod_tables_ready ) then
       ! the trainer's carried set as closure seeds: the closure of a
       ! downward-closed set is the set itself, in the same graded order
       allocate( seeds(d0*NUM_alpha) )
       do ia = 1, NUM_alpha
          seeds((ia-1)*d0+1:ia*d0) = alpha_list(1:d0,ia)
       end do
       rc = dnnf90_tabset_closure( int(d0,c_int), int(HOD_KMAX,c_int), int(NUM_alpha,c_int), seeds )
       call capi_check( rc, "lnet_init" )
       if ( dnnf90_tabset_nderiv( rc ) /= NUM_alpha ) then
          write(*,*) "lnet_init: the library's carried set differs from the trainer's"
          stop
       end if
       do ia = 1, NUM_alpha
          call capi_check( dnnf90_alpha( int(ia,c_int), a ), "lnet_init" )
          if ( any( a(1:d0) /= alpha_list(1:d0,ia) ) ) then
             write(*,*) "lnet_init: the library's slot order differs from the trainer's at slot", ia
             stop
          end if
       end do
    else if ( iswitch_out_deriv /= 0 ) then
       rc = dnnf90_tabset_dense( int(d0,c_int), 1_c_int )
    else
       allocate( seeds(d0) )
       seeds = 0
       rc = dnnf90_tabset_closure( int(d0,c_int), 1_c_int, 1_c_int, seeds )
    end if
    call capi_check( rc, "lnet_init" )
    tid_l = rc
    select case ( trim(Activation_type) )
    case ( "SIN" );     iact = 1
    case ( "ERF" );     iact = 2
    case ( "BESSEL" );  iact = 3
    case ( "BESSEL1" ); iact = 4
    case default;       iact = 0
    end select
    call capi_check( dnnf90_tabset_activation( int(tid_l,c_int), int(iact,c_int) ), "lnet_init" )
    rc = dnnf90_net_create_ts( int(tid_l,c_int), int(Nlayer,c_int), int(ndim(1:Nlayer),c_int) )
    call capi_check( rc, "lnet_init" )
    nid_l = rc
    rc = dnnf90_bwork_create( int(nid_l,c_int), 1_c_int, 0_c_int )
    call capi_check( rc, "lnet_init" )
    bid_l = rc
    rc = dnnf90_grad_create( int(nid_l,c_int) )
    call capi_check( rc, "lnet_init" )
    gid_l = rc
    na_l = dnnf90_tabset_nderiv( int(tid_l,c_int) )
    nout_l = ndim(Nlayer)
    nw_l = dnnf90_net_nweights( int(nid_l,c_int) )
    allocate( flat_l(nw_l), seedp_l(nout_l*na_l), tbuf_l(nout_l*na_l) )
    net_ready = .true.
  END SUBROUTINE lnet_init

  ! the live slice of a (Nlayer, ndim_max, 0:ndim_max) cube and the flat
  ! ABI vector (layer blocks, rows of ndim(l-1)+1, bias first)
  SUBROUTINE to_flat( c, v )
    real(8),intent(IN) :: c(:,:,0:)
    real(8),intent(OUT) :: v(:)
    integer :: l, j, p
    p = 0
    do l = 2, Nlayer
       do j = 1, ndim(l)
          v(p+1:p+ndim(l-1)+1) = c(l,j,0:ndim(l-1))
          p = p + ndim(l-1) + 1
       end do
    end do
  END SUBROUTINE to_flat

  SUBROUTINE add_flat( v, c )
    real(8),intent(IN) :: v(:)
    real(8),intent(INOUT) :: c(:,:,0:)
    integer :: l, j, p
    p = 0
    do l = 2, Nlayer
       do j = 1, ndim(l)
          c(l,j,0:ndim(l-1)) = c(l,j,0:ndim(l-1)) + v(p+1:p+ndim(l-1)+1)
          p = p + ndim(l-1) + 1
       end do
    end do
  END SUBROUTINE add_flat

  SUBROUTINE put_flat( v, c )
    real(8),intent(IN) :: v(:)
    real(8),intent(OUT) :: c(:,:,0:)
    integer :: l, j, p
    c = 0.d0
    p = 0
    do l = 2, Nlayer
       do j = 1, ndim(l)
          c(l,j,0:ndim(l-1)) = v(p+1:p+ndim(l-1)+1)
          p = p + ndim(l-1) + 1
       end do
    end do
  END SUBROUTINE put_flat

  SUBROUTINE lnet_sync_weights( w )
    real(8),intent(IN) :: w(:,:,0:)
    if ( .not. net_ready ) call lnet_init
    call to_flat( w, flat_l )
    call capi_check( dnnf90_net_set_weights( int(nid_l,c_int), flat_l ), "lnet_sync_weights" )
    synced_gen = weight_gen
    synced_src = 1
  END SUBROUTINE lnet_sync_weights

  SUBROUTINE lnet_sync_best( w )
    real(8),intent(IN) :: w(:,:,0:)
    if ( .not. net_ready ) call lnet_init
    call to_flat( w, flat_l )
    call capi_check( dnnf90_net_set_weights( int(nid_l,c_int), flat_l ), "lnet_sync_best" )
    synced_gen = weight_best_gen
    synced_src = 2
  END SUBROUTINE lnet_sync_best

  !> The network must hold the trainer's weights as they are now
  !! (app/lib_net_module.f90:149-175).
  SUBROUTINE fresh_check( cname )
    character(*),intent(IN) :: cname
    if ( synced_src == 0 ) then
       write(*,*) cname, ": the library network has never been given"
       write(*,*) "  any weights.  Call lnet_sync_weights first."
       stop
    end if
    if ( ( synced_src == 1 .and. synced_gen /= weight_gen ) .or. &
         ( synced_src == 2 .and. synced_gen /= weight_best_gen ) ) then
       write(*,*) cname, ": the weights have been written since the"
       write(*,*) "  library network was last synchronized (generation", synced_gen, ")."
       write(*,*) "  Call lnet_sync_weights with the array you intend to"
       write(*,*) "  evaluate before asking for a value or a gradient."
       stop
    end if
  END SUBROUTINE fresh_check

  SUBROUTINE seed_check( n, cname )
    integer,intent(IN) :: n
    character(*),intent(IN) :: cname
    if ( n /= na_l ) then
       write(*,*) cname, ": seed length", n, " but the tables carry", na_l
       stop
    end if
  END SUBROUTINE seed_check

  ! forward of one point on the one-point work space: tbuf_l = T [nout][na]
  SUBROUTINE forward( x, cname )
    real(8),intent(IN) :: x(:)
    character(*),intent(IN) :: cname
    call capi_check( dnnf90_forward_batch( int(nid_l,c_int), int(bid_l,c_int), 1_c_int, x, tbuf_l, 0_c_int ), cname )
  END SUBROUTINE forward

  ! adjoint of the seed plane seedp_l [nout][na] after forward(): the
  ! point's weight gradient in flat_l
  SUBROUTINE adjoint( cname )
    character(*),intent(IN) :: cname
    call capi_check( dnnf90_grad_zero( int(gid_l,c_int) ), cname )
    call capi_check( dnnf90_backward_batch( int(nid_l,c_int), int(bid_l,c_int), seedp_l, int(gid_l,c_int), 0_c_int ), &
                     cname )
    call capi_check( dnnf90_grad_get( int(gid_l,c_int), flat_l ), cname )
  END SUBROUTINE adjoint

  ! seed matrix sm(nout, na) as the seed plane
  SUBROUTINE seed_from_matrix( sm )
    real(8),intent(IN) :: sm(:,:)
    integer :: k
    do k = 1, nout_l
       seedp_l((k-1)*na_l+1:k*na_l) = sm(k,1:na_l)
    end do
  END SUBROUTINE seed_from_matrix

  SUBROUTINE lnet_forward_value( x, val )
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: val
    if ( .not. net_ready ) call lnet_init
    call fresh_check( "lnet_forward_value" )
    call forward( x, "lnet_forward_value" )
    val = tbuf_l(1)
  END SUBROUTINE lnet_forward_value

  SUBROUTINE lnet_forward_hod( x, t )
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: t(:)
    if ( .not. net_ready ) call lnet_init
    call fresh_check( "lnet_forward_hod" )
    if ( size(t) < na_l ) then
       write(*,*) "lnet_forward_hod: the buffer holds", size(t), " values but the tables carry", na_l
       stop
    end if
    call forward( x, "lnet_forward_hod" )
    t(1:na_l) = tbuf_l(1:na_l)
  END SUBROUTINE lnet_forward_hod

  SUBROUTINE lnet_forward_hod_multi( x, tm )
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: tm(:,:)
    integer :: k
    if ( .not. net_ready ) call lnet_init
    call fresh_check( "lnet_forward_hod_multi" )
    if ( size(tm,1) < nout_l .or. size(tm,2) < na_l ) then
       write(*,*) "lnet_forward_hod_multi: the buffer is", size(tm,1), " by", size(tm,2), " but", nout_l, &
                  " by", na_l, " is needed"
       stop
    end if
    call forward( x, "lnet_forward_hod_multi" )
    do k = 1, nout_l
       tm(k,1:na_l) = tbuf_l((k-1)*na_l+1:k*na_l)
    end do
  END SUBROUTINE lnet_forward_hod_multi

  INTEGER FUNCTION lnet_nalpha()
    if ( .not. net_ready ) call lnet_init
    lnet_nalpha = na_l
  END FUNCTION lnet_nalpha

  !> Adjoint of seed = dL/dT (output 1) after lnet_forward_hod of the same
  !! point, added to the trainer's gradient cube.
  SUBROUTINE lnet_seed_grad( seed, nabla )
    real(8),intent(IN) :: seed(:)
    real(8),intent(INOUT) :: nabla(:,:,0:)
    call fresh_check( "lnet_seed_grad" )
    call seed_check( size(seed), "lnet_seed_grad" )
    seedp_l = 0.d0
    seedp_l(1:na_l) = seed(1:na_l)
    call adjoint( "lnet_seed_grad" )
    call add_flat( flat_l, nabla )
  END SUBROUTINE lnet_seed_grad

  SUBROUTINE lnet_seed_grad_multi( sm, nabla )
    real(8),intent(IN) :: sm(:,:)
    real(8),intent(INOUT) :: nabla(:,:,0:)
    if ( .not. net_ready ) call lnet_init
    call fresh_check( "lnet_seed_grad_multi" )
    call seed_from_matrix( sm )
    call adjoint( "lnet_seed_grad_multi" )
    call add_flat( flat_l, nabla )
  END SUBROUTINE lnet_seed_grad_multi

  !> The adjoint of a seed written (not added) as a Gauss-Newton row.
  SUBROUTINE lnet_seed_row( seed, row )
    real(8),intent(IN) :: seed(:)
    real(8),intent(OUT) :: row(:,:,0:)
    call fresh_check( "lnet_seed_row" )
    call seed_check( size(seed), "lnet_seed_row" )
    seedp_l = 0.d0
    seedp_l(1:na_l) = seed(1:na_l)
    call adjoint( "lnet_seed_row" )
    call put_flat( flat_l, row )
  END SUBROUTINE lnet_seed_row

  SUBROUTINE lnet_seed_row_multi( sm, row )
    real(8),intent(IN) :: sm(:,:)
    real(8),intent(OUT) :: row(:,:,0:)
    if ( .not. net_ready ) call lnet_init
    call fresh_check( "lnet_seed_row_multi" )
    call seed_from_matrix( sm )
    call adjoint( "lnet_seed_row_multi" )
    call put_flat( flat_l, row )
  END SUBROUTINE lnet_seed_row_multi

  !> Row of the value output: unit seed on the value, zero elsewhere.
  SUBROUTINE lnet_value_row( row )
    real(8),intent(OUT) :: row(:,:,0:)
    call fresh_check( "lnet_value_row" )
    seedp_l = 0.d0
    seedp_l(1) = 1.d0
    call adjoint( "lnet_value_row" )
    call put_flat( flat_l, row )
  END SUBROUTINE lnet_value_row

  !> Gradient of (coef/2)(u - y)^2 at one point, added to nabla.
  SUBROUTINE lnet_value_grad( x, y, coef, nabla )
    real(8),intent(IN) :: x(:)
    real(8),intent(IN) :: y
    real(8),intent(IN) :: coef
    real(8),intent(INOUT) :: nabla(:,:,0:)
    if ( .not. net_ready ) call lnet_init
    call fresh_check( "lnet_value_grad" )
    call forward( x, "lnet_value_grad" )
    seedp_l = 0.d0
    seedp_l(1) = coef*( tbuf_l(1) - y )
    call adjoint( "lnet_value_grad" )
    call add_flat( flat_l, nabla )
  END SUBROUTINE lnet_value_grad

  !> The same per component, each weighted by sys_wcomp.
  SUBROUTINE lnet_value_grad_multi( x, y, coef, nabla )
    real(8),intent(IN) :: x(:)
    real(8),intent(IN) :: y(:)
    real(8),intent(IN) :: coef
    real(8),intent(INOUT) :: nabla(:,:,0:)
    integer :: i
    if ( .not. net_ready ) call lnet_init
    call fresh_check( "lnet_value_grad_multi" )
    if ( size(y) < nout_l ) then
       write(*,*) "lnet_value_grad_multi: ", size(y), " targets for ", nout_l, " components"
       stop
    end if
    call forward( x, "lnet_value_grad_multi" )
    seedp_l = 0.d0
    do i = 1, nout_l
       seedp_l((i-1)*na_l+1) = coef*sys_wcomp(i)*( tbuf_l((i-1)*na_l+1) - y(i) )
    end do
    call adjoint( "lnet_value_grad_multi" )
    call add_flat( flat_l, nabla )
  END SUBROUTINE lnet_value_grad_multi

  !> The value-only batch gradient of (coef/2) sum_n (u_n - y_n)^2, added
  !! to nabla (net_grad_batch on a batched GPU work space).
  SUBROUTINE lnet_batch_grad( X, Y, coef, nabla )
    real(8),intent(IN) :: X(:,:)
    real(8),intent(IN) :: Y(:)
    real(8),intent(IN) :: coef
    real(8),intent(INOUT) :: nabla(:,:,0:)
    integer(c_int) :: rc
    if ( .not. net_ready ) call lnet_init
    call fresh_check( "lnet_batch_grad" )
    if ( nb_l /= size(Y) ) then
       if ( bwb_l > 0 ) call capi_check( dnnf90_bwork_free( int(bwb_l,c_int) ), "lnet_batch_grad" )
       rc = dnnf90_bwork_create( int(nid_l,c_int), int(max(size(Y),1),c_int), 0_c_int )
       call capi_check( rc, "lnet_batch_grad" )
       bwb_l = rc
       nb_l = size(Y)
    end if
    call capi_check( dnnf90_grad_zero( int(gid_l,c_int) ), "lnet_batch_grad" )
    call capi_check( dnnf90_grad_batch( int(nid_l,c_int), int(bwb_l,c_int), int(size(Y),c_int), X, Y, coef, &
                                        int(gid_l,c_int), 0_c_int ), "lnet_batch_grad" )
    call capi_check( dnnf90_grad_get( int(gid_l,c_int), flat_l ), "lnet_batch_grad" )
    call add_flat( flat_l, nabla )
  END SUBROUTINE lnet_batch_grad

  ! ---------------- extended Kalman filter, on the GPU ----------------

  SUBROUTINE lnet_kalman_init( p0, lam, lam0 )
    real(8),intent(IN) :: p0, lam, lam0
    real(c_double) :: info(10)
    integer(c_int) :: rc
    if ( .not. net_ready ) call lnet_init
    if ( kf_ready ) call capi_check( dnnf90_kf_free( int(kid_l,c_int) ), "lnet_kalman_init" )
    kf_ready = .false.
    if ( .not. kalman_decoupled .and. dble(nw_l)**2*8.d0 > 2.d9 ) then
       write(*,*) "Kalman: the covariance of", nw_l, " weights would need", dble(nw_l)**2*8.d-9, " GB."
       write(*,*) "  This method is for small networks updated few times"
       write(*,*) "  per pattern; use a gradient method at this size."
       stop
    end if
    rc = dnnf90_kf_create( int(nid_l,c_int), p0, lam, lam0, merge(1_c_int, 0_c_int, kalman_decoupled) )
    call capi_check( rc, "lnet_kalman_init" )
    kid_l = rc
    call capi_check( dnnf90_kf_set( int(kid_l,c_int), kalman_gate, kalman_q ), "lnet_kalman_init" )
    if ( kalman_decoupled ) then
       call capi_check( dnnf90_kf_get( int(kid_l,c_int), c_null_ptr, info ), "lnet_kalman_init" )
       write(*,'(a,i0,a)') "### Kalman: decoupled (node-wise) filter, ", nint(info(7)), " blocks"
    end if
    if ( kalman_q > 0.d0 ) write(*,'(a,es10.2)') "### Kalman process noise q = ", kalman_q
    if ( kalman_gate > 0.d0 ) write(*,'(a,f8.3)') "### Kalman innovation gate: ", kalman_gate
    kf_ready = .true.
  END SUBROUTINE lnet_kalman_init

  INTEGER FUNCTION lnet_kalman_ngated()
    real(c_double) :: info(10)
    lnet_kalman_ngated = 0
    if ( .not. kf_ready ) return
    call capi_check( dnnf90_kf_get( int(kid_l,c_int), c_null_ptr, info ), "lnet_kalman_ngated" )
    lnet_kalman_ngated = nint( info(5) )
  END FUNCTION lnet_kalman_ngated

  LOGICAL FUNCTION lnet_kalman_active()
    lnet_kalman_active = kf_ready
  END FUNCTION lnet_kalman_active

  INTEGER FUNCTION lnet_nweights()
    if ( .not. net_ready ) call lnet_init
    lnet_nweights = nw_l
  END FUNCTION lnet_nweights

  SUBROUTINE kf_needed( cname )
    character(*),intent(IN) :: cname
    if ( .not. kf_ready ) then
       write(*,*) cname, ": filter not initialized"
       stop
    end if
  END SUBROUTINE kf_needed

  !> One observation of carried slot islot against target (kf_update).
  SUBROUTINE lnet_kalman_slot( x, islot, target )
    real(8),intent(IN) :: x(:)
    integer,intent(IN) :: islot
    real(8),intent(IN) :: target
    real(8) :: tv(1)
    call fresh_check( "lnet_kalman_slot" )
    call kf_needed( "lnet_kalman_slot" )
    seedp_l = 0.d0
    seedp_l(islot) = 1.d0
    tv(1) = target
    call capi_check( dnnf90_kf_update_obs( int(nid_l,c_int), int(kid_l,c_int), int(bid_l,c_int), 1_c_int, x, &
                                           seedp_l, tv, c_null_ptr, 0_c_int ), "lnet_kalman_slot" )
  END SUBROUTINE lnet_kalman_slot

  !> One residual observation: seed = dR/dT of output 1, resid = R.
  SUBROUTINE lnet_kalman_resid( x, seed, resid )
    real(8),intent(IN) :: x(:)
    real(8),intent(IN) :: seed(:)
    real(8),intent(IN) :: resid
    real(8) :: rv(1), rn(1)
    call fresh_check( "lnet_kalman_resid" )
    call seed_check( size(seed), "lnet_kalman_resid" )
    call kf_needed( "lnet_kalman_resid" )
    seedp_l = 0.d0
    seedp_l(1:na_l) = seed(1:na_l)
    rv(1) = resid
    rn(1) = 1.d0
    call capi_check( dnnf90_kf_update_resid( int(nid_l,c_int), int(kid_l,c_int), int(bid_l,c_int), 1_c_int, x, &
                                             seedp_l, rv, rn, 0_c_int ), "lnet_kalman_resid" )
  END SUBROUTINE lnet_kalman_resid

  SUBROUTINE lnet_kalman_resid_multi( x, sm, resid, rnoise )
    real(8),intent(IN) :: x(:)
    real(8),intent(IN) :: sm(:,:)
    real(8),intent(IN) :: resid
    real(8),intent(IN),optional :: rnoise
    real(8) :: rv(1), rn(1)
    call fresh_check( "lnet_kalman_resid_multi" )
    call kf_needed( "lnet_kalman_resid_multi" )
    call seed_from_matrix( sm )
    rv(1) = resid
    rn(1) = 1.d0
    if ( present(rnoise) ) rn(1) = rnoise
    call capi_check( dnnf90_kf_update_resid( int(nid_l,c_int), int(kid_l,c_int), int(bid_l,c_int), 1_c_int, x, &
                                             seedp_l, rv, rn, 0_c_int ), "lnet_kalman_resid_multi" )
  END SUBROUTINE lnet_kalman_resid_multi

  SUBROUTINE lnet_kalman_iekf_begin()
    if ( .not. kf_ready ) then
       write(*,*) "lnet_kalman_iekf_begin: filter not initialized"
       stop
    end if
    call capi_check( dnnf90_kf_iekf_begin( int(nid_l,c_int), int(kid_l,c_int) ), "lnet_kalman_iekf_begin" )
  END SUBROUTINE lnet_kalman_iekf_begin

  SUBROUTINE lnet_kalman_iekf_iter( x, sm, resid, rnoise, final )
    real(8),intent(IN) :: x(:)
    real(8),intent(IN) :: sm(:,:)
    real(8),intent(IN) :: resid
    real(8),intent(IN) :: rnoise
    logical,intent(IN) :: final
    call fresh_check( "lnet_kalman_iekf_iter" )
    call kf_needed( "lnet_kalman_iekf_iter" )
    call seed_from_matrix( sm )
    call capi_check( dnnf90_kf_iekf_iter( int(nid_l,c_int), int(kid_l,c_int), int(bid_l,c_int), x, seedp_l, &
                                          resid, rnoise, merge(1_c_int, 0_c_int, final), 0_c_int ), &
                     "lnet_kalman_iekf_iter" )
  END SUBROUTINE lnet_kalman_iekf_iter

  !> The library network's weights into the trainer's array.
  SUBROUTINE lnet_export_weights( w )
    real(8),intent(OUT) :: w(:,:,0:)
    integer :: l, j, p
    call capi_check( dnnf90_net_get_weights( int(nid_l,c_int), flat_l ), "lnet_export_weights" )
    p = 0
    do l = 2, Nlayer
       do j = 1, ndim(l)
          w(l,j,0:ndim(l-1)) = flat_l(p+1:p+ndim(l-1)+1)
          p = p + ndim(l-1) + 1
       end do
    end do
  END SUBROUTINE lnet_export_weights

  SUBROUTINE lnet_free
    if ( .not. net_ready ) return
    if ( kf_ready ) call capi_check( dnnf90_kf_free( int(kid_l,c_int) ), "lnet_free" )
    kf_ready = .false.
    if ( bwb_l > 0 ) call capi_check( dnnf90_bwork_free( int(bwb_l,c_int) ), "lnet_free" )
    call capi_check( dnnf90_grad_free( int(gid_l,c_int) ), "lnet_free" )
    call capi_check( dnnf90_bwork_free( int(bid_l,c_int) ), "lnet_free" )
    call capi_check( dnnf90_net_free( int(nid_l,c_int) ), "lnet_free" )
    call capi_check( dnnf90_tabset_free( int(tid_l,c_int) ), "lnet_free" )
    bwb_l = 0;  nb_l = 0
    deallocate( flat_l, seedp_l, tbuf_l )
    synced_gen = -1
    synced_src = 0
    net_ready = .false.
  END SUBROUTINE lnet_free

END MODULE lib_net_module
