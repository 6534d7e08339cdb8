! net_module.f90 -- the reference's network module (lib/net_module.f90:48-60,
! 92-408) over the MI355X library's C ABI.
!
! net_t keeps the reference's public components: nlayer, ndmax, ndim(:) and
! the weight cube w(nlayer, ndmax, 0:ndmax) (bias at i = 0), which a host
! reads and writes directly (finite-difference checks perturb w(l,j,i) one
! entry at a time).  The network itself lives on the GPU; before every
! evaluation the live slice of w is compared with what was last copied
! there and copied again when it changed.  Evaluations run on the GPU; the
! routines keep the reference's names, argument order and stops.
MODULE net_module
  use iso_c_binding
  use dnnf90_capi
  use multi_index_bell_module, only: tabset_t, tabset_from_current, tabset_free

  implicit none

  PRIVATE
  PUBLIC :: net_t, work_t
  PUBLIC :: net_init, net_load, net_save, net_free
  PUBLIC :: work_init, work_free
  PUBLIC :: net_eval_hod, net_eval_hod_multi
  ! for train_module
  PUBLIC :: net_sync, net_pull, net_nweights, cube_to_flat, flat_to_cube

  TYPE :: net_t
     integer :: nlayer = 0
     integer :: ndmax  = 0
     integer,allocatable :: ndim(:)      ! (nlayer)
     real(8),allocatable :: w(:,:,:)     ! (nlayer, ndmax, 0:ndmax)
     type(tabset_t) :: tab               ! this network's tables (owns no handle)
     integer :: nid = 0                  ! the library's network handle
  END TYPE net_t

  !> Evaluation work space: one per thread (lib/net_module.f90:73-87).
  TYPE :: work_t
     integer :: wid = 0     ! per-point evaluation (dnnf90_eval)
     integer :: bid = 0     ! one-point batched work space (every output)
  END TYPE work_t

  ! the weights last copied to each network on the GPU, by handle
  TYPE :: wcopy_t
     real(8),allocatable :: w(:)
  END TYPE wcopy_t
  type(wcopy_t),save :: wlast(64)

CONTAINS

  PURE INTEGER FUNCTION net_nweights( nt )
    type(net_t),intent(IN) :: nt
    integer :: l
    net_nweights = 0
    do l = 2, nt%nlayer
       net_nweights = net_nweights + nt%ndim(l)*(nt%ndim(l-1)+1)
    end do
  END FUNCTION net_nweights

  !> The live slice of a (nlayer, ndmax, 0:ndmax) cube as the ABI's flat
  !! vector: for l = 2..nlayer, ndim(l) rows of ndim(l-1)+1, bias first.
  SUBROUTINE cube_to_flat( nlayer, ndim, c, v )
    integer,intent(IN) :: nlayer, ndim(:)
    real(8),intent(IN) :: c(:,:,0:)
    real(8),intent(OUT) :: v(:)
    integer :: l, j, p
    p = 0
    do l = 2, nlayer
       do j = 1, ndim(l)
          v(p+1:p+ndim(l-1)+1) = c(l,j,0:ndim(l-1))
          p = p + ndim(l-1) + 1
       end do
    end do
  END SUBROUTINE cube_to_flat

  SUBROUTINE flat_to_cube( nlayer, ndim, v, c )
    integer,intent(IN) :: nlayer, ndim(:)
    real(8),intent(IN) :: v(:)
    real(8),intent(INOUT) :: c(:,:,0:)
    integer :: l, j, p
    p = 0
    do l = 2, nlayer
       do j = 1, ndim(l)
          c(l,j,0:ndim(l-1)) = v(p+1:p+ndim(l-1)+1)
          p = p + ndim(l-1) + 1
       end do
    end do
  END SUBROUTINE flat_to_cube

  !> Copy w to the GPU when it differs from what was copied last.
  SUBROUTINE net_sync( nt )
    type(net_t),intent(IN) :: nt
    real(8) :: v(net_nweights(nt))
    call cube_to_flat( nt%nlayer, nt%ndim, nt%w, v )
    if ( allocated(wlast(nt%nid)%w) ) then
       if ( size(wlast(nt%nid)%w) == size(v) ) then
          if ( all( wlast(nt%nid)%w == v ) ) return
       end if
       deallocate( wlast(nt%nid)%w )
    end if
    call capi_check( dnnf90_net_set_weights( int(nt%nid,c_int), v ), "net_sync" )
    allocate( wlast(nt%nid)%w(size(v)) )
    wlast(nt%nid)%w = v
  END SUBROUTINE net_sync

  !> Read w back after the GPU changed the weights (the Adam step).
  SUBROUTINE net_pull( nt )
    type(net_t),intent(INOUT) :: nt
    real(8) :: v(net_nweights(nt))
    call capi_check( dnnf90_net_get_weights( int(nt%nid,c_int), v ), "net_pull" )
    call flat_to_cube( nt%nlayer, nt%ndim, v, nt%w )
    if ( allocated(wlast(nt%nid)%w) ) deallocate( wlast(nt%nid)%w )
    allocate( wlast(nt%nid)%w(size(v)) )
    wlast(nt%nid)%w = v
  END SUBROUTINE net_pull

  !> Build an empty network of the given shape (lib/net_module.f90:92-143),
  !! on the table set ts or on the current tables.
  SUBROUTINE net_init( nt, nlayer, ndim, ts )
    type(net_t),intent(OUT) :: nt
    integer,intent(IN) :: nlayer, ndim(nlayer)
    type(tabset_t),intent(IN),optional :: ts
    type(tabset_t) :: cur
    integer(c_int) :: rc
    if ( present(ts) ) then
       nt%tab = ts
    else
       call tabset_from_current( cur )
       nt%tab = cur
    end if
    nt%tab%tid = 0
    if ( nlayer < 2 ) then
       write(*,*) "net_init: nlayer must be at least 2, got", nlayer
       stop 1
    end if
    if ( ndim(1) /= nt%tab%d0 ) then
       write(*,*) "net_init: input width", ndim(1), " does not match the table set d0 =", nt%tab%d0
       write(*,*) "  build the tables with the same number of input variables as the network takes"
       stop 1
    end if
    if ( ndim(nlayer) < 1 ) then
       write(*,*) "net_init: output width", ndim(nlayer), " must be >= 1"
       stop 1
    end if
    if ( present(ts) ) then
       if ( ts%tid <= 0 ) then
          write(*,*) "net_init: the table set was not built by tabset_init or tabset_from_current"
          stop 1
       end if
       call capi_check( dnnf90_tabset_activation( int(ts%tid,c_int), int(ts%iact,c_int) ), "net_init" )
       rc = dnnf90_net_create_ts( int(ts%tid,c_int), int(nlayer,c_int), int(ndim,c_int) )
    else
       call capi_check( dnnf90_tabset_activation( int(cur%tid,c_int), int(nt%tab%iact,c_int) ), "net_init" )
       rc = dnnf90_net_create_ts( int(cur%tid,c_int), int(nlayer,c_int), int(ndim,c_int) )
       call tabset_free( cur )
    end if
    call capi_check( rc, "net_init" )
    nt%nid = rc
    nt%nlayer = nlayer
    nt%ndmax = maxval( ndim(1:nlayer) )
    allocate( nt%ndim(nlayer) )
    nt%ndim = ndim
    allocate( nt%w(nlayer,nt%ndmax,0:nt%ndmax) )
    nt%w = 0.d0
    if ( allocated(wlast(nt%nid)%w) ) deallocate( wlast(nt%nid)%w )
  END SUBROUTINE net_init

  !> Read a network in nn_weight.dat format (lib/net_module.f90:146-215);
  !! a missing file, an unknown activation, a nonlinear output layer or an
  !! activation other than the table set's stop with the library's message.
  SUBROUTINE net_load( nt, filename, ts )
    type(net_t),intent(OUT) :: nt
    character(len=*),intent(IN) :: filename
    type(tabset_t),intent(IN),optional :: ts
    type(tabset_t) :: cur
    integer(c_int) :: rc, nl, d(64)
    if ( present(ts) ) then
       call capi_check( dnnf90_tabset_activation( int(ts%tid,c_int), int(ts%iact,c_int) ), "net_load" )
       rc = dnnf90_net_load_ts( int(ts%tid,c_int), c_string(filename) )
       nt%tab = ts
    else
       call tabset_from_current( cur )
       rc = dnnf90_net_load_ts( int(cur%tid,c_int), c_string(filename) )
       nt%tab = cur
       call tabset_free( cur )
    end if
    nt%tab%tid = 0
    call capi_check( rc, "net_load" )
    nt%nid = rc
    nl = dnnf90_net_nlayer( rc )
    call capi_check( nl, "net_load" )
    call capi_check( dnnf90_net_dims( rc, d ), "net_load" )
    nt%nlayer = nl
    allocate( nt%ndim(nl) )
    nt%ndim = d(1:nl)
    nt%ndmax = maxval( nt%ndim )
    allocate( nt%w(nl,nt%ndmax,0:nt%ndmax) )
    nt%w = 0.d0
    if ( allocated(wlast(nt%nid)%w) ) deallocate( wlast(nt%nid)%w )
    call net_pull( nt )
  END SUBROUTINE net_load

  !> Write nn_weight.dat (lib/net_module.f90:219-243).
  SUBROUTINE net_save( nt, filename )
    type(net_t),intent(IN) :: nt
    character(len=*),intent(IN) :: filename
    call net_sync( nt )
    call capi_check( dnnf90_net_save( int(nt%nid,c_int), c_string(filename) ), "net_save" )
  END SUBROUTINE net_save

  SUBROUTINE net_free( nt )
    type(net_t),intent(INOUT) :: nt
    if ( nt%nid > 0 ) then
       call capi_check( dnnf90_net_free( int(nt%nid,c_int) ), "net_free" )
       if ( allocated(wlast(nt%nid)%w) ) deallocate( wlast(nt%nid)%w )
    end if
    nt%nid = 0
    if ( allocated(nt%ndim) ) deallocate( nt%ndim )
    if ( allocated(nt%w) )    deallocate( nt%w )
    call tabset_free( nt%tab )
    nt%nlayer = 0
    nt%ndmax  = 0
  END SUBROUTINE net_free

  SUBROUTINE work_init( wk, nt )
    type(work_t),intent(OUT) :: wk
    type(net_t),intent(IN) :: nt
    integer(c_int) :: rc
    rc = dnnf90_work_create( int(nt%nid,c_int) )
    call capi_check( rc, "work_init" )
    wk%wid = rc
    rc = dnnf90_bwork_create( int(nt%nid,c_int), 1_c_int, 0_c_int )
    call capi_check( rc, "work_init" )
    wk%bid = rc
  END SUBROUTINE work_init

  SUBROUTINE work_free( wk )
    type(work_t),intent(INOUT) :: wk
    if ( wk%wid > 0 ) call capi_check( dnnf90_work_free( int(wk%wid,c_int) ), "work_free" )
    if ( wk%bid > 0 ) call capi_check( dnnf90_bwork_free( int(wk%bid,c_int) ), "work_free" )
    wk%wid = 0
    wk%bid = 0
  END SUBROUTINE work_free

  !> The value and every carried derivative of output 1
  !! (lib/net_module.f90:295-375): t(1:na) in slot order.
  SUBROUTINE net_eval_hod( nt, wk, x, t )
    type(net_t),intent(IN) :: nt
    type(work_t),intent(INOUT) :: wk
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: t(:)
    real(c_double) :: tv(nt%tab%na)
    call net_sync( nt )
    call capi_check( dnnf90_eval( int(nt%nid,c_int), int(wk%wid,c_int), x, tv ), "net_eval_hod" )
    t(1:nt%tab%na) = tv
  END SUBROUTINE net_eval_hod

  !> Every output's carried derivatives, tm(1:nout, 1:na)
  !! (lib/net_module.f90:389-408).
  SUBROUTINE net_eval_hod_multi( nt, wk, x, tm )
    type(net_t),intent(IN) :: nt
    type(work_t),intent(INOUT) :: wk
    real(8),intent(IN) :: x(:)
    real(8),intent(OUT) :: tm(:,:)
    integer :: nout, na, k
    real(c_double) :: flat(nt%ndim(nt%nlayer)*nt%tab%na)
    nout = nt%ndim(nt%nlayer)
    na = nt%tab%na
    call net_sync( nt )
    call capi_check( dnnf90_forward_batch( int(nt%nid,c_int), int(wk%bid,c_int), 1_c_int, x, flat, 0_c_int ), &
                     "net_eval_hod_multi" )
    do k = 1, nout
       tm(k,1:na) = flat((k-1)*na+1:k*na)
    end do
  END SUBROUTINE net_eval_hod_multi

END MODULE net_module
