! multi_index_bell_module.f90 -- the reference's table-set module
! (lib/multi_index_bell_module.f90:40-78) over the MI355X library's C ABI.
!
! The tables themselves are built by the library (the same lists in the same
! order); this module keeps the reference's Fortran surface: tabset_t with
! its public arrays, tabset_init / tabset_from_current / tabset_free, and the
! module-level "current" family of init_hod_tables (NUM_alpha, alpha_list,
! alpha_deg, ind_e1 and the pair and forward lists), 1-based as there.
MODULE multi_index_bell_module
  use iso_c_binding
  use dnnf90_capi
  implicit none

  PRIVATE
  PUBLIC :: init_hod_tables, write_alpha_order, alpha_index, hod_tables_ready
  PUBLIC :: tabset_t, tabset_init, tabset_from_current, tabset_free
  PUBLIC :: HOD_D0, HOD_KMAX, NUM_alpha
  PUBLIC :: alpha_list, alpha_deg, ind_e1
  PUBLIC :: pair_start, pair_num, pair_ia, pair_id, pair_binom
  PUBLIC :: fq_start, fq_num, fq_ib, fq_id, fq_c

  !> One table family (cf. lib/multi_index_bell_module.f90:56-78).  tid is
  !! the library's table-set handle; a copy held by a network owns none
  !! (tid = 0), so freeing it releases only the arrays.
  TYPE :: tabset_t
     integer :: d0 = 0, kmax = 0, na = 0
     integer,allocatable :: alpha_list(:,:)   ! (d0, na)
     integer,allocatable :: alpha_deg(:)      ! (na)
     integer,allocatable :: ind_e1(:)         ! (d0), 0 when e_i is not carried
     integer,allocatable :: pair_start(:), pair_num(:), pair_ia(:), pair_id(:)
     real(8),allocatable :: pair_binom(:)
     integer,allocatable :: fq_start(:,:), fq_num(:,:)   ! (na, 2:kmax)
     integer,allocatable :: fq_ib(:), fq_id(:)
     real(8),allocatable :: fq_c(:)
     !> 0 tanh, 1 sin, 2 erf, 3 J0, 4 J1 (5: erf-GELU, this library only)
     integer :: iact = 0
     integer :: tid = 0
  END TYPE tabset_t

  logical :: hod_tables_ready = .false.
  integer :: HOD_D0 = 0, HOD_KMAX = 0, NUM_alpha = 0
  integer,allocatable :: alpha_list(:,:), alpha_deg(:), ind_e1(:)
  integer,allocatable :: pair_start(:), pair_num(:), pair_ia(:), pair_id(:)
  real(8),allocatable :: pair_binom(:)
  integer,allocatable :: fq_start(:,:), fq_num(:,:), fq_ib(:), fq_id(:)
  real(8),allocatable :: fq_c(:)

  ! the current family's arguments, to build handles of it
  integer :: cur_nseed = 0
  integer,allocatable :: cur_seeds(:)

CONTAINS

  !> Build the current family (lib/multi_index_bell_module.f90:128-351).
  SUBROUTINE init_hod_tables( D0, K, nseed, seed_alpha )
    integer,intent(IN) :: D0, K, nseed
    integer,intent(IN) :: seed_alpha(D0,*)
    type(tabset_t) :: t
    integer :: s
    if ( allocated(cur_seeds) ) deallocate( cur_seeds )
    allocate( cur_seeds(max(D0*nseed,1)) )
    cur_seeds = 0
    do s = 1, nseed
       cur_seeds((s-1)*D0+1:s*D0) = seed_alpha(1:D0,s)
    end do
    cur_nseed = nseed
    if ( nseed > 0 ) then
       call capi_check( dnnf90_tables_init_closure( int(D0,c_int), int(K,c_int), int(nseed,c_int), &
                                                    int(cur_seeds,c_int) ), "init_hod_tables" )
    else
       call capi_check( dnnf90_tables_init_dense( int(D0,c_int), int(K,c_int) ), "init_hod_tables" )
    end if
    HOD_D0 = D0
    HOD_KMAX = K
    ! the module-level lists, from a handle of the family
    call tabset_from_current( t )
    NUM_alpha = t%na
    call move_alloc( t%alpha_list, alpha_list )
    call move_alloc( t%alpha_deg, alpha_deg )
    call move_alloc( t%ind_e1, ind_e1 )
    call move_alloc( t%pair_start, pair_start )
    call move_alloc( t%pair_num, pair_num )
    call move_alloc( t%pair_ia, pair_ia )
    call move_alloc( t%pair_id, pair_id )
    call move_alloc( t%pair_binom, pair_binom )
    call move_alloc( t%fq_start, fq_start )
    call move_alloc( t%fq_num, fq_num )
    call move_alloc( t%fq_ib, fq_ib )
    call move_alloc( t%fq_id, fq_id )
    call move_alloc( t%fq_c, fq_c )
    call tabset_free( t )
    hod_tables_ready = .true.
  END SUBROUTINE init_hod_tables

  !> An independent table set (lib/multi_index_bell_module.f90:692-699).
  SUBROUTINE tabset_init( ts, D0, K, nseed, seed_alpha )
    type(tabset_t),intent(OUT) :: ts
    integer,intent(IN) :: D0, K, nseed
    integer,intent(IN) :: seed_alpha(D0,*)
    call init_hod_tables( D0, K, nseed, seed_alpha )
    call tabset_from_current( ts )
  END SUBROUTINE tabset_init

  !> A table set of the current family, with its own library handle.
  SUBROUTINE tabset_from_current( ts )
    type(tabset_t),intent(OUT) :: ts
    integer(c_int) :: cnt(4), rc, a(64)
    integer(c_int),allocatable :: ps(:), pn(:), pia(:), pid(:), fs(:), fn(:), fib(:), fid(:)
    real(c_double),allocatable :: pb(:), fc(:)
    integer :: ia, i, q, na, np, nf, k, d0
    if ( HOD_D0 <= 0 ) then
       write(*,*) "tabset_from_current: no current tables (call init_hod_tables first)"
       stop 1
    end if
    d0 = HOD_D0
    if ( cur_nseed > 0 ) then
       rc = dnnf90_tabset_closure( int(d0,c_int), int(HOD_KMAX,c_int), int(cur_nseed,c_int), &
                                   int(cur_seeds,c_int) )
    else
       rc = dnnf90_tabset_dense( int(d0,c_int), int(HOD_KMAX,c_int) )
    end if
    call capi_check( rc, "tabset_from_current" )
    ts%tid = rc
    call capi_check( dnnf90_tabset_counts( rc, cnt ), "tabset_from_current" )
    na = cnt(1);  np = cnt(2);  nf = cnt(3);  k = cnt(4)
    ts%d0 = d0;  ts%kmax = k;  ts%na = na;  ts%iact = 0
    allocate( ts%alpha_list(d0,na), ts%alpha_deg(na), ts%ind_e1(d0) )
    ts%ind_e1 = 0
    do ia = 1, na   ! building the handle made the family current again
       call capi_check( dnnf90_alpha( int(ia,c_int), a ), "tabset_from_current" )
       ts%alpha_list(1:d0,ia) = a(1:d0)
       ts%alpha_deg(ia) = sum( a(1:d0) )
       if ( ts%alpha_deg(ia) == 1 ) then
          do i = 1, d0
             if ( a(i) == 1 ) ts%ind_e1(i) = ia
          end do
       end if
    end do
    allocate( ps(na), pn(na), pia(max(np,1)), pid(max(np,1)), pb(max(np,1)) )
    allocate( fs(na*(k+1)), fn(na*(k+1)), fib(max(nf,1)), fid(max(nf,1)), fc(max(nf,1)) )
    call capi_check( dnnf90_tabset_lists( rc, ps, pn, pia, pid, pb, fs, fn, fib, fid, fc ), &
                     "tabset_from_current" )
    ! 0-based slots and list positions in the ABI, 1-based here
    allocate( ts%pair_start(na), ts%pair_num(na), ts%pair_ia(np), ts%pair_id(np), ts%pair_binom(np) )
    ts%pair_start = ps + 1;  ts%pair_num = pn
    ts%pair_ia = pia(1:np) + 1;  ts%pair_id = pid(1:np) + 1;  ts%pair_binom = pb(1:np)
    allocate( ts%fq_start(na,2:max(k,2)), ts%fq_num(na,2:max(k,2)) )
    ts%fq_start = 0;  ts%fq_num = 0
    do ia = 1, na
       do q = 2, k
          ts%fq_start(ia,q) = fs((ia-1)*(k+1)+q+1) + 1
          ts%fq_num(ia,q) = fn((ia-1)*(k+1)+q+1)
       end do
    end do
    allocate( ts%fq_ib(nf), ts%fq_id(nf), ts%fq_c(nf) )
    ts%fq_ib = fib(1:nf) + 1;  ts%fq_id = fid(1:nf) + 1;  ts%fq_c = fc(1:nf)
  END SUBROUTINE tabset_from_current

  SUBROUTINE tabset_free( ts )
    type(tabset_t),intent(INOUT) :: ts
    if ( ts%tid > 0 ) call capi_check( dnnf90_tabset_free( int(ts%tid,c_int) ), "tabset_free" )
    ts%tid = 0
    if ( allocated(ts%alpha_list) ) deallocate( ts%alpha_list )
    if ( allocated(ts%alpha_deg) )  deallocate( ts%alpha_deg )
    if ( allocated(ts%ind_e1) )     deallocate( ts%ind_e1 )
    if ( allocated(ts%pair_start) ) deallocate( ts%pair_start, ts%pair_num, ts%pair_ia, ts%pair_id, ts%pair_binom )
    if ( allocated(ts%fq_start) )   deallocate( ts%fq_start, ts%fq_num, ts%fq_ib, ts%fq_id, ts%fq_c )
    ts%d0 = 0;  ts%kmax = 0;  ts%na = 0
  END SUBROUTINE tabset_free

  !> hod_alpha_order.dat of the current family (lib/multi_index_bell_module.f90:452-471).
  SUBROUTINE write_alpha_order( fname )
    character(*),intent(IN) :: fname
    type(tabset_t) :: t
    call tabset_from_current( t )
    call capi_check( dnnf90_tabset_write_order( int(t%tid,c_int), c_string(fname) ), "write_alpha_order" )
    call tabset_free( t )
  END SUBROUTINE write_alpha_order

  !> Slot of a multi-index in the current family, 0 if not carried
  !! (lib/multi_index_bell_module.f90:475-483).
  INTEGER FUNCTION alpha_index( a )
    integer,intent(IN) :: a(HOD_D0)
    integer :: ia
    alpha_index = 0
    if ( any(a < 0) .or. sum(a) > HOD_KMAX ) return
    do ia = 1, NUM_alpha
       if ( all( alpha_list(1:HOD_D0,ia) == a(1:HOD_D0) ) ) then
          alpha_index = ia
          return
       end if
    end do
  END FUNCTION alpha_index

END MODULE multi_index_bell_module
