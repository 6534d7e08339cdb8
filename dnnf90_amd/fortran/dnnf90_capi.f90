! dnnf90_capi.f90 -- iso_c_binding declarations of the library's C ABI
! (include/dnnf90.h, the reference's lib/dnnf90.h entries under their own
! names, and the include/dnnf90_amd.h entries the Fortran modules need),
! plus the error check the modules share.  The reference's Fortran module
! API (multi_index_bell_module, net_module, train_module) is built on this
! in the files of the same names.
MODULE dnnf90_capi
  use iso_c_binding
  implicit none

  PUBLIC

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
     integer(c_int) function dnnf90_tabset_free(tid) bind(C)
       import :: c_int
       integer(c_int), value :: tid
     end function
     integer(c_int) function dnnf90_tabset_counts(tid, counts) bind(C)
       import :: c_int
       integer(c_int), value :: tid
       integer(c_int), intent(out) :: counts(*)
     end function
     integer(c_int) function dnnf90_tabset_lists(tid, pair_start, pair_num, pair_ia, pair_id, pair_binom, &
                                                 fq_start, fq_num, fq_ib, fq_id, fq_c) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: tid
       integer(c_int), intent(out) :: pair_start(*), pair_num(*), pair_ia(*), pair_id(*)
       real(c_double), intent(out) :: pair_binom(*)
       integer(c_int), intent(out) :: fq_start(*), fq_num(*), fq_ib(*), fq_id(*)
       real(c_double), intent(out) :: fq_c(*)
     end function
     integer(c_int) function dnnf90_tabset_write_order(tid, filename) bind(C)
       import :: c_char, c_int
       integer(c_int), value :: tid
       character(kind=c_char), intent(in) :: filename(*)
     end function
     integer(c_int) function dnnf90_net_create(nlayer, dims) bind(C)
       import :: c_int
       integer(c_int), value :: nlayer
       integer(c_int), intent(in) :: dims(*)
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
     integer(c_int) function dnnf90_net_save(nid, filename) bind(C)
       import :: c_char, c_int
       integer(c_int), value :: nid
       character(kind=c_char), intent(in) :: filename(*)
     end function
     integer(c_int) function dnnf90_net_free(nid) bind(C)
       import :: c_int
       integer(c_int), value :: nid
     end function
     integer(c_int) function dnnf90_net_nweights(nid) bind(C)
       import :: c_int
       integer(c_int), value :: nid
     end function
     integer(c_int) function dnnf90_net_nlayer(nid) bind(C)
       import :: c_int
       integer(c_int), value :: nid
     end function
     integer(c_int) function dnnf90_net_dims(nid, dims) bind(C)
       import :: c_int
       integer(c_int), value :: nid
       integer(c_int), intent(out) :: dims(*)
     end function
     integer(c_int) function dnnf90_net_get_weights(nid, w) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid
       real(c_double), intent(out) :: w(*)
     end function
     integer(c_int) function dnnf90_net_set_weights(nid, w) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid
       real(c_double), intent(in) :: w(*)
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
     integer(c_int) function dnnf90_grad_get(gid, g) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: gid
       real(c_double), intent(out) :: g(*)
     end function
     integer(c_int) function dnnf90_grad_set(gid, g) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: gid
       real(c_double), intent(in) :: g(*)
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
     integer(c_int) function dnnf90_grad_batch(nid, bid, npt, x, y, coef, gid, flags) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, bid, npt, gid, flags
       real(c_double), intent(in) :: x(*), y(*)
       real(c_double), value :: coef
     end function
     integer(c_int) function dnnf90_kf_create(nid, p0, lam_ini, lam0, decoupled) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, decoupled
       real(c_double), value :: p0, lam_ini, lam0
     end function
     integer(c_int) function dnnf90_kf_set(kid, gate, q) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: kid
       real(c_double), value :: gate, q
     end function
     integer(c_int) function dnnf90_kf_free(kid) bind(C)
       import :: c_int
       integer(c_int), value :: kid
     end function
     integer(c_int) function dnnf90_kf_update_obs(nid, kid, bid, nobs, x, seed, target, pred, flags) bind(C)
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: nid, kid, bid, nobs, flags
       real(c_double), intent(in) :: x(*), seed(*), target(*)
       type(c_ptr), value :: pred
     end function
     integer(c_int) function dnnf90_kf_update_resid(nid, kid, bid, nobs, x, seed, resid, rnoise, flags) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, kid, bid, nobs, flags
       real(c_double), intent(in) :: x(*), seed(*), resid(*), rnoise(*)
     end function
     integer(c_int) function dnnf90_kf_iekf_begin(nid, kid) bind(C)
       import :: c_int
       integer(c_int), value :: nid, kid
     end function
     integer(c_int) function dnnf90_kf_iekf_iter(nid, kid, bid, x, seed, resid, rnoise, final, flags) bind(C)
       import :: c_int, c_double
       integer(c_int), value :: nid, kid, bid, final, flags
       real(c_double), intent(in) :: x(*), seed(*)
       real(c_double), value :: resid, rnoise
     end function
     integer(c_int) function dnnf90_kf_get(kid, P, info) bind(C)
       import :: c_int, c_ptr, c_double
       integer(c_int), value :: kid
       type(c_ptr), value :: P
       real(c_double), intent(out) :: info(*)
     end function
     type(c_ptr) function dnnf90_last_error() bind(C)
       import :: c_ptr
     end function
  end interface

CONTAINS

  !> A failed library call stops the program with the library's message, as
  !! the reference stops on a contract violation.
  SUBROUTINE capi_check( rc, where )
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
    stop 1
  END SUBROUTINE capi_check

  !> A Fortran string as a NUL-terminated C string.
  FUNCTION c_string( s ) result( cs )
    character(*),intent(IN) :: s
    character(kind=c_char) :: cs(len_trim(s)+1)
    integer :: i
    do i = 1, len_trim(s)
       cs(i) = s(i:i)
    end do
    cs(len_trim(s)+1) = c_null_char
  END FUNCTION c_string

END MODULE dnnf90_capi
