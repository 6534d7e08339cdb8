! example_fd.f90 -- the adjoint check of tools/dbg_backward_fd.f90:10-66 run
! through the Fortran binding on the GPU: D0 = 2, K = 3 dense, net [2,5,1],
! w(l,j,i) = 0.3 sin(17 l + 7 j + 3 i + 0.5), seed_i = cos(2.7 i),
! x = (0.37, -0.61), central differences with h = 1e-6.  Also checks the
! evaluation entry against the training forward (:135-140 there).
! Exit status 0 when the reference's tolerances hold.
program example_fd
  use dnnf90_gpu
  implicit none
  integer,parameter :: D0=2, K=3, NL=3
  integer :: ndim(NL), l, j, i, na, p, nw
  type(net_t) :: nt
  type(twork_t) :: tw
  type(work_t) :: wk
  type(grad_t) :: g
  real(8),allocatable :: w(:,:,:), seed(:), tout(:), tp(:), tm(:), tev(:), gv(:)
  real(8) :: x(D0), h, num, ana, emax, wsave
  integer :: ifail

  ndim = (/ D0, 5, 1 /)
  call tables_dense( D0, K )
  call net_init( nt, NL, ndim )
  na = nt%na
  allocate( w(NL,5,0:5) )
  w = 0.d0
  do l=2,NL
     do j=1,ndim(l)
        do i=0,ndim(l-1)
           w(l,j,i) = 0.3d0*sin( 17.d0*l + 7.d0*j + 3.d0*i + 0.5d0 )
        end do
     end do
     call net_set_layer( nt, l, w(l,1:ndim(l),0:ndim(l-1)) )
  end do
  call twork_init( tw, nt )
  call work_init( wk, nt )
  call grad_init( g, nt )
  allocate( seed(na), tout(na), tp(na), tm(na), tev(na), gv(nt%nw) )
  do i=1,na
     seed(i) = cos( 2.7d0*i )
  end do
  x = (/ 0.37d0, -0.61d0 /)
  ifail = 0

  call net_forward_point( nt, tw, x, tout )
  call grad_zero( g )
  call net_backward_point( nt, tw, seed, g )
  call grad_get( g, nt, gv )

  h = 1.d-6
  emax = 0.d0
  p = 0
  do l=2,NL
     do j=1,ndim(l)
        do i=0,ndim(l-1)
           p = p + 1
           wsave = w(l,j,i)
           w(l,j,i) = wsave + h
           call net_set_layer( nt, l, w(l,1:ndim(l),0:ndim(l-1)) )
           call net_forward_point( nt, tw, x, tp )
           w(l,j,i) = wsave - h
           call net_set_layer( nt, l, w(l,1:ndim(l),0:ndim(l-1)) )
           call net_forward_point( nt, tw, x, tm )
           w(l,j,i) = wsave
           call net_set_layer( nt, l, w(l,1:ndim(l),0:ndim(l-1)) )
           num = sum( seed(1:na)*(tp(1:na)-tm(1:na)) )/(2.d0*h)
           ana = gv(p)
           emax = max( emax, abs(num-ana)/max(abs(num),1.d-12) )
        end do
     end do
  end do
  write(*,'(a,e12.4)') "point adjoint vs FD          : ", emax
  if ( emax > 1.d-6 ) ifail = ifail + 1

  call net_forward_point( nt, tw, x, tout )
  call net_eval_hod( nt, wk, x, tev )
  emax = maxval( abs( tev(1:na) - tout(1:na) ) )/max( maxval(abs(tout(1:na))), 1.d-12 )
  write(*,'(a,e12.4)') "net_eval_hod vs forward      : ", emax
  if ( emax > 1.d-12 ) ifail = ifail + 1

  call grad_free( g )
  call work_free( wk )
  call twork_free( tw )
  call net_free( nt )
  if ( ifail > 0 ) then
     write(*,*) "FAILED checks:", ifail
     stop 1
  end if
  write(*,'(a)') "all checks passed"
end program example_fd
