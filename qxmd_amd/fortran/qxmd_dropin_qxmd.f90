! qxmd_dropin_qxmd.f90 -- drop-ins by name for the three routines of the path
! whose inputs and outputs live in QXMD's modules:
!
!   schmidt (Sources/schmidt.f90:71)   Gram-Schmidt of cgjr_(2 npnod, nband)
!   pwpchg  (Sources/chgdns.f90:5)     density of gdcr into glocal and rhgr
!   eigen   (Sources/eigen.f90:52)     the BKL CG band update
!
! Unlike qxmd_dropin.f90 this file is compiled inside the QXMD build, after
! its modules (it reads param, pwlda_variables, pwlda_pw, ifcholesky,
! eigen_variables), and linked with libqxmd_amd_f.so; the reference's own
! bodies of schmidt / schmidt2, pwpchg / pwpchg2 and eigen are left out of
! that build.  Scope: one rank, the Gamma point, no spin, norm-conserving
! pseudopotentials (what the path covers); anything else stops with a
! message.  The device context (qxmd_gpu_init*), the potential's KB table
! (stscfn_pw / qxmd_gpu_setnlc_g) are set up as for the other drop-ins.

subroutine schmidt( nfile, myid, nodes, ct0,  &
& cgjr_, iod_, iodg_, ltimecnt, lout, leig_start, leig_end )
use param, only: nband
use pwlda_variables, only: npnod
use ifcholesky, only: lcholesky, lstop
use qxmd_amd_gpu, only: qxmd_schmidt
implicit none
integer :: nfile(*), myid, nodes
real*8  :: ct0
real*8  :: cgjr_(2*npnod, nband)
integer :: iod_(*), iodg_(*)
logical :: ltimecnt, lout, leig_start, leig_end

if( nodes /= 1 ) call qxmd_stop( 'schmidt: one rank only' )
!--- Cholesky of the overlap, then psi <- L^-1 psi; the bands are left
!--- untouched when the overlap is not positive definite (lcholesky = .false.)
call qxmd_schmidt( cgjr_, 2*npnod, nband, lcholesky )
if( .not.lcholesky ) then
    write(*,*) '*** the matrix <phi_i|phi_j> is not positive definite.'
    if( lstop ) call qxmd_stop( 'in schmidt' )
end if

return
end subroutine


subroutine pwpchg( nfile, myid, nodes, ltimecnt, lfssh_updtcg )
use param, only: nband, lspin, lgamma
use param_atom, only: lvand
use pwlda_variables, only: nplwex, nplw, nplw2, nplw2ex, kfft0, rvol
use pwlda_pw, only: gdcr, occ, glocal, rhgr, mfd2ft_c, ntotfd_c, nd1vks_c, ijkg, iods,  &
&                   fft3x, fft3y, fftwork
implicit none
integer :: nfile(*), myid, nodes
logical :: ltimecnt, lfssh_updtcg

if( nodes /= 1 .or. lspin .or. .not.lgamma .or. lvand )  &
&   call qxmd_stop( 'pwpchg: one rank, Gamma, no spin, norm-conserving only' )
!--- density of the soft part on the coarse grid (softchg), then rho(G) on
!--- the nplw2 sphere (outchg's chgr2g), as pwpchg2 does for this case
call softchg( nfile, myid, nodes,  &
& glocal, occ, gdcr, nplwex, nplw, nband, 1, nband, nband,  &
& iods, rvol, mfd2ft_c, ntotfd_c, nd1vks_c, kfft0, ijkg, nplw2,  &
& fft3x, fft3y, fftwork )
rhgr(1:nplw2ex) = 0.d0
call chgr2g( nfile, myid, nodes,  &
& glocal, rhgr, nplw2ex, nplw2, mfd2ft_c, ntotfd_c, nd1vks_c, kfft0,  &
& fft3x, fft3y, fftwork, ijkg, nplw2 )

return
end subroutine


subroutine eigen( nfile, myid, nodes,  &
& eig, gdcr, cgjr, rhcr, hcsr, nplwex, nplw, nplw7ex,  &
& dkgnrm, prcd, ekib,  &
& gnk, hnk, sck, nband, nbnod1, nbnod2, nbnod, iod,  &
& glocal, mfd2ft, ntotfd, nd1vks, kfft0, hfv, hfv_c,  &
& fft3x, fft3y, fftwork, ijkgd, ngenh,  &
& eigr, eigi, apk, apki, scwr, scwi, rdelv, rvol,  &
& ntype, nhk1, nhk2, natom, iatoit, lvand,  &
& lkbpp_r, lvand_r, lkbpp_g, lvand_g, lkbppi, lvandi, lking,  &
& prod, prodr, w1r, w1i, iter, istdot, tolera, threinn, wdinner, lortho,  &
& bzan, nriter, ihyoji, itermx, noband, iteremx,  &
& lhcunt, ltimecnt, leig_start, leig_end,  &
& lmax, mxl, lchk, lclno, nylmmx, ycos, ysin, nplwcs, method,  &
& hfrhcr, jhybrid, lefield_islts )
use, intrinsic :: iso_c_binding
use eigen_variables, only: roughfeig, roughfzan
use qxmd_amd_c
use qxmd_amd_gpu, only: qxctx, qxmd_gpu_check, qxmd_gpu_need_ctx, qxmd_gpu_set_glocal
implicit none
integer :: nfile(*), myid, nodes
integer :: nplwex, nplw, nplw7ex
real*8  :: eig(*), dkgnrm(*)
real*8  :: gdcr(nplwex,*), rhcr(nplwex,*), hcsr(*)
real*8  :: cgjr(nplwex,*)
real*8  :: prcd(*), ekib(*), gnk(*), hnk(*), sck(*)
integer :: nband, nbnod1, nbnod2, nbnod
integer :: iod(*)
real*8  :: glocal(*), hfv(*), hfv_c(*)
integer :: mfd2ft(*), nd1vks(*), ntotfd, kfft0
real*8  :: fft3x(*), fft3y(*)
complex*16 :: fftwork(*)
integer :: ngenh
integer :: ijkgd(-ngenh:ngenh)
real*8  :: eigr(0:*), eigi(*), apk(*), apki(*), scwr(*), scwi(*)
real*8  :: prod(*), prodr(*), w1r(*), w1i(*)
real*8  :: rdelv, rvol
integer :: nhk1(*), nhk2(*), iatoit(*), ntype, natom
logical :: lvand, lkbpp_r, lvand_r, lkbpp_g, lvand_g
logical :: lkbppi(ntype), lvandi(ntype), lking(ntype)
integer :: iter, istdot
real*8  :: tolera, threinn, wdinner
integer :: ihyoji, itermx, noband, iteremx
logical :: lortho
logical :: lhcunt, ltimecnt, leig_start, leig_end
real*8  :: bzan(*)
integer :: nriter(*)
integer :: lmax(*), lclno(*)
integer :: mxl
logical :: lchk(0:mxl,*)
real*8  :: ycos(*), ysin(*)
integer :: nylmmx, nplwcs
integer :: method
real*8  :: hfrhcr(*)
integer :: jhybrid
logical :: lefield_islts

real*8, allocatable :: vfft(:)
integer :: ifd, ib

if( nodes /= 1 .or. nbnod /= nband ) call qxmd_stop( 'eigen: one rank only' )
if( method == 1 ) call qxmd_stop( 'eigen: line minimization (method 1) is not provided; BKL only' )
if( lvand .or. lkbpp_r .or. jhybrid /= 0 .or. lefield_islts .or. .not.lortho )  &
&   call qxmd_stop( 'eigen: norm-conserving KB in G space, no hybrid or field terms only' )
call qxmd_gpu_need_ctx( 'eigen' )
!--- the local potential of this SCF iteration (glocal, VECTOR order)
allocate( vfft(kfft0) )
vfft = 0.d0
do ifd = 1, ntotfd
   vfft(mfd2ft(ifd)) = glocal(ifd)
end do
call qxmd_gpu_set_glocal( vfft )
deallocate( vfft )
!--- psi and H psi into the device slots, the BKL update (the reference's
!--- band-by-band schedule), both back
call qxmd_gpu_check( qx_slot_put( qxctx, 0_c_int, gdcr, int(nband, c_int) ), 'eigen: qx_slot_put' )
call qxmd_gpu_check( qx_slot_put( qxctx, 1_c_int, rhcr, int(nband, c_int) ), 'eigen: qx_slot_put' )
call qxmd_gpu_check( qx_eigen_bkl_slot( qxctx, 0_c_int, 1_c_int, int(nband, c_int), ekib, eig,  &
&    int(noband, c_int), int(itermx, c_int), int(iteremx, c_int), tolera, threinn, wdinner,  &
&    roughfzan, roughfeig, bzan, nriter, QX_FP64 ), 'eigen: qx_eigen_bkl_slot' )
call qxmd_gpu_check( qx_slot_get( qxctx, 0_c_int, gdcr, int(nband, c_int) ), 'eigen: qx_slot_get' )
call qxmd_gpu_check( qx_slot_get( qxctx, 1_c_int, rhcr, int(nband, c_int) ), 'eigen: qx_slot_get' )
!--- the unified copy the reference keeps in step (one rank: the same bands)
do ib = 1, nband
   cgjr(1:nplwex,ib) = gdcr(1:nplwex,ib)
end do

return
end subroutine


subroutine qxmd_stop( what )
implicit none
character(*) :: what
write(*,*) 'qxmd_amd: ', what
error stop 1
end subroutine
