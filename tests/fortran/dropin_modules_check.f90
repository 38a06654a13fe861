!-----------------------------------------------------------------------
! Check of the drop-ins that read QXMD's modules (qxmd_dropin_qxmd.f90):
! schmidt, pwpchg and eigen are called by their reference names and argument
! lists, with the reference's own module data (param, pwlda_variables,
! pwlda_pw, ifcholesky, eigen_variables from the reference build) filled as
! QXMD fills them for one rank at the Gamma point, on an 8-atom Si cell; each
! result is compared with the explicit-argument form of the same routine and
! with a closed form.  Prints one line per check and "MODULE DROPIN OK" at
! the end; stops with status 1 on failure.
!-----------------------------------------------------------------------
program dropin_modules_check
use param, only: nband, lspin, lgamma
use param_atom, only: lvand
use pwlda_variables, only: npnod, nplwex, nplw, nplw2, nplw2ex, kfft0, rvol
use pwlda_pw, only: gdcr, occ, glocal, rhgr, mfd2ft_c, ntotfd_c, nd1vks_c, ijkg, iods,  &
&                   fft3x, fft3y, fftwork
use ifcholesky, only: lcholesky, lstop
use eigen_variables, only: roughfeig, roughfzan
use qxmd_amd_gpu
use qxmd_amd_c
implicit none
integer, parameter :: nb = 8, n5 = 3
real*8 :: hcell(3, 3), pi, err, tel
integer :: nd(3), n1, n2, n3, ntot, i, j, ig, ib
integer :: gv(3, 0:n5), nfile(1), iodum(nb)
real*8, allocatable :: cg(:,:), cg0(:,:), rho_ref(:), rhg_ref(:), rh(:,:), ek(:), eg(:), bz(:)
real*8, allocatable :: eg2(:), bz2(:), cg2(:,:), rh2(:,:), vpot(:)
integer, allocatable :: nri(:), nri2(:), mf(:)
logical :: ok
integer :: seed(64)

pi = 4.d0 * atan(1.d0)
seed = 20261203
call random_seed(put=seed(1:size(seed)))
hcell = 0.d0
hcell(1, 1) = 10.26d0
hcell(2, 2) = 10.26d0
hcell(3, 3) = 10.26d0
call qxmd_gpu_init_cell(hcell, 20.d0, 80.d0, 16, nd, nplw)
n1 = nd(1); n2 = nd(2); n3 = nd(3); ntot = n1 * n2 * n3

! ---- module data as QXMD holds it for one rank at Gamma
nband = nb
lspin = .false.
lgamma = .true.
lvand = .false.
npnod = nplw + 1
nplwex = 2 * (nplw + 1)
rvol = 10.26d0**3
kfft0 = ntot
ntotfd_c = ntot
nd1vks_c = nd
nplw2 = n5
nplw2ex = 2 * (n5 + 1)
allocate(gdcr(nplwex * nb), occ(nb), glocal(ntot), rhgr(nplw2ex), mfd2ft_c(ntot),  &
&        ijkg(-n5:n5), iods(nb), fft3x(ntot), fft3y(ntot), fftwork(1))
do i = 1, ntot
  mfd2ft_c(i) = i
end do
do ib = 1, nb
  iods(ib) = ib
  iodum(ib) = ib
end do
! a small density sphere: G = 0 and three G vectors (with mirrors)
gv(:, 0) = (/0, 0, 0/)
gv(:, 1) = (/1, 0, 0/)
gv(:, 2) = (/0, 2, 1/)
gv(:, 3) = (/1, -1, 2/)
do ig = 0, n5
  ijkg(ig) = 1 + modulo(gv(1, ig), n1) + n1 * (modulo(gv(2, ig), n2) + n2 * modulo(gv(3, ig), n3))
  ijkg(-ig) = 1 + modulo(-gv(1, ig), n1) + n1 * (modulo(-gv(2, ig), n2) +  &
&             n2 * modulo(-gv(3, ig), n3))
end do
occ = 2.d0
occ(nb) = 0.5d0

! ---- schmidt by name against the explicit form and packed orthonormality
allocate(cg(nplwex, nb), cg0(nplwex, nb))
call random_number(cg0)
cg0 = cg0 - 0.5d0
cg0(2, :) = 0.d0  ! Im psi(G = 0) = 0
cg = cg0
call qxmd_schmidt(cg, nplwex, nb, ok)
call expect(ok, 'qxmd_schmidt lcholesky')
gdcr = reshape(cg0, (/nplwex * nb/))
lstop = .true.
lcholesky = .false.
call schmidt(nfile, 0, 1, 0.d0, gdcr, iodum, iodum, .false., .false., .false., .false.)
call expect(lcholesky, 'schmidt sets lcholesky')
err = maxval(abs(gdcr - reshape(cg, (/nplwex * nb/))))
call report('schmidt by name = qxmd_schmidt', err, 0.d0)
err = 0.d0
do j = 1, nb
  do i = 1, nb
    err = max(err, abs(pdot(cg(:, i), cg(:, j)) - merge(1.d0, 0.d0, i == j)))
  end do
end do
call report('schmidt: packed overlap = identity', err, 1.d-12)

! ---- pwpchg by name: glocal and rhgr against softchg + chgr2g called
! directly, and the electron count
call pwpchg(nfile, 0, 1, .false., .false.)
allocate(rho_ref(ntot), rhg_ref(nplw2ex), mf(ntot))
mf = mfd2ft_c
call softchg(nfile, 0, 1, rho_ref, occ, cg, nplwex, nplw, nb, 1, nb, nb, iods, rvol, mf,  &
&            ntot, nd, ntot, ijkg, n5, fft3x, fft3y, fftwork)
rhg_ref = 0.d0
call chgr2g(nfile, 0, 1, rho_ref, rhg_ref, nplw2ex, n5, mf, ntot, nd, ntot, fft3x, fft3y,  &
&           fftwork, ijkg, n5)
call report('pwpchg glocal = softchg', maxval(abs(glocal - rho_ref)), 0.d0)
call report('pwpchg rhgr = chgr2g(softchg)', maxval(abs(rhgr - rhg_ref)), 0.d0)
tel = sum(glocal) * rvol / ntot
call report('pwpchg: integral of rho = sum occ', abs(tel - sum(occ)) / sum(occ), 1.d-12)
call report('pwpchg: rhgr(G=0) = nel / volume', abs(rhgr(1) - sum(occ) / rvol) * rvol, 1.d-12)

! ---- eigen by name against the explicit call on the same inputs, and the
! returned eigenvalues are the Rayleigh quotients of the returned pairs
allocate(rh(nplwex, nb), ek(nb), eg(nb), bz(nb), nri(nb), vpot(ntot))
allocate(cg2(nplwex, nb), rh2(nplwex, nb), eg2(nb), bz2(nb), nri2(nb))
call random_number(vpot)
vpot = -0.3d0 + 0.05d0 * vpot
call qxmd_gpu_set_glocal(vpot)
gdcr = reshape(cg, (/nplwex * nb/))
call qxmd_gpu_check(qx_slot_put(qxctx, 0_c_int, cg, int(nb, c_int)), 'put')
call qxmd_gpu_check(qx_hpsi_slot(qxctx, 0_c_int, 1_c_int, int(nb, c_int), QX_FP64), 'hpsi')
call qxmd_gpu_check(qx_slot_get(qxctx, 1_c_int, rh, int(nb, c_int)), 'get')
call qxmd_gpu_check(qx_ekib_slot(qxctx, 0_c_int, int(nb, c_int), ek), 'ekib')
do ib = 1, nb
  eg(ib) = pdot(cg(:, ib), rh(:, ib))
end do
roughfeig = 1.d-3
roughfzan = 1.d-1
cg2 = cg
rh2 = rh
eg2 = eg
call qxmd_gpu_check(qx_slot_put(qxctx, 0_c_int, cg2, int(nb, c_int)), 'put')
call qxmd_gpu_check(qx_slot_put(qxctx, 1_c_int, rh2, int(nb, c_int)), 'put')
call qxmd_gpu_check(qx_eigen_bkl_slot(qxctx, 0_c_int, 1_c_int, int(nb, c_int), ek, eg2,  &
&    int(nb, c_int), 4_c_int, 4_c_int, 1.d-10, 0.d0, 0.d0, roughfzan, roughfeig, bz2, nri2,  &
&    QX_FP64), 'qx_eigen_bkl_slot')
call qxmd_gpu_check(qx_slot_get(qxctx, 0_c_int, cg2, int(nb, c_int)), 'get')
call qxmd_gpu_check(qx_slot_get(qxctx, 1_c_int, rh2, int(nb, c_int)), 'get')
call run_eigen()
err = max(maxval(abs(gdcr - reshape(cg2, (/nplwex * nb/)))), maxval(abs(eg - eg2)),  &
&         maxval(abs(rh - rh2)), maxval(abs(bz - bz2)))
call report('eigen by name = qx_eigen_bkl_slot', err, 0.d0)
call expect(all(nri == nri2), 'eigen nriter')
err = 0.d0
do ib = 1, nb
  cg(:, ib) = gdcr((ib - 1) * nplwex + 1:ib * nplwex)
  err = max(err, abs(pdot(cg(:, ib), rh(:, ib)) / pdot(cg(:, ib), cg(:, ib)) - eg(ib)))
end do
call report('eigen: eig = Rayleigh quotients of the returned bands', err, 1.d-10)

call qxmd_gpu_finalize()
write(*, '(a)') 'MODULE DROPIN OK'

contains

subroutine run_eigen()
  real*8 :: dkg(nplwex), prcd(nplwex), gnk(nplwex), hnk(nplwex), sck(1), hcsr(1), hfv(1)
  real*8 :: eigr(0:1), ycos(1), ysin(1), prod(1), rdum(1), hfrhcr(1)
  integer :: iter, ijkgd(-1:1), lmax(1), lclno(1), nhk(1), iatoit(1)
  logical :: lkb(1), lchk(0:0, 1)
  complex*16 :: fw(1)
  dkg = 0.d0
  call eigen(nfile, 0, 1, eg, gdcr, gdcr, rh, hcsr, nplwex, nplw, nplwex,  &
&   dkg, prcd, ek, gnk, hnk, sck, nb, 1, nb, nb, iodum,  &
&   vpot, mfd2ft_c, ntot, nd, ntot, hfv, hfv,  &
&   fft3x, fft3y, fw, ijkgd, 1,  &
&   eigr, rdum, rdum, rdum, rdum, rdum, rvol / ntot, rvol,  &
&   1, nhk, nhk, 1, iatoit, .false.,  &
&   .false., .false., .true., .false., lkb, lkb, lkb,  &
&   prod, prod, rdum, rdum, iter, 0, 1.d-10, 0.d0, 0.d0, .true.,  &
&   bz, nri, 0, 4, nb, 4,  &
&   .false., .false., .true., .true.,  &
&   lmax, 0, lchk, lclno, 1, ycos, ysin, nplw, 2,  &
&   hfrhcr, 0, .false.)
end subroutine

real*8 function pdot(x, y)
  real*8, intent(in) :: x(:), y(:)
  pdot = 2.d0 * sum(x * y) - x(1) * y(1)
end function

subroutine report(what, e, tol)
  character(*), intent(in) :: what
  real*8, intent(in) :: e, tol
  write(*, '(a, a, es10.2)') what, ': ', e
  if (.not. (e <= tol)) then
    write(*, '(a, a)') 'FAILED: ', what
    error stop 1
  end if
end subroutine

subroutine expect(cond, what)
  logical, intent(in) :: cond
  character(*), intent(in) :: what
  if (.not. cond) then
    write(*, '(a, a)') 'FAILED: ', what
    error stop 1
  end if
end subroutine

end program
