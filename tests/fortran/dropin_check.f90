!-----------------------------------------------------------------------
! Fortran host check of the drop-in layer: every replaced QXMD routine is
! called by its reference name and argument list through the ISO_C_BINDING
! layer, on an 8-atom Si cell, and checked against closed forms.  Prints one
! line per check and "DROPIN OK" at the end; stops with status 1 on failure.
!-----------------------------------------------------------------------
program dropin_check
use qxmd_amd_gpu
implicit none
integer, parameter :: nb = 8, nmat = 12
real*8 :: hcell(3, 3), pi, rvol, err, worst
integer :: nd(3), nplw, nplwex, n1, n2, n3, ntot, ierror, i, j, k, l, ib, ia
real*8, allocatable :: f(:), fx(:), fy(:), f0(:), cg(:,:), rh(:,:), rho(:), occ(:)
real*8, allocatable :: a(:,:), a0(:,:), w(:), s(:,:), s0(:,:), t(:,:), eq(:,:,:), vcl(:,:)
real*8, allocatable :: slm(:,:,:), ref(:,:), glocal(:), ycos(:,:), ysin(:,:)
complex*16 :: work(1), fftwork(1), acc
integer, allocatable :: lmx(:), mfd2ft(:), iod(:), nfile(:), ijk(:), nhk(:)
logical, allocatable :: lk(:)
logical :: lchol
character*5 :: chk
real*8 :: w1r(nmat), w1i(nmat), wk(nmat)
integer :: seed(64)

pi = 4.d0 * atan(1.d0)
worst = 0.d0
seed = 20261202
call random_seed(put=seed(1:size(seed)))
hcell = 0.d0
hcell(1, 1) = 10.26d0
hcell(2, 2) = 10.26d0
hcell(3, 3) = 10.26d0
rvol = 10.26d0**3
call qxmd_gpu_init_cell(hcell, 20.d0, 80.d0, 16, nd, nplw)
nplwex = 2 * (nplw + 1)
n1 = nd(1); n2 = nd(2); n3 = nd(3); ntot = n1 * n2 * n3
write(*, '(a, 3i4, a, i6)') 'grid', nd, '  nplw', nplw

! ---- rfft3: one coefficient against a direct sum, then the round trip
allocate(f(ntot), fx(ntot), fy(ntot), f0(ntot))
call random_number(f0)
f0 = f0 - 0.5d0
fx = f0
fy = 0.d0
call rfft3(1, fx, fy, work, n1, n2, n3, ntot, ierror)
call expect(ierror == 0, 'rfft3 forward ierror')
! F(1, 2, 3) = (1/N) sum_r f(r) exp(-2 pi i (x/n1 + 2 y/n2 + 3 z/n3))
acc = (0.d0, 0.d0)
do k = 0, n3 - 1
  do j = 0, n2 - 1
    do i = 0, n1 - 1
      acc = acc + f0(1 + i + n1 * (j + n2 * k)) *  &
&           exp(cmplx(0.d0, -2.d0 * pi * (dble(i) / n1 + 2.d0 * j / n2 + 3.d0 * k / n3), kind=8))
    end do
  end do
end do
acc = acc / ntot
l = 1 + 1 + n1 * (2 + n2 * 3)
err = abs(cmplx(fx(l), fy(l), kind=8) - acc) / abs(acc)
call report('rfft3 coefficient vs direct sum', err, 1.d-10)
call rfft3(-1, fx, fy, work, n1, n2, n3, ntot, ierror)
call expect(ierror == 0, 'rfft3 backward ierror')
err = maxval(abs(fx - f0)) / maxval(abs(f0))
call report('rfft3 round trip', err, 1.d-12)

! ---- fft3: complex round trip
call random_number(fx)
call random_number(fy)
f = fx
f0 = fy
call fft3(1, fx, fy, work, n1, n2, n3, ntot, ierror)
call fft3(-1, fx, fy, work, n1, n2, n3, ntot, ierror)
err = max(maxval(abs(fx - f)), maxval(abs(fy - f0)))
call report('fft3 round trip', err, 1.d-12)

! ---- RsEIGQR: residual, orthonormality, ascending order
allocate(a(nmat, nmat), a0(nmat, nmat), w(nmat))
call random_number(a0)
a0 = a0 + transpose(a0)
a = a0
call RsEIGQR(a, w, nmat, nmat, 1.d-15, 1, 2, chk, w1r, w1i, wk)
err = 0.d0
do j = 1, nmat
  err = max(err, maxval(abs(matmul(a0, a(:, j)) - w(j) * a(:, j))))
  if (j > 1) call expect(w(j) >= w(j - 1), 'RsEIGQR ascending')
end do
err = max(err, maxval(abs(matmul(transpose(a), a) - eye(nmat))))
call report('RsEIGQR residual + orthonormality', err, 1.d-12)

! ---- icholesky: L^-1 S L^-T = I, upper triangle zero
allocate(s(nmat, nmat), s0(nmat, nmat), t(nmat, nmat))
call random_number(t)
s0 = matmul(transpose(t), t) + 0.1d0 * eye(nmat)
s = s0
allocate(nfile(1), iod(1))
call icholesky(nfile, 0, 1, s, nmat, 1, nmat, nmat, t, t, iod, iod, 1, iod, iod, iod, iod, iod,  &
&              t, t)
err = maxval(abs(matmul(matmul(s, s0), transpose(s)) - eye(nmat)))
do j = 2, nmat
  err = max(err, maxval(abs(s(1:j - 1, j))))
end do
call report('icholesky L^-1 S L^-T = I', err, 1.d-12)

! ---- hcnonloc with a constant potential: rhcr = v0 * cgjr exactly
allocate(cg(nplwex, nb), rh(nplwex, nb), glocal(ntot))
call random_number(cg)
cg = cg - 0.5d0
cg(2, :) = 0.d0  ! Im psi(G = 0) = 0
glocal = 0.7d0
call qxmd_gpu_set_glocal(glocal)
rh = 0.d0
call qxmd_hcnonloc(cg(:, 1), rh(:, 1))
err = maxval(abs(rh(:, 1) - 0.7d0 * cg(:, 1))) / maxval(abs(cg(:, 1)))
call report('hcnonloc constant potential', err, 1.d-12)

! ---- schmidt: packed overlap 2 sum - x0 y0 = identity afterwards
call qxmd_schmidt(cg, nplwex, nb, lchol)
call expect(lchol, 'schmidt lcholesky')
err = 0.d0
do j = 1, nb
  do i = 1, nb
    err = max(err, abs(2.d0 * dot_product(cg(:, i), cg(:, j)) - cg(1, i) * cg(1, j) -  &
&                      merge(1.d0, 0.d0, i == j)))
  end do
end do
call report('schmidt orthonormality', err, 1.d-12)

! ---- softchg: integral of rho = sum of occupations
allocate(rho(ntot), occ(nb), mfd2ft(ntot), ijk(1), nhk(1), lk(1))
deallocate(iod)
allocate(iod(nb))
do i = 1, nb
  occ(i) = 2.d0 - 0.1d0 * i
  iod(i) = i
end do
do i = 1, ntot
  mfd2ft(i) = i
end do
call softchg(nfile, 0, 1, rho, occ, cg, nplwex, nplw, nb, 1, nb, nb, iod, rvol, mfd2ft, ntot,  &
&            nd, ntot, ijk, 0, fx, fy, fftwork)
err = abs(sum(rho) * rvol / ntot - sum(occ)) / sum(occ)
call report('softchg integral = sum occ', err, 1.d-12)

! ---- hckbpp_g: the KB sum 2 sum_{k>=3} + x0 y0 and its back-projection
ia = 2
allocate(eq(nplwex, 4, ia), vcl(4, ia), lmx(ia), slm(4, ia, nb), ref(nplwex, nb))
allocate(ycos(0:1, ia), ysin(0:1, ia))
call random_number(eq)
eq = eq - 0.5d0
call random_number(vcl)
lmx = 4
call qxmd_gpu_set_kb(eq, nplwex, 4, ia, lmx, vcl)
rh = 0.d0
call hckbpp_g(nfile, 0, 1, cg, rh, nplwex, nplw, nb, 1, nb, nb, 1, nhk, nhk, ia, lk, lk,  &
&             ycos, ysin, 1)
ref = 0.d0
do ib = 1, nb
  do k = 1, ia
    do l = 1, 4
      slm(l, k, ib) = 2.d0 * dot_product(eq(3:, l, k), cg(3:, ib)) + eq(1, l, k) * cg(1, ib)
      ref(:, ib) = ref(:, ib) + vcl(l, k) * slm(l, k, ib) * eq(:, l, k)
    end do
  end do
end do
err = maxval(abs(rh - ref)) / maxval(abs(ref))
err = max(err, maxval(abs(qx_slmir(:, :, 1:nb) - slm)) / maxval(abs(slm)))
call report('hckbpp_g projection + back-projection', err, 1.d-12)
rh(:, 1) = 0.d0
call hckbpp_ib_g(nfile, 0, 1, cg(:, 3), rh(:, 1), nplwex, nplw, 1, nhk, nhk, ia, lk, lk,  &
&                ycos, ysin, 1)
err = maxval(abs(rh(:, 1) - ref(:, 3))) / maxval(abs(ref(:, 3)))
call report('hckbpp_ib_g one band', err, 1.d-12)
call check_forces()

! ---- stscfn_pw: 3 atoms, 40 G indices in -3..3, against cos/sin(2 pi G . x)
call check_stscfn()

! ---- SCF glue: chgg2r / harting / chgr2g on a three-vector density, xclda,
! ---- simple and Pulay mixing through mxchgg_alloc / mxchgg
call check_scf_glue()

! ---- flocal_g: two atoms, three G vectors, against the sums written out
call check_flocal()

! ---- frcrec (Ewald reciprocal part): two ions, four G, against the sums
call check_frcrec()

! ---- sbalinaspc: bands turning inside one subspace are predicted as they are
call check_aspc()

! ---- velocity Verlet (normalized_a, vkick, updtps, updtps_pbc): a harmonic
! ---- crystal run forward and back returns to its start
call check_verlet()

call qxmd_gpu_finalize()
write(*, '(a, es10.2)') 'DROPIN OK worst', worst

contains

function eye(n) result(m)
  integer, intent(in) :: n
  real*8 :: m(n, n)
  integer :: q
  m = 0.d0
  do q = 1, n
    m(q, q) = 1.d0
  end do
end function

subroutine report(what, e, tol)
  character(*), intent(in) :: what
  real*8, intent(in) :: e, tol
  write(*, '(a, a, es10.2)') what, ': ', e
  worst = max(worst, e)
  if (.not. (e <= tol)) then
    write(*, '(a, a)') 'FAILED ', what
    error stop 1
  end if
end subroutine

! kbnlpp_frc2 against its sums written out, with G = (1, 2, 3) scaled per index
subroutine check_forces()
  real*8 :: gxv(0:nplw), gyv(0:nplw), gzv(0:nplw), fn(3, ia), fr(3, ia), z, sx, sy, sz, wgt
  real*8 :: oc(nb), st(6)
  integer :: io(nb), ig, jb, ja, ll
  do ig = 0, nplw
    gxv(ig) = 0.01d0 * ig
    gyv(ig) = 0.02d0 * ig
    gzv(ig) = -0.03d0 * ig
  end do
  do jb = 1, nb
    oc(jb) = 2.d0 - 0.25d0 * jb
    io(jb) = jb
  end do
  fr = 0.d0
  do ja = 1, ia
    do ll = 1, 4
      do jb = 1, nb
        sx = 0.d0; sy = 0.d0; sz = 0.d0
        do ig = 1, nplw
          z = eq(2 * ig + 1, ll, ja) * cg(2 * ig + 2, jb) - eq(2 * ig + 2, ll, ja) * cg(2 * ig + 1, jb)
          sx = sx + z * gxv(ig); sy = sy + z * gyv(ig); sz = sz + z * gzv(ig)
        end do
        wgt = 2.d0 * slm(ll, ja, jb) * vcl(ll, ja) * oc(jb) * 2.d0
        fr(1, ja) = fr(1, ja) + sx * wgt
        fr(2, ja) = fr(2, ja) + sy * wgt
        fr(3, ja) = fr(3, ja) + sz * wgt
      end do
    end do
  end do
  fn = 1.d0
  st = 0.d0
  call kbnlpp_frc2(nfile, 0, 1, fn, st, .false., cg, nplwex, nplw, gxv, gyv, gzv, nb, 1, nb, nb,  &
&     oc, io, ia, nhk, nhk, ia, lk, lk, ycos, ysin, 1)
  call report('kbnlpp_frc2 forces', maxval(abs(fn - 1.d0 - fr)) / maxval(abs(fr)), 1.d-11)
end subroutine

subroutine check_stscfn()
  integer, parameter :: na = 3, ng = 40, km = 3
  real*8 :: ra(3, na), yc(0:ng, na), ys(0:ng, na), ph, e2
  real*8 :: c1(-km:km, na), c2(-km:km, na), c3(-km:km, na)
  real*8 :: s1(-km:km, na), s2(-km:km, na), s3(-km:km, na), h(3, 3)
  integer :: g1(0:ng), g2(0:ng), g3(0:ng), d(3), ig, ja, nf(2)
  call random_number(ra)
  g1(0) = 0; g2(0) = 0; g3(0) = 0
  do ig = 1, ng
    g1(ig) = mod(ig, 7) - 3
    g2(ig) = mod(3 * ig, 7) - 3
    g3(ig) = mod(5 * ig + 1, 7) - 3
  end do
  d = 30
  h = 0.d0
  nf = 6
  call stscfn_pw(nf, 0, 1, yc, ys, c1, c2, c3, s1, s2, s3, km, km, km, na, ng, ng, g1, g2, g3,  &
&     .false., ra, h, d, d, .true., .false., .false., .false.)
  e2 = 0.d0
  do ja = 1, na
    do ig = 0, ng
      ph = 2.d0 * pi * (g1(ig) * ra(1, ja) + g2(ig) * ra(2, ja) + g3(ig) * ra(3, ja))
      e2 = max(e2, abs(yc(ig, ja) - cos(ph)), abs(ys(ig, ja) - sin(ph)))
    end do
    e2 = max(e2, abs(c2(-2, ja) - cos(4.d0 * pi * ra(2, ja))), abs(s3(3, ja) - sin(6.d0 * pi * ra(3, ja))))
  end do
  call report('stscfn_pw structure factors vs cos/sin(G.x)', e2, 1.d-13)
end subroutine

! A density of G = 0 and three G vectors (and mirrors): chgg2r, harting and
! chgr2g against the Fourier sums written out; xclda's potential against the
! derivative of its energy density; simple mixing and the first two Pulay
! steps (Kerker, then the 2 x 2 normal equation) against their closed forms.
subroutine check_scf_glue()
  integer, parameter :: n5 = 3, n5x = 2 * (n5 + 1)
  integer :: gv(3, 0:n5), ijg(-n5:n5), ig, ix, iy, iz, q, nf(1)
  real*8 :: rx(0:n5), g2(0:n5), rg(0:n5x - 1), rg2(0:n5x - 1), r0(ntot), vh(ntot), gl(ntot)
  real*8 :: rref, vref, ph, e1, e2, tpl, vx(ntot), ee(ntot), vxp(ntot), eep(ntot), vxm(ntot)
  real*8 :: eem(ntot), eexc, evexc, ecor, evcor, ex2, ev2, ec2, evc2, hh(ntot), dkg(n5x)
  real*8 :: rin(n5x), out(n5x), in1(n5x), f1(n5x), f2(n5x), wq(n5x), fk(n5x), a11, a12, a22
  real*8 :: c1, c2, q1, mixed(n5x), fx_(ntot), fy_(ntot), amem
  logical :: lres
  complex*16 :: fw(1)
  tpl = 2.d0 * pi / 10.26d0
  gv(:, 0) = (/0, 0, 0/)
  gv(:, 1) = (/1, 0, 0/)
  gv(:, 2) = (/0, 2, 1/)
  gv(:, 3) = (/1, -1, 2/)
  do ig = 0, n5
    ijg(ig) = 1 + modulo(gv(1, ig), n1) + n1 * (modulo(gv(2, ig), n2) + n2 * modulo(gv(3, ig), n3))
    ijg(-ig) = 1 + modulo(-gv(1, ig), n1) + n1 * (modulo(-gv(2, ig), n2) +  &
&              n2 * modulo(-gv(3, ig), n3))
    g2(ig) = tpl**2 * sum(gv(:, ig)**2)
    rx(ig) = 0.d0
    if (ig > 0) rx(ig) = 4.d0 * pi / g2(ig)
  end do
  rg = (/0.03d0, 0.d0, 0.004d0, -0.002d0, -0.003d0, 0.001d0, 0.0015d0, 0.0025d0/)
  nf = 6
  call chgg2r(nf, 0, 1, r0, rg, n5x, n5, mfd2ft, ntot, nd, ntot, fx_, fy_, fw, ijg, n5)
  call harting(nf, 0, 1, vh, rg, n5x, n5, rx, mfd2ft, ntot, nd, ntot, fx_, fy_, fw, ijg, n5)
  e1 = 0.d0
  e2 = 0.d0
  do iz = 0, n3 - 1
    do iy = 0, n2 - 1
      do ix = 0, n1 - 1
        rref = rg(0)
        vref = 0.d0
        do ig = 1, n5
          ph = 2.d0 * pi * (gv(1, ig) * dble(ix) / n1 + gv(2, ig) * dble(iy) / n2 +  &
&              gv(3, ig) * dble(iz) / n3)
          rref = rref + 2.d0 * (rg(2 * ig) * cos(ph) - rg(2 * ig + 1) * sin(ph))
          vref = vref + 2.d0 * rx(ig) * 2.d0 * (rg(2 * ig) * cos(ph) - rg(2 * ig + 1) * sin(ph))
        end do
        q = 1 + ix + n1 * (iy + n2 * iz)
        e1 = max(e1, abs(r0(q) - rref))
        e2 = max(e2, abs(vh(q) - vref))
      end do
    end do
  end do
  call report('chgg2r vs Fourier sum', e1 / 0.03d0, 1.d-12)
  call report('harting vs 8 pi/G^2 Fourier sum', e2 / maxval(abs(vh)), 1.d-12)
  call chgr2g(nf, 0, 1, r0, rg2, n5x, n5, mfd2ft, ntot, nd, ntot, fx_, fy_, fw, ijg, n5)
  call report('chgr2g(chgg2r(rhgr)) = rhgr', maxval(abs(rg2 - rg)) / 0.03d0, 1.d-13)
  ! xclda: vexc = d(rho eeexc)/d rho by central differences
  call xclda(r0, -r0, ntot, rvol / ntot, vx, ee, eexc, evexc, ecor, evcor)
  hh = 1.d-6 * r0
  call xclda(r0 + hh, -r0, ntot, rvol / ntot, vxp, eep, ex2, ev2, ec2, evc2)
  call xclda(r0 - hh, -r0, ntot, rvol / ntot, vxm, eem, ex2, ev2, ec2, evc2)
  e1 = maxval(abs(vx - ((r0 + hh) * eep - (r0 - hh) * eem) / (2.d0 * hh))) / maxval(abs(vx))
  call report('xclda vexc = d(rho eeexc)/d rho', e1, 1.d-7)
  e2 = abs(sum(ee * r0) * rvol / ntot - (eexc + ecor)) / abs(eexc + ecor)
  call report('xclda energy sums', e2, 1.d-12)
  ! simple mixing
  do ig = 0, n5
    dkg(2 * ig + 1) = g2(ig)
    dkg(2 * ig + 2) = g2(ig)
  end do
  rin = rg
  out = rg * 1.3d0
  out(2) = 0.d0
  call mxchgg_alloc(nf, 0, 1, amem, .false., n5x, 3, 4, 0, 0, 2.d0, 1.d0, 1.d0, 1.d0, .false.)
  call set_g2ming2max_in_chgdns(g2(1), g2(n5))
  lres = .true.
  mixed = out
  call mxchgg(nf, 0, 1, lres, mixed, rin, n5x, n5, dkg, rvol, 0.3d0, 0.5d0, gl, fx_, fx_,  &
&     mfd2ft, ntot, nd, ntot, fx_, fy_, fw, ijg, n5, rvol / ntot, 8, .false.)
  e1 = maxval(abs(mixed - (0.7d0 * rg + 0.3d0 * out)))
  e1 = max(e1, maxval(abs(rin - mixed)))
  call chgg2r(nf, 0, 1, r0, mixed, n5x, n5, mfd2ft, ntot, nd, ntot, fx_, fy_, fw, ijg, n5)
  e1 = max(e1, maxval(abs(gl - r0)))
  call report('mxchgg simple mixing', e1 / 0.03d0, 1.d-13)
  ! Pulay (cmetric = 2): Kerker first, then the two-residual combination
  call mxchgg_alloc(nf, 0, 1, amem, .false., n5x, 1, 4, 0, 0, 2.d0, 1.d0, 1.d0, 1.d0, .false.)
  call set_g2ming2max_in_chgdns(g2(1), g2(n5))
  q1 = (2.d0 - 1.d0) * g2(n5) * g2(1) / (g2(n5) - 2.d0 * g2(1))
  do q = 1, n5x
    fk(q) = 0.8d0 * dkg(q) / (dkg(q) + 0.6d0)
    wq(q) = 0.d0
    if (q >= 3) wq(q) = 2.d0 * (dkg(q) + q1) / dkg(q)
  end do
  wq(1) = (dkg(3) + q1) / dkg(3)
  in1 = rg
  out = rg
  out(3:) = rg(2:) * 1.2d0 + 0.0005d0
  f1 = out - in1
  rin = in1
  mixed = out
  lres = .false.
  call mxchgg(nf, 0, 1, lres, mixed, rin, n5x, n5, dkg, rvol, 0.8d0, 0.6d0, gl, fx_, fx_,  &
&     mfd2ft, ntot, nd, ntot, fx_, fy_, fw, ijg, n5, rvol / ntot, 8, .false.)
  e1 = maxval(abs(mixed - (in1 + fk * f1)))
  out = mixed
  out(3:) = out(3:) * 0.9d0 - 0.0003d0
  f2 = out - mixed
  a11 = rvol * sum(wq * f1 * f1)
  a12 = rvol * sum(wq * f1 * f2)
  a22 = rvol * sum(wq * f2 * f2)
  c1 = (a22 - a12) / (a11 + a22 - 2.d0 * a12)
  c2 = (a11 - a12) / (a11 + a22 - 2.d0 * a12)
  in1 = in1 + 0.d0
  rg2 = mixed
  call mxchgg(nf, 0, 1, lres, out, rin, n5x, n5, dkg, rvol, 0.8d0, 0.6d0, gl, fx_, fx_,  &
&     mfd2ft, ntot, nd, ntot, fx_, fy_, fw, ijg, n5, rvol / ntot, 8, .false.)
  e2 = maxval(abs(out - (c1 * (in1 + fk * f1) + c2 * (rg2 + fk * f2))))
  call report('mxchgg Pulay: Kerker step, then 2 x 2 normal equation', max(e1, e2) / 0.03d0,  &
&     1.d-11)
end subroutine

! flocal_g on the density sphere of check_scf_glue's three G vectors, two
! atoms of two types (iatmpt swapped), axis tables cos/sin(2 pi k s)
subroutine check_flocal()
  integer, parameter :: n5 = 3, na = 2, km = 2
  integer :: g1(0:n5), g2(0:n5), g3(0:n5), nh1(2), nh2(2), iat(na), it_(na), nf(1), ig, ja, q
  real*8 :: sv(n5 + 1, 2), gxv(0:n5), gyv(0:n5), gzv(0:n5), rg(2 * (n5 + 1)), s(3, na)
  real*8 :: c1(-km:km, na), c2(-km:km, na), c3(-km:km, na), d1(-km:km, na), d2(-km:km, na)
  real*8 :: d3(-km:km, na), fl(3, na), fr(3, na), st(3, 3), ph, ftmp, tpl
  logical :: llk(2)
  g1 = (/0, 1, 0, 1/)
  g2 = (/0, 0, 2, -1/)
  g3 = (/0, 0, 1, 2/)
  tpl = 2.d0 * pi / 10.26d0
  gxv = tpl * g1
  gyv = tpl * g2
  gzv = tpl * g3
  do ig = 0, n5
    sv(ig + 1, 1) = 1.d0 / (1.d0 + ig)
    sv(ig + 1, 2) = -0.5d0 * ig
  end do
  rg = (/0.03d0, 0.d0, 0.004d0, -0.002d0, -0.003d0, 0.001d0, 0.0015d0, 0.0025d0/)
  s(:, 1) = (/0.1d0, 0.25d0, 0.7d0/)
  s(:, 2) = (/0.55d0, 0.8d0, 0.05d0/)
  do ja = 1, na
    do q = -km, km
      c1(q, ja) = cos(2.d0 * pi * q * s(1, ja)); d1(q, ja) = sin(2.d0 * pi * q * s(1, ja))
      c2(q, ja) = cos(2.d0 * pi * q * s(2, ja)); d2(q, ja) = sin(2.d0 * pi * q * s(2, ja))
      c3(q, ja) = cos(2.d0 * pi * q * s(3, ja)); d3(q, ja) = sin(2.d0 * pi * q * s(3, ja))
    end do
  end do
  ! ion 1 -> atom 2 (type 1), ion 2 -> atom 1 (type 2)
  nh1 = (/1, 2/)
  nh2 = (/1, 2/)
  iat = (/2, 1/)
  it_ = (/2, 1/)
  fr = 0.d0
  do ja = 1, na
    do ig = 1, n5
      ph = 2.d0 * pi * (g1(ig) * s(1, ja) + g2(ig) * s(2, ja) + g3(ig) * s(3, ja))
      ftmp = sv(ig + 1, it_(ja)) * (rg(2 * ig + 1) * sin(ph) + rg(2 * ig + 2) * cos(ph))
      fr(1, ja) = fr(1, ja) + 2.d0 * ftmp * gxv(ig)
      fr(2, ja) = fr(2, ja) + 2.d0 * ftmp * gyv(ig)
      fr(3, ja) = fr(3, ja) + 2.d0 * ftmp * gzv(ig)
    end do
  end do
  call qxmd_gpu_set_lopp(2, n5, sv, gxv, gyv, gzv)
  llk = .false.
  fl = 0.d0
  st = 0.d0
  nf = 6
  call flocal_g(nf, 0, 1, fl, st, .false., 2, nh1, nh2, it_, iat, na, na, llk, n5, 2 * (n5 + 1),  &
&     n5, g1, g2, g3, gxv, gyv, gzv, rg, 0.d0, km, km, km, c1, c2, c3, d1, d2, d3, .false.)
  call report('flocal_g vs its sums written out', maxval(abs(fl - fr)) / maxval(abs(fr)), 1.d-13)
end subroutine

! frcrec on four Ewald G vectors (units without 2 pi) and two ions of two
! types, iatmpt swapped: forces, total and per-atom energies written out
subroutine check_frcrec()
  integer, parameter :: ng = 4, na = 2, km = 2
  integer :: g1(ng), g2(ng), g3(ng), nh1(2), nh2(2), iat(na), imp(na), m, k, ja, q
  real*8 :: gxv(ng), gyv(ng), gzv(ng), rn(ng), rx(0:ng), zv2(2), s(3, na), gam, vol
  real*8 :: c1(-km:km, na), c2(-km:km, na), c3(-km:km, na), d1(-km:km, na), d2(-km:km, na)
  real*8 :: d3(-km:km, na), fl(3, na), fr(3, na), ae(na), aer(na), ec, er, hep, sc, ss, ph
  real*8 :: yc(na), ys(na), cst
  g1 = (/1, 0, 1, 2/)
  g2 = (/0, 1, -1, 1/)
  g3 = (/0, 1, 2, -2/)
  do m = 1, ng
    gxv(m) = g1(m) / 10.26d0
    gyv(m) = g2(m) / 10.26d0
    gzv(m) = g3(m) / 10.26d0
    rn(m) = gxv(m)**2 + gyv(m)**2 + gzv(m)**2
  end do
  rx = 0.d0
  gam = 0.4d0
  vol = 10.26d0**3
  zv2 = (/4.d0, 3.d0/)
  s(:, 1) = (/0.1d0, 0.25d0, 0.7d0/)
  s(:, 2) = (/0.55d0, 0.8d0, 0.05d0/)
  nh1 = (/1, 2/)
  nh2 = (/1, 2/)
  iat = (/2, 1/)   ! atom 1 is type 2, atom 2 type 1
  imp = (/2, 1/)   ! ion 1 is atom 2, ion 2 atom 1
  do k = 1, na
    ja = imp(k)
    do q = -km, km
      c1(q, k) = cos(2.d0 * pi * q * s(1, ja)); d1(q, k) = sin(2.d0 * pi * q * s(1, ja))
      c2(q, k) = cos(2.d0 * pi * q * s(2, ja)); d2(q, k) = sin(2.d0 * pi * q * s(2, ja))
      c3(q, k) = cos(2.d0 * pi * q * s(3, ja)); d3(q, k) = sin(2.d0 * pi * q * s(3, ja))
    end do
  end do
  fr = 0.d0
  aer = 0.d0
  er = 0.d0
  cst = 2.d0 / (vol * pi)
  do m = 1, ng
    sc = 0.d0
    ss = 0.d0
    do ja = 1, na
      ph = 2.d0 * pi * (g1(m) * s(1, ja) + g2(m) * s(2, ja) + g3(m) * s(3, ja))
      yc(ja) = zv2(iat(ja)) * cos(ph)
      ys(ja) = zv2(iat(ja)) * sin(ph)
      sc = sc + yc(ja)
      ss = ss + ys(ja)
    end do
    hep = exp(-(pi / gam)**2 * rn(m)) / rn(m)
    er = er + cst * hep * (sc * sc + ss * ss)
    do ja = 1, na
      fr(1, ja) = fr(1, ja) + 8.d0 / vol * hep * gxv(m) * (sc * ys(ja) - ss * yc(ja))
      fr(2, ja) = fr(2, ja) + 8.d0 / vol * hep * gyv(m) * (sc * ys(ja) - ss * yc(ja))
      fr(3, ja) = fr(3, ja) + 8.d0 / vol * hep * gzv(m) * (sc * ys(ja) - ss * yc(ja))
      aer(ja) = aer(ja) + cst * hep * (yc(ja) * sc + ys(ja) * ss)
    end do
  end do
  er = er - 49.d0 / vol * pi / gam**2 - 2.d0 * gam * 25.d0 / sqrt(pi)
  do ja = 1, na
    aer(ja) = aer(ja) - zv2(iat(ja)) * 7.d0 / vol * pi / gam**2  &
&             - 2.d0 * gam * zv2(iat(ja))**2 / sqrt(pi)
  end do
  call qxmd_gpu_set_ewald(ng, g1, g2, g3, gxv, gyv, gzv, rn, rx, gam, vol)
  fl = 0.d0
  ae = 0.d0
  ec = 0.d0
  call qxmd_frcrec(fl, 2, nh1, nh2, zv2, 7, ec, ae, iat, imp, na, na, km, km, km, c1, c2, c3,  &
&     d1, d2, d3)
  call report('frcrec forces vs sums written out', maxval(abs(fl - fr)) / maxval(abs(fr)), 1.d-12)
  call report('frcrec energies vs sums written out',  &
&     max(abs(ec - er) / abs(er), maxval(abs(ae - aer)) / maxval(abs(aer))), 1.d-12)
end subroutine

! five MD steps of the orthonormal cg (from the schmidt check) turned by
! Givens rotations inside their span: once aligned, every history set equals
! the current bands and the ASPC prediction returns them
subroutine check_aspc()
  real*8 :: g(nplwex, nb), g0(nplwex, nb), th, e2
  integer :: step, ia_, jb
  logical :: lc
  call qxmd_aspc_init(3, 2)
  e2 = 0.d0
  do step = 1, 5
    th = 0.3d0 * step
    g0 = cg
    do jb = 1, nb - 1, 2
      g0(:, jb) = cos(th) * cg(:, jb) - sin(th) * cg(:, jb + 1)
      g0(:, jb + 1) = sin(th) * cg(:, jb) + cos(th) * cg(:, jb + 1)
    end do
    g = g0
    call qxmd_sbalinaspc(g, nplwex, nb, step, 0, ia_, lc)
    call expect(lc, 'sbalinaspc lcholesky')
    if (step > 1) e2 = max(e2, maxval(abs(g - g0)))
  end do
  call report('sbalinaspc keeps bands turning in their span', e2, 1.d-12)
end subroutine

subroutine check_verlet()
  integer, parameter :: na = 4
  real*8 :: h(3, 3, 0:1), x(3, na), x0(3, na), xs(3, na), v(3, na), a(3, na), acon(2), e
  integer :: isp(na), step, dir
  logical :: lfix(2)
  h = 0.d0
  h(1, 1, 0) = 10.d0
  h(2, 2, 0) = 11.d0
  h(3, 3, 0) = 9.5d0
  h(1, 2, 0) = 0.4d0
  isp = (/ 1, 2, 1, 2 /)
  lfix = .false.
  acon = 0.5d0 * 40.d0**2 / (/ 51196.d0, 1837.d0 /)
  x0 = reshape((/ 0.3d0, 0.4d0, 0.5d0, 0.6d0, 0.35d0, 0.45d0, 0.55d0, 0.65d0, 0.5d0,  &
  &               0.42d0, 0.58d0, 0.37d0 /), (/ 3, na /))
  x = x0 + 0.01d0
  xs = x
  v = 0.d0
  do dir = 1, 2
    call spring(x, x0, h, a)
    call normalized_a(nfile, 0, 1, na, 2, a, isp, h, acon)
    do step = 1, 100
      call vkick(nfile, 0, 1, v, a, isp, na, 2, lfix)
      call updtps(nfile, 0, 1, x, v, isp, na, 2, lfix)
      call updtps_pbc(nfile, 0, 1, x, isp, na, 2)
      call spring(x, x0, h, a)
      call normalized_a(nfile, 0, 1, na, 2, a, isp, h, acon)
      call vkick(nfile, 0, 1, v, a, isp, na, 2, lfix)
    end do
    if (dir == 1) then
      call expect(maxval(abs(x - xs)) > 1.d-3, 'velocity Verlet moves the atoms')
      v = -v
    end if
  end do
  e = maxval(abs(x - xs))
  call report('velocity Verlet forward and back', e, 1.d-12)
end subroutine

! Cartesian forces -0.2 (r - r0) of independent springs, r = h x
subroutine spring(x, x0, h, f)
  real*8, intent(in) :: x(:, :), x0(:, :), h(3, 3, 0:1)
  real*8, intent(out) :: f(:, :)
  integer :: q
  do q = 1, size(x, 2)
    f(:, q) = -0.2d0 * matmul(h(:, :, 0), x(:, q) - x0(:, q))
  end do
end subroutine

subroutine expect(cond, what)
  logical, intent(in) :: cond
  character(*), intent(in) :: what
  if (.not. cond) then
    write(*, '(a, a)') 'FAILED ', what
    error stop 1
  end if
end subroutine

end program dropin_check
