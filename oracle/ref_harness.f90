!-----------------------------------------------------------------------
! ref_harness.f90 -- C-callable wrappers around the reference QXMD
! routines that need no FFT, so the CPU oracle can be pinned against the
! reference's own code (test infrastructure only).
!
! Built by oracle/ref.mk together with the reference sources compiled
! where they lie; nothing from the reference is copied.  Every wrapper
! sets the reference's module state the way the single-rank serial build
! (no_comm) would, then calls the reference subroutine unchanged.
!-----------------------------------------------------------------------

subroutine qxref_get_mesh(hcell, ecutdens, nd) bind(C, name="qxref_get_mesh")
  use iso_c_binding
  implicit none
  real(c_double), intent(in) :: hcell(3,3)
  real(c_double), value :: ecutdens
  integer(c_int) :: nd(3)
  integer :: nfile(2), nd1v(3), nd1vks(3)
  real*8 :: h(3,3), ec
  nfile = 6
  h = hcell
  ec = ecutdens
  call get_mesh(nfile, 0, 1, ec, h, nd1v, nd1vks)
  nd = nd1vks
end subroutine

! strvec + set_rciprl + reclat: G list in sorted order and the ijkgd map.
subroutine qxref_gsphere(rba_in, ecut_in, ecutdens_in, cap, counts, nga_o, ngb_o, ngc_o, &
&                        recnrm_o, nd, ijkgd_o) bind(C, name="qxref_gsphere")
  use iso_c_binding
  use pwlda_pw, only: nga, ngb, ngc, recnrm
  use reclat_variables, only: rv_lnc => lnoncollinear
  implicit none
  real(c_double), intent(in) :: rba_in(3,3)
  real(c_double), value :: ecut_in, ecutdens_in
  integer(c_int), value :: cap
  integer(c_int) :: counts(6), nga_o(0:cap-1), ngb_o(0:cap-1), ngc_o(0:cap-1), nd(3)
  real(c_double) :: recnrm_o(0:cap-1)
  integer(c_int) :: ijkgd_o(*)
  integer :: nfile(2), nplw5, nplw, nplw2, nplwc, nplw7, nplwcs, nplw3
  integer :: nplw5ex, nplwex, nplw2ex, nplw3ex, nplwcex, nplw7ex, kfft0d
  integer :: kfft1, kfft2, kfft3, kfft0, kfft1b, kfft2b, kfft3b, kfft0b
  integer :: kmax1, kmax2, kmax3, kmax1d, kmax2d, kmax3d, kmax1b, kmax2b, kmax3b
  integer :: kmax1cs, kmax2cs, kmax3cs, lplh, i
  real*8 :: rba(3,3), rvol, cross(3,3), rk1d(3), pwscale, ecut, ecutdens, ecutsoft, ecutc, ecutorth
  integer, allocatable :: ijkgd(:), ijkg(:)
  real*8 :: gbox(0:0), abg(0:0)
  integer :: ngbox(0:0), ijkgb(0:0)
  nfile = 6
  rba = rba_in
  ecut = ecut_in
  ecutdens = ecutdens_in
  ecutsoft = ecutdens
  ecutc = ecut
  ecutorth = ecut
  pwscale = 1.2d0
  rv_lnc = .false.
  nplw5 = 0
  call strvec(nfile, 0, 1, nplw5, nplw, nplw2, nplwc, nplw7, nplwcs, &
&             ecutdens, ecut, ecutsoft, ecutc, ecutorth, rba, rvol, cross, rk1d, &
&             .true., pwscale, .false., .false.)
  counts(1) = 0
  counts(2) = nplw5
  counts(3) = nplw
  counts(4) = nplw2
  counts(5) = nplwc
  counts(6) = nplw7
  if (nplw5 + 1 > cap) then
    counts(1) = -1
    return
  end if
  do i = 0, nplw5
    nga_o(i) = nga(i)
    ngb_o(i) = ngb(i)
    ngc_o(i) = ngc(i)
    recnrm_o(i) = recnrm(i)
  end do
  nplw3 = 0
  nplw5ex = 2*(nplw5+1); nplwex = 2*(nplw+1); nplw2ex = 2*(nplw2+1)
  nplw3ex = 2; nplwcex = 2*(nplwc+1); nplw7ex = 2*(nplw7+1)
  kfft0d = nd(1)*nd(2)*nd(3)
  lplh = nplw
  allocate(ijkgd(-nplw5:nplw5), ijkg(-nplw2:nplw2))
  kfft1b = 0; kfft2b = 0; kfft3b = 0; kfft0b = 0
  call reclat(nfile, 0, 1, nd, rba, ecut, ecutdens, ecutsoft, ecutorth, &
&     nplw5, nplw, nplw2, nplw3, nplwc, nplw7, nplwcs, &
&     nplw5ex, nplwex, nplw2ex, nplw3ex, nplwcex, nplw7ex, kfft0d, &
&     nga, ngb, ngc, recnrm, ijkgd, nplw5, lplh, &
&     kfft1, kfft2, kfft3, kfft0, ijkg, &
&     kfft1b, kfft2b, kfft3b, kfft0b, &
&     gbox, gbox, gbox, ngbox, ngbox, ngbox, abg, ijkgb, &
&     kmax1, kmax2, kmax3, kmax1d, kmax2d, kmax3d, &
&     kmax1b, kmax2b, kmax3b, kmax1cs, kmax2cs, kmax3cs)
  do i = -nplw5, nplw5
    ijkgd_o(i + nplw5 + 1) = ijkgd(i)
  end do
  counts(1) = kfft0
  deallocate(ijkgd, ijkg)
end subroutine

! clmrcp (Sources/bulk.f90:556) on its own: dkmax2 set in para_ewald as ewaldini
! leaves it, no double grid; the list it builds in para_ewald copied out.
subroutine qxref_clmrcp(b_in, dkmax2_in, cap, out_n, nga_o, ngb_o, ngc_o, gx_o, gy_o, gz_o,  &
&                       recnrm_o, kcsma_o) bind(C, name="qxref_clmrcp")
  use iso_c_binding
  use para_ewald, only: dkmax2, numk2, nga, ngb, ngc, gx, gy, gz, recnrm, KCSMA1, KCSMA2, KCSMA3
  implicit none
  real(c_double), intent(in) :: b_in(3,3)
  real(c_double), value :: dkmax2_in
  integer(c_int), value :: cap
  integer(c_int) :: out_n, nga_o(cap), ngb_o(cap), ngc_o(cap), kcsma_o(3)
  real(c_double) :: gx_o(cap), gy_o(cap), gz_o(cap), recnrm_o(cap)
  integer :: nfile(2), i
  real*8 :: b(3,3)
  nfile = 6
  b = b_in
  dkmax2 = dkmax2_in
  call clmrcp(nfile, 0, 1, b, 1, 1, 1, 1, .false., 0.d0, .false., 1.2d0, .false.)
  out_n = numk2
  kcsma_o(1) = KCSMA1
  kcsma_o(2) = KCSMA2
  kcsma_o(3) = KCSMA3
  if (numk2 > cap) return
  do i = 1, numk2
    nga_o(i) = nga(i)
    ngb_o(i) = ngb(i)
    ngc_o(i) = ngc(i)
    gx_o(i) = gx(i)
    gy_o(i) = gy(i)
    gz_o(i) = gz(i)
    recnrm_o(i) = recnrm(i)
  end do
end subroutine

subroutine qxref_rnd00(seed, n, out) bind(C, name="qxref_rnd00")
  use iso_c_binding
  implicit none
  real(c_double) :: seed
  integer(c_int), value :: n
  real(c_double) :: out(n)
  integer :: i
  real*8 :: rn, s
  s = seed
  do i = 1, n
    call rnd00(rn, s)
    out(i) = rn
  end do
  seed = s
end subroutine

subroutine qxref_rseigqr(a, n, w) bind(C, name="qxref_rseigqr")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  real(c_double) :: a(n,n), w(n)
  real*8, allocatable :: w1r(:), w1i(:), work(:)
  character*5 :: check
  allocate(w1r(n), w1i(n), work(n))
  check = 'OK   '
  call RsEIGQR(a, w, n, n, 1.0d-15, 1, 2, check, w1r, w1i, work)
  deallocate(w1r, w1i, work)
end subroutine

subroutine qxref_icholesky(s, n) bind(C, name="qxref_icholesky")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  real(c_double) :: s(n,n)
  real*8, allocatable :: work(:)
  allocate(work(n))
  call rsdecmp(s, n, n)
  call rsinvcd(s, work, n, n)
  deallocate(work)
end subroutine

subroutine qxref_schmidt2(cgjr, nplwex, nband, ok) bind(C, name="qxref_schmidt2")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nplwex, nband
  real(c_double) :: cgjr(nplwex, nband)
  integer(c_int) :: ok
  integer :: nfile(2), npnod, i, nbncnt(1), nbndsp(1), lcnt(1), ldsp(1), mcnt(1), mdsp(1), jd(1)
  integer, allocatable :: iod(:)
  real*8, allocatable :: prod(:), prod2(:), dbuf(:), dmtrxr(:,:)
  real*8 :: ct0, bijr(1), bm1r(1), rvol
  integer :: nhk1(1), nhk2(1)
  logical :: lvandi(1), lc
  nfile = 6
  npnod = nplwex/2
  allocate(iod(nband), prod(nband), prod2(nband), dbuf(nplwex), dmtrxr(nband, nband))
  do i = 1, nband
    iod(i) = i
  end do
  nbncnt(1) = nband; nbndsp(1) = 0
  ct0 = 0.d0; rvol = 1.d0; nhk1 = 1; nhk2 = 1; lvandi = .false.
  call setlstop(nfile, 0, 1, .false.)
  call schmidt2(nfile, 0, 1, ct0, .false., cgjr, 1, npnod, npnod, nband, 1, nband, nband, &
&     nbncnt, nbndsp, iod, iod, prod, prod2, dbuf, .false., .false., .false., .false., &
&     dmtrxr, nband, 1, lcnt, ldsp, mcnt, mdsp, jd, bijr, bm1r, rvol, 1, nhk1, nhk2, 1, lvandi)
  call getlcholesky(nfile, 0, 1, lc)
  ok = 0
  if (lc) ok = 1
  deallocate(iod, prod, prod2, dbuf, dmtrxr)
end subroutine

subroutine qxref_hmatrix2(cgjr, rhcr, nplwex, nband, dmtrxr, eig) bind(C, name="qxref_hmatrix2")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nplwex, nband
  real(c_double) :: cgjr(nplwex, nband), rhcr(nplwex, nband), dmtrxr(nband, nband), eig(nband)
  integer :: nfile(2), nbncnt(1), nbndsp(1), npnod
  integer, allocatable :: iod(:)
  real*8, allocatable :: prod(:), prodr(:), bx0r(:)
  nfile = 6
  npnod = nplwex/2
  allocate(iod(nband), prod(nband), prodr(nband), bx0r(nband))
  nbncnt(1) = nband; nbndsp(1) = 0
  call hmatrix2(nfile, 0, 1, eig, cgjr, rhcr, 1, npnod, npnod, nband, 1, nband, nband, &
&     nbncnt, nbndsp, iod, bx0r, dmtrxr, nband, prod, prodr)
  deallocate(iod, prod, prodr, bx0r)
end subroutine

subroutine qxref_untrycore(gdcr, nplwex, nband, u) bind(C, name="qxref_untrycore")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nplwex, nband
  real(c_double) :: gdcr(nplwex, nband), u(nband, nband)
  integer :: nfile(2), nbncnt(1), nbndsp(1), npnod, iblock
  real*8, allocatable :: prod(:), prodr(:), w1i(:), tmpr(:,:)
  nfile = 6
  npnod = nplwex/2
  iblock = int(1.d+07/8/nband)
  if (iblock < 100) iblock = 100
  allocate(prod(nband), prodr(nband), w1i(nband), tmpr(iblock, nband))
  nbncnt(1) = nband; nbndsp(1) = 0
  call untrycore(nfile, 0, 1, gdcr, 1, npnod, npnod, nband, 1, nband, nband, nbncnt, nbndsp, &
&     prod, prodr, w1i, tmpr, iblock, u, nband)
  deallocate(prod, prodr, w1i, tmpr)
end subroutine

subroutine qxref_prpcg(ekib, dkgnrm, nplwex, prcd) bind(C, name="qxref_prpcg")
  use iso_c_binding
  implicit none
  real(c_double), value :: ekib
  integer(c_int), value :: nplwex
  real(c_double) :: dkgnrm(nplwex), prcd(nplwex)
  integer :: nfile(2)
  real*8 :: ek(1)
  nfile = 6
  ek(1) = ekib
  call prpcg(nfile, 0, 1, 1, prcd, ek, dkgnrm, nplwex)
end subroutine

subroutine qxref_ekinetic2(cgjr, rhcr, nplwex, nband, occ, dkgnrm, ekib, ekint) &
&   bind(C, name="qxref_ekinetic2")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nplwex, nband
  real(c_double) :: cgjr(nplwex, nband), rhcr(nplwex, nband), occ(nband), dkgnrm(nplwex)
  real(c_double) :: ekib(nband), ekint
  integer :: nfile(2), i
  integer, allocatable :: iod(:)
  real*8 :: ek
  nfile = 6
  allocate(iod(nband))
  do i = 1, nband
    iod(i) = i
  end do
  ek = 0.d0
  call ekinetic2(nfile, 0, 1, ek, ekib, occ, cgjr, rhcr, nplwex, nplwex/2 - 1, &
&     nband, 1, nband, nband, iod, dkgnrm, 1)
  ekint = ek
  deallocate(iod)
end subroutine

! KB state: one species per atom so that lmx and vclalp may vary per atom.
subroutine qxref_kb_setup(eqlmr, nplwex, lmmax_in, natom, lmx_atom, vclalp_atom, clalp_atom, &
&   nbmax) bind(C, name="qxref_kb_setup")
  use iso_c_binding
  use nlpp_variables, only: lmx, clalp, lmmax
  use nl_variables, only: eqlmr_, vclalp, slmir, zzzr, zzzi, lbigmm2
  implicit none
  integer(c_int), value :: nplwex, lmmax_in, natom, nbmax
  real(c_double) :: eqlmr(nplwex, lmmax_in, natom)
  integer(c_int) :: lmx_atom(natom)
  real(c_double) :: vclalp_atom(lmmax_in, natom), clalp_atom(lmmax_in, natom)
  if (allocated(eqlmr_)) deallocate(eqlmr_)
  if (allocated(vclalp)) deallocate(vclalp)
  if (allocated(clalp)) deallocate(clalp)
  if (allocated(lmx)) deallocate(lmx)
  if (allocated(slmir)) deallocate(slmir)
  if (allocated(zzzr)) deallocate(zzzr)
  if (allocated(zzzi)) deallocate(zzzi)
  lmmax = lmmax_in
  allocate(eqlmr_(nplwex, lmmax_in, natom), vclalp(lmmax_in, natom), clalp(lmmax_in, natom))
  allocate(lmx(natom), slmir(lmmax_in, natom, nbmax), zzzr(0:nplwex/2), zzzi(0:nplwex/2))
  eqlmr_ = eqlmr
  vclalp = vclalp_atom
  clalp = clalp_atom
  lmx = lmx_atom
  slmir = 0.d0
  lbigmm2 = .true.
end subroutine

subroutine qxref_hckbpp_g(cgjr, rhcr, nplwex, nband, natom, slmir_o) bind(C, name="qxref_hckbpp_g")
  use iso_c_binding
  use nlpp_variables, only: lmmax
  use nl_variables, only: slmir
  implicit none
  integer(c_int), value :: nplwex, nband, natom
  real(c_double) :: cgjr(nplwex, nband), rhcr(nplwex, nband)
  real(c_double) :: slmir_o(lmmax, natom, nband)
  integer :: nfile(2), nhk(natom), i
  logical :: lkb(natom), lkg(natom)
  real*8 :: yc(1), ys(1)
  nfile = 6
  do i = 1, natom
    nhk(i) = i
  end do
  lkb = .true.
  lkg = .false.
  call hckbpp_g(nfile, 0, 1, cgjr, rhcr, nplwex, nplwex/2 - 1, nband, 1, nband, nband, &
&     natom, nhk, nhk, natom, lkb, lkg, yc, ys, 0)
  slmir_o = slmir(:, :, 1:nband)
end subroutine

subroutine qxref_hckbpp_ib_g(cgjr, rhcr, nplwex, natom) bind(C, name="qxref_hckbpp_ib_g")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nplwex, natom
  real(c_double) :: cgjr(nplwex), rhcr(nplwex)
  integer :: nfile(2), nhk(natom), i
  logical :: lkb(natom), lkg(natom)
  real*8 :: yc(1), ys(1)
  nfile = 6
  do i = 1, natom
    nhk(i) = i
  end do
  lkb = .true.
  lkg = .false.
  call hckbpp_ib_g(nfile, 0, 1, cgjr, rhcr, nplwex, nplwex/2 - 1, &
&     natom, nhk, nhk, natom, lkb, lkg, yc, ys, 0)
end subroutine

function qxref_enkbpp_g2(nband, natom, occ, rvol) bind(C, name="qxref_enkbpp_g2")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nband, natom
  real(c_double) :: occ(nband)
  real(c_double), value :: rvol
  real(c_double) :: qxref_enkbpp_g2
  integer :: nfile(2), nhk(natom), i, iod(nband)
  logical :: lkb(natom), lkg(natom)
  real*8 :: e, rv
  nfile = 6
  do i = 1, natom
    nhk(i) = i
  end do
  do i = 1, nband
    iod(i) = i
  end do
  lkb = .true.
  lkg = .false.
  e = 0.d0
  rv = rvol
  call enkbpp_g2(nfile, 0, 1, e, nband, 1, nband, nband, occ, iod, natom, nhk, nhk, natom, &
&     lkb, lkg, rv)
  qxref_enkbpp_g2 = e
end function

! stscfn_pw (Sources/pwlda.f90:5711): structure factors for natom atoms, with
! the recurrence work arrays sized here; lnlpp_g = .true. so ycos/ysin are set.
subroutine qxref_stscfn_pw(ratm, hcell, nd1v, nd1vks, lclust_i, natom, kmax1, kmax2, kmax3, &
&   nplwcs, nplw5, nga, ngb, ngc, ycos, ysin) bind(C, name="qxref_stscfn_pw")
  use iso_c_binding
  implicit none
  integer(c_int), value :: lclust_i, natom, kmax1, kmax2, kmax3, nplwcs, nplw5
  real(c_double) :: ratm(3, natom), hcell(3, 3)
  integer(c_int) :: nd1v(3), nd1vks(3), nga(0:nplw5), ngb(0:nplw5), ngc(0:nplw5)
  real(c_double) :: ycos(0:nplwcs, natom), ysin(0:nplwcs, natom)
  real*8, allocatable :: dcq1(:,:), dcq2(:,:), dcq3(:,:), dsq1(:,:), dsq2(:,:), dsq3(:,:)
  integer :: nfile(2), nd1(3), nd2(3)
  logical :: lcl
  nfile = 6
  nd1 = nd1v
  nd2 = nd1vks
  lcl = lclust_i /= 0
  allocate(dcq1(-kmax1:kmax1, natom), dsq1(-kmax1:kmax1, natom))
  allocate(dcq2(-kmax2:kmax2, natom), dsq2(-kmax2:kmax2, natom))
  allocate(dcq3(-kmax3:kmax3, natom), dsq3(-kmax3:kmax3, natom))
  call stscfn_pw(nfile, 0, 1, ycos, ysin, dcq1, dcq2, dcq3, dsq1, dsq2, dsq3, &
&     kmax1, kmax2, kmax3, natom, nplwcs, nplw5, nga, ngb, ngc, &
&     lcl, ratm, hcell, nd1, nd2, .true., .false., .false., .false.)
  deallocate(dcq1, dsq1, dcq2, dsq2, dcq3, dsq3)
end subroutine

! setnlc_g (Sources/nlkbpp.f90:1236) for ntype types: the nlpp_variables /
! nl_variables state it reads (tables of the nonlocal term, angular factors,
! abgvec = sqrt(recnrm)) is set from the arguments, then the reference routine
! runs unchanged; eqlmr_ and vclalp are copied out.
subroutine qxref_setnlc_g(nplw, ntype, nhk1, nhk2, natom, lkbppi_i, lking_i, ycos, ysin, &
&   nplwcs, rvol, lmmax_in, lnref_in, lmx_t, clalp_t, tbqlm_t, tbqlma_t, dlqlm_t, itola_t, &
&   itolm_t, itolmf_t, litoll_i, nqlmmx_in, qlmr_t, qlmi_t, recnrm, eq_out, vcl_out) &
&   bind(C, name="qxref_setnlc_g")
  use iso_c_binding
  use nlpp_parameters, only: lmmax, lnref, nqlmmx
  use nlpp_variables, only: lmx, clalp, tbqlm, tbqlma, dlqlm, itola, itolm, itolmf, litoll, &
&     qlmr, qlmi, abgvec, mxqlm
  use nl_variables, only: eqlmr_, vclalp, zzzr, zzzi, lbigmm2
  implicit none
  integer(c_int), value :: nplw, ntype, natom, nplwcs, lmmax_in, lnref_in, nqlmmx_in
  integer(c_int) :: nhk1(ntype), nhk2(ntype), lkbppi_i(ntype), lking_i(ntype), lmx_t(ntype)
  real(c_double) :: ycos(0:nplwcs, natom), ysin(0:nplwcs, natom)
  real(c_double), value :: rvol
  real(c_double) :: clalp_t(lmmax_in, ntype), dlqlm_t(ntype)
  real(c_double) :: tbqlm_t(0:550, lnref_in, ntype), tbqlma_t(0:550, lnref_in, ntype)
  integer(c_int) :: itola_t(lmmax_in, ntype), itolm_t(lmmax_in, ntype), itolmf_t(lmmax_in, ntype)
  integer(c_int) :: litoll_i(lmmax_in, ntype)
  real(c_double) :: qlmr_t(0:nplw, nqlmmx_in), qlmi_t(0:nplw, nqlmmx_in), recnrm(0:nplw)
  real(c_double) :: eq_out(2 * (nplw + 1), lmmax_in, natom), vcl_out(lmmax_in, ntype)
  integer :: nfile(2), n1(ntype), n2(ntype), i
  logical :: lkb(ntype), lkg(ntype)
  real*8 :: rv
  nfile = 6
  lmmax = lmmax_in
  lnref = lnref_in
  nqlmmx = nqlmmx_in
  if (allocated(lmx)) deallocate(lmx)
  if (allocated(clalp)) deallocate(clalp)
  if (allocated(tbqlm)) deallocate(tbqlm)
  if (allocated(tbqlma)) deallocate(tbqlma)
  if (allocated(dlqlm)) deallocate(dlqlm)
  if (allocated(itola)) deallocate(itola)
  if (allocated(itolm)) deallocate(itolm)
  if (allocated(itolmf)) deallocate(itolmf)
  if (allocated(litoll)) deallocate(litoll)
  if (allocated(qlmr)) deallocate(qlmr)
  if (allocated(qlmi)) deallocate(qlmi)
  if (allocated(abgvec)) deallocate(abgvec)
  if (allocated(eqlmr_)) deallocate(eqlmr_)
  if (allocated(vclalp)) deallocate(vclalp)
  if (allocated(zzzr)) deallocate(zzzr)
  if (allocated(zzzi)) deallocate(zzzi)
  allocate(lmx(ntype), clalp(lmmax, ntype), dlqlm(ntype))
  allocate(tbqlm(0:mxqlm, lnref, ntype), tbqlma(0:mxqlm, lnref, ntype))
  allocate(itola(lmmax, ntype), itolm(lmmax, ntype), itolmf(lmmax, ntype), litoll(lmmax, ntype))
  allocate(qlmr(0:nplw, nqlmmx, 1, 1), qlmi(0:nplw, nqlmmx, 1, 1), abgvec(0:nplw, 1))
  allocate(eqlmr_(2 * (nplw + 1), lmmax, natom), vclalp(lmmax, ntype))
  allocate(zzzr(0:nplw), zzzi(0:nplw))
  lmx = lmx_t
  clalp = clalp_t
  dlqlm = dlqlm_t
  tbqlm = tbqlm_t
  tbqlma = tbqlma_t
  itola = itola_t
  itolm = itolm_t
  itolmf = itolmf_t
  litoll = litoll_i /= 0
  qlmr(:, :, 1, 1) = qlmr_t
  qlmi(:, :, 1, 1) = qlmi_t
  do i = 0, nplw
    abgvec(i, 1) = sqrt(recnrm(i))
  end do
  eqlmr_ = 0.d0
  vclalp = 0.d0
  lbigmm2 = .true.
  n1 = nhk1
  n2 = nhk2
  lkb = lkbppi_i /= 0
  lkg = lking_i /= 0
  rv = rvol
  call setnlc_g(nfile, 0, 1, nplw, ntype, n1, n2, natom, lkb, lkg, ycos, ysin, nplwcs, rv)
  eq_out = eqlmr_
  vcl_out = vclalp
end subroutine

! kbnlpp_frc2 (Sources/nlkbpp.f90:1878), no stress, one type per atom as in
! qxref_kb_setup; slmir is the module state the last qxref_hckbpp_g left for
! the same cgjr.  fnlc is accumulated into, as in the reference.
subroutine qxref_kbnlpp_frc2(cgjr, nplwex, nband, natom, occ, gx, gy, gz, fnlc) &
&   bind(C, name="qxref_kbnlpp_frc2")
  use iso_c_binding
  use nlpp_parameters, only: lmmax
  use nl_variables, only: fslxr, fslyr, fslzr, dslmir, vclweg
  implicit none
  integer(c_int), value :: nplwex, nband, natom
  real(c_double) :: cgjr(nplwex, nband), occ(nband), fnlc(3, natom)
  real(c_double) :: gx(0:nplwex/2 - 1), gy(0:nplwex/2 - 1), gz(0:nplwex/2 - 1)
  integer :: nfile(2), nhk(natom), i, iod(nband)
  logical :: lkb(natom), lkg(natom)
  real*8 :: yc(1), ys(1), str(6)
  nfile = 6
  do i = 1, natom
    nhk(i) = i
  end do
  do i = 1, nband
    iod(i) = i
  end do
  lkb = .true.
  lkg = .false.
  str = 0.d0
  ! work arrays of the VECTOR branch, lmx(it) * nbnod long
  if (allocated(fslxr)) deallocate(fslxr, fslyr, fslzr, dslmir, vclweg)
  allocate(fslxr(lmmax * nband), fslyr(lmmax * nband), fslzr(lmmax * nband))
  allocate(dslmir(lmmax * nband), vclweg(lmmax * nband))
  call kbnlpp_frc2(nfile, 0, 1, fnlc, str, .false., cgjr, nplwex, nplwex/2 - 1, gx, gy, gz, &
&     nband, 1, nband, nband, occ, iod, natom, nhk, nhk, natom, lkb, lkg, yc, ys, 0)
end subroutine

! xclda (Sources/vxc.f90:961): LDA exchange-correlation on n mesh points
subroutine qxref_xclda(rho, rhocore, n, rdelv, vexc, eeexc, e4) bind(C, name="qxref_xclda")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  real(c_double) :: rho(n), rhocore(n), vexc(n), eeexc(n), e4(4)
  real(c_double), value :: rdelv
  real*8 :: a, b, c, d, rd
  rd = rdelv
  call xclda(rho, rhocore, n, rd, vexc, eeexc, a, b, c, d)
  e4(1) = a
  e4(2) = b
  e4(3) = c
  e4(4) = d
end subroutine

! INVERS (Sources/funcs.f90:5): cofactor inverse of an n x n matrix
subroutine qxref_invers(w, wi, n, ierr) bind(C, name="qxref_invers")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  real(c_double) :: w(n, n), wi(n, n)
  integer(c_int) :: ierr
  real*8, allocatable :: work(:,:)
  integer, allocatable :: ip(:)
  integer :: ie
  allocate(work(n, n), ip(n))
  call INVERS(w, wi, n, work, ip, n, ie)
  ierr = ie
  deallocate(work, ip)
end subroutine

! flocal_g (Sources/nlkbpp.f90:1037) without stress on one plane-wave rank:
! svco of lopp_variables and the rank layout of planewave_decomp_variables set
! here; DCQ*/DSQ* (2 kmax + 1, natom) per axis
subroutine qxref_flocal_g(ntype, nhk1, nhk2, iatmpt, llk, natom, nplw5, nga, ngb, ngc, gx, gy,  &
&    gz, rhgr, svco_in, kmax1, kmax2, kmax3, dcq1, dcq2, dcq3, dsq1, dsq2, dsq3, floc)  &
&    bind(C, name="qxref_flocal_g")
  use iso_c_binding
  use lopp_variables
  use planewave_decomp_variables
  implicit none
  integer(c_int), value :: ntype, natom, nplw5, kmax1, kmax2, kmax3
  integer(c_int) :: nhk1(ntype), nhk2(ntype), iatmpt(natom), llk(ntype)
  integer(c_int) :: nga(0:nplw5), ngb(0:nplw5), ngc(0:nplw5)
  real(c_double) :: gx(0:nplw5), gy(0:nplw5), gz(0:nplw5), rhgr(2 * (nplw5 + 1))
  real(c_double) :: svco_in(nplw5 + 1, ntype)
  real(c_double) :: dcq1(-kmax1:kmax1, natom), dcq2(-kmax2:kmax2, natom), dcq3(-kmax3:kmax3, natom)
  real(c_double) :: dsq1(-kmax1:kmax1, natom), dsq2(-kmax2:kmax2, natom), dsq3(-kmax3:kmax3, natom)
  real(c_double) :: floc(3, natom)
  integer :: nf(2), na, ni, n5, ni5, nh1(ntype), nh2(ntype), iat(natom), it1(natom)
  integer :: k1, k2, k3, it
  integer, allocatable :: g1(:), g2(:), g3(:)
  logical :: llking(ntype)
  real*8 :: st(3, 3), elclv
  if (allocated(svco)) deallocate(svco)
  allocate(svco(nplw5 + 1, ntype))
  svco = svco_in
  myid_pw = 0
  nodes_pw = 1
  nplw5nod1 = 1
  nplw5nod = nplw5 + 1
  nplw5nod2 = nplw5 + 1
  nf = 6
  na = natom
  ni = natom
  n5 = nplw5
  ni5 = 2 * (nplw5 + 1)
  nh1 = nhk1
  nh2 = nhk2
  iat = iatmpt
  it1 = 1
  allocate(g1(0:nplw5), g2(0:nplw5), g3(0:nplw5))
  g1 = nga
  g2 = ngb
  g3 = ngc
  do it = 1, ntype
    llking(it) = llk(it) /= 0
  end do
  k1 = kmax1
  k2 = kmax2
  k3 = kmax3
  st = 0.d0
  elclv = 0.d0
  call flocal_g(nf, 0, 1, floc, st, .false., ntype, nh1, nh2, it1, iat, na, ni, llking, n5, ni5,  &
&    n5, g1, g2, g3, gx, gy, gz, rhgr, elclv, k1, k2, k3, dcq1, dcq2, dcq3, dsq1, dsq2, dsq3,  &
&    .false.)
  deallocate(g1, g2, g3)
end subroutine

! frcrec (Sources/bulk.f90:1880, VECTOR form) without stress: the Ewald G list,
! its axis tables per ion and the parameters of para_ewald set here
subroutine qxref_frcrec(ntype, nhk1, nhk2, zv, nel, natom, nion, iatoit, iatmpt, numk2_, nga_,  &
&    ngb_, ngc_, gx_, gy_, gz_, recnrm_, recnrmex_, gamma_, volume_, km1, km2, km3, c1, c2, c3,  &
&    s1, s2, s3, fclm, ecolmb, atom_ecolmb) bind(C, name="qxref_frcrec")
  use iso_c_binding
  use para_ewald
  implicit none
  integer(c_int), value :: ntype, nel, natom, nion, numk2_, km1, km2, km3
  integer(c_int) :: nhk1(ntype), nhk2(ntype), iatoit(natom), iatmpt(nion)
  integer(c_int) :: nga_(numk2_), ngb_(numk2_), ngc_(numk2_)
  real(c_double) :: zv(ntype), gx_(numk2_), gy_(numk2_), gz_(numk2_), recnrm_(numk2_)
  real(c_double) :: recnrmex_(0:numk2_)
  real(c_double), value :: gamma_, volume_
  real(c_double) :: c1(-km1:km1, nion), c2(-km2:km2, nion), c3(-km3:km3, nion)
  real(c_double) :: s1(-km1:km1, nion), s2(-km2:km2, nion), s3(-km3:km3, nion)
  real(c_double) :: fclm(3, natom), ecolmb, atom_ecolmb(natom)
  integer :: nt, nl, na, ni, h1(ntype), h2(ntype), iat(natom), imp(nion)
  real*8 :: z(ntype), st(3, 3), e
  real*8, allocatable :: dyc(:), dys(:)
  if (allocated(nga)) deallocate(nga, ngb, ngc, gx, gy, gz, recnrm, recnrmex, sumtmp, buff)
  if (allocated(DCQ1)) deallocate(DCQ1, DCQ2, DCQ3, DSQ1, DSQ2, DSQ3)
  numk2 = numk2_
  allocate(nga(numk2), ngb(numk2), ngc(numk2), gx(numk2), gy(numk2), gz(numk2), recnrm(numk2),  &
&    recnrmex(0:numk2), sumtmp(4 * numk2), buff(4 * numk2))
  nga = nga_
  ngb = ngb_
  ngc = ngc_
  gx = gx_
  gy = gy_
  gz = gz_
  recnrm = recnrm_
  recnrmex = recnrmex_
  KCSMA1 = km1
  KCSMA2 = km2
  KCSMA3 = km3
  allocate(DCQ1(-km1:km1, nion), DCQ2(-km2:km2, nion), DCQ3(-km3:km3, nion),  &
&    DSQ1(-km1:km1, nion), DSQ2(-km2:km2, nion), DSQ3(-km3:km3, nion))
  DCQ1 = c1
  DCQ2 = c2
  DCQ3 = c3
  DSQ1 = s1
  DSQ2 = s2
  DSQ3 = s3
  gamma = gamma_
  volume = volume_
  pi = acos(-1.d0)
  allocate(dyc(3 * natom), dys(3 * natom))
  nt = ntype
  nl = nel
  na = natom
  ni = nion
  h1 = nhk1
  h2 = nhk2
  iat = iatoit
  imp = iatmpt
  z = zv
  st = 0.d0
  e = ecolmb
  call frcrec(fclm, st, nt, h1, h2, z, nl, e, atom_ecolmb, iat, imp, na, ni, dyc, dys, .false.)
  ecolmb = e
  deallocate(dyc, dys)
end subroutine

! roteig2 (Sources/sbalin.f90:772) on one rank: cur(2 npnod, nband), prev(2
! npnod, nband, nprvmx) rotated in place
subroutine qxref_roteig(cur, prev, npnod, nband, nprvmx, kpwvx) bind(C, name="qxref_roteig")
  use iso_c_binding
  implicit none
  integer(c_int), value :: npnod, nband, nprvmx, kpwvx
  real(c_double) :: cur(2 * npnod, nband), prev(2 * npnod, nband, nprvmx)
  integer :: nf(2), nb, nn, np, nbx, ib, nprv, kpw, cnt(1), dsp(1), iw(1)
  real*8, allocatable :: prod(:), prodr(:), aij(:,:), bij(:,:), bm1(:,:), bx0(:,:)
  real*8, allocatable :: w1r(:), w1i(:), ew1(:), tmpr(:,:)
  real*8 :: rv, tc
  nf = 6
  nb = nband
  nn = npnod
  np = npnod
  nbx = nband
  ib = 64
  nprv = nprvmx
  kpw = kpwvx
  cnt(1) = nband
  dsp(1) = 0
  iw(1) = 0
  rv = 1.d0
  tc = 0.d0
  allocate(prod(2 * nband + 2 * npnod), prodr(2 * nband + 2 * npnod), aij(nband, nband),  &
&    bij(nband, nband), bm1(nband, nband), bx0(nband, nband), w1r(2 * nband), w1i(2 * nband),  &
&    ew1(nband), tmpr(ib, nband))
  call roteig2(nf, 0, 1, cur, prev, nb, 1, nb, nb, cnt, dsp, 1, np, nn, prod, prodr, .false.,  &
&    rv, 1, nprv, kpw, aij, bij, bm1, bx0, nbx, w1r, w1i, ew1, tmpr, ib, 1, iw, iw, iw, iw, iw,  &
&    tc, .false.)
  deallocate(prod, prodr, aij, bij, bm1, bx0, w1r, w1i, ew1, tmpr)
end subroutine

! set_coef_ASPC (Sources/sbalin.f90:77): coefaspc of module aspc, kaspc^2
subroutine qxref_coef_aspc(ihest, coef) bind(C, name="qxref_coef_aspc")
  use iso_c_binding
  use aspc
  implicit none
  integer(c_int), value :: ihest
  real(c_double) :: coef(*)
  integer :: nf(2), ich, ihe, nb, kx
  real*8 :: corr
  nf = 6
  ich = 0
  ihe = ihest
  nb = 1
  kx = 1
  corr = 0.d0
  if (allocated(coefaspc)) deallocate(coefaspc)
  if (allocated(pspmat)) deallocate(pspmat)
  if (allocated(omegaaspc)) deallocate(omegaaspc)
  call set_coef_ASPC(nf, 0, 1, ich, ihe, .false., .true., nb, corr, .false., kx)
  coef(1:kaspc * kaspc) = reshape(coefaspc, (/kaspc * kaspc/))
end subroutine

! one-rank communicator state for the reference's file writers
subroutine qxref_serial_world()
  use communicator
  implicit none
  myid = 0
  nodes = 1
  myid_qm = 0
  nodes_qm = 1
  myid_qm_un = 0
  nodes_qm_un = 1
  myid_kd = 0
  nodes_kd = 1
  nkd = 0
  myid_lr = 0
  nodes_lr = 1
  myid_pw = 0
  nodes_pw = 1
end subroutine

! svwfns (Sources/savedd.f90:2441) for one spin, ncscale = 1: writes gdcr
subroutine qxref_svwfns(fname, nlen, gdcr, nplwex, nplw, nband, eig, occ, lsr)  &
&    bind(C, name="qxref_svwfns")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nlen, nplwex, nplw, nband, lsr
  character(kind=c_char) :: fname(nlen)
  real(c_double) :: gdcr(nplwex, nband), eig(nband), occ(nband)
  character(len=256) :: fn
  integer :: nf(2), i, nb, nex, np, cnt(1), dsp(1), nsp, nspn
  real*8, allocatable :: work(:,:), dummy(:)
  logical :: l8
  real*8 :: wv(2)
  fn = ' '
  do i = 1, nlen
    fn(i:i) = fname(i)
  end do
  call qxref_serial_world()
  l8 = lsr /= 0
  nf = 6
  call set_lsreal8(nf, 0, 1, l8)
  nb = nband
  nex = nplwex
  np = nplw
  cnt(1) = nband
  dsp(1) = 0
  nsp = 1
  nspn = 1
  allocate(work(nplwex, nband), dummy(1))
  call svwfns(nf, 1, trim(fn), work, nex, np, nb, 1, nb, nb, cnt, dsp, eig, dummy, occ, dummy,  &
&    gdcr, l8, .false., nspn, dummy, nsp, .false., 1, wv)
  deallocate(work, dummy)
end subroutine

! svacon (Sources/savedd.f90:3058), file version -2
subroutine qxref_svacon(fname, nlen, ntype, nhk1, nhk2, nstepcg, nstepmd, lclust, ratm, ifmd,  &
&    frc, prevr, hcell, natom) bind(C, name="qxref_svacon")
  use iso_c_binding
  implicit none
  integer(c_int), value :: nlen, ntype, nstepcg, nstepmd, lclust, ifmd, natom
  character(kind=c_char) :: fname(nlen)
  integer(c_int) :: nhk1(ntype), nhk2(ntype)
  real(c_double) :: ratm(3, natom), frc(3, natom), prevr(3, natom, 3), hcell(3, 3)
  character(len=256) :: fn
  integer :: nf(2), i, nt, na, ncg, nmd, imd, h1(ntype), h2(ntype)
  logical :: lc
  fn = ' '
  do i = 1, nlen
    fn(i:i) = fname(i)
  end do
  call qxref_serial_world()
  nf = 6
  nt = ntype
  na = natom
  ncg = nstepcg
  nmd = nstepmd
  imd = ifmd
  h1 = nhk1
  h2 = nhk2
  lc = lclust /= 0
  call svacon(nf, 1, trim(fn), prevr, imd, ncg, nmd, ratm, frc, nt, h1, h2, lc, na, hcell)
end subroutine

! ---- velocity-Verlet updates of the MD loop (Sources/remd.f90) -------------
subroutine qxref_md_kick(mode, x, v, is, n, ntype, lfix) bind(C, name="qxref_md_kick")
  ! mode 0: updtps (x += v), 1: vkick (x += v as v += a)
  use iso_c_binding
  implicit none
  integer(c_int), value :: mode, n, ntype
  real(c_double) :: x(3 * n), v(3 * n)
  integer(c_int) :: is(n), lfix(ntype)
  integer :: nf(2), nn, nt, iss(n)
  logical :: lf(ntype)
  nf = 6
  nn = n
  nt = ntype
  iss = is
  lf = lfix /= 0
  if (mode == 0) then
    call updtps(nf, 0, 1, x, v, iss, nn, nt, lf)
  else
    call vkick(nf, 0, 1, x, v, iss, nn, nt, lf)
  end if
end subroutine

subroutine qxref_updtps_pbc(x, n) bind(C, name="qxref_updtps_pbc")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n
  real(c_double) :: x(3 * n)
  integer :: nf(2), nn, iss(n)
  nf = 6
  nn = n
  iss = 1
  call updtps_pbc(nf, 0, 1, x, iss, nn, 1)
end subroutine

subroutine qxref_normalized_a(n, ntype, a, is, h, acon) bind(C, name="qxref_normalized_a")
  use iso_c_binding
  implicit none
  integer(c_int), value :: n, ntype
  real(c_double) :: a(3 * n), h(3, 3, 0:1), acon(ntype)
  integer(c_int) :: is(n)
  integer :: nf(2), nn, nt, iss(n)
  nf = 6
  nn = n
  nt = ntype
  iss = is
  call normalized_a(nf, 0, 1, nn, nt, a, iss, h, acon)
end subroutine

subroutine qxref_rciprl(a, b, v) bind(C, name="qxref_rciprl")
  use iso_c_binding
  implicit none
  real(c_double) :: a(3, 3), b(3, 3), v
  call rciprl(a, b, v)
end subroutine

!-----------------------------------------------------------------------
! Norm-conserving pseudopotential of one species, set up as the serial
! reference's initialisation does it: read_pp (pptmpp, mxl_alloc,
! set_nlpp_parameters), then setpp's steps for a KB pseudopotential in
! reciprocal space without core correction or smoothing (rdtmpp, lotmpp,
! lotmpp_g, nlkbpp_g; Sources/input.f90:5240-5330, Sources/ppkb.f90:816-1120).
! The files are opened by the reference's openpp under the current
! directory (control/NCPP/<element>/ for LDA).  Inputs: the wave-function
! sphere gx, gy, gz, recnrm(0:nplw) and the density sphere recnrm5,
! recnrmex(0:nplw5).  Outputs, in the caller's capacities (lmcap projector
! rows, lncap radial rows, nqcap angular rows): the nlpp_variables tables
! setqlm reads, svco(0:nplw5) of lopp_variables and lotmpp_g's radial table
! rhocr/rhocra(0:1000).  Call once per process: nlkbpp_g keeps the table
! range of its first call.
!-----------------------------------------------------------------------
subroutine qxref_ncpp(zatom_in, jgga_in, ecutdens_in, nplw, gx, gy, gz, recnrm, nplw5, recnrm5, &
&   recnrmex, lmcap, lncap, nqcap, dims, zv_o, clalp_o, tbqlm_o, tbqlma_o, dlqlm_o, itola_o, &
&   itolm_o, itolmf_o, litoll_o, qlmr_o, qlmi_o, svco_o, rhocr_o, rhocra_o) &
&   bind(C, name="qxref_ncpp")
  use iso_c_binding
  use param, only: ntype, mxl, jgga, lrela, mx1, mx1loc, pwscale
  use param_atom, only: zatom, lkbpp, lvand, lkbppi, lvandi, lvflag, lking, llking, rctflc, &
&     lpcci, rpcc, lpking, rpking, gpkgmax, gpkgexct, zv, lmax, lclno, lchk, lsomax
  use constants, only: ms, aname
  use nlpp_parameters, only: lmmax, lnref, nqlmmx
  use nlpp_variables, only: lmx, clalp, tbqlm, tbqlma, dlqlm, itola, itolm, itolmf, litoll, &
&     qlmr, qlmi, mxqlm
  use lopp_variables, only: svco
  use for_lotmpp_g, only: rhocr, rhocra, mxqab_lc => mxqab
  use outfile, only: loutfile
  implicit none
  real(c_double), value :: zatom_in, ecutdens_in
  integer(c_int), value :: jgga_in, nplw, nplw5, lmcap, lncap, nqcap
  real(c_double) :: gx(0:nplw), gy(0:nplw), gz(0:nplw), recnrm(0:nplw)
  real(c_double) :: recnrm5(0:nplw5), recnrmex(0:nplw5)
  integer(c_int) :: dims(8)
  real(c_double) :: zv_o, clalp_o(lmcap), dlqlm_o
  real(c_double) :: tbqlm_o(0:550, lncap), tbqlma_o(0:550, lncap)
  integer(c_int) :: itola_o(lmcap), itolm_o(lmcap), itolmf_o(lmcap), litoll_o(lmcap)
  real(c_double) :: qlmr_o(0:nplw, nqcap), qlmi_o(0:nplw, nqcap)
  real(c_double) :: svco_o(0:nplw5), rhocr_o(0:1000), rhocra_o(0:1000)
  integer :: nfile(2), myid, nodes, nvlcl, nplw_dummy(1), ierror, nplw5ex
  real*8 :: alloc_mem, ecutdens
  real*8, allocatable :: vlocli(:), xitgrd(:), rr(:)
  real*8, allocatable :: tablc(:, :), tablca(:, :), tbflc(:, :), tbflca(:, :)
  real*8 :: dltlc(1), rmxlc(1), dltflc(1), rmxflc(1)
  nfile = 6
  myid = 0
  nodes = 1
  loutfile(1) = .true.   ! the reference's log lines go to the child's stdout
  call qxref_serial_world()
  alloc_mem = 0.d0
  ecutdens = ecutdens_in
  ntype = 1
  jgga = jgga_in
  lrela = .false.
  lkbpp = .true.
  lvand = .false.
  allocate(zatom(1), lkbppi(1), lvandi(1), lvflag(1), lking(1), llking(1), lpcci(1), rpcc(1), &
&          lpking(1), rpking(1), gpkgmax(1), gpkgexct(1), zv(1), lmax(1), lclno(1), lsomax(1))
  zatom(1) = zatom_in
  lkbppi = .true.
  lvandi = .false.
  lvflag = .false.
  lking = .false.
  llking = .false.
  lpcci = .false.
  lpking = .false.
  rpcc = 0.d0
  rpking = 0.d0
  gpkgmax = 0.d0
  gpkgexct = 0.d0
  call psvariables_alloc(nfile, myid, nodes, alloc_mem, ntype)
  call read_pp(nfile, myid, nodes)
  call planewave_decomp_variables_alloc(nfile, alloc_mem)
  nplw5ex = 2 * (nplw5 + 1)
  call planewave_decomp(nfile, nplw5ex, nplw5)
  call nlpp_variables_alloc(nfile, myid, nodes, alloc_mem, ntype, mxl, nplw, pwscale)
  call nlppg_variables_alloc(nfile, myid, nodes, alloc_mem, ntype, mxl, nplw, recnrm, pwscale)
  nvlcl = max(mx1, mx1loc)
  allocate(vlocli(0:nvlcl), xitgrd(0:nvlcl), rr(0:nvlcl))
  allocate(tablc(0:mx1loc, 1), tablca(0:mx1loc, 1), tbflc(0:mx1loc, 1), tbflca(0:mx1loc, 1))
  call rdtmpp(nfile, myid, nodes, jgga, lrela, ntype, zatom, zv, lmax, lclno, lchk, ecutdens, &
&   lkbppi, lpcci, rpcc, lpking, rpking, gpkgmax, gpkgexct, ms, aname, mxl, mx1, vlocli, xitgrd, &
&   rr, mx1)
  call lotmpp(nfile, myid, nodes, ntype, zv, rctflc, llking, tablc, tablca, dltlc, rmxlc, &
&   tbflc, tbflca, dltflc, rmxflc, mx1loc)
  call lopp_variables_alloc(nfile, myid, nodes, alloc_mem, ntype, nplw5, .false., pwscale)
  call lotmpp_g(nfile, myid, nodes, ntype, zv, llking, dltlc, rmxlc, recnrm5, recnrmex, nplw5, &
&   vlocli, xitgrd, rr, nvlcl, .false., .true.)
  nplw_dummy(1) = nplw
  call nlkbpp_g(nfile, myid, nodes, jgga, lrela, ntype, zatom, zv, lmax, lclno, lchk, lkbppi, &
&   lking, ms, aname, mxl, nplw_dummy, gx, gy, gz, nplw, vlocli, xitgrd, rr, mx1, ierror)
  dims = 0
  dims(1) = lmmax
  dims(2) = lnref
  dims(3) = nqlmmx
  dims(4) = lmx(1)
  dims(5) = mxl
  dims(6) = lmax(1)
  dims(7) = lclno(1)
  dims(8) = ierror
  if (lmmax > lmcap .or. lnref > lncap .or. nqlmmx > nqcap .or. mxqlm /= 550 .or. &
&     mxqab_lc /= 1000) then
    dims(8) = -1
    return
  end if
  zv_o = zv(1)
  clalp_o(1:lmmax) = clalp(1:lmmax, 1)
  tbqlm_o(:, 1:lnref) = tbqlm(:, 1:lnref, 1)
  tbqlma_o(:, 1:lnref) = tbqlma(:, 1:lnref, 1)
  dlqlm_o = dlqlm(1)
  itola_o(1:lmmax) = itola(1:lmmax, 1)
  itolm_o(1:lmmax) = itolm(1:lmmax, 1)
  itolmf_o(1:lmmax) = itolmf(1:lmmax, 1)
  litoll_o(1:lmmax) = merge(1, 0, litoll(1:lmmax, 1))
  qlmr_o(:, 1:nqlmmx) = qlmr(0:nplw, 1:nqlmmx, 1, 1)
  qlmi_o(:, 1:nqlmmx) = qlmi(0:nplw, 1:nqlmmx, 1, 1)
  svco_o(0:nplw5) = svco(1:nplw5 + 1, 1)
  rhocr_o = rhocr(0:1000)
  rhocra_o = rhocra(0:1000)
end subroutine
