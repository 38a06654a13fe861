!-----------------------------------------------------------------------
! qxmd_amd_gpu -- Fortran host side of the MI355X path.
!
! The module holds the one device context of the process (G sphere, FFT
! grid, local potential, KB tables, band storage) and gives the QXMD
! operations whose inputs live in QXMD modules an explicit-argument form.
! After it, the external subroutines with the reference's names and
! argument lists replace the reference's own:
!   rfft3, fft3          Sources/fft3.f90:113, 237
!   RsEIGQR              Sources/funcs.f90:380
!   icholesky            Sources/schmidt.f90:383
!   softchg              Sources/chgdns.f90:388
!   hckbpp_g             Sources/nlkbpp.f90:1284
!   hckbpp_ib_g          Sources/nlkbpp.f90:1623
! A failing library call stops the program with the library's message,
! the role of fstop in the reference.
!-----------------------------------------------------------------------
module qxmd_amd_gpu
use, intrinsic :: iso_c_binding
use qxmd_amd_c
implicit none
private :: c_msg

type(c_ptr), save :: qxctx = c_null_ptr
integer, save :: qx_nplw = -1, qx_nbmax = 0, qx_lmmax = 0, qx_natom = 0
integer, save :: qx_nd(3) = 0
! slmir(lmmax, natom, nband) of the last hckbpp_g / hckbpp_ib_g call (the
! module output of nl_variables in the reference)
real(c_double), allocatable, target, save :: qx_slmir(:,:,:)
! mixer settings of mxchgg_alloc / set_g2ming2max_in_chgdns, applied to the
! device mixer at the next mxchgg (which brings dkgnrm and rvol)
integer, save :: mx_imxchg = -1, mx_itratn = 0, mx_nplwcex = 0
real*8, save :: mx_cmetric = 1.d0, mx_wkerker = 1.d0, mx_g2min = 0.d0, mx_g2max = 0.d0
logical, save :: mx_pending = .false.

contains

function c_msg() result(s)
  character(len=256) :: s
  character(kind=c_char), pointer :: p(:)
  type(c_ptr) :: cp
  integer :: i
  s = ' '
  cp = qx_last_error()
  if (.not. c_associated(cp)) return
  call c_f_pointer(cp, p, [256])
  do i = 1, 256
    if (p(i) == c_null_char) exit
    s(i:i) = p(i)
  end do
end function

subroutine qxmd_gpu_check(rc, what)
  integer(c_int), intent(in) :: rc
  character(*), intent(in) :: what
  if (rc /= 0) then
    write(*, '(a, a, i0, a, a)') what, ': qxmd_amd error ', rc, ': ', trim(c_msg())
    error stop 1
  end if
end subroutine

subroutine qxmd_gpu_need_ctx(what)
  character(*), intent(in) :: what
  if (.not. c_associated(qxctx)) then
    write(*, '(a, a)') what, ': qxmd_gpu_init has not been called'
    error stop 1
  end if
end subroutine

! Context from the reference's own set_rciprl arrays for G index 0..nplw
! (Sources/modules.f90:1402-1603) and FFT grid nd (Sources/input.f90:1082).
subroutine qxmd_gpu_init(nd, nplw, nga, ngb, ngc, recnrm, nbmax)
  integer, intent(in) :: nd(3), nplw, nbmax
  integer, intent(in) :: nga(0:nplw), ngb(0:nplw), ngc(0:nplw)
  real*8, intent(in) :: recnrm(0:nplw)
  integer(c_int) :: ia(0:nplw), ib(0:nplw), ic(0:nplw)
  call qxmd_gpu_finalize()
  ia = nga
  ib = ngb
  ic = ngc
  call qxmd_gpu_check(qx_create(qxctx, int(nd(1), c_int), int(nd(2), c_int),  &
&      int(nd(3), c_int), int(nplw, c_int), ia, ib, ic, recnrm, int(nbmax, c_int)),  &
&      'qxmd_gpu_init')
  qx_nplw = nplw
  qx_nbmax = nbmax
  qx_nd = nd
end subroutine

! Context for a cell alone: get_mesh + strvec + set_rciprl in the library.
! hcell(3,3): lattice vectors as columns (bohr); cutoffs in Ry.
subroutine qxmd_gpu_init_cell(hcell, ecut, ecutdens, nbmax, nd, nplw)
  real*8, intent(in) :: hcell(3, 3), ecut, ecutdens
  integer, intent(in) :: nbmax
  integer, intent(out) :: nd(3), nplw
  integer(c_int) :: ndc(3), counts(6), cap, rc, dummy(1)
  integer(c_int), allocatable :: nga(:), ngb(:), ngc(:)
  real(c_double), allocatable :: rec(:)
  real(c_double) :: dr(1)
  call qxmd_gpu_check(qx_get_mesh(hcell, ecutdens, ndc), 'qx_get_mesh')
  rc = qx_strvec(hcell, ecutdens, ecut, 1_c_int, counts, dummy, dummy, dummy, dr,  &
&                c_null_ptr, c_null_ptr, c_null_ptr)
  cap = -rc
  allocate(nga(cap), ngb(cap), ngc(cap), rec(cap))
  rc = qx_strvec(hcell, ecutdens, ecut, cap, counts, nga, ngb, ngc, rec,  &
&                c_null_ptr, c_null_ptr, c_null_ptr)
  if (rc <= 0) call qxmd_gpu_check(-1_c_int, 'qx_strvec')
  nd = ndc
  nplw = counts(3)
  call qxmd_gpu_init(nd, nplw, nga(1:nplw + 1), ngb(1:nplw + 1), ngc(1:nplw + 1),  &
&                    rec(1:nplw + 1), nbmax)
end subroutine

! releases the context and the drop-ins' process-wide device resources
subroutine qxmd_gpu_finalize()
  integer(c_int) :: rc
  if (c_associated(qxctx)) rc = qx_destroy(qxctx)
  qxctx = c_null_ptr
  qx_nplw = -1
  rc = qx_shutdown()
end subroutine

! glocal = V_H + V_xc + V_loc on the FFT grid, Ry (Sources/engrad.f90:487-495)
subroutine qxmd_gpu_set_glocal(glocal)
  real*8, intent(in) :: glocal(*)
  call qxmd_gpu_need_ctx('qxmd_gpu_set_glocal')
  call qxmd_gpu_check(qx_set_glocal(qxctx, glocal), 'qx_set_glocal')
end subroutine

! KB tables as setnlc_g leaves them (Sources/nlkbpp.f90:1236-1279)
subroutine qxmd_gpu_set_kb(eqlmr_, nplwex, lmmax, natom, lmx, vclalp)
  integer, intent(in) :: nplwex, lmmax, natom, lmx(natom)
  real*8, intent(in) :: eqlmr_(nplwex, lmmax, natom), vclalp(lmmax, natom)
  integer(c_int) :: lmxc(natom)
  call qxmd_gpu_need_ctx('qxmd_gpu_set_kb')
  lmxc = lmx
  call qxmd_gpu_check(qx_set_kb(qxctx, eqlmr_, int(lmmax, c_int), int(natom, c_int), lmxc,  &
&      vclalp), 'qx_set_kb')
  qx_lmmax = lmmax
  qx_natom = natom
end subroutine

! Tables of the nonlocal term (the nlpp_variables arrays setqlm reads,
! Sources/ppkb.f90:108-191) registered once per cell / pseudopotential set.
subroutine qxmd_gpu_set_nlpp_tables(ntype, lmmax, lnref, mxqlm, lmx, clalp, tbqlm, tbqlma,  &
&   dlqlm, itola, itolm, itolmf, litoll, nqlmmx, lplh, qlmr, qlmi)
  integer, intent(in) :: ntype, lmmax, lnref, mxqlm, nqlmmx, lplh, lmx(ntype)
  real*8, intent(in) :: clalp(lmmax, ntype), dlqlm(ntype)
  real*8, intent(in) :: tbqlm(0:mxqlm, lnref, ntype), tbqlma(0:mxqlm, lnref, ntype)
  integer, intent(in) :: itola(lmmax, ntype), itolm(lmmax, ntype), itolmf(lmmax, ntype)
  logical, intent(in) :: litoll(lmmax, ntype)
  real*8, intent(in) :: qlmr(0:lplh, nqlmmx), qlmi(0:lplh, nqlmmx)
  integer(c_int) :: lmxc(ntype), ia(lmmax, ntype), im(lmmax, ntype), imf(lmmax, ntype)
  integer(c_int) :: il(lmmax, ntype)
  call qxmd_gpu_need_ctx('qxmd_gpu_set_nlpp_tables')
  lmxc = lmx
  ia = itola
  im = itolm
  imf = itolmf
  il = merge(1, 0, litoll)
  call qxmd_gpu_check(qx_set_nlpp_tables(qxctx, int(ntype, c_int), int(lmmax, c_int),  &
&      int(lnref, c_int), int(mxqlm, c_int), lmxc, clalp, tbqlm, tbqlma, dlqlm, ia, im, imf, il, &
&      int(nqlmmx, c_int), int(lplh, c_int), qlmr, qlmi), 'qx_set_nlpp_tables')
end subroutine

! setnlc_g (Sources/nlkbpp.f90:1236-1279): eqlmr_ and vclalp built on the
! device from the registered tables and the device copy of the structure
! factors of the last stscfn_pw (the ycos/ysin arguments are not read).
subroutine qxmd_gpu_setnlc_g(nplw, ntype, nhk1, nhk2, natom, lkbppi, lking, ycos, ysin, nplwcs,  &
&   rvol, lmmax, vclalp)
  integer, intent(in) :: nplw, ntype, natom, nplwcs, lmmax, nhk1(ntype), nhk2(ntype)
  logical, intent(in) :: lkbppi(ntype), lking(ntype)
  real*8, intent(in) :: ycos(0:nplwcs, natom), ysin(0:nplwcs, natom), rvol
  real*8, target, intent(out) :: vclalp(lmmax, ntype)
  integer(c_int) :: n1(ntype), n2(ntype), lk(ntype), lg(ntype)
  call qxmd_gpu_need_ctx('qxmd_gpu_setnlc_g')
  n1 = nhk1
  n2 = nhk2
  lk = merge(1, 0, lkbppi)
  lg = merge(1, 0, lking)
  call qxmd_gpu_check(qx_setnlc_g(qxctx, int(nplw, c_int), int(ntype, c_int), n1, n2,  &
&      int(natom, c_int), lk, lg, c_null_ptr, c_null_ptr, int(nplwcs, c_int), rvol,  &
&      c_loc(vclalp)), 'qx_setnlc_g')
  qx_lmmax = lmmax
  qx_natom = natom
end subroutine

! hcnonloc (Sources/engrad.f90:411-606): rhcr += V_loc cgjr for one band
subroutine qxmd_hcnonloc(cgjr, rhcr)
  real*8, intent(in) :: cgjr(*)
  real*8, intent(inout) :: rhcr(*)
  call qxmd_gpu_need_ctx('qxmd_hcnonloc')
  call qxmd_gpu_check(qx_hcnonloc(qxctx, cgjr, rhcr), 'qx_hcnonloc')
end subroutine

! schmidt (Sources/schmidt.f90:71-378) on cgjr_(nplwex, nband)
subroutine qxmd_schmidt(cgjr_, nplwex, nband, lcholesky)
  integer, intent(in) :: nplwex, nband
  real*8, intent(inout) :: cgjr_(nplwex, nband)
  logical, intent(out) :: lcholesky
  integer(c_int) :: ok
  call qxmd_gpu_need_ctx('qxmd_schmidt')
  call qxmd_gpu_check(qx_schmidt(qxctx, cgjr_, int(nband, c_int), ok), 'qx_schmidt')
  lcholesky = ok /= 0
end subroutine

! One SCF band step with frozen potential on gdcr(nplwex, nband): energy
! pass, rotation, CG (batched or band by band), Gram-Schmidt, density
! (Sources/pwlda.f90:2087-3574).  rho on the FFT grid.
subroutine qxmd_scf_band_step(gdcr, nplwex, nband, occ, noband, rvol, eig, rho, lcholesky,  &
&                             mixed, batched)
  integer, intent(in) :: nplwex, nband, noband
  real*8, intent(inout) :: gdcr(nplwex, nband)
  real*8, intent(in) :: occ(nband), rvol
  real*8, intent(out) :: eig(nband)
  real*8, intent(out), target :: rho(*)
  logical, intent(out) :: lcholesky
  logical, intent(in) :: mixed, batched
  integer(c_int) :: ok, prec, bat
  call qxmd_gpu_need_ctx('qxmd_scf_band_step')
  prec = QX_FP64
  if (mixed) prec = QX_MIXED
  bat = 0
  if (batched) bat = 1
  call qxmd_gpu_check(qx_slot_put(qxctx, 0_c_int, gdcr, int(nband, c_int)), 'qx_slot_put')
  call qxmd_gpu_check(qx_scf_band_step_ex(qxctx, 0_c_int, 1_c_int, int(nband, c_int), occ,  &
&      int(noband, c_int), rvol, eig, c_loc(rho), ok, prec, bat), 'qx_scf_band_step_ex')
  call qxmd_gpu_check(qx_slot_get(qxctx, 0_c_int, gdcr, int(nband, c_int)), 'qx_slot_get')
  lcholesky = ok /= 0
end subroutine

! svco(nplw5+1, ntype) of lopp_variables (Sources/ppkb.f90:66-79) and the
! Cartesian G of the density sphere, once per cell, for flocal_g
! vlocal_g on the device (Sources/nlkbpp.f90:862-1032): the local
! pseudopotential from the svco of qxmd_gpu_set_lopp and the device structure
! factors of the last stscfn_pw over the density sphere, into the SCF glue's
! external potential (vext += ... when accumulate, as the reference's
! vext(i) = vext(i) + fft3x(i)); host copies of vext (FFT order) and thrhgr.
subroutine qxmd_vlocal_g(ntype, nhk1, nhk2, iatmpt, llking, natom, rvol, accumulate, vext,  &
&                        thrhgr)
  integer, intent(in) :: ntype, natom, nhk1(ntype), nhk2(ntype), iatmpt(natom)
  logical, intent(in) :: llking(ntype), accumulate
  real*8, intent(in) :: rvol
  real*8, intent(inout) :: vext(*), thrhgr(*)
  integer(c_int) :: lk(ntype)
  call qxmd_gpu_need_ctx('qxmd_vlocal_g')
  lk = merge(1_c_int, 0_c_int, llking)
  call qxmd_gpu_check(qx_vlocal_g(qxctx, int(ntype, c_int), nhk1, nhk2, iatmpt, lk,  &
&      int(natom, c_int), rvol, merge(1_c_int, 0_c_int, accumulate), vext, thrhgr),  &
&      'qxmd_vlocal_g')
end subroutine

subroutine qxmd_gpu_set_lopp(ntype, nplw5, svco, gx, gy, gz)
  integer, intent(in) :: ntype, nplw5
  real*8, intent(in) :: svco(nplw5 + 1, ntype), gx(0:nplw5), gy(0:nplw5), gz(0:nplw5)
  call qxmd_gpu_need_ctx('qxmd_gpu_set_lopp')
  call qxmd_gpu_check(qx_set_lopp(qxctx, int(ntype, c_int), int(nplw5, c_int), svco, gx, gy,  &
&      gz), 'qxmd_gpu_set_lopp')
end subroutine

! the Ewald G list of clmrcp and para_ewald's parameters, once per cell
! (Sources/bulk.f90:556-1068): arrays over m = 1..numk2, recnrmex(0:numk2)
subroutine qxmd_gpu_set_ewald(numk2, nga, ngb, ngc, gx, gy, gz, recnrm, recnrmex, gamma, volume)
  integer, intent(in) :: numk2, nga(numk2), ngb(numk2), ngc(numk2)
  real*8, intent(in) :: gx(numk2), gy(numk2), gz(numk2), recnrm(numk2), recnrmex(0:numk2)
  real*8, intent(in) :: gamma, volume
  integer(c_int) :: g1(max(numk2, 1)), g2(max(numk2, 1)), g3(max(numk2, 1))
  call qxmd_gpu_need_ctx('qxmd_gpu_set_ewald')
  g1(1:numk2) = nga
  g2(1:numk2) = ngb
  g3(1:numk2) = ngc
  call qxmd_gpu_check(qx_set_ewald(qxctx, int(numk2, c_int), g1, g2, g3, gx, gy, gz, recnrm,  &
&      recnrmex, gamma, volume), 'qxmd_gpu_set_ewald')
end subroutine

! frcrec (Sources/bulk.f90:1880-2026) without stress, explicit-argument form
! (the reference reads the G list and the tables from module para_ewald):
! call it in frcrec's body with para_ewald's KCSMA1..3 and DCQ1..DSQ3.
subroutine qxmd_frcrec(fclm, ntype, nhk1, nhk2, zv, nel, ecolmb, atom_ecolmb, iatoit, iatmpt,  &
&                      natom, nion, kcsma1, kcsma2, kcsma3, dcq1, dcq2, dcq3, dsq1, dsq2, dsq3)
  integer, intent(in) :: ntype, nel, natom, nion, kcsma1, kcsma2, kcsma3
  integer, intent(in) :: nhk1(ntype), nhk2(ntype), iatoit(natom), iatmpt(nion)
  real*8, intent(inout) :: fclm(3, natom), ecolmb, atom_ecolmb(natom)
  real*8, intent(in) :: zv(ntype)
  real*8, intent(in) :: dcq1(-kcsma1:kcsma1, nion), dsq1(-kcsma1:kcsma1, nion)
  real*8, intent(in) :: dcq2(-kcsma2:kcsma2, nion), dsq2(-kcsma2:kcsma2, nion)
  real*8, intent(in) :: dcq3(-kcsma3:kcsma3, nion), dsq3(-kcsma3:kcsma3, nion)
  integer(c_int) :: h1(ntype), h2(ntype), it(natom), ia(nion)
  call qxmd_gpu_need_ctx('qxmd_frcrec')
  h1 = nhk1
  h2 = nhk2
  it = iatoit
  ia = iatmpt
  call qxmd_gpu_check(qx_frcrec(qxctx, fclm, int(ntype, c_int), h1, h2, zv, int(nel, c_int),  &
&      ecolmb, atom_ecolmb, it, ia, int(natom, c_int), int(nion, c_int), int(kcsma1, c_int),  &
&      int(kcsma2, c_int), int(kcsma3, c_int), dcq1, dcq2, dcq3, dsq1, dsq2, dsq3), 'qxmd_frcrec')
end subroutine

! set_coef_ASPC (Sources/sbalin.f90:77) with ichest = 0 and the device
! history of nprvmx sets; once per run
subroutine qxmd_aspc_init(nprvmx, ihest)
  integer, intent(in) :: nprvmx, ihest
  call qxmd_gpu_need_ctx('qxmd_aspc_init')
  call qxmd_gpu_check(qx_aspc_init(qxctx, int(nprvmx, c_int), int(ihest, c_int)),  &
&      'qxmd_aspc_init')
end subroutine

! sbalinaspc (Sources/sbalin.f90:369-712), one spin, lvand false: gdcr in
! (current bands), out (the extrapolated bands of the next step); the
! history stays on the device between calls
subroutine qxmd_sbalinaspc(gdcr, nplwex, nband, nstep, nstep_ini, iaspc, lcholesky)
  integer, intent(in) :: nplwex, nband, nstep, nstep_ini
  real*8, intent(inout) :: gdcr(nplwex, nband)
  integer, intent(out) :: iaspc
  logical, intent(out) :: lcholesky
  integer(c_int) :: info(2)
  call qxmd_gpu_need_ctx('qxmd_sbalinaspc')
  call qxmd_gpu_check(qx_slot_put(qxctx, 0_c_int, gdcr, int(nband, c_int)), 'qxmd_sbalinaspc')
  call qxmd_gpu_check(qx_sbalinaspc_slot(qxctx, 0_c_int, int(nband, c_int), int(nstep, c_int),  &
&      int(nstep_ini, c_int), info), 'qxmd_sbalinaspc')
  call qxmd_gpu_check(qx_slot_get(qxctx, 0_c_int, gdcr, int(nband, c_int)), 'qxmd_sbalinaspc')
  iaspc = info(1)
  lcholesky = info(2) /= 0
end subroutine

end module qxmd_amd_gpu


!-----------------------------------------------------------------------
! Drop-in externals: the reference's names and argument lists.
!-----------------------------------------------------------------------

! rfft3 (Sources/fft3.f90:113-232): key = 1 real-to-complex x1/N with the full
! Hermitian grid written back, otherwise complex-to-real from the half grid.
! work is not used (the transform runs on the device).
subroutine rfft3(key, fft3x, fft3y, work, kfft1, kfft2, kfft3, kfft0, ierror)
use qxmd_amd_c
implicit none
integer :: key, kfft1, kfft2, kfft3, kfft0, ierror
real*8 :: fft3x(kfft0), fft3y(kfft0)
complex*16 :: work(*)
integer(c_int) :: ie, rc
rc = qx_rfft3(int(key, c_int), fft3x, fft3y, int(kfft1, c_int), int(kfft2, c_int),  &
&             int(kfft3, c_int), ie)
ierror = ie
if (rc /= 0) ierror = 1
end subroutine

! fft3 (Sources/fft3.f90:237-333): complex 3-D transform, key = 1 forward x1/N.
subroutine fft3(key, fft3x, fft3y, work, kfft1, kfft2, kfft3, kfft0, ierror)
use qxmd_amd_c
implicit none
integer :: key, kfft1, kfft2, kfft3, kfft0, ierror
real*8 :: fft3x(kfft0), fft3y(kfft0)
complex*16 :: work(*)
integer(c_int) :: ie, rc
rc = qx_fft3(int(key, c_int), fft3x, fft3y, int(kfft1, c_int), int(kfft2, c_int),  &
&            int(kfft3, c_int), ie)
ierror = ie
if (rc /= 0) ierror = 1
end subroutine

! RsEIGQR (Sources/funcs.f90:380-764): eigenvalues of the symmetric AR (lower
! triangle used) ascending in EVR; with IFVEC = 1 the eigenvectors replace AR
! column by column.  Sorting is always ascending (IFSRT = 1 returns the
! identity sorting list in W1I).  Runs in FP64 on the device of the context
! (N <= nbmax).  CHECK is left untouched.
SUBROUTINE RsEIGQR(AR, EVR, N, MTRX, EPSI, IFVEC, IFSRT, CHECK, W1R, W1I, work)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: N, MTRX
real*8 :: AR(MTRX, N), EVR(N)
real*8 :: EPSI
integer :: IFVEC, IFSRT
character*5 :: CHECK
real*8 :: W1R(N), W1I(N), work(N)
real*8, allocatable :: a(:,:)
integer :: j
call qxmd_gpu_need_ctx('RsEIGQR')
allocate(a(N, N))
a = AR(1:N, 1:N)
call qxmd_gpu_check(qx_rseigqr(qxctx, a, int(N, c_int), EVR), 'RsEIGQR')
if (IFVEC == 1) AR(1:N, 1:N) = a
if (IFSRT == 1) then
  do j = 1, N
    W1I(j) = dble(j)
  end do
end if
deallocate(a)
END SUBROUTINE

! icholesky (Sources/schmidt.f90:383-423), one rank: dmtrxr(1:nband, 1:nband)
! lower triangle -> L^-1 (lower), upper zeroed (rsdecmp + rsinvcd).
subroutine icholesky(nfile, myid, nodes, dmtrxr, nband, nbnod1, nbnod2, nbxxxx, prod, prod2,  &
& nbncnt, nbndsp, node_c, lbncnt, lbndsp, mbncnt, mbndsp, jdstnd, bijr, bm1r)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nband, nbnod1, nbnod2, nbxxxx, node_c
integer :: nbncnt(*), nbndsp(*), lbncnt(*), lbndsp(*), mbncnt(*), mbndsp(*), jdstnd(*)
real*8 :: dmtrxr(nbxxxx, *), prod(*), prod2(*), bijr(*), bm1r(*)
real*8, allocatable :: s(:,:)
integer(c_int) :: ok
call qxmd_gpu_need_ctx('icholesky')
allocate(s(nband, nband))
s = dmtrxr(1:nband, 1:nband)
call qxmd_gpu_check(qx_icholesky(qxctx, s, int(nband, c_int), ok), 'icholesky')
dmtrxr(1:nband, 1:nband) = s
deallocate(s)
end subroutine

! softchg (Sources/chgdns.f90:388-454), one rank, VECTOR layout: the density
! sum occ/rvol psi(r)^2 on the FFT grid in fft3y, mapped to the FD mesh in
! glocal(ifd) = fft3y(mfd2ft(ifd)).
subroutine softchg(nfile, myid, nodes, glocal, occ, cgjr, nplwex, nplw, nband, nbnod1, nbnod2,  &
& nbnod, iod, rvol, mfd2ft, ntotfd, nd1vks, kfft0, ijkgd, ngenh, fft3x, fft3y, fftwork)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes
integer :: nplwex, nplw, nband, nbnod1, nbnod2, nbnod, iod(*)
real*8 :: glocal(*), occ(*), cgjr(nplwex, *), rvol
integer :: mfd2ft(*), ntotfd, nd1vks(*), kfft0, ngenh
integer :: ijkgd(-ngenh:ngenh)
real*8 :: fft3x(*), fft3y(*)
complex*16 :: fftwork(*)
real*8, allocatable :: occl(:)
integer :: i, ifd
call qxmd_gpu_need_ctx('softchg')
allocate(occl(nbnod))
do i = 1, nbnod
  occl(i) = occ(iod(i))
end do
call qxmd_gpu_check(qx_softchg(qxctx, cgjr, int(nbnod, c_int), occl, rvol, fft3y), 'softchg')
do ifd = 1, ntotfd
  glocal(ifd) = fft3y(mfd2ft(ifd))
end do
deallocate(occl)
end subroutine

! hckbpp_g (Sources/nlkbpp.f90:1284-1393): rhcr(:, 1:nbnod) += V_NL cgjr with
! the projector tables registered by qxmd_gpu_set_kb; slmir goes to
! qx_slmir(lmmax, natom, nbnod).  ycos/ysin are already folded into eqlmr_.
subroutine hckbpp_g(nfile, myid, nodes, cgjr, rhcr, nplwex, nplw, nband, nbnod1, nbnod2, nbnod,  &
& ntype, nhk1, nhk2, natom, lkbppi, lking, ycos, ysin, nplwcs)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nplwex, nplw, nband, nbnod1, nbnod2, nbnod
integer :: ntype, natom, nhk1(ntype), nhk2(ntype), nplwcs
real*8 :: cgjr(nplwex, nbnod), rhcr(nplwex, nbnod)
logical :: lkbppi(ntype), lking(ntype)
real*8 :: ycos(0:nplwcs, natom), ysin(0:nplwcs, natom)
call qxmd_gpu_need_ctx('hckbpp_g')
if (allocated(qx_slmir)) then
  if (size(qx_slmir, 3) < nbnod) deallocate(qx_slmir)
end if
if (.not. allocated(qx_slmir)) allocate(qx_slmir(max(qx_lmmax, 1), max(qx_natom, 1), nbnod))
call qxmd_gpu_check(qx_hckbpp_g(qxctx, cgjr, rhcr, int(nbnod, c_int), c_loc(qx_slmir)),  &
&    'hckbpp_g')
end subroutine

! kbnlpp_frc2 (Sources/nlkbpp.f90:1878-2060) with the reference's argument
! list: fnlc(3, natom) += KB forces of the nbnod local bands (occupation
! occ(iod(ibb))).  The stress term (lcalstr) is not computed on the device.
subroutine kbnlpp_frc2(nfile, myid, nodes, fnlc, str, lcalstr, cgjr, nplwex, nplw, gx, gy, gz,  &
& nband, nbnod1, nbnod2, nbnod, occ, iod, ntype, nhk1, nhk2, natom, lkbppi, lking, ycos, ysin,  &
& nplwcs)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nplwex, nplw, nband, nbnod1, nbnod2, nbnod
integer :: ntype, natom, nhk1(ntype), nhk2(ntype), nplwcs, iod(nbnod)
real*8 :: fnlc(3, natom), str(6), cgjr(nplwex, nbnod), occ(nband)
real*8 :: gx(0:nplw), gy(0:nplw), gz(0:nplw)
logical :: lcalstr, lkbppi(ntype), lking(ntype)
real*8 :: ycos(0:nplwcs, natom), ysin(0:nplwcs, natom)
real*8 :: occb(nbnod)
integer :: ibb
call qxmd_gpu_need_ctx('kbnlpp_frc2')
if (lcalstr) call qxmd_gpu_check(-1_c_int, 'kbnlpp_frc2: stress (lcalstr) is not on the device')
do ibb = 1, nbnod
  occb(ibb) = occ(iod(ibb))
end do
call qxmd_gpu_check(qx_kbnlpp_frc2(qxctx, cgjr, int(nbnod, c_int), occb, gx, gy, gz, fnlc),  &
&    'kbnlpp_frc2')
end subroutine

! stscfn_pw (Sources/pwlda.f90:5711-5854) with the reference's argument list:
! the 1-D tables DCQ*/DSQ* are filled by the reference's recurrence here, and
! ycos/ysin come from the device (qx_stscfn_pw), which keeps them for setnlc_g.
subroutine stscfn_pw(nfile, myid, nodes, ycos, ysin, DCQ1, DCQ2, DCQ3, DSQ1, DSQ2, DSQ3,  &
& kmax1, kmax2, kmax3, natom, nplwcs, nplw5, nga, ngb, ngc, lclust, ratm, hcell, nd1v, nd1vks,  &
& lnlpp_g, llclpp_g, lpcc_g, lsphexp)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, natom, nplwcs, nplw5, kmax1, kmax2, kmax3
real*8, target :: ycos(0:nplwcs, natom), ysin(0:nplwcs, natom)
real*8 :: DCQ1(-kmax1:kmax1, natom), DSQ1(-kmax1:kmax1, natom)
real*8 :: DCQ2(-kmax2:kmax2, natom), DSQ2(-kmax2:kmax2, natom)
real*8 :: DCQ3(-kmax3:kmax3, natom), DSQ3(-kmax3:kmax3, natom)
integer :: nga(0:nplw5), ngb(0:nplw5), ngc(0:nplw5)
logical :: lclust, lnlpp_g, llclpp_g, lpcc_g, lsphexp
real*8 :: ratm(3, natom), hcell(3, 3)
integer :: nd1v(3), nd1vks(3)
real*8, parameter :: dpi = 3.141592653589793d0 * 2.d0
real*8 :: fact(3)
integer(c_int) :: d1(3), d2(3), ga(0:nplw5), gb(0:nplw5), gc(0:nplw5)
integer :: k, a
do a = 1, 3
  if (lclust) then
    fact(a) = 1.d0 / hcell(a, a) * dble(nd1v(a)) / dble(nd1vks(a))
  else
    fact(a) = dble(nd1v(a)) / dble(nd1vks(a))
  end if
  fact(a) = fact(a) * dpi
end do
do k = 1, natom
  call axis_table(DCQ1(:, k), DSQ1(:, k), kmax1, ratm(1, k) * fact(1))
  call axis_table(DCQ2(:, k), DSQ2(:, k), kmax2, ratm(2, k) * fact(2))
  call axis_table(DCQ3(:, k), DSQ3(:, k), kmax3, ratm(3, k) * fact(3))
end do
if (lnlpp_g .or. lpcc_g .or. lsphexp) then
  call qxmd_gpu_need_ctx('stscfn_pw')
  d1 = nd1v
  d2 = nd1vks
  ga = nga
  gb = ngb
  gc = ngc
  call qxmd_gpu_check(qx_stscfn_pw(qxctx, ratm, int(natom, c_int), hcell, d1, d2,  &
&      merge(1_c_int, 0_c_int, lclust), int(kmax1, c_int), int(kmax2, c_int), int(kmax3, c_int), &
&      int(nplwcs, c_int), int(nplw5, c_int), ga, gb, gc, c_loc(ycos), c_loc(ysin)), 'stscfn_pw')
end if
contains
  ! c + i s over k = -km..km: powers of exp(i ph) by the recurrence of 5806-5834
  subroutine axis_table(c, s, km, ph)
    integer, intent(in) :: km
    real*8, intent(out) :: c(-km:km), s(-km:km)
    real*8, intent(in) :: ph
    integer :: m
    c(0) = 1.d0
    s(0) = 0.d0
    c(1) = cos(ph)
    s(1) = sin(ph)
    c(-1) = c(1)
    s(-1) = -s(1)
    do m = 2, km
      c(m) = c(1) * c(m - 1) - s(1) * s(m - 1)
      s(m) = s(1) * c(m - 1) + c(1) * s(m - 1)
      c(-m) = c(m)
      s(-m) = -s(m)
    end do
  end subroutine
end subroutine

! hckbpp_ib_g (Sources/nlkbpp.f90:1623-1721): one band.
subroutine hckbpp_ib_g(nfile, myid, nodes, cgjr, rhcr, nplwex, nplw, ntype, nhk1, nhk2, natom,  &
& lkbppi, lking, ycos, ysin, nplwcs)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nplwex, nplw
integer :: ntype, natom, nhk1(ntype), nhk2(ntype), nplwcs
real*8 :: cgjr(nplwex), rhcr(nplwex)
logical :: lkbppi(ntype), lking(ntype)
real*8 :: ycos(0:nplwcs, natom), ysin(0:nplwcs, natom)
call qxmd_gpu_need_ctx('hckbpp_ib_g')
if (allocated(qx_slmir)) then
  if (size(qx_slmir, 3) < 1) deallocate(qx_slmir)
end if
if (.not. allocated(qx_slmir)) allocate(qx_slmir(max(qx_lmmax, 1), max(qx_natom, 1), 1))
call qxmd_gpu_check(qx_hckbpp_g(qxctx, cgjr, rhcr, 1_c_int, c_loc(qx_slmir)), 'hckbpp_ib_g')
end subroutine


! xclda (Sources/vxc.f90:961-1017): LDA exchange-correlation on the device.
! ---- velocity-Verlet updates (Sources/remd.f90), same argument lists
subroutine updtps(nfile, myid, nodes, x, v, is, n, ntype, lfixion)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, n, ntype
real*8 :: x(3*n), v(3*n)
integer :: is(n)
logical :: lfixion(ntype)
integer(c_int) :: lf(ntype)
lf = merge(1_c_int, 0_c_int, lfixion)
call qxmd_gpu_check(qx_updtps(x, v, is, int(n, c_int), int(ntype, c_int), lf), 'updtps')
end subroutine

subroutine vkick(nfile, myid, nodes, v, a, is, n, ntype, lfixion)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, n, ntype
real*8 :: v(3*n), a(3*n)
integer :: is(n)
logical :: lfixion(ntype)
integer(c_int) :: lf(ntype)
lf = merge(1_c_int, 0_c_int, lfixion)
call qxmd_gpu_check(qx_vkick(v, a, is, int(n, c_int), int(ntype, c_int), lf), 'vkick')
end subroutine

subroutine updtps_pbc(nfile, myid, nodes, x, is, n, ntype)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, n, ntype
real*8 :: x(3*n)
integer :: is(n)
call qxmd_gpu_check(qx_updtps_pbc(x, int(n, c_int)), 'updtps_pbc')
end subroutine

subroutine normalized_a(nfile, myid, nodes, n, ntype, a, is, h, acon)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, n, ntype
real*8 :: a(3*n), h(3,3,0:1), acon(ntype)
integer :: is(n)
call qxmd_gpu_check(qx_normalized_a(int(n, c_int), int(ntype, c_int), a, is, h(:,:,0), acon),  &
&                   'normalized_a')
end subroutine

subroutine xclda(rho, rhocore, mshnod, rdelv, vexc, eeexc, eexc, evexc, ecor, evcor)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: mshnod
real*8 :: rho(mshnod), rhocore(mshnod), vexc(mshnod), eeexc(mshnod)
real*8 :: rdelv, eexc, evexc, ecor, evcor
real(c_double) :: e4(4)
call qxmd_gpu_check(qx_xclda(rho, rhocore, int(mshnod, c_int), rdelv, vexc, eeexc, e4), 'xclda')
eexc = e4(1)
evexc = e4(2)
ecor = e4(3)
evcor = e4(4)
end subroutine

! harting (Sources/engrad.f90:1673-1743): vhar(ntotfd) in FD order.
subroutine harting(nfile, myid, nodes, vhar, rhgr, nplw5ex, nplw5, recnrmex, mfd2ft, ntotfd,  &
& nd1vks, kfft0, fft3x, fft3y, fftwork, ijkgd, ngenh)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nplw5ex, nplw5, ntotfd, kfft0, ngenh
real*8 :: vhar(*), rhgr(*), recnrmex(0:*), fft3x(*), fft3y(*)
integer :: mfd2ft(*), nd1vks(3), ijkgd(-ngenh:ngenh)
complex*16 :: fftwork(*)
integer(c_int) :: d(3), ij(-nplw5:nplw5), mf(ntotfd)
d = nd1vks
ij = ijkgd(-nplw5:nplw5)
mf = mfd2ft(1:ntotfd)
call qxmd_gpu_check(qx_harting(vhar, rhgr, int(nplw5, c_int), recnrmex, mf,  &
&    int(ntotfd, c_int), d, ij), 'harting')
end subroutine

! chgr2g (Sources/engrad.f90:1253-1297)
subroutine chgr2g(nfile, myid, nodes, rho, rhgr, nplw5ex, nplw5, mfd2ft, ntotfd, nd1vks, kfft0,  &
& fft3x, fft3y, fftwork, ijkgd, ngenh)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nplw5ex, nplw5, ntotfd, kfft0, ngenh
real*8 :: rho(*), rhgr(*), fft3x(*), fft3y(*)
integer :: mfd2ft(*), nd1vks(3), ijkgd(-ngenh:ngenh)
complex*16 :: fftwork(*)
integer(c_int) :: d(3), ij(-nplw5:nplw5), mf(ntotfd)
d = nd1vks
ij = ijkgd(-nplw5:nplw5)
mf = mfd2ft(1:ntotfd)
call qxmd_gpu_check(qx_chgr2g(rho, rhgr, int(nplw5, c_int), mf, int(ntotfd, c_int), d, ij),  &
&    'chgr2g')
end subroutine

! chgg2r (Sources/engrad.f90:1302-1353)
subroutine chgg2r(nfile, myid, nodes, rho, rhgr, nplw5ex, nplw5, mfd2ft, ntotfd, nd1vks, kfft0,  &
& fft3x, fft3y, fftwork, ijkgd, ngenh)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nplw5ex, nplw5, ntotfd, kfft0, ngenh
real*8 :: rho(*), rhgr(*), fft3x(*), fft3y(*)
integer :: mfd2ft(*), nd1vks(3), ijkgd(-ngenh:ngenh)
complex*16 :: fftwork(*)
integer(c_int) :: d(3), ij(-nplw5:nplw5), mf(ntotfd)
d = nd1vks
ij = ijkgd(-nplw5:nplw5)
mf = mfd2ft(1:ntotfd)
call qxmd_gpu_check(qx_chgg2r(rho, rhgr, int(nplw5, c_int), mf, int(ntotfd, c_int), d, ij),  &
&    'chgg2r')
end subroutine

! set_g2ming2max_in_chgdns (Sources/chgdns.f90:529-539)
subroutine set_g2ming2max_in_chgdns(g2min_, g2max_)
use qxmd_amd_gpu
implicit none
real*8 :: g2min_, g2max_
mx_g2min = g2min_
mx_g2max = g2max_
mx_pending = mx_imxchg >= 0
end subroutine

! mxchgg_alloc (Sources/chgdns.f90:559-690): imxchg 0 (none), 1 (Pulay) or 3
! (simple); spin and noncollinear mixing are not on the device.
subroutine mxchgg_alloc(nfile, myid, nodes, alloc_mem, lspin, nplwcex_, imxchg_, itratn_,  &
& imxspin_, nmxspin_, cmetric, spinmetric, wkerker, pwscale, lnoncollinear)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nplwcex_, imxchg_, itratn_, imxspin_, nmxspin_
real*8 :: alloc_mem, cmetric, spinmetric, wkerker, pwscale
logical :: lspin, lnoncollinear
if (lspin .or. lnoncollinear) call qxmd_gpu_check(-1_c_int,  &
&    'mxchgg_alloc: spin-polarised mixing is not on the device')
mx_imxchg = imxchg_
mx_itratn = itratn_
mx_nplwcex = nplwcex_
mx_cmetric = cmetric
mx_wkerker = wkerker
mx_pending = .true.
end subroutine

! mxchgg (Sources/chgdns.f90:796-902): mixing of rhgr with rinr on the
! device (history kept there), glocal = chgg2r(mixed rhgr) in FD order.
subroutine mxchgg(nfile, myid, nodes, lreset, rhgr, rinr, nplw5ex, nplw5, dkgnrm, rvol, aslh,  &
& bslh, glocal, apk, eigr, mfd2ft, ntotfd, nd1vks, kfft0, fft3x, fft3y, fftwork, ijkgd, ngenh,  &
& rdelv, nel, lmixhold)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, nplw5ex, nplw5, ntotfd, kfft0, ngenh, nel
logical :: lreset, lmixhold
real*8 :: rhgr(nplw5ex), rinr(nplw5ex), dkgnrm(*), rvol, aslh, bslh, rdelv
real*8 :: glocal(ntotfd), apk(*), eigr(*), fft3x(*), fft3y(*)
integer :: mfd2ft(ntotfd), nd1vks(3), ijkgd(-ngenh:ngenh)
complex*16 :: fftwork(*)
integer(c_int) :: ij(-nplw5:nplw5), mf(ntotfd), lr
call qxmd_gpu_need_ctx('mxchgg')
if (mx_pending) then
  call qxmd_gpu_check(qx_mxchgg_alloc(qxctx, int(mx_nplwcex, c_int), int(mx_imxchg, c_int),  &
&      int(mx_itratn, c_int), mx_cmetric, mx_wkerker, mx_g2min, mx_g2max, dkgnrm, rvol),  &
&      'mxchgg_alloc')
  mx_pending = .false.
end if
ij = ijkgd(-nplw5:nplw5)
mf = mfd2ft
lr = merge(1_c_int, 0_c_int, lreset)
call qxmd_gpu_check(qx_mxchgg(qxctx, lr, rhgr, rinr, int(nplw5, c_int), aslh, bslh, glocal, mf,  &
&    int(ntotfd, c_int), ij, merge(1_c_int, 0_c_int, lmixhold)), 'mxchgg')
lreset = .false.
end subroutine


! flocal_g (Sources/nlkbpp.f90:1037-1231) with the reference's argument list:
! local-pseudopotential forces added to floc (and doubled with it, as there);
! svco and gx/gy/gz from qxmd_gpu_set_lopp.  Stress (lcstress) stops with an
! error.
subroutine flocal_g(nfile, myid, nodes, floc, strloc, lcstress, ntype, nhk1, nhk2, iatoit,  &
& iatmpt, natom, nion, llking, nplw5, nplw5ex, nplw, nga, ngb, ngc, gx, gy, gz, rhgr, elclv,  &
& kmax1, kmax2, kmax3, DCQ1, DCQ2, DCQ3, DSQ1, DSQ2, DSQ3, ltimecnt)
use qxmd_amd_c
use qxmd_amd_gpu
implicit none
integer :: nfile(*), myid, nodes, ntype, natom, nion, nplw5, nplw5ex, nplw
integer :: kmax1, kmax2, kmax3
real*8 :: floc(3, natom), strloc(3, 3), elclv
logical :: lcstress, ltimecnt, llking(ntype)
integer :: nhk1(ntype), nhk2(ntype), iatoit(natom), iatmpt(nion)
integer :: nga(0:nplw5), ngb(0:nplw5), ngc(0:nplw5)
real*8 :: gx(0:nplw5), gy(0:nplw5), gz(0:nplw5), rhgr(nplw5ex)
real*8 :: DCQ1(-kmax1:kmax1, natom), DSQ1(-kmax1:kmax1, natom)
real*8 :: DCQ2(-kmax2:kmax2, natom), DSQ2(-kmax2:kmax2, natom)
real*8 :: DCQ3(-kmax3:kmax3, natom), DSQ3(-kmax3:kmax3, natom)
integer(c_int) :: h1(ntype), h2(ntype), ia(nion), lk(ntype), g1(0:nplw5), g2(0:nplw5), g3(0:nplw5)
integer :: it
call qxmd_gpu_need_ctx('flocal_g')
if (lcstress) call qxmd_gpu_check(-1_c_int, 'flocal_g: stress (lcstress) is not on the device')
h1 = nhk1
h2 = nhk2
ia = iatmpt
do it = 1, ntype
  lk(it) = merge(1_c_int, 0_c_int, llking(it))
end do
g1 = nga
g2 = ngb
g3 = ngc
call qxmd_gpu_check(qx_flocal_g(qxctx, floc, int(ntype, c_int), h1, h2, ia, int(natom, c_int),  &
&    lk, int(nplw5, c_int), g1, g2, g3, rhgr, int(kmax1, c_int), int(kmax2, c_int),  &
&    int(kmax3, c_int), DCQ1, DCQ2, DCQ3, DSQ1, DSQ2, DSQ3), 'flocal_g')
end subroutine
