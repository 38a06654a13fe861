!-----------------------------------------------------------------------
! qxmd_amd_c -- ISO_C_BINDING interfaces of the qxmd_amd C-ABI
! (include/qxmd_amd.h).  Every function returns 0 on success and a
! negative code on failure; qx_last_error() describes the failure.
! Arrays follow the reference's Fortran layouts (cgjr(nplwex, nb),
! fft3x(kfft0), eqlmr_(nplwex, lmmax, natom), dmtrxr(n, n)).
!-----------------------------------------------------------------------
module qxmd_amd_c
use, intrinsic :: iso_c_binding
implicit none

integer(c_int), parameter :: QX_FP64 = 0, QX_MIXED = 1

interface

type(c_ptr) function qx_last_error() bind(C, name="qx_last_error")
  import :: c_ptr
end function

integer(c_int) function qx_version() bind(C, name="qx_version")
  import :: c_int
end function

integer(c_int) function qx_get_mesh(hcell, ecutdens, nd) bind(C, name="qx_get_mesh")
  import :: c_int, c_double
  real(c_double), intent(in) :: hcell(3, 3)
  real(c_double), value :: ecutdens
  integer(c_int), intent(out) :: nd(3)
end function

integer(c_int) function qx_strvec(rba, ecutdens, ecut, cap, counts, nga, ngb, ngc, recnrm,  &
&                                 gx, gy, gz) bind(C, name="qx_strvec")
  import :: c_int, c_double, c_ptr
  real(c_double), intent(in) :: rba(3, 3)
  real(c_double), value :: ecutdens, ecut
  integer(c_int), value :: cap
  integer(c_int), intent(out) :: counts(6)
  integer(c_int), intent(out) :: nga(*), ngb(*), ngc(*)
  real(c_double), intent(out) :: recnrm(*)
  type(c_ptr), value :: gx, gy, gz
end function

integer(c_int) function qx_ijkgd(nplw5, nga, ngb, ngc, nd, ijkgd) bind(C, name="qx_ijkgd")
  import :: c_int
  integer(c_int), value :: nplw5
  integer(c_int), intent(in) :: nga(*), ngb(*), ngc(*), nd(3)
  integer(c_int), intent(out) :: ijkgd(*)
end function

integer(c_int) function qx_create(ctx, n1, n2, n3, nplw, nga, ngb, ngc, recnrm, nbmax)  &
&                                 bind(C, name="qx_create")
  import :: c_int, c_double, c_ptr
  type(c_ptr), intent(out) :: ctx
  integer(c_int), value :: n1, n2, n3, nplw, nbmax
  integer(c_int), intent(in) :: nga(*), ngb(*), ngc(*)
  real(c_double), intent(in) :: recnrm(*)
end function

integer(c_int) function qx_destroy(ctx) bind(C, name="qx_destroy")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
end function

integer(c_int) function qx_shutdown() bind(C, name="qx_shutdown")
  import :: c_int
end function

integer(c_int) function qx_sync(ctx) bind(C, name="qx_sync")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
end function

integer(c_int) function qx_set_glocal(ctx, glocal) bind(C, name="qx_set_glocal")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(in) :: glocal(*)
end function

integer(c_int) function qx_set_kb(ctx, eqlmr, lmmax, natom, lmx_atom, vclalp_atom)  &
&                                 bind(C, name="qx_set_kb")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(in) :: eqlmr(*)
  integer(c_int), value :: lmmax, natom
  integer(c_int), intent(in) :: lmx_atom(*)
  real(c_double), intent(in) :: vclalp_atom(*)
end function

integer(c_int) function qx_set_nlpp_tables(ctx, ntype, lmmax, lnref, mxqlm, lmx, clalp, tbqlm,  &
&     tbqlma, dlqlm, itola, itolm, itolmf, litoll, nqlmmx, lplh, qlmr, qlmi)  &
&     bind(C, name="qx_set_nlpp_tables")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: ntype, lmmax, lnref, mxqlm, nqlmmx, lplh
  integer(c_int), intent(in) :: lmx(*), itola(*), itolm(*), itolmf(*), litoll(*)
  real(c_double), intent(in) :: clalp(*), tbqlm(*), tbqlma(*), dlqlm(*), qlmr(*), qlmi(*)
end function

integer(c_int) function qx_stscfn_pw(ctx, ratm, natom, hcell, nd1v, nd1vks, lclust, kmax1,  &
&     kmax2, kmax3, nplwcs, nplw5, nga, ngb, ngc, ycos, ysin) bind(C, name="qx_stscfn_pw")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(in) :: ratm(*), hcell(*)
  integer(c_int), value :: natom, lclust, kmax1, kmax2, kmax3, nplwcs, nplw5
  integer(c_int), intent(in) :: nd1v(3), nd1vks(3), nga(*), ngb(*), ngc(*)
  type(c_ptr), value :: ycos, ysin
end function

integer(c_int) function qx_setnlc_g(ctx, nplw, ntype, nhk1, nhk2, natom, lkbppi, lking, ycos,  &
&     ysin, nplwcs, rvol, vclalp) bind(C, name="qx_setnlc_g")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: nplw, ntype, natom, nplwcs
  integer(c_int), intent(in) :: nhk1(*), nhk2(*), lkbppi(*), lking(*)
  type(c_ptr), value :: ycos, ysin, vclalp
  real(c_double), value :: rvol
end function

integer(c_int) function qx_get_kb(ctx, eqlmr, vclalp_atom) bind(C, name="qx_get_kb")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(out) :: eqlmr(*)
  type(c_ptr), value :: vclalp_atom
end function

integer(c_int) function qx_info(ctx, out) bind(C, name="qx_info")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), intent(out) :: out(5)
end function

integer(c_int) function qx_slot_put(ctx, slot, gdcr, nb) bind(C, name="qx_slot_put")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot, nb
  real(c_double), intent(in) :: gdcr(*)
end function

integer(c_int) function qx_slot_get(ctx, slot, gdcr, nb) bind(C, name="qx_slot_get")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot, nb
  real(c_double), intent(out) :: gdcr(*)
end function

integer(c_int) function qx_slot_copy(ctx, src, dst, nb) bind(C, name="qx_slot_copy")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: src, dst, nb
end function

integer(c_int) function qx_hpsi_slot(ctx, src, dst, nb, prec) bind(C, name="qx_hpsi_slot")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: src, dst, nb, prec
end function

integer(c_int) function qx_local_slot(ctx, src, dst, nb, prec, accumulate)  &
&                                     bind(C, name="qx_local_slot")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: src, dst, nb, prec, accumulate
end function

integer(c_int) function qx_kb_slot(ctx, src, dst, nb, slmir) bind(C, name="qx_kb_slot")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: src, dst, nb
  type(c_ptr), value :: slmir
end function

integer(c_int) function qx_schmidt_slot(ctx, slot, nb, lcholesky) bind(C, name="qx_schmidt_slot")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot, nb
  integer(c_int), intent(out) :: lcholesky
end function

integer(c_int) function qx_rotate_slot(ctx, psi, hpsi, nb, eig) bind(C, name="qx_rotate_slot")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: psi, hpsi, nb
  real(c_double), intent(out) :: eig(*)
end function

integer(c_int) function qx_density_slot(ctx, slot, nb, occ, rvol, rho, prec)  &
&                                       bind(C, name="qx_density_slot")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot, nb, prec
  real(c_double), intent(in) :: occ(*)
  real(c_double), value :: rvol
  type(c_ptr), value :: rho
end function

integer(c_int) function qx_eigen_bkl_slot(ctx, psi, hpsi, nb, ekib, eig, noband, itermx,  &
&     iteremx, tolera, threinn, wdinner, roughfzan, roughfeig, bzan, nriter, prec)  &
&     bind(C, name="qx_eigen_bkl_slot")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: psi, hpsi, nb, noband, itermx, iteremx, prec
  real(c_double), intent(in) :: ekib(*)
  real(c_double), intent(inout) :: eig(*)
  real(c_double), value :: tolera, threinn, wdinner, roughfzan, roughfeig
  real(c_double), intent(out) :: bzan(*)
  integer(c_int), intent(out) :: nriter(*)
end function

integer(c_int) function qx_eigen_bkl_batched_slot(ctx, psi, hpsi, nb, ekib, eig, noband,  &
&     itermx, iteremx, tolera, roughfzan, roughfeig, bzan, nriter, prec)  &
&     bind(C, name="qx_eigen_bkl_batched_slot")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: psi, hpsi, nb, noband, itermx, iteremx, prec
  real(c_double), intent(in) :: ekib(*)
  real(c_double), intent(inout) :: eig(*)
  real(c_double), value :: tolera, roughfzan, roughfeig
  real(c_double), intent(out) :: bzan(*)
  integer(c_int), intent(out) :: nriter(*)
end function

integer(c_int) function qx_ekib_slot(ctx, slot, nb, ekib) bind(C, name="qx_ekib_slot")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot, nb
  real(c_double), intent(out) :: ekib(*)
end function

integer(c_int) function qx_scf_band_step(ctx, psi, hpsi, nb, occ, noband, rvol, eig, rho,  &
&     lcholesky, prec) bind(C, name="qx_scf_band_step")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: psi, hpsi, nb, noband, prec
  real(c_double), intent(in) :: occ(*)
  real(c_double), value :: rvol
  real(c_double), intent(out) :: eig(*)
  type(c_ptr), value :: rho
  integer(c_int), intent(out) :: lcholesky
end function

integer(c_int) function qx_scf_band_step_ex(ctx, psi, hpsi, nb, occ, noband, rvol, eig, rho,  &
&     lcholesky, prec, batched) bind(C, name="qx_scf_band_step_ex")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: psi, hpsi, nb, noband, prec, batched
  real(c_double), intent(in) :: occ(*)
  real(c_double), value :: rvol
  real(c_double), intent(out) :: eig(*)
  type(c_ptr), value :: rho
  integer(c_int), intent(out) :: lcholesky
end function

integer(c_int) function qx_event_record(ctx, slot) bind(C, name="qx_event_record")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot
end function

integer(c_int) function qx_event_elapsed(ctx, slot0, slot1, ms) bind(C, name="qx_event_elapsed")
  import :: c_int, c_float, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot0, slot1
  real(c_float), intent(out) :: ms
end function

integer(c_int) function qx_scf_phase_ms(ctx, ms) bind(C, name="qx_scf_phase_ms")
  import :: c_int, c_float, c_ptr
  type(c_ptr), value :: ctx
  real(c_float), intent(out) :: ms(5)
end function

integer(c_int) function qx_launch_timer(ctx, id, mode, ms, count) bind(C, name="qx_launch_timer")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: id, mode
  real(c_double), intent(out) :: ms
  integer(c_int), intent(out) :: count
end function

integer(c_int) function qx_rfft3(key, fft3x, fft3y, k1, k2, k3, ierror) bind(C, name="qx_rfft3")
  import :: c_int, c_double
  integer(c_int), value :: key, k1, k2, k3
  real(c_double), intent(inout) :: fft3x(*), fft3y(*)
  integer(c_int), intent(out) :: ierror
end function

integer(c_int) function qx_fft3(key, fft3x, fft3y, k1, k2, k3, ierror) bind(C, name="qx_fft3")
  import :: c_int, c_double
  integer(c_int), value :: key, k1, k2, k3
  real(c_double), intent(inout) :: fft3x(*), fft3y(*)
  integer(c_int), intent(out) :: ierror
end function

integer(c_int) function qx_hcnonloc(ctx, cgjr, rhcr) bind(C, name="qx_hcnonloc")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(in) :: cgjr(*)
  real(c_double), intent(inout) :: rhcr(*)
end function

integer(c_int) function qx_hckbpp_g(ctx, cgjr, rhcr, nb, slmir) bind(C, name="qx_hckbpp_g")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(in) :: cgjr(*)
  real(c_double), intent(inout) :: rhcr(*)
  integer(c_int), value :: nb
  type(c_ptr), value :: slmir
end function

integer(c_int) function qx_kbnlpp_frc2(ctx, cgjr, nb, occ, gx, gy, gz, fnlc)  &
&     bind(C, name="qx_kbnlpp_frc2")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(in) :: cgjr(*), occ(*), gx(*), gy(*), gz(*)
  integer(c_int), value :: nb
  real(c_double), intent(inout) :: fnlc(*)
end function

integer(c_int) function qx_hckbpp_ib_g(ctx, cgjr, rhcr) bind(C, name="qx_hckbpp_ib_g")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(in) :: cgjr(*)
  real(c_double), intent(inout) :: rhcr(*)
end function

integer(c_int) function qx_schmidt(ctx, cgjr, nband, lcholesky) bind(C, name="qx_schmidt")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(inout) :: cgjr(*)
  integer(c_int), value :: nband
  integer(c_int), intent(out) :: lcholesky
end function

integer(c_int) function qx_rseigqr(ctx, a, n, w) bind(C, name="qx_rseigqr")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(inout) :: a(*)
  integer(c_int), value :: n
  real(c_double), intent(out) :: w(*)
end function

integer(c_int) function qx_icholesky(ctx, s, n, ok) bind(C, name="qx_icholesky")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(inout) :: s(*)
  integer(c_int), value :: n
  integer(c_int), intent(out) :: ok
end function

integer(c_int) function qx_softchg(ctx, cgjr, nb, occ, rvol, rho) bind(C, name="qx_softchg")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(in) :: cgjr(*)
  integer(c_int), value :: nb
  real(c_double), intent(in) :: occ(*)
  real(c_double), value :: rvol
  real(c_double), intent(out) :: rho(*)
end function

! ---- local-pseudopotential forces
integer(c_int) function qx_set_lopp(ctx, ntype, nplw5, svco, gx, gy, gz)  &
&     bind(C, name="qx_set_lopp")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: ntype, nplw5
  real(c_double), intent(in) :: svco(*), gx(*), gy(*), gz(*)
end function

integer(c_int) function qx_flocal_g(ctx, floc, ntype, nhk1, nhk2, iatmpt, natom, llking, nplw5,  &
&     nga, ngb, ngc, rhgr, kmax1, kmax2, kmax3, dcq1, dcq2, dcq3, dsq1, dsq2, dsq3)  &
&     bind(C, name="qx_flocal_g")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(inout) :: floc(*)
  integer(c_int), value :: ntype, natom, nplw5, kmax1, kmax2, kmax3
  integer(c_int), intent(in) :: nhk1(*), nhk2(*), iatmpt(*), llking(*), nga(*), ngb(*), ngc(*)
  real(c_double), intent(in) :: rhgr(*), dcq1(*), dcq2(*), dcq3(*), dsq1(*), dsq2(*), dsq3(*)
end function

integer(c_int) function qx_flocal_g_dev(ctx, ntype, nhk1, nhk2, iatmpt, llking, natom, floc)  &
&     bind(C, name="qx_flocal_g_dev")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: ntype, natom
  integer(c_int), intent(in) :: nhk1(*), nhk2(*), iatmpt(*), llking(*)
  real(c_double), intent(inout) :: floc(*)
end function

integer(c_int) function qx_vlocal_g(ctx, ntype, nhk1, nhk2, iatmpt, llking, natom, rvol,  &
&     accumulate, vext, thrhgr) bind(C, name="qx_vlocal_g")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: ntype, natom, accumulate
  integer(c_int), intent(in) :: nhk1(*), nhk2(*), iatmpt(*), llking(*)
  real(c_double), value :: rvol
  real(c_double) :: vext(*), thrhgr(*)
end function

! ---- velocity-Verlet updates of the MD loop (host arithmetic)
integer(c_int) function qx_rciprl(a, b, v) bind(C, name="qx_rciprl")
  import :: c_int, c_double
  real(c_double) :: a(3, 3), b(3, 3), v
end function

integer(c_int) function qx_normalized_a(n, ntype, a, is, h, acon) bind(C, name="qx_normalized_a")
  import :: c_int, c_double
  integer(c_int), value :: n, ntype
  real(c_double) :: a(*), h(3, 3), acon(*)
  integer(c_int) :: is(*)
end function

integer(c_int) function qx_vkick(v, a, is, n, ntype, lfixion) bind(C, name="qx_vkick")
  import :: c_int, c_double
  real(c_double) :: v(*), a(*)
  integer(c_int) :: is(*), lfixion(*)
  integer(c_int), value :: n, ntype
end function

integer(c_int) function qx_updtps(x, v, is, n, ntype, lfixion) bind(C, name="qx_updtps")
  import :: c_int, c_double
  real(c_double) :: x(*), v(*)
  integer(c_int) :: is(*), lfixion(*)
  integer(c_int), value :: n, ntype
end function

integer(c_int) function qx_updtps_pbc(x, n) bind(C, name="qx_updtps_pbc")
  import :: c_int, c_double
  real(c_double) :: x(*)
  integer(c_int), value :: n
end function

! ---- wave-function extrapolation
integer(c_int) function qx_aspc_init(ctx, nprvmx, ihest) bind(C, name="qx_aspc_init")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: nprvmx, ihest
end function

integer(c_int) function qx_sbalinaspc_slot(ctx, slot, nb, nstep, nstep_ini, info)  &
&     bind(C, name="qx_sbalinaspc_slot")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot, nb, nstep, nstep_ini
  integer(c_int), intent(out) :: info(2)
end function

! ---- Ewald reciprocal-space forces
integer(c_int) function qx_set_ewald(ctx, numk2, nga, ngb, ngc, gx, gy, gz, recnrm, recnrmex,  &
&     gamma, volume) bind(C, name="qx_set_ewald")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: numk2
  integer(c_int), intent(in) :: nga(*), ngb(*), ngc(*)
  real(c_double), intent(in) :: gx(*), gy(*), gz(*), recnrm(*), recnrmex(*)
  real(c_double), value :: gamma, volume
end function

integer(c_int) function qx_frcrec(ctx, fclm, ntype, nhk1, nhk2, zv, nel, ecolmb, atom_ecolmb,  &
&     iatoit, iatmpt, natom, nion, kcsma1, kcsma2, kcsma3, dcq1, dcq2, dcq3, dsq1, dsq2, dsq3)  &
&     bind(C, name="qx_frcrec")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(inout) :: fclm(*), ecolmb, atom_ecolmb(*)
  integer(c_int), value :: ntype, nel, natom, nion, kcsma1, kcsma2, kcsma3
  integer(c_int), intent(in) :: nhk1(*), nhk2(*), iatoit(*), iatmpt(*)
  real(c_double), intent(in) :: zv(*), dcq1(*), dcq2(*), dcq3(*), dsq1(*), dsq2(*), dsq3(*)
end function

integer(c_int) function qx_frcrec_dev(ctx, fclm, ntype, nhk1, nhk2, zv, nel, ecolmb,  &
&     atom_ecolmb, iatoit, iatmpt, natom, nion) bind(C, name="qx_frcrec_dev")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(inout) :: fclm(*), ecolmb, atom_ecolmb(*)
  integer(c_int), value :: ntype, nel, natom, nion
  integer(c_int), intent(in) :: nhk1(*), nhk2(*), iatoit(*), iatmpt(*)
  real(c_double), intent(in) :: zv(*)
end function

! ---- SCF glue
integer(c_int) function qx_xclda(rho, rhocore, mshnod, rdelv, vexc, eeexc, e4)  &
&     bind(C, name="qx_xclda")
  import :: c_int, c_double
  real(c_double), intent(in) :: rho(*), rhocore(*)
  integer(c_int), value :: mshnod
  real(c_double), value :: rdelv
  real(c_double), intent(out) :: vexc(*), eeexc(*), e4(4)
end function

integer(c_int) function qx_harting(vhar, rhgr, nplw5, recnrmex, mfd2ft, ntotfd, nd1vks, ijkgd) &
&     bind(C, name="qx_harting")
  import :: c_int, c_double
  real(c_double), intent(out) :: vhar(*)
  real(c_double), intent(in) :: rhgr(*), recnrmex(*)
  integer(c_int), value :: nplw5, ntotfd
  integer(c_int), intent(in) :: mfd2ft(*), nd1vks(3), ijkgd(*)
end function

integer(c_int) function qx_chgr2g(rho, rhgr, nplw5, mfd2ft, ntotfd, nd1vks, ijkgd)  &
&     bind(C, name="qx_chgr2g")
  import :: c_int, c_double
  real(c_double), intent(in) :: rho(*)
  real(c_double), intent(out) :: rhgr(*)
  integer(c_int), value :: nplw5, ntotfd
  integer(c_int), intent(in) :: mfd2ft(*), nd1vks(3), ijkgd(*)
end function

integer(c_int) function qx_chgg2r(rho, rhgr, nplw5, mfd2ft, ntotfd, nd1vks, ijkgd)  &
&     bind(C, name="qx_chgg2r")
  import :: c_int, c_double
  real(c_double), intent(out) :: rho(*)
  real(c_double), intent(in) :: rhgr(*)
  integer(c_int), value :: nplw5, ntotfd
  integer(c_int), intent(in) :: mfd2ft(*), nd1vks(3), ijkgd(*)
end function

integer(c_int) function qx_mxchgg_alloc(ctx, nplwcex, imxchg, itratn, cmetric, wkerker, g2min,  &
&     g2max, dkgnrm, rvol) bind(C, name="qx_mxchgg_alloc")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: nplwcex, imxchg, itratn
  real(c_double), value :: cmetric, wkerker, g2min, g2max, rvol
  real(c_double), intent(in) :: dkgnrm(*)
end function

integer(c_int) function qx_mxchgg(ctx, lreset, rhgr, rinr, nplw5, aslh, bslh, glocal, mfd2ft,  &
&     ntotfd, ijkgd, lmixhold) bind(C, name="qx_mxchgg")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), intent(inout) :: lreset
  real(c_double), intent(inout) :: rhgr(*), rinr(*)
  integer(c_int), value :: nplw5, ntotfd, lmixhold
  real(c_double), value :: aslh, bslh
  real(c_double), intent(out) :: glocal(*)
  integer(c_int), intent(in) :: mfd2ft(*), ijkgd(*)
end function

integer(c_int) function qx_scf_set_grid(ctx, nplw5, ijkgd, recnrmex, vext, rhocore, rdelv)  &
&     bind(C, name="qx_scf_set_grid")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: nplw5
  integer(c_int), intent(in) :: ijkgd(*)
  real(c_double), intent(in) :: recnrmex(*)
  type(c_ptr), value :: vext, rhocore
  real(c_double), value :: rdelv
end function

integer(c_int) function qx_pwpchg_slot(ctx, slot, nb, occ, rvol, prec)  &
&     bind(C, name="qx_pwpchg_slot")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: slot, nb, prec
  real(c_double), intent(in) :: occ(*)
  real(c_double), value :: rvol
end function

integer(c_int) function qx_outchg(ctx) bind(C, name="qx_outchg")
  import :: c_int, c_ptr
  type(c_ptr), value :: ctx
end function

integer(c_int) function qx_pwpmxchgg(ctx, lreset, aslh, bslh, lmixhold, info)  &
&     bind(C, name="qx_pwpmxchgg")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: lreset, lmixhold
  real(c_double), value :: aslh, bslh
  integer(c_int), intent(out) :: info(2)
end function

integer(c_int) function qx_scf_potential(ctx, ene) bind(C, name="qx_scf_potential")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  real(c_double), intent(out) :: ene(6)
end function

integer(c_int) function qx_scf_get(ctx, what, host) bind(C, name="qx_scf_get")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: what
  real(c_double), intent(out) :: host(*)
end function

integer(c_int) function qx_scf_put(ctx, what, host) bind(C, name="qx_scf_put")
  import :: c_int, c_double, c_ptr
  type(c_ptr), value :: ctx
  integer(c_int), value :: what
  real(c_double), intent(in) :: host(*)
end function

end interface

end module qxmd_amd_c
