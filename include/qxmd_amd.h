/*
 * qxmd_amd.h -- C-ABI of the MI355X-native QXMD Gamma-point H|psi> /
 * band-update path.  Plain pointers and sizes only; every entry point returns
 * 0 on success and a negative code on failure (qx_last_error() describes it).
 *
 * Layout conventions follow the reference's Fortran arrays:
 *   - wavefunctions cgjr/gdcr/rhcr(nplwex, nb): real-packed half sphere,
 *     (Re, Im) interleaved for G index 0..nplw, nplwex = 2*(nplw+1)
 *     (Sources/reclat.f90:90), band-major;
 *   - FFT grids fft3x/fft3y/glocal(kfft0), x fastest (Sources/fft3.f90:166-224);
 *   - eqlmr_(nplwex, lmmax, natom), vclalp(lmmax, natom) (Sources/nlkbpp.f90:5-73);
 *   - occupations occ(nb), cell volume rvol (bohr^3).
 * Energies are in Rydberg, as in the reference.
 *
 * Host-array entry points (the Fortran-shaped drop-ins) copy to the device,
 * run, and copy back before returning.  The qx_slot_* / *_slot entry points
 * work on device-resident band storage and are the throughput path.
 */
#ifndef QXMD_AMD_H
#define QXMD_AMD_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct qx_ctx qx_ctx;

enum { QX_FP64 = 0, QX_MIXED = 1 };

const char *qx_last_error(void);
int qx_version(void);

/* ---------------- plane-wave set-up (host) --------------------------------- */
/* get_mesh: FFT grid for a cutoff (Sources/input.f90:1082-1134). hcell(3,3)
 * column-major, lattice vectors as columns. */
int qx_get_mesh(const double *hcell, double ecutdens, int *nd);
/* strvec + set_rciprl: G vectors sorted by |G| with the reference's stable key
 * order (Sources/reclat.f90:500-875, Sources/modules.f90:1518-1540).
 * counts = {num, nplw5, nplw, nplw2, nplwc, nplw7}; arrays need cap >= num+1
 * entries (index 0 is G = 0).  Returns num or -(num+1) if cap is too small. */
int qx_strvec(const double *rba, double ecutdens, double ecut, int cap, int *counts, int *nga,
              int *ngb, int *ngc, double *recnrm, double *gx, double *gy, double *gz);
/* clmrcp (Sources/bulk.f90:556-1050, no double grid): the Ewald reciprocal
 * lattice list for |G|^2 <= dkmax2, b(3,3) the reciprocal vectors without
 * 2 pi (rciprl's output, columns); arrays of cap entries m = 1..numk2 in the
 * reference's order, kcsma(3) the largest |nga|, |ngb|, |ngc|.  Returns
 * numk2, or -numk2 when cap is smaller. */
int qx_clmrcp(const double *b, double dkmax2, int cap, int *nga, int *ngb, int *ngc, double *gx,
              double *gy, double *gz, double *recnrm, int *kcsma);
/* reclat: ijkgd(-nplw5:nplw5), 1-based grid indices (Sources/reclat.f90:259-276). */
int qx_ijkgd(int nplw5, const int *nga, const int *ngb, const int *ngc, const int *nd,
             int *ijkgd);

/* ---------------- context ------------------------------------------------- */
/* One cell / G sphere / FFT grid on the current HIP device.  nga..recnrm are
 * the set_rciprl arrays for G index 0..nplw; nbmax bounds the bands held. */
int qx_create(qx_ctx **ctx, int n1, int n2, int n3, int nplw, const int *nga, const int *ngb,
              const int *ngc, const double *recnrm, int nbmax);
int qx_destroy(qx_ctx *ctx);
int qx_sync(qx_ctx *ctx);
/* releases the process-wide device resources of the context-free drop-ins
 * (rfft3, fft3, xclda, harting, chgr2g, chgg2r, mxchgg: their grid and
 * stream); call after the last qx_destroy and before the process exits, as
 * QXMD's end of run would (the Fortran qxmd_gpu_finalize does).  Later
 * drop-in calls set them up again.  No reference counterpart. */
int qx_shutdown(void);
/* local potential glocal on the FFT grid, Ry (VECTOR build, Sources/engrad.f90:487-495) */
int qx_set_glocal(qx_ctx *ctx, const double *glocal);
/* KB projector tables as setnlc_g leaves them (Sources/nlkbpp.f90:1236-1279):
 * eqlmr_(nplwex, lmmax, natom), lmx per atom, vclalp(lmmax, natom). */
int qx_set_kb(qx_ctx *ctx, const double *eqlmr, int lmmax, int natom, const int *lmx_atom,
              const double *vclalp_atom);
int qx_info(qx_ctx *ctx, int *out /* {ncol, npair, nH, npj, nplwex} */);

/* ---------------- KB inputs built on the device (once per MD step) --------- */
/* The nlpp_variables tables setqlm reads (Sources/ppkb.f90:108-191; allocated
 * at 636-644 and 736-762), column-major with the reference's shapes:
 * lmx(ntype), clalp(lmmax, ntype), tbqlm/tbqlma(0:mxqlm, lnref, ntype),
 * dlqlm(ntype), itola/itolm/itolmf/litoll(lmmax, ntype) (litoll as 0/1),
 * qlmr/qlmi(0:lplh, nqlmmx) for k point 1.  Registered once per cell and
 * pseudopotential set; fails (-2) when the radial table ends below the
 * sphere's largest |G|. */
int qx_set_nlpp_tables(qx_ctx *ctx, int ntype, int lmmax, int lnref, int mxqlm, const int *lmx,
                       const double *clalp, const double *tbqlm, const double *tbqlma,
                       const double *dlqlm, const int *itola, const int *itolm,
                       const int *itolmf, const int *litoll, int nqlmmx, int lplh,
                       const double *qlmr, const double *qlmi);
/* stscfn_pw (Sources/pwlda.f90:5711-5854): structure factors ycos/ysin(0:nplwcs,
 * natom) from ratm(3, natom) (scaled coordinates; real ones when lclust != 0),
 * hcell(3,3), nd1v/nd1vks(3), kmax1..3 and nga/ngb/ngc(0:nplw5).  The result
 * stays on the device for qx_setnlc_g; ycos/ysin (host, may be NULL) receive a
 * copy. */
int qx_stscfn_pw(qx_ctx *ctx, const double *ratm, int natom, const double *hcell,
                 const int *nd1v, const int *nd1vks, int lclust, int kmax1, int kmax2,
                 int kmax3, int nplwcs, int nplw5, const int *nga, const int *ngb,
                 const int *ngc, double *ycos, double *ysin);
/* setnlc_g (Sources/nlkbpp.f90:1236-1279) with setbgq / setqlm (2135-2247):
 * builds eqlmr_ and vclalp = clalp / rvol for the KB types (lkbppi and not
 * lking, as 0/1 ints) on the device, in place of qx_set_kb's upload.  ycos and
 * ysin NULL take the device copy of the last qx_stscfn_pw; vclalp(lmmax, ntype)
 * (may be NULL) receives the scaled energies. */
int qx_setnlc_g(qx_ctx *ctx, int nplw, int ntype, const int *nhk1, const int *nhk2, int natom,
                const int *lkbppi, const int *lking, const double *ycos, const double *ysin,
                int nplwcs, double rvol, double *vclalp);
/* the context's KB table in the reference layout: eqlmr_(nplwex, lmmax, natom)
 * and vclalp per atom (lmmax, natom); zero where no projector is held. */
int qx_get_kb(qx_ctx *ctx, double *eqlmr, double *vclalp_atom);

/* ---------------- device-resident band storage ----------------------------- */
int qx_slot_put(qx_ctx *ctx, int slot, const double *gdcr, int nb);
int qx_slot_get(qx_ctx *ctx, int slot, double *gdcr, int nb);
int qx_slot_copy(qx_ctx *ctx, int src, int dst, int nb);

/* dst = H src for nb bands: kinetic |G|^2 + local (two pruned FFTs) + KB;
 * prec QX_FP64, or QX_MIXED: FP32 FFT stage and FP32 KB products (split
 * bf16 products by default, plain FP32 with QX_MIXED_GEMM=f32), FP64
 * accumulation of the result.  Replaces ekinetic2 +
 * hcnonloc + hckbpp_g (Sources/engrad.f90:66-127, 411-606;
 * Sources/nlkbpp.f90:1284-1393). */
int qx_hpsi_slot(qx_ctx *ctx, int src, int dst, int nb, int prec);
/* local part only: dst (+)= (T + V) src */
int qx_local_slot(qx_ctx *ctx, int src, int dst, int nb, int prec, int accumulate);
/* KB part only: dst += B D B^T W src; slmir(lmmax... packed per projector) optional */
int qx_kb_slot(qx_ctx *ctx, int src, int dst, int nb, double *slmir_host);
/* schmidt on a slot (Sources/schmidt.f90:71-378); *lcholesky = 1 if S was
 * positive definite (psi updated), 0 otherwise (psi untouched). */
int qx_schmidt_slot(qx_ctx *ctx, int slot, int nb, int *lcholesky);
/* subspace rotation: hmatrix2 + RsEIGQR + untrycore on psi and hpsi
 * (Sources/engrad.f90:845-956, Sources/eigen.f90:1171-1514); eig ascending. */
int qx_rotate_slot(qx_ctx *ctx, int psi, int hpsi, int nb, double *eig);
/* rho(r) = sum occ/rvol psi(r)^2 on the FFT grid (softchg, Sources/chgdns.f90:388-454) */
int qx_density_slot(qx_ctx *ctx, int slot, int nb, const double *occ, double rvol, double *rho,
                    int prec);
/* BKL CG band update, lortho=.false. / lhcunt=.true. as the SCF driver calls it
 * (Sources/eigen.f90:199-938, Sources/pwlda.f90:3121).  psi/hpsi are slots 0
 * and 1 (either order; slots 2-4 are the update's work space), updated in
 * place; ekib, eig in; eig, bzan, nriter out (host arrays). */
int qx_eigen_bkl_slot(qx_ctx *ctx, int psi, int hpsi, int nb, const double *ekib, double *eig,
                      int noband, int itermx, int iteremx, double tolera, double threinn,
                      double wdinner, double roughfzan, double roughfeig, double *bzan,
                      int *nriter, int prec);
/* the batched CG captures one iteration as a HIP graph on its first call for
 * a (bands, precision) and replays it for every later iteration (on by
 * default; off = every iteration launched eagerly, the graphs released);
 * qx_cg_graph_count returns the number of captured iterations held. */
int qx_set_cg_graph(qx_ctx *ctx, int on);
int qx_cg_graph_count(qx_ctx *ctx);
/* the same CG update with all bands advanced together: every iteration is one
 * batched H|psi> over the bands still active, and the projection against the
 * other bands (subgs3, Sources/engrad.f90:1072-1142) is done against all bands
 * as two GEMMs.  Same arguments as qx_eigen_bkl_slot without the inner-loop
 * controls threinn / wdinner. */
int qx_eigen_bkl_batched_slot(qx_ctx *ctx, int psi, int hpsi, int nb, const double *ekib,
                              double *eig, int noband, int itermx, int iteremx, double tolera,
                              double roughfzan, double roughfeig, double *bzan, int *nriter,
                              int prec);
/* per-band kinetic energy ekib = 2 <psi|T|psi> (ekinetic2, Sources/engrad.f90:66-127),
 * the preconditioner scale of prpcg */
int qx_ekib_slot(qx_ctx *ctx, int slot, int nb, double *ekib);
/* one SCF band step with frozen potential (energy pass, rotation, CG,
 * Gram-Schmidt, density) -- see Sources/pwlda.f90:2087-3574; psi/hpsi are
 * slots 0 and 1 as for qx_eigen_bkl_slot.  prec QX_MIXED: FP32 FFTs and
 * split bf16x3 products for H psi, the subspace-H, rotation, overlap and CG
 * products; the eigensolver and the Cholesky of the overlap stay FP64. */
int qx_scf_band_step(qx_ctx *ctx, int psi, int hpsi, int nb, const double *occ, int noband,
                     double rvol, double *eig, double *rho, int *lcholesky, int prec);
/* qx_scf_band_step with the CG form chosen: batched != 0 uses
 * qx_eigen_bkl_batched_slot, 0 the band-by-band qx_eigen_bkl_slot */
int qx_scf_band_step_ex(qx_ctx *ctx, int psi, int hpsi, int nb, const double *occ, int noband,
                        double rvol, double *eig, double *rho, int *lcholesky, int prec,
                        int batched);

/* ---------------- timing on the library stream ---------------------------- */
int qx_event_record(qx_ctx *ctx, int slot);
int qx_event_elapsed(qx_ctx *ctx, int slot0, int slot1, float *ms);
/* phase times of the last SCF band step (ms): energy pass, rotation, CG,
 * Gram-Schmidt, density -- the role of the ltimecnt phase timers
 * (Sources/pwlda.f90:1194-1196) */
int qx_scf_phase_ms(qx_ctx *ctx, float *ms5);
/* launch timers with HIP events on the library stream: id 0 = the KB GEMM
 * launches, 1 = the pencil FFT stage of a band batch.  mode 1 clears and
 * starts the timer; mode 0 stops it and returns the summed ms and the number
 * of timed launches.  (Measurement only; no counterpart in the reference.) */
int qx_launch_timer(qx_ctx *ctx, int id, int mode, double *ms, int *count);

/* which kernels run the sizable contractions (M, N >= 256), process-wide:
 * 0 = ours (FP64 MFMA; split bf16x3 products for the mixed form, the
 * default), 1 = hipBLASLt (a cross-check of ours), 2 = ours with plain FP32
 * MFMA for the mixed form.  The environment's QX_GEMM_PATH sets the
 * default.  (No counterpart in the reference: its contractions are loops.) */
int qx_set_gemm_path(int path);
int qx_gemm_path(void);

/* work schedules of our contraction kernels, process-wide: 0 = each product
 * shape takes the faster of stream-K and K-sliced on its first use (the
 * default), 1 = stream-K only, 2 = K-sliced only.  Results of the schedules
 * agree to rounding of the partial-sum order; tests run each.  (No
 * counterpart in the reference.) */
int qx_set_gemm_schedule(int mode);
int qx_gemm_schedule(void);

/* ---------------- Fortran-shaped drop-ins (host arrays) -------------------- */
/* rfft3 (Sources/fft3.f90:113-232): key=1 r2c x1/N with full Hermitian grid
 * out, else c2r from the half grid; ierror = 0 on success. */
int qx_rfft3(int key, double *fft3x, double *fft3y, int kfft1, int kfft2, int kfft3,
             int *ierror);
/* rfft3 on device arrays, allocating nothing: the same transform as qx_rfft3
 * with d_fft3x / d_fft3y (kfft1*kfft2*kfft3 doubles each) in device memory,
 * run on `stream` (a hipStream_t; NULL = the null stream) and waited for.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (tight serving loops, harnesses that forbid hipMalloc); it waits for its
 * stream, so it is not meant for graph capture.  Its complex work grid lives
 * in a fixed device array of the library, one per GPU, held until the
 * process ends: grids up to qx_noalloc_limits() points (2^21 = 128^3) and
 * axis length; larger grids return -2.  Calls that overlap on one GPU share
 * that array and race. */
int qx_rfft3_dev_noalloc(int key, double *d_fft3x, double *d_fft3y, int kfft1, int kfft2,
                         int kfft3, void *stream, int *ierror);
int qx_noalloc_limits(long *max_points, int *max_axis);
/* fft3 (Sources/fft3.f90:237-333): complex, key=1 forward x1/N. */
int qx_fft3(int key, double *fft3x, double *fft3y, int kfft1, int kfft2, int kfft3, int *ierror);
/* hcnonloc for one band with the context's glocal (Sources/engrad.f90:411-606):
 * rhcr += local part of H cgjr (no kinetic). */
int qx_hcnonloc(qx_ctx *ctx, const double *cgjr, double *rhcr);
/* kbnlpp_frc2 without stress (Sources/nlkbpp.f90:1878-2060): the KB
 * Hellmann-Feynman forces of nb bands with occupations occ(nb) are added to
 * fnlc(3, natom) (Ry/bohr); gx/gy/gz(0:nplw) are the Cartesian G of the
 * sphere in the reference's order.  The slot form reads device-resident bands
 * (slot 0, 1 or 5; slots 2-4 are used as work), the host form cgjr(nplwex, nb). */
int qx_kbnlpp_frc2_slot(qx_ctx *ctx, int slot, int nb, const double *occ, const double *gx,
                        const double *gy, const double *gz, double *fnlc);
int qx_kbnlpp_frc2(qx_ctx *ctx, const double *cgjr, int nb, const double *occ, const double *gx,
                   const double *gy, const double *gz, double *fnlc);
/* hckbpp_g (Sources/nlkbpp.f90:1284-1393): rhcr(:, 1:nb) += KB cgjr; slmir
 * (lmmax, natom, nb) written if non-null. */
int qx_hckbpp_g(qx_ctx *ctx, const double *cgjr, double *rhcr, int nb, double *slmir);
/* hckbpp_ib_g (Sources/nlkbpp.f90:1623-1721): one band. */
int qx_hckbpp_ib_g(qx_ctx *ctx, const double *cgjr, double *rhcr);
/* schmidt (Sources/schmidt.f90:71-112) on host cgjr(nplwex, nband). */
int qx_schmidt(qx_ctx *ctx, double *cgjr, int nband, int *lcholesky);
/* RsEIGQR with ifvec=1, ifsrt=2 (Sources/funcs.f90:380-764): a(n,n) column-major,
 * lower triangle used; eigenvectors returned in the columns of a, w ascending.
 * Runs in FP64 on the device; n <= nbmax of the context. */
int qx_rseigqr(qx_ctx *ctx, double *a, int n, double *w);
/* icholesky = rsdecmp + rsinvcd (Sources/schmidt.f90:383-544): s(n,n) column-major
 * lower triangle -> L^-1 (lower, upper zeroed); ok = 1 when positive definite. */
int qx_icholesky(qx_ctx *ctx, double *s, int n, int *ok);
/* softchg (Sources/chgdns.f90:388-454) from host cgjr(nplwex, nb). */
int qx_softchg(qx_ctx *ctx, const double *cgjr, int nb, const double *occ, double rvol,
               double *rho);

/* ---------------- local-pseudopotential forces ----------------------------- */
/* svco(ig, it) of lopp_variables (Sources/ppkb.f90:66-79): local form
 * factors over ig = 0..nplw5 of the density sphere, (nplw5+1) x ntype
 * column-major, and its Cartesian G gx/gy/gz(0:nplw5). */
int qx_set_lopp(qx_ctx *ctx, int ntype, int nplw5, const double *svco, const double *gx,
                const double *gy, const double *gz);
/* flocal_g without stress (Sources/nlkbpp.f90:1037-1125): for the ions
 * k = nhk1(it)..nhk2(it) of the types with llking(it) = 0, ia = iatmpt(k)
 * (1-based), floc(3, ia) = 2 (floc(3, ia) + sum over G != 0 of svco (Re rhgr
 * sin + Im rhgr cos)(G.tau) G).  Host form: the caller's axis tables
 * DCQ1..3 and DSQ1..3 (-kmax:kmax, natom) of stscfn_pw, its G list
 * nga/ngb/ngc(0:nplw5)
 * and rhgr; gx/gy/gz and svco from qx_set_lopp. */
int qx_flocal_g(qx_ctx *ctx, double *floc, int ntype, const int *nhk1, const int *nhk2,
                const int *iatmpt, int natom, const int *llking, int nplw5, const int *nga,
                const int *ngb, const int *ngc, const double *rhgr, int kmax1, int kmax2,
                int kmax3, const double *dcq1, const double *dcq2, const double *dcq3,
                const double *dsq1, const double *dsq2, const double *dsq3);
/* device-resident form: the density sphere of qx_scf_set_grid (the mixed
 * rhgr of the last qx_pwpmxchgg, or as put), the axis tables of the last
 * qx_stscfn_pw (called with the nplw5 G list), svco of qx_set_lopp. */
int qx_flocal_g_dev(qx_ctx *ctx, int ntype, const int *nhk1, const int *nhk2,
                    const int *iatmpt, const int *llking, int natom, double *floc);
/* vlocal_g (Sources/nlkbpp.f90:862-1032, caller Sources/pwlda.f90:1495): the
 * local pseudopotential on the grid from the svco of qx_set_lopp and the
 * device axis tables of the last qx_stscfn_pw over the density sphere, into
 * the external potential the SCF glue adds (qx_scf_set_grid's vext):
 * vext = (accumulate ? vext : 0) + FFT^-1[ sum_type svco(G) / rvol
 * conj(S_type(G)) ], S_type the structure factor of the type's ions (llking
 * types skipped), G = 0 term nhk svco(0) / rvol.  Optional host copies: vext
 * (FFT order, N) and thrhgr (2 (nplw5 + 1), its G = 0 entry zero). */
int qx_vlocal_g(qx_ctx *ctx, int ntype, const int *nhk1, const int *nhk2, const int *iatmpt,
                const int *llking, int natom, double rvol, int accumulate, double *vext,
                double *thrhgr);

/* ---------------- restart files (the reference's own layout) ---------------- */
/* svwfns (Sources/savedd.f90:2441-2810), one spin, ncscale = 1: QM_eig with
 * gdcr(nplwex, nband) in the reference's G order, eig, occ; lsreal8 = 0 stores
 * 16-bit integers with a per-band scale, as the reference does.  Byte-equal to
 * the reference's file. */
int qx_svwfns(const char *fname, const double *gdcr, int nplwex, int nplw, int nband,
              const double *eig, const double *occ, int lsreal8);
/* rdwfns / rdwfn22 (Sources/savedd.f90:268-1172): the file's nplw must match;
 * min(nband, file's nband) bands read; *nband_file = the file's band count. */
int qx_rdwfns(const char *fname, double *gdcr, int nplwex, int nplw, int nband, double *eig,
              double *occ, int *nband_file);
/* the same from / into a device slot (restart without host band arrays) */
int qx_svwfns_slot(qx_ctx *ctx, const char *fname, int slot, int nb, const double *eig,
                   const double *occ, int lsreal8);
int qx_rdwfns_slot(qx_ctx *ctx, const char *fname, int slot, int nb, double *eig, double *occ,
                   int *nband_file);
/* svacon / rdacon (Sources/savedd.f90:3058-3143, 40-265), file version -2:
 * QM_ion with nhk1/nhk2(ntype), ratm and frc (3, natom), prevr (3, natom, 3),
 * hcell (3, 3), column-major; natom = nhk2(ntype).  rdacon checks ntype and
 * nhk1/nhk2 against the file as the reference does. */
int qx_svacon(const char *fname, int ntype, const int *nhk1, const int *nhk2, int nstepCG,
              int nstepMD, int lclust, const double *ratm, int ifmd, const double *frc,
              const double *prevr, const double *hcell);
int qx_rdacon(const char *fname, int ntype, const int *nhk1, const int *nhk2, int *nstepCG,
              int *nstepMD, int *lclust, double *ratm, int *ifmd, double *frc, double *prevr,
              double *hcell);

/* ---------------- velocity-Verlet updates of the MD loop ------------------- */
/* Host arithmetic (O(natom)), bit-identical to the reference.  Scaled
 * coordinates: x(3, n) in cell units, v the displacement per MD step, a the
 * scaled acceleration with acon = dt^2 / (2 m) per species
 * (Sources/remd.f90:3568-3571); is(n) the 1-based species of each atom;
 * lfixion(ntype) != 0 fixes a species.  One constant-energy step is
 * qx_normalized_a + qx_vkick after the forces (remd_semifinal,
 * Sources/remd.f90:633-790), then qx_vkick + qx_updtps + qx_updtps_pbc at the
 * next step's start (540-616). */
/* rciprl (Sources/reclat.f90:1089-1128): b = reciprocal vectors / 2 pi of the
 * cell a (3 x 3 column-major, a_i . b_j = delta_ij), *v = |det a| */
int qx_rciprl(const double *a, double *b, double *v);
/* normalized_a (Sources/remd.f90:6789-6830): forces a(3, n) -> acon(is) hi^T a
 * in place, hi = rciprl(h), h the cell (3 x 3 column-major) */
int qx_normalized_a(int n, int ntype, double *a, const int *is, const double *h,
                    const double *acon);
/* vkick (Sources/remd.f90:5408-5430): v += a for unfixed species */
int qx_vkick(double *v, const double *a, const int *is, int n, int ntype, const int *lfixion);
/* updtps (Sources/remd.f90:5755-5783): x += v for unfixed species */
int qx_updtps(double *x, const double *v, const int *is, int n, int ntype, const int *lfixion);
/* updtps_pbc (Sources/remd.f90:5786-5806): x into [0, 1); -2 if some |x| does
 * not fit the reference's default INT */
int qx_updtps_pbc(double *x, int n);

/* ---------------- wave-function extrapolation between MD steps ------------- */
/* set_coef_ASPC with ichest = 0 (Sources/sbalin.f90:77-170): predictor of
 * order ihest (kaspc = ihest + 1), history of nprvmx sets (<= 10) in HBM,
 * keypwv = 0.  Resets the history. */
int qx_aspc_init(qx_ctx *ctx, int nprvmx, int ihest);
/* sbalinaspc for one spin without ultrasoft terms (Sources/sbalin.f90:
 * 369-712) on the bands of slot 0 or 1: roteig2, the ASPC combination, the
 * history shift, and Gram-Schmidt when ihest > 0.  info = {iaspc (0: the
 * history was only started), lcholesky} if non-null. */
int qx_sbalinaspc_slot(qx_ctx *ctx, int slot, int nb, int nstep, int nstep_ini, int *info);
/* roteig2 (Sources/sbalin.f90:772-1373): history sets 1..min(nprvmx, kpwvx)
 * turned to the bands of the slot. */
int qx_roteig_slot(qx_ctx *ctx, int slot, int nb, int kpwvx);
/* read / write history set k (0-based) in the reference's G order; put sets
 * keypwv. */
int qx_aspc_get(qx_ctx *ctx, int k, int nb, double *gdcr);
int qx_aspc_put(qx_ctx *ctx, int k, int nb, const double *gdcr, int keypwv);

/* ---------------- Ewald reciprocal-space forces ----------------------------- */
/* The Ewald G list of clmrcp (Sources/bulk.f90:556-1068) and para_ewald:
 * nga/ngb/ngc, gx/gy/gz, recnrm over m = 1..numk2 (arrays from index 0),
 * recnrmex(0:numk2), gamma, volume.  Once per cell. */
int qx_set_ewald(qx_ctx *ctx, int numk2, const int *nga, const int *ngb, const int *ngc,
                 const double *gx, const double *gy, const double *gz, const double *recnrm,
                 const double *recnrmex, double gamma, double volume);
/* frcrec without stress (Sources/bulk.f90:1880-2026): fclm(3, natom),
 * *ecolmb and atom_ecolmb(natom) accumulate the reciprocal-space Ewald
 * forces and energies of the ions k = 1..nion (atom iatmpt(k), type
 * iatoit(atom), all 1-based).  Host form: the per-ion axis tables
 * DCQ1..3 and DSQ1..3 (-kcsma:kcsma, nion) of the Ewald stscfn
 * (Sources/bulk.f90:2086). */
int qx_frcrec(qx_ctx *ctx, double *fclm, int ntype, const int *nhk1, const int *nhk2,
              const double *zv, int nel, double *ecolmb, double *atom_ecolmb, const int *iatoit,
              const int *iatmpt, int natom, int nion, int kcsma1, int kcsma2, int kcsma3,
              const double *dcq1, const double *dcq2, const double *dcq3, const double *dsq1,
              const double *dsq2, const double *dsq3);
/* device-resident form: the per-atom axis tables of the last qx_stscfn_pw
 * (its kmax must cover the Ewald G list). */
int qx_frcrec_dev(qx_ctx *ctx, double *fclm, int ntype, const int *nhk1, const int *nhk2,
                  const double *zv, int nel, double *ecolmb, double *atom_ecolmb,
                  const int *iatoit, const int *iatmpt, int natom, int nion);

/* ---------------- SCF glue between band updates ----------------------------
 * Sphere arrays: rhgr(0:2*nplw5+1) real-packed (Re, Im per G, G index
 * 0..nplw5); ijkgd points at ijkgd(-nplw5) of the array ijkgd(-nplw5:nplw5)
 * (1-based grid offsets, as set_rciprl builds it); recnrmex(0:nplw5).
 * Grid arrays in FD order through mfd2ft(ntotfd) (1-based FFT offsets; null
 * means FFT order, ntotfd = kfft1 kfft2 kfft3). */
/* xclda (Sources/vxc.f90:961-1017): LDA (Perdew-Zunger) vexc(mshnod) and
 * eeexc(mshnod) in Ry; e4 = {eexc, evexc, ecor, evcor}.  rhocore may be null;
 * negative rhocore entries mean no core charge there. */
int qx_xclda(const double *rho, const double *rhocore, int mshnod, double rdelv, double *vexc,
             double *eeexc, double *e4);
/* harting (Sources/engrad.f90:1673-1743): vhar = c2r of 2 recnrmex rhgr. */
int qx_harting(double *vhar, const double *rhgr, int nplw5, const double *recnrmex,
               const int *mfd2ft, int ntotfd, const int *nd1vks, const int *ijkgd);
/* chgr2g (Sources/engrad.f90:1253-1297): rho -> rhgr (forward x1/N). */
int qx_chgr2g(const double *rho, double *rhgr, int nplw5, const int *mfd2ft, int ntotfd,
              const int *nd1vks, const int *ijkgd);
/* chgg2r (Sources/engrad.f90:1302-1353): rhgr -> rho. */
int qx_chgg2r(double *rho, const double *rhgr, int nplw5, const int *mfd2ft, int ntotfd,
              const int *nd1vks, const int *ijkgd);
/* mxchgg_alloc + set_g2ming2max_in_chgdns (Sources/chgdns.f90:529-690):
 * imxchg 0 (none), 1 (Pulay, itratn <= 32) or 3 (simple); dkgnrm(nplwcex);
 * rvol is the cell volume (the reference's rvol).
 * Resets the mixing history of the context. */
int qx_mxchgg_alloc(qx_ctx *ctx, int nplwcex, int imxchg, int itratn, double cmetric,
                    double wkerker, double g2min, double g2max, const double *dkgnrm,
                    double rvol);
/* mxchgg (Sources/chgdns.f90:796-1210) on host arrays: rhgr (output density
 * in, mixed out), rinr (input density in, = rhgr out), glocal = chgg2r(rhgr);
 * *lreset cleared.  The history stays on the device between calls. */
int qx_mxchgg(qx_ctx *ctx, int *lreset, double *rhgr, double *rinr, int nplw5, double aslh,
              double bslh, double *glocal, const int *mfd2ft, int ntotfd, const int *ijkgd,
              int lmixhold);
/* Device-resident SCF glue.  qx_scf_set_grid fixes the density sphere, the
 * Hartree kernel recnrmex, the external (local pseudopotential) potential
 * vext and the core charge (null: none), grid arrays in FFT order. */
int qx_scf_set_grid(qx_ctx *ctx, int nplw5, const int *ijkgd, const double *recnrmex,
                    const double *vext, const double *rhocore, double rdelv);
/* softchg + outchg (Sources/chgdns.f90:29-383): density occ/rvol psi^2 of nb
 * bands of a slot (rvol = cell volume), then its rhgr, kept on the device. */
int qx_pwpchg_slot(qx_ctx *ctx, int slot, int nb, const double *occ, double rvol, int prec);
/* outchg (Sources/chgdns.f90:248-383): rhgr of the density the context holds
 * (left by qx_scf_band_step, qx_density_slot or qx_pwpchg_slot). */
int qx_outchg(qx_ctx *ctx);
/* pwpmxchgg2 (Sources/chgdns.f90:724-790): mix the device rhgr with rinr,
 * then rho = chgg2r(rhgr); info = {kntchg, history resets} if non-null. */
int qx_pwpmxchgg(qx_ctx *ctx, int lreset, double aslh, double bslh, int lmixhold, int *info);
/* harting + xclda + pwpunifylc2 (Sources/pwlda.f90:1640-2030): the local
 * potential of the next band update, V = vhar + vext + vexc, written into the
 * context (as qx_set_glocal does); ene = {ehar, elcl, eexc, evexc, ecor,
 * evcor} (Sources/pwlda.f90:1868-1873, 1963-1968). */
int qx_scf_potential(qx_ctx *ctx, double *ene);
/* read (write) the device SCF arrays: 0 rho, 1 rhgr, 2 rinr, 3 vhar, 4 vexc,
 * 5 eeexc, 6 glocal (FFT order), 7 Pulay coefficients (itratn, read only);
 * only 0-2 can be written. */
int qx_scf_get(qx_ctx *ctx, int what, double *host);
int qx_scf_put(qx_ctx *ctx, int what, const double *host);

/* ---------------- Bloch bands at several k-points ------------------------- */
/* The reference's band algebra is Gamma-only (Sources/kpoints.f90:33-39; its
 * k-point input block, Sources/input.f90:2016-2126, is commented out); these
 * extend its H|psi> (hcnonloc + ekinetic, Sources/engrad.f90:411-606, 66-127;
 * hckbpp_g, Sources/nlkbpp.f90:1284-1393) to complex Bloch bands c_nk(G),
 * |k+G|^2 <= ecut, on the context's FFT grid and local potential
 * (qx_set_glocal).  No reference interface exists for them.
 *
 * qx_kpoints_set: nk k-points kfrac (3 x nk, units of the reciprocal vectors),
 * bvec the reciprocal vectors (3 x 3 column-major, columns b_i with 2 pi),
 * ecut in Ry, nb bands per k-point; npw(nk) receives the sphere sizes.
 * Allocates four k-slots of nk x nb bands. */
int qx_kpoints_set(qx_ctx *ctx, int nk, const double *kfrac, const double *bvec, double ecut,
                   int nb, int *npw);
/* The sphere of k-point kfrac on grid nd(3) without a context: hkl(3, npw),
 * kin(npw) as qx_kpoints_sphere; returns npw (-npw when cap is smaller; -1 on
 * error). */
int qx_ksphere(const double *bvec, const double *kfrac, double ecut, const int *nd, int cap,
               int *hkl, double *kin);
/* The sphere of k-point ik in its storage order (points grouped by (h, k)
 * column, columns in lexicographic order, l rising): hkl(3, npw), kin(npw) =
 * |k+G|^2. */
int qx_kpoints_sphere(qx_ctx *ctx, int ik, int *hkl, double *kin);
/* Bands 0..nb-1 of k-point ik in / out of a k-slot: psi(2 npw_ik, nb),
 * complex interleaved. */
int qx_kslot_put(qx_ctx *ctx, int slot, int ik, const double *psi, int nb);
int qx_kslot_get(qx_ctx *ctx, int slot, int ik, double *psi, int nb);
/* KB projectors of k-point ik: beta(2 npw_ik, npj) the coefficients
 * beta_p(k+G) (complex interleaved), D(npj) their energies (vclalp);
 * <beta_p|c> = sum_G conj(beta_p(G)) c(G).  Every k-point takes the same npj. */
int qx_kset_kb(qx_ctx *ctx, int ik, int npj, const double *beta, const double *D);
/* H_k c for bands 0..nb-1 of every k-point, k-slot src -> k-slot dst.
 * per_band = 0: each stage over all (k, band) blocks in one launch; 1: one
 * launch per (k, band) block (the per-band loop the batched form replaces). */
int qx_khpsi_slot(qx_ctx *ctx, int src, int dst, int nb, int per_band);
/* Lowest nb eigenpairs of H_k at every k-point in the context's local
 * potential (the band update of the k-point form; the reference has none):
 * block Rayleigh-Ritz steps on [psi, W] with W the residuals preconditioned
 * as prpcg does (Sources/eigen.f90:1093-1113) and scaled to unit norm, the
 * small generalized Hermitian problems of all k-points in one batched
 * rocSOLVER solve.  Starts from the bands of k-slot src; leaves the Ritz
 * vectors in src, H of them in dst, eig(nb, nk) the Ritz values (Ry), *iters
 * the steps taken; stops when no Ritz value moves by more than tol or after
 * max_iter steps.  Returns -21 if a basis overlap is not positive definite. */
int qx_ksolve_slot(qx_ctx *ctx, int src, int dst, int max_iter, double tol, double *eig,
                   int *iters);

#ifdef __cplusplus
}
#endif
#endif
