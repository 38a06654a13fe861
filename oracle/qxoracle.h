/*
 * qxoracle.h -- CPU FP64 oracle for the QXMD Gamma-point H|psi> / band-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product (qxmd_amd/) never links or calls it.
 *
 * Conventions: all arrays use the reference's Fortran layouts, 0-based in C.
 *   - Wavefunctions are real-packed half-sphere vectors of nplwex = 2*(nplw+1)
 *     doubles, (Re, Im) interleaved for G index 0..nplw (Sources/reclat.f90:90).
 *     Band-major blocks: psi(ig, ib) at psi[ig + nplwex*ib].
 *   - FFT grids are x-fastest: ijk = i + n1*(j + n2*k) (Sources/fft3.f90:166-224).
 *   - ijkgd maps are the reference's 1-based values; ijkgd_c points at the m=0
 *     entry so ijkgd_c[m] is valid for -nplw5 <= m <= nplw5 (Sources/reclat.f90:259-276).
 */
#ifndef QXORACLE_H
#define QXORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- lattice / grid rule ---------------------------------------------- */
void qxo_rciprl(const double a[9], double b[9], double *v);
int  qxo_fft_size(int n, int first_dim);
void qxo_get_mesh(const double hcell[9], double ecutdens, int nd[3]);

/* ---- G sphere (strvec + vsgar/vsoe + set_rciprl) ------------------------ */
int  qxo_strvec_capacity(const double rba[9], double ecutdens);
int  qxo_strvec(const double rba[9], double ecutdens, double ecut, double ecutsoft,
                double ecutc, double ecutorth, int cap, int counts[6],
                int *nga, int *ngb, int *ngc, double *gx, double *gy, double *gz,
                double *recnrm, double cross[9], double *rvol);
void qxo_ijkgd(int nplw5, const int *nga, const int *ngb, const int *ngc,
               const int nd[3], int *ijkgd /* 2*nplw5+1 entries */);
void qxo_dkgnrm(int nplwc, const double *recnrm, double *dkgnrm);
double qxo_rnd00(double *seed);

/* ---- FFT (FFTW3 semantics of rfft3 / fft3) ------------------------------ */
void qxo_cfft3(int sign, double *xy /* complex interleaved, N */, int n1, int n2, int n3);
void qxo_rfft3(int key, double *fft3x, double *fft3y, int n1, int n2, int n3);
void qxo_fft3(int key, double *fft3x, double *fft3y, int n1, int n2, int n3);

/* ---- local H|psi>, kinetic ---------------------------------------------- */
void qxo_wvg2r(const double *cgjr, int nplw, const int *ijkgd_c, int n1, int n2, int n3,
               double *fft3x, double *fft3y);
void qxo_hcnonloc(const double *glocal, const double *cgjr, double *rhcr, int nplw,
                  const int *ijkgd_c, int n1, int n2, int n3, double *fft3x, double *fft3y);
void qxo_ekinetic(const double *dkgnrm, const double *cgjr, double *rhcr, int nplwex, int nb,
                  const double *occ, double *ekib, double *ekint);

/* ---- Kleinman-Bylander nonlocal term ------------------------------------ */
void qxo_hckbpp_g(const double *eqlmr, int nplwex, int lmmax, int natom, const int *lmx_atom,
                  const double *vclalp_atom, const double *cgjr, double *rhcr, int nb,
                  double *slmir);
void qxo_stscfn_pw(const double *ratm, const double *hcell, const int *nd1v, const int *nd1vks,
                   int lclust, int natom, int kmax1, int kmax2, int kmax3, int nplwcs, int nplw5,
                   const int *nga, const int *ngb, const int *ngc, double *ycos, double *ysin);
void qxo_setnlc_g(int nplw, int ntype, const int *nhk1, const int *nhk2, int natom,
                  const int *lkbppi, const int *lking, const double *ycos, const double *ysin,
                  int nplwcs, double rvol, int lmmax, int lnref, int mxqlm, const int *lmx,
                  const double *clalp, const double *tbqlm, const double *tbqlma,
                  const double *dlqlm, const int *itola, const int *itolm, const int *itolmf,
                  const int *litoll, int nqlmmx, const double *qlmr, const double *qlmi,
                  const double *recnrm, double *eqlmr, double *vclalp);
void qxo_kbnlpp_frc2(const double *eqlmr, int nplwex, int lmmax, int natom, const int *lmx_atom,
                     const double *vclalp_atom, const double *cgjr, int nb, const double *occ,
                     const double *slmir, const double *gx, const double *gy, const double *gz,
                     double *fnlc);
void qxo_xclda(const double *rho, const double *rhocore, int n, double rdelv, double *vexc,
               double *eeexc, double *e4);
void qxo_harting(const double *rhgr, int nplw5, const double *recnrmex, const int *ijkgd_c,
                 int n1, int n2, int n3, double *vhar);
void qxo_chgr2g(const double *rho, int nplw5, const int *ijkgd_c, int n1, int n2, int n3,
                double *rhgr);
void qxo_chgg2r(const double *rhgr, int nplw5, const int *ijkgd_c, int n1, int n2, int n3,
                double *rho);
int qxo_invers(const double *w, double *wi, int n, int mx);
int qxo_mxchgg_pulay(double *rhgr, double *rinr, int nplwcex, int nplw5ex, const double *dkgnrm,
                     double rvol, double aslh, double bslh, double factq1, double wkerker,
                     int lreset, int lmixhold, int itratn, int *kntchg_io, double *unr,
                     double *dfmr, double *aij, double *cmk_out);
void qxo_mxchgg_simple(double *rhgr, double *rinr, int nplw5ex, double aslh, int none);
double qxo_vlocal_g(int ntype, const int *nhk1, const int *nhk2, const int *iatmpt,
                    const int *llking, int natom, int nplw5, const int *nga, const int *ngb,
                    const int *ngc, const double *svco, double rvol, const int *kmax,
                    const double *const *dcq, const double *const *dsq, double *thrhgr);
void qxo_flocal_g(int ntype, const int *nhk1, const int *nhk2, const int *iatmpt,
                  const int *llking, int natom, int nplw5, const int *nga, const int *ngb,
                  const int *ngc, const double *gx, const double *gy, const double *gz,
                  const double *rhgr, const double *svco, const int *kmax,
                  const double *const *dcq, const double *const *dsq, double *floc);
void qxo_frcrec(int ntype, const int *nhk1, const int *nhk2, const double *zv, int nel,
                int natom, int nion, const int *iatoit, const int *iatmpt, int numk2,
                const int *nga, const int *ngb, const int *ngc, const double *gx, const double *gy,
                const double *gz, const double *recnrm, const double *recnrmex, double gamma,
                double volume, const int *kcsma, const double *const *dcq,
                const double *const *dsq, double *fclm, double *ecolmb, double *atom_ecolmb);
int qxo_roteig(const double *cur, double *prev, int nplwex, int nband, int nprvmx, int kpwvx,
               double *u_out);
void qxo_coef_aspc(int kaspc, double *coef);
int qxo_sbalinaspc(double *gdcr, int nplwex, int nband, double *prveig, int nprvmx,
                   const double *coefaspc, int kaspc, int ihest, int nstep, int nstep_ini,
                   int *keypwv);
double qxo_enkbpp_g(const double *slmir, int lmmax, int natom, const int *lmx_atom,
                    const double *clalp_atom, const double *occ, int nb, double rvol);

/* ---- dense band algebra -------------------------------------------------- */
void qxo_rsdecmp(double *sr, int n, int lda);
void qxo_rsinvcd(double *sr, int n, int lda, double *work);
int  qxo_schmidt(double *cgjr, int nplwex, int nband, double *smat /* nband*nband */);
void qxo_hmatrix(const double *cgjr, const double *rhcr, int nplwex, int nband,
                 double *dmtrxr, double *eig);
int  qxo_eigh(double *a, int n, int lda, double *w);
void qxo_untrycore(double *gdcr, int nplwex, int nband, const double *u, int ldu, double *work);
void qxo_softchg(const double *cgjr, int nplwex, int nplw, int nband, const double *occ,
                 double rvol, const int *ijkgd_c, int n1, int n2, int n3, double *rho);
void qxo_prpcg(double ekib, const double *dkgnrm, int nplwex, double *prcd);

/* ---- BKL conjugate-gradient band solver ---------------------------------- */
typedef struct {
  int nplw, nplwex, n1, n2, n3;
  const int *ijkgd_c;        /* centered map, valid for -nplw..nplw        */
  const double *glocal;      /* local potential on the FFT grid (Ry)       */
  const double *dkgnrm;      /* |G|^2 duplicated for (Re, Im)              */
  const double *eqlmr;       /* KB projectors (nplwex, lmmax, natom) or 0  */
  int lmmax, natom;
  const int *lmx_atom;
  const double *vclalp_atom; /* (lmmax, natom): clalp/Omega              */
} qxo_ham;

void qxo_hpsi(const qxo_ham *h, const double *psi, double *hpsi, int nb, double *slmir);
/* test hook: per-iteration trace of qxo_eigen_bkl (see qxoracle.c) */
#define QXO_BKL_TRACE_W 17
void qxo_bkl_trace(double *rec, double *vec, int cap);
int qxo_bkl_trace_count(void);
void qxo_eigen_bkl(const qxo_ham *h, double *eig, double *gdcr, double *rhcr,
                   const double *ekib, int nband, int noband, int itermx, int iteremx,
                   double tolera, double threinn, double wdinner, double roughfzan,
                   double roughfeig, double *bzan, int *nriter);

/* One SCF iteration of the band part with a frozen local potential:
 * energy pass, subspace rotation, BKL CG, Gram-Schmidt, density. */
int qxo_scf_band_step(const qxo_ham *h, double *gdcr, double *rhcr, double *eig,
                      const double *occ, int nband, int noband, double rvol,
                      double *rho, double *ekin, double *enl, const double *clalp_atom);

#ifdef __cplusplus
}
#endif
#endif
