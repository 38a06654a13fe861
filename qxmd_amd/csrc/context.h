// context.h -- device-resident plane-wave context of the qxmd_amd library.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>
#include "common.h"

namespace qx {

enum Prec { PREC_FP64 = 0, PREC_FP32 = 1 };
constexpr int kBbChunk = 256;  // vector elements per chunk of the single-band CG passes
// bf16 split work arrays of the GEMM layer: 0 psi of the fused mixed
// band-apply, 1 its KB projections, 2 and 3 band slots of the mixed band step,
// 4 the nb x nb factor of a band product
constexpr int kU16Scratch = 5;

// Pencil layout of the half sphere.  Columns are the (h,k) lines along z that
// the full sphere (G and -G) touches; they come in mirror pairs (c, -c), with
// the self-mirror column (0,0) alone as pair 0.  Internal point order:
// pair 0 (G=0 first), then for each pair q>=1 the points of column 2q-1 and
// then those of column 2q.  Pair q>=1 owns columns 2q-1 and 2q.
struct Pencil {
  int ncol = 0, npair = 0, nH = 0, hmin = 0;
  std::vector<int> perm;        // internal point p -> reference index g
  std::vector<int> pt_col;      // column holding +G of point p
  std::vector<int> pt_z;        // l mod n3 of point p
  std::vector<int> pair_p0;     // first point of pair q (size npair+1)
  std::vector<int> col_hi;      // h - hmin of column c
  std::vector<int> col_kmod;    // k mod n2 of column c
};

// dense complex 3-D FFT over the whole grid (fullgrid.hip): lines along x,
// y and z transformed in LDS; d_g holds the N complex values, x fastest
struct FullGrid {
  int n1 = 0, n2 = 0, n3 = 0;
  FftPlan p[3] = {};
  double* d_tw[3] = {nullptr, nullptr, nullptr};
  double2* d_g = nullptr;
  double *d_x = nullptr, *d_y = nullptr;  // host-array staging of the drop-ins
};
int fg_prepare(FullGrid& f, int n1, int n2, int n3, bool staging);
void fg_release(FullGrid& f);
// sign -1: forward (exp(-i)), +1: backward; scale applied on the last pass
int fg_transform(FullGrid& f, hipStream_t st, int sign, double scale);
// the grid and stream of the context-free drop-ins (rfft3, fft3, harting, ...)
int dropin_grid(int n1, int n2, int n3, FullGrid** f, hipStream_t* st);
int dropin_stream(hipStream_t* st);
void dropin_release();

struct Ctx {
  int dev = 0;
  int n1 = 0, n2 = 0, n3 = 0;
  long N = 0;
  int nplw = 0, nplwex = 0;
  int nbmax = 0;
  hipStream_t stream = nullptr;
  Pencil pc;
  FftPlan plx{}, ply{}, plz{};
  // device arrays
  int *d_perm = nullptr, *d_pt_col = nullptr, *d_pt_z = nullptr, *d_pair_p0 = nullptr;
  int *d_col_hi = nullptr, *d_col_kmod = nullptr;
  double *d_kin = nullptr;        // |G_p|^2 per internal point
  double *d_tw = nullptr;         // twiddles: x (n1), y (n2), z (n3), exp(+2pi i k/n)
  double *d_V = nullptr;          // local potential, FFT order (N)
  float *d_Vf = nullptr;          // FP32 copy
  double *d_Vp = nullptr;         // slot-permuted copies for the fast pencil path
  float *d_Vpf = nullptr;
  // KB projectors, internal order: B[p*nplwex + i], p = projector
  int npj = 0;
  double *d_B = nullptr;
  double *d_D = nullptr;          // vclalp per projector
  std::vector<int> proj_atom, proj_l;
  int lmmax = 0, natom = 0;
  // band storage slots, internal order: slot s, band b at d_slot[s] + b*nplwex
  static constexpr int kSlots = 6;
  double *d_slot[kSlots] = {};
  // scratch
  void *d_t1 = nullptr, *d_t2 = nullptr;  // pencil intermediates
  size_t t_bytes = 0;
  int batch_bp = 0;                       // band pairs per pencil batch
  double *d_P = nullptr;                  // projections npj x nbmax
  double *d_small = nullptr;              // nbmax x nbmax x 4 dense work
  double *d_rho_part = nullptr;           // density partial planes
  int rho_groups = 0;
  double *d_rho = nullptr;                // density (N)
  double *d_scal = nullptr;               // scalar work (device)
  int* d_sync = nullptr;                  // grid-barrier counter and error flags
  double* d_dwork = nullptr;              // dense eigensolver work (8 nbmax^2 + ...)
  double* d_bb = nullptr;                 // band-by-band CG: scalars, then chunk sums
  int bb_nch = 0;                         // chunks of kBbChunk the chunk-sum space holds
  int *h_flag = nullptr, *d_hflag = nullptr;  // pinned host flag and its device address
  size_t dwork_len = 0;
  hipEvent_t ev[8] = {};
  hipEvent_t pev[6] = {};                 // SCF band-step phase timers
  float phase_ms[5] = {};
  // launch timers (HIP events on the context stream), off unless enabled
  static constexpr int kTimers = 2;       // 0: KB GEMM launches, 1: pencil FFT stage
  struct LaunchTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;           // begin/end pairs
    std::vector<unsigned char> cont;      // per pair: 1 = continues the previous launch
    bool next_cont = false;
    size_t used = 0;
  } lt[kTimers];
  // nonlocal-pseudopotential tables of setqlm (qx_set_nlpp_tables) and the
  // structure factors of the last qx_stscfn_pw, on the device
  struct Nlpp {
    int ntype = 0, lmmax = 0, lnref = 0, mxqlm = 0, nqlmmx = 0, lplh = 0;
    std::vector<int> lmx, itola, itolm, itolmf, litoll;  // (lmmax x ntype) column-major
    std::vector<double> clalp, dlqlm;
    double *d_tbqlm = nullptr, *d_tbqlma = nullptr;      // (mxqlm+1) x lnref x ntype
    double *d_qlm = nullptr;                             // qlmr then qlmi, (lplh+1) x nqlmmx each
  } nl;
  double *d_ycos = nullptr, *d_ysin = nullptr;  // (nplwcs+1) x natom
  int sf_natom = 0, sf_nplwcs = 0;
  double gmax2 = 0.0;                           // largest |G|^2 of the sphere
  // axis tables c_k + i s_k of the last qx_stscfn_pw (per atom and axis,
  // k = -km..km at (ia * 3 + a) * ld + km_a + k; sines after 3 natom ld
  // cosines) and its G list nga | ngb | ngc over 0..ngmax
  struct AxisTables {
    double* d_tab = nullptr;
    int* d_g = nullptr;
    int ld = 0, km[3] = {0, 0, 0}, ngmax = -1, natom = 0;
  } sf;
  // local pseudopotential form factors svco(ig, it), ig = 0..nplw5, and the
  // Cartesian G of the density sphere, for flocal_g
  // factors of a device-built KB table for the one-band apply (kb1.hip):
  // phases (natom x nplw1 complex), z rows per (type, l), projector ranges
  struct Kb1 {
    bool valid = false;
    double *d_ph = nullptr, *d_z = nullptr;
    int *d_p0 = nullptr, *d_pcls = nullptr, *d_ppar = nullptr;
    int ncls = 0, natom = 0, nplw1 = 0, npj = 0, apg = 1;
  } kb1;
  struct Lopp {
    int ntype = 0, nplw5 = -1;
    double *d_svco = nullptr, *d_g = nullptr;  // d_g: gx | gy | gz
    double *d_thr = nullptr, *d_zin = nullptr; // vlocal_g: thrhgr and the transform input
    void* d_ions = nullptr;                    // vlocal_g: ion list (capacity natom)
    int ion_cap = 0;
  } lopp;
  // Ewald reciprocal-space G list (clmrcp) with hep = (exp(-(pi/gamma)^2 g^2)
  // - recnrmex) / g^2 per G, for frcrec
  struct Ewald {
    int numk2 = -1, km[3] = {0, 0, 0};
    int* d_g = nullptr;                         // nga | ngb | ngc
    double *d_gv = nullptr, *d_hep = nullptr;   // gx | gy | gz, hep
    double gamma = 0.0, volume = 0.0, recnrmex0 = 0.0;
  } ew;
  // wave-function extrapolation (sbalinaspc): nprvmx history sets of
  // nplwex x nbmax in internal order, logical set i at d_prv[i]
  struct Aspc {
    static constexpr int kMax = 10;
    int nprvmx = 0, kaspc = 0, ihest = 0, keypwv = 0;
    std::vector<double> coef;                 // coefaspc, kaspc x kaspc column-major
    double* d_prv[kMax] = {};
    double* d_work = nullptr;                 // 8 nbmax^2 + nbmax + 64
  } aspc;
  // SCF glue on the dense grid (scfglue.hip): the density sphere, Hartree
  // kernel, fixed potentials, the charge-density vectors and the mixer history
  FullGrid fg;
  struct Scf {
    int nplw5 = 0;
    int* d_ijk = nullptr;         // ijkgd(-nplw5:nplw5) as 0-based grid offsets
    double* d_p4 = nullptr;       // 2 recnrmex(ig), ig = 0..nplw5
    double *d_vext = nullptr, *d_core = nullptr;  // FFT order; d_core null: no core
    double rdelv = 0.0;
    double *d_rhgr = nullptr, *d_rinr = nullptr;  // 2 (nplw5+1)
    double *d_vhar = nullptr, *d_vexc = nullptr, *d_eeexc = nullptr;
    double* d_red = nullptr;      // reduction partials and results
    bool z_vhar = false;          // fg.d_g holds rho + i vhar of the current rhgr
    // mixer (module mxchgg_variables)
    int imxchg = -1, itratn = 0, nplwcex = 0, kntchg = 0;
    double rvol = 0.0, factq1 = 0.0, wkerker = 1.0;
    double *d_unr = nullptr, *d_dfmr = nullptr;   // nplwcex x itratn physical columns
    double *d_dkg = nullptr, *d_wq = nullptr;     // dkgnrm and metric weights (nplwcex)
    double* d_cmk = nullptr;
    std::vector<int> col;         // logical history column -> physical column
    std::vector<double> aij;      // itratn x itratn, column-major
    std::vector<double> cmk;
    int nreset = 0;
  } scf;
  int lt_gemm_id = -1;                    // timer the GEMM launches report to
  void* blaslt = nullptr;                 // hipBLASLt state (cross-check only)
  void* mg = nullptr;                     // own-GEMM state (workspace, tuned grids, FP32 copies)
  void* cg_graphs = nullptr;              // captured batched-CG iterations (cg.hip)
  void* kp = nullptr;                     // Bloch bands at several k-points (kpoint.hip)
  bool cg_graph_on = true;                // replay them (qx_set_cg_graph)
  // bumped whenever a device buffer a captured graph may read or write is
  // (re)allocated or its contents' shape changes (KB table, GEMM workspace,
  // FP32 copies, potential copies): captured graphs of older generations are
  // dropped and captured again
  unsigned long buffers_gen = 0;
};
inline void buffers_changed(Ctx* c) { ++c->buffers_gen; }

// record the begin (end) event of a timed launch of timer id, if enabled; a
// begin with cont = true opens a further segment of the same launch
void lt_mark(Ctx* c, int id, bool end, bool cont = false);

int pencil_build(Ctx* c, const int* nga, const int* ngb, const int* ngc, const double* recnrm);

// band-apply: out = (T + V_loc) psi for bands [b0, b0+nb) of slot src into
// slot dst (kinetic + local); KB handled separately by the GEMM path.
int local_apply(Ctx* c, const double* psi, double* out, int nb, int prec, int accumulate);
int density(Ctx* c, const double* psi, int nb, const double* occ_dev, double scale, int prec);
bool fast_ok(const Ctx* c);
int fast_permute_v(Ctx* c);
int local_apply_fast(Ctx* c, const double* psi, double* out, int nb, int prec, int accumulate);
int density_fast(Ctx* c, const double* psi, int nb, const double* w, int prec);
int kb_apply(Ctx* c, const double* psi, double* out, int nb, int accumulate_into_out,
             double* slmir_out, int prec);
int lt_dgemm(Ctx* c, bool ta, int M, int N, int K, double alpha, const double* A, int lda,
             const double* B, int ldb, double beta, double* C, int ldc);
int lt_sgemm_mixed(Ctx* c, bool ta, int M, int N, int K, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc);
void lt_destroy(Ctx* c);
// which kernels run the sizable contractions: GEMM_OWN (our MFMA kernels,
// split bf16x3 for the mixed form), GEMM_LIB (hipBLASLt, a cross-check of
// ours) or GEMM_OWN_F32 (ours, plain FP32 MFMA for the mixed form).  Set by
// qx_set_gemm_path; the environment's QX_GEMM_PATH (0, 1, 2) is the default.
enum { GEMM_OWN = 0, GEMM_LIB = 1, GEMM_OWN_F32 = 2 };
int gemm_path();
void set_gemm_path(int p);
bool lt_enabled();
int lt_sgemm_f32(Ctx* c, bool ta, int M, int N, int K, const float* A, int lda, const float* B,
                 int ldb, float* C, int ldc);
// our MFMA contractions (mfma_gemm.hip); -3 = shape not taken (odd leading
// dimension), the caller falls back to the small-tile kernels
void mg_destroy(Ctx* c);
// schedules the first-use timing of a product shape may choose from:
// 0 stream-K and K-sliced, 1 stream-K only, 2 K-sliced only
void mg_set_schedule(int mode);
int mg_schedule();
void cg_graphs_destroy(Ctx* c);
// k-points (kpoint.hip)
void kp_destroy(Ctx* c);
int kp_set(Ctx* c, int nk, const double* kfrac, const double* bvec, double ecut, int nb,
           int* npw_out);
int kp_sphere(Ctx* c, int ik, int* hkl, double* kin);
int kp_sphere_host(const double* bvec, const double* kf, double ecut, int n1, int n2, int n3,
                   int cap, int* hkl, double* kin);
int kp_slot_put(Ctx* c, int slot, int ik, const double* psi, int nb);
int kp_slot_get(Ctx* c, int slot, int ik, double* psi, int nb);
int kp_set_kb(Ctx* c, int ik, int npj, const double* beta, const double* D);
int kp_hpsi(Ctx* c, int src, int dst, int nb, int per_band);
int kp_solve(Ctx* c, int src, int dst, int max_iter, double tol, double* eig, int* iters);
int cg_graph_entries(Ctx* c);
int mg_dgemm_tn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc, int corr);
int mg_dgemm_nn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc);
template <typename T, typename TC>
int mg_bf3_tn(Ctx* c, int M, int N, int K, double alpha, const T* A, int lda, const T* B, int ldb,
              double beta, TC* C, int ldc, int corr);
template <typename T, typename TC>
int mg_bf3_nn(Ctx* c, int M, int N, int K, double alpha, const T* A, int lda, const T* B, int ldb,
              double beta, TC* C, int ldc);
// FP32 pieces of the fused mixed band-apply: the KB table's FP32 copy (made
// once per table) and FP32 work buffers i = 0..3 kept by the context
const float* mg_kb_b32(Ctx* c);
float* mg_f32_scratch(Ctx* c, int i, size_t n);
const unsigned short* mg_kb_split(Ctx* c, int* ld, int* rows_pad, long* lo);
unsigned short* mg_u16_scratch(Ctx* c, int i, size_t n);
int mg_ps3_tn(Ctx* c, int M, int N, int K, const unsigned short* A, int lda, long alo,
              const unsigned short* B, int ldb, long blo, float* C, int ldc);
int mg_ps3_nn(Ctx* c, int M, int N, int K, const unsigned short* A, int lda, long alo,
              const unsigned short* B, int ldb, long blo, float* C, int ldc);
// FP64-output forms (alpha, beta; no packed-sum correction)
int mg_ps3_tn_d(Ctx* c, int M, int N, int K, double alpha, const unsigned short* A, int lda,
                long alo, const unsigned short* B, int ldb, long blo, double beta, double* C,
                int ldc);
int mg_ps3_nn_d(Ctx* c, int M, int N, int K, double alpha, const unsigned short* A, int lda,
                long alo, const unsigned short* B, int ldb, long blo, double beta, double* C,
                int ldc);
int mg_split_rows(Ctx* c, const double* src, int rows, int cols, int lds, unsigned short* dst,
                  int ld, long lo);
void mg_forget_kb(Ctx* c);
// a band slot split into bf16 hi | lo planes (band stride ld, lo = offset of
// the lo plane) for the mixed band step's products
struct BandSplit {
  unsigned short* p = nullptr;
  int ld = 0;
  long lo = 0;
};
bool band_split_wanted(Ctx* c, int nb, int prec);
int band_split_buf(Ctx* c, int which, BandSplit* s);
int band_split(Ctx* c, int which, const double* src, int nb, BandSplit* s);
int band_tn(Ctx* c, int M, int N, double alpha, const BandSplit& A, const double* a,
            const BandSplit& B, const double* b, double* C, int ldc);
int band_nn(Ctx* c, int nb, double alpha, const BandSplit& A, const double* Z, int ldz,
            double beta, double* C);
int mg_tuned_shapes(Ctx* c);
bool use_big(int M, int N);
// KB term of the mixed band-apply from the FP32 copy psi32 of psi: kb32 =
// B diag(D) (B^T W psi) in FP32 (nplwex x nb)
int kb_mixed_f32(Ctx* c, const double* psi, const float* psi32, const unsigned short* psis,
                 int ldk, long slo, int nb, float* kb32);
// mixed band-apply with the KB contractions between the pencil FFT passes:
// the FP32 copy of psi comes out of the z inverse pass and the KB term goes
// into the z forward pass.  Returns -3 when the shape is not covered.
int hpsi_mixed_fused(Ctx* c, const double* psi, double* out, int nb);
// per-MD-step KB inputs on the device (nlsetup.hip)
int nl_set_tables(Ctx* c, int ntype, int lmmax, int lnref, int mxqlm, const int* lmx,
                  const double* clalp, const double* tbqlm, const double* tbqlma,
                  const double* dlqlm, const int* itola, const int* itolm, const int* itolmf,
                  const int* litoll, int nqlmmx, int lplh, const double* qlmr,
                  const double* qlmi);
int nl_stscfn(Ctx* c, const double* ratm, int natom, const double* hcell, const int* nd1v,
              const int* nd1vks, int lclust, int kmax1, int kmax2, int kmax3, int nplwcs,
              int nplw5, const int* nga, const int* ngb, const int* ngc, double* ycos_out,
              double* ysin_out);
int nl_setnlc_g(Ctx* c, int nplw, int ntype, const int* nhk1, const int* nhk2, int natom,
                const int* lkbppi, const int* lking, const double* ycos, const double* ysin,
                int nplwcs, double rvol, double* vclalp_out);
void nl_destroy(Ctx* c);
void scf_destroy(Ctx* c);
// SCF glue (scfglue.hip)
int scf_set_grid(Ctx* c, int nplw5, const int* ijkgd, const double* recnrmex, const double* vext,
                 const double* rhocore, double rdelv);
int scf_mixer_alloc(Ctx* c, int nplwcex, int imxchg, int itratn, double cmetric, double wkerker,
                    double g2min, double g2max, const double* dkgnrm, double rvol);
int scf_pwpchg(Ctx* c, const double* psi, int nb, const double* occ, double rvol, int prec);
int scf_outchg(Ctx* c);
int scf_mix(Ctx* c, int lreset, double aslh, double bslh, int lmixhold, int* info);
int scf_potential(Ctx* c, double* ene);
int scf_io(Ctx* c, int what, double* host, bool put);
int dropin_xclda(const double* rho, const double* rhocore, int n, double rdelv, double* vexc,
                 double* eeexc, double* e4);
int dropin_sphere(int mode, double* grid, double* rhgr, int nplw5, const double* recnrmex,
                  const int* mfd2ft, int ntotfd, const int* nd, const int* ijkgd);
int dropin_mxchgg(Ctx* c, int* lreset, double* rhgr, double* rinr, int nplw5, double aslh,
                  double bslh, double* glocal, const int* mfd2ft, int ntotfd, const int* ijkgd,
                  int lmixhold);
// KB Hellmann-Feynman forces (forces.hip): fnlc(3, natom) += kbnlpp_frc2
int kb_forces(Ctx* c, const double* psi, int nb, const double* occ_host, const double* gx,
              const double* gy, const double* gz, double* fnlc_host);
// local-pseudopotential forces (flocal.hip)
int lopp_set(Ctx* c, int ntype, int nplw5, const double* svco, const double* gx,
             const double* gy, const double* gz);
void lopp_destroy(Ctx* c);
int schmidt_dev(Ctx* c, double* psi, int nb, int* ok_host, int prec, bool split_ready = false);
// wave-function extrapolation between MD steps (aspc.hip)
void aspc_destroy(Ctx* c);
int aspc_init(Ctx* c, int nprvmx, int ihest);
int aspc_roteig(Ctx* c, const double* cur, int nb, int kpwvx);
int aspc_step(Ctx* c, double* psi, int nb, int nstep, int nstep_ini, int* iaspc, int* lchol);
// Ewald reciprocal-space forces (ewald.hip)
void ewald_destroy(Ctx* c);
int ewald_set(Ctx* c, int numk2, const int* nga, const int* ngb, const int* ngc, const double* gx,
              const double* gy, const double* gz, const double* recnrm, const double* recnrmex,
              double gamma, double volume);
int ewald_frcrec_host(Ctx* c, double* fclm, int ntype, const int* nhk1, const int* nhk2,
                      const double* zv, int nel, double* ecolmb, double* atom_ecolmb,
                      const int* iatoit, const int* iatmpt, int natom, int nion, const int* kcsma,
                      const double* const* dcq, const double* const* dsq);
int ewald_frcrec_dev(Ctx* c, double* fclm, int ntype, const int* nhk1, const int* nhk2,
                     const double* zv, int nel, double* ecolmb, double* atom_ecolmb,
                     const int* iatoit, const int* iatmpt, int natom, int nion);
int flocal_dev(Ctx* c, int ntype, const int* nhk1, const int* nhk2, const int* iatmpt,
               const int* llking, int natom, const double* d_rhgr, int nplw5, double* floc);
// vlocal_g: thrhgr (lopp.d_thr) and the transform input lopp.d_zin (thrhgr
// with the G = 0 grid value) on the density sphere, from the device tables
int vlocal_sphere(Ctx* c, int ntype, const int* nhk1, const int* nhk2, const int* iatmpt,
                  const int* llking, int natom, double rvol, int nplw5);
// one-band KB apply from the factors of a device-built table (kb1.hip)
void kb1_destroy(Ctx* c);
int kb1_build(Ctx* c, const int* type, int natom, const int* lkbppi, const int* lking,
              const double* dyc, const double* dys, int ldy);
size_t kb1_scratch(const Ctx* c);
int kb1_project(Ctx* c, const double* v, double* pn);
// vext (scf) = (accumulate ? vext : 0) + the inverse transform of zin
int vext_from_sphere(Ctx* c, const double* d_zin, int accumulate);
int flocal_host(Ctx* c, double* floc, int ntype, const int* nhk1, const int* nhk2,
                const int* iatmpt, int natom, const int* llking, int nplw5, const int* nga,
                const int* ngb, const int* ngc, const double* rhgr, const int* kmax,
                const double* const* dcq, const double* const* dsq);
int coop_launch(Ctx* c, const void* fn, int want, size_t lds, void** args, bool exact);
int dense_chol_inv(Ctx* c, double* s, int n, double* x, int* ok, int* flag);
int gemm_tn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
            int ldb, double beta, double* C, int ldc, int corr);
int gemm_nn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
            int ldb, double beta, double* C, int ldc);
// the same contractions with a precision: PREC_FP32 -> FP32 MFMA (operands
// rounded to FP32, FP32 accumulation per K chunk), otherwise FP64
int gemm_tn_p(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc, int corr, int prec);
int gemm_nn_p(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc, int prec);

}  // namespace qx
