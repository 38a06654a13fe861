// api.hip -- the C-ABI of qxmd_amd (declared in include/qxmd_amd.h).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/qxmd_amd.h"
#include "context.h"

namespace qx {
int strvec(const double* rba, double ecutdens, double ecut, int cap, int* counts, int* nga,
           int* ngb, int* ngc, double* recnrm, double* gx, double* gy, double* gz);
void get_mesh(const double* hcell, double ecutdens, int* nd);
void ijkgd_map(int nplw5, const int* nga, const int* ngb, const int* ngc, const int* nd, int* ijkgd);
int clmrcp(const double* B, double dkmax2, int cap, int* nga, int* ngb, int* ngc, double* gx,
           double* gy, double* gz, double* recnrm, int* kcsma);
int fullgrid_rfft3(int key, double* fx, double* fy, int n1, int n2, int n3);
int fullgrid_fft3(int key, double* fx, double* fy, int n1, int n2, int n3);
int fullgrid_rfft3_dev(int key, double* dx, double* dy, int n1, int n2, int n3, hipStream_t st);
int noalloc_limits(long* max_points, int* max_axis);
int dense_chol_inv(Ctx* c, double* s, int n, double* x, int* ok, int* flag);
int dense_eigh(Ctx* c, double* a, int n, double* z, double* w, double* work, int* info);
int dense_symmetrize(Ctx* c, const double* C, int n, double* H, double* eig);
int pdot_batched(Ctx* c, const double* x, const double* y, int nb, double* out_dev);
int bkl_band_by_band(Ctx* c, double* psi, double* hpsi, int nb, const double* ekib, double* eig,
                     int noband, int itermx, int iteremx, double tolera, double threinn,
                     double wdinner, double roughfzan, double roughfeig, double* bzan, int* nriter,
                     int prec);
int bkl_batched(Ctx* c, double* psi, double* hpsi, int nb, const double* ekib_host,
                double* eig_host, int noband, int itermx, int iteremx, double tolera,
                double roughfzan, double roughfeig, double* bzan, int* nriter, int prec);

static thread_local std::string g_err;
static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// ---- internal <-> reference order ------------------------------------------
__global__ void k_to_internal(const double* __restrict__ ref, double* __restrict__ in,
                              const int* __restrict__ perm, int np, int nb) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)np * nb) return;
  int b = (int)(i / np), p = (int)(i % np);
  size_t n = 2 * (size_t)np;
  int g = perm[p];
  in[b * n + 2 * p] = ref[b * n + 2 * (size_t)g];
  in[b * n + 2 * p + 1] = ref[b * n + 2 * (size_t)g + 1];
}
__global__ void k_to_reference(const double* __restrict__ in, double* __restrict__ ref,
                               const int* __restrict__ perm, int np, int nb) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)np * nb) return;
  int b = (int)(i / np), p = (int)(i % np);
  size_t n = 2 * (size_t)np;
  int g = perm[p];
  ref[b * n + 2 * (size_t)g] = in[b * n + 2 * p];
  ref[b * n + 2 * (size_t)g + 1] = in[b * n + 2 * p + 1];
}

static unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Argument check shared by the slot entry points, run before anything touches
// the device: every slot index names one of the kSlots band arrays, 1 <= nb <=
// nbmax, the slots a routine reads and writes (a, b) are distinct, `staged`
// refuses slot 5 (the routine stages its result there) and `cg_work` refuses
// slots 2-4 (the CG's g, h and H h).
constexpr int kSlotStaged = 1, kSlotCgWork = 2, kSlotPair01 = 4;
static int check_slots(const Ctx* c, const char* who, int nb, int a, int b, int rules) {
  auto bad_slot = [&](int s) {
    if (s < 0 || s >= Ctx::kSlots) return true;
    if ((rules & kSlotStaged) && s == 5) return true;
    if ((rules & kSlotCgWork) && s >= 2 && s <= 4) return true;
    if ((rules & kSlotPair01) && s > 1) return true;
    return false;
  };
  if (bad_slot(a) || (b != -1 && bad_slot(b)))
    return fail(-1, "%s: slot out of range or reserved (slots 0..%d%s%s%s)", who, Ctx::kSlots - 1,
                (rules & kSlotPair01) ? "; only 0 and 1 here" : "",
                (rules & kSlotCgWork) ? "; 2-4 are the CG work space" : "",
                (rules & kSlotStaged) ? "; 5 is the staging slot" : "");
  if (b != -1 && a == b) return fail(-1, "%s: the two slots must differ", who);
  if (nb < 1 || nb > c->nbmax) return fail(-1, "%s: need 1 <= nb <= nbmax (%d)", who, c->nbmax);
  return 0;
}

// acc += (T+V) psi - T psi  (the local part alone, hcnonloc)
__global__ void k_add_minus_kin(double* acc, const double* tv, const double* psi, const double* kin,
                                int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) acc[i] += tv[i] - kin[i >> 1] * psi[i];
}

// stage a host reference-order array (nb bands) into dst (internal order) via slot 5
static int put_internal(Ctx* c, double* dst, const double* host, int nb) {
  const size_t bytes = sizeof(double) * (size_t)c->nplwex * nb;
  double* stage = c->d_slot[5];
  QX_CHECK(hipMemcpyAsync(stage, host, bytes, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_to_internal, dim3(cdiv((long)(c->nplw + 1) * nb, 256)), dim3(256), 0,
                     c->stream, stage, dst, c->d_perm, c->nplw + 1, nb);
  QX_CHECK(hipGetLastError());
  return 0;
}
static int get_reference(Ctx* c, const double* src, double* host, int nb) {
  const size_t bytes = sizeof(double) * (size_t)c->nplwex * nb;
  double* stage = c->d_slot[5];
  hipLaunchKernelGGL(k_to_reference, dim3(cdiv((long)(c->nplw + 1) * nb, 256)), dim3(256), 0,
                     c->stream, src, stage, c->d_perm, c->nplw + 1, nb);
  QX_CHECK(hipMemcpyAsync(host, stage, bytes, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int hpsi_batched(Ctx* c, const double* x, double* hx, int nb, int prec) {
  if (prec == PREC_FP32) {
    const int rf = hpsi_mixed_fused(c, x, hx, nb);
    if (rf != -3) return rf;  // -3: shape not covered by the fused order
  }
  int rc = local_apply(c, x, hx, nb, prec, 0);
  if (rc) return rc;
  return kb_apply(c, x, hx, nb, 1, nullptr, prec);
}

// Gram-Schmidt by Cholesky of the packed overlap (schmidt2).  split_ready:
// the mixed form's split planes of psi (band split 2) are current.
int schmidt_dev(Ctx* c, double* psi, int nb, int* ok_host, int prec, bool split_ready) {
  const int n = c->nplwex;
  double* S = c->d_small;                         // nb x nb overlap, then L
  double* Linv = c->d_small + (size_t)(nb + 1) * nb;  // row-major L^-1
  int* ok = reinterpret_cast<int*>(c->d_scal + 62 * (size_t)c->nbmax);
  int* flag = ok + 2;
  // S[i][j] = <psi_j|psi_i> (C[n=i][m=j]); symmetric, so its storage is also
  // the column-major S(i,j).
  // the overlap product in the given precision (split bf16x3 in the mixed
  // form); the Cholesky and the inverse stay FP64
  const bool split = band_split_wanted(c, nb, prec);
  BandSplit ps;
  int rc = 0;
  if (split) {
    rc = split_ready ? band_split_buf(c, 2, &ps) : band_split(c, 2, psi, nb, &ps);
    if (!rc) rc = band_tn(c, nb, nb, 2.0, ps, psi, ps, psi, S, nb);
  } else {
    rc = gemm_tn_p(c, nb, nb, n, 2.0, psi, n, psi, n, 0.0, S, nb, 1, prec);
  }
  if (rc) return rc;
  if ((rc = dense_chol_inv(c, S, nb, Linv, ok, flag))) return rc;
  int okh[2] = {0, 0};
  QX_CHECK(hipMemcpyAsync(okh, ok, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (okh[1]) return -20;  // a grid barrier of the Cholesky gave up
  *ok_host = okh[0];
  if (!okh[0]) return 0;
  // psi'_ib = sum_jb Linv(ib,jb) psi_jb: gemm_nn's B[n*ldb + k] = Linv(n,k)
  if (split) return band_nn(c, nb, 1.0, ps, Linv, nb, 0.0, psi);  // reads the split planes
  double* out = c->d_slot[5];
  if ((rc = gemm_nn_p(c, n, nb, nb, 1.0, psi, n, Linv, nb, 0.0, out, n, prec))) return rc;
  QX_CHECK(hipMemcpyAsync(psi, out, sizeof(double) * (size_t)n * nb, hipMemcpyDeviceToDevice,
                          c->stream));
  return 0;
}

// subspace rotation: H_sub, eigenvectors U, psi <- psi U, hpsi <- hpsi U
// (prec: the three large products; the eigensolver stays FP64)
static int rotate_dev(Ctx* c, double* psi, double* hpsi, int nb, double* eig_host,
                      int prec = PREC_FP64) {
  const int n = c->nplwex;
  double* C = c->d_small;
  double* H = C + (size_t)nb * nb;
  double* Z = H + (size_t)nb * nb;
  double* work = Z + (size_t)nb * nb;
  double* w = c->d_scal + 40 * (size_t)c->nbmax;
  int* info = reinterpret_cast<int*>(c->d_scal + 63 * (size_t)c->nbmax);
  // the mixed form's products on split band slots (split 2: psi, 3: hpsi)
  const bool split = band_split_wanted(c, nb, prec);
  BandSplit ps, hs;
  int rc = 0;
  if (split) {
    rc = band_split(c, 2, psi, nb, &ps);
    if (!rc) rc = band_split(c, 3, hpsi, nb, &hs);
    if (!rc) rc = band_tn(c, nb, nb, 2.0, ps, psi, hs, hpsi, C, nb);
  } else {
    rc = gemm_tn_p(c, nb, nb, n, 2.0, psi, n, hpsi, n, 0.0, C, nb, 1, prec);
  }
  if (rc) return rc;
  if ((rc = dense_symmetrize(c, C, nb, H, nullptr))) return rc;
  if ((rc = dense_eigh(c, H, nb, Z, w, work, info))) return rc;
  // psi' = psi U: B(n=b, k=j) = U(j,b) = Z[j + nb*b] -> ldb = nb with B[n*ldb+k]
  if (split) {  // the products read the split planes: written in place
    if ((rc = band_nn(c, nb, 1.0, ps, Z, nb, 0.0, psi))) return rc;
    if ((rc = band_nn(c, nb, 1.0, hs, Z, nb, 0.0, hpsi))) return rc;
  } else {
    double* out = c->d_slot[5];
    if ((rc = gemm_nn_p(c, n, nb, nb, 1.0, psi, n, Z, nb, 0.0, out, n, prec))) return rc;
    QX_CHECK(hipMemcpyAsync(psi, out, sizeof(double) * (size_t)n * nb, hipMemcpyDeviceToDevice,
                            c->stream));
    if ((rc = gemm_nn_p(c, n, nb, nb, 1.0, hpsi, n, Z, nb, 0.0, out, n, prec))) return rc;
    QX_CHECK(hipMemcpyAsync(hpsi, out, sizeof(double) * (size_t)n * nb, hipMemcpyDeviceToDevice,
                            c->stream));
  }
  if (eig_host) {
    QX_CHECK(hipMemcpyAsync(eig_host, w, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    QX_CHECK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

// ekib: 2 <psi|T|psi> per band (ekinetic2, Sources/engrad.f90:117-122)
__global__ void k_kin_apply(const double* psi, double* out, const double* kin, int n, int nb) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * nb) return;
  int k = (int)(i % n);
  out[i] = kin[k >> 1] * psi[i];
}

}  // namespace qx

using namespace qx;

struct qx_ctx {
  Ctx c;
};

extern "C" {

const char* qx_last_error(void) { return g_err.c_str(); }
int qx_version(void) { return 1; }

int qx_get_mesh(const double* hcell, double ecutdens, int* nd) {
  get_mesh(hcell, ecutdens, nd);
  return 0;
}

int qx_strvec(const double* rba, double ecutdens, double ecut, int cap, int* counts, int* nga,
              int* ngb, int* ngc, double* recnrm, double* gx, double* gy, double* gz) {
  return strvec(rba, ecutdens, ecut, cap, counts, nga, ngb, ngc, recnrm, gx, gy, gz);
}

int qx_ijkgd(int nplw5, const int* nga, const int* ngb, const int* ngc, const int* nd, int* ijkgd) {
  ijkgd_map(nplw5, nga, ngb, ngc, nd, ijkgd);
  return 0;
}

int qx_clmrcp(const double* b, double dkmax2, int cap, int* nga, int* ngb, int* ngc, double* gx,
              double* gy, double* gz, double* recnrm, int* kcsma) {
  if (!b || !(dkmax2 > 0.0) || cap < 0) return fail(-1, "qx_clmrcp: bad arguments");
  return clmrcp(b, dkmax2, cap, nga, ngb, ngc, gx, gy, gz, recnrm, kcsma);
}

int qx_create(qx_ctx** out, int n1, int n2, int n3, int nplw, const int* nga, const int* ngb,
              const int* ngc, const double* recnrm, int nbmax) {
  if (!out || n1 <= 0 || n2 <= 0 || n3 <= 0 || nplw < 0 || nbmax < 1)
    return fail(-1, "qx_create: bad arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(-9, "qx_create: no HIP device");
  qx_ctx* h = new qx_ctx();
  Ctx* c = &h->c;
  QX_CHECK(hipGetDevice(&c->dev));
  c->n1 = n1; c->n2 = n2; c->n3 = n3;
  c->N = (long)n1 * n2 * n3;
  c->nplw = nplw;
  c->nplwex = 2 * (nplw + 1);
  c->nbmax = std::max(2, nbmax + (nbmax & 1));
  c->plx = make_plan(n1); c->ply = make_plan(n2); c->plz = make_plan(n3);
  if (c->plx.npass < 0 || c->ply.npass < 0 || c->plz.npass < 0 || n1 > 1024 || n2 > 1024 ||
      n3 > 1024) {
    delete h;
    return fail(-2, "qx_create: FFT sizes must be 2^a 3^b 5^c 7^d <= 1024");
  }
  QX_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  int rc = pencil_build(c, nga, ngb, ngc, recnrm);
  if (rc) {
    delete h;
    return fail(rc, "qx_create: pencil layout failed (sphere wraps the grid?)");
  }
  const size_t vbytes = sizeof(double) * (size_t)c->nplwex * c->nbmax;
  for (int s = 0; s < Ctx::kSlots; ++s) {
    QX_CHECK(hipMalloc(&c->d_slot[s], vbytes));
    QX_CHECK(hipMemset(c->d_slot[s], 0, vbytes));
  }
  // pencil intermediates: a batch of band pairs, FP64 complex
  const size_t per_bp = 16 * (size_t)n3 * c->pc.ncol;
  const int nbp = c->nbmax / 2;
  // pencil intermediates per buffer: 4 GiB holds every band pair of the
  // 512-atom cell (1280 bands) in one batch, so the fused mixed band-apply
  // (one batch) covers it; two buffers of that size are 3% of the HBM
  size_t cap_bytes = (size_t)4096 << 20;
  if (const char* e = std::getenv("QX_PENCIL_MB")) cap_bytes = (size_t)std::max(1, std::atoi(e)) << 20;
  c->batch_bp = std::max(1, std::min(nbp, (int)std::max<size_t>(1, cap_bytes / per_bp)));
  c->t_bytes = per_bp * c->batch_bp;
  QX_CHECK(hipMalloc(&c->d_t1, c->t_bytes));
  QX_CHECK(hipMalloc(&c->d_t2, c->t_bytes));
  QX_CHECK(hipMalloc(&c->d_V, sizeof(double) * c->N));
  QX_CHECK(hipMalloc(&c->d_Vf, sizeof(float) * c->N));
  QX_CHECK(hipMemset(c->d_V, 0, sizeof(double) * c->N));
  QX_CHECK(hipMemset(c->d_Vf, 0, sizeof(float) * c->N));
  const size_t nsmall = 4 * (size_t)(c->nbmax + 1) * c->nbmax + 16 * (size_t)c->nbmax;
  QX_CHECK(hipMalloc(&c->d_small, sizeof(double) * nsmall));
  // 64 scalar arrays, then partial sums of up to 3 inner products per band and 4096-chunk
  const size_t nsc = 64 * (size_t)c->nbmax + 3 * (size_t)c->nbmax * (c->nplwex / 4096 + 2);
  QX_CHECK(hipMalloc(&c->d_scal, sizeof(double) * nsc));
  QX_CHECK(hipMemset(c->d_scal, 0, sizeof(double) * nsc));
  c->rho_groups = std::max(1, std::min(nbp, 32));
  QX_CHECK(hipMalloc(&c->d_rho_part, sizeof(double) * (size_t)c->rho_groups * c->N));
  QX_CHECK(hipMalloc(&c->d_rho, sizeof(double) * c->N));
  QX_CHECK(hipMalloc(&c->d_sync, sizeof(int) * 16));
  QX_CHECK(hipMemset(c->d_sync, 0, sizeof(int) * 16));
  c->dwork_len = 8 * (size_t)c->nbmax * c->nbmax + 16 * (size_t)c->nbmax + 64;
  QX_CHECK(hipMalloc(&c->d_dwork, sizeof(double) * c->dwork_len));
  c->bb_nch = c->nplwex / kBbChunk + 1;
  QX_CHECK(hipMalloc(&c->d_bb, sizeof(double) * (32 + 5 * (size_t)(c->bb_nch + 1))));
  QX_CHECK(hipMemset(c->d_bb, 0, sizeof(double) * (32 + 5 * (size_t)(c->bb_nch + 1))));
  QX_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_flag), sizeof(int) * 4, hipHostMallocDefault));
  QX_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_hflag), c->h_flag, 0));
  for (auto& e : c->ev) QX_CHECK(hipEventCreate(&e));
  for (auto& e : c->pev) QX_CHECK(hipEventCreate(&e));
  *out = h;
  return 0;
}

int qx_destroy(qx_ctx* h) {
  if (!h) return 0;
  Ctx* c = &h->c;
  (void)hipStreamSynchronize(c->stream);
  void* ptrs[] = {c->d_perm, c->d_pt_col, c->d_pt_z, c->d_pair_p0, c->d_col_hi, c->d_col_kmod,
                  c->d_kin, c->d_tw, c->d_V, c->d_Vf, c->d_Vp, c->d_Vpf, c->d_B, c->d_D, c->d_t1, c->d_t2, c->d_P,
                  c->d_small, c->d_rho_part, c->d_rho, c->d_scal, c->d_sync, c->d_dwork, c->d_bb};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_flag) (void)hipHostFree(c->h_flag);
  for (auto* p : c->d_slot)
    if (p) (void)hipFree(p);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->pev)
    if (e) (void)hipEventDestroy(e);
  for (auto& t : c->lt)
    for (auto& e : t.ev) (void)hipEventDestroy(e);
  lt_destroy(c);
  cg_graphs_destroy(c);
  kp_destroy(c);
  mg_destroy(c);
  nl_destroy(c);
  scf_destroy(c);
  lopp_destroy(c);
  kb1_destroy(c);
  ewald_destroy(c);
  aspc_destroy(c);
  (void)hipStreamDestroy(c->stream);
  delete h;
  return 0;
}

int qx_shutdown(void) {
  dropin_release();
  return 0;
}

int qx_sync(qx_ctx* h) {
  QX_CHECK(hipStreamSynchronize(h->c.stream));
  return 0;
}

int qx_info(qx_ctx* h, int* out) {
  out[0] = h->c.pc.ncol;
  out[1] = h->c.pc.npair;
  out[2] = h->c.pc.nH;
  out[3] = h->c.npj;
  out[4] = h->c.nplwex;
  return 0;
}

__global__ void k_d2f(const double* a, float* b, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

int qx_set_glocal(qx_ctx* h, const double* glocal) {
  Ctx* c = &h->c;
  QX_CHECK(hipMemcpyAsync(c->d_V, glocal, sizeof(double) * c->N, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_d2f, dim3(cdiv(c->N, 256)), dim3(256), 0, c->stream, c->d_V, c->d_Vf, c->N);
  int rc = fast_permute_v(c);
  if (rc) return rc;
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int qx_set_kb(qx_ctx* h, const double* eqlmr, int lmmax, int natom, const int* lmx_atom,
              const double* vclalp_atom) {
  Ctx* c = &h->c;
  std::vector<int> pa, pl;
  for (int ia = 0; ia < natom; ++ia)
    for (int l = 0; l < std::min(lmx_atom[ia], lmmax); ++l) {
      pa.push_back(ia);
      pl.push_back(l);
    }
  const int npj = (int)pa.size();
  const size_t n = (size_t)c->nplwex;
  mg_forget_kb(c);
  buffers_changed(c);
  kb1_destroy(c);  // a host table has no factors
  if (c->d_B) { (void)hipFree(c->d_B); c->d_B = nullptr; }
  if (c->d_D) { (void)hipFree(c->d_D); c->d_D = nullptr; }
  if (c->d_P) { (void)hipFree(c->d_P); c->d_P = nullptr; }
  c->npj = npj;
  c->lmmax = lmmax;
  c->natom = natom;
  c->proj_atom = pa;
  c->proj_l = pl;
  if (npj == 0) return 0;
  // projectors in reference order, then permuted to internal order on device
  std::vector<double> B(n * npj), D(npj);
  for (int p = 0; p < npj; ++p) {
    std::memcpy(&B[n * p], eqlmr + n * ((size_t)pl[p] + (size_t)lmmax * pa[p]), sizeof(double) * n);
    D[p] = vclalp_atom[pl[p] + lmmax * pa[p]];
  }
  QX_CHECK(hipMalloc(&c->d_B, sizeof(double) * n * npj));
  QX_CHECK(hipMalloc(&c->d_D, sizeof(double) * npj));
  QX_CHECK(hipMalloc(&c->d_P, sizeof(double) * (size_t)npj * c->nbmax));
  double* stage = nullptr;
  QX_CHECK(hipMalloc(&stage, sizeof(double) * n * npj));
  QX_CHECK(hipMemcpy(stage, B.data(), sizeof(double) * n * npj, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_internal, dim3(cdiv((long)(c->nplw + 1) * npj, 256)), dim3(256), 0,
                     c->stream, stage, c->d_B, c->d_perm, c->nplw + 1, npj);
  QX_CHECK(hipMemcpyAsync(c->d_D, D.data(), sizeof(double) * npj, hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  QX_CHECK(hipFree(stage));
  return 0;
}

int qx_set_nlpp_tables(qx_ctx* h, int ntype, int lmmax, int lnref, int mxqlm, const int* lmx,
                       const double* clalp, const double* tbqlm, const double* tbqlma,
                       const double* dlqlm, const int* itola, const int* itolm, const int* itolmf,
                       const int* litoll, int nqlmmx, int lplh, const double* qlmr,
                       const double* qlmi) {
  const int rc = nl_set_tables(&h->c, ntype, lmmax, lnref, mxqlm, lmx, clalp, tbqlm, tbqlma,
                               dlqlm, itola, itolm, itolmf, litoll, nqlmmx, lplh, qlmr, qlmi);
  if (rc == -2) return fail(rc, "qx_set_nlpp_tables: the radial table ends below the sphere's |G|");
  if (rc) return fail(rc, "qx_set_nlpp_tables: bad table shapes or indices");
  return 0;
}

int qx_stscfn_pw(qx_ctx* h, const double* ratm, int natom, const double* hcell, const int* nd1v,
                 const int* nd1vks, int lclust, int kmax1, int kmax2, int kmax3, int nplwcs,
                 int nplw5, const int* nga, const int* ngb, const int* ngc, double* ycos,
                 double* ysin) {
  const int rc = nl_stscfn(&h->c, ratm, natom, hcell, nd1v, nd1vks, lclust, kmax1, kmax2, kmax3,
                           nplwcs, nplw5, nga, ngb, ngc, ycos, ysin);
  if (rc == -2) return fail(rc, "qx_stscfn_pw: a G index exceeds kmax");
  if (rc) return fail(rc, "qx_stscfn_pw: bad arguments");
  return 0;
}

int qx_setnlc_g(qx_ctx* h, int nplw, int ntype, const int* nhk1, const int* nhk2, int natom,
                const int* lkbppi, const int* lking, const double* ycos, const double* ysin,
                int nplwcs, double rvol, double* vclalp) {
  const int rc = nl_setnlc_g(&h->c, nplw, ntype, nhk1, nhk2, natom, lkbppi, lking, ycos, ysin,
                             nplwcs, rvol, vclalp);
  if (rc == -4) return fail(rc, "qx_setnlc_g: nonlocal tables not registered for ntype types");
  if (rc == -5) return fail(rc, "qx_setnlc_g: no device structure factors of this shape");
  if (rc) return fail(rc, "qx_setnlc_g: bad arguments");
  return 0;
}

// the KB table in the reference layout: eqlmr(nplwex, lmmax, natom) (zero for
// projectors the context does not hold) and vclalp per atom (lmmax, natom)
int qx_get_kb(qx_ctx* h, double* eqlmr, double* vclalp_atom) {
  Ctx* c = &h->c;
  const size_t n = (size_t)c->nplwex;
  std::memset(eqlmr, 0, sizeof(double) * n * c->lmmax * c->natom);
  if (vclalp_atom) std::memset(vclalp_atom, 0, sizeof(double) * c->lmmax * c->natom);
  if (c->npj == 0) return 0;
  double* stage = nullptr;
  QX_CHECK(hipMalloc(&stage, sizeof(double) * n * c->npj));
  hipLaunchKernelGGL(k_to_reference, dim3(cdiv((long)(c->nplw + 1) * c->npj, 256)), dim3(256), 0,
                     c->stream, c->d_B, stage, c->d_perm, c->nplw + 1, c->npj);
  std::vector<double> B(n * c->npj), D(c->npj);
  QX_CHECK(hipMemcpyAsync(B.data(), stage, sizeof(double) * B.size(), hipMemcpyDeviceToHost,
                          c->stream));
  QX_CHECK(hipMemcpyAsync(D.data(), c->d_D, sizeof(double) * D.size(), hipMemcpyDeviceToHost,
                          c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  QX_CHECK(hipFree(stage));
  for (int p = 0; p < c->npj; ++p) {
    const size_t slot = (size_t)c->proj_l[p] + (size_t)c->lmmax * c->proj_atom[p];
    std::memcpy(eqlmr + n * slot, &B[n * p], sizeof(double) * n);
    if (vclalp_atom) vclalp_atom[slot] = D[p];
  }
  return 0;
}

int qx_slot_put(qx_ctx* h, int slot, const double* gdcr, int nb) {
  if (!h) return fail(-1, "qx_slot_put: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_slot_put", nb, slot, -1, kSlotPair01)) return e;
  int rc = put_internal(c, c->d_slot[slot], gdcr, nb);
  if (rc) return rc;
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int qx_slot_get(qx_ctx* h, int slot, double* gdcr, int nb) {
  if (!h) return fail(-1, "qx_slot_get: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_slot_get", nb, slot, -1, kSlotStaged)) return e;
  return get_reference(c, c->d_slot[slot], gdcr, nb);
}

__global__ void k_slot_copy(const double2* __restrict__ in, double2* __restrict__ out, long n2) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (long)gridDim.x * blockDim.x)
    out[i] = in[i];
}

int qx_slot_copy(qx_ctx* h, int src, int dst, int nb) {
  if (!h) return fail(-1, "qx_slot_copy: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_slot_copy", nb, src, dst, 0)) return e;
  // one dispatch of a plain streaming copy (16 bytes per lane): each byte of
  // the bands is read once and written once, which also makes this launch the
  // known-size reference of the HBM byte counters
  const long n2 = (long)c->nplwex * nb / 2;  // nplwex = 2 (nplw + 1) is even
  const int grid = (int)std::min<long>(cdiv(n2, 256), 8192);
  hipLaunchKernelGGL(k_slot_copy, dim3(grid), dim3(256), 0, c->stream,
                     reinterpret_cast<const double2*>(c->d_slot[src]),
                     reinterpret_cast<double2*>(c->d_slot[dst]), n2);
  QX_CHECK(hipGetLastError());
  return 0;
}

int qx_hpsi_slot(qx_ctx* h, int src, int dst, int nb, int prec) {
  if (!h) return fail(-1, "qx_hpsi_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_hpsi_slot", nb, src, dst, 0)) return e;
  return hpsi_batched(c, c->d_slot[src], c->d_slot[dst], nb, prec);
}

int qx_local_slot(qx_ctx* h, int src, int dst, int nb, int prec, int accumulate) {
  if (!h) return fail(-1, "qx_local_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_local_slot", nb, src, dst, 0)) return e;
  return local_apply(c, c->d_slot[src], c->d_slot[dst], nb, prec, accumulate);
}

int qx_kb_slot(qx_ctx* h, int src, int dst, int nb, double* slmir_host) {
  if (!h) return fail(-1, "qx_kb_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_kb_slot", nb, src, dst, 0)) return e;
  int rc = kb_apply(c, c->d_slot[src], c->d_slot[dst], nb, 1, nullptr, PREC_FP64);
  if (rc) return rc;
  (void)slmir_host;
  return 0;
}

int qx_schmidt_slot(qx_ctx* h, int slot, int nb, int* lcholesky) {
  if (!h) return fail(-1, "qx_schmidt_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_schmidt_slot", nb, slot, -1, kSlotStaged)) return e;
  return schmidt_dev(c, c->d_slot[slot], nb, lcholesky, PREC_FP64);
}

int qx_rotate_slot(qx_ctx* h, int psi, int hpsi, int nb, double* eig) {
  if (!h) return fail(-1, "qx_rotate_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_rotate_slot", nb, psi, hpsi, kSlotStaged)) return e;
  return rotate_dev(c, c->d_slot[psi], c->d_slot[hpsi], nb, eig);
}

int qx_density_slot(qx_ctx* h, int slot, int nb, const double* occ, double rvol, double* rho,
                    int prec) {
  if (!h) return fail(-1, "qx_density_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_density_slot", nb, slot, -1, 0)) return e;
  double* w = c->d_scal + 41 * (size_t)c->nbmax;
  std::vector<double> wh(nb + 1, 0.0);
  for (int b = 0; b < nb; ++b) wh[b] = occ[b] / rvol;
  QX_CHECK(hipMemcpyAsync(w, wh.data(), sizeof(double) * (nb + 1), hipMemcpyHostToDevice, c->stream));
  int rc = density(c, c->d_slot[slot], nb, w, 1.0, prec);
  if (rc) return rc;
  if (rho) {
    QX_CHECK(hipMemcpyAsync(rho, c->d_rho, sizeof(double) * c->N, hipMemcpyDeviceToHost, c->stream));
    QX_CHECK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int qx_eigen_bkl_slot(qx_ctx* h, int psi, int hpsi, int nb, const double* ekib, double* eig,
                      int noband, int itermx, int iteremx, double tolera, double threinn,
                      double wdinner, double roughfzan, double roughfeig, double* bzan, int* nriter,
                      int prec) {
  Ctx* c = &h->c;
  if (psi < 0 || psi > 1 || hpsi < 0 || hpsi > 1 || psi == hpsi || nb < 1 || nb > c->nbmax)
    return fail(-1, "qx_eigen_bkl_slot: psi/hpsi must be slots 0 and 1, 1 <= nb <= nbmax");
  int rc = bkl_band_by_band(c, c->d_slot[psi], c->d_slot[hpsi], nb, ekib, eig, noband, itermx,
                            iteremx, tolera, threinn, wdinner, roughfzan, roughfeig, bzan, nriter,
                            prec);
  if (rc) return rc;
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// batched BKL (all bands per iteration)
int qx_eigen_bkl_batched_slot(qx_ctx* h, int psi, int hpsi, int nb, const double* ekib, double* eig,
                              int noband, int itermx, int iteremx, double tolera, double roughfzan,
                              double roughfeig, double* bzan, int* nriter, int prec) {
  Ctx* c = &h->c;
  if (psi < 0 || psi > 1 || hpsi < 0 || hpsi > 1 || psi == hpsi || nb < 1 || nb > c->nbmax)
    return fail(-1, "qx_eigen_bkl_batched_slot: psi/hpsi must be slots 0 and 1, 1 <= nb <= nbmax");
  return bkl_batched(c, c->d_slot[psi], c->d_slot[hpsi], nb, ekib, eig, noband, itermx, iteremx,
                     tolera, roughfzan, roughfeig, bzan, nriter, prec);
}

int qx_ekib_slot(qx_ctx* h, int slot, int nb, double* ekib_host) {
  if (!h) return fail(-1, "qx_ekib_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_ekib_slot", nb, slot, -1, kSlotStaged)) return e;
  const int n = c->nplwex;
  double* t = c->d_slot[5];
  double* d = c->d_scal + 42 * (size_t)c->nbmax;
  hipLaunchKernelGGL(k_kin_apply, dim3(cdiv((long)n * nb, 256)), dim3(256), 0, c->stream,
                     c->d_slot[slot], t, c->d_kin, n, nb);
  pdot_batched(c, t, c->d_slot[slot], nb, d);
  std::vector<double> e(nb);
  QX_CHECK(hipMemcpyAsync(e.data(), d, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < nb; ++b) ekib_host[b] = 2.0 * e[b];
  return 0;
}

int qx_scf_band_step_ex(qx_ctx* h, int psi, int hpsi, int nb, const double* occ, int noband,
                        double rvol, double* eig, double* rho, int* lcholesky, int prec,
                        int batched) {
  Ctx* c = &h->c;
  // slots 2-4 are the CG's work space
  if (psi < 0 || psi > 1 || hpsi < 0 || hpsi > 1 || psi == hpsi || nb < 1 || nb > c->nbmax)
    return fail(-1, "qx_scf_band_step: psi/hpsi must be slots 0 and 1, 1 <= nb <= nbmax");
  double* P = c->d_slot[psi];
  double* HP = c->d_slot[hpsi];
  std::vector<double> ekib(nb), bz(nb);
  std::vector<int> nr(nb);
  // phase timers on the context stream (the role of the ltimecnt timers)
  QX_CHECK(hipEventRecord(c->pev[0], c->stream));
  // energy pass: H psi, ekib
  int rc = hpsi_batched(c, P, HP, nb, prec);
  if (rc) return rc;
  if ((rc = qx_ekib_slot(h, psi, nb, ekib.data()))) return rc;
  QX_CHECK(hipEventRecord(c->pev[1], c->stream));
  // subspace rotation
  // the mixed form runs the overlap, subspace-H and rotation products as
  // split bf16x3 products (the eigensolver and Cholesky stay FP64)
  if ((rc = rotate_dev(c, P, HP, nb, eig, prec))) return rc;
  QX_CHECK(hipEventRecord(c->pev[2], c->stream));
  // CG
  if (batched)
    rc = bkl_batched(c, P, HP, nb, ekib.data(), eig, noband, 4, 4, 1e-10, 0.3, 0.3, bz.data(),
                     nr.data(), prec);
  else
    rc = bkl_band_by_band(c, P, HP, nb, ekib.data(), eig, noband, 4, 4, 1e-10, 0.0, 0.9, 0.3, 0.3,
                          bz.data(), nr.data(), prec);
  if (rc) return rc;
  QX_CHECK(hipEventRecord(c->pev[3], c->stream));
  // the batched CG leaves the split planes of psi current for the overlap
  if ((rc = schmidt_dev(c, P, nb, lcholesky, prec, batched != 0))) return rc;
  QX_CHECK(hipEventRecord(c->pev[4], c->stream));
  if ((rc = qx_density_slot(h, psi, nb, occ, rvol, rho, prec))) return rc;
  QX_CHECK(hipEventRecord(c->pev[5], c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < 5; ++i) QX_CHECK(hipEventElapsedTime(&c->phase_ms[i], c->pev[i], c->pev[i + 1]));
  return 0;
}

}  // extern "C"

namespace qx {
void lt_mark(Ctx* c, int id, bool end, bool cont) {
  if (id < 0 || id >= Ctx::kTimers || !c->lt[id].on) return;
  auto& t = c->lt[id];
  if (!end && t.used + 2 > t.ev.size()) {
    if (t.ev.size() >= 200000) return;  // bounded: stop recording past 100k launches
    for (int k = 0; k < 64; ++k) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      t.ev.push_back(e);
    }
  }
  if (!end) {
    (void)hipEventRecord(t.ev[t.used], c->stream);
    t.next_cont = cont;
  } else if (t.used + 1 < t.ev.size()) {
    (void)hipEventRecord(t.ev[t.used + 1], c->stream);
    t.cont.resize(t.used / 2 + 1);
    t.cont[t.used / 2] = t.next_cont ? 1 : 0;
    t.used += 2;
  }
}
}  // namespace qx

extern "C" {

// launch timers: mode 1 clears and starts timer id, mode 0 stops it and
// returns the summed duration (ms) and the number of timed launches
int qx_launch_timer(qx_ctx* h, int id, int mode, double* ms, int* count) {
  Ctx* c = &h->c;
  if (id < 0 || id >= Ctx::kTimers) return fail(-1, "qx_launch_timer: bad timer id");
  auto& t = c->lt[id];
  if (mode == 1) {
    t.on = true;
    t.used = 0;
    return 0;
  }
  t.on = false;
  QX_CHECK(hipStreamSynchronize(c->stream));
  double sum = 0.0;
  int launches = 0;
  for (size_t k = 0; k + 1 < t.used; k += 2) {
    float e = 0.f;
    QX_CHECK(hipEventElapsedTime(&e, t.ev[k], t.ev[k + 1]));
    sum += e;
    if (k / 2 >= t.cont.size() || !t.cont[k / 2]) ++launches;
  }
  if (ms) *ms = sum;
  if (count) *count = launches;
  return 0;
}

// phases of the last SCF band step, ms: energy pass, rotation, CG, schmidt, density
int qx_scf_phase_ms(qx_ctx* h, float* ms) {
  for (int i = 0; i < 5; ++i) ms[i] = h->c.phase_ms[i];
  return 0;
}

int qx_scf_band_step(qx_ctx* h, int psi, int hpsi, int nb, const double* occ, int noband,
                     double rvol, double* eig, double* rho, int* lcholesky, int prec) {
  return qx_scf_band_step_ex(h, psi, hpsi, nb, occ, noband, rvol, eig, rho, lcholesky, prec, 0);
}

int qx_event_record(qx_ctx* h, int slot) {
  QX_CHECK(hipEventRecord(h->c.ev[slot & 7], h->c.stream));
  return 0;
}

int qx_event_elapsed(qx_ctx* h, int s0, int s1, float* ms) {
  QX_CHECK(hipEventSynchronize(h->c.ev[s1 & 7]));
  QX_CHECK(hipEventElapsedTime(ms, h->c.ev[s0 & 7], h->c.ev[s1 & 7]));
  return 0;
}

// ---------------- drop-ins --------------------------------------------------
int qx_rfft3(int key, double* fft3x, double* fft3y, int k1, int k2, int k3, int* ierror) {
  int rc = fullgrid_rfft3(key, fft3x, fft3y, k1, k2, k3);
  if (ierror) *ierror = rc ? 1 : 0;
  return rc ? fail(rc, "qx_rfft3: failed (%d)", rc) : 0;
}

int qx_rfft3_dev_noalloc(int key, double* d_fft3x, double* d_fft3y, int k1, int k2, int k3,
                         void* stream, int* ierror) {
  long maxp = 0;
  int maxa = 0;
  noalloc_limits(&maxp, &maxa);
  if (k1 < 1 || k2 < 1 || k3 < 1 || k1 > maxa || k2 > maxa || k3 > maxa ||
      (long)k1 * k2 * k3 > maxp) {
    if (ierror) *ierror = 1;
    return fail(-2, "qx_rfft3_dev_noalloc: grid %d x %d x %d exceeds the scratch limit "
                "(%ld points, %d per axis)", k1, k2, k3, maxp, maxa);
  }
  int rc = fullgrid_rfft3_dev(key, d_fft3x, d_fft3y, k1, k2, k3,
                              static_cast<hipStream_t>(stream));
  if (ierror) *ierror = rc ? 1 : 0;
  return rc ? fail(rc, "qx_rfft3_dev_noalloc: failed (%d)", rc) : 0;
}

int qx_set_gemm_path(int path) {
  if (path < GEMM_OWN || path > GEMM_OWN_F32) return fail(-1, "qx_set_gemm_path: bad path %d", path);
  set_gemm_path(path);
  return 0;
}

int qx_gemm_path(void) { return gemm_path(); }

int qx_set_gemm_schedule(int mode) {
  if (mode < 0 || mode > 2) return fail(-1, "qx_set_gemm_schedule: bad mode %d", mode);
  mg_set_schedule(mode);
  return 0;
}

int qx_gemm_schedule(void) { return mg_schedule(); }

int qx_set_cg_graph(qx_ctx* h, int on) {
  Ctx* c = &h->c;
  c->cg_graph_on = on != 0;
  if (!on) cg_graphs_destroy(c);
  return 0;
}

int qx_cg_graph_count(qx_ctx* h) {
  return h->c.cg_graphs ? cg_graph_entries(&h->c) : 0;
}

int qx_noalloc_limits(long* max_points, int* max_axis) {
  return noalloc_limits(max_points, max_axis);
}

int qx_fft3(int key, double* fft3x, double* fft3y, int k1, int k2, int k3, int* ierror) {
  int rc = fullgrid_fft3(key, fft3x, fft3y, k1, k2, k3);
  if (ierror) *ierror = rc ? 1 : 0;
  return rc ? fail(rc, "qx_fft3: failed (%d)", rc) : 0;
}

int qx_hcnonloc(qx_ctx* h, const double* cgjr, double* rhcr) {
  Ctx* c = &h->c;
  int rc = put_internal(c, c->d_slot[0], cgjr, 1);
  if (rc) return rc;
  if ((rc = put_internal(c, c->d_slot[1], rhcr, 1))) return rc;
  // slot 2 = (T + V) psi; slot 1 += slot 2 - T psi
  if ((rc = local_apply(c, c->d_slot[0], c->d_slot[2], 1, PREC_FP64, 0))) return rc;
  hipLaunchKernelGGL(k_add_minus_kin, dim3(cdiv(c->nplwex, 256)), dim3(256), 0, c->stream,
                     c->d_slot[1], c->d_slot[2], c->d_slot[0], c->d_kin, c->nplwex);
  return get_reference(c, c->d_slot[1], rhcr, 1);
}

int qx_hckbpp_g(qx_ctx* h, const double* cgjr, double* rhcr, int nb, double* slmir) {
  Ctx* c = &h->c;
  if (nb > c->nbmax) return fail(-1, "qx_hckbpp_g: nb > nbmax");
  int rc = put_internal(c, c->d_slot[0], cgjr, nb);
  if (rc) return rc;
  if ((rc = put_internal(c, c->d_slot[1], rhcr, nb))) return rc;
  double* P = c->d_slot[2];  // projections staged here (npj x nb fits a slot)
  if ((size_t)c->npj * nb > (size_t)c->nplwex * c->nbmax) return fail(-1, "too many projectors");
  if ((rc = kb_apply(c, c->d_slot[0], c->d_slot[1], nb, 1, slmir ? P : nullptr, PREC_FP64)))
    return rc;
  if (slmir) {
    std::vector<double> ph((size_t)c->npj * nb);
    QX_CHECK(hipMemcpyAsync(ph.data(), P, sizeof(double) * ph.size(), hipMemcpyDeviceToHost,
                            c->stream));
    QX_CHECK(hipStreamSynchronize(c->stream));
    std::memset(slmir, 0, sizeof(double) * (size_t)c->lmmax * c->natom * nb);
    for (int b = 0; b < nb; ++b)
      for (int p = 0; p < c->npj; ++p)
        slmir[c->proj_l[p] + (size_t)c->lmmax * (c->proj_atom[p] + (size_t)c->natom * b)] =
            ph[(size_t)b * c->npj + p];
  }
  return get_reference(c, c->d_slot[1], rhcr, nb);
}

int qx_kbnlpp_frc2_slot(qx_ctx* h, int slot, int nb, const double* occ, const double* gx,
                        const double* gy, const double* gz, double* fnlc) {
  Ctx* c = &h->c;
  if (slot < 0 || slot >= Ctx::kSlots || (slot >= 2 && slot <= 4))
    return fail(-1, "qx_kbnlpp_frc2_slot: slot 0, 1 or 5 (2-4 hold the G-weighted bands)");
  if (nb < 1 || nb > c->nbmax) return fail(-1, "qx_kbnlpp_frc2_slot: 1 <= nb <= nbmax");
  const int rc = kb_forces(c, c->d_slot[slot], nb, occ, gx, gy, gz, fnlc);
  if (rc) return fail(rc, "qx_kbnlpp_frc2_slot failed");
  return 0;
}

int qx_kbnlpp_frc2(qx_ctx* h, const double* cgjr, int nb, const double* occ, const double* gx,
                   const double* gy, const double* gz, double* fnlc) {
  Ctx* c = &h->c;
  if (nb < 1 || nb > c->nbmax) return fail(-1, "qx_kbnlpp_frc2: 1 <= nb <= nbmax");
  const int rc = put_internal(c, c->d_slot[0], cgjr, nb);
  if (rc) return rc;
  return qx_kbnlpp_frc2_slot(h, 0, nb, occ, gx, gy, gz, fnlc);
}

int qx_hckbpp_ib_g(qx_ctx* h, const double* cgjr, double* rhcr) {
  return qx_hckbpp_g(h, cgjr, rhcr, 1, nullptr);
}

int qx_schmidt(qx_ctx* h, double* cgjr, int nband, int* lcholesky) {
  Ctx* c = &h->c;
  if (nband > c->nbmax) return fail(-1, "qx_schmidt: nband > nbmax");
  int rc = put_internal(c, c->d_slot[0], cgjr, nband);
  if (rc) return rc;
  if ((rc = schmidt_dev(c, c->d_slot[0], nband, lcholesky, PREC_FP64))) return rc;
  if (!*lcholesky) return 0;
  return get_reference(c, c->d_slot[0], cgjr, nband);
}

int qx_softchg(qx_ctx* h, const double* cgjr, int nb, const double* occ, double rvol, double* rho) {
  Ctx* c = &h->c;
  if (nb > c->nbmax) return fail(-1, "qx_softchg: nb > nbmax");
  int rc = put_internal(c, c->d_slot[0], cgjr, nb);
  if (rc) return rc;
  return qx_density_slot(h, 0, nb, occ, rvol, rho, PREC_FP64);
}


// RsEIGQR with ifvec=1, ifsrt=2 (Sources/funcs.f90:380-764) on a host matrix:
// a(n,n) column-major (lower triangle used) -> eigenvectors in its columns,
// w ascending.
int qx_rseigqr(qx_ctx* h, double* a, int n, double* w) {
  Ctx* c = &h->c;
  if (n < 1 || n > c->nbmax) return fail(-1, "qx_rseigqr: need 1 <= n <= nbmax");
  double* C = c->d_small;
  double* H = C + (size_t)n * n;
  double* Z = H + (size_t)n * n;
  double* work = Z + (size_t)n * n;
  double* wd = c->d_scal + 40 * (size_t)c->nbmax;
  // symmetric fill from the lower triangle: C[hi*n+lo] = a(hi, lo)
  std::vector<double> ct((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) ct[(size_t)(i > j ? i : j) * n + (i > j ? j : i)] = a[i + (size_t)n * j];
  QX_CHECK(hipMemcpyAsync(C, ct.data(), sizeof(double) * ct.size(), hipMemcpyHostToDevice, c->stream));
  int rc = dense_symmetrize(c, C, n, H, nullptr);
  if (rc) return rc;
  if ((rc = dense_eigh(c, H, n, Z, wd, work, nullptr))) return rc;
  QX_CHECK(hipMemcpyAsync(a, Z, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipMemcpyAsync(w, wd, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// icholesky (Sources/schmidt.f90:383-423): s(n,n) column-major, lower triangle
// of a symmetric positive-definite matrix -> L^-1 in the lower triangle (upper
// zeroed); *ok = 1 when the matrix was positive definite.
int qx_icholesky(qx_ctx* h, double* s, int n, int* ok) {
  Ctx* c = &h->c;
  if (n < 1 || n > c->nbmax) return fail(-1, "qx_icholesky: need 1 <= n <= nbmax");
  double* S = c->d_small;
  double* X = c->d_small + (size_t)(n + 1) * n;
  int* okd = reinterpret_cast<int*>(c->d_scal + 62 * (size_t)c->nbmax);
  QX_CHECK(hipMemcpyAsync(S, s, sizeof(double) * (size_t)n * n, hipMemcpyHostToDevice, c->stream));
  int rc = dense_chol_inv(c, S, n, X, okd, okd + 2);
  if (rc) return rc;
  std::vector<double> xr((size_t)n * n);
  int okh[2] = {0, 0};
  QX_CHECK(hipMemcpyAsync(xr.data(), X, sizeof(double) * xr.size(), hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipMemcpyAsync(okh, okd, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (okh[1]) return fail(-20, "qx_icholesky: a grid barrier of the Cholesky gave up");
  *ok = okh[0];
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) s[i + (size_t)n * j] = xr[(size_t)i * n + j];  // row-major -> col-major
  return 0;
}

// ---------------- local-pseudopotential forces ------------------------------
int qx_set_lopp(qx_ctx* h, int ntype, int nplw5, const double* svco, const double* gx,
                const double* gy, const double* gz) {
  int rc = lopp_set(&h->c, ntype, nplw5, svco, gx, gy, gz);
  return rc ? fail(rc, "qx_set_lopp: bad sizes (%d)", rc) : 0;
}

int qx_flocal_g(qx_ctx* h, double* floc, int ntype, const int* nhk1, const int* nhk2,
                const int* iatmpt, int natom, const int* llking, int nplw5, const int* nga,
                const int* ngb, const int* ngc, const double* rhgr, int kmax1, int kmax2,
                int kmax3, const double* dcq1, const double* dcq2, const double* dcq3,
                const double* dsq1, const double* dsq2, const double* dsq3) {
  const int km[3] = {kmax1, kmax2, kmax3};
  const double* dc[3] = {dcq1, dcq2, dcq3};
  const double* ds[3] = {dsq1, dsq2, dsq3};
  int rc = flocal_host(&h->c, floc, ntype, nhk1, nhk2, iatmpt, natom, llking, nplw5, nga, ngb,
                       ngc, rhgr, km, dc, ds);
  if (rc == -4) return fail(rc, "qx_flocal_g: call qx_set_lopp with this ntype and nplw5 first");
  return rc ? fail(rc, "qx_flocal_g: bad ion ranges or G beyond kmax (%d)", rc) : 0;
}

int qx_flocal_g_dev(qx_ctx* h, int ntype, const int* nhk1, const int* nhk2, const int* iatmpt,
                    const int* llking, int natom, double* floc) {
  Ctx* c = &h->c;
  if (!c->scf.d_rhgr || c->scf.nplw5 != c->lopp.nplw5)
    return fail(-2, "qx_flocal_g_dev: the density sphere (qx_scf_set_grid) and qx_set_lopp "
                    "must have the same nplw5");
  int rc = flocal_dev(c, ntype, nhk1, nhk2, iatmpt, llking, natom, c->scf.d_rhgr, c->scf.nplw5,
                      floc);
  if (rc == -5) return fail(rc, "qx_flocal_g_dev: call qx_stscfn_pw over the density sphere "
                                "(nplw5) for these atoms first");
  if (rc == -4) return fail(rc, "qx_flocal_g_dev: call qx_set_lopp first");
  return rc ? fail(rc, "qx_flocal_g_dev: bad ion ranges (%d)", rc) : 0;
}

int qx_vlocal_g(qx_ctx* h, int ntype, const int* nhk1, const int* nhk2, const int* iatmpt,
                const int* llking, int natom, double rvol, int accumulate, double* vext,
                double* thrhgr) {
  Ctx* c = &h->c;
  if (!c->scf.d_ijk || !c->scf.d_vext || c->scf.nplw5 != c->lopp.nplw5)
    return fail(-2, "qx_vlocal_g: the density sphere (qx_scf_set_grid) and qx_set_lopp must "
                    "have the same nplw5");
  int rc = vlocal_sphere(c, ntype, nhk1, nhk2, iatmpt, llking, natom, rvol, c->scf.nplw5);
  if (rc == -5) return fail(rc, "qx_vlocal_g: call qx_stscfn_pw over the density sphere "
                                "(nplw5) for these atoms first");
  if (rc == -4) return fail(rc, "qx_vlocal_g: call qx_set_lopp first");
  if (rc) return fail(rc, "qx_vlocal_g: bad ion ranges or volume (%d)", rc);
  if ((rc = vext_from_sphere(c, c->lopp.d_zin, accumulate))) return fail(rc, "qx_vlocal_g: FFT");
  if (vext)
    QX_CHECK(hipMemcpyAsync(vext, c->scf.d_vext, sizeof(double) * c->N, hipMemcpyDeviceToHost,
                            c->stream));
  if (thrhgr)
    QX_CHECK(hipMemcpyAsync(thrhgr, c->lopp.d_thr, sizeof(double) * 2 * (c->scf.nplw5 + 1),
                            hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------- restart files ---------------------------------------------
int qx_svwfns_slot(qx_ctx* h, const char* fname, int slot, int nb, const double* eig,
                   const double* occ, int lsreal8) {
  Ctx* c = &h->c;
  if (slot < 0 || slot > 4 || nb < 1 || nb > c->nbmax)
    return fail(-1, "qx_svwfns_slot: bad slot or band count");
  std::vector<double> g((size_t)c->nplwex * nb);
  int rc = get_reference(c, c->d_slot[slot], g.data(), nb);
  if (rc) return rc;
  rc = qx_svwfns(fname, g.data(), c->nplwex, c->nplw, nb, eig, occ, lsreal8);
  return rc ? fail(rc, "qx_svwfns_slot: cannot write %s", fname ? fname : "") : 0;
}

int qx_rdwfns_slot(qx_ctx* h, const char* fname, int slot, int nb, double* eig, double* occ,
                   int* nband_file) {
  Ctx* c = &h->c;
  if (slot < 0 || slot > 1 || nb < 1 || nb > c->nbmax)
    return fail(-1, "qx_rdwfns_slot: slot 0 or 1, nb <= nbmax");
  std::vector<double> g((size_t)c->nplwex * nb, 0.0);
  int nf = 0;
  int rc = qx_rdwfns(fname, g.data(), c->nplwex, c->nplw, nb, eig, occ, &nf);
  if (rc == -2) return fail(rc, "qx_rdwfns_slot: %s holds another G sphere", fname);
  if (rc) return fail(rc, "qx_rdwfns_slot: cannot read %s", fname ? fname : "");
  if ((rc = put_internal(c, c->d_slot[slot], g.data(), nb))) return rc;
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (nband_file) *nband_file = nf;
  return 0;
}

// ---------------- wave-function extrapolation -------------------------------
int qx_aspc_init(qx_ctx* h, int nprvmx, int ihest) {
  int rc = aspc_init(&h->c, nprvmx, ihest);
  return rc ? fail(rc, "qx_aspc_init: need 1 <= nprvmx <= 10 and 0 <= ihest <= nprvmx (%d)", rc)
            : 0;
}

int qx_sbalinaspc_slot(qx_ctx* h, int slot, int nb, int nstep, int nstep_ini, int* info) {
  if (!h) return fail(-1, "qx_sbalinaspc_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_sbalinaspc_slot", nb, slot, -1, kSlotPair01)) return e;
  int iaspc = 0, lchol = 1;
  int rc = aspc_step(c, c->d_slot[slot], nb, nstep, nstep_ini, &iaspc, &lchol);
  if (rc == -21)
    return fail(rc, "qx_sbalinaspc_slot: the previous bands do not overlap the current ones");
  if (rc) return fail(rc, "qx_sbalinaspc_slot: call qx_aspc_init first, nb <= nbmax (%d)", rc);
  if (info) {
    info[0] = iaspc;
    info[1] = lchol;
  }
  return 0;
}

int qx_roteig_slot(qx_ctx* h, int slot, int nb, int kpwvx) {
  if (!h) return fail(-1, "qx_roteig_slot: null context");
  Ctx* c = &h->c;
  if (int e = check_slots(c, "qx_roteig_slot", nb, slot, -1, kSlotPair01)) return e;
  int rc = aspc_roteig(c, c->d_slot[slot], nb, kpwvx);
  if (rc == -21) return fail(rc, "qx_roteig_slot: A^T A is not positive definite");
  if (rc) return fail(rc, "qx_roteig_slot: call qx_aspc_init first (%d)", rc);
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int qx_aspc_get(qx_ctx* h, int k, int nb, double* gdcr) {
  Ctx* c = &h->c;
  if (k < 0 || k >= c->aspc.nprvmx || nb < 1 || nb > c->nbmax)
    return fail(-1, "qx_aspc_get: bad set or band count");
  return get_reference(c, c->aspc.d_prv[k], gdcr, nb);
}

int qx_aspc_put(qx_ctx* h, int k, int nb, const double* gdcr, int keypwv) {
  Ctx* c = &h->c;
  if (k < 0 || k >= c->aspc.nprvmx || nb < 1 || nb > c->nbmax || keypwv < 0)
    return fail(-1, "qx_aspc_put: bad set, band count or keypwv");
  int rc = put_internal(c, c->aspc.d_prv[k], gdcr, nb);
  if (rc) return rc;
  c->aspc.keypwv = keypwv;
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------- Ewald reciprocal-space forces ------------------------------
int qx_set_ewald(qx_ctx* h, int numk2, const int* nga, const int* ngb, const int* ngc,
                 const double* gx, const double* gy, const double* gz, const double* recnrm,
                 const double* recnrmex, double gamma, double volume) {
  int rc = ewald_set(&h->c, numk2, nga, ngb, ngc, gx, gy, gz, recnrm, recnrmex, gamma, volume);
  return rc ? fail(rc, "qx_set_ewald: bad sizes or parameters (%d)", rc) : 0;
}

int qx_frcrec(qx_ctx* h, double* fclm, int ntype, const int* nhk1, const int* nhk2,
              const double* zv, int nel, double* ecolmb, double* atom_ecolmb, const int* iatoit,
              const int* iatmpt, int natom, int nion, int kcsma1, int kcsma2, int kcsma3,
              const double* dcq1, const double* dcq2, const double* dcq3, const double* dsq1,
              const double* dsq2, const double* dsq3) {
  const int km[3] = {kcsma1, kcsma2, kcsma3};
  const double* dc[3] = {dcq1, dcq2, dcq3};
  const double* ds[3] = {dsq1, dsq2, dsq3};
  int rc = ewald_frcrec_host(&h->c, fclm, ntype, nhk1, nhk2, zv, nel, ecolmb, atom_ecolmb, iatoit,
                             iatmpt, natom, nion, km, dc, ds);
  if (rc == -4) return fail(rc, "qx_frcrec: call qx_set_ewald first");
  return rc ? fail(rc, "qx_frcrec: bad ion map or tables narrower than the G list (%d)", rc) : 0;
}

int qx_frcrec_dev(qx_ctx* h, double* fclm, int ntype, const int* nhk1, const int* nhk2,
                  const double* zv, int nel, double* ecolmb, double* atom_ecolmb,
                  const int* iatoit, const int* iatmpt, int natom, int nion) {
  int rc = ewald_frcrec_dev(&h->c, fclm, ntype, nhk1, nhk2, zv, nel, ecolmb, atom_ecolmb, iatoit,
                            iatmpt, natom, nion);
  if (rc == -4) return fail(rc, "qx_frcrec_dev: call qx_set_ewald first");
  if (rc == -5) return fail(rc, "qx_frcrec_dev: call qx_stscfn_pw for these atoms with kmax "
                                "covering the Ewald G list first");
  return rc ? fail(rc, "qx_frcrec_dev: bad ion map (%d)", rc) : 0;
}

// ---------------- SCF glue --------------------------------------------------
// xclda (Sources/vxc.f90:961-1017) on host arrays; rhocore may be null
int qx_xclda(const double* rho, const double* rhocore, int mshnod, double rdelv, double* vexc,
             double* eeexc, double* e4) {
  int rc = dropin_xclda(rho, rhocore, mshnod, rdelv, vexc, eeexc, e4);
  return rc ? fail(rc, "qx_xclda: failed (%d)", rc) : 0;
}

int qx_harting(double* vhar, const double* rhgr, int nplw5, const double* recnrmex,
               const int* mfd2ft, int ntotfd, const int* nd1vks, const int* ijkgd) {
  int rc = dropin_sphere(1, vhar, const_cast<double*>(rhgr), nplw5, recnrmex, mfd2ft, ntotfd,
                         nd1vks, ijkgd);
  return rc ? fail(rc, "qx_harting: bad sizes or index maps (%d)", rc) : 0;
}

int qx_chgr2g(const double* rho, double* rhgr, int nplw5, const int* mfd2ft, int ntotfd,
              const int* nd1vks, const int* ijkgd) {
  int rc = dropin_sphere(2, const_cast<double*>(rho), rhgr, nplw5, nullptr, mfd2ft, ntotfd,
                         nd1vks, ijkgd);
  return rc ? fail(rc, "qx_chgr2g: bad sizes or index maps (%d)", rc) : 0;
}

int qx_chgg2r(double* rho, const double* rhgr, int nplw5, const int* mfd2ft, int ntotfd,
              const int* nd1vks, const int* ijkgd) {
  int rc = dropin_sphere(0, rho, const_cast<double*>(rhgr), nplw5, nullptr, mfd2ft, ntotfd,
                         nd1vks, ijkgd);
  return rc ? fail(rc, "qx_chgg2r: bad sizes or index maps (%d)", rc) : 0;
}

int qx_mxchgg_alloc(qx_ctx* h, int nplwcex, int imxchg, int itratn, double cmetric,
                    double wkerker, double g2min, double g2max, const double* dkgnrm,
                    double rvol) {
  int rc = scf_mixer_alloc(&h->c, nplwcex, imxchg, itratn, cmetric, wkerker, g2min, g2max, dkgnrm,
                           rvol);
  return rc ? fail(rc, "qx_mxchgg_alloc: imxchg must be 0, 1 or 3 and 1 <= itratn <= 32 (%d)", rc)
            : 0;
}

int qx_mxchgg(qx_ctx* h, int* lreset, double* rhgr, double* rinr, int nplw5, double aslh,
              double bslh, double* glocal, const int* mfd2ft, int ntotfd, const int* ijkgd,
              int lmixhold) {
  int rc = dropin_mxchgg(&h->c, lreset, rhgr, rinr, nplw5, aslh, bslh, glocal, mfd2ft, ntotfd,
                         ijkgd, lmixhold);
  return rc ? fail(rc, "qx_mxchgg: mixer not set up or bad sizes (%d)", rc) : 0;
}

int qx_scf_set_grid(qx_ctx* h, int nplw5, const int* ijkgd, const double* recnrmex,
                    const double* vext, const double* rhocore, double rdelv) {
  int rc = scf_set_grid(&h->c, nplw5, ijkgd, recnrmex, vext, rhocore, rdelv);
  return rc ? fail(rc, "qx_scf_set_grid: bad sphere or index map (%d)", rc) : 0;
}

int qx_pwpchg_slot(qx_ctx* h, int slot, int nb, const double* occ, double rvol, int prec) {
  Ctx* c = &h->c;
  if (slot < 0 || slot >= Ctx::kSlots || nb < 1 || nb > c->nbmax)
    return fail(-1, "qx_pwpchg_slot: bad slot or band count");
  int rc = scf_pwpchg(c, c->d_slot[slot], nb, occ, rvol, prec);
  return rc ? fail(rc, "qx_pwpchg_slot: call qx_scf_set_grid first (%d)", rc) : 0;
}

int qx_outchg(qx_ctx* h) {
  int rc = scf_outchg(&h->c);
  return rc ? fail(rc, "qx_outchg: call qx_scf_set_grid first (%d)", rc) : 0;
}

int qx_pwpmxchgg(qx_ctx* h, int lreset, double aslh, double bslh, int lmixhold, int* info) {
  int rc = scf_mix(&h->c, lreset, aslh, bslh, lmixhold, info);
  return rc ? fail(rc, "qx_pwpmxchgg: call qx_scf_set_grid and qx_mxchgg_alloc first (%d)", rc)
            : 0;
}

int qx_scf_potential(qx_ctx* h, double* ene) {
  int rc = scf_potential(&h->c, ene);
  return rc ? fail(rc, "qx_scf_potential: call qx_scf_set_grid first (%d)", rc) : 0;
}

int qx_scf_get(qx_ctx* h, int what, double* host) {
  int rc = scf_io(&h->c, what, host, false);
  return rc ? fail(rc, "qx_scf_get: unknown or unset array %d", what) : 0;
}

int qx_scf_put(qx_ctx* h, int what, const double* host) {
  int rc = scf_io(&h->c, what, const_cast<double*>(host), true);
  return rc ? fail(rc, "qx_scf_put: only rho (0), rhgr (1) and rinr (2) can be set (%d)", what)
            : 0;
}

// ---- Bloch bands at several k-points (kpoint.hip)
int qx_ksphere(const double* bvec, const double* kfrac, double ecut, const int* nd, int cap,
               int* hkl, double* kin) {
  if (!bvec || !kfrac || !nd || !(ecut > 0.0) || cap < 0)
    return fail(-1, "qx_ksphere: bad arguments");
  int n = kp_sphere_host(bvec, kfrac, ecut, nd[0], nd[1], nd[2], cap, hkl, kin);
  if (n < -100000000) return fail(-1, "qx_ksphere: the sphere does not fit the grid");
  return n;
}

int qx_kpoints_set(qx_ctx* h, int nk, const double* kfrac, const double* bvec, double ecut,
                   int nb, int* npw) {
  if (!h || !kfrac || !bvec) return fail(-1, "qx_kpoints_set: null argument");
  int rc = kp_set(&h->c, nk, kfrac, bvec, ecut, nb, npw);
  if (rc == -1) return fail(rc, "qx_kpoints_set: bad arguments");
  if (rc == -2)
    return fail(rc, "qx_kpoints_set: a k-point sphere does not fit the FFT grid (or the grid "
                    "exceeds the plane kernel: n1 n2 <= 8192, n3 <= 128)");
  return rc ? fail(rc, "qx_kpoints_set: device error %d", rc) : 0;
}

int qx_kpoints_sphere(qx_ctx* h, int ik, int* hkl, double* kin) {
  int rc = h ? kp_sphere(&h->c, ik, hkl, kin) : -1;
  return rc ? fail(rc, "qx_kpoints_sphere: no k-points set or ik out of range") : 0;
}

int qx_kslot_put(qx_ctx* h, int slot, int ik, const double* psi, int nb) {
  int rc = h && psi ? kp_slot_put(&h->c, slot, ik, psi, nb) : -1;
  return rc ? fail(rc, "qx_kslot_put: bad slot, k-point or band count (%d)", rc) : 0;
}

int qx_kslot_get(qx_ctx* h, int slot, int ik, double* psi, int nb) {
  int rc = h && psi ? kp_slot_get(&h->c, slot, ik, psi, nb) : -1;
  return rc ? fail(rc, "qx_kslot_get: bad slot, k-point or band count (%d)", rc) : 0;
}

int qx_kset_kb(qx_ctx* h, int ik, int npj, const double* beta, const double* D) {
  int rc = h && (npj == 0 || (beta && D)) ? kp_set_kb(&h->c, ik, npj, beta, D) : -1;
  return rc ? fail(rc, "qx_kset_kb: no k-points set, bad k-point or projector count (%d)", rc)
            : 0;
}

int qx_ksolve_slot(qx_ctx* h, int src, int dst, int max_iter, double tol, double* eig,
                   int* iters) {
  int rc = h ? kp_solve(&h->c, src, dst, max_iter, tol, eig, iters) : -1;
  if (rc == -21)
    return fail(rc, "qx_ksolve_slot: the overlap of a Rayleigh-Ritz basis is not positive definite");
  return rc ? fail(rc, "qx_ksolve_slot: bad slots, no k-points set, or a solver error (%d)", rc)
            : 0;
}

int qx_khpsi_slot(qx_ctx* h, int src, int dst, int nb, int per_band) {
  int rc = h ? kp_hpsi(&h->c, src, dst, nb, per_band) : -1;
  return rc ? fail(rc, "qx_khpsi_slot: bad slots or band count, or no k-points set (%d)", rc) : 0;
}

}  // extern "C"

