// aspc.hip -- wave-function extrapolation between MD steps, device-resident
// (sbalinaspc, Sources/sbalin.f90:369-712, one spin, no ultrasoft part):
//   roteig2 (772-1373): the previous sets are turned to the current bands by
//     the orthogonal factor of their overlap, A = <prev_1|cur>, B = A^T A =
//     V diag(w) V^T, P_k <- P_k (A V diag(w)^-1/2 V^T) -- two FP64 GEMMs of
//     nplwex x nb x nb per set, three of nb^3 and the device eigensolver;
//   the always-stable predictor-corrector combination with the coefficients
//     of set_coef_ASPC (77-170), the history shift, and Gram-Schmidt.
// The history lives in HBM (nprvmx x 189 MB at Si216); shifting it rotates
// the set pointers.
#include <algorithm>
#include <cmath>
#include <vector>
#include "context.h"

namespace qx {

int dense_eigh(Ctx* c, double* a, int n, double* z, double* w, double* work, int* info);

namespace {

inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

// T2t(k, j) = T1(j, k) / sqrt(w_k), Vt(k, i) = V(i, k); flag[0] = 1 if some w <= 0
__global__ void k_aspc_tscale(const double* __restrict__ T1, const double* __restrict__ V,
                              const double* __restrict__ w, int nb, double* __restrict__ T2t,
                              double* __restrict__ Vt, int* __restrict__ flag) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nb * nb) return;
  const int k = (int)(i % nb), j = (int)(i / nb);  // T2t(k, j) at k + nb j
  const double wk = w[k];
  if (j == 0 && !(wk > 0.0)) flag[0] = 1;
  T2t[i] = T1[j + (size_t)nb * k] / sqrt(wk);
  Vt[i] = V[j + (size_t)nb * k];
}

struct SetPtrs {
  const double* p[Ctx::Aspc::kMax];
  double c[Ctx::Aspc::kMax + 1];
};

// out = c0 g + sum_{j < nset} c_{j+1} p_j, in the reference's order
__global__ void k_aspc_combine(const double* __restrict__ g, SetPtrs sp, int nset, long n,
                               double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = sp.c[0] * g[i];
  for (int j = 0; j < nset; ++j) v = v + sp.c[j + 1] * sp.p[j][i];
  out[i] = v;
}

}  // namespace

void aspc_destroy(Ctx* c) {
  auto& S = c->aspc;
  for (auto& p : S.d_prv)
    if (p) (void)hipFree(p);
  if (S.d_work) (void)hipFree(S.d_work);
  S = Ctx::Aspc{};
}

// set_coef_ASPC with ichest = 0 (kaspc = ihest + 1) and the history buffers
int aspc_init(Ctx* c, int nprvmx, int ihest) {
  if (nprvmx < 1 || nprvmx > Ctx::Aspc::kMax || ihest < 0 || ihest > nprvmx) return -2;
  aspc_destroy(c);
  auto& S = c->aspc;
  S.nprvmx = nprvmx;
  S.ihest = ihest;
  S.kaspc = ihest + 1;
  S.keypwv = 0;
  const int K = S.kaspc;
  S.coef.assign((size_t)K * K, 0.0);
  for (int k = 2; k <= K; ++k)
    for (int j = 1; j <= k; ++j) {
      double a = 1.0, b = 1.0;
      for (int i = k - j + 1; i <= k - 1; ++i) a = a * (double)i;
      for (int i = k; i <= k + j; ++i) b = b * (double)i;
      S.coef[(j - 1) + (size_t)K * (k - 1)] =
          (double)(((j + 1) % 2) ? -1 : 1) * (double)(j * (2 * k) * (2 * k - 1)) * a / b;
    }
  const size_t set = (size_t)c->nplwex * c->nbmax;
  for (int i = 0; i < nprvmx; ++i) {
    QX_CHECK(hipMalloc(&S.d_prv[i], sizeof(double) * set));
    QX_CHECK(hipMemsetAsync(S.d_prv[i], 0, sizeof(double) * set, c->stream));
  }
  const size_t nb2 = (size_t)c->nbmax * c->nbmax;
  QX_CHECK(hipMalloc(&S.d_work, sizeof(double) * (8 * nb2 + c->nbmax + 64)));
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// roteig2 on the device history: sets 1..min(nprvmx, kpwvx) turned to cur
int aspc_roteig(Ctx* c, const double* cur, int nb, int kpwvx) {
  auto& S = c->aspc;
  if (!S.d_work || nb < 1 || nb > c->nbmax) return -2;
  const int n = c->nplwex;
  const size_t nb2 = (size_t)nb * nb;
  double* A = S.d_work;
  double* B = A + nb2;
  double* Z = B + nb2;
  double* T1 = Z + nb2;
  double* T2t = T1 + nb2;
  double* Vt = T2t + nb2;
  double* Y = Vt + nb2;
  double* ework = Y + nb2;                 // nb^2 for the eigensolver's fallback path
  double* w = ework + nb2;                 // nb
  int* flag = reinterpret_cast<int*>(w + nb);
  int* info = flag + 2;
  // A(i, j) = <prev1_i|cur_j>, packed product
  int rc = gemm_tn(c, nb, nb, n, 2.0, S.d_prv[0], n, cur, n, 0.0, A, nb, 1);
  if (rc) return rc;
  if ((rc = gemm_tn(c, nb, nb, nb, 1.0, A, nb, A, nb, 0.0, B, nb, 0))) return rc;  // A^T A
  if ((rc = dense_eigh(c, B, nb, Z, w, ework, info))) return rc;
  if ((rc = gemm_nn(c, nb, nb, nb, 1.0, A, nb, Z, nb, 0.0, T1, nb))) return rc;   // A V
  QX_CHECK(hipMemsetAsync(flag, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_aspc_tscale, dim3(cdivu((long)nb2, 256)), dim3(256), 0, c->stream, T1, Z,
                     w, nb, T2t, Vt, flag);
  QX_CHECK(hipGetLastError());
  // Y = A V diag(w)^-1/2 V^T = U^T, the transpose of roteig2's aijr
  if ((rc = gemm_tn(c, nb, nb, nb, 1.0, T2t, nb, Vt, nb, 0.0, Y, nb, 0))) return rc;
  int fh = 0;
  QX_CHECK(hipMemcpyAsync(&fh, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (fh) return -21;  // A^T A not positive definite: the sets do not overlap the bands
  for (int k = 0; k < S.nprvmx && k < kpwvx; ++k) {
    double* out = c->d_slot[5];
    if ((rc = gemm_nn(c, n, nb, nb, 1.0, S.d_prv[k], n, Y, nb, 0.0, out, n))) return rc;
    // the turned set takes the staging buffer's place; the old set becomes staging
    std::swap(S.d_prv[k], c->d_slot[5]);
  }
  return 0;
}

// one call of sbalinaspc (369-712) for the bands psi (lvand false):
// *iaspc = 0 when only the history was started; *lchol = Gram-Schmidt result
int aspc_step(Ctx* c, double* psi, int nb, int nstep, int nstep_ini, int* iaspc_out,
              int* lchol) {
  auto& S = c->aspc;
  if (!S.d_work || nb < 1 || nb > c->nbmax) return -2;
  if (psi == c->d_slot[5]) return -2;
  const size_t set = (size_t)c->nplwex * nb;
  *lchol = 1;
  if (nstep == nstep_ini + 1 || S.keypwv == 0) {
    S.keypwv = 1;
    QX_CHECK(hipMemcpyAsync(S.d_prv[0], psi, sizeof(double) * set, hipMemcpyDeviceToDevice,
                            c->stream));
    QX_CHECK(hipStreamSynchronize(c->stream));
    *iaspc_out = 0;
    return 0;
  }
  int iaspc = 0;
  for (int i = 0; i <= S.nprvmx; ++i)
    if (S.ihest == i || nstep == nstep_ini + 1 + i || S.keypwv == i) {
      iaspc = i + 1;
      break;
    }
  if (iaspc > S.kaspc || iaspc - 1 > S.nprvmx) return -2;
  S.keypwv = std::min(S.keypwv + 1, S.nprvmx);
  int rc = aspc_roteig(c, psi, nb, std::max(iaspc - 1, 1));
  if (rc) return rc;
  if (iaspc > 1) {
    SetPtrs sp{};
    sp.c[0] = S.coef[(size_t)S.kaspc * (iaspc - 1)];
    for (int j = 2; j <= iaspc; ++j) {
      sp.p[j - 2] = S.d_prv[j - 2];
      sp.c[j - 1] = S.coef[(j - 1) + (size_t)S.kaspc * (iaspc - 1)];
    }
    hipLaunchKernelGGL(k_aspc_combine, dim3(cdivu((long)set, 256)), dim3(256), 0, c->stream, psi,
                       sp, iaspc - 1, (long)set, c->d_slot[5]);
    QX_CHECK(hipGetLastError());
  }
  // shift: prv(i+1) <- prv(i), prv(1) <- psi (the last set's buffer is reused)
  double* last = S.d_prv[S.nprvmx - 1];
  for (int i = S.nprvmx - 1; i >= 1; --i) S.d_prv[i] = S.d_prv[i - 1];
  S.d_prv[0] = last;
  QX_CHECK(hipMemcpyAsync(S.d_prv[0], psi, sizeof(double) * set, hipMemcpyDeviceToDevice,
                          c->stream));
  if (iaspc > 1)
    QX_CHECK(hipMemcpyAsync(psi, c->d_slot[5], sizeof(double) * set, hipMemcpyDeviceToDevice,
                            c->stream));
  *iaspc_out = iaspc;
  if (S.ihest != 0) {
    if ((rc = schmidt_dev(c, psi, nb, lchol))) return rc;
  }
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace qx
