// aspc.hip -- wave-function extrapolation between MD steps, device-resident
// (sbalinaspc, Sources/sbalin.f90:369-712, one spin, no ultrasoft part):
//   roteig2 (772-1373): the previous sets are turned to the current bands by
//     the orthogonal factor of their overlap, A = <prev_1|cur>, P_k <- P_k Y
//     with Y = A (A^T A)^-1/2 -- the reference forms Y from the eigenvalues
//     of B = A^T A = V diag(w) V^T as A V diag(w)^-1/2 V^T; here Y comes from
//     a Newton-Schulz iteration (nb^3 GEMMs only) when A is near orthogonal,
//     as it is between MD steps, and from the eigenvalue route otherwise;
//   the always-stable predictor-corrector combination with the coefficients
//     of set_coef_ASPC (77-170), the history shift, and Gram-Schmidt.
// The history lives in HBM (nprvmx x 189 MB at Si216); shifting it rotates
// the set pointers.
#include <algorithm>
#include <cmath>
#include <cstdlib>
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

// Newton-Schulz step prep: err_part[block] = max |M - I| over the block's
// entries, M <- 1.5 I - 0.5 M
__global__ void __launch_bounds__(256) k_ns_prep(double* __restrict__ M, int nb,
                                                 double* __restrict__ err_part) {
  __shared__ double red[4];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (i < (long)nb * nb) {
    const bool diag = (i % nb) == (i / nb);
    const double m = M[i];
    e = fabs(diag ? m - 1.0 : m);
    M[i] = (diag ? 1.5 : 0.0) - 0.5 * m;
  }
  for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_down(e, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) err_part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ void k_ns_err(const double* __restrict__ err_part, int nblk, double* __restrict__ err) {
  double e = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) e = fmax(e, err_part[b]);
  for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_down(e, o, 64));
  if (threadIdx.x == 0) *err = e;
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

// The orthogonal polar factor Y = A (A^T A)^-1/2 of A by Newton-Schulz,
// Y <- Y (1.5 I - 0.5 Y^T Y) from Y = A: GEMMs only, quadratic convergence
// when the singular values of A are near 1 (consecutive MD steps' bands).
// Returns 0 with *Y set after the step taken at max |Y^T Y - I| < 1e-10 (so
// ||Y^T Y - I||_2 < 1e-10 nb before it and at round-off after it), 1 if that
// is not reached in a few steps (the caller then takes the eigenvalue route,
// which also detects a singular A).
static int polar_newton_schulz(Ctx* c, const double* A, int nb, double* Y, double* M, double* T,
                               double* err_part, double* err) {
  const size_t nb2 = (size_t)nb * nb;
  const unsigned nblk = cdivu((long)nb2, 256);
  QX_CHECK(hipMemcpyAsync(Y, A, sizeof(double) * nb2, hipMemcpyDeviceToDevice, c->stream));
  double prev = 1e300;
  for (int it = 0; it < 12; ++it) {
    int rc = gemm_tn(c, nb, nb, nb, 1.0, Y, nb, Y, nb, 0.0, M, nb, 0);  // Y^T Y
    if (rc) return rc;
    hipLaunchKernelGGL(k_ns_prep, dim3(nblk), dim3(256), 0, c->stream, M, nb, err_part);
    hipLaunchKernelGGL(k_ns_err, dim3(1), dim3(64), 0, c->stream, err_part, (int)nblk, err);
    QX_CHECK(hipGetLastError());
    double eh = 0.0;
    QX_CHECK(hipMemcpyAsync(&eh, err, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    QX_CHECK(hipStreamSynchronize(c->stream));
    // far from orthogonal (the iteration may not converge) or not improving
    if (!(eh < 0.5) || !(eh < prev)) return 1;
    if ((rc = gemm_nn(c, nb, nb, nb, 1.0, Y, nb, M, nb, 0.0, T, nb))) return rc;
    QX_CHECK(hipMemcpyAsync(Y, T, sizeof(double) * nb2, hipMemcpyDeviceToDevice, c->stream));
    if (eh < 1e-10) return 0;
    prev = eh;
  }
  return 1;
}

// P_k <- P_k Y for the sets 1..min(nprvmx, kpwvx)
static int aspc_turn_sets(Ctx* c, const double* Y, int nb, int kpwvx) {
  auto& S = c->aspc;
  const int n = c->nplwex;
  for (int k = 0; k < S.nprvmx && k < kpwvx; ++k) {
    double* out = c->d_slot[5];
    int rc = gemm_nn(c, n, nb, nb, 1.0, S.d_prv[k], n, Y, nb, 0.0, out, n);
    if (rc) return rc;
    // the turned set takes the staging buffer's place; the old set becomes staging
    std::swap(S.d_prv[k], c->d_slot[5]);
  }
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
  // Y = A (A^T A)^-1/2: Newton-Schulz when A is near orthogonal, else (and
  // to detect a singular A) through the eigenvalues of A^T A as roteig2
  rc = std::getenv("QX_ROTEIG_EIGH") ? 1 : polar_newton_schulz(c, A, nb, Y, B, Z, T1, T2t);
  if (rc < 0) return rc;
  if (rc == 0) return aspc_turn_sets(c, Y, nb, kpwvx);
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
  return aspc_turn_sets(c, Y, nb, kpwvx);
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
    if ((rc = schmidt_dev(c, psi, nb, lchol, PREC_FP64))) return rc;
  }
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace qx
