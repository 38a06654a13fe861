// forces.hip -- Hellmann-Feynman forces of the KB nonlocal term (kbnlpp_frc2,
// Sources/nlkbpp.f90:1878-2060, lbigmm2 / VECTOR branch), as contractions:
//   P      = slmir = 2 B^T W psi - (G = 0 terms)       (the projection of hckbpp_g)
//   F_x    = B^T (J_x psi),  (J_x psi)(G) = i G_x psi(G) in the real-packed form,
//            i.e. (Re, Im) -> (G_x Im, -G_x Re), a plain sum over the sphere
//   f(ia) += 2 sum_{p of ia, b} F_x(p, b) P(p, b) 2 occ(b) vclalp(p)
// The three F products run on the context's FP64 GEMM path (hipBLASLt for
// large shapes); J is applied by one pass over the bands.
#include <algorithm>
#include <vector>
#include "context.h"

namespace qx {

namespace {

inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

// w_c[b][2p] = g_c[p] psi[b][2p+1], w_c[b][2p+1] = -g_c[p] psi[b][2p], c = x, y, z
__global__ void k_apply_j(const double* __restrict__ psi, const double* __restrict__ g, int np,
                          int nb, double* __restrict__ wx, double* __restrict__ wy,
                          double* __restrict__ wz) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)np * nb) return;
  const int b = (int)(i / np), p = (int)(i % np);
  const size_t o = (size_t)b * 2 * np + 2 * (size_t)p;
  const double re = psi[o], im = psi[o + 1];
  const double gx = g[p], gy = g[np + p], gz = g[2 * np + p];
  wx[o] = gx * im;
  wx[o + 1] = -gx * re;
  wy[o] = gy * im;
  wy[o + 1] = -gy * re;
  wz[o] = gz * im;
  wz[o + 1] = -gz * re;
}

// one workgroup per atom: f(ia, c) += 2 sum over its projectors and the bands
__global__ void __launch_bounds__(256) k_force_reduce(const double* __restrict__ F,
                                                      const double* __restrict__ P,
                                                      const double* __restrict__ D,
                                                      const double* __restrict__ occ,
                                                      const int* __restrict__ p0, int npj, int nb,
                                                      double* __restrict__ f) {
  __shared__ double red[3][4];
  const int ia = blockIdx.x;
  const int pa = p0[ia], pe = p0[ia + 1], cnt = pe - pa;
  const size_t plane = (size_t)npj * nb;
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < cnt * nb; i += blockDim.x) {
    const int b = i / cnt, p = pa + i % cnt;
    const size_t o = (size_t)b * npj + p;
    const double w = P[o] * (D[p] * occ[b] * 2.0);
    s[0] += F[o] * w;
    s[1] += F[plane + o] * w;
    s[2] += F[2 * plane + o] * w;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int c = 0; c < 3; ++c) {
    double v = s[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[c][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int c = threadIdx.x;
    f[3 * ia + c] += 2.0 * (red[c][0] + red[c][1] + red[c][2] + red[c][3]);
  }
}

}  // namespace

int kb_forces(Ctx* c, const double* psi, int nb, const double* occ_host, const double* gx,
              const double* gy, const double* gz, double* fnlc_host) {
  if (c->npj == 0 || c->natom == 0) return 0;
  const int n = c->nplwex, np = c->nplw + 1, npj = c->npj;
  if (nb < 1 || nb > c->nbmax) return -1;
  if (psi == c->d_slot[2] || psi == c->d_slot[3] || psi == c->d_slot[4]) return -1;
  // projections slmir = 2 B^T psi - G0 terms
  int rc = gemm_tn(c, npj, nb, n, 2.0, c->d_B, n, psi, n, 0.0, c->d_P, npj, 2);
  if (rc) return rc;
  // G components in internal order
  std::vector<int> perm(np);
  QX_CHECK(hipMemcpyAsync(perm.data(), c->d_perm, sizeof(int) * np, hipMemcpyDeviceToHost,
                          c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  std::vector<double> g(3 * (size_t)np);
  for (int p = 0; p < np; ++p) {
    g[p] = gx[perm[p]];
    g[np + p] = gy[perm[p]];
    g[2 * np + p] = gz[perm[p]];
  }
  // projector range of each atom (projectors are atom-major)
  std::vector<int> p0(c->natom + 1, 0);
  for (int p = 0; p < npj; ++p) p0[c->proj_atom[p] + 1]++;
  for (int ia = 0; ia < c->natom; ++ia) p0[ia + 1] += p0[ia];
  double *d_g = nullptr, *d_F = nullptr, *d_occ = nullptr, *d_f = nullptr;
  int* d_p0 = nullptr;
  QX_CHECK(hipMalloc(&d_g, sizeof(double) * g.size()));
  QX_CHECK(hipMalloc(&d_F, sizeof(double) * 3 * (size_t)npj * nb));
  QX_CHECK(hipMalloc(&d_occ, sizeof(double) * nb));
  QX_CHECK(hipMalloc(&d_f, sizeof(double) * 3 * (size_t)c->natom));
  QX_CHECK(hipMalloc(&d_p0, sizeof(int) * p0.size()));
  QX_CHECK(hipMemcpyAsync(d_g, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice,
                          c->stream));
  QX_CHECK(hipMemcpyAsync(d_occ, occ_host, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipMemcpyAsync(d_f, fnlc_host, sizeof(double) * 3 * c->natom, hipMemcpyHostToDevice,
                          c->stream));
  QX_CHECK(hipMemcpyAsync(d_p0, p0.data(), sizeof(int) * p0.size(), hipMemcpyHostToDevice,
                          c->stream));
  double* w[3] = {c->d_slot[2], c->d_slot[3], c->d_slot[4]};
  hipLaunchKernelGGL(k_apply_j, dim3(cdivu((long)np * nb, 256)), dim3(256), 0, c->stream, psi,
                     d_g, np, nb, w[0], w[1], w[2]);
  for (int k = 0; k < 3 && !rc; ++k)
    rc = gemm_tn(c, npj, nb, n, 1.0, c->d_B, n, w[k], n, 0.0, d_F + (size_t)k * npj * nb, npj, 0);
  if (!rc) {
    hipLaunchKernelGGL(k_force_reduce, dim3(c->natom), dim3(256), 0, c->stream, d_F, c->d_P,
                       c->d_D, d_occ, d_p0, npj, nb, d_f);
    QX_CHECK(hipGetLastError());
    QX_CHECK(hipMemcpyAsync(fnlc_host, d_f, sizeof(double) * 3 * c->natom, hipMemcpyDeviceToHost,
                            c->stream));
    QX_CHECK(hipStreamSynchronize(c->stream));
  }
  (void)hipFree(d_g);
  (void)hipFree(d_F);
  (void)hipFree(d_occ);
  (void)hipFree(d_f);
  (void)hipFree(d_p0);
  return rc;
}

}  // namespace qx
