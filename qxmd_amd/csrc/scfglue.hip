// scfglue.hip -- the SCF glue of the Gamma-point loop on the device, between
// two band updates (Sources/pwlda.f90:1640-2030, Sources/chgdns.f90:29-243,
// 796-1210):
//   chgr2g  (Sources/engrad.f90:1253)  rho(r) -> rhgr on the density sphere
//   mxchgg  (Sources/chgdns.f90:796)   Pulay / simple / no mixing of rhgr
//   chgg2r  (Sources/engrad.f90:1302)  mixed rhgr -> rho(r)
//   harting (Sources/engrad.f90:1673)  vhar(r) from 8 pi / G^2 rhgr
//   xclda   (Sources/vxc.f90:961)      Perdew-Zunger LDA vexc and energies
//   pwpunifylc2 (Sources/commwv.f90:1290) V = vhar + vext + vexc
// The density sphere spans the whole grid, so the transforms are full-grid
// complex FFTs (fullgrid.hip).  chgg2r and harting share one inverse
// transform: both spectra are Hermitian, so z(G) = rho(G) + i 2 recnrmex(G)
// rho(G) transforms to rho(r) + i vhar(r).  The per-point XC, the potential
// assembly and the six energy sums are one fused pass; sums go through
// per-block partials and a fixed-order final sum, so they are deterministic.
// The Pulay history lives on the device; its small normal-equation solve
// (INVERS, Sources/funcs.f90:5) runs on the host between the dot products
// and the combination, as the reference's node 0 does.
#include <algorithm>
#include <cmath>
#include <vector>
#include "context.h"

namespace qx {

namespace {

constexpr int kRedBlocks = 512;   // fixed reduction grid: deterministic sum order
constexpr int kRedThreads = 256;
constexpr int kNe = 6;            // ehar, elcl, eexc, evexc, ecor, evcor
constexpr int kPlMax = 32;        // largest Pulay history (itratn) on the device

inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

// VXC (Sources/vxc.f90:2063-2104): Ceperley-Alder in the Perdew-Zunger fit,
// Hartree units.  B2, B3 and C2..C4 are single-precision literals in the
// reference's DATA statement; their float values are used here as there.
__device__ inline void pz_vxc(double rs, double& ex, double& vx, double& ec, double& vc) {
  const double A1 = -0.4582, B1 = -0.1423, B2 = (double)1.0529f, B3 = (double)0.3334f;
  const double C1 = -0.0480, C2 = (double)0.0311f, C3 = (double)-0.0116f, C4 = (double)0.0020f;
  ex = A1 / rs;
  vx = 4.0 * ex / 3.0;
  if (rs >= 1.0) {
    const double sr = sqrt(rs);
    const double bunb = 1.0 + B2 * sr + B3 * rs;
    ec = B1 / bunb;
    vc = ec * (1.0 + (B2 * sr / 6.0 + B3 * rs / 3.0) / bunb);
  } else {
    const double lr = log(rs);
    ec = C1 + C2 * lr + C3 * rs + C4 * rs * lr;
    vc = ec - (C2 + (C3 + C4) * rs + C4 * rs * lr) / 3.0;
  }
}

template <int K>
__device__ inline void block_sum(double (&s)[K], double* out) {
  __shared__ double red[K][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double v = s[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[k][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double v = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) v += red[threadIdx.x][w];
    out[threadIdx.x] = v;
  }
}

__global__ void k_sg_zero(double2* __restrict__ z, long N) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) z[i] = make_double2(0.0, 0.0);
}

__global__ void k_sg_real(const double* __restrict__ r, double2* __restrict__ z, long N) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) z[i] = make_double2(r[i], 0.0);
}

__global__ void k_sg_re(const double2* __restrict__ z, double* __restrict__ re,
                        double* __restrict__ im, long N) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double2 v = z[i];
  if (re) re[i] = v.x;
  if (im) im[i] = v.y;
}

// chgr2g gather: rhgr(2 ig, 2 ig + 1) = z at ijkgd(ig), ig = 0..nplw5
__global__ void k_sg_gather(const double2* __restrict__ z, const int* __restrict__ ijk0,
                            int nplw5, double* __restrict__ rhgr) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig > nplw5) return;
  const double2 v = z[ijk0[ig]];
  rhgr[2 * ig] = v.x;
  rhgr[2 * ig + 1] = v.y;
}

// chgg2r + harting scatter: with rho(G) = a + i b and q = 2 recnrmex(G) (0
// when p4 is null), z(+G) = rho(G) + i q rho(G), z(-G) = conj rho(G) + i q
// conj rho(G).  G = 0 and points that are their own mirror on the grid are
// real (the Hermitian part, which is all the c2r of rfft3 keeps of them).
__global__ void k_sg_scatter(const double* __restrict__ rhgr, const double* __restrict__ p4,
                             const int* __restrict__ ijk0, int nplw5, double2* __restrict__ z) {
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig > nplw5) return;
  const double a = rhgr[2 * ig], b = rhgr[2 * ig + 1];
  const double q = p4 ? p4[ig] : 0.0;
  const int ip = ijk0[ig];
  const int im = ijk0[-ig];
  if (ig == 0 || ip == im) {
    z[ip] = make_double2(a, q * a);
    return;
  }
  z[ip] = make_double2(a - q * b, b + q * a);
  z[im] = make_double2(a + q * b, q * a - b);
}

// xclda + pwpunifylc2 + the Hartree and external energies, one pass over the
// grid.  rho: density; z (optional): Im = vhar; vext (optional); core
// (optional, negative entries = no core); outputs optional except vexc,
// eeexc.  part[blk * 8 + k]: sums of vhar rho, vext rho, ex rhot, vx rho,
// ec rhot, vc rho.
__global__ void __launch_bounds__(kRedThreads)
    k_sg_potential(const double* __restrict__ rho, const double2* __restrict__ z,
                   const double* __restrict__ vext, const double* __restrict__ core, long N,
                   double* __restrict__ V, float* __restrict__ Vf, double* __restrict__ vhar,
                   double* __restrict__ vexc, double* __restrict__ eeexc,
                   double* __restrict__ part) {
  const double pi4 = 12.56637061436;  // the reference's constant, not 4 pi
  double s[kNe] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (long)gridDim.x * blockDim.x) {
    const double r = rho[i];
    const double c = core ? core[i] : -1.0;
    const double rhot = c >= 0.0 ? r + c : r;
    double vxc = 0.0, exc = 0.0;
    if (!(rhot < 1e-30)) {
      const double rs = pow(3.0 / (pi4 * rhot), 1.0 / 3.0);
      double ex, vx, ec, vc;
      pz_vxc(rs, ex, vx, ec, vc);
      vxc = 2.0 * (vx + vc);
      exc = 2.0 * (ex + ec);
      s[2] += ex * rhot;
      s[3] += vx * r;
      s[4] += ec * rhot;
      s[5] += vc * r;
    }
    vexc[i] = vxc;
    eeexc[i] = exc;
    const double vh = z ? z[i].y : 0.0;
    const double ve = vext ? vext[i] : 0.0;
    s[0] += vh * r;
    s[1] += ve * r;
    if (vhar) vhar[i] = vh;
    if (V) {
      // hdiag = vhar + vhshdp + vext + vexc, vhshdp = 0 for a periodic cell
      const double v = (vh + ve) + vxc;
      V[i] = v;
      Vf[i] = (float)v;
    }
  }
  block_sum<kNe>(s, part + (size_t)blockIdx.x * 8);
}

// out[k] = scale(k) * sum_b part[b * stride + k] for k < K, one block of
// kRedThreads: thread t takes b = t, t + kRedThreads, ... then a fixed wave
// and block tree, so the order of the sum is the same on every run.
// mode 0: scale 1 x s; mode 1: the energies (ehar = s0 rdelv / 2, elcl =
// s1 rdelv (Sources/pwlda.f90:1868-1873, 1963-1968), XC sums x 2 rdelv
// (Sources/vxc.f90:1009-1012)).
__global__ void __launch_bounds__(kRedThreads)
    k_red_final(const double* __restrict__ part, int nblk, int stride, int K, int mode, double s,
                double* __restrict__ out) {
  __shared__ double red[kRedThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < K; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kRedThreads) v += part[(size_t)b * stride + k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kRedThreads / 64; ++w) t += red[w];
      if (mode == 1) t = k == 0 ? 0.5 * t * s : (k == 1 ? t * s : t * 2.0 * s);
      else t = t * s;
      out[k] = t;
    }
    __syncthreads();
  }
}

struct Cols {
  int c[kPlMax];
};
struct Coefs {
  double c[kPlMax];
};

// mxchgg_pulay: unr(:, k) = rinr, dfmr(:, k) = rhgr - rinr
__global__ void k_pl_store(const double* __restrict__ rhgr, const double* __restrict__ rinr,
                           int n, double* __restrict__ u, double* __restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u[i] = rinr[i];
  d[i] = rhgr[i] - rinr[i];
}

// partial metric products of the new residual with the k logical history
// columns: part[blk * kPlMax + j] = sum_ig dfmr(ig, col j) dfmr(ig, new) wq(ig)
__global__ void __launch_bounds__(kRedThreads)
    k_pl_dots(const double* __restrict__ dfmr, int ld, Cols cols, int k, int knew,
              const double* __restrict__ wq, int n, double* __restrict__ part) {
  double s[kPlMax];
#pragma unroll
  for (int j = 0; j < kPlMax; ++j) s[j] = 0.0;
  const double* dn = dfmr + (size_t)knew * ld;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double x = dn[i] * wq[i];
#pragma unroll
    for (int j = 0; j < kPlMax; ++j)
      if (j < k) s[j] += dfmr[(size_t)cols.c[j] * ld + i] * x;
  }
  block_sum<kPlMax>(s, part + (size_t)blockIdx.x * kPlMax);
}

__device__ inline double kerker(double g2, double aslh, double bslh, double wk) {
  const double f = g2 / (g2 + bslh);
  return aslh * (1.0 - wk + wk * f);
}

// first step of a history: rhgr = rinr + f(G) dfmr(:, k)
__global__ void k_pl_kerker(const double* __restrict__ rinr, const double* __restrict__ d,
                            const double* __restrict__ dkg, double aslh, double bslh, double wk,
                            int n, double* __restrict__ rhgr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rhgr[i] = rinr[i] + kerker(dkg[i], aslh, bslh, wk) * d[i];
}

// rhgr = sum_ii cmk(ii) (unr(:, ii) + f(G) dfmr(:, ii)), summed in ii order
__global__ void k_pl_combine(const double* __restrict__ unr, const double* __restrict__ dfmr,
                             int ld, Cols cols, Coefs cmk, int k, const double* __restrict__ dkg,
                             double aslh, double bslh, double wk, int n,
                             double* __restrict__ rhgr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = kerker(dkg[i], aslh, bslh, wk);
  double acc = 0.0;
  for (int j = 0; j < k; ++j) {
    const size_t o = (size_t)cols.c[j] * ld + i;
    acc = acc + cmk.c[j] * (unr[o] + f * dfmr[o]);
  }
  rhgr[i] = acc;
}

// mxchgg_simple: rhgr = (1 - ph) rinr + ph rhgr
__global__ void k_mx_simple(double* __restrict__ rhgr, const double* __restrict__ rinr, int n,
                            double ph) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rhgr[i] = (1.0 - ph) * rinr[i] + ph * rhgr[i];
}

// DETMAT (Sources/funcs.f90:66-128): determinant by row scaling and Gaussian
// elimination over a row permutation; w (n x n, leading dimension ld) is
// overwritten.  The inverse below uses the same cofactor scheme as INVERS so
// that the Pulay coefficients follow the reference's rounding.
double detmat(std::vector<double>& w, int n, int ld, std::vector<int>& ip) {
  const double dzero = 1e-30;
  double fugo = 1.0, det = 1.0;
  auto W = [&](int r, int c) -> double& { return w[(size_t)r + (size_t)ld * c]; };
  for (int i = 0; i < n; ++i) ip[i] = i;
  for (int i = 0; i < n; ++i) {
    double pivot = 0.0;
    for (int j = 0; j < n; ++j) pivot = std::max(pivot, std::fabs(W(i, j)));
    if (pivot < dzero) return 0.0;
    for (int j = 0; j < n; ++j) W(i, j) = W(i, j) / pivot;
    det = det * pivot;
  }
  for (int i = 0; i < n - 1; ++i) {
    double pivot = 0.0;
    int jpv = i;
    for (int j = i; j < n; ++j) {
      const double a = std::fabs(W(ip[j], i));
      if (a > pivot) {
        pivot = a;
        jpv = j;
      }
    }
    if (pivot < dzero) return 0.0;
    if (jpv != i) {
      std::swap(ip[jpv], ip[i]);
      fugo = -fugo;
    }
    const int ipv = ip[i];
    for (int j = i + 1; j < n; ++j) {
      const int jj = ip[j];
      const double dmj = W(jj, i) / W(ipv, i);
      W(jj, i) = 0.0;
      for (int j2 = i + 1; j2 < n; ++j2) W(jj, j2) = W(jj, j2) - dmj * W(ipv, j2);
    }
  }
  for (int i = 0; i < n; ++i) det = det * W(ip[i], i);
  return det * fugo;
}

// INVERS (Sources/funcs.f90:5-61): wi = w^-1 by cofactors of w / max|w|;
// returns 1 when the scaled determinant is below 1e-30.
int invers(const std::vector<double>& w, std::vector<double>& wi, int n, int ld) {
  std::vector<double> work((size_t)ld * ld);
  std::vector<int> ip(ld);
  double elmax = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) elmax = std::max(elmax, std::fabs(w[i + (size_t)ld * j]));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) work[i + (size_t)ld * j] = w[i + (size_t)ld * j] / elmax;
  const double det = detmat(work, n, ld, ip);
  if (std::fabs(det) < 1e-30) return 1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      int isum = 0;
      for (int ii = 0; ii < n; ++ii) {
        if (ii == j) continue;
        int jsum = 0;
        for (int jj = 0; jj < n; ++jj) {
          if (jj == i) continue;
          work[isum + (size_t)ld * jsum] = w[ii + (size_t)ld * jj] / elmax;
          ++jsum;
        }
        ++isum;
      }
      const double detij = detmat(work, n - 1, ld, ip);
      wi[i + (size_t)ld * j] = (((i + j) % 2) ? -1.0 : 1.0) * detij / det;
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) wi[i + (size_t)ld * j] = wi[i + (size_t)ld * j] / elmax;
  return 0;
}

template <typename T>
int dev_alloc(T** p, size_t n) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  QX_CHECK(hipMalloc(p, sizeof(T) * std::max<size_t>(n, 1)));
  return 0;
}

// 0-based grid offsets of ijkgd(-nplw5:nplw5) (1-based in the reference)
std::vector<int> zero_based(const int* ijkgd, int nplw5) {
  std::vector<int> v(2 * (size_t)nplw5 + 1);
  for (size_t i = 0; i < v.size(); ++i) v[i] = ijkgd[i] - 1;
  return v;
}

int check_ijk(const int* ijkgd, int nplw5, long N) {
  for (long i = 0; i < 2L * nplw5 + 1; ++i)
    if (ijkgd[i] < 1 || ijkgd[i] > N) return -2;
  return 0;
}

// the sphere of the context: index map and the charge-density vectors
int scf_sphere(Ctx* c, int nplw5, const int* ijkgd) {
  auto& S = c->scf;
  if (nplw5 < 0) return -2;
  if (check_ijk(ijkgd, nplw5, c->N)) return -2;
  int rc = fg_prepare(c->fg, c->n1, c->n2, c->n3, false);
  if (rc) return rc;
  if (S.nplw5 != nplw5 || !S.d_ijk) {
    if ((rc = dev_alloc(&S.d_ijk, 2 * (size_t)nplw5 + 1))) return rc;
    if ((rc = dev_alloc(&S.d_rhgr, 2 * ((size_t)nplw5 + 1)))) return rc;
    if ((rc = dev_alloc(&S.d_rinr, 2 * ((size_t)nplw5 + 1)))) return rc;
    QX_CHECK(hipMemsetAsync(S.d_rhgr, 0, sizeof(double) * 2 * (nplw5 + 1), c->stream));
    QX_CHECK(hipMemsetAsync(S.d_rinr, 0, sizeof(double) * 2 * (nplw5 + 1), c->stream));
    if (S.d_p4) {
      (void)hipFree(S.d_p4);
      S.d_p4 = nullptr;
    }
    S.nplw5 = nplw5;
  }
  const std::vector<int> ij = zero_based(ijkgd, nplw5);
  QX_CHECK(hipMemcpyAsync(S.d_ijk, ij.data(), sizeof(int) * ij.size(), hipMemcpyHostToDevice,
                          c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (!S.d_red) {
    if ((rc = dev_alloc(&S.d_red, (size_t)kRedBlocks * kPlMax + 64))) return rc;
  }
  if (!S.d_vhar) {
    if ((rc = dev_alloc(&S.d_vhar, c->N))) return rc;
    if ((rc = dev_alloc(&S.d_vexc, c->N))) return rc;
    if ((rc = dev_alloc(&S.d_eeexc, c->N))) return rc;
  }
  S.z_vhar = false;
  return 0;
}

// rho (device, FFT order) -> rhgr (device) through the context's grid
int chgr2g_dev(Ctx* c, const double* rho, double* rhgr) {
  auto& S = c->scf;
  const long N = c->N;
  hipLaunchKernelGGL(k_sg_real, dim3(cdivu(N, 256)), dim3(256), 0, c->stream, rho, c->fg.d_g, N);
  int rc = fg_transform(c->fg, c->stream, -1, 1.0 / (double)N);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sg_gather, dim3(cdivu(S.nplw5 + 1, 256)), dim3(256), 0, c->stream,
                     c->fg.d_g, S.d_ijk + S.nplw5, S.nplw5, rhgr);
  QX_CHECK(hipGetLastError());
  S.z_vhar = false;
  return 0;
}

// rhgr (device) -> fg.d_g = rho(r) + i vhar(r) (vhar only if the Hartree kernel is set)
int g2r_dev(Ctx* c, const double* rhgr) {
  auto& S = c->scf;
  const long N = c->N;
  hipLaunchKernelGGL(k_sg_zero, dim3(cdivu(N, 256)), dim3(256), 0, c->stream, c->fg.d_g, N);
  hipLaunchKernelGGL(k_sg_scatter, dim3(cdivu(S.nplw5 + 1, 256)), dim3(256), 0, c->stream, rhgr,
                     S.d_p4, S.d_ijk + S.nplw5, S.nplw5, c->fg.d_g);
  QX_CHECK(hipGetLastError());
  int rc = fg_transform(c->fg, c->stream, +1, 1.0);
  if (rc) return rc;
  S.z_vhar = S.d_p4 != nullptr;
  return 0;
}

__global__ void k_sg_vext(const double2* __restrict__ z, double* __restrict__ vext, long N,
                          int accumulate) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  vext[i] = accumulate ? vext[i] + z[i].x : z[i].x;
}

// mixing part of mxchgg (Sources/chgdns.f90:796-1210) on d_rhgr / d_rinr
int mix_dev(Ctx* c, int lreset, double aslh, double bslh, int lmixhold) {
  auto& S = c->scf;
  const int n5x = 2 * (S.nplw5 + 1);
  if (S.imxchg == 1) {
    const int n = S.nplwcex, itr = S.itratn;
    if (n > n5x) return -2;
    const double small = 1e-15;
    int kntchg = S.kntchg + 1;
    if (kntchg > itr || lreset) kntchg = 1;
    S.nreset = 0;
    std::vector<double> a(kPlMax), betakl((size_t)itr * itr);
    auto A = [&](int r, int col) -> double& { return S.aij[(size_t)r + (size_t)itr * col]; };
    auto B = [&](int r, int col) -> double& { return betakl[(size_t)r + (size_t)itr * col]; };
    for (;;) {
      const int k = kntchg - 1;
      const int pk = S.col[k];
      hipLaunchKernelGGL(k_pl_store, dim3(cdivu(n, 256)), dim3(256), 0, c->stream, S.d_rhgr,
                         S.d_rinr, n, S.d_unr + (size_t)pk * n, S.d_dfmr + (size_t)pk * n);
      Cols cols{};
      for (int j = 0; j < kntchg; ++j) cols.c[j] = S.col[j];
      const int nblk = std::min<int>(kRedBlocks, (int)cdivu(n, kRedThreads));
      hipLaunchKernelGGL(k_pl_dots, dim3(nblk), dim3(kRedThreads), 0, c->stream, S.d_dfmr, n,
                         cols, kntchg, pk, S.d_wq, n, S.d_red);
      double* d_out = S.d_red + (size_t)kRedBlocks * kPlMax;
      hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kRedThreads), 0, c->stream, S.d_red, nblk,
                         kPlMax, kntchg, 0, S.rvol, d_out);
      QX_CHECK(hipGetLastError());
      QX_CHECK(hipMemcpyAsync(a.data(), d_out, sizeof(double) * kntchg, hipMemcpyDeviceToHost,
                              c->stream));
      QX_CHECK(hipStreamSynchronize(c->stream));
      for (int i1 = 0; i1 < kntchg; ++i1) {
        A(i1, k) = a[i1];
        A(k, i1) = a[i1];
      }
      int ierr = 0;
      if (kntchg == 1) {
        hipLaunchKernelGGL(k_pl_kerker, dim3(cdivu(n, 256)), dim3(256), 0, c->stream, S.d_rinr,
                           S.d_dfmr + (size_t)pk * n, S.d_dkg, aslh, bslh, S.wkerker, n, S.d_rhgr);
        std::fill(S.cmk.begin(), S.cmk.end(), 0.0);
      } else {
        const int m = kntchg, icntr = m / 2 + m % 2;
        const double amax = std::fabs(A(icntr - 1, icntr - 1));
        if (amax < small) {
          kntchg = 1;
          ++S.nreset;
          continue;
        }
        for (int j = 0; j < m; ++j)
          for (int i = 0; i < m; ++i) A(i, j) = A(i, j) / amax;
        ierr = invers(S.aij, betakl, m, itr);
        if (ierr == 0) {
          for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) A(i, j) = A(i, j) * amax;
          for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) B(i, j) = B(i, j) / amax;
          double bbnb = 0.0;
          Coefs cm{};
          for (int ii = 0; ii < m; ++ii) {
            double s = 0.0;
            for (int jj = 0; jj < m; ++jj) {
              bbnb = bbnb + B(ii, jj);
              s = s + B(ii, jj);
            }
            cm.c[ii] = s;
          }
          std::fill(S.cmk.begin(), S.cmk.end(), 0.0);
          for (int ii = 0; ii < m; ++ii) {
            cm.c[ii] = cm.c[ii] / bbnb;
            S.cmk[ii] = cm.c[ii];
          }
          hipLaunchKernelGGL(k_pl_combine, dim3(cdivu(n, 256)), dim3(256), 0, c->stream, S.d_unr,
                             S.d_dfmr, n, cols, cm, m, S.d_dkg, aslh, bslh, S.wkerker, n,
                             S.d_rhgr);
        } else {
          // drop the oldest entry (the reference copies the columns down;
          // here the logical -> physical map rotates)
          ++S.nreset;
          kntchg = std::max(1, kntchg - 1);
          std::rotate(S.col.begin(), S.col.begin() + 1, S.col.end());
          for (int j = 0; j < kntchg - 1; ++j)
            for (int i = 0; i < kntchg - 1; ++i) A(i, j) = A(i + 1, j + 1) * amax;
        }
      }
      if (ierr == 0) break;
    }
    if (kntchg == itr) {
      std::rotate(S.col.begin(), S.col.begin() + 1, S.col.end());
      for (int j = 0; j < kntchg - 1; ++j)
        for (int i = 0; i < kntchg - 1; ++i) A(i, j) = A(i + 1, j + 1);
      if (lmixhold) kntchg = kntchg - 1;
    }
    S.kntchg = kntchg;
  } else if (S.imxchg == 3) {
    hipLaunchKernelGGL(k_mx_simple, dim3(cdivu(n5x, 256)), dim3(256), 0, c->stream, S.d_rhgr,
                       S.d_rinr, n5x, aslh);
  } else if (S.imxchg != 0) {
    return -2;
  }
  QX_CHECK(hipGetLastError());
  QX_CHECK(hipMemcpyAsync(S.d_rinr, S.d_rhgr, sizeof(double) * n5x, hipMemcpyDeviceToDevice,
                          c->stream));
  return 0;
}

int potential_dev(Ctx* c, double* ene) {
  auto& S = c->scf;
  if (!S.d_ijk || !S.d_p4 || !S.d_vext) return -2;
  int rc;
  if (!S.z_vhar && (rc = g2r_dev(c, S.d_rhgr))) return rc;
  hipLaunchKernelGGL(k_sg_potential, dim3(kRedBlocks), dim3(kRedThreads), 0, c->stream, c->d_rho,
                     c->fg.d_g, S.d_vext, S.d_core, c->N, c->d_V, c->d_Vf, S.d_vhar, S.d_vexc,
                     S.d_eeexc, S.d_red);
  double* d_out = S.d_red + (size_t)kRedBlocks * kPlMax;
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kRedThreads), 0, c->stream, S.d_red, kRedBlocks, 8,
                     kNe, 1, S.rdelv, d_out);
  QX_CHECK(hipGetLastError());
  if ((rc = fast_permute_v(c))) return rc;
  double e[kNe];
  QX_CHECK(hipMemcpyAsync(e, d_out, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (ene)
    for (int k = 0; k < kNe; ++k) ene[k] = e[k];
  return 0;
}

}  // namespace

// vlocal_g's grid step (Sources/nlkbpp.f90:974-1027): zin scattered to +-G
// (G = 0 real), the unnormalised inverse transform, vext (+)= its real part
int vext_from_sphere(Ctx* c, const double* d_zin, int accumulate) {
  auto& S = c->scf;
  if (!S.d_ijk || !S.d_vext) return -4;
  const long N = c->N;
  hipLaunchKernelGGL(k_sg_zero, dim3(cdivu(N, 256)), dim3(256), 0, c->stream, c->fg.d_g, N);
  hipLaunchKernelGGL(k_sg_scatter, dim3(cdivu(S.nplw5 + 1, 256)), dim3(256), 0, c->stream, d_zin,
                     nullptr, S.d_ijk + S.nplw5, S.nplw5, c->fg.d_g);
  QX_CHECK(hipGetLastError());
  int rc = fg_transform(c->fg, c->stream, +1, 1.0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sg_vext, dim3(cdivu(N, 256)), dim3(256), 0, c->stream, c->fg.d_g, S.d_vext,
                     N, accumulate);
  QX_CHECK(hipGetLastError());
  S.z_vhar = false;  // fg.d_g no longer holds rho + i vhar
  return 0;
}


void scf_destroy(Ctx* c) {
  auto& S = c->scf;
  for (double* p : {S.d_p4, S.d_vext, S.d_core, S.d_rhgr, S.d_rinr, S.d_vhar, S.d_vexc, S.d_eeexc,
                    S.d_red, S.d_unr, S.d_dfmr, S.d_dkg, S.d_wq, S.d_cmk})
    if (p) (void)hipFree(p);
  if (S.d_ijk) (void)hipFree(S.d_ijk);
  S = Ctx::Scf{};
  fg_release(c->fg);
}

int scf_set_grid(Ctx* c, int nplw5, const int* ijkgd, const double* recnrmex, const double* vext,
                 const double* rhocore, double rdelv) {
  auto& S = c->scf;
  int rc = scf_sphere(c, nplw5, ijkgd);
  if (rc) return rc;
  std::vector<double> p4(nplw5 + 1);
  for (int ig = 0; ig <= nplw5; ++ig) p4[ig] = 2.0 * recnrmex[ig];
  if ((rc = dev_alloc(&S.d_p4, p4.size()))) return rc;
  QX_CHECK(hipMemcpyAsync(S.d_p4, p4.data(), sizeof(double) * p4.size(), hipMemcpyHostToDevice,
                          c->stream));
  if (!S.d_vext && (rc = dev_alloc(&S.d_vext, c->N))) return rc;
  if (vext)
    QX_CHECK(hipMemcpyAsync(S.d_vext, vext, sizeof(double) * c->N, hipMemcpyHostToDevice,
                            c->stream));
  else
    QX_CHECK(hipMemsetAsync(S.d_vext, 0, sizeof(double) * c->N, c->stream));
  if (rhocore) {
    if (!S.d_core && (rc = dev_alloc(&S.d_core, c->N))) return rc;
    QX_CHECK(hipMemcpyAsync(S.d_core, rhocore, sizeof(double) * c->N, hipMemcpyHostToDevice,
                            c->stream));
  } else if (S.d_core) {
    (void)hipFree(S.d_core);
    S.d_core = nullptr;
  }
  S.rdelv = rdelv;
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int scf_mixer_alloc(Ctx* c, int nplwcex, int imxchg, int itratn, double cmetric, double wkerker,
                    double g2min, double g2max, const double* dkgnrm, double rvol) {
  auto& S = c->scf;
  if (imxchg != 0 && imxchg != 1 && imxchg != 3) return -2;
  if (imxchg == 1 && (itratn < 1 || itratn > kPlMax || nplwcex < 3)) return -2;
  S.imxchg = imxchg;
  S.itratn = itratn;
  S.nplwcex = nplwcex;
  S.kntchg = 0;
  S.rvol = rvol;
  S.wkerker = wkerker;
  // factq1 (Sources/chgdns.f90:661-667)
  const double bbb = g2max - cmetric * g2min;
  S.factq1 = bbb > 1.0e-5 ? (cmetric - 1.0) * g2max * g2min / bbb : 0.0;
  if (imxchg != 1) return 0;
  int rc;
  if ((rc = dev_alloc(&S.d_unr, (size_t)nplwcex * itratn))) return rc;
  if ((rc = dev_alloc(&S.d_dfmr, (size_t)nplwcex * itratn))) return rc;
  if ((rc = dev_alloc(&S.d_dkg, nplwcex))) return rc;
  if ((rc = dev_alloc(&S.d_wq, nplwcex))) return rc;
  // metric weights of the real-packed vector: 2 factq for G != 0 (the
  // reference doubles the sum over ig >= 3), factq of index 3 for Re G0,
  // nothing for Im G0 (Sources/chgdns.f90:1017-1027)
  std::vector<double> wq(nplwcex);
  wq[0] = (dkgnrm[2] + S.factq1) / dkgnrm[2];
  wq[1] = 0.0;
  for (int ig = 2; ig < nplwcex; ++ig) wq[ig] = 2.0 * ((dkgnrm[ig] + S.factq1) / dkgnrm[ig]);
  QX_CHECK(hipMemcpy(S.d_dkg, dkgnrm, sizeof(double) * nplwcex, hipMemcpyHostToDevice));
  QX_CHECK(hipMemcpy(S.d_wq, wq.data(), sizeof(double) * nplwcex, hipMemcpyHostToDevice));
  S.col.resize(itratn);
  for (int j = 0; j < itratn; ++j) S.col[j] = j;
  S.aij.assign((size_t)itratn * itratn, 0.0);
  S.cmk.assign(itratn, 0.0);
  return 0;
}

// softchg + outchg (Sources/chgdns.f90:29-383): density of a slot, then rhgr
int scf_pwpchg(Ctx* c, const double* psi, int nb, const double* occ, double rvol, int prec) {
  auto& S = c->scf;
  if (!S.d_ijk) return -2;
  double* w = c->d_scal + 41 * (size_t)c->nbmax;
  std::vector<double> wh(nb + 1, 0.0);
  for (int b = 0; b < nb; ++b) wh[b] = occ[b] / rvol;
  QX_CHECK(hipMemcpyAsync(w, wh.data(), sizeof(double) * (nb + 1), hipMemcpyHostToDevice,
                          c->stream));
  int rc = density(c, psi, nb, w, 1.0, prec);
  if (rc) return rc;
  return chgr2g_dev(c, c->d_rho, S.d_rhgr);
}

// outchg (Sources/chgdns.f90:248-383): rhgr of the density already in the
// context (the one the last band step or density call left there)
int scf_outchg(Ctx* c) {
  auto& S = c->scf;
  if (!S.d_ijk) return -2;
  return chgr2g_dev(c, c->d_rho, S.d_rhgr);
}

// pwpmxchgg2 (Sources/chgdns.f90:724-790): mix, then rho = chgg2r(rhgr)
int scf_mix(Ctx* c, int lreset, double aslh, double bslh, int lmixhold, int* info) {
  auto& S = c->scf;
  if (!S.d_ijk || S.imxchg < 0) return -2;
  int rc = mix_dev(c, lreset, aslh, bslh, lmixhold);
  if (rc) return rc;
  if ((rc = g2r_dev(c, S.d_rhgr))) return rc;
  hipLaunchKernelGGL(k_sg_re, dim3(cdivu(c->N, 256)), dim3(256), 0, c->stream, c->fg.d_g,
                     c->d_rho, nullptr, c->N);
  QX_CHECK(hipGetLastError());
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (info) {
    info[0] = S.kntchg;
    info[1] = S.nreset;
  }
  return 0;
}

int scf_potential(Ctx* c, double* ene) { return potential_dev(c, ene); }

// 0 rho, 1 rhgr, 2 rinr, 3 vhar, 4 vexc, 5 eeexc, 6 V (glocal), 7 cmk
int scf_io(Ctx* c, int what, double* host, bool put) {
  auto& S = c->scf;
  double* p = nullptr;
  size_t n = (size_t)c->N;
  switch (what) {
    case 0: p = c->d_rho; break;
    case 1: p = S.d_rhgr; n = 2 * ((size_t)S.nplw5 + 1); break;
    case 2: p = S.d_rinr; n = 2 * ((size_t)S.nplw5 + 1); break;
    case 3: p = S.d_vhar; break;
    case 4: p = S.d_vexc; break;
    case 5: p = S.d_eeexc; break;
    case 6: p = c->d_V; break;
    case 7:
      if (put) return -2;
      for (size_t i = 0; i < S.cmk.size(); ++i) host[i] = S.cmk[i];
      return 0;
    default: return -2;
  }
  if (!p) return -2;
  if (put && what > 2) return -2;
  if (put) {
    QX_CHECK(hipMemcpyAsync(p, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if (what == 1) S.z_vhar = false;
  } else {
    QX_CHECK(hipMemcpyAsync(host, p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  }
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// host-array drop-ins (no context): the drop-in grid and its stream

namespace {

// FD order <-> FFT order (mfd2ft 1-based, null = identity); returns -2 on a bad map
int fd_check(const int* mfd2ft, int ntotfd, long N) {
  if (!mfd2ft) return ntotfd == N ? 0 : -2;
  for (int i = 0; i < ntotfd; ++i)
    if (mfd2ft[i] < 1 || mfd2ft[i] > N) return -2;
  return 0;
}

struct DropinBufs {
  double *d_a = nullptr, *d_b = nullptr, *d_c = nullptr, *d_d = nullptr, *d_red = nullptr;
  int* d_ijk = nullptr;
  size_t na = 0, nijk = 0;
};
DropinBufs g_db;

int dropin_bufs(size_t n, size_t nijk) {
  if (n > g_db.na) {
    int rc;
    if ((rc = dev_alloc(&g_db.d_a, n))) return rc;
    if ((rc = dev_alloc(&g_db.d_b, n))) return rc;
    if ((rc = dev_alloc(&g_db.d_c, n))) return rc;
    if ((rc = dev_alloc(&g_db.d_d, n))) return rc;
    g_db.na = n;
  }
  if (nijk > g_db.nijk) {
    int rc = dev_alloc(&g_db.d_ijk, nijk);
    if (rc) return rc;
    g_db.nijk = nijk;
  }
  if (!g_db.d_red) return dev_alloc(&g_db.d_red, (size_t)kRedBlocks * 8 + 64);
  return 0;
}

}  // namespace

int dropin_xclda(const double* rho, const double* rhocore, int n, double rdelv, double* vexc,
                 double* eeexc, double* e4) {
  if (n < 0) return -2;
  if (n == 0) {
    for (int k = 0; k < 4; ++k) e4[k] = 0.0;
    return 0;
  }
  hipStream_t st = nullptr;
  int rc = dropin_stream(&st);
  if (rc) return rc;
  if ((rc = dropin_bufs((size_t)n, 0))) return rc;
  QX_CHECK(hipMemcpyAsync(g_db.d_a, rho, sizeof(double) * n, hipMemcpyHostToDevice, st));
  if (rhocore)
    QX_CHECK(hipMemcpyAsync(g_db.d_b, rhocore, sizeof(double) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_sg_potential, dim3(kRedBlocks), dim3(kRedThreads), 0, st, g_db.d_a,
                     nullptr, nullptr, rhocore ? g_db.d_b : nullptr, (long)n, nullptr, nullptr,
                     nullptr, g_db.d_c, g_db.d_d, g_db.d_red);
  double* d_out = g_db.d_red + (size_t)kRedBlocks * 8;
  hipLaunchKernelGGL(k_red_final, dim3(1), dim3(kRedThreads), 0, st, g_db.d_red, kRedBlocks, 8,
                     kNe, 1, rdelv, d_out);
  QX_CHECK(hipGetLastError());
  double e[kNe];
  QX_CHECK(hipMemcpyAsync(vexc, g_db.d_c, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipMemcpyAsync(eeexc, g_db.d_d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipMemcpyAsync(e, d_out, sizeof(e), hipMemcpyDeviceToHost, st));
  QX_CHECK(hipStreamSynchronize(st));
  for (int k = 0; k < 4; ++k) e4[k] = e[2 + k];
  return 0;
}

// mode 0: chgg2r (rho out), 1: harting (vhar out), 2: chgr2g (rhgr out)
int dropin_sphere(int mode, double* grid, double* rhgr, int nplw5, const double* recnrmex,
                  const int* mfd2ft, int ntotfd, const int* nd, const int* ijkgd) {
  const int n1 = nd[0], n2 = nd[1], n3 = nd[2];
  const long N = (long)n1 * n2 * n3;
  if (nplw5 < 0 || fd_check(mfd2ft, ntotfd, N) || check_ijk(ijkgd, nplw5, N)) return -2;
  FullGrid* f = nullptr;
  hipStream_t st = nullptr;
  int rc = dropin_grid(n1, n2, n3, &f, &st);
  if (rc) return rc;
  const size_t n5x = 2 * ((size_t)nplw5 + 1);
  if ((rc = dropin_bufs(std::max<size_t>(n5x, nplw5 + 1), 2 * (size_t)nplw5 + 1))) return rc;
  const std::vector<int> ij = zero_based(ijkgd, nplw5);
  QX_CHECK(hipMemcpyAsync(g_db.d_ijk, ij.data(), sizeof(int) * ij.size(), hipMemcpyHostToDevice,
                          st));
  std::vector<double> h(N, 0.0);
  if (mode == 2) {
    if (mfd2ft)
      for (int i = 0; i < ntotfd; ++i) h[mfd2ft[i] - 1] = grid[i];
    else
      std::copy(grid, grid + N, h.begin());
    QX_CHECK(hipMemcpyAsync(f->d_x, h.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_sg_real, dim3(cdivu(N, 256)), dim3(256), 0, st, f->d_x, f->d_g, N);
    if ((rc = fg_transform(*f, st, -1, 1.0 / (double)N))) return rc;
    hipLaunchKernelGGL(k_sg_gather, dim3(cdivu(nplw5 + 1, 256)), dim3(256), 0, st, f->d_g,
                       g_db.d_ijk + nplw5, nplw5, g_db.d_a);
    QX_CHECK(hipGetLastError());
    QX_CHECK(hipMemcpyAsync(rhgr, g_db.d_a, sizeof(double) * n5x, hipMemcpyDeviceToHost, st));
    QX_CHECK(hipStreamSynchronize(st));
    return 0;
  }
  QX_CHECK(hipMemcpyAsync(g_db.d_a, rhgr, sizeof(double) * n5x, hipMemcpyHostToDevice, st));
  const double* p4 = nullptr;
  if (mode == 1) {
    std::vector<double> q(nplw5 + 1);
    for (int ig = 0; ig <= nplw5; ++ig) q[ig] = 2.0 * recnrmex[ig];
    QX_CHECK(hipMemcpyAsync(g_db.d_b, q.data(), sizeof(double) * q.size(), hipMemcpyHostToDevice,
                            st));
    p4 = g_db.d_b;
  }
  hipLaunchKernelGGL(k_sg_zero, dim3(cdivu(N, 256)), dim3(256), 0, st, f->d_g, N);
  hipLaunchKernelGGL(k_sg_scatter, dim3(cdivu(nplw5 + 1, 256)), dim3(256), 0, st, g_db.d_a, p4,
                     g_db.d_ijk + nplw5, nplw5, f->d_g);
  QX_CHECK(hipGetLastError());
  if ((rc = fg_transform(*f, st, +1, 1.0))) return rc;
  hipLaunchKernelGGL(k_sg_re, dim3(cdivu(N, 256)), dim3(256), 0, st, f->d_g,
                     mode == 0 ? f->d_x : nullptr, mode == 1 ? f->d_x : nullptr, N);
  QX_CHECK(hipGetLastError());
  QX_CHECK(hipMemcpyAsync(h.data(), f->d_x, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipStreamSynchronize(st));
  if (mfd2ft)
    for (int i = 0; i < ntotfd; ++i) grid[i] = h[mfd2ft[i] - 1];
  else
    std::copy(h.begin(), h.end(), grid);
  return 0;
}

// mxchgg (Sources/chgdns.f90:796) on host arrays, with the context's mixer
int dropin_mxchgg(Ctx* c, int* lreset, double* rhgr, double* rinr, int nplw5, double aslh,
                  double bslh, double* glocal, const int* mfd2ft, int ntotfd, const int* ijkgd,
                  int lmixhold) {
  auto& S = c->scf;
  if (S.imxchg < 0) return -2;
  if (fd_check(mfd2ft, ntotfd, c->N)) return -2;
  if (S.nplw5 != nplw5 || !S.d_ijk) {
    int rc = scf_sphere(c, nplw5, ijkgd);
    if (rc) return rc;
  }
  const size_t n5x = 2 * ((size_t)nplw5 + 1);
  QX_CHECK(hipMemcpyAsync(S.d_rhgr, rhgr, sizeof(double) * n5x, hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipMemcpyAsync(S.d_rinr, rinr, sizeof(double) * n5x, hipMemcpyHostToDevice, c->stream));
  int rc = scf_mix(c, *lreset, aslh, bslh, lmixhold, nullptr);
  if (rc) return rc;
  std::vector<double> h(c->N);
  QX_CHECK(hipMemcpyAsync(rhgr, S.d_rhgr, sizeof(double) * n5x, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipMemcpyAsync(rinr, S.d_rinr, sizeof(double) * n5x, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipMemcpyAsync(h.data(), c->d_rho, sizeof(double) * c->N, hipMemcpyDeviceToHost,
                          c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (mfd2ft)
    for (int i = 0; i < ntotfd; ++i) glocal[i] = h[mfd2ft[i] - 1];
  else
    std::copy(h.begin(), h.end(), glocal);
  *lreset = 0;
  return 0;
}

}  // namespace qx
