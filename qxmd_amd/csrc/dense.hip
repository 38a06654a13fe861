// dense.hip -- FP64 small dense linear algebra on the GPU for the band update:
//   * Cholesky S = L L^T and the explicit inverse L^-1 (the steps of rsdecmp
//     and rsinvcd, Sources/schmidt.f90:428-544), with the positive-definite
//     check of schmidt2 (Sources/schmidt.f90:256-267);
//   * the symmetric eigenproblem of the subspace Hamiltonian (role of
//     RsEIGQR, Sources/funcs.f90:380-764): Householder tridiagonalisation,
//     implicit QL with Wilkinson shifts, back-transformation, and the
//     reference's stable ascending sort.
// Matrices here are nb x nb with nb <= a few thousand; each routine runs in a
// single 1024-thread workgroup on L2-resident data (latency-bound stage).
#include <cmath>
#include "context.h"

namespace qx {

constexpr int kDenseThreads = 1024;

__device__ inline double block_sum(double v, double* red) {
  // wave reduction then across waves
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    red[32] = s;
  }
  __syncthreads();
  return red[32];
}

// S (column-major, lower triangle valid) -> L^-1 in the lower triangle, upper
// zeroed.  ok[0] = 1 if every diagonal of L^-1 exceeds 1e-30 (lcholesky).
// S(i,j) at s[i + n*j].
__global__ void __launch_bounds__(kDenseThreads) k_chol_inv(double* s, int n, int* ok) {
  __shared__ double red[40];
  const double zero = 1e-30;
#define S_(i, j) s[(size_t)(i) + (size_t)n * (j)]
  // ---- rsdecmp: column i of L from the already-finished columns k < i
  for (int i = 0; i < n; ++i) {
    double part = 0.0;
    for (int k = threadIdx.x; k < i; k += blockDim.x) part += S_(i, k) * S_(i, k);
    double d = S_(i, i) - block_sum(part, red);
    const bool pos = d > zero;
    double lii = pos ? sqrt(d) : 0.0;
    for (int j = i + 1 + threadIdx.x; j < n; j += blockDim.x) {
      if (pos) {
        double t = S_(j, i);
        for (int k = 0; k < i; ++k) t -= S_(j, k) * S_(i, k);
        S_(j, i) = t / lii;
      } else {
        S_(j, i) = 0.0;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) S_(i, i) = lii;
    __syncthreads();
  }
  // ---- rsinvcd: scale columns, then columns of the inverse from the right
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double d = S_(i, i);
    if (fabs(d) > zero) {
      double r = 1.0 / d;
      for (int j = i + 1; j < n; ++j) S_(j, i) *= r;
      S_(i, i) = r;
    }
  }
  __syncthreads();
  for (int j = n - 2; j >= 0; --j) {
    // work(i) = sum_{k=j+1}^{n-1} S(i,k) S(k,j), i > j; S(i,k) = 0 for k > i
    for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) {
      double w = 0.0;
      for (int k = j + 1; k <= i; ++k) w += S_(i, k) * S_(k, j);
      // columns k > j are final; S(k,j) for k > j is read by other threads
      // before this column is overwritten below
      s[(size_t)n * n + i] = w;  // scratch column after the matrix
    }
    __syncthreads();
    for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) S_(i, j) = -s[(size_t)n * n + i];
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    int i = idx % n, j = idx / n;
    if (i < j) S_(i, j) = 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int good = 1;
    for (int b = 0; b < n; ++b)
      if (!(S_(b, b) > zero)) good = 0;
    ok[0] = good;
  }
#undef S_
}

// Symmetric eigensolver.  a: n x n column-major, full symmetric (both halves);
// on exit eigenvectors in the columns of z (n x n), eigenvalues ascending in w.
// work: >= 4n doubles.  Householder vectors are kept in a's lower triangle.
__global__ void __launch_bounds__(kDenseThreads) k_eigh(double* a, int n, double* z, double* w,
                                                        double* work, int* info) {
  __shared__ double red[40];
  __shared__ double sh[8];
  double* d = work;          // diagonal
  double* e = work + n;      // off-diagonal (e[i] couples i, i+1)
  double* p = work + 2 * n;  // matvec
  double* v = work + 3 * n;  // Householder vector
#define A_(i, j) a[(size_t)(i) + (size_t)n * (j)]
#define Z_(i, j) z[(size_t)(i) + (size_t)n * (j)]
  // ---- tridiagonalisation: column k annihilated below k+1
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;  // length of x = A[k+1:, k]
    double part = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      double x = A_(k + 1 + i, k);
      part += x * x;
    }
    double nx2 = block_sum(part, red);
    double x0 = A_(k + 1, k);
    double alpha = nx2 > 0.0 ? (x0 >= 0.0 ? -sqrt(nx2) : sqrt(nx2)) : 0.0;
    // v = x - alpha e1, |v|^2 = nx2 - 2 alpha x0 + alpha^2
    double vn2 = nx2 - 2.0 * alpha * x0 + alpha * alpha;
    const bool skip = !(vn2 > 1e-300);
    double vinv = skip ? 0.0 : 1.0 / sqrt(vn2);
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      double x = A_(k + 1 + i, k);
      v[i] = (i == 0 ? x - alpha : x) * vinv;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      d[k] = A_(k, k);
      e[k] = skip ? x0 : alpha;
    }
    if (!skip) {
      // p = A_sub v  (A_sub = A[k+1:, k+1:], symmetric, full storage)
      for (int i = threadIdx.x; i < m; i += blockDim.x) {
        double t = 0.0;
        for (int j = 0; j < m; ++j) t += A_(k + 1 + j, k + 1 + i) * v[j];
        p[i] = t;
      }
      __syncthreads();
      part = 0.0;
      for (int i = threadIdx.x; i < m; i += blockDim.x) part += v[i] * p[i];
      double K = block_sum(part, red);
      // w = 2 (p - K v);  A_sub -= v w^T + w v^T
      for (int i = threadIdx.x; i < m; i += blockDim.x) p[i] = 2.0 * (p[i] - K * v[i]);
      __syncthreads();
      for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
        int i = idx % m, j = idx / m;
        A_(k + 1 + i, k + 1 + j) -= v[i] * p[j] + p[i] * v[j];
      }
    }
    __syncthreads();
    // keep v (unit) in column k below the diagonal for the back-transformation
    for (int i = threadIdx.x; i < m; i += blockDim.x) A_(k + 1 + i, k) = skip ? 0.0 : v[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (n >= 2) {
      d[n - 2] = A_(n - 2, n - 2);
      e[n - 2] = A_(n - 1, n - 2);
      // last column: no reflection, mark vector zero
    }
    d[n - 1] = A_(n - 1, n - 1);
    e[n - 1] = 0.0;
  }
  if (n >= 2)
    for (int i = threadIdx.x; i < 1; i += blockDim.x) A_(n - 1, n - 2) = 0.0;
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    int i = idx % n, j = idx / n;
    Z_(i, j) = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  // ---- implicit QL on (d, e); rotations of each sweep applied to rows of Z
  __shared__ double cs[2 * 4096];
  int fail = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      // find small subdiagonal
      __shared__ int msh;
      if (threadIdx.x == 0) {
        int m = l;
        for (; m < n - 1; ++m) {
          double dd = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= 1e-16 * dd) break;
        }
        msh = m;
      }
      __syncthreads();
      const int m = msh;
      if (m == l) break;
      if (++iter > 60) { fail = 1; break; }
      // thread 0: one QL sweep from m-1 down to l, record rotations
      if (threadIdx.x == 0) {
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, pp = 0.0;
        int i;
        int early = 0;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= pp;
            e[m] = 0.0;
            early = 1;
            cs[2 * (i - l)] = 1.0;
            cs[2 * (i - l) + 1] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - pp;
          r = (d[i] - g) * s + 2.0 * c * b;
          pp = s * r;
          d[i + 1] = g + pp;
          g = c * r - b;
          cs[2 * (i - l)] = c;
          cs[2 * (i - l) + 1] = s;
        }
        if (!early) {
          d[l] -= pp;
          e[l] = g;
          e[m] = 0.0;
        }
        sh[0] = early ? (double)(i) : (double)(l - 1);  // last rotation index applied
      }
      __syncthreads();
      const int ilast = (int)sh[0];
      // apply rotations i = m-1 .. ilast+1 (plus the r==0 one as identity) to Z rows
      for (int row = threadIdx.x; row < n; row += blockDim.x) {
        for (int i = m - 1; i > ilast; --i) {
          double c = cs[2 * (i - l)], s = cs[2 * (i - l) + 1];
          double f = Z_(row, i + 1);
          double zi = Z_(row, i);
          Z_(row, i + 1) = s * zi + c * f;
          Z_(row, i) = c * zi - s * f;
        }
      }
      __syncthreads();
    }
    if (fail) break;
  }
  // ---- back-transformation Z <- H_0 ... H_{n-3} Z
  for (int k = n - 3; k >= 0; --k) {
    const int m = n - k - 1;
    for (int col = threadIdx.x; col < n; col += blockDim.x) {
      double t = 0.0;
      for (int i = 0; i < m; ++i) t += A_(k + 1 + i, k) * Z_(k + 1 + i, col);
      t *= 2.0;
      if (t != 0.0)
        for (int i = 0; i < m; ++i) Z_(k + 1 + i, col) -= t * A_(k + 1 + i, k);
    }
    __syncthreads();
  }
  // ---- stable ascending sort of (d, Z columns): ranks, then permute via a
  if (threadIdx.x == 0) info[0] = fail;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double di = d[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      double dj = d[j];
      if (dj < di || (dj == di && j < i)) ++rank;
    }
    p[i] = (double)rank;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    int row = idx % n, col = idx / n;
    A_(row, (int)p[col]) = Z_(row, col);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) w[(int)p[i]] = d[i];
  __syncthreads();
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) z[idx] = a[idx];
#undef A_
#undef Z_
}

// H[i][j] for i >= j taken from C[n=i][m=j] and mirrored (hmatrix2 stores the
// lower-triangle products in both halves, Sources/engrad.f90:872-888).
__global__ void k_symmetrize_lower(const double* C, int n, double* H, double* eig) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  int i = (int)(idx % n), j = (int)(idx / n);
  int hi = i > j ? i : j, lo = i > j ? j : i;
  H[(size_t)i + (size_t)n * j] = C[(size_t)hi * n + lo];
  if (i == j && eig) eig[i] = C[(size_t)i * n + i];
}

int dense_chol_inv(Ctx* c, double* s, int n, int* ok_dev) {
  hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(kDenseThreads), 0, c->stream, s, n, ok_dev);
  QX_CHECK(hipGetLastError());
  return 0;
}

int dense_eigh(Ctx* c, double* a, int n, double* z, double* w, double* work, int* info) {
  if (n > 4096) return -5;
  hipLaunchKernelGGL(k_eigh, dim3(1), dim3(kDenseThreads), 0, c->stream, a, n, z, w, work, info);
  QX_CHECK(hipGetLastError());
  return 0;
}

int dense_symmetrize(Ctx* c, const double* C, int n, double* H, double* eig) {
  hipLaunchKernelGGL(k_symmetrize_lower, dim3((unsigned)(((long)n * n + 255) / 256)), dim3(256), 0,
                     c->stream, C, n, H, eig);
  QX_CHECK(hipGetLastError());
  return 0;
}

}  // namespace qx
