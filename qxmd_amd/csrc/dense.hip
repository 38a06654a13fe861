// dense.hip -- FP64 dense linear algebra of the band update, on the GPU.
//
//   * Cholesky S = L L^T and the explicit inverse L^-1 (the role of rsdecmp /
//     rsinvcd, Sources/schmidt.f90:428-544) with the positive-definite check of
//     schmidt2 (Sources/schmidt.f90:256-267), blocked by 64: the diagonal block
//     is factored and inverted in LDS, the panel below and the trailing matrix
//     are updated by many workgroups, and L^-1 is assembled block row by block
//     row, X_ij = -X_ii sum_{k=j}^{i-1} L_ik X_kj.
//   * The symmetric eigenproblem of the subspace Hamiltonian (the role of
//     RsEIGQR, Sources/funcs.f90:380-764): Householder tridiagonalisation,
//     eigenvalues by bisection on Sturm counts (one thread per eigenvalue),
//     eigenvectors by inverse iteration with modified Gram-Schmidt inside each
//     cluster of close eigenvalues (one workgroup per cluster), and the
//     back-transformation by the stored reflectors (parallel over columns).
//     Eigenvalues come out ascending with their vectors, as RsEIGQR's sort.
#include <cmath>
#include <cstdlib>
#include <vector>
#include "context.h"
#include "coop.h"

namespace qx {

constexpr int NBK = 64;           // block size of the Cholesky / inverse
constexpr int kDenseThreads = 1024;

__device__ inline double block_sum(double v, double* red) {
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

// ===========================================================================
// Cholesky and inverse.  S(i,j) at s[i + n*j] (lower triangle used).
// X = L^-1 is written row-major: X(i,j) at x[i*n + j], zero above the diagonal.
// ===========================================================================

// factor the diagonal block [k0, k0+kb): L_kk into S (lower), L_kk^-1 into X
__global__ void __launch_bounds__(256) k_potrf_blk(double* s, int n, int k0, int kb, double* x,
                                                   int* fail) {
  __shared__ double a[NBK][NBK + 1];
  __shared__ double xi[NBK][NBK + 1];
  const int t = threadIdx.x;
  for (int idx = t; idx < kb * kb; idx += blockDim.x) {
    int i = idx % kb, j = idx / kb;
    a[i][j] = (i >= j) ? s[(size_t)(k0 + i) + (size_t)n * (k0 + j)] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    if (t == 0) {
      double d = a[j][j];
      if (d > 1e-30) {
        a[j][j] = sqrt(d);
      } else {
        a[j][j] = 0.0;
        fail[0] = 1;
      }
    }
    __syncthreads();
    const double ljj = a[j][j];
    for (int i = j + 1 + t; i < kb; i += blockDim.x) a[i][j] = ljj > 0.0 ? a[i][j] / ljj : 0.0;
    __syncthreads();
    const int m = kb - j - 1;
    for (int idx = t; idx < m * m; idx += blockDim.x) {
      int i = j + 1 + idx % m, c = j + 1 + idx / m;
      if (i >= c) a[i][c] -= a[i][j] * a[c][j];
    }
    __syncthreads();
  }
  // inverse of the lower factor: column c by forward substitution
  if (t < kb) {
    const int c = t;
    for (int i = 0; i < kb; ++i) {
      if (i < c) {
        xi[i][c] = 0.0;
        continue;
      }
      double v = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) v -= a[i][k] * xi[k][c];
      xi[i][c] = a[i][i] > 0.0 ? v / a[i][i] : 0.0;
    }
  }
  __syncthreads();
  for (int idx = t; idx < kb * kb; idx += blockDim.x) {
    int i = idx % kb, j = idx / kb;
    if (i >= j) s[(size_t)(k0 + i) + (size_t)n * (k0 + j)] = a[i][j];
    x[(size_t)(k0 + i) * n + (k0 + j)] = xi[i][j];
  }
}

// panel: L(r, k0+c) = sum_{t<=c} S(r, k0+t) X_kk(c, t), rows r in [k0+kb, n)
__global__ void __launch_bounds__(256) k_trsm_panel(double* s, int n, int k0, int kb,
                                                    const double* x) {
  __shared__ double di[NBK][NBK + 1];  // X_kk(c, t)
  __shared__ double rs[NBK][NBK + 1];  // S(r, k0+t)
  const int r0 = k0 + kb + blockIdx.x * NBK;
  const int nr = min(NBK, n - r0);
  const int t = threadIdx.x;
  for (int idx = t; idx < kb * kb; idx += blockDim.x) {
    int c = idx / kb, tt = idx % kb;
    di[c][tt] = x[(size_t)(k0 + c) * n + (k0 + tt)];
  }
  for (int idx = t; idx < nr * kb; idx += blockDim.x) {
    int r = idx % nr, tt = idx / nr;
    rs[r][tt] = s[(size_t)(r0 + r) + (size_t)n * (k0 + tt)];
  }
  __syncthreads();
  for (int idx = t; idx < nr * kb; idx += blockDim.x) {
    int r = idx % nr, c = idx / nr;
    double v = 0.0;
    for (int tt = 0; tt <= c; ++tt) v += rs[r][tt] * di[c][tt];
    s[(size_t)(r0 + r) + (size_t)n * (k0 + c)] = v;
  }
}

// trailing update of the lower triangle: S(i,j) -= sum_t L(i,k0+t) L(j,k0+t)
__global__ void __launch_bounds__(256) k_syrk_update(double* s, int n, int k0, int kb) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bj > bi) return;
  __shared__ double li[NBK][NBK + 1];
  __shared__ double lj[NBK][NBK + 1];
  const int base = k0 + kb;
  const int i0 = base + bi * NBK, j0 = base + bj * NBK;
  const int ni = min(NBK, n - i0), nj = min(NBK, n - j0);
  const int t = threadIdx.x;
  for (int idx = t; idx < ni * kb; idx += blockDim.x) {
    int r = idx % ni, tt = idx / ni;
    li[r][tt] = s[(size_t)(i0 + r) + (size_t)n * (k0 + tt)];
  }
  for (int idx = t; idx < nj * kb; idx += blockDim.x) {
    int r = idx % nj, tt = idx / nj;
    lj[r][tt] = s[(size_t)(j0 + r) + (size_t)n * (k0 + tt)];
  }
  __syncthreads();
  for (int idx = t; idx < ni * nj; idx += blockDim.x) {
    int r = idx % ni, c = idx / ni;
    if (i0 + r < j0 + c) continue;
    double v = 0.0;
    for (int tt = 0; tt < kb; ++tt) v += li[r][tt] * lj[c][tt];
    s[(size_t)(i0 + r) + (size_t)n * (j0 + c)] -= v;
  }
}

// block row i of X: X_ij = -X_ii * sum_{k=j}^{i-1} L_ik X_kj for every j < i
__global__ void __launch_bounds__(256) k_trtri_row(const double* s, double* x, int n, int bi) {
  __shared__ double la[NBK][NBK + 1];
  __shared__ double xb[NBK][NBK + 1];
  const int bj = blockIdx.x;
  if (bj >= bi) return;
  const int i0 = bi * NBK, j0 = bj * NBK;
  const int ni = min(NBK, n - i0), nj = min(NBK, n - j0);
  const int t = threadIdx.x;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  for (int bk = bj; bk < bi; ++bk) {
    const int k0 = bk * NBK, nk = min(NBK, n - k0);
    __syncthreads();
    for (int idx = t; idx < ni * nk; idx += blockDim.x) {
      int r = idx % ni, c = idx / ni;
      la[r][c] = s[(size_t)(i0 + r) + (size_t)n * (k0 + c)];
    }
    for (int idx = t; idx < nk * nj; idx += blockDim.x) {
      int r = idx / nj, c = idx % nj;
      xb[r][c] = x[(size_t)(k0 + r) * n + (j0 + c)];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      int idx = t + q * 256;
      int r = idx / NBK, c = idx % NBK;
      if (r < ni && c < nj) {
        double v = 0.0;
        for (int kk = 0; kk < nk; ++kk) v += la[r][kk] * xb[kk][c];
        acc[q] += v;
      }
    }
  }
  // multiply by -X_ii
  __syncthreads();
  for (int idx = t; idx < ni * ni; idx += blockDim.x) {
    int r = idx / ni, c = idx % ni;
    la[r][c] = x[(size_t)(i0 + r) * n + (i0 + c)];
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    int idx = t + q * 256;
    int r = idx / NBK, c = idx % NBK;
    if (r < ni && c < nj) xb[r][c] = acc[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    int idx = t + q * 256;
    int r = idx / NBK, c = idx % NBK;
    if (r < ni && c < nj) {
      double v = 0.0;
      for (int kk = 0; kk <= r; ++kk) v += la[r][kk] * xb[kk][c];
      x[(size_t)(i0 + r) * n + (j0 + c)] = -v;
    }
  }
}

__global__ void k_zero_upper_rm(double* x, int n) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n) return;
  int i = (int)(idx / n), j = (int)(idx % n);
  if (j > i) x[idx] = 0.0;
}

__global__ void k_flag_clear(int* f) { f[0] = 0; }
__global__ void k_flag_to_ok(const int* fail, int* ok) { ok[0] = fail[0] ? 0 : 1; }

// ---- cooperative Cholesky + inverse: one persistent launch -----------------
// With X = L^-1, X_ij = -X_ii T_ij and T_ij = sum_{k=j}^{i-1} L_ik X_kj.  T is
// accumulated in place in the lower blocks of X as the factorisation moves
// right, so every tile product has K = 64.  Block column b takes two phases
// separated by grid barriers:
//   (B) the panel, L_rb = S_rb X_bb^T (r > b), and block row b of the
//       inverse, X_bj = -X_bb T_bj (j < b);
//   (C) the trailing update S_rc -= L_rb L_cb^T (b < c <= r) and the
//       accumulation T_ij += L_ib X_bj (i > b, j <= b); the workgroup that
//       owns the next diagonal tile factors and inverts it in LDS right after
//       its update (look-ahead), so block b+1 is ready at the barrier.
// All matrix traffic goes through device-scope accesses (the tiles move
// between workgroups on different XCDs).

// diagonal block in LDS: P (lower part filled) -> L_bb, Q -> L_bb^-1 (column
// sweep: row k of Q is final once divided by L_kk, then the rows below
// subtract L_ik Q_k).  Written to s (lower) and x (full block, zeros above).
__device__ void chol_diag_block(double* s, int n, int k0, int kb, double* x, int* fail,
                                double (*P)[NBK + 1], double (*Q)[NBK + 1]) {
  const int t = threadIdx.x;
  for (int idx = t; idx < NBK * NBK; idx += blockDim.x) {
    const int i = idx & (NBK - 1), j = idx / NBK;
    if (!(i < kb && j < kb && i >= j)) P[i][j] = 0.0;
    Q[i][j] = (i == j && i < kb) ? 1.0 : 0.0;
  }
  __syncthreads();
  const int rr = t & 63, part = t >> 6;
  for (int k = 0; k < kb; ++k) {
    const double dkk = P[k][k];
    const bool bad = !(dkk > 1e-30);
    const double lkk = bad ? 0.0 : sqrt(dkk), il = bad ? 0.0 : 1.0 / lkk;
    __syncthreads();
    if (t == 0) {
      P[k][k] = lkk;
      if (bad) fail[0] = 1;
    }
    if (t > k && t < kb) P[t][k] *= il;
    if (t <= k) Q[k][t] *= il;
    __syncthreads();
    const int i = k + 1 + rr;
    if (i < kb) {
      const double lik = P[i][k];
      for (int c = k + 1 + part; c <= i; c += 4) P[i][c] -= lik * P[c][k];
      for (int c = part; c <= k; c += 4) Q[i][c] -= lik * Q[k][c];
    }
    __syncthreads();
  }
  for (int idx = t; idx < kb * kb; idx += blockDim.x) {
    const int i = idx % kb, j = idx / kb;
    if (i >= j) st_agent(&s[(size_t)(k0 + i) + (size_t)n * (k0 + j)], P[i][j]);
    st_agent(&x[(size_t)(k0 + i) * n + (k0 + j)], Q[i][j]);
  }
  __syncthreads();
}

constexpr size_t kCholLds = sizeof(double) * (32 * kTileLd + 2 * NBK * (NBK + 1));

__global__ void __launch_bounds__(256) k_chol_inv_coop(double* __restrict__ s, int n,
                                                       double* __restrict__ x, int* fail, int* bar,
                                                       int* err) {
  extern __shared__ double smem[];
  auto As = reinterpret_cast<double (*)[kTileLd]>(smem);
  auto Bs = reinterpret_cast<double (*)[kTileLd]>(smem + 16 * kTileLd);
  auto P = reinterpret_cast<double (*)[NBK + 1]>(smem + 32 * kTileLd);
  auto Q = reinterpret_cast<double (*)[NBK + 1]>(smem + 32 * kTileLd + NBK * (NBK + 1));
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int nblk = (n + NBK - 1) / NBK;
  int nbar = 0;
  auto S_ = [&](int i, int j) -> double* { return &s[(size_t)i + (size_t)n * j]; };
  auto X_ = [&](int i, int j) -> double* { return &x[(size_t)i * n + j]; };
  for (size_t i = (size_t)g * blockDim.x + t; i < (size_t)n * n; i += (size_t)G * blockDim.x)
    st_agent(&x[i], 0.0);
  if (g == 0) {
    const int kb = min(NBK, n);
    for (int idx = t; idx < kb * kb; idx += blockDim.x) {
      const int i = idx % kb, j = idx / kb;
      if (i >= j) P[i][j] = *S_(i, j);
    }
    __syncthreads();
  }
  grid_barrier(bar, ++nbar * G, err);  // X zeroed before any block of it is written
  if (g == 0) chol_diag_block(s, n, 0, min(NBK, n), x, fail, P, Q);
  grid_barrier(bar, ++nbar * G, err);
  for (int b = 0; b < nblk; ++b) {
    const int k0 = b * NBK, kb = min(NBK, n - k0);
    const int nrb = nblk - b - 1;
    // (B) panel and block row b of the inverse
    for (int task = g; task < nrb + b; task += G) {
      Tile64 T;
      T.zero();
      if (task < nrb) {
        const int r0 = k0 + kb + task * NBK, nr = min(NBK, n - r0);
        T.run(
            kb, [&](int m, int k) {
              const double v = ld_agent(S_(r0 + min(m, nr - 1), k0 + k));
              return m < nr ? v : 0.0;
            },
            [&](int k, int c) {
              const double v = ld_agent(X_(k0 + min(c, kb - 1), k0 + k));
              return c < kb ? v : 0.0;
            }, As, Bs);
        T.each([&](int m, int c, double v) {
          if (m < nr && c < kb) st_agent(S_(r0 + m, k0 + c), v);
        });
      } else {
        const int j0 = (task - nrb) * NBK;  // a full block column left of b
        T.run(
            kb, [&](int m, int k) {
              const double v = ld_agent(X_(k0 + min(m, kb - 1), k0 + k));
              return m < kb ? v : 0.0;
            },
            [&](int k, int c) { return ld_agent(X_(k0 + k, j0 + c)); }, As, Bs);
        T.each([&](int m, int c, double v) {
          if (m < kb) st_agent(X_(k0 + m, j0 + c), -v);
        });
      }
    }
    if (nrb == 0) break;
    grid_barrier(bar, ++nbar * G, err);
    // (C) trailing update with the look-ahead factorisation of block b+1, and
    // the accumulation of T for the block rows below b
    const int ntile = nrb * (nrb + 1) / 2;
    for (int task = g; task < ntile + nrb * (b + 1); task += G) {
      Tile64 T;
      T.zero();
      if (task < ntile) {
        int ri = (int)((sqrt(8.0 * task + 1.0) - 1.0) * 0.5);
        while ((ri + 1) * (ri + 2) / 2 <= task) ++ri;
        while (ri * (ri + 1) / 2 > task) --ri;
        const int ci = task - ri * (ri + 1) / 2;
        const int r0 = k0 + kb + ri * NBK, c0 = k0 + kb + ci * NBK;
        const int nr = min(NBK, n - r0), nc = min(NBK, n - c0);
        T.run(
            kb, [&](int m, int k) {
              const double v = ld_agent(S_(r0 + min(m, nr - 1), k0 + k));
              return m < nr ? v : 0.0;
            },
            [&](int k, int c) {
              const double v = ld_agent(S_(c0 + min(c, nc - 1), k0 + k));
              return c < nc ? v : 0.0;
            }, As, Bs);
        if (task == 0) {
          // next diagonal block: updated values straight into LDS, then factor
          T.each([&](int m, int c, double v) {
            if (m < nr && c < nc && m >= c) P[m][c] = ld_agent(S_(r0 + m, c0 + c)) - v;
          });
          __syncthreads();
          chol_diag_block(s, n, r0, nr, x, fail, P, Q);
        } else {
          T.each([&](int m, int c, double v) {
            if (m < nr && c < nc) st_agent(S_(r0 + m, c0 + c), ld_agent(S_(r0 + m, c0 + c)) - v);
          });
        }
      } else {
        const int tt = task - ntile;
        const int i0 = k0 + kb + (tt / (b + 1)) * NBK, j0 = (tt % (b + 1)) * NBK;
        const int ni = min(NBK, n - i0), nj = j0 == k0 ? kb : NBK;
        T.run(
            kb, [&](int m, int k) {
              const double v = ld_agent(S_(i0 + min(m, ni - 1), k0 + k));
              return m < ni ? v : 0.0;
            },
            [&](int k, int c) {
              const double v = ld_agent(X_(k0 + k, j0 + min(c, nj - 1)));
              return c < nj ? v : 0.0;
            }, As, Bs);
        T.each([&](int m, int c, double v) {
          if (m < ni && c < nj) st_agent(X_(i0 + m, j0 + c), ld_agent(X_(i0 + m, j0 + c)) + v);
        });
      }
    }
    grid_barrier(bar, ++nbar * G, err);
  }
}

// sync[0] barrier counter, sync[1] barrier error of the current launch,
// sync[2] sticky error of all launches since the host last looked
__global__ void k_coop_reset(int* sync) {
  sync[2] |= sync[1];
  sync[0] = 0;
  sync[1] = 0;
}
__global__ void k_chol_flags(const int* fail, const int* sync, int* ok) {
  ok[0] = (fail[0] || sync[1]) ? 0 : 1;
  ok[1] = sync[1];
}

// exact: launch exactly `want` groups or nothing (kernels whose data layout
// depends on the group count); otherwise fewer groups are fine.
int coop_launch(Ctx* c, const void* fn, int want, size_t lds, void** args, bool exact) {
  if (std::getenv("QX_DENSE_LEGACY")) return -1;
  int dev = 0, ncu = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
  if (lds > 65536 &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return -1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 256, lds) != hipSuccess || per < 1)
    return -1;
  const int G = std::max(1, std::min(want, per * ncu));
  if (exact && G != want) return -1;
  hipLaunchKernelGGL(k_coop_reset, dim3(1), dim3(1), 0, c->stream, c->d_sync);
  if (hipLaunchCooperativeKernel(fn, dim3(G), dim3(256), args, (unsigned)lds, c->stream) !=
      hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  return G;
}

// legacy multi-launch Cholesky + inverse (used when a cooperative launch is
// not available)
static int chol_inv_legacy(Ctx* c, double* s, int n, double* x, int* flag) {
  hipStream_t st = c->stream;
  const int nblk = (n + NBK - 1) / NBK;
  for (int b = 0; b < nblk; ++b) {
    const int k0 = b * NBK, kb = std::min(NBK, n - k0);
    hipLaunchKernelGGL(k_potrf_blk, dim3(1), dim3(256), 0, st, s, n, k0, kb, x, flag);
    const int rem = n - k0 - kb;
    if (rem > 0) {
      const int rb = (rem + NBK - 1) / NBK;
      hipLaunchKernelGGL(k_trsm_panel, dim3(rb), dim3(256), 0, st, s, n, k0, kb, x);
      hipLaunchKernelGGL(k_syrk_update, dim3(rb, rb), dim3(256), 0, st, s, n, k0, kb);
    }
  }
  hipLaunchKernelGGL(k_zero_upper_rm, dim3((unsigned)(((long)n * n + 255) / 256)), dim3(256), 0, st,
                     x, n);
  for (int bi = 1; bi < nblk; ++bi)
    hipLaunchKernelGGL(k_trtri_row, dim3(bi), dim3(256), 0, st, s, x, n, bi);
  return 0;
}

// s: n x n column-major (lower valid; modified), x: n x n row-major output
// L^-1; ok[0] = lcholesky, ok[1] = 1 if a grid barrier gave up.  flag:
// device int scratch.
int dense_chol_inv(Ctx* c, double* s, int n, double* x, int* ok, int* flag) {
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(1), 0, st, flag);
  const int nblk = (n + NBK - 1) / NBK;
  int want = 1;
  for (int b = 0; b < nblk; ++b) {
    const int nrb = nblk - b - 1;
    want = std::max(want, std::max(nrb * (nrb + 1) / 2 + nrb * (b + 1), nrb + b));
  }
  int* bar = c->d_sync;
  int* err = c->d_sync + 1;
  void* args[] = {&s, &n, &x, &flag, &bar, &err};
  if (coop_launch(c, reinterpret_cast<const void*>(k_chol_inv_coop), want, kCholLds, args, false) < 0) {
    hipLaunchKernelGGL(k_coop_reset, dim3(1), dim3(1), 0, st, c->d_sync);
    chol_inv_legacy(c, s, n, x, flag);
  }
  hipLaunchKernelGGL(k_chol_flags, dim3(1), dim3(1), 0, st, flag, c->d_sync, ok);
  QX_CHECK(hipGetLastError());
  return 0;
}

// ===========================================================================
// Symmetric eigensolver
// ===========================================================================

// Householder tridiagonalisation in one workgroup.  a: n x n column-major,
// full symmetric storage.  On exit d (diag), e (e[i] couples i, i+1) and the
// unit reflector of step k in a(k+1:n, k) (zero when skipped).  Each step
// makes one coalesced pass over the trailing block: it applies the rank-2
// update of the step and accumulates the next step's matrix-vector product.
__global__ void __launch_bounds__(kDenseThreads) k_tridiag(double* a, int n, double* d, double* e,
                                                           double* vwork) {
  __shared__ double red[40];
  double* v = vwork;            // current reflector (n)
  double* w = vwork + n;        // rank-2 partner (n)
  double* p = vwork + 2 * n;    // matvec accumulators (n)
  double* vn = vwork + 3 * n;   // next reflector (n)
#define A_(i, j) a[(size_t)(i) + (size_t)n * (j)]
  const int t = threadIdx.x;
  // reflector for column k from the current column values (already updated)
  auto make_reflector = [&](int k, double* vout, double& alpha, bool& skip) {
    const int m = n - k - 1;
    double part = 0.0;
    for (int i = t; i < m; i += blockDim.x) {
      double xv = A_(k + 1 + i, k);
      part += xv * xv;
    }
    double nx2 = block_sum(part, red);
    double x0 = A_(k + 1, k);
    alpha = nx2 > 0.0 ? (x0 >= 0.0 ? -sqrt(nx2) : sqrt(nx2)) : 0.0;
    double vn2 = nx2 - 2.0 * alpha * x0 + alpha * alpha;
    skip = !(vn2 > 1e-300);
    double inv = skip ? 0.0 : 1.0 / sqrt(vn2);
    for (int i = t; i < m; i += blockDim.x) {
      double xv = A_(k + 1 + i, k);
      vout[i] = (i == 0 ? xv - alpha : xv) * inv;
    }
    __syncthreads();
  };
  if (n < 3) {
    if (t == 0) {
      for (int i = 0; i < n; ++i) d[i] = A_(i, i);
      if (n == 2) e[0] = A_(1, 0);
      e[n - 1] = 0.0;
    }
    return;
  }
  double alpha;
  bool skip;
  make_reflector(0, v, alpha, skip);
  // p = A(1:, 1:) v
  {
    const int m = n - 1;
    for (int i = t; i < m; i += blockDim.x) {
      double acc = 0.0;
      for (int j = 0; j < m; ++j) acc += A_(1 + i, 1 + j) * v[j];
      p[i] = acc;
    }
    __syncthreads();
  }
  for (int k = 0; k + 2 < n; ++k) {
    const int m = n - k - 1;
    if (t == 0) {
      d[k] = A_(k, k);
      e[k] = skip ? A_(k + 1, k) : alpha;
    }
    if (!skip) {
      double part = 0.0;
      for (int i = t; i < m; i += blockDim.x) part += v[i] * p[i];
      double K = block_sum(part, red);
      for (int i = t; i < m; i += blockDim.x) w[i] = 2.0 * (p[i] - K * v[i]);
      __syncthreads();
      // column k+1 below the diagonal after this step's update (it feeds the
      // next reflector, or e[n-2] on the last step)
      for (int i = t; i < m - 1; i += blockDim.x)
        A_(k + 2 + i, k + 1) -= v[1 + i] * w[0] + w[1 + i] * v[0];
      __syncthreads();
    }
    double alpha_n = 0.0;
    bool skip_n = true;
    if (k + 3 < n) make_reflector(k + 1, vn, alpha_n, skip_n);
    // fused pass: update the trailing block (columns k+2.. and the diagonal
    // element k+1) and accumulate p_next = A'(k+2:, k+2:) vn
    const int mn = m - 1;  // size of the next trailing block
    if (!skip) {
      if (t == 0) A_(k + 1, k + 1) -= 2.0 * v[0] * w[0];
      for (int i = t; i < mn; i += blockDim.x) {
        // column k+1 rows below (symmetric copy) handled above; row k+1 of
        // the next block is not needed later
        double acc = 0.0;
        const double vi = v[1 + i], wi = w[1 + i];
        for (int j = 0; j < mn; ++j) {
          double aij = A_(k + 2 + i, k + 2 + j) - vi * w[1 + j] - wi * v[1 + j];
          A_(k + 2 + i, k + 2 + j) = aij;
          acc += aij * vn[j];
        }
        p[i] = acc;
      }
    } else if (mn > 0) {
      for (int i = t; i < mn; i += blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < mn; ++j) acc += A_(k + 2 + i, k + 2 + j) * vn[j];
        p[i] = acc;
      }
    }
    __syncthreads();
    // keep v in column k, move vn -> v
    for (int i = t; i < m; i += blockDim.x) A_(k + 1 + i, k) = skip ? 0.0 : v[i];
    __syncthreads();
    for (int i = t; i < mn; i += blockDim.x) v[i] = vn[i];
    __syncthreads();
    alpha = alpha_n;
    skip = skip_n;
  }
  if (t == 0) {
    d[n - 2] = A_(n - 2, n - 2);
    e[n - 2] = A_(n - 1, n - 2);
    d[n - 1] = A_(n - 1, n - 1);
    e[n - 1] = 0.0;
  }
#undef A_
}

// Sturm count: number of eigenvalues of T below sigma
__device__ inline int sturm_count(const double* d, const double* e, int n, double sigma,
                                  double pivmin) {
  int cnt = 0;
  double q = d[0] - sigma;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++cnt;
  for (int i = 1; i < n; ++i) {
    q = d[i] - sigma - e[i - 1] * e[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// one thread per eigenvalue index j: bisection to full precision
__global__ void k_bisect(const double* d, const double* e, int n, double* w, double lo0, double hi0,
                         double pivmin) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double lo = lo0, hi = hi0;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (hi - lo <= 4.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    if (sturm_count(d, e, n, mid, pivmin) > j) hi = mid;
    else lo = mid;
  }
  w[j] = 0.5 * (lo + hi);
}

// Inverse iteration for the eigenvalues of one cluster [c0, c1) in one
// workgroup; vectors are orthogonalised against the cluster's earlier ones.
// z: n x n column-major (vector j in column j).  lu: 5n doubles per block.
__global__ void __launch_bounds__(256) k_inviter(const double* d, const double* e, int n,
                                                 const double* w, const int* cstart, int ncl,
                                                 double* z, double* luwork, double tnorm) {
  __shared__ double red[40];
  const int cl = blockIdx.x;
  if (cl >= ncl) return;
  const int c0 = cstart[cl], c1 = cstart[cl + 1];
  double* lu = luwork + (size_t)cl * 5 * n;
  double* dd = lu;             // U diagonal
  double* du = lu + n;         // U first superdiagonal
  double* du2 = lu + 2 * n;    // U second superdiagonal (from pivoting)
  double* dl = lu + 3 * n;     // multipliers
  double* piv = lu + 4 * n;    // 1 if rows i, i+1 swapped
  const double eps = 2.220446049250313e-16;
  const double pert = eps * tnorm;
  for (int j = c0; j < c1; ++j) {
    double* x = z + (size_t)n * j;
    const double lam = w[j];
    // LU with partial pivoting of T - lam I (one thread; O(n))
    if (threadIdx.x == 0) {
      for (int i = 0; i < n; ++i) {
        dd[i] = d[i] - lam;
        du[i] = i < n - 1 ? e[i] : 0.0;
        du2[i] = 0.0;
      }
      for (int i = 0; i < n - 1; ++i) {
        double sub = e[i];
        if (fabs(dd[i]) >= fabs(sub)) {
          piv[i] = 0.0;
          double f = dd[i] != 0.0 ? sub / dd[i] : 0.0;
          dl[i] = f;
          dd[i + 1] -= f * du[i];
        } else {
          piv[i] = 1.0;
          double f = dd[i] / sub;
          dl[i] = f;
          dd[i] = sub;
          double tmp = du[i];
          du[i] = dd[i + 1];
          dd[i + 1] = tmp - f * dd[i + 1];
          if (i < n - 2) {
            du2[i] = du[i + 1];
            du[i + 1] = -f * du[i + 1];
          }
        }
      }
      for (int i = 0; i < n; ++i)
        if (fabs(dd[i]) < pert) dd[i] = dd[i] < 0.0 ? -pert : pert;
    }
    // deterministic start vector
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      unsigned h = (unsigned)(i * 2654435761u) ^ (unsigned)(j * 40503u + 12345u);
      h ^= h >> 13;
      h *= 0x5bd1e995u;
      h ^= h >> 15;
      x[i] = 0.5 + (double)(h & 0xffff) / 65536.0;
    }
    __syncthreads();
    for (int it = 0; it < 3; ++it) {
      if (threadIdx.x == 0) {
        // solve L U x = P b: forward with the recorded row swaps, then back
        for (int i = 0; i < n - 1; ++i) {
          if (piv[i] == 0.0) {
            x[i + 1] -= dl[i] * x[i];
          } else {
            double tmp = x[i];
            x[i] = x[i + 1];
            x[i + 1] = tmp - dl[i] * x[i + 1];
          }
        }
        // scale to avoid overflow: track the largest entry
        double big = 0.0;
        for (int i = n - 1; i >= 0; --i) {
          double v = x[i];
          if (i + 1 < n) v -= du[i] * x[i + 1];
          if (i + 2 < n) v -= du2[i] * x[i + 2];
          x[i] = v / dd[i];
          big = fmax(big, fabs(x[i]));
          if (big > 1e150) {  // rescale the whole system (solved part and rhs)
            for (int q = 0; q < n; ++q) x[q] *= 1e-150;
            big *= 1e-150;
          }
        }
      }
      __syncthreads();
      // orthogonalise against the cluster's finished vectors (MGS)
      for (int k = c0; k < j; ++k) {
        const double* zk = z + (size_t)n * k;
        double part = 0.0;
        for (int i = threadIdx.x; i < n; i += blockDim.x) part += zk[i] * x[i];
        double dot = block_sum(part, red);
        for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] -= dot * zk[i];
        __syncthreads();
      }
      double part = 0.0;
      for (int i = threadIdx.x; i < n; i += blockDim.x) part += x[i] * x[i];
      double nrm = sqrt(block_sum(part, red));
      double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
      for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] *= inv;
      __syncthreads();
    }
  }
}

// back-transformation Z <- H_0 ... H_{n-3} Z, 16 columns per block, 16 lanes
// per column over the rows
__global__ void __launch_bounds__(256) k_backtrans(const double* a, int n, double* z) {
  const int col = blockIdx.x * 16 + (threadIdx.x >> 4);
  const int lane = threadIdx.x & 15;
  const bool active = col < n;
  double* zc = z + (size_t)n * (active ? col : 0);
  for (int k = n - 3; k >= 0; --k) {
    const double* vk = a + (size_t)n * k + (k + 1);  // reflector rows k+1..n-1
    const int m = n - k - 1;
    double part = 0.0;
    if (active)
      for (int i = lane; i < m; i += 16) part += vk[i] * zc[k + 1 + i];
    for (int o = 8; o > 0; o >>= 1) part += __shfl_xor(part, o, 16);
    const double tv = 2.0 * part;
    if (active && tv != 0.0)
      for (int i = lane; i < m; i += 16) zc[k + 1 + i] -= tv * vk[i];
  }
}

// host driver.  a: n x n column-major (symmetric, destroyed), z: eigenvectors
// (columns), w: eigenvalues ascending; work >= 12n doubles; iwork >= n+1 ints
// (device).  Cluster boundaries are found on the host from w.
int dense_eigh_legacy(Ctx* c, double* a, int n, double* z, double* w, double* work) {
  hipStream_t st = c->stream;
  double* d = work;
  double* e = work + n;
  double* vw = work + 2 * n;  // 4n
  hipLaunchKernelGGL(k_tridiag, dim3(1), dim3(kDenseThreads), 0, st, a, n, d, e, vw);
  std::vector<double> hd(n), he(n);
  QX_CHECK(hipMemcpyAsync(hd.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipMemcpyAsync(he.data(), e, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipStreamSynchronize(st));
  double lo = 0.0, hi = 0.0, tnorm = 0.0, emax2 = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = (i > 0 ? std::fabs(he[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(he[i]) : 0.0);
    lo = i == 0 ? hd[i] - r : std::min(lo, hd[i] - r);
    hi = i == 0 ? hd[i] + r : std::max(hi, hd[i] + r);
    tnorm = std::max(tnorm, std::fabs(hd[i]) + r);
    if (i < n - 1) emax2 = std::max(emax2, he[i] * he[i]);
  }
  const double eps = 2.220446049250313e-16;
  const double safmin = 2.2250738585072014e-308;
  const double pivmin = std::max(safmin * std::max(1.0, emax2), safmin);
  lo -= 2.0 * eps * tnorm + 4.0 * pivmin;
  hi += 2.0 * eps * tnorm + 4.0 * pivmin;
  hipLaunchKernelGGL(k_bisect, dim3((n + 127) / 128), dim3(128), 0, st, d, e, n, w, lo, hi, pivmin);
  std::vector<double> hw(n);
  QX_CHECK(hipMemcpyAsync(hw.data(), w, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipStreamSynchronize(st));
  // clusters: consecutive eigenvalues closer than 1e-3 ||T|| are handled
  // together (the orthogonality tolerance of the classical inverse iteration)
  std::vector<int> cs;
  cs.push_back(0);
  for (int j = 1; j < n; ++j)
    if (hw[j] - hw[j - 1] > 1e-3 * tnorm) cs.push_back(j);
  cs.push_back(n);
  const int ncl = (int)cs.size() - 1;
  int* dcs = reinterpret_cast<int*>(work + 6 * n);  // n+1 ints
  QX_CHECK(hipMemcpyAsync(dcs, cs.data(), sizeof(int) * cs.size(), hipMemcpyHostToDevice, st));
  // 5 n * ncl <= 5 n^2 doubles of scratch from the dense workspace
  if (c->dwork_len < 5 * (size_t)n * n) return -8;
  double* lu = c->d_dwork;
  hipLaunchKernelGGL(k_inviter, dim3(ncl), dim3(256), 0, st, d, e, n, w, dcs, ncl, z, lu, tnorm);
  hipLaunchKernelGGL(k_backtrans, dim3((n + 15) / 16), dim3(256), 0, st, a, n, z);
  QX_CHECK(hipGetLastError());
  return 0;
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

int dense_symmetrize(Ctx* c, const double* C, int n, double* H, double* eig) {
  hipLaunchKernelGGL(k_symmetrize_lower, dim3((unsigned)(((long)n * n + 255) / 256)), dim3(256), 0,
                     c->stream, C, n, H, eig);
  QX_CHECK(hipGetLastError());
  return 0;
}

}  // namespace qx
