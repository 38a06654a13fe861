// gemm.hip -- FP64 MFMA contractions of the band algebra.
//
// Every dense contraction on the path is one of two shapes over the
// real-packed plane-wave index (length nplwex, contiguous in memory):
//   gemm_tn : C[n][m] = alpha * sum_k A[m][k] B[n][k] + beta C   (K = plane waves)
//             projections P = B^T Psi (hckbpp_g, Sources/nlkbpp.f90:1358-1364),
//             overlap S = Psi^T Psi (schmidt2, Sources/schmidt.f90:199-216),
//             subspace H = Psi^T H Psi (hmatrix2, Sources/engrad.f90:869-882)
//   gemm_nn : C[n][m] = alpha * sum_k A[k][m] B[n][k] + beta C   (M = plane waves)
//             back-projection H Psi += B D P (Sources/nlkbpp.f90:1376-1379),
//             rotations Psi U (untrycore, Sources/eigen.f90:1456-1514) and
//             Psi L^-T (schmidt2, Sources/schmidt.f90:284-313).
// The packed inner product 2 sum - x0 y0 (G=0 counted once) is applied as
// alpha = 2 plus a rank-1 correction of row 0 (and row 1 for the KB sum,
// which skips Im c(G=0): Sources/nlkbpp.f90:1360-1363).
//
// Sizable outputs (M, N >= 256) run on the 128 x 128 kernels of
// mfma_gemm.hip.  The kernels here take the skinny shapes: 64x64 per
// 256-thread block, 4 wavefronts of 32x32, each a 2x2 grid of
// v_mfma_f64_16x16x4 blocks, K staged through LDS 16 at a time, the long-K
// shape split over blocks and reduced with FP64 atomics; at most 4 columns,
// matrix-vector kernels.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include "context.h"

namespace qx {

static std::atomic<int> g_gemm_path{-1};

int gemm_path() {
  int p = g_gemm_path.load(std::memory_order_relaxed);
  if (p < 0) {
    const char* e = std::getenv("QX_GEMM_PATH");
    p = e ? std::atoi(e) : GEMM_OWN;
    if (p < GEMM_OWN || p > GEMM_OWN_F32) p = GEMM_OWN;
    g_gemm_path.store(p, std::memory_order_relaxed);
  }
  return p;
}

void set_gemm_path(int p) { g_gemm_path.store(p, std::memory_order_relaxed); }

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, KT = 16, LDT = BM + 2;

template <bool AKC, bool SPLIT>
__global__ void __launch_bounds__(256) k_gemm_f64(int M, int N, int K, int kchunk, double alpha,
                                                  const double* __restrict__ A, int lda,
                                                  const double* __restrict__ B, int ldb,
                                                  double* __restrict__ C, int ldc) {
  __shared__ double As[KT][LDT];
  __shared__ double Bs[KT][LDT];
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  // operands of a stage: unconditional loads from clamped (valid) addresses,
  // then a select, so all of them are in flight together; the next stage is
  // fetched into registers while the current one is on the MFMA pipes
  double ra[4], rb[4];
  auto fetch = [&](int k0) {
    if (AKC) {
      const int m = tid >> 2, kk = (tid & 3) * 4, gm = m0 + m, gmc = min(gm, M - 1);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int gk = k0 + kk + u;
        const double v = A[(size_t)gmc * lda + min(gk, ke - 1)];
        ra[u] = (gm < M && gk < ke) ? v : 0.0;
      }
    } else {
      const int kk = tid >> 4, m = (tid & 15) * 4, gk = k0 + kk, gkc = min(gk, ke - 1);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int gm = m0 + m + u;
        const double v = A[(size_t)gkc * lda + min(gm, M - 1)];
        ra[u] = (gm < M && gk < ke) ? v : 0.0;
      }
    }
    const int n = tid >> 2, kk = (tid & 3) * 4, gn = n0 + n, gnc = min(gn, N - 1);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int gk = k0 + kk + u;
      const double v = B[(size_t)gnc * ldb + min(gk, ke - 1)];
      rb[u] = (gn < N && gk < ke) ? v : 0.0;
    }
  };
  fetch(kb);
  for (int k0 = kb; k0 < ke; k0 += KT) {
    if (AKC) {
      const int m = tid >> 2, kk = (tid & 3) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u) As[kk + u][m] = ra[u];
    } else {
      const int kk = tid >> 4, m = (tid & 15) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u) As[kk][m + u] = ra[u];
    }
    {
      const int n = tid >> 2, kk = (tid & 3) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u) Bs[kk + u][n] = rb[u];
    }
    __syncthreads();
    if (k0 + KT < ke) fetch(k0 + KT);
#pragma unroll
    for (int k4 = 0; k4 < KT; k4 += 4) {
      const int kr = k4 + (lane >> 4), c = lane & 15;
      double a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
      double b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  // ---- epilogue: v_mfma_f64_16x16x4 leaves D(i, j) in lane j + 16*(i%4),
  // register i/4, i.e. acc[i][j][v] = C(m = .. + lane/16 + 4*v, n = .. + lane%16)
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      int n = n0 + wn + j * 16 + (lane & 15);
      if (n >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int m = m0 + wm + i * 16 + (lane >> 4) + 4 * v;
        if (m >= M) continue;
        double val = alpha * acc[i][j][v];
        if (SPLIT) atomicAdd(&C[(size_t)n * ldc + m], val);
        else C[(size_t)n * ldc + m] += val;
      }
    }
}

__global__ void k_scale(double* C, int M, int N, int ldc, double beta) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  int m = (int)(i % M), n = (int)(i / M);
  double* p = C + (size_t)n * ldc + m;
  *p = beta == 0.0 ? 0.0 : beta * *p;
}

// packed-inner-product corrections: C[n][m] -= A[m][0] B[n][0]  (+ 2 A[m][1] B[n][1])
__global__ void k_corr(double* C, int M, int N, int ldc, const double* A, int lda,
                       const double* B, int ldb, double alpha, int skip_im0) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)M * N) return;
  int m = (int)(i % M), n = (int)(i / M);
  double c = A[(size_t)m * lda] * B[(size_t)n * ldb];
  if (skip_im0) c += 2.0 * A[(size_t)m * lda + 1] * B[(size_t)n * ldb + 1];
  C[(size_t)n * ldc + m] -= alpha * 0.5 * c;
}

static inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// corr: 0 none, 1 subtract the G=0 term, 2 also the Im(G=0) term (KB)
static int gemv_tn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc);
static int gemv_nn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc);
// products with at most kGemvCols columns and more than kGemvMinRows rows take
// the GEMV kernels (tiny products keep the tile kernels' summation order)
constexpr int kGemvCols = 4, kGemvMinRows = 16;
static bool gemv_shape(int M, int N) {
  return N <= kGemvCols && M > kGemvMinRows && !std::getenv("QX_NO_GEMV");
}

static int gemm_tn_small(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                         const double* B, int ldb, double beta, double* C, int ldc, int corr) {
  if (M <= 0 || N <= 0) return 0;
  if (gemv_shape(M, N)) {
    gemv_tn(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    if (corr)
      hipLaunchKernelGGL(k_corr, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N,
                         ldc, A, lda, B, ldb, alpha, corr == 2 ? 1 : 0);
    QX_CHECK(hipGetLastError());
    return 0;
  }
  hipLaunchKernelGGL(k_scale, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N, ldc,
                     beta);
  const int tiles = (int)(cdiv(M, BM) * cdiv(N, BN));
  int nsplit = std::max<int>(1, std::min<int>((int)cdiv(K, 1024), std::max(1, 2048 / tiles)));
  int kchunk = (int)(cdiv(cdiv(K, nsplit), KT) * KT);
  nsplit = (int)cdiv(K, kchunk);
  dim3 grid(cdiv(M, BM), cdiv(N, BN), nsplit);
  lt_mark(c, c->lt_gemm_id, false);
  if (nsplit > 1)
    hipLaunchKernelGGL((k_gemm_f64<true, true>), grid, dim3(256), 0, c->stream, M, N, K, kchunk,
                       alpha, A, lda, B, ldb, C, ldc);
  else
    hipLaunchKernelGGL((k_gemm_f64<true, false>), grid, dim3(256), 0, c->stream, M, N, K, kchunk,
                       alpha, A, lda, B, ldb, C, ldc);
  lt_mark(c, c->lt_gemm_id, true);
  if (corr)
    hipLaunchKernelGGL(k_corr, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N, ldc,
                       A, lda, B, ldb, alpha, corr == 2 ? 1 : 0);
  QX_CHECK(hipGetLastError());
  return 0;
}

static int gemm_nn_small(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                         const double* B, int ldb, double beta, double* C, int ldc) {
  if (M <= 0 || N <= 0) return 0;
  if (gemv_shape(M, N)) {
    gemv_nn(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    QX_CHECK(hipGetLastError());
    return 0;
  }
  if (beta != 1.0)
    hipLaunchKernelGGL(k_scale, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N,
                       ldc, beta);
  dim3 grid(cdiv(M, BM), cdiv(N, BN), 1);
  lt_mark(c, c->lt_gemm_id, false);
  hipLaunchKernelGGL((k_gemm_f64<false, false>), grid, dim3(256), 0, c->stream, M, N, K, K, alpha,
                     A, lda, B, ldb, C, ldc);
  lt_mark(c, c->lt_gemm_id, true);
  QX_CHECK(hipGetLastError());
  return 0;
}

// ---- skinny shapes (N <= 4 columns): matrix-vector products at HBM speed ---
// The single-band projections of the band-by-band schedule read all bands once
// per product; a 64 x 64 tile kernel would waste 63/64 of its work there.
// C(m, n) = alpha sum_k A(k, m) B(k, n) + beta C(m, n), n < NV: one workgroup
// per row m of C (A column m is contiguous), K spread over its threads
template <int NV>
__global__ void __launch_bounds__(256) k_gemv_t(int M, int K, double alpha,
                                                const double* __restrict__ A, int lda,
                                                const double* __restrict__ B, int ldb, double beta,
                                                double* __restrict__ C, int ldc) {
  __shared__ double red[NV][4];
  const int m = blockIdx.x;
  const double* a = A + (size_t)m * lda;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  // eight independent loads of A per thread in flight: the row is streamed
  // from HBM, B (shared by every row) comes from L2
  constexpr int U = 8;
  int k = threadIdx.x;
  for (; k + (U - 1) * 256 < K; k += U * 256) {
    double av[U];
#pragma unroll
    for (int u = 0; u < U; ++u) av[u] = a[k + u * 256];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] += av[u] * B[(size_t)v * ldb + k + u * 256];
  }
  for (; k < K; k += 256) {
    const double av = a[k];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += av * B[(size_t)v * ldb + k];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double x = acc[v];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    if (lane == 0) red[v][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int v = threadIdx.x;
    const double sum = red[v][0] + red[v][1] + red[v][2] + red[v][3];
    double* cp = C + (size_t)v * ldc + m;
    *cp = beta == 0.0 ? alpha * sum : beta * *cp + alpha * sum;
  }
}

// C(m, n) += alpha sum_k A(m, k) B(k, n), n < NV (C already scaled by beta):
// threads over m (coalesced along the columns of A), K split over blockIdx.y
// and combined with FP64 atomics
template <int NV>
__global__ void __launch_bounds__(256) k_gemv_n(int M, int K, int kchunk, double alpha,
                                                const double* __restrict__ A, int lda,
                                                const double* __restrict__ B, int ldb,
                                                double* __restrict__ C, int ldc) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  const int k0 = blockIdx.y * kchunk, k1 = min(K, k0 + kchunk);
  if (m >= M) return;
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  constexpr int U = 8;
  int k = k0;
  for (; k + U <= k1; k += U) {
    double av[U];
#pragma unroll
    for (int u = 0; u < U; ++u) av[u] = A[(size_t)(k + u) * lda + m];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] += av[u] * B[(size_t)v * ldb + k + u];
  }
  for (; k < k1; ++k) {
    const double av = A[(size_t)k * lda + m];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += av * B[(size_t)v * ldb + k];
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) atomicAdd(&C[(size_t)v * ldc + m], alpha * acc[v]);
}

template <int NV>
static void launch_gemv_t(Ctx* c, int M, int K, double alpha, const double* A, int lda,
                          const double* B, int ldb, double beta, double* C, int ldc) {
  hipLaunchKernelGGL(k_gemv_t<NV>, dim3(M), dim3(256), 0, c->stream, M, K, alpha, A, lda, B, ldb,
                     beta, C, ldc);
}

template <int NV>
static void launch_gemv_n(Ctx* c, int M, int K, double alpha, const double* A, int lda,
                          const double* B, int ldb, double* C, int ldc) {
  // about 4 workgroups per CU in total
  const int mb = (int)cdiv(M, 256);
  const int ks = std::max(1, std::min(K, 1024 / std::max(1, mb)));
  const int kchunk = (int)cdiv(K, ks);
  hipLaunchKernelGGL(k_gemv_n<NV>, dim3(mb, cdiv(K, kchunk)), dim3(256), 0, c->stream, M, K, kchunk,
                     alpha, A, lda, B, ldb, C, ldc);
}

static int gemv_tn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc) {
  switch (N) {
    case 1: launch_gemv_t<1>(c, M, K, alpha, A, lda, B, ldb, beta, C, ldc); break;
    case 2: launch_gemv_t<2>(c, M, K, alpha, A, lda, B, ldb, beta, C, ldc); break;
    case 3: launch_gemv_t<3>(c, M, K, alpha, A, lda, B, ldb, beta, C, ldc); break;
    default: launch_gemv_t<4>(c, M, K, alpha, A, lda, B, ldb, beta, C, ldc); break;
  }
  return 0;
}

static int gemv_nn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc) {
  if (beta != 1.0)
    hipLaunchKernelGGL(k_scale, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N,
                       ldc, beta);
  switch (N) {
    case 1: launch_gemv_n<1>(c, M, K, alpha, A, lda, B, ldb, C, ldc); break;
    case 2: launch_gemv_n<2>(c, M, K, alpha, A, lda, B, ldb, C, ldc); break;
    case 3: launch_gemv_n<3>(c, M, K, alpha, A, lda, B, ldb, C, ldc); break;
    default: launch_gemv_n<4>(c, M, K, alpha, A, lda, B, ldb, C, ldc); break;
  }
  return 0;
}

// ---- mixed precision: FP64 operands rounded to FP32 while staged into LDS,
// v_mfma_f32_16x16x4f32 products with FP32 accumulation over each K chunk,
// chunks combined in FP64 (split-K atomics or the FP64 C update).
typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int KT32 = 32, LDT32 = BM + 4;

template <bool AKC, bool SPLIT>
__global__ void __launch_bounds__(256) k_gemm_f32(int M, int N, int K, int kchunk, double alpha,
                                                  const double* __restrict__ A, int lda,
                                                  const double* __restrict__ B, int ldb,
                                                  double* __restrict__ C, int ldc) {
  __shared__ float As[KT32][LDT32];
  __shared__ float Bs[KT32][LDT32];
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  // per thread: 8 A values and 8 B values of the 64 x 32 stage
  float ra[8], rb[8];
  auto fetch = [&](int k0) {
    if (AKC) {
      const int m = tid >> 2, kk = (tid & 3) * 8, gm = m0 + m, gmc = min(gm, M - 1);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int gk = k0 + kk + u;
        const double v = A[(size_t)gmc * lda + min(gk, ke - 1)];
        ra[u] = (gm < M && gk < ke) ? (float)v : 0.f;
      }
    } else {
      const int kk = tid >> 3, m = (tid & 7) * 8, gk = k0 + kk, gkc = min(gk, ke - 1);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int gm = m0 + m + u;
        const double v = A[(size_t)gkc * lda + min(gm, M - 1)];
        ra[u] = (gm < M && gk < ke) ? (float)v : 0.f;
      }
    }
    const int n = tid >> 2, kk = (tid & 3) * 8, gn = n0 + n, gnc = min(gn, N - 1);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int gk = k0 + kk + u;
      const double v = B[(size_t)gnc * ldb + min(gk, ke - 1)];
      rb[u] = (gn < N && gk < ke) ? (float)v : 0.f;
    }
  };
  fetch(kb);
  for (int k0 = kb; k0 < ke; k0 += KT32) {
    if (AKC) {
      const int m = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
      for (int u = 0; u < 8; ++u) As[kk + u][m] = ra[u];
    } else {
      const int kk = tid >> 3, m = (tid & 7) * 8;
#pragma unroll
      for (int u = 0; u < 8; ++u) As[kk][m + u] = ra[u];
    }
    {
      const int n = tid >> 2, kk = (tid & 3) * 8;
#pragma unroll
      for (int u = 0; u < 8; ++u) Bs[kk + u][n] = rb[u];
    }
    __syncthreads();
    if (k0 + KT32 < ke) fetch(k0 + KT32);  // next stage in flight during the MFMAs
#pragma unroll
    for (int k4 = 0; k4 < KT32; k4 += 4) {
      const int kr = k4 + (lane >> 4), c = lane & 15;
      const float a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
      const float b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  // v_mfma_f32_16x16x4f32 leaves D(i, j) in lane j + 16*(i/4), register i%4
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + j * 16 + (lane & 15);
      if (n >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + v;
        if (m >= M) continue;
        const double val = alpha * (double)acc[i][j][v];
        if (SPLIT) atomicAdd(&C[(size_t)n * ldc + m], val);
        else C[(size_t)n * ldc + m] += val;
      }
    }
}

// ---- 128 x 128 tiles: four wavefronts of 64 x 64 (4 x 4 MFMA blocks each),
// the next K stage fetched into registers while the current one is on the
// MFMA pipes.  T = double (v_mfma_f64_16x16x4f64, KT 16) or float
// (v_mfma_f32_16x16x4f32 on FP64 operands rounded while staged, KT 32).
template <typename T>
struct BigCfg;
template <>
struct BigCfg<double> {
  static constexpr int KT = 16;
  typedef d4 acc_t;
  __device__ static acc_t mfma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // D(i, j): lane j + 16 (i % 4), register i / 4
  __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct BigCfg<float> {
  static constexpr int KT = 32;
  typedef f4 acc_t;
  __device__ static acc_t mfma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // D(i, j): lane j + 16 (i / 4), register i % 4
  __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

constexpr int BB = 128;

template <typename T, bool AKC, bool SPLIT>
__global__ void __launch_bounds__(256) k_gemm_big(int M, int N, int K, int kchunk, double alpha,
                                                  const double* __restrict__ A, int lda,
                                                  const double* __restrict__ B, int ldb,
                                                  double* __restrict__ C, int ldc) {
  typedef BigCfg<T> Cfg;
  constexpr int KT = Cfg::KT, LD = BB + 4;
  constexpr int PER = BB * KT / 256;  // operand values per thread per stage
  __shared__ T As[KT][LD];
  __shared__ T Bs[KT][LD];
  const int m0 = blockIdx.x * BB, n0 = blockIdx.y * BB;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  typename Cfg::acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[i][j][v] = (T)0;
  T ra[PER], rb[PER];
  // k-contiguous operands: thread -> (row r = tid / (KT/PER), k run of PER)
  constexpr int KRUN = PER, RPT = KT / KRUN;  // threads per row
  auto fetch = [&](int k0) {
    if (AKC) {
      const int r = tid / RPT, kk = (tid % RPT) * KRUN, gm = m0 + r, gmc = min(gm, M - 1);
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int gk = k0 + kk + u;
        const double v = A[(size_t)gmc * lda + min(gk, ke - 1)];
        ra[u] = (gm < M && gk < ke) ? (T)v : (T)0;
      }
    } else {
      // m-contiguous: thread -> (k row, m run of PER)
      const int kk = tid / (BB / PER), m = (tid % (BB / PER)) * PER, gk = k0 + kk;
      const int gkc = min(gk, ke - 1);
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int gm = m0 + m + u;
        const double v = A[(size_t)gkc * lda + min(gm, M - 1)];
        ra[u] = (gm < M && gk < ke) ? (T)v : (T)0;
      }
    }
    const int r = tid / RPT, kk = (tid % RPT) * KRUN, gn = n0 + r, gnc = min(gn, N - 1);
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int gk = k0 + kk + u;
      const double v = B[(size_t)gnc * ldb + min(gk, ke - 1)];
      rb[u] = (gn < N && gk < ke) ? (T)v : (T)0;
    }
  };
  fetch(kb);
  for (int k0 = kb; k0 < ke; k0 += KT) {
    if (AKC) {
      const int r = tid / RPT, kk = (tid % RPT) * KRUN;
#pragma unroll
      for (int u = 0; u < PER; ++u) As[kk + u][r] = ra[u];
    } else {
      const int kk = tid / (BB / PER), m = (tid % (BB / PER)) * PER;
#pragma unroll
      for (int u = 0; u < PER; ++u) As[kk][m + u] = ra[u];
    }
    {
      const int r = tid / RPT, kk = (tid % RPT) * KRUN;
#pragma unroll
      for (int u = 0; u < PER; ++u) Bs[kk + u][r] = rb[u];
    }
    __syncthreads();
    if (k0 + KT < ke) fetch(k0 + KT);
#pragma unroll
    for (int k4 = 0; k4 < KT; k4 += 4) {
      const int kr = k4 + (lane >> 4), c = lane & 15;
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = As[kr][wm + 16 * i + c];
        b[i] = Bs[kr][wn + 16 * i + c];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Cfg::mfma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int n = n0 + wn + j * 16 + (lane & 15);
      if (n >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int m = m0 + wm + i * 16 + Cfg::row(lane, v);
        if (m >= M) continue;
        const double val = alpha * (double)acc[i][j][v];
        if (SPLIT) atomicAdd(&C[(size_t)n * ldc + m], val);
        else C[(size_t)n * ldc + m] += val;
      }
    }
}

// use the 128 x 128 kernels when the output has enough tiles of that size
bool use_big(int M, int N) {
  static const int env = std::getenv("QX_GEMM_SMALL") ? 1 : 0;
  return !env && M >= 256 && N >= 256;
}

template <typename T>
static void launch_big_tn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                          const double* B, int ldb, double* C, int ldc) {
  constexpr int KT = BigCfg<T>::KT;
  const int tiles = (int)(cdiv(M, BB) * cdiv(N, BB));
  int nsplit = std::max<int>(1, std::min<int>((int)cdiv(K, 512), std::max(1, 2048 / tiles)));
  int kchunk = (int)(cdiv(cdiv(K, nsplit), KT) * KT);
  nsplit = (int)cdiv(K, kchunk);
  dim3 grid(cdiv(M, BB), cdiv(N, BB), nsplit);
  if (nsplit > 1)
    hipLaunchKernelGGL((k_gemm_big<T, true, true>), grid, dim3(256), 0, c->stream, M, N, K, kchunk,
                       alpha, A, lda, B, ldb, C, ldc);
  else
    hipLaunchKernelGGL((k_gemm_big<T, true, false>), grid, dim3(256), 0, c->stream, M, N, K,
                       kchunk, alpha, A, lda, B, ldb, C, ldc);
}

template <typename T>
static void launch_big_nn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                          const double* B, int ldb, double* C, int ldc) {
  dim3 grid(cdiv(M, BB), cdiv(N, BB), 1);
  hipLaunchKernelGGL((k_gemm_big<T, false, false>), grid, dim3(256), 0, c->stream, M, N, K, K,
                     alpha, A, lda, B, ldb, C, ldc);
}

// FP32-class contractions (the mixed form): our split bf16x3 MFMA kernels
// for sizable outputs, FP32 MFMA tiles for skinny ones; hipBLASLt only as a
// cross-check (QX_GEMM_LIB=hipblaslt)
int gemm_tn_p(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
              int ldb, double beta, double* C, int ldc, int corr, int prec) {
  if (prec != PREC_FP32) return gemm_tn(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, corr);
  if (M <= 0 || N <= 0) return 0;
  if (use_big(M, N)) {
    if (lt_enabled()) {
      if (lt_sgemm_mixed(c, true, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc) == 0) {
        if (corr)
          hipLaunchKernelGGL(k_corr, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C,
                             M, N, ldc, A, lda, B, ldb, alpha, corr == 2 ? 1 : 0);
        QX_CHECK(hipGetLastError());
        return 0;
      }
    } else if (gemm_path() == GEMM_OWN) {
      const int rc = mg_bf3_tn<double, double>(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                               corr);
      if (rc != -3) return rc;
    }
  }

  hipLaunchKernelGGL(k_scale, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N, ldc,
                     beta);
  lt_mark(c, c->lt_gemm_id, false);
  if (use_big(M, N)) {
    launch_big_tn<float>(c, M, N, K, alpha, A, lda, B, ldb, C, ldc);
  } else {
    const int tiles = (int)(cdiv(M, BM) * cdiv(N, BN));
    int nsplit = std::max<int>(1, std::min<int>((int)cdiv(K, 1024), std::max(1, 2048 / tiles)));
    int kchunk = (int)(cdiv(cdiv(K, nsplit), KT32) * KT32);
    nsplit = (int)cdiv(K, kchunk);
    dim3 grid(cdiv(M, BM), cdiv(N, BN), nsplit);
    if (nsplit > 1)
      hipLaunchKernelGGL((k_gemm_f32<true, true>), grid, dim3(256), 0, c->stream, M, N, K, kchunk,
                         alpha, A, lda, B, ldb, C, ldc);
    else
      hipLaunchKernelGGL((k_gemm_f32<true, false>), grid, dim3(256), 0, c->stream, M, N, K,
                         kchunk, alpha, A, lda, B, ldb, C, ldc);
  }
  lt_mark(c, c->lt_gemm_id, true);
  if (corr)
    hipLaunchKernelGGL(k_corr, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N, ldc,
                       A, lda, B, ldb, alpha, corr == 2 ? 1 : 0);
  QX_CHECK(hipGetLastError());
  return 0;
}

int gemm_nn_p(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
              int ldb, double beta, double* C, int ldc, int prec) {
  if (prec != PREC_FP32) return gemm_nn(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  if (M <= 0 || N <= 0) return 0;
  if (use_big(M, N)) {
    if (lt_enabled()) {
      if (lt_sgemm_mixed(c, false, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc) == 0) return 0;
    } else if (gemm_path() == GEMM_OWN) {
      const int rc = mg_bf3_nn<double, double>(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
      if (rc != -3) return rc;
    }
  }

  if (beta != 1.0)
    hipLaunchKernelGGL(k_scale, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N,
                       ldc, beta);
  lt_mark(c, c->lt_gemm_id, false);
  if (use_big(M, N)) {
    launch_big_nn<float>(c, M, N, K, alpha, A, lda, B, ldb, C, ldc);
  } else {
    dim3 grid(cdiv(M, BM), cdiv(N, BN), 1);
    hipLaunchKernelGGL((k_gemm_f32<false, false>), grid, dim3(256), 0, c->stream, M, N, K, K,
                       alpha, A, lda, B, ldb, C, ldc);
  }
  lt_mark(c, c->lt_gemm_id, true);
  QX_CHECK(hipGetLastError());
  return 0;
}

// FP64 contractions: our 128 x 128 MFMA kernels (mfma_gemm.hip) for sizable
// outputs, 64 x 64 tiles and matrix-vector kernels for skinny ones
int gemm_tn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
            int ldb, double beta, double* C, int ldc, int corr) {
  if (!use_big(M, N)) return gemm_tn_small(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, corr);
  if (lt_enabled()) {
    if (lt_dgemm(c, true, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc) != 0)
      return gemm_tn_small(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, corr);
    if (corr)
      hipLaunchKernelGGL(k_corr, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N,
                         ldc, A, lda, B, ldb, alpha, corr == 2 ? 1 : 0);
    QX_CHECK(hipGetLastError());
    return 0;
  }
  const int rc = mg_dgemm_tn(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, corr);
  if (rc != -3) return rc;
  return gemm_tn_small(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, corr);
}

int gemm_nn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
            int ldb, double beta, double* C, int ldc) {
  if (!use_big(M, N)) return gemm_nn_small(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  if (lt_enabled()) {
    if (lt_dgemm(c, false, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc) != 0)
      return gemm_nn_small(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
    QX_CHECK(hipGetLastError());
    return 0;
  }
  const int rc = mg_dgemm_nn(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  if (rc != -3) return rc;
  return gemm_nn_small(c, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// P[b][p] -> slmir (optional copy) and DP[b][p] = D[p] P[b][p]
__global__ void k_scale_proj(double* P, const double* D, int npj, int nb) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)npj * nb) return;
  P[i] *= D[i % npj];
}

// Kleinman-Bylander apply for nb bands: out += B diag(D) (B^T W psi).
// prec PREC_FP32 runs both contractions on FP32 MFMA (mixed form).
int kb_apply(Ctx* c, const double* psi, double* out, int nb, int, double* slmir_out, int prec) {
  if (c->npj == 0) return 0;
  struct TimeGemms {  // the KB contractions are what launch timer 0 measures
    Ctx* c;
    explicit TimeGemms(Ctx* cc) : c(cc) { c->lt_gemm_id = 0; }
    ~TimeGemms() { c->lt_gemm_id = -1; }
  } tg(c);
  int rc = gemm_tn_p(c, c->npj, nb, c->nplwex, 2.0, c->d_B, c->nplwex, psi, c->nplwex, 0.0,
                     c->d_P, c->npj, 2, prec);
  if (rc) return rc;
  if (slmir_out)
    QX_CHECK(hipMemcpyAsync(slmir_out, c->d_P, sizeof(double) * (size_t)c->npj * nb,
                            hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_scale_proj, dim3(cdiv((long)c->npj * nb, 256)), dim3(256), 0, c->stream,
                     c->d_P, c->d_D, c->npj, nb);
  return gemm_nn_p(c, c->nplwex, nb, c->npj, 1.0, c->d_B, c->nplwex, c->d_P, c->npj, 1.0, out,
                   c->nplwex, prec);
}

// Pd32 = (float) (D (2 P32 - c)) with c the G = 0 correction of the packed
// KB sum (as k_corr with the Im term, alpha 2): the projection scaled for the
// back-projection, from the FP32 product P32 = B^T psi
__global__ void k_proj_fix(const float* __restrict__ P32, const double* __restrict__ B, int ldb,
                           const double* __restrict__ psi, int ldp, const double* __restrict__ D,
                           int npj, int nb, float* __restrict__ Pd32) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)npj * nb) return;
  const int m = (int)(i % npj), n = (int)(i / npj);
  const double c = B[(size_t)m * ldb] * psi[(size_t)n * ldp] +
                   2.0 * B[(size_t)m * ldb + 1] * psi[(size_t)n * ldp + 1];
  Pd32[i] = (float)((2.0 * (double)P32[i] - c) * D[m]);
}

// the same scaled projection written as its bf16 hi | lo split for the
// split bf16x3 back-projection: rows of ldp (zero pads kept), lo at plo; one
// thread per pair of projectors (npj even)
__global__ void k_proj_fix_split(const float* __restrict__ P32, const double* __restrict__ B,
                                 int ldb, const double* __restrict__ psi, int ldp_psi,
                                 const double* __restrict__ D, int npj, int nb,
                                 unsigned short* __restrict__ Pds, int ldp, long plo) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int hp = npj / 2;
  if (i >= (long)hp * nb) return;
  const int m = 2 * (int)(i % hp), n = (int)(i / hp);
  float v[2];
  for (int u = 0; u < 2; ++u) {
    const double c = B[(size_t)(m + u) * ldb] * psi[(size_t)n * ldp_psi] +
                     2.0 * B[(size_t)(m + u) * ldb + 1] * psi[(size_t)n * ldp_psi + 1];
    v[u] = (float)((2.0 * (double)P32[(size_t)n * npj + m + u] - c) * D[m + u]);
  }
  unsigned h, l;
  split_bf16x2(v[0], v[1], h, l);
  *reinterpret_cast<unsigned*>(Pds + (size_t)n * ldp + m) = h;
  *reinterpret_cast<unsigned*>(Pds + plo + (size_t)n * ldp + m) = l;
}

int kb_mixed_f32(Ctx* c, const double* psi, const float* psi32, const unsigned short* psis,
                 int ldk, long slo, int nb, float* kb32) {
  const int n = c->nplwex, npj = c->npj;
  if (psis) {  // split bf16x3 on operands split beforehand
    int ld = 0, rows = 0;
    long tlo = 0;
    const unsigned short* Ts = mg_kb_split(c, &ld, &rows, &tlo);
    float* P32 = mg_f32_scratch(c, 2, (size_t)npj * nb);
    const long plo = (long)c->nbmax * rows;
    unsigned short* Pds = mg_u16_scratch(c, 1, 2 * (size_t)plo);
    if (!Ts || !P32 || !Pds || ld != ldk || (npj & 1)) return -3;
    struct TimeGemms {
      Ctx* c;
      explicit TimeGemms(Ctx* cc) : c(cc) { c->lt_gemm_id = 0; }
      ~TimeGemms() { c->lt_gemm_id = -1; }
    } tg(c);
    int rc = mg_ps3_tn(c, npj, nb, ldk, Ts, ld, tlo, psis, ldk, slo, P32, npj);
    if (rc) return rc;
    hipLaunchKernelGGL(k_proj_fix_split, dim3(cdiv((long)(npj / 2) * nb, 256)), dim3(256), 0,
                       c->stream, P32, c->d_B, n, psi, n, c->d_D, npj, nb, Pds, rows, plo);
    return mg_ps3_nn(c, n, nb, rows, Ts, ld, tlo, Pds, rows, plo, kb32, n);
  }
  const float* B32 = mg_kb_b32(c);
  float* P32 = mg_f32_scratch(c, 2, (size_t)npj * nb);
  float* Pd32 = mg_f32_scratch(c, 3, (size_t)npj * nb);
  if (!B32 || !P32 || !Pd32) return -3;
  struct TimeGemms {  // the KB contractions are what launch timer 0 measures
    Ctx* c;
    explicit TimeGemms(Ctx* cc) : c(cc) { c->lt_gemm_id = 0; }
    ~TimeGemms() { c->lt_gemm_id = -1; }
  } tg(c);
  int rc = lt_enabled() ? lt_sgemm_f32(c, true, npj, nb, n, B32, n, psi32, n, P32, npj)
                        : mg_bf3_tn<float, float>(c, npj, nb, n, 1.0, B32, n, psi32, n, 0.0, P32,
                                                  npj, 0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_proj_fix, dim3(cdiv((long)npj * nb, 256)), dim3(256), 0, c->stream, P32,
                     c->d_B, n, psi, n, c->d_D, npj, nb, Pd32);
  return lt_enabled() ? lt_sgemm_f32(c, false, n, nb, npj, B32, n, Pd32, npj, kb32, n)
                      : mg_bf3_nn<float, float>(c, n, nb, npj, 1.0, B32, n, Pd32, npj, 0.0, kb32,
                                                n);
}

// ---- the mixed band step's products on split band slots ---------------------
// A band slot (nb bands of nplwex FP64) is split into bf16 hi | lo planes:
// band stride round_up(nplwex, 64), round_up(nbmax, 64) bands, zeros in
// every pad, each plane followed by a margin of 128 elements (the NN edge
// tiles read into it).  The overlap, subspace-H, projection and rotation
// products then run the split bf16x3 MFMA kernel on operands it reads at 2 x 2
// bytes per element, split once per producer instead of once per tile.
bool band_split_wanted(Ctx* c, int nb, int prec) {
  static const bool off = std::getenv("QX_BAND_SPLIT_OFF") != nullptr;
  return !off && prec == PREC_FP32 && gemm_path() == GEMM_OWN && !lt_enabled() &&
         use_big(nb, nb);
}

int band_split_buf(Ctx* c, int which, BandSplit* s) {
  const int ld = (int)(cdiv(c->nplwex, 64) * 64);
  const long rows = (long)cdiv(c->nbmax, 64) * 64;
  s->ld = ld;
  s->lo = rows * ld + 128;
  s->p = mg_u16_scratch(c, which, 2 * (size_t)s->lo);
  return s->p ? 0 : -3;
}

int band_split(Ctx* c, int which, const double* src, int nb, BandSplit* s) {
  int rc = band_split_buf(c, which, s);
  if (rc) return rc;
  return mg_split_rows(c, src, nb, c->nplwex, c->nplwex, s->p, s->ld, s->lo);
}

// C(m, n) = alpha (<a_m|b_n>_packed) for m < M, n < N: the split product
// over the whole padded band, then the G = 0 term of the packed sum taken out
// with the FP64 values of the slots (corr = 1 of gemm_tn)
int band_tn(Ctx* c, int M, int N, double alpha, const BandSplit& A, const double* a,
            const BandSplit& B, const double* b, double* C, int ldc) {
  int rc = mg_ps3_tn_d(c, M, N, A.ld, alpha, A.p, A.ld, A.lo, B.p, B.ld, B.lo, 0.0, C, ldc);
  if (rc) return rc;
  const int n = c->nplwex;
  hipLaunchKernelGGL(k_corr, dim3(cdiv((long)M * N, 256)), dim3(256), 0, c->stream, C, M, N, ldc,
                     a, n, b, n, alpha, 0);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

// C(:, j) = alpha sum_k A_k Z(k, j) + beta C(:, j) for j < nb (bands of
// nplwex).  The product reads A's split planes only, so C may be the slot A
// was split from.
int band_nn(Ctx* c, int nb, double alpha, const BandSplit& A, const double* Z, int ldz,
            double beta, double* C) {
  const int kp = (int)(cdiv(nb, 64) * 64);
  const long zlo = (long)nb * kp;
  unsigned short* Zs = mg_u16_scratch(c, 4, 2 * (size_t)zlo);
  if (!Zs) return -3;
  int rc = mg_split_rows(c, Z, nb, nb, ldz, Zs, kp, zlo);
  if (rc) return rc;
  return mg_ps3_nn_d(c, c->nplwex, nb, kp, alpha, A.p, A.ld, A.lo, Zs, kp, zlo, beta, C,
                     c->nplwex);
}

}  // namespace qx
