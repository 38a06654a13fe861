// mfma_gemm.hip -- the large contractions of the band algebra on our own
// MFMA kernels (gfx950), in the two shapes of gemm.hip (column-major C):
//   TN : C(m, n) = alpha sum_k A[m*lda + k] B[n*ldb + k] + beta C   (K = plane waves)
//        KB projection P = B^T psi (Sources/nlkbpp.f90:1358-1364), overlap
//        S = psi^T psi (Sources/schmidt.f90:199-216), subspace H
//        (Sources/engrad.f90:869-882), CG projections (Sources/engrad.f90:1122-1142)
//   NN : C(m, n) = alpha sum_k A[k*lda + m] B[n*ldb + k] + beta C   (M = plane waves)
//        back-projection H psi += B D P (Sources/nlkbpp.f90:1376-1379), rotation
//        psi U (Sources/eigen.f90:1485-1509), psi L^-T (Sources/schmidt.f90:284-313).
//
// Tiles: 128 x 128 outputs per 256-thread workgroup, four wavefronts of
// 64 x 64.  FP64: a 4 x 4 grid of v_mfma_f64_16x16x4f64 blocks per wavefront,
// K staged 16 at a time through double-buffered LDS (one barrier per stage),
// the next stage's global loads in flight during the current stage's MFMAs.
//
// Split bf16x3 (the mixed form's FP32-class products): every FP32 (or FP64)
// operand x is split while staged into x_hi = bf16(x) and x_lo = bf16(x - x_hi);
// a product is x_lo y_hi + x_hi y_lo + x_hi y_hi on v_mfma_f32_16x16x32_bf16
// with FP32 accumulation (relative error of order 2^-16 per product, the
// dropped x_lo y_lo term).  K staged 32 at a time.
//
// Work distribution (stream-K): the shapes of the path have few output tiles
// (35 for the KB projection of the 216-atom cell) and long K, or many tiles
// and short K, so neither a tile per workgroup nor a fixed K split fills 256
// CUs evenly.  A persistent grid of G workgroups splits the (tile, K stage)
// iteration space into G equal contiguous ranges; a workgroup accumulates
// each tile segment of its range, writes a tile it covered whole straight to
// C (alpha, beta and the packed-sum correction applied), and leaves the
// partial first / last segments of its range in a workspace; one fix-up pass
// sums the partials of each shared tile in workgroup order (deterministic).
// Workgroups are numbered so that each XCD (dispatch round-robin over 8)
// works a contiguous stretch of the iteration space, sharing its L2.  G is
// chosen per shape by timing the candidates on first use (into a scratch
// output) and cached.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <tuple>
#include <vector>
#include "context.h"

namespace qx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int TE = 128;                    // output tile edge
constexpr int KT64 = 16, LP64 = TE + 4;    // FP64 K stage, LDS row pitch (doubles)
constexpr int KT16 = 32, LP16 = KT16 + 8;  // bf16 K stage, LDS row pitch (bf16)
constexpr int kXcd = 8;                    // XCDs the dispatcher round-robins over

__host__ __device__ inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

// the stream-K decomposition of one product
struct Sk {
  int M, N, K;
  int tiles_m, nk;  // tiles along M; K stages per tile
  long L;           // iterations (tile, K stage) per workgroup
  int G;            // workgroups
  double alpha, beta;
  int ldc, corr;
};

// logical workgroup: XCD x = blockIdx % 8 takes the x-th contiguous eighth
__device__ inline int logical_wg(int G) {
  const int b = (int)blockIdx.x;
  return (b % kXcd) * (G / kXcd) + b / kXcd;
}

// packed-sum correction of C(m, n) (TN only): 1 -> 0.5 A0 B0, 2 -> + A1 B1
template <typename T>
__device__ inline double corr_term(int corr, const T* A, int lda, const T* B, int ldb, int m,
                                   int n) {
  if (!corr) return 0.0;
  double cc = (double)A[(size_t)m * lda] * (double)B[(size_t)n * ldb];
  if (corr == 2) cc += 2.0 * (double)A[(size_t)m * lda + 1] * (double)B[(size_t)n * ldb + 1];
  return 0.5 * cc;
}

// ---------------------------------------------------------------- FP64 ----
template <bool AK>
__global__ void __launch_bounds__(256, 2)
    k_mg_f64(Sk sk, const double* __restrict__ A, int lda, const double* __restrict__ B, int ldb,
             double* __restrict__ C, double* __restrict__ Wp) {
  __shared__ double As[2][KT64][LP64];
  __shared__ double Bs[2][KT64][LP64];
  const int M = sk.M, N = sk.N, K = sk.K;
  const int lb = logical_wg(sk.G);
  const long I = (long)sk.tiles_m * cdivu(N, TE) * sk.nk;
  const long g0 = lb * sk.L, g1 = min(I, g0 + sk.L);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  // k-contiguous operand: thread -> row tid >> 1, k run (tid & 1) * 8
  const int kr_r = tid >> 1, kr_k = (tid & 1) * 8;
  // m-contiguous operand: thread -> k row tid >> 4, m run (tid & 15) * 8
  const int mr_k = tid >> 4, mr_m = (tid & 15) * 8;
  for (long g = g0; g < g1;) {
    const int t = (int)(g / sk.nk);
    const int s0 = (int)(g - (long)t * sk.nk);
    const int s1 = (int)min((long)sk.nk, g1 - (long)t * sk.nk);
    const int m0 = (t % sk.tiles_m) * TE, n0 = (t / sk.tiles_m) * TE;
    const int kb = s0 * KT64, ke = min(K, s1 * KT64);
    d4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[8], rb[8];
    const double* Ar = A + (AK ? (size_t)min(m0 + kr_r, M - 1) * lda : 0);
    const double* Br = B + (size_t)min(n0 + kr_r, N - 1) * ldb;
    auto fetch = [&](int k0) {
      if (k0 + KT64 <= ke) {
        if (AK) {
          const double2* p = reinterpret_cast<const double2*>(Ar + k0 + kr_k);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double2 v = p[u];
            ra[2 * u] = v.x;
            ra[2 * u + 1] = v.y;
          }
        } else {
          const int gm = m0 + mr_m;
          const double* p = A + (size_t)(k0 + mr_k) * lda + gm;
          if (gm + 8 <= M) {
            const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const double2 v = q[u];
              ra[2 * u] = v.x;
              ra[2 * u + 1] = v.y;
            }
          } else {
#pragma unroll
            for (int u = 0; u < 8; ++u) ra[u] = (gm + u < M) ? p[u] : 0.0;
          }
        }
        const double2* q = reinterpret_cast<const double2*>(Br + k0 + kr_k);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double2 v = q[u];
          rb[2 * u] = v.x;
          rb[2 * u + 1] = v.y;
        }
      } else {  // the last, partial K stage
        if (AK) {
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int k = k0 + kr_k + u;
            ra[u] = k < ke ? Ar[k] : 0.0;
          }
        } else {
          const int gm = m0 + mr_m, k = k0 + mr_k;
          const double* p = A + (size_t)min(k, ke - 1) * lda;
#pragma unroll
          for (int u = 0; u < 8; ++u) ra[u] = (k < ke && gm + u < M) ? p[gm + u] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int k = k0 + kr_k + u;
          rb[u] = k < ke ? Br[k] : 0.0;
        }
      }
    };
    auto stash = [&](int buf) {
      if (AK) {
#pragma unroll
        for (int u = 0; u < 8; ++u) As[buf][kr_k + u][kr_r] = ra[u];
      } else {
#pragma unroll
        for (int u = 0; u < 8; ++u) As[buf][mr_k][mr_m + u] = ra[u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) Bs[buf][kr_k + u][kr_r] = rb[u];
    };
    int buf = 0;
    fetch(kb);
    stash(0);
    __syncthreads();
    for (int k0 = kb; k0 < ke; k0 += KT64) {
      const bool more = k0 + KT64 < ke;
      if (more) fetch(k0 + KT64);
#pragma unroll
      for (int k4 = 0; k4 < KT64; k4 += 4) {
        const int kr = k4 + (lane >> 4), c = lane & 15;
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = As[buf][kr][wm + 16 * i + c];
          b[i] = Bs[buf][kr][wn + 16 * i + c];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
      if (more) stash(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
    // v_mfma_f64_16x16x4f64 leaves D(i, j) in lane j + 16 (i % 4), register i / 4
    const bool whole = s0 == 0 && s1 == sk.nk;
    double* P = Wp + ((size_t)lb * 2 + (g == g0 ? 0 : 1)) * TE * TE;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int nl = wn + 16 * j + (lane & 15), n = n0 + nl;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ml = wm + 16 * i + (lane >> 4) + 4 * v, m = m0 + ml;
          if (!whole) {
            P[nl * TE + ml] = acc[i][j][v];
          } else if (m < M && n < N) {
            double* p = C + (size_t)n * sk.ldc + m;
            const double x = acc[i][j][v] - corr_term(sk.corr, A, lda, B, ldb, m, n);
            *p = sk.alpha * x + (sk.beta == 0.0 ? 0.0 : sk.beta * *p);
          }
        }
      }
    g = (long)t * sk.nk + s1;
  }
}

// fix-up of the tiles shared by workgroups: C = alpha (sum of the partials in
// workgroup order - corr) + beta C.  One thread per tile element; a tile that
// one workgroup covered whole was written by it already.
template <typename TW, typename T, typename TC>
__global__ void k_mg_fixup(Sk sk, const TW* __restrict__ Wp, const T* __restrict__ A, int lda,
                           const T* __restrict__ B, int ldb, TC* __restrict__ C) {
  const int t = (int)blockIdx.y;
  const long lo = (long)t * sk.nk, hi = lo + sk.nk - 1;
  const int b_lo = (int)(lo / sk.L), b_hi = (int)(hi / sk.L);
  if (b_lo == b_hi) return;
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= TE * TE) return;
  const int ml = e % TE, nl = e / TE;
  const int m = (t % sk.tiles_m) * TE + ml, n = (t / sk.tiles_m) * TE + nl;
  if (m >= sk.M || n >= sk.N) return;
  double s = 0.0;
  for (int b = b_lo; b <= b_hi; ++b) {
    const int first = (int)((long)b * sk.L / sk.nk);  // the workgroup's first tile
    const size_t slot = (size_t)b * 2 + (first == t ? 0 : 1);
    s += (double)Wp[slot * TE * TE + (size_t)nl * TE + ml];
  }
  s -= corr_term(sk.corr, A, lda, B, ldb, m, n);
  TC* p = C + (size_t)n * sk.ldc + m;
  *p = (TC)(sk.alpha * s + (sk.beta == 0.0 ? 0.0 : sk.beta * (double)*p));
}

// ------------------------------------------------------------ bf16x3 ------
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

// two values split at once: hi = bf16(x) (round to nearest even, one packed
// conversion for the pair), lo = bf16(x - hi); returned as packed bf16 pairs
__device__ inline void split2(float a, float b, unsigned& hi, unsigned& lo) {
  const bf2 h = {(__bf16)a, (__bf16)b};
  const bf2 l = {(__bf16)(a - (float)h[0]), (__bf16)(b - (float)h[1])};
  hi = __builtin_bit_cast(unsigned, h);
  lo = __builtin_bit_cast(unsigned, l);
}

// 16 consecutive elements of a k-contiguous row (k0 + 16 <= end), as floats
template <typename T>
__device__ inline void load16(const T* p, float* r);
template <>
__device__ inline void load16<float>(const float* p, float* r) {
  const float2* q = reinterpret_cast<const float2*>(p);
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const float2 v = q[u];
    r[2 * u] = v.x;
    r[2 * u + 1] = v.y;
  }
}
template <>
__device__ inline void load16<double>(const double* p, float* r) {
  const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const double2 v = q[u];
    r[2 * u] = (float)v.x;
    r[2 * u + 1] = (float)v.y;
  }
}
// 4 consecutive elements (m-contiguous row), as floats
template <typename T>
__device__ inline void load4(const T* p, float* r);
template <>
__device__ inline void load4<float>(const float* p, float* r) {
  const float2* q = reinterpret_cast<const float2*>(p);
  const float2 a = q[0], b = q[1];
  r[0] = a.x; r[1] = a.y; r[2] = b.x; r[3] = b.y;
}
template <>
__device__ inline void load4<double>(const double* p, float* r) {
  const double2* q = reinterpret_cast<const double2*>(p);
  const double2 a = q[0], b = q[1];
  r[0] = (float)a.x; r[1] = (float)a.y; r[2] = (float)b.x; r[3] = (float)b.y;
}

// AK: TN (A rows k-contiguous), else NN (A m-contiguous); T the operands'
// type, TC the output's; partial segments go to Wp as floats
template <typename T, typename TC, bool AK>
__global__ void __launch_bounds__(256, 2)
    k_mg_bf3(Sk sk, const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb,
             TC* __restrict__ C, float* __restrict__ Wp) {
  __shared__ unsigned short Ah[TE][LP16], Al[TE][LP16], Bh[TE][LP16], Bl[TE][LP16];
  const int M = sk.M, N = sk.N, K = sk.K;
  const int lb = logical_wg(sk.G);
  const long I = (long)sk.tiles_m * cdivu(N, TE) * sk.nk;
  const long g0 = lb * sk.L, g1 = min(I, g0 + sk.L);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  // k-contiguous rows: thread -> row tid >> 1, k run (tid & 1) * 16
  const int lr = tid >> 1, lk = (tid & 1) * 16;
  // m-contiguous A (NN): thread -> m run (tid & 31) * 4, k rows (tid >> 5) * 4 .. + 3
  const int am = (tid & 31) * 4, ak = (tid >> 5) * 4;
  for (long g = g0; g < g1;) {
    const int t = (int)(g / sk.nk);
    const int s0 = (int)(g - (long)t * sk.nk);
    const int s1 = (int)min((long)sk.nk, g1 - (long)t * sk.nk);
    const int m0 = (t % sk.tiles_m) * TE, n0 = (t / sk.tiles_m) * TE;
    const int kb = s0 * KT16, ke = min(K, s1 * KT16);
    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const T* Ar = A + (AK ? (size_t)min(m0 + lr, M - 1) * lda : 0);
    const T* Br = B + (size_t)min(n0 + lr, N - 1) * ldb;
    float ra[16], rb[16];
    auto fetch = [&](int k0) {
      const bool full = k0 + KT16 <= ke;
      if (AK) {
        if (full) {
          load16<T>(Ar + k0 + lk, ra);
        } else {
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            const int k = k0 + lk + u;
            ra[u] = k < ke ? (float)Ar[k] : 0.f;
          }
        }
      } else {
        const int gm = m0 + am;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = k0 + ak + r;
          const T* p = A + (size_t)min(k, ke - 1) * lda + gm;
          if (k < ke && gm + 4 <= M) {
            load4<T>(p, ra + 4 * r);
          } else {
#pragma unroll
            for (int u = 0; u < 4; ++u)
              ra[4 * r + u] = (k < ke && gm + u < M) ? (float)p[u] : 0.f;
          }
        }
      }
      if (full) {
        load16<T>(Br + k0 + lk, rb);
      } else {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int k = k0 + lk + u;
          rb[u] = k < ke ? (float)Br[k] : 0.f;
        }
      }
    };
    // 16 k-consecutive values -> 8 packed pairs each of hi and lo, as two
    // 16-byte rows of 8 bf16 per part
    auto split16 = [](const float* v, uint4& h0, uint4& h1, uint4& l0, uint4& l1) {
      unsigned h[8], l[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) split2(v[2 * q], v[2 * q + 1], h[q], l[q]);
      h0 = make_uint4(h[0], h[1], h[2], h[3]);
      h1 = make_uint4(h[4], h[5], h[6], h[7]);
      l0 = make_uint4(l[0], l[1], l[2], l[3]);
      l1 = make_uint4(l[4], l[5], l[6], l[7]);
    };
    auto stash = [&]() {
      uint4 h0, h1, l0, l1;
      if (AK) {
        split16(ra, h0, h1, l0, l1);
        *reinterpret_cast<uint4*>(&Ah[lr][lk]) = h0;
        *reinterpret_cast<uint4*>(&Ah[lr][lk + 8]) = h1;
        *reinterpret_cast<uint4*>(&Al[lr][lk]) = l0;
        *reinterpret_cast<uint4*>(&Al[lr][lk + 8]) = l1;
      } else {
        // ra[4 r + u] = A(k = ak + r, m = am + u): per m, 4 consecutive k
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          unsigned h[2], l[2];
          split2(ra[u], ra[4 + u], h[0], l[0]);
          split2(ra[8 + u], ra[12 + u], h[1], l[1]);
          *reinterpret_cast<uint2*>(&Ah[am + u][ak]) = make_uint2(h[0], h[1]);
          *reinterpret_cast<uint2*>(&Al[am + u][ak]) = make_uint2(l[0], l[1]);
        }
      }
      split16(rb, h0, h1, l0, l1);
      *reinterpret_cast<uint4*>(&Bh[lr][lk]) = h0;
      *reinterpret_cast<uint4*>(&Bh[lr][lk + 8]) = h1;
      *reinterpret_cast<uint4*>(&Bl[lr][lk]) = l0;
      *reinterpret_cast<uint4*>(&Bl[lr][lk + 8]) = l1;
    };
    fetch(kb);
    for (int k0 = kb; k0 < ke; k0 += KT16) {
      stash();
      __syncthreads();
      if (k0 + KT16 < ke) fetch(k0 + KT16);
      // v_mfma_f32_16x16x32_bf16 operands: lane l holds row l % 16, k 8 (l / 16) .. + 7
      const int r = lane & 15, kk = 8 * (lane >> 4);
      bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ah[i] = *reinterpret_cast<const bf8*>(&Ah[wm + 16 * i + r][kk]);
        al[i] = *reinterpret_cast<const bf8*>(&Al[wm + 16 * i + r][kk]);
        bh[i] = *reinterpret_cast<const bf8*>(&Bh[wn + 16 * i + r][kk]);
        bl[i] = *reinterpret_cast<const bf8*>(&Bl[wn + 16 * i + r][kk]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
      __syncthreads();
    }
    // v_mfma_f32_16x16x32_bf16 leaves D(i, j) in lane j + 16 (i / 4), register i % 4
    const bool whole = s0 == 0 && s1 == sk.nk;
    float* P = Wp + ((size_t)lb * 2 + (g == g0 ? 0 : 1)) * TE * TE;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int nl = wn + 16 * j + (lane & 15), n = n0 + nl;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ml = wm + 16 * i + 4 * (lane >> 4) + v, m = m0 + ml;
          if (!whole) {
            P[nl * TE + ml] = acc[i][j][v];
          } else if (m < M && n < N) {
            TC* p = C + (size_t)n * sk.ldc + m;
            const double x = (double)acc[i][j][v] - corr_term(sk.corr, A, lda, B, ldb, m, n);
            *p = (TC)(sk.alpha * x + (sk.beta == 0.0 ? 0.0 : sk.beta * (double)*p));
          }
        }
      }
    g = (long)t * sk.nk + s1;
  }
}

// ------------------------------------------------------------ host side ---
struct MgState {
  void* ws = nullptr;  // partial tiles of the stream-K workgroups
  size_t ws_bytes = 0;
  // tuned workgroup count per (kind, M, N, K, lda, ldb)
  std::map<std::tuple<int, int, int, int, int, int>, int> grid;
  float* f32[4] = {};  // FP32 work buffers of the fused mixed band-apply
  size_t f32_len[4] = {};
  float* B32 = nullptr;  // FP32 copy of the KB table
  const double* B32_src = nullptr;
  size_t B32_len = 0;
};

__global__ void k_to_f32(const double* __restrict__ x, float* __restrict__ y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

MgState* state(Ctx* c) {
  if (!c->mg) c->mg = new MgState();
  return static_cast<MgState*>(c->mg);
}

void* workspace(Ctx* c, size_t bytes) {
  MgState* s = state(c);
  if (s->ws_bytes >= bytes) return s->ws;
  buffers_changed(c);
  if (s->ws) (void)hipFree(s->ws);
  s->ws = nullptr;
  s->ws_bytes = 0;
  if (hipMalloc(&s->ws, bytes) != hipSuccess) return nullptr;
  s->ws_bytes = bytes;
  return s->ws;
}

Sk make_sk(int M, int N, int K, int kt, int G, double alpha, double beta, int ldc, int corr) {
  Sk sk;
  sk.M = M;
  sk.N = N;
  sk.K = K;
  sk.tiles_m = (int)cdivu(M, TE);
  sk.nk = (int)cdivu(K, kt);
  const long I = (long)sk.tiles_m * cdivu(N, TE) * sk.nk;
  sk.G = G;
  sk.L = (long)cdivu(I, G);
  sk.alpha = alpha;
  sk.beta = beta;
  sk.ldc = ldc;
  sk.corr = corr;
  return sk;
}

bool tuning_off() {
  static const bool off = std::getenv("QX_GEMM_NOTUNE") != nullptr;
  return off;
}

// workgroup count for a product: candidates of 1 .. per_cu_max workgroups
// per CU, timed on first use of a shape with `trial(G)` (which writes a
// scratch output), then cached
template <typename F>
int choose_grid(Ctx* c, int kind, int M, int N, int K, int lda, int ldb, int per_cu_max,
                size_t part_elem, F trial) {
  MgState* s = state(c);
  const auto key = std::make_tuple(kind, M, N, K, lda, ldb);
  auto it = s->grid.find(key);
  if (it != s->grid.end()) return it->second;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->dev);
  cus = std::max(kXcd, cus / kXcd * kXcd);
  int best_g = 2 * cus;
  if (!tuning_off()) {
    hipEvent_t e0, e1;
    QX_CHECK(hipEventCreate(&e0));
    QX_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    for (int k = 1; k <= per_cu_max; ++k) {
      const int G = k * cus;
      if (!workspace(c, part_elem * 2 * TE * TE * (size_t)G)) return -3;
      if (trial(G)) return -3;  // warm
      QX_CHECK(hipEventRecord(e0, c->stream));
      for (int r = 0; r < 2; ++r)
        if (trial(G)) return -3;
      QX_CHECK(hipEventRecord(e1, c->stream));
      QX_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      QX_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_g = G;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  s->grid[key] = best_g;
  return best_g;
}

// scratch output of a tuning trial (M x N of TC), released after the trials
template <typename TC>
struct TrialOut {
  TC* p = nullptr;
  explicit TrialOut(size_t n) {
    if (hipMalloc(&p, sizeof(TC) * n) != hipSuccess) p = nullptr;
  }
  ~TrialOut() {
    if (p) (void)hipFree(p);
  }
};

template <typename TW, typename T, typename TC, typename K1>
int launch_sk(Ctx* c, const Sk& sk, K1 kernel, const T* A, int lda, const T* B, int ldb, TC* C) {
  TW* W = static_cast<TW*>(state(c)->ws);
  hipLaunchKernelGGL(kernel, dim3(sk.G), dim3(256), 0, c->stream, sk, A, lda, B, ldb, C, W);
  const int tiles = sk.tiles_m * (int)cdivu(sk.N, TE);
  hipLaunchKernelGGL((k_mg_fixup<TW, T, TC>), dim3(TE * TE / 256, tiles), dim3(256), 0,
                     c->stream, sk, W, A, lda, B, ldb, C);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

template <typename TW, typename T, typename TC, typename K1>
int run_product(Ctx* c, int kind, int kt, int per_cu_max, K1 kernel, int M, int N, int K,
                double alpha, const T* A, int lda, const T* B, int ldb, double beta, TC* C,
                int ldc, int corr) {
  if (M <= 0 || N <= 0) return 0;
  if ((lda | ldb) & 1) return -3;  // 8 / 16-byte row loads need even leading dimensions
  MgState* s = state(c);
  const bool known = s->grid.count(std::make_tuple(kind, M, N, K, lda, ldb)) != 0;
  int G;
  if (known || tuning_off()) {
    G = choose_grid(c, kind, M, N, K, lda, ldb, per_cu_max, sizeof(TW), [](int) { return 0; });
  } else {
    TrialOut<TC> out((size_t)M * N);
    if (!out.p) return -3;
    G = choose_grid(c, kind, M, N, K, lda, ldb, per_cu_max, sizeof(TW), [&](int g) {
      const Sk tk = make_sk(M, N, K, kt, g, alpha, 0.0, M, corr);
      return launch_sk<TW>(c, tk, kernel, A, lda, B, ldb, out.p);
    });
    QX_CHECK(hipStreamSynchronize(c->stream));  // trials done before the scratch goes
  }
  if (G < 0) return G;
  if (!workspace(c, sizeof(TW) * 2 * TE * TE * (size_t)G)) return -3;
  const Sk sk = make_sk(M, N, K, kt, G, alpha, beta, ldc, corr);
  lt_mark(c, c->lt_gemm_id, false);
  const int rc = launch_sk<TW>(c, sk, kernel, A, lda, B, ldb, C);
  lt_mark(c, c->lt_gemm_id, true);
  return rc;
}

}  // namespace

void mg_destroy(Ctx* c) {
  if (!c->mg) return;
  MgState* s = static_cast<MgState*>(c->mg);
  if (s->ws) (void)hipFree(s->ws);
  for (float* p : s->f32)
    if (p) (void)hipFree(p);
  if (s->B32) (void)hipFree(s->B32);
  delete s;
  c->mg = nullptr;
}

// the KB table changed in place: its FP32 copy is made again on next use
void mg_forget_kb(Ctx* c) {
  if (c->mg) state(c)->B32_src = nullptr;
}

const float* mg_kb_b32(Ctx* c) {
  if (!c->d_B) return nullptr;
  MgState* s = state(c);
  const size_t len = (size_t)c->nplwex * c->npj;
  if (s->B32_src != c->d_B || s->B32_len != len || !s->B32) {
    buffers_changed(c);
    if (s->B32_len != len || !s->B32) {
      if (s->B32) (void)hipFree(s->B32);
      s->B32 = nullptr;
      s->B32_len = 0;
      if (hipMalloc(&s->B32, sizeof(float) * len) != hipSuccess) return nullptr;
    }
    hipLaunchKernelGGL(k_to_f32, dim3(cdivu((long)len, 256)), dim3(256), 0, c->stream, c->d_B,
                       s->B32, (long)len);
    s->B32_src = c->d_B;
    s->B32_len = len;
  }
  return s->B32;
}

float* mg_f32_scratch(Ctx* c, int i, size_t n) {
  if (i < 0 || i >= 4) return nullptr;
  MgState* s = state(c);
  if (s->f32_len[i] < n) {
    buffers_changed(c);
    if (s->f32[i]) (void)hipFree(s->f32[i]);
    s->f32[i] = nullptr;
    s->f32_len[i] = 0;
    if (hipMalloc(&s->f32[i], sizeof(float) * n) != hipSuccess) return nullptr;
    s->f32_len[i] = n;
  }
  return s->f32[i];
}

int mg_tuned_shapes(Ctx* c) { return c->mg ? (int)state(c)->grid.size() : 0; }

// FP64 TN: C = alpha (A^T B - corr) + beta C
int mg_dgemm_tn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc, int corr) {
  return run_product<double>(c, 0, KT64, 2, k_mg_f64<true>, M, N, K, alpha, A, lda, B, ldb, beta,
                             C, ldc, corr);
}

// FP64 NN: C = alpha A B + beta C
int mg_dgemm_nn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc) {
  return run_product<double>(c, 1, KT64, 2, k_mg_f64<false>, M, N, K, alpha, A, lda, B, ldb,
                             beta, C, ldc, 0);
}

// split bf16x3 TN on T operands into TC: C = alpha (A^T B - corr) + beta C
template <typename T, typename TC>
int mg_bf3_tn(Ctx* c, int M, int N, int K, double alpha, const T* A, int lda, const T* B, int ldb,
              double beta, TC* C, int ldc, int corr) {
  return run_product<float>(c, 2 + (sizeof(T) == 8 ? 1 : 0), KT16, 2, k_mg_bf3<T, TC, true>, M,
                            N, K, alpha, A, lda, B, ldb, beta, C, ldc, corr);
}

// split bf16x3 NN on T operands into TC: C = alpha A B + beta C
template <typename T, typename TC>
int mg_bf3_nn(Ctx* c, int M, int N, int K, double alpha, const T* A, int lda, const T* B, int ldb,
              double beta, TC* C, int ldc) {
  return run_product<float>(c, 4 + (sizeof(T) == 8 ? 1 : 0), KT16, 2, k_mg_bf3<T, TC, false>, M,
                            N, K, alpha, A, lda, B, ldb, beta, C, ldc, 0);
}

template int mg_bf3_tn<double, double>(Ctx*, int, int, int, double, const double*, int,
                                       const double*, int, double, double*, int, int);
template int mg_bf3_tn<float, float>(Ctx*, int, int, int, double, const float*, int, const float*,
                                     int, double, float*, int, int);
template int mg_bf3_nn<double, double>(Ctx*, int, int, int, double, const double*, int,
                                       const double*, int, double, double*, int);
template int mg_bf3_nn<float, float>(Ctx*, int, int, int, double, const float*, int, const float*,
                                     int, double, float*, int);

}  // namespace qx
