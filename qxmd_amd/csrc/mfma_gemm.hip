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
// works a contiguous stretch of the iteration space, sharing its L2.
//
// K-sliced alternative: one workgroup per (K-slice, tile), slices numbered
// slowest, so an XCD's workgroups sweep the same K-slice of every tile in
// step and fetch each operand stage once into its L2 (stream-K fetched it
// once per tile); the fix-up sums the slices in order.  With one slice it is
// one workgroup per tile and no fix-up.  Per shape the stream-K grid or the
// slice count is chosen by timing the candidates on first use (into a
// scratch output) and cached.
#include <algorithm>
#include <atomic>
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

// the work decomposition of one product: stream-K (S == 0) or K-sliced (S > 0)
struct Sk {
  int M, N, K;
  int tiles_m, tiles_n, nk;  // tiles along M and N; K stages per tile
  long L;                    // stream-K: iterations (tile, K stage) per workgroup
  int S, nks;                // K-sliced: slices, K stages per slice
  int nfast;                 // tiles numbered with the N index fastest (NN products)
  int G;                     // workgroups
  double alpha, beta;
  int ldc, corr;
  long alo, blo;  // pre-split operands: element offset of the lo half after the hi half
  int tm, tn;     // output tile (TE x TE, or 96 along an edge for the FP64 kernels)
};

// logical workgroup: XCD x = blockIdx % 8 takes the x-th contiguous eighth
__device__ inline int logical_wg(int G) {
  const int b = (int)blockIdx.x;
  return (b % kXcd) * (G / kXcd) + b / kXcd;
}

// The (tile, K stage) iterations of logical workgroup lb, [g0, g1) with
// iteration g = tile * nk + stage.  Stream-K: a contiguous range of L
// iterations, tile-major.  K-sliced: the single segment of tile lb % T over
// slice lb / T, stages [s nks, min(nk, (s + 1) nks)); the slices are
// numbered slowest, so the workgroups of one XCD (a contiguous stretch of
// logical workgroups) run the same K-slice of every tile at the same time and
// share each K stage of both operands in their L2.  Empty for workgroups past
// the last non-empty slice.
__device__ inline void sk_range(const Sk& sk, int lb, long& g0, long& g1) {
  const int T = sk.tiles_m * sk.tiles_n;
  if (sk.S > 0) {
    const int s = lb / T, t = lb - s * T;
    const int a = s * sk.nks, b = min(sk.nk, a + sk.nks);
    if (s >= sk.S || a >= b) {
      g0 = g1 = 0;
      return;
    }
    g0 = (long)t * sk.nk + a;
    g1 = (long)t * sk.nk + b;
  } else {
    g0 = (long)lb * sk.L;
    g1 = min((long)T * sk.nk, g0 + sk.L);
  }
}

// origin of output tile t: M index fastest, or N index fastest (NN products:
// consecutive tiles share the rows of the long operand A)
__device__ inline void tile_origin(const Sk& sk, int t, int& m0, int& n0) {
  if (sk.nfast) {
    m0 = (t / sk.tiles_n) * sk.tm;
    n0 = (t % sk.tiles_n) * sk.tn;
  } else {
    m0 = (t % sk.tiles_m) * sk.tm;
    n0 = (t / sk.tiles_m) * sk.tn;
  }
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
// Output tile 32 BI x 32 BJ (BI, BJ = 4 or 3: 128 or 96 along each edge, so
// that 540 bands or 864 projectors waste little of the last tile), four
// wavefronts of 16 BI x 16 BJ in a 2 x 2 layout.  Operands move in 16-byte
// chunks: a k-contiguous row of the K stage is 8 chunks (8 threads read 128
// consecutive bytes), an m-contiguous k row is TM / 2 chunks.
template <bool AK, int BI, int BJ>
__global__ void __launch_bounds__(256, 2)
    k_mg_f64(Sk sk, const double* __restrict__ A, int lda, const double* __restrict__ B, int ldb,
             double* __restrict__ C, double* __restrict__ Wp) {
  constexpr int TM = 32 * BI, TN = 32 * BJ;
  constexpr int LPM = TM + 4, LPN = TN + 4;
  constexpr int CA = TM * KT64 / 2 / 256, CB = TN * KT64 / 2 / 256;  // chunks per thread
  static_assert(TM * KT64 / 2 % 256 == 0 && TN * KT64 / 2 % 256 == 0, "chunk split");
  __shared__ double As[2][KT64][LPM];
  __shared__ double Bs[2][KT64][LPN];
  const int M = sk.M, N = sk.N, K = sk.K;
  const int lb = logical_wg(sk.G);
  long g0, g1;
  sk_range(sk, lb, g0, g1);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * (TM / 2), wn = (wave >> 1) * (TN / 2);
  for (long g = g0; g < g1;) {
    const int t = (int)(g / sk.nk);
    const int s0 = (int)(g - (long)t * sk.nk);
    const int s1 = (int)min((long)sk.nk, g1 - (long)t * sk.nk);
    int m0, n0;
    tile_origin(sk, t, m0, n0);
    const int kb = s0 * KT64, ke = min(K, s1 * KT64);
    d4 acc[BI][BJ];
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    double2 ra[CA], rb[CB];
    // chunk u of this thread: k-contiguous -> row c / 8, k 2 (c % 8);
    // m-contiguous -> k row c / (TM / 2), m 2 (c % (TM / 2))
    auto fetch = [&](int k0) {
      const bool full = k0 + KT64 <= ke;
#pragma unroll
      for (int u = 0; u < CA; ++u) {
        const int c = tid + 256 * u;
        if (AK) {
          const int row = c >> 3, k = k0 + 2 * (c & 7);
          const double* p = A + (size_t)min(m0 + row, M - 1) * lda + k;
          if (full) {
            ra[u] = *reinterpret_cast<const double2*>(p);
          } else {
            ra[u].x = k < ke ? p[0] : 0.0;
            ra[u].y = k + 1 < ke ? p[1] : 0.0;
          }
        } else {
          const int kr = c / (TM / 2), m = m0 + 2 * (c % (TM / 2)), k = k0 + kr;
          const double* p = A + (size_t)min(k, K - 1) * lda + m;
          if (k < ke && m + 2 <= M) {
            ra[u] = *reinterpret_cast<const double2*>(p);
          } else {
            ra[u].x = (k < ke && m < M) ? p[0] : 0.0;
            ra[u].y = (k < ke && m + 1 < M) ? p[1] : 0.0;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        const int c = tid + 256 * u;
        const int row = c >> 3, k = k0 + 2 * (c & 7);
        const double* p = B + (size_t)min(n0 + row, N - 1) * ldb + k;
        if (full) {
          rb[u] = *reinterpret_cast<const double2*>(p);
        } else {
          rb[u].x = k < ke ? p[0] : 0.0;
          rb[u].y = k + 1 < ke ? p[1] : 0.0;
        }
      }
    };
    auto stash = [&](int buf) {
#pragma unroll
      for (int u = 0; u < CA; ++u) {
        const int c = tid + 256 * u;
        if (AK) {
          const int row = c >> 3, k = 2 * (c & 7);
          As[buf][k][row] = ra[u].x;
          As[buf][k + 1][row] = ra[u].y;
        } else {
          const int kr = c / (TM / 2), m = 2 * (c % (TM / 2));
          *reinterpret_cast<double2*>(&As[buf][kr][m]) = ra[u];
        }
      }
#pragma unroll
      for (int u = 0; u < CB; ++u) {
        const int c = tid + 256 * u;
        const int row = c >> 3, k = 2 * (c & 7);
        Bs[buf][k][row] = rb[u].x;
        Bs[buf][k + 1][row] = rb[u].y;
      }
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
        double a[BI], b[BJ];
#pragma unroll
        for (int i = 0; i < BI; ++i) a[i] = As[buf][kr][wm + 16 * i + c];
#pragma unroll
        for (int j = 0; j < BJ; ++j) b[j] = Bs[buf][kr][wn + 16 * j + c];
#pragma unroll
        for (int i = 0; i < BI; ++i)
#pragma unroll
          for (int j = 0; j < BJ; ++j)
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
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
        const int nl = wn + 16 * j + (lane & 15), n = n0 + nl;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ml = wm + 16 * i + (lane >> 4) + 4 * v, m = m0 + ml;
          if (!whole) {
            P[nl * TM + ml] = acc[i][j][v];
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
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= sk.tm * sk.tn) return;
  const int ml = e % sk.tm, nl = e / sk.tm;
  int m0, n0;
  tile_origin(sk, t, m0, n0);
  const int m = m0 + ml, n = n0 + nl;
  if (m >= sk.M || n >= sk.N) return;
  double s = 0.0;
  if (sk.S > 0) {
    // K-sliced: the first partial slot of workgroup (slice q, tile t), slices in order
    const int nt = sk.tiles_m * sk.tiles_n, ns = (int)cdivu(sk.nk, sk.nks);
    if (ns <= 1) return;  // one slice covered the tile whole
    const TW* w0 = Wp + (size_t)t * 2 * TE * TE + (size_t)nl * sk.tm + ml;
    const size_t qs = (size_t)nt * 2 * TE * TE;  // slice to slice
    int q = 0;
    for (; q + 8 <= ns; q += 8) {  // eight loads in flight, summed in slice order
      TW v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = w0[(size_t)(q + u) * qs];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += (double)v[u];
    }
    for (; q < ns; ++q) s += (double)w0[(size_t)q * qs];
  } else {
    const long lo = (long)t * sk.nk, hi = lo + sk.nk - 1;
    const int b_lo = (int)(lo / sk.L), b_hi = (int)(hi / sk.L);
    if (b_lo == b_hi) return;
    for (int b = b_lo; b <= b_hi; ++b) {
      const int first = (int)((long)b * sk.L / sk.nk);  // the workgroup's first tile
      const size_t slot = (size_t)b * 2 + (first == t ? 0 : 1);
      s += (double)Wp[slot * TE * TE + (size_t)nl * sk.tm + ml];
    }
  }
  s -= corr_term(sk.corr, A, lda, B, ldb, m, n);
  TC* p = C + (size_t)n * sk.ldc + m;
  *p = (TC)(sk.alpha * s + (sk.beta == 0.0 ? 0.0 : sk.beta * (double)*p));
}

// ------------------------------------------------------------ bf16x3 ------
// two values split at once (split_bf16x2: hi = bf16(x), lo = bf16(x - hi))
__device__ inline void split2(float a, float b, unsigned& hi, unsigned& lo) {
  split_bf16x2(a, b, hi, lo);
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

// NN: the 16-byte k-chunk c of staged row r sits at chunk c ^ ((r >> 4) & 3)
// (four chunks per 32-deep stage), which spreads the transposing stores of
// 32 lanes (rows 4j + u) over the banks (8-way -> 2-way conflicts in a
// 64-bank model) and keeps the fragment reads of 16 rows conflict-free
__device__ inline int nn_col4(int r, int k) { return ((((k >> 3) ^ (r >> 4)) & 3) << 3) | (k & 7); }

// AK: TN (A rows k-contiguous), else NN (A m-contiguous); T the operands'
// type, TC the output's; partial segments go to Wp as floats
template <typename T, typename TC, bool AK>
__global__ void __launch_bounds__(256, 2)
    k_mg_bf3(Sk sk, const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb,
             TC* __restrict__ C, float* __restrict__ Wp) {
  __shared__ unsigned short Ah[TE][LP16], Al[TE][LP16], Bh[TE][LP16], Bl[TE][LP16];
  const int M = sk.M, N = sk.N, K = sk.K;
  const int lb = logical_wg(sk.G);
  long g0, g1;
  sk_range(sk, lb, g0, g1);
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
    int m0, n0;
    tile_origin(sk, t, m0, n0);
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
          const int c = nn_col4(am + u, ak);
          *reinterpret_cast<uint2*>(&Ah[am + u][c]) = make_uint2(h[0], h[1]);
          *reinterpret_cast<uint2*>(&Al[am + u][c]) = make_uint2(l[0], l[1]);
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
        const int ra = wm + 16 * i + r, ka = AK ? kk : nn_col4(ra, kk);
        ah[i] = *reinterpret_cast<const bf8*>(&Ah[ra][ka]);
        al[i] = *reinterpret_cast<const bf8*>(&Al[ra][ka]);
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

// Split bf16x3 on operands split beforehand: each operand is a bf16 hi array
// followed (sk.alo / sk.blo elements later) by its bf16 lo array, with K
// padded by zeros to a multiple of KP, so the K loop has no guards and no
// conversion work; the m-contiguous A of NN is read up to TE elements past
// its last row (the producers allocate that margin).  TN: A rows k-contiguous.
// The K stage is 64 deep, so that a stage's 96 MFMAs per wavefront cover the
// latency of the next stage's loads.
typedef unsigned short u16;
constexpr int KP = 64, LPP = KP + 8;  // pre-split K stage, LDS row pitch (bf16)

// NN products stage A transposed (k-contiguous rows of m): the 16-byte
// k-chunk c of row r sits at chunk c ^ ((r >> 4) & 7), which spreads the
// transposing stores of 32 lanes (rows 4j + u) over the banks (8-way -> 2-way
// conflicts in a 64-bank model) and keeps the MFMA fragment reads, whose 16
// rows share one swizzle, conflict-free
__device__ inline int nn_col(int r, int k) { return ((((k >> 3) ^ (r >> 4)) & 7) << 3) | (k & 7); }

// 4 k rows of 4 m values (each row a pair of packed bf16 pairs) -> the run
// of 4 k values of m = u (16-bit transpose)
__device__ inline uint2 t4(uint2 r0, uint2 r1, uint2 r2, uint2 r3, int u) {
  const unsigned w0 = u < 2 ? r0.x : r0.y, w1 = u < 2 ? r1.x : r1.y;
  const unsigned w2 = u < 2 ? r2.x : r2.y, w3 = u < 2 ? r3.x : r3.y;
  const int sh = (u & 1) * 16;
  return make_uint2(((w0 >> sh) & 0xffffu) | (((w1 >> sh) & 0xffffu) << 16),
                    ((w2 >> sh) & 0xffffu) | (((w3 >> sh) & 0xffffu) << 16));
}

template <bool AK, typename TC>
__global__ void __launch_bounds__(256, 2)
    k_mg_ps3(Sk sk, const u16* __restrict__ A, int lda, const u16* __restrict__ B, int ldb,
             TC* __restrict__ C, float* __restrict__ Wp) {
  __shared__ u16 Ah[TE][LPP], Al[TE][LPP], Bh[TE][LPP], Bl[TE][LPP];
  const int M = sk.M, N = sk.N;
  const int lb = logical_wg(sk.G);
  long g0, g1;
  sk_range(sk, lb, g0, g1);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  // k-contiguous rows: thread -> row tid >> 1, k run (tid & 1) * 32 .. + 31
  const int lr = tid >> 1, lk = (tid & 1) * 32;
  // m-contiguous A (NN): thread -> m run (tid & 31) * 4, k rows (tid >> 5) * 8 .. + 7
  const int am = (tid & 31) * 4, ak = (tid >> 5) * 8;
  for (long g = g0; g < g1;) {
    const int t = (int)(g / sk.nk);
    const int s0 = (int)(g - (long)t * sk.nk);
    const int s1 = (int)min((long)sk.nk, g1 - (long)t * sk.nk);
    int m0, n0;
    tile_origin(sk, t, m0, n0);
    const int kb = s0 * KP, ke = s1 * KP;
    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    const u16* Ar = A + (AK ? (size_t)min(m0 + lr, M - 1) * lda : (size_t)m0 + am);
    const u16* Br = B + (size_t)min(n0 + lr, N - 1) * ldb;
    // operand registers of the next stage: named scalars (arrays of vectors
    // would be placed in scratch memory)
    uint4 ah0, ah1, ah2, ah3, al0, al1, al2, al3, bh0, bh1, bh2, bh3, bl0, bl1, bl2, bl3;
    uint2 nh0, nh1, nh2, nh3, nh4, nh5, nh6, nh7, nl0, nl1, nl2, nl3, nl4, nl5, nl6, nl7;
    auto fetch = [&](int k0) {
      if constexpr (AK) {
        const uint4* p = reinterpret_cast<const uint4*>(Ar + k0 + lk);
        const uint4* q = reinterpret_cast<const uint4*>(Ar + sk.alo + k0 + lk);
        ah0 = p[0]; ah1 = p[1]; ah2 = p[2]; ah3 = p[3];
        al0 = q[0]; al1 = q[1]; al2 = q[2]; al3 = q[3];
      } else {
        const u16* p = Ar + (size_t)(k0 + ak) * lda;
        const u16* q = p + sk.alo;
        const size_t d = (size_t)lda;
        nh0 = *reinterpret_cast<const uint2*>(p + 0 * d);
        nh1 = *reinterpret_cast<const uint2*>(p + 1 * d);
        nh2 = *reinterpret_cast<const uint2*>(p + 2 * d);
        nh3 = *reinterpret_cast<const uint2*>(p + 3 * d);
        nh4 = *reinterpret_cast<const uint2*>(p + 4 * d);
        nh5 = *reinterpret_cast<const uint2*>(p + 5 * d);
        nh6 = *reinterpret_cast<const uint2*>(p + 6 * d);
        nh7 = *reinterpret_cast<const uint2*>(p + 7 * d);
        nl0 = *reinterpret_cast<const uint2*>(q + 0 * d);
        nl1 = *reinterpret_cast<const uint2*>(q + 1 * d);
        nl2 = *reinterpret_cast<const uint2*>(q + 2 * d);
        nl3 = *reinterpret_cast<const uint2*>(q + 3 * d);
        nl4 = *reinterpret_cast<const uint2*>(q + 4 * d);
        nl5 = *reinterpret_cast<const uint2*>(q + 5 * d);
        nl6 = *reinterpret_cast<const uint2*>(q + 6 * d);
        nl7 = *reinterpret_cast<const uint2*>(q + 7 * d);
      }
      const uint4* p = reinterpret_cast<const uint4*>(Br + k0 + lk);
      const uint4* q = reinterpret_cast<const uint4*>(Br + sk.blo + k0 + lk);
      bh0 = p[0]; bh1 = p[1]; bh2 = p[2]; bh3 = p[3];
      bl0 = q[0]; bl1 = q[1]; bl2 = q[2]; bl3 = q[3];
    };
    auto stash = [&]() {
      if constexpr (AK) {
        *reinterpret_cast<uint4*>(&Ah[lr][lk + 0]) = ah0;
        *reinterpret_cast<uint4*>(&Ah[lr][lk + 8]) = ah1;
        *reinterpret_cast<uint4*>(&Ah[lr][lk + 16]) = ah2;
        *reinterpret_cast<uint4*>(&Ah[lr][lk + 24]) = ah3;
        *reinterpret_cast<uint4*>(&Al[lr][lk + 0]) = al0;
        *reinterpret_cast<uint4*>(&Al[lr][lk + 8]) = al1;
        *reinterpret_cast<uint4*>(&Al[lr][lk + 16]) = al2;
        *reinterpret_cast<uint4*>(&Al[lr][lk + 24]) = al3;
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int c = nn_col(am + u, ak);
          *reinterpret_cast<uint2*>(&Ah[am + u][c]) = t4(nh0, nh1, nh2, nh3, u);
          *reinterpret_cast<uint2*>(&Ah[am + u][c + 4]) = t4(nh4, nh5, nh6, nh7, u);
          *reinterpret_cast<uint2*>(&Al[am + u][c]) = t4(nl0, nl1, nl2, nl3, u);
          *reinterpret_cast<uint2*>(&Al[am + u][c + 4]) = t4(nl4, nl5, nl6, nl7, u);
        }
      }
      *reinterpret_cast<uint4*>(&Bh[lr][lk + 0]) = bh0;
      *reinterpret_cast<uint4*>(&Bh[lr][lk + 8]) = bh1;
      *reinterpret_cast<uint4*>(&Bh[lr][lk + 16]) = bh2;
      *reinterpret_cast<uint4*>(&Bh[lr][lk + 24]) = bh3;
      *reinterpret_cast<uint4*>(&Bl[lr][lk + 0]) = bl0;
      *reinterpret_cast<uint4*>(&Bl[lr][lk + 8]) = bl1;
      *reinterpret_cast<uint4*>(&Bl[lr][lk + 16]) = bl2;
      *reinterpret_cast<uint4*>(&Bl[lr][lk + 24]) = bl3;
    };
    fetch(kb);
    for (int k0 = kb; k0 < ke; k0 += KP) {
      stash();
      __syncthreads();
      if (k0 + KP < ke) fetch(k0 + KP);
      // v_mfma_f32_16x16x32_bf16 operands: lane l holds row l % 16, k 8 (l / 16) .. + 7
      const int r = lane & 15;
#pragma unroll
      for (int ks = 0; ks < KP; ks += 32) {
        const int kk = ks + 8 * (lane >> 4);
        bf8 xh[4], xl[4], yh[4], yl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int ra = wm + 16 * i + r, ka = AK ? kk : nn_col(ra, kk);
          xh[i] = *reinterpret_cast<const bf8*>(&Ah[ra][ka]);
          xl[i] = *reinterpret_cast<const bf8*>(&Al[ra][ka]);
          yh[i] = *reinterpret_cast<const bf8*>(&Bh[wn + 16 * i + r][kk]);
          yl[i] = *reinterpret_cast<const bf8*>(&Bl[wn + 16 * i + r][kk]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[i], yh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[i], yl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[i], yh[j], acc[i][j], 0, 0, 0);
          }
      }
      __syncthreads();
    }
    const bool whole = s0 == 0 && s1 == sk.nk;
    float* P = Wp + ((size_t)lb * 2 + (g == g0 ? 0 : 1)) * TE * TE;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int nl2 = wn + 16 * j + (lane & 15), n = n0 + nl2;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ml = wm + 16 * i + 4 * (lane >> 4) + v, m = m0 + ml;
          if (!whole) {
            P[nl2 * TE + ml] = acc[i][j][v];
          } else if (m < M && n < N) {
            TC* p = C + (size_t)n * sk.ldc + m;
            *p = (TC)(sk.alpha * (double)acc[i][j][v] +
                      (sk.beta == 0.0 ? 0.0 : sk.beta * (double)*p));
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
  // FP64 output tile per (kind, M, N, K, lda, ldb): index into the variants
  std::map<std::tuple<int, int, int, int, int, int>, int> tile;
  float* f32[4] = {};  // FP32 work buffers of the fused mixed band-apply
  size_t f32_len[4] = {};
  float* B32 = nullptr;  // FP32 copy of the KB table
  const double* B32_src = nullptr;
  size_t B32_len = 0;
  // the KB table split into bf16 hi | lo (rows padded to 32, K to 32)
  u16* Ts = nullptr;
  const double* Ts_src = nullptr;
  size_t Ts_len = 0;
  u16* u16s[kU16Scratch] = {};  // zero-initialised split work arrays
  size_t u16_len[kU16Scratch] = {};
};

__global__ void k_split_rows(const double* __restrict__ src, int rows, int cols, int lds,
                             u16* __restrict__ dst, int ld, int rows_pad, long lo) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;  // pair index
  const long npair = lo / 2;
  if (i >= npair) return;
  const long e = 2 * i;
  const long r = e / ld;
  const int col = (int)(e - r * ld);
  float a = 0.f, b = 0.f;
  if (r < rows && col < cols) a = (float)src[r * lds + col];
  if (r < rows && col + 1 < cols) b = (float)src[r * lds + col + 1];
  (void)rows_pad;
  unsigned h, l;
  split_bf16x2(a, b, h, l);
  *reinterpret_cast<unsigned*>(dst + e) = h;
  *reinterpret_cast<unsigned*>(dst + lo + e) = l;
}

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
  bytes = std::max(bytes, (size_t)256);  // a product with no partial tiles still gets a pointer
  if (s->ws_bytes >= bytes) return s->ws;
  buffers_changed(c);
  if (s->ws) (void)hipFree(s->ws);
  s->ws = nullptr;
  s->ws_bytes = 0;
  if (hipMalloc(&s->ws, bytes) != hipSuccess) return nullptr;
  s->ws_bytes = bytes;
  return s->ws;
}

// A schedule code: > 0 = stream-K on that many workgroups, < 0 = K-sliced
// with -code slices (on the multiple of 8 workgroups that covers slices x
// tiles; -1 is one workgroup per output tile)
Sk make_sk(int M, int N, int K, int kt, int code, double alpha, double beta, int ldc, int corr,
           int tm = TE, int tn = TE, bool nfast = false) {
  Sk sk;
  sk.M = M;
  sk.N = N;
  sk.K = K;
  sk.tm = tm;
  sk.tn = tn;
  sk.tiles_m = (int)cdivu(M, tm);
  sk.tiles_n = (int)cdivu(N, tn);
  sk.nk = (int)cdivu(K, kt);
  sk.nfast = nfast ? 1 : 0;
  const long T = (long)sk.tiles_m * sk.tiles_n;
  if (code > 0) {
    sk.S = 0;
    sk.nks = 0;
    sk.G = code;
    sk.L = (long)cdivu(T * sk.nk, code);
  } else {
    sk.S = std::min(-code, std::max(sk.nk, 1));
    sk.nks = (int)cdivu(sk.nk, sk.S);
    sk.G = (int)(cdivu(sk.S * T, kXcd) * kXcd);
    sk.L = 0;
  }
  sk.alpha = alpha;
  sk.beta = beta;
  sk.ldc = ldc;
  sk.corr = corr;
  sk.alo = sk.blo = 0;
  return sk;
}

// partial-tile slots of the workspace a product needs (none when every
// tile is covered whole by one workgroup)
size_t sk_slots(const Sk& sk) {
  if (sk.S > 0) return cdivu(sk.nk, sk.nks) <= 1 ? 0 : 2 * (size_t)sk.G;
  return 2 * (size_t)sk.G;
}

// which schedules the timed choice may take: 0 both, 1 stream-K only,
// 2 K-sliced only (qx_set_gemm_schedule; tests run each on its own)
std::atomic<int> g_sched{0};

// candidate schedule codes of a product with T output tiles on `cus` CUs:
// stream-K on one and two workgroups per CU; K-sliced with 8 slices (one per
// XCD) and with as many slices as fill one or two workgroups per CU, all
// slices x tiles resident at once; one workgroup per tile where the tiles
// alone fill the chip
std::vector<int> sched_candidates(long T, int nk, int cus, int per_cu_max) {
  std::vector<int> v;
  const int mode = g_sched.load(std::memory_order_relaxed);
  if (mode != 2)
    for (int k = 1; k <= per_cu_max; ++k) v.push_back(k * cus);
  if (mode != 1) {
    const int before = (int)v.size();
    auto add = [&](long S) {
      if (S < 1 || S > nk) return;
      const int code = -(int)S;
      if (std::find(v.begin() + before, v.end(), code) == v.end()) v.push_back(code);
    };
    if (8 * T <= (long)per_cu_max * cus) add(8);
    for (int k = 1; k <= per_cu_max; ++k)
      if ((long)k * cus / T >= 2) add((long)k * cus / T);
    if (T >= cus || (int)v.size() == before) add(1);
  }
  return v;
}

bool tuning_off() {
  static const bool off = std::getenv("QX_GEMM_NOTUNE") != nullptr;
  return off;
}

// schedule of a product (a code of make_sk): the candidates of
// sched_candidates, timed on first use of a shape with `trial(code)` (which
// writes a scratch output), then cached; untimed, stream-K on two
// workgroups per CU; 0 on failure
template <typename F>
int choose_grid(Ctx* c, int kind, int M, int N, int K, int kt, int tm, int tn, int lda, int ldb,
                int per_cu_max, size_t part_elem, F trial) {
  MgState* s = state(c);
  const int mode = g_sched.load(std::memory_order_relaxed);
  const auto key = std::make_tuple(kind + 1000 * mode, M, N, K, lda, ldb);
  auto it = s->grid.find(key);
  if (it != s->grid.end()) return it->second;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->dev);
  cus = std::max(kXcd, cus / kXcd * kXcd);
  const long T = (long)cdivu(M, tm) * cdivu(N, tn);
  const std::vector<int> cand = sched_candidates(T, (int)cdivu(K, kt), cus, per_cu_max);
  int best_code = cand.empty() ? 2 * cus : cand.back();
  if (mode == 0 && !cand.empty()) best_code = cand[per_cu_max > 1 ? 1 : 0];
  if (!tuning_off() && cand.size() > 1) {
    hipEvent_t e0, e1;
    QX_CHECK(hipEventCreate(&e0));
    QX_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    for (const int code : cand) {
      const Sk probe = make_sk(M, N, K, kt, code, 1.0, 0.0, M, 0, tm, tn);
      if (!workspace(c, part_elem * TE * TE * sk_slots(probe))) return 0;
      if (trial(code)) return 0;  // warm
      QX_CHECK(hipEventRecord(e0, c->stream));
      for (int r = 0; r < 2; ++r)
        if (trial(code)) return 0;
      QX_CHECK(hipEventRecord(e1, c->stream));
      QX_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      QX_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        best_code = code;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  s->grid[key] = best_code;
  return best_code;
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
  const int tiles = sk.tiles_m * sk.tiles_n;
  if (sk_slots(sk))  // else every tile was written whole by its workgroup
    hipLaunchKernelGGL((k_mg_fixup<TW, T, TC>), dim3(TE * TE / 256, tiles), dim3(256), 0,
                       c->stream, sk, W, A, lda, B, ldb, C);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

// nfast: number the output tiles with the N index fastest (the NN products,
// whose long operand A is shared by the tiles of one row of tiles)
template <typename TW, typename T, typename TC, typename K1>
int run_product(Ctx* c, int kind, int kt, int per_cu_max, K1 kernel, int M, int N, int K,
                double alpha, const T* A, int lda, const T* B, int ldb, double beta, TC* C,
                int ldc, int corr, long alo = 0, long blo = 0, int tm = TE, int tn = TE,
                bool nfast = false) {
  if (M <= 0 || N <= 0) return 0;
  if ((lda | ldb) & 1) return -3;  // 8 / 16-byte row loads need even leading dimensions
  MgState* s = state(c);
  const int mode = g_sched.load(std::memory_order_relaxed);
  const bool known = s->grid.count(std::make_tuple(kind + 1000 * mode, M, N, K, lda, ldb)) != 0;
  int code;
  if (known || tuning_off()) {
    code = choose_grid(c, kind, M, N, K, kt, tm, tn, lda, ldb, per_cu_max, sizeof(TW),
                       [](int) { return 0; });
  } else {
    TrialOut<TC> out((size_t)M * N);
    if (!out.p) return -3;
    code = choose_grid(c, kind, M, N, K, kt, tm, tn, lda, ldb, per_cu_max, sizeof(TW),
                       [&](int cd) {
                         Sk tk = make_sk(M, N, K, kt, cd, alpha, 0.0, M, corr, tm, tn, nfast);
                         tk.alo = alo;
                         tk.blo = blo;
                         return launch_sk<TW>(c, tk, kernel, A, lda, B, ldb, out.p);
                       });
    QX_CHECK(hipStreamSynchronize(c->stream));  // trials done before the scratch goes
  }
  if (code == 0) return -3;
  Sk sk = make_sk(M, N, K, kt, code, alpha, beta, ldc, corr, tm, tn, nfast);
  if (!workspace(c, sizeof(TW) * TE * TE * sk_slots(sk))) return -3;
  sk.alo = alo;
  sk.blo = blo;
  lt_mark(c, c->lt_gemm_id, false);
  const int rc = launch_sk<TW>(c, sk, kernel, A, lda, B, ldb, C);
  lt_mark(c, c->lt_gemm_id, true);
  return rc;
}

}  // namespace

void mg_set_schedule(int mode) { g_sched.store(mode, std::memory_order_relaxed); }
int mg_schedule() { return g_sched.load(std::memory_order_relaxed); }

void mg_destroy(Ctx* c) {
  if (!c->mg) return;
  MgState* s = static_cast<MgState*>(c->mg);
  if (s->ws) (void)hipFree(s->ws);
  for (float* p : s->f32)
    if (p) (void)hipFree(p);
  for (u16* p : s->u16s)
    if (p) (void)hipFree(p);
  if (s->B32) (void)hipFree(s->B32);
  if (s->Ts) (void)hipFree(s->Ts);
  delete s;
  c->mg = nullptr;
}

// the KB table changed in place: its FP32 copy is made again on next use
void mg_forget_kb(Ctx* c) {
  if (c->mg) {
    state(c)->B32_src = nullptr;
    state(c)->Ts_src = nullptr;
  }
}

// the KB table (npj rows of nplwex) split into bf16 hi | lo: row stride
// round_up(nplwex, 64), rows padded to round_up(npj, 64), zeros in every pad,
// a margin of TE elements after each half (the NN edge tiles read into it);
// lo = the lo half's offset
const unsigned short* mg_kb_split(Ctx* c, int* ld, int* rows_pad, long* lo) {
  if (!c->d_B) return nullptr;
  MgState* s = state(c);
  *ld = (int)(cdivu(c->nplwex, KP) * KP);
  *rows_pad = (int)(cdivu(c->npj, KP) * KP);
  *lo = (long)*rows_pad * *ld + TE;
  const size_t len = 2 * (size_t)*lo;
  if (s->Ts_src != c->d_B || s->Ts_len != len || !s->Ts) {
    buffers_changed(c);
    if (s->Ts_len != len || !s->Ts) {
      if (s->Ts) (void)hipFree(s->Ts);
      s->Ts = nullptr;
      s->Ts_len = 0;
      if (hipMalloc(&s->Ts, sizeof(u16) * len) != hipSuccess) return nullptr;
    }
    hipLaunchKernelGGL(k_split_rows, dim3(cdivu(*lo / 2, 256)), dim3(256), 0, c->stream, c->d_B,
                       c->npj, c->nplwex, c->nplwex, s->Ts, *ld, *rows_pad, *lo);
    s->Ts_src = c->d_B;
    s->Ts_len = len;
  }
  return s->Ts;
}

// zero-initialised u16 work array i (0 .. kU16Scratch - 1) of at least n
// elements; contents persist between calls (the producers write only inside
// their pads)
unsigned short* mg_u16_scratch(Ctx* c, int i, size_t n) {
  if (i < 0 || i >= kU16Scratch) return nullptr;
  MgState* s = state(c);
  if (s->u16_len[i] < n) {
    buffers_changed(c);
    if (s->u16s[i]) (void)hipFree(s->u16s[i]);
    s->u16s[i] = nullptr;
    s->u16_len[i] = 0;
    if (hipMalloc(&s->u16s[i], sizeof(u16) * n) != hipSuccess) return nullptr;
    if (hipMemsetAsync(s->u16s[i], 0, sizeof(u16) * n, c->stream) != hipSuccess) return nullptr;
    s->u16_len[i] = n;
  }
  return s->u16s[i];
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

namespace {
// FP64 product on the output tile chosen for its shape: the four tiles
// (128 | 96) x (128 | 96) are timed on the first product of a shape (into a
// scratch output; each with its own tuned grid) and the fastest is kept;
// QX_GEMM_NOTUNE=1 takes the tile that pads the output least
template <bool AK>
int f64_product(Ctx* c, int kind, int M, int N, int K, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc, int corr) {
  using KF = void (*)(Sk, const double*, int, const double*, int, double*, double*);
  static const KF ks[4] = {k_mg_f64<AK, 4, 4>, k_mg_f64<AK, 4, 3>, k_mg_f64<AK, 3, 4>,
                           k_mg_f64<AK, 3, 3>};
  static const int tms[4] = {128, 128, 96, 96}, tns[4] = {128, 96, 128, 96};
  if (M <= 0 || N <= 0) return 0;
  if ((lda | ldb) & 1) return -3;
  MgState* s = state(c);
  const auto key = std::make_tuple(kind, M, N, K, lda, ldb);
  auto run = [&](int v, double b, double* out, int ldo) {
    return run_product<double>(c, kind + 10 * (v + 1), KT64, 2, ks[v], M, N, K, alpha, A, lda, B,
                               ldb, b, out, ldo, corr, 0, 0, tms[v], tns[v], !AK);
  };
  auto it = s->tile.find(key);
  int best = 0;
  if (it != s->tile.end()) {
    best = it->second;
  } else if (tuning_off()) {
    long least = -1;
    for (int v = 0; v < 4; ++v) {
      const long area = (long)cdivu(M, tms[v]) * tms[v] * (long)cdivu(N, tns[v]) * tns[v];
      if (least < 0 || area < least) {
        least = area;
        best = v;
      }
    }
    s->tile[key] = best;
  } else {
    TrialOut<double> out((size_t)M * N);
    if (!out.p) return -3;
    hipEvent_t e0, e1;
    QX_CHECK(hipEventCreate(&e0));
    QX_CHECK(hipEventCreate(&e1));
    float bms = 1e30f;
    for (int v = 0; v < 4; ++v) {
      if (int rc = run(v, 0.0, out.p, M)) return rc;  // tunes the grid of this tile
      QX_CHECK(hipEventRecord(e0, c->stream));
      for (int r = 0; r < 2; ++r)
        if (int rc = run(v, 0.0, out.p, M)) return rc;
      QX_CHECK(hipEventRecord(e1, c->stream));
      QX_CHECK(hipEventSynchronize(e1));
      float ms = 0.f;
      QX_CHECK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < bms) {
        bms = ms;
        best = v;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->tile[key] = best;
  }
  return run(best, beta, C, ldc);
}
}  // namespace

// FP64 TN: C = alpha (A^T B - corr) + beta C
int mg_dgemm_tn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc, int corr) {
  return f64_product<true>(c, 0, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, corr);
}

// FP64 NN: C = alpha A B + beta C
int mg_dgemm_nn(Ctx* c, int M, int N, int K, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc) {
  return f64_product<false>(c, 1, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 0);
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
                            N, K, alpha, A, lda, B, ldb, beta, C, ldc, 0, 0, 0, TE, TE, true);
}

// split bf16x3 on pre-split operands (hi then lo, K a multiple of 64 with
// zero padding, leading dimensions and lo offsets multiples of 64): TN / NN
// into FP32 C
int mg_ps3_tn(Ctx* c, int M, int N, int K, const unsigned short* A, int lda, long alo,
              const unsigned short* B, int ldb, long blo, float* C, int ldc) {
  if (K % KP || (lda | ldb) % KP || alo % KP || blo % KP) return -3;
  return run_product<float>(c, 6, KP, 2, k_mg_ps3<true, float>, M, N, K, 1.0, A, lda, B, ldb,
                            0.0, C, ldc, 0, alo, blo);
}

int mg_ps3_nn(Ctx* c, int M, int N, int K, const unsigned short* A, int lda, long alo,
              const unsigned short* B, int ldb, long blo, float* C, int ldc) {
  if (K % KP || (lda | ldb) % KP || alo % KP || blo % KP) return -3;
  return run_product<float>(c, 7, KP, 2, k_mg_ps3<false, float>, M, N, K, 1.0, A, lda, B, ldb,
                            0.0, C, ldc, 0, alo, blo, TE, TE, true);
}

// the same on FP64 C: C = alpha A^T B + beta C (TN), C = alpha A B + beta C (NN)
int mg_ps3_tn_d(Ctx* c, int M, int N, int K, double alpha, const unsigned short* A, int lda,
                long alo, const unsigned short* B, int ldb, long blo, double beta, double* C,
                int ldc) {
  if (K % KP || (lda | ldb) % KP || alo % KP || blo % KP) return -3;
  return run_product<float>(c, 8, KP, 2, k_mg_ps3<true, double>, M, N, K, alpha, A, lda, B, ldb,
                            beta, C, ldc, 0, alo, blo);
}

int mg_ps3_nn_d(Ctx* c, int M, int N, int K, double alpha, const unsigned short* A, int lda,
                long alo, const unsigned short* B, int ldb, long blo, double beta, double* C,
                int ldc) {
  if (K % KP || (lda | ldb) % KP || alo % KP || blo % KP) return -3;
  return run_product<float>(c, 9, KP, 2, k_mg_ps3<false, double>, M, N, K, alpha, A, lda, B, ldb,
                            beta, C, ldc, 0, alo, blo, TE, TE, true);
}

// rows x cols FP64 (row stride lds) -> bf16 hi at dst, lo at dst + lo: row
// stride ld, zeros in every element of [0, lo) outside rows x cols
int mg_split_rows(Ctx* c, const double* src, int rows, int cols, int lds, unsigned short* dst,
                  int ld, long lo) {
  if (ld % 2 || lo % 2 || cols > ld || (long)rows * ld > lo) return -3;
  hipLaunchKernelGGL(k_split_rows, dim3(cdivu(lo / 2, 256)), dim3(256), 0, c->stream, src, rows,
                     cols, lds, dst, ld, 0, lo);
  return hipGetLastError() == hipSuccess ? 0 : -4;
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
