// pencil_fast.hip -- the band-apply / density pencil FFTs for cubic grids whose
// edge N factors as NA*NB with both factors <= 16 (30 = 6*5, 60 = 6*10,
// 90 = 9*10, 120 = 12*10, 64 = 8*8, 96 = 12*8, 128 = 16*8).
//
// A line transform of length N is done in two register steps with a single
// LDS exchange between them:
//   NP ("natural -> permuted"): step A, thread b < NB: DFT_NA over the NB-strided
//       elements b + NB*a, twiddle W_N^(b*ka); step B, thread ka < NA: DFT_NB over
//       the contiguous run NB*ka + b.  Index ka + NA*kb ends up at slot kb + NB*ka.
//   PN ("permuted -> natural"): the same two steps in reverse order.
// Every inverse transform of the band-apply is NP and every forward one PN, so
// the digit-transposed intermediate order never needs a reordering pass: the
// z planes of T1/T2 and the y/x slots of a plane are kept in slot order and
// V(r) is stored pre-permuted.  In x the inverse step B, the V(r) multiply and
// the forward step B' act on the same registers.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include "context.h"

namespace qx {

constexpr int kFastThreads = 256;
// column pairs per z block: 16, or 8 where 32 complex lines of length N would
// pass 48 KiB of static LDS
__host__ __device__ constexpr int zpairs(int N, int elem_bytes) {
  return 2 * 16 * N * elem_bytes <= 49152 ? 16 : 8;
}

__host__ __device__ inline int slot_phys(int s, int NA, int NB) { return s / NB + NA * (s % NB); }

struct FastArgs {
  int N, nH, hmax, npair, ncol, nplwex, nb, bp0;
  int sa, sr;  // row strides (elements) of the plane kernel's sphere tile and x strips
  const int *pt_col, *pt_z, *pair_p0, *col_hi, *col_kmod;
  const double* kin;
  const double* tw;  // exp(+2 pi i k / N), k < N
};

template <typename T>
__device__ inline cplx<T> twid(const cplx<T>* tw, int k, int sign) {
  cplx<T> w = tw[k];
  if (sign < 0) w.y = -w.y;
  return w;
}

// ---- line steps (one thread, line in LDS with unit stride) ----------------
template <typename T, int NA, int NB>
__device__ inline void np_stepA(cplx<T>* L, int b, const cplx<T>* tw, int sign) {
  cplx<T> v[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) v[a] = L[b + NB * a];
  Dft<T, NA>::run(v, sign);
  int k = 0;
#pragma unroll
  for (int ka = 1; ka < NA; ++ka) {
    k += b;  // b * ka
    v[ka] = cmul(v[ka], twid(tw, k, sign));
  }
#pragma unroll
  for (int ka = 0; ka < NA; ++ka) L[b + NB * ka] = v[ka];
}
template <typename T, int NA, int NB>
__device__ inline void np_stepB(cplx<T>* L, int ka, int sign) {
  cplx<T> v[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) v[b] = L[NB * ka + b];
  Dft<T, NB>::run(v, sign);
#pragma unroll
  for (int kb = 0; kb < NB; ++kb) L[NB * ka + kb] = v[kb];
}
template <typename T, int NA, int NB>
__device__ inline void pn_stepB(cplx<T>* L, int ka, const cplx<T>* tw, int sign) {
  cplx<T> v[NB];
#pragma unroll
  for (int kb = 0; kb < NB; ++kb) v[kb] = L[NB * ka + kb];
  Dft<T, NB>::run(v, sign);
  int k = 0;
#pragma unroll
  for (int b = 1; b < NB; ++b) {
    k += ka;  // b * ka
    v[b] = cmul(v[b], twid(tw, k, sign));
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) L[NB * ka + b] = v[b];
}
template <typename T, int NA, int NB>
__device__ inline void pn_stepA(cplx<T>* L, int b, int sign) {
  cplx<T> v[NA];
#pragma unroll
  for (int ka = 0; ka < NA; ++ka) v[ka] = L[b + NB * ka];
  Dft<T, NA>::run(v, sign);
#pragma unroll
  for (int a = 0; a < NA; ++a) L[b + NB * a] = v[a];
}

__device__ inline int mirror_col_f(int col) {
  return col == 0 ? 0 : ((col & 1) ? col + 1 : col - 1);
}

// ---- z inverse: sphere -> T1[bp][z slot][col] ------------------------------
// psi32 (optional): receives the FP32 copy of the coefficients read;
// psis (optional): their bf16 hi | lo split instead (row stride ldk, lo half
// at offset slo), the operand of the split bf16x3 KB projection
template <typename T, int NA, int NB>
__global__ void __launch_bounds__(kFastThreads) k_zinv_f(FastArgs a, const double* __restrict__ psi,
                                                         cplx<T>* __restrict__ t1,
                                                         float* __restrict__ psi32,
                                                         unsigned short* __restrict__ psis,
                                                         int ldk, long slo) {
  constexpr int N = NA * NB;
  constexpr int PB = zpairs(N, (int)sizeof(cplx<T>));
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> lines[2 * PB][N];
  const int q0 = blockIdx.x * PB, q1 = min(q0 + PB, a.npair);
  const int bp = blockIdx.y;
  const int ba = 2 * (a.bp0 + bp), bb = ba + 1;
  const int c0 = q0 == 0 ? 0 : 2 * q0 - 1, c1 = 2 * q1 - 1, ncl = c1 - c0;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = mk<T>((T)a.tw[2 * k], (T)a.tw[2 * k + 1]);
  for (int i = threadIdx.x; i < ncl * N; i += blockDim.x) lines[i / N][i % N] = mk<T>(0, 0);
  __syncthreads();
  const double* pa = psi + (size_t)ba * a.nplwex;
  const double* pb = psi + (size_t)bb * a.nplwex;
  const bool hasb = bb < a.nb;
  const int p0 = a.pair_p0[q0], p1 = a.pair_p0[q1];
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    double ar = pa[2 * p], ai = pa[2 * p + 1];
    double br = hasb ? pb[2 * p] : 0.0, bi = hasb ? pb[2 * p + 1] : 0.0;
    if (psi32) {
      float* qa = psi32 + (size_t)ba * a.nplwex;
      qa[2 * p] = (float)ar;
      qa[2 * p + 1] = (float)ai;
      if (hasb) {
        qa[a.nplwex + 2 * p] = (float)br;
        qa[a.nplwex + 2 * p + 1] = (float)bi;
      }
    }
    if (psis) {
      unsigned h, l;
      unsigned short* qa = psis + (size_t)ba * ldk + 2 * p;
      split_bf16x2((float)ar, (float)ai, h, l);
      *reinterpret_cast<unsigned*>(qa) = h;
      *reinterpret_cast<unsigned*>(qa + slo) = l;
      if (hasb) {
        split_bf16x2((float)br, (float)bi, h, l);
        *reinterpret_cast<unsigned*>(qa + ldk) = h;
        *reinterpret_cast<unsigned*>(qa + ldk + slo) = l;
      }
    }
    int col = a.pt_col[p] - c0, z = a.pt_z[p];
    if (p == 0) {
      lines[col][0] = mk<T>((T)ar, (T)br);
    } else {
      lines[col][z] = mk<T>((T)(ar - bi), (T)(ai + br));
      lines[mirror_col_f(col + c0) - c0][z == 0 ? 0 : N - z] = mk<T>((T)(ar + bi), (T)(br - ai));
    }
  }
  __syncthreads();
  for (int it = threadIdx.x; it < ncl * NB; it += blockDim.x)
    np_stepA<T, NA, NB>(lines[it / NB], it % NB, tw, +1);
  __syncthreads();
  for (int it = threadIdx.x; it < ncl * NA; it += blockDim.x)
    np_stepB<T, NA, NB>(lines[it / NA], it % NA, +1);
  __syncthreads();
  cplx<T>* out = t1 + (size_t)bp * N * a.ncol;
  const unsigned mc = div_magic((unsigned)ncl);
  for (int i = threadIdx.x; i < ncl * N; i += blockDim.x) {
    const int zs = fdiv(i, mc), cl = i - zs * ncl;
    out[(size_t)zs * a.ncol + c0 + cl] = lines[cl][zs];
  }
}

// ---- plane: y inverse, x inverse, x V, x forward, y forward ----------------
// Row offsets in the sphere tile of the natural-x inputs b + NB a of x step A
// (lane b): row h + hmax for |h| <= hmax, else row 2 hmax + 1, which is kept
// zero, so every lane loads unconditionally; bit a of *inside is set for the
// slots inside the sphere, the ones step A' writes back.
template <int NA, int NB>
__device__ inline void x_offsets(int b, int hmax, int sa, int (&off)[NA], unsigned* inside) {
  constexpr int N = NA * NB;
  unsigned m = 0;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    const int n = b + NB * a;
    const bool lo = n <= hmax, hi = n >= N - hmax;
    off[a] = (lo ? n + hmax : (hi ? n - N + hmax : 2 * hmax + 1)) * sa;
    m |= (lo || hi) ? 1u << a : 0u;
  }
  *inside = m;
}

// twiddles of an x-pass lane: the lane's index j is the same for every row it
// transforms (b of step A, ka of step B'), so W^(j m), m = 1 .. max(NA, NB) - 1,
// are read from the table once and kept in registers
template <int NA, int NB>
__host__ __device__ constexpr int x_ntw() {
  return (NA > NB ? NA : NB) - 1;
}
template <typename T, int NA, int NB>
__device__ inline void x_twiddles(const cplx<T>* tw, int j, cplx<T> (&w)[x_ntw<NA, NB>()]) {
  constexpr int N = NA * NB;
  int k = 0;
#pragma unroll
  for (int m = 0; m < x_ntw<NA, NB>(); ++m) {
    k += j;
    if (k >= N) k -= N;  // only lanes idle in both steps get here
    w[m] = tw[k];
  }
}

// x step A for row slot ys, inputs gathered from the sphere tile; xw: the
// lane's twiddles W^(b ka)
template <typename T, int NA, int NB>
__device__ inline void x_stepA(const cplx<T>* A, const int (&off)[NA], cplx<T>* R, int ys, int b,
                               const cplx<T> (&xw)[x_ntw<NA, NB>()]) {
  cplx<T> v[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) v[a] = A[off[a] + ys];
  Dft<T, NA>::run(v, +1);
#pragma unroll
  for (int ka = 1; ka < NA; ++ka) v[ka] = cmul(v[ka], xw[ka - 1]);
#pragma unroll
  for (int ka = 0; ka < NA; ++ka) R[b + NB * ka] = v[ka];
}

// x step A': back to natural x, the slots inside the sphere into the tile
template <typename T, int NA, int NB>
__device__ inline void x_stepAp(cplx<T>* A, const int (&off)[NA], unsigned inside,
                                const cplx<T>* R, int ys, int b) {
  cplx<T> v[NA];
#pragma unroll
  for (int ka = 0; ka < NA; ++ka) v[ka] = R[b + NB * ka];
  Dft<T, NA>::run(v, -1);
#pragma unroll
  for (int a = 0; a < NA; ++a)
    if (inside & (1u << a)) A[off[a] + ys] = v[a];
}

// LDS accesses of one wavefront are performed in issue order, so lanes of a
// wave can exchange data through LDS without a workgroup barrier; the fences
// only keep the compiler from moving LDS accesses across this point.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kPlaneMaxThreads = 512;
constexpr int kPlaneLoads = 4;  // sphere columns held in registers per thread

// x rows a wavefront transforms at once: step A runs NB lanes per row, step B
// NA lanes per row
__host__ __device__ constexpr int plane_rows_per_wave(int NA, int NB) {
  return 64 / (NA > NB ? NA : NB);
}

// One z plane of one band pair: y inverse over the nH sphere rows, then for
// every y slot the x line (inverse, V(r) multiply, forward) and y forward.
// The x lines are owned by wavefronts (a wave takes plane_rows_per_wave rows
// at a time through its own LDS strip), so the x pass needs no workgroup
// barrier; V(r) of a wave's next rows is loaded while the current rows are
// transformed.  LDS bank conflicts are kept down by the row strides a.sa (sphere
// tile) and a.sr (x strips), chosen on the host, and by lane maps that put the
// line (y pass) or row (x pass) index fastest across lanes.
template <typename T, int NA, int NB>
__global__ void __launch_bounds__(kPlaneMaxThreads) k_plane_f(FastArgs a, const T* __restrict__ Vp,
                                                            const cplx<T>* __restrict__ t1,
                                                            cplx<T>* __restrict__ t2) {
  constexpr int N = NA * NB;
  constexpr int RW = plane_rows_per_wave(NA, NB);
  constexpr int NG = (N + RW - 1) / RW;
  extern __shared__ unsigned char smem_raw[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int sa = a.sa, sr = a.sr, nH = a.nH;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* A = tw + N;                               // [nH + 1][sa], y in slot order
  cplx<T>* Rw = A + (nH + 1) * sa + wave * RW * sr;  // this wave's [RW][sr] strip
  const int zs = blockIdx.x, bp = blockIdx.y;

  // lane roles of the x pass (row fastest)
  const int rx = lane % RW, jx = lane / RW;
  const bool actA = jx < NB, actB = jx < NA;
  int xoff[NA];
  unsigned xin;
  x_offsets<NA, NB>(min(jx, NB - 1), a.hmax, sa, xoff, &xin);
  // V(r) row pieces of step B; rows past the grid and idle lanes load a valid
  // row that is not used
  const T* Vz = Vp + (size_t)zs * N * N + NB * min(jx, NA - 1);
  T vv[NB];  // V(r) of the wave's current rows; the next rows' are loaded once it is used
  {
    const T* row = Vz + (size_t)min(wave * RW + rx, N - 1) * N;
#pragma unroll
    for (int kb = 0; kb < NB; ++kb) vv[kb] = row[kb];
  }

  // sphere columns of this plane: loads issued before the LDS fill
  const cplx<T>* in = t1 + ((size_t)bp * N + zs) * a.ncol;
  cplx<T> cv[kPlaneLoads];
  int ci[kPlaneLoads];
#pragma unroll
  for (int j = 0; j < kPlaneLoads; ++j) {
    const int cl = threadIdx.x + j * blockDim.x;
    const int cc = cl < a.ncol ? cl : 0;
    cv[j] = in[cc];
    ci[j] = cl < a.ncol ? a.col_hi[cc] * sa + a.col_kmod[cc] : -1;
  }
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = mk<T>((T)a.tw[2 * k], (T)a.tw[2 * k + 1]);
  for (int i = threadIdx.x; i < (nH + 1) * sa; i += blockDim.x) A[i] = mk<T>(0, 0);
  __syncthreads();
  cplx<T> xw[x_ntw<NA, NB>()];
  x_twiddles<T, NA, NB>(tw, jx, xw);
#pragma unroll
  for (int j = 0; j < kPlaneLoads; ++j)
    if (ci[j] >= 0) A[ci[j]] = cv[j];
  for (int cl = threadIdx.x + kPlaneLoads * blockDim.x; cl < a.ncol; cl += blockDim.x)
    A[a.col_hi[cl] * sa + a.col_kmod[cl]] = in[cl];
  __syncthreads();

  // y inverse (line fastest across lanes)
  const unsigned mH = div_magic((unsigned)nH);
  for (int it = threadIdx.x; it < nH * NB; it += blockDim.x) {
    const int q = fdiv(it, mH);
    np_stepA<T, NA, NB>(A + (it - q * nH) * sa, q, tw, +1);
  }
  __syncthreads();
  for (int it = threadIdx.x; it < nH * NA; it += blockDim.x) {
    const int q = fdiv(it, mH);
    np_stepB<T, NA, NB>(A + (it - q * nH) * sa, q, +1);
  }
  __syncthreads();

  for (int g = wave; g < NG; g += nwave) {
    const int y0 = g * RW;
    const bool row_ok = y0 + rx < N;
    cplx<T>* Rr = Rw + rx * sr;
    if (actA && row_ok) x_stepA<T, NA, NB>(A, xoff, Rr, y0 + rx, jx, xw);
    wave_lds_sync();
    if (actB && row_ok) {
      cplx<T>* L = Rr + NB * jx;
      cplx<T> v[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) v[b] = L[b];
      Dft<T, NB>::run(v, +1);
#pragma unroll
      for (int kb = 0; kb < NB; ++kb) {
        v[kb].x *= vv[kb];
        v[kb].y *= vv[kb];
      }
      Dft<T, NB>::run(v, -1);
#pragma unroll
      for (int b = 1; b < NB; ++b) v[b] = cmul(v[b], mk<T>(xw[b - 1].x, -xw[b - 1].y));  // W^-(b jx)
#pragma unroll
      for (int b = 0; b < NB; ++b) L[b] = v[b];
    }
    if (g + nwave < NG) {
      const T* row = Vz + (size_t)min((g + nwave) * RW + rx, N - 1) * N;
#pragma unroll
      for (int kb = 0; kb < NB; ++kb) vv[kb] = row[kb];
    }
    wave_lds_sync();
    if (actA && row_ok) x_stepAp<T, NA, NB>(A, xoff, xin, Rr, y0 + rx, jx);
    wave_lds_sync();
  }
  __syncthreads();
  // y forward (line fastest across lanes)
  for (int it = threadIdx.x; it < nH * NA; it += blockDim.x) {
    const int q = fdiv(it, mH);
    pn_stepB<T, NA, NB>(A + (it - q * nH) * sa, q, tw, -1);
  }
  __syncthreads();
  for (int it = threadIdx.x; it < nH * NB; it += blockDim.x) {
    const int q = fdiv(it, mH);
    pn_stepA<T, NA, NB>(A + (it - q * nH) * sa, q, -1);
  }
  __syncthreads();
  cplx<T>* out = t2 + ((size_t)bp * N + zs) * a.ncol;
  for (int cl = threadIdx.x; cl < a.ncol; cl += blockDim.x)
    out[cl] = A[a.col_hi[cl] * sa + a.col_kmod[cl]];
}

// ---- z forward: T2 -> split bands, x1/N, + |G|^2 psi --------------------
// kb32 (optional): an FP32 term added to the result (the KB part of H psi)
template <typename T, int NA, int NB>
__global__ void __launch_bounds__(kFastThreads) k_zfwd_f(FastArgs a, const double* __restrict__ psi,
                                                         const cplx<T>* __restrict__ t2,
                                                         double* __restrict__ out, int accumulate,
                                                         double invN,
                                                         const float* __restrict__ kb32) {
  constexpr int N = NA * NB;
  constexpr int PB = zpairs(N, (int)sizeof(cplx<T>));
  __shared__ cplx<T> tw[N];
  __shared__ cplx<T> lines[2 * PB][N];
  const int q0 = blockIdx.x * PB, q1 = min(q0 + PB, a.npair);
  const int bp = blockIdx.y;
  const int ba = 2 * (a.bp0 + bp), bb = ba + 1;
  const int c0 = q0 == 0 ? 0 : 2 * q0 - 1, c1 = 2 * q1 - 1, ncl = c1 - c0;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = mk<T>((T)a.tw[2 * k], (T)a.tw[2 * k + 1]);
  const cplx<T>* in = t2 + (size_t)bp * N * a.ncol;
  const unsigned mc = div_magic((unsigned)ncl);
  for (int i = threadIdx.x; i < ncl * N; i += blockDim.x) {
    const int zs = fdiv(i, mc), cl = i - zs * ncl;
    lines[cl][zs] = in[(size_t)zs * a.ncol + c0 + cl];
  }
  __syncthreads();
  for (int it = threadIdx.x; it < ncl * NA; it += blockDim.x)
    pn_stepB<T, NA, NB>(lines[it / NA], it % NA, tw, -1);
  __syncthreads();
  for (int it = threadIdx.x; it < ncl * NB; it += blockDim.x)
    pn_stepA<T, NA, NB>(lines[it / NB], it % NB, -1);
  __syncthreads();
  const double* pa = psi + (size_t)ba * a.nplwex;
  const double* pb = psi + (size_t)bb * a.nplwex;
  double* oa = out + (size_t)ba * a.nplwex;
  double* ob = out + (size_t)bb * a.nplwex;
  const bool hasb = bb < a.nb;
  const int p0 = a.pair_p0[q0], p1 = a.pair_p0[q1];
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    int col = a.pt_col[p] - c0, z = a.pt_z[p];
    cplx<T> zp = lines[col][z];
    cplx<T> zm = lines[mirror_col_f(col + c0) - c0][z == 0 ? 0 : N - z];
    double Ar = 0.5 * ((double)zp.x + (double)zm.x) * invN;
    double Ai = 0.5 * ((double)zp.y - (double)zm.y) * invN;
    double Br = 0.5 * ((double)zp.y + (double)zm.y) * invN;
    double Bi = -0.5 * ((double)zp.x - (double)zm.x) * invN;
    double kin = a.kin[p];
    double r0 = kin * pa[2 * p] + Ar, r1 = kin * pa[2 * p + 1] + Ai;
    if (accumulate) { r0 += oa[2 * p]; r1 += oa[2 * p + 1]; }
    if (kb32) {
      r0 += (double)kb32[(size_t)ba * a.nplwex + 2 * p];
      r1 += (double)kb32[(size_t)ba * a.nplwex + 2 * p + 1];
    }
    oa[2 * p] = r0;
    oa[2 * p + 1] = r1;
    if (hasb) {
      double s0 = kin * pb[2 * p] + Br, s1 = kin * pb[2 * p + 1] + Bi;
      if (accumulate) { s0 += ob[2 * p]; s1 += ob[2 * p + 1]; }
      if (kb32) {
        s0 += (double)kb32[(size_t)bb * a.nplwex + 2 * p];
        s1 += (double)kb32[(size_t)bb * a.nplwex + 2 * p + 1];
      }
      ob[2 * p] = s0;
      ob[2 * p + 1] = s1;
    }
  }
}

// ---- density: y inverse, x inverse, accumulate |psi|^2 in slot order --------
// Same plane layout and lane maps as k_plane_f.  A block (z plane, group of
// band pairs) keeps its share of the plane's density in registers over all its
// band pairs (RG row groups per wavefront at most) and writes it once.
constexpr int kRhoGroups = 6;  // largest RG instantiated (2, 4 and 6 are)

// more than four row groups run with at most four wavefronts per block, which
// leaves each lane the register file's full width for the accumulators
constexpr int rho_max_threads(int RG) { return RG > 4 ? 256 : kPlaneMaxThreads; }

template <typename T, int NA, int NB, int RG>
__global__ void __launch_bounds__(rho_max_threads(RG)) k_plane_rho_f(FastArgs a,
                                                                const cplx<T>* __restrict__ t1,
                                                                int nbp_batch, int bp_per_group,
                                                                const double* __restrict__ w,
                                                                double* __restrict__ rho_part) {
  constexpr int N = NA * NB;
  constexpr int RW = plane_rows_per_wave(NA, NB);
  constexpr int NG = (N + RW - 1) / RW;
  extern __shared__ unsigned char smem_raw[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int sa = a.sa, sr = a.sr, nH = a.nH;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* A = tw + N;                               // [nH + 1][sa], last row zero
  cplx<T>* Rw = A + (nH + 1) * sa + wave * RW * sr;
  const int zs = blockIdx.x, grp = blockIdx.y;
  const int rx = lane % RW, jx = lane / RW;
  const bool actA = jx < NB, actB = jx < NA;
  int xoff[NA];
  unsigned xin;
  x_offsets<NA, NB>(min(jx, NB - 1), a.hmax, sa, xoff, &xin);
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = mk<T>((T)a.tw[2 * k], (T)a.tw[2 * k + 1]);
  __syncthreads();
  cplx<T> xw[x_ntw<NA, NB>()];
  x_twiddles<T, NA, NB>(tw, jx, xw);
  double acc[RG][NB];
#pragma unroll
  for (int gi = 0; gi < RG; ++gi)
#pragma unroll
    for (int kb = 0; kb < NB; ++kb) acc[gi][kb] = 0.0;
  const int bpe = min(nbp_batch, (grp + 1) * bp_per_group);
  for (int bp = grp * bp_per_group; bp < bpe; ++bp) {
    const int ba = 2 * (a.bp0 + bp), bb = ba + 1;
    const double wa = w[ba], wb = bb < a.nb ? w[bb] : 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < (nH + 1) * sa; i += blockDim.x) A[i] = mk<T>(0, 0);
    __syncthreads();
    const cplx<T>* in = t1 + ((size_t)bp * N + zs) * a.ncol;
    for (int cl = threadIdx.x; cl < a.ncol; cl += blockDim.x)
      A[a.col_hi[cl] * sa + a.col_kmod[cl]] = in[cl];
    __syncthreads();
    const unsigned mH = div_magic((unsigned)nH);
    for (int it = threadIdx.x; it < nH * NB; it += blockDim.x) {
      const int q = fdiv(it, mH);
      np_stepA<T, NA, NB>(A + (it - q * nH) * sa, q, tw, +1);
    }
    __syncthreads();
    for (int it = threadIdx.x; it < nH * NA; it += blockDim.x) {
      const int q = fdiv(it, mH);
      np_stepB<T, NA, NB>(A + (it - q * nH) * sa, q, +1);
    }
    __syncthreads();
#pragma unroll
    for (int gi = 0; gi < RG; ++gi) {
      const int g = wave + gi * nwave;
      if (g < NG) {
        const int y0 = g * RW;
        const bool row_ok = y0 + rx < N;
        cplx<T>* Rr = Rw + rx * sr;
        if (actA && row_ok) x_stepA<T, NA, NB>(A, xoff, Rr, y0 + rx, jx, xw);
        wave_lds_sync();
        if (actB && row_ok) {
          const cplx<T>* L = Rr + NB * jx;
          cplx<T> v[NB];
#pragma unroll
          for (int b = 0; b < NB; ++b) v[b] = L[b];
          Dft<T, NB>::run(v, +1);
#pragma unroll
          for (int kb = 0; kb < NB; ++kb) {
            const double re = (double)v[kb].x, im = (double)v[kb].y;
            acc[gi][kb] += re * re * wa + im * im * wb;
          }
        }
        wave_lds_sync();
      }
    }
  }
  double* rp = rho_part + ((size_t)grp * N + zs) * N * N;
#pragma unroll
  for (int gi = 0; gi < RG; ++gi) {
    const int g = wave + gi * nwave, y = g * RW + rx;
    if (g < NG && y < N && actB) {
      double* dst = rp + (size_t)y * N + NB * jx;
#pragma unroll
      for (int kb = 0; kb < NB; ++kb) dst[kb] = acc[gi][kb];
    }
  }
}

// sum the group partials and undo the slot permutation of all three axes
__global__ void k_rho_reduce_f(const double* __restrict__ part, int ngrp, int N, int NA, int NB,
                               double* __restrict__ rho) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long NN = (long)N * N * N;
  if (i >= NN) return;
  int xs = (int)(i % N);
  long r = i / N;
  int ys = (int)(r % N), zs = (int)(r / N);
  double s = 0.0;
  for (int g = 0; g < ngrp; ++g) s += part[(size_t)g * NN + i];
  long phys = slot_phys(xs, NA, NB) + (long)N * (slot_phys(ys, NA, NB) + (long)N * slot_phys(zs, NA, NB));
  rho[phys] = s;
}

// V_perm[zs][ys][xs] = V[phys(zs)][phys(ys)][phys(xs)]
__global__ void k_permute_v(const double* __restrict__ V, double* __restrict__ Vp,
                            float* __restrict__ Vpf, int N, int NA, int NB) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long NN = (long)N * N * N;
  if (i >= NN) return;
  int xs = (int)(i % N);
  long r = i / N;
  int ys = (int)(r % N), zs = (int)(r / N);
  long phys = slot_phys(xs, NA, NB) + (long)N * (slot_phys(ys, NA, NB) + (long)N * slot_phys(zs, NA, NB));
  double v = V[phys];
  Vp[i] = v;
  Vpf[i] = (float)v;
}

// --------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------
bool fast_split(int n, int* na, int* nb) {
  switch (n) {
    case 30: *na = 6; *nb = 5; return true;
    case 60: *na = 6; *nb = 10; return true;
    case 64: *na = 8; *nb = 8; return true;
    case 90: *na = 9; *nb = 10; return true;
    case 96: *na = 12; *nb = 8; return true;
    case 120: *na = 12; *nb = 10; return true;
    case 128: *na = 16; *nb = 8; return true;
    default: return false;
  }
}

bool fast_ok(const Ctx* c) {
  int na, nb;
  // the x pass folds h and h - N onto one natural index, so the sphere's x
  // extent must fit in one period
  return c->n1 == c->n2 && c->n2 == c->n3 && fast_split(c->n1, &na, &nb) && c->d_Vp != nullptr &&
         c->pc.nH <= c->n1 && std::getenv("QX_GENERIC_FFT") == nullptr;
}

int fast_permute_v(Ctx* c) {
  int na, nb;
  if (!(c->n1 == c->n2 && c->n2 == c->n3 && fast_split(c->n1, &na, &nb))) return 0;
  if (!c->d_Vp) {
    buffers_changed(c);
    QX_CHECK(hipMalloc(&c->d_Vp, sizeof(double) * c->N));
    QX_CHECK(hipMalloc(&c->d_Vpf, sizeof(float) * c->N));
  }
  hipLaunchKernelGGL(k_permute_v, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, c->d_V,
                     c->d_Vp, c->d_Vpf, c->n1, na, nb);
  QX_CHECK(hipGetLastError());
  return 0;
}

static FastArgs fast_args(Ctx* c, int nb, int bp0) {
  FastArgs a{};
  a.N = c->n1;
  a.nH = c->pc.nH;
  a.hmax = (c->pc.nH - 1) / 2;
  a.npair = c->pc.npair;
  a.ncol = c->pc.ncol;
  a.nplwex = c->nplwex;
  a.nb = nb;
  a.bp0 = bp0;
  a.pt_col = c->d_pt_col;
  a.pt_z = c->d_pt_z;
  a.pair_p0 = c->d_pair_p0;
  a.col_hi = c->d_col_hi;
  a.col_kmod = c->d_col_kmod;
  a.kin = c->d_kin;
  a.tw = c->d_tw;  // first n1 entries: exp(+2 pi i k/N)
  return a;
}

// LDS of the band-apply plane kernel with nwave wavefronts and row strides sa, sr
template <typename T>
static size_t plane_lds(int N, int nH, int sa, int sr, int rw, int nwave) {
  return sizeof(cplx<T>) * ((size_t)N + (size_t)(nH + 1) * sa + (size_t)nwave * rw * sr);
}

constexpr size_t kCuLds = 160 * 1024;

// wavefronts per plane block: 8 unless the LDS strips would not fit
template <typename T>
static int plane_waves(int N, int nH, int rw) {
  int w = kPlaneMaxThreads / 64;
  if (const char* e = std::getenv("QX_PLANE_WAVES")) w = std::max(1, std::min(w, std::atoi(e)));
  while (w > 1 && plane_lds<T>(N, nH, N, N, rw, w) > kCuLds) w /= 2;
  return plane_lds<T>(N, nH, N, N, rw, w) <= kCuLds ? w : -1;
}

// Passes one wavefront LDS access takes when each lane moves el dwords: 64
// banks of 4 bytes serve 64/el lanes per pass, and a pass lasts as many
// cycles as the most distinct addresses one bank is asked for (lanes reading
// the same address share it).  e[l] < 0 marks an idle lane.
// (banks = 64; the variant with 32 banks serves 32/el lanes per pass)
static int lds_passes(const int* e, int el, int banks = 64) {
  int tot = 0;
  const int lpp = banks / el;
  for (int g = 0; g < 64; g += lpp) {
    int d[64], nd = 0;
    for (int l = g; l < g + lpp; ++l)
      if (e[l] >= 0)
        for (int k = 0; k < el; ++k) d[nd++] = el * e[l] + k;
    if (!nd) continue;
    std::sort(d, d + nd);
    nd = (int)(std::unique(d, d + nd) - d);
    int cnt[64] = {0}, mx = 0;
    for (int i = 0; i < nd; ++i) mx = std::max(mx, ++cnt[d[i] & (banks - 1)]);
    tot += mx;
  }
  return tot;
}

// modelled LDS passes of the accesses that depend on the sphere tile stride
// sa (y passes and the x pass's column reads/writes) ...
static long plane_cost_sa(int NA, int NB, int nH, int sa, int el, int banks) {
  const int N = NA * NB, hmax = (nH - 1) / 2, RW = plane_rows_per_wave(NA, NB);
  long c = 0;
  int e[64];
  for (int step = 0; step < 2; ++step) {
    const int K = step == 0 ? NB : NA, Q = step == 0 ? NA : NB, n = nH * K;
    for (int q = 0; q < Q; ++q)
      for (int w0 = 0; w0 < n; w0 += 64) {
        for (int l = 0; l < 64; ++l) {
          const int it = w0 + l, line = it % nH, j = it / nH;
          e[l] = it < n ? line * sa + (step == 0 ? j + NB * q : NB * j + q) : -1;
        }
        c += 4 * lds_passes(e, el, banks);  // inverse and forward, read and write
      }
  }
  for (int y0 = 0; y0 < N; y0 += RW)
    for (int q = 0; q < NA; ++q) {
      int e2[64];
      for (int l = 0; l < 64; ++l) {
        const int rx = l % RW, jx = l / RW, n = jx + NB * q;
        const int h = n <= hmax ? n : (n >= N - hmax ? n - N : N);
        const bool ok = jx < NB && y0 + rx < N;
        // step A reads every slot (the zero row nH outside the sphere), step A'
        // writes the sphere's slots only
        e[l] = ok ? (h < N ? h + hmax : nH) * sa + y0 + rx : -1;
        e2[l] = ok && h < N ? (h + hmax) * sa + y0 + rx : -1;
      }
      c += lds_passes(e, el, banks) + lds_passes(e2, el, banks);
    }
  return c;
}

// ... and on the x strip stride sr
static long plane_cost_sr(int NA, int NB, int sr, int el, int banks) {
  const int N = NA * NB, RW = plane_rows_per_wave(NA, NB);
  long c = 0;
  int e[64];
  for (int y0 = 0; y0 < N; y0 += RW) {
    for (int q = 0; q < NA; ++q) {
      for (int l = 0; l < 64; ++l) {
        const int rx = l % RW, jx = l / RW;
        e[l] = (jx < NB && y0 + rx < N) ? rx * sr + jx + NB * q : -1;
      }
      c += 2 * lds_passes(e, el, banks);
    }
    for (int q = 0; q < NB; ++q) {
      for (int l = 0; l < 64; ++l) {
        const int rx = l % RW, jx = l / RW;
        e[l] = (jx < NA && y0 + rx < N) ? rx * sr + NB * jx + q : -1;
      }
      c += 2 * lds_passes(e, el, banks);
    }
  }
  return c;
}

// Row strides of the plane kernel: pairs in [N, N+32]^2 ranked by the
// modelled LDS passes, among those that keep as many blocks per CU as
// unpadded rows do; the best `keep` pairs are returned (the first is the
// model's choice).  Cached per shape.
template <typename T>
static std::vector<std::pair<int, int>> plane_stride_ranks(int NA, int NB, int nH, int nwave,
                                                           int keep, int banks = 64) {
  static std::mutex mu;
  static std::map<std::array<int, 6>, std::vector<std::pair<int, int>>> cache;
  const std::array<int, 6> key{NA * 100 + NB, nH, (int)sizeof(T), nwave, keep, banks};
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it == cache.end()) {
    const int N = NA * NB, RW = plane_rows_per_wave(NA, NB), el = (int)sizeof(cplx<T>) / 4;
    const size_t occ0 = kCuLds / plane_lds<T>(N, nH, N, N, RW, nwave);
    long csr[33];
    for (int d = 0; d <= 32; ++d) csr[d] = plane_cost_sr(NA, NB, N + d, el, banks);
    std::vector<std::pair<long, std::pair<int, int>>> all;
    for (int da = 0; da <= 32; ++da) {
      const long ca = plane_cost_sa(NA, NB, nH, N + da, el, banks);
      for (int dr = 0; dr <= 32; ++dr) {
        const size_t l = plane_lds<T>(N, nH, N + da, N + dr, RW, nwave);
        if (l > kCuLds || kCuLds / l < occ0) continue;
        all.push_back({ca + csr[dr], {N + da, N + dr}});
      }
    }
    std::stable_sort(all.begin(), all.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    std::vector<std::pair<int, int>> r;
    for (const auto& x : all) {
      if ((int)r.size() >= keep) break;
      r.push_back(x.second);
    }
    if (r.empty()) r.push_back({N, N});
    it = cache.emplace(key, r).first;
  }
  return it->second;
}

// the model's stride pair (QX_PLANE_SA / QX_PLANE_SR override)
template <typename T>
static void plane_strides(int NA, int NB, int nH, int nwave, int* psa, int* psr) {
  const auto r = plane_stride_ranks<T>(NA, NB, nH, nwave, 1);
  *psa = r[0].first;
  *psr = r[0].second;
  if (const char* e = std::getenv("QX_PLANE_SA")) *psa = std::max(NA * NB, std::atoi(e));
  if (const char* e = std::getenv("QX_PLANE_SR")) *psr = std::max(NA * NB, std::atoi(e));
}

// Stride pairs the plane tuner times: the three best of the bank model with
// 64 banks and with 32, and a coarse grid of odd paddings (the model's
// conflict counts have not matched the measured ones, so time decides), all
// keeping the blocks per CU of unpadded rows.
template <typename T>
static std::vector<std::pair<int, int>> plane_stride_candidates(int NA, int NB, int nH,
                                                                int nwave) {
  const int N = NA * NB, RW = plane_rows_per_wave(NA, NB);
  std::vector<std::pair<int, int>> v = plane_stride_ranks<T>(NA, NB, nH, nwave, 3, 64);
  for (const auto& p : plane_stride_ranks<T>(NA, NB, nH, nwave, 3, 32))
    if (std::find(v.begin(), v.end(), p) == v.end()) v.push_back(p);
  const size_t occ0 = kCuLds / plane_lds<T>(N, nH, N, N, RW, nwave);
  for (int da : {1, 3, 5, 9, 17})
    for (int dr : {1, 5, 9, 17}) {
      const std::pair<int, int> p{N + da, N + dr};
      const size_t l = plane_lds<T>(N, nH, p.first, p.second, RW, nwave);
      if (l > kCuLds || kCuLds / l < occ0) continue;
      if (std::find(v.begin(), v.end(), p) == v.end()) v.push_back(p);
    }
  return v;
}

// mid (optional): work launched between the plane pass and the z forward pass,
// e.g. the KB contractions of the fused mixed band-apply (needs one batch)
struct FastMid {
  float* psi32;        // FP32 copy of psi, written by the z inverse pass, or
  unsigned short* psis;  // its bf16 hi | lo split (row stride ldk, lo at slo)
  int ldk;
  long slo;
  const float* kb32;   // FP32 term the z forward pass adds
  int (*run)(Ctx*, const double*, int, const FastMid&);
};

// Launch shape of the band-apply plane kernel: wavefronts per block and the
// two LDS row strides.  On the first band-apply of a shape the candidates
// (8 and 4 wavefronts, the model's three best stride pairs for each) are
// timed on the real z-pencils (k_zinv_f fills T1 first; the trial launches
// write only T2) and the fastest is kept for the process; QX_FFT_NOTUNE=1
// takes the model's choice.
struct PlaneCfg {
  int nwave = 0, sa = 0, sr = 0;
  float ms = 0.f;
};

template <typename T, int NA, int NB>
static PlaneCfg plane_config(Ctx* c, const double* psi, int nb, const T* Vp, cplx<T>* t1,
                             cplx<T>* t2, int nbb) {
  constexpr int N = NA * NB;
  constexpr int RW = plane_rows_per_wave(NA, NB);
  static std::mutex mu;
  static std::map<std::array<int, 3>, PlaneCfg> tuned;
  const std::array<int, 3> key{N, c->pc.nH, (int)sizeof(T)};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = tuned.find(key);
    if (it != tuned.end()) return it->second;
  }
  PlaneCfg best;
  best.nwave = plane_waves<T>(N, c->pc.nH, RW);
  if (best.nwave < 1) return best;
  plane_strides<T>(NA, NB, c->pc.nH, best.nwave, &best.sa, &best.sr);
  const bool fixed = std::getenv("QX_FFT_NOTUNE") || std::getenv("QX_PLANE_WAVES") ||
                     std::getenv("QX_PLANE_SA") || std::getenv("QX_PLANE_SR");
  if (!fixed) {
    FastArgs a = fast_args(c, nb, 0);
    constexpr int PB = zpairs(N, (int)sizeof(cplx<T>));
    const unsigned gz = (unsigned)((c->pc.npair + PB - 1) / PB);
    hipLaunchKernelGGL((k_zinv_f<T, NA, NB>), dim3(gz, nbb), dim3(kFastThreads), 0, c->stream, a,
                       psi, t1, nullptr, nullptr, 0, 0L);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
      float bms = 1e30f;
      for (int nw = kPlaneMaxThreads / 64; nw >= 4; nw /= 2) {
        if (plane_lds<T>(N, c->pc.nH, N, N, RW, nw) > kCuLds) continue;
        for (const auto& st : plane_stride_candidates<T>(NA, NB, c->pc.nH, nw)) {
          const size_t lp = plane_lds<T>(N, c->pc.nH, st.first, st.second, RW, nw);
          if (lp > kCuLds) continue;
          if (lp > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_plane_f<T, NA, NB>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lp);
          a.sa = st.first;
          a.sr = st.second;
          hipLaunchKernelGGL((k_plane_f<T, NA, NB>), dim3(N, nbb), dim3(64 * nw), lp, c->stream,
                             a, Vp, t1, t2);
          (void)hipEventRecord(e0, c->stream);
          for (int r = 0; r < 2; ++r)
            hipLaunchKernelGGL((k_plane_f<T, NA, NB>), dim3(N, nbb), dim3(64 * nw), lp,
                               c->stream, a, Vp, t1, t2);
          (void)hipEventRecord(e1, c->stream);
          (void)hipEventSynchronize(e1);
          float ms = 0.f;
          (void)hipEventElapsedTime(&ms, e0, e1);
          if (std::getenv("QX_FFT_VERBOSE"))
            fprintf(stderr, "plane tune N=%d nH=%d T%zu: %d waves, sa %d sr %d: %.4f ms\n", N,
                    c->pc.nH, sizeof(T), nw, st.first, st.second, ms / 2);
          if (hipGetLastError() == hipSuccess && ms > 0.f && ms < bms) {
            bms = ms;
            best.nwave = nw;
            best.sa = st.first;
            best.sr = st.second;
            best.ms = ms / 2;
          }
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
  }
  std::lock_guard<std::mutex> lk(mu);
  tuned[key] = best;
  return best;
}

template <typename T, int NA, int NB>
static int local_fast_nn(Ctx* c, const double* psi, double* out, int nb, int accumulate,
                         const FastMid* mid = nullptr) {
  constexpr int N = NA * NB;
  constexpr int RW = plane_rows_per_wave(NA, NB);
  const int nbp = (nb + 1) / 2;
  const size_t per_bp = sizeof(cplx<T>) * (size_t)N * c->pc.ncol;
  const int batch = (int)std::max<size_t>(1, std::min<size_t>(c->t_bytes / per_bp, (size_t)nbp));
  constexpr int PB = zpairs(N, (int)sizeof(cplx<T>));
  const unsigned gz = (unsigned)((c->pc.npair + PB - 1) / PB);
  cplx<T>* t1 = reinterpret_cast<cplx<T>*>(c->d_t1);
  cplx<T>* t2 = reinterpret_cast<cplx<T>*>(c->d_t2);
  const T* Vp = sizeof(T) == 8 ? reinterpret_cast<const T*>(c->d_Vp) : reinterpret_cast<const T*>(c->d_Vpf);
  if (mid && batch < nbp) return -3;
  const PlaneCfg cfg = plane_config<T, NA, NB>(c, psi, nb, Vp, t1, t2, std::min(batch, nbp));
  const int nwave = cfg.nwave;
  if (nwave < 1) return -3;
  const int sa = cfg.sa, sr = cfg.sr;
  const size_t lp = plane_lds<T>(N, c->pc.nH, sa, sr, RW, nwave);
  if (lp > kCuLds) return -3;
  if (lp > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_plane_f<T, NA, NB>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lp);
  for (int bp0 = 0; bp0 < nbp; bp0 += batch) {
    const int nbb = std::min(batch, nbp - bp0);
    FastArgs a = fast_args(c, nb, bp0);
    a.sa = sa;
    a.sr = sr;
    lt_mark(c, 1, false);
    hipLaunchKernelGGL((k_zinv_f<T, NA, NB>), dim3(gz, nbb), dim3(kFastThreads), 0, c->stream, a,
                       psi, t1, mid ? mid->psi32 : nullptr, mid ? mid->psis : nullptr,
                       mid ? mid->ldk : 0, mid ? mid->slo : 0L);
    hipLaunchKernelGGL((k_plane_f<T, NA, NB>), dim3(N, nbb), dim3(64 * nwave), lp, c->stream, a,
                       Vp, t1, t2);
    if (mid) {  // the mid work is not part of the FFT stage's launch time
      lt_mark(c, 1, true);
      const int rc = mid->run(c, psi, nb, *mid);
      if (rc) return rc;
      lt_mark(c, 1, false, true);
    }
    hipLaunchKernelGGL((k_zfwd_f<T, NA, NB>), dim3(gz, nbb), dim3(kFastThreads), 0, c->stream, a,
                       psi, t2, out, accumulate, 1.0 / (double)c->N, mid ? mid->kb32 : nullptr);
    lt_mark(c, 1, true);
  }
  QX_CHECK(hipGetLastError());
  return 0;
}

template <typename T, int NA, int NB>
static int density_fast_nn(Ctx* c, const double* psi, int nb, const double* w) {
  constexpr int N = NA * NB;
  constexpr int RW = plane_rows_per_wave(NA, NB);
  const int nbp = (nb + 1) / 2;
  const int nwave = plane_waves<T>(N, c->pc.nH, RW);
  // every wavefront holds at most kRhoGroups row groups in registers
  if (nwave < 1 || nwave * kRhoGroups < (N + RW - 1) / RW) return -3;
  int sa, sr;
  plane_strides<T>(NA, NB, c->pc.nH, nwave, &sa, &sr);
  const size_t per_bp = sizeof(cplx<T>) * (size_t)N * c->pc.ncol;
  const int batch = (int)std::max<size_t>(1, std::min<size_t>(c->t_bytes / per_bp, (size_t)nbp));
  constexpr int PB = zpairs(N, (int)sizeof(cplx<T>));
  const unsigned gz = (unsigned)((c->pc.npair + PB - 1) / PB);
  const size_t lp = plane_lds<T>(N, c->pc.nH, sa, sr, RW, nwave);
  if (lp > kCuLds) return -3;
  // the fewest row groups per wavefront that cover the plane's rows
  const int ng_rows = (N + RW - 1) / RW;
  const int rg = nwave * 2 >= ng_rows ? 2 : nwave * 4 >= ng_rows ? 4 : kRhoGroups;
  if (64 * nwave > rho_max_threads(rg)) return -3;
  const void* kfn = rg == 2   ? reinterpret_cast<const void*>(k_plane_rho_f<T, NA, NB, 2>)
                    : rg == 4 ? reinterpret_cast<const void*>(k_plane_rho_f<T, NA, NB, 4>)
                              : reinterpret_cast<const void*>(k_plane_rho_f<T, NA, NB, kRhoGroups>);
  if (lp > 65536)
    (void)hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lp);
  cplx<T>* t1 = reinterpret_cast<cplx<T>*>(c->d_t1);
  const int nbatch = (nbp + batch - 1) / batch;
  const int groups_per_batch = std::max(1, c->rho_groups / nbatch);
  if (groups_per_batch * nbatch > c->rho_groups) return -4;
  int nparts = 0;
  for (int bp0 = 0; bp0 < nbp; bp0 += batch) {
    const int nbb = std::min(batch, nbp - bp0);
    FastArgs a = fast_args(c, nb, bp0);
    a.sa = sa;
    a.sr = sr;
    hipLaunchKernelGGL((k_zinv_f<T, NA, NB>), dim3(gz, nbb), dim3(kFastThreads), 0, c->stream, a,
                       psi, t1, nullptr, nullptr, 0, 0L);
    int ng = std::min(groups_per_batch, nbb);
    int per = (nbb + ng - 1) / ng;
    ng = (nbb + per - 1) / per;
    double* part = c->d_rho_part + (size_t)nparts * c->N;
    if (rg == 2)
      hipLaunchKernelGGL((k_plane_rho_f<T, NA, NB, 2>), dim3(N, ng), dim3(64 * nwave), lp,
                         c->stream, a, t1, nbb, per, w, part);
    else if (rg == 4)
      hipLaunchKernelGGL((k_plane_rho_f<T, NA, NB, 4>), dim3(N, ng), dim3(64 * nwave), lp,
                         c->stream, a, t1, nbb, per, w, part);
    else
      hipLaunchKernelGGL((k_plane_rho_f<T, NA, NB, kRhoGroups>), dim3(N, ng), dim3(64 * nwave), lp,
                         c->stream, a, t1, nbb, per, w, part);
    nparts += ng;
  }
  hipLaunchKernelGGL(k_rho_reduce_f, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream,
                     c->d_rho_part, nparts, N, NA, NB, c->d_rho);
  QX_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
static int local_fast_t(Ctx* c, const double* psi, double* out, int nb, int accumulate,
                        const FastMid* mid = nullptr) {
  switch (c->n1) {
    case 30: return local_fast_nn<T, 6, 5>(c, psi, out, nb, accumulate, mid);
    case 60: return local_fast_nn<T, 6, 10>(c, psi, out, nb, accumulate, mid);
    case 64: return local_fast_nn<T, 8, 8>(c, psi, out, nb, accumulate, mid);
    case 90: return local_fast_nn<T, 9, 10>(c, psi, out, nb, accumulate, mid);
    case 96: return local_fast_nn<T, 12, 8>(c, psi, out, nb, accumulate, mid);
    case 120: return local_fast_nn<T, 12, 10>(c, psi, out, nb, accumulate, mid);
    case 128: return local_fast_nn<T, 16, 8>(c, psi, out, nb, accumulate, mid);
    default: return -10;
  }
}

static int fused_kb_mid(Ctx* c, const double* psi, int nb, const FastMid& m) {
  return kb_mixed_f32(c, psi, m.psi32, m.psis, m.ldk, m.slo, nb, const_cast<float*>(m.kb32));
}

int hpsi_mixed_fused(Ctx* c, const double* psi, double* out, int nb) {
  if (!fast_ok(c) || c->npj == 0 || std::getenv("QX_HPSI_UNFUSED") ||
      gemm_path() == GEMM_OWN_F32 || !use_big(c->npj, nb) ||
      !use_big(c->nplwex, nb) || (c->nplwex & 1) || (c->npj & 1))
    return -3;
  FastMid m{};
  if (gemm_path() == GEMM_OWN) {
    // the bf16 split of psi, band rows of round_up(nplwex, 64) (zero pads),
    // the lo half after room for nbmax bands, so the pads never move
    m.ldk = (c->nplwex + 63) / 64 * 64;
    m.slo = (long)c->nbmax * m.ldk;
    m.psis = mg_u16_scratch(c, 0, 2 * (size_t)m.slo);
    if (!m.psis) return -3;
  } else {
    m.psi32 = mg_f32_scratch(c, 0, (size_t)c->nplwex * nb);
    if (!m.psi32) return -3;
  }
  m.kb32 = mg_f32_scratch(c, 1, (size_t)c->nplwex * nb);
  m.run = fused_kb_mid;
  if (!m.kb32) return -3;
  return local_fast_t<float>(c, psi, out, nb, 0, &m);
}

template <typename T>
static int density_fast_t(Ctx* c, const double* psi, int nb, const double* w) {
  switch (c->n1) {
    case 30: return density_fast_nn<T, 6, 5>(c, psi, nb, w);
    case 60: return density_fast_nn<T, 6, 10>(c, psi, nb, w);
    case 64: return density_fast_nn<T, 8, 8>(c, psi, nb, w);
    case 90: return density_fast_nn<T, 9, 10>(c, psi, nb, w);
    case 96: return density_fast_nn<T, 12, 8>(c, psi, nb, w);
    case 120: return density_fast_nn<T, 12, 10>(c, psi, nb, w);
    case 128: return density_fast_nn<T, 16, 8>(c, psi, nb, w);
    default: return -10;
  }
}

int local_apply_fast(Ctx* c, const double* psi, double* out, int nb, int prec, int accumulate) {
  if (prec == PREC_FP32) return local_fast_t<float>(c, psi, out, nb, accumulate);
  return local_fast_t<double>(c, psi, out, nb, accumulate);
}

int density_fast(Ctx* c, const double* psi, int nb, const double* w, int prec) {
  if (prec == PREC_FP32) return density_fast_t<float>(c, psi, nb, w);
  return density_fast_t<double>(c, psi, nb, w);
}

}  // namespace qx
