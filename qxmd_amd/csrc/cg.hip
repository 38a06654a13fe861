// cg.hip -- BKL conjugate-gradient band update (Sources/eigen.f90:199-938).
//
// Two schedules share the vector kernels below:
//   * band-by-band (the reference's order): band ib's search direction is made
//     orthogonal to the already-updated bands 1..ib-1 (subgs3) and to itself
//     (subgs4); its H h is one single-vector band-apply.  Scalars of the 2x2
//     Rayleigh-Ritz step are evaluated on the host exactly as the reference.
//   * batched: every band advances one BKL step per iteration; directions are
//     projected against all current bands with two GEMMs, H h is one batched
//     band-apply over all bands, and the per-band 2x2 step runs in a device
//     kernel, so an iteration never waits on the host.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>
#include "context.h"

namespace qx {

int local_apply(Ctx* c, const double* psi, double* out, int nb, int prec, int accumulate);

constexpr int kDotThreads = 256;
constexpr int kDotChunk = 4096;

// partial[b*nch + ch] = sum over chunk of x_b y_b (plain sum)
__global__ void __launch_bounds__(kDotThreads) k_dot_partial(const double* __restrict__ x,
                                                             const double* __restrict__ y, int n,
                                                             int ldx, int ldy, int nch,
                                                             double* __restrict__ part) {
  __shared__ double red[kDotThreads / 64];
  const int b = blockIdx.y, ch = blockIdx.x;
  const double* xb = x + (size_t)b * ldx;
  const double* yb = y + (size_t)b * ldy;
  const int i0 = ch * kDotChunk, i1 = min(n, i0 + kDotChunk);
  double s = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) s += xb[i] * yb[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kDotThreads / 64; ++w) t += red[w];
    part[(size_t)b * nch + ch] = t;
  }
}

// out[b] = 2 * sum - x_b[0] y_b[0]  (packed inner product)
__global__ void k_dot_final(const double* __restrict__ part, int nch, int nb,
                            const double* __restrict__ x, const double* __restrict__ y, int ldx,
                            int ldy, double* __restrict__ out) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s = 0.0;
  for (int c = 0; c < nch; ++c) s += part[(size_t)b * nch + c];
  out[b] = 2.0 * s - x[(size_t)b * ldx] * y[(size_t)b * ldy];
}

// out_b = a_b * x_b + c_b * y_b (coefficient arrays may be null: 1 / 0);
// `mask` (optional) skips bands whose flag is zero.
__global__ void k_axpby(double* out, const double* x, const double* y, const double* a,
                        const double* cc, double sa, double sc, int n, int ld, int nb,
                        const int* mask) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * nb) return;
  int b = (int)(i / n), k = (int)(i % n);
  if (mask && !mask[b]) return;
  double av = a ? a[b] : sa, cv = cc ? cc[b] : sc;
  size_t o = (size_t)b * ld + k;
  double yv = y ? y[o] : 0.0;
  out[o] = av * x[o] + cv * yv;
}

// prcd_b = Teter-Payne-Allan preconditioner from ekib[b] (prpcg, Sources/eigen.f90:1093-1113)
__global__ void k_precond_mul(double* g, const double* kin2, const double* ekib, int n, int ld,
                              int nb, const int* mask) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * nb) return;
  int b = (int)(i / n), k = (int)(i % n);
  if (mask && !mask[b]) return;
  double e15 = 1.5 * ekib[b];
  double x = kin2[k >> 1] / e15;
  double tesu = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
  double x2 = x * x, x4 = x2 * x2;
  tesu = tesu / (tesu + 16.0 * x4);
  g[(size_t)b * ld + k] *= tesu * 2.0 / e15;
}

static inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

// ---- fused passes of the batched schedule: each reads its vectors once and
// leaves per-chunk partial sums in the layout of k_dot_partial (NQ sums per
// band and chunk: part[(b * nch + ch) * NQ + q]) --------------------------------
template <int NQ>
__device__ inline void dot_store(double (&s)[NQ], double* part, int b, int ch, int nch) {
  __shared__ double red[NQ][kDotThreads / 64];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double v = s[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    double t = 0.0;
    for (int w = 0; w < kDotThreads / 64; ++w) t += red[threadIdx.x][w];
    part[((size_t)b * nch + ch) * NQ + threadIdx.x] = t;
  }
}

// g <- preconditioned g for the active bands (as k_precond_mul); sums of g.g
template <int CH>
__global__ void __launch_bounds__(kDotThreads) k_precond_dot(double* __restrict__ g,
                                                             const double* __restrict__ kin2,
                                                             const double* __restrict__ ekib,
                                                             int n, const int* __restrict__ mask,
                                                             int nch, double* __restrict__ part) {
  const int b = blockIdx.y, ch = blockIdx.x;
  const bool act = !mask || mask[b];
  const double e15 = 1.5 * ekib[b];
  double* gb = g + (size_t)b * n;
  const int i0 = ch * CH, i1 = min(n, i0 + CH);
  double s[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double v = gb[i];
    if (act) {
      const double x = kin2[i >> 1] / e15;
      double tesu = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
      const double x2 = x * x, x4 = x2 * x2;
      tesu = tesu / (tesu + 16.0 * x4);
      v *= tesu * 2.0 / e15;
      gb[i] = v;
    }
    s[0] += v * v;
  }
  dot_store<1>(s, part, b, ch, nch);
}

// sums of h.hpsi, h.hh, h.h
template <int CH>
__global__ void __launch_bounds__(kDotThreads) k_dot3(const double* __restrict__ h,
                                                      const double* __restrict__ hp,
                                                      const double* __restrict__ hh, int n,
                                                      int nch, double* __restrict__ part) {
  const int b = blockIdx.y, ch = blockIdx.x;
  const size_t o = (size_t)b * n;
  const int i0 = ch * CH, i1 = min(n, i0 + CH);
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double hv = h[o + i];
    s[0] += hv * hp[o + i];
    s[1] += hv * hh[o + i];
    s[2] += hv * hv;
  }
  dot_store<3>(s, part, b, ch, nch);
}

// active bands: x <- a x + t h, hx <- a hx + t hh, g <- eig x - hx (t = a dt,
// as the three k_axpby passes it replaces); every band: sums of g.g
__global__ void __launch_bounds__(kDotThreads) k_cg_update(double* __restrict__ x,
                                                           double* __restrict__ hx,
                                                           double* __restrict__ g,
                                                           const double* __restrict__ h,
                                                           const double* __restrict__ hh,
                                                           const double* __restrict__ a,
                                                           const double* __restrict__ t,
                                                           const double* __restrict__ eig, int n,
                                                           const int* __restrict__ mask, int nch,
                                                           double* __restrict__ part) {
  const int b = blockIdx.y, ch = blockIdx.x;
  const bool act = !mask || mask[b];
  const double av = a[b], tv = t[b], ev = eig[b];
  const size_t o = (size_t)b * n;
  const int i0 = ch * kDotChunk, i1 = min(n, i0 + kDotChunk);
  double s[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double gv;
    if (act) {
      const double xv = av * x[o + i] + tv * h[o + i];
      const double hv = av * hx[o + i] + tv * hh[o + i];
      x[o + i] = xv;
      hx[o + i] = hv;
      gv = ev * xv + -1.0 * hv;
      g[o + i] = gv;
    } else {
      gv = g[o + i];
    }
    s[0] += gv * gv;
  }
  dot_store<1>(s, part, b, ch, nch);
}

// one value split into bf16 hi and lo = bf16(a - hi) (as split_bf16x2)
__device__ inline void split_bf16(float a, unsigned short& hi, unsigned short& lo) {
  const __bf16 h = (__bf16)a;
  const __bf16 l = (__bf16)(a - (float)h);
  hi = __builtin_bit_cast(unsigned short, h);
  lo = __builtin_bit_cast(unsigned short, l);
}

// k_cg_update of the mixed form's split products: the updated x of the
// active bands is also written to its bf16 hi | lo planes (xs: band stride
// ldk, lo plane at xs + slo), which the next projection and the overlap read
__global__ void __launch_bounds__(kDotThreads) k_cg_update_split(
    double* __restrict__ x, double* __restrict__ hx, double* __restrict__ g,
    const double* __restrict__ h, const double* __restrict__ hh, const double* __restrict__ a,
    const double* __restrict__ t, const double* __restrict__ eig, int n,
    const int* __restrict__ mask, int nch, double* __restrict__ part,
    unsigned short* __restrict__ xs, int ldk, long slo) {
  const int b = blockIdx.y, ch = blockIdx.x;
  const bool act = !mask || mask[b];
  const double av = a[b], tv = t[b], ev = eig[b];
  const size_t o = (size_t)b * n, os = (size_t)b * ldk;
  const int i0 = ch * kDotChunk, i1 = min(n, i0 + kDotChunk);
  double s[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double gv;
    if (act) {
      const double xv = av * x[o + i] + tv * h[o + i];
      const double hv = av * hx[o + i] + tv * hh[o + i];
      x[o + i] = xv;
      hx[o + i] = hv;
      gv = ev * xv + -1.0 * hv;
      g[o + i] = gv;
      unsigned short sh, sl;
      split_bf16((float)xv, sh, sl);
      xs[os + i] = sh;
      xs[slo + os + i] = sl;
    } else {
      gv = g[o + i];
    }
    s[0] += gv * gv;
  }
  dot_store<1>(s, part, b, ch, nch);
}

// CG direction of the active bands, h = g (beta null) or g + beta h, and its
// bf16 hi | lo planes (hs, as xs above)
__global__ void k_cg_dir_split(double* __restrict__ h, const double* __restrict__ g,
                               const double* __restrict__ beta, int n, int nb,
                               const int* __restrict__ mask, unsigned short* __restrict__ hs,
                               int ldk, long slo) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * nb) return;
  const int b = (int)(i / n), k = (int)(i % n);
  if (mask && !mask[b]) return;
  const size_t o = (size_t)b * n + k, os = (size_t)b * ldk + k;
  const double v = beta ? g[o] + beta[b] * h[o] : g[o];
  h[o] = v;
  unsigned short sh, sl;
  split_bf16((float)v, sh, sl);
  hs[os] = sh;
  hs[slo + os] = sl;
}

// out_q[b] = 2 sum_q - x_b[0] y_qb[0] for the NQ sums of the fused passes
__global__ void k_dotq_final(const double* __restrict__ part, int nch, int nb, int nq,
                             const double* __restrict__ x, const double* y0,
                             const double* y1, const double* y2, int ld,
                             double* out0, double* out1, double* out2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double* ys[3] = {y0, y1, y2};
  double* outs[3] = {out0, out1, out2};
  for (int q = 0; q < nq; ++q) {
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += part[((size_t)b * nch + c) * nq + q];
    outs[q][b] = 2.0 * s - x[(size_t)b * ld] * ys[q][(size_t)b * ld];
  }
}

int pdot_batched(Ctx* c, const double* x, const double* y, int nb, double* out_dev) {
  const int n = c->nplwex, nch = (n + kDotChunk - 1) / kDotChunk;
  double* part = c->d_scal + 64 * (size_t)c->nbmax;  // scratch after the scalar arrays
  hipLaunchKernelGGL(k_dot_partial, dim3(nch, nb), dim3(kDotThreads), 0, c->stream, x, y, n, n, n,
                     nch, part);
  hipLaunchKernelGGL(k_dot_final, dim3(cdivu(nb, 256)), dim3(256), 0, c->stream, part, nch, nb, x, y,
                     n, n, out_dev);
  return 0;
}

static void axpby(Ctx* c, double* out, const double* x, const double* y, const double* a,
                  const double* cc, double sa, double sc, int nb, const int* mask) {
  const int n = c->nplwex;
  hipLaunchKernelGGL(k_axpby, dim3(cdivu((long)n * nb, 256)), dim3(256), 0, c->stream, out, x, y, a,
                     cc, sa, sc, n, n, nb, mask);
}


// ---- band-by-band schedule: the scalars of the band being updated stay on
// the device; every inner product is a chunked pass whose partial sums the
// next kernel finishes itself, and the host waits once per CG iteration for
// the convergence flag.  part[0..nch) are chunk sums, part[nch] is the
// x[0] y[0] correction of the packed product (stored by the pass, so a later
// pass may overwrite either vector).
constexpr int kSvChunk = 1024;   // vector elements per chunk of the skinny projection
constexpr int kProjRows = 8;     // columns per workgroup of the skinny projection
constexpr int kProjBands = 64;   // columns per workgroup of the skinny back-projection

struct BbScal {
  double eigv, zansa, zansab, z000re, x000re, eigb, dinnerx, betk, zansap, alpsq, dtnew;
  int iter, izcnt, lreset, active;
};

// fixed-order total of part[ch * nq + q] over ch < nch by one wavefront: lane
// l sums ch = l, l + 64, ...; the butterfly leaves the same bits in every lane
// (each level adds the same two values in both partners), so every wavefront
// that calls it gets the same total
__device__ inline double bb_total(const double* part, int nch, int nq = 1, int q = 0) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int ch = lane; ch < nch; ch += 64) s += part[ch * nq + q];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

__global__ void __launch_bounds__(kDotThreads) k_bb_dot(const double* __restrict__ x,
                                                        const double* __restrict__ y, int n,
                                                        int nch, double* __restrict__ part) {
  const int ch = blockIdx.x;
  const int i0 = ch * kBbChunk, i1 = min(n, i0 + kBbChunk);
  double s[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) s[0] += x[i] * y[i];
  dot_store<1>(s, part, 0, ch, nch);
  if (ch == 0 && threadIdx.x == 0) part[nch] = x[0] * y[0];
}

// Products of one vector with a skinny matrix A (M columns of length n,
// column j at A + j n), each pass reading A once, in fixed summation order:
//   projection       pt[j nch + ch] = sum over chunk ch of A_j . v
//   back-projection  pn[s n + i] = sum_{j in column chunk s} A_j[i] coef_j,
//                    coef_j = f_j (2 sum_ch pt[j nch + ch] - A_j[0] v[0]
//                             - im0 2 A_j[1] v[1]),  f_j = scale ? scale[j] : -1
// (the packed product's G = 0 corrections; im0 also drops the Im part of the
// G = 0 term, as the KB sums do).  The caller adds sum_s pn[s] to its vector.
// subgs3 (Sources/engrad.f90:1072-1142): A = the updated bands, f = -1;
// hckbpp_ib_g (Sources/nlkbpp.f90:1623): A = the KB projectors, f = vclalp, im0.
__global__ void __launch_bounds__(kDotThreads) k_sv_proj_t(const double* __restrict__ A, int n,
                                                           int M, const double* __restrict__ v,
                                                           int nch, double* __restrict__ pt) {
  constexpr int PT = kSvChunk / kDotThreads;
  const int ch = blockIdx.x, j0 = blockIdx.y * kProjRows;
  const int i0 = ch * kSvChunk + threadIdx.x;
  double hv[PT];
#pragma unroll
  for (int u = 0; u < PT; ++u) {
    const int i = i0 + u * kDotThreads;
    hv[u] = i < n ? v[i] : 0.0;
  }
  // every load of the workgroup's columns is issued before the first reduction
  double av[kProjRows][PT];
#pragma unroll
  for (int r = 0; r < kProjRows; ++r) {
    const double* a = A + (size_t)min(j0 + r, M - 1) * n;
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int i = i0 + u * kDotThreads;
      av[r][u] = (j0 + r < M && i < n) ? a[i] : 0.0;
    }
  }
  __shared__ double red[kProjRows][kDotThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < kProjRows; ++r) {
    double sum = 0.0;
#pragma unroll
    for (int u = 0; u < PT; ++u) sum += av[r][u] * hv[u];
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_down(sum, o, 64);
    if (lane == 0) red[r][wv] = sum;
  }
  __syncthreads();
  if (threadIdx.x < kProjRows && j0 + threadIdx.x < M) {
    double t = 0.0;
    for (int w = 0; w < kDotThreads / 64; ++w) t += red[threadIdx.x][w];
    pt[(size_t)(j0 + threadIdx.x) * nch + ch] = t;
  }
}

__global__ void __launch_bounds__(kDotThreads) k_sv_proj_n(const double* __restrict__ A, int n,
                                                           int M, const double* __restrict__ v,
                                                           int nch, const double* __restrict__ pt,
                                                           const double* __restrict__ scale,
                                                           int im0, double* __restrict__ pn) {
  const int sidx = blockIdx.y, j0 = sidx * kProjBands, nj = min(M - j0, kProjBands);
  // coef_j by kSub consecutive lanes per column: strided chunk sums, then a
  // butterfly over the kSub lanes (fixed order)
  constexpr int kSub = kDotThreads / kProjBands;
  __shared__ double coef[kProjBands];
  {
    const int jj = threadIdx.x / kSub, sub = threadIdx.x % kSub;
    const int j = j0 + min(jj, nj - 1);
    double t = 0.0;
    for (int ch = sub; ch < nch; ch += kSub) t += pt[(size_t)j * nch + ch];
    for (int o = kSub / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (sub == 0 && jj < nj) {
      const double* a = A + (size_t)j * n;
      double c0 = a[0] * v[0];
      if (im0) c0 = c0 + 2.0 * a[1] * v[1];
      coef[jj] = (2.0 * t - c0) * (scale ? scale[j] : -1.0);
    }
  }
  __syncthreads();
  const int i = blockIdx.x * kDotThreads + threadIdx.x;
  if (i >= n) return;
  const double* a = A + (size_t)j0 * n + i;
  double acc = 0.0;
  if (nj == kProjBands) {
#pragma unroll
    for (int j = 0; j < kProjBands; ++j) acc += a[(size_t)j * n] * coef[j];
  } else {
    for (int j = 0; j < nj; ++j) acc += a[(size_t)j * n] * coef[j];
  }
  pn[(size_t)sidx * n + i] = acc;
}

// h <- h + sum_s pn[s] (subgs3's back-projection); chunk sums of x.h for
// subgs4 (as k_bb_dot)
__global__ void __launch_bounds__(kDotThreads) k_bb_dot_proj(const double* __restrict__ x,
                                                             double* __restrict__ h,
                                                             const double* __restrict__ pn,
                                                             int ns, int n, int nch,
                                                             double* __restrict__ part) {
  const int ch = blockIdx.x;
  const int i0 = ch * kBbChunk, i1 = min(n, i0 + kBbChunk);
  double s[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double t = 0.0;
    for (int k = 0; k < ns; ++k) t += pn[(size_t)k * n + i];
    const double hv = h[i] + t;
    h[i] = hv;
    s[0] += x[i] * hv;
  }
  dot_store<1>(s, part, 0, ch, nch);
  if (ch == 0 && threadIdx.x == 0) part[nch] = x[0] * h[0];
}

// eigv = <x|Hx>; g = eigv x - Hx; chunk sums of g.g (Sources/eigen.f90:380-400)
__global__ void __launch_bounds__(kDotThreads) k_bb_start(const double* __restrict__ x,
                                                          const double* __restrict__ hx,
                                                          double* __restrict__ g, int n, int nch,
                                                          const double* __restrict__ pin,
                                                          BbScal* S, double* __restrict__ pout) {
  const double eigv = 2.0 * bb_total(pin, nch) - pin[nch];
  const int ch = blockIdx.x;
  const int i0 = ch * kBbChunk, i1 = min(n, i0 + kBbChunk);
  double s[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double gv = eigv * x[i] + -1.0 * hx[i];
    g[i] = gv;
    s[0] += gv * gv;
  }
  dot_store<1>(s, pout, 0, ch, nch);
  if (ch == 0 && threadIdx.x == 0) {
    S->eigv = eigv;
    pout[nch] = g[0] * g[0];
  }
}

__global__ void k_bb_init(BbScal* S, const double* __restrict__ part, int nch) {
  const double zansa = 2.0 * bb_total(part, nch) - part[nch];
  if (threadIdx.x) return;
  S->zansa = S->zansab = S->z000re = zansa;
  S->x000re = 1e10;
  S->eigb = S->eigv;
  S->dinnerx = 0.0;
  S->betk = 0.0;
  S->iter = 0;
  S->izcnt = 0;
  S->lreset = 1;
  S->active = 1;
}

// CG direction h = g + beta h, beta = zansap / betk (h = g after a reset)
// from the chunk sums of the preconditioned g.g (Sources/eigen.f90:440-452)
__global__ void __launch_bounds__(kDotThreads) k_bb_dir(const double* __restrict__ g,
                                                        double* __restrict__ h, int n, int nch,
                                                        const double* __restrict__ part,
                                                        BbScal* S) {
  const double zansap = 2.0 * bb_total(part, nch) - g[0] * g[0];
  const bool reset = S->lreset != 0;
  const double beta = reset ? 0.0 : zansap / S->betk;
  const int i0 = blockIdx.x * kBbChunk, i1 = min(n, i0 + kBbChunk);
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) h[i] = reset ? g[i] : g[i] + beta * h[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) S->zansap = zansap;
}

// subgs4: h <- h - <x|h> x; chunk sums of h.Hx for cai (Sources/engrad.f90:1232-1239,
// eigen.f90:486-491)
__global__ void __launch_bounds__(kDotThreads) k_bb_orth(const double* __restrict__ x,
                                                         double* __restrict__ h,
                                                         const double* __restrict__ hx, int n,
                                                         int nch, const double* __restrict__ pin,
                                                         double* __restrict__ pout) {
  const double hhx = 2.0 * bb_total(pin, nch) - pin[nch];
  const int ch = blockIdx.x;
  const int i0 = ch * kBbChunk, i1 = min(n, i0 + kBbChunk);
  double s[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double hv = h[i] - hhx * x[i];
    h[i] = hv;
    s[0] += hv * hx[i];
  }
  dot_store<1>(s, pout, 0, ch, nch);
  if (ch == 0 && threadIdx.x == 0) pout[nch] = h[0] * hx[0];
}

// g <- g + sum_s pn[s] (the KB back-projection of H h, ns = 0 when H h is
// complete); chunk sums of h.g, h.x, h.h (the layout of k_dot3)
__global__ void __launch_bounds__(kDotThreads) k_bb_dot3(const double* __restrict__ h,
                                                         double* __restrict__ g,
                                                         const double* __restrict__ x,
                                                         const double* __restrict__ pn, int ns,
                                                         int n, int nch, double* __restrict__ part) {
  const int ch = blockIdx.x;
  const int i0 = ch * kBbChunk, i1 = min(n, i0 + kBbChunk);
  double s[3] = {0.0, 0.0, 0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double gv = g[i];
    if (ns) {
      double t = 0.0;
      for (int k = 0; k < ns; ++k) t += pn[(size_t)k * n + i];
      gv = gv + t;
      g[i] = gv;
    }
    const double hv = h[i];
    s[0] += hv * gv;
    s[1] += hv * x[i];
    s[2] += hv * hv;
  }
  dot_store<3>(s, part, 0, ch, nch);
}

// the 2x2 step of BKL, or the line minimisation when h.h <= 0
// (Sources/eigen.f90:575-693); pc: cai, pd: h.Hh | h.x | h.h
__global__ void k_bb_step(BbScal* S, const double* __restrict__ pc, const double* __restrict__ pd,
                          const double* __restrict__ h, const double* __restrict__ hh,
                          const double* __restrict__ x, int nch, double threinn, double wdinner) {
  const double cai = 2.0 * bb_total(pc, nch) - pc[nch];
  const double cbi = 2.0 * bb_total(pd, nch, 3, 0) - h[0] * hh[0];
  const double hhx = 2.0 * bb_total(pd, nch, 3, 1) - h[0] * x[0];
  const double hhi = 2.0 * bb_total(pd, nch, 3, 2) - h[0] * h[0];
  if (threadIdx.x) return;
  double eigv = S->eigv, dtnew, alpsq;
  if (hhi > 0.0) {
    const double h11 = eigv, h12 = cai / sqrt(hhi), h21 = h12, h22 = cbi / hhi;
    const double enew =
        0.5 * (h11 + h22 - sqrt((h11 + h22) * (h11 + h22) - 4.0 * (h11 * h22 - h12 * h21)));
    dtnew = (enew - h11) / h12 / sqrt(hhi);
    alpsq = sqrt(1.0 / (1.0 + dtnew * (0.0 + dtnew * hhi)));
    eigv = enew;
  } else {
    const double c0 = eigv, c1 = cai, c2 = cbi, c3 = hhi, c4 = 2.0 * hhx;
    const double s0 = c0 * c4 - c1, s1 = c0 * c3 - c2, s2 = c1 * c3 - c2 * c4;
    const double sr = s1 * s1 - s2 * s0;
    dtnew = sr > 0.0 ? (-s1 - sqrt(sr)) / s2 : -1.0;
    if (dtnew > 0.0) {
      const double dinner = sqrt(1.0 / (1.0 + dtnew * dtnew * hhi));
      if (dinner < threinn) dtnew = 0.0;
      if (dinner < S->dinnerx * wdinner) dtnew = 0.0;
      S->dinnerx = fmax(S->dinnerx, dinner);
    }
    const double dt2 = dtnew * dtnew, alpsq2 = 1.0 / (1.0 + dtnew * c4 + dt2 * c3);
    eigv = (c0 + dtnew * c1 + dt2 * c2) * alpsq2;
    alpsq = sqrt(alpsq2);
  }
  S->eigv = eigv;
  S->dtnew = dtnew;
  S->alpsq = alpsq;
  S->iter += 1;
  S->lreset = 0;
  S->betk = S->zansap;
}

// x <- alpsq (x + dt h), Hx <- alpsq (Hx + dt Hh), g <- eig x - Hx; chunk
// sums of g.g.  hh_g holds Hh on entry and the new g on exit.
__global__ void __launch_bounds__(kDotThreads) k_bb_update(double* __restrict__ x,
                                                           double* __restrict__ hx,
                                                           double* hh_g,
                                                           const double* __restrict__ h, int n,
                                                           int nch, const BbScal* S,
                                                           double* __restrict__ pout) {
  const double a = S->alpsq, dt = S->dtnew, e = S->eigv;
  const int ch = blockIdx.x;
  const int i0 = ch * kBbChunk, i1 = min(n, i0 + kBbChunk);
  double s[1] = {0.0};
  for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double xv = a * (x[i] + dt * h[i]);
    const double hv = a * (hx[i] + dt * hh_g[i]);
    const double gv = e * xv - hv;
    x[i] = xv;
    hx[i] = hv;
    hh_g[i] = gv;
    s[0] += gv * gv;
  }
  dot_store<1>(s, pout, 0, ch, nch);
  if (ch == 0 && threadIdx.x == 0) pout[nch] = hh_g[0] * hh_g[0];
}

// convergence test (Sources/eigen.f90:755-778); the band's results go to
// eig_out / bz_out / it_out and the flag the host waits for to *flag
__global__ void k_bb_conv(BbScal* S, const double* __restrict__ part, int nch, int ib, int noband,
                          int itermx, int iteremx, double tolera, double roughfzan,
                          double roughfeig, double* eig_out, double* bz_out, int* it_out,
                          int* flag) {
  const double zansa = 2.0 * bb_total(part, nch) - part[nch];
  if (threadIdx.x) return;
  const double eig = S->eigv, eigb = S->eigb;
  const int iter = S->iter;
  S->zansa = zansa;
  *eig_out = eig;
  *bz_out = zansa;
  *it_out = iter;
  const bool stop = (ib <= noband && iter >= itermx) || (ib > noband && iter >= iteremx) ||
                    zansa < tolera || fabs(eigb - eig) < tolera || zansa < S->z000re * roughfzan ||
                    (iter > 1 && fabs(eigb - eig) < S->x000re * roughfeig);
  if (stop) {
    S->active = 0;
  } else {
    if (iter == 1) S->x000re = fabs(eigb - eig);
    if (zansa > S->zansab) {
      if (++S->izcnt > 20) {
        S->izcnt = 0;
        S->lreset = 1;
      }
    }
    S->zansab = zansa;
    S->eigb = eig;
  }
  *flag = S->active;
}

// both passes of a single-vector product with A (M columns of length nplwex);
// the back-projection's column-chunk partials go to pn (returned: their
// count), the projection's chunk sums right after them
static int sv_project(Ctx* c, const double* A, int M, const double* v, const double* scale,
                      int im0, double* pn) {
  const int n = c->nplwex, ns = (M + kProjBands - 1) / kProjBands;
  const int nch = (n + kSvChunk - 1) / kSvChunk;
  double* pt = pn + (size_t)ns * n;
  hipLaunchKernelGGL(k_sv_proj_t, dim3(nch, (M + kProjRows - 1) / kProjRows), dim3(kDotThreads), 0,
                     c->stream, A, n, M, v, nch, pt);
  hipLaunchKernelGGL(k_sv_proj_n, dim3((n + kDotThreads - 1) / kDotThreads, ns), dim3(kDotThreads),
                     0, c->stream, A, n, M, v, nch, pt, scale, im0, pn);
  return ns;
}

int bkl_band_by_band(Ctx* c, double* psi, double* hpsi, int nb, const double* ekib, double* eig,
                     int noband, int itermx, int iteremx, double tolera, double threinn,
                     double wdinner, double roughfzan, double roughfeig, double* bzan, int* nriter,
                     int prec) {
  const int n = c->nplwex, nch = (n + kBbChunk - 1) / kBbChunk;
  if (!c->d_bb || !c->h_flag || nch > c->bb_nch) return -2;
  double* gnk = c->d_slot[4];
  double* hnk = c->d_slot[4] + n;
  const size_t L = (size_t)c->nbmax;
  double* prod = c->d_scal + 8 * L;
  double* ek = c->d_scal + 9 * L;
  double* eig_d = c->d_scal + 24 * L;
  double* bz_d = c->d_scal + 25 * L;
  int* it_d = reinterpret_cast<int*>(c->d_scal + 26 * L);
  BbScal* S = reinterpret_cast<BbScal*>(c->d_bb);
  double* p0 = c->d_bb + 32;                 // nch + 1
  double* p1 = p0 + (nch + 1);               // nch + 1
  double* p2 = p1 + (nch + 1);               // 3 nch
  // subgs3 scratch after gnk and hnk in the same slot: the back-projection's
  // band-chunk partials (ns x n), then the projection's chunk sums (ibb x nch)
  double* pn = c->d_slot[4] + 2 * (size_t)n;
  const size_t slot_len = (size_t)n * c->nbmax;
  auto proj_fits = [&](int ibb) {
    const size_t ns = (size_t)(ibb + kProjBands - 1) / kProjBands;
    return (2 + ns) * n + (size_t)ibb * ((n + kSvChunk - 1) / kSvChunk) <= slot_len;
  };
  const bool use_kb1 = c->kb1.valid && c->npj > 0 && !std::getenv("QX_KB1_OFF") &&
                       2 * (size_t)n + kb1_scratch(c) <= slot_len;
  const int maxit = std::max(1, std::max(itermx, iteremx));
  const dim3 grid(nch), blk(kDotThreads);
  QX_CHECK(hipMemcpyAsync(ek, ekib, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
  for (int ibb = 0; ibb < nb; ++ibb) {
    const int ib = ibb + 1;
    double* x = psi + (size_t)n * ibb;
    double* hx = hpsi + (size_t)n * ibb;
    hipLaunchKernelGGL(k_bb_dot, grid, blk, 0, c->stream, x, hx, n, nch, p0);
    hipLaunchKernelGGL(k_bb_start, grid, blk, 0, c->stream, x, hx, gnk, n, nch, p0, S, p1);
    hipLaunchKernelGGL(k_bb_init, dim3(1), dim3(64), 0, c->stream, S, p1, nch);
    for (int it = 0; it < maxit; ++it) {
      hipLaunchKernelGGL(k_precond_dot<kBbChunk>, grid, blk, 0, c->stream, gnk, c->d_kin, ek + ibb, n,
                         nullptr, nch, p0);
      hipLaunchKernelGGL(k_bb_dir, grid, blk, 0, c->stream, gnk, hnk, n, nch, p0, S);
      // subgs3 against the updated bands 1..ib-1, then the chunk sums of x.h
      if (ibb > 0 && proj_fits(ibb)) {
        const int ns = sv_project(c, psi, ibb, hnk, nullptr, 0, pn);
        hipLaunchKernelGGL(k_bb_dot_proj, grid, blk, 0, c->stream, x, hnk, pn, ns, n, nch, p0);
      } else {
        if (ibb > 0) {
          int rc = gemm_tn(c, ibb, 1, n, 2.0, psi, n, hnk, n, 0.0, prod, ibb, 1);
          if (rc) return rc;
          rc = gemm_nn(c, n, 1, ibb, -1.0, psi, n, prod, ibb, 1.0, hnk, n);
          if (rc) return rc;
        }
        hipLaunchKernelGGL(k_bb_dot, grid, blk, 0, c->stream, x, hnk, n, nch, p0);
      }
      hipLaunchKernelGGL(k_bb_orth, grid, blk, 0, c->stream, x, hnk, hx, n, nch, p0, p1);
      // H h: local part, then the KB part as a single-vector projection pair
      // whose back-projection k_bb_dot3 adds (the FP32 form keeps kb_apply)
      int rc = local_apply(c, hnk, gnk, 1, prec, 0);
      if (rc) return rc;
      int ns_kb = 0;
      if (prec == PREC_FP64 && use_kb1) {
        ns_kb = kb1_project(c, hnk, pn);  // from the device-built table's factors
      } else if (prec == PREC_FP64 && proj_fits(c->npj)) {
        if (c->npj) ns_kb = sv_project(c, c->d_B, c->npj, hnk, c->d_D, 1, pn);
      } else if ((rc = kb_apply(c, hnk, gnk, 1, 1, nullptr, prec))) {
        return rc;
      }
      hipLaunchKernelGGL(k_bb_dot3, grid, blk, 0, c->stream, hnk, gnk, x, pn, ns_kb, n, nch, p2);
      hipLaunchKernelGGL(k_bb_step, dim3(1), dim3(64), 0, c->stream, S, p1, p2, hnk, gnk, x, nch,
                         threinn, wdinner);
      hipLaunchKernelGGL(k_bb_update, grid, blk, 0, c->stream, x, hx, gnk, hnk, n, nch, S, p0);
      hipLaunchKernelGGL(k_bb_conv, dim3(1), dim3(64), 0, c->stream, S, p0, nch, ib, noband,
                         itermx, iteremx, tolera, roughfzan, roughfeig, eig_d + ibb, bz_d + ibb,
                         it_d + ibb, c->d_hflag);
      QX_CHECK(hipGetLastError());
      QX_CHECK(hipStreamSynchronize(c->stream));
      if (!*static_cast<volatile int*>(c->h_flag)) break;
    }
  }
  QX_CHECK(hipMemcpyAsync(eig, eig_d, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  std::vector<double> bz(nb);
  std::vector<int> itv(nb);
  QX_CHECK(hipMemcpyAsync(bz.data(), bz_d, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipMemcpyAsync(itv.data(), it_d, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  for (int b = 0; b < nb; ++b) {
    if (bzan) bzan[b] = bz[b];
    if (nriter) nriter[b] = itv[b];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// batched form: per-band scalars live on the device
// ---------------------------------------------------------------------------
struct BklState {
  double *eig, *eigb, *zansa, *zansab, *z000re, *x000re, *betk, *cai, *cbi, *hhi, *zansap;
  double *alpha, *dt, *beta, *one;
  int *active, *iter, *izcnt, *lreset;
};

__global__ void k_bkl_init(BklState s, int nb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  s.zansab[b] = s.zansa[b];
  s.z000re[b] = s.zansa[b];
  s.x000re[b] = 1e10;
  s.eigb[b] = s.eig[b];
  s.izcnt[b] = 0;
  s.lreset[b] = 1;
  s.iter[b] = 0;
  s.active[b] = 1;
  s.one[b] = 1.0;
}

__global__ void k_bkl_beta(BklState s, int nb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || !s.active[b]) return;
  s.beta[b] = s.lreset[b] ? 0.0 : s.zansap[b] / s.betk[b];
  s.lreset[b] = 0;
  s.betk[b] = s.zansap[b];
}

// 2x2 Rayleigh-Ritz of BKL (Sources/eigen.f90:593-622); inactive bands get
// alpha = 1, dt = 0 so the update leaves them unchanged.
__global__ void k_bkl_step(BklState s, int nb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (!s.active[b]) {
    s.alpha[b] = 1.0;
    s.dt[b] = 0.0;
    return;
  }
  double hhi = s.hhi[b], cai = s.cai[b], cbi = s.cbi[b], h11 = s.eig[b];
  if (hhi > 0.0) {
    double h12 = cai / sqrt(hhi), h22 = cbi / hhi;
    double enew = 0.5 * (h11 + h22 - sqrt((h11 + h22) * (h11 + h22) - 4.0 * (h11 * h22 - h12 * h12)));
    double dt = (enew - h11) / h12 / sqrt(hhi);
    double a2 = 1.0 / (1.0 + dt * (0.0 + dt * hhi));
    s.alpha[b] = sqrt(a2);
    s.dt[b] = dt;
    s.eig[b] = enew;
  } else {
    s.alpha[b] = 1.0;
    s.dt[b] = 0.0;
  }
  s.iter[b] += 1;
}

// convergence test after the update (Sources/eigen.f90:755-778)
__global__ void k_bkl_conv(BklState s, int nb, int noband, int itermx, int iteremx, double tolera,
                           double roughfzan, double roughfeig) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || !s.active[b]) return;
  const int ib = b + 1, iter = s.iter[b];
  double zansa = s.zansa[b], eig = s.eig[b], eigb = s.eigb[b];
  bool stop = (ib <= noband && iter >= itermx) || (ib > noband && iter >= iteremx) ||
              zansa < tolera || fabs(eigb - eig) < tolera || zansa < s.z000re[b] * roughfzan ||
              (iter > 1 && fabs(eigb - eig) < s.x000re[b] * roughfeig);
  if (stop) {
    s.active[b] = 0;
    return;
  }
  if (iter == 1) s.x000re[b] = fabs(eigb - eig);
  if (zansa > s.zansab[b]) {
    if (++s.izcnt[b] > 20) { s.izcnt[b] = 0; s.lreset[b] = 1; }
  }
  s.zansab[b] = zansa;
  s.eigb[b] = eig;
}

__global__ void k_neg(double* out, const double* in, int nb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) out[b] = -in[b];
}
__global__ void k_mul(double* out, const double* a, const double* b2, int nb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) out[b] = a[b] * b2[b];
}
__global__ void k_copy(double* out, const double* in, int nb) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) out[b] = in[b];
}

// One batched-CG iteration captured as a HIP graph, per (bands, precision,
// psi, H psi) of a context: every later iteration (all but the first, which
// runs eagerly and sets up whatever the products need on first use) is one
// graph launch.  The graph holds kernel launches only (no memset or host
// copy nodes).
struct CgGraphs {
  struct Entry {
    int nb, prec, path;
    const double *psi, *hpsi;
    unsigned long gen;  // Ctx::buffers_gen at capture
    hipGraphExec_t exec;
  };
  std::vector<Entry> e;
};

void cg_graphs_destroy(Ctx* c) {
  if (!c->cg_graphs) return;
  auto* g = static_cast<CgGraphs*>(c->cg_graphs);
  for (auto& x : g->e) (void)hipGraphExecDestroy(x.exec);
  delete g;
  c->cg_graphs = nullptr;
}

int cg_graph_entries(Ctx* c) {
  return c->cg_graphs ? (int)static_cast<CgGraphs*>(c->cg_graphs)->e.size() : 0;
}

static bool cg_graph_wanted(Ctx* c) {
  static const bool off = std::getenv("QX_CG_GRAPH") && std::atoi(std::getenv("QX_CG_GRAPH")) == 0;
  if (off || !c->cg_graph_on || gemm_path() == GEMM_LIB) return false;
  for (auto& t : c->lt)
    if (t.on) return false;  // launch timers record events between launches
  return true;
}

int hpsi_batched(Ctx* c, const double* x, double* hx, int nb, int prec);

int bkl_batched(Ctx* c, double* psi, double* hpsi, int nb, const double* ekib_host, double* eig_host,
                int noband, int itermx, int iteremx, double tolera, double roughfzan,
                double roughfeig, double* bzan, int* nriter, int prec) {
  const int n = c->nplwex;
  double* S = c->d_scal;
  const size_t L = (size_t)c->nbmax;
  BklState s;
  s.eig = S; s.eigb = S + L; s.zansa = S + 2 * L; s.zansab = S + 3 * L; s.z000re = S + 4 * L;
  s.x000re = S + 5 * L; s.betk = S + 6 * L; s.cai = S + 7 * L; s.cbi = S + 10 * L;
  s.hhi = S + 11 * L; s.zansap = S + 12 * L; s.alpha = S + 13 * L; s.dt = S + 14 * L;
  s.beta = S + 15 * L; s.one = S + 16 * L;
  double* ek = S + 9 * L;
  double* tmp = S + 17 * L;
  double* tmp2 = S + 18 * L;
  int* I = reinterpret_cast<int*>(S + 20 * L);
  s.active = I; s.iter = I + L; s.izcnt = I + 2 * L; s.lreset = I + 3 * L;
  double* gnk = c->d_slot[2];
  double* hnk = c->d_slot[3];
  double* hh = c->d_slot[4];
  double* proj = c->d_small;  // nb x nb
  const unsigned gb = cdivu(nb, 256);
  QX_CHECK(hipMemcpyAsync(ek, ekib_host, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
  // eig = <x|Hx>, g = eig x - Hx, zansa = <g|g>
  pdot_batched(c, psi, hpsi, nb, s.eig);
  axpby(c, gnk, psi, hpsi, s.eig, nullptr, 0.0, -1.0, nb, nullptr);
  pdot_batched(c, gnk, gnk, nb, s.zansa);
  hipLaunchKernelGGL(k_bkl_init, dim3(gb), dim3(256), 0, c->stream, s, nb);
  const int maxit = std::max(itermx, iteremx);
  const int nch = (n + kDotChunk - 1) / kDotChunk;
  double* part = c->d_scal + 64 * (size_t)c->nbmax;  // 3 x nch x nb partial sums
  const dim3 dgrid(nch, nb);
  // the mixed form's products on split operands: psi split here, kept current
  // by the update pass (the overlap of the Gram-Schmidt step that follows
  // reads it too), h split by the direction pass
  const bool split = band_split_wanted(c, nb, prec);
  BandSplit xs, hs;
  if (split) {
    int rc = band_split(c, 2, psi, nb, &xs);
    if (!rc) rc = band_split_buf(c, 3, &hs);
    if (rc) return rc;
  }
  // one iteration; the first takes h = P g (beta = 0 and h may hold anything)
  auto iteration = [&](bool first) -> int {
    // precondition, CG direction
    hipLaunchKernelGGL(k_precond_dot<kDotChunk>, dgrid, dim3(kDotThreads), 0, c->stream, gnk,
                       c->d_kin, ek, n, s.active, nch, part);
    hipLaunchKernelGGL(k_dotq_final, dim3(gb), dim3(256), 0, c->stream, part, nch, nb, 1, gnk, gnk,
                       nullptr, nullptr, n, s.zansap, nullptr, nullptr);
    hipLaunchKernelGGL(k_bkl_beta, dim3(gb), dim3(256), 0, c->stream, s, nb);
    int rc = 0;
    if (split) {
      hipLaunchKernelGGL(k_cg_dir_split, dim3(cdivu((long)n * nb, 256)), dim3(256), 0, c->stream,
                         hnk, gnk, first ? nullptr : s.beta, n, nb, s.active, hs.p, hs.ld, hs.lo);
      // project out all current bands: h <- h - Psi (Psi^T W h)
      if ((rc = band_tn(c, nb, nb, 2.0, xs, psi, hs, hnk, proj, nb))) return rc;
      if ((rc = band_nn(c, nb, -1.0, xs, proj, nb, 1.0, hnk))) return rc;
    } else {
      if (first) axpby(c, hnk, gnk, nullptr, nullptr, nullptr, 1.0, 0.0, nb, s.active);
      else axpby(c, hnk, gnk, hnk, nullptr, s.beta, 1.0, 0.0, nb, s.active);
      // project out all current bands: h <- h - Psi (Psi^T W h)
      rc = gemm_tn_p(c, nb, nb, n, 2.0, psi, n, hnk, n, 0.0, proj, nb, 1, prec);
      if (rc) return rc;
      rc = gemm_nn_p(c, n, nb, nb, -1.0, psi, n, proj, nb, 1.0, hnk, n, prec);
      if (rc) return rc;
    }
    rc = hpsi_batched(c, hnk, hh, nb, prec);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dot3<kDotChunk>, dgrid, dim3(kDotThreads), 0, c->stream, hnk, hpsi, hh,
                       n, nch, part);
    hipLaunchKernelGGL(k_dotq_final, dim3(gb), dim3(256), 0, c->stream, part, nch, nb, 3, hnk,
                       hpsi, hh, hnk, n, s.cai, s.cbi, s.hhi);
    hipLaunchKernelGGL(k_bkl_step, dim3(gb), dim3(256), 0, c->stream, s, nb);
    // x = a (x + dt h); Hx = a (Hx + dt Hh); g = eig x - Hx
    hipLaunchKernelGGL(k_mul, dim3(gb), dim3(256), 0, c->stream, tmp, s.alpha, s.dt, nb);
    if (split)
      hipLaunchKernelGGL(k_cg_update_split, dgrid, dim3(kDotThreads), 0, c->stream, psi, hpsi,
                         gnk, hnk, hh, s.alpha, tmp, s.eig, n, s.active, nch, part, xs.p, xs.ld,
                         xs.lo);
    else
      hipLaunchKernelGGL(k_cg_update, dgrid, dim3(kDotThreads), 0, c->stream, psi, hpsi, gnk,
                         hnk, hh, s.alpha, tmp, s.eig, n, s.active, nch, part);
    hipLaunchKernelGGL(k_dotq_final, dim3(gb), dim3(256), 0, c->stream, part, nch, nb, 1, gnk,
                       gnk, nullptr, nullptr, n, tmp2, nullptr, nullptr);
    hipLaunchKernelGGL(k_copy, dim3(gb), dim3(256), 0, c->stream, s.zansa, tmp2, nb);
    hipLaunchKernelGGL(k_bkl_conv, dim3(gb), dim3(256), 0, c->stream, s, nb, noband, itermx,
                       iteremx, tolera, roughfzan, roughfeig);
    return hipGetLastError() == hipSuccess ? 0 : -4;
  };
  int rc = iteration(true);
  if (rc) return rc;
  hipGraphExec_t exec = nullptr;
  if (maxit > 1 && cg_graph_wanted(c)) {
    if (!c->cg_graphs) c->cg_graphs = new CgGraphs();
    auto* G = static_cast<CgGraphs*>(c->cg_graphs);
    // graphs captured before a buffer they use moved are stale: drop them
    for (size_t i = 0; i < G->e.size();) {
      if (G->e[i].gen != c->buffers_gen) {
        (void)hipGraphExecDestroy(G->e[i].exec);
        G->e.erase(G->e.begin() + (long)i);
      } else {
        ++i;
      }
    }
    const int path = gemm_path();
    for (auto& x : G->e)
      if (x.nb == nb && x.prec == prec && x.path == path && x.psi == psi && x.hpsi == hpsi)
        exec = x.exec;
    if (!exec) {
      hipGraph_t graph = nullptr;
      QX_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rcc = iteration(false);
      const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
      if (rcc == 0 && ec == hipSuccess && graph &&
          hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
        G->e.push_back({nb, prec, path, psi, hpsi, c->buffers_gen, exec});
      } else {
        exec = nullptr;
        (void)hipGetLastError();
      }
      if (graph) (void)hipGraphDestroy(graph);
      if (!exec) c->cg_graph_on = false;  // capture not possible: stay eager from now on
    }
  }
  for (int it = 1; it < maxit; ++it) {
    if (exec) {
      QX_CHECK(hipGraphLaunch(exec, c->stream));
    } else if ((rc = iteration(false))) {
      return rc;
    }
  }
  QX_CHECK(hipMemcpyAsync(eig_host, s.eig, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  if (bzan)
    QX_CHECK(hipMemcpyAsync(bzan, s.zansa, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  std::vector<int> itv(nb);
  QX_CHECK(hipMemcpyAsync(itv.data(), s.iter, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (nriter)
    for (int b = 0; b < nb; ++b) nriter[b] = itv[b];
  return 0;
}

}  // namespace qx
