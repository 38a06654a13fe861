// eigh.hip -- symmetric eigensolver of the subspace Hamiltonian on the GPU
// (the role of RsEIGQR, Sources/funcs.f90:380-764: eigenvalues ascending with
// their vectors), built for many workgroups:
//
//   1. Householder tridiagonalisation in one cooperative launch: the rows of
//      A live in the LDS of the workgroups (row i in workgroup i mod G); each
//      step publishes a column, every workgroup forms the reflector from it,
//      the row owners form p = A v and publish it, and every workgroup applies
//      the rank-2 update A -= v w^T + w v^T to its rows.  Two grid barriers
//      per column, no pass over A in HBM.
//   2. Eigenvalues by multisection on Sturm counts: one wavefront per
//      eigenvalue, its 64 lanes probe 64 points of the bracket at once.
//   3. Eigenvectors by inverse iteration, one thread per eigenvalue (LU with
//      partial pivoting of T - lambda I), followed by a CholeskyQR
//      re-orthonormalisation of all vectors (Gram matrix and Z L^-T on MFMA):
//      from random start vectors with a shifted first pass (shifted
//      CholeskyQR3), then once more from the orthonormal vectors (CholeskyQR2).
//      Numerically equal eigenvalues get shifts pulled apart by 10 eps |lambda|
//      as in LAPACK's dstein.
//   4. Back-transformation by the stored reflectors, 8 (4 for large n)
//      columns of Z per workgroup held in LDS, reflectors staged through LDS.
#include <cmath>
#include <cstdlib>
#include <vector>
#include "context.h"
#include "coop.h"

namespace qx {

int dense_eigh_legacy(Ctx* c, double* a, int n, double* z, double* w, double* work);

__device__ inline double block_sum256(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  const double s = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return s;
}

// ---- 1. tridiagonalisation --------------------------------------------------
// a: n x n column-major symmetric (full storage).  On exit d, e (e[i] couples
// i and i+1, e[n-1] = 0) and the unit reflector of step k in a(k+1:n, k).
//
// One exchange per step.  At the top of step k every workgroup holds the
// current column k (x_j = A_k(j, k), j > k) and forms the reflector v from it.
// Row owners publish p_i = (A_k v)_i for their rows with their workgroup's
// share of K = v.p, and the owner of row k+1 publishes that row, A_k(k+1, :).
// Every thread then forms w_j = 2 (p_j - K v_j) for its columns, updates them
// in the owned rows, A_{k+1} = A_k - v w^T - w v^T, and forms the next column
// from the published row: x_j = A_k(k+1, j) - v_{k+1} w_j - w_{k+1} v_j.
// Column j is owned by thread j mod 256 throughout (x, v and w of it stay in
// that thread's registers).
//
// The published values carry the step as a tag (st_tagged), and readers wait
// for the tags of exactly the values they need instead of a grid barrier, so
// a step costs one store-to-load trip through the device-coherent level.  The
// buffers are double-buffered by step parity: a workgroup writes parity q
// again two steps later, which needs the step in between from every
// workgroup, and each workgroup publishes that step only after it has read all
// of its inputs of the step before.  Waits are bounded (err raised, as in
// grid_barrier), so a launch that cannot complete still ends.
// M: columns per thread, n <= 256 M, and KQ: K shares per lane, G <= 64 KQ;
// instantiated for the smallest that cover n and G (M = 3, 5, 8; KQ = 3, 4),
// so a step issues and polls only the loads of columns and shares that exist
constexpr int kTrdM = 8;   // largest M: n <= 256 * kTrdM
constexpr int kTrdKq = 4;  // K shares read per lane: groups <= 64 * kTrdKq

template <int M, int KQ>
__global__ void __launch_bounds__(256) k_tridiag_coop(double* __restrict__ a, int n,
                                                      double* __restrict__ d,
                                                      double* __restrict__ e, double* rowbuf,
                                                      unsigned long long* pub, int* bar,
                                                      int* err) {
  extern __shared__ double sm[];
  const int G = gridDim.x, g = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int rpw = (n + G - 1) / G;
  double* R = sm;                      // R[r * n + j] = A(g + G r, j)
  double* v = R + (size_t)rpw * n;     // reflector, indexed by global column
  double* prow = v + n;                // p of the owned rows
  double* red = prow + rpw;            // 2 x 8 reduction slots (step parity)
  double* kw = red + 16;               // K shares of the 4 wavefronts
  // tagged buffers, 2 words per value: p [2][n], row k+1 [2][n], K shares [2][G]
  unsigned long long* pb = pub;
  unsigned long long* rb = pb + 4 * (size_t)n;
  unsigned long long* kb = rb + 4 * (size_t)n;
  for (int r = 0; r < rpw; ++r) {
    const int i = g + G * r;
    if (i >= n) break;
    for (int j = t; j < n; j += 256) R[(size_t)r * n + j] = a[(size_t)i + (size_t)n * j];
  }
  __syncthreads();
  if (g == 0)  // row 0 is column 0: publish it (row 0 lives in workgroup 0)
    for (int j = t; j < n; j += 256) st_agent(&rowbuf[j], R[j]);
  grid_barrier(bar, G, err);
  double x[M], vr[M], w[M];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int j = t + 256 * m;
    x[m] = j < n ? ld_agent(&rowbuf[j]) : 0.0;
    vr[m] = 0.0;
    w[m] = 0.0;
  }
  if (g == 0 && t == 0) d[0] = ld_agent(&rowbuf[0]);
  bool failed = false;
  for (int k = 0; k + 2 < n; ++k) {
    const int q = k & 1;
    const unsigned tag = (unsigned)k + 1u;
    unsigned long long* pq = pb + 2 * (size_t)q * n;
    unsigned long long* rq = rb + 2 * (size_t)q * n;
    unsigned long long* kq = kb + 2 * (size_t)q * G;
    double* rd = red + 8 * q;
    // reflector of column k, formed redundantly by every workgroup
    double part = 0.0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int j = t + 256 * m;
      if (j > k && j < n) part += x[m] * x[m];
      if (j == k + 1) rd[4] = x[m];
    }
    part = wave_sum(part);
    if (lane == 0) rd[wv] = part;
    __syncthreads();
    const double nx2 = rd[0] + rd[1] + rd[2] + rd[3];
    const double x0 = rd[4];
    const double alpha = nx2 > 0.0 ? (x0 >= 0.0 ? -sqrt(nx2) : sqrt(nx2)) : 0.0;
    const double vn2 = nx2 - 2.0 * alpha * x0 + alpha * alpha;
    const bool skip = !(vn2 > 1e-300);
    const double inv = skip ? 0.0 : 1.0 / sqrt(vn2);
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int j = t + 256 * m;
      if (j > k && j < n) {
        vr[m] = (j == k + 1 ? x[m] - alpha : x[m]) * inv;
        v[j] = vr[m];
        if (g == 0) a[(size_t)j + (size_t)n * k] = vr[m];
      }
    }
    if (g == 0 && t == 0) e[k] = skip ? x0 : alpha;
    __syncthreads();
    // publish p and the K share for the owned rows of the trailing block, and row k+1
    double kp = 0.0;
    for (int r = wv; r < rpw; r += 4) {
      const int i = g + G * r;
      if (i <= k || i >= n) continue;
      double acc = 0.0;
      if (!skip)
        for (int j = k + 1 + lane; j < n; j += 64) acc += R[(size_t)r * n + j] * v[j];
      acc = wave_sum(acc);
      if (lane == 0) {
        st_tagged(pq + 2 * i, acc, tag);
        prow[r] = acc;
      }
      kp += v[i] * acc;
    }
    if (lane == 0) kw[wv] = kp;
    if ((k + 1) % G == g) {
      const double* Rr = R + (size_t)((k + 1) / G) * n;
      for (int j = k + 1 + t; j < n; j += 256) st_tagged(rq + 2 * j, Rr[j], tag);
    }
    __syncthreads();
    if (t == 0) st_tagged(kq + 2 * g, kw[0] + kw[1] + kw[2] + kw[3], tag);
    // wait for this step's values: every load is issued before any tag is
    // tested (entries clamped to published ones), so their latencies overlap
    unsigned long long kw0[KQ], kw1[KQ], pw0[M], pw1[M], rw0[M], rw1[M];
    unsigned long long p10, p11, r10, r11;
    for (int spins = 0;; ++spins) {
#pragma unroll
      for (int u = 0; u < KQ; ++u) {
        const unsigned long long* s = kq + 2 * min(lane + 64 * u, G - 1);
        kw0[u] = ld_u64_agent(s);
        kw1[u] = ld_u64_agent(s + 1);
      }
      p10 = ld_u64_agent(pq + 2 * (k + 1));
      p11 = ld_u64_agent(pq + 2 * (k + 1) + 1);
      r10 = ld_u64_agent(rq + 2 * (k + 1));
      r11 = ld_u64_agent(rq + 2 * (k + 1) + 1);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int j = min(max(t + 256 * m, k + 1), n - 1);
        pw0[m] = ld_u64_agent(pq + 2 * j);
        pw1[m] = ld_u64_agent(pq + 2 * j + 1);
        rw0[m] = ld_u64_agent(rq + 2 * j);
        rw1[m] = ld_u64_agent(rq + 2 * j + 1);
      }
      bool all = tag_ok(p10, p11, tag) && tag_ok(r10, r11, tag);
#pragma unroll
      for (int u = 0; u < KQ; ++u) all = all && tag_ok(kw0[u], kw1[u], tag);
#pragma unroll
      for (int m = 0; m < M; ++m)
        all = all && tag_ok(pw0[m], pw1[m], tag) && tag_ok(rw0[m], rw1[m], tag);
      if (all || failed) break;
      if ((spins & 1023) == 1023 &&
          (spins > (1 << 20) ||
           __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        failed = true;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    const double p1 = tag_value(p10, p11), r1 = tag_value(r10, r11);
    double ks = 0.0;
#pragma unroll
    for (int u = 0; u < KQ; ++u) ks += lane + 64 * u < G ? tag_value(kw0[u], kw1[u]) : 0.0;
    const double K = wave_sum(ks);
    if (!skip) {
      const double v1 = v[k + 1];
      const double w1 = 2.0 * (p1 - K * v1);
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int j = t + 256 * m;
        if (j > k && j < n) w[m] = 2.0 * (tag_value(pw0[m], pw1[m]) - K * vr[m]);
        if (j > k + 1 && j < n) x[m] = tag_value(rw0[m], rw1[m]) - v1 * w[m] - w1 * vr[m];
      }
      if (g == 0 && t == 0) d[k + 1] = r1 - 2.0 * v1 * w1;
      for (int r = 0; r < rpw; ++r) {
        const int i = g + G * r;
        if (i <= k + 1 || i >= n) continue;  // row k+1 is finished
        const double vi = v[i], wi = 2.0 * (prow[r] - K * vi);
        double* Rr = R + (size_t)r * n;
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int j = t + 256 * m;
          if (j > k + 1 && j < n) Rr[j] -= vi * w[m] + wi * vr[m];
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int j = t + 256 * m;
        if (j > k + 1 && j < n) x[m] = tag_value(rw0[m], rw1[m]);
      }
      if (g == 0 && t == 0) d[k + 1] = r1;
    }
  }
  if (g == 0) {
#pragma unroll
    for (int m = 0; m < M; ++m)
      if (t + 256 * m == n - 1) e[n - 2] = x[m];
  }
  // R(n-1, n-1) was last updated by the thread owning column n-1: it reads it
  if ((n - 1) % G == g && t == (n - 1) % 256) {
    d[n - 1] = R[(size_t)((n - 1) / G) * n + n - 1];
    e[n - 1] = 0.0;
  }
}

__global__ void k_zero_u64(unsigned long long* p, long cnt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = 0ull;
}

// ---- 2. eigenvalues ----------------------------------------------------------
// Gershgorin bracket and scales: bounds = {lo, hi, pivmin, ||T||_inf}
__global__ void __launch_bounds__(256) k_gersh(const double* d, const double* e, int n,
                                               double* bounds) {
  __shared__ double rlo[256], rhi[256], rnm[256], re2[256];
  double lo = 1e300, hi = -1e300, nm = 0.0, e2 = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    nm = fmax(nm, fabs(d[i]) + r);
    if (i < n - 1) e2 = fmax(e2, e[i] * e[i]);
  }
  rlo[threadIdx.x] = lo;
  rhi[threadIdx.x] = hi;
  rnm[threadIdx.x] = nm;
  re2[threadIdx.x] = e2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      rlo[threadIdx.x] = fmin(rlo[threadIdx.x], rlo[threadIdx.x + o]);
      rhi[threadIdx.x] = fmax(rhi[threadIdx.x], rhi[threadIdx.x + o]);
      rnm[threadIdx.x] = fmax(rnm[threadIdx.x], rnm[threadIdx.x + o]);
      re2[threadIdx.x] = fmax(re2[threadIdx.x], re2[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double eps = 2.220446049250313e-16, safmin = 2.2250738585072014e-308;
    const double pivmin = fmax(safmin * fmax(1.0, re2[0]), safmin);
    const double tn = rnm[0];
    bounds[0] = rlo[0] - 2.0 * eps * tn - 4.0 * pivmin;
    bounds[1] = rhi[0] + 2.0 * eps * tn + 4.0 * pivmin;
    bounds[2] = pivmin;
    bounds[3] = tn;
  }
}

__device__ inline int sturm_count_ms(const double* d, const double* e, int n, double sigma,
                                     double pivmin) {
  int cnt = 0;
  double q = d[0] - sigma;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0.0;
  for (int i = 1; i < n; ++i) {
    q = d[i] - sigma - e[i - 1] * e[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// the same count with d and the squared couplings e2[i] = e[i]^2 in LDS (the
// same arithmetic: e[i-1] * e[i-1] rounds to e2[i-1])
__device__ inline int sturm_count_lds(const double* d, const double* e2, int n, double sigma,
                                      double pivmin) {
  int cnt = 0;
  double q = d[0] - sigma;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0.0;
#pragma unroll 8
  for (int i = 1; i < n; ++i) {
    q = d[i] - sigma - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

constexpr int kBisectLds = 2048;  // d and e^2 staged in LDS up to this order

// one wavefront per eigenvalue j (0-based, ascending): the bracket [lo, hi]
// holds lambda_j; lane l tests lo + (l+1)(hi-lo)/65, the first point with more
// than j eigenvalues below it closes the new bracket.
__global__ void __launch_bounds__(256) k_bisect_ms(const double* __restrict__ d,
                                                   const double* __restrict__ e, int n,
                                                   const double* __restrict__ bounds,
                                                   double* __restrict__ w) {
  __shared__ double ds[kBisectLds], e2s[kBisectLds];
  const bool staged = n <= kBisectLds;
  if (staged) {
    for (int i = threadIdx.x; i < n; i += 256) {
      ds[i] = d[i];
      e2s[i] = i < n - 1 ? e[i] * e[i] : 0.0;
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const double eps = 2.220446049250313e-16;
  double lo = bounds[0], hi = bounds[1];
  const double pivmin = bounds[2];
  for (int it = 0; it < 64; ++it) {
    if (hi - lo <= 4.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double step = (hi - lo) / 65.0;
    const double x = lo + step * (double)(lane + 1);
    const int cnt = staged ? sturm_count_lds(ds, e2s, n, x, pivmin)
                           : sturm_count_ms(d, e, n, x, pivmin);
    const unsigned long long m = __ballot(cnt > j);
    const int l = m ? __ffsll((long long)m) - 1 : 64;
    const double nlo = l > 0 ? lo + step * (double)l : lo;
    const double nhi = l < 64 ? lo + step * (double)(l + 1) : hi;
    if (!(nhi - nlo < hi - lo)) break;
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) w[j] = 0.5 * (lo + hi);
}

// ---- 3. eigenvectors ---------------------------------------------------------
// One thread per eigenvalue j: factor T - lambda_j I (partial pivoting, rows i
// and i+1 swapped when the subdiagonal dominates), then solve with the column
// j of X (row-major, X[i*n + j]) as right-hand side, in place, and normalise.
// first != 0: the right-hand side is a fixed pseudo-random vector instead.
// The running row of the factorisation and of the forward sweep stay in
// registers, so every global access is independent of the recurrence.
// lu: 5 n^2 doubles, entry (q, i) of thread j at lu[(q*n + i)*n + j].
__device__ inline double inviter_start(int i, int j) {
  unsigned h = (unsigned)(i * 2654435761u) ^ (unsigned)(j * 40503u + 12345u);
  h ^= h >> 13;
  h *= 0x5bd1e995u;
  h ^= h >> 15;
  return 0.5 + (double)(h & 0xffff) / 65536.0;
}

// Loads are gathered 32 steps at a time into registers ahead of the
// recurrence (the solution overwrites the right-hand side in place, so the
// compiler cannot hoist them on its own).
constexpr int kIvChunk = 32;

__global__ void __launch_bounds__(64) k_inviter_par(const double* __restrict__ d,
                                                    const double* __restrict__ e, int n,
                                                    const double* __restrict__ w,
                                                    const double* __restrict__ bounds,
                                                    double* __restrict__ lu,
                                                    double* __restrict__ X, int first) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const size_t nn = (size_t)n * n;
  double* dd = lu + j;           // U diagonal (tiny pivots replaced by +-pert)
  double* du = dd + nn;          // U superdiagonal
  double* du2 = du + nn;         // U second superdiagonal
  double* dl = du2 + nn;         // multipliers
  double* pv = dl + nn;          // 1 when rows i, i+1 were swapped
#define L_(arr, i) arr[(size_t)(i) * n]
  if (first) {
    // factor T - lambda I once; the second pass reuses it
    const double eps = 2.220446049250313e-16;
    const double pert = eps * bounds[3];
    double lam = w[j];
    int r = 0;
    for (int q = j; q > 0 && w[q] - w[q - 1] < 10.0 * eps * fabs(w[q]); --q) ++r;
    lam += (double)r * 10.0 * eps * fabs(lam);
    auto guard = [&](double v) { return fabs(v) < pert ? (v < 0.0 ? -pert : pert) : v; };
    double dc = d[0] - lam, uc = n > 1 ? e[0] : 0.0;  // current row: diagonal, superdiagonal
    for (int i = 0; i < n - 1; ++i) {
      const double sub = e[i];
      const double dn = d[i + 1] - lam;
      const double un = i + 2 < n ? e[i + 1] : 0.0;
      if (fabs(dc) >= fabs(sub)) {
        const double f = dc != 0.0 ? sub / dc : 0.0;
        L_(dd, i) = guard(dc);
        L_(du, i) = uc;
        L_(du2, i) = 0.0;
        L_(dl, i) = f;
        L_(pv, i) = 0.0;
        dc = dn - f * uc;
        uc = un;
      } else {
        const double f = dc / sub;
        L_(dd, i) = guard(sub);
        L_(du, i) = dn;
        L_(du2, i) = un;
        L_(dl, i) = f;
        L_(pv, i) = 1.0;
        dc = uc - f * dn;
        uc = -f * un;
      }
    }
    L_(dd, n - 1) = guard(dc);
    L_(du, n - 1) = 0.0;
    L_(du2, n - 1) = 0.0;
  }
  double* x = X + j;
  // forward sweep L y = P b, the row being eliminated carried in xc
  double xc = first ? inviter_start(0, j) : L_(x, 0);
  for (int i0 = 0; i0 < n - 1; i0 += kIvChunk) {
    double xr[kIvChunk], fr[kIvChunk], pr[kIvChunk];
#pragma unroll
    for (int c = 0; c < kIvChunk; ++c) {
      const int i = i0 + c;
      if (i < n - 1) {
        xr[c] = first ? inviter_start(i + 1, j) : L_(x, i + 1);
        fr[c] = L_(dl, i);
        pr[c] = L_(pv, i);
      }
    }
#pragma unroll
    for (int c = 0; c < kIvChunk; ++c) {
      const int i = i0 + c;
      if (i < n - 1) {
        if (pr[c] == 0.0) {
          L_(x, i) = xc;
          xc = xr[c] - fr[c] * xc;
        } else {
          L_(x, i) = xr[c];
          xc = xc - fr[c] * xr[c];
        }
      }
    }
  }
  L_(x, n - 1) = xc;
  // back substitution U x = y, with the squared norm accumulated on the way
  double x1 = 0.0, x2 = 0.0, big = 0.0, s2 = 0.0;  // x1 = x[i+1], x2 = x[i+2]
  for (int i1 = n - 1; i1 >= 0; i1 -= kIvChunk) {
    double yr[kIvChunk], ur[kIvChunk], u2r[kIvChunk], dr[kIvChunk];
#pragma unroll
    for (int c = 0; c < kIvChunk; ++c) {
      const int i = i1 - c;
      if (i >= 0) {
        yr[c] = L_(x, i);
        ur[c] = L_(du, i);
        u2r[c] = L_(du2, i);
        dr[c] = L_(dd, i);
      }
    }
#pragma unroll
    for (int c = 0; c < kIvChunk; ++c) {
      const int i = i1 - c;
      if (i >= 0) {
        const double v = (yr[c] - ur[c] * x1 - u2r[c] * x2) / dr[c];
        L_(x, i) = v;
        x2 = x1;
        x1 = v;
        s2 += v * v;
        big = fmax(big, fabs(v));
        if (big > 1e150) {
          // rescale the solved part, the right-hand side still in memory and
          // the part of it already in registers
          for (int q = 0; q < n; ++q) L_(x, q) *= 1e-150;
#pragma unroll
          for (int c2 = 0; c2 < kIvChunk; ++c2)
            if (c2 > c) yr[c2] *= 1e-150;
          x1 *= 1e-150;
          x2 *= 1e-150;
          s2 *= 1e-300;
          big *= 1e-150;
        }
      }
    }
  }
  const double sc = s2 > 0.0 ? 1.0 / sqrt(s2) : 0.0;
  for (int i0 = 0; i0 < n; i0 += kIvChunk) {
    double xr[kIvChunk];
#pragma unroll
    for (int c = 0; c < kIvChunk; ++c)
      if (i0 + c < n) xr[c] = L_(x, i0 + c);
#pragma unroll
    for (int c = 0; c < kIvChunk; ++c)
      if (i0 + c < n) L_(x, i0 + c) = xr[c] * sc;
  }
#undef L_
}

// ---- 3b. re-orthonormalisation within clusters ----------------------------------
// Eigenvalues within kClusterTol * ||T|| of their neighbour form a cluster (the
// grouping of LAPACK dstein).  Inverse-iteration vectors of different clusters
// are orthogonal to O(eps ||T|| / gap) already, so only the vectors of one
// cluster are orthonormalised against each other.
constexpr double kClusterTol = 1e-3;
constexpr int kClusterMax = 64;  // larger clusters take the CholeskyQR path

// cl[j] = first index of the cluster holding eigenvalue j; cl[n] = size of the
// largest cluster.  One workgroup of 1024 threads: the cluster starts are an
// inclusive max-scan of the indices that open a cluster (n <= 4096), one
// thread for larger n.
constexpr int kScanMax = 4096;
__global__ void __launch_bounds__(1024) k_cluster_scan(const double* __restrict__ w, int n,
                                                       const double* __restrict__ bounds,
                                                       int* __restrict__ cl) {
  __shared__ int st[kScanMax];
  __shared__ int wmx[16];
  const int t = threadIdx.x;
  const double tol = kClusterTol * bounds[3];
  if (n > kScanMax) {
    if (t != 0) return;
    int s = 0, mx = 1;
    cl[0] = 0;
    for (int j = 1; j < n; ++j) {
      if (w[j] - w[j - 1] > tol) s = j;
      cl[j] = s;
      mx = max(mx, j - s + 1);
    }
    cl[n] = mx;
    return;
  }
  for (int j = t; j < n; j += 1024) st[j] = (j == 0 || w[j] - w[j - 1] > tol) ? j : 0;
  __syncthreads();
  for (int off = 1; off < n; off <<= 1) {
    int v[kScanMax / 1024];
#pragma unroll
    for (int u = 0; u < kScanMax / 1024; ++u) {
      const int j = t + 1024 * u;
      v[u] = j < n ? (j >= off ? max(st[j], st[j - off]) : st[j]) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kScanMax / 1024; ++u)
      if (t + 1024 * u < n) st[t + 1024 * u] = v[u];
    __syncthreads();
  }
  int mx = 1;
  for (int j = t; j < n; j += 1024) {
    cl[j] = st[j];
    mx = max(mx, j - st[j] + 1);
  }
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
  if ((t & 63) == 0) wmx[t >> 6] = mx;
  __syncthreads();
  if (t == 0) {
    int m = wmx[0];
    for (int q = 1; q < 16; ++q) m = max(m, wmx[q]);
    cl[n] = m;
  }
}

// One workgroup per cluster start: classical Gram-Schmidt applied twice to
// each vector against the earlier ones of its cluster, then normalised.
// z: n x n column-major, vectors in columns.  Clusters hold <= kClusterMax.
__global__ void __launch_bounds__(256) k_cluster_orth(double* __restrict__ z, int n,
                                                      const int* __restrict__ cl) {
  const int j0 = blockIdx.x;
  if (cl[j0] != j0) return;
  int k = 1;
  while (j0 + k < n && cl[j0 + k] == j0) ++k;
  if (k == 1 || k > kClusterMax) return;
  __shared__ double dots[kClusterMax];
  __shared__ double red[4];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double* zj = z + (size_t)j0 * n;
  for (int c = 0; c < k; ++c) {
    double* zc = zj + (size_t)c * n;
    for (int pass = 0; pass < 2; ++pass) {
      for (int p = wv; p < c; p += 4) {  // one wavefront per inner product
        const double* zp = zj + (size_t)p * n;
        double acc = 0.0;
        for (int i = lane; i < n; i += 64) acc += zp[i] * zc[i];
        acc = wave_sum(acc);
        if (lane == 0) dots[p] = acc;
      }
      __syncthreads();
      for (int i = t; i < n; i += 256) {
        double v = zc[i];
        for (int p = 0; p < c; ++p) v -= dots[p] * zj[(size_t)p * n + i];
        zc[i] = v;
      }
      __syncthreads();
    }
    double acc = 0.0;
    for (int i = t; i < n; i += 256) acc += zc[i] * zc[i];
    acc = wave_sum(acc);
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    const double nrm = red[0] + red[1] + red[2] + red[3];
    const double sc = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
    for (int i = t; i < n; i += 256) zc[i] *= sc;
    __syncthreads();
  }
}

// out[j*n + i] = in[i*n + j]
__global__ void __launch_bounds__(256) k_transpose_sq(const double* __restrict__ in, int n,
                                                      double* __restrict__ out) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows per pass
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, j = bx + tx;
    if (i < n && j < n) tile[r][tx] = in[(size_t)i * n + j];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int j = bx + r, i = by + tx;
    if (i < n && j < n) out[(size_t)j * n + i] = tile[tx][r];
  }
}

__global__ void k_add_diag(double* g, int n, double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[(size_t)i * n + i] += s;
}

// ---- 4. back-transformation ----------------------------------------------------
// COLS columns of Z per workgroup, 256 / COLS lanes per column; nref
// reflectors staged in LDS at a time.  LDS = (COLS + nref) n doubles.
template <int COLS>
__global__ void __launch_bounds__(256) k_backtrans_lds(const double* __restrict__ a, int n,
                                                       double* __restrict__ z, int nref) {
  constexpr int LANES = 256 / COLS;
  extern __shared__ double sm[];
  double* zs = sm;                         // [COLS][n]
  double* vs = sm + (size_t)COLS * n;      // [nref][n]
  const int c0 = blockIdx.x * COLS;
  const int ncol = min(COLS, n - c0);
  const int cl = threadIdx.x / LANES, lane = threadIdx.x % LANES;
  for (int idx = threadIdx.x; idx < COLS * n; idx += 256) {
    const int cc = idx / n, i = idx - cc * n;
    zs[idx] = cc < ncol ? z[(size_t)(c0 + cc) * n + i] : 0.0;
  }
  double* zc = zs + (size_t)cl * n;
  for (int kt = n - 3; kt >= 0; kt -= nref) {
    const int kl = max(0, kt - nref + 1);  // reflectors kl..kt
    __syncthreads();
    for (int idx = threadIdx.x; idx < (kt - kl + 1) * n; idx += 256) {
      const int q = idx / n, i = idx - q * n;
      vs[idx] = i > kl + q ? a[(size_t)(kl + q) * n + i] : 0.0;
    }
    __syncthreads();
    for (int k = kt; k >= kl; --k) {
      const double* vk = vs + (size_t)(k - kl) * n;
      // element i belongs to lane i % LANES for every k: no cross-lane LDS traffic
      const int i0 = k + 1 + ((lane - (k + 1)) % LANES + LANES) % LANES;
      double part = 0.0;
      for (int i = i0; i < n; i += LANES) part += vk[i] * zc[i];
      for (int o = LANES / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, LANES);
      const double tv = 2.0 * part;
      for (int i = i0; i < n; i += LANES) zc[i] -= tv * vk[i];
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < COLS * n; idx += 256) {
    const int cc = idx / n, i = idx - cc * n;
    if (cc < ncol) z[(size_t)(c0 + cc) * n + i] = zs[idx];
  }
}

template <int COLS>
static int backtrans(Ctx* c, const double* a, int n, double* z) {
  const size_t cap = 150 * 1024 / sizeof(double);
  const long nref = std::min<long>(32, (long)(cap / (size_t)n) - COLS);
  if (nref < 1) return -22;
  const size_t lds = sizeof(double) * (size_t)(COLS + nref) * n;
  if (lds > 65536)
    QX_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_backtrans_lds<COLS>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_backtrans_lds<COLS>, dim3((n + COLS - 1) / COLS), dim3(256), lds, c->stream,
                     a, n, z, (int)nref);
  return 0;
}

// ---- 4b. blocked back-transformation on MFMA ----------------------------------
// The reflectors in blocks of 16: Q_b = H_k0 ... H_k0+15 = I - V T V^T with T
// upper triangular, T(j, j) = 2 and T(0:j, j) = -2 T(0:j, 0:j) V(:, 0:j)^T v_j
// (the forward recurrence of LAPACK's dlarft for unit reflectors, tau = 2).
// Z = Q_0 Q_1 ... Z_T is formed from the last block to the first as
// Z <- Z - V (T (V^T Z)), three chains of v_mfma_f64_16x16x4f64.
constexpr int kWyB = 16;

// Vt[b][i][r] = v_{16 b + r}(i): zero outside the reflector (rows <= k, rows
// >= n, blocks past the last reflector n - 3)
__global__ void k_wy_pack(const double* __restrict__ a, int n, int npad, int nblk,
                          double* __restrict__ Vt) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)nblk * npad * kWyB) return;
  const int r = (int)(idx % kWyB);
  const long q = idx / kWyB;
  const int i = (int)(q % npad), b = (int)(q / npad);
  const int k = b * kWyB + r;
  Vt[idx] = (k <= n - 3 && i > k && i < n) ? a[(size_t)k * n + i] : 0.0;
}

// T of block b (one workgroup each): Gram matrix G = V^T V, then row i of T
// by thread i (row i of the recurrence needs row i only)
__global__ void __launch_bounds__(256) k_wy_t(const double* __restrict__ Vt, int npad,
                                              double* __restrict__ Tb) {
  __shared__ double G[kWyB][kWyB + 1], T[kWyB][kWyB + 1], Vs[64][kWyB + 1];
  const int b = blockIdx.x, t = threadIdx.x, r = t / kWyB, s = t % kWyB;
  const double* V = Vt + (size_t)b * npad * kWyB;
  double acc = 0.0;
  for (int i0 = 0; i0 < npad; i0 += 64) {  // 64 rows of the block at a time through LDS
    for (int idx = t; idx < 64 * kWyB; idx += 256) {
      const int ri = idx / kWyB, ci = idx % kWyB;
      Vs[ri][ci] = i0 + ri < npad ? V[(size_t)(i0 + ri) * kWyB + ci] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int i = 0; i < 64; ++i) acc += Vs[i][r] * Vs[i][s];
    __syncthreads();
  }
  G[r][s] = acc;
  T[r][s] = 0.0;
  __syncthreads();
  if (t < kWyB) {
    T[t][t] = 2.0;
    for (int j = t + 1; j < kWyB; ++j) {
      double x = 0.0;
      for (int l = t; l < j; ++l) x += T[t][l] * G[l][j];
      T[t][j] = -2.0 * x;
    }
  }
  __syncthreads();
  Tb[(size_t)b * kWyB * kWyB + t] = T[r][s];
}

// One workgroup per 16 columns of Z, held in registers for the whole pass as
// 16 x 16 tiles in the MFMA accumulator layout (D(i, j) in lane j + 16 (i % 4),
// register i / 4).  That layout is also the B-operand layout of four k steps
// (register s = rows 4 s .. 4 s + 3), so V^T Z reads Z from registers, and
// T W and V W' take their B operand straight from the previous product.
// Tile t of the column panel belongs to wavefront t % 4; the four partial
// V^T Z are summed through LDS in wavefront order.  TPW: tiles per wavefront.
template <int TPW>
__global__ void __launch_bounds__(256) k_backtrans_wy(const double* __restrict__ Vt,
                                                      const double* __restrict__ Tb, int n,
                                                      int npad, int nblk, double* __restrict__ z) {
  __shared__ double red[4][4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, cl = lane & 15;
  const int col = blockIdx.x * kWyB + cl;
  const int ntile = npad / kWyB;
  dv4 zt[TPW];
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = wave + 4 * q;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = kWyB * t + g + 4 * v;
      zt[q][v] = (t < ntile && row < n && col < n) ? z[(size_t)col * n + row] : 0.0;
    }
  }
  for (int b = nblk - 1; b >= 0; --b) {
    const int k0 = b * kWyB;
    const double* V = Vt + (size_t)b * npad * kWyB;
    // W = V^T Z over this wavefront's tiles (tiles wholly at rows <= k0 hold
    // no reflector entry)
    dv4 w = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int t = wave + 4 * q;
      if (t >= ntile || kWyB * t + kWyB - 1 <= k0) continue;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double av = V[(size_t)(kWyB * t + 4 * s + g) * kWyB + cl];
        w = __builtin_amdgcn_mfma_f64_16x16x4f64(av, zt[q][s], w, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) red[wave][v][lane] = w[v];
    __syncthreads();
    dv4 W;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      W[v] = red[0][v][lane] + red[1][v][lane] + red[2][v][lane] + red[3][v][lane];
    __syncthreads();
    // W' = T W
    const double* T = Tb + (size_t)b * kWyB * kWyB;
    dv4 wp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4; ++s)
      wp = __builtin_amdgcn_mfma_f64_16x16x4f64(T[cl * kWyB + 4 * s + g], W[s], wp, 0, 0, 0);
    // Z <- Z - V W'
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
      const int t = wave + 4 * q;
      if (t >= ntile || kWyB * t + kWyB - 1 <= k0) continue;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const double av = -V[(size_t)(kWyB * t + cl) * kWyB + 4 * s + g];
        zt[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, wp[s], zt[q], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < TPW; ++q) {
    const int t = wave + 4 * q;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = kWyB * t + g + 4 * v;
      if (t < ntile && row < n && col < n) z[(size_t)col * n + row] = zt[q][v];
    }
  }
}

// the blocked back-transformation for 64 <= n <= 2048 (ws: n^2 + 32 n + 512
// doubles at least); -3 outside that range
static int backtrans_wy(Ctx* c, const double* a, int n, double* z, double* ws) {
  if (n < 64 || n > 2048 || std::getenv("QX_BACKTRANS_LDS")) return -3;
  const int npad = (n + kWyB - 1) / kWyB * kWyB;
  const int nblk = (n - 2 + kWyB - 1) / kWyB;
  double* Vt = ws;
  double* Tb = ws + (size_t)nblk * npad * kWyB;
  const long nv = (long)nblk * npad * kWyB;
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(k_wy_pack, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, a, n, npad,
                     nblk, Vt);
  hipLaunchKernelGGL(k_wy_t, dim3(nblk), dim3(256), 0, st, Vt, npad, Tb);
  const int tpw = (npad / kWyB + 3) / 4;
  const dim3 grid((n + kWyB - 1) / kWyB);
#define QX_WY(T_)                                                                               \
  if (tpw <= T_) {                                                                              \
    hipLaunchKernelGGL(k_backtrans_wy<T_>, grid, dim3(256), 0, st, Vt, Tb, n, npad, nblk, z); \
    return hipGetLastError() == hipSuccess ? 0 : -4;                                            \
  }
  QX_WY(3)
  QX_WY(5)
  QX_WY(9)
  QX_WY(13)
  QX_WY(20)
  QX_WY(32)
#undef QX_WY
  return -3;
}

__global__ void k_sync_take(int* sync, int* out) {
  out[0] = sync[1] | sync[2];
  sync[2] = 0;
}

// ---- host driver ----------------------------------------------------------------
// workgroups of the tridiagonalisation: four rows each up to n = 640 (at
// n = 540 on one MI355X: 135 groups 7.68 ms per rotation phase, 68 groups
// 7.83 ms, 256 groups 8.0 ms), five above (n = 1280: 256 groups 22.3 ms per
// eigensolve, 160 groups of eight rows 23.3 ms)
static int tridiag_groups(int n) {
  if (const char* s = std::getenv("QX_TRID_WG")) return std::max(1, std::atoi(s));
  return std::max(1, n <= 640 ? (n + 3) / 4 : (n + 4) / 5);
}

// z (n x n, column-major, vectors in columns) <- QR factor of z by CholeskyQR
// passes; with shifted != 0 the first pass factors Z^T Z + s I (shifted
// CholeskyQR, for a Z that may be far from orthonormal), then two plain ones.
// ok: 2 ints per pass + 1 scratch.
static int chol_qr(Ctx* c, double* z, int n, double* gram, double* linv, double* tmp, int* ok,
                   int shifted) {
  const double u = 1.1102230246251565e-16;
  const double shift = 11.0 * (2.0 * n * (double)n + n) * u * (double)n;
  const int npass = shifted ? 3 : 2;
  for (int pass = 0; pass < npass; ++pass) {
    int rc = gemm_tn(c, n, n, n, 1.0, z, n, z, n, 0.0, gram, n, 0);
    if (rc) return rc;
    if (shifted && pass == 0)
      hipLaunchKernelGGL(k_add_diag, dim3((n + 255) / 256), dim3(256), 0, c->stream, gram, n, shift);
    if ((rc = dense_chol_inv(c, gram, n, linv, ok + 2 * pass, ok + 7))) return rc;
    if ((rc = gemm_nn(c, n, n, n, 1.0, z, n, linv, n, 0.0, tmp, n))) return rc;
    QX_CHECK(hipMemcpyAsync(z, tmp, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice,
                            c->stream));
  }
  return 0;
}

int dense_eigh(Ctx* c, double* a, int n, double* z, double* w, double* work, int* info) {
  (void)info;
  if (n < 3 || std::getenv("QX_DENSE_LEGACY") ||
      c->dwork_len < 8 * (size_t)n * n + 16 * (size_t)n + 64)
    return dense_eigh_legacy(c, a, n, z, w, work);
  hipStream_t st = c->stream;
  double* dw = c->d_dwork;
  double* d = dw;
  double* e = d + n;
  double* rowbuf = e + n;                                       // 2n
  double* pbuf = rowbuf + 2 * (size_t)n;                        // 2n (spare)
  double* bounds = pbuf + 2 * (size_t)n;                        // 8
  int* ok = reinterpret_cast<int*>(bounds + 8);                 // 2 x 8 ints
  int* sflag = ok + 16;                                         // 1 int
  int* cl = reinterpret_cast<int*>(dw + 8 * (size_t)n + 64);    // n + 1 ints
  double* X = dw + 16 * (size_t)n + 64;                         // n^2, row-major vectors
  double* gram = X + (size_t)n * n;                             // n^2
  double* linv = gram + (size_t)n * n;                          // n^2
  double* lu = linv + (size_t)n * n;                            // 5 n^2
  // 1. tridiagonalisation: the row distribution follows the group count, so
  // the launch must get exactly the groups its LDS is sized for
  int dev = 0, ncu = 0;
  QX_CHECK(hipGetDevice(&dev));
  QX_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int G = std::min(tridiag_groups(n), std::min(n, ncu));
  const size_t rpw = (size_t)((n + G - 1) / G);
  const size_t lds = sizeof(double) * (rpw * n + (size_t)n + rpw + 24);
  if (lds > 160 * 1024 - 1024 || n > 256 * kTrdM || G > 64 * kTrdKq)
    return dense_eigh_legacy(c, a, n, z, w, work);
  int* bar = c->d_sync;
  int* err = c->d_sync + 1;
  // the tagged exchange buffers (8 n + 4 G words) use the eigenvector scratch,
  // idle until step 3; tags must start cleared
  unsigned long long* pub = reinterpret_cast<unsigned long long*>(X);
  const long npub = 8 * (long)n + 4 * (long)G;
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((npub + 255) / 256)), dim3(256), 0, st, pub, npub);
  void* args[] = {&a, &n, &d, &e, &rowbuf, &pub, &bar, &err};
  const void* trd =
      n <= 256 * 3 ? (G <= 64 * 3 ? reinterpret_cast<const void*>(k_tridiag_coop<3, 3>)
                                  : reinterpret_cast<const void*>(k_tridiag_coop<3, kTrdKq>))
      : n <= 256 * 5 ? reinterpret_cast<const void*>(k_tridiag_coop<5, kTrdKq>)
                     : reinterpret_cast<const void*>(k_tridiag_coop<kTrdM, kTrdKq>);
  if (coop_launch(c, trd, G, lds, args, true) < 0) return dense_eigh_legacy(c, a, n, z, w, work);
  // 2. eigenvalues
  hipLaunchKernelGGL(k_gersh, dim3(1), dim3(256), 0, st, d, e, n, bounds);
  hipLaunchKernelGGL(k_bisect_ms, dim3((n + 3) / 4), dim3(256), 0, st, d, e, n, bounds, w);
  // 3. eigenvectors: inverse iteration + re-orthonormalisation, twice; the LU
  // factors of pass 0 are reused by pass 1.  Vectors are orthonormalised
  // within eigenvalue clusters, or, when a cluster is too large for that, all
  // together by CholeskyQR (scratch X).
  hipLaunchKernelGGL(k_cluster_scan, dim3(1), dim3(1024), 0, st, w, n, bounds, cl);
  int maxcl = 0;
  QX_CHECK(hipMemcpyAsync(&maxcl, cl + n, sizeof(int), hipMemcpyDeviceToHost, st));
  QX_CHECK(hipStreamSynchronize(st));
  const bool local = maxcl <= kClusterMax && std::getenv("QX_EIGH_CHOLQR") == nullptr;
  const dim3 tgrid((n + 31) / 32, (n + 31) / 32);
  for (int pass = 0; pass < 2; ++pass) {
    if (pass > 0) hipLaunchKernelGGL(k_transpose_sq, tgrid, dim3(256), 0, st, z, n, X);
    hipLaunchKernelGGL(k_inviter_par, dim3((n + 63) / 64), dim3(64), 0, st, d, e, n, w, bounds, lu,
                       X, pass == 0 ? 1 : 0);
    hipLaunchKernelGGL(k_transpose_sq, tgrid, dim3(256), 0, st, X, n, z);
    if (local) {
      hipLaunchKernelGGL(k_cluster_orth, dim3(n), dim3(256), 0, st, z, n, cl);
      continue;
    }
    // pass 0 starts from random vectors (shifted first CholeskyQR); pass 1
    // from orthonormal ones, which the solve keeps near orthonormal
    int rc = chol_qr(c, z, n, gram, linv, X, ok + 8 * pass, pass == 0);
    if (rc) return rc;
  }
  // 4. back-transformation (the LU scratch is free by now)
  {
    int rc = backtrans_wy(c, a, n, z, lu);
    if (rc == -3) rc = n <= 2000 ? backtrans<8>(c, a, n, z) : backtrans<4>(c, a, n, z);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_sync_take, dim3(1), dim3(1), 0, st, c->d_sync, sflag);
  QX_CHECK(hipGetLastError());
  int hflags[17];
  QX_CHECK(hipMemcpyAsync(hflags, ok, sizeof(int) * 17, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipStreamSynchronize(st));
  if (hflags[16]) return -20;  // a grid barrier gave up
  // a Gram matrix of the re-orthonormalisation was not positive definite
  if (!local && (!hflags[0] || !hflags[2] || !hflags[4] || !hflags[8] || !hflags[10])) return -23;
  return 0;
}

}  // namespace qx
