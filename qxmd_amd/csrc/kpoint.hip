// kpoint.hip -- H_k|psi> for Bloch bands at several k-points of the cell, all
// (k-point, band) blocks of a band set through one launch per stage.
//
// The reference's band algebra is Gamma-only (Sources/kpoints.f90:33-39 fixes
// the single k = 0; the k-point input block of Sources/input.f90:2016-2126 is
// commented out), so this extends its Hamiltonian apply -- hcnonloc + ekinetic
// (Sources/engrad.f90:411-606, 66-127) and hckbpp_g (Sources/nlkbpp.f90:
// 1284-1393) -- from real Gamma bands to complex Bloch bands:
//
//   psi_nk(r) = sum_G c_nk(G) e^{i (k+G) r},  |k+G|^2 <= E_cut (Ry, bohr^-2)
//   (H_k c)(G) = |k+G|^2 c(G) + [V_loc u](G) + sum_p beta_p(G) D_p <beta_p|c>
//
// with u(r) = sum_G c(G) e^{iGr} the periodic part on the context's FFT grid
// (V_loc multiplies u as it multiplies psi), beta_p(G) the projector's
// coefficients at k+G and <beta_p|c> = sum_G conj(beta_p(G)) c(G).
//
// Sphere of each k-point (our own order, the reference has none): the points
// grouped by (h, k) column, columns in lexicographic (h, k) order, l rising
// within a column.  Bands are stored complex-interleaved, [k][band][2 npw_max]
// doubles with a common stride npw_max (the largest sphere) and zeros beyond
// npw_k, so every stage addresses all k-points by one (k, band) index.
//
// Local part, three kernels over all (k, band) blocks of a batch:
//   k_kzinv  : the column points -> z lines in LDS, inverse FFT along z  -> T
//   k_kplane : per z plane, the columns into a dense plane in LDS, inverse y
//              and x, x V(r), forward x and y, back to the columns        -> T
//   k_kzfwd  : forward z, x 1/N at the sphere points, + |k+G|^2 c (the
//              kinetic term of ekinetic at k+G); zeros the padding tail.
// T holds only the columns of each sphere ([kb][z][column]), in place.
//
// Nonlocal part: two real products per k-point on the stacked table
// A_k = [beta_k ; i beta_k] (2 npj rows of 2 npw_max interleaved reals):
//   P = A_k c  gives [Re <beta|c> ; Im <beta|c>]  (TN, K = 2 npw_max)
//   H c += A_k^T (D P)                            (NN, the same table)
// so the complex projection and back-projection run on the FP64 MFMA
// kernels of the Gamma path with no complex GEMM.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>
#include <rocsolver/rocsolver.h>
#include "context.h"

namespace qx {

namespace {

constexpr int kKThreads = 256;
constexpr int kKCols = 16;  // columns per block of the z kernels

struct KArgs {
  int n1, n2, n3, nb, npw_max, ncol_max, kb0;
  int rpk, r0;        // band rows of a k-point in the src/dst arrays, first row used
  const int* npw;     // [k]
  const int* ncol;    // [k]
  const int* col_p0;  // [k][ncol_max + 1]: first point of column c
  const int* col_xy;  // [k][ncol_max]: x + n1 y of column c
  const int* pt_z;    // [k][npw_max]: z index of point p
  const double* kin;  // [k][npw_max]: |k+G|^2
  const double* tw;   // exp(+2 pi i j / n) for x (n1), y (n2), z (n3)
  FftPlan px, py, pz;
};

inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

__global__ void __launch_bounds__(kKThreads)
    k_kzinv(KArgs a, const double* __restrict__ psi, double2* __restrict__ T) {
  __shared__ cplx<double> tw[1024];
  __shared__ cplx<double> lines[kKCols * 128];
  const int kb = a.kb0 + (int)blockIdx.y;
  const int k = kb / a.nb;
  const int nc = a.ncol[k];
  const int c0 = (int)blockIdx.x * kKCols;
  if (c0 >= nc) return;
  const int cn = min(kKCols, nc - c0);
  const int n3 = a.n3;
  load_tw<double>(tw, a.tw + 2 * (a.n1 + a.n2), n3);
  for (int i = threadIdx.x; i < cn * n3; i += blockDim.x) lines[i] = mk<double>(0.0, 0.0);
  __syncthreads();
  const int* cp = a.col_p0 + (size_t)k * (a.ncol_max + 1);
  const int* pz = a.pt_z + (size_t)k * a.npw_max;
  const double* src = psi + ((size_t)k * a.rpk + a.r0 + (kb - k * a.nb)) * 2 * a.npw_max;
  const int p0 = cp[c0], p1 = cp[c0 + cn];
  for (int p = p0 + (int)threadIdx.x; p < p1; p += blockDim.x) {
    // the column of p: the last c with cp[c] <= p (columns are short)
    int c = c0;
    while (cp[c + 1] <= p) ++c;
    lines[(c - c0) * n3 + pz[p]] = mk<double>(src[2 * p], src[2 * p + 1]);
  }
  __syncthreads();
  lds_fft<double, 4>(lines, cn, n3, 1, a.pz, tw, +1);
  double2* out = T + (size_t)blockIdx.y * n3 * a.ncol_max;
  for (int i = threadIdx.x; i < cn * n3; i += blockDim.x) {
    const int z = i / cn, cl = i % cn;
    const cplx<double> v = lines[cl * n3 + z];
    out[(size_t)z * a.ncol_max + c0 + cl] = make_double2(v.x, v.y);
  }
}

// one z plane of one (k, band) block; dynamic LDS: twiddles x | y, plane n2 x n1
__global__ void __launch_bounds__(kKThreads)
    k_kplane(KArgs a, const double* __restrict__ V, double2* __restrict__ T) {
  extern __shared__ unsigned char smem_raw[];
  cplx<double>* twx = reinterpret_cast<cplx<double>*>(smem_raw);
  cplx<double>* twy = twx + a.n1;
  cplx<double>* pl = twy + a.n2;
  const int n1 = a.n1, n2 = a.n2;
  const int z = (int)blockIdx.x;
  const int kb = a.kb0 + (int)blockIdx.y;
  const int k = kb / a.nb;
  const int nc = a.ncol[k];
  load_tw<double>(twx, a.tw, n1);
  load_tw<double>(twy, a.tw + 2 * n1, n2);
  for (int i = threadIdx.x; i < n1 * n2; i += blockDim.x) pl[i] = mk<double>(0.0, 0.0);
  __syncthreads();
  const int* cxy = a.col_xy + (size_t)k * a.ncol_max;
  double2* row = T + ((size_t)blockIdx.y * a.n3 + z) * a.ncol_max;
  for (int c = threadIdx.x; c < nc; c += blockDim.x) {
    const double2 v = row[c];
    pl[cxy[c]] = mk<double>(v.x, v.y);
  }
  __syncthreads();
  lds_fft<double, 4>(pl, n1, 1, n1, a.py, twy, +1);  // along y, one line per x
  lds_fft<double, 4>(pl, n2, n1, 1, a.px, twx, +1);  // along x, one line per y
  const double* vz = V + (size_t)z * n1 * n2;
  for (int i = threadIdx.x; i < n1 * n2; i += blockDim.x) {
    const double v = vz[i];
    pl[i] = mk<double>(pl[i].x * v, pl[i].y * v);
  }
  __syncthreads();
  lds_fft<double, 4>(pl, n2, n1, 1, a.px, twx, -1);
  lds_fft<double, 4>(pl, n1, 1, n1, a.py, twy, -1);
  for (int c = threadIdx.x; c < nc; c += blockDim.x) {
    const cplx<double> v = pl[cxy[c]];
    row[c] = make_double2(v.x, v.y);
  }
}

__global__ void __launch_bounds__(kKThreads)
    k_kzfwd(KArgs a, const double* __restrict__ psi, const double2* __restrict__ T,
            double* __restrict__ out, double scale, int accumulate) {
  __shared__ cplx<double> tw[1024];
  __shared__ cplx<double> lines[kKCols * 128];
  const int kb = a.kb0 + (int)blockIdx.y;
  const int k = kb / a.nb;
  const int nc = a.ncol[k], np = a.npw[k];
  const size_t row = (size_t)k * a.rpk + a.r0 + (kb - k * a.nb);
  double* dst = out + row * 2 * a.npw_max;
  if (blockIdx.x == 0)  // the padding beyond npw_k stays zero
    for (int i = 2 * np + (int)threadIdx.x; i < 2 * a.npw_max; i += blockDim.x) dst[i] = 0.0;
  const int c0 = (int)blockIdx.x * kKCols;
  if (c0 >= nc) return;
  const int cn = min(kKCols, nc - c0);
  const int n3 = a.n3;
  load_tw<double>(tw, a.tw + 2 * (a.n1 + a.n2), n3);
  const double2* in = T + (size_t)blockIdx.y * n3 * a.ncol_max;
  for (int i = threadIdx.x; i < cn * n3; i += blockDim.x) {
    const int z = i / cn, cl = i % cn;
    const double2 v = in[(size_t)z * a.ncol_max + c0 + cl];
    lines[cl * n3 + z] = mk<double>(v.x, v.y);
  }
  __syncthreads();
  lds_fft<double, 4>(lines, cn, n3, 1, a.pz, tw, -1);
  const int* cp = a.col_p0 + (size_t)k * (a.ncol_max + 1);
  const int* pz = a.pt_z + (size_t)k * a.npw_max;
  const double* kin = a.kin + (size_t)k * a.npw_max;
  const double* src = psi + row * 2 * a.npw_max;
  const int p0 = cp[c0], p1 = cp[c0 + cn];
  for (int p = p0 + (int)threadIdx.x; p < p1; p += blockDim.x) {
    int c = c0;
    while (cp[c + 1] <= p) ++c;
    const cplx<double> v = lines[(c - c0) * n3 + pz[p]];
    double re = v.x * scale + kin[p] * src[2 * p];
    double im = v.y * scale + kin[p] * src[2 * p + 1];
    if (accumulate) {
      re += dst[2 * p];
      im += dst[2 * p + 1];
    }
    dst[2 * p] = re;
    dst[2 * p + 1] = im;
  }
}

// P[b][r] *= D[r % npj] for the stacked [Re; Im] projections (2 npj rows)
__global__ void k_kscale(double* P, const double* D, int npj, int nb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2L * npj * nb) return;
  P[i] *= D[(int)(i % (2 * npj)) % npj];
}

// ---- band solver kernels (rows of 2 L doubles, L = npw_max) -----------------
// copy nb rows per k-point: dst row k drpk + dr0 + b <- src row k srpk + sr0 + b
__global__ void k_krows(const double* __restrict__ src, int srpk, int sr0, double* __restrict__ dst,
                        int drpk, int dr0, int nb, long L2) {
  const int kb = blockIdx.y, k = kb / nb, b = kb % nb;
  const double* s = src + ((size_t)k * srpk + sr0 + b) * L2;
  double* d = dst + ((size_t)k * drpk + dr0 + b) * L2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < L2; i += (long)gridDim.x * blockDim.x)
    d[i] = s[i];
}

// J c = i c: rows [0, nbs) of each k-point's block -> rows [nbs, 2 nbs)
__global__ void k_kjrows(double* __restrict__ Z, int rpk, int nbs, long L2) {
  const int kb = blockIdx.y, k = kb / nbs, b = kb % nbs;
  const double2* s = reinterpret_cast<const double2*>(Z + ((size_t)k * rpk + b) * L2);
  double2* d = reinterpret_cast<double2*>(Z + ((size_t)k * rpk + nbs + b) * L2);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < L2 / 2; i += (long)gridDim.x * blockDim.x) {
    const double2 v = s[i];
    d[i] = make_double2(-v.y, v.x);
  }
}

// one block per (k, band): ekin = <c|T|c> / <c|c> over the full sphere (the
// kinetic energy the preconditioner scales with, the ekib of a real Gamma
// band) -- for bands of rows [0, nb)
__global__ void k_kekin(const double* __restrict__ Z, int rpk, int nb, int L,
                        const double* __restrict__ kin, double* __restrict__ ekin) {
  const int kb = blockIdx.x, k = kb / nb, b = kb % nb;
  const double* z = Z + ((size_t)k * rpk + b) * 2 * L;
  const double* q = kin + (size_t)k * L;
  double t = 0.0, n = 0.0;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    const double a = z[2 * i] * z[2 * i] + z[2 * i + 1] * z[2 * i + 1];
    t += q[i] * a;
    n += a;
  }
  __shared__ double st[256], sn[256];
  st[threadIdx.x] = t;
  sn[threadIdx.x] = n;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      st[threadIdx.x] += st[threadIdx.x + o];
      sn[threadIdx.x] += sn[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) ekin[kb] = st[0] / sn[0];
}

// W = K (H psi - lambda psi), the Teter-Payne-Allan preconditioner of prpcg
// (Sources/eigen.f90:1093-1113) on |k+G|^2, then W scaled to unit norm; one
// block per (k, band); rows: psi b, H psi b in Z / HZ, W -> Z row nb + b
__global__ void k_kresid(double* __restrict__ Z, const double* __restrict__ HZ, int rpk, int nb,
                         int L, const int* __restrict__ npw, const double* __restrict__ kin,
                         const double* __restrict__ lam, const double* __restrict__ ekin) {
  const int kb = blockIdx.x, k = kb / nb, b = kb % nb;
  const double* z = Z + ((size_t)k * rpk + b) * 2 * L;
  const double* h = HZ + ((size_t)k * rpk + b) * 2 * L;
  double* w = Z + ((size_t)k * rpk + nb + b) * 2 * L;
  const double* q = kin + (size_t)k * L;
  const double l = lam[kb], e15 = 1.5 * ekin[kb];
  const int np = npw[k];
  double s = 0.0;
  for (int i = threadIdx.x; i < L; i += blockDim.x) {
    double re = 0.0, im = 0.0;
    if (i < np) {
      const double x = q[i] / e15;
      double tesu = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
      const double x2 = x * x;
      tesu = tesu / (tesu + 16.0 * x2 * x2) * 2.0 / e15;
      re = tesu * (h[2 * i] - l * z[2 * i]);
      im = tesu * (h[2 * i + 1] - l * z[2 * i + 1]);
    }
    w[2 * i] = re;
    w[2 * i + 1] = im;
    s += re * re + im * im;
  }
  __shared__ double sr[256];
  sr[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sr[threadIdx.x] += sr[threadIdx.x + o];
    __syncthreads();
  }
  const double f = sr[0] > 0.0 ? 1.0 / sqrt(sr[0]) : 0.0;
  for (int i = threadIdx.x; i < 2 * L; i += blockDim.x) w[i] *= f;
}

// the stacked products G[n][m] (m < 2 nbs: rows of [Y; J Y]) -> complex
// column-major A(m, n) = G[n][m] + i G[n][nbs + m], m, n < nbs
__global__ void k_kpack(const double* __restrict__ G, int ldg, int nbs, int ldc,
                        rocblas_double_complex* __restrict__ A, long strideG, long strideA) {
  const int k = blockIdx.y;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nbs * nbs) return;
  const int m = (int)(i % nbs), n = (int)(i / nbs);
  const double* g = G + k * strideG + (size_t)n * ldg;
  A[k * strideA + m + (size_t)n * ldc] = rocblas_double_complex(g[m], g[nbs + m]);
}

// the lowest nb eigenvectors (columns of A, nbs long) -> Cs[n][j] = [Re; Im]
// (4 nb rows per column, the first 2 nbs used)
__global__ void k_kunpack(const rocblas_double_complex* __restrict__ A, int lda, int nbs, int nb,
                          double* __restrict__ Cs, int ldcs, long strideA, long strideC) {
  const int k = blockIdx.y;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nbs * nb) return;
  const int j = (int)(i % nbs), n = (int)(i / nbs);
  const rocblas_double_complex v = A[k * strideA + j + (size_t)n * lda];
  double* c = Cs + k * strideC + (size_t)n * ldcs;
  c[j] = v.real();
  c[nbs + j] = v.imag();
}

}  // namespace

struct KPoints {
  int nk = 0, nb = 0, npw_max = 0, ncol_max = 0;
  std::vector<double> kfrac;               // 3 x nk, units of the reciprocal vectors
  std::vector<int> npw, ncol;
  std::vector<std::vector<int>> hkl;       // per k: 3 x npw (h, k, l)
  std::vector<std::vector<double>> kin;    // per k: |k+G|^2
  int *d_npw = nullptr, *d_ncol = nullptr, *d_col_p0 = nullptr, *d_col_xy = nullptr,
      *d_pt_z = nullptr;
  double* d_kin = nullptr;
  static constexpr int kSlots = 4;
  double* d_slot[kSlots] = {};
  double2* d_T = nullptr;                  // pencil columns of a batch, in place
  int batch = 0;                           // (k, band) blocks per batch
  // KB: stacked tables A_k (2 npj x 2 npw_max) per k, D duplicated (2 npj)
  int npj = 0;
  double* d_A = nullptr;
  double* d_D = nullptr;
  double* d_P = nullptr;                   // 2 npj x nb
  // band solver: per k-point the basis Y = [psi, W] and its partner J Y (J c
  // = i c in the interleaved layout), [k][4 nb][2 npw_max], and H of it
  double *d_Z = nullptr, *d_HZ = nullptr;
  double* d_G = nullptr;                   // [k][2 nb][4 nb]: the stacked products
  rocblas_double_complex *d_Hs = nullptr, *d_Ss = nullptr;  // [k][2 nb][2 nb]
  double *d_w = nullptr, *d_E = nullptr;   // eigenvalues, solver work
  int* d_info = nullptr;
  double* d_Cs = nullptr;                  // [k][nb][4 nb]: [Re C; Im C]
  double *d_ekin = nullptr, *d_lam = nullptr;  // per (k, band)
  rocblas_handle rb = nullptr;
};

namespace {

void kp_free(KPoints* k) {
  void* ptrs[] = {k->d_npw, k->d_ncol, k->d_col_p0, k->d_col_xy, k->d_pt_z, k->d_kin,
                  k->d_T, k->d_A, k->d_D, k->d_P, k->d_Z, k->d_HZ, k->d_G, k->d_Hs, k->d_Ss,
                  k->d_w, k->d_E, k->d_info, k->d_Cs, k->d_ekin, k->d_lam};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (k->rb) (void)rocblas_destroy_handle(k->rb);
  for (auto* p : k->d_slot)
    if (p) (void)hipFree(p);
}

KArgs kargs(const Ctx* c, const KPoints* k, int kb0, int rpk, int r0) {
  KArgs a{};
  a.n1 = c->n1; a.n2 = c->n2; a.n3 = c->n3;
  a.nb = k->nb; a.npw_max = k->npw_max; a.ncol_max = k->ncol_max; a.kb0 = kb0;
  a.rpk = rpk; a.r0 = r0;
  a.npw = k->d_npw; a.ncol = k->d_ncol; a.col_p0 = k->d_col_p0; a.col_xy = k->d_col_xy;
  a.pt_z = k->d_pt_z; a.kin = k->d_kin; a.tw = c->d_tw;
  a.px = c->plx; a.py = c->ply; a.pz = c->plz;
  return a;
}

size_t plane_lds(const Ctx* c) { return sizeof(double) * 2 * ((size_t)c->n1 * c->n2 + c->n1 + c->n2); }

}  // namespace

void kp_destroy(Ctx* c) {
  auto* k = static_cast<KPoints*>(c->kp);
  if (!k) return;
  kp_free(k);
  delete k;
  c->kp = nullptr;
}

// The sphere of k: integer (h, k, l) with |k+G|^2 <= ecut, G = h b1 + k b2 +
// l b3 (b columns of bvec, 2 pi included), in column order.
static int k_sphere(const double* bvec, const double* kf, double ecut, int n1, int n2, int n3,
                    std::vector<int>& hkl, std::vector<double>& kin, int& ncol) {
  // |G . a_i| / 2 pi = |n_i| with a_i the real-space vectors: |n_i| <= |k+G| |a_i| / 2 pi + |k_i|
  double a[9];
  {
    // a_i = 2 pi (b^-1)^T rows: solve via the 3 x 3 inverse of bvec (column-major)
    const double* b = bvec;
    const double det = b[0] * (b[4] * b[8] - b[7] * b[5]) - b[3] * (b[1] * b[8] - b[7] * b[2]) +
                       b[6] * (b[1] * b[5] - b[4] * b[2]);
    if (!(std::fabs(det) > 0.0)) return -1;
    // rows of inv(B): the real-space vectors over 2 pi
    const double inv[9] = {(b[4] * b[8] - b[5] * b[7]) / det, (b[7] * b[2] - b[1] * b[8]) / det,
                           (b[1] * b[5] - b[2] * b[4]) / det, (b[6] * b[5] - b[3] * b[8]) / det,
                           (b[0] * b[8] - b[6] * b[2]) / det, (b[3] * b[2] - b[0] * b[5]) / det,
                           (b[3] * b[7] - b[6] * b[4]) / det, (b[6] * b[1] - b[0] * b[7]) / det,
                           (b[0] * b[4] - b[3] * b[1]) / det};
    std::memcpy(a, inv, sizeof a);  // inv[3 i + j] = (B^-1)(i, j)
  }
  const double gm = std::sqrt(ecut);
  int lim[3];
  for (int i = 0; i < 3; ++i) {
    const double r = std::sqrt(a[3 * i] * a[3 * i] + a[3 * i + 1] * a[3 * i + 1] +
                               a[3 * i + 2] * a[3 * i + 2]);
    lim[i] = (int)std::floor(gm * r + std::fabs(kf[i])) + 1;
  }
  if (2 * lim[0] + 1 > 4 * n1 || 2 * lim[1] + 1 > 4 * n2 || 2 * lim[2] + 1 > 4 * n3) return -2;
  std::map<std::pair<int, int>, std::vector<std::pair<int, double>>> cols;
  for (int h = -lim[0]; h <= lim[0]; ++h)
    for (int kk = -lim[1]; kk <= lim[1]; ++kk)
      for (int l = -lim[2]; l <= lim[2]; ++l) {
        const double f0 = h + kf[0], f1 = kk + kf[1], f2 = l + kf[2];
        const double gx = bvec[0] * f0 + bvec[3] * f1 + bvec[6] * f2;
        const double gy = bvec[1] * f0 + bvec[4] * f1 + bvec[7] * f2;
        const double gz = bvec[2] * f0 + bvec[5] * f1 + bvec[8] * f2;
        const double g2 = gx * gx + gy * gy + gz * gz;
        if (g2 <= ecut) cols[{h, kk}].emplace_back(l, g2);
      }
  hkl.clear();
  kin.clear();
  ncol = (int)cols.size();
  for (auto& kv : cols) {
    // a column must not wrap onto itself along z, nor two columns onto one
    // grid column (the sphere's extent stays below the grid edge)
    if (2 * std::abs(kv.first.first) >= n1 || 2 * std::abs(kv.first.second) >= n2) return -2;
    for (auto& pl : kv.second) {
      if (2 * std::abs(pl.first) >= n3) return -2;
      hkl.push_back(kv.first.first);
      hkl.push_back(kv.first.second);
      hkl.push_back(pl.first);
      kin.push_back(pl.second);
    }
  }
  return 0;
}

// the sphere alone (no context): returns npw, or -npw when cap < npw, or a
// negative error of k_sphere offset by -1e9
int kp_sphere_host(const double* bvec, const double* kf, double ecut, int n1, int n2, int n3,
                   int cap, int* hkl, double* kin) {
  std::vector<int> h;
  std::vector<double> e;
  int nc = 0;
  if (int rc = k_sphere(bvec, kf, ecut, n1, n2, n3, h, e, nc)) return rc - 1000000000;
  const int n = (int)e.size();
  if (cap < n) return -n;
  if (hkl) std::memcpy(hkl, h.data(), sizeof(int) * h.size());
  if (kin) std::memcpy(kin, e.data(), sizeof(double) * e.size());
  return n;
}

int kp_set(Ctx* c, int nk, const double* kfrac, const double* bvec, double ecut, int nb,
           int* npw_out) {
  if (nk < 1 || nb < 1 || !(ecut > 0.0)) return -1;
  if (c->n3 > 128 || c->n1 * c->n2 > 8192) return -2;  // LDS of the z and plane kernels
  kp_destroy(c);
  auto* k = new KPoints();
  c->kp = k;
  k->nk = nk;
  k->nb = nb;
  k->kfrac.assign(kfrac, kfrac + 3 * nk);
  k->hkl.resize(nk);
  k->kin.resize(nk);
  k->npw.resize(nk);
  k->ncol.resize(nk);
  for (int i = 0; i < nk; ++i) {
    int nc = 0;
    if (int rc = k_sphere(bvec, kfrac + 3 * i, ecut, c->n1, c->n2, c->n3, k->hkl[i], k->kin[i], nc))
      return rc;
    k->npw[i] = (int)k->kin[i].size();
    k->ncol[i] = nc;
    k->npw_max = std::max(k->npw_max, k->npw[i]);
    k->ncol_max = std::max(k->ncol_max, nc);
    if (npw_out) npw_out[i] = k->npw[i];
  }
  const int pm = k->npw_max, cm = k->ncol_max;
  std::vector<int> col_p0((size_t)nk * (cm + 1), 0), col_xy((size_t)nk * cm, 0),
      pt_z((size_t)nk * pm, 0);
  std::vector<double> kin((size_t)nk * pm, 0.0);
  for (int i = 0; i < nk; ++i) {
    const auto& g = k->hkl[i];
    int col = -1, ph = 0, pk = 0;
    for (int p = 0; p < k->npw[i]; ++p) {
      const int h = g[3 * p], kk = g[3 * p + 1], l = g[3 * p + 2];
      if (col < 0 || h != ph || kk != pk) {
        ++col;
        col_p0[(size_t)i * (cm + 1) + col] = p;
        col_xy[(size_t)i * cm + col] = ((h % c->n1 + c->n1) % c->n1) +
                                       c->n1 * ((kk % c->n2 + c->n2) % c->n2);
        ph = h;
        pk = kk;
      }
      pt_z[(size_t)i * pm + p] = (l % c->n3 + c->n3) % c->n3;
      kin[(size_t)i * pm + p] = k->kin[i][p];
    }
    for (int cc = col + 1; cc <= cm; ++cc) col_p0[(size_t)i * (cm + 1) + cc] = k->npw[i];
  }
  auto up = [&](auto** d, const auto& v) -> int {
    QX_CHECK(hipMalloc(d, sizeof(v[0]) * std::max<size_t>(v.size(), 1)));
    QX_CHECK(hipMemcpy(*d, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice));
    return 0;
  };
  if (int rc = up(&k->d_npw, k->npw)) return rc;
  if (int rc = up(&k->d_ncol, k->ncol)) return rc;
  if (int rc = up(&k->d_col_p0, col_p0)) return rc;
  if (int rc = up(&k->d_col_xy, col_xy)) return rc;
  if (int rc = up(&k->d_pt_z, pt_z)) return rc;
  if (int rc = up(&k->d_kin, kin)) return rc;
  const size_t vbytes = sizeof(double) * 2 * (size_t)pm * nb * nk;
  for (auto*& s : k->d_slot) {
    QX_CHECK(hipMalloc(&s, vbytes));
    QX_CHECK(hipMemset(s, 0, vbytes));
  }
  // pencil columns of up to all (k, band) blocks, at most 1 GiB
  const size_t per = sizeof(double2) * (size_t)c->n3 * cm;
  k->batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nk * nb, (1ull << 30) / per));
  QX_CHECK(hipMalloc(&k->d_T, per * k->batch));
  const size_t lds = plane_lds(c);
  if (lds > 65536)
    QX_CHECK(hipFuncSetAttribute((const void*)k_kplane, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds));
  buffers_changed(c);
  return 0;
}

int kp_sphere(Ctx* c, int ik, int* hkl, double* kin) {
  auto* k = static_cast<KPoints*>(c->kp);
  if (!k || ik < 0 || ik >= k->nk) return -1;
  if (hkl) std::memcpy(hkl, k->hkl[ik].data(), sizeof(int) * k->hkl[ik].size());
  if (kin) std::memcpy(kin, k->kin[ik].data(), sizeof(double) * k->kin[ik].size());
  return 0;
}

static int kp_slot_check(Ctx* c, int slot, int ik, int nb) {
  auto* k = static_cast<KPoints*>(c->kp);
  if (!k) return -1;
  if (slot < 0 || slot >= KPoints::kSlots || ik < 0 || ik >= k->nk || nb < 0 || nb > k->nb) return -1;
  return 0;
}

// bands 0..nb-1 of k-point ik, complex interleaved with stride 2 npw_k
int kp_slot_put(Ctx* c, int slot, int ik, const double* psi, int nb) {
  if (int rc = kp_slot_check(c, slot, ik, nb)) return rc;
  auto* k = static_cast<KPoints*>(c->kp);
  double* d = k->d_slot[slot] + (size_t)ik * k->nb * 2 * k->npw_max;
  QX_CHECK(hipMemcpy2DAsync(d, sizeof(double) * 2 * k->npw_max, psi,
                            sizeof(double) * 2 * k->npw[ik], sizeof(double) * 2 * k->npw[ik], nb,
                            hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

int kp_slot_get(Ctx* c, int slot, int ik, double* psi, int nb) {
  if (int rc = kp_slot_check(c, slot, ik, nb)) return rc;
  auto* k = static_cast<KPoints*>(c->kp);
  const double* d = k->d_slot[slot] + (size_t)ik * k->nb * 2 * k->npw_max;
  QX_CHECK(hipMemcpy2DAsync(psi, sizeof(double) * 2 * k->npw[ik], d,
                            sizeof(double) * 2 * k->npw_max, sizeof(double) * 2 * k->npw[ik], nb,
                            hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  return 0;
}

// projector table of k-point ik: beta (npj x 2 npw_k, complex interleaved),
// D (npj); every k-point takes the same projector count
int kp_set_kb(Ctx* c, int ik, int npj, const double* beta, const double* D) {
  auto* k = static_cast<KPoints*>(c->kp);
  if (!k || ik < 0 || ik >= k->nk || npj < 0) return -1;
  if (k->npj != npj) {
    if (k->d_A) (void)hipFree(k->d_A);
    if (k->d_D) (void)hipFree(k->d_D);
    if (k->d_P) (void)hipFree(k->d_P);
    k->d_A = k->d_D = k->d_P = nullptr;
    k->npj = npj;
    if (npj == 0) return 0;
    const size_t abytes = sizeof(double) * 2 * (size_t)npj * 2 * k->npw_max * k->nk;
    QX_CHECK(hipMalloc(&k->d_A, abytes));
    QX_CHECK(hipMemset(k->d_A, 0, abytes));
    QX_CHECK(hipMalloc(&k->d_D, sizeof(double) * 2 * npj));
    QX_CHECK(hipMalloc(&k->d_P, sizeof(double) * 2 * (size_t)npj * k->nb));
    buffers_changed(c);
  }
  if (npj == 0) return 0;
  const int np = k->npw[ik], ld = 2 * k->npw_max;
  // rows 0..npj-1: beta; rows npj..2npj-1: i beta = (-Im, Re)
  std::vector<double> A((size_t)2 * npj * ld, 0.0);
  for (int p = 0; p < npj; ++p)
    for (int g = 0; g < np; ++g) {
      const double re = beta[(size_t)p * 2 * np + 2 * g], im = beta[(size_t)p * 2 * np + 2 * g + 1];
      A[(size_t)p * ld + 2 * g] = re;
      A[(size_t)p * ld + 2 * g + 1] = im;
      A[(size_t)(npj + p) * ld + 2 * g] = -im;
      A[(size_t)(npj + p) * ld + 2 * g + 1] = re;
    }
  QX_CHECK(hipMemcpy(k->d_A + (size_t)ik * 2 * npj * ld, A.data(), sizeof(double) * A.size(),
                     hipMemcpyHostToDevice));
  std::vector<double> D2(2 * (size_t)npj);
  for (int p = 0; p < npj; ++p) D2[p] = D2[npj + p] = D[p];
  QX_CHECK(hipMemcpy(k->d_D, D2.data(), sizeof(double) * D2.size(), hipMemcpyHostToDevice));
  return 0;
}

// local part of H for blocks kb0 .. kb0 + nkb - 1 (all within one batch of T);
// band b of k-point k is row k rpk + r0 + b of src and dst
static int kp_local_run(Ctx* c, KPoints* k, const double* src, double* dst, int kb0, int nkb,
                        int rpk, int r0) {
  const KArgs a = kargs(c, k, kb0, rpk, r0);
  const dim3 gz(cdivu(k->ncol_max, kKCols), nkb);
  hipLaunchKernelGGL(k_kzinv, gz, dim3(kKThreads), 0, c->stream, a, src, k->d_T);
  hipLaunchKernelGGL(k_kplane, dim3(c->n3, nkb), dim3(kKThreads), plane_lds(c), c->stream, a,
                     c->d_V, k->d_T);
  hipLaunchKernelGGL(k_kzfwd, gz, dim3(kKThreads), 0, c->stream, a, src, k->d_T, dst,
                     1.0 / (double)c->N, 0);
  QX_CHECK(hipGetLastError());
  return 0;
}

// H_k for the first nb bands of every k-point: slot src -> slot dst.
// per_band = 0: every stage over all (k, band) blocks of a batch in one launch
// (one launch per batch; one batch when all blocks fit), the KB products one
// pair per k-point; per_band = 1: the comparison loop, one (k, band) block per
// launch of each stage and the KB products per band.
static int kp_apply(Ctx* c, KPoints* k, const double* S, double* H, int rpk, int r0, int nb,
                    int per_band) {
  const size_t bstride = (size_t)2 * k->npw_max;
  const int total = k->nk * k->nb;
  if (per_band) {
    for (int ik = 0; ik < k->nk; ++ik)
      for (int b = 0; b < nb; ++b)
        if (int rc = kp_local_run(c, k, S, H, ik * k->nb + b, 1, rpk, r0)) return rc;
  } else if (nb == k->nb) {
    for (int kb0 = 0; kb0 < total; kb0 += k->batch)
      if (int rc = kp_local_run(c, k, S, H, kb0, std::min(k->batch, total - kb0), rpk, r0))
        return rc;
  } else {
    for (int ik = 0; ik < k->nk; ++ik)
      for (int b0 = 0; b0 < nb; b0 += k->batch)
        if (int rc = kp_local_run(c, k, S, H, ik * k->nb + b0, std::min(k->batch, nb - b0), rpk,
                                  r0))
          return rc;
  }
  if (k->npj == 0) return 0;
  const int npj2 = 2 * k->npj, K = (int)bstride;
  for (int ik = 0; ik < k->nk; ++ik) {
    const double* A = k->d_A + (size_t)ik * npj2 * bstride;
    const double* Sk = S + ((size_t)ik * rpk + r0) * bstride;
    double* Hk = H + ((size_t)ik * rpk + r0) * bstride;
    const int step = per_band ? 1 : nb;
    for (int b0 = 0; b0 < nb; b0 += step) {
      const int n = std::min(step, nb - b0);
      if (int rc = gemm_tn(c, npj2, n, K, 1.0, A, K, Sk + b0 * bstride, K, 0.0, k->d_P, npj2, 0))
        return rc;
      hipLaunchKernelGGL(k_kscale, dim3(cdivu((long)npj2 * n, 256)), dim3(256), 0, c->stream,
                         k->d_P, k->d_D, k->npj, n);
      if (int rc = gemm_nn(c, K, n, npj2, 1.0, A, K, k->d_P, npj2, 1.0, Hk + b0 * bstride, K))
        return rc;
    }
  }
  QX_CHECK(hipGetLastError());
  return 0;
}

// ---- band solver ------------------------------------------------------------
static int kp_solver_alloc(Ctx* c, KPoints* k) {
  if (k->d_Z) return 0;
  const size_t L2 = 2 * (size_t)k->npw_max, nb = k->nb, nk = k->nk;
  const size_t zbytes = sizeof(double) * nk * 4 * nb * L2;
  QX_CHECK(hipMalloc(&k->d_Z, zbytes));
  QX_CHECK(hipMalloc(&k->d_HZ, zbytes));
  QX_CHECK(hipMemset(k->d_Z, 0, zbytes));
  QX_CHECK(hipMemset(k->d_HZ, 0, zbytes));
  QX_CHECK(hipMalloc(&k->d_G, sizeof(double) * nk * 2 * (2 * nb) * (4 * nb)));
  QX_CHECK(hipMalloc(&k->d_Hs, sizeof(rocblas_double_complex) * nk * 4 * nb * nb));
  QX_CHECK(hipMalloc(&k->d_Ss, sizeof(rocblas_double_complex) * nk * 4 * nb * nb));
  QX_CHECK(hipMalloc(&k->d_w, sizeof(double) * nk * 2 * nb));
  QX_CHECK(hipMalloc(&k->d_E, sizeof(double) * nk * 2 * nb));
  QX_CHECK(hipMalloc(&k->d_info, sizeof(int) * nk));
  QX_CHECK(hipMalloc(&k->d_Cs, sizeof(double) * nk * nb * 4 * nb));
  QX_CHECK(hipMalloc(&k->d_ekin, sizeof(double) * nk * nb));
  QX_CHECK(hipMalloc(&k->d_lam, sizeof(double) * nk * nb));
  if (rocblas_create_handle(&k->rb) != rocblas_status_success) return -3;
  buffers_changed(c);
  return 0;
}

// Rayleigh-Ritz on the basis Y = rows [0, nbs) of each k-point's Z block (its
// J Y in rows [nbs, 2 nbs), H Y in HZ likewise): the generalized Hermitian
// problem (Y^H H Y) x = lambda (Y^H Y) x of every k-point in one batched
// solve; the lowest nb Ritz vectors and H of them go to rows [0, nb) of the
// k-slots psi / hpsi (nb rows per k-point), the Ritz values to d_lam and lam.
static int kp_rayleigh_ritz(Ctx* c, KPoints* k, int nbs, double* psi, double* hpsi,
                            double* lam_host) {
  const int nb = k->nb, rpk = 4 * nb, ldg = 4 * nb, ld2 = 2 * nb;
  const long L2 = 2L * k->npw_max;
  const int K = (int)L2;
  // J Y and J H Y
  hipLaunchKernelGGL(k_kjrows, dim3(64, k->nk * nbs), dim3(256), 0, c->stream, k->d_Z, rpk, nbs, L2);
  hipLaunchKernelGGL(k_kjrows, dim3(64, k->nk * nbs), dim3(256), 0, c->stream, k->d_HZ, rpk, nbs,
                     L2);
  const long strideG = 2L * ld2 * ldg, strideA = (long)ld2 * ld2;
  for (int ik = 0; ik < k->nk; ++ik) {
    const double* Zk = k->d_Z + (size_t)ik * rpk * L2;
    const double* HZk = k->d_HZ + (size_t)ik * rpk * L2;
    double* Gs = k->d_G + ik * strideG;
    double* Gh = Gs + (size_t)ld2 * ldg;
    // G[n][m] = <z_m, y_n> over [Y; J Y] x Y: Re and Im of Y^H Y (and Y^H H Y)
    if (int rc = gemm_tn(c, 2 * nbs, nbs, K, 1.0, Zk, K, Zk, K, 0.0, Gs, ldg, 0)) return rc;
    if (int rc = gemm_tn(c, 2 * nbs, nbs, K, 1.0, Zk, K, HZk, K, 0.0, Gh, ldg, 0)) return rc;
  }
  const dim3 gp(cdivu((long)nbs * nbs, 256), k->nk);
  hipLaunchKernelGGL(k_kpack, gp, dim3(256), 0, c->stream, k->d_G, ldg, nbs, ld2, k->d_Ss,
                     strideG, strideA);
  hipLaunchKernelGGL(k_kpack, gp, dim3(256), 0, c->stream, k->d_G + (size_t)ld2 * ldg, ldg, nbs,
                     ld2, k->d_Hs, strideG, strideA);
  QX_CHECK(hipGetLastError());
  if (rocblas_set_stream(k->rb, c->stream) != rocblas_status_success) return -3;
  if (rocsolver_zhegvd_strided_batched(k->rb, rocblas_eform_ax, rocblas_evect_original,
                                       rocblas_fill_lower, nbs, k->d_Hs, ld2, strideA, k->d_Ss,
                                       ld2, strideA, k->d_w, ld2, k->d_E, ld2, k->d_info,
                                       k->nk) != rocblas_status_success)
    return -3;
  std::vector<int> info(k->nk);
  QX_CHECK(hipMemcpyAsync(info.data(), k->d_info, sizeof(int) * k->nk, hipMemcpyDeviceToHost,
                          c->stream));
  std::vector<double> w((size_t)k->nk * ld2);
  QX_CHECK(hipMemcpyAsync(w.data(), k->d_w, sizeof(double) * w.size(), hipMemcpyDeviceToHost,
                          c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  for (int ik = 0; ik < k->nk; ++ik)
    if (info[ik] != 0) return -21;  // the overlap of the basis is not positive definite
  std::vector<double> lam((size_t)k->nk * nb);
  for (int ik = 0; ik < k->nk; ++ik)
    for (int b = 0; b < nb; ++b) lam[(size_t)ik * nb + b] = w[(size_t)ik * ld2 + b];
  if (lam_host) std::memcpy(lam_host, lam.data(), sizeof(double) * lam.size());
  QX_CHECK(hipMemcpyAsync(k->d_lam, lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice,
                          c->stream));
  // psi = Y C, H psi = (H Y) C: real NN products on [Y; J Y] with [Re C; Im C]
  const long strideC = (long)nb * ldg;
  hipLaunchKernelGGL(k_kunpack, dim3(cdivu((long)nbs * nb, 256), k->nk), dim3(256), 0, c->stream,
                     k->d_Hs, ld2, nbs, nb, k->d_Cs, ldg, strideA, strideC);
  for (int ik = 0; ik < k->nk; ++ik) {
    const double* Zk = k->d_Z + (size_t)ik * rpk * L2;
    const double* HZk = k->d_HZ + (size_t)ik * rpk * L2;
    const double* Ck = k->d_Cs + ik * strideC;
    double* Pk = psi + (size_t)ik * nb * L2;
    double* Hk = hpsi + (size_t)ik * nb * L2;
    if (int rc = gemm_nn(c, K, nb, 2 * nbs, 1.0, Zk, K, Ck, ldg, 0.0, Pk, K)) return rc;
    if (int rc = gemm_nn(c, K, nb, 2 * nbs, 1.0, HZk, K, Ck, ldg, 0.0, Hk, K)) return rc;
  }
  QX_CHECK(hipGetLastError());
  return 0;
}

// Lowest nb eigenpairs of H_k at every k-point in the fixed local potential:
// block Rayleigh-Ritz on [psi, W], W the preconditioned residuals (Teter-
// Payne-Allan, as prpcg) scaled to unit norm, all k-points per step.  Starts
// from the bands in k-slot src (nb per k-point), leaves the Ritz vectors in
// src and H of them in dst; eig (nk x nb) the Ritz values; stops when no Ritz
// value moves by more than tol (Ry) or after max_iter steps.
int kp_solve(Ctx* c, int src, int dst, int max_iter, double tol, double* eig, int* iters) {
  auto* k = static_cast<KPoints*>(c->kp);
  if (!k || src < 0 || dst < 0 || src >= KPoints::kSlots || dst >= KPoints::kSlots || src == dst ||
      max_iter < 0)
    return -1;
  if (int rc = kp_solver_alloc(c, k)) return rc;
  const int nb = k->nb, rpk = 4 * nb;
  const long L2 = 2L * k->npw_max;
  double* P = k->d_slot[src];
  double* HP = k->d_slot[dst];
  const dim3 gr(64, k->nk * nb);
  // Ritz vectors of the starting bands
  if (int rc = kp_apply(c, k, P, HP, nb, 0, nb, 0)) return rc;
  hipLaunchKernelGGL(k_krows, gr, dim3(256), 0, c->stream, P, nb, 0, k->d_Z, rpk, 0, nb, L2);
  hipLaunchKernelGGL(k_krows, gr, dim3(256), 0, c->stream, HP, nb, 0, k->d_HZ, rpk, 0, nb, L2);
  std::vector<double> lam((size_t)k->nk * nb), prev;
  if (int rc = kp_rayleigh_ritz(c, k, nb, P, HP, lam.data())) return rc;
  int it = 0;
  for (; it < max_iter; ++it) {
    hipLaunchKernelGGL(k_krows, gr, dim3(256), 0, c->stream, P, nb, 0, k->d_Z, rpk, 0, nb, L2);
    hipLaunchKernelGGL(k_krows, gr, dim3(256), 0, c->stream, HP, nb, 0, k->d_HZ, rpk, 0, nb, L2);
    hipLaunchKernelGGL(k_kekin, dim3(k->nk * nb), dim3(256), 0, c->stream, k->d_Z, rpk, nb,
                       k->npw_max, k->d_kin, k->d_ekin);
    hipLaunchKernelGGL(k_kresid, dim3(k->nk * nb), dim3(256), 0, c->stream, k->d_Z, k->d_HZ, rpk,
                       nb, k->npw_max, k->d_npw, k->d_kin, k->d_lam, k->d_ekin);
    if (int rc = kp_apply(c, k, k->d_Z, k->d_HZ, rpk, nb, nb, 0)) return rc;
    prev = lam;
    if (int rc = kp_rayleigh_ritz(c, k, 2 * nb, P, HP, lam.data())) return rc;
    double dmax = 0.0;
    for (size_t i = 0; i < lam.size(); ++i) dmax = std::max(dmax, std::fabs(lam[i] - prev[i]));
    if (dmax < tol) {
      ++it;
      break;
    }
  }
  if (eig) std::memcpy(eig, lam.data(), sizeof(double) * lam.size());
  if (iters) *iters = it;
  return 0;
}

int kp_hpsi(Ctx* c, int src, int dst, int nb, int per_band) {
  auto* k = static_cast<KPoints*>(c->kp);
  if (!k || src < 0 || dst < 0 || src >= KPoints::kSlots || dst >= KPoints::kSlots || src == dst ||
      nb < 1 || nb > k->nb)
    return -1;
  return kp_apply(c, k, k->d_slot[src], k->d_slot[dst], k->nb, 0, nb, per_band);
}

}  // namespace qx
