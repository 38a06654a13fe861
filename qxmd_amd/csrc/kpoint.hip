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
#include "context.h"

namespace qx {

namespace {

constexpr int kKThreads = 256;
constexpr int kKCols = 16;  // columns per block of the z kernels

struct KArgs {
  int n1, n2, n3, nb, npw_max, ncol_max, kb0;
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
  const double* src = psi + (size_t)kb * 2 * a.npw_max;
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
  double* dst = out + (size_t)kb * 2 * a.npw_max;
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
  const double* src = psi + (size_t)kb * 2 * a.npw_max;
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
};

namespace {

void kp_free(KPoints* k) {
  void* ptrs[] = {k->d_npw, k->d_ncol, k->d_col_p0, k->d_col_xy, k->d_pt_z, k->d_kin,
                  k->d_T, k->d_A, k->d_D, k->d_P};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto* p : k->d_slot)
    if (p) (void)hipFree(p);
}

KArgs kargs(const Ctx* c, const KPoints* k, int kb0) {
  KArgs a{};
  a.n1 = c->n1; a.n2 = c->n2; a.n3 = c->n3;
  a.nb = k->nb; a.npw_max = k->npw_max; a.ncol_max = k->ncol_max; a.kb0 = kb0;
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

// local part of H for blocks kb0 .. kb0 + nkb - 1 (all within one batch of T)
static int kp_local_run(Ctx* c, KPoints* k, const double* src, double* dst, int kb0, int nkb) {
  const KArgs a = kargs(c, k, kb0);
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
int kp_hpsi(Ctx* c, int src, int dst, int nb, int per_band) {
  auto* k = static_cast<KPoints*>(c->kp);
  if (!k || src < 0 || dst < 0 || src >= KPoints::kSlots || dst >= KPoints::kSlots || src == dst ||
      nb < 1 || nb > k->nb)
    return -1;
  const double* S = k->d_slot[src];
  double* H = k->d_slot[dst];
  const size_t bstride = (size_t)2 * k->npw_max;
  const int total = k->nk * k->nb;
  if (per_band) {
    for (int ik = 0; ik < k->nk; ++ik)
      for (int b = 0; b < nb; ++b)
        if (int rc = kp_local_run(c, k, S, H, ik * k->nb + b, 1)) return rc;
  } else if (nb == k->nb) {
    for (int kb0 = 0; kb0 < total; kb0 += k->batch)
      if (int rc = kp_local_run(c, k, S, H, kb0, std::min(k->batch, total - kb0))) return rc;
  } else {
    for (int ik = 0; ik < k->nk; ++ik)
      for (int b0 = 0; b0 < nb; b0 += k->batch)
        if (int rc = kp_local_run(c, k, S, H, ik * k->nb + b0, std::min(k->batch, nb - b0)))
          return rc;
  }
  if (k->npj == 0) return 0;
  const int npj2 = 2 * k->npj, K = (int)bstride;
  for (int ik = 0; ik < k->nk; ++ik) {
    const double* A = k->d_A + (size_t)ik * npj2 * bstride;
    const double* Sk = S + (size_t)ik * k->nb * bstride;
    double* Hk = H + (size_t)ik * k->nb * bstride;
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

}  // namespace qx
