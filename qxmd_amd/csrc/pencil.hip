// pencil.hip -- sphere-pruned batched 3-D FFTs for the Gamma-point local
// H|psi> (QXMD hcnonloc + ekinetic, Sources/engrad.f90:411-606, 66-127) and the
// soft charge density (softchg, Sources/chgdns.f90:388-454).
//
// Two real bands a, b travel as one complex field psi_a(r) + i psi_b(r).  One
// band-apply batch runs three kernels over all band pairs of the batch:
//   zinv   : gather the half-sphere coefficients of a column pair, rebuild the
//            -G half by conjugation, inverse FFT along z in LDS      -> T1
//   plane  : per z-plane, inverse y (pruned to |h| <= hmax), inverse x, x V(r),
//            forward x, forward y, all in LDS                         -> T2
//   zfwd   : forward FFT along z, split the two bands, x1/N, add |G|^2 psi.
// T1/T2 hold only the columns the sphere touches ([bp][z][col]), so the HBM
// stream per band pair is a few times the sphere size instead of the grid.
#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>
#include "context.h"

namespace qx {

// --------------------------------------------------------------------------
// host: pencil layout
// --------------------------------------------------------------------------
int pencil_build(Ctx* c, const int* nga, const int* ngb, const int* ngc, const double* recnrm) {
  Pencil& P = c->pc;
  const int nplw = c->nplw;
  // canonical key of a column pair: the lexicographically larger of (h,k), (-h,-k)
  auto canon = [](int h, int k) {
    if (h > 0 || (h == 0 && k >= 0)) return std::make_pair(h, k);
    return std::make_pair(-h, -k);
  };
  std::map<std::pair<int, int>, int> pairs;  // canonical -> pair id
  pairs[{0, 0}] = 0;
  for (int g = 1; g <= nplw; ++g) {
    auto key = canon(nga[g], ngb[g]);
    if (!pairs.count(key)) pairs[key] = 0;
  }
  // ids in sorted order (pair 0 stays the (0,0) column)
  int q = 1;
  for (auto& kv : pairs)
    if (!(kv.first.first == 0 && kv.first.second == 0)) kv.second = q++;
  P.npair = q;
  P.ncol = 2 * P.npair - 1;
  std::vector<int> colh(P.ncol), colk(P.ncol);
  for (auto& kv : pairs) {
    int id = kv.second;
    if (id == 0) {
      colh[0] = 0; colk[0] = 0;
    } else {
      colh[2 * id - 1] = kv.first.first;
      colk[2 * id - 1] = kv.first.second;
      colh[2 * id] = -kv.first.first;
      colk[2 * id] = -kv.first.second;
    }
  }
  // bucket the points: (pair, side, l, g)
  std::vector<std::tuple<int, int, int, int>> pts;
  pts.reserve(nplw + 1);
  pts.emplace_back(0, -1, 0, 0);  // G = 0 first
  for (int g = 1; g <= nplw; ++g) {
    int h = nga[g], k = ngb[g];
    auto key = canon(h, k);
    int id = pairs[key];
    int side = (id == 0) ? 0 : ((key.first == h && key.second == k) ? 0 : 1);
    pts.emplace_back(id, side, ngc[g], g);
  }
  std::sort(pts.begin() + 1, pts.end());
  P.perm.resize(nplw + 1);
  P.pt_col.resize(nplw + 1);
  P.pt_z.resize(nplw + 1);
  P.pair_p0.assign(P.npair + 1, 0);
  for (int p = 0; p <= nplw; ++p) {
    int id, side, l, g;
    std::tie(id, side, l, g) = pts[p];
    P.perm[p] = g;
    P.pt_col[p] = (id == 0) ? 0 : (2 * id - 1 + (side < 0 ? 0 : side));
    int z = l % c->n3;
    if (z < 0) z += c->n3;
    P.pt_z[p] = z;
    P.pair_p0[id + 1]++;
  }
  for (int i = 0; i < P.npair; ++i) P.pair_p0[i + 1] += P.pair_p0[i];
  int hmin = 0, hmax = 0;
  for (int i = 0; i < P.ncol; ++i) {
    hmin = std::min(hmin, colh[i]);
    hmax = std::max(hmax, colh[i]);
  }
  if (hmax - hmin + 1 > c->n1 || 2 * std::max(-hmin, hmax) >= c->n1) return -2;  // sphere wraps
  P.hmin = hmin;
  P.nH = hmax - hmin + 1;
  P.col_hi.resize(P.ncol);
  P.col_kmod.resize(P.ncol);
  for (int i = 0; i < P.ncol; ++i) {
    P.col_hi[i] = colh[i] - hmin;
    int km = colk[i] % c->n2;
    if (km < 0) km += c->n2;
    P.col_kmod[i] = km;
  }
  // device copies
  auto up = [&](int** d, const std::vector<int>& h) -> int {
    QX_CHECK(hipMalloc(d, sizeof(int) * h.size()));
    QX_CHECK(hipMemcpy(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    return 0;
  };
  int rc;
  if ((rc = up(&c->d_perm, P.perm))) return rc;
  if ((rc = up(&c->d_pt_col, P.pt_col))) return rc;
  if ((rc = up(&c->d_pt_z, P.pt_z))) return rc;
  if ((rc = up(&c->d_pair_p0, P.pair_p0))) return rc;
  if ((rc = up(&c->d_col_hi, P.col_hi))) return rc;
  if ((rc = up(&c->d_col_kmod, P.col_kmod))) return rc;
  std::vector<double> kin(nplw + 1);
  for (int p = 0; p <= nplw; ++p) kin[p] = recnrm[P.perm[p]];
  c->gmax2 = 0.0;
  for (int p = 0; p <= nplw; ++p) c->gmax2 = std::max(c->gmax2, recnrm[p]);
  QX_CHECK(hipMalloc(&c->d_kin, sizeof(double) * kin.size()));
  QX_CHECK(hipMemcpy(c->d_kin, kin.data(), sizeof(double) * kin.size(), hipMemcpyHostToDevice));
  std::vector<double> tw(2 * (c->n1 + c->n2 + c->n3));
  int o = 0;
  for (int n : {c->n1, c->n2, c->n3})
    for (int k = 0; k < n; ++k, ++o) {
      double a = 2.0 * M_PI * (double)k / (double)n;
      tw[2 * o] = std::cos(a);
      tw[2 * o + 1] = std::sin(a);
    }
  QX_CHECK(hipMalloc(&c->d_tw, sizeof(double) * tw.size()));
  QX_CHECK(hipMemcpy(c->d_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
  return 0;
}

// --------------------------------------------------------------------------
// device kernels
// --------------------------------------------------------------------------
constexpr int kZThreads = 256;
constexpr int kPlaneThreads = 512;
constexpr int kPairsPerBlock = 16;

__device__ inline int mirror_col(int col) {
  return col == 0 ? 0 : ((col & 1) ? col + 1 : col - 1);
}

struct PencilArgs {
  int n1, n2, n3, npair, ncol, nH, nplwex, nb, bp0;
  int rows;  // rows per chunk in the plane kernel
  FftPlan plx, ply, plz;
  const int *pt_col, *pt_z, *pair_p0, *col_hi, *col_kmod;
  const double* kin;
  const double* tw;  // x | y | z
};

// columns [c0, c1) owned by pairs [q0, q1)
__device__ inline void pair_cols(int q0, int q1, int& c0, int& c1) {
  c0 = q0 == 0 ? 0 : 2 * q0 - 1;
  c1 = 2 * q1 - 1;
}

template <typename T>
__global__ void __launch_bounds__(kZThreads) k_zinv(PencilArgs a, const double* __restrict__ psi,
                                                     cplx<T>* __restrict__ t1) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* lines = tw + a.n3;
  const int q0 = blockIdx.x * kPairsPerBlock;
  const int q1 = min(q0 + kPairsPerBlock, a.npair);
  const int bp = blockIdx.y;  // band pair within the batch
  const int ba = 2 * (a.bp0 + bp), bb = ba + 1;
  int c0, c1;
  pair_cols(q0, q1, c0, c1);
  const int ncl = c1 - c0;
  load_tw<T>(tw, a.tw + 2 * (a.n1 + a.n2), a.n3);
  for (int i = threadIdx.x; i < ncl * a.n3; i += blockDim.x) lines[i] = mk<T>(0, 0);
  __syncthreads();
  const double* pa = psi + (size_t)ba * a.nplwex;
  const double* pb = psi + (size_t)bb * a.nplwex;
  const bool hasb = bb < a.nb;
  const int p0 = a.pair_p0[q0], p1 = a.pair_p0[q1];
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    double ar = pa[2 * p], ai = pa[2 * p + 1];
    double br = hasb ? pb[2 * p] : 0.0, bi = hasb ? pb[2 * p + 1] : 0.0;
    int col = a.pt_col[p] - c0, z = a.pt_z[p];
    if (p == 0) {  // G = 0: Im c(0) is not used (wvg2r, Sources/engrad.f90:1527-1530)
      lines[col * a.n3] = mk<T>((T)ar, (T)br);
    } else {
      lines[col * a.n3 + z] = mk<T>((T)(ar - bi), (T)(ai + br));
      int mz = z == 0 ? 0 : a.n3 - z;
      lines[(mirror_col(col + c0) - c0) * a.n3 + mz] = mk<T>((T)(ar + bi), (T)(br - ai));
    }
  }
  __syncthreads();
  lds_fft<T, 4>(lines, ncl, a.n3, 1, a.plz, tw, +1);
  cplx<T>* out = t1 + (size_t)bp * a.n3 * a.ncol;
  for (int i = threadIdx.x; i < ncl * a.n3; i += blockDim.x) {
    int z = i / ncl, cl = i % ncl;
    out[(size_t)z * a.ncol + c0 + cl] = lines[cl * a.n3 + z];
  }
}

template <typename T, typename VT>
__global__ void __launch_bounds__(kPlaneThreads) k_plane(PencilArgs a, const VT* __restrict__ V,
                                                         const cplx<T>* __restrict__ t1,
                                                         cplx<T>* __restrict__ t2) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* twx = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* twy = twx + a.n1;
  cplx<T>* A = twy + a.n2;                 // [nH][n2]
  cplx<T>* R = A + a.nH * a.n2;            // [rows][n1]
  const int z = blockIdx.x, bp = blockIdx.y;
  load_tw<T>(twx, a.tw, a.n1);
  load_tw<T>(twy, a.tw + 2 * a.n1, a.n2);
  for (int i = threadIdx.x; i < a.nH * a.n2; i += blockDim.x) A[i] = mk<T>(0, 0);
  __syncthreads();
  const cplx<T>* in = t1 + ((size_t)bp * a.n3 + z) * a.ncol;
  for (int cl = threadIdx.x; cl < a.ncol; cl += blockDim.x)
    A[a.col_hi[cl] * a.n2 + a.col_kmod[cl]] = in[cl];
  __syncthreads();
  lds_fft<T, 4>(A, a.nH, a.n2, 1, a.ply, twy, +1);
  const int hmin = -((a.nH - 1) / 2);  // columns are symmetric in h
  const VT* Vz = V + (size_t)z * a.n1 * a.n2;
  for (int y0 = 0; y0 < a.n2; y0 += a.rows) {
    const int rows = min(a.rows, a.n2 - y0);
    for (int i = threadIdx.x; i < rows * a.n1; i += blockDim.x) R[i] = mk<T>(0, 0);
    __syncthreads();
    for (int i = threadIdx.x; i < rows * a.nH; i += blockDim.x) {
      int r = i / a.nH, hi = i % a.nH;
      int x = hmin + hi;
      if (x < 0) x += a.n1;
      R[r * a.n1 + x] = A[hi * a.n2 + y0 + r];
    }
    __syncthreads();
    lds_fft<T, 4>(R, rows, a.n1, 1, a.plx, twx, +1);
    for (int i = threadIdx.x; i < rows * a.n1; i += blockDim.x) {
      T v = (T)Vz[(size_t)y0 * a.n1 + i];
      R[i].x *= v;
      R[i].y *= v;
    }
    __syncthreads();
    lds_fft<T, 4>(R, rows, a.n1, 1, a.plx, twx, -1);
    for (int i = threadIdx.x; i < rows * a.nH; i += blockDim.x) {
      int r = i / a.nH, hi = i % a.nH;
      int x = hmin + hi;
      if (x < 0) x += a.n1;
      A[hi * a.n2 + y0 + r] = R[r * a.n1 + x];
    }
    __syncthreads();
  }
  lds_fft<T, 4>(A, a.nH, a.n2, 1, a.ply, twy, -1);
  cplx<T>* out = t2 + ((size_t)bp * a.n3 + z) * a.ncol;
  for (int cl = threadIdx.x; cl < a.ncol; cl += blockDim.x)
    out[cl] = A[a.col_hi[cl] * a.n2 + a.col_kmod[cl]];
}

template <typename T>
__global__ void __launch_bounds__(kZThreads) k_zfwd(PencilArgs a, const double* __restrict__ psi,
                                                     const cplx<T>* __restrict__ t2,
                                                     double* __restrict__ out, int accumulate,
                                                     double invN) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* lines = tw + a.n3;
  const int q0 = blockIdx.x * kPairsPerBlock;
  const int q1 = min(q0 + kPairsPerBlock, a.npair);
  const int bp = blockIdx.y;
  const int ba = 2 * (a.bp0 + bp), bb = ba + 1;
  int c0, c1;
  pair_cols(q0, q1, c0, c1);
  const int ncl = c1 - c0;
  load_tw<T>(tw, a.tw + 2 * (a.n1 + a.n2), a.n3);
  const cplx<T>* in = t2 + (size_t)bp * a.n3 * a.ncol;
  for (int i = threadIdx.x; i < ncl * a.n3; i += blockDim.x) {
    int z = i / ncl, cl = i % ncl;
    lines[cl * a.n3 + z] = in[(size_t)z * a.ncol + c0 + cl];
  }
  __syncthreads();
  lds_fft<T, 4>(lines, ncl, a.n3, 1, a.plz, tw, -1);
  const double* pa = psi + (size_t)ba * a.nplwex;
  const double* pb = psi + (size_t)bb * a.nplwex;
  double* oa = out + (size_t)ba * a.nplwex;
  double* ob = out + (size_t)bb * a.nplwex;
  const bool hasb = bb < a.nb;
  const int p0 = a.pair_p0[q0], p1 = a.pair_p0[q1];
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    int col = a.pt_col[p] - c0, z = a.pt_z[p];
    int mz = z == 0 ? 0 : a.n3 - z;
    cplx<T> zp = lines[col * a.n3 + z];
    cplx<T> zm = lines[(mirror_col(col + c0) - c0) * a.n3 + mz];
    // A = (Zp + conj Zm)/2, B = (Zp - conj Zm)/(2i), both x 1/N
    double Ar = 0.5 * ((double)zp.x + (double)zm.x) * invN;
    double Ai = 0.5 * ((double)zp.y - (double)zm.y) * invN;
    double Br = 0.5 * ((double)zp.y + (double)zm.y) * invN;
    double Bi = -0.5 * ((double)zp.x - (double)zm.x) * invN;
    double kin = a.kin[p];
    double r0 = kin * pa[2 * p] + Ar, r1 = kin * pa[2 * p + 1] + Ai;
    if (accumulate) { r0 += oa[2 * p]; r1 += oa[2 * p + 1]; }
    oa[2 * p] = r0;
    oa[2 * p + 1] = r1;
    if (hasb) {
      double s0 = kin * pb[2 * p] + Br, s1 = kin * pb[2 * p + 1] + Bi;
      if (accumulate) { s0 += ob[2 * p]; s1 += ob[2 * p + 1]; }
      ob[2 * p] = s0;
      ob[2 * p + 1] = s1;
    }
  }
}

// density: rho_part[g][z-plane] += sum over the group's band pairs of
// w_a psi_a(r)^2 + w_b psi_b(r)^2 (softchg, Sources/chgdns.f90:421-439)
template <typename T>
__global__ void __launch_bounds__(kPlaneThreads) k_plane_rho(PencilArgs a, const cplx<T>* __restrict__ t1,
                                                             int nbp_batch, int bp_per_group,
                                                             const double* __restrict__ w,
                                                             double* __restrict__ rho_part) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T>* twx = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* twy = twx + a.n1;
  cplx<T>* A = twy + a.n2;
  cplx<T>* R = A + a.nH * a.n2;
  const int z = blockIdx.x, grp = blockIdx.y;
  load_tw<T>(twx, a.tw, a.n1);
  load_tw<T>(twy, a.tw + 2 * a.n1, a.n2);
  double* rp = rho_part + ((size_t)grp * a.n3 + z) * a.n1 * a.n2;
  const int hmin = -((a.nH - 1) / 2);
  const int bpe = min(nbp_batch, (grp + 1) * bp_per_group);
  for (int bp = grp * bp_per_group; bp < bpe; ++bp) {
    const int ba = 2 * (a.bp0 + bp), bb = ba + 1;
    const double wa = w[ba], wb = bb < a.nb ? w[bb] : 0.0;
    __syncthreads();
    for (int i = threadIdx.x; i < a.nH * a.n2; i += blockDim.x) A[i] = mk<T>(0, 0);
    __syncthreads();
    const cplx<T>* in = t1 + ((size_t)bp * a.n3 + z) * a.ncol;
    for (int cl = threadIdx.x; cl < a.ncol; cl += blockDim.x)
      A[a.col_hi[cl] * a.n2 + a.col_kmod[cl]] = in[cl];
    __syncthreads();
    lds_fft<T, 4>(A, a.nH, a.n2, 1, a.ply, twy, +1);
    for (int y0 = 0; y0 < a.n2; y0 += a.rows) {
      const int rows = min(a.rows, a.n2 - y0);
      for (int i = threadIdx.x; i < rows * a.n1; i += blockDim.x) R[i] = mk<T>(0, 0);
      __syncthreads();
      for (int i = threadIdx.x; i < rows * a.nH; i += blockDim.x) {
        int r = i / a.nH, hi = i % a.nH;
        int x = hmin + hi;
        if (x < 0) x += a.n1;
        R[r * a.n1 + x] = A[hi * a.n2 + y0 + r];
      }
      __syncthreads();
      lds_fft<T, 4>(R, rows, a.n1, 1, a.plx, twx, +1);
      for (int i = threadIdx.x; i < rows * a.n1; i += blockDim.x) {
        double re = (double)R[i].x, im = (double)R[i].y;
        double v = (bp == grp * bp_per_group) ? 0.0 : rp[(size_t)y0 * a.n1 + i];
        rp[(size_t)y0 * a.n1 + i] = v + re * re * wa + im * im * wb;
      }
      __syncthreads();
    }
  }
}

__global__ void k_rho_reduce(const double* __restrict__ part, int ngrp, long N,
                             double* __restrict__ rho) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (int g = 0; g < ngrp; ++g) s += part[(size_t)g * N + i];
  rho[i] = s;
}

// --------------------------------------------------------------------------
// host drivers
// --------------------------------------------------------------------------
static PencilArgs make_args(Ctx* c, int nb, int bp0) {
  PencilArgs a{};
  a.n1 = c->n1; a.n2 = c->n2; a.n3 = c->n3;
  a.npair = c->pc.npair; a.ncol = c->pc.ncol; a.nH = c->pc.nH;
  a.nplwex = c->nplwex; a.nb = nb; a.bp0 = bp0;
  a.plx = c->plx; a.ply = c->ply; a.plz = c->plz;
  a.pt_col = c->d_pt_col; a.pt_z = c->d_pt_z; a.pair_p0 = c->d_pair_p0;
  a.col_hi = c->d_col_hi; a.col_kmod = c->d_col_kmod;
  a.kin = c->d_kin; a.tw = c->d_tw;
  return a;
}

template <typename T>
static size_t plane_lds(Ctx* c, int rows) {
  return sizeof(cplx<T>) * ((size_t)c->n1 + c->n2 + (size_t)c->pc.nH * c->n2 + (size_t)rows * c->n1);
}

template <typename T>
static int plane_rows(Ctx* c) {
  const size_t limit = 160 * 1024 - 1024;
  int rows = c->n2;
  while (rows > 1 && plane_lds<T>(c, rows) > limit) rows = (rows + 1) / 2;
  if (plane_lds<T>(c, rows) > limit) return -1;
  return rows;
}

template <typename T>
static size_t zlines_lds(Ctx* c) {
  return sizeof(cplx<T>) * ((size_t)c->n3 + (size_t)(2 * kPairsPerBlock) * c->n3);
}

// allow dynamic LDS beyond the 64 KiB default (gfx950 has 160 KiB per CU)
template <typename K>
static void lds_optin(K kernel, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename T>
static int local_apply_t(Ctx* c, const double* psi, double* out, int nb, int accumulate) {
  const int nbp = (nb + 1) / 2;
  const int rows = plane_rows<T>(c);
  if (rows < 1) return -3;
  const size_t per_bp = sizeof(cplx<T>) * (size_t)c->n3 * c->pc.ncol;
  const int batch = (int)std::max<size_t>(1, std::min<size_t>(c->t_bytes / per_bp, (size_t)nbp));
  const dim3 gz((c->pc.npair + kPairsPerBlock - 1) / kPairsPerBlock, 1);
  const size_t lz = zlines_lds<T>(c), lp = plane_lds<T>(c, rows);
  cplx<T>* t1 = reinterpret_cast<cplx<T>*>(c->d_t1);
  cplx<T>* t2 = reinterpret_cast<cplx<T>*>(c->d_t2);
  lds_optin(k_zinv<T>, lz);
  lds_optin(k_zfwd<T>, lz);
  if (sizeof(T) == 8) lds_optin(k_plane<T, double>, lp);
  else lds_optin(k_plane<T, float>, lp);
  for (int bp0 = 0; bp0 < nbp; bp0 += batch) {
    const int nbb = std::min(batch, nbp - bp0);
    PencilArgs a = make_args(c, nb, bp0);
    a.rows = rows;
    hipLaunchKernelGGL(k_zinv<T>, dim3(gz.x, nbb), dim3(kZThreads), lz, c->stream, a, psi, t1);
    if (sizeof(T) == 8)
      hipLaunchKernelGGL((k_plane<T, double>), dim3(c->n3, nbb), dim3(kPlaneThreads), lp, c->stream,
                         a, c->d_V, t1, t2);
    else
      hipLaunchKernelGGL((k_plane<T, float>), dim3(c->n3, nbb), dim3(kPlaneThreads), lp, c->stream,
                         a, c->d_Vf, t1, t2);
    hipLaunchKernelGGL(k_zfwd<T>, dim3(gz.x, nbb), dim3(kZThreads), lz, c->stream, a, psi, t2, out,
                       accumulate, 1.0 / (double)c->N);
  }
  QX_CHECK(hipGetLastError());
  return 0;
}

int local_apply(Ctx* c, const double* psi, double* out, int nb, int prec, int accumulate) {
  if (fast_ok(c)) {
    const int rc = local_apply_fast(c, psi, out, nb, prec, accumulate);
    if (rc != -3) return rc;  // -3: the fast plane kernel does not fit this grid
  }
  if (prec == PREC_FP32) return local_apply_t<float>(c, psi, out, nb, accumulate);
  return local_apply_t<double>(c, psi, out, nb, accumulate);
}

template <typename T>
static int density_t(Ctx* c, const double* psi, int nb, const double* w) {
  const int nbp = (nb + 1) / 2;
  const int rows = plane_rows<T>(c);
  if (rows < 1) return -3;
  const size_t per_bp = sizeof(cplx<T>) * (size_t)c->n3 * c->pc.ncol;
  const int batch = (int)std::max<size_t>(1, std::min<size_t>(c->t_bytes / per_bp, (size_t)nbp));
  const dim3 gz((c->pc.npair + kPairsPerBlock - 1) / kPairsPerBlock, 1);
  const size_t lz = zlines_lds<T>(c), lp = plane_lds<T>(c, rows);
  cplx<T>* t1 = reinterpret_cast<cplx<T>*>(c->d_t1);
  lds_optin(k_zinv<T>, lz);
  lds_optin(k_plane_rho<T>, lp);
  int nparts = 0;
  const int nbatch = (nbp + batch - 1) / batch;
  const int groups_per_batch = std::max(1, c->rho_groups / nbatch);
  if (groups_per_batch * nbatch > c->rho_groups) return -4;
  for (int bp0 = 0; bp0 < nbp; bp0 += batch) {
    const int nbb = std::min(batch, nbp - bp0);
    PencilArgs a = make_args(c, nb, bp0);
    a.rows = rows;
    hipLaunchKernelGGL(k_zinv<T>, dim3(gz.x, nbb), dim3(kZThreads), lz, c->stream, a, psi, t1);
    int ng = std::min(groups_per_batch, nbb);
    int per = (nbb + ng - 1) / ng;
    ng = (nbb + per - 1) / per;
    hipLaunchKernelGGL(k_plane_rho<T>, dim3(c->n3, ng), dim3(kPlaneThreads), lp, c->stream, a, t1,
                       nbb, per, w, c->d_rho_part + (size_t)nparts * c->N);
    nparts += ng;
  }
  hipLaunchKernelGGL(k_rho_reduce, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream,
                     c->d_rho_part, nparts, c->N, c->d_rho);
  QX_CHECK(hipGetLastError());
  return 0;
}

int density(Ctx* c, const double* psi, int nb, const double* w, double, int prec) {
  if (fast_ok(c)) {
    const int rc = density_fast(c, psi, nb, w, prec);
    if (rc != -3) return rc;  // -3: the fast plane kernel does not fit this grid
  }
  if (prec == PREC_FP32) return density_t<float>(c, psi, nb, w);
  return density_t<double>(c, psi, nb, w);
}

}  // namespace qx
