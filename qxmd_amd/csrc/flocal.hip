// flocal.hip -- Hellmann-Feynman forces of the local pseudopotential in G
// space (flocal_g without stress, Sources/nlkbpp.f90:1037-1125):
//   floc(:, ia) = 2 (floc(:, ia) + sum_{G != 0 of the density sphere}
//                 svco(G, type) (Re rho_G sin(G.tau) + Im rho_G cos(G.tau)) G)
// The phase exp(i G.tau) of every (G, atom) pair is rebuilt from the per-axis
// tables c_k + i s_k of stscfn_pw with the reference's three complex products,
// so no natom x nplw5 structure-factor array is stored.  One block takes a
// chunk of G for a group of atoms: the density, svco and G of the chunk are
// read once per group, the axis tables come from L1/L2.  Per-block partial
// sums and a fixed-order final sum keep the result deterministic.
#include <algorithm>
#include <vector>
#include "context.h"

namespace qx {

namespace {

constexpr int kFlThreads = 256;
constexpr int kFlAtoms = 4;      // atoms per block
constexpr int kFlChunk = 4096;   // G per block

inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

struct Ion {
  int ia, it;
};

// part[((chunk * nion) + j) * 3 + c]: sum over the chunk's G of ion j
__global__ void __launch_bounds__(kFlThreads)
    k_flocal(const Ion* __restrict__ ions, int nion, const double* __restrict__ tab, int ld,
             int km1, int km2, int km3, int natom_tab, const int* __restrict__ nga,
             const int* __restrict__ ngb, const int* __restrict__ ngc, int nplw5,
             const double* __restrict__ rhgr, const double* __restrict__ svco,
             const double* __restrict__ g3, double* __restrict__ part) {
#pragma clang fp contract(off)
  const int chunk = blockIdx.x;
  const int j0 = blockIdx.y * kFlAtoms;
  const double* dc = tab;
  const double* ds = tab + 3 * (size_t)natom_tab * ld;
  double s[kFlAtoms][3];
#pragma unroll
  for (int q = 0; q < kFlAtoms; ++q) s[q][0] = s[q][1] = s[q][2] = 0.0;
  const int ig_end = min(nplw5, (chunk + 1) * kFlChunk);
  for (int ig = 1 + chunk * kFlChunk + threadIdx.x; ig <= ig_end; ig += kFlThreads) {
    const int k1 = nga[ig], k2 = ngb[ig], k3 = ngc[ig];
    const double a = rhgr[2 * ig], b = rhgr[2 * ig + 1];
    const double gx = g3[ig], gy = g3[nplw5 + 1 + ig], gz = g3[2 * (nplw5 + 1) + ig];
#pragma unroll
    for (int q = 0; q < kFlAtoms; ++q) {
      const int j = j0 + q;
      if (j >= nion) break;
      const Ion io = ions[j];
      const size_t r = (size_t)(3 * io.ia) * ld;
      const double c1 = dc[r + km1 + k1], s1 = ds[r + km1 + k1];
      const double c2 = dc[r + ld + km2 + k2], s2 = ds[r + ld + km2 + k2];
      const double c3 = dc[r + 2 * ld + km3 + k3], s3 = ds[r + 2 * ld + km3 + k3];
      const double dcc12 = c1 * c2 - s1 * s2;
      const double dsc12 = s1 * c2 + c1 * s2;
      const double zcosk = dcc12 * c3 - dsc12 * s3;
      const double zsink = dsc12 * c3 + dcc12 * s3;
      const double ftmp = svco[(size_t)io.it * (nplw5 + 1) + ig] * (a * zsink + b * zcosk);
      s[q][0] += ftmp * gx;
      s[q][1] += ftmp * gy;
      s[q][2] += ftmp * gz;
    }
  }
  __shared__ double red[kFlAtoms * 3][kFlThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kFlAtoms; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = s[q][c];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) red[q * 3 + c][wv] = v;
    }
  __syncthreads();
  if (threadIdx.x < kFlAtoms * 3) {
    const int q = threadIdx.x / 3, c = threadIdx.x % 3;
    if (j0 + q < nion) {
      double v = 0.0;
      for (int w = 0; w < kFlThreads / 64; ++w) v += red[threadIdx.x][w];
      part[((size_t)chunk * nion + j0 + q) * 3 + c] = v;
    }
  }
}

// vlocal_g, reciprocal-space part (Sources/nlkbpp.f90:928-972), one thread per
// G of the density sphere: the ions of each type in the reference's order
// (ions are grouped by type) give rhmur = sum (cos, sin)(G.tau), then
// thrhgr += svco / rvol (Re, -Im) rhmur, type by type.  zin = thrhgr with the
// G = 0 grid value v0 (912-924) in its G = 0 slot.
__global__ void __launch_bounds__(kFlThreads)
    k_vlocal(const Ion* __restrict__ ions, int nion, const double* __restrict__ tab, int ld,
             int km1, int km2, int km3, int natom_tab, const int* __restrict__ nga,
             const int* __restrict__ ngb, const int* __restrict__ ngc, int nplw5,
             const double* __restrict__ svco, double rvolr, double v0, double* __restrict__ thr,
             double* __restrict__ zin) {
#pragma clang fp contract(off)
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  if (ig > nplw5) return;
  if (ig == 0) {
    thr[0] = thr[1] = 0.0;
    zin[0] = v0;
    zin[1] = 0.0;
    return;
  }
  const double* dc = tab;
  const double* ds = tab + 3 * (size_t)natom_tab * ld;
  const int k1 = nga[ig], k2 = ngb[ig], k3 = ngc[ig];
  double tr = 0.0, ti = 0.0;
  int j = 0;
  while (j < nion) {
    const int it = ions[j].it;
    double rc = 0.0, rs = 0.0;
    for (; j < nion && ions[j].it == it; ++j) {
      const size_t r = (size_t)(3 * ions[j].ia) * ld;
      const double c1 = dc[r + km1 + k1], s1 = ds[r + km1 + k1];
      const double c2 = dc[r + ld + km2 + k2], s2 = ds[r + ld + km2 + k2];
      const double c3 = dc[r + 2 * ld + km3 + k3], s3 = ds[r + 2 * ld + km3 + k3];
      const double dcc12 = c1 * c2 - s1 * s2;
      const double dsc12 = s1 * c2 + c1 * s2;
      rc = rc + (dcc12 * c3 - dsc12 * s3);
      rs = rs + (dsc12 * c3 + dcc12 * s3);
    }
    const double dggg = svco[(size_t)it * (nplw5 + 1) + ig] * rvolr;
    const double phar = rc * dggg;
    const double phai = -rs * dggg;
    tr = tr + phar;
    ti = ti + phai;
  }
  thr[2 * ig] = zin[2 * ig] = tr;
  thr[2 * ig + 1] = zin[2 * ig + 1] = ti;
}

// out[j * 3 + c] = sum over chunks, in chunk order
__global__ void k_flocal_sum(const double* __restrict__ part, int nchunk, int nion,
                             double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * nion) return;
  double v = 0.0;
  for (int ch = 0; ch < nchunk; ++ch) v += part[(size_t)ch * 3 * nion + i];
  out[i] = v;
}

// the ions of the types with llking false, in the reference's order
int ion_list(int ntype, const int* nhk1, const int* nhk2, const int* iatmpt, const int* llking,
             int natom, std::vector<Ion>& ions) {
  ions.clear();
  for (int it = 0; it < ntype; ++it) {
    if (llking[it]) continue;
    if (nhk1[it] < 1 || nhk2[it] > natom) return -2;
    for (int k = nhk1[it]; k <= nhk2[it]; ++k) {
      const int ia = iatmpt[k - 1] - 1;
      if (ia < 0 || ia >= natom) return -2;
      ions.push_back({ia, it});
    }
  }
  return 0;
}

// sums of the ions for a density on the device; tables, G list and svco given
int flocal_sums(Ctx* c, const std::vector<Ion>& ions, const double* tab, int ld, const int* km,
                int natom_tab, const int* d_g, int ngmax, const double* d_rhgr, int nplw5,
                std::vector<double>& sums) {
  const int nion = (int)ions.size();
  sums.assign(3 * (size_t)nion, 0.0);
  if (nion == 0 || nplw5 < 1) return 0;
  const int nchunk = (int)cdivu(nplw5, kFlChunk);
  Ion* d_ions = nullptr;
  double *d_part = nullptr, *d_out = nullptr;
  QX_CHECK(hipMalloc(&d_ions, sizeof(Ion) * nion));
  QX_CHECK(hipMalloc(&d_part, sizeof(double) * 3 * (size_t)nion * nchunk));
  QX_CHECK(hipMalloc(&d_out, sizeof(double) * 3 * (size_t)nion));
  QX_CHECK(hipMemcpyAsync(d_ions, ions.data(), sizeof(Ion) * nion, hipMemcpyHostToDevice,
                          c->stream));
  hipLaunchKernelGGL(k_flocal, dim3(nchunk, cdivu(nion, kFlAtoms)), dim3(kFlThreads), 0,
                     c->stream, d_ions, nion, tab, ld, km[0], km[1], km[2], natom_tab, d_g,
                     d_g + ngmax + 1, d_g + 2 * (ngmax + 1), nplw5, d_rhgr, c->lopp.d_svco,
                     c->lopp.d_g, d_part);
  hipLaunchKernelGGL(k_flocal_sum, dim3(cdivu(3L * nion, 256)), dim3(256), 0, c->stream, d_part,
                     nchunk, nion, d_out);
  QX_CHECK(hipGetLastError());
  QX_CHECK(hipMemcpyAsync(sums.data(), d_out, sizeof(double) * 3 * nion, hipMemcpyDeviceToHost,
                          c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  (void)hipFree(d_ions);
  (void)hipFree(d_part);
  (void)hipFree(d_out);
  return 0;
}

// floc(:, ia) = 2 (floc(:, ia) + sum) per type, as 1103-1116
void flocal_finish(int ntype, const int* nhk1, const int* nhk2, const int* iatmpt,
                   const int* llking, const std::vector<double>& sums, double* floc) {
  size_t j = 0;
  for (int it = 0; it < ntype; ++it) {
    if (llking[it]) continue;
    for (int k = nhk1[it]; k <= nhk2[it]; ++k, ++j) {
      const int ia = iatmpt[k - 1] - 1;
      for (int c = 0; c < 3; ++c) floc[3 * ia + c] = floc[3 * ia + c] + sums[3 * j + c];
    }
    for (int k = nhk1[it]; k <= nhk2[it]; ++k) {
      const int ia = iatmpt[k - 1] - 1;
      for (int c = 0; c < 3; ++c) floc[3 * ia + c] = floc[3 * ia + c] * 2.0;
    }
  }
}

}  // namespace

void lopp_destroy(Ctx* c) {
  for (void* p : {(void*)c->lopp.d_svco, (void*)c->lopp.d_g, (void*)c->lopp.d_thr,
                  (void*)c->lopp.d_zin, c->lopp.d_ions})
    if (p) (void)hipFree(p);
  c->lopp = Ctx::Lopp{};
}

int vlocal_sphere(Ctx* c, int ntype, const int* nhk1, const int* nhk2, const int* iatmpt,
                  const int* llking, int natom, double rvol, int nplw5) {
  auto& L = c->lopp;
  if (!L.d_svco || L.ntype != ntype || L.nplw5 != nplw5) return -4;
  if (!c->sf.d_tab || c->sf.ngmax < nplw5 || c->sf.natom != natom) return -5;
  if (!(rvol > 0.0)) return -2;
  std::vector<Ion> ions;
  int rc = ion_list(ntype, nhk1, nhk2, iatmpt, llking, natom, ions);
  if (rc) return rc;
  const size_t n5x = 2 * ((size_t)nplw5 + 1);
  if (!L.d_thr) {
    QX_CHECK(hipMalloc(&L.d_thr, sizeof(double) * n5x));
    QX_CHECK(hipMalloc(&L.d_zin, sizeof(double) * n5x));
  }
  if (L.ion_cap < (int)ions.size() || !L.d_ions) {
    if (L.d_ions) (void)hipFree(L.d_ions);
    L.ion_cap = std::max<int>(1, (int)ions.size());
    QX_CHECK(hipMalloc(&L.d_ions, sizeof(Ion) * L.ion_cap));
  }
  if (!ions.empty())
    QX_CHECK(hipMemcpyAsync(L.d_ions, ions.data(), sizeof(Ion) * ions.size(),
                            hipMemcpyHostToDevice, c->stream));
  // G = 0 grid value, type by type (nhk = nhk2 - nhk1 + 1 atoms of the type)
  std::vector<double> svco0(ntype);
  for (int it = 0; it < ntype; ++it)
    QX_CHECK(hipMemcpyAsync(&svco0[it], L.d_svco + (size_t)it * (nplw5 + 1), sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  const double rvolr = 1.0 / rvol;
  double v0 = 0.0;
  for (int it = 0; it < ntype; ++it) {
    if (llking[it]) continue;
    const double dcsn = (double)(nhk2[it] - nhk1[it] + 1) * svco0[it];
    v0 = v0 + dcsn * rvolr;
  }
  const int* g = c->sf.d_g;
  const int ng1 = c->sf.ngmax + 1;
  hipLaunchKernelGGL(k_vlocal, dim3(cdivu((long)nplw5 + 1, kFlThreads)), dim3(kFlThreads), 0,
                     c->stream, static_cast<const Ion*>(L.d_ions), (int)ions.size(), c->sf.d_tab,
                     c->sf.ld, c->sf.km[0], c->sf.km[1], c->sf.km[2], c->sf.natom, g, g + ng1,
                     g + 2 * ng1, nplw5, L.d_svco, rvolr, v0, L.d_thr, L.d_zin);
  QX_CHECK(hipGetLastError());
  return 0;
}

int lopp_set(Ctx* c, int ntype, int nplw5, const double* svco, const double* gx,
             const double* gy, const double* gz) {
  if (ntype < 1 || nplw5 < 0) return -2;
  lopp_destroy(c);
  const size_t n = (size_t)nplw5 + 1;
  QX_CHECK(hipMalloc(&c->lopp.d_svco, sizeof(double) * n * ntype));
  QX_CHECK(hipMalloc(&c->lopp.d_g, sizeof(double) * 3 * n));
  QX_CHECK(hipMemcpy(c->lopp.d_svco, svco, sizeof(double) * n * ntype, hipMemcpyHostToDevice));
  QX_CHECK(hipMemcpy(c->lopp.d_g, gx, sizeof(double) * n, hipMemcpyHostToDevice));
  QX_CHECK(hipMemcpy(c->lopp.d_g + n, gy, sizeof(double) * n, hipMemcpyHostToDevice));
  QX_CHECK(hipMemcpy(c->lopp.d_g + 2 * n, gz, sizeof(double) * n, hipMemcpyHostToDevice));
  c->lopp.ntype = ntype;
  c->lopp.nplw5 = nplw5;
  return 0;
}

// device-resident: the density sphere d_rhgr, the axis tables and G list of
// the last stscfn_pw, svco of lopp_set
int flocal_dev(Ctx* c, int ntype, const int* nhk1, const int* nhk2, const int* iatmpt,
               const int* llking, int natom, const double* d_rhgr, int nplw5, double* floc) {
  if (!c->lopp.d_svco || c->lopp.ntype != ntype || c->lopp.nplw5 != nplw5) return -4;
  if (!c->sf.d_tab || c->sf.ngmax < nplw5 || c->sf.natom != natom) return -5;
  if (!d_rhgr) return -2;
  std::vector<Ion> ions;
  int rc = ion_list(ntype, nhk1, nhk2, iatmpt, llking, natom, ions);
  if (rc) return rc;
  std::vector<double> sums;
  if ((rc = flocal_sums(c, ions, c->sf.d_tab, c->sf.ld, c->sf.km, c->sf.natom, c->sf.d_g,
                        c->sf.ngmax, d_rhgr, nplw5, sums)))
    return rc;
  flocal_finish(ntype, nhk1, nhk2, iatmpt, llking, sums, floc);
  return 0;
}

// host arrays with the reference's argument list: the caller's axis tables
// DCQ/DSQ (2 kmax_a + 1 per atom), G list and density; svco of lopp_set
int flocal_host(Ctx* c, double* floc, int ntype, const int* nhk1, const int* nhk2,
                const int* iatmpt, int natom, const int* llking, int nplw5, const int* nga,
                const int* ngb, const int* ngc, const double* rhgr, const int* kmax,
                const double* const* dcq, const double* const* dsq) {
  if (!c->lopp.d_svco || c->lopp.ntype != ntype || c->lopp.nplw5 != nplw5) return -4;
  for (int ig = 0; ig <= nplw5; ++ig)
    if (std::abs(nga[ig]) > kmax[0] || std::abs(ngb[ig]) > kmax[1] || std::abs(ngc[ig]) > kmax[2])
      return -2;
  std::vector<Ion> ions;
  int rc = ion_list(ntype, nhk1, nhk2, iatmpt, llking, natom, ions);
  if (rc) return rc;
  const int ld = 2 * std::max(kmax[0], std::max(kmax[1], kmax[2])) + 1;
  std::vector<double> tab(2 * 3 * (size_t)natom * ld, 0.0);
  for (int ia = 0; ia < natom; ++ia)
    for (int a = 0; a < 3; ++a) {
      const int w = 2 * kmax[a] + 1;
      for (int k = 0; k < w; ++k) {
        tab[(size_t)(3 * ia + a) * ld + k] = dcq[a][(size_t)ia * w + k];
        tab[3 * (size_t)natom * ld + (size_t)(3 * ia + a) * ld + k] = dsq[a][(size_t)ia * w + k];
      }
    }
  std::vector<int> g(3 * ((size_t)nplw5 + 1));
  std::copy(nga, nga + nplw5 + 1, g.begin());
  std::copy(ngb, ngb + nplw5 + 1, g.begin() + nplw5 + 1);
  std::copy(ngc, ngc + nplw5 + 1, g.begin() + 2 * (nplw5 + 1));
  double *d_tab = nullptr, *d_rh = nullptr;
  int* d_g = nullptr;
  QX_CHECK(hipMalloc(&d_tab, sizeof(double) * tab.size()));
  QX_CHECK(hipMalloc(&d_rh, sizeof(double) * 2 * ((size_t)nplw5 + 1)));
  QX_CHECK(hipMalloc(&d_g, sizeof(int) * g.size()));
  QX_CHECK(hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice,
                          c->stream));
  QX_CHECK(hipMemcpyAsync(d_rh, rhgr, sizeof(double) * 2 * ((size_t)nplw5 + 1),
                          hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipMemcpyAsync(d_g, g.data(), sizeof(int) * g.size(), hipMemcpyHostToDevice,
                          c->stream));
  std::vector<double> sums;
  const int km[3] = {kmax[0], kmax[1], kmax[2]};
  rc = flocal_sums(c, ions, d_tab, ld, km, natom, d_g, nplw5, d_rh, nplw5, sums);
  (void)hipFree(d_tab);
  (void)hipFree(d_rh);
  (void)hipFree(d_g);
  if (rc) return rc;
  flocal_finish(ntype, nhk1, nhk2, iatmpt, llking, sums, floc);
  return 0;
}

}  // namespace qx
