// ewald.hip -- reciprocal-space Ewald forces and energies (frcrec without
// stress, Sources/bulk.f90:1880-2026, the form the reference compiles with
// VECTOR).  Per Ewald G the ion sum S(G) = sum_k zv (cos, sin)(G.tau_k) is
// formed by one thread in the reference's ion order (so S is the reference's
// value for the same tables); a second pass over (G chunk, ion group) blocks
// accumulates the forces 8/V hep G (Re S sin_k - Im S cos_k) zv and the
// per-atom energies, with per-block partials summed in a fixed order.  The
// phases come from per-axis tables c_q + i s_q: the caller's per-ion tables
// of stscfn (Sources/bulk.f90:2086), or the per-atom tables the device
// stscfn_pw left in the context.
#include <algorithm>
#include <cmath>
#include <vector>
#include "context.h"

namespace qx {

namespace {

constexpr int kEwThreads = 256;
constexpr int kEwIons = 4;      // ions per block of the force pass
constexpr int kEwChunk = 1024;  // G per block of the force pass

inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

struct EwIon {
  int ia, row;  // atom (output slot) and table row
  double zv;
};

struct Tab {
  const double* t;  // cosines, then sines after 3 nrow ld
  int ld, nrow, km1, km2, km3;
};

__device__ inline void ew_phase(const Tab& tb, int row, int k1, int k2, int k3, double& yc,
                                double& ys) {
#pragma clang fp contract(off)
  const double* dc = tb.t;
  const double* ds = tb.t + 3 * (size_t)tb.nrow * tb.ld;
  const size_t r = (size_t)(3 * row) * tb.ld;
  const double c1 = dc[r + tb.km1 + k1], s1 = ds[r + tb.km1 + k1];
  const double c2 = dc[r + tb.ld + tb.km2 + k2], s2 = ds[r + tb.ld + tb.km2 + k2];
  const double c3 = dc[r + 2 * tb.ld + tb.km3 + k3], s3 = ds[r + 2 * tb.ld + tb.km3 + k3];
  const double b1 = c1 * c2 - s1 * s2;
  const double b2 = s1 * c2 + c1 * s2;
  yc = b1 * c3 - b2 * s3;
  ys = b2 * c3 + b1 * s3;
}

// S(G) and hep |S|^2 per G, ions summed in order
__global__ void k_ewald_sf(const EwIon* __restrict__ ions, int nion, Tab tb,
                           const int* __restrict__ g, int numk2, const double* __restrict__ hep,
                           double* __restrict__ sf, double* __restrict__ eterm) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= numk2) return;
  const int k1 = g[m], k2 = g[numk2 + m], k3 = g[2 * numk2 + m];
  double sc = 0.0, ss = 0.0;
  for (int j = 0; j < nion; ++j) {
    double yc, ys;
    ew_phase(tb, ions[j].row, k1, k2, k3, yc, ys);
    sc = sc + ions[j].zv * yc;
    ss = ss + ions[j].zv * ys;
  }
  sf[m] = sc;
  sf[numk2 + m] = ss;
  eterm[m] = hep[m] * (sc * sc + ss * ss);
}

// part[((chunk * nion) + j) * 4 + {fx, fy, fz, e}]
__global__ void __launch_bounds__(kEwThreads)
    k_ewald_force(const EwIon* __restrict__ ions, int nion, Tab tb, const int* __restrict__ g,
                  const double* __restrict__ gv, int numk2, const double* __restrict__ hep,
                  const double* __restrict__ sf, double vol8, double cst,
                  double* __restrict__ part) {
#pragma clang fp contract(off)
  const int chunk = blockIdx.x;
  const int j0 = blockIdx.y * kEwIons;
  double s[kEwIons][4];
#pragma unroll
  for (int q = 0; q < kEwIons; ++q) s[q][0] = s[q][1] = s[q][2] = s[q][3] = 0.0;
  const int m_end = min(numk2, (chunk + 1) * kEwChunk);
  for (int m = chunk * kEwChunk + threadIdx.x; m < m_end; m += kEwThreads) {
    const int k1 = g[m], k2 = g[numk2 + m], k3 = g[2 * numk2 + m];
    const double dckm = sf[m], dskm = sf[numk2 + m], h = hep[m];
    const double hep8 = h * vol8, hepc = h * cst;
    const double hc8 = hep8 * dckm, hs8 = hep8 * dskm, hcc = hepc * dckm, hsc = hepc * dskm;
    const double gx = gv[m], gy = gv[numk2 + m], gz = gv[2 * numk2 + m];
#pragma unroll
    for (int q = 0; q < kEwIons; ++q) {
      const int j = j0 + q;
      if (j >= nion) break;
      double yc, ys;
      ew_phase(tb, ions[j].row, k1, k2, k3, yc, ys);
      yc = ions[j].zv * yc;
      ys = ions[j].zv * ys;
      s[q][0] += hc8 * gx * ys - hs8 * gx * yc;
      s[q][1] += hc8 * gy * ys - hs8 * gy * yc;
      s[q][2] += hc8 * gz * ys - hs8 * gz * yc;
      s[q][3] += yc * hcc + ys * hsc;
    }
  }
  __shared__ double red[kEwIons * 4][kEwThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kEwIons; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double v = s[q][c];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) red[q * 4 + c][wv] = v;
    }
  __syncthreads();
  if (threadIdx.x < kEwIons * 4) {
    const int q = threadIdx.x / 4, c = threadIdx.x % 4;
    if (j0 + q < nion) {
      double v = 0.0;
      for (int w = 0; w < kEwThreads / 64; ++w) v += red[threadIdx.x][w];
      part[((size_t)chunk * nion + j0 + q) * 4 + c] = v;
    }
  }
}

// out[i] = sum over chunks in chunk order (i < 4 nion); out[4 nion] = sum of
// the per-G energy terms in a fixed tree order
__global__ void __launch_bounds__(kEwThreads)
    k_ewald_sum(const double* __restrict__ part, int nchunk, int nion,
                const double* __restrict__ eterm, int numk2, double* __restrict__ out) {
  if (blockIdx.x < gridDim.x - 1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * nion) return;
    double v = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) v += part[(size_t)ch * 4 * nion + i];
    out[i] = v;
    return;
  }
  __shared__ double red[kEwThreads / 64];
  double v = 0.0;
  for (int m = threadIdx.x; m < numk2; m += kEwThreads) v += eterm[m];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kEwThreads / 64; ++w) t += red[w];
    out[4 * nion] = t;
  }
}

int ewald_run(Ctx* c, const std::vector<EwIon>& ions, const Tab& tb, std::vector<double>& acc,
              double& recolmb) {
  auto& E = c->ew;
  const int nion = (int)ions.size(), numk2 = E.numk2;
  acc.assign(4 * (size_t)nion, 0.0);
  recolmb = 0.0;
  if (nion == 0 || numk2 == 0) return 0;
  const int nchunk = (int)cdivu(numk2, kEwChunk);
  EwIon* d_ions = nullptr;
  double *d_sf = nullptr, *d_et = nullptr, *d_part = nullptr, *d_out = nullptr;
  QX_CHECK(hipMalloc(&d_ions, sizeof(EwIon) * nion));
  QX_CHECK(hipMalloc(&d_sf, sizeof(double) * 2 * numk2));
  QX_CHECK(hipMalloc(&d_et, sizeof(double) * numk2));
  QX_CHECK(hipMalloc(&d_part, sizeof(double) * 4 * (size_t)nion * nchunk));
  QX_CHECK(hipMalloc(&d_out, sizeof(double) * (4 * (size_t)nion + 1)));
  QX_CHECK(hipMemcpyAsync(d_ions, ions.data(), sizeof(EwIon) * nion, hipMemcpyHostToDevice,
                          c->stream));
  const double pi = std::acos(-1.0);
  const double vol8 = 8.0 / E.volume, cst = 2.0 / (E.volume * pi);
  hipLaunchKernelGGL(k_ewald_sf, dim3(cdivu(numk2, 256)), dim3(256), 0, c->stream, d_ions, nion, tb,
                     E.d_g, numk2, E.d_hep, d_sf, d_et);
  hipLaunchKernelGGL(k_ewald_force, dim3(nchunk, cdivu(nion, kEwIons)), dim3(kEwThreads), 0,
                     c->stream, d_ions, nion, tb, E.d_g, E.d_gv, numk2, E.d_hep, d_sf, vol8, cst,
                     d_part);
  hipLaunchKernelGGL(k_ewald_sum, dim3(cdivu(4L * nion, kEwThreads) + 1), dim3(kEwThreads), 0,
                     c->stream, d_part, nchunk, nion, d_et, numk2, d_out);
  QX_CHECK(hipGetLastError());
  std::vector<double> out(4 * (size_t)nion + 1);
  QX_CHECK(hipMemcpyAsync(out.data(), d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost,
                          c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  std::copy(out.begin(), out.begin() + 4 * nion, acc.begin());
  recolmb = out[4 * nion];
  (void)hipFree(d_ions);
  (void)hipFree(d_sf);
  (void)hipFree(d_et);
  (void)hipFree(d_part);
  (void)hipFree(d_out);
  return 0;
}

// the reference's closing terms (1995-2024) and the adds into the outputs
void ewald_finish(Ctx* c, int ntype, const int* nhk1, const int* nhk2, const double* zv, int nel,
                  const std::vector<EwIon>& ions, const std::vector<int>& ityp,
                  const std::vector<double>& acc, double recolmb, double* fclm, double* ecolmb,
                  double* atom_ecolmb) {
  auto& E = c->ew;
  const double pi = std::acos(-1.0);
  const double cst = 2.0 / (E.volume * pi);
  for (size_t j = 0; j < ions.size(); ++j) {
    const int ia = ions[j].ia;
    fclm[3 * ia] = fclm[3 * ia] + acc[4 * j];
    fclm[3 * ia + 1] = fclm[3 * ia + 1] + acc[4 * j + 1];
    fclm[3 * ia + 2] = fclm[3 * ia + 2] + acc[4 * j + 2];
    atom_ecolmb[ia] = atom_ecolmb[ia] + acc[4 * j + 3];
  }
  *ecolmb = *ecolmb + cst * recolmb;
  double ztot2 = 0.0;
  for (int it = 0; it < ntype; ++it)
    ztot2 = ztot2 + (double)(nhk2[it] - nhk1[it] + 1) * zv[it] * zv[it];
  const double g2 = E.gamma * E.gamma;
  *ecolmb = *ecolmb - (double)(nel * nel) / E.volume * (pi / g2 - E.recnrmex0) -
            2.0 * E.gamma * ztot2 / std::sqrt(pi);
  for (size_t j = 0; j < ions.size(); ++j) {
    const int ia = ions[j].ia;
    const double z = zv[ityp[j]];
    atom_ecolmb[ia] = atom_ecolmb[ia] - z * (double)nel / E.volume * (pi / g2 - E.recnrmex0) -
                      2.0 * E.gamma * z * z / std::sqrt(pi);
  }
}

int ewald_ions(int ntype, const double* zv, int natom, int nion, const int* iatoit,
               const int* iatmpt, bool per_ion_rows, std::vector<EwIon>& ions,
               std::vector<int>& ityp) {
  ions.resize(nion);
  ityp.resize(nion);
  for (int k = 0; k < nion; ++k) {
    const int ia = iatmpt[k] - 1;
    if (ia < 0 || ia >= natom) return -2;
    const int it = iatoit[ia] - 1;
    if (it < 0 || it >= ntype) return -2;
    ions[k] = {ia, per_ion_rows ? k : ia, zv[it]};
    ityp[k] = it;
  }
  return 0;
}

}  // namespace

void ewald_destroy(Ctx* c) {
  auto& E = c->ew;
  if (E.d_g) (void)hipFree(E.d_g);
  if (E.d_gv) (void)hipFree(E.d_gv);
  if (E.d_hep) (void)hipFree(E.d_hep);
  E = Ctx::Ewald{};
}

// clmrcp's G list (Sources/bulk.f90:556-1068) and the parameters of para_ewald
int ewald_set(Ctx* c, int numk2, const int* nga, const int* ngb, const int* ngc, const double* gx,
              const double* gy, const double* gz, const double* recnrm, const double* recnrmex,
              double gamma, double volume) {
  if (numk2 < 0 || !(gamma > 0.0) || !(volume > 0.0)) return -2;
  ewald_destroy(c);
  auto& E = c->ew;
  const size_t n = (size_t)std::max(numk2, 1);
  std::vector<int> g(3 * n, 0);
  std::vector<double> gv(3 * n, 0.0), hep(n, 0.0);
  const double pi = std::acos(-1.0);
  const double const2 = -((pi / gamma) * (pi / gamma));
  E.km[0] = E.km[1] = E.km[2] = 0;
  for (int m = 0; m < numk2; ++m) {
    g[m] = nga[m];
    g[numk2 + m] = ngb[m];
    g[2 * numk2 + m] = ngc[m];
    E.km[0] = std::max(E.km[0], std::abs(nga[m]));
    E.km[1] = std::max(E.km[1], std::abs(ngb[m]));
    E.km[2] = std::max(E.km[2], std::abs(ngc[m]));
    gv[m] = gx[m];
    gv[numk2 + m] = gy[m];
    gv[2 * numk2 + m] = gz[m];
    // hep of 1970-1971, host libm exp as the reference evaluates it
    const double recr = 1.0 / recnrm[m];
    hep[m] = (std::exp(const2 * recnrm[m]) - recnrmex[m + 1]) * recr;
  }
  QX_CHECK(hipMalloc(&E.d_g, sizeof(int) * g.size()));
  QX_CHECK(hipMalloc(&E.d_gv, sizeof(double) * gv.size()));
  QX_CHECK(hipMalloc(&E.d_hep, sizeof(double) * hep.size()));
  QX_CHECK(hipMemcpy(E.d_g, g.data(), sizeof(int) * g.size(), hipMemcpyHostToDevice));
  QX_CHECK(hipMemcpy(E.d_gv, gv.data(), sizeof(double) * gv.size(), hipMemcpyHostToDevice));
  QX_CHECK(hipMemcpy(E.d_hep, hep.data(), sizeof(double) * hep.size(), hipMemcpyHostToDevice));
  E.numk2 = numk2;
  E.gamma = gamma;
  E.volume = volume;
  E.recnrmex0 = recnrmex[0];
  return 0;
}

// host tables per ion k (the reference's DCQ*(-kcsma:kcsma, nion))
int ewald_frcrec_host(Ctx* c, double* fclm, int ntype, const int* nhk1, const int* nhk2,
                      const double* zv, int nel, double* ecolmb, double* atom_ecolmb,
                      const int* iatoit, const int* iatmpt, int natom, int nion, const int* kcsma,
                      const double* const* dcq, const double* const* dsq) {
  auto& E = c->ew;
  if (E.numk2 < 0 || !E.d_g) return -4;
  for (int a = 0; a < 3; ++a)
    if (kcsma[a] < E.km[a]) return -2;
  std::vector<EwIon> ions;
  std::vector<int> ityp;
  int rc = ewald_ions(ntype, zv, natom, nion, iatoit, iatmpt, true, ions, ityp);
  if (rc) return rc;
  const int ld = 2 * std::max(kcsma[0], std::max(kcsma[1], kcsma[2])) + 1;
  std::vector<double> tab(2 * 3 * (size_t)std::max(nion, 1) * ld, 0.0);
  for (int k = 0; k < nion; ++k)
    for (int a = 0; a < 3; ++a) {
      const int w = 2 * kcsma[a] + 1;
      for (int q = 0; q < w; ++q) {
        tab[(size_t)(3 * k + a) * ld + q] = dcq[a][(size_t)k * w + q];
        tab[3 * (size_t)nion * ld + (size_t)(3 * k + a) * ld + q] = dsq[a][(size_t)k * w + q];
      }
    }
  double* d_tab = nullptr;
  QX_CHECK(hipMalloc(&d_tab, sizeof(double) * tab.size()));
  QX_CHECK(hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice,
                          c->stream));
  const Tab tb{d_tab, ld, nion, kcsma[0], kcsma[1], kcsma[2]};
  std::vector<double> acc;
  double recolmb = 0.0;
  rc = ewald_run(c, ions, tb, acc, recolmb);
  (void)hipFree(d_tab);
  if (rc) return rc;
  ewald_finish(c, ntype, nhk1, nhk2, zv, nel, ions, ityp, acc, recolmb, fclm, ecolmb, atom_ecolmb);
  return 0;
}

// device-resident: the per-atom axis tables of the last stscfn_pw
int ewald_frcrec_dev(Ctx* c, double* fclm, int ntype, const int* nhk1, const int* nhk2,
                     const double* zv, int nel, double* ecolmb, double* atom_ecolmb,
                     const int* iatoit, const int* iatmpt, int natom, int nion) {
  auto& E = c->ew;
  if (E.numk2 < 0 || !E.d_g) return -4;
  if (!c->sf.d_tab || c->sf.natom != natom) return -5;
  for (int a = 0; a < 3; ++a)
    if (c->sf.km[a] < E.km[a]) return -5;
  std::vector<EwIon> ions;
  std::vector<int> ityp;
  int rc = ewald_ions(ntype, zv, natom, nion, iatoit, iatmpt, false, ions, ityp);
  if (rc) return rc;
  const Tab tb{c->sf.d_tab, c->sf.ld, c->sf.natom, c->sf.km[0], c->sf.km[1], c->sf.km[2]};
  std::vector<double> acc;
  double recolmb = 0.0;
  if ((rc = ewald_run(c, ions, tb, acc, recolmb))) return rc;
  ewald_finish(c, ntype, nhk1, nhk2, zv, nel, ions, ityp, acc, recolmb, fclm, ecolmb, atom_ecolmb);
  return 0;
}

}  // namespace qx
