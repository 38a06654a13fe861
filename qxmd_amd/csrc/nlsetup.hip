// nlsetup.hip -- the per-MD-step inputs of the KB contractions built on the
// device: structure factors (stscfn_pw, Sources/pwlda.f90:5711-5854) and the KB
// projector table eqlmr_ with vclalp (setnlc_g / setbgq / setqlm,
// Sources/nlkbpp.f90:1236-1279, 2135-2247), written straight into the
// context's internal G order so no table crosses PCIe when the atoms move.
//
// Both kernels evaluate the reference's expressions term for term with FP
// contraction off, so with the same first cos/sin (taken from the host libm,
// as the reference does) they reproduce the reference's values.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "context.h"

namespace qx {

namespace {

inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

// per atom and axis: c_k + i s_k = (c_1 + i s_1)^k by the recurrence of
// stscfn_pw (5806-5834), k = -kmax..kmax; table[(ia * 3 + a) * ld + kmax_a + k]
__global__ void k_stscfn_axis(const double* __restrict__ cs1, int natom, int km1, int km2,
                              int km3, int ld, double* __restrict__ dc, double* __restrict__ ds) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * natom) return;
  const int a = i % 3;
  const int km = a == 0 ? km1 : (a == 1 ? km2 : km3);
  double* c = dc + (size_t)i * ld + km;
  double* s = ds + (size_t)i * ld + km;
  const double c1 = cs1[2 * i], s1 = cs1[2 * i + 1];
  c[0] = 1.0;
  s[0] = 0.0;
  if (km >= 1) {
    c[1] = c1;
    s[1] = s1;
    c[-1] = c1;
    s[-1] = -s1;
  }
  for (int k = 2; k <= km; ++k) {
    const double ck = c1 * c[k - 1] - s1 * s[k - 1];
    const double sk = s1 * c[k - 1] + c1 * s[k - 1];
    c[k] = ck;
    s[k] = sk;
    c[-k] = ck;
    s[-k] = -sk;
  }
}

// ycos/ysin(ig, ia) for ig = 0..ngmax (5836-5848)
__global__ void k_stscfn(const double* __restrict__ dc, const double* __restrict__ ds, int ld,
                         int km1, int km2, int km3, const int* __restrict__ nga,
                         const int* __restrict__ ngb, const int* __restrict__ ngc, int ngmax,
                         int natom, int ldy, double* __restrict__ ycos, double* __restrict__ ysin) {
#pragma clang fp contract(off)
  const int ig = blockIdx.x * blockDim.x + threadIdx.x;
  const int ia = blockIdx.y;
  if (ig > ngmax || ia >= natom) return;
  double* yc = ycos + (size_t)ia * ldy;
  double* ys = ysin + (size_t)ia * ldy;
  if (ig == 0) {
    yc[0] = 1.0;
    ys[0] = 0.0;
    return;
  }
  const double* c1 = dc + (size_t)(3 * ia) * ld + km1;
  const double* s1 = ds + (size_t)(3 * ia) * ld + km1;
  const double* c2 = dc + (size_t)(3 * ia + 1) * ld + km2;
  const double* s2 = ds + (size_t)(3 * ia + 1) * ld + km2;
  const double* c3 = dc + (size_t)(3 * ia + 2) * ld + km3;
  const double* s3 = ds + (size_t)(3 * ia + 2) * ld + km3;
  const int k1 = nga[ig], k2 = ngb[ig], k3 = ngc[ig];
  const double c12 = c1[k1] * c2[k2] - s1[k1] * s2[k2];
  const double s12 = s1[k1] * c2[k2] + c1[k1] * s2[k2];
  yc[ig] = c12 * c3[k3] - s12 * s3[k3];
  ys[ig] = s12 * c3[k3] + c12 * s3[k3];
}

// One KB projector p = (atom ia, projector la of its type): the offset of its
// radial table row (L, type), of its angular row (qlmr or qlmi, lm), its parity
// (litoll) and the type's table step.
struct ProjSpec {
  int ia, tb, q, par;
  double dl;
};

// eqlmr_ of setbgq (2171-2183) with zzzr/zzzi of setqlm (2207-2244), written at
// the internal position of each G: B[p * nplwex + 2 ip + {0, 1}]
__global__ void k_setbgq(const ProjSpec* __restrict__ ps, int np, const int* __restrict__ perm,
                         const double* __restrict__ kin, const double* __restrict__ ycos,
                         const double* __restrict__ ysin, int ldy,
                         const double* __restrict__ tbqlm, const double* __restrict__ tbqlma,
                         const double* __restrict__ qlm, double* __restrict__ B, int nplwex) {
#pragma clang fp contract(off)
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (ip >= np) return;
  const ProjSpec s = ps[p];
  const int ig = perm[ip];
  const double g = sqrt(kin[ip]);  // abgvec = sqrt(recnrm), Sources/ppkb.f90:785-787
  int m = (int)(g / (2.0 * s.dl));
  m = 2 * m;
  const double d = 0.5 * (g / s.dl - (double)m);
  const double* tb = tbqlm + s.tb;
  const double* ta = tbqlma + s.tb;
  const double zz = d * ((d - 1.0) * ta[m] + tb[m + 2] - tb[m]) + tb[m];
  const double z = zz * qlm[s.q + ig];
  const double yc = ycos[(size_t)s.ia * ldy + ig], ys = ysin[(size_t)s.ia * ldy + ig];
  double* out = B + (size_t)p * nplwex + 2 * (size_t)ip;
  if (s.par) {
    out[0] = yc * z;
    out[1] = -ys * z;
  } else {
    out[0] = -ys * z;
    out[1] = -yc * z;
  }
}

template <typename T>
T* upload(const T* h, size_t n, hipStream_t st) {
  T* d = nullptr;
  if (hipMalloc(&d, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
  if (n && hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, st) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  return d;
}

void free_nl(Ctx* c) {
  if (c->nl.d_tbqlm) (void)hipFree(c->nl.d_tbqlm);
  if (c->nl.d_tbqlma) (void)hipFree(c->nl.d_tbqlma);
  if (c->nl.d_qlm) (void)hipFree(c->nl.d_qlm);
  c->nl.d_tbqlm = c->nl.d_tbqlma = c->nl.d_qlm = nullptr;
}

}  // namespace

void nl_destroy(Ctx* c) {
  free_nl(c);
  if (c->sf.d_tab) (void)hipFree(c->sf.d_tab);
  if (c->sf.d_g) (void)hipFree(c->sf.d_g);
  c->sf = Ctx::AxisTables{};
  if (c->d_ycos) (void)hipFree(c->d_ycos);
  if (c->d_ysin) (void)hipFree(c->d_ysin);
  c->d_ycos = c->d_ysin = nullptr;
}

int nl_set_tables(Ctx* c, int ntype, int lmmax, int lnref, int mxqlm, const int* lmx,
                  const double* clalp, const double* tbqlm, const double* tbqlma,
                  const double* dlqlm, const int* itola, const int* itolm, const int* itolmf,
                  const int* litoll, int nqlmmx, int lplh, const double* qlmr,
                  const double* qlmi) {
  if (ntype < 1 || lmmax < 1 || lnref < 1 || mxqlm < 2 || nqlmmx < 1 || lplh < c->nplw)
    return -1;
  for (int it = 0; it < ntype; ++it) {
    if (lmx[it] < 0 || lmx[it] > lmmax || !(dlqlm[it] > 0.0)) return -1;
    for (int l = 0; l < lmx[it]; ++l) {
      const int k = l + lmmax * it;
      if (itola[k] < 1 || itola[k] > lnref || itolm[k] < 1 || itolm[k] > nqlmmx) return -1;
    }
    // the quadratic interpolation reads rows m..m+2 with m = 2 int(|G| / (2 dl))
    const int mtop = 2 * (int)(std::sqrt(c->gmax2) / (2.0 * dlqlm[it])) + 2;
    if (mtop > mxqlm) return -2;
  }
  free_nl(c);
  auto& n = c->nl;
  n.ntype = ntype;
  n.lmmax = lmmax;
  n.lnref = lnref;
  n.mxqlm = mxqlm;
  n.nqlmmx = nqlmmx;
  n.lplh = lplh;
  n.lmx.assign(lmx, lmx + ntype);
  n.clalp.assign(clalp, clalp + (size_t)lmmax * ntype);
  n.dlqlm.assign(dlqlm, dlqlm + ntype);
  const size_t nl2 = (size_t)lmmax * ntype;
  n.itola.assign(itola, itola + nl2);
  n.itolm.assign(itolm, itolm + nl2);
  n.itolmf.assign(itolmf, itolmf + nl2);
  n.litoll.assign(litoll, litoll + nl2);
  const size_t ntb = (size_t)(mxqlm + 1) * lnref * ntype;
  n.d_tbqlm = upload(tbqlm, ntb, c->stream);
  n.d_tbqlma = upload(tbqlma, ntb, c->stream);
  const size_t nq = (size_t)(lplh + 1) * nqlmmx;
  QX_CHECK(hipMalloc(&n.d_qlm, sizeof(double) * 2 * nq));
  QX_CHECK(hipMemcpyAsync(n.d_qlm, qlmr, sizeof(double) * nq, hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipMemcpyAsync(n.d_qlm + nq, qlmi, sizeof(double) * nq, hipMemcpyHostToDevice,
                          c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  if (!n.d_tbqlm || !n.d_tbqlma) return -3;
  return 0;
}

int nl_stscfn(Ctx* c, const double* ratm, int natom, const double* hcell, const int* nd1v,
              const int* nd1vks, int lclust, int kmax1, int kmax2, int kmax3, int nplwcs,
              int nplw5, const int* nga, const int* ngb, const int* ngc, double* ycos_out,
              double* ysin_out) {
  if (natom < 1 || kmax1 < 1 || kmax2 < 1 || kmax3 < 1 || nplwcs < 0 || nplw5 < 0) return -1;
  const int ngmax = std::min(nplwcs, nplw5);
  for (int ig = 0; ig <= ngmax; ++ig)
    if (std::abs(nga[ig]) > kmax1 || std::abs(ngb[ig]) > kmax2 || std::abs(ngc[ig]) > kmax3)
      return -2;
  // first cos/sin per atom and axis from the host libm, as the reference
  // evaluates them (5750-5792)
  const double dpi = 3.141592653589793 * 2.0;
  double fact[3];
  for (int a = 0; a < 3; ++a) {
    fact[a] = lclust ? 1.0 / hcell[4 * a] * (double)nd1v[a] / (double)nd1vks[a]
                     : (double)nd1v[a] / (double)nd1vks[a];
    fact[a] = fact[a] * dpi;
  }
  std::vector<double> cs1(6 * (size_t)natom);
  for (int ia = 0; ia < natom; ++ia)
    for (int a = 0; a < 3; ++a) {
      const double ph = ratm[a + 3 * ia] * fact[a];
      cs1[2 * (3 * ia + a)] = std::cos(ph);
      cs1[2 * (3 * ia + a) + 1] = std::sin(ph);
    }
  const int ld = 2 * std::max(kmax1, std::max(kmax2, kmax3)) + 1;
  const size_t ny = (size_t)(nplwcs + 1) * natom;
  if (c->sf_natom != natom || c->sf_nplwcs != nplwcs) {
    if (c->d_ycos) (void)hipFree(c->d_ycos);
    if (c->d_ysin) (void)hipFree(c->d_ysin);
    c->d_ycos = c->d_ysin = nullptr;
    QX_CHECK(hipMalloc(&c->d_ycos, sizeof(double) * ny));
    QX_CHECK(hipMalloc(&c->d_ysin, sizeof(double) * ny));
    c->sf_natom = natom;
    c->sf_nplwcs = nplwcs;
  }
  QX_CHECK(hipMemsetAsync(c->d_ycos, 0, sizeof(double) * ny, c->stream));
  QX_CHECK(hipMemsetAsync(c->d_ysin, 0, sizeof(double) * ny, c->stream));
  double* d_cs1 = upload(cs1.data(), cs1.size(), c->stream);
  // the G list kept with the axis tables covers the whole density sphere
  // when kmax does (what flocal_g needs), else the structure-factor range
  int nkeep = nplw5;
  for (int ig = ngmax + 1; ig <= nplw5 && nkeep == nplw5; ++ig)
    if (std::abs(nga[ig]) > kmax1 || std::abs(ngb[ig]) > kmax2 || std::abs(ngc[ig]) > kmax3)
      nkeep = ngmax;
  const size_t ng1 = (size_t)nkeep + 1;
  int* d_g = nullptr;
  double* d_tab = nullptr;
  QX_CHECK(hipMalloc(&d_g, sizeof(int) * 3 * ng1));
  QX_CHECK(hipMalloc(&d_tab, sizeof(double) * 2 * 3 * (size_t)natom * ld));
  if (!d_cs1) return -3;
  QX_CHECK(hipMemcpyAsync(d_g, nga, sizeof(int) * ng1, hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipMemcpyAsync(d_g + ng1, ngb, sizeof(int) * ng1, hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipMemcpyAsync(d_g + 2 * ng1, ngc, sizeof(int) * ng1, hipMemcpyHostToDevice,
                          c->stream));
  double* dc = d_tab;
  double* ds = d_tab + 3 * (size_t)natom * ld;
  hipLaunchKernelGGL(k_stscfn_axis, dim3(cdivu(3L * natom, 64)), dim3(64), 0, c->stream, d_cs1,
                     natom, kmax1, kmax2, kmax3, ld, dc, ds);
  hipLaunchKernelGGL(k_stscfn, dim3(cdivu(ngmax + 1, 256), natom), dim3(256), 0, c->stream, dc, ds,
                     ld, kmax1, kmax2, kmax3, d_g, d_g + ng1, d_g + 2 * ng1, ngmax,
                     natom, nplwcs + 1, c->d_ycos, c->d_ysin);
  QX_CHECK(hipGetLastError());
  if (ycos_out)
    QX_CHECK(hipMemcpyAsync(ycos_out, c->d_ycos, sizeof(double) * ny, hipMemcpyDeviceToHost,
                            c->stream));
  if (ysin_out)
    QX_CHECK(hipMemcpyAsync(ysin_out, c->d_ysin, sizeof(double) * ny, hipMemcpyDeviceToHost,
                            c->stream));
  QX_CHECK(hipStreamSynchronize(c->stream));
  (void)hipFree(d_cs1);
  // the axis tables and the G list stay for flocal_g (local-pseudopotential forces)
  if (c->sf.d_tab) (void)hipFree(c->sf.d_tab);
  if (c->sf.d_g) (void)hipFree(c->sf.d_g);
  c->sf.d_tab = d_tab;
  c->sf.d_g = d_g;
  c->sf.ld = ld;
  c->sf.km[0] = kmax1;
  c->sf.km[1] = kmax2;
  c->sf.km[2] = kmax3;
  c->sf.ngmax = nkeep;
  c->sf.natom = natom;
  return 0;
}

int nl_setnlc_g(Ctx* c, int nplw, int ntype, const int* nhk1, const int* nhk2, int natom,
                const int* lkbppi, const int* lking, const double* ycos, const double* ysin,
                int nplwcs, double rvol, double* vclalp_out) {
  auto& n = c->nl;
  if (!n.d_tbqlm || n.ntype != ntype) return -4;  // tables not registered
  if (nplw != c->nplw || nplwcs < nplw || natom < 1 || !(rvol > 0.0)) return -1;
  const double* dyc = c->d_ycos;
  const double* dys = c->d_ysin;
  double* up_c = nullptr;
  double* up_s = nullptr;
  if (ycos && ysin) {
    const size_t ny = (size_t)(nplwcs + 1) * natom;
    up_c = upload(ycos, ny, c->stream);
    up_s = upload(ysin, ny, c->stream);
    if (!up_c || !up_s) return -3;
    dyc = up_c;
    dys = up_s;
  } else if (!c->d_ycos || c->sf_natom != natom || c->sf_nplwcs != nplwcs) {
    return -5;  // no structure factors of this shape on the device
  }
  // type of each atom from the nhk1..nhk2 runs (1-based)
  std::vector<int> type(natom, -1);
  for (int it = 0; it < ntype; ++it)
    for (int ia = nhk1[it] - 1; ia < nhk2[it]; ++ia)
      if (ia >= 0 && ia < natom) type[ia] = it;
  // vclalp(l, it) = clalp(l, it) / rvol for the KB types (1258-1265)
  std::vector<double> vcl((size_t)n.lmmax * ntype, 0.0);
  for (int it = 0; it < ntype; ++it)
    if (lkbppi[it] && !lking[it])
      for (int l = 0; l < n.lmx[it]; ++l)
        vcl[l + (size_t)n.lmmax * it] = n.clalp[l + (size_t)n.lmmax * it] / rvol;
  if (vclalp_out) std::memcpy(vclalp_out, vcl.data(), sizeof(double) * vcl.size());
  // projectors: atom-major, l fastest, KB types only (the order of hckbpp_g)
  std::vector<ProjSpec> ps;
  std::vector<int> pa, pl;
  std::vector<double> D;
  const size_t ntb = (size_t)n.mxqlm + 1, nq = (size_t)(n.lplh + 1) * n.nqlmmx;
  for (int ia = 0; ia < natom; ++ia) {
    const int it = type[ia];
    if (it < 0 || !lkbppi[it] || lking[it]) continue;
    for (int l = 0; l < n.lmx[it]; ++l) {
      const int k = l + n.lmmax * it;
      ProjSpec s;
      s.ia = ia;
      s.tb = (int)(ntb * ((size_t)(n.itola[k] - 1) + (size_t)n.lnref * it));
      s.q = (int)((n.itolmf[k] == 1 ? 0 : nq) + (size_t)(n.lplh + 1) * (n.itolm[k] - 1));
      s.par = n.litoll[k] ? 1 : 0;
      s.dl = n.dlqlm[it];
      ps.push_back(s);
      pa.push_back(ia);
      pl.push_back(l);
      D.push_back(vcl[k]);
    }
  }
  const int npj = (int)ps.size();
  const size_t nex = (size_t)c->nplwex;
  mg_forget_kb(c);
  buffers_changed(c);
  if (c->d_B) { (void)hipFree(c->d_B); c->d_B = nullptr; }
  if (c->d_D) { (void)hipFree(c->d_D); c->d_D = nullptr; }
  if (c->d_P) { (void)hipFree(c->d_P); c->d_P = nullptr; }
  c->npj = npj;
  c->lmmax = n.lmmax;
  c->natom = natom;
  c->proj_atom = pa;
  c->proj_l = pl;
  if (npj > 0) {
    QX_CHECK(hipMalloc(&c->d_B, sizeof(double) * nex * npj));
    QX_CHECK(hipMalloc(&c->d_D, sizeof(double) * npj));
    QX_CHECK(hipMalloc(&c->d_P, sizeof(double) * (size_t)npj * c->nbmax));
    ProjSpec* d_ps = upload(ps.data(), ps.size(), c->stream);
    if (!d_ps) return -3;
    QX_CHECK(hipMemcpyAsync(c->d_D, D.data(), sizeof(double) * npj, hipMemcpyHostToDevice,
                            c->stream));
    hipLaunchKernelGGL(k_setbgq, dim3(cdivu(nplw + 1, 256), npj), dim3(256), 0, c->stream, d_ps,
                       nplw + 1, c->d_perm, c->d_kin, dyc, dys, nplwcs + 1, n.d_tbqlm,
                       n.d_tbqlma, n.d_qlm, c->d_B, c->nplwex);
    QX_CHECK(hipGetLastError());
    QX_CHECK(hipStreamSynchronize(c->stream));
    (void)hipFree(d_ps);
  }
  // the table's factors for the one-band apply (before the structure factors
  // uploaded for this call are freed)
  if (int rc = kb1_build(c, type.data(), natom, lkbppi, lking, dyc, dys, nplwcs + 1)) return rc;
  if (up_c) (void)hipFree(up_c);
  if (up_s) (void)hipFree(up_s);
  return 0;
}

}  // namespace qx
