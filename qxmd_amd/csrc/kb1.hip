// kb1.hip -- the KB apply of one band (hckbpp_ib_g, Sources/nlkbpp.f90:
// 1623-1721) from the factors of the projector table.  When the table comes
// from setnlc_g on the device, its row of projector p = (atom ia, l) is
//   B_p(G) = (yc z, -ys z) (litoll) or (-ys z, -yc z),  (yc, ys) = the
//   structure factor of ia at G, z = the radial x angular factor of (type, l)
// (k_setbgq, Sources/nlkbpp.f90:2171-2183).  One band's projection and
// back-projection then read the natom phases and the few z rows -- 16 B per
// (G, atom) -- instead of the lmx rows of 16 B per (G, atom) each: a quarter
// of the bytes for s + p projectors.  B_p(G) is rebuilt with the operations
// of k_setbgq, so the terms are those of the table path; the sums run in
// another order.  Deterministic: per-chunk partial sums, fixed-order totals.
#include <algorithm>
#include <vector>
#include "context.h"

namespace qx {

namespace {

constexpr int kK1Threads = 256;
constexpr int kK1Chunk = 1024;  // G per projection workgroup (4 per thread)
constexpr int kK1MaxL = 16;     // projectors per atom
constexpr int kK1Coef = 64;     // projectors per back-projection workgroup

inline unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

struct ClassSpec {
  int tb, q;
  double dl;
};

// z of every (type, l) class and the phase of every atom, at each internal G
__global__ void k_kb1_tables(const ClassSpec* __restrict__ cs, int ncls, const int* __restrict__ perm,
                             const double* __restrict__ kin, const double* __restrict__ ycos,
                             const double* __restrict__ ysin, int ldy, int natom,
                             const double* __restrict__ tbqlm, const double* __restrict__ tbqlma,
                             const double* __restrict__ qlm, int nplw1, double2* __restrict__ ph,
                             double* __restrict__ z) {
#pragma clang fp contract(off)
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= nplw1) return;
  const int ig = perm[ip];
  const double g = sqrt(kin[ip]);
  for (int k = 0; k < ncls; ++k) {
    const ClassSpec s = cs[k];
    int m = (int)(g / (2.0 * s.dl));
    m = 2 * m;
    const double d = 0.5 * (g / s.dl - (double)m);
    const double* tb = tbqlm + s.tb;
    const double* ta = tbqlma + s.tb;
    const double zz = d * ((d - 1.0) * ta[m] + tb[m + 2] - tb[m]) + tb[m];
    z[(size_t)k * nplw1 + ip] = zz * qlm[s.q + ig];
  }
  for (int ia = 0; ia < natom; ++ia)
    ph[(size_t)ia * nplw1 + ip] = make_double2(ycos[(size_t)ia * ldy + ig], ysin[(size_t)ia * ldy + ig]);
}

// B_p(G) of the table, from the phase and z as k_setbgq forms it
__device__ inline double2 kb1_entry(double2 h, double zz, int par) {
#pragma clang fp contract(off)
  return par ? make_double2(h.x * zz, -h.y * zz) : make_double2(-h.y * zz, -h.x * zz);
}

// part[p * nch + ch] = sum over the chunk's G of B_p . v, for the projectors
// of atom blockIdx.y
__global__ void __launch_bounds__(kK1Threads)
    k_kb1_proj(const double* __restrict__ v, const double2* __restrict__ ph,
               const double* __restrict__ z, int nplw1, const int* __restrict__ p0,
               const int* __restrict__ pcls, const int* __restrict__ ppar, int nch,
               double* __restrict__ part) {
  const int ch = blockIdx.x, ia = blockIdx.y;
  const int pb = p0[ia], np = min(p0[ia + 1] - pb, kK1MaxL);
  if (np <= 0) return;  // uniform over the workgroup
  double acc[kK1MaxL];
#pragma unroll
  for (int j = 0; j < kK1MaxL; ++j) acc[j] = 0.0;
#pragma unroll
  for (int u = 0; u < kK1Chunk / kK1Threads; ++u) {
    const int ip = ch * kK1Chunk + threadIdx.x + u * kK1Threads;
    if (ip < nplw1) {
      const double vr = v[2 * ip], vi = v[2 * ip + 1];
      const double2 h = ph[(size_t)ia * nplw1 + ip];
#pragma unroll
      for (int j = 0; j < kK1MaxL; ++j) {
        if (j < np) {
          const double2 b = kb1_entry(h, z[(size_t)pcls[pb + j] * nplw1 + ip], ppar[pb + j]);
          acc[j] += b.x * vr + b.y * vi;
        }
      }
    }
  }
  __shared__ double red[kK1MaxL][kK1Threads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kK1MaxL; ++j) {
    if (j < np) {
      double s = acc[j];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
      if (lane == 0) red[j][wv] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < np) {
    double t = 0.0;
    for (int w = 0; w < kK1Threads / 64; ++w) t += red[threadIdx.x][w];
    part[(size_t)(pb + threadIdx.x) * nch + ch] = t;
  }
}

// pn[grp nex + 2 ip + {0, 1}] = sum over the group's atoms and projectors of
// B_p(G) coef_p, coef_p = D_p (2 sum_ch part[p nch + ch] - B_p(0).x v[0]
// - 2 B_p(0).y v[1]) (the packed product with the Im G = 0 term dropped)
__global__ void __launch_bounds__(kK1Threads)
    k_kb1_back(const double* __restrict__ v, const double2* __restrict__ ph,
               const double* __restrict__ z, int nplw1, const int* __restrict__ p0,
               const int* __restrict__ pcls, const int* __restrict__ ppar,
               const double* __restrict__ D, int natom, int apg, int nch,
               const double* __restrict__ part, int nex, double* __restrict__ pn) {
  const int grp = blockIdx.y, a0 = grp * apg, a1 = min(natom, a0 + apg);
  const int pb = p0[a0], npr = p0[a1] - pb;
  __shared__ double coef[kK1Coef];
  __shared__ int aof[kK1Coef];  // atom of each projector
  {
    constexpr int kSub = kK1Threads / kK1Coef;
    const int jj = threadIdx.x / kSub, sub = threadIdx.x % kSub;
    const int p = pb + min(jj, max(npr - 1, 0));
    double t = 0.0;
    if (npr > 0)
      for (int c = sub; c < nch; c += kSub) t += part[(size_t)p * nch + c];
    for (int o = kSub / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (sub == 0 && jj < npr) {
      int ia = a0;
      while (p0[ia + 1] <= p) ++ia;
      const double2 b0 = kb1_entry(ph[(size_t)ia * nplw1], z[(size_t)pcls[p] * nplw1], ppar[p]);
      const double c0 = b0.x * v[0] + 2.0 * b0.y * v[1];
      coef[jj] = (2.0 * t - c0) * D[p];
      aof[jj] = ia;
    }
  }
  __syncthreads();
  const int ip = blockIdx.x * kK1Threads + threadIdx.x;
  if (ip >= nplw1) return;
  double re = 0.0, im = 0.0;
  double zc[kK1MaxL];
  int cls0 = -1, ncached = 0;
  for (int j = 0; j < npr;) {  // atom by atom
    const int ia = aof[j];
    const int pa = pb + j, na = min(p0[ia + 1] - pa, kK1MaxL);
    if (pcls[pa] != cls0 || na != ncached) {
      cls0 = pcls[pa];
      ncached = na;
#pragma unroll
      for (int q = 0; q < kK1MaxL; ++q)
        zc[q] = q < na ? z[(size_t)pcls[pa + q] * nplw1 + ip] : 0.0;
    }
    const double2 h = ph[(size_t)ia * nplw1 + ip];
#pragma unroll
    for (int q = 0; q < kK1MaxL; ++q) {
      if (q < na) {
        const double2 b = kb1_entry(h, zc[q], ppar[pa + q]);
        re += b.x * coef[j + q];
        im += b.y * coef[j + q];
      }
    }
    j += na;
  }
  pn[(size_t)grp * nex + 2 * ip] = re;
  pn[(size_t)grp * nex + 2 * ip + 1] = im;
}

}  // namespace

void kb1_destroy(Ctx* c) {
  auto& K = c->kb1;
  for (void* p : {(void*)K.d_ph, (void*)K.d_z, (void*)K.d_p0, (void*)K.d_pcls, (void*)K.d_ppar})
    if (p) (void)hipFree(p);
  K = Ctx::Kb1{};
}

// the factors of the table k_setbgq just built (same projector order)
int kb1_build(Ctx* c, const int* type, int natom, const int* lkbppi, const int* lking,
              const double* dyc, const double* dys, int ldy) {
  auto& n = c->nl;
  auto& K = c->kb1;
  K.valid = false;
  const size_t ntb = (size_t)n.mxqlm + 1, nq = (size_t)(n.lplh + 1) * n.nqlmmx;
  std::vector<ClassSpec> cs((size_t)n.lmmax * n.ntype);
  for (int it = 0; it < n.ntype; ++it)
    for (int l = 0; l < n.lmmax; ++l) {
      const int k = l + n.lmmax * it;
      ClassSpec s{0, 0, n.dlqlm[it]};
      if (l < n.lmx[it]) {
        s.tb = (int)(ntb * ((size_t)(n.itola[k] - 1) + (size_t)n.lnref * it));
        s.q = (int)((n.itolmf[k] == 1 ? 0 : nq) + (size_t)(n.lplh + 1) * (n.itolm[k] - 1));
      }
      cs[k] = s;
    }
  // projectors atom-major, l fastest, KB types only: the order of nl_setnlc_g
  std::vector<int> p0(natom + 1, 0), pcls, ppar;
  int maxl = 0;
  for (int ia = 0; ia < natom; ++ia) {
    p0[ia] = (int)pcls.size();
    const int it = type[ia];
    if (it >= 0 && lkbppi[it] && !lking[it]) {
      for (int l = 0; l < n.lmx[it]; ++l) {
        const int k = l + n.lmmax * it;
        pcls.push_back(k);
        ppar.push_back(n.litoll[k] ? 1 : 0);
      }
      maxl = std::max(maxl, n.lmx[it]);
    }
  }
  p0[natom] = (int)pcls.size();
  if (pcls.empty() || maxl > kK1MaxL || (int)pcls.size() != c->npj) {  // table path only
    kb1_destroy(c);
    return 0;
  }
  const int nplw1 = c->nplw + 1;
  // the arrays of the previous MD step are reused when the shapes repeat
  if (!K.d_ph || K.natom != natom || K.nplw1 != nplw1 || K.ncls != (int)cs.size() ||
      K.npj != (int)pcls.size()) {
    kb1_destroy(c);
    QX_CHECK(hipMalloc(&K.d_ph, sizeof(double2) * (size_t)natom * nplw1));
    QX_CHECK(hipMalloc(&K.d_z, sizeof(double) * cs.size() * nplw1));
    QX_CHECK(hipMalloc(&K.d_p0, sizeof(int) * p0.size()));
    QX_CHECK(hipMalloc(&K.d_pcls, sizeof(int) * pcls.size()));
    QX_CHECK(hipMalloc(&K.d_ppar, sizeof(int) * ppar.size()));
  }
  K.ncls = (int)cs.size();
  K.natom = natom;
  K.nplw1 = nplw1;
  K.npj = (int)pcls.size();
  K.apg = std::max(1, kK1Coef / maxl);
  ClassSpec* d_cs = nullptr;
  QX_CHECK(hipMalloc(&d_cs, sizeof(ClassSpec) * cs.size()));
  QX_CHECK(hipMemcpyAsync(d_cs, cs.data(), sizeof(ClassSpec) * cs.size(), hipMemcpyHostToDevice,
                          c->stream));
  QX_CHECK(hipMemcpyAsync(K.d_p0, p0.data(), sizeof(int) * p0.size(), hipMemcpyHostToDevice,
                          c->stream));
  QX_CHECK(hipMemcpyAsync(K.d_pcls, pcls.data(), sizeof(int) * pcls.size(),
                          hipMemcpyHostToDevice, c->stream));
  QX_CHECK(hipMemcpyAsync(K.d_ppar, ppar.data(), sizeof(int) * ppar.size(),
                          hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_kb1_tables, dim3(cdivu(nplw1, 256)), dim3(256), 0, c->stream, d_cs,
                     K.ncls, c->d_perm, c->d_kin, dyc, dys, ldy, natom, n.d_tbqlm, n.d_tbqlma,
                     n.d_qlm, nplw1, reinterpret_cast<double2*>(K.d_ph), K.d_z);
  QX_CHECK(hipGetLastError());
  QX_CHECK(hipStreamSynchronize(c->stream));
  (void)hipFree(d_cs);
  K.valid = true;
  return 0;
}

// scratch doubles kb1_project needs at pn (partials, then chunk sums)
size_t kb1_scratch(const Ctx* c) {
  const auto& K = c->kb1;
  if (!K.valid) return 0;
  const size_t ngrp = (K.natom + K.apg - 1) / K.apg;
  const size_t nch = (K.nplw1 + kK1Chunk - 1) / kK1Chunk;
  return ngrp * c->nplwex + (size_t)c->npj * nch;
}

// both passes for the vector v; returns the number of partial vectors at pn
// (their sum is the KB term of H v), 0 if the factors are not available
int kb1_project(Ctx* c, const double* v, double* pn) {
  const auto& K = c->kb1;
  if (!K.valid) return 0;
  const int ngrp = (K.natom + K.apg - 1) / K.apg;
  const int nch = (K.nplw1 + kK1Chunk - 1) / kK1Chunk;
  double* part = pn + (size_t)ngrp * c->nplwex;
  const double2* ph = reinterpret_cast<const double2*>(K.d_ph);
  hipLaunchKernelGGL(k_kb1_proj, dim3(nch, K.natom), dim3(kK1Threads), 0, c->stream, v, ph, K.d_z,
                     K.nplw1, K.d_p0, K.d_pcls, K.d_ppar, nch, part);
  hipLaunchKernelGGL(k_kb1_back, dim3(cdivu(K.nplw1, kK1Threads), ngrp), dim3(kK1Threads), 0,
                     c->stream, v, ph, K.d_z, K.nplw1, K.d_p0, K.d_pcls, K.d_ppar, c->d_D,
                     K.natom, K.apg, nch, part, c->nplwex, pn);
  return ngrp;
}

}  // namespace qx
