// restart.cpp -- the reference's restart files, read and written on the host
// in their own layout (Fortran sequential unformatted records: a 4-byte
// little-endian length before and after each record):
//   QM_eig, svwfns / rdwfns (Sources/savedd.f90:2441-2810, 268-1300): wave
//     functions with eigenvalues and occupations, one spin, ncscale = 1,
//     either FP64 records (lsreal8) or per band a scale factor and 16-bit
//     integers;
//   QM_ion, svacon / rdacon (Sources/savedd.f90:3058-3143, 40-265): atom
//     configuration, file version -2.
// Files written here are byte-identical to the reference's for the same data.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct RecWriter {
  FILE* f = nullptr;
  bool ok = true;
  void rec(const void* p, size_t n) {
    if (!f || n > 0x7fffffffu) {
      ok = false;
      return;
    }
    const int32_t len = (int32_t)n;
    ok = ok && fwrite(&len, 4, 1, f) == 1;
    if (n) ok = ok && fwrite(p, 1, n, f) == n;
    ok = ok && fwrite(&len, 4, 1, f) == 1;
  }
  template <typename T>
  void scalar(T v) { rec(&v, sizeof(T)); }
};

struct RecReader {
  FILE* f = nullptr;
  bool ok = true;
  // reads one record; its length must be n (exact == true) or at least n
  bool rec(void* p, size_t n, bool exact = true) {
    int32_t len = 0, len2 = 0;
    if (!ok || fread(&len, 4, 1, f) != 1 || len < 0) return ok = false;
    if ((exact && (size_t)len != n) || (size_t)len < n) return ok = false;
    if (n && fread(p, 1, n, f) != n) return ok = false;
    if (fseek(f, (long)(len - (int32_t)n), SEEK_CUR) != 0) return ok = false;
    if (fread(&len2, 4, 1, f) != 1 || len2 != len) return ok = false;
    return true;
  }
};

// Fortran LOGICAL(4) as flang stores it
inline int32_t flogical(bool v) { return v ? 1 : 0; }

}  // namespace

extern "C" {

// svwfns for one spin and ncscale = 1: gdcr(nplwex, nband) column-major in the
// reference's G order, eig(nband), occ(nband).  lsreal8 = 0 stores each band
// as efactr = max|x| / 32767 and the 16-bit integers trunc(x * 32767 / max|x|).
int qx_svwfns(const char* fname, const double* gdcr, int nplwex, int nplw, int nband,
              const double* eig, const double* occ, int lsreal8) {
  if (!fname || nplwex != 2 * (nplw + 1) || nband < 1) return -1;
  RecWriter w;
  w.f = fopen(fname, "wb");
  if (!w.f) return -7;
  w.scalar<int32_t>(-1);  // -ncscale
  w.scalar<int32_t>(nband);
  w.scalar<int32_t>(nplw);
  w.scalar<int32_t>(flogical(lsreal8 != 0));
  w.scalar<int32_t>(1);   // nspnmx
  w.rec(eig, sizeof(double) * nband);
  w.rec(occ, sizeof(double) * nband);
  std::vector<int16_t> q(nplwex);
  for (int n = 0; n < nband && w.ok; ++n) {
    const double* x = gdcr + (size_t)n * nplwex;
    if (lsreal8) {
      w.rec(x, sizeof(double) * nplwex);
      continue;
    }
    double eigmax = 0.0;
    for (int i = 0; i < nplwex; ++i) eigmax = std::max(eigmax, std::fabs(x[i]));
    if (!(eigmax > 0.0)) {  // the reference divides by zero here
      w.ok = false;
      break;
    }
    const double efact = 32767.0 / eigmax, efactr = eigmax / 32767.0;
    for (int i = 0; i < nplwex; ++i) q[i] = (int16_t)(efact * x[i]);  // Fortran INT: toward 0
    w.scalar<double>(efactr);
    w.rec(q.data(), sizeof(int16_t) * nplwex);
  }
  const bool ok = w.ok;
  if (fclose(w.f) != 0 || !ok) return -7;
  return 0;
}

// rdwfns for one spin (the path of 330-470 and rdwfn22, 1027-1172): the
// header's nband / nplw / storage kind, eig and occ, then the bands.  The
// file's nplw must equal nplw; min(nband, file's nband) bands are read;
// *nband_file receives the file's band count.  A file without the ncscale
// record (positive first record) is read as the older layout without
// occupations (occ left unchanged).
int qx_rdwfns(const char* fname, double* gdcr, int nplwex, int nplw, int nband, double* eig,
              double* occ, int* nband_file) {
  if (!fname || nplwex != 2 * (nplw + 1) || nband < 1) return -1;
  RecReader r;
  r.f = fopen(fname, "rb");
  if (!r.f) return -7;
  int32_t first = 0, nbo = 0, nplwo = 0, lsr = 0, nspin = 0;
  r.rec(&first, 4);
  bool readocc = first < 0;
  if (readocc) {
    if (first != -1) r.ok = false;  // ncscale = 2 (noncollinear) is not read here
    r.rec(&nbo, 4);
  } else {
    nbo = first;
  }
  r.rec(&nplwo, 4);
  r.rec(&lsr, 4);
  r.rec(&nspin, 4);
  if (!r.ok || nplwo != nplw || nbo < 1 || nspin < 1) {
    fclose(r.f);
    return r.ok ? -2 : -7;
  }
  const int nb = std::min(nband, (int)nbo);
  std::vector<double> tmp((size_t)nbo);
  // per spin eig (and occ); the first spin's are returned
  for (int s = 0; s < nspin && r.ok; ++s) {
    r.rec(tmp.data(), sizeof(double) * nbo);
    if (s == 0) std::memcpy(eig, tmp.data(), sizeof(double) * nb);
    if (readocc) {
      r.rec(tmp.data(), sizeof(double) * nbo);
      if (s == 0) std::memcpy(occ, tmp.data(), sizeof(double) * nb);
    }
  }
  std::vector<int16_t> q(nplwex);
  for (int n = 0; n < nb && r.ok; ++n) {
    double* x = gdcr + (size_t)n * nplwex;
    if (lsr) {
      r.rec(x, sizeof(double) * nplwex);
    } else {
      double efactr = 0.0;
      r.rec(&efactr, sizeof(double));
      r.rec(q.data(), sizeof(int16_t) * nplwex);
      for (int i = 0; i < nplwex; ++i) x[i] = (double)q[i] * efactr;
    }
  }
  const bool ok = r.ok;
  fclose(r.f);
  if (!ok) return -7;
  if (nband_file) *nband_file = nbo;
  return 0;
}

// svacon, file version -2: nhk1/nhk2(ntype), ratm/frc(3, natom),
// prevr(3, natom, 3), hcell(3, 3), all column-major as in the reference;
// natom = nhk2(ntype).
int qx_svacon(const char* fname, int ntype, const int* nhk1, const int* nhk2, int nstepCG,
              int nstepMD, int lclust, const double* ratm, int ifmd, const double* frc,
              const double* prevr, const double* hcell) {
  if (!fname || ntype < 1) return -1;
  const int natom = nhk2[ntype - 1];
  if (natom < 0) return -1;
  RecWriter w;
  w.f = fopen(fname, "wb");
  if (!w.f) return -7;
  w.scalar<int32_t>(-2);
  w.scalar<int32_t>(ntype);
  std::vector<int32_t> hk(2 * (size_t)ntype);
  for (int it = 0; it < ntype; ++it) {
    hk[2 * it] = nhk1[it];
    hk[2 * it + 1] = nhk2[it];
  }
  w.rec(hk.data(), sizeof(int32_t) * hk.size());
  const int32_t steps[2] = {nstepCG, nstepMD};
  w.rec(steps, sizeof(steps));
  w.scalar<int32_t>(flogical(lclust != 0));
  w.rec(ratm, sizeof(double) * 3 * (size_t)natom);
  w.scalar<int32_t>(ifmd);
  w.rec(frc, sizeof(double) * 3 * (size_t)natom);
  w.rec(prevr, sizeof(double) * 9 * (size_t)natom);
  w.rec(hcell, sizeof(double) * 9);
  const bool ok = w.ok;
  if (fclose(w.f) != 0 || !ok) return -7;
  return 0;
}

// rdacon of a version -2 file written for the same ntype / nhk1 / nhk2 (the
// reference's check); outputs as qx_svacon's inputs.
int qx_rdacon(const char* fname, int ntype, const int* nhk1, const int* nhk2, int* nstepCG,
              int* nstepMD, int* lclust, double* ratm, int* ifmd, double* frc, double* prevr,
              double* hcell) {
  if (!fname || ntype < 1) return -1;
  const int natom = nhk2[ntype - 1];
  RecReader r;
  r.f = fopen(fname, "rb");
  if (!r.f) return -7;
  int32_t ver = 0, nt = 0, lc = 0, imd = 0;
  r.rec(&ver, 4);
  if (r.ok && ver != -2) r.ok = false;
  r.rec(&nt, 4);
  if (!r.ok || nt != ntype) {
    fclose(r.f);
    return r.ok ? -2 : -7;
  }
  std::vector<int32_t> hk(2 * (size_t)ntype);
  r.rec(hk.data(), sizeof(int32_t) * hk.size());
  for (int it = 0; it < ntype && r.ok; ++it)
    if (hk[2 * it] != nhk1[it] || hk[2 * it + 1] != nhk2[it]) {
      fclose(r.f);
      return -2;
    }
  int32_t steps[2] = {0, 0};
  r.rec(steps, sizeof(steps));
  r.rec(&lc, 4);
  r.rec(ratm, sizeof(double) * 3 * (size_t)natom);
  r.rec(&imd, 4);
  r.rec(frc, sizeof(double) * 3 * (size_t)natom);
  r.rec(prevr, sizeof(double) * 9 * (size_t)natom);
  r.rec(hcell, sizeof(double) * 9);
  const bool ok = r.ok;
  fclose(r.f);
  if (!ok) return -7;
  *nstepCG = steps[0];
  *nstepMD = steps[1];
  *lclust = lc != 0;
  *ifmd = imd;
  return 0;
}

}  // extern "C"
