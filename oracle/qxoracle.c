/*
 * qxoracle.c -- CPU FP64 oracle for the QXMD Gamma-point H|psi> / band-update path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference algorithm that
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use as the
 * checker.  Nothing in the product links or calls this file.
 *
 * Every routine cites the reference file:line it restates.  Floating-point
 * expressions that feed integer decisions (the G-vector sort key) are written
 * in the reference's evaluation order and this file is compiled with
 * -ffp-contract=off so that the G maps come out bit-identical.
 *
 * The FFT used here is a textbook mixed-radix decimation-in-time transform
 * (radix 4/2/3/5, naive DFT for other primes) with FFTW's sign and storage
 * conventions; it stands in for FFTW 3.3.8, which the reference calls
 * (Sources/fft3.f90:113-333) but does not vendor.
 */
#include "qxoracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 1-based column-major 3x3 accessor: M(a, r, c) = a(r, c) */
#define M3(a, r, c) (a)[((r) - 1) + 3 * ((c) - 1)]

/* ======================================================================= */
/* Lattice and FFT-grid rule                                               */
/* ======================================================================= */

/* rciprl: volume and reciprocal vectors / 2pi (Sources/reclat.f90:1089-1132). */
void qxo_rciprl(const double a[9], double b[9], double *v)
{
  double vr;
  M3(b, 1, 1) = M3(a, 2, 2) * M3(a, 3, 3) - M3(a, 3, 2) * M3(a, 2, 3);
  M3(b, 2, 1) = M3(a, 3, 2) * M3(a, 1, 3) - M3(a, 1, 2) * M3(a, 3, 3);
  M3(b, 3, 1) = M3(a, 1, 2) * M3(a, 2, 3) - M3(a, 2, 2) * M3(a, 1, 3);
  *v = M3(a, 1, 1) * M3(b, 1, 1) + M3(a, 2, 1) * M3(b, 2, 1) + M3(a, 3, 1) * M3(b, 3, 1);
  vr = 1.0 / *v;
  M3(b, 1, 1) *= vr;
  M3(b, 2, 1) *= vr;
  M3(b, 3, 1) *= vr;
  M3(b, 1, 2) = (M3(a, 2, 3) * M3(a, 3, 1) - M3(a, 3, 3) * M3(a, 2, 1)) * vr;
  M3(b, 2, 2) = (M3(a, 3, 3) * M3(a, 1, 1) - M3(a, 1, 3) * M3(a, 3, 1)) * vr;
  M3(b, 3, 2) = (M3(a, 1, 3) * M3(a, 2, 1) - M3(a, 2, 3) * M3(a, 1, 1)) * vr;
  M3(b, 1, 3) = (M3(a, 2, 1) * M3(a, 3, 2) - M3(a, 3, 1) * M3(a, 2, 2)) * vr;
  M3(b, 2, 3) = (M3(a, 3, 1) * M3(a, 1, 2) - M3(a, 1, 1) * M3(a, 3, 2)) * vr;
  M3(b, 3, 3) = (M3(a, 1, 1) * M3(a, 2, 2) - M3(a, 2, 1) * M3(a, 1, 2)) * vr;
  *v = fabs(*v);
}

/* FFT sizes FFTW handles well: products of 2,3,5,7 up to 4900, i.e. the
 * LIBFFTW3 table of Sources/reclat.f90:889-918 (regenerated here, not copied). */
static int is_fftw_size(int n)
{
  if (n < 1 || n > 4900) return 0;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  while (n % 7 == 0) n /= 7;
  return n == 1;
}

/* chkftm_org with imev=1 (Sources/reclat.f90:1017-1084): smallest table size
 * >= n; the 2nd and 3rd dimensions start their search at table entry 2. */
int qxo_fft_size(int n, int first_dim)
{
  int m = first_dim ? 1 : 2;
  for (; m <= 4900; ++m)
    if (is_fftw_size(m) && m >= n) return m;
  return n;
}

/* get_mesh (Sources/input.f90:1082-1134). */
void qxo_get_mesh(const double hcell[9], double ecutdens, int nd[3])
{
  static const int IIP1[3] = {2, 3, 1}, IIP2[3] = {3, 1, 2}, IIP3[3] = {1, 2, 3};
  double B[9], rvol, dpi, gx, a1x, a1y, a1z, tri, aaa, pk3;
  int i, kmaxm, raw[3];
  qxo_rciprl(hcell, B, &rvol);
  dpi = 2.0 * acos(-1.0);
  gx = sqrt(ecutdens) / dpi;
  for (i = 0; i < 3; ++i) {
    int ip1 = IIP1[i], ip2 = IIP2[i], ip3 = IIP3[i];
    a1x = M3(B, 2, ip1) * M3(B, 3, ip2) - M3(B, 3, ip1) * M3(B, 2, ip2);
    a1y = M3(B, 3, ip1) * M3(B, 1, ip2) - M3(B, 1, ip1) * M3(B, 3, ip2);
    a1z = M3(B, 1, ip1) * M3(B, 2, ip2) - M3(B, 2, ip1) * M3(B, 1, ip2);
    tri = M3(B, 1, ip3) * a1x + M3(B, 2, ip3) * a1y + M3(B, 3, ip3) * a1z;
    aaa = a1x * a1x + a1y * a1y + a1z * a1z;
    pk3 = tri / sqrt(aaa);
    kmaxm = (int)(gx / pk3);
    raw[i] = 2 * kmaxm + 1;
  }
  nd[0] = qxo_fft_size(raw[0], 1);
  nd[1] = (raw[1] == raw[0]) ? nd[0] : qxo_fft_size(raw[1], 0);
  nd[2] = (raw[2] == raw[0]) ? nd[0] : qxo_fft_size(raw[2], 0);
}

/* ======================================================================= */
/* G sphere: strvec (Sources/reclat.f90:500-875), set_rciprl               */
/* (Sources/modules.f90:1402-1603)                                         */
/* ======================================================================= */

typedef struct { int kmaxm, kmaxm1[3]; double cross[9], rvol; } sv_setup;

static void strvec_setup(const double rba[9], double ecutdens, sv_setup *s)
{
  static const int IIP1[3] = {2, 3, 1}, IIP2[3] = {3, 1, 2}, IIP3[3] = {1, 2, 3};
  double B[9], dpi, gxfft, a1x, a1y, a1z, aaa, pk3, dmmz;
  int i, j;
  qxo_rciprl(rba, B, &s->rvol);
  dpi = 2.0 * acos(-1.0);
  for (i = 0; i < 9; ++i) s->cross[i] = dpi * B[i];
  gxfft = sqrt(ecutdens) / dpi;
  s->kmaxm = 0;
  for (i = 0; i < 3; ++i) {
    int ip1 = IIP1[i], ip2 = IIP2[i], ip3 = IIP3[i];
    a1x = M3(B, 2, ip1) * M3(B, 3, ip2) - M3(B, 3, ip1) * M3(B, 2, ip2);
    a1y = M3(B, 3, ip1) * M3(B, 1, ip2) - M3(B, 1, ip1) * M3(B, 3, ip2);
    a1z = M3(B, 1, ip1) * M3(B, 2, ip2) - M3(B, 2, ip1) * M3(B, 1, ip2);
    aaa = a1x * a1x + a1y * a1y + a1z * a1z;
    aaa = 1.0 / sqrt(aaa);
    a1x *= aaa; a1y *= aaa; a1z *= aaa;
    pk3 = M3(B, 1, ip3) * a1x + M3(B, 2, ip3) * a1y + M3(B, 3, ip3) * a1z;
    dmmz = gxfft / pk3;
    j = (int)fabs(dmmz);
    s->kmaxm1[i] = j;
    if (j > s->kmaxm) s->kmaxm = j;
  }
}

/* The half-space generator of strvec (Sources/reclat.f90:626-763): for every
 * 0 <= kh <= jh <= ih (ih >= 1) it emits the signed permutations of (ih,jh,kh)
 * that represent one of each +-G pair, in a fixed order keyed by which of the
 * three magnitudes coincide or vanish.  The emission order decides the tie
 * order of the stable |G| sort, so it is kept identical. */
static int strvec_enumerate(const sv_setup *s, int *k1, int *k2, int *k3)
{
  int numg = 0, i, j, k, mkk;
  for (i = 2; i <= s->kmaxm + 1; ++i) {
    int ih = i - 1;
    for (j = 1; j <= i; ++j) {
      int jh = j - 1;
      for (k = 1; k <= j; ++k) {
        int kh = k - 1, num = 0, kji = 1;
        int t1[24], t2[24], t3[24];
        if (k == 1) kji = 2;
        if (k == j) kji += 2;
        if (i == j) kji += 4;
#define EMIT(a, b, c) do { t1[num] = (a); t2[num] = (b); t3[num] = (c); ++num; } while (0)
        EMIT(ih, jh, kh);
        if (kji != 7) { EMIT(kh, ih, jh); EMIT(jh, kh, ih); }
        if (kji != 4) {
          EMIT(-ih, jh, kh); EMIT(kh, -ih, jh); EMIT(jh, kh, -ih);
          if (kji != 7 && kji != 6) {
            if (kji != 2) {
              EMIT(ih, -jh, kh); EMIT(kh, ih, -jh); EMIT(-jh, kh, ih);
              EMIT(ih, jh, -kh); EMIT(-kh, ih, jh); EMIT(jh, -kh, ih);
            }
            if (kji != 3 && kji != 5) {
              EMIT(ih, kh, jh); EMIT(jh, ih, kh); EMIT(kh, jh, ih);
              EMIT(-ih, kh, jh); EMIT(jh, -ih, kh); EMIT(kh, jh, -ih);
              if (kji != 2) {
                EMIT(ih, kh, -jh); EMIT(-jh, ih, kh); EMIT(kh, -jh, ih);
                EMIT(ih, -kh, jh); EMIT(jh, ih, -kh); EMIT(-kh, jh, ih);
              }
            }
          }
        }
#undef EMIT
        for (mkk = 0; mkk < num; ++mkk) {
          if (abs(t1[mkk]) <= s->kmaxm1[0] + 1 && abs(t2[mkk]) <= s->kmaxm1[1] + 1 &&
              abs(t3[mkk]) <= s->kmaxm1[2] + 1) {
            if (k1) { k1[numg] = t1[mkk]; k2[numg] = t2[mkk]; k3[numg] = t3[mkk]; }
            ++numg;
          }
        }
      }
    }
  }
  return numg;
}

int qxo_strvec_capacity(const double rba[9], double ecutdens)
{
  sv_setup s;
  strvec_setup(rba, ecutdens, &s);
  return strvec_enumerate(&s, NULL, NULL, NULL) + 1;
}

typedef struct { int key, ix; } keyrec;
static int keyrec_cmp(const void *a, const void *b)
{
  const keyrec *x = (const keyrec *)a, *y = (const keyrec *)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->ix < y->ix ? -1 : (x->ix > y->ix);
}

/* strvec: returns num (all generated vectors).  The sort of
 * Sources/funcs.f90:1211-1359 (radix on a monotone bucket of the key, then
 * odd-even transposition with strict comparisons) is a stable sort by key of
 * the emission order; it is computed here as a stable sort on (key, index). */
int qxo_strvec(const double rba[9], double ecutdens, double ecut, double ecutsoft,
               double ecutc, double ecutorth, int cap, int counts[6],
               int *nga, int *ngb, int *ngc, double *gx, double *gy, double *gz,
               double *recnrm, double cross[9], double *rvol)
{
  sv_setup s;
  int num, mkk, i;
  int *k1, *k2, *k3;
  double *g1, *g2, *g3, *gss, ddmx, fctmax;
  keyrec *kr;
  strvec_setup(rba, ecutdens, &s);
  num = strvec_enumerate(&s, NULL, NULL, NULL);
  if (cap < num + 1) return -(num + 1);
  k1 = (int *)malloc(sizeof(int) * 3 * (size_t)num);
  k2 = k1 + num; k3 = k2 + num;
  g1 = (double *)malloc(sizeof(double) * 4 * (size_t)num);
  g2 = g1 + num; g3 = g2 + num; gss = g3 + num;
  kr = (keyrec *)malloc(sizeof(keyrec) * (size_t)num);
  strvec_enumerate(&s, k1, k2, k3);
  for (mkk = 0; mkk < num; ++mkk) {
    double d1h = k1[mkk], d2h = k2[mkk], d3h = k3[mkk];
    g1[mkk] = M3(s.cross, 1, 1) * d1h + M3(s.cross, 1, 2) * d2h + M3(s.cross, 1, 3) * d3h;
    g2[mkk] = M3(s.cross, 2, 1) * d1h + M3(s.cross, 2, 2) * d2h + M3(s.cross, 2, 3) * d3h;
    g3[mkk] = M3(s.cross, 3, 1) * d1h + M3(s.cross, 3, 2) * d2h + M3(s.cross, 3, 3) * d3h;
    gss[mkk] = g1[mkk] * g1[mkk] + g2[mkk] * g2[mkk] + g3[mkk] * g3[mkk];
  }
  ddmx = 0.0;
  for (mkk = 0; mkk < num; ++mkk) ddmx = fmax(ddmx, fabs(gss[mkk]));
  fctmax = 1000000000.0 / ddmx;
  for (mkk = 0; mkk < num; ++mkk) {
    kr[mkk].key = (int)(fctmax * gss[mkk]);
    kr[mkk].ix = mkk;
  }
  qsort(kr, (size_t)num, sizeof(keyrec), keyrec_cmp);
  counts[0] = num;
  counts[1] = counts[2] = counts[3] = counts[4] = counts[5] = 0;
  for (i = 0; i < num; ++i) {
    if (gss[i] <= ecutdens) ++counts[1];
    if (gss[i] <= ecut) ++counts[2];
    if (gss[i] <= ecutsoft) ++counts[3];
    if (gss[i] <= ecutc) ++counts[4];
    if (gss[i] <= ecutorth) ++counts[5];
  }
  nga[0] = ngb[0] = ngc[0] = 0;
  gx[0] = gy[0] = gz[0] = recnrm[0] = 0.0;
  for (i = 1; i <= num; ++i) {
    int j = kr[i - 1].ix;
    nga[i] = k1[j]; ngb[i] = k2[j]; ngc[i] = k3[j];
    gx[i] = g1[j]; gy[i] = g2[j]; gz[i] = g3[j];
    recnrm[i] = gss[j];
  }
  for (i = 0; i < 9; ++i) cross[i] = s.cross[i];
  *rvol = s.rvol;
  free(k1); free(g1); free(kr);
  return num;
}

/* reclat: ijkgd(+-m) (Sources/reclat.f90:259-276). */
void qxo_ijkgd(int nplw5, const int *nga, const int *ngb, const int *ngc,
               const int nd[3], int *ijkgd)
{
  int *c = ijkgd + nplw5, m;
  c[0] = 1;
  for (m = 1; m <= nplw5; ++m) {
    int a = nga[m], b = ngb[m], d = ngc[m];
    if (a < 0) a += nd[0];
    if (b < 0) b += nd[1];
    if (d < 0) d += nd[2];
    c[m] = 1 + a + (b + d * nd[1]) * nd[0];
    a = -nga[m]; b = -ngb[m]; d = -ngc[m];
    if (a < 0) a += nd[0];
    if (b < 0) b += nd[1];
    if (d < 0) d += nd[2];
    c[-m] = 1 + a + (b + d * nd[1]) * nd[0];
  }
}

/* dkgnrm(2i+1) = dkgnrm(2i+2) = |G_i|^2 (Sources/modules.f90:1535-1540). */
void qxo_dkgnrm(int nplwc, const double *recnrm, double *dkgnrm)
{
  int i;
  for (i = 0; i <= nplwc; ++i) dkgnrm[2 * i] = dkgnrm[2 * i + 1] = recnrm[i];
}

/* rnd00: Lehmer generator in FP64 (Sources/funcs.f90:1673-1689). */
double qxo_rnd00(double *seed)
{
  double rnd = 16807.0 * *seed;
  *seed = fmod(rnd, 2147483647.0);
  return *seed / 2147483647.0;
}

/* ======================================================================= */
/* FFT                                                                     */
/* ======================================================================= */

typedef struct { int n, sign; double *w; } twtab;
#define NTW 16
static twtab g_tw[NTW];
static int g_ntw = 0;

static const double *twiddles(int n, int sign)
{
  int i;
  double *w;
  for (i = 0; i < g_ntw; ++i)
    if (g_tw[i].n == n && g_tw[i].sign == sign) return g_tw[i].w;
  w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (i = 0; i < n; ++i) {
    double ang = 2.0 * acos(-1.0) * (double)i / (double)n;
    w[2 * i] = cos(ang);
    w[2 * i + 1] = sign * sin(ang);
  }
#ifdef _OPENMP
#pragma omp critical(qxo_tw)
#endif
  {
    if (g_ntw < NTW) { g_tw[g_ntw].n = n; g_tw[g_ntw].sign = sign; g_tw[g_ntw].w = w; ++g_ntw; }
  }
  return w;
}

static int pick_radix(int n)
{
  int p;
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

/* Recursive decimation-in-time DFT of length n with element stride is;
 * out is contiguous.  w is the exp(sign*2*pi*i*k/N0) table. */
static void fft_rec(int n, const double *in, int is, double *out, const double *w, int N0)
{
  int p, m, r, k, q;
  double t[2 * 64];
  if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
  p = pick_radix(n);
  m = n / p;
  for (r = 0; r < p; ++r) fft_rec(m, in + 2 * (size_t)r * is, is * p, out + 2 * (size_t)r * m, w, N0);
  if (p > 64) {
    /* large prime: plain DFT through a heap buffer */
    double *tt = (double *)malloc(sizeof(double) * 2 * (size_t)p);
    for (k = 0; k < m; ++k) {
      for (r = 0; r < p; ++r) {
        long e = ((long)(N0 / n) * r * k) % N0;
        double wr = w[2 * e], wi = w[2 * e + 1];
        double xr = out[2 * (r * m + k)], xi = out[2 * (r * m + k) + 1];
        tt[2 * r] = xr * wr - xi * wi;
        tt[2 * r + 1] = xr * wi + xi * wr;
      }
      for (q = 0; q < p; ++q) {
        double sr = 0.0, si = 0.0;
        for (r = 0; r < p; ++r) {
          long e = ((long)(N0 / p) * ((r * q) % p)) % N0;
          sr += tt[2 * r] * w[2 * e] - tt[2 * r + 1] * w[2 * e + 1];
          si += tt[2 * r] * w[2 * e + 1] + tt[2 * r + 1] * w[2 * e];
        }
        out[2 * (q * m + k)] = sr;
        out[2 * (q * m + k) + 1] = si;
      }
    }
    free(tt);
    return;
  }
  for (k = 0; k < m; ++k) {
    for (r = 0; r < p; ++r) {
      long e = ((long)(N0 / n) * r * k) % N0;
      double wr = w[2 * e], wi = w[2 * e + 1];
      double xr = out[2 * (r * m + k)], xi = out[2 * (r * m + k) + 1];
      t[2 * r] = xr * wr - xi * wi;
      t[2 * r + 1] = xr * wi + xi * wr;
    }
    for (q = 0; q < p; ++q) {
      double sr = 0.0, si = 0.0;
      for (r = 0; r < p; ++r) {
        long e = ((long)(N0 / p) * ((r * q) % p)) % N0;
        sr += t[2 * r] * w[2 * e] - t[2 * r + 1] * w[2 * e + 1];
        si += t[2 * r] * w[2 * e + 1] + t[2 * r + 1] * w[2 * e];
      }
      out[2 * (q * m + k)] = sr;
      out[2 * (q * m + k) + 1] = si;
    }
  }
}

/* Transform all lines along one axis of a complex grid. */
static void fft_axis(double *xy, int n1, int n2, int n3, int axis, int sign)
{
  int len = axis == 0 ? n1 : (axis == 1 ? n2 : n3);
  long stride = axis == 0 ? 1 : (axis == 1 ? n1 : (long)n1 * n2);
  long nlines = (long)n1 * n2 * n3 / len;
  const double *w = twiddles(len, sign);
  long l;
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    double *a = (double *)malloc(sizeof(double) * 4 * (size_t)len);
    double *b = a + 2 * len;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (l = 0; l < nlines; ++l) {
      long base;
      int t;
      if (axis == 0) base = l * n1;
      else if (axis == 1) base = (l % n1) + (l / n1) * (long)n1 * n2;
      else base = l;
      for (t = 0; t < len; ++t) {
        a[2 * t] = xy[2 * (base + t * stride)];
        a[2 * t + 1] = xy[2 * (base + t * stride) + 1];
      }
      fft_rec(len, a, 1, b, w, len);
      for (t = 0; t < len; ++t) {
        xy[2 * (base + t * stride)] = b[2 * t];
        xy[2 * (base + t * stride) + 1] = b[2 * t + 1];
      }
    }
    free(a);
  }
}

/* Unnormalised 3-D complex DFT, sign -1 forward / +1 backward, x fastest. */
void qxo_cfft3(int sign, double *xy, int n1, int n2, int n3)
{
  /* build the twiddle tables outside the parallel regions */
  twiddles(n1, sign); twiddles(n2, sign); twiddles(n3, sign);
  fft_axis(xy, n1, n2, n3, 0, sign);
  fft_axis(xy, n1, n2, n3, 1, sign);
  fft_axis(xy, n1, n2, n3, 2, sign);
}

/* rfft3 (Sources/fft3.f90:113-232).
 * key=1: r2c of fft3x, x1/N, full Hermitian grid written back (Re->x, Im->y),
 *        the upper x-half copied from the mirror points exactly as 186-205.
 * else : c2r from the half grid i < n1/2+1 of (fft3x, fft3y), unnormalised,
 *        real result in fft3x and fft3y cleared. */
void qxo_rfft3(int key, double *fft3x, double *fft3y, int n1, int n2, int n3)
{
  long N = (long)n1 * n2 * n3, t;
  int n1h = n1 / 2 + 1, i, j, k;
  double *z = (double *)malloc(sizeof(double) * 2 * (size_t)N);
  if (key == 1) {
    double fin = 1.0 / (double)N;
    for (t = 0; t < N; ++t) { z[2 * t] = fft3x[t]; z[2 * t + 1] = 0.0; }
    qxo_cfft3(-1, z, n1, n2, n3);
    for (k = 0; k < n3; ++k)
      for (j = 0; j < n2; ++j)
        for (i = 0; i < n1h; ++i) {
          long ijk = i + (long)n1 * (j + (long)n2 * k);
          fft3x[ijk] = z[2 * ijk] * fin;
          fft3y[ijk] = z[2 * ijk + 1] * fin;
        }
    for (k = 0; k < n3; ++k) {
      int mk = (n3 - k) % n3;
      for (j = 0; j < n2; ++j) {
        int mj = (n2 - j) % n2;
        for (i = n1h; i < n1; ++i) {
          long ijk = i + (long)n1 * (j + (long)n2 * k);
          long mijk = (n1 - i) + (long)n1 * (mj + (long)n2 * mk);
          fft3x[ijk] = fft3x[mijk];
          fft3y[ijk] = -fft3y[mijk];
        }
      }
    }
  } else {
    for (k = 0; k < n3; ++k) {
      int mk = (n3 - k) % n3;
      for (j = 0; j < n2; ++j) {
        int mj = (n2 - j) % n2;
        for (i = 0; i < n1; ++i) {
          long ijk = i + (long)n1 * (j + (long)n2 * k);
          if (i < n1h) {
            z[2 * ijk] = fft3x[ijk];
            z[2 * ijk + 1] = fft3y[ijk];
          } else {
            long mijk = (n1 - i) + (long)n1 * (mj + (long)n2 * mk);
            z[2 * ijk] = fft3x[mijk];
            z[2 * ijk + 1] = -fft3y[mijk];
          }
        }
      }
    }
    qxo_cfft3(+1, z, n1, n2, n3);
    for (t = 0; t < N; ++t) { fft3x[t] = z[2 * t]; fft3y[t] = 0.0; }
  }
  free(z);
}

/* fft3, complex (Sources/fft3.f90:237-333): key=1 forward x1/N, else backward. */
void qxo_fft3(int key, double *fft3x, double *fft3y, int n1, int n2, int n3)
{
  long N = (long)n1 * n2 * n3, t;
  double *z = (double *)malloc(sizeof(double) * 2 * (size_t)N);
  for (t = 0; t < N; ++t) { z[2 * t] = fft3x[t]; z[2 * t + 1] = fft3y[t]; }
  qxo_cfft3(key == 1 ? -1 : +1, z, n1, n2, n3);
  if (key == 1) {
    double fin = 1.0 / (double)N;
    for (t = 0; t < N; ++t) { z[2 * t] *= fin; z[2 * t + 1] *= fin; }
  }
  for (t = 0; t < N; ++t) { fft3x[t] = z[2 * t]; fft3y[t] = z[2 * t + 1]; }
  free(z);
}

/* ======================================================================= */
/* Local H|psi> and kinetic                                               */
/* ======================================================================= */

/* wvg2r (Sources/engrad.f90:1492-1559): psi(G) half sphere -> psi(r) in fft3x. */
void qxo_wvg2r(const double *cgjr, int nplw, const int *ijkgd_c, int n1, int n2, int n3,
               double *fft3x, double *fft3y)
{
  long N = (long)n1 * n2 * n3;
  int ig;
  memset(fft3x, 0, sizeof(double) * (size_t)N);
  memset(fft3y, 0, sizeof(double) * (size_t)N);
  fft3x[ijkgd_c[0] - 1] = cgjr[0];
  fft3y[ijkgd_c[0] - 1] = 0.0;
  for (ig = 1; ig <= nplw; ++ig) {
    long ijk = ijkgd_c[ig] - 1;
    fft3x[ijk] = cgjr[2 * ig];
    fft3y[ijk] = cgjr[2 * ig + 1];
  }
  for (ig = 1; ig <= nplw; ++ig) {
    long ijk = ijkgd_c[-ig] - 1;
    fft3x[ijk] = cgjr[2 * ig];
    fft3y[ijk] = -cgjr[2 * ig + 1];
  }
  qxo_rfft3(2, fft3x, fft3y, n1, n2, n3);
}

/* hcnonloc, VECTOR build, no real-space KB (Sources/engrad.f90:411-606):
 * rhcr += gather_G[ rfft3_fwd( glocal * psi(r) ) ]. */
void qxo_hcnonloc(const double *glocal, const double *cgjr, double *rhcr, int nplw,
                  const int *ijkgd_c, int n1, int n2, int n3, double *fft3x, double *fft3y)
{
  long N = (long)n1 * n2 * n3, t;
  int ig;
  qxo_wvg2r(cgjr, nplw, ijkgd_c, n1, n2, n3, fft3x, fft3y);
  for (t = 0; t < N; ++t) { fft3x[t] = glocal[t] * fft3x[t]; fft3y[t] = 0.0; }
  qxo_rfft3(1, fft3x, fft3y, n1, n2, n3);
  for (ig = 0; ig <= nplw; ++ig) {
    long ijk = ijkgd_c[ig] - 1;
    rhcr[2 * ig] += fft3x[ijk];
    rhcr[2 * ig + 1] += fft3y[ijk];
  }
}

/* ekinetic2 (Sources/engrad.f90:66-127): rhcr = dkgnrm*cgjr, per-band 2<T>. */
void qxo_ekinetic(const double *dkgnrm, const double *cgjr, double *rhcr, int nplwex, int nb,
                  const double *occ, double *ekib, double *ekint)
{
  int i, ig;
  double ekin = 0.0;
  for (i = 0; i < nb; ++i) {
    const double *c = cgjr + (size_t)nplwex * i;
    double *r = rhcr + (size_t)nplwex * i, e = 0.0;
    for (ig = 0; ig < nplwex; ++ig) r[ig] = dkgnrm[ig] * c[ig];
    for (ig = 0; ig < nplwex; ++ig) e += r[ig] * c[ig];
    e = 2.0 * e - r[0] * c[0];
    if (occ) ekin += occ[i] * e;
    if (ekib) ekib[i] = 2.0 * e;
  }
  if (ekint) *ekint += ekin;
}

/* ======================================================================= */
/* Kleinman-Bylander                                                       */
/* ======================================================================= */

/* hckbpp_g with lbigmm2 (Sources/nlkbpp.f90:1284-1393):
 *   slmir(l,ia,ib) = 2*sum_{k>=3} eqlmr_(k,l,ia) psi(k) + eqlmr_(1,l,ia) psi(1)
 *   rhcr(:,ib)    += vclalp(l) * eqlmr_(:,l,ia) * slmir(l,ia,ib)
 * Note the Im(G=0) element k=2 is excluded from the sum. */
void qxo_hckbpp_g(const double *eqlmr, int nplwex, int lmmax, int natom, const int *lmx_atom,
                  const double *vclalp_atom, const double *cgjr, double *rhcr, int nb,
                  double *slmir)
{
  int ia, l, ib, ig;
#ifdef _OPENMP
#pragma omp parallel for private(ia, l, ig) schedule(static)
#endif
  for (ib = 0; ib < nb; ++ib) {
    const double *c = cgjr + (size_t)nplwex * ib;
    double *r = rhcr + (size_t)nplwex * ib;
    for (ia = 0; ia < natom; ++ia)
      for (l = 0; l < lmx_atom[ia]; ++l) {
        const double *e = eqlmr + (size_t)nplwex * (l + (size_t)lmmax * ia);
        double v = vclalp_atom[l + lmmax * ia], t = 0.0;
        for (ig = 2; ig < nplwex; ++ig) t += e[ig] * c[ig];
        t = 2.0 * t + e[0] * c[0];
        if (slmir) slmir[l + (size_t)lmmax * (ia + (size_t)natom * ib)] = t;
        for (ig = 0; ig < nplwex; ++ig) r[ig] += (e[ig] * v) * t;
      }
  }
}

/* stscfn_pw (Sources/pwlda.f90:5711-5854): structure factors
 *   ycos/ysin(ig, ia) = cos/sin(G_ig . tau_ia)
 * by the reference's 1-D recurrences.  Per atom and axis a (5754-5834):
 *   phase = ratm(a, ia) * fact_a, fact_a = 2 pi nd1v(a)/nd1vks(a) (periodic)
 *   or 2 pi (1/hcell(a,a)) nd1v(a)/nd1vks(a) (cluster);
 *   c_0 = 1, s_0 = 0, c_1 + i s_1 = cos + i sin(phase),
 *   c_k = c_1 c_{k-1} - s_1 s_{k-1}, s_k = s_1 c_{k-1} + c_1 s_{k-1},
 *   c_{-k} = c_k, s_{-k} = -s_k;
 * then per G (5836-5848): (c1 + i s1)(c2 + i s2) first, times (c3 + i s3).
 * ycos, ysin: (nplwcs + 1) x natom, column-major; entries past
 * min(nplwcs, nplw5) are left as they are, like the reference. */
void qxo_stscfn_pw(const double *ratm, const double *hcell, const int *nd1v, const int *nd1vks,
                   int lclust, int natom, int kmax1, int kmax2, int kmax3, int nplwcs, int nplw5,
                   const int *nga, const int *ngb, const int *ngc, double *ycos, double *ysin)
{
  const double dpi = 3.141592653589793 * 2.0;
  const int kmax[3] = {kmax1, kmax2, kmax3};
  double fact[3];
  int a, ia, k, ig;
  for (a = 0; a < 3; ++a) {
    if (lclust)
      fact[a] = 1.0 / hcell[4 * a] * (double)nd1v[a] / (double)nd1vks[a];
    else
      fact[a] = (double)nd1v[a] / (double)nd1vks[a];
    fact[a] = fact[a] * dpi;
  }
  const int ngmax = nplwcs < nplw5 ? nplwcs : nplw5;
  for (ia = 0; ia < natom; ++ia) {
    double *dc[3], *ds[3];
    for (a = 0; a < 3; ++a) {
      const int km = kmax[a];
      double *c = (double *)malloc(sizeof(double) * (2 * km + 1)) + km;
      double *sn = (double *)malloc(sizeof(double) * (2 * km + 1)) + km;
      const double ph = ratm[a + 3 * ia] * fact[a];
      c[0] = 1.0;
      sn[0] = 0.0;
      if (km >= 1) {
        c[1] = cos(ph);
        sn[1] = sin(ph);
        c[-1] = c[1];
        sn[-1] = -sn[1];
      }
      for (k = 2; k <= km; ++k) {
        c[k] = c[1] * c[k - 1] - sn[1] * sn[k - 1];
        sn[k] = sn[1] * c[k - 1] + c[1] * sn[k - 1];
        c[-k] = c[k];
        sn[-k] = -sn[k];
      }
      dc[a] = c;
      ds[a] = sn;
    }
    double *yc = ycos + (size_t)(nplwcs + 1) * ia, *ys = ysin + (size_t)(nplwcs + 1) * ia;
    yc[0] = 1.0;
    ys[0] = 0.0;
    for (ig = 1; ig <= ngmax; ++ig) {
      const int k1 = nga[ig], k2 = ngb[ig], k3 = ngc[ig];
      const double c12 = dc[0][k1] * dc[1][k2] - ds[0][k1] * ds[1][k2];
      const double s12 = ds[0][k1] * dc[1][k2] + dc[0][k1] * ds[1][k2];
      yc[ig] = c12 * dc[2][k3] - s12 * ds[2][k3];
      ys[ig] = s12 * dc[2][k3] + c12 * ds[2][k3];
    }
    for (a = 0; a < 3; ++a) {
      free(dc[a] - kmax[a]);
      free(ds[a] - kmax[a]);
    }
  }
}

/* setnlc_g (Sources/nlkbpp.f90:1236-1279) with setbgq (2135-2192) and setqlm
 * (2197-2247), Gamma point (nk = ikng = 1):
 *   vclalp(l, it) = clalp(l, it) / rvol for the KB types (lkbppi, not lking);
 *   per type it, l = 1..lmx(it), G index ig = 0..nplw:
 *     g = sqrt(recnrm(ig)) (abgvec, Sources/ppkb.f90:785-787),
 *     m = 2 int(g / (2 dlqlm(it))), d = (g / dlqlm(it) - m) / 2,
 *     z = d ((d - 1) tbqlma(m, L, it) + tbqlm(m+2, L, it) - tbqlm(m, L, it))
 *         + tbqlm(m, L, it)                    with L = itola(l, it),
 *     z *= qlmr(ig, lm) if itolmf(l, it) == 1, else qlmi(ig, lm), lm = itolm(l, it);
 *   and for the atoms ia = nhk1(it)..nhk2(it):
 *     litoll(l, it): eqlmr_(2ig+1, l, ia) = ycos(ig, ia) z, eqlmr_(2ig+2, l, ia) = -ysin(ig, ia) z
 *     otherwise:     eqlmr_(2ig+1, l, ia) = -ysin(ig, ia) z, eqlmr_(2ig+2, l, ia) = -ycos(ig, ia) z.
 * Arrays column-major with the reference's shapes: tbqlm/tbqlma(0:mxqlm, lnref,
 * ntype), itola/itolm/itolmf/litoll(lmmax, ntype), clalp/vclalp(lmmax, ntype),
 * qlmr/qlmi(0:nplw, nqlmmx), ycos/ysin(0:nplwcs, natom), eqlmr(2(nplw+1),
 * lmmax, natom).  nhk1/nhk2 are 1-based atom indices. */
void qxo_setnlc_g(int nplw, int ntype, const int *nhk1, const int *nhk2, int natom,
                  const int *lkbppi, const int *lking, const double *ycos, const double *ysin,
                  int nplwcs, double rvol, int lmmax, int lnref, int mxqlm, const int *lmx,
                  const double *clalp, const double *tbqlm, const double *tbqlma,
                  const double *dlqlm, const int *itola, const int *itolm, const int *itolmf,
                  const int *litoll, int nqlmmx, const double *qlmr, const double *qlmi,
                  const double *recnrm, double *eqlmr, double *vclalp)
{
  const size_t nex = 2 * (size_t)(nplw + 1), ntb = (size_t)mxqlm + 1;
  double *z = (double *)malloc(sizeof(double) * (nplw + 1));
  int it, l, ia, ig;
  (void)natom;
  for (it = 0; it < ntype; ++it) {
    if (!lkbppi[it] || lking[it]) continue;
    for (l = 0; l < lmx[it]; ++l) vclalp[l + (size_t)lmmax * it] = clalp[l + (size_t)lmmax * it] / rvol;
    for (l = 0; l < lmx[it]; ++l) {
      const int L = itola[l + lmmax * it], lm = itolm[l + lmmax * it];
      const int lmf = itolmf[l + lmmax * it], par = litoll[l + lmmax * it];
      const double *tb = tbqlm + ntb * ((size_t)(L - 1) + (size_t)lnref * it);
      const double *ta = tbqlma + ntb * ((size_t)(L - 1) + (size_t)lnref * it);
      const double *q = (lmf == 1 ? qlmr : qlmi) + (size_t)(nplw + 1) * (lm - 1);
      const double dl = dlqlm[it];
      for (ig = 0; ig <= nplw; ++ig) {
        const double g = sqrt(recnrm[ig]);
        int m = (int)(g / (2.0 * dl));
        m = 2 * m;
        const double d = 0.5 * (g / dl - (double)m);
        const double zz = d * ((d - 1.0) * ta[m] + tb[m + 2] - tb[m]) + tb[m];
        z[ig] = zz * q[ig];
      }
      (void)nqlmmx;
      for (ia = nhk1[it] - 1; ia < nhk2[it]; ++ia) {
        double *e = eqlmr + nex * ((size_t)l + (size_t)lmmax * ia);
        const double *yc = ycos + (size_t)(nplwcs + 1) * ia, *ys = ysin + (size_t)(nplwcs + 1) * ia;
        for (ig = 0; ig <= nplw; ++ig) {
          if (par) {
            e[2 * ig] = yc[ig] * z[ig];
            e[2 * ig + 1] = -ys[ig] * z[ig];
          } else {
            e[2 * ig] = -ys[ig] * z[ig];
            e[2 * ig + 1] = -yc[ig] * z[ig];
          }
        }
      }
    }
  }
  free(z);
}

/* kbnlpp_frc2 with lbigmm2, no stress (Sources/nlkbpp.f90:1878-2060, VECTOR
 * branch 1980-2030):
 *   fsl(l, ia, ib) = sum_{ig=1..nplw} (e(2ig+1) psi(2ig+2) - e(2ig+2) psi(2ig+1)) G_ig
 *   fnlc(:, ia)   += 2 sum_{l, ib} fsl(l, ia, ib) slmir(l, ia, ib) 2 occ(ib) vclalp(l)
 * for the three Cartesian components of G (gx, gy, gz over G index 0..nplw);
 * slmir as qxo_hckbpp_g leaves it.  fnlc (natom x 3, row-major) is added to. */
void qxo_kbnlpp_frc2(const double *eqlmr, int nplwex, int lmmax, int natom, const int *lmx_atom,
                     const double *vclalp_atom, const double *cgjr, int nb, const double *occ,
                     const double *slmir, const double *gx, const double *gy, const double *gz,
                     double *fnlc)
{
  const int nplw = nplwex / 2 - 1;
  int ia, l, ib, ig;
  for (ia = 0; ia < natom; ++ia) {
    double f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (l = 0; l < lmx_atom[ia]; ++l) {
      const double *e = eqlmr + (size_t)nplwex * (l + (size_t)lmmax * ia);
      for (ib = 0; ib < nb; ++ib) {
        const double *c = cgjr + (size_t)nplwex * ib;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (ig = 1; ig <= nplw; ++ig) {
          const double z = e[2 * ig] * c[2 * ig + 1] - e[2 * ig + 1] * c[2 * ig];
          sx += z * gx[ig];
          sy += z * gy[ig];
          sz += z * gz[ig];
        }
        const double w = slmir[l + (size_t)lmmax * (ia + (size_t)natom * ib)] *
                         (vclalp_atom[l + lmmax * ia] * occ[ib] * 2.0);
        f1 += sx * w;
        f2 += sy * w;
        f3 += sz * w;
      }
    }
    fnlc[3 * ia] += 2.0 * f1;
    fnlc[3 * ia + 1] += 2.0 * f2;
    fnlc[3 * ia + 2] += 2.0 * f3;
  }
}

/* ======================================================================= */
/* SCF glue: LDA exchange-correlation and the Hartree potential              */
/* ======================================================================= */

/* VXC (Sources/vxc.f90:2063-2104): Ceperley-Alder as parametrised by Perdew
 * and Zunger, Hartree units.  B2, B3, C2..C4 are default-real (single
 * precision) constants in the reference's DATA statement, so they carry the
 * float rounding of their decimal values. */
static void pz_vxc(double rs, double *ex, double *vx, double *ec, double *vc)
{
  const double A1 = -0.4582, B1 = -0.1423, B2 = (double)1.0529f, B3 = (double)0.3334f;
  const double C1 = -0.0480, C2 = (double)0.0311f, C3 = (double)-0.0116f, C4 = (double)0.0020f;
  *ex = A1 / rs;
  *vx = 4.0 * *ex / 3.0;
  if (rs >= 1.0) {
    const double bunb = 1.0 + B2 * sqrt(rs) + B3 * rs;
    *ec = B1 / bunb;
    *vc = *ec * (1.0 + (B2 * sqrt(rs) / 6.0 + B3 * rs / 3.0) / bunb);
  } else {
    *ec = C1 + C2 * log(rs) + C3 * rs + C4 * rs * log(rs);
    *vc = *ec - (C2 + (C3 + C4) * rs + C4 * rs * log(rs)) / 3.0;
  }
}

/* xclda (Sources/vxc.f90:961-1017): per mesh point rhot = rho + rhocore
 * (rhocore < 0 or NULL: rho only); rhot < 1e-30 gives 0; else
 * rs = (3 / (pi4 rhot))^(1/3) with pi4 = 12.56637061436 as in the reference,
 * vexc = 2 (vx + vc), eeexc = 2 (ex + ec) (Ry); e4 = {eexc, evexc, ecor, evcor}
 * = 2 rdelv * {sum ex rhot, sum vx rho, sum ec rhot, sum vc rho}. */
void qxo_xclda(const double *rho, const double *rhocore, int n, double rdelv, double *vexc,
               double *eeexc, double *e4)
{
  const double pi4 = 12.56637061436;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int m;
  for (m = 0; m < n; ++m) {
    const double rhot = (rhocore && rhocore[m] >= 0.0) ? rho[m] + rhocore[m] : rho[m];
    if (rhot < 1e-30) {
      vexc[m] = 0.0;
      eeexc[m] = 0.0;
      continue;
    }
    const double rs = pow(3.0 / (pi4 * rhot), 1.0 / 3.0);
    double ex, vx, ec, vc;
    pz_vxc(rs, &ex, &vx, &ec, &vc);
    vexc[m] = 2.0 * (vx + vc);
    eeexc[m] = 2.0 * (ex + ec);
    s0 += ex * rhot;
    s1 += vx * rho[m];
    s2 += ec * rhot;
    s3 += vc * rho[m];
  }
  e4[0] = s0 * 2.0 * rdelv;
  e4[1] = s1 * 2.0 * rdelv;
  e4[2] = s2 * 2.0 * rdelv;
  e4[3] = s3 * 2.0 * rdelv;
}

/* harting (Sources/engrad.f90:1673-1743): the density rhgr on the G sphere
 * (index 0..nplw5, real-packed) times 2 recnrmex(ig) onto the full grid at
 * ijkgd(ig) and its conjugate at ijkgd(-ig) (G = 0: real part only), then the
 * c2r transform of rfft3 (key = inverse).  ijkgd_c points at ijkgd(0) of the
 * array ijkgd(-nplw5:nplw5) (1-based grid indices); vhar (n1 n2 n3) in FFT
 * order. */
void qxo_harting(const double *rhgr, int nplw5, const double *recnrmex, const int *ijkgd_c,
                 int n1, int n2, int n3, double *vhar)
{
  const size_t N = (size_t)n1 * n2 * n3;
  double *fx = (double *)calloc(N, sizeof(double)), *fy = (double *)calloc(N, sizeof(double));
  int ig;
  fx[ijkgd_c[0] - 1] = 2.0 * recnrmex[0] * rhgr[0];
  /* the VECTOR order (1720-1732): all +G first, then all -G */
  for (ig = 1; ig <= nplw5; ++ig) {
    const double p4 = 2.0 * recnrmex[ig];
    const int ijk = ijkgd_c[ig] - 1;
    fx[ijk] = p4 * rhgr[2 * ig];
    fy[ijk] = p4 * rhgr[2 * ig + 1];
  }
  for (ig = 1; ig <= nplw5; ++ig) {
    const double p4 = 2.0 * recnrmex[ig];
    const int ijkm = ijkgd_c[-ig] - 1;
    fx[ijkm] = p4 * rhgr[2 * ig];
    fy[ijkm] = -(p4 * rhgr[2 * ig + 1]);
  }
  qxo_rfft3(-1, fx, fy, n1, n2, n3);
  memcpy(vhar, fx, sizeof(double) * N);
  free(fx);
  free(fy);
}

/* chgr2g (Sources/engrad.f90:1253-1297): rho on the grid (FFT order) -> the
 * forward rfft3 (x 1/N) -> rhgr(2ig..2ig+1) = value at ijkgd(ig), ig = 0..nplw5. */
void qxo_chgr2g(const double *rho, int nplw5, const int *ijkgd_c, int n1, int n2, int n3,
                double *rhgr)
{
  const size_t N = (size_t)n1 * n2 * n3;
  double *fx = (double *)malloc(N * sizeof(double)), *fy = (double *)calloc(N, sizeof(double));
  int ig;
  memcpy(fx, rho, N * sizeof(double));
  qxo_rfft3(1, fx, fy, n1, n2, n3);
  for (ig = 0; ig <= nplw5; ++ig) {
    const int ijk = ijkgd_c[ig] - 1;
    rhgr[2 * ig] = fx[ijk];
    rhgr[2 * ig + 1] = fy[ijk];
  }
  free(fx);
  free(fy);
}

/* chgg2r (Sources/engrad.f90:1302-1353): rhgr at ijkgd(ig), ig = 0..nplw5, then
 * its conjugate at ijkgd(-ig), ig = 1..nplw5, and the inverse rfft3. */
void qxo_chgg2r(const double *rhgr, int nplw5, const int *ijkgd_c, int n1, int n2, int n3,
                double *rho)
{
  const size_t N = (size_t)n1 * n2 * n3;
  double *fx = (double *)calloc(N, sizeof(double)), *fy = (double *)calloc(N, sizeof(double));
  int ig;
  for (ig = 0; ig <= nplw5; ++ig) {
    const int ijk = ijkgd_c[ig] - 1;
    fx[ijk] = rhgr[2 * ig];
    fy[ijk] = rhgr[2 * ig + 1];
  }
  for (ig = 1; ig <= nplw5; ++ig) {
    const int ijk = ijkgd_c[-ig] - 1;
    fx[ijk] = rhgr[2 * ig];
    fy[ijk] = -rhgr[2 * ig + 1];
  }
  qxo_rfft3(-1, fx, fy, n1, n2, n3);
  memcpy(rho, fx, N * sizeof(double));
  free(fx);
  free(fy);
}

/* DETMAT (Sources/funcs.f90:66-128): determinant by row scaling and Gaussian
 * elimination with a row-permutation vector; w (column-major, leading
 * dimension mx) is overwritten. */
static double detmat(double *w, int n, int *ip, int mx)
{
  const double dzero = 1e-30;
  double fugo = 1.0, det = 1.0;
  int i, j, j2, jpv = 0;
#define W_(r, c) w[(r) + (size_t)mx * (c)]
  for (i = 0; i < n; ++i) ip[i] = i;
  for (i = 0; i < n; ++i) {
    double pivot = 0.0;
    for (j = 0; j < n; ++j) pivot = fmax(pivot, fabs(W_(i, j)));
    if (pivot < dzero) return 0.0;
    for (j = 0; j < n; ++j) W_(i, j) = W_(i, j) / pivot;
    det = det * pivot;
  }
  for (i = 0; i < n - 1; ++i) {
    double pivot = 0.0;
    int ipv;
    for (j = i; j < n; ++j) {
      const int jj = ip[j];
      if (fabs(W_(jj, i)) > pivot) {
        pivot = fabs(W_(jj, i));
        jpv = j;
      }
    }
    if (pivot < dzero) return 0.0;
    if (jpv != i) {
      const int t = ip[jpv];
      ip[jpv] = ip[i];
      ip[i] = t;
      fugo = fugo * (-1.0);
    }
    ipv = ip[i];
    for (j = i + 1; j < n; ++j) {
      const int jj = ip[j];
      const double dmj = W_(jj, i) / W_(ipv, i);
      W_(jj, i) = 0.0;
      for (j2 = i + 1; j2 < n; ++j2) W_(jj, j2) = W_(jj, j2) - dmj * W_(ipv, j2);
    }
  }
  for (i = 0; i < n; ++i) det = det * W_(ip[i], i);
#undef W_
  return det * fugo;
}

/* INVERS (Sources/funcs.f90:5-61): inverse by cofactors over determinants of
 * the matrix scaled by its largest element.  w, wi: n x n in arrays of
 * leading dimension mx.  Returns ierr (1 = singular). */
int qxo_invers(const double *w, double *wi, int n, int mx)
{
  double *work = (double *)malloc(sizeof(double) * (size_t)mx * mx);
  int *ip = (int *)malloc(sizeof(int) * (size_t)mx);
  double elmax = 0.0, det;
  int i, j, ii, jj;
  for (i = 0; i < n; ++i)
    for (j = 0; j < n; ++j) elmax = fmax(elmax, fabs(w[i + (size_t)mx * j]));
  for (i = 0; i < n; ++i)
    for (j = 0; j < n; ++j) work[i + (size_t)mx * j] = w[i + (size_t)mx * j] / elmax;
  det = detmat(work, n, ip, mx);
  if (fabs(det) < 1e-30) {
    free(work);
    free(ip);
    return 1;
  }
  for (i = 0; i < n; ++i)
    for (j = 0; j < n; ++j) {
      int isum = 0;
      double detij;
      for (ii = 0; ii < n; ++ii) {
        int jsum = 0;
        if (ii == j) continue;
        for (jj = 0; jj < n; ++jj) {
          if (jj == i) continue;
          work[isum + (size_t)mx * jsum] = w[ii + (size_t)mx * jj] / elmax;
          ++jsum;
        }
        ++isum;
      }
      detij = detmat(work, n - 1, ip, mx);
      /* (-1)**(I+J) with 1-based I, J */
      wi[i + (size_t)mx * j] = (((i + j) % 2) ? -1.0 : 1.0) * detij / det;
    }
  for (i = 0; i < n; ++i)
    for (j = 0; j < n; ++j) wi[i + (size_t)mx * j] = wi[i + (size_t)mx * j] / elmax;
  free(work);
  free(ip);
  return 0;
}

/* mxchgg_pulay (Sources/chgdns.f90:954-1150), the mixing part on node 0 (the
 * chgg2r at its end is qxo_chgg2r).  State of module mxchgg_variables passed
 * explicitly: kntchg, unr/dfmr (nplwcex x itratn, column-major), aij (itratn x
 * itratn).  rhgr (in: output density, out: mixed density in its first nplwcex
 * entries, the rest untouched) and rinr (in: input density; out: = rhgr over
 * nplw5ex).  cmk_out (itratn) receives the Pulay coefficients of the last
 * successful combination (zeros on a Kerker step).  Returns the number of
 * history resets taken (INVERS failures and the small-amax trap). */
int qxo_mxchgg_pulay(double *rhgr, double *rinr, int nplwcex, int nplw5ex, const double *dkgnrm,
                     double rvol, double aslh, double bslh, double factq1, double wkerker,
                     int lreset, int lmixhold, int itratn, int *kntchg_io, double *unr,
                     double *dfmr, double *aij, double *cmk_out)
{
  const double small = 1e-15;
  const size_t ld = (size_t)nplwcex;
  int kntchg = *kntchg_io + 1, ierr, nreset = 0, ig, i1, ii, jj, i, j;
  double *betakl = (double *)calloc((size_t)itratn * itratn, sizeof(double));
  double *cmk = (double *)calloc((size_t)itratn, sizeof(double));
#define A_(r, c) aij[(r) + (size_t)itratn * (c)]
#define B_(r, c) betakl[(r) + (size_t)itratn * (c)]
  for (i = 0; i < itratn; ++i) cmk_out[i] = 0.0;
  if (kntchg > itratn || lreset) kntchg = 1;
  for (;;) {
    const int k = kntchg - 1;
    for (ig = 0; ig < nplwcex; ++ig) {
      unr[ig + ld * k] = rinr[ig];
      dfmr[ig + ld * k] = rhgr[ig] - rinr[ig];
    }
    for (i1 = 0; i1 < kntchg; ++i1) {
      double dfmsq = 0.0, factq;
      for (ig = 2; ig < nplwcex; ++ig) {
        factq = (dkgnrm[ig] + factq1) / dkgnrm[ig];
        dfmsq = dfmsq + dfmr[ig + ld * i1] * dfmr[ig + ld * k] * factq;
      }
      factq = (dkgnrm[2] + factq1) / dkgnrm[2];
      dfmsq = 2.0 * dfmsq + dfmr[ld * i1] * dfmr[ld * k] * factq;
      dfmsq = dfmsq * rvol;
      A_(i1, k) = dfmsq;
      A_(k, i1) = dfmsq;
    }
    ierr = 0;
    if (kntchg == 1) {
      for (ig = 0; ig < nplwcex; ++ig) {
        double f = dkgnrm[ig] / (dkgnrm[ig] + bslh);
        f = aslh * (1.0 - wkerker + wkerker * f);
        rhgr[ig] = rinr[ig] + f * dfmr[ig + ld * k];
      }
    } else {
      const int n = kntchg, icntr = n / 2 + n % 2;
      const double amax = fabs(A_(icntr - 1, icntr - 1));
      if (amax < small) {
        kntchg = 1;
        ++nreset;
        continue;
      }
      for (j = 0; j < n; ++j)
        for (i = 0; i < n; ++i) A_(i, j) = A_(i, j) / amax;
      ierr = qxo_invers(aij, betakl, n, itratn);
      if (ierr == 0) {
        double bbnb = 0.0;
        for (j = 0; j < n; ++j)
          for (i = 0; i < n; ++i) A_(i, j) = A_(i, j) * amax;
        for (j = 0; j < n; ++j)
          for (i = 0; i < n; ++i) B_(i, j) = B_(i, j) / amax;
        for (ii = 0; ii < n; ++ii) {
          cmk[ii] = 0.0;
          for (jj = 0; jj < n; ++jj) {
            bbnb = bbnb + B_(ii, jj);
            cmk[ii] = cmk[ii] + B_(ii, jj);
          }
        }
        for (ii = 0; ii < n; ++ii) cmk[ii] = cmk[ii] / bbnb;
        for (ig = 0; ig < nplwcex; ++ig) rhgr[ig] = 0.0;
        for (ii = 0; ii < n; ++ii)
          for (ig = 0; ig < nplwcex; ++ig) {
            double f = dkgnrm[ig] / (dkgnrm[ig] + bslh);
            f = aslh * (1.0 - wkerker + wkerker * f);
            rhgr[ig] = rhgr[ig] + cmk[ii] * (unr[ig + ld * ii] + f * dfmr[ig + ld * ii]);
          }
        for (ii = 0; ii < n; ++ii) cmk_out[ii] = cmk[ii];
      } else {
        const int nhi = 1;
        ++nreset;
        kntchg = (kntchg - nhi > 1) ? kntchg - nhi : 1;
        for (ii = 0; ii < kntchg - 1; ++ii)
          for (ig = 0; ig < nplwcex; ++ig) {
            unr[ig + ld * ii] = unr[ig + ld * (ii + nhi)];
            dfmr[ig + ld * ii] = dfmr[ig + ld * (ii + nhi)];
          }
        for (j = 0; j < kntchg - 1; ++j)
          for (i = 0; i < kntchg - 1; ++i) A_(i, j) = A_(i + nhi, j + nhi) * amax;
      }
    }
    if (ierr == 0) break;
  }
  if (kntchg == itratn) {
    for (i = 0; i < kntchg - 1; ++i)
      for (ig = 0; ig < nplwcex; ++ig) {
        unr[ig + ld * i] = unr[ig + ld * (i + 1)];
        dfmr[ig + ld * i] = dfmr[ig + ld * (i + 1)];
      }
    for (j = 0; j < kntchg - 1; ++j)
      for (i = 0; i < kntchg - 1; ++i) A_(i, j) = A_(i + 1, j + 1);
    if (lmixhold) kntchg = kntchg - 1;
  }
#undef A_
#undef B_
  for (ig = 0; ig < nplw5ex; ++ig) rinr[ig] = rhgr[ig];
  *kntchg_io = kntchg;
  free(betakl);
  free(cmk);
  return nreset;
}

/* mxchgg_simple (Sources/chgdns.f90:1155-1210) and mxchgg_none (907-949),
 * mixing part: ph = aslh (simple) or 1 (none), rhgr = (1-ph) rinr + ph rhgr
 * over nplw5ex; then rinr = rhgr. */
void qxo_mxchgg_simple(double *rhgr, double *rinr, int nplw5ex, double aslh, int none)
{
  int ig;
  if (!none) {
    const double ph = aslh;
    for (ig = 0; ig < nplw5ex; ++ig) rhgr[ig] = (1.0 - ph) * rinr[ig] + ph * rhgr[ig];
  }
  for (ig = 0; ig < nplw5ex; ++ig) rinr[ig] = rhgr[ig];
}

/* flocal_g without stress (Sources/nlkbpp.f90:1037-1125), one plane-wave
 * rank: for the ions k = nhk1(it)..nhk2(it) of each type with llking(it)
 * false, atom ia = iatmpt(k), floc(:, ia) = 2 (floc(:, ia) + sum_{ig=1..nplw5}
 * svco(ig+1, it) (Re rhgr_G sin(G.tau) + Im rhgr_G cos(G.tau)) G), the phase
 * from the axis tables DCQ/DSQ (k = -kmax..kmax per atom) as 1091-1098 does.
 * nhk1/nhk2/iatmpt 1-based; svco (nplw5+1) x ntype column-major; dcq[a] and
 * dsq[a] (2 kmax_a + 1) x natom; floc 3 x natom. */
void qxo_flocal_g(int ntype, const int *nhk1, const int *nhk2, const int *iatmpt,
                  const int *llking, int natom, int nplw5, const int *nga, const int *ngb,
                  const int *ngc, const double *gx, const double *gy, const double *gz,
                  const double *rhgr, const double *svco, const int *kmax,
                  const double *const *dcq, const double *const *dsq, double *floc)
{
  int it, k, ig;
  for (it = 0; it < ntype; ++it) {
    if (llking[it]) continue;
    for (k = nhk1[it]; k <= nhk2[it]; ++k) {
      const int ia = iatmpt[k - 1] - 1;
      const double *c1 = dcq[0] + (size_t)ia * (2 * kmax[0] + 1) + kmax[0];
      const double *s1 = dsq[0] + (size_t)ia * (2 * kmax[0] + 1) + kmax[0];
      const double *c2 = dcq[1] + (size_t)ia * (2 * kmax[1] + 1) + kmax[1];
      const double *s2 = dsq[1] + (size_t)ia * (2 * kmax[1] + 1) + kmax[1];
      const double *c3 = dcq[2] + (size_t)ia * (2 * kmax[2] + 1) + kmax[2];
      const double *s3 = dsq[2] + (size_t)ia * (2 * kmax[2] + 1) + kmax[2];
      double b0 = 0.0, b1 = 0.0, b2 = 0.0;
      for (ig = 1; ig <= nplw5; ++ig) {
        const int k1 = nga[ig], k2 = ngb[ig], k3 = ngc[ig];
        const double dcc12 = c1[k1] * c2[k2] - s1[k1] * s2[k2];
        const double dsc12 = s1[k1] * c2[k2] + c1[k1] * s2[k2];
        const double zcosk = dcc12 * c3[k3] - dsc12 * s3[k3];
        const double zsink = dsc12 * c3[k3] + dcc12 * s3[k3];
        const double ftmp = svco[ig + (size_t)(nplw5 + 1) * it] *
                            (rhgr[2 * ig] * zsink + rhgr[2 * ig + 1] * zcosk);
        b0 = b0 + ftmp * gx[ig];
        b1 = b1 + ftmp * gy[ig];
        b2 = b2 + ftmp * gz[ig];
      }
      floc[3 * ia] = floc[3 * ia] + b0;
      floc[3 * ia + 1] = floc[3 * ia + 1] + b1;
      floc[3 * ia + 2] = floc[3 * ia + 2] + b2;
    }
    for (k = nhk1[it]; k <= nhk2[it]; ++k) {
      const int ia = iatmpt[k - 1] - 1;
      floc[3 * ia] = floc[3 * ia] * 2.0;
      floc[3 * ia + 1] = floc[3 * ia + 1] * 2.0;
      floc[3 * ia + 2] = floc[3 * ia + 2] * 2.0;
    }
  }
}

/* vlocal_g, reciprocal-space part (Sources/nlkbpp.f90:908-972): per type
 * (llking false) the structure factor rhmur(G) = sum_k (cos, sin)(G.tau_k)
 * from the axis tables (the three complex products of flocal_g), then
 * thrhgr(G) += svco(G, type) / rvol (Re, -Im) rhmur(G) for G >= 1; thrhgr over
 * ig = 0..nplw5, its G = 0 entry left zero.  Returns the G = 0 grid value
 * sum_type nhk(type) svco(0, type) / rvol (912-924). */
double qxo_vlocal_g(int ntype, const int *nhk1, const int *nhk2, const int *iatmpt,
                    const int *llking, int natom, int nplw5, const int *nga, const int *ngb,
                    const int *ngc, const double *svco, double rvol, const int *kmax,
                    const double *const *dcq, const double *const *dsq, double *thrhgr)
{
  int it, k, ig;
  const double rvolr = 1.0 / rvol;
  double v0 = 0.0;
  double *rhmur = (double *)malloc(sizeof(double) * 2 * ((size_t)nplw5 + 1));
  (void)natom;
  for (it = 0; it < ntype; ++it) {
    if (llking[it]) continue;
    const double dcsn = (double)(nhk2[it] - nhk1[it] + 1) * svco[(size_t)(nplw5 + 1) * it];
    v0 = v0 + dcsn * rvolr;
  }
  for (ig = 0; ig < 2 * (nplw5 + 1); ++ig) thrhgr[ig] = 0.0;
  for (it = 0; it < ntype; ++it) {
    if (llking[it]) continue;
    for (ig = 0; ig < 2 * (nplw5 + 1); ++ig) rhmur[ig] = 0.0;
    for (k = nhk1[it]; k <= nhk2[it]; ++k) {
      const int ia = iatmpt[k - 1] - 1;
      const double *c1 = dcq[0] + (size_t)ia * (2 * kmax[0] + 1) + kmax[0];
      const double *s1 = dsq[0] + (size_t)ia * (2 * kmax[0] + 1) + kmax[0];
      const double *c2 = dcq[1] + (size_t)ia * (2 * kmax[1] + 1) + kmax[1];
      const double *s2 = dsq[1] + (size_t)ia * (2 * kmax[1] + 1) + kmax[1];
      const double *c3 = dcq[2] + (size_t)ia * (2 * kmax[2] + 1) + kmax[2];
      const double *s3 = dsq[2] + (size_t)ia * (2 * kmax[2] + 1) + kmax[2];
      for (ig = 1; ig <= nplw5; ++ig) {
        const int k1 = nga[ig], k2 = ngb[ig], k3 = ngc[ig];
        const double dcc12 = c1[k1] * c2[k2] - s1[k1] * s2[k2];
        const double dsc12 = s1[k1] * c2[k2] + c1[k1] * s2[k2];
        rhmur[2 * ig] = rhmur[2 * ig] + (dcc12 * c3[k3] - dsc12 * s3[k3]);
        rhmur[2 * ig + 1] = rhmur[2 * ig + 1] + (dsc12 * c3[k3] + dcc12 * s3[k3]);
      }
    }
    for (ig = 1; ig <= nplw5; ++ig) {
      const double dggg = svco[ig + (size_t)(nplw5 + 1) * it] * rvolr;
      const double phar = rhmur[2 * ig] * dggg;
      const double phai = -rhmur[2 * ig + 1] * dggg;
      thrhgr[2 * ig] = thrhgr[2 * ig] + phar;
      thrhgr[2 * ig + 1] = thrhgr[2 * ig + 1] + phai;
    }
  }
  free(rhmur);
  return v0;
}

/* frcrec without stress, VECTOR form (Sources/bulk.f90:1880-2026): Ewald
 * reciprocal-space forces, energy and per-atom energies.  Per G (m =
 * 1..numk2, arrays at [m-1]) the ion sum S = sum_k zv (cos, sin)(G.tau_k)
 * from the per-ion axis tables (k = -kcsma..kcsma, dcq[a] + k_ion * (2
 * kcsma_a + 1)), hep = (exp(-(pi/gamma)^2 |g|^2) - recnrmex) / |g|^2;
 * forces += 8/V hep G (Re S sin_k - Im S cos_k) zv, energies as there.
 * iatoit, iatmpt, nhk1/nhk2 1-based. */
void qxo_frcrec(int ntype, const int *nhk1, const int *nhk2, const double *zv, int nel,
                int natom, int nion, const int *iatoit, const int *iatmpt, int numk2,
                const int *nga, const int *ngb, const int *ngc, const double *gx, const double *gy,
                const double *gz, const double *recnrm, const double *recnrmex, double gamma,
                double volume, const int *kcsma, const double *const *dcq,
                const double *const *dsq, double *fclm, double *ecolmb, double *atom_ecolmb)
{
  const double pi = acos(-1.0);
  const double vol8 = 8.0 / volume, const2 = -((pi / gamma) * (pi / gamma));
  double cst = 2.0 / (volume * pi), recolmb = 0.0, ztot2 = 0.0;
  double *sc = (double *)calloc((size_t)numk2, sizeof(double));
  double *ss = (double *)calloc((size_t)numk2, sizeof(double));
  double *hp = (double *)malloc(sizeof(double) * (size_t)numk2);
  double *acc = (double *)calloc(4 * (size_t)nion, sizeof(double));
  double *dyc = (double *)malloc(sizeof(double) * (size_t)nion);
  double *dys = (double *)malloc(sizeof(double) * (size_t)nion);
  int k, m, it;
#define TAB(t, a, kk, q) ((t)[a][(size_t)(kk) * (2 * kcsma[a] + 1) + kcsma[a] + (q)])
  for (k = 0; k < nion; ++k) {
    const int ia = iatmpt[k] - 1, ty = iatoit[ia] - 1;
    for (m = 0; m < numk2; ++m) {
      const int k1 = nga[m], k2 = ngb[m], k3 = ngc[m];
      const double b1 = TAB(dcq, 0, k, k1) * TAB(dcq, 1, k, k2) - TAB(dsq, 0, k, k1) * TAB(dsq, 1, k, k2);
      const double b2 = TAB(dsq, 0, k, k1) * TAB(dcq, 1, k, k2) + TAB(dcq, 0, k, k1) * TAB(dsq, 1, k, k2);
      double yc = b1 * TAB(dcq, 2, k, k3) - b2 * TAB(dsq, 2, k, k3);
      double ys = b2 * TAB(dcq, 2, k, k3) + b1 * TAB(dsq, 2, k, k3);
      yc = zv[ty] * yc;
      ys = zv[ty] * ys;
      sc[m] = sc[m] + yc;
      ss[m] = ss[m] + ys;
    }
  }
  for (m = 0; m < numk2; ++m) {
    const double recr = 1.0 / recnrm[m];
    hp[m] = (exp(const2 * recnrm[m]) - recnrmex[m + 1]) * recr;
  }
  for (m = 0; m < numk2; ++m) {
    const int k1 = nga[m], k2 = ngb[m], k3 = ngc[m];
    const double dckm = sc[m], dskm = ss[m], hep = hp[m];
    double hep8, hckm, hskm, dckmx, dckmy, dckmz, dskmx, dskmy, dskmz, hepc;
    for (k = 0; k < nion; ++k) {
      const int ia = iatmpt[k] - 1, ty = iatoit[ia] - 1;
      const double d12 = TAB(dcq, 0, k, k1) * TAB(dcq, 1, k, k2) - TAB(dsq, 0, k, k1) * TAB(dsq, 1, k, k2);
      const double s12 = TAB(dsq, 0, k, k1) * TAB(dcq, 1, k, k2) + TAB(dcq, 0, k, k1) * TAB(dsq, 1, k, k2);
      double yc = d12 * TAB(dcq, 2, k, k3) - s12 * TAB(dsq, 2, k, k3);
      double ys = s12 * TAB(dcq, 2, k, k3) + d12 * TAB(dsq, 2, k, k3);
      dyc[k] = zv[ty] * yc;
      dys[k] = zv[ty] * ys;
    }
    recolmb = recolmb + hep * (dckm * dckm + dskm * dskm);
    hep8 = hep * vol8;
    hckm = hep8 * dckm;
    hskm = hep8 * dskm;
    dckmx = hckm * gx[m];
    dckmy = hckm * gy[m];
    dckmz = hckm * gz[m];
    dskmx = hskm * gx[m];
    dskmy = hskm * gy[m];
    dskmz = hskm * gz[m];
    hepc = hep * cst;
    hckm = hepc * dckm;
    hskm = hepc * dskm;
    for (k = 0; k < nion; ++k) {
      acc[4 * k] = acc[4 * k] + dckmx * dys[k] - dskmx * dyc[k];
      acc[4 * k + 1] = acc[4 * k + 1] + dckmy * dys[k] - dskmy * dyc[k];
      acc[4 * k + 2] = acc[4 * k + 2] + dckmz * dys[k] - dskmz * dyc[k];
      acc[4 * k + 3] = acc[4 * k + 3] + dyc[k] * hckm + dys[k] * hskm;
    }
  }
#undef TAB
  for (k = 0; k < nion; ++k) {
    const int ia = iatmpt[k] - 1;
    fclm[3 * ia] = fclm[3 * ia] + acc[4 * k];
    fclm[3 * ia + 1] = fclm[3 * ia + 1] + acc[4 * k + 1];
    fclm[3 * ia + 2] = fclm[3 * ia + 2] + acc[4 * k + 2];
    atom_ecolmb[ia] = atom_ecolmb[ia] + acc[4 * k + 3];
  }
  *ecolmb = *ecolmb + cst * recolmb;
  for (it = 0; it < ntype; ++it) ztot2 = ztot2 + (double)(nhk2[it] - nhk1[it] + 1) * zv[it] * zv[it];
  *ecolmb = *ecolmb - (double)(nel * nel) / volume * (pi / (gamma * gamma) - recnrmex[0]) -
            2.0 * gamma * ztot2 / sqrt(pi);
  for (k = 0; k < nion; ++k) {
    const int ia = iatmpt[k] - 1, ty = iatoit[ia] - 1;
    atom_ecolmb[ia] = atom_ecolmb[ia] -
                      zv[ty] * (double)nel / volume * (pi / (gamma * gamma) - recnrmex[0]) -
                      2.0 * gamma * zv[ty] * zv[ty] / sqrt(pi);
  }
  free(sc);
  free(ss);
  free(hp);
  free(acc);
  free(dyc);
  free(dys);
}

/* roteig2, single rank, PCHOLESKY = 0 (Sources/sbalin.f90:772-1373): the
 * previous wave functions are turned to the current ones by the orthogonal
 * factor of their overlap.  A(i,j) = <prev1_i|cur_j> (packed product), B =
 * A^T A = V diag(ew) V^T (eigenvectors by qxo_eigh), U = V diag(ew)^-1/2 V^T
 * A^T, then every history set k = 1..min(nprvmx, kpwvx) becomes
 * prev_k(:, ib) = sum_jb U(ib, jb) prev_k(:, jb).  cur, prev: band-major
 * (band ib at ib * nplwex), prev sets nband * nplwex apart.  u_out (nband^2,
 * column-major U) if non-null.  Returns 0, or 1 if B has a non-positive
 * eigenvalue (then nothing is changed). */
int qxo_roteig(const double *cur, double *prev, int nplwex, int nband, int nprvmx, int kpwvx,
               double *u_out)
{
  const size_t nb = (size_t)nband, n = (size_t)nplwex;
  double *A = (double *)malloc(sizeof(double) * nb * nb);
  double *B = (double *)malloc(sizeof(double) * nb * nb);
  double *ew = (double *)malloc(sizeof(double) * nb);
  double *X = (double *)malloc(sizeof(double) * nb * nb);
  double *U = (double *)malloc(sizeof(double) * nb * nb);
  double *tmp = (double *)malloc(sizeof(double) * nb * n);
  size_t i, j, k, m;
  int key, rc = 0;
#define M_(a, r, c) (a)[(r) + nb * (c)]
  for (i = 0; i < nb; ++i)
    for (j = 0; j < nb; ++j) {
      const double *x = cur + j * n, *y = prev + i * n;
      double sum = 0.0;
      for (m = 0; m < n; ++m) sum = sum + x[m] * y[m];
      M_(A, i, j) = sum * 2.0 - x[0] * y[0];
    }
  for (i = 0; i < nb; ++i)
    for (j = 0; j <= i; ++j) {
      double sum = 0.0;
      for (k = 0; k < nb; ++k) sum = sum + M_(A, k, i) * M_(A, k, j);
      M_(B, i, j) = sum;
      M_(B, j, i) = sum;
    }
  qxo_eigh(B, nband, nband, ew);
  for (i = 0; i < nb; ++i)
    if (!(ew[i] > 0.0)) rc = 1;
  if (!rc) {
    /* X = diag(ew)^-1/2 V^T A^T, U = V X */
    for (i = 0; i < nb; ++i) {
      const double cosi = 1.0 / sqrt(ew[i]);
      for (j = 0; j < nb; ++j) {
        double sum = 0.0;
        for (k = 0; k < nb; ++k) sum = sum + M_(B, k, i) * M_(A, j, k);
        M_(X, i, j) = sum * cosi;
      }
    }
    for (i = 0; i < nb; ++i)
      for (j = 0; j < nb; ++j) {
        double sum = 0.0;
        for (k = 0; k < nb; ++k) sum = sum + M_(B, i, k) * M_(X, k, j);
        M_(U, i, j) = sum;
      }
    for (key = 0; key < nprvmx && key < kpwvx; ++key) {
      double *p = prev + (size_t)key * nb * n;
      for (i = 0; i < nb; ++i)
        for (m = 0; m < n; ++m) {
          double sum = 0.0;
          for (j = 0; j < nb; ++j) sum = sum + M_(U, i, j) * p[j * n + m];
          tmp[i * n + m] = sum;
        }
      memcpy(p, tmp, sizeof(double) * nb * n);
    }
    if (u_out) memcpy(u_out, U, sizeof(double) * nb * nb);
  }
#undef M_
  free(A);
  free(B);
  free(ew);
  free(X);
  free(U);
  free(tmp);
  return rc;
}

/* set_coef_ASPC (Sources/sbalin.f90:77-170): the predictor coefficients
 * coefaspc(j, k), j <= k = 2..kaspc, kaspc x kaspc column-major, zero
 * elsewhere (always-stable predictor-corrector of Kolafa). */
void qxo_coef_aspc(int kaspc, double *coef)
{
  int k, j, i;
  for (i = 0; i < kaspc * kaspc; ++i) coef[i] = 0.0;
  for (k = 2; k <= kaspc; ++k)
    for (j = 1; j <= k; ++j) {
      double a = 1.0, b = 1.0;
      for (i = k - j + 1; i <= k - 1; ++i) a = a * (double)i;
      for (i = k; i <= k + j; ++i) b = b * (double)i;
      coef[(j - 1) + (size_t)kaspc * (k - 1)] =
          (double)(((j + 1) % 2) ? -1 : 1) * (double)(j * (2 * k) * (2 * k - 1)) * a / b;
    }
}

/* sbalinaspc for one spin, lvand false (Sources/sbalin.f90:369-712): state
 * keypwv in and out; gdcr (band-major) is replaced by the ASPC prediction
 * from the current and the aligned previous wave functions, the history
 * prveig (nprvmx sets) shifted; with ihest > 0 Gram-Schmidt (qxo_schmidt)
 * follows.  Returns the iaspc used (0 when only the history was started). */
int qxo_sbalinaspc(double *gdcr, int nplwex, int nband, double *prveig, int nprvmx,
                   const double *coefaspc, int kaspc, int ihest, int nstep, int nstep_ini,
                   int *keypwv)
{
  const size_t set = (size_t)nplwex * nband;
  int i, iaspc = 0;
  if (nstep == nstep_ini + 1 || *keypwv == 0) {
    *keypwv = 1;
    memcpy(prveig, gdcr, sizeof(double) * set);
    return 0;
  }
  for (i = 0; i <= nprvmx; ++i)
    if (ihest == i || nstep == nstep_ini + 1 + i || *keypwv == i) {
      iaspc = i + 1;
      break;
    }
  *keypwv = (*keypwv + 1 < nprvmx) ? *keypwv + 1 : nprvmx;
  qxo_roteig(gdcr, prveig, nplwex, nband, nprvmx, (iaspc - 1 > 1) ? iaspc - 1 : 1, NULL);
  {
    double *prcd = (double *)malloc(sizeof(double) * (size_t)nplwex);
    int ib, m, j;
    for (ib = 0; ib < nband; ++ib) {
      double *g = gdcr + (size_t)ib * nplwex;
      if (iaspc > 1) {
        const double c1 = coefaspc[(size_t)kaspc * (iaspc - 1)];
        for (m = 0; m < nplwex; ++m) prcd[m] = c1 * g[m];
        for (j = 2; j <= iaspc; ++j) {
          const double cj = coefaspc[(j - 1) + (size_t)kaspc * (iaspc - 1)];
          const double *p = prveig + (size_t)(j - 2) * set + (size_t)ib * nplwex;
          for (m = 0; m < nplwex; ++m) prcd[m] = prcd[m] + cj * p[m];
        }
      }
      for (j = nprvmx - 1; j >= 1; --j)
        memcpy(prveig + (size_t)j * set + (size_t)ib * nplwex,
               prveig + (size_t)(j - 1) * set + (size_t)ib * nplwex, sizeof(double) * nplwex);
      memcpy(prveig + (size_t)ib * nplwex, g, sizeof(double) * nplwex);
      if (iaspc > 1) memcpy(g, prcd, sizeof(double) * nplwex);
    }
    free(prcd);
  }
  if (ihest != 0) {
    double *s = (double *)malloc(sizeof(double) * (size_t)nband * nband);
    qxo_schmidt(gdcr, nplwex, nband, s);
    free(s);
  }
  return iaspc;
}

/* enkbpp_g2 (Sources/nlkbpp.f90:1763-1824), Gamma point. */
double qxo_enkbpp_g(const double *slmir, int lmmax, int natom, const int *lmx_atom,
                    const double *clalp_atom, const double *occ, int nb, double rvol)
{
  double ecenl = 0.0;
  int ib, ia, l;
  for (ib = 0; ib < nb; ++ib) {
    double e = 0.0;
    if (!(fabs(occ[ib]) > 1e-15)) continue;
    for (ia = 0; ia < natom; ++ia)
      for (l = 0; l < lmx_atom[ia]; ++l) {
        double s = slmir[l + (size_t)lmmax * (ia + (size_t)natom * ib)];
        e += s * s * clalp_atom[l + lmmax * ia];
      }
    ecenl += occ[ib] * e / rvol;
  }
  return ecenl;
}

/* ======================================================================= */
/* Dense band algebra                                                      */
/* ======================================================================= */

#define SR(i, j) sr[((i) - 1) + (size_t)lda * ((j) - 1)]

/* rsdecmp: Cholesky S = L L^T, lower triangle (Sources/schmidt.f90:428-489). */
void qxo_rsdecmp(double *sr, int n, int lda)
{
  int i, j, k;
  const double zero = 1e-30;
  for (i = 1; i <= n; ++i) {
    for (k = 1; k <= i - 1; ++k) SR(i, i) -= SR(i, k) * SR(i, k);
    if (SR(i, i) > zero) {
      double r;
      SR(i, i) = sqrt(SR(i, i));
      for (k = 1; k <= i - 1; ++k) {
        double srik = SR(i, k);
        for (j = i + 1; j <= n; ++j) SR(j, i) -= SR(j, k) * srik;
      }
      r = 1.0 / SR(i, i);
      for (j = i + 1; j <= n; ++j) SR(j, i) *= r;
    } else {
      for (j = i; j <= n; ++j) SR(j, i) = 0.0;
    }
  }
  for (i = 1; i <= n - 1; ++i)
    for (j = i + 1; j <= n; ++j) SR(i, j) = 0.0;
}

/* rsinvcd: L^-1 of the Cholesky factor in place (Sources/schmidt.f90:494-544). */
void qxo_rsinvcd(double *sr, int n, int lda, double *work)
{
  int i, j, k;
  const double zero = 1e-30;
  for (i = 1; i <= n; ++i) {
    if (fabs(SR(i, i)) > zero) {
      double r = 1.0 / SR(i, i);
      for (j = i + 1; j <= n; ++j) SR(j, i) *= r;
      SR(i, i) = r;
    }
  }
  for (j = n - 1; j >= 1; --j) {
    for (i = j + 1; i <= n; ++i) work[i - 1] = 0.0;
    for (k = j + 1; k <= n; ++k) {
      double srkj = SR(k, j);
      for (i = j + 1; i <= n; ++i) work[i - 1] += SR(i, k) * srkj;
    }
    for (i = j + 1; i <= n; ++i) SR(i, j) = -work[i - 1];
  }
}
#undef SR

/* schmidt2, single rank (Sources/schmidt.f90:117-378): S = Psi^T Psi (packed
 * inner product), S = L L^T, Psi <- Psi L^-T.  Returns lcholesky. */
int qxo_schmidt(double *cgjr, int nplwex, int nband, double *smat)
{
  int i, j, ig, ib, jb;
  double *work = (double *)malloc(sizeof(double) * ((size_t)nband + (size_t)nplwex));
  double *dbuf = work + nband;
  for (i = 0; i < nband * nband; ++i) smat[i] = 0.0;
#ifdef _OPENMP
#pragma omp parallel for private(j, ig) schedule(dynamic, 4)
#endif
  for (i = 0; i < nband; ++i) {
    const double *ci = cgjr + (size_t)nplwex * i;
    for (j = 0; j <= i; ++j) {
      const double *cj = cgjr + (size_t)nplwex * j;
      double t = 0.0;
      for (ig = 0; ig < nplwex; ++ig) t += cj[ig] * ci[ig];
      t = t * 2.0 - cj[0] * ci[0];
      smat[i + (size_t)nband * j] = t;
    }
  }
  qxo_rsdecmp(smat, nband, nband);
  qxo_rsinvcd(smat, nband, nband, work);
  for (ib = 0; ib < nband; ++ib)
    if (!(smat[ib + (size_t)nband * ib] > 1e-30)) { free(work); return 0; }
  for (ib = nband - 1; ib >= 0; --ib) {
    memset(dbuf, 0, sizeof(double) * (size_t)nplwex);
    for (jb = 0; jb <= ib; ++jb) {
      double p = smat[ib + (size_t)nband * jb];
      const double *cj = cgjr + (size_t)nplwex * jb;
      for (ig = 0; ig < nplwex; ++ig) dbuf[ig] += p * cj[ig];
    }
    memcpy(cgjr + (size_t)nplwex * ib, dbuf, sizeof(double) * (size_t)nplwex);
  }
  free(work);
  return 1;
}

/* hmatrix2, single rank (Sources/engrad.f90:845-956). */
void qxo_hmatrix(const double *cgjr, const double *rhcr, int nplwex, int nband,
                 double *dmtrxr, double *eig)
{
  int ib, jb, ig;
#ifdef _OPENMP
#pragma omp parallel for private(jb, ig) schedule(dynamic, 4)
#endif
  for (ib = 0; ib < nband; ++ib) {
    const double *r = rhcr + (size_t)nplwex * ib;
    for (jb = 0; jb <= ib; ++jb) {
      const double *c = cgjr + (size_t)nplwex * jb;
      double p = 0.0;
      for (ig = 0; ig < nplwex; ++ig) p += r[ig] * c[ig];
      p = 2.0 * p - r[0] * c[0];
      dmtrxr[jb + (size_t)nband * ib] = p;
      dmtrxr[ib + (size_t)nband * jb] = p;
    }
  }
  for (ib = 0; ib < nband; ++ib) eig[ib] = dmtrxr[ib + (size_t)nband * ib];
}

/* Symmetric eigensolver standing in for RsEIGQR (Sources/funcs.f90:380-764):
 * cyclic Jacobi in FP64; eigenvalues ascending (stable insertion order as in
 * the reference's sort, 713-741) with eigenvectors in the columns of a. */
int qxo_eigh(double *a, int n, int lda, double *w)
{
  double *v = (double *)calloc((size_t)n * n, sizeof(double));
  int i, j, k, sweep;
#define A_(r, c) a[(r) + (size_t)lda * (c)]
#define V_(r, c) v[(r) + (size_t)n * (c)]
  for (i = 0; i < n; ++i) V_(i, i) = 1.0;
  for (j = 0; j < n; ++j)
    for (i = 0; i < j; ++i) A_(i, j) = A_(j, i);
  for (sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (j = 0; j < n; ++j)
      for (i = 0; i < n; ++i) {
        double x = A_(i, j) * A_(i, j);
        tot += x;
        if (i != j) off += x;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (i = 0; i < n - 1; ++i)
      for (j = i + 1; j < n; ++j) {
        double aij = A_(i, j), aii, ajj, theta, t, c, s;
        if (fabs(aij) < 1e-300) continue;
        aii = A_(i, i); ajj = A_(j, j);
        theta = (ajj - aii) / (2.0 * aij);
        t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        c = 1.0 / sqrt(t * t + 1.0);
        s = t * c;
        for (k = 0; k < n; ++k) {
          double aki = A_(k, i), akj = A_(k, j);
          A_(k, i) = c * aki - s * akj;
          A_(k, j) = s * aki + c * akj;
        }
        for (k = 0; k < n; ++k) {
          double aik = A_(i, k), ajk = A_(j, k);
          A_(i, k) = c * aik - s * ajk;
          A_(j, k) = s * aik + c * ajk;
        }
        for (k = 0; k < n; ++k) {
          double vki = V_(k, i), vkj = V_(k, j);
          V_(k, i) = c * vki - s * vkj;
          V_(k, j) = s * vki + c * vkj;
        }
      }
  }
  {
    int *ord = (int *)malloc(sizeof(int) * (size_t)n);
    double *ev = (double *)malloc(sizeof(double) * (size_t)n);
    for (i = 0; i < n; ++i) { ev[i] = A_(i, i); ord[i] = i; }
    for (i = 1; i < n; ++i) {          /* stable insertion sort */
      int oi = ord[i];
      double e = ev[oi];
      j = i - 1;
      while (j >= 0 && ev[ord[j]] > e) { ord[j + 1] = ord[j]; --j; }
      ord[j + 1] = oi;
    }
    for (j = 0; j < n; ++j) {
      w[j] = ev[ord[j]];
      for (i = 0; i < n; ++i) A_(i, j) = V_(i, ord[j]);
    }
    free(ord); free(ev);
  }
  free(v);
#undef A_
#undef V_
  return 0;
}

/* untrycore, single block (Sources/eigen.f90:1456-1514): Psi <- Psi U. */
void qxo_untrycore(double *gdcr, int nplwex, int nband, const double *u, int ldu, double *work)
{
  int ib, jb, ig;
#ifdef _OPENMP
#pragma omp parallel for private(jb, ig) schedule(static)
#endif
  for (ib = 0; ib < nband; ++ib) {
    double *t = work + (size_t)nplwex * ib;
    for (ig = 0; ig < nplwex; ++ig) t[ig] = 0.0;
    for (jb = 0; jb < nband; ++jb) {
      double p = u[jb + (size_t)ldu * ib];
      const double *g = gdcr + (size_t)nplwex * jb;
      for (ig = 0; ig < nplwex; ++ig) t[ig] += p * g[ig];
    }
  }
  memcpy(gdcr, work, sizeof(double) * (size_t)nplwex * nband);
}

/* softchg, VECTOR build on one bulk rank where FD and FFT meshes coincide
 * (Sources/chgdns.f90:388-454; mshglb identity, Sources/ba.f90:727-733). */
void qxo_softchg(const double *cgjr, int nplwex, int nplw, int nband, const double *occ,
                 double rvol, const int *ijkgd_c, int n1, int n2, int n3, double *rho)
{
  long N = (long)n1 * n2 * n3, t;
  int i;
  double *fx = (double *)malloc(sizeof(double) * 2 * (size_t)N), *fy = fx + N;
  memset(rho, 0, sizeof(double) * (size_t)N);
  for (i = 0; i < nband; ++i) {
    double wr = occ[i] / rvol;
    qxo_wvg2r(cgjr + (size_t)nplwex * i, nplw, ijkgd_c, n1, n2, n3, fx, fy);
    for (t = 0; t < N; ++t) rho[t] = rho[t] + fx[t] * fx[t] * wr;
  }
  free(fx);
}

/* prpcg: modified Teter-Payne-Allan preconditioner (Sources/eigen.f90:1093-1113). */
void qxo_prpcg(double ekib, const double *dkgnrm, int nplwex, double *prcd)
{
  int ig;
  for (ig = 0; ig < nplwex; ++ig) {
    double e15 = 1.5 * ekib, x = dkgnrm[ig] / e15, tesu, x2, x4;
    tesu = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
    x2 = x * x;
    x4 = x2 * x2;
    tesu = tesu / (tesu + 16.0 * x4);
    prcd[ig] = tesu * 2.0 / e15;
  }
}

/* ======================================================================= */
/* Band solver                                                             */
/* ======================================================================= */

static double pdot(const double *x, const double *y, int n)
{
  double s = 0.0;
  int i;
  for (i = 0; i < n; ++i) s += x[i] * y[i];
  return 2.0 * s - x[0] * y[0];
}

/* H|psi> for nb vectors: kinetic, KB (all bands), local (per band).
 * Order follows ekinetic2 -> hcnonloc -> hckbpp_g of enonloc2. */
void qxo_hpsi(const qxo_ham *h, const double *psi, double *hpsi, int nb, double *slmir)
{
  int ib;
  long N = (long)h->n1 * h->n2 * h->n3;
  qxo_ekinetic(h->dkgnrm, psi, hpsi, h->nplwex, nb, NULL, NULL, NULL);
  {
    double *fx = (double *)malloc(sizeof(double) * 2 * (size_t)N), *fy = fx + N;
    for (ib = 0; ib < nb; ++ib)
      qxo_hcnonloc(h->glocal, psi + (size_t)h->nplwex * ib, hpsi + (size_t)h->nplwex * ib,
                   h->nplw, h->ijkgd_c, h->n1, h->n2, h->n3, fx, fy);
    free(fx);
  }
  if (h->eqlmr)
    qxo_hckbpp_g(h->eqlmr, h->nplwex, h->lmmax, h->natom, h->lmx_atom, h->vclalp_atom,
                 psi, hpsi, nb, slmir);
}

/* H|h> for one vector in the CG inner loop: kinetic, hckbpp_ib_g, hcnonloc
 * (Sources/eigen.f90:516-551). */
static void hpsi_one(const qxo_ham *h, const double *x, double *hx, double *fx, double *fy)
{
  int ig;
  for (ig = 0; ig < h->nplwex; ++ig) hx[ig] = h->dkgnrm[ig] * x[ig];
  if (h->eqlmr)
    qxo_hckbpp_g(h->eqlmr, h->nplwex, h->lmmax, h->natom, h->lmx_atom, h->vclalp_atom,
                 x, hx, 1, NULL);
  qxo_hcnonloc(h->glocal, x, hx, h->nplw, h->ijkgd_c, h->n1, h->n2, h->n3, fx, fy);
}

/* Trace of the BKL schedule (test hook): when a buffer is registered,
 * qxo_eigen_bkl appends one record of QXO_BKL_TRACE_W doubles per CG iteration
 * -- ib, iter, h11 (eig before the step), zansap, beta, cai, cbi, hhi, eig
 * after the step, dtnew, alpsq, zansa, the ifhhi branch (1 BKL, 0 line
 * minimisation), the exit flag, eigb, z000re, x000re -- and, if `vec` is
 * given, the vectors x (before the update) and h (after subgs3/subgs4) of that
 * iteration, 2 nplwex doubles per record.  Records past `cap` are dropped. */
static double *g_bkl_rec = NULL, *g_bkl_vec = NULL;
static int g_bkl_cap = 0, g_bkl_n = 0;

void qxo_bkl_trace(double *rec, double *vec, int cap)
{
  g_bkl_rec = rec;
  g_bkl_vec = vec;
  g_bkl_cap = rec ? cap : 0;
  g_bkl_n = 0;
}

int qxo_bkl_trace_count(void) { return g_bkl_n; }

/* eigen_BKL with lortho=.false. and lhcunt=.true., the form the SCF driver
 * calls (Sources/pwlda.f90:2940, 3121; Sources/eigen.f90:199-938).  gdcr and
 * the all-band copy cgjr coincide on one FP64 rank, so one array serves both. */
void qxo_eigen_bkl(const qxo_ham *h, double *eig, double *gdcr, double *rhcr,
                   const double *ekib, int nband, int noband, int itermx, int iteremx,
                   double tolera, double threinn, double wdinner, double roughfzan,
                   double roughfeig, double *bzan, int *nriter)
{
  const int n = h->nplwex;
  long N = (long)h->n1 * h->n2 * h->n3;
  double *prcd = (double *)malloc(sizeof(double) * (5 * (size_t)n + 2 * (size_t)N + nband));
  double *gnk = prcd + n, *hnk = gnk + n, *fx = hnk + n, *fy = fx + N, *prod = fy + N;
  int ibb, ig, jb;
  for (ibb = 0; ibb < nband; ++ibb) {
    const int ib = ibb + 1; /* 1-based band number */
    double *x = gdcr + (size_t)n * ibb, *hx = rhcr + (size_t)n * ibb;
    double zansa, zansab, z000re, x000re, eigb, dinnerx, zansap, beta, betk = 0.0;
    double cai, cbi, hhi, hhx, dtnew = 0.0, alpsq = 1.0;
    int iter, izcnt, lreset;
    qxo_prpcg(ekib[ibb], h->dkgnrm, n, prcd);
    eig[ibb] = pdot(x, hx, n);
    for (ig = 0; ig < n; ++ig) gnk[ig] = eig[ibb] * x[ig] - hx[ig];
    zansa = pdot(gnk, gnk, n);
    iter = 0; zansab = zansa; z000re = zansa; x000re = 1e10; eigb = eig[ibb];
    izcnt = 0; lreset = 1; dinnerx = 0.0;
    for (;;) {
      ++iter;
      for (ig = 0; ig < n; ++ig) gnk[ig] = prcd[ig] * gnk[ig];
      zansap = pdot(gnk, gnk, n);
      if (lreset) { beta = 0.0; memset(hnk, 0, sizeof(double) * (size_t)n); }
      else beta = zansap / betk;
      lreset = 0;
      for (ig = 0; ig < n; ++ig) hnk[ig] = gnk[ig] + beta * hnk[ig];
      betk = zansap;
      /* subgs3 against bands 1..ib-1 (Sources/engrad.f90:1072-1191, ifnorm=0) */
      for (jb = 0; jb < ibb; ++jb) {
        const double *c = gdcr + (size_t)n * jb;
        double p = 0.0;
        for (ig = 0; ig < n; ++ig) p += c[ig] * hnk[ig];
        prod[jb] = 2.0 * p - c[0] * hnk[0];
      }
      for (jb = 0; jb < ibb; ++jb) {
        const double *c = gdcr + (size_t)n * jb;
        for (ig = 0; ig < n; ++ig) hnk[ig] -= prod[jb] * c[ig];
      }
      /* subgs4 against the current band (Sources/engrad.f90:1196-1248) */
      hhx = pdot(x, hnk, n);
      for (ig = 0; ig < n; ++ig) hnk[ig] -= hhx * x[ig];
      cai = pdot(hnk, hx, n);
      hpsi_one(h, hnk, gnk, fx, fy);
      cbi = pdot(hnk, gnk, n);
      hhi = pdot(hnk, hnk, n);
      const double h11_trace = eig[ibb];
      double *rec = NULL;
      if (g_bkl_rec && g_bkl_n < g_bkl_cap) {
        rec = g_bkl_rec + (size_t)QXO_BKL_TRACE_W * g_bkl_n;
        if (g_bkl_vec) {
          memcpy(g_bkl_vec + 2 * (size_t)n * g_bkl_n, x, sizeof(double) * (size_t)n);
          memcpy(g_bkl_vec + (2 * (size_t)g_bkl_n + 1) * n, hnk, sizeof(double) * (size_t)n);
        }
        ++g_bkl_n;
      }
      if (hhi > 0.0) {
        double h11 = eig[ibb], h12 = cai / sqrt(hhi), h21 = cai / sqrt(hhi), h22 = cbi / hhi;
        double enew = 0.5 * (h11 + h22 - sqrt((h11 + h22) * (h11 + h22) - 4.0 * (h11 * h22 - h12 * h21)));
        double cona3, cona4, alpsq2;
        hhx = 0.0;
        dtnew = (enew - h11) / h12 / sqrt(hhi);
        cona3 = hhi;
        cona4 = 2.0 * hhx;
        alpsq2 = 1.0 / (1.0 + dtnew * (cona4 + dtnew * cona3));
        alpsq = sqrt(alpsq2);
        eig[ibb] = enew;
      } else {
        double cona0, cona1, cona2, cona3, cona4, s0, s1, s2, s, dtex2, alpsq2;
        hhx = pdot(hnk, x, n);
        cona0 = eig[ibb]; cona1 = cai; cona2 = cbi; cona3 = hhi; cona4 = 2.0 * hhx;
        s0 = cona0 * cona4 - cona1;
        s1 = cona0 * cona3 - cona2;
        s2 = cona1 * cona3 - cona2 * cona4;
        s = s1 * s1 - s2 * s0;
        dtnew = (s > 0.0) ? (-s1 - sqrt(s)) / s2 : -1.0;
        if (dtnew > 0.0) {
          double dinner = sqrt(1.0 / (1.0 + dtnew * dtnew * hhi));
          if (dinner < threinn) dtnew = 0.0;
          if (dinner < dinnerx * wdinner) dtnew = 0.0;
          dinnerx = fmax(dinnerx, dinner);
        }
        dtex2 = dtnew * dtnew;
        alpsq2 = 1.0 / (1.0 + dtnew * cona4 + dtex2 * cona3);
        eig[ibb] = (cona0 + dtnew * cona1 + dtex2 * cona2) * alpsq2;
        alpsq = sqrt(alpsq2);
      }
      for (ig = 0; ig < n; ++ig) x[ig] = alpsq * (x[ig] + dtnew * hnk[ig]);
      for (ig = 0; ig < n; ++ig) hx[ig] = alpsq * (hx[ig] + dtnew * gnk[ig]);
      for (ig = 0; ig < n; ++ig) gnk[ig] = eig[ibb] * x[ig] - hx[ig];
      zansa = pdot(gnk, gnk, n);
      {
        const int stop = (ib <= noband && iter >= itermx) || (ib > noband && iter >= iteremx) ||
                         zansa < tolera || fabs(eigb - eig[ibb]) < tolera ||
                         zansa < z000re * roughfzan ||
                         (iter > 1 && fabs(eigb - eig[ibb]) < x000re * roughfeig);
        if (rec) {
          const double r[QXO_BKL_TRACE_W] = {ib, iter, h11_trace, zansap, beta, cai, cbi, hhi,
                                             eig[ibb], dtnew, alpsq, zansa, hhi > 0.0, stop,
                                             eigb, z000re, x000re};
          memcpy(rec, r, sizeof(r));
        }
        if (stop) break;
      }
      if (iter == 1) x000re = fabs(eigb - eig[ibb]);
      if (zansa > zansab) {
        ++izcnt;
        if (izcnt > 20) { izcnt = 0; lreset = 1; }
      }
      zansab = zansa;
      eigb = eig[ibb];
    }
    bzan[ibb] = zansa;
    nriter[ibb] = iter;
  }
  free(prcd);
}

/* One SCF iteration of the band part with a frozen local potential, in the
 * order of Sources/pwlda.f90: energy pass (2087-2326: ekinetic, enonloc,
 * hmatrix), subspace rotation (2700-2905: untryt + untryhc), BKL CG with
 * default parameters (2975-3285, Sources/modules.f90:173-179), Gram-Schmidt
 * (3225) and the soft charge density (3574).  Returns lcholesky. */
int qxo_scf_band_step(const qxo_ham *h, double *gdcr, double *rhcr, double *eig,
                      const double *occ, int nband, int noband, double rvol,
                      double *rho, double *ekin, double *enl, const double *clalp_atom)
{
  const int n = h->nplwex;
  double *ekib = (double *)malloc(sizeof(double) * (size_t)nband * (4 + (size_t)nband));
  double *bzan = ekib + nband, *hm = bzan + nband, *ev = hm + (size_t)nband * nband;
  int *nriter = (int *)malloc(sizeof(int) * (size_t)nband);
  double *work = (double *)malloc(sizeof(double) * (size_t)n * nband);
  double *slmir = NULL;
  long N = (long)h->n1 * h->n2 * h->n3;
  int ib, ok;
  *ekin = 0.0;
  qxo_ekinetic(h->dkgnrm, gdcr, rhcr, n, nband, occ, ekib, ekin);
  {
    double *fx = (double *)malloc(sizeof(double) * 2 * (size_t)N), *fy = fx + N;
    for (ib = 0; ib < nband; ++ib)
      qxo_hcnonloc(h->glocal, gdcr + (size_t)n * ib, rhcr + (size_t)n * ib, h->nplw,
                   h->ijkgd_c, h->n1, h->n2, h->n3, fx, fy);
    free(fx);
  }
  *enl = 0.0;
  if (h->eqlmr) {
    slmir = (double *)malloc(sizeof(double) * (size_t)h->lmmax * h->natom * nband);
    qxo_hckbpp_g(h->eqlmr, n, h->lmmax, h->natom, h->lmx_atom, h->vclalp_atom, gdcr, rhcr,
                 nband, slmir);
    if (clalp_atom)
      *enl = qxo_enkbpp_g(slmir, h->lmmax, h->natom, h->lmx_atom, clalp_atom, occ, nband, rvol);
  }
  qxo_hmatrix(gdcr, rhcr, n, nband, hm, eig);
  qxo_eigh(hm, nband, nband, ev);
  qxo_untrycore(gdcr, n, nband, hm, nband, work);
  qxo_untrycore(rhcr, n, nband, hm, nband, work);
  qxo_eigen_bkl(h, eig, gdcr, rhcr, ekib, nband, noband, 4, 4, 1e-10, 0.0, 0.9, 0.3, 0.3,
                bzan, nriter);
  ok = qxo_schmidt(gdcr, n, nband, hm);
  if (rho) qxo_softchg(gdcr, n, h->nplw, nband, occ, rvol, h->ijkgd_c, h->n1, h->n2, h->n3, rho);
  free(ekib); free(nriter); free(work); free(slmir);
  return ok;
}
