// mdstep.cpp -- the velocity-Verlet updates of the reference's MD loop at
// constant energy (ifmd = 2), on the host: they are O(natom) arithmetic
// between device calls, and the total force they read is assembled on the
// host (the real-space Ewald sum stays there).  Scaled coordinates as in the
// reference: x in units of the cell vectors, v the displacement per MD step,
// a = acon H^-1 f with acon = dt^2 / (2 m) per species
// (Sources/remd.f90:3568-3571), so that one step is
//   remd_semifinal:  a = normalized_a(f);  v += a          (second half kick)
//   next step:       v += a;  x += v;  wrap x into [0, 1)  (first half kick)
// (Sources/remd.f90:633-790, 540-616).  Every expression keeps the
// reference's operand order, so results are bit-identical to it.
#include <cmath>
#include <cstdint>

extern "C" {

// rciprl (Sources/reclat.f90:1089-1128): b = reciprocal vectors / 2 pi of the
// cell a (3 x 3 column-major, a_i . b_j = delta_ij), v = |det a|
int qx_rciprl(const double* a, double* b, double* v) {
  if (!a || !b || !v) return -1;
  auto A = [&](int i, int j) { return a[(i - 1) + 3 * (j - 1)]; };
  double B[3][3];
  B[0][0] = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3);
  B[1][0] = A(3, 2) * A(1, 3) - A(1, 2) * A(3, 3);
  B[2][0] = A(1, 2) * A(2, 3) - A(2, 2) * A(1, 3);
  double vol = A(1, 1) * B[0][0] + A(2, 1) * B[1][0] + A(3, 1) * B[2][0];
  const double vr = 1.0 / vol;
  B[0][0] = B[0][0] * vr;
  B[1][0] = B[1][0] * vr;
  B[2][0] = B[2][0] * vr;
  B[0][1] = (A(2, 3) * A(3, 1) - A(3, 3) * A(2, 1)) * vr;
  B[1][1] = (A(3, 3) * A(1, 1) - A(1, 3) * A(3, 1)) * vr;
  B[2][1] = (A(1, 3) * A(2, 1) - A(2, 3) * A(1, 1)) * vr;
  B[0][2] = (A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2)) * vr;
  B[1][2] = (A(3, 1) * A(1, 2) - A(1, 1) * A(3, 2)) * vr;
  B[2][2] = (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) * vr;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) b[i + 3 * j] = B[i][j];
  *v = std::fabs(vol);
  return 0;
}

// normalized_a (Sources/remd.f90:6789-6830): a(3, n) Cartesian forces in,
// scaled accelerations out: with hi = rciprl(h), a_i <- acon(is_i) *
// (hi(:,1) . a_i, hi(:,2) . a_i, hi(:,3) . a_i).  is: species of each atom,
// 1-based; h: the cell h(:,:,0), column-major.
int qx_normalized_a(int n, int ntype, double* a, const int* is, const double* h,
                    const double* acon) {
  if (n < 0 || ntype < 1 || !a || !is || !h || !acon) return -1;
  for (int i = 0; i < n; ++i)
    if (is[i] < 1 || is[i] > ntype) return -1;
  double hi[9], vtemp;
  qx_rciprl(h, hi, &vtemp);
  auto HI = [&](int i, int j) { return hi[(i - 1) + 3 * (j - 1)]; };
  for (int i = 0; i < n; ++i) {
    double* ai = a + 3 * (size_t)i;
    const double arx = ai[0], ary = ai[1], arz = ai[2];
    const double ax = HI(1, 1) * arx + HI(2, 1) * ary + HI(3, 1) * arz;
    const double ay = HI(1, 2) * arx + HI(2, 2) * ary + HI(3, 2) * arz;
    const double az = HI(1, 3) * arx + HI(2, 3) * ary + HI(3, 3) * arz;
    const double c = acon[is[i] - 1];
    ai[0] = ax * c;
    ai[1] = ay * c;
    ai[2] = az * c;
  }
  return 0;
}

// vkick (Sources/remd.f90:5408-5430): v += a for atoms of species not fixed
// (lfixion(it) != 0 fixes species it)
int qx_vkick(double* v, const double* a, const int* is, int n, int ntype, const int* lfixion) {
  if (n < 0 || ntype < 1 || !v || !a || !is || !lfixion) return -1;
  for (int i = 0; i < n; ++i) {
    if (is[i] < 1 || is[i] > ntype) return -1;
    if (lfixion[is[i] - 1]) continue;
    for (int k = 0; k < 3; ++k) v[3 * (size_t)i + k] = v[3 * (size_t)i + k] + a[3 * (size_t)i + k];
  }
  return 0;
}

// updtps (Sources/remd.f90:5755-5783): x += v for atoms of species not fixed
int qx_updtps(double* x, const double* v, const int* is, int n, int ntype, const int* lfixion) {
  return qx_vkick(x, v, is, n, ntype, lfixion);
}

// updtps_pbc (Sources/remd.f90:5786-5806): every scaled coordinate into
// [0, 1): x < 0 -> x - int(x) + 1, then x >= 1 -> x - int(x) (int truncates
// toward zero, as Fortran's INT)
int qx_updtps_pbc(double* x, int n) {
  if (n < 0 || !x) return -1;
  for (size_t i = 0; i < 3 * (size_t)n; ++i)  // INT of default kind must hold x
    if (!(std::fabs(x[i]) < 2147483647.0)) return -2;
  for (size_t i = 0; i < 3 * (size_t)n; ++i) {
    double xi = x[i];
    if (xi < 0.0) xi = xi - (double)(int32_t)xi + 1.0;
    if (xi >= 1.0) xi = xi - (double)(int32_t)xi;
    x[i] = xi;
  }
  return 0;
}

}  // extern "C"
