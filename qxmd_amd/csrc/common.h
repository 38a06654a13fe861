// common.h -- shared device helpers for the qxmd_amd HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#define QX_CHECK(expr)                                                                 \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      std::fprintf(stderr, "qxmd_amd: %s failed: %s (%s:%d)\n", #expr,                 \
                   hipGetErrorString(_e), __FILE__, __LINE__);                          \
      return -(int)_e - 1000;                                                          \
    }                                                                                  \
  } while (0)

namespace qx {

template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
__host__ __device__ inline cplx<T> mk(T a, T b) {
  cplx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by sign*i
template <typename T>
__host__ __device__ inline cplx<T> cmul_si(cplx<T> a, int sign) {
  return sign > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// Mixed-radix FFT plan for one line length: radices in pass order (Stockham).
// ---------------------------------------------------------------------------
constexpr int kMaxPasses = 10;
struct FftPlan {
  int n;
  int npass;
  int radix[kMaxPasses];
};

inline FftPlan make_plan(int n) {
  FftPlan p{};
  p.n = n;
  p.npass = 0;
  int m = n;
  while (m % 4 == 0) { p.radix[p.npass++] = 4; m /= 4; }
  while (m % 2 == 0) { p.radix[p.npass++] = 2; m /= 2; }
  while (m % 3 == 0) { p.radix[p.npass++] = 3; m /= 3; }
  while (m % 5 == 0) { p.radix[p.npass++] = 5; m /= 5; }
  while (m % 7 == 0) { p.radix[p.npass++] = 7; m /= 7; }
  if (m != 1) p.npass = -1;  // unsupported length
  return p;
}

// Small DFT codelets, y_q = sum_r x_r w^(rq), w = exp(sign*2*pi*i/R).
template <typename T, int R>
struct Dft;

template <typename T>
struct Dft<T, 2> {
  __host__ __device__ static void run(cplx<T>* v, int) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <typename T>
struct Dft<T, 4> {
  __host__ __device__ static void run(cplx<T>* v, int sign) {
    cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    cplx<T> t2 = cadd(v[1], v[3]), t3 = cmul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};
template <typename T>
struct Dft<T, 3> {
  __host__ __device__ static void run(cplx<T>* v, int sign) {
    const T c1 = T(-0.5), s1 = T(0.86602540378443864676372317075294);
    cplx<T> a = v[0], b = v[1], c = v[2];
    cplx<T> s = cadd(b, c), d = csub(b, c);
    cplx<T> t = mk<T>(a.x + c1 * s.x, a.y + c1 * s.y);
    cplx<T> u = cmul_si(mk<T>(s1 * d.x, s1 * d.y), sign);
    v[0] = cadd(a, s);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  }
};
template <typename T>
struct Dft<T, 5> {
  __host__ __device__ static void run(cplx<T>* v, int sign) {
    const T c1 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
    const T s1 = T(0.95105651629515357211643933337938);   // sin(2pi/5)
    const T s2 = T(0.58778525229247312916870595463907);   // sin(4pi/5)
    cplx<T> a = v[0];
    cplx<T> b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    cplx<T> d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    cplx<T> t1 = mk<T>(a.x + c1 * b1.x + c2 * b2.x, a.y + c1 * b1.y + c2 * b2.y);
    cplx<T> t2 = mk<T>(a.x + c2 * b1.x + c1 * b2.x, a.y + c2 * b1.y + c1 * b2.y);
    cplx<T> u1 = cmul_si(mk<T>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), sign);
    cplx<T> u2 = cmul_si(mk<T>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), sign);
    v[0] = cadd(a, cadd(b1, b2));
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
  }
};
template <typename T>
struct Dft<T, 7> {
  __host__ __device__ static void run(cplx<T>* v, int sign) {
    const T C[3] = {T(0.62348980185873353053), T(-0.22252093395631440429),
                    T(-0.90096886790241912624)};
    const T S[3] = {T(0.78183148246802980871), T(0.97492791218182360702),
                    T(0.43388373911755812048)};
    cplx<T> a = v[0];
    cplx<T> b[3], d[3];
    for (int r = 0; r < 3; ++r) {
      b[r] = cadd(v[r + 1], v[6 - r]);
      d[r] = csub(v[r + 1], v[6 - r]);
    }
    cplx<T> y[7];
    y[0] = cadd(a, cadd(b[0], cadd(b[1], b[2])));
    for (int q = 1; q <= 3; ++q) {
      T tr = a.x, ti = a.y, ur = 0, ui = 0;
      for (int r = 1; r <= 3; ++r) {
        int e = (r * q) % 7;
        T c = e <= 3 ? C[e - 1] : C[7 - e - 1];
        T s = e <= 3 ? S[e - 1] : -S[7 - e - 1];
        tr += c * b[r - 1].x;
        ti += c * b[r - 1].y;
        ur += s * d[r - 1].x;
        ui += s * d[r - 1].y;
      }
      cplx<T> u = cmul_si(mk<T>(ur, ui), sign);
      y[q] = cadd(mk<T>(tr, ti), u);
      y[7 - q] = csub(mk<T>(tr, ti), u);
    }
    for (int q = 0; q < 7; ++q) v[q] = y[q];
  }
};


// ---------------------------------------------------------------------------
// Compile-time twiddles and composite codelets (N = P*M <= 16) for the
// two-step line transforms.  cos/sin are evaluated by constexpr Taylor series
// on [-pi, pi], so every constant is folded at compile time.
// ---------------------------------------------------------------------------
constexpr double kPiCE = 3.14159265358979323846264338327950288;
constexpr double ce_cos_red(double x) {
  double term = 1.0, sum = 1.0, x2 = x * x;
  for (int i = 1; i < 30; ++i) {
    term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
constexpr double ce_sin_red(double x) {
  double term = x, sum = x, x2 = x * x;
  for (int i = 1; i < 30; ++i) {
    term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double ce_angle(int k, int n) {
  int kk = ((k % n) + n) % n;
  double x = 2.0 * kPiCE * (double)kk / (double)n;
  return x > kPiCE ? x - 2.0 * kPiCE : x;
}
template <int N>
struct TwTable {
  double c[N], s[N];
  constexpr TwTable() : c(), s() {
    for (int k = 0; k < N; ++k) {
      c[k] = ce_cos_red(ce_angle(k, N));
      s[k] = ce_sin_red(ce_angle(k, N));
    }
  }
};
template <int N>
struct TwHolder {
  static constexpr TwTable<N> tab = TwTable<N>();
};

// Cooley-Tukey in registers: N = P*M, natural order in and out
template <typename T, int P, int M>
struct DftCT {
  __host__ __device__ static void run(cplx<T>* x, int sign) {
    constexpr int N = P * M;
    cplx<T> y[P][M];
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int m = 0; m < M; ++m) y[p][m] = x[p + P * m];
      Dft<T, M>::run(y[p], sign);
    }
#pragma unroll
    for (int p = 1; p < P; ++p)
#pragma unroll
      for (int k = 1; k < M; ++k) {
        const T c = (T)TwHolder<N>::tab.c[(p * k) % N];
        const T sn = (T)TwHolder<N>::tab.s[(p * k) % N];
        y[p][k] = cmul(y[p][k], mk<T>(c, sign > 0 ? sn : -sn));
      }
#pragma unroll
    for (int k = 0; k < M; ++k) {
      cplx<T> z[P];
#pragma unroll
      for (int p = 0; p < P; ++p) z[p] = y[p][k];
      Dft<T, P>::run(z, sign);
#pragma unroll
      for (int q = 0; q < P; ++q) x[k + M * q] = z[q];
    }
  }
};
template <typename T> struct Dft<T, 6> { __host__ __device__ static void run(cplx<T>* v, int s) { DftCT<T, 2, 3>::run(v, s); } };
template <typename T> struct Dft<T, 8> { __host__ __device__ static void run(cplx<T>* v, int s) { DftCT<T, 2, 4>::run(v, s); } };
template <typename T> struct Dft<T, 9> { __host__ __device__ static void run(cplx<T>* v, int s) { DftCT<T, 3, 3>::run(v, s); } };
template <typename T> struct Dft<T, 10> { __host__ __device__ static void run(cplx<T>* v, int s) { DftCT<T, 2, 5>::run(v, s); } };
template <typename T> struct Dft<T, 12> { __host__ __device__ static void run(cplx<T>* v, int s) { DftCT<T, 3, 4>::run(v, s); } };
template <typename T> struct Dft<T, 15> { __host__ __device__ static void run(cplx<T>* v, int s) { DftCT<T, 3, 5>::run(v, s); } };
template <typename T> struct Dft<T, 16> { __host__ __device__ static void run(cplx<T>* v, int s) { DftCT<T, 4, 4>::run(v, s); } };

// ---------------------------------------------------------------------------
// In-LDS batched Stockham FFT.  Lines start at buf + line*lstride with
// element stride estride; tw holds exp(+2*pi*i*k/n), k < n (conjugated for
// sign < 0).  Each pass reads every butterfly of a group of lines into
// registers, synchronises, then writes in place, so a pass needs no second
// buffer; MAXB bounds the butterflies one thread holds per pass, and lines
// are processed in groups small enough for that bound.  All threads of the
// block must call this.
// ---------------------------------------------------------------------------
template <typename T, int R, int MAXB>
__device__ inline void fft_pass(cplx<T>* buf, int line0, int nl, int lstride, int estride,
                                int n, int Ns, const cplx<T>* tw, int sign) {
  const int nbf = n / R;
  const int total = nl * nbf;
  const int step = n / (Ns * R);
  cplx<T> v[MAXB][R];
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    int t = threadIdx.x + c * blockDim.x;
    if (t < total) {
      int line = line0 + t / nbf, j = t % nbf, k = j % Ns;
      cplx<T>* L = buf + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cplx<T> x = L[(j + r * nbf) * estride];
        if (r > 0 && k > 0) {
          cplx<T> w = tw[(k * r * step) % n];
          if (sign < 0) w.y = -w.y;
          x = cmul(x, w);
        }
        v[c][r] = x;
      }
      Dft<T, R>::run(v[c], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    int t = threadIdx.x + c * blockDim.x;
    if (t < total) {
      int line = line0 + t / nbf, j = t % nbf, k = j % Ns;
      cplx<T>* L = buf + line * lstride;
      int idx = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) L[(idx + r * Ns) * estride] = v[c][r];
    }
  }
  __syncthreads();
}

template <typename T, int MAXB>
__device__ inline void lds_fft(cplx<T>* buf, int nlines, int lstride, int estride,
                               const FftPlan& pl, const cplx<T>* tw, int sign) {
  const int n = pl.n;
  // lines per group so that the radix-2 pass (most butterflies) fits MAXB
  int gl = (2 * MAXB * (int)blockDim.x) / n;
  if (gl < 1) gl = 1;
  for (int l0 = 0; l0 < nlines; l0 += gl) {
    int nl = nlines - l0 < gl ? nlines - l0 : gl;
    int Ns = 1;
    for (int s = 0; s < pl.npass; ++s) {
      int R = pl.radix[s];
      switch (R) {
        case 4: fft_pass<T, 4, (MAXB + 1) / 2>(buf, l0, nl, lstride, estride, n, Ns, tw, sign); break;
        case 2: fft_pass<T, 2, MAXB>(buf, l0, nl, lstride, estride, n, Ns, tw, sign); break;
        case 3: fft_pass<T, 3, (2 * MAXB + 2) / 3>(buf, l0, nl, lstride, estride, n, Ns, tw, sign); break;
        case 5: fft_pass<T, 5, (2 * MAXB + 4) / 5>(buf, l0, nl, lstride, estride, n, Ns, tw, sign); break;
        default: fft_pass<T, 7, (2 * MAXB + 6) / 7>(buf, l0, nl, lstride, estride, n, Ns, tw, sign); break;
      }
      Ns *= R;
    }
  }
}

// Quotient by a divisor known only at run time, without the integer divide
// sequence: n / d = (n * m) >> 21 with m = ceil(2^21 / d), exact for
// 0 <= n < 8192 and 1 <= d <= 256 (checked over that whole range).
__host__ __device__ inline unsigned div_magic(unsigned d) { return ((1u << 21) + d - 1) / d; }
__device__ inline int fdiv(int n, unsigned m) {
  return (int)(((unsigned long long)(unsigned)n * m) >> 21);
}

// Split of two FP32 values into bf16 parts for the split bf16x3 products:
// hi = bf16(x) (round to nearest even), lo = bf16(x - hi); each returned as a
// packed pair (element 0 in the low half).
typedef __bf16 qx_bf2 __attribute__((ext_vector_type(2)));
__device__ inline void split_bf16x2(float a, float b, unsigned& hi, unsigned& lo) {
  const qx_bf2 h = {(__bf16)a, (__bf16)b};
  const qx_bf2 l = {(__bf16)(a - (float)h[0]), (__bf16)(b - (float)h[1])};
  hi = __builtin_bit_cast(unsigned, h);
  lo = __builtin_bit_cast(unsigned, l);
}

// load the twiddle table exp(2*pi*i*k/n) from global into LDS
template <typename T>
__device__ inline void load_tw(cplx<T>* dst, const double* src, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) dst[k] = mk<T>((T)src[2 * k], (T)src[2 * k + 1]);
}

}  // namespace qx
