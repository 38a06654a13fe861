// fullgrid.hip -- dense 3-D FFT on the full grid, behind the rfft3/fft3
// drop-ins (Sources/fft3.f90:113-333): x1/N on the forward transform, the r2c
// result returned as the full Hermitian grid, c2r reading only i < n1/2+1.
// Lines along each axis are transformed in LDS by the same Stockham engine as
// the pencil path; blocks take runs of lines adjacent in x so that strided
// (y, z) lines are still read in coalesced rows.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <mutex>
#include "context.h"

namespace qx {

constexpr int kFgThreads = 256;

// axis 0: lines along x, block = `nl` consecutive lines (rows)
// axis 1/2: lines along y or z; block = `nl` consecutive x values of one
//           (z) resp. (y) index.
__global__ void __launch_bounds__(kFgThreads) k_fg_lines(double2* __restrict__ g, int n1, int n2,
                                                         int n3, int axis, int nl, FftPlan pl,
                                                         const double* __restrict__ twtab,
                                                         int sign, double scale) {
  extern __shared__ unsigned char smem_raw[];
  cplx<double>* tw = reinterpret_cast<cplx<double>*>(smem_raw);
  const int n = pl.n;
  cplx<double>* lines = tw + n;
  load_tw<double>(tw, twtab, n);
  long base, estride;
  int lcount;
  if (axis == 0) {
    long l0 = (long)blockIdx.x * nl;
    long nlines = (long)n2 * n3;
    lcount = (int)std::min<long>(nl, nlines - l0);
    base = l0 * n1;
    estride = 1;
    for (int i = threadIdx.x; i < lcount * n; i += blockDim.x) {
      double2 v = g[base + i];
      lines[i] = mk<double>(v.x, v.y);
    }
  } else {
    int x0 = (blockIdx.x % ((n1 + nl - 1) / nl)) * nl;
    int other = blockIdx.x / ((n1 + nl - 1) / nl);  // z for axis 1, y for axis 2
    lcount = std::min(nl, n1 - x0);
    estride = axis == 1 ? n1 : (long)n1 * n2;
    base = axis == 1 ? x0 + (long)other * n1 * n2 : x0 + (long)other * n1;
    for (int i = threadIdx.x; i < lcount * n; i += blockDim.x) {
      int t = i / lcount, x = i % lcount;
      double2 v = g[base + (long)t * estride + x];
      lines[x * n + t] = mk<double>(v.x, v.y);
    }
  }
  __syncthreads();
  lds_fft<double, 4>(lines, lcount, n, 1, pl, tw, sign);
  if (axis == 0) {
    for (int i = threadIdx.x; i < lcount * n; i += blockDim.x) {
      cplx<double> v = lines[i];
      g[base + i] = make_double2(v.x * scale, v.y * scale);
    }
  } else {
    for (int i = threadIdx.x; i < lcount * n; i += blockDim.x) {
      int t = i / lcount, x = i % lcount;
      cplx<double> v = lines[x * n + t];
      g[base + (long)t * estride + x] = make_double2(v.x * scale, v.y * scale);
    }
  }
}

__global__ void k_fg_pack(const double* __restrict__ x, const double* __restrict__ y, double2* g,
                          long N) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) g[i] = make_double2(x[i], y ? y[i] : 0.0);
}

// c2r input: Hermitian completion from the half grid i < n1/2+1
__global__ void k_fg_hermitian(const double* __restrict__ x, const double* __restrict__ y,
                               double2* g, int n1, int n2, int n3) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long N = (long)n1 * n2 * n3;
  if (i >= N) return;
  int ix = (int)(i % n1);
  long r = i / n1;
  int iy = (int)(r % n2), iz = (int)(r / n2);
  const int n1h = n1 / 2 + 1;
  if (ix < n1h) {
    g[i] = make_double2(x[i], y[i]);
  } else {
    int mx = n1 - ix, my = (n2 - iy) % n2, mz = (n3 - iz) % n3;
    long m = mx + (long)n1 * (my + (long)n2 * mz);
    g[i] = make_double2(x[m], -y[m]);
  }
}

__global__ void k_fg_unpack(const double2* __restrict__ g, double* x, double* y, long N,
                            int real_only) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double2 v = g[i];
  x[i] = v.x;
  y[i] = real_only ? 0.0 : v.y;
}

int fg_prepare(FullGrid& f, int n1, int n2, int n3, bool staging) {
  if (f.n1 == n1 && f.n2 == n2 && f.n3 == n3 && (!staging || f.d_x)) return 0;
  fg_release(f);
  const int dims[3] = {n1, n2, n3};
  for (int a = 0; a < 3; ++a) {
    f.p[a] = make_plan(dims[a]);
    if (f.p[a].npass < 0 || dims[a] > 2 * 4 * kFgThreads) return -6;
    std::vector<double> tw(2 * dims[a]);
    for (int k = 0; k < dims[a]; ++k) {
      double ang = 2.0 * M_PI * (double)k / (double)dims[a];
      tw[2 * k] = std::cos(ang);
      tw[2 * k + 1] = std::sin(ang);
    }
    QX_CHECK(hipMalloc(&f.d_tw[a], sizeof(double) * tw.size()));
    QX_CHECK(hipMemcpy(f.d_tw[a], tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
  }
  long N = (long)n1 * n2 * n3;
  QX_CHECK(hipMalloc(&f.d_g, sizeof(double2) * N));
  if (staging) {
    QX_CHECK(hipMalloc(&f.d_x, sizeof(double) * N));
    QX_CHECK(hipMalloc(&f.d_y, sizeof(double) * N));
  }
  f.n1 = n1;
  f.n2 = n2;
  f.n3 = n3;
  return 0;
}

void fg_release(FullGrid& f) {
  if (f.d_g) (void)hipFree(f.d_g);
  if (f.d_x) (void)hipFree(f.d_x);
  if (f.d_y) (void)hipFree(f.d_y);
  for (auto& t : f.d_tw)
    if (t) (void)hipFree(t);
  f = FullGrid{};
}

// lines per block: small enough that a 90^3 grid gives ~500 blocks per pass
// on 256 CUs, large enough that the strided passes read rows of nl x 16 bytes
static int fg_lines_per_block(int n) {
  static const int cap = [] {
    const char* e = std::getenv("QX_FG_NL");
    const int v = e ? std::atoi(e) : 16;
    return v >= 1 && v <= 64 ? v : 16;
  }();
  return std::max(1, std::min(cap, 4096 / n));
}

int fg_transform(FullGrid& f, hipStream_t st, int sign, double scale) {
  const int n1 = f.n1, n2 = f.n2, n3 = f.n3;
  const int dims[3] = {n1, n2, n3};
  for (int a = 0; a < 3; ++a) {
    const int n = dims[a];
    const int nl = fg_lines_per_block(n);
    unsigned blocks;
    if (a == 0) blocks = (unsigned)(((long)n2 * n3 + nl - 1) / nl);
    else blocks = (unsigned)(((n1 + nl - 1) / nl) * (a == 1 ? n3 : n2));
    size_t lds = sizeof(cplx<double>) * ((size_t)n + (size_t)nl * n);
    double s = (a == 2) ? scale : 1.0;
    if (lds > 65536)
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k_fg_lines),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_fg_lines, dim3(blocks), dim3(kFgThreads), lds, st, f.d_g, n1, n2, n3, a,
                       nl, f.p[a], f.d_tw[a], sign, s);
  }
  QX_CHECK(hipGetLastError());
  return 0;
}

// the drop-ins' own grid and stream (no context involved)
static FullGrid g_fg;
static hipStream_t g_fg_stream = nullptr;

int dropin_stream(hipStream_t* st) {
  if (!g_fg_stream) QX_CHECK(hipStreamCreateWithFlags(&g_fg_stream, hipStreamNonBlocking));
  *st = g_fg_stream;
  return 0;
}

// releases the drop-ins' grid and stream (qx_shutdown); the next drop-in
// call creates them again
void dropin_release() {
  if (g_fg_stream) (void)hipStreamSynchronize(g_fg_stream);
  fg_release(g_fg);
  if (g_fg_stream) (void)hipStreamDestroy(g_fg_stream);
  g_fg_stream = nullptr;
}

int dropin_grid(int n1, int n2, int n3, FullGrid** f, hipStream_t* st) {
  int rc = dropin_stream(st);
  if (rc) return rc;
  rc = fg_prepare(g_fg, n1, n2, n3, true);
  if (rc) return rc;
  *f = &g_fg;
  return 0;
}

int fullgrid_rfft3(int key, double* fx, double* fy, int n1, int n2, int n3) {
  FullGrid* f = nullptr;
  hipStream_t st = nullptr;
  int rc = dropin_grid(n1, n2, n3, &f, &st);
  if (rc) return rc;
  const long N = (long)n1 * n2 * n3;
  const unsigned nb = (unsigned)((N + 255) / 256);
  QX_CHECK(hipMemcpyAsync(f->d_x, fx, sizeof(double) * N, hipMemcpyHostToDevice, st));
  if (key == 1) {
    hipLaunchKernelGGL(k_fg_pack, dim3(nb), dim3(256), 0, st, f->d_x, nullptr, f->d_g, N);
    if ((rc = fg_transform(*f, st, -1, 1.0 / (double)N))) return rc;
    hipLaunchKernelGGL(k_fg_unpack, dim3(nb), dim3(256), 0, st, f->d_g, f->d_x,
                       f->d_y, N, 0);
  } else {
    QX_CHECK(hipMemcpyAsync(f->d_y, fy, sizeof(double) * N, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fg_hermitian, dim3(nb), dim3(256), 0, st, f->d_x, f->d_y,
                       f->d_g, n1, n2, n3);
    if ((rc = fg_transform(*f, st, +1, 1.0))) return rc;
    hipLaunchKernelGGL(k_fg_unpack, dim3(nb), dim3(256), 0, st, f->d_g, f->d_x,
                       f->d_y, N, 1);
  }
  QX_CHECK(hipMemcpyAsync(fx, f->d_x, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipMemcpyAsync(fy, f->d_y, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipStreamSynchronize(st));
  return 0;
}

// ---- rfft3 on device arrays without allocation ----------------------------
// The complex work grid and the three twiddle tables live in one file-scope
// device array, sized for grids of up to kNaMaxPoints points with each axis
// at most kNaMaxAxis.  Every call on a GPU uses that GPU's copy; calls that
// overlap on one GPU race.  The array's device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and kept in
// g_na_ptr: in host code the array's name is not a device address.
constexpr long kNaMaxPoints = 1L << 21;  // 128^3
constexpr int kNaMaxAxis = 2 * 4 * kFgThreads;
constexpr long kNaDoubles = 2 * kNaMaxPoints + 3L * 2 * kNaMaxAxis;
static __device__ __align__(256) double scratch[kNaDoubles];

constexpr int kNaMaxDev = 64;
static double* g_na_ptr[kNaMaxDev] = {};
static std::mutex g_na_mu;

static int noalloc_scratch(double** p) {
  int dev = 0;
  QX_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNaMaxDev) return -1;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void* a = nullptr;
    QX_CHECK(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = static_cast<double*>(a);
  }
  *p = g_na_ptr[dev];
  return 0;
}

int noalloc_limits(long* max_points, int* max_axis) {
  *max_points = kNaMaxPoints;
  *max_axis = kNaMaxAxis;
  return 0;
}

int fullgrid_rfft3_dev(int key, double* dx, double* dy, int n1, int n2, int n3, hipStream_t st) {
  const int dims[3] = {n1, n2, n3};
  for (int a = 0; a < 3; ++a)
    if (dims[a] < 1 || dims[a] > kNaMaxAxis) return -2;
  const long N = (long)n1 * n2 * n3;
  if (N > kNaMaxPoints) return -2;
  if (!dx || !dy) return -1;
  FullGrid f;  // views into the scratch array; nothing here is freed
  for (int a = 0; a < 3; ++a) {
    f.p[a] = make_plan(dims[a]);
    if (f.p[a].npass < 0) return -6;
  }
  double* base = nullptr;
  int rc = noalloc_scratch(&base);
  if (rc) return rc;
  f.d_g = reinterpret_cast<double2*>(base);
  // the twiddles of fg_prepare, cos/sin of 2 pi k / n from the host libm,
  // staged through a static host table (the call waits for the copies)
  static double tw_host[3 * 2 * kNaMaxAxis];
  std::lock_guard<std::mutex> lk(g_na_mu);
  for (int a = 0; a < 3; ++a) {
    f.d_tw[a] = base + 2 * kNaMaxPoints + (long)a * 2 * kNaMaxAxis;
    double* t = tw_host + (long)a * 2 * kNaMaxAxis;
    for (int k = 0; k < dims[a]; ++k) {
      const double ang = 2.0 * M_PI * (double)k / (double)dims[a];
      t[2 * k] = std::cos(ang);
      t[2 * k + 1] = std::sin(ang);
    }
    QX_CHECK(hipMemcpyAsync(f.d_tw[a], t, sizeof(double) * 2 * dims[a], hipMemcpyHostToDevice,
                            st));
  }
  f.n1 = n1;
  f.n2 = n2;
  f.n3 = n3;
  const unsigned nb = (unsigned)((N + 255) / 256);
  if (key == 1) {
    hipLaunchKernelGGL(k_fg_pack, dim3(nb), dim3(256), 0, st, dx, nullptr, f.d_g, N);
    if ((rc = fg_transform(f, st, -1, 1.0 / (double)N))) return rc;
    hipLaunchKernelGGL(k_fg_unpack, dim3(nb), dim3(256), 0, st, f.d_g, dx, dy, N, 0);
  } else {
    hipLaunchKernelGGL(k_fg_hermitian, dim3(nb), dim3(256), 0, st, dx, dy, f.d_g, n1, n2, n3);
    if ((rc = fg_transform(f, st, +1, 1.0))) return rc;
    hipLaunchKernelGGL(k_fg_unpack, dim3(nb), dim3(256), 0, st, f.d_g, dx, dy, N, 1);
  }
  QX_CHECK(hipGetLastError());
  QX_CHECK(hipStreamSynchronize(st));
  return 0;
}

int fullgrid_fft3(int key, double* fx, double* fy, int n1, int n2, int n3) {
  FullGrid* f = nullptr;
  hipStream_t st = nullptr;
  int rc = dropin_grid(n1, n2, n3, &f, &st);
  if (rc) return rc;
  const long N = (long)n1 * n2 * n3;
  const unsigned nb = (unsigned)((N + 255) / 256);
  QX_CHECK(hipMemcpyAsync(f->d_x, fx, sizeof(double) * N, hipMemcpyHostToDevice, st));
  QX_CHECK(hipMemcpyAsync(f->d_y, fy, sizeof(double) * N, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_fg_pack, dim3(nb), dim3(256), 0, st, f->d_x, f->d_y, f->d_g, N);
  if ((rc = fg_transform(*f, st, key == 1 ? -1 : +1, key == 1 ? 1.0 / (double)N : 1.0))) return rc;
  hipLaunchKernelGGL(k_fg_unpack, dim3(nb), dim3(256), 0, st, f->d_g, f->d_x, f->d_y,
                     N, 0);
  QX_CHECK(hipMemcpyAsync(fx, f->d_x, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipMemcpyAsync(fy, f->d_y, sizeof(double) * N, hipMemcpyDeviceToHost, st));
  QX_CHECK(hipStreamSynchronize(st));
  return 0;
}

}  // namespace qx
