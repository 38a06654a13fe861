// coop.h -- helpers of the persistent (cooperatively launched) dense kernels:
// a bounded grid-wide barrier, cross-workgroup loads/stores, and a 64 x 64
// FP64 MFMA tile product staged through LDS.
#pragma once
#include <hip/hip_runtime.h>

namespace qx {

// Cross-workgroup data published between barriers goes through agent-scope
// atomics so that it is coherent across the L2s of the XCDs.
__device__ inline void st_agent(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double ld_agent(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A double published with a tag, for readers that wait on the data itself
// instead of a barrier: two 64-bit words, each 32 bits of the value and the
// 32-bit tag.  Each word is stored and loaded single-copy atomically, so a
// reader that finds the tag in both words has the whole value of that
// publication (the buffer must not hold the tag from earlier use).
__device__ inline void st_tagged(unsigned long long* p, double v, unsigned tag) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  __hip_atomic_store(p, (bits & 0xffffffff00000000ull) | tag, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(p + 1, (bits << 32) | tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long ld_u64_agent(const unsigned long long* p) {
  return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool tag_ok(unsigned long long w0, unsigned long long w1, unsigned tag) {
  return (unsigned)w0 == tag && (unsigned)w1 == tag;
}
__device__ inline double tag_value(unsigned long long w0, unsigned long long w1) {
  return __longlong_as_double((long long)((w0 & 0xffffffff00000000ull) | (w1 >> 32)));
}

// Grid barrier for a kernel launched with hipLaunchCooperativeKernel (every
// workgroup resident).  Data that crosses workgroups must go through
// st_agent / ld_agent (device-scope accesses, coherent across the XCDs'
// L2s): the barrier then only has to wait until this workgroup's stores are
// acknowledged before it arrives, with no L2 write-back or invalidation.
// bar[0] counts arrivals and only grows: the b-th barrier of a launch waits
// until it reaches b * gridDim.x (the counter is cleared by a fill kernel
// before the launch).  The wait is bounded: after about 2^20 polls the
// barrier raises err[0], and every barrier after that returns at once, so a
// launch that cannot complete still ends.  Returns false once err is raised.
__device__ inline bool grid_barrier(int* bar, int target, int* err) {
  __shared__ int s_ok;
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);  // this thread's stores acknowledged
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    if (ok) {
      __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int spins = 0;
      while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(1);
        if ((++spins & 1023) == 0 &&
            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ok = 0;
          break;
        }
        if (spins > (1 << 20)) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
      }
    }
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ---- 64 x 64 FP64 MFMA tile: acc += A (64 x K) * B (K x 64) ---------------
// 256 threads = 4 wavefronts, each a 32 x 32 quadrant as 2 x 2 blocks of
// v_mfma_f64_16x16x4f64.  fa(m, k) and fb(k, n) return the operands, 0 out of
// range, and must read a valid address for every m, n < 64 and k < K (they
// load unconditionally from a clamped index and select, so that the loads of
// a stage are all in flight together); K is walked 16 at a time through the
// LDS tiles As/Bs.
typedef double dv4 __attribute__((ext_vector_type(4)));
constexpr int kTileLd = 66;

struct Tile64 {
  dv4 acc[2][2];
  __device__ void zero() {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = dv4{0.0, 0.0, 0.0, 0.0};
  }
  template <typename FA, typename FB>
  __device__ void run(int K, FA fa, FB fb, double (*As)[kTileLd], double (*Bs)[kTileLd]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int m = tid >> 2, kk = (tid & 3) * 4;
    double ra[4], rb[4];
    auto fetch = [&](int k0) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + kk + u, kc = min(k, K - 1);
        const double va = fa(m, kc), vb = fb(kc, m);
        ra[u] = k < K ? va : 0.0;
        rb[u] = k < K ? vb : 0.0;
      }
    };
    fetch(0);
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        As[kk + u][m] = ra[u];
        Bs[kk + u][m] = rb[u];
      }
      __syncthreads();
      if (k0 + 16 < K) fetch(k0 + 16);  // next chunk in flight during the MFMAs
#pragma unroll
      for (int k4 = 0; k4 < 16; k4 += 4) {
        const int kr = k4 + (lane >> 4), c = lane & 15;
        const double a0 = As[kr][wm + c], a1 = As[kr][wm + 16 + c];
        const double b0 = Bs[kr][wn + c], b1 = Bs[kr][wn + 16 + c];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
      __syncthreads();
    }
  }
  // visit every (m, n, value) of the tile held by this thread
  template <typename F>
  __device__ void each(F f) const {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          f(wm + i * 16 + (lane >> 4) + 4 * v, wn + j * 16 + (lane & 15), acc[i][j][v]);
  }
};

}  // namespace qx
