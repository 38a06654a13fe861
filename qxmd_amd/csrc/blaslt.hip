// blaslt.hip -- the large plain contractions on hipBLASLt, kept as a
// cross-check of our own MFMA kernels (mfma_gemm.hip): it runs only when the
// environment sets QX_GEMM_LIB=hipblaslt (tests compare the two).
//
// For each shape the library's candidate algorithms are timed once per
// context on the first call and the fastest is cached.  FP64 operands of the
// FP32 modes are rounded to FP32 copies; mode 2 is the library's fast FP32
// compute (split products on the matrix cores).
#include <hipblaslt/hipblaslt.h>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include "context.h"

namespace qx {

namespace {

struct LtPlan {
  bool ok = false;
  hipblasLtMatmulDesc_t desc = nullptr;
  hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
  hipblasLtMatmulAlgo_t algo{};
};

typedef std::tuple<int, int, int, int, int, int, int, int> PlanKey;  // ta M N K lda ldb ldc mode

constexpr int kScratch = 3;

struct LtState {
  hipblasLtHandle_t h = nullptr;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  std::map<PlanKey, LtPlan> plans;
  float* f32[kScratch] = {};  // operand copies of lt_sgemm_mixed
  size_t f32_len[kScratch] = {};
};

constexpr size_t kWorkspace = 128u << 20;

LtState* state(Ctx* c) {
  if (!c->blaslt) {
    LtState* s = new LtState();
    if (hipblasLtCreate(&s->h) != HIPBLAS_STATUS_SUCCESS) {
      delete s;
      return nullptr;
    }
    if (hipMalloc(&s->ws, kWorkspace) == hipSuccess) s->ws_bytes = kWorkspace;
    c->blaslt = s;
  }
  return static_cast<LtState*>(c->blaslt);
}

float* scratch(LtState* s, int i, size_t n) {
  if (s->f32_len[i] < n) {
    if (s->f32[i]) (void)hipFree(s->f32[i]);
    s->f32[i] = nullptr;
    s->f32_len[i] = 0;
    if (hipMalloc(&s->f32[i], sizeof(float) * n) != hipSuccess) return nullptr;
    s->f32_len[i] = n;
  }
  return s->f32[i];
}

unsigned cdivu(long a, long b) { return (unsigned)((a + b - 1) / b); }

// mode 0: FP64; 1: FP32 with FP32 accumulation; 2: FP32 inputs, the
// library's fast FP32 compute (split products on the matrix cores)
LtPlan* plan(Ctx* c, LtState* s, bool ta, int M, int N, int K, int lda, int ldb, int ldc,
             int mode, const void* A, const void* B, void* C) {
  PlanKey key{ta, M, N, K, lda, ldb, ldc, mode};
  auto it = s->plans.find(key);
  if (it != s->plans.end()) return it->second.ok ? &it->second : nullptr;
  LtPlan& p = s->plans[key];
  const hipDataType dt = mode == 0 ? HIP_R_64F : HIP_R_32F;
  const hipblasComputeType_t ct = mode == 0 ? HIPBLAS_COMPUTE_64F
                                  : mode == 1 ? HIPBLAS_COMPUTE_32F
                                              : HIPBLAS_COMPUTE_32F_FAST_TF32;
  if (hipblasLtMatmulDescCreate(&p.desc, ct, dt) != HIPBLAS_STATUS_SUCCESS) return nullptr;
  hipblasOperation_t opA = ta ? HIPBLAS_OP_T : HIPBLAS_OP_N, opB = HIPBLAS_OP_N;
  hipblasLtMatmulDescSetAttribute(p.desc, HIPBLASLT_MATMUL_DESC_TRANSA, &opA, sizeof(opA));
  hipblasLtMatmulDescSetAttribute(p.desc, HIPBLASLT_MATMUL_DESC_TRANSB, &opB, sizeof(opB));
  if (hipblasLtMatrixLayoutCreate(&p.la, dt, ta ? K : M, ta ? M : K, lda) != HIPBLAS_STATUS_SUCCESS ||
      hipblasLtMatrixLayoutCreate(&p.lb, dt, K, N, ldb) != HIPBLAS_STATUS_SUCCESS ||
      hipblasLtMatrixLayoutCreate(&p.lc, dt, M, N, ldc) != HIPBLAS_STATUS_SUCCESS)
    return nullptr;
  hipblasLtMatmulPreference_t pref;
  if (hipblasLtMatmulPreferenceCreate(&pref) != HIPBLAS_STATUS_SUCCESS) return nullptr;
  size_t wsz = s->ws_bytes;
  hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &wsz,
                                        sizeof(wsz));
  hipblasLtMatmulHeuristicResult_t heur[8];
  int nres = 0;
  hipblasStatus_t hs = hipblasLtMatmulAlgoGetHeuristic(s->h, p.desc, p.la, p.lb, p.lc, p.lc, pref,
                                                       8, heur, &nres);
  hipblasLtMatmulPreferenceDestroy(pref);
  if (hs != HIPBLAS_STATUS_SUCCESS || nres < 1) return nullptr;
  // time the candidates on the caller's operands, into a temporary result
  // (the caller's C may hold data to accumulate into)
  const bool want_tune = nres > 1 && !std::getenv("QX_GEMM_NOTUNE");
  void* tmp = nullptr;
  if (want_tune &&
      hipMalloc(&tmp, (mode == 0 ? sizeof(double) : sizeof(float)) * (size_t)ldc * N) != hipSuccess)
    tmp = nullptr;
  (void)C;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const double one64 = 1.0, zero64 = 0.0;
  const float one32 = 1.0f, zero32 = 0.0f;
  const void* alpha = mode == 0 ? (const void*)&one64 : (const void*)&one32;
  const void* beta = mode == 0 ? (const void*)&zero64 : (const void*)&zero32;
  float best = 1e30f;
  int besti = -1;
  const bool tune = want_tune && tmp != nullptr;
  for (int a = 0; a < nres; ++a) {
    if (heur[a].state != HIPBLAS_STATUS_SUCCESS) continue;
    if (!tune) {
      besti = a;
      break;
    }
    if (hipblasLtMatmul(s->h, p.desc, alpha, A, p.la, B, p.lb, beta, tmp, p.lc, tmp, p.lc,
                        &heur[a].algo, s->ws, s->ws_bytes, c->stream) != HIPBLAS_STATUS_SUCCESS)
      continue;
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < 2; ++r)
      hipblasLtMatmul(s->h, p.desc, alpha, A, p.la, B, p.lb, beta, tmp, p.lc, tmp, p.lc,
                      &heur[a].algo, s->ws, s->ws_bytes, c->stream);
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) {
      best = ms;
      besti = a;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (tmp) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
  }
  if (besti < 0) return nullptr;
  p.algo = heur[besti].algo;
  p.ok = true;
  return &p;
}

// dst(rows x cols, ldd) = (float) src(rows x cols, lds), column-major
__global__ void k_to_f32(const double* __restrict__ src, int lds, float* __restrict__ dst, int ldd,
                         int rows, long cols) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const int r = (int)(i % rows);
  const long col = i / rows;
  dst[(size_t)col * ldd + r] = (float)src[(size_t)col * lds + r];
}

// C(m, n) = beta C + alpha R(m, n) in FP64; beta == 0 overwrites
__global__ void k_f32_axpby(const float* __restrict__ R, int M, long N, double alpha, double beta,
                            double* __restrict__ C, int ldc) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  const int m = (int)(i % M);
  const long n = i / M;
  double* p = C + (size_t)n * ldc + m;
  const double v = alpha * (double)R[i];
  *p = beta == 0.0 ? v : beta * *p + v;
}

}  // namespace

bool lt_enabled() { return gemm_path() == GEMM_LIB; }

// FP64: C = alpha op(A) B + beta C.  Returns < 0 when hipBLASLt cannot take
// the shape (the caller falls back to its own kernels).
int lt_dgemm(Ctx* c, bool ta, int M, int N, int K, double alpha, const double* A, int lda,
             const double* B, int ldb, double beta, double* C, int ldc) {
  if (!lt_enabled()) return -1;
  LtState* s = state(c);
  if (!s) return -1;
  LtPlan* p = plan(c, s, ta, M, N, K, lda, ldb, ldc, 0, A, B, C);
  if (!p) return -1;
  lt_mark(c, c->lt_gemm_id, false);
  hipblasStatus_t st = hipblasLtMatmul(s->h, p->desc, &alpha, A, p->la, B, p->lb, &beta, C, p->lc,
                                       C, p->lc, &p->algo, s->ws, s->ws_bytes, c->stream);
  lt_mark(c, c->lt_gemm_id, true);
  return st == HIPBLAS_STATUS_SUCCESS ? 0 : -2;
}

// the FP32 compute mode of the mixed products: QX_MIXED_GEMM=f32 keeps plain
// FP32 products; the default is the library's fast FP32 mode (split products
// on the matrix cores)
static int mixed_mode() {
  const char* mode_env = std::getenv("QX_MIXED_GEMM");
  return mode_env && std::strcmp(mode_env, "f32") == 0 ? 1 : 2;
}

// FP32 in and out: C = op(A) B
int lt_sgemm_f32(Ctx* c, bool ta, int M, int N, int K, const float* A, int lda, const float* B,
                 int ldb, float* C, int ldc) {
  if (!lt_enabled()) return -1;
  LtState* s = state(c);
  if (!s) return -1;
  LtPlan* p = plan(c, s, ta, M, N, K, lda, ldb, ldc, mixed_mode(), A, B, C);
  if (!p) return -1;
  const float one = 1.0f, zero = 0.0f;
  lt_mark(c, c->lt_gemm_id, false);
  hipblasStatus_t st = hipblasLtMatmul(s->h, p->desc, &one, A, p->la, B, p->lb, &zero, C, p->lc, C,
                                       p->lc, &p->algo, s->ws, s->ws_bytes, c->stream);
  lt_mark(c, c->lt_gemm_id, true);
  return st == HIPBLAS_STATUS_SUCCESS ? 0 : -2;
}

// mixed: FP32 copies of A and B, FP32 product R, then C = beta C + alpha R
int lt_sgemm_mixed(Ctx* c, bool ta, int M, int N, int K, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc) {
  if (!lt_enabled()) return -1;
  LtState* s = state(c);
  if (!s) return -1;
  const int arows = ta ? K : M, acols = ta ? M : K;
  const float* A32;
  int lda32 = arows;
  if (A == c->d_B && lda == c->nplwex && c->d_B) {
    A32 = mg_kb_b32(c);
    if (!A32) return -3;
    lda32 = c->nplwex;
  } else {
    float* a = scratch(s, 0, (size_t)arows * acols);
    if (!a) return -3;
    hipLaunchKernelGGL(k_to_f32, dim3(cdivu((long)arows * acols, 256)), dim3(256), 0, c->stream, A,
                       lda, a, arows, arows, (long)acols);
    A32 = a;
  }
  float* b = scratch(s, 1, (size_t)K * N);
  float* r = scratch(s, 2, (size_t)M * N);
  if (!b || !r) return -3;
  hipLaunchKernelGGL(k_to_f32, dim3(cdivu((long)K * N, 256)), dim3(256), 0, c->stream, B, ldb, b,
                     K, K, (long)N);
  LtPlan* p = plan(c, s, ta, M, N, K, lda32, K, M, mixed_mode(), A32, b, r);
  if (!p) return -1;
  const float one = 1.0f, zero = 0.0f;
  lt_mark(c, c->lt_gemm_id, false);
  hipblasStatus_t st = hipblasLtMatmul(s->h, p->desc, &one, A32, p->la, b, p->lb, &zero, r, p->lc,
                                       r, p->lc, &p->algo, s->ws, s->ws_bytes, c->stream);
  lt_mark(c, c->lt_gemm_id, true);
  if (st != HIPBLAS_STATUS_SUCCESS) return -2;
  hipLaunchKernelGGL(k_f32_axpby, dim3(cdivu((long)M * N, 256)), dim3(256), 0, c->stream, r, M,
                     (long)N, alpha, beta, C, ldc);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

void lt_destroy(Ctx* c) {
  if (!c->blaslt) return;
  LtState* s = static_cast<LtState*>(c->blaslt);
  for (auto& kv : s->plans) {
    LtPlan& p = kv.second;
    if (p.la) hipblasLtMatrixLayoutDestroy(p.la);
    if (p.lb) hipblasLtMatrixLayoutDestroy(p.lb);
    if (p.lc) hipblasLtMatrixLayoutDestroy(p.lc);
    if (p.desc) hipblasLtMatmulDescDestroy(p.desc);
  }
  for (int i = 0; i < kScratch; ++i)
    if (s->f32[i]) (void)hipFree(s->f32[i]);
  if (s->ws) (void)hipFree(s->ws);
  if (s->h) hipblasLtDestroy(s->h);
  delete s;
  c->blaslt = nullptr;
}

}  // namespace qx
