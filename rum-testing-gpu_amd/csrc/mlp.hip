// Learned warm start: a three-layer MLP (in -> hidden -> hidden -> out, ReLU) trained and run on
// the GPU with bf16 matrix cores (v_mfma_f32_16x16x16_bf16: bf16 operands, float32 accumulation).
// No reference counterpart: it maps a bootstrap draw's change of p on the observed cells to a
// correction of the ADMM state there; every draw it touches is still decided by its float64
// certificates.
//
// Parameters live in one float32 vector (the master copy: W1 in x hidden row-major, b1, W2, b2,
// W3 hidden x out, b3); the GEMM kernel rounds both operands to bf16 (round to nearest even) as it
// loads them, so the products are the bf16 products of those values, accumulated in float32.
// Training: mean squared error 1/(2N) sum ||Y - T||^2, backward by the chain rule, Adam on the
// float32 master copy.
#include "rumgpu_internal.hpp"

#include <cmath>
#include <vector>

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);        // round to nearest, ties to even (finite inputs)
  return (short)(u >> 16);
}

enum { EPI_ACC = 0, EPI_BIAS = 1, EPI_BIAS_RELU = 2, EPI_MASK = 3 };

struct GemmArgs {
  const float* A;
  i64 sa0, sa1;          // A(i, k) = A[i sa0 + k sa1], i < M, k < K
  const float* B;
  i64 sb0, sb1;          // B(k, j) = B[k sb0 + j sb1]
  float* C;
  i64 ldc;               // C(i, j) = C[i ldc + j]
  int M, N, K;
  const float* bias;     // per column (EPI_BIAS, EPI_BIAS_RELU)
  const float* mask;     // EPI_MASK: C = acc * (mask(i, j) > 0), mask row stride ldc
  int epi;
};

// One wavefront per 16 x 16 tile of C. Lane l holds A(row l % 16, k0 + 4 (l / 16) + t) and
// B(k0 + 4 (l / 16) + t, col l % 16), t < 4, and gets C(4 (l / 16) + t, l % 16). Out-of-range
// rows, columns and k are read as zero, so any M, N, K work.
__global__ __launch_bounds__(64) void mlp_gemm(GemmArgs g) {
  const int lane = threadIdx.x;
  const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 16;
  const int r = lane & 15, q = lane >> 4;
  const int ia = i0 + r, jb = j0 + r;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = k0 + 4 * q + t;
      const bool kin = k < g.K;
      const float av = (kin && ia < g.M) ? g.A[(i64)ia * g.sa0 + (i64)k * g.sa1] : 0.0f;
      const float bv = (kin && jb < g.N) ? g.B[(i64)k * g.sb0 + (i64)jb * g.sb1] : 0.0f;
      a[t] = f2bf(av);
      b[t] = f2bf(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = i0 + 4 * q + t, j = j0 + r;
    if (i < g.M && j < g.N) {
      float v = acc[t];
      if (g.epi == EPI_BIAS || g.epi == EPI_BIAS_RELU) v = v + g.bias[j];
      if (g.epi == EPI_BIAS_RELU) v = v > 0.0f ? v : 0.0f;
      if (g.epi == EPI_MASK) v = g.mask[(i64)i * g.ldc + j] > 0.0f ? v : 0.0f;
      g.C[(i64)i * g.ldc + j] = v;
    }
  }
}

// dY = (Y - T) / N and the per-row loss 1/(2N) ||Y_i - T_i||^2 (one block per row).
__global__ __launch_bounds__(256) void mlp_loss(const float* Y, const float* T, float* dY, float* row_loss, int out,
                                                float inv_n) {
  __shared__ float sh[256];
  const int i = blockIdx.x;
  float s = 0.0f;
  for (int j = threadIdx.x; j < out; j += 256) {
    const float d = Y[(i64)i * out + j] - T[(i64)i * out + j];
    dY[(i64)i * out + j] = d * inv_n;
    s += d * d;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) row_loss[i] = 0.5f * inv_n * sh[0];
}

// db[j] = sum_i dA(i, j) (one thread per column, rows in order).
__global__ void mlp_colsum(const float* dA, float* db, int N, int cols) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.0f;
  for (int i = 0; i < N; ++i) s += dA[(i64)i * cols + j];
  db[j] = s;
}

// Adam (Kingma and Ba) on the float32 master copy: bias-corrected first and second moments.
__global__ void mlp_adam(float* p, const float* g, float* m, float* v, i64 n, float lr, float b1, float b2, float eps,
                         float c1, float c2) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ void mlp_zero(float* a, i64 n) {
  for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) a[i] = 0.0f;
}

struct rumgpu_mlp {
  rumgpu_ctx* ctx = nullptr;
  int in = 0, hid = 0, out = 0, nmax = 0;
  i64 np = 0;                        // parameter count
  i64 oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0;
  float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  float *X = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *T = nullptr, *dY = nullptr, *dH2 = nullptr,
        *dH1 = nullptr, *rl = nullptr;
  int step = 0;
  int grad_valid = 0;
};

static void mlp_free_all(rumgpu_mlp* q) {
  float* bufs[] = {q->theta, q->grad, q->m, q->v, q->X, q->H1, q->H2, q->Y, q->T, q->dY, q->dH2, q->dH1, q->rl};
  for (float* b : bufs) rg_free(b);
}

static int gemm(hipStream_t st, const float* A, i64 sa0, i64 sa1, const float* B, i64 sb0, i64 sb1, float* C, i64 ldc,
                int M, int N, int K, int epi, const float* bias, const float* mask) {
  GemmArgs g{A, sa0, sa1, B, sb0, sb1, C, ldc, M, N, K, bias, mask, epi};
  const dim3 grid((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16));
  hipLaunchKernelGGL(mlp_gemm, grid, dim3(64), 0, st, g);
  RG_LAUNCH_CHECK();
  return RG_OK;
}

// forward of N rows already in q->X
static int forward(rumgpu_mlp* q, int N) {
  hipStream_t st = q->ctx->stream;
  const float* th = q->theta;
  const int in = q->in, H = q->hid, out = q->out;
  RG_TRY(gemm(st, q->X, in, 1, th + q->oW1, H, 1, q->H1, H, N, H, in, EPI_BIAS_RELU, th + q->ob1, nullptr));
  RG_TRY(gemm(st, q->H1, H, 1, th + q->oW2, H, 1, q->H2, H, N, H, H, EPI_BIAS_RELU, th + q->ob2, nullptr));
  RG_TRY(gemm(st, q->H2, H, 1, th + q->oW3, out, 1, q->Y, out, N, out, H, EPI_BIAS, th + q->ob3, nullptr));
  return RG_OK;
}

extern "C" int rumgpu_mlp_create(rumgpu_ctx* ctx, int in_dim, int hidden, int out_dim, int max_batch,
                                 rumgpu_mlp** out) {
  RG_CHECK(ctx && out, RG_EINVAL, "mlp_create: null argument");
  RG_CHECK(in_dim >= 1 && hidden >= 1 && out_dim >= 1 && max_batch >= 1 && in_dim <= (1 << 20) &&
               hidden <= 8192 && out_dim <= (1 << 20) && max_batch <= (1 << 20),
           RG_EINVAL, "mlp_create: sizes out of range");
  RG_HIP(hipSetDevice(ctx->device));
  rumgpu_mlp* q = new rumgpu_mlp();
  q->ctx = ctx;
  q->in = in_dim, q->hid = hidden, q->out = out_dim, q->nmax = max_batch;
  q->oW1 = 0;
  q->ob1 = q->oW1 + (i64)in_dim * hidden;
  q->oW2 = q->ob1 + hidden;
  q->ob2 = q->oW2 + (i64)hidden * hidden;
  q->oW3 = q->ob2 + hidden;
  q->ob3 = q->oW3 + (i64)hidden * out_dim;
  q->np = q->ob3 + out_dim;
  const i64 N = max_batch;
  struct {
    float** p;
    i64 n;
    const char* what;
  } al[] = {{&q->theta, q->np, "mlp parameters"}, {&q->grad, q->np, "mlp gradient"}, {&q->m, q->np, "adam m"},
            {&q->v, q->np, "adam v"},             {&q->X, N * in_dim, "mlp input"},  {&q->H1, N * hidden, "h1"},
            {&q->H2, N * hidden, "h2"},            {&q->Y, N * out_dim, "mlp output"}, {&q->T, N * out_dim, "target"},
            {&q->dY, N * out_dim, "dY"},           {&q->dH2, N * hidden, "dH2"},       {&q->dH1, N * hidden, "dH1"},
            {&q->rl, N, "row loss"}};
  for (auto& a : al) {
    int rc = rg_malloc((void**)a.p, (size_t)a.n * 4, a.what);
    if (rc != RG_OK) {
      mlp_free_all(q);
      delete q;
      return rc;
    }
  }
  hipStream_t st = ctx->stream;
  hipLaunchKernelGGL(mlp_zero, dim3(256), dim3(256), 0, st, q->theta, q->np);
  hipLaunchKernelGGL(mlp_zero, dim3(256), dim3(256), 0, st, q->m, q->np);
  hipLaunchKernelGGL(mlp_zero, dim3(256), dim3(256), 0, st, q->v, q->np);
  RG_HIP(hipStreamSynchronize(st));
  *out = q;
  return RG_OK;
}

extern "C" int rumgpu_mlp_destroy(rumgpu_mlp* q) {
  if (!q) return RG_OK;
  (void)hipStreamSynchronize(q->ctx->stream);
  mlp_free_all(q);
  delete q;
  return RG_OK;
}

extern "C" int rumgpu_mlp_param_count(rumgpu_mlp* q, int64_t* count) {
  RG_CHECK(q && count, RG_EINVAL, "mlp_param_count: null argument");
  *count = q->np;
  return RG_OK;
}

extern "C" int rumgpu_mlp_set_params(rumgpu_mlp* q, const float* theta) {
  RG_CHECK(q && theta, RG_EINVAL, "mlp_set_params: null argument");
  hipStream_t st = q->ctx->stream;
  RG_HIP(rg_copy(st, q->theta, theta, (size_t)q->np * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mlp_zero, dim3(256), dim3(256), 0, st, q->m, q->np);
  hipLaunchKernelGGL(mlp_zero, dim3(256), dim3(256), 0, st, q->v, q->np);
  RG_LAUNCH_CHECK();
  RG_HIP(hipStreamSynchronize(st));
  q->step = 0;
  q->grad_valid = 0;
  return RG_OK;
}

extern "C" int rumgpu_mlp_get_params(rumgpu_mlp* q, float* theta) {
  RG_CHECK(q && theta, RG_EINVAL, "mlp_get_params: null argument");
  RG_HIP(rg_copy(q->ctx->stream, theta, q->theta, (size_t)q->np * 4, hipMemcpyDeviceToHost));
  return RG_OK;
}

extern "C" int rumgpu_mlp_forward(rumgpu_mlp* q, int N, const float* X, float* Y) {
  RG_CHECK(q && X && Y, RG_EINVAL, "mlp_forward: null argument");
  RG_CHECK(N >= 1 && N <= q->nmax, RG_EINVAL, "mlp_forward: batch %d outside 1..%d", N, q->nmax);
  RG_HIP(hipSetDevice(q->ctx->device));
  hipStream_t st = q->ctx->stream;
  RG_HIP(hipMemcpyAsync(q->X, X, (size_t)N * q->in * 4, hipMemcpyHostToDevice, st));
  RG_TRY(forward(q, N));
  RG_HIP(rg_copy(st, Y, q->Y, (size_t)N * q->out * 4, hipMemcpyDeviceToHost));
  return RG_OK;
}

// Forward and backward of one batch: loss (mean over rows of 1/2 ||Y - T||^2) and the gradient
// of every parameter, kept on the device for rumgpu_mlp_adam and copied to grad when given.
extern "C" int rumgpu_mlp_backward(rumgpu_mlp* q, int N, const float* X, const float* T, float* loss, float* grad) {
  RG_CHECK(q && X && T && loss, RG_EINVAL, "mlp_backward: null argument");
  RG_CHECK(N >= 1 && N <= q->nmax, RG_EINVAL, "mlp_backward: batch %d outside 1..%d", N, q->nmax);
  RG_HIP(hipSetDevice(q->ctx->device));
  hipStream_t st = q->ctx->stream;
  const int in = q->in, H = q->hid, out = q->out;
  RG_HIP(hipMemcpyAsync(q->X, X, (size_t)N * in * 4, hipMemcpyHostToDevice, st));
  RG_HIP(hipMemcpyAsync(q->T, T, (size_t)N * out * 4, hipMemcpyHostToDevice, st));
  RG_TRY(forward(q, N));
  hipLaunchKernelGGL(mlp_loss, dim3(N), dim3(256), 0, st, q->Y, q->T, q->dY, q->rl, out, 1.0f / (float)N);
  float* g = q->grad;
  const float* th = q->theta;
  // layer 3: dW3 = H2^T dY, db3 = colsum dY, dH2 = (dY W3^T) * [H2 > 0]
  RG_TRY(gemm(st, q->H2, 1, H, q->dY, out, 1, g + q->oW3, out, H, out, N, EPI_ACC, nullptr, nullptr));
  hipLaunchKernelGGL(mlp_colsum, dim3((out + 255) / 256), dim3(256), 0, st, q->dY, g + q->ob3, N, out);
  RG_TRY(gemm(st, q->dY, out, 1, th + q->oW3, 1, out, q->dH2, H, N, H, out, EPI_MASK, nullptr, q->H2));
  // layer 2: dW2 = H1^T dA2, db2, dH1 = (dA2 W2^T) * [H1 > 0]
  RG_TRY(gemm(st, q->H1, 1, H, q->dH2, H, 1, g + q->oW2, H, H, H, N, EPI_ACC, nullptr, nullptr));
  hipLaunchKernelGGL(mlp_colsum, dim3((H + 255) / 256), dim3(256), 0, st, q->dH2, g + q->ob2, N, H);
  RG_TRY(gemm(st, q->dH2, H, 1, th + q->oW2, 1, H, q->dH1, H, N, H, H, EPI_MASK, nullptr, q->H1));
  // layer 1: dW1 = X^T dA1, db1
  RG_TRY(gemm(st, q->X, 1, in, q->dH1, H, 1, g + q->oW1, H, in, H, N, EPI_ACC, nullptr, nullptr));
  hipLaunchKernelGGL(mlp_colsum, dim3((H + 255) / 256), dim3(256), 0, st, q->dH1, g + q->ob1, N, H);
  RG_LAUNCH_CHECK();
  std::vector<float> rl(N);
  RG_HIP(rg_copy(st, rl.data(), q->rl, (size_t)N * 4, hipMemcpyDeviceToHost));
  double s = 0.0;
  for (int i = 0; i < N; ++i) s += rl[i];
  *loss = (float)s;
  if (grad) RG_HIP(rg_copy(st, grad, q->grad, (size_t)q->np * 4, hipMemcpyDeviceToHost));
  q->grad_valid = 1;
  return RG_OK;
}

extern "C" int rumgpu_mlp_adam(rumgpu_mlp* q, float lr, float beta1, float beta2, float eps) {
  RG_CHECK(q, RG_EINVAL, "mlp_adam: null handle");
  RG_CHECK(q->grad_valid, RG_EINVAL, "mlp_adam: no gradient (call rumgpu_mlp_backward first)");
  RG_CHECK(lr > 0 && beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1 && eps > 0, RG_EINVAL,
           "mlp_adam: bad hyper-parameters");
  hipStream_t st = q->ctx->stream;
  q->step += 1;
  const float c1 = 1.0f - (float)std::pow((double)beta1, (double)q->step);
  const float c2 = 1.0f - (float)std::pow((double)beta2, (double)q->step);
  hipLaunchKernelGGL(mlp_adam, dim3(grid_for(q->np, 256) > 4096 ? 4096 : grid_for(q->np, 256)), dim3(256), 0, st,
                     q->theta, q->grad, q->m, q->v, q->np, lr, beta1, beta2, eps, c1, c2);
  RG_LAUNCH_CHECK();
  RG_HIP(hipStreamSynchronize(st));
  q->grad_valid = 0;
  return RG_OK;
}

// (A B) with bf16 operands and float32 accumulation, for tests of the matrix-core tile:
// A M x K, B K x N, C M x N, all row-major host arrays.
extern "C" int rumgpu_mlp_gemm_test(rumgpu_ctx* ctx, int M, int N, int K, const float* A, const float* B, float* C) {
  RG_CHECK(ctx && A && B && C && M >= 1 && N >= 1 && K >= 1, RG_EINVAL, "mlp_gemm_test: bad argument");
  hipStream_t st = ctx->stream;
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  RG_TRY(rg_malloc((void**)&dA, (size_t)M * K * 4, "gemm A"));
  int rc = rg_malloc((void**)&dB, (size_t)K * N * 4, "gemm B");
  if (rc == RG_OK) rc = rg_malloc((void**)&dC, (size_t)M * N * 4, "gemm C");
  if (rc == RG_OK && (rg_copy(st, dA, A, (size_t)M * K * 4, hipMemcpyHostToDevice) != hipSuccess ||
                      rg_copy(st, dB, B, (size_t)K * N * 4, hipMemcpyHostToDevice) != hipSuccess))
    rc = RG_EHIP;
  if (rc == RG_OK) rc = gemm(st, dA, K, 1, dB, N, 1, dC, N, M, N, K, EPI_ACC, nullptr, nullptr);
  if (rc == RG_OK && rg_copy(st, C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = RG_EHIP;
  rg_free(dA), rg_free(dB), rg_free(dC);
  if (rc == RG_EHIP) rg_set_error("mlp_gemm_test: HIP error");
  return rc;
}
