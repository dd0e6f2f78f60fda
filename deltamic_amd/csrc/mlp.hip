// Per-vertex proposal MLP (15 -> 128 -> 128 -> 128 -> 3, ReLU) for gfx950:
// inference, backpropagation and Adam, with bf16 weights and FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x16_bf16).
//
// No counterpart in the reference: the learned mesh-displacement proposal
// is new in this build.  Every GEMM has one wave per 16x16 output tile that
// walks K in order; the weight gradients (K = number of vertices, few output
// tiles) split K into 256-row chunks whose partial tiles are added in chunk
// order, so every result is reduced in a fixed order without atomics.
#include "common.h"

namespace dmic {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f32_to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
  return __uint_as_float(((unsigned)h) << 16);
}

template <bool BF16>
__device__ __forceinline__ float load_elem(const void* p, int64_t i) {
  if (BF16) return bf16_to_f32(static_cast<const unsigned short*>(p)[i]);
  return static_cast<const float*>(p)[i];
}

enum Epilogue { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_RELU = 2, EPI_RELU_MASK = 3 };

// C[M,N] = A(M,K) B(K,N) with A(i,k) = A[i*sa0 + k*sa1], B(k,j) = B[k*sb0 + j*sb1].
// MFMA 16x16x16 bf16 lane layout (wave64):
//   A operand: lane l holds A(l%16, 4*(l/16) + e), e = 0..3
//   B operand: lane l holds B(4*(l/16) + e, l%16)
//   C result : lane l holds C(4*(l/16) + e, l%16)
template <bool ABF, bool BBF>
__global__ __launch_bounds__(64) void gemm_bf16_kernel(const void* __restrict__ A, int64_t sa0,
                                                       int64_t sa1, const void* __restrict__ B,
                                                       int64_t sb0, int64_t sb1,
                                                       float* __restrict__ C, int64_t ldc,
                                                       int64_t M, int64_t N, int64_t K,
                                                       const float* __restrict__ bias,
                                                       const float* __restrict__ mask,
                                                       int64_t ldm, int epi) {
  const int l = threadIdx.x;
  // rows (one per vertex: the large dimension) on grid.x, columns on grid.y
  const int64_t m0 = (int64_t)blockIdx.x * 16, n0 = (int64_t)blockIdx.y * 16;
  const int64_t ar = m0 + (l & 15);
  const int64_t bc = n0 + (l & 15);
  const int kq = 4 * (l >> 4);
  floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t k = k0 + kq + e;
      const float av = (ar < M && k < K) ? load_elem<ABF>(A, ar * sa0 + k * sa1) : 0.0f;
      const float bv = (bc < N && k < K) ? load_elem<BBF>(B, k * sb0 + bc * sb1) : 0.0f;
      a[e] = f32_to_bf16(av);
      b[e] = f32_to_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int64_t r = m0 + kq + e, c = bc;
    if (r >= M || c >= N) continue;
    float v = acc[e];
    if (epi == EPI_BIAS || epi == EPI_BIAS_RELU) v += bias[c];
    if (epi == EPI_BIAS_RELU) v = v > 0.0f ? v : 0.0f;
    if (epi == EPI_RELU_MASK) v = (mask[r * ldm + c] > 0.0f) ? v : 0.0f;
    C[r * ldc + c] = v;
  }
}

// Split-K partial of C = A B for the weight gradients (K = rows, large;
// few output tiles): wave (x, y, z) computes the 16x16 tile (x, y) over
// k in [z kchunk, (z + 1) kchunk) into P[z][M][N]; splitk_reduce_kernel adds
// the partials in z order.  Same operand layout as gemm_bf16_kernel.
template <bool ABF, bool BBF>
__global__ __launch_bounds__(64) void gemm_bf16_splitk_kernel(
    const void* __restrict__ A, int64_t sa0, int64_t sa1, const void* __restrict__ B, int64_t sb0,
    int64_t sb1, float* __restrict__ P, int64_t M, int64_t N, int64_t K, int64_t kchunk) {
  const int l = threadIdx.x;
  const int64_t m0 = (int64_t)blockIdx.x * 16, n0 = (int64_t)blockIdx.y * 16;
  const int64_t kb = (int64_t)blockIdx.z * kchunk;
  const int64_t ke = (kb + kchunk < K) ? kb + kchunk : K;
  const int64_t ar = m0 + (l & 15);
  const int64_t bc = n0 + (l & 15);
  const int kq = 4 * (l >> 4);
  floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t k0 = kb; k0 < ke; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t k = k0 + kq + e;
      const float av = (ar < M && k < ke) ? load_elem<ABF>(A, ar * sa0 + k * sa1) : 0.0f;
      const float bv = (bc < N && k < ke) ? load_elem<BBF>(B, k * sb0 + bc * sb1) : 0.0f;
      a[e] = f32_to_bf16(av);
      b[e] = f32_to_bf16(bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  float* Pz = P + (int64_t)blockIdx.z * M * N;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int64_t r = m0 + kq + e, c = bc;
    if (r < M && c < N) Pz[r * N + c] = acc[e];
  }
}

__global__ __launch_bounds__(256) void splitk_reduce_kernel(const float* __restrict__ P, int nz,
                                                            int64_t M, int64_t N,
                                                            float* __restrict__ C, int64_t ldc) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= M * N) return;
  float s = 0.0f;
  for (int z = 0; z < nz; ++z) s += P[(int64_t)z * M * N + i];
  C[(i / N) * ldc + i % N] = s;
}

// Column sums out[j] = sum_r X[r*ld + j] in two fixed-order stages.
// colsum_part_kernel: block (x, y) covers 32 columns and rows
// [y kColChunk, (y + 1) kColChunk): 8 row lanes per column (coalesced rows)
// each sum every 8th row in order, then the 8 lanes are added in order into
// part[y][j] (or straight into out[j] when there is one chunk);
// colsum_final_kernel adds the chunks in order.
constexpr int64_t kColChunk = 256;
__global__ __launch_bounds__(256) void colsum_part_kernel(const float* __restrict__ X,
                                                          int64_t rows, int64_t cols, int64_t ld,
                                                          double* __restrict__ part,
                                                          float* __restrict__ out) {
  __shared__ double sh[8][32];
  const int lane = threadIdx.x & 31, rl = threadIdx.x >> 5;
  const int64_t j = (int64_t)blockIdx.x * 32 + lane;
  const int64_t rb = (int64_t)blockIdx.y * kColChunk;
  const int64_t re = (rb + kColChunk < rows) ? rb + kColChunk : rows;
  double s = 0.0;
  if (j < cols)
    for (int64_t r = rb + rl; r < re; r += 8) s += X[r * ld + j];
  sh[rl][lane] = s;
  __syncthreads();
  if (rl == 0 && j < cols) {
    double t = 0.0;
    for (int k = 0; k < 8; ++k) t += sh[k][lane];
    if (out)
      out[j] = (float)t;
    else
      part[(int64_t)blockIdx.y * cols + j] = t;
  }
}

__global__ __launch_bounds__(128) void colsum_final_kernel(const double* __restrict__ part,
                                                           int64_t nchunk, int64_t cols,
                                                           float* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * 128 + threadIdx.x;
  if (j >= cols) return;
  double t = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) t += part[c * cols + j];
  out[j] = (float)t;
}

__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ w,
                                                   unsigned short* __restrict__ wb,
                                                   const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v,
                                                   int64_t n, float lr, float b1, float b2,
                                                   float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (unsigned short)f32_to_bf16(wi);
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float* __restrict__ x, int64_t n,
                                                      unsigned short* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = (unsigned short)f32_to_bf16(x[i]);
}

// ------------------------------------------------------------------ MLP layout
constexpr int64_t kIn = 15, kHid = 128, kOut = 3;
// weights, row-major [in][out]: W1 [15][128], W2 [128][128], W3 [128][128], W4 [128][3]
constexpr int64_t kW1 = 0, kW2 = kW1 + kIn * kHid, kW3 = kW2 + kHid * kHid,
                  kW4 = kW3 + kHid * kHid, kNW = kW4 + kHid * kOut;
// biases: b1 [128], b2 [128], b3 [128], b4 [3]
constexpr int64_t kB1 = 0, kB2 = kHid, kB3 = 2 * kHid, kB4 = 3 * kHid, kNB = 3 * kHid + kOut;

template <bool ABF, bool BBF>
static int gemm(const void* A, int64_t sa0, int64_t sa1, const void* B, int64_t sb0, int64_t sb1,
                float* C, int64_t ldc, int64_t M, int64_t N, int64_t K, const float* bias,
                const float* mask, int64_t ldm, int epi, hipStream_t st) {
  if (M == 0 || N == 0) return DMIC_OK;
  DMIC_REQUIRE((M + 15) / 16 <= 2147483647 && (N + 15) / 16 <= 65535, DMIC_ERR_ARG,
               "gemm too large: %lld x %lld output tiles (at most 2^31-1 row tiles and 65535 "
               "column tiles)", (long long)((M + 15) / 16), (long long)((N + 15) / 16));
  dim3 grid((unsigned)((M + 15) / 16), (unsigned)((N + 15) / 16));
  hipLaunchKernelGGL((gemm_bf16_kernel<ABF, BBF>), grid, dim3(64), 0, st, A, sa0, sa1, B, sb0, sb1,
                     C, ldc, M, N, K, bias, mask, ldm, epi);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

// Weight gradient C[M,N] = A B over K = rows with split K when the output
// has few tiles and K is long; part: at least ceil(K / kSplitChunk) M N
// floats of free scratch.
constexpr int64_t kSplitChunk = 256;
template <bool ABF, bool BBF>
static int gemm_wgrad(const void* A, int64_t sa0, int64_t sa1, const void* B, int64_t sb0,
                      int64_t sb1, float* C, int64_t ldc, int64_t M, int64_t N, int64_t K,
                      float* part, hipStream_t st) {
  const int64_t tiles = ((M + 15) / 16) * ((N + 15) / 16);
  if (K < 4 * kSplitChunk || tiles >= 1024 || part == nullptr)
    return gemm<ABF, BBF>(A, sa0, sa1, B, sb0, sb1, C, ldc, M, N, K, nullptr, nullptr, 0,
                          EPI_NONE, st);
  const int64_t nz = (K + kSplitChunk - 1) / kSplitChunk;
  DMIC_REQUIRE(nz <= 65535, DMIC_ERR_ARG, "too many split-K chunks");
  dim3 grid((unsigned)((M + 15) / 16), (unsigned)((N + 15) / 16), (unsigned)nz);
  hipLaunchKernelGGL((gemm_bf16_splitk_kernel<ABF, BBF>), grid, dim3(64), 0, st, A, sa0, sa1, B,
                     sb0, sb1, part, M, N, K, kSplitChunk);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3(ceil_div(M * N, 256)), dim3(256), 0, st, part,
                     (int)nz, M, N, C, ldc);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

// part: free scratch of ceil(rows / kColChunk) * cols doubles, used only
// when rows > kColChunk (then at most rows cols floats).
static int colsum(const float* X, int64_t rows, int64_t cols, int64_t ld, float* out,
                  void* part, hipStream_t st) {
  if (cols == 0) return DMIC_OK;
  const int64_t nchunk = (rows + kColChunk - 1) / kColChunk;
  DMIC_REQUIRE(nchunk <= 65535, DMIC_ERR_ARG, "too many rows for the column sum");
  double* pd = static_cast<double*>(part);
  if (nchunk == 1) {
    hipLaunchKernelGGL(colsum_part_kernel, dim3(ceil_div(cols, 32), 1), dim3(256), 0, st, X, rows,
                       cols, ld, pd, out);
    DMIC_LAUNCH_CHECK();
    return DMIC_OK;
  }
  if (nchunk > 1) {
    hipLaunchKernelGGL(colsum_part_kernel, dim3(ceil_div(cols, 32), (unsigned)nchunk), dim3(256),
                       0, st, X, rows, cols, ld, pd, (float*)nullptr);
    DMIC_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(colsum_final_kernel, dim3(ceil_div(cols, 128)), dim3(128), 0, st, pd, nchunk,
                     cols, out);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

#define DMIC_TRY(x)          \
  do {                       \
    int rc__ = (x);          \
    if (rc__ != DMIC_OK) return rc__; \
  } while (0)

}  // namespace dmic

using namespace dmic;

extern "C" {

int dmic_gemm_bf16(const void* A, int a_is_bf16, int64_t sa0, int64_t sa1, const void* B,
                   int b_is_bf16, int64_t sb0, int64_t sb1, float* C, int64_t ldc, int64_t M,
                   int64_t N, int64_t K, const float* bias, const float* mask, int64_t ldm,
                   int epilogue, hipStream_t stream) {
  DMIC_REQUIRE(M >= 0 && N >= 0 && K >= 0, DMIC_ERR_ARG, "negative size");
  DMIC_REQUIRE(epilogue >= 0 && epilogue <= 3, DMIC_ERR_ARG, "bad epilogue");
  DMIC_REQUIRE(M == 0 || N == 0 || (A && B && C), DMIC_ERR_ARG, "null pointer");
  DMIC_REQUIRE(!(epilogue == 1 || epilogue == 2) || bias, DMIC_ERR_ARG, "null bias");
  DMIC_REQUIRE(epilogue != 3 || mask, DMIC_ERR_ARG, "null mask");
  if (a_is_bf16 && b_is_bf16)
    return gemm<true, true>(A, sa0, sa1, B, sb0, sb1, C, ldc, M, N, K, bias, mask, ldm, epilogue, stream);
  if (a_is_bf16)
    return gemm<true, false>(A, sa0, sa1, B, sb0, sb1, C, ldc, M, N, K, bias, mask, ldm, epilogue, stream);
  if (b_is_bf16)
    return gemm<false, true>(A, sa0, sa1, B, sb0, sb1, C, ldc, M, N, K, bias, mask, ldm, epilogue, stream);
  return gemm<false, false>(A, sa0, sa1, B, sb0, sb1, C, ldc, M, N, K, bias, mask, ldm, epilogue, stream);
}

int64_t dmic_mlp_weight_count(void) { return kNW; }
int64_t dmic_mlp_bias_count(void) { return kNB; }

// y[rows][3] = MLP(x[rows][15]); acts (optional) receives H1, H2, H3 as
// [3][rows][128] fp32 for dmic_mlp_backward.  w_bf16: packed weights (bf16).
int dmic_mlp_forward(const uint16_t* w_bf16, const float* bias, const float* x, int64_t rows,
                     float* y, float* acts, float* scratch, hipStream_t stream) {
  DMIC_REQUIRE(rows >= 0, DMIC_ERR_ARG, "negative rows");
  if (rows == 0) return DMIC_OK;
  DMIC_REQUIRE(w_bf16 && bias && x && y, DMIC_ERR_ARG, "null pointer");
  float* H1 = acts ? acts : scratch;
  float* H2 = acts ? acts + rows * kHid : scratch + rows * kHid;
  float* H3 = acts ? acts + 2 * rows * kHid : scratch;
  DMIC_REQUIRE(H1, DMIC_ERR_ARG, "need acts or scratch [2*rows*128]");
  DMIC_TRY((gemm<false, true>(x, kIn, 1, w_bf16 + kW1, kHid, 1, H1, kHid, rows, kHid, kIn,
                              bias + kB1, nullptr, 0, EPI_BIAS_RELU, stream)));
  DMIC_TRY((gemm<false, true>(H1, kHid, 1, w_bf16 + kW2, kHid, 1, H2, kHid, rows, kHid, kHid,
                              bias + kB2, nullptr, 0, EPI_BIAS_RELU, stream)));
  DMIC_TRY((gemm<false, true>(H2, kHid, 1, w_bf16 + kW3, kHid, 1, H3, kHid, rows, kHid, kHid,
                              bias + kB3, nullptr, 0, EPI_BIAS_RELU, stream)));
  DMIC_TRY((gemm<false, true>(H3, kHid, 1, w_bf16 + kW4, kOut, 1, y, kOut, rows, kOut, kHid,
                              bias + kB4, nullptr, 0, EPI_BIAS, stream)));
  return DMIC_OK;
}

// Gradients of sum(dy . y) w.r.t. weights (grad_w [kNW] fp32) and biases
// (grad_b [kNB]); scratch: 2*rows*128 floats.
int dmic_mlp_backward(const uint16_t* w_bf16, const float* x, const float* acts,
                      const float* dy, int64_t rows, float* grad_w, float* grad_b, float* scratch,
                      hipStream_t stream) {
  DMIC_REQUIRE(rows >= 0, DMIC_ERR_ARG, "negative rows");
  DMIC_REQUIRE(w_bf16 && x && acts && dy && grad_w && grad_b && scratch, DMIC_ERR_ARG,
               "null pointer");
  const float* H1 = acts;
  const float* H2 = acts + rows * kHid;
  const float* H3 = acts + 2 * rows * kHid;
  float* D1 = scratch;
  float* D2 = scratch + rows * kHid;
  // the split-K partials of each weight gradient (ceil(rows/256) x 128 x 128
  // floats, <= rows x 128 where split K is used) and the column-sum partials
  // go to whichever half of scratch is free at that point
  // layer 4: dW4 = H3^T dy, db4 = colsum(dy), dH3 = (dy W4^T) * [H3 > 0]
  DMIC_TRY((gemm_wgrad<false, false>(H3, 1, kHid, dy, kOut, 1, grad_w + kW4, kOut, kHid, kOut,
                                     rows, D2, stream)));
  DMIC_TRY(colsum(dy, rows, kOut, kOut, grad_b + kB4, D1, stream));
  DMIC_TRY((gemm<false, true>(dy, kOut, 1, w_bf16 + kW4, 1, kOut, D1, kHid, rows, kHid, kOut,
                              nullptr, H3, kHid, EPI_RELU_MASK, stream)));
  // layer 3
  DMIC_TRY((gemm_wgrad<false, false>(H2, 1, kHid, D1, kHid, 1, grad_w + kW3, kHid, kHid, kHid,
                                     rows, D2, stream)));
  DMIC_TRY(colsum(D1, rows, kHid, kHid, grad_b + kB3, D2, stream));
  DMIC_TRY((gemm<false, true>(D1, kHid, 1, w_bf16 + kW3, 1, kHid, D2, kHid, rows, kHid, kHid,
                              nullptr, H2, kHid, EPI_RELU_MASK, stream)));
  // layer 2
  DMIC_TRY((gemm_wgrad<false, false>(H1, 1, kHid, D2, kHid, 1, grad_w + kW2, kHid, kHid, kHid,
                                     rows, D1, stream)));
  DMIC_TRY(colsum(D2, rows, kHid, kHid, grad_b + kB2, D1, stream));
  DMIC_TRY((gemm<false, true>(D2, kHid, 1, w_bf16 + kW2, 1, kHid, D1, kHid, rows, kHid, kHid,
                              nullptr, H1, kHid, EPI_RELU_MASK, stream)));
  // layer 1
  DMIC_TRY((gemm_wgrad<false, false>(x, 1, kIn, D1, kHid, 1, grad_w + kW1, kHid, kIn, kHid, rows,
                                     D2, stream)));
  DMIC_TRY(colsum(D1, rows, kHid, kHid, grad_b + kB1, D2, stream));
  return DMIC_OK;
}

// Adam (bias-corrected) on fp32 master parameters; wb (optional) receives the bf16 copy.
int dmic_adam_step(float* w, uint16_t* wb, const float* g, float* m, float* v, int64_t n,
                   float lr, float beta1, float beta2, float eps, int64_t step,
                   hipStream_t stream) {
  DMIC_REQUIRE(n >= 0 && step >= 1, DMIC_ERR_ARG, "bad size/step");
  if (n == 0) return DMIC_OK;
  DMIC_REQUIRE(w && g && m && v, DMIC_ERR_ARG, "null pointer");
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, w, wb, g, m, v, n,
                     lr, beta1, beta2, eps, c1, c2);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_to_bf16(const float* x, int64_t n, uint16_t* y, hipStream_t stream) {
  DMIC_REQUIRE(n >= 0, DMIC_ERR_ARG, "negative size");
  if (n == 0) return DMIC_OK;
  DMIC_REQUIRE(x && y, DMIC_ERR_ARG, "null pointer");
  hipLaunchKernelGGL(to_bf16_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, stream, x, n, y);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

}  // extern "C"
