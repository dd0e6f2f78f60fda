// Render chain around the meshFT layer for gfx950: the modulus of the
// inverse FFT, its min/max and normalize_tensor, forward and backward.
//
// render_image_from_ftmesh (deltamic/render.py:177-180) ends in torch.abs of
// the complex box and normalize_tensor (render.py:5-6) computes
// ((t - min) / (max - min) + c) / (1 + c) with the gradient flowing through
// min and max.  In torch that is one pass for abs, two reductions and four
// elementwise passes for normalize, and as many again backwards.  Here:
//   forward : abs (reads 8 B, writes 4 B per voxel); min/max (reads 4 B);
//             normalize (reads 4 B, writes 4 B);
//   backward: one reduction pass over (grad, t) for the min/max terms and
//             one elementwise pass for dL/dt; abs backward reads the complex
//             box once more (dL/dz = g z / |z|, 0 at z = 0, torch's sgn).
// All HBM-bound grid-stride loops.  Reductions: per-block partials, then
// one block reduces them in a fixed tree (bitwise reproducible).  NaN propagates through min and
// max as in torch.  Ties at the minimum or maximum share its gradient
// evenly (torch's evenly_distribute_backward for full reductions).
#include "common.h"

namespace dmic {

constexpr int kRThreads = 256;
constexpr int kRMaxBlocks = 2048;  // partial slots of the reductions

__device__ __forceinline__ float nan_min(float a, float b) {
  return (a != a || b != b) ? __builtin_nanf("") : (a < b ? a : b);
}
__device__ __forceinline__ float nan_max(float a, float b) {
  return (a != a || b != b) ? __builtin_nanf("") : (a > b ? a : b);
}

__device__ __forceinline__ void block_minmax(float& mn, float& mx, float* smn, float* smx) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mn = nan_min(mn, __shfl_xor(mn, o));
    mx = nan_max(mx, __shfl_xor(mx, o));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smn[w] = mn;
    smx[w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kRThreads / 64; ++i) {
      mn = nan_min(mn, smn[i]);
      mx = nan_max(mx, smx[i]);
    }
  }
}

// img = |z| (hypot, as torch's complex abs) and per-block min / max.
__global__ __launch_bounds__(kRThreads) void abs_minmax_kernel(const float2* __restrict__ z,
                                                                int64_t n, float* __restrict__ img,
                                                                float* __restrict__ part) {
  __shared__ float smn[kRThreads / 64], smx[kRThreads / 64];
  float mn = __builtin_inff(), mx = -__builtin_inff();
  const int64_t stride = (int64_t)gridDim.x * kRThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRThreads + threadIdx.x; i < n; i += stride) {
    const float2 v = z[i];
    const float a = hypotf(v.x, v.y);
    img[i] = a;
    mn = nan_min(mn, a);
    mx = nan_max(mx, a);
  }
  block_minmax(mn, mx, smn, smx);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = mn;
    part[2 * blockIdx.x + 1] = mx;
  }
}

__global__ __launch_bounds__(kRThreads) void minmax_kernel(const float* __restrict__ t, int64_t n,
                                                            float* __restrict__ part) {
  __shared__ float smn[kRThreads / 64], smx[kRThreads / 64];
  float mn = __builtin_inff(), mx = -__builtin_inff();
  const int64_t stride = (int64_t)gridDim.x * kRThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRThreads + threadIdx.x; i < n; i += stride) {
    const float a = t[i];
    mn = nan_min(mn, a);
    mx = nan_max(mx, a);
  }
  block_minmax(mn, mx, smn, smx);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = mn;
    part[2 * blockIdx.x + 1] = mx;
  }
}

constexpr int kFinal = 1024;  // threads of the one-block final reductions

// minmax[0..1] = (min, max) over the block partials: one block, each thread
// a fixed strided subset, then a fixed tree.
__global__ __launch_bounds__(kFinal) void minmax_final_kernel(const float* __restrict__ part,
                                                              int nblocks,
                                                              float* __restrict__ minmax) {
  __shared__ float smn[kFinal], smx[kFinal];
  float mn = __builtin_inff(), mx = -__builtin_inff();
  for (int b = threadIdx.x; b < nblocks; b += kFinal) {
    mn = nan_min(mn, part[2 * b]);
    mx = nan_max(mx, part[2 * b + 1]);
  }
  smn[threadIdx.x] = mn;
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int w = kFinal / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smn[threadIdx.x] = nan_min(smn[threadIdx.x], smn[threadIdx.x + w]);
      smx[threadIdx.x] = nan_max(smx[threadIdx.x], smx[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    minmax[0] = smn[0];
    minmax[1] = smx[0];
  }
}

// out = ((t - min) / (max - min) + c) / (1 + c), FP32 in torch's operation order.
__global__ __launch_bounds__(kRThreads) void normalize_kernel(const float* __restrict__ t,
                                                               int64_t n,
                                                               const float* __restrict__ minmax,
                                                               float c, float* __restrict__ out) {
  const float mn = minmax[0], rng = minmax[1] - minmax[0], den = 1.0f + c;
  const int64_t stride = (int64_t)gridDim.x * kRThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRThreads + threadIdx.x; i < n; i += stride)
    out[i] = ((t[i] - mn) / rng + c) / den;
}

// Per block: S1 = sum g, S2 = sum g u (u = (t - min)/(max - min)) in FP64,
// and the counts of t == min and t == max.
__global__ __launch_bounds__(kRThreads) void normalize_bwd_reduce_kernel(
    const float* __restrict__ g, const float* __restrict__ t, int64_t n,
    const float* __restrict__ minmax, double* __restrict__ part) {
  __shared__ double sh[4][kRThreads];
  const float mn = minmax[0], mx = minmax[1];
  const double inv = 1.0 / ((double)mx - (double)mn);
  double s1 = 0.0, s2 = 0.0, cmin = 0.0, cmax = 0.0;
  const int64_t stride = (int64_t)gridDim.x * kRThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRThreads + threadIdx.x; i < n; i += stride) {
    const double gi = g[i];
    const float ti = t[i];
    s1 += gi;
    s2 += gi * (((double)ti - (double)mn) * inv);
    cmin += (ti == mn) ? 1.0 : 0.0;
    cmax += (ti == mx) ? 1.0 : 0.0;
  }
  sh[0][threadIdx.x] = s1;
  sh[1][threadIdx.x] = s2;
  sh[2][threadIdx.x] = cmin;
  sh[3][threadIdx.x] = cmax;
  __syncthreads();
  for (int w = kRThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// coef = {r, dmin / count_min, dmax / count_max} with r = 1/((max - min)(1 + c)),
// dmin = r sum g (u - 1), dmax = -r sum g u.
__global__ __launch_bounds__(kFinal) void normalize_bwd_final_kernel(
    const double* __restrict__ part, int nblocks, const float* __restrict__ minmax, float c,
    double* __restrict__ coef) {
  __shared__ double sh[4][kFinal];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += kFinal)
    for (int k = 0; k < 4; ++k) a[k] += part[4 * b + k];
  for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int w = kFinal / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double s1 = sh[0][0], s2 = sh[1][0], cmin = sh[2][0], cmax = sh[3][0];
  const double r = 1.0 / (((double)minmax[1] - (double)minmax[0]) * (1.0 + (double)c));
  coef[0] = r;
  coef[1] = (cmin > 0.0) ? r * (s2 - s1) / cmin : 0.0;
  coef[2] = (cmax > 0.0) ? -r * s2 / cmax : 0.0;
}

__global__ __launch_bounds__(kRThreads) void normalize_bwd_kernel(
    const float* __restrict__ g, const float* __restrict__ t, int64_t n,
    const float* __restrict__ minmax, const double* __restrict__ coef,
    float* __restrict__ grad_t) {
  const float mn = minmax[0], mx = minmax[1];
  const double r = coef[0], cmn = coef[1], cmx = coef[2];
  const int64_t stride = (int64_t)gridDim.x * kRThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRThreads + threadIdx.x; i < n; i += stride) {
    const float ti = t[i];
    double v = (double)g[i] * r;
    if (ti == mn) v += cmn;
    if (ti == mx) v += cmx;
    grad_t[i] = (float)v;
  }
}

// dL/dz = g z / |z| (0 where z = 0): torch's abs backward, grad * sgn(z).
__global__ __launch_bounds__(kRThreads) void abs_bwd_kernel(const float* __restrict__ g,
                                                             const float2* __restrict__ z,
                                                             int64_t n, float2* __restrict__ gz) {
  const int64_t stride = (int64_t)gridDim.x * kRThreads;
  for (int64_t i = (int64_t)blockIdx.x * kRThreads + threadIdx.x; i < n; i += stride) {
    const float2 v = z[i];
    const float a = hypotf(v.x, v.y);
    const float s = (a > 0.0f) ? g[i] / a : 0.0f;
    gz[i] = make_float2(s * v.x, s * v.y);
  }
}

static unsigned render_blocks(int64_t n) {
  int64_t b = (n + kRThreads * 4 - 1) / (kRThreads * 4);  // ~4 elements per thread
  if (b < 1) b = 1;
  if (b > kRMaxBlocks) b = kRMaxBlocks;
  return (unsigned)b;
}

}  // namespace dmic

using namespace dmic;

extern "C" {

size_t dmic_render_workspace(int64_t n) {
  if (n < 0) return 0;
  Carver c(nullptr);
  c.take<double>(4 * kRMaxBlocks);
  c.take<double>(4);
  return c.bytes();
}

int dmic_abs_minmax(const float* z, int64_t n, float* img, float* minmax, void* workspace,
                    size_t workspace_bytes, hipStream_t stream) {
  DMIC_REQUIRE(n > 0, DMIC_ERR_ARG, "empty box");
  DMIC_REQUIRE(z && img && minmax && workspace, DMIC_ERR_ARG, "null pointer");
  DMIC_REQUIRE(workspace_bytes >= dmic_render_workspace(n), DMIC_ERR_WORKSPACE,
               "render workspace too small");
  float* part = static_cast<float*>(workspace);
  const unsigned nb = render_blocks(n);
  hipLaunchKernelGGL(abs_minmax_kernel, dim3(nb), dim3(kRThreads), 0, stream,
                     reinterpret_cast<const float2*>(z), n, img, part);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(kFinal), 0, stream, part, (int)nb, minmax);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_minmax(const float* t, int64_t n, float* minmax, void* workspace, size_t workspace_bytes,
                hipStream_t stream) {
  DMIC_REQUIRE(n > 0, DMIC_ERR_ARG, "empty tensor");
  DMIC_REQUIRE(t && minmax && workspace, DMIC_ERR_ARG, "null pointer");
  DMIC_REQUIRE(workspace_bytes >= dmic_render_workspace(n), DMIC_ERR_WORKSPACE,
               "render workspace too small");
  float* part = static_cast<float*>(workspace);
  const unsigned nb = render_blocks(n);
  hipLaunchKernelGGL(minmax_kernel, dim3(nb), dim3(kRThreads), 0, stream, t, n, part);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(minmax_final_kernel, dim3(1), dim3(kFinal), 0, stream, part, (int)nb, minmax);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_normalize(const float* t, int64_t n, const float* minmax, float c, float* out,
                   hipStream_t stream) {
  DMIC_REQUIRE(n >= 0, DMIC_ERR_ARG, "negative size");
  if (n == 0) return DMIC_OK;
  DMIC_REQUIRE(t && minmax && out, DMIC_ERR_ARG, "null pointer");
  hipLaunchKernelGGL(normalize_kernel, dim3(render_blocks(n)), dim3(kRThreads), 0, stream, t, n,
                     minmax, c, out);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_normalize_backward(const float* g, const float* t, int64_t n, const float* minmax,
                            float c, float* grad_t, void* workspace, size_t workspace_bytes,
                            hipStream_t stream) {
  DMIC_REQUIRE(n > 0, DMIC_ERR_ARG, "empty tensor");
  DMIC_REQUIRE(g && t && minmax && grad_t && workspace, DMIC_ERR_ARG, "null pointer");
  Carver cv(workspace);
  double* part = cv.take<double>(4 * kRMaxBlocks);
  double* coef = cv.take<double>(4);
  DMIC_REQUIRE(workspace_bytes >= cv.bytes(), DMIC_ERR_WORKSPACE, "render workspace too small");
  const unsigned nb = render_blocks(n);
  hipLaunchKernelGGL(normalize_bwd_reduce_kernel, dim3(nb), dim3(kRThreads), 0, stream, g, t, n,
                     minmax, part);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(normalize_bwd_final_kernel, dim3(1), dim3(kFinal), 0, stream, part, (int)nb,
                     minmax, c, coef);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(normalize_bwd_kernel, dim3(nb), dim3(kRThreads), 0, stream, g, t, n, minmax,
                     coef, grad_t);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_abs_backward(const float* g, const float* z, int64_t n, float* grad_z,
                      hipStream_t stream) {
  DMIC_REQUIRE(n >= 0, DMIC_ERR_ARG, "negative size");
  if (n == 0) return DMIC_OK;
  DMIC_REQUIRE(g && z && grad_z, DMIC_ERR_ARG, "null pointer");
  hipLaunchKernelGGL(abs_bwd_kernel, dim3(render_blocks(n)), dim3(kRThreads), 0, stream, g,
                     reinterpret_cast<const float2*>(z), n, reinterpret_cast<float2*>(grad_z));
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

}  // extern "C"
