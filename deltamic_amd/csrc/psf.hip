// Learnable Gaussian PSFs on a separable grid, forward and parameter gradient.
//
// The reference builds these PSFs with torch meshgrids (deltamic/psf.py:551-665,
// Gaussian_psf, Gaussian_psf_logsigma, Anisotropic_gaussian_psf) and lets
// autograd differentiate them; every iteration of its fit loop evaluates the
// PSF, takes its FFT as the OTF and steps the PSF parameters
// (Examples/Gradient_Descent.ipynb cell 2).  Here the forward is one pass that
// writes the box (4 bytes per voxel, the three 1-D axes stay in cache) and the
// backward is one pass that reads the upstream gradient and the saved PSF
// (8 bytes per voxel), forms per-row sums, applies the row's x0/x1 factors
// and reduces the few parameter sums per workgroup in FP64; a single
// workgroup then adds the workgroup partials in a fixed order, so gradients
// are bitwise reproducible.  Both are HBM-bound.
//
// Full 3x3 form:    psf = exp(-q/2),  q = x^T S x,            x = (x0[i], x1[j], x2[k])
//   dpsf/dS_ab = psf * (-x_a x_b / 2)
// Anisotropic form: psf = exp(-Q/(2D)) / D,  Q = (x0, x1) S2 (x0, x1)^T,
//                   D = eps + (|x2| sz)^2
//   dpsf/dS2_ab = psf * (-x_a x_b / (2D)),  dpsf/dD = psf * (Q/(2D^2) - 1/D),
//   dD/dsz = 2 x2^2 sz,  dD/deps = 1.
// Parameters are read from device memory (they are learnable tensors), so
// no host synchronisation is needed.
#include "common.h"

namespace dmic {

namespace {

constexpr int kPsfThreads = 256;
constexpr int kWaves = kPsfThreads / 64;
constexpr int kRowsPerWave = 4;
constexpr int kRowsPerBlock = kWaves * kRowsPerWave;
constexpr int kGaussSums = 6;  // x0x0 x0x1 x0x2 x1x1 x1x2 x2x2
constexpr int kAnisoSums = 5;  // T00 T01 T11 U V

// Launch shape shared by all kernels: the box is a list of n0*n1 rows of n2
// voxels (row r = i*n1 + j, x0 = x0[i], x1 = x1[j]); each wave takes
// kRowsPerWave rows and its 64 lanes stride along x2, so stores are
// coalesced, every lane is busy and the per-row factors of the gradient sums
// are applied once per row.
struct Grid3 {
  const float* x0;
  const float* x1;
  const float* x2;
  int64_t n0, n1, n2;
};

__device__ __forceinline__ int wave_id() { return (int)threadIdx.x >> 6; }
__device__ __forceinline__ int lane_id() { return (int)threadIdx.x & 63; }

__global__ __launch_bounds__(kPsfThreads) void psf_gauss_fwd_kernel(const float* __restrict__ S,
                                                                    Grid3 g,
                                                                    float* __restrict__ out) {
  const int64_t rows = g.n0 * g.n1;
  for (int rr = 0; rr < kRowsPerWave; ++rr) {
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + rr * kWaves + wave_id();
    if (r >= rows) break;
    const float a = g.x0[r / g.n1], b = g.x1[r % g.n1];
    float* o = out + r * g.n2;
    for (int64_t k = lane_id(); k < g.n2; k += 64) {
      const float c = g.x2[k];
      // psf.py:573-576 operation order
      const float q = a * (S[0] * a + S[1] * b + S[2] * c) +
                      b * (S[3] * a + S[4] * b + S[5] * c) + c * (S[6] * a + S[7] * b + S[8] * c);
      o[k] = expf(-q / 2.0f);
    }
  }
}

// prm = {S2_00, S2_01, S2_10, S2_11, sz, eps}
__global__ __launch_bounds__(kPsfThreads) void psf_aniso_fwd_kernel(const float* __restrict__ prm,
                                                                    Grid3 g,
                                                                    float* __restrict__ out) {
  const int64_t rows = g.n0 * g.n1;
  for (int rr = 0; rr < kRowsPerWave; ++rr) {
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + rr * kWaves + wave_id();
    if (r >= rows) break;
    const float a = g.x0[r / g.n1], b = g.x1[r % g.n1];
    // psf.py:656-658 operation order
    const float Q = a * (prm[0] * a + prm[1] * b) + b * (prm[2] * a + prm[3] * b);
    float* o = out + r * g.n2;
    for (int64_t k = lane_id(); k < g.n2; k += 64) {
      const float t = __builtin_fabsf(g.x2[k]) * prm[4];
      const float D = prm[5] + t * t;
      o[k] = expf(-Q / (2.0f * D)) / D;
    }
  }
}

// Sum over the 64 lanes of a wave (every lane gets the total; fixed order).
__device__ __forceinline__ double wave_sum(double v) {
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// K per-lane FP64 sums -> workgroup total (fixed order) -> part[block][K].
template <int K>
__device__ __forceinline__ void block_partial(double (&P)[K], double* __restrict__ part) {
  __shared__ double red[kWaves][K];
#pragma unroll
  for (int k = 0; k < K; ++k) P[k] = wave_sum(P[k]);
  if (lane_id() == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) red[wave_id()][k] = P[k];
  __syncthreads();
  if ((int)threadIdx.x < K) {
    double t = 0.0;
    for (int w = 0; w < kWaves; ++w) t += red[w][threadIdx.x];
    part[(int64_t)blockIdx.x * K + threadIdx.x] = t;
  }
}

// Per row and lane: R0 = sum w, R1 = sum w x2, R2 = sum w x2^2 (w = G psf,
// FP32 over the lane's few voxels); the six pair sums are x0^2 R0, x0 x1 R0,
// x0 R1, x1^2 R0, x1 R1, R2 (FP64).
__global__ __launch_bounds__(kPsfThreads) void psf_gauss_grad_kernel(
    const float* __restrict__ G, const float* __restrict__ psf, Grid3 g,
    double* __restrict__ part) {
  double P[kGaussSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t rows = g.n0 * g.n1;
  for (int rr = 0; rr < kRowsPerWave; ++rr) {
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + rr * kWaves + wave_id();
    if (r >= rows) break;
    const double a = g.x0[r / g.n1], b = g.x1[r % g.n1];
    float R0 = 0.f, R1 = 0.f, R2 = 0.f;
    for (int64_t k = lane_id(); k < g.n2; k += 64) {
      const float c = g.x2[k];
      const float w = G[r * g.n2 + k] * psf[r * g.n2 + k];
      R0 += w;
      R1 = __builtin_fmaf(w, c, R1);
      R2 = __builtin_fmaf(w * c, c, R2);
    }
    P[0] += a * a * R0;
    P[1] += a * b * R0;
    P[2] += a * R1;
    P[3] += b * b * R0;
    P[4] += b * R1;
    P[5] += R2;
  }
  block_partial<kGaussSums>(P, part);
}

// Per row and lane: A = sum w/D, B = sum w/D^2, C = sum w x2^2/D^2,
// E = sum w x2^2/D; with the row's Q: T_ab = x_a x_b A, U = Q/2 B - A,
// V = Q/2 C - E.
__global__ __launch_bounds__(kPsfThreads) void psf_aniso_grad_kernel(
    const float* __restrict__ prm, const float* __restrict__ G, const float* __restrict__ psf,
    Grid3 g, double* __restrict__ part) {
  double P[kAnisoSums] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t rows = g.n0 * g.n1;
  for (int rr = 0; rr < kRowsPerWave; ++rr) {
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + rr * kWaves + wave_id();
    if (r >= rows) break;
    const float af = g.x0[r / g.n1], bf = g.x1[r % g.n1];
    const double a = af, b = bf;
    const double Q = (double)(af * (prm[0] * af + prm[1] * bf) + bf * (prm[2] * af + prm[3] * bf));
    float A = 0.f, B = 0.f, C = 0.f, E = 0.f;
    for (int64_t k = lane_id(); k < g.n2; k += 64) {
      const float c = g.x2[k];
      const float t = __builtin_fabsf(c) * prm[4];
      const float rD = 1.0f / (prm[5] + t * t);
      const float w = G[r * g.n2 + k] * psf[r * g.n2 + k];
      const float wd = w * rD, c2 = c * c;
      A += wd;
      B = __builtin_fmaf(wd, rD, B);
      C = __builtin_fmaf(wd * rD, c2, C);
      E = __builtin_fmaf(wd, c2, E);
    }
    P[0] += a * a * A;
    P[1] += a * b * A;
    P[2] += b * b * A;
    P[3] += 0.5 * Q * B - A;
    P[4] += 0.5 * Q * C - E;
  }
  block_partial<kAnisoSums>(P, part);
}

// One workgroup: fixed-order sum of the workgroup partials (each thread a
// strided subset in order, then a fixed tree over the threads).
template <int K>
__device__ __forceinline__ void finish_sums(const double* __restrict__ part, int64_t nblocks,
                                            double (&red)[K][kPsfThreads]) {
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int64_t b = threadIdx.x; b < nblocks; b += kPsfThreads)
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += part[b * K + k];
#pragma unroll
  for (int k = 0; k < K; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = kPsfThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int k = 0; k < K; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kPsfThreads) void psf_gauss_finish_kernel(
    const double* __restrict__ part, int64_t nblocks, float* __restrict__ grad_S) {
  __shared__ double red[kGaussSums][kPsfThreads];
  finish_sums<kGaussSums>(part, nblocks, red);
  if (threadIdx.x == 0) {
    // dL/dS_ab = -1/2 sum G psf x_a x_b (S_ab and S_ba get the same sum)
    const int pair[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int e = 0; e < 9; ++e) grad_S[e] = (float)(-0.5 * red[pair[e]][0]);
  }
}

__global__ __launch_bounds__(kPsfThreads) void psf_aniso_finish_kernel(
    const float* __restrict__ prm, const double* __restrict__ part, int64_t nblocks,
    float* __restrict__ grad_prm) {
  __shared__ double red[kAnisoSums][kPsfThreads];
  finish_sums<kAnisoSums>(part, nblocks, red);
  if (threadIdx.x == 0) {
    grad_prm[0] = (float)(-0.5 * red[0][0]);
    grad_prm[1] = (float)(-0.5 * red[1][0]);
    grad_prm[2] = (float)(-0.5 * red[1][0]);
    grad_prm[3] = (float)(-0.5 * red[2][0]);
    grad_prm[4] = (float)(2.0 * (double)prm[4] * red[4][0]);  // dD/dsz = 2 x2^2 sz
    grad_prm[5] = (float)red[3][0];                           // dD/deps = 1
  }
}

__global__ void psf_zero_kernel(float* __restrict__ p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0f;
}

int64_t psf_blocks(int64_t n0, int64_t n1) {
  return (n0 * n1 + kRowsPerBlock - 1) / kRowsPerBlock;
}

}  // namespace

}  // namespace dmic

using namespace dmic;

extern "C" {

int dmic_psf_gaussian(const float* S, const float* x0, const float* x1, const float* x2,
                      int64_t n0, int64_t n1, int64_t n2, float* out, hipStream_t stream) {
  DMIC_REQUIRE(n0 >= 0 && n1 >= 0 && n2 >= 0, DMIC_ERR_ARG, "negative box shape");
  if (n0 * n1 * n2 == 0) return DMIC_OK;
  DMIC_REQUIRE(S && x0 && x1 && x2 && out, DMIC_ERR_ARG, "null pointer");
  const Grid3 g{x0, x1, x2, n0, n1, n2};
  const unsigned nblk = (unsigned)psf_blocks(n0, n1);
  hipLaunchKernelGGL(psf_gauss_fwd_kernel, dim3(nblk), dim3(kPsfThreads), 0, stream, S, g, out);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

size_t dmic_psf_grad_workspace(int64_t n0, int64_t n1, int64_t n2) {
  if (n0 < 0 || n1 < 0 || n2 < 0) return 0;
  const int64_t nb = psf_blocks(n0, n1);
  return (size_t)(nb > 0 ? nb : 1) * 8 * sizeof(double);
}

int dmic_psf_gaussian_grad(const float* grad_out, const float* psf, const float* x0,
                           const float* x1, const float* x2, int64_t n0, int64_t n1, int64_t n2,
                           float* grad_S, void* workspace, size_t workspace_bytes,
                           hipStream_t stream) {
  DMIC_REQUIRE(n0 >= 0 && n1 >= 0 && n2 >= 0, DMIC_ERR_ARG, "negative box shape");
  DMIC_REQUIRE(grad_S, DMIC_ERR_ARG, "null gradient output");
  if (n0 * n1 * n2 == 0) {
    hipLaunchKernelGGL(psf_zero_kernel, dim3(1), dim3(64), 0, stream, grad_S, 9);
    DMIC_LAUNCH_CHECK();
    return DMIC_OK;
  }
  DMIC_REQUIRE(grad_out && psf && x0 && x1 && x2, DMIC_ERR_ARG, "null pointer");
  DMIC_REQUIRE(workspace && workspace_bytes >= dmic_psf_grad_workspace(n0, n1, n2),
               DMIC_ERR_WORKSPACE, "workspace too small");
  double* part = static_cast<double*>(workspace);
  const Grid3 g{x0, x1, x2, n0, n1, n2};
  const unsigned nblk = (unsigned)psf_blocks(n0, n1);
  hipLaunchKernelGGL(psf_gauss_grad_kernel, dim3(nblk), dim3(kPsfThreads), 0, stream,
                     grad_out, psf, g, part);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(psf_gauss_finish_kernel, dim3(1), dim3(kPsfThreads), 0, stream, part,
                     (int64_t)nblk, grad_S);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_psf_aniso(const float* prm, const float* x0, const float* x1, const float* x2,
                   int64_t n0, int64_t n1, int64_t n2, float* out, hipStream_t stream) {
  DMIC_REQUIRE(n0 >= 0 && n1 >= 0 && n2 >= 0, DMIC_ERR_ARG, "negative box shape");
  if (n0 * n1 * n2 == 0) return DMIC_OK;
  DMIC_REQUIRE(prm && x0 && x1 && x2 && out, DMIC_ERR_ARG, "null pointer");
  const Grid3 g{x0, x1, x2, n0, n1, n2};
  const unsigned nblk = (unsigned)psf_blocks(n0, n1);
  hipLaunchKernelGGL(psf_aniso_fwd_kernel, dim3(nblk), dim3(kPsfThreads), 0, stream,
                     prm, g, out);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_psf_aniso_grad(const float* prm, const float* grad_out, const float* psf,
                        const float* x0, const float* x1, const float* x2, int64_t n0, int64_t n1,
                        int64_t n2, float* grad_prm, void* workspace, size_t workspace_bytes,
                        hipStream_t stream) {
  DMIC_REQUIRE(n0 >= 0 && n1 >= 0 && n2 >= 0, DMIC_ERR_ARG, "negative box shape");
  DMIC_REQUIRE(grad_prm && prm, DMIC_ERR_ARG, "null pointer");
  if (n0 * n1 * n2 == 0) {
    hipLaunchKernelGGL(psf_zero_kernel, dim3(1), dim3(64), 0, stream, grad_prm, 6);
    DMIC_LAUNCH_CHECK();
    return DMIC_OK;
  }
  DMIC_REQUIRE(grad_out && psf && x0 && x1 && x2, DMIC_ERR_ARG, "null pointer");
  DMIC_REQUIRE(workspace && workspace_bytes >= dmic_psf_grad_workspace(n0, n1, n2),
               DMIC_ERR_WORKSPACE, "workspace too small");
  double* part = static_cast<double*>(workspace);
  const Grid3 g{x0, x1, x2, n0, n1, n2};
  const unsigned nblk = (unsigned)psf_blocks(n0, n1);
  hipLaunchKernelGGL(psf_aniso_grad_kernel, dim3(nblk), dim3(kPsfThreads), 0, stream,
                     prm, grad_out, psf, g, part);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(psf_aniso_finish_kernel, dim3(1), dim3(kPsfThreads), 0, stream, prm, part,
                     (int64_t)nblk, grad_prm);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

}  // extern "C"
