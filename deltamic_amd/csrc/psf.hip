// Learnable Gaussian PSFs on a separable grid, forward and parameter gradient.
//
// The reference builds these PSFs with torch meshgrids (deltamic/psf.py:551-665,
// Gaussian_psf, Gaussian_psf_logsigma, Anisotropic_gaussian_psf) and lets
// autograd differentiate them; every iteration of its fit loop evaluates the
// PSF, takes its FFT as the OTF and steps the PSF parameters
// (Examples/Gradient_Descent.ipynb cell 2).  Here the forward is one pass that
// writes the box (4 bytes per voxel, the three 1-D axes stay in cache) and the
// backward is one pass that reads the upstream gradient and the saved PSF
// (8 bytes per voxel) and reduces the few parameter sums per workgroup in
// FP64; a single workgroup then adds the workgroup partials in a fixed order,
// so gradients are bitwise reproducible.  Both are HBM-bound.
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
constexpr int kGaussSums = 6;  // x0x0 x0x1 x0x2 x1x1 x1x2 x2x2
constexpr int kAnisoSums = 5;  // T00 T01 T11 U V

// Launch shape shared by all kernels: blockIdx.z = i (x0), blockIdx.y = j
// (x1), blockIdx.x * 256 + threadIdx.x = k (x2).  Each workgroup covers part
// of one row of the box, so the per-row factors of the gradient sums are
// applied once per workgroup.
struct Grid3 {
  const float* x0;
  const float* x1;
  const float* x2;
  int64_t n1, n2;
};

__device__ __forceinline__ bool voxel(const Grid3& g, int64_t& v, float& a, float& b, float& c) {
  const int64_t k = (int64_t)blockIdx.x * kPsfThreads + threadIdx.x;
  a = g.x0[blockIdx.z];
  b = g.x1[blockIdx.y];
  v = ((int64_t)blockIdx.z * g.n1 + blockIdx.y) * g.n2 + k;
  if (k >= g.n2) return false;
  c = g.x2[k];
  return true;
}

__device__ __forceinline__ int64_t block_id() {
  return ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
}

__global__ __launch_bounds__(kPsfThreads) void psf_gauss_fwd_kernel(const float* __restrict__ S,
                                                                    Grid3 g,
                                                                    float* __restrict__ out) {
  int64_t v;
  float a, b, c;
  if (!voxel(g, v, a, b, c)) return;
  // psf.py:573-576 operation order
  const float q = a * (S[0] * a + S[1] * b + S[2] * c) + b * (S[3] * a + S[4] * b + S[5] * c) +
                  c * (S[6] * a + S[7] * b + S[8] * c);
  out[v] = expf(-q / 2.0f);
}

// prm = {S2_00, S2_01, S2_10, S2_11, sz, eps}
__global__ __launch_bounds__(kPsfThreads) void psf_aniso_fwd_kernel(const float* __restrict__ prm,
                                                                    Grid3 g,
                                                                    float* __restrict__ out) {
  int64_t v;
  float a, b, c;
  if (!voxel(g, v, a, b, c)) return;
  // psf.py:656-658 operation order
  const float Q = a * (prm[0] * a + prm[1] * b) + b * (prm[2] * a + prm[3] * b);
  const float t = __builtin_fabsf(c) * prm[4];
  const float D = prm[5] + t * t;
  out[v] = expf(-Q / (2.0f * D)) / D;
}

// Workgroup sum of R per-thread values in FP64 (fixed tree order); the result
// is in red[r][0] for thread 0.
template <int R>
__device__ __forceinline__ void block_reduce(double (&red)[R][kPsfThreads], const double (&s)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) red[r][threadIdx.x] = s[r];
  __syncthreads();
  for (int w = kPsfThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int r = 0; r < R; ++r) red[r][threadIdx.x] += red[r][threadIdx.x + w];
    }
    __syncthreads();
  }
}

// Row sums R0 = sum w, R1 = sum w x2, R2 = sum w x2^2 (w = G psf); the six
// pair sums are x0^2 R0, x0 x1 R0, x0 R1, x1^2 R0, x1 R1, R2.
__global__ __launch_bounds__(kPsfThreads) void psf_gauss_grad_kernel(
    const float* __restrict__ G, const float* __restrict__ psf, Grid3 g,
    double* __restrict__ part) {
  __shared__ double red[3][kPsfThreads];
  int64_t v;
  float a, b, c;
  double s[3] = {0.0, 0.0, 0.0};
  if (voxel(g, v, a, b, c)) {
    const double w = (double)(G[v] * psf[v]);
    s[0] = w;
    s[1] = w * c;
    s[2] = w * c * c;
  }
  block_reduce<3>(red, s);
  if (threadIdx.x == 0) {
    const double A = a, B = b, R0 = red[0][0], R1 = red[1][0], R2 = red[2][0];
    double* o = part + block_id() * kGaussSums;
    o[0] = A * A * R0;
    o[1] = A * B * R0;
    o[2] = A * R1;
    o[3] = B * B * R0;
    o[4] = B * R1;
    o[5] = R2;
  }
}

// Row sums A = sum w/D, B = sum w/D^2, C = sum w x2^2/D^2, E = sum w x2^2/D;
// with the row's Q: T_ab = x_a x_b A, U = Q/2 B - A, V = Q/2 C - E.
__global__ __launch_bounds__(kPsfThreads) void psf_aniso_grad_kernel(
    const float* __restrict__ prm, const float* __restrict__ G, const float* __restrict__ psf,
    Grid3 g, double* __restrict__ part) {
  __shared__ double red[4][kPsfThreads];
  int64_t v;
  float a, b, c;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  if (voxel(g, v, a, b, c)) {
    const float t = __builtin_fabsf(c) * prm[4];
    const double rD = 1.0 / (double)(prm[5] + t * t);
    const double w = (double)(G[v] * psf[v]);
    const double c2 = (double)c * c;
    s[0] = w * rD;
    s[1] = w * rD * rD;
    s[2] = w * c2 * rD * rD;
    s[3] = w * c2 * rD;
  }
  block_reduce<4>(red, s);
  if (threadIdx.x == 0) {
    const double A0 = a, B0 = b;
    const double Q = (double)(a * (prm[0] * a + prm[1] * b) + b * (prm[2] * a + prm[3] * b));
    const double A = red[0][0], B = red[1][0], C = red[2][0], E = red[3][0];
    double* o = part + block_id() * kAnisoSums;
    o[0] = A0 * A0 * A;
    o[1] = A0 * B0 * A;
    o[2] = B0 * B0 * A;
    o[3] = 0.5 * Q * B - A;
    o[4] = 0.5 * Q * C - E;
  }
}

// One workgroup: fixed-order sum of the workgroup partials.
template <int K>
__device__ __forceinline__ void finish_sums(const double* __restrict__ part, int64_t nblocks,
                                            double (&red)[K][kPsfThreads]) {
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (int64_t b = threadIdx.x; b < nblocks; b += kPsfThreads)
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += part[b * K + k];
  block_reduce<K>(red, acc);
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

dim3 psf_grid(int64_t n0, int64_t n1, int64_t n2) {
  return dim3(ceil_div(n2, kPsfThreads), (unsigned)n1, (unsigned)n0);
}

int64_t psf_blocks(int64_t n0, int64_t n1, int64_t n2) {
  return (int64_t)ceil_div(n2, kPsfThreads) * n1 * n0;
}

// grid y and z are limited to 65535 on this target
constexpr int64_t kMaxGridYZ = 65535;

}  // namespace

}  // namespace dmic

using namespace dmic;

extern "C" {

int dmic_psf_gaussian(const float* S, const float* x0, const float* x1, const float* x2,
                      int64_t n0, int64_t n1, int64_t n2, float* out, hipStream_t stream) {
  DMIC_REQUIRE(n0 >= 0 && n1 >= 0 && n2 >= 0, DMIC_ERR_ARG, "negative box shape");
  DMIC_REQUIRE(n0 <= kMaxGridYZ && n1 <= kMaxGridYZ, DMIC_ERR_ARG, "box dimension above 65535");
  if (n0 * n1 * n2 == 0) return DMIC_OK;
  DMIC_REQUIRE(S && x0 && x1 && x2 && out, DMIC_ERR_ARG, "null pointer");
  const Grid3 g{x0, x1, x2, n1, n2};
  hipLaunchKernelGGL(psf_gauss_fwd_kernel, psf_grid(n0, n1, n2), dim3(kPsfThreads), 0, stream, S,
                     g, out);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

size_t dmic_psf_grad_workspace(int64_t n0, int64_t n1, int64_t n2) {
  if (n0 < 0 || n1 < 0 || n2 < 0) return 0;
  const int64_t nb = psf_blocks(n0, n1, n2);
  return (size_t)(nb > 0 ? nb : 1) * 8 * sizeof(double);
}

int dmic_psf_gaussian_grad(const float* grad_out, const float* psf, const float* x0,
                           const float* x1, const float* x2, int64_t n0, int64_t n1, int64_t n2,
                           float* grad_S, void* workspace, size_t workspace_bytes,
                           hipStream_t stream) {
  DMIC_REQUIRE(n0 >= 0 && n1 >= 0 && n2 >= 0, DMIC_ERR_ARG, "negative box shape");
  DMIC_REQUIRE(n0 <= kMaxGridYZ && n1 <= kMaxGridYZ, DMIC_ERR_ARG, "box dimension above 65535");
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
  const Grid3 g{x0, x1, x2, n1, n2};
  hipLaunchKernelGGL(psf_gauss_grad_kernel, psf_grid(n0, n1, n2), dim3(kPsfThreads), 0, stream,
                     grad_out, psf, g, part);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(psf_gauss_finish_kernel, dim3(1), dim3(kPsfThreads), 0, stream, part,
                     psf_blocks(n0, n1, n2), grad_S);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_psf_aniso(const float* prm, const float* x0, const float* x1, const float* x2,
                   int64_t n0, int64_t n1, int64_t n2, float* out, hipStream_t stream) {
  DMIC_REQUIRE(n0 >= 0 && n1 >= 0 && n2 >= 0, DMIC_ERR_ARG, "negative box shape");
  DMIC_REQUIRE(n0 <= kMaxGridYZ && n1 <= kMaxGridYZ, DMIC_ERR_ARG, "box dimension above 65535");
  if (n0 * n1 * n2 == 0) return DMIC_OK;
  DMIC_REQUIRE(prm && x0 && x1 && x2 && out, DMIC_ERR_ARG, "null pointer");
  const Grid3 g{x0, x1, x2, n1, n2};
  hipLaunchKernelGGL(psf_aniso_fwd_kernel, psf_grid(n0, n1, n2), dim3(kPsfThreads), 0, stream,
                     prm, g, out);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

int dmic_psf_aniso_grad(const float* prm, const float* grad_out, const float* psf,
                        const float* x0, const float* x1, const float* x2, int64_t n0, int64_t n1,
                        int64_t n2, float* grad_prm, void* workspace, size_t workspace_bytes,
                        hipStream_t stream) {
  DMIC_REQUIRE(n0 >= 0 && n1 >= 0 && n2 >= 0, DMIC_ERR_ARG, "negative box shape");
  DMIC_REQUIRE(n0 <= kMaxGridYZ && n1 <= kMaxGridYZ, DMIC_ERR_ARG, "box dimension above 65535");
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
  const Grid3 g{x0, x1, x2, n1, n2};
  hipLaunchKernelGGL(psf_aniso_grad_kernel, psf_grid(n0, n1, n2), dim3(kPsfThreads), 0, stream,
                     prm, grad_out, psf, g, part);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(psf_aniso_finish_kernel, dim3(1), dim3(kPsfThreads), 0, stream, prm, part,
                     psf_blocks(n0, n1, n2), grad_prm);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

}  // extern "C"
