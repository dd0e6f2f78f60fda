// Large-steps preconditioner of the fit loop for gfx950: x = (I + lambda L)^-power g,
// L = D - A the combinatorial Laplacian of the mesh's vertex graph.
//
// The reference factorises M^2 = (I + lambda L)^2 once with a sparse
// Cholesky solver and back-substitutes the vertex gradient every step
// (Examples/Gradient_Descent.ipynb cell 2: CholeskySolver(M @ M),
// solver.solve(Verts.grad)).  Here M stays sparse (CSR adjacency, O(nnz)
// memory) and each application runs `power` conjugate-gradient solves with
// M itself (condition number <= 1 + 2 lambda max_degree, a few hundred at
// lambda = 50, instead of its square for M^2), Jacobi preconditioned, for
// the three coordinate columns at once.
//
// One workgroup of 1024 threads runs the whole iteration: the vectors
// (FP64, [nv][3]) stay in L2, the dot products are fixed-order LDS
// reductions, so results are bitwise reproducible, and every thread leaves
// the loop on the same reduced values (or at max_iter).
#include "common.h"

namespace dmic {

constexpr int kCgThreads = 1024;

// Sum of v over the workgroup into red[0..2] (three columns), fixed order.
__device__ __forceinline__ void block_sum3(double a, double b, double c, double* red) {
  const int t = threadIdx.x;
  red[t] = a;
  red[kCgThreads + t] = b;
  red[2 * kCgThreads + t] = c;
  __syncthreads();
  for (int w = kCgThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      red[t] += red[t + w];
      red[kCgThreads + t] += red[kCgThreads + t + w];
      red[2 * kCgThreads + t] += red[2 * kCgThreads + t + w];
    }
    __syncthreads();
  }
}

// y = M p, M = I + lambda (D - A), rows i = threadIdx.x + k * 1024.
__device__ __forceinline__ void apply_m(const int* __restrict__ row, const int* __restrict__ col,
                                        int64_t nv, double lambda, const double* __restrict__ p,
                                        double* __restrict__ y) {
  for (int64_t i = threadIdx.x; i < nv; i += kCgThreads) {
    const int b = row[i], e = row[i + 1];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int k = b; k < e; ++k) {
      const int64_t j = col[k];
      s0 += p[3 * j];
      s1 += p[3 * j + 1];
      s2 += p[3 * j + 2];
    }
    const double d = 1.0 + lambda * (double)(e - b);
    y[3 * i] = d * p[3 * i] - lambda * s0;
    y[3 * i + 1] = d * p[3 * i + 1] - lambda * s1;
    y[3 * i + 2] = d * p[3 * i + 2] - lambda * s2;
  }
}

// `power` successive solves M x_k = x_{k-1}, x_0 = g.  Workspace: x, r, z
// (reused as the preconditioned residual), p, Ap, each FP64 [nv][3].
__global__ __launch_bounds__(kCgThreads) void laplacian_cg_kernel(
    const int* __restrict__ row, const int* __restrict__ col, int64_t nv, double lambda,
    int power, const float* __restrict__ g, float* __restrict__ out, double tol, int max_iter,
    double* __restrict__ ws, int* __restrict__ iters_out) {
  __shared__ double red[3 * kCgThreads];
  double* x = ws;
  double* r = ws + 3 * nv;
  double* z = ws + 6 * nv;
  double* p = ws + 9 * nv;
  double* Ap = ws + 12 * nv;
  double* rhs = ws + 15 * nv;
  const int t = threadIdx.x;
  for (int64_t i = t; i < 3 * nv; i += kCgThreads) rhs[i] = (double)g[i];
  __syncthreads();
  int last_iters = 0;
  for (int pass = 0; pass < power; ++pass) {
    // x = 0, r = rhs, z = D_M^-1 r, p = z
    double bb0 = 0, bb1 = 0, bb2 = 0, rz0 = 0, rz1 = 0, rz2 = 0;
    for (int64_t i = t; i < nv; i += kCgThreads) {
      const double dinv = 1.0 / (1.0 + lambda * (double)(row[i + 1] - row[i]));
      for (int c = 0; c < 3; ++c) {
        const double b = rhs[3 * i + c];
        x[3 * i + c] = 0.0;
        r[3 * i + c] = b;
        z[3 * i + c] = dinv * b;
        p[3 * i + c] = dinv * b;
      }
      bb0 += rhs[3 * i] * rhs[3 * i];
      bb1 += rhs[3 * i + 1] * rhs[3 * i + 1];
      bb2 += rhs[3 * i + 2] * rhs[3 * i + 2];
      rz0 += rhs[3 * i] * z[3 * i];
      rz1 += rhs[3 * i + 1] * z[3 * i + 1];
      rz2 += rhs[3 * i + 2] * z[3 * i + 2];
    }
    block_sum3(bb0, bb1, bb2, red);
    const double lim[3] = {tol * tol * red[0], tol * tol * red[kCgThreads],
                           tol * tol * red[2 * kCgThreads]};
    __syncthreads();
    block_sum3(rz0, rz1, rz2, red);
    double rz[3] = {red[0], red[kCgThreads], red[2 * kCgThreads]};
    __syncthreads();
    int it = 0;
    bool done = (rz[0] <= 0.0 && rz[1] <= 0.0 && rz[2] <= 0.0);
    while (!done && it < max_iter) {
      apply_m(row, col, nv, lambda, p, Ap);
      __syncthreads();
      double q0 = 0, q1 = 0, q2 = 0;
      for (int64_t i = t; i < nv; i += kCgThreads) {
        q0 += p[3 * i] * Ap[3 * i];
        q1 += p[3 * i + 1] * Ap[3 * i + 1];
        q2 += p[3 * i + 2] * Ap[3 * i + 2];
      }
      block_sum3(q0, q1, q2, red);
      // a column whose residual is already zero keeps alpha = 0
      double alpha[3];
      for (int c = 0; c < 3; ++c) {
        const double q = red[c * kCgThreads];
        alpha[c] = (q > 0.0) ? rz[c] / q : 0.0;
      }
      __syncthreads();
      double rr0 = 0, rr1 = 0, rr2 = 0, n0 = 0, n1 = 0, n2 = 0;
      for (int64_t i = t; i < nv; i += kCgThreads) {
        const double dinv = 1.0 / (1.0 + lambda * (double)(row[i + 1] - row[i]));
        for (int c = 0; c < 3; ++c) {
          const int64_t k = 3 * i + c;
          x[k] += alpha[c] * p[k];
          r[k] -= alpha[c] * Ap[k];
          z[k] = dinv * r[k];
        }
        rr0 += r[3 * i] * r[3 * i];
        rr1 += r[3 * i + 1] * r[3 * i + 1];
        rr2 += r[3 * i + 2] * r[3 * i + 2];
        n0 += r[3 * i] * z[3 * i];
        n1 += r[3 * i + 1] * z[3 * i + 1];
        n2 += r[3 * i + 2] * z[3 * i + 2];
      }
      block_sum3(rr0, rr1, rr2, red);
      const double rr[3] = {red[0], red[kCgThreads], red[2 * kCgThreads]};
      __syncthreads();
      block_sum3(n0, n1, n2, red);
      double beta[3];
      for (int c = 0; c < 3; ++c) {
        const double rzn = red[c * kCgThreads];
        beta[c] = (rz[c] > 0.0) ? rzn / rz[c] : 0.0;
        rz[c] = rzn;
      }
      __syncthreads();
      for (int64_t i = t; i < nv; i += kCgThreads)
        for (int c = 0; c < 3; ++c) {
          const int64_t k = 3 * i + c;
          p[k] = z[k] + beta[c] * p[k];
        }
      __syncthreads();
      ++it;
      done = (rr[0] <= lim[0]) && (rr[1] <= lim[1]) && (rr[2] <= lim[2]);
    }
    last_iters = it;
    // this pass's solution is the next pass's right-hand side
    for (int64_t i = t; i < 3 * nv; i += kCgThreads) rhs[i] = x[i];
    __syncthreads();
  }
  for (int64_t i = t; i < 3 * nv; i += kCgThreads) out[i] = (float)rhs[i];
  if (t == 0 && iters_out) iters_out[0] = last_iters;
}

}  // namespace dmic

using namespace dmic;

extern "C" {

size_t dmic_laplacian_solve_workspace(int64_t nv) {
  if (nv < 0) return 0;
  Carver c(nullptr);
  c.take<double>((size_t)18 * (nv > 0 ? nv : 1));
  return c.bytes();
}

int dmic_laplacian_solve(const int* row_ptr, const int* col, int64_t nv, float lambda, int power,
                         const float* g, float* x, float tol, int max_iter, int* iters_out,
                         void* workspace, size_t workspace_bytes, hipStream_t stream) {
  DMIC_REQUIRE(nv >= 0 && power >= 0 && max_iter >= 0, DMIC_ERR_ARG, "negative size");
  DMIC_REQUIRE(nv < (int64_t(1) << 30), DMIC_ERR_ARG, "size too large");
  DMIC_REQUIRE(lambda >= 0.0f && tol >= 0.0f, DMIC_ERR_ARG,
               "lambda and tol must be non-negative (M = I + lambda L is then SPD)");
  if (nv == 0) return DMIC_OK;
  DMIC_REQUIRE(row_ptr && col && g && x, DMIC_ERR_ARG, "null pointer");
  Carver c(workspace);
  double* ws = c.take<double>((size_t)18 * nv);
  DMIC_REQUIRE(workspace && workspace_bytes >= c.bytes(), DMIC_ERR_WORKSPACE,
               "laplacian solve workspace too small: %zu < %zu bytes", workspace_bytes, c.bytes());
  hipLaunchKernelGGL(laplacian_cg_kernel, dim3(1), dim3(kCgThreads), 0, stream, row_ptr, col, nv,
                     (double)lambda, power, g, x, (double)tol, max_iter, ws, iters_out);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

}  // extern "C"
