// meshFT forward and adjoint kernels for gfx950, and their C-ABI launchers.
//
// Forward (replaces Mesh_NUFT_kernel, deltamic/cpp/fourier3d_cuda_kernel.cu:282-311):
//   one lane per band frequency; faces are staged through LDS in tiles of
//   256 packed 64-byte records and read as wave-uniform broadcasts.  Each
//   lane sums w_t F_t over a tile in FP32 and folds every 64 faces into an
//   FP64 accumulator.  Small meshes or bands split the face range over
//   blockIdx.y and a fixed-order reduction adds the FP64 partials.
//
// Adjoint (replaces Grad_Mesh_NUFT_kernel + Additions_kernel, .cu:315-416):
//   one lane per face; band frequencies (xi, conj of the upstream gradient)
//   are staged through LDS.  Each lane accumulates ten per-face sums over a
//   frequency chunk,
//     P_a = sum Re(conj g dF/dx_a) xi,  P_b = sum Re(conj g dF/dx_b) xi,
//     T   = sum Re(conj g (-iF)) xi,    R   = sum Re(conj g F),
//   (P_c = T - P_a - P_b because the three corner derivatives sum to -iF),
//   written per chunk.  A per-face kernel adds the chunks in order and forms
//   the corner gradients including the area-weight terms; a CSR
//   face->vertex gather (rows sorted by face) sums them per vertex.  No
//   floating-point atomics anywhere: results are bitwise reproducible.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "common.h"
#include "face_math.h"

namespace dmic {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// Packed per-face record of the forward tiles (64 bytes, 4 x 16B LDS reads).
struct __align__(16) FaceRec {
  double a[3];   // vertex a / (2 pi) in FP64: xi . a[] is the phase of a in revolutions
  float eb[3];   // v_b - v_a (FP32, correctly rounded)
  float ec[3];   // v_c - v_a
  float w;       // 2 c_t A_t / sum c A
  float pad[3];
};
static_assert(sizeof(FaceRec) == 64, "FaceRec layout");

// Per-frequency record of the adjoint tiles (48 bytes, 3 x 16B LDS reads).
struct __align__(16) FreqRec {
  double X[3];  // xi in FP64
  float x[3];   // xi in FP32
  float gr, gi; // upstream gradient g = dL/dRe + i dL/dIm
  float pad;
};
static_assert(sizeof(FreqRec) == 48, "FreqRec layout");

constexpr int kTile = 256;        // faces (forward) / frequencies (adjoint) per LDS tile
constexpr int kFold = 64;         // FP32 partial sums fold into FP64 every kFold pairs
#ifndef DMIC_TARGET_BLOCKS
#define DMIC_TARGET_BLOCKS 16384
#endif
// Workgroups a launch is split into (frequency blocks x face splits, face
// blocks x frequency chunks).  The resident capacity is 256 CUs x 2-3
// workgroups; 16384 makes the last partial round of workgroups a few percent
// of the kernel (on the 1.08M-frequency band of a 20480-face sphere: 2048 -> 64.7
// ms per forward+adjoint, 8192 -> 53.8 ms, 16384 -> 52.6 ms on MI355X).
constexpr int kTargetBlocks = DMIC_TARGET_BLOCKS;

// -------------------------------------------------------------------- prep

template <typename IdxT>
__device__ __forceinline__ void load_face_verts(const float* __restrict__ V, int64_t nv,
                                                const IdxT* __restrict__ Fc, int64_t t,
                                                float (&P)[3][3], int* flag) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int64_t i = (int64_t)Fc[3 * t + j];
    if (i >= 0 && i < nv) {
      P[j][0] = V[3 * i]; P[j][1] = V[3 * i + 1]; P[j][2] = V[3 * i + 2];
    } else {
      P[j][0] = P[j][1] = P[j][2] = 0.0f;
      flag[0] = 1;
    }
  }
}

// c_t A_t (FP64, cross-product norm as .cu:9-12,38-40) and the fixed-order
// sum S = sum_t c_t A_t, in one workgroup.
template <typename IdxT>
__global__ __launch_bounds__(1024) void face_area_sum_kernel(
    const float* __restrict__ V, int64_t nv, const IdxT* __restrict__ Fc, int64_t nf,
    const float* __restrict__ coeff, double* __restrict__ cA, double* __restrict__ S_out,
    int* flag) {
  __shared__ double red[1024];
  double s = 0.0;
  for (int64_t t = threadIdx.x; t < nf; t += 1024) {
    float P[3][3];
    load_face_verts(V, nv, Fc, t, P, flag);
    double u[3], q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      u[d] = (double)P[1][d] - (double)P[0][d];
      q[d] = (double)P[2][d] - (double)P[0][d];
    }
    const double n0 = u[1] * q[2] - u[2] * q[1];
    const double n1 = u[2] * q[0] - u[0] * q[2];
    const double n2 = u[0] * q[1] - u[1] * q[0];
    const double ca = (double)coeff[t] * 0.5 * sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    cA[t] = ca;
    s += ca;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) S_out[0] = red[0];
}

template <typename IdxT>
__global__ __launch_bounds__(256) void face_pack_kernel(
    const float* __restrict__ V, int64_t nv, const IdxT* __restrict__ Fc, int64_t nf,
    const double* __restrict__ cA, const double* __restrict__ S_in, FaceRec* __restrict__ out,
    int* flag) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nf) return;
  float P[3][3];
  load_face_verts(V, nv, Fc, t, P, flag);
  FaceRec r;
  const double S = S_in[0];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r.a[d] = (double)P[0][d] * kInv2Pi;
    r.eb[d] = P[1][d] - P[0][d];
    r.ec[d] = P[2][d] - P[0][d];
  }
  r.w = (S > 0.0) ? (float)(2.0 * cA[t] / S) : 0.0f;
  r.pad[0] = r.pad[1] = r.pad[2] = 0.0f;
  out[t] = r;
}

// -------------------------------------------------------------------- forward

__global__ __launch_bounds__(256) void meshft_fwd_kernel(
    const FaceRec* __restrict__ faces, int64_t nf, int64_t faces_per_split,
    const float* __restrict__ xi, int64_t n, float2* __restrict__ out,
    double2* __restrict__ part) {
  __shared__ FaceRec tile[kTile];
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool active = m < n;
  float x = 0.0f, y = 0.0f, z = 0.0f;
  if (active) {
    x = xi[3 * m];
    y = xi[3 * m + 1];
    z = xi[3 * m + 2];
  }
  const double X = x, Y = y, Z = z;
  const int64_t f0 = (int64_t)blockIdx.y * faces_per_split;
  const int64_t f1 = (f0 + faces_per_split < nf) ? f0 + faces_per_split : nf;
  double accr = 0.0, acci = 0.0;
  for (int64_t base = f0; base < f1; base += kTile) {
    const int cnt = (int)((f1 - base < kTile) ? f1 - base : kTile);
    __syncthreads();
    if ((int)threadIdx.x < cnt) tile[threadIdx.x] = faces[base + threadIdx.x];
    __syncthreads();
    for (int k0 = 0; k0 < cnt; k0 += kFold) {
      const int k1 = (k0 + kFold < cnt) ? k0 + kFold : cnt;
      float sr = 0.0f, si = 0.0f;
      for (int k = k0; k < k1; ++k) {
        const FaceRec& r = tile[k];
        const double rev_a = fma(r.a[2], Z, fma(r.a[1], Y, r.a[0] * X));
        const float p = __builtin_fmaf(r.eb[2], z, __builtin_fmaf(r.eb[1], y, r.eb[0] * x));
        const float q = __builtin_fmaf(r.ec[2], z, __builtin_fmaf(r.ec[1], y, r.ec[0] * x));
        const Sorted so = sort3(p, q);
        cf F;   // F = e_mid N, face value = m F
        float m;
        if (so.D <= kWidePhase)
          F = cmul(expmi_rev(mid_rev<false>(rev_a, so.mid)), face_H<false>(so, m));
        else
          F = cmul(expmi_rev(mid_rev<true>(rev_a, so.mid)), face_H<true>(so, m));
        const float wm = r.w * m;
        sr = __builtin_fmaf(wm, F.re, sr);
        si = __builtin_fmaf(wm, F.im, si);
      }
      accr += (double)sr;
      acci += (double)si;
    }
  }
  if (active) {
    if (part)
      part[(int64_t)blockIdx.y * n + m] = make_double2(accr, acci);
    else
      out[m] = make_float2((float)accr, (float)acci);
  }
}

__global__ __launch_bounds__(256) void fwd_split_reduce_kernel(const double2* __restrict__ part,
                                                               int nsplit, int64_t n,
                                                               float2* __restrict__ out) {
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= n) return;
  double re = 0.0, im = 0.0;
  for (int s = 0; s < nsplit; ++s) {
    const double2 v = part[(int64_t)s * n + m];
    re += v.x;
    im += v.y;
  }
  out[m] = make_float2((float)re, (float)im);
}

// -------------------------------------------------------------------- adjoint

__global__ __launch_bounds__(256) void meshft_bwd_kernel(
    const FaceRec* __restrict__ faces, int64_t nf, const float* __restrict__ xi,
    const float* __restrict__ grad, int64_t n, int64_t freqs_per_chunk,
    double* __restrict__ part /* [nchunk][10][nf] */) {
  __shared__ FreqRec tile[kTile];
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool active = t < nf;
  double A0 = 0.0, A1 = 0.0, A2 = 0.0;
  float b0 = 0.0f, b1 = 0.0f, b2 = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  if (active) {
    const FaceRec r = faces[t];
    A0 = r.a[0]; A1 = r.a[1]; A2 = r.a[2];
    b0 = r.eb[0]; b1 = r.eb[1]; b2 = r.eb[2];
    c0 = r.ec[0]; c1 = r.ec[1]; c2 = r.ec[2];
  }
  const int64_t m0 = (int64_t)blockIdx.y * freqs_per_chunk;
  const int64_t m1 = (m0 + freqs_per_chunk < n) ? m0 + freqs_per_chunk : n;
  double acc[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) acc[i] = 0.0;
  for (int64_t base = m0; base < m1; base += kTile) {
    const int cnt = (int)((m1 - base < kTile) ? m1 - base : kTile);
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
      const int64_t m = base + threadIdx.x;
      FreqRec fr;
      fr.x[0] = xi[3 * m]; fr.x[1] = xi[3 * m + 1]; fr.x[2] = xi[3 * m + 2];
      fr.X[0] = fr.x[0]; fr.X[1] = fr.x[1]; fr.X[2] = fr.x[2];
      fr.gr = grad[2 * m];
      fr.gi = grad[2 * m + 1];
      fr.pad = 0.0f;
      tile[threadIdx.x] = fr;
    }
    __syncthreads();
    for (int k0 = 0; k0 < cnt; k0 += kFold) {
      const int k1 = (k0 + kFold < cnt) ? k0 + kFold : cnt;
      float a[10];
#pragma unroll
      for (int i = 0; i < 10; ++i) a[i] = 0.0f;
      for (int k = k0; k < k1; ++k) {
        const FreqRec& fr = tile[k];
        const double rev_a = fma(A2, fr.X[2], fma(A1, fr.X[1], A0 * fr.X[0]));
        const float p = __builtin_fmaf(b2, fr.x[2], __builtin_fmaf(b1, fr.x[1], b0 * fr.x[0]));
        const float q = __builtin_fmaf(c2, fr.x[2], __builtin_fmaf(c1, fr.x[1], c0 * fr.x[0]));
        const Sorted so = sort3(p, q);
        cf em, d, pP, pQ;
        float rD;
        if (so.D <= kWidePhase) {
          em = expmi_rev(mid_rev<false>(rev_a, so.mid));
          face_H_grad<false>(so, d, pP, pQ, rD);
        } else {
          em = expmi_rev(mid_rev<true>(rev_a, so.mid));
          face_H_grad<true>(so, d, pP, pQ, rD);
        }
        // Contract with conj(g) first: z = conj(g) e_mid, then with H = d rD
        //   rho = Re(conj g F) = Re(z H),  gamma = Re(conj g (-iF)) = Im(z H),
        //   Re(conj g dF/dx_lo) = (Re(z pP) + rho) rD,
        //   Re(conj g dF/dx_hi) = -(Re(z pQ) + rho) rD,
        //   Re(conj g dF/dx_mid) = gamma - (lo) - (hi).
        const cf z = cmul({fr.gr, -fr.gi}, em);
        const float rh = __builtin_fmaf(z.re, d.re, -z.im * d.im) * rD;
        const float ga = __builtin_fmaf(z.re, d.im, z.im * d.re) * rD;
        const float alo = (__builtin_fmaf(z.re, pP.re, -z.im * pP.im) + rh) * rD;
        const float ahi = -(__builtin_fmaf(z.re, pQ.re, -z.im * pQ.im) + rh) * rD;
        const float amid = ga - alo - ahi;
        const int ra = (p < 0.0f) + (q < 0.0f);
        const int rb = (p >= 0.0f) + (q < p);
        const float al = (ra == 0) ? alo : ((ra == 2) ? ahi : amid);
        const float be = (rb == 0) ? alo : ((rb == 2) ? ahi : amid);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          a[j] = __builtin_fmaf(al, fr.x[j], a[j]);
          a[3 + j] = __builtin_fmaf(be, fr.x[j], a[3 + j]);
          a[6 + j] = __builtin_fmaf(ga, fr.x[j], a[6 + j]);
        }
        a[9] += rh;
      }
#pragma unroll
      for (int i = 0; i < 10; ++i) acc[i] += a[i];
    }
  }
  if (active) {
    double* dst = part + (int64_t)blockIdx.y * 10 * nf + t;
#pragma unroll
    for (int i = 0; i < 10; ++i) dst[(int64_t)i * nf] = acc[i];
  }
}

// Per-face: add the chunk partials in order, then form the corner gradients
//   G_tj  = w_t P_tj + (2 c_t / S) R_t dA_t/dv_j
//   cgA_tj = c_t dA_t/dv_j
// and wR_t = w_t R_t (for K = sum_t w_t R_t = sum_m Re(conj g_m ftmesh_m)).
// dA/dv_j = 1/2 n_hat x (v_{j+2} - v_{j+1}); a zero-area face has no normal
// and contributes no area gradient (the reference divides by zero, .cu:24).
template <typename IdxT>
__global__ __launch_bounds__(256) void bwd_face_finalize_kernel(
    const float* __restrict__ V, int64_t nv, const IdxT* __restrict__ Fc, int64_t nf,
    const float* __restrict__ coeff, const double* __restrict__ cA, const double* __restrict__ S_in,
    const double* __restrict__ part, int nchunk, double* __restrict__ G, double* __restrict__ cgA,
    double* __restrict__ wR, int* flag) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nf) return;
  double s[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) s[i] = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    const double* src = part + (int64_t)c * 10 * nf + t;
#pragma unroll
    for (int i = 0; i < 10; ++i) s[i] += src[(int64_t)i * nf];
  }
  float Pf[3][3];
  load_face_verts(V, nv, Fc, t, Pf, flag);
  double P[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int d = 0; d < 3; ++d) P[j][d] = (double)Pf[j][d];
  double u[3], q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u[d] = P[1][d] - P[0][d];
    q[d] = P[2][d] - P[0][d];
  }
  double nr[3] = {u[1] * q[2] - u[2] * q[1], u[2] * q[0] - u[0] * q[2], u[0] * q[1] - u[1] * q[0]};
  const double nn = sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
  const double inv = (nn > 0.0) ? 1.0 / nn : 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) nr[d] *= inv;
  const double S = S_in[0];
  const double ct = (double)coeff[t];
  const double w = (S > 0.0) ? 2.0 * cA[t] / S : 0.0;
  const double aw = (S > 0.0) ? 2.0 * ct / S * s[9] : 0.0;
  double Pj[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    Pj[0][d] = s[d];
    Pj[1][d] = s[3 + d];
    Pj[2][d] = s[6 + d] - s[d] - s[3 + d];
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double* e2 = P[(j + 2) % 3];
    const double* e1 = P[(j + 1) % 3];
    const double e[3] = {e2[0] - e1[0], e2[1] - e1[1], e2[2] - e1[2]};
    const double gA[3] = {0.5 * (nr[1] * e[2] - nr[2] * e[1]), 0.5 * (nr[2] * e[0] - nr[0] * e[2]),
                          0.5 * (nr[0] * e[1] - nr[1] * e[0])};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      G[9 * t + 3 * j + d] = w * Pj[j][d] + aw * gA[d];
      cgA[9 * t + 3 * j + d] = ct * gA[d];
    }
  }
  wR[t] = w * s[9];
}

// Fixed-order sum of `count` doubles in one workgroup.
__global__ __launch_bounds__(1024) void sum_f64_kernel(const double* __restrict__ x, int64_t count,
                                                       double* __restrict__ out) {
  __shared__ double red[1024];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += 1024) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// -------------------------------------------------------------------- CSR face->vertex

__global__ void fill_i32_kernel(int* __restrict__ p, int64_t count, int v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < count) p[i] = v;
}

template <typename IdxT>
__global__ __launch_bounds__(256) void csr_count_kernel(const IdxT* __restrict__ Fc, int64_t nf,
                                                        int64_t nv, int* __restrict__ cnt) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= 3 * nf) return;
  const int64_t v = (int64_t)Fc[e];
  if (v >= 0 && v < nv) atomicAdd(&cnt[v], 1);  // integer: order-independent
}

// Exclusive scan of cnt[0..count) into off[0..count], off[count] = total, in one workgroup.
__global__ __launch_bounds__(1024) void scan_i32_kernel(const int* __restrict__ cnt, int64_t count,
                                                        int* __restrict__ off) {
  __shared__ int64_t sums[1024];
  const int64_t per = (count + 1023) / 1024;
  const int64_t b = threadIdx.x * per;
  const int64_t e = (b + per < count) ? b + per : count;
  int64_t s = 0;
  for (int64_t i = b; i < e; ++i) s += cnt[i];
  sums[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int i = 0; i < 1024; ++i) {
      const int64_t v = sums[i];
      sums[i] = run;
      run += v;
    }
    off[count] = (int)run;
  }
  __syncthreads();
  int64_t run = sums[threadIdx.x];
  for (int64_t i = b; i < e; ++i) {
    off[i] = (int)run;
    run += cnt[i];
  }
}

template <typename IdxT>
__global__ __launch_bounds__(256) void csr_fill_kernel(const IdxT* __restrict__ Fc, int64_t nf,
                                                       int64_t nv, const int* __restrict__ row,
                                                       int* __restrict__ fillc,
                                                       int* __restrict__ ent) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= 3 * nf) return;
  const int64_t v = (int64_t)Fc[e];
  if (v < 0 || v >= nv) return;
  const int slot = atomicAdd(&fillc[v], 1);
  ent[row[v] + slot] = (int)e;
}

// Sorts each row by incidence id (3 t + j): the fill order above is
// arbitrary, the sorted order makes every vertex sum reproducible.
__global__ __launch_bounds__(256) void csr_sort_rows_kernel(const int* __restrict__ row, int64_t nv,
                                                            int* __restrict__ ent) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const int b = row[v], e = row[v + 1];
  for (int i = b + 1; i < e; ++i) {
    const int key = ent[i];
    int j = i - 1;
    while (j >= b && ent[j] > key) {
      ent[j + 1] = ent[j];
      --j;
    }
    ent[j + 1] = key;
  }
}

__global__ __launch_bounds__(256) void vertex_gather_kernel(
    const int* __restrict__ row, const int* __restrict__ ent, int64_t nv,
    const double* __restrict__ G, const double* __restrict__ cgA, const double* __restrict__ K_in,
    const double* __restrict__ S_in, float* __restrict__ grad) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const double S = S_in[0];
  const double kS = (S > 0.0) ? K_in[0] / S : 0.0;
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  for (int i = row[v]; i < row[v + 1]; ++i) {
    const int64_t e = ent[i];  // = 3 t + j, so G[3e + d] = G[t][j][d]
    g0 += G[3 * e] - kS * cgA[3 * e];
    g1 += G[3 * e + 1] - kS * cgA[3 * e + 1];
    g2 += G[3 * e + 2] - kS * cgA[3 * e + 2];
  }
  grad[3 * v] = (float)g0;
  grad[3 * v + 1] = (float)g1;
  grad[3 * v + 2] = (float)g2;
}

// -------------------------------------------------------------------- host side

// Optional per-kernel timing (dmic_profile_*): hipEvents recorded on the
// launch stream around meshft_fwd_kernel and meshft_bwd_kernel only.
struct KernelTimer {
  bool enabled = false;
  hipEvent_t ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  double total_ms[2] = {0.0, 0.0};
  long count[2] = {0, 0};
  bool pending[2] = {false, false};
};
static KernelTimer g_timer;

static int timer_begin(int k, hipStream_t st) {
  if (!g_timer.enabled) return DMIC_OK;
  if (!g_timer.ev[k][0]) {
    DMIC_HIP_CHECK(hipEventCreate(&g_timer.ev[k][0]));
    DMIC_HIP_CHECK(hipEventCreate(&g_timer.ev[k][1]));
  }
  if (g_timer.pending[k]) {  // fold the previous launch before reusing the events
    DMIC_HIP_CHECK(hipEventSynchronize(g_timer.ev[k][1]));
    float ms = 0.0f;
    DMIC_HIP_CHECK(hipEventElapsedTime(&ms, g_timer.ev[k][0], g_timer.ev[k][1]));
    g_timer.total_ms[k] += ms;
    g_timer.count[k] += 1;
    g_timer.pending[k] = false;
  }
  DMIC_HIP_CHECK(hipEventRecord(g_timer.ev[k][0], st));
  return DMIC_OK;
}

static int timer_end(int k, hipStream_t st) {
  if (!g_timer.enabled) return DMIC_OK;
  DMIC_HIP_CHECK(hipEventRecord(g_timer.ev[k][1], st));
  g_timer.pending[k] = true;
  return DMIC_OK;
}


struct FwdPlan {
  int64_t faces_per_split;
  int nsplit;
};

static FwdPlan fwd_plan(int64_t nf, int64_t n) {
  FwdPlan p{nf > 0 ? nf : 1, 1};
  const int64_t fblocks = (n + 255) / 256;
  if (fblocks <= 0 || nf <= 0) return p;
  int64_t want = (kTargetBlocks + fblocks - 1) / fblocks;
  const int64_t maxsplit = (nf + kFold - 1) / kFold;
  if (want > maxsplit) want = maxsplit;
  if (want < 1) want = 1;
  int64_t per = (nf + want - 1) / want;
  per = (per + kFold - 1) / kFold * kFold;
  p.faces_per_split = per;
  p.nsplit = (int)((nf + per - 1) / per);
  return p;
}

struct BwdPlan {
  int64_t freqs_per_chunk;
  int nchunk;
};

static BwdPlan bwd_plan(int64_t nf, int64_t n) {
  BwdPlan p{n > 0 ? n : 1, 1};
  const int64_t tblocks = (nf + 255) / 256;
  if (tblocks <= 0 || n <= 0) return p;
  int64_t want = (kTargetBlocks + tblocks - 1) / tblocks;
  const int64_t maxc = (n + kFold - 1) / kFold;
  if (want > maxc) want = maxc;
  if (want < 1) want = 1;
  int64_t per = (n + want - 1) / want;
  per = (per + kFold - 1) / kFold * kFold;
  p.freqs_per_chunk = per;
  p.nchunk = (int)((n + per - 1) / per);
  return p;
}

struct FwdWs {
  int* flag;
  double* cA;
  double* S;
  FaceRec* recs;
  double2* part;
};

static FwdWs carve_fwd(Carver& c, int64_t nf, int64_t n) {
  FwdWs w;
  w.flag = c.take<int>(4);
  w.cA = c.take<double>(nf > 0 ? nf : 1);
  w.S = c.take<double>(2);
  w.recs = c.take<FaceRec>(nf > 0 ? nf : 1);
  const FwdPlan p = fwd_plan(nf, n);
  w.part = (p.nsplit > 1) ? c.take<double2>((size_t)p.nsplit * n) : nullptr;
  return w;
}

struct BwdWs {
  int* flag;
  double* cA;
  double* S;
  FaceRec* recs;
  double* part;
  double* G;
  double* cgA;
  double* wR;
  double* K;
  int* cnt;
  int* fillc;
  int* row;
  int* ent;
};

static BwdWs carve_bwd(Carver& c, int64_t nv, int64_t nf, int64_t n) {
  BwdWs w;
  const int64_t f1 = nf > 0 ? nf : 1, v1 = nv > 0 ? nv : 1;
  w.flag = c.take<int>(4);
  w.cA = c.take<double>(f1);
  w.S = c.take<double>(2);
  w.recs = c.take<FaceRec>(f1);
  const BwdPlan p = bwd_plan(nf, n);
  w.part = c.take<double>((size_t)p.nchunk * 10 * f1);
  w.G = c.take<double>(9 * f1);
  w.cgA = c.take<double>(9 * f1);
  w.wR = c.take<double>(f1);
  w.K = c.take<double>(2);
  w.cnt = c.take<int>(v1);
  w.fillc = c.take<int>(v1);
  w.row = c.take<int>(v1 + 1);
  w.ent = c.take<int>(3 * f1);
  return w;
}

template <typename IdxT>
static int prep_faces(const float* V, int64_t nv, const IdxT* Fc, int64_t nf, const float* coeff,
                      int* flag, double* cA, double* S, FaceRec* recs, hipStream_t st) {
  hipLaunchKernelGGL(face_area_sum_kernel<IdxT>, dim3(1), dim3(1024), 0, st, V, nv, Fc, nf, coeff,
                     cA, S, flag);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(face_pack_kernel<IdxT>, dim3(ceil_div(nf, 256)), dim3(256), 0, st, V, nv, Fc,
                     nf, cA, S, recs, flag);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

template <typename IdxT>
static int forward_impl(const float* V, int64_t nv, const IdxT* Fc, int64_t nf, const float* coeff,
                        const float* xi, int64_t n, float* out, void* ws, size_t ws_bytes,
                        hipStream_t st) {
  Carver c(ws);
  FwdWs w = carve_fwd(c, nf, n);
  DMIC_REQUIRE(ws_bytes >= c.bytes(), DMIC_ERR_WORKSPACE,
               "forward workspace too small: %zu < %zu bytes", ws_bytes, c.bytes());
  hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, st, w.flag, 4, 0);
  DMIC_LAUNCH_CHECK();
  if (n == 0) return DMIC_OK;
  if (nf == 0) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(2 * n, 256)), dim3(256), 0, st,
                       reinterpret_cast<int*>(out), 2 * n, 0);
    DMIC_LAUNCH_CHECK();
    return DMIC_OK;
  }
  int rc = prep_faces<IdxT>(V, nv, Fc, nf, coeff, w.flag, w.cA, w.S, w.recs, st);
  if (rc != DMIC_OK) return rc;
  const FwdPlan p = fwd_plan(nf, n);
  dim3 grid(ceil_div(n, 256), p.nsplit);
  rc = timer_begin(0, st);
  if (rc != DMIC_OK) return rc;
  hipLaunchKernelGGL(meshft_fwd_kernel, grid, dim3(256), 0, st, w.recs, nf, p.faces_per_split, xi,
                     n, reinterpret_cast<float2*>(out), p.nsplit > 1 ? w.part : nullptr);
  DMIC_LAUNCH_CHECK();
  rc = timer_end(0, st);
  if (rc != DMIC_OK) return rc;
  if (p.nsplit > 1) {
    hipLaunchKernelGGL(fwd_split_reduce_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, st, w.part,
                       p.nsplit, n, reinterpret_cast<float2*>(out));
    DMIC_LAUNCH_CHECK();
  }
  return DMIC_OK;
}

template <typename IdxT>
static int backward_impl(const float* V, int64_t nv, const IdxT* Fc, int64_t nf,
                         const float* coeff, const float* grad, const float* xi, int64_t n,
                         float* gverts, void* ws, size_t ws_bytes, hipStream_t st) {
  Carver c(ws);
  BwdWs w = carve_bwd(c, nv, nf, n);
  DMIC_REQUIRE(ws_bytes >= c.bytes(), DMIC_ERR_WORKSPACE,
               "backward workspace too small: %zu < %zu bytes", ws_bytes, c.bytes());
  hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, st, w.flag, 4, 0);
  DMIC_LAUNCH_CHECK();
  if (nv == 0) return DMIC_OK;
  if (nf == 0) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(3 * nv, 256)), dim3(256), 0, st,
                       reinterpret_cast<int*>(gverts), 3 * nv, 0);
    DMIC_LAUNCH_CHECK();
    return DMIC_OK;
  }
  int rc = prep_faces<IdxT>(V, nv, Fc, nf, coeff, w.flag, w.cA, w.S, w.recs, st);
  if (rc != DMIC_OK) return rc;
  const BwdPlan p = bwd_plan(nf, n);
  if (n > 0) {
    dim3 grid(ceil_div(nf, 256), p.nchunk);
    rc = timer_begin(1, st);
    if (rc != DMIC_OK) return rc;
    hipLaunchKernelGGL(meshft_bwd_kernel, grid, dim3(256), 0, st, w.recs, nf, xi, grad, n,
                       p.freqs_per_chunk, w.part);
    DMIC_LAUNCH_CHECK();
    rc = timer_end(1, st);
    if (rc != DMIC_OK) return rc;
  } else {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(20 * nf, 256)), dim3(256), 0, st,
                       reinterpret_cast<int*>(w.part), 20 * nf, 0);
    DMIC_LAUNCH_CHECK();
  }
  const int nchunk = (n > 0) ? p.nchunk : 1;
  hipLaunchKernelGGL(bwd_face_finalize_kernel<IdxT>, dim3(ceil_div(nf, 256)), dim3(256), 0, st, V,
                     nv, Fc, nf, coeff, w.cA, w.S, w.part, nchunk, w.G, w.cgA, w.wR, w.flag);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(1024), 0, st, w.wR, nf, w.K);
  DMIC_LAUNCH_CHECK();
  // CSR face -> vertex (rows sorted by incidence id 3t+j)
  hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(nv, 256)), dim3(256), 0, st, w.cnt, nv, 0);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(nv, 256)), dim3(256), 0, st, w.fillc, nv, 0);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(csr_count_kernel<IdxT>, dim3(ceil_div(3 * nf, 256)), dim3(256), 0, st, Fc, nf,
                     nv, w.cnt);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(scan_i32_kernel, dim3(1), dim3(1024), 0, st, w.cnt, nv, w.row);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(csr_fill_kernel<IdxT>, dim3(ceil_div(3 * nf, 256)), dim3(256), 0, st, Fc, nf,
                     nv, w.row, w.fillc, w.ent);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(csr_sort_rows_kernel, dim3(ceil_div(nv, 256)), dim3(256), 0, st, w.row, nv,
                     w.ent);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(vertex_gather_kernel, dim3(ceil_div(nv, 256)), dim3(256), 0, st, w.row, w.ent,
                     nv, w.G, w.cgA, w.K, w.S, gverts);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

}  // namespace dmic

using namespace dmic;

extern "C" {

int dmic_version(void) { return 100; }

const char* dmic_last_error(void) { return g_err; }

size_t dmic_meshft_forward_workspace(int64_t nv, int64_t nf, int64_t n) {
  (void)nv;
  if (nf < 0 || n < 0) return 0;
  Carver c(nullptr);
  carve_fwd(c, nf, n);
  return c.bytes();
}

size_t dmic_meshft_backward_workspace(int64_t nv, int64_t nf, int64_t n) {
  if (nv < 0 || nf < 0 || n < 0) return 0;
  Carver c(nullptr);
  carve_bwd(c, nv, nf, n);
  return c.bytes();
}

int dmic_meshft_forward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                        int64_t nf, const float* faces_coeff, const float* xi_list, int64_t n,
                        float* ftmesh_list, void* workspace, size_t workspace_bytes,
                        hipStream_t stream) {
  g_err[0] = 0;
  DMIC_REQUIRE(nv >= 0 && nf >= 0 && n >= 0, DMIC_ERR_ARG, "negative size");
  DMIC_REQUIRE(nf < (int64_t(1) << 31) / 3 && n < (int64_t(1) << 40), DMIC_ERR_ARG, "size too large");
  DMIC_REQUIRE(n == 0 || (xi_list && ftmesh_list), DMIC_ERR_ARG, "null xi_list/ftmesh_list");
  DMIC_REQUIRE(nf == 0 || n == 0 || (verts && faces && faces_coeff), DMIC_ERR_ARG,
               "null mesh pointer");
  DMIC_REQUIRE(workspace || workspace_bytes == 0, DMIC_ERR_ARG, "null workspace");
  if (faces_is_int64)
    return forward_impl<int64_t>(verts, nv, static_cast<const int64_t*>(faces), nf, faces_coeff,
                                 xi_list, n, ftmesh_list, workspace, workspace_bytes, stream);
  return forward_impl<int32_t>(verts, nv, static_cast<const int32_t*>(faces), nf, faces_coeff,
                               xi_list, n, ftmesh_list, workspace, workspace_bytes, stream);
}

int dmic_meshft_backward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                         int64_t nf, const float* faces_coeff, const float* grad_output_list,
                         const float* ftmesh_list, const float* xi_list, int64_t n,
                         float* grad_verts, void* workspace, size_t workspace_bytes,
                         hipStream_t stream) {
  (void)ftmesh_list;
  g_err[0] = 0;
  DMIC_REQUIRE(nv >= 0 && nf >= 0 && n >= 0, DMIC_ERR_ARG, "negative size");
  DMIC_REQUIRE(nf < (int64_t(1) << 31) / 3 && nv < (int64_t(1) << 31) && n < (int64_t(1) << 40),
               DMIC_ERR_ARG, "size too large");
  DMIC_REQUIRE(nv == 0 || grad_verts, DMIC_ERR_ARG, "null grad_verts");
  DMIC_REQUIRE(nf == 0 || (verts && faces && faces_coeff), DMIC_ERR_ARG, "null mesh pointer");
  DMIC_REQUIRE(n == 0 || nf == 0 || (xi_list && grad_output_list), DMIC_ERR_ARG,
               "null xi_list/grad_output_list");
  DMIC_REQUIRE(workspace || workspace_bytes == 0, DMIC_ERR_ARG, "null workspace");
  if (faces_is_int64)
    return backward_impl<int64_t>(verts, nv, static_cast<const int64_t*>(faces), nf, faces_coeff,
                                  grad_output_list, xi_list, n, grad_verts, workspace,
                                  workspace_bytes, stream);
  return backward_impl<int32_t>(verts, nv, static_cast<const int32_t*>(faces), nf, faces_coeff,
                                grad_output_list, xi_list, n, grad_verts, workspace,
                                workspace_bytes, stream);
}

int dmic_profile_enable(int enable) {
  g_timer.enabled = enable != 0;
  g_timer.total_ms[0] = g_timer.total_ms[1] = 0.0;
  g_timer.count[0] = g_timer.count[1] = 0;
  g_timer.pending[0] = g_timer.pending[1] = false;
  return DMIC_OK;
}

int dmic_profile_read(double* fwd_ms, long* fwd_count, double* bwd_ms, long* bwd_count) {
  for (int k = 0; k < 2; ++k) {
    if (g_timer.pending[k]) {
      DMIC_HIP_CHECK(hipEventSynchronize(g_timer.ev[k][1]));
      float ms = 0.0f;
      DMIC_HIP_CHECK(hipEventElapsedTime(&ms, g_timer.ev[k][0], g_timer.ev[k][1]));
      g_timer.total_ms[k] += ms;
      g_timer.count[k] += 1;
      g_timer.pending[k] = false;
    }
  }
  if (fwd_ms) *fwd_ms = g_timer.total_ms[0];
  if (fwd_count) *fwd_count = g_timer.count[0];
  if (bwd_ms) *bwd_ms = g_timer.total_ms[1];
  if (bwd_count) *bwd_count = g_timer.count[1];
  return DMIC_OK;
}

int dmic_check_index_flag(void* workspace, hipStream_t stream) {
  g_err[0] = 0;
  DMIC_REQUIRE(workspace, DMIC_ERR_ARG, "null workspace");
  int flag = 0;
  // the flag is the first region of both forward and backward workspaces
  DMIC_HIP_CHECK(hipMemcpyAsync(&flag, workspace, sizeof(int), hipMemcpyDeviceToHost, stream));
  DMIC_HIP_CHECK(hipStreamSynchronize(stream));
  if (flag) {
    set_error("face index outside [0, nv)");
    return DMIC_ERR_INDEX;
  }
  return DMIC_OK;
}

}  // extern "C"
