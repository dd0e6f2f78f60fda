// meshFT forward and adjoint kernels for gfx950, and their C-ABI launchers.
//
// Forward (replaces Mesh_NUFT_kernel, deltamic/cpp/fourier3d_cuda_kernel.cu:282-311):
//   one lane per face; a one-wave workgroup takes 8 consecutive band
//   frequencies (records in LDS, read as broadcasts) and walks every face in
//   blocks of 64 packed 64-byte records read from L2.  Each lane sums w_t F_t
//   per frequency in FP32; every 64 face blocks the sums are added over the
//   lanes by a fixed butterfly and folded into FP64.  No partials, no second
//   kernel, deterministic.
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
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>

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
  float w;       // -2 x the face weight 2 c_t A_t / sum c A (the forward's -2 folded in)
  float an[3];   // autofluorescence variant: -2 (regint / regvol) (v_b - v_a) x (v_c - v_a); else 0
};
static_assert(sizeof(FaceRec) == 64, "FaceRec layout");

// Per-frequency record of the band tiles (48 bytes).  Every pair of the
// adjoint reads the first 20 bytes (a 16-byte and a 4-byte LDS broadcast),
// every pair of the forward the first 12; x0, x1, X0 and X1 are read only
// where a band row starts (the phases' xi_0 and xi_1 terms are constant
// along a row).
struct __align__(16) FreqRec {
  double X2;       // xi_2 in FP64
  float x2;        // xi_2 in FP32
  float gph;       // adjoint: arg(g) / (2 pi) in [-1, 0) revolutions, of the
  float gm;        //   upstream gradient g = dL/dRe + i dL/dIm, and 2 |g|
  float x0, x1;    // xi_0, xi_1 in FP32
  float inv2;      // autofluorescence variant: 1 / |xi|^2 (0 at xi = 0); else 0
  double X0, X1;   // xi_0, xi_1 in FP64
};
static_assert(sizeof(FreqRec) == 48, "FreqRec layout");
// The first 16 bytes of a FreqRec, read as one LDS broadcast.
struct __align__(16) FreqHead {
  double X2;
  float x2, gph;
};


// -------------------------------------------------------------------- prep

// Vertices are float or double (T); the corners come back in FP64 (exact).
template <typename IdxT, typename T>
__device__ __forceinline__ void load_face_verts(const T* __restrict__ V, int64_t nv,
                                                const IdxT* __restrict__ Fc, int64_t t,
                                                double (&P)[3][3], int* flag) {
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int64_t i = (int64_t)Fc[3 * t + j];
    if (i >= 0 && i < nv) {
      P[j][0] = V[3 * i]; P[j][1] = V[3 * i + 1]; P[j][2] = V[3 * i + 2];
    } else {
      P[j][0] = P[j][1] = P[j][2] = 0.0;
      flag[0] = 1;
    }
  }
}

// c_t A_t (FP64, cross-product norm as .cu:9-12,38-40) per face and, per
// workgroup, the partial sums of c_t A_t and of a . (b x c) (for
// regvol = -sum_t a . (b x c) / 6, compute_volume_manifold, render.py:221-226;
// used by the autofluorescence variant).  face_sum_final_kernel adds the
// partials in a fixed tree: S[0] = sum_t c_t A_t, S[1] = regvol.
template <typename IdxT, typename T>
__global__ __launch_bounds__(256) void face_area_sum_kernel(
    const T* __restrict__ V, int64_t nv, const IdxT* __restrict__ Fc, int64_t nf,
    const T* __restrict__ coeff, double* __restrict__ cA, double* __restrict__ part,
    int* flag) {
  __shared__ double red[256];
  __shared__ double redv[256];
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double s = 0.0, vs = 0.0;
  if (t < nf) {
    double P[3][3];
    load_face_verts(V, nv, Fc, t, P, flag);
    double u[3], q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      u[d] = P[1][d] - P[0][d];
      q[d] = P[2][d] - P[0][d];
    }
    const double n0 = u[1] * q[2] - u[2] * q[1];
    const double n1 = u[2] * q[0] - u[0] * q[2];
    const double n2 = u[0] * q[1] - u[1] * q[0];
    const double ca = (double)coeff[t] * 0.5 * sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    cA[t] = ca;
    s = ca;
    vs = P[0][0] * (P[1][1] * P[2][2] - P[1][2] * P[2][1]) +
         P[0][1] * (P[1][2] * P[2][0] - P[1][0] * P[2][2]) +
         P[0][2] * (P[1][0] * P[2][1] - P[1][1] * P[2][0]);
  }
  red[threadIdx.x] = s;
  redv[threadIdx.x] = vs;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
      redv[threadIdx.x] += redv[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = red[0];
    part[2 * blockIdx.x + 1] = redv[0];
  }
}

__global__ __launch_bounds__(1024) void face_sum_final_kernel(const double* __restrict__ part,
                                                              int64_t nblocks,
                                                              double* __restrict__ S_out) {
  __shared__ double red[1024];
  __shared__ double redv[1024];
  double s = 0.0, vs = 0.0;
  for (int64_t b = threadIdx.x; b < nblocks; b += 1024) {
    s += part[2 * b];
    vs += part[2 * b + 1];
  }
  red[threadIdx.x] = s;
  redv[threadIdx.x] = vs;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
      redv[threadIdx.x] += redv[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    S_out[0] = red[0];
    S_out[1] = -redv[0] / 6.0;
  }
}

// Squared bound on a face's phase gap per unit |xi|:
// D = max(|xi.eb|, |xi.ec|, |xi.(ec - eb)|) <= |xi| max(|eb|, |ec|, |ec - eb|).
__device__ __forceinline__ float face_extent2(float b0, float b1, float b2, float c0, float c1,
                                              float c2) {
  const float d0 = c0 - b0, d1 = c1 - b1, d2 = c2 - b2;
  const float nb = b0 * b0 + b1 * b1 + b2 * b2;
  const float nc = c0 * c0 + c1 * c1 + c2 * c2;
  const float nd = d0 * d0 + d1 * d1 + d2 * d2;
  return __builtin_fmaxf(nb, __builtin_fmaxf(nc, nd));
}

// Packed face records are padded to whole blocks of 64 faces.
__host__ __device__ inline int64_t face_blocks64(int64_t nf) { return (nf + 63) / 64; }

template <typename IdxT, typename T>
__global__ __launch_bounds__(256) void face_pack_kernel(
    const T* __restrict__ V, int64_t nv, const IdxT* __restrict__ Fc, int64_t nf,
    const double* __restrict__ cA, const double* __restrict__ S_in, float regint, double wscale,
    FaceRec* __restrict__ out, int* flag) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= face_blocks64(nf) * 64) return;
  FaceRec r;
  if (t >= nf) {
    // padding of the last 64-face block: zero weight, all phases 0 (a finite
    // Taylor-branch pair that adds nothing)
    r.a[0] = r.a[1] = r.a[2] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) r.eb[d] = r.ec[d] = r.an[d] = 0.0f;
    r.w = 0.0f;
    out[t] = r;
    return;
  }
  double P[3][3];
  load_face_verts(V, nv, Fc, t, P, flag);
  const double S = S_in[0];
  const double regvol = S_in[1];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r.a[d] = P[0][d] * kInv2Pi;
    r.eb[d] = (float)(P[1][d] - P[0][d]);  // exact in FP64, rounded once
    r.ec[d] = (float)(P[2][d] - P[0][d]);
  }
  r.w = (S > 0.0) ? (float)(-4.0 * wscale * cA[t] / S) : 0.0f;
  // autofluorescence: nor2_t = nor_t regint / regvol, nor_t = a x b + b x c + c x a
  // = (b - a) x (c - a) (render.py:208,232)
  // the forward's factor -2 is folded in here too, as in r.w
  const double sc =
      (regint != 0.0f && regvol != 0.0) ? -2.0 * wscale * regint / regvol : 0.0;
  double u[3], q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u[d] = (double)P[1][d] - (double)P[0][d];
    q[d] = (double)P[2][d] - (double)P[0][d];
  }
  r.an[0] = (float)(sc * (u[1] * q[2] - u[2] * q[1]));
  r.an[1] = (float)(sc * (u[2] * q[0] - u[0] * q[2]));
  r.an[2] = (float)(sc * (u[0] * q[1] - u[1] * q[0]));
  out[t] = r;
  // largest squared face extent of the mesh (flag[1]) and largest |a / 2 pi|^2
  // (flag[2], rounded up), zeroed with the index flag; non-negative floats
  // order like their bit patterns
  atomicMax(flag + 1, __float_as_int(face_extent2(r.eb[0], r.eb[1], r.eb[2], r.ec[0], r.ec[1],
                                                  r.ec[2])));
  atomicMax(flag + 2, __float_as_int(__double2float_ru(r.a[0] * r.a[0] + r.a[1] * r.a[1] +
                                                       r.a[2] * r.a[2])));
}

// -------------------------------------------------------------------- forward

// Maximum of v over the 64 lanes of the wave, returned to every lane.
__device__ __forceinline__ float wave_max64(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = __builtin_fmaxf(v, __shfl_xor(v, o));
  return v;
}

// Orders this wave's LDS writes before its reads of other lanes' records (and
// the reads of one tile before the writes of the next): LDS operations of a
// wave complete in order, so only the compiler has to be kept from moving
// them across.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A pair can exceed kWidePhase only if |xi|^2 times the face's squared extent
// does.  Where (max |xi| of the wave's frequencies)^2 times (max extent over
// the mesh, from face_pack_kernel)^2 cannot, the per-pair wave check is
// skipped (may_be_wide = false, a wave-uniform branch); elsewhere every pair
// takes it (wave_is_wide).
constexpr float kWidePhase2 = kWidePhase * kWidePhase;

// Forward: one wave per workgroup, on kFwdFreqs<kAF> consecutive band frequencies
// (their records in LDS, read as broadcasts) and every face, one face per
// lane: face block b is faces [64 b, 64 b + 64).  Each lane sums w_t F_t of
// its faces per frequency in FP32 registers; every kFwdFold face blocks and at
// the end each sum is added over the 64 lanes (a fixed butterfly) and folded
// into an FP64 total.  No LDS beyond the frequency records, so the kernel runs
// at 7 waves per SIMD (with FP64 sums kept per lane in LDS it ran at 5, and 6%
// slower).  Consecutive band entries mostly lie on one band row (equal xi_0,
// xi_1), so the row terms of the vertex phase (FP64) and of the edge phases
// are formed once per row and face.
// 8 frequencies per wave; 4 for the autofluorescence variant, whose four sums
// per frequency would otherwise double the registers per wave
template <bool kAF>
constexpr int kFwdFreqs = kAF ? 4 : 8;
constexpr int kFwdFold = 64;

// Per-frequency record of the forward (48 bytes).  Every pair reads the
// first 12 bytes (the autofluorescence variant 16); the rest is read where a
// band row starts.  The edge phases are formed in half revolutions from
// xi / (4 pi) (face_H_halfrev); the unscaled xi serves the wide-phase
// variant, which forms them in radians as before.
struct __align__(16) FwdFreqRec {
  double X2;       // xi_2 in FP64
  float h2;        // xi_2 / (4 pi) in FP32 (from FP64)
  float x2;        // xi_2 in FP32
  float h0, h1;    // xi_0 / (4 pi), xi_1 / (4 pi)
  float x0, x1;    // xi_0, xi_1 in FP32
  double X0, X1;   // xi_0, xi_1 in FP64
};
static_assert(sizeof(FwdFreqRec) == 48, "FwdFreqRec layout");

template <bool kAF>
struct FwdLds {
  static constexpr int kV = kAF ? 4 : 2;  // sums per frequency: F (re, im) [+ (xi . an) F]
  FwdFreqRec rec[kFwdFreqs<kAF>];
};

// kAF: autofluorescence variant (fourier_from_mesh_narrowband_autofluorescence,
// render.py:228-318): each face weighs w_t + i (xi . an_t) / |xi|^2, so the
// lane also sums (xi . an_t) F_t and i/|xi|^2 times it is added at the end;
// at xi = 0 that term is its limit -regint (the solid's transform at 0 over
// regvol; the reference divides by zero there).
// Work split (FwdPlan): the first `head` workgroups take kF band entries each
// (whole rounds of the device's resident waves); the rest of the band, less
// than one round's worth, is spread evenly over the `tail` last workgroups
// (at most kF entries each), so the launch does not end on a partly filled
// round of full-length waves.
template <bool kAF, typename T>
__global__ __launch_bounds__(64) void meshft_fwd_kernel(
    const FaceRec* __restrict__ faces, int64_t nf, const T* __restrict__ xi, int64_t n,
    int64_t head, int64_t tail, T* __restrict__ out /* [n][2] */,
    const double* __restrict__ S_in, const int* __restrict__ e2bits, float regint) {
  constexpr int kV = FwdLds<kAF>::kV;
  constexpr int kF = kFwdFreqs<kAF>;
  __shared__ FwdLds<kAF> sm;
  const int lane = (int)threadIdx.x;
  const int64_t g = blockIdx.x;
  int64_t m0;
  int cntf;
  if (g < head) {
    m0 = g * kF;
    cntf = kF;
  } else {
    const int64_t base = head * kF, rem = n - base, j = g - head;
    m0 = base + rem * j / tail;
    cntf = (int)(base + rem * (j + 1) / tail - m0);
  }
  bool same_row = false;
  float xx = 0.0f;
  if (lane < cntf) {
    const int64_t m = m0 + lane;
    const T* xm = xi + 3 * m;
    FwdFreqRec fr;
    fr.X0 = xm[0]; fr.X1 = xm[1]; fr.X2 = xm[2];
    fr.x0 = (float)fr.X0; fr.x1 = (float)fr.X1; fr.x2 = (float)fr.X2;
    fr.h0 = (float)(fr.X0 * (0.5 * kInv2Pi));
    fr.h1 = (float)(fr.X1 * (0.5 * kInv2Pi));
    fr.h2 = (float)(fr.X2 * (0.5 * kInv2Pi));
    sm.rec[lane] = fr;
    xx = fr.x0 * fr.x0 + fr.x1 * fr.x1 + fr.x2 * fr.x2;
    same_row = lane > 0 && xm[-3] == xm[0] && xm[-2] == xm[1];
  }
  const uint64_t cont = __builtin_amdgcn_ballot_w64(same_row);
  const float xxmax = wave_max64(xx);
  const bool may_be_wide = xxmax * __int_as_float(e2bits[0]) > kWidePhase2;
  // vertex phases below 2^18 revolutions: reduced without FP64 fract/convert
  const bool small_rev = xxmax * __int_as_float(e2bits[1]) < kRevMagicLimit2;
  wave_lds_sync();
  // lane v * kF + k keeps the FP64 total of sum v of frequency k
  double tot = 0.0;
  float acc[kV][kF];
#pragma unroll
  for (int v = 0; v < kV; ++v)
#pragma unroll
    for (int k = 0; k < kF; ++k) acc[v][k] = 0.0f;
  const int64_t nblk = (nf + 63) / 64;
  // the face loop in two forms, chosen once per workgroup: kMagic reduces
  // the vertex phase through rev_frac_scaled (the row term carries
  // kRevMagic), else through the FP64 fract
  auto face_loop = [&](auto magic) {
    constexpr bool kMagic = decltype(magic)::value;
    for (int64_t b = 0; b < nblk; ++b) {
      // the frequency records are re-read from LDS per face block: without this
      // compiler barrier they would be hoisted into ~80 registers for the launch
      asm volatile("" ::: "memory");
      // the records are padded to whole blocks with zero-weight faces; a
      // uniform block address plus the lane's offset needs no vector address
      // arithmetic
      const FaceRec* blk = faces + b * 64;
      const FaceRec r = blk[lane];
      double row = 0.0;
      float prow = 0.0f, qrow = 0.0f, nrow = 0.0f;
#pragma unroll
      for (int k = 0; k < kF; ++k) {
        if (k < cntf) {  // wave-uniform
          const FwdFreqRec& fr = sm.rec[k];
          if (!((cont >> k) & 1)) {  // a band row starts (wave-uniform)
            row = fma(r.a[1], fr.X1, fma(r.a[0], fr.X0, kMagic ? kRevMagic : 0.0));
            prow = __builtin_fmaf(r.eb[1], fr.h1, r.eb[0] * fr.h0);
            qrow = __builtin_fmaf(r.ec[1], fr.h1, r.ec[0] * fr.h0);
            if (kAF) nrow = __builtin_fmaf(r.an[1], fr.x1, r.an[0] * fr.x0);
          }
          // the same operations, in the same order, as a full evaluation;
          // edge phases in half revolutions
          const double rev_a = fma(r.a[2], fr.X2, row);  // + kRevMagic if kMagic
          const float p = __builtin_fmaf(r.eb[2], fr.h2, prow);
          const float q = __builtin_fmaf(r.ec[2], fr.h2, qrow);
          const Sorted so = sort3(p, q);
          cf F;  // F = e_mid N, face value = -2 m F / c^2 (r.w carries -2 c^2)
          float m;
          if (!may_be_wide || !wave_is_wide(so.D * k4PiF)) {
            // median phase: 2 mid' revolutions from vertex a
            const float ua = kMagic ? rev_frac_scaled(rev_a) * kTwoM32 : reduce_rev(rev_a);
            const float u = __builtin_fmaf(so.mid, 2.0f, ua);
            F = cmul(expmi_rev(u), face_H_halfrev(so, m));
          } else {
            // wide variant: edge phases in radians from the unscaled xi (the
            // products y c of the reduction are then exact fmas), m / c^2
            const float pr = __builtin_fmaf(r.eb[2], fr.x2,
                                            __builtin_fmaf(r.eb[1], fr.x1, r.eb[0] * fr.x0));
            const float qr = __builtin_fmaf(r.ec[2], fr.x2,
                                            __builtin_fmaf(r.ec[1], fr.x1, r.ec[0] * fr.x0));
            const Sorted sr = sort3(pr, qr);
            // (the offset removed exactly: both lie in [2^20, 2^21))
            const double ra = kMagic ? rev_a - kRevMagic : rev_a;
            F = cmul(expmi_rev(mid_rev<true>(ra, sr.mid)), face_H<true>(sr, m));
            m *= kFwdInvWScale;
          }
          const float wm = r.w * m;
          acc[0][k] = __builtin_fmaf(wm, F.re, acc[0][k]);
          acc[1][k] = __builtin_fmaf(wm, F.im, acc[1][k]);
          if (kAF) {
            const float tm = __builtin_fmaf(r.an[2], fr.x2, nrow) * m;
            acc[kV - 2][k] = __builtin_fmaf(tm, F.re, acc[kV - 2][k]);
            acc[kV - 1][k] = __builtin_fmaf(tm, F.im, acc[kV - 1][k]);
          }
        }
      }
      if ((b + 1) % kFwdFold == 0 || b + 1 == nblk) {
        // fold: each FP32 sum is added over the 64 lanes (butterfly, fixed
        // order) and into its owner lane's FP64 total
#pragma unroll
        for (int v = 0; v < kV; ++v)
#pragma unroll
          for (int k = 0; k < kF; ++k) {
            float x = acc[v][k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
            if (lane == v * kF + k) tot += (double)x;
            acc[v][k] = 0.0f;
          }
      }
    }
  };
  // (the autofluorescence variant keeps the FP64 form only: both forms
  // together would cost it a wave per SIMD)
  if constexpr (!kAF) {
    if (small_rev)
      face_loop(std::true_type{});
    else
      face_loop(std::false_type{});
  } else {
    face_loop(std::false_type{});
  }
  // lane k gathers the kV totals of frequency k
  double fin[kV];
#pragma unroll
  for (int v = 0; v < kV; ++v) fin[v] = __shfl(tot, v * kF + (lane & (kF - 1)));
  if (lane >= cntf) return;
  const int64_t m = m0 + lane;
  double accr = fin[0], acci = fin[1];
  if (kAF) {
    const FwdFreqRec& fr = sm.rec[lane];
    const double q = fr.X0 * fr.X0 + fr.X1 * fr.X1 + fr.X2 * fr.X2;
    if (q > 0.0) {
      accr -= fin[kV - 1] / q;
      acci += fin[kV - 2] / q;
    } else if (S_in[1] != 0.0) {
      accr -= (double)regint;
    }
  }
  out[2 * m] = (T)accr;
  out[2 * m + 1] = (T)acci;
}

// -------------------------------------------------------------------- adjoint

// Per-face sums of the adjoint: 10, or 13 for the autofluorescence variant,
// whose face weight is W_t = w_t + i s / |xi|^2 with s = xi . an_t:
//   P_a, P_b, T use conj(g) W_t (the face weight inside the sum),
//   R = sum Re(conj g F) (area term, as in the plain transform),
//   N = sum Re(conj g i F / |xi|^2) xi (derivative w.r.t. an_t).
template <bool kAF>
constexpr int kFaceSums = kAF ? 13 : 10;
// Type of the chunk partials of those sums: each is rounded once to FP32
// (below the rounding the FP32 tile sums already carry) and the chunks are
// added in FP64.  The autofluorescence variant keeps FP64 partials: with the
// conversion its adjoint spills registers inside the frequency loop.
template <bool kAF>
using PartT = typename std::conditional<kAF, double, float>::type;

// Adjoint workgroup: kBwdWaves waves on the same 64 faces (one per lane),
// each wave on its own share of the workgroup's frequency chunk with its
// own LDS tile of 64 frequency records; the waves' sums are added in LDS in
// a fixed order at the end, so one partial per face and chunk is written.
constexpr int kBwdWaves = 4;
constexpr int kBwdFaces = 64;

// The row-start data of a workgroup's 64 faces, kept in LDS: read only where
// a band row starts, so those values need no registers across the pair loop.
struct __align__(16) FaceRow {
  double A0, A1;               // vertex a / (2 pi), first two coordinates
  float b0, b1, c0, c1;        // edge vectors, first two coordinates
};
static_assert(sizeof(FaceRow) == 32, "FaceRow layout");

// This lane's face in the adjoint (zeros for lanes past the last face).
struct FaceLane {
  double A0, A1, A2;            // vertex a / (2 pi)
  float b0, b1, b2, c0, c1, c2;  // edge vectors v_b - v_a, v_c - v_a
  float fw;                      // face weight
  float n0, n1, n2;              // autofluorescence normal term
};

// One LDS tile of the adjoint: this lane's face against cnt frequencies,
// adding to the face's K sums.  Bit k of `cont` is set when record k lies on
// the same band row as record k - 1 (same xi_0 and xi_1), so the row terms of
// the vertex phase (FP64 A0 X0 + A1 X1) and of the edge phases (b0 x0 + b1 x1,
// c0 x0 + c1 x1) are reused; bit 0 is never set.  Along a row the xi_0 and
// xi_1 parts of the vertex sums sum_m (.) xi_m are xi_0 and xi_1 times the
// row's sum of (.): those are kept per row and added when the row ends
// (not in the autofluorescence variant, whose extra registers would cost
// it a wave per SIMD).
// may_be_wide: some pair of the tile may have a phase gap above kWidePhase
// (then every pair takes the wave-wide check of wave_is_wide).
// The FP32 sums of a tile (at most 64 frequencies) are folded into the FP64
// sums at its end.
// kMagic: the vertex phase is reduced through rev_frac_scaled (the row term
// carries kRevMagic; the caller has checked |a/2pi| |xi| < 2^18 for the tile).
template <bool kAF, bool kMagic>
__device__ __forceinline__ void bwd_tile(const FreqRec* tile, int cnt, uint64_t cont,
                                         bool may_be_wide, const FaceLane& L,
                                         const FaceRow& frow,
                                         double (&acc)[kFaceSums<kAF>]) {
  constexpr int K = kFaceSums<kAF>;
  constexpr bool kRowSums = !kAF;
  // sums with a xi factor: 3 (vertex a, vertex b, all vertices) or 4 (+ N)
  constexpr int R = kAF ? 4 : 3;
  float a[K];
#pragma unroll
  for (int i = 0; i < K; ++i) a[i] = 0.0f;
  float rs[R];  // this row's sums of the xi factors' coefficients
#pragma unroll
  for (int i = 0; i < R; ++i) rs[i] = 0.0f;
  float rx0 = 0.0f, rx1 = 0.0f;  // this row's xi_0, xi_1
  auto end_row = [&]() {
    if (!kRowSums) return;
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int o = i < 3 ? 3 * i : 10;
      a[o] = __builtin_fmaf(rs[i], rx0, a[o]);
      a[o + 1] = __builtin_fmaf(rs[i], rx1, a[o + 1]);
      rs[i] = 0.0f;
    }
  };
  // adds v xi to the sums of factor i
  auto add_xi = [&](int i, float v, const FreqRec& fr, float x2) {
    const int o = i < 3 ? 3 * i : 10;
    if (kRowSums) {
      rs[i] += v;
    } else {
      a[o] = __builtin_fmaf(v, fr.x0, a[o]);
      a[o + 1] = __builtin_fmaf(v, fr.x1, a[o + 1]);
    }
    a[o + 2] = __builtin_fmaf(v, x2, a[o + 2]);
  };
  double row = 0.0;    // A0 X0 + A1 X1 of the current band row
  float prow = 0.0f;   // b0 x0 + b1 x1
  float qrow = 0.0f;   // c0 x0 + c1 x1
  // where a band row starts at record k
  auto start_row = [&](int k) __attribute__((always_inline)) {
    const FreqRec& fr = tile[k];
    // the face's row data is read from LDS here, every time: without this
    // compiler barrier it would be held in registers for the whole tile
    asm volatile("" ::: "memory");
    const FaceRow fw = frow;
    row = fma(fw.A1, fr.X1, fma(fw.A0, fr.X0, kMagic ? kRevMagic : 0.0));
    prow = __builtin_fmaf(fw.b1, fr.x1, fw.b0 * fr.x0);
    qrow = __builtin_fmaf(fw.c1, fr.x1, fw.c0 * fr.x0);
    rx0 = fr.x0;
    rx1 = fr.x1;
  };
  auto pair = [&](int k) __attribute__((always_inline)) {
    const FreqRec& fr = tile[k];
    const FreqHead fh = *reinterpret_cast<const FreqHead*>(&fr);
    // the same operations, in the same order, as a full evaluation
    const double rev_a = fma(L.A2, fh.X2, row);
    const float p = __builtin_fmaf(L.b2, fh.x2, prow);
    const float q = __builtin_fmaf(L.c2, fh.x2, qrow);
    const Sorted so = sort3(p, q);
    cf d, pP, pQ;
    float rD, u;  // u: phase of the median vertex plus arg(g) / (2 pi), in revolutions
    if (!may_be_wide || !wave_is_wide(so.D)) {
      const float ua = kMagic ? __builtin_fmaf(rev_frac_scaled(rev_a), kTwoM32, fh.gph)
                              : reduce_rev(rev_a) + fh.gph;
      u = __builtin_fmaf(so.mid, kInv2PiF, ua);
      face_H_grad<false>(so, d, pP, pQ, rD);
    } else {
      // (the offset removed exactly: both lie in [2^20, 2^21))
      u = mid_rev<true>(kMagic ? rev_a - kRevMagic : rev_a, so.mid) + fh.gph;
      face_H_grad<true>(so, d, pP, pQ, rD);
    }
    // Contract with conj(g) first, scaled by rD: z = 2 conj(g) e_mid rD
    // = 2 |g| rD exp(-2 pi i u); with the halved d, pP, pQ (H = 2 d rD)
    //   rho = Re(conj g F) = Re(z d),  gamma = Re(conj g (-iF)) = Im(z d),
    //   Re(conj g dF/dx_lo) = Re(z pP) + rho rD,
    //   Re(conj g dF/dx_hi) = -(Re(z pQ) + rho rD),
    //   Re(conj g dF/dx_mid) = gamma - (lo) - (hi).
    const cf em = expmi_rev(u);
    const float zm = fr.gm * rD;
    const cf z = {zm * em.re, zm * em.im};
    const float rh = __builtin_fmaf(z.re, d.re, -z.im * d.im);
    cf zt = z;      // z times the face weight when it sits inside the sum
    float rt = rh;  // Re(zt d)
    if (kAF) {
      // z2 = i z / |xi|^2; zt = w z + s z2 with s = xi . an
      const cf z2 = {-z.im * fr.inv2, z.re * fr.inv2};
      const float sn = __builtin_fmaf(L.n2, fh.x2, __builtin_fmaf(L.n1, fr.x1, L.n0 * fr.x0));
      zt = {__builtin_fmaf(L.fw, z.re, sn * z2.re), __builtin_fmaf(L.fw, z.im, sn * z2.im)};
      rt = __builtin_fmaf(zt.re, d.re, -zt.im * d.im);
      const float nn = __builtin_fmaf(z2.re, d.re, -z2.im * d.im);
      add_xi(3, nn, fr, fh.x2);
    }
    const float rtD = rt * rD;
    const float ga = __builtin_fmaf(zt.re, d.im, zt.im * d.re);
    const float alo = __builtin_fmaf(zt.re, pP.re, __builtin_fmaf(-zt.im, pP.im, rtD));
    const float ahi = -__builtin_fmaf(zt.re, pQ.re, __builtin_fmaf(-zt.im, pQ.im, rtD));
    const float amid = ga - alo - ahi;
    // roles: vertex a (phase 0) is lo, mid or hi as ra = 0, 1, 2; vertex b likewise
    const bool pn = p < 0.0f;
    const int ra = pn + (q < 0.0f);
    const int rb = !pn + (q < p);
    const float al = (ra == 0) ? alo : ((ra == 2) ? ahi : amid);
    const float be = (rb == 0) ? alo : ((rb == 2) ? ahi : amid);
    add_xi(0, al, fr, fh.x2);
    add_xi(1, be, fr, fh.x2);
    add_xi(2, ga, fr, fh.x2);
    a[9] += rh;
  };
  // Row by row: the set bits of ~cont are the row starts (bit 0 always; the
  // bits from cnt on are clear in cont), so a row's pairs run with no
  // per-pair test.  One pair per iteration: two per iteration (one LDS
  // address for both) need registers that would spill at 6 waves per SIMD.
  const uint64_t starts = ~cont;
  for (int k = 0; k < cnt;) {
    start_row(k);
    const uint64_t next = (k < 63) ? (starts & (~0ull << (k + 1))) : 0ull;
    int k1 = next ? (int)__builtin_ctzll(next) : cnt;
    if (k1 > cnt) k1 = cnt;
    for (; k < k1; ++k) pair(k);
    end_row();
  }
#pragma unroll
  for (int i = 0; i < K; ++i) acc[i] += a[i];
}

// The adjoint's upstream gradient in polar form, once per band frequency:
// (2 |g|, arg(g) / (2 pi) in [-1, 0) revolutions), from FP64 hypot and
// atan2, each rounded once to FP32.  conj(g) exp(-i x) is then
// |g| exp(-2 pi i (x / (2 pi) + arg(g) / (2 pi))): one add to the phase
// instead of a complex product per pair.
template <typename T>
__global__ __launch_bounds__(256) void grad_polar_kernel(const T* __restrict__ grad, int64_t n,
                                                         float2* __restrict__ gpol) {
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= n) return;
  const double gr = grad[2 * m], gi = grad[2 * m + 1];
  double ph = atan2(gi, gr) * kInv2Pi;  // [-1/2, 1/2]
  if (ph >= 0.0) ph -= 1.0;
  gpol[m] = make_float2((float)(2.0 * hypot(gr, gi)), (float)ph);
}

// Work decomposition of the adjoint (BwdPlan): workgroup w takes face block
// fb (64 faces) and one frequency chunk.  Workgroups are dispatched
// round-robin over the 8 XCDs, so with `levels` > 0 the workgroups with
// w % 8 == x take the x-th eighth of the band, which then streams through
// that XCD's L2 once.  Each eighth is cut into `levels` geometric levels
// (1/2, 1/4, ..., the last two equal) and each level into `parts` equal
// parts; the face block varies fastest and the large pieces come first, so
// the last workgroups of the launch are short and the launch ends with
// little idle time.  levels == 0: nchunk equal chunks, no XCD grouping.
struct BwdChunk {
  int64_t m0, m1;  // frequency range
  int chunk;       // partial slot
};
__device__ __forceinline__ BwdChunk bwd_chunk(int w, int64_t n, int nfb, int nchunk, int levels,
                                              int parts, int& fb) {
  BwdChunk c;
  if (levels == 0) {
    fb = w % nfb;
    c.chunk = w / nfb;
    c.m0 = n * c.chunk / nchunk;
    c.m1 = n * (c.chunk + 1) / nchunk;
    return c;
  }
  const int x = w & 7, s = w >> 3;
  fb = s % nfb;
  const int idx = s / nfb, j = idx / parts, i = idx % parts;
  const int64_t X0 = n * x / 8, len = n * (x + 1) / 8 - X0;
  const int64_t L0 = X0 + len - (len >> j);
  const int64_t L1 = (j + 1 < levels) ? X0 + len - (len >> (j + 1)) : X0 + len;
  c.m0 = L0 + (L1 - L0) * i / parts;
  c.m1 = L0 + (L1 - L0) * (i + 1) / parts;
  c.chunk = x * levels * parts + idx;
  return c;
}

// Compiled for 6 waves per SIMD (79 VGPRs, no spills; the autofluorescence
// variant 80 with 24 B spilled once per wave), which hides more of the pair
// loop's latency than 5 waves at 92 VGPRs (5 % fewer cycles).  To fit, the
// faces' row-start data sits in LDS (FaceRow) and the pair loop takes one
// pair per iteration (with two, spills stored per tile wrote ~1 GB per launch).
template <bool kAF, typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6))) void meshft_bwd_kernel(
    const FaceRec* __restrict__ faces, int64_t nf, const T* __restrict__ xi,
    const float2* __restrict__ gpol, int64_t n, int nfb, int nchunk, int levels, int parts,
    PartT<kAF>* __restrict__ part /* [nchunk][kFaceSums][nf] */) {
  constexpr int K = kFaceSums<kAF>;
  constexpr size_t kTileBytes = sizeof(FreqRec) * kBwdWaves * 64;
  constexpr size_t kRowBytes = sizeof(FaceRow) * kBwdFaces;
  constexpr size_t kRedBytes = sizeof(double) * (kBwdWaves - 1) * K * kBwdFaces;
  constexpr size_t kLoopBytes = kTileBytes + kRowBytes;
  __shared__ __align__(16) unsigned char smem[kLoopBytes > kRedBytes ? kLoopBytes : kRedBytes];
  // the wave index as a scalar: the compiler then keeps the frequency loop's
  // bounds and the tile address in scalar registers
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int lane = (int)threadIdx.x & 63;
  FreqRec* tile = reinterpret_cast<FreqRec*>(smem) + wv * 64;
  int fb;
  const BwdChunk ck = bwd_chunk((int)blockIdx.x, n, nfb, nchunk, levels, parts, fb);
  const int chunk = ck.chunk;
  const int t = fb * kBwdFaces + lane;  // nf < 2^31 / 3 (checked at the boundary)
  const bool active = t < nf;
  FaceLane L = {};
  if (active) {
    const FaceRec r = faces[t];
    L.A0 = r.a[0]; L.A1 = r.a[1]; L.A2 = r.a[2];
    L.b0 = r.eb[0]; L.b1 = r.eb[1]; L.b2 = r.eb[2];
    L.c0 = r.ec[0]; L.c1 = r.ec[1]; L.c2 = r.ec[2];
    L.fw = -0.5f * r.w;  // the face weight 2 c_t A_t / sum c A (exact)
    L.n0 = -0.5f * r.an[0]; L.n1 = -0.5f * r.an[1]; L.n2 = -0.5f * r.an[2];  // exact
  }
  // (wave-uniform: kept in a scalar register)
  const float e2max = __builtin_bit_cast(
      float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(
                 int, wave_max64(face_extent2(L.b0, L.b1, L.b2, L.c0, L.c1, L.c2)))));
  // largest |a / 2 pi|^2 of the block's faces (wave-uniform), for the bound
  // on the vertex phases that allows the kMagic reduction
  const float amax2 = __builtin_bit_cast(
      float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(
                 int, wave_max64(__double2float_ru(L.A0 * L.A0 + L.A1 * L.A1 + L.A2 * L.A2)))));
  // the four waves share the face block: wave 0 stores its row data
  FaceRow* rows = reinterpret_cast<FaceRow*>(smem + kTileBytes);
  if (wv == 0) rows[lane] = FaceRow{L.A0, L.A1, L.b0, L.b1, L.c0, L.c1};
  __syncthreads();
  // this wave's share of the chunk
  const int64_t q0 = ck.m0 + (ck.m1 - ck.m0) * wv / kBwdWaves;
  const int64_t q1 = ck.m0 + (ck.m1 - ck.m0) * (wv + 1) / kBwdWaves;
  double acc[K];
#pragma unroll
  for (int i = 0; i < K; ++i) acc[i] = 0.0;
  for (int64_t base = q0; base < q1; base += 64) {
    const int cnt = (int)((q1 - base < 64) ? q1 - base : 64);
    bool wide = false, same_row = false, big = false;
    if (lane < cnt) {
      const int64_t m = base + lane;
      const T* xm = xi + 3 * m;
      FreqRec fr;
      fr.X0 = xm[0]; fr.X1 = xm[1]; fr.X2 = xm[2];
      fr.x0 = (float)fr.X0; fr.x1 = (float)fr.X1; fr.x2 = (float)fr.X2;
      const float2 gp = gpol[m];
      fr.gm = gp.x;
      fr.gph = gp.y;
      if (kAF) {
        const double q = fr.X0 * fr.X0 + fr.X1 * fr.X1 + fr.X2 * fr.X2;
        fr.inv2 = (q > 0.0) ? (float)(1.0 / q) : 0.0f;
      } else {
        fr.inv2 = 0.0f;
      }
      tile[lane] = fr;
      const float q32 = fr.x0 * fr.x0 + fr.x1 * fr.x1 + fr.x2 * fr.x2;
      wide = q32 * e2max > kWidePhase2;
      big = !(q32 * amax2 < kRevMagicLimit2);
      same_row = lane > 0 && xm[-3] == xm[0] && xm[-2] == xm[1];
    }
    const uint64_t cont = __builtin_amdgcn_ballot_w64(same_row);
    const bool may_be_wide = __builtin_amdgcn_ballot_w64(wide) != 0;
    const bool small_rev = __builtin_amdgcn_ballot_w64(big) == 0;
    wave_lds_sync();
    if (small_rev)
      bwd_tile<kAF, true>(tile, cnt, cont, may_be_wide, L, rows[lane], acc);
    else
      bwd_tile<kAF, false>(tile, cnt, cont, may_be_wide, L, rows[lane], acc);
    wave_lds_sync();
  }
  // add the waves' sums in a fixed order (the tiles' LDS is reused)
  __syncthreads();
  double* red = reinterpret_cast<double*>(smem);
  if (wv > 0) {
#pragma unroll
    for (int i = 0; i < K; ++i) red[((wv - 1) * K + i) * kBwdFaces + lane] = acc[i];
  }
  __syncthreads();
  // the face index formed again from an opaque copy of the lane id, so that
  // the compiler does not hold t (spilled at 6 waves) through the loop
  int lane_c;
  asm volatile("v_mov_b32 %0, %1" : "=v"(lane_c) : "v"(lane));
  const int tc = fb * kBwdFaces + lane_c;
  if (wv == 0 && tc < nf) {
    PartT<kAF>* dst = part + (int64_t)chunk * K * nf + tc;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      double s = acc[i];
#pragma unroll
      for (int w = 0; w < kBwdWaves - 1; ++w) s += red[(w * K + i) * kBwdFaces + lane];
      dst[(int64_t)i * nf] = (PartT<kAF>)s;
    }
  }
}

// Per-face: add the chunk partials in order, then form the corner gradients
//   G_tj  = w_t P_tj + (2 c_t / S) R_t dA_t/dv_j
//   cgA_tj = c_t dA_t/dv_j
// and wR_t = w_t R_t (for K = sum_t w_t R_t = sum_m Re(conj g_m ftmesh_m)).
// dA/dv_j = 1/2 n_hat x (v_{j+2} - v_{j+1}); a zero-area face has no normal
// and contributes no area gradient (the reference divides by zero, .cu:24).
// Autofluorescence variant (P_tj already carries the face weight):
//   G_tj  = P_tj + (2 c_t / S) R_t dA_t/dv_j + (regint / regvol) (v_{j+1} - v_{j+2}) x N_t
//   cgV_tj = d regvol / dv_j = -(v_{j+1} x v_{j+2}) / 6,  qN_t = nor_t . N_t
// (nor_t . N_t is the derivative w.r.t. nor_t; its regvol dependence is the
// global term -(regint / regvol^2) sum_t qN_t cgV_tj, added in the gather).
template <typename IdxT, bool kAF, typename T>
__global__ __launch_bounds__(256) void bwd_face_finalize_kernel(
    const T* __restrict__ V, int64_t nv, const IdxT* __restrict__ Fc, int64_t nf,
    const T* __restrict__ coeff, const double* __restrict__ cA, const double* __restrict__ S_in,
    const PartT<kAF>* __restrict__ part, int nchunk, float regint, double* __restrict__ G,
    double* __restrict__ cgA, double* __restrict__ wR, double* __restrict__ cgV,
    double* __restrict__ qN, int* flag) {
  constexpr int K = kFaceSums<kAF>;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nf) return;
  double s[K];
#pragma unroll
  for (int i = 0; i < K; ++i) s[i] = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    const PartT<kAF>* src = part + (int64_t)c * K * nf + t;
#pragma unroll
    for (int i = 0; i < K; ++i) s[i] += (double)src[(int64_t)i * nf];
  }
  double P[3][3];
  load_face_verts(V, nv, Fc, t, P, flag);
  double u[3], q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u[d] = P[1][d] - P[0][d];
    q[d] = P[2][d] - P[0][d];
  }
  double nr[3] = {u[1] * q[2] - u[2] * q[1], u[2] * q[0] - u[0] * q[2], u[0] * q[1] - u[1] * q[0]};
  if (kAF) {
    qN[t] = nr[0] * s[10] + nr[1] * s[11] + nr[2] * s[12];
  }
  const double nn = sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
  const double inv = (nn > 0.0) ? 1.0 / nn : 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) nr[d] *= inv;
  const double S = S_in[0];
  const double ct = (double)coeff[t];
  const double w = (S > 0.0) ? 2.0 * cA[t] / S : 0.0;
  const double aw = (S > 0.0) ? 2.0 * ct / S * s[9] : 0.0;
  const double regvol = S_in[1];
  const double sc = (kAF && regint != 0.0f && regvol != 0.0) ? (double)regint / regvol : 0.0;
  const double pw = kAF ? 1.0 : w;  // the autofluorescence sums carry w_t already
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
    double gN[3] = {0.0, 0.0, 0.0};
    if (kAF) {
      // (v_{j+1} - v_{j+2}) x N and d regvol / dv_j = -(v_{j+1} x v_{j+2}) / 6
      const double f[3] = {-e[0], -e[1], -e[2]};
      gN[0] = sc * (f[1] * s[12] - f[2] * s[11]);
      gN[1] = sc * (f[2] * s[10] - f[0] * s[12]);
      gN[2] = sc * (f[0] * s[11] - f[1] * s[10]);
      cgV[9 * t + 3 * j + 0] = -(e1[1] * e2[2] - e1[2] * e2[1]) / 6.0;
      cgV[9 * t + 3 * j + 1] = -(e1[2] * e2[0] - e1[0] * e2[2]) / 6.0;
      cgV[9 * t + 3 * j + 2] = -(e1[0] * e2[1] - e1[1] * e2[0]) / 6.0;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      G[9 * t + 3 * j + d] = pw * Pj[j][d] + aw * gA[d] + gN[d];
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

// Exclusive scan of cnt[0..count) into off[0..count], off[count] = total, in
// one workgroup: each thread sums a contiguous range, the 1024 range sums
// are scanned in LDS (Hillis-Steele), then each thread writes its range.
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
  for (int d = 1; d < 1024; d <<= 1) {
    const int64_t add = ((int)threadIdx.x >= d) ? sums[threadIdx.x - d] : 0;
    __syncthreads();
    sums[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t run = sums[threadIdx.x] - s;  // exclusive prefix of this thread's range
  if (threadIdx.x == 1023) off[count] = (int)sums[1023];
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

// grad_v = sum over incidences (t, j) of G_tj - (K/S) cgA_tj, and for the
// autofluorescence variant (cgV != null) - (regint Kv / regvol^2) cgV_tj
// with Kv = sum_t qN_t.
template <typename T>
__global__ __launch_bounds__(256) void vertex_gather_kernel(
    const int* __restrict__ row, const int* __restrict__ ent, int64_t nv,
    const double* __restrict__ G, const double* __restrict__ cgA, const double* __restrict__ K_in,
    const double* __restrict__ S_in, const double* __restrict__ cgV,
    const double* __restrict__ Kv_in, float regint, T* __restrict__ grad) {
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const double S = S_in[0];
  const double kS = (S > 0.0) ? K_in[0] / S : 0.0;
  const double regvol = S_in[1];
  const double kv = (cgV && regvol != 0.0) ? (double)regint * Kv_in[0] / (regvol * regvol) : 0.0;
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  for (int i = row[v]; i < row[v + 1]; ++i) {
    const int64_t e = ent[i];  // = 3 t + j, so G[3e + d] = G[t][j][d]
    g0 += G[3 * e] - kS * cgA[3 * e];
    g1 += G[3 * e + 1] - kS * cgA[3 * e + 1];
    g2 += G[3 * e + 2] - kS * cgA[3 * e + 2];
    if (cgV) {
      g0 -= kv * cgV[3 * e];
      g1 -= kv * cgV[3 * e + 1];
      g2 -= kv * cgV[3 * e + 2];
    }
  }
  grad[3 * v] = (T)g0;
  grad[3 * v + 1] = (T)g1;
  grad[3 * v + 2] = (T)g2;
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


// Workgroups of `kernel` (`threads` threads each) resident on the whole
// current device: CUs x workgroups per CU from the runtime's occupancy
// calculator, cached per device and kernel.  Without a device (workspace
// queries on a host with no GPU) the MI355X figure, 256 CUs x
// fallback_per_cu, is used.
static int resident_wgs(const void* kernel, int fallback_per_cu, int threads = 256) {
  constexpr int kSlots = 16;
  static std::mutex mu;
  static int dev_of[kSlots], val_of[kSlots];
  static const void* ker_of[kSlots];
  static int used = 0;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 256 * fallback_per_cu;
  }
  std::lock_guard<std::mutex> lock(mu);
  for (int i = 0; i < used; ++i)
    if (dev_of[i] == dev && ker_of[i] == kernel) return val_of[i];
  int cus = 0, per = 0;
  int v = 256 * fallback_per_cu;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) == hipSuccess &&
      cus > 0 && per > 0)
    v = cus * per;
  else
    (void)hipGetLastError();
  if (used < kSlots) {
    dev_of[used] = dev;
    ker_of[used] = kernel;
    val_of[used] = v;
    ++used;
  }
  return v;
}

// Geometric levels per XCD eighth of the band (DMIC_BWD_LEVELS overrides).
static int bwd_levels() {
  static const int r = [] {
    const char* e = getenv("DMIC_BWD_LEVELS");
    const int v = e ? atoi(e) : -1;
    return (v >= 0 && v <= 20) ? v : 6;
  }();
  return r;
}

struct BwdPlan {
  int nfb;     // face blocks of kBwdFaces
  int nchunk;  // frequency chunks (partial slots)
  int levels;  // 0: nchunk equal chunks; else geometric levels per XCD eighth
  int parts;   // equal parts per level
};

// The float kernel's occupancy sizes the plan for both boundary types, so the
// float and double entry points share one workspace size.
template <bool kAF>
static BwdPlan bwd_plan(int64_t nf, int64_t n) {
  BwdPlan p{(int)ceil_div(nf > 0 ? nf : 1, kBwdFaces), 1, 0, 1};
  if (nf <= 0 || n <= 0) return p;
  const int cap = resident_wgs((const void*)meshft_bwd_kernel<kAF, float>, 6);
  constexpr int64_t kMinPiece = kBwdWaves * 64;  // one full tile per wave
  // parts per level: enough workgroups per level to fill the device
  const int64_t parts = (cap + 8 * (int64_t)p.nfb - 1) / (8 * (int64_t)p.nfb);
  const int64_t len = n / 8;
  int levels = bwd_levels();
  while (levels > 0 && (len >> (levels > 1 ? levels - 1 : 0)) / parts < kMinPiece) --levels;
  if (levels > 0) {
    p.levels = levels;
    p.parts = (int)parts;
    p.nchunk = 8 * levels * (int)parts;
    return p;
  }
  // small bands: equal chunks, about 8 rounds of resident workgroups
  int64_t c = (8 * (int64_t)cap + p.nfb - 1) / p.nfb;
  const int64_t cmax = n / kMinPiece > 1 ? n / kMinPiece : 1;
  if (c > cmax) c = cmax;
  if (c < 1) c = 1;
  p.nchunk = (int)c;
  return p;
}

struct FwdWs {
  int* flag;      // [0] face index out of range, [1] largest squared face extent (float bits)
  double* cA;
  double* S;
  double* spart;  // per-workgroup partials of the area and volume sums
  FaceRec* recs;
};

static FwdWs carve_fwd(Carver& c, int64_t nf) {
  FwdWs w;
  w.flag = c.take<int>(4);
  w.cA = c.take<double>(nf > 0 ? nf : 1);
  w.S = c.take<double>(2);
  w.spart = c.take<double>(2 * (size_t)ceil_div(nf > 0 ? nf : 1, 256));
  w.recs = c.take<FaceRec>(64 * (size_t)face_blocks64(nf > 0 ? nf : 1));
  return w;
}

struct BwdWs {
  int* flag;
  double* cA;
  double* S;
  double* spart;
  FaceRec* recs;
  float2* gpol;  // upstream gradient in polar form, per band frequency
  void* part;    // chunk partials of the per-face sums, PartT<kAF>
  double* G;
  double* cgA;
  double* wR;
  double* K;
  double* cgV;  // autofluorescence variant only
  double* qN;
  double* Kv;
  int* cnt;
  int* fillc;
  int* row;
  int* ent;
};

template <bool kAF>
static BwdWs carve_bwd(Carver& c, int64_t nv, int64_t nf, int64_t n) {
  constexpr bool af = kAF;
  BwdWs w;
  const int64_t f1 = nf > 0 ? nf : 1, v1 = nv > 0 ? nv : 1;
  w.flag = c.take<int>(4);
  w.cA = c.take<double>(f1);
  w.S = c.take<double>(2);
  w.spart = c.take<double>(2 * (size_t)ceil_div(f1, 256));
  w.recs = c.take<FaceRec>(64 * (size_t)face_blocks64(f1));
  w.gpol = c.take<float2>(n > 0 ? n : 1);
  const BwdPlan p = bwd_plan<kAF>(nf, n);
  w.part = c.take<PartT<kAF>>((size_t)p.nchunk * kFaceSums<kAF> * f1);
  w.G = c.take<double>(9 * f1);
  w.cgA = c.take<double>(9 * f1);
  w.wR = c.take<double>(f1);
  w.K = c.take<double>(2);
  w.cgV = af ? c.take<double>(9 * f1) : nullptr;
  w.qN = af ? c.take<double>(f1) : nullptr;
  w.Kv = af ? c.take<double>(2) : nullptr;
  w.cnt = c.take<int>(v1);
  w.fillc = c.take<int>(v1);
  w.row = c.take<int>(v1 + 1);
  w.ent = c.take<int>(3 * f1);
  return w;
}

template <typename IdxT, typename T>
static int prep_faces(const T* V, int64_t nv, const IdxT* Fc, int64_t nf, const T* coeff,
                      float regint, double wscale, int* flag, double* cA, double* S,
                      double* spart, FaceRec* recs, hipStream_t st) {
  const unsigned nb = ceil_div(nf, 256);
  hipLaunchKernelGGL((face_area_sum_kernel<IdxT, T>), dim3(nb), dim3(256), 0, st, V, nv, Fc, nf,
                     coeff, cA, spart, flag);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(face_sum_final_kernel, dim3(1), dim3(1024), 0, st, spart, (int64_t)nb, S);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL((face_pack_kernel<IdxT, T>), dim3(ceil_div(64 * face_blocks64(nf), 256)),
                     dim3(256), 0, st, V, nv,
                     Fc, nf, cA, S, regint, wscale, recs, flag);
  DMIC_LAUNCH_CHECK();
  return DMIC_OK;
}

// T: the boundary's floating type (float, or double for the _f64 entry points).
template <typename IdxT, bool kAF, typename T>
static int forward_impl(const T* V, int64_t nv, const IdxT* Fc, int64_t nf, const T* coeff,
                        float regint, const T* xi, int64_t n, T* out, void* ws,
                        size_t ws_bytes, hipStream_t st) {
  Carver c(ws);
  FwdWs w = carve_fwd(c, nf);
  DMIC_REQUIRE(ws_bytes >= c.bytes(), DMIC_ERR_WORKSPACE,
               "forward workspace too small: %zu < %zu bytes", ws_bytes, c.bytes());
  // about one 64-thread workgroup per kFwdFreqs band entries on a 1-D grid,
  // whose size in work-items must fit in 32 bits
  constexpr int64_t kMaxFwdGroups = 0xffffffffLL / 64 - 65536;
  DMIC_REQUIRE(n / kFwdFreqs<kAF> <= kMaxFwdGroups, DMIC_ERR_ARG,
               "band of %lld frequencies exceeds the forward's grid (%lld)", (long long)n,
               (long long)(kMaxFwdGroups * kFwdFreqs<kAF>));
  hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, st, w.flag, 4, 0);
  DMIC_LAUNCH_CHECK();
  if (n == 0) return DMIC_OK;
  constexpr int64_t kWords = sizeof(T) / sizeof(int);  // 32-bit words per value
  if (nf == 0) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(2 * kWords * n, 256)), dim3(256), 0, st,
                       reinterpret_cast<int*>(out), 2 * kWords * n, 0);
    DMIC_LAUNCH_CHECK();
    return DMIC_OK;
  }
  int rc = prep_faces<IdxT, T>(V, nv, Fc, nf, coeff, kAF ? regint : 0.0f, kFwdWScale, w.flag,
                               w.cA, w.S, w.spart, w.recs, st);
  if (rc != DMIC_OK) return rc;
  // whole rounds of the resident waves with kF entries per wave, then one
  // round over which the remainder is spread (kFwdFreqs 8: 7 waves per SIMD)
  constexpr int64_t kF = kFwdFreqs<kAF>;
  const int64_t W = resident_wgs((const void*)meshft_fwd_kernel<kAF, T>, kAF ? 32 : 28, 64);
  const int64_t head = (n / (kF * W)) * W;
  const int64_t rem = n - head * kF;
  const int64_t tail = rem > 0 ? (rem < W ? rem : W) : 0;
  const dim3 grid((unsigned)(head + tail));
  rc = timer_begin(0, st);
  if (rc != DMIC_OK) return rc;
  hipLaunchKernelGGL((meshft_fwd_kernel<kAF, T>), grid, dim3(64), 0, st, w.recs, nf, xi, n, head,
                     tail, out, w.S, w.flag + 1, regint);
  DMIC_LAUNCH_CHECK();
  rc = timer_end(0, st);
  if (rc != DMIC_OK) return rc;
  return DMIC_OK;
}

template <typename IdxT, bool kAF, typename T>
static int backward_impl(const T* V, int64_t nv, const IdxT* Fc, int64_t nf, const T* coeff,
                         float regint, const T* grad, const T* xi, int64_t n, T* gverts, void* ws,
                         size_t ws_bytes, hipStream_t st) {
  constexpr int K = kFaceSums<kAF>;
  Carver c(ws);
  BwdWs w = carve_bwd<kAF>(c, nv, nf, n);
  DMIC_REQUIRE(ws_bytes >= c.bytes(), DMIC_ERR_WORKSPACE,
               "backward workspace too small: %zu < %zu bytes", ws_bytes, c.bytes());
  const BwdPlan p = bwd_plan<kAF>(nf, n);
  // the adjoint's grid size in work-items must fit in 32 bits
  DMIC_REQUIRE((int64_t)p.nfb * p.nchunk * 256 <= 0xffffffffLL, DMIC_ERR_ARG,
               "mesh of %lld faces exceeds the adjoint's grid (%d face blocks x %d chunks)",
               (long long)nf, p.nfb, p.nchunk);
  hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(256), 0, st, w.flag, 4, 0);
  DMIC_LAUNCH_CHECK();
  if (nv == 0) return DMIC_OK;
  constexpr int64_t kWords = sizeof(T) / sizeof(int);  // 32-bit words per value
  if (nf == 0) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(3 * kWords * nv, 256)), dim3(256), 0, st,
                       reinterpret_cast<int*>(gverts), 3 * kWords * nv, 0);
    DMIC_LAUNCH_CHECK();
    return DMIC_OK;
  }
  int rc = prep_faces<IdxT, T>(V, nv, Fc, nf, coeff, kAF ? regint : 0.0f, 1.0, w.flag, w.cA,
                               w.S, w.spart, w.recs, st);
  if (rc != DMIC_OK) return rc;
  if (n > 0) {
    hipLaunchKernelGGL((grad_polar_kernel<T>), dim3(ceil_div(n, 256)), dim3(256), 0, st, grad, n,
                       w.gpol);
    DMIC_LAUNCH_CHECK();
    const dim3 grid((unsigned)((int64_t)p.nfb * p.nchunk));
    rc = timer_begin(1, st);
    if (rc != DMIC_OK) return rc;
    hipLaunchKernelGGL((meshft_bwd_kernel<kAF, T>), grid, dim3(256), 0, st, w.recs, nf, xi,
                       w.gpol, n, p.nfb, p.nchunk, p.levels, p.parts,
                       static_cast<PartT<kAF>*>(w.part));
    DMIC_LAUNCH_CHECK();
    rc = timer_end(1, st);
    if (rc != DMIC_OK) return rc;
  } else {
    const int64_t words = K * nf * (int64_t)(sizeof(PartT<kAF>) / 4);
    hipLaunchKernelGGL(fill_i32_kernel, dim3(ceil_div(words, 256)), dim3(256), 0, st,
                       static_cast<int*>(w.part), words, 0);
    DMIC_LAUNCH_CHECK();
  }
  const int nchunk = (n > 0) ? p.nchunk : 1;
  hipLaunchKernelGGL((bwd_face_finalize_kernel<IdxT, kAF, T>), dim3(ceil_div(nf, 256)), dim3(256), 0,
                     st, V, nv, Fc, nf, coeff, w.cA, w.S, static_cast<const PartT<kAF>*>(w.part),
                     nchunk, regint, w.G, w.cgA, w.wR,
                     w.cgV, w.qN, w.flag);
  DMIC_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(1024), 0, st, w.wR, nf, w.K);
  DMIC_LAUNCH_CHECK();
  if (kAF) {
    hipLaunchKernelGGL(sum_f64_kernel, dim3(1), dim3(1024), 0, st, w.qN, nf, w.Kv);
    DMIC_LAUNCH_CHECK();
  }
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
  hipLaunchKernelGGL(vertex_gather_kernel<T>, dim3(ceil_div(nv, 256)), dim3(256), 0, st, w.row,
                     w.ent, nv, w.G, w.cgA, w.K, w.S, w.cgV, w.Kv, regint, gverts);
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
  (void)n;
  Carver c(nullptr);
  carve_fwd(c, nf);
  return c.bytes();
}

size_t dmic_meshft_backward_workspace(int64_t nv, int64_t nf, int64_t n) {
  if (nv < 0 || nf < 0 || n < 0) return 0;
  Carver c(nullptr);
  carve_bwd<false>(c, nv, nf, n);
  return c.bytes();
}

size_t dmic_meshft_af_backward_workspace(int64_t nv, int64_t nf, int64_t n) {
  if (nv < 0 || nf < 0 || n < 0) return 0;
  Carver c(nullptr);
  carve_bwd<true>(c, nv, nf, n);
  return c.bytes();
}

}  // extern "C"

namespace {

template <bool kAF, typename T>
int forward_entry(const T* verts, int64_t nv, const void* faces, int faces_is_int64,
                  int64_t nf, const T* faces_coeff, float regint, const T* xi_list,
                  int64_t n, T* ftmesh_list, void* workspace, size_t workspace_bytes,
                  hipStream_t stream) {
  g_err[0] = 0;
  DMIC_REQUIRE(nv >= 0 && nf >= 0 && n >= 0, DMIC_ERR_ARG, "negative size");
  DMIC_REQUIRE(nf < (int64_t(1) << 31) / 3 && n < (int64_t(1) << 40), DMIC_ERR_ARG, "size too large");
  DMIC_REQUIRE(n == 0 || (xi_list && ftmesh_list), DMIC_ERR_ARG, "null xi_list/ftmesh_list");
  DMIC_REQUIRE(nf == 0 || n == 0 || (verts && faces && faces_coeff), DMIC_ERR_ARG,
               "null mesh pointer");
  DMIC_REQUIRE(workspace || workspace_bytes == 0, DMIC_ERR_ARG, "null workspace");
  if (faces_is_int64)
    return forward_impl<int64_t, kAF, T>(verts, nv, static_cast<const int64_t*>(faces), nf,
                                         faces_coeff, regint, xi_list, n, ftmesh_list, workspace,
                                         workspace_bytes, stream);
  return forward_impl<int32_t, kAF, T>(verts, nv, static_cast<const int32_t*>(faces), nf,
                                       faces_coeff, regint, xi_list, n, ftmesh_list, workspace,
                                       workspace_bytes, stream);
}

template <bool kAF, typename T>
int backward_entry(const T* verts, int64_t nv, const void* faces, int faces_is_int64,
                   int64_t nf, const T* faces_coeff, float regint, const T* grad_output_list,
                   const T* xi_list, int64_t n, T* grad_verts, void* workspace,
                   size_t workspace_bytes, hipStream_t stream) {
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
    return backward_impl<int64_t, kAF, T>(verts, nv, static_cast<const int64_t*>(faces), nf,
                                          faces_coeff, regint, grad_output_list, xi_list, n,
                                          grad_verts, workspace, workspace_bytes, stream);
  return backward_impl<int32_t, kAF, T>(verts, nv, static_cast<const int32_t*>(faces), nf,
                                        faces_coeff, regint, grad_output_list, xi_list, n,
                                        grad_verts, workspace, workspace_bytes, stream);
}

}  // namespace

// ------------------------------------------------------- no-allocation forward
// dmic_meshft_forward_noalloc keeps its scratch memory (the packed face
// records and the area sums) in this file-scope device array: one copy per
// GPU, held until the process ends.
// Host code never passes `scratch` itself (its host-side name is not a device
// address); it uses the per-device address from hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = size_t(20) << 20;
constexpr int64_t kNoAllocMaxFaces = 262144;  // 72 B per face: 18.9 MB + the area-sum partials
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Address of this device's copy of `scratch`, looked up once per device with
// that device current.
int scratch_for_current_device(void** out) {
  int dev = 0;
  DMIC_HIP_CHECK(hipGetDevice(&dev));
  DMIC_REQUIRE(dev >= 0 && dev < kMaxDevices, DMIC_ERR_ARG, "device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    DMIC_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DMIC_REQUIRE(p != nullptr, DMIC_ERR_HIP, "hipGetSymbolAddress returned null for scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return DMIC_OK;
}

}  // namespace

extern "C" {

int dmic_meshft_forward_noalloc(const float* verts, int64_t nv, const void* faces,
                                int faces_is_int64, int64_t nf, const float* faces_coeff,
                                const float* xi_list, int64_t n, float* ftmesh_list,
                                hipStream_t stream) {
  g_err[0] = 0;
  DMIC_REQUIRE(nv >= 0 && nf >= 0 && n >= 0, DMIC_ERR_ARG, "negative size");
  DMIC_REQUIRE(nf <= kNoAllocMaxFaces, DMIC_ERR_ARG,
               "nf = %lld exceeds the %lld faces the no-allocation forward's %zu-byte scratch "
               "holds; use dmic_meshft_forward with a workspace",
               (long long)nf, (long long)kNoAllocMaxFaces, kNoAllocScratchBytes);
  DMIC_REQUIRE(n < (int64_t(1) << 40), DMIC_ERR_ARG, "size too large");
  DMIC_REQUIRE(n == 0 || (xi_list && ftmesh_list), DMIC_ERR_ARG, "null xi_list/ftmesh_list");
  DMIC_REQUIRE(nf == 0 || n == 0 || (verts && faces && faces_coeff), DMIC_ERR_ARG,
               "null mesh pointer");
  void* ws = nullptr;
  const int rc = scratch_for_current_device(&ws);
  if (rc != DMIC_OK) return rc;
  if (faces_is_int64)
    return forward_impl<int64_t, false, float>(verts, nv, static_cast<const int64_t*>(faces), nf,
                                               faces_coeff, 0.0f, xi_list, n, ftmesh_list, ws,
                                               kNoAllocScratchBytes, stream);
  return forward_impl<int32_t, false, float>(verts, nv, static_cast<const int32_t*>(faces), nf,
                                             faces_coeff, 0.0f, xi_list, n, ftmesh_list, ws,
                                             kNoAllocScratchBytes, stream);
}

int dmic_meshft_forward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                        int64_t nf, const float* faces_coeff, const float* xi_list, int64_t n,
                        float* ftmesh_list, void* workspace, size_t workspace_bytes,
                        hipStream_t stream) {
  return forward_entry<false, float>(verts, nv, faces, faces_is_int64, nf, faces_coeff, 0.0f,
                                     xi_list, n, ftmesh_list, workspace, workspace_bytes, stream);
}

int dmic_meshft_forward_f64(const double* verts, int64_t nv, const void* faces,
                            int faces_is_int64, int64_t nf, const double* faces_coeff,
                            const double* xi_list, int64_t n, double* ftmesh_list,
                            void* workspace, size_t workspace_bytes, hipStream_t stream) {
  return forward_entry<false, double>(verts, nv, faces, faces_is_int64, nf, faces_coeff, 0.0f,
                                      xi_list, n, ftmesh_list, workspace, workspace_bytes, stream);
}

int dmic_meshft_backward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                         int64_t nf, const float* faces_coeff, const float* grad_output_list,
                         const float* ftmesh_list, const float* xi_list, int64_t n,
                         float* grad_verts, void* workspace, size_t workspace_bytes,
                         hipStream_t stream) {
  (void)ftmesh_list;
  return backward_entry<false, float>(verts, nv, faces, faces_is_int64, nf, faces_coeff, 0.0f,
                                      grad_output_list, xi_list, n, grad_verts, workspace,
                                      workspace_bytes, stream);
}

int dmic_meshft_backward_f64(const double* verts, int64_t nv, const void* faces,
                             int faces_is_int64, int64_t nf, const double* faces_coeff,
                             const double* grad_output_list, const double* ftmesh_list,
                             const double* xi_list, int64_t n, double* grad_verts,
                             void* workspace, size_t workspace_bytes, hipStream_t stream) {
  (void)ftmesh_list;
  return backward_entry<false, double>(verts, nv, faces, faces_is_int64, nf, faces_coeff, 0.0f,
                                       grad_output_list, xi_list, n, grad_verts, workspace,
                                       workspace_bytes, stream);
}

int dmic_meshft_af_forward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                           int64_t nf, const float* faces_coeff, float regint,
                           const float* xi_list, int64_t n, float* ftmesh_list, void* workspace,
                           size_t workspace_bytes, hipStream_t stream) {
  return forward_entry<true, float>(verts, nv, faces, faces_is_int64, nf, faces_coeff, regint,
                                    xi_list, n, ftmesh_list, workspace, workspace_bytes, stream);
}

int dmic_meshft_af_backward(const float* verts, int64_t nv, const void* faces, int faces_is_int64,
                            int64_t nf, const float* faces_coeff, float regint,
                            const float* grad_output_list, const float* xi_list, int64_t n,
                            float* grad_verts, void* workspace, size_t workspace_bytes,
                            hipStream_t stream) {
  return backward_entry<true, float>(verts, nv, faces, faces_is_int64, nf, faces_coeff, regint,
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
