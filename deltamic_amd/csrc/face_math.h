// Per-face closed form of the narrow-band mesh Fourier transform, FP32 with
// FP64 phases, for gfx950.
//
// For a face (a, b, c) and frequency xi, with phases x_j = xi . v_j, the
// reference evaluates (fourier_xi_tri_f, deltamic/cpp/fourier3d_cuda_kernel.cu:239-278)
//     F = e_a/(d_ab d_ca) + e_b/(d_bc d_ab) + e_c/(d_ca d_bc),  e_j = exp(-i x_j),
// the second divided difference f[x_a, x_b, x_c] of f(x) = -exp(-ix).  In
// FP32 that sum cancels catastrophically: its error grows like eps/d^2.
//
// Here the same function is evaluated in a shifted form.  Let the three
// phases, sorted, be lo <= mid <= hi, P = lo - mid <= 0, Q = hi - mid >= 0,
// D = Q - P (the largest phase gap).  Then
//     F = exp(-i x_mid) * H(P, Q),  H = (phi(P) - phi(Q)) / D,
//     phi(y) = (exp(-iy) - 1)/y = -(sin(y/2)/(y/2)) (sin(y/2) + i cos(y/2)),
// where phi is evaluated without cancellation, so the error is ~eps/D.
// When D < TAYLOR_D every phase is within 0.1 rad of the others and H comes
// from its Taylor series in P and Q (complete homogeneous symmetric
// polynomials h_j(P, Q)), exact to < 1e-9.
//
// Derivatives (adjoint, grad_fourier_xi_tri_{a,b,c}_f, .cu:74-233):
//     dF/dx_lo = e_m dH/dP,  dF/dx_hi = e_m dH/dQ,
//     dF/dx_mid = -iF - dF/dx_lo - dF/dx_hi,
//     dH/dP = (phi'(P) + H)/D,  dH/dQ = -(phi'(Q) + H)/D,
// with Im phi'(y) from its series for |y| < 0.5 and in closed form otherwise.
//
// The phase of vertex a is computed in revolutions in FP64 from xi (FP32,
// exact in FP64) and a / (2 pi) (FP64, relative error 1e-16), and reduced
// exactly before it is rounded to FP32.  Phase differences p = xi.(v_b - v_a) and
// q = xi.(v_c - v_a) come from FP32 edge vectors.
//
// sin and cos are the hardware v_sin_f32 / v_cos_f32, which take revolutions
// (measured on MI355X: 1.2e-7 absolute, 8.7e-7 relative down to 1e-9
// revolutions).  Their input domain is |u| <= 256 revolutions.  Without a
// reduction the FP32 arguments lose absolute precision as the phase gap D
// grows (|mid| <= D, so |x_mid| < 1 + D/2pi and |y/4pi| <= D/4pi
// revolutions, each rounded to half an ulp), so a wave in which any pair has
// D > kWidePhase takes the "wide" variant for all its lanes: the
// median-vertex phase in FP64 and every sin/cos argument reduced to
// [-1/2, 1/2] revolutions first.  Below kWidePhase = 16 rad the unreduced
// arguments stay under 3.6 revolutions, i.e. at most 1.2e-7 revolutions of
// rounding, the size of the hardware sin/cos error.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dmic {

constexpr double kInv2Pi = 0.159154943091895335768883763372514362;
constexpr float kInv4PiF = 0.0795774715459476678844418816862571810f;
constexpr float kInv4PiLoF = 3.2103191632826267e-09f;  // 1/(4 pi) - kInv4PiF
constexpr float kTaylorD = 0.1f;    // largest phase gap handled by the Taylor branch
constexpr float kPhiSeries = 0.5f;  // |y| below which Im phi'(y) uses its series
constexpr float kWidePhase = 16.0f;  // phase gap above which a wave range-reduces its arguments
// P and Q are kept at least kTinyY away from 0 (phi moves by < 7e-9 there),
// so one reciprocal of P Q D gives 1/P, 1/Q and 1/D.
constexpr float kTinyY = 1.3e-8f;
constexpr float kFltMax = 3.402823466e38f;

// P <= -kTinyY and Q >= kTinyY, each one v_med3_f32 (fminf/fmaxf would add a
// NaN-quieting canonicalize per operand).
__device__ __forceinline__ float clamp_P(float P) {
  return __builtin_amdgcn_fmed3f(P, -kFltMax, -kTinyY);
}
__device__ __forceinline__ float clamp_Q(float Q) {
  return __builtin_amdgcn_fmed3f(Q, kTinyY, kFltMax);
}

struct cf {
  float re, im;
};

// 1/x with v_rcp_f32 (1 ulp): the IEEE division sequence costs ~10 VALU ops.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ cf cmul(cf a, cf b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

constexpr float kInv2PiF = 0.159154943091895335768883763372514362f;

// A phase given in revolutions (FP64), reduced to [0, 1) (v_fract_f64) and
// rounded to FP32 (absolute error <= 3e-8 revolutions).
__device__ __forceinline__ float reduce_rev(double r) {
  return (float)__builtin_amdgcn_fract(r);
}

// The same reduction without the FP64 fract and convert.  When the FP64
// phase is formed as x = kRevMagic + rev (the offset added to the row term,
// the FMA of the last term rounding x once) and |rev| < 2^19, x lies in
// [2^20, 2^21), where the FP64 spacing is 2^-32 revolutions and kRevMagic is
// an integer: the low 32 bits of x are frac(rev) 2^32.  Read as a signed
// integer they are the phase in [-1/2, 1/2) revolutions times 2^32, which
// one conversion rounds to FP32 (the caller applies 2^-32 in its next fma).
// Callers use it where (max |a/2pi|^2)(max |xi|^2) < kRevMagicLimit2, i.e.
// |rev| < 2^18, and the FP64 path otherwise.
constexpr double kRevMagic = 1572864.0;             // 1.5 * 2^20
constexpr float kRevMagicLimit2 = 6.8719476736e10f;  // (2^18)^2
constexpr float kTwoM32 = 2.3283064365386963e-10f;   // 2^-32
__device__ __forceinline__ float rev_frac_scaled(double x) {
  return (float)__double2loint(x);
}

// exp(-2 pi i u) for u in revolutions (|u| <= 256) with the hardware sin/cos.
// Only the common phase factor of a face uses it, where an absolute error
// just rotates that face's contribution.
__device__ __forceinline__ cf expmi_rev(float u) {
  return {__builtin_amdgcn_cosf(u), -__builtin_amdgcn_sinf(u)};
}
// Phase of the median vertex in revolutions, reduced: FP32 sum for ordinary
// waves, FP64 sum and exact reduction for wide ones (|mid| up to any size).
// Whether a wave is wide is decided for the whole wave (wave_is_wide), so
// the two variants never run side by side under an execution mask.
__device__ __forceinline__ bool wave_is_wide(float D) {
  return __builtin_amdgcn_ballot_w64(D > kWidePhase) != 0;
}
template <bool kWide>
__device__ __forceinline__ float mid_rev(double rev_a, float mid) {
  if (kWide) return reduce_rev(fma((double)mid, kInv2Pi, rev_a));
  return __builtin_fmaf(mid, kInv2PiF, reduce_rev(rev_a));
}
// phi'(y)/2, from s = sin(y/2), c = cos(y/2), h = s/y (= k/2 with
// k = sin(y/2)/(y/2)), t = h c (which the caller shares with H) and ry = 1/y:
//     phi'(y) = [(-sin y + k s) + i(-cos y + k c)] / y.
// Real part / 2 = h^2 - t = h (h - c): no cancellation for any y (2h, c -> 1
// as y -> 0).
// Imaginary part / 2 = (t + s^2 - 1/2)/y (half the spherical Bessel j1(y))
// cancels for small y: the sin/cos errors (1.2e-7) give it an absolute
// error of ~1.2e-7/y^2, i.e. ~5e-7/y^2 of |phi'/2| (>= 1/4).  Below
// kPhiSeries = 0.5 it comes from its series y (1/6 - y^2/60 + y^4/1680)
// (truncation y^7/90720 < 9e-8, 3.5e-7 of |phi'/2|, at |y| = 0.5, where
// the closed form is within 1.9e-6 of it).  The threshold trades that error for
// fewer waves whose lanes split between the two forms: on the 20,480-face
// sphere 53 % of the waves split at 0.5 against 67 % at 0.85, where the
// series needed a fourth term.
// A wave whose lanes all take the same form evaluates only that one
// (wave-uniform branches); at 0.5, 46 % of those waves take the closed form
// on every lane.
__device__ __forceinline__ float dphi_im_series(float y) {
  const float y2 = y * y;
  const float si = __builtin_fmaf(y2, 5.9523809524e-4f, -1.6666666667e-2f);
  return __builtin_fmaf(y2, si, 1.6666666667e-1f) * y;
}
__device__ __forceinline__ cf dphi_half(float y, float s, float h, float t, float ry) {
  const float re = __builtin_fmaf(h, h, -t);
  const bool small = __builtin_fabsf(y) < kPhiSeries;
  const uint64_t ns = __builtin_amdgcn_ballot_w64(small);
  float im;
  if (ns == 0) {
    im = (t + __builtin_fmaf(s, s, -0.5f)) * ry;
  } else if (ns == __builtin_amdgcn_ballot_w64(true)) {
    im = dphi_im_series(y);
  } else {
    // both forms, selected by the ballot itself (the compare is not redone)
    const float di = (t + __builtin_fmaf(s, s, -0.5f)) * ry;
    im = __builtin_amdgcn_inverse_ballot_w64(ns) ? dphi_im_series(y) : di;
  }
  return {re, im};
}
// Taylor branch: H(P,Q) = sum_k -(-i)^k h_{k-2}(P,Q)/k!, k = 2..8
// (kHalf: H/2, the same series with every coefficient halved, exactly).
template <bool kHalf = false>
__device__ __forceinline__ cf taylor_H(float P, float Q) {
  constexpr float c = kHalf ? 0.5f : 1.0f;
  const float s1 = P + Q, t1 = P * Q;
  const float h1 = s1;
  const float h2 = s1 * h1 - t1;
  const float h3 = s1 * h2 - t1 * h1;
  const float h4 = s1 * h3 - t1 * h2;
  const float h5 = s1 * h4 - t1 * h3;
  const float h6 = s1 * h5 - t1 * h4;
  const float re = c * 0.5f - h2 * (c * 4.1666666667e-2f) + h4 * (c * 1.3888888889e-3f) -
                   h6 * (c * 2.4801587302e-5f);
  const float im = -h1 * (c * 1.6666666667e-1f) + h3 * (c * 8.3333333333e-3f) -
                   h5 * (c * 1.9841269841e-4f);
  return {re, im};
}

// Taylor branch derivative dH/dP = sum_k -(-i)^k g_{k-3}/k!, k = 3..8, where
// g_j = h_j(P, P, Q) (the derivative of h_j(P,Q) w.r.t. P duplicates P).
// dH/dQ is the same with P and Q exchanged.  The adjoint's only caller
// takes half of it (coefficients halved, exactly).
__device__ __forceinline__ cf taylor_dH_half(float P, float Q) {
  const float E1 = 2.0f * P + Q, E2 = P * (P + 2.0f * Q), E3 = P * P * Q;
  const float g1 = E1;
  const float g2 = E1 * g1 - E2;
  const float g3 = E1 * g2 - E2 * g1 + E3;
  const float g4 = E1 * g3 - E2 * g2 + E3 * g1;
  const float g5 = E1 * g4 - E2 * g3 + E3 * g2;
  const float re = -g1 * (0.5f * 4.1666666667e-2f) + g3 * (0.5f * 1.3888888889e-3f) -
                   g5 * (0.5f * 2.4801587302e-5f);
  const float im = -(0.5f * 1.6666666667e-1f) + g2 * (0.5f * 8.3333333333e-3f) -
                   g4 * (0.5f * 1.9841269841e-4f);
  return {re, im};
}

// Sorted relative phases of a face: phases of (a, b, c) relative to a are
// (0, p, q).  mid is the median phase offset from a, P = lo - mid, Q = hi - mid.
struct Sorted {
  float mid, P, Q, D;
};

__device__ __forceinline__ Sorted sort3(float p, float q) {
  const float lo = __builtin_fminf(0.0f, __builtin_fminf(p, q));
  const float hi = __builtin_fmaxf(0.0f, __builtin_fmaxf(p, q));
  const float mid = __builtin_amdgcn_fmed3f(0.0f, p, q);
  return {mid, lo - mid, hi - mid, hi - lo};
}

// sin(y/2), cos(y/2) with the hardware sin/cos (kWide: reduced argument).
template <bool kWide>
__device__ __forceinline__ void half_sincos(float y, float& s, float& c) {
  float u = y * kInv4PiF;
  if (kWide) {
    // y/(4 pi) - r, r the nearest integer, from the exact product in one
    // fma (|result| <= 1/2 + ulp, so one rounding of <= 3e-8), plus the
    // error of the FP32 constant: accurate to ~6e-8 revolutions for any y.
    // Written as explicit fmas: with the product rounded first and its error
    // added separately, contraction of (u - r) into an fma would count that
    // error twice.
    const float r = __builtin_rintf(u);
    u = __builtin_fmaf(y, kInv4PiLoF, __builtin_fmaf(y, kInv4PiF, -r));
  }
  s = __builtin_amdgcn_sinf(u);
  c = __builtin_amdgcn_cosf(u);
}
// H(P, Q) for the forward, returned as H = -2 m N (the kernel folds m into
// the face weight, which carries the factor -2):
//     H = (phi(P) - phi(Q))/D = (Q A_P - P A_Q) / (P Q D),
//     A_y = exp(-iy) - 1 = -2 s (s + i c),  s = sin(y/2), c = cos(y/2),
// so N = sP Q (sP + i cP) - sQ P (sQ + i cQ) and m = 1/(P Q D): one
// reciprocal per pair.  The Taylor branch returns N = H and m = -1/2.
template <bool kWide>
__device__ __forceinline__ cf face_H(const Sorted& so, float& m) {
  cf N;
  if (so.D < kTaylorD) {
    N = taylor_H(so.P, so.Q);
    m = -0.5f;
  } else {
    const float P = clamp_P(so.P), Q = clamp_Q(so.Q);
    float sP, cP, sQ, cQ;
    half_sincos<kWide>(P, sP, cP);
    half_sincos<kWide>(Q, sQ, cQ);
    const float u = sP * Q, v = sQ * P;
    N = {__builtin_fmaf(u, sP, -v * sQ), __builtin_fmaf(u, cP, -v * cQ)};
    m = rcp(P * Q * so.D);
  }
  return N;
}
// The forward's form of face_H, with the phases in half revolutions
// (y' = c y, c = 1/(4 pi)): y' is then the argument of the hardware
// sin/cos of y/2 as it stands, which saves the two products y c per pair.
// N and m are homogeneous in the phases: N' = c N and m' = 1/(P'Q'D') =
// m / c^3, so the returned N' m' is N m / c^2; the caller's face weight
// carries the factor c^2 (kFwdWScale).  The Taylor branch scales P', Q'
// back to radians and returns m = -1/(2 c^2).  Only for waves whose phase
// gaps stay below kWidePhase (no argument reduction).
constexpr float k4PiF = 12.566370614359172954f;
// c^2 = 1/(16 pi^2) in FP64: the packing kernel forms the weight in FP64 and
// rounds it once, so the forward's values carry no scale error of their own
constexpr double kFwdWScale = (0.5 * kInv2Pi) * (0.5 * kInv2Pi);
constexpr float kFwdInvWScale = 157.91367041742973f;  // 16 pi^2
__device__ __forceinline__ cf face_H_halfrev(const Sorted& so, float& m) {
  cf N;
  if (so.D < kTaylorD * kInv4PiF) {
    N = taylor_H(so.P * k4PiF, so.Q * k4PiF);
    m = -0.5f * kFwdInvWScale;
  } else {
    const float P = __builtin_amdgcn_fmed3f(so.P, -kFltMax, -kTinyY * kInv4PiF);
    const float Q = __builtin_amdgcn_fmed3f(so.Q, kTinyY * kInv4PiF, kFltMax);
    const float sP = __builtin_amdgcn_sinf(P), cP = __builtin_amdgcn_cosf(P);
    const float sQ = __builtin_amdgcn_sinf(Q), cQ = __builtin_amdgcn_cosf(Q);
    const float u = sP * Q, v = sQ * P;
    N = {__builtin_fmaf(u, sP, -v * sQ), __builtin_fmaf(u, cP, -v * cQ)};
    m = rcp(P * Q * so.D);
  }
  return N;
}
// H and its derivatives for the adjoint, returned halved as (d, pP, pQ, rD) with
//     H/2 = d rD,  (dH/dP)/2 = (pP + d rD) rD,  (dH/dQ)/2 = -(pQ + d rD) rD,
// i.e. d = (phi(P) - phi(Q))/2, pP = phi'(P)/2, pQ = phi'(Q)/2, rD = 1/D in
// the closed form (phi(y) = -2 h (s + i c), h = sin(y/2)/y), and d = H/2,
// rD = 1 in the Taylor branch.  The halving saves the factor 2 of
// phi = -k (s + i c), k = 2h; the adjoint carries 2 |g| instead of |g|.  The
// kernel contracts with the upstream gradient before scaling by rD.
template <bool kWide>
__device__ __forceinline__ void face_H_grad(const Sorted& so, cf& d, cf& pP, cf& pQ, float& rD) {
  if (so.D < kTaylorD) {
    d = taylor_H<true>(so.P, so.Q);
    const cf tP = taylor_dH_half(so.P, so.Q);
    const cf tQ = taylor_dH_half(so.Q, so.P);
    pP = {tP.re - d.re, tP.im - d.im};
    pQ = {-tQ.re - d.re, -tQ.im - d.im};
    rD = 1.0f;
  } else {
    const float P = clamp_P(so.P), Q = clamp_Q(so.Q);
    float sP, cP, sQ, cQ;
    half_sincos<kWide>(P, sP, cP);
    half_sincos<kWide>(Q, sQ, cQ);
    const float PQ = P * Q;
    const float r = rcp(PQ * so.D);
    const float Dr = so.D * r;
    const float rP = Q * Dr, rQ = P * Dr;   // 1/P, 1/Q
    rD = PQ * r;
    const float hP = sP * rP, hQ = sQ * rQ;
    const float tP = hP * cP, tQ = hQ * cQ;  // shared by d and phi'
    d = {hQ * sQ - hP * sP, tQ - tP};
    pP = dphi_half(P, sP, hP, tP, rP);
    pQ = dphi_half(Q, sQ, hQ, tQ, rQ);
  }
}
}  // namespace dmic
