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
// with phi'(y) from its series for |y| < 0.25 and in closed form otherwise.
//
// The phase of vertex a is computed in revolutions in FP64 from xi (FP32,
// exact in FP64) and a / (2 pi) (FP64, relative error 1e-16), and reduced
// exactly before it is rounded to FP32.  Phase differences p = xi.(v_b - v_a) and
// q = xi.(v_c - v_a) come from FP32 edge vectors.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dmic {

constexpr double kInv2Pi = 0.159154943091895335768883763372514362;
constexpr float kInv4PiF = 0.0795774715459476678844418816862571810f;
constexpr float kTaylorD = 0.1f;    // largest phase gap handled by the Taylor branch
constexpr float kPhiSeries = 0.25f;  // |y| below which phi'(y) uses its series

struct cf {
  float re, im;
};

// 1/x with v_rcp_f32 (1 ulp): the IEEE division sequence costs ~10 VALU ops.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ cf cmul(cf a, cf b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// sin(2 pi u) and cos(2 pi u) for u in revolutions.  The octant reduction
// u = n/4 + f, |f| <= 1/8, is exact in FP32.  sin(2 pi f) = f P(f^2) and
// cos(2 pi f) = Q(f^2) are near-minimax fits on |f| <= 1/8 (relative error
// 3.2e-9 and 5.6e-11; fitted for this file by weighted least squares on
// Chebyshev nodes), so small arguments keep full relative accuracy.
__device__ __forceinline__ void sincos_rev(float u, float& s, float& c) {
  const float n = __builtin_rintf(u * 4.0f);
  const float f = __builtin_fmaf(n, -0.25f, u);
  const float f2 = f * f;
  float sp = __builtin_fmaf(f2, -75.39373016f, 81.59236145f);
  sp = __builtin_fmaf(f2, sp, -41.34166336f);
  sp = __builtin_fmaf(f2, sp, 6.283185482f);
  sp *= f;
  float cp = __builtin_fmaf(f2, 59.20288849f, -85.44229889f);
  cp = __builtin_fmaf(f2, cp, 64.93930817f);
  cp = __builtin_fmaf(f2, cp, -19.73920822f);
  cp = __builtin_fmaf(f2, cp, 1.0f);
  const int q = ((int)n) & 3;
  const float ss = (q & 1) ? cp : sp;
  const float cc = (q & 1) ? sp : cp;
  s = (q & 2) ? -ss : ss;
  c = ((q + 1) & 2) ? -cc : cc;
}

constexpr float kInv2PiF = 0.159154943091895335768883763372514362f;

// A phase given in revolutions (FP64), reduced to [-1/2, 1/2] and rounded to FP32.
__device__ __forceinline__ float reduce_rev(double r) {
  return (float)(r - __builtin_rint(r));
}

// exp(-2 pi i u) for u in revolutions (|u| < 2 here) with the hardware
// v_sin_f32 / v_cos_f32, which take revolutions and are accurate to ~1.2e-7
// absolute: enough for the common phase factor of a face, whose error only
// rotates that face's contribution.  Where relative accuracy near zero
// matters (phi) the polynomial sincos_rev is used instead.
__device__ __forceinline__ cf expmi_rev(float u) {
  return {__builtin_amdgcn_cosf(u), -__builtin_amdgcn_sinf(u)};
}

// phi(y) = (exp(-iy) - 1)/y.  Also returns s = sin(y/2), c = cos(y/2),
// k = sin(y/2)/(y/2) and ry = 1/y (0 for y ~ 0) for phi'(y).
__device__ __forceinline__ cf phi_of(float y, float& s, float& c, float& k, float& ry) {
  sincos_rev(y * kInv4PiF, s, c);
  const float h = 0.5f * y;
  const bool nz = __builtin_fabsf(h) > 1e-30f;
  const float rh = nz ? rcp(h) : 0.0f;
  k = nz ? s * rh : 1.0f;
  ry = 0.5f * rh;
  return {-k * s, -k * c};
}

// phi'(y): series for |y| < kPhiSeries, closed form otherwise.
__device__ __forceinline__ cf dphi_of(float y, float s, float c, float k, float ry) {
  const float y2 = y * y;
  // series: [-1/2 + y^2/8 - y^4/144 + y^6/5760] + i y [1/3 - y^2/30 + y^4/840 - y^6/45360]
  float sr = __builtin_fmaf(y2, 1.7361111111e-4f, -6.9444444444e-3f);
  sr = __builtin_fmaf(y2, sr, 0.125f);
  sr = __builtin_fmaf(y2, sr, -0.5f);
  float si = __builtin_fmaf(y2, -2.2045855379e-5f, 1.1904761905e-3f);
  si = __builtin_fmaf(y2, si, -3.3333333333e-2f);
  si = __builtin_fmaf(y2, si, 3.3333333333e-1f);
  si *= y;
  // closed form: [(-sin y + k s) + i(-cos y + k c)] / y,
  // sin y = 2 s c, cos y = 1 - 2 s^2
  const float dr = (__builtin_fmaf(-2.0f * s, c, k * s)) * ry;
  const float di = (__builtin_fmaf(2.0f * s, s, -1.0f) + k * c) * ry;
  const bool ser = __builtin_fabsf(y) < kPhiSeries;
  return {ser ? sr : dr, ser ? si : di};
}

// Taylor branch: H(P,Q) = sum_k -(-i)^k h_{k-2}(P,Q)/k!, k = 2..8.
__device__ __forceinline__ cf taylor_H(float P, float Q) {
  const float s1 = P + Q, t1 = P * Q;
  const float h1 = s1;
  const float h2 = s1 * h1 - t1;
  const float h3 = s1 * h2 - t1 * h1;
  const float h4 = s1 * h3 - t1 * h2;
  const float h5 = s1 * h4 - t1 * h3;
  const float h6 = s1 * h5 - t1 * h4;
  const float re = 0.5f - h2 * 4.1666666667e-2f + h4 * 1.3888888889e-3f - h6 * 2.4801587302e-5f;
  const float im = -h1 * 1.6666666667e-1f + h3 * 8.3333333333e-3f - h5 * 1.9841269841e-4f;
  return {re, im};
}

// Taylor branch derivative dH/dP = sum_k -(-i)^k g_{k-3}/k!, k = 3..8, where
// g_j = h_j(P, P, Q) (the derivative of h_j(P,Q) w.r.t. P duplicates P).
// dH/dQ is the same with P and Q exchanged.
__device__ __forceinline__ cf taylor_dH(float P, float Q) {
  const float E1 = 2.0f * P + Q, E2 = P * (P + 2.0f * Q), E3 = P * P * Q;
  const float g1 = E1;
  const float g2 = E1 * g1 - E2;
  const float g3 = E1 * g2 - E2 * g1 + E3;
  const float g4 = E1 * g3 - E2 * g2 + E3 * g1;
  const float g5 = E1 * g4 - E2 * g3 + E3 * g2;
  const float re = -g1 * 4.1666666667e-2f + g3 * 1.3888888889e-3f - g5 * 2.4801587302e-5f;
  const float im = -1.6666666667e-1f + g2 * 8.3333333333e-3f - g4 * 1.9841269841e-4f;
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

// H(P, Q) for the forward.
__device__ __forceinline__ cf face_H(const Sorted& so) {
  cf H;
  if (so.D < kTaylorD) {
    H = taylor_H(so.P, so.Q);
  } else {
    float sP, cP, kP, rP, sQ, cQ, kQ, rQ;
    const cf fP = phi_of(so.P, sP, cP, kP, rP);
    const cf fQ = phi_of(so.Q, sQ, cQ, kQ, rQ);
    const float rD = rcp(so.D);
    H = {(fP.re - fQ.re) * rD, (fP.im - fQ.im) * rD};
  }
  return H;
}

// H and its derivatives for the adjoint.
__device__ __forceinline__ void face_H_grad(const Sorted& so, cf& H, cf& dP, cf& dQ) {
  if (so.D < kTaylorD) {
    H = taylor_H(so.P, so.Q);
    dP = taylor_dH(so.P, so.Q);
    dQ = taylor_dH(so.Q, so.P);
  } else {
    float sP, cP, kP, rP, sQ, cQ, kQ, rQ;
    const cf fP = phi_of(so.P, sP, cP, kP, rP);
    const cf fQ = phi_of(so.Q, sQ, cQ, kQ, rQ);
    const float rD = rcp(so.D);
    H = {(fP.re - fQ.re) * rD, (fP.im - fQ.im) * rD};
    const cf pP = dphi_of(so.P, sP, cP, kP, rP);
    const cf pQ = dphi_of(so.Q, sQ, cQ, kQ, rQ);
    dP = {(pP.re + H.re) * rD, (pP.im + H.im) * rD};
    dQ = {-(pQ.re + H.re) * rD, -(pQ.im + H.im) * rD};
  }
}

}  // namespace dmic
