// meshFT CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// This file is the checker for the HIP product path: only tests/, the smoke
// check and the CPU-baseline leg of bench.py may load it.  The product path
// (deltamic_amd/) never links, loads or calls it.
//
// Two restatements of the reference's narrow-band mesh Fourier transform
// (deltamic/cpp/fourier3d_cuda_kernel.cu) live here:
//
//  1. "accurate" (orc_acc_*): the same mathematics evaluated to ~1e-12 in
//     FP64, with a centroid Taylor series when all three phases of a face
//     are close and an __float128 evaluation when one pair of phases is
//     nearly coincident.  Every error bound of the HIP path is stated against
//     this oracle.  It is pinned by tests/test_oracle.py against mpmath
//     quadrature of single triangles, the closed-form transform of a cube
//     surface, invariants (sum of corner derivatives = -iF, translation
//     phase, Hermitian symmetry, coefficient-scale invariance) and central
//     finite differences of the forward.
//
//  2. "reference semantics" (orc_ref_*): a line-by-line restatement of the
//     reference's branch structure (fourier_xi_tri_f, .cu:239-278, all
//     comparisons <= sqrt(eps)), of its per-frequency face loop
//     (Mesh_NUFT_kernel, .cu:282-311), of its per-(vertex, chunk) adjoint
//     with l = 256 chunks (Grad_Mesh_NUFT_kernel, .cu:315-397) and of the
//     fixed-order chunk sum (Additions_kernel, .cu:400-416), templated on
//     float / double like AT_DISPATCH_FLOATING_TYPES (.cu:447,508).  It is the
//     timed CPU baseline and the stand-in for the reference gradient.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp, links libquadmath).

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>
#include <omp.h>
#include <quadmath.h>

typedef std::complex<double> cd;

namespace {

// ---------------------------------------------------------------- helpers
inline cd expmi(double x) { return cd(std::cos(x), -std::sin(x)); }  // e^{-ix}

struct Quad { __float128 re, im; };
inline Quad qmul(Quad a, Quad b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Quad qadd(Quad a, Quad b) { return {a.re + b.re, a.im + b.im}; }
inline Quad qsub(Quad a, Quad b) { return {a.re - b.re, a.im - b.im}; }
inline Quad qscale(Quad a, __float128 s) { return {a.re * s, a.im * s}; }
inline Quad qexpmi(__float128 x) { return {cosq(x), -sinq(x)}; }
inline cd q2d(Quad a) { return cd((double)a.re, (double)a.im); }

// Divided difference of E(y) = exp(-i y) over the nodes y[0..n-1]
// (repeated nodes allowed), by its Taylor series
//     E[y_0..y_{n-1}] = sum_{k >= n-1} (-i)^k / k! * h_{k-n+1}(y)
// where h_j is the complete homogeneous symmetric polynomial of degree j,
// generated from the elementary symmetric polynomials e_i by Newton's
// recursion h_j = sum_i (-1)^{i+1} e_i h_{j-i}.  Used only when every
// |y_i| <= 1, where the series converges to double precision in < 30 terms.
cd dd_exp_taylor(const double* y, int n) {
  double e[6] = {1, 0, 0, 0, 0, 0};  // elementary symmetric polys, n <= 5
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j >= 1; --j) e[j] += e[j - 1] * y[i];
  double h[64];
  h[0] = 1.0;
  const int kmax = 26;  // |y_i| <= 2/3: the first omitted term is < 1e-24
  cd sum = 0.0;
  // coefficient (-i)^k / k!
  double fact = 1.0;
  for (int k = 1; k <= n - 1; ++k) fact *= k;
  for (int j = 0; j + n - 1 <= kmax; ++j) {
    if (j > 0) {
      double hj = 0.0;
      for (int i = 1; i <= std::min(j, n); ++i)
        hj += ((i & 1) ? 1.0 : -1.0) * e[i] * h[j - i];
      h[j] = hj;
    }
    int k = j + n - 1;
    if (j > 0) fact *= k;
    double c = h[j] / fact;
    switch (k & 3) {  // (-i)^k
      case 0: sum += cd(c, 0); break;
      case 1: sum += cd(0, -c); break;
      case 2: sum += cd(-c, 0); break;
      case 3: sum += cd(0, c); break;
    }
  }
  return sum;
}

// Accurate F = f[xa,xb,xc] and its corner derivatives dF/dx_j for
// f(x) = -exp(-ix): the face integral of fourier_xi_tri_f (.cu:239-278)
// and grad_fourier_xi_tri_{a,b,c}_f (.cu:74-233) without the FP32/FP64
// cancellation of their generic branch.
void face_accurate(double xa, double xb, double xc, cd* F, cd* dF /*[3] or null*/) {
  const double dab = xb - xa, dbc = xc - xb, dca = xa - xc;
  const double aab = std::fabs(dab), abc = std::fabs(dbc), aca = std::fabs(dca);
  const double mx = std::max(aab, std::max(abc, aca));
  const double mn = std::min(aab, std::min(abc, aca));
  if (mx < 1.0) {
    // all three phases within 1 rad: Taylor series about the centroid
    const double m = (xa + xb + xc) / 3.0;
    const cd em = expmi(m);
    double y[4] = {xa - m, xb - m, xc - m, 0.0};
    *F = -em * dd_exp_taylor(y, 3);
    if (dF) {
      for (int j = 0; j < 3; ++j) {
        double y4[4] = {y[0], y[1], y[2], y[j]};
        dF[j] = -em * dd_exp_taylor(y4, 4);
      }
    }
    return;
  }
  if (mn >= 1e-3) {
    // generic closed form (.cu:274 and .cu:122,176,230) in double
    const cd ea = expmi(xa), eb = expmi(xb), ec = expmi(xc);
    *F = ea / (dab * dca) + eb / (dbc * dab) + ec / (dca * dbc);
    if (dF) {
      const double rab = 1.0 / dab, rbc = 1.0 / dbc, rca = 1.0 / dca;
      const cd I(0, 1);
      dF[0] = (rab - rca - I) * (rca * rab) * ea + rab * rab * rbc * eb - rca * rca * rbc * ec;
      dF[1] = (rbc - rab - I) * (rab * rbc) * eb + rbc * rbc * rca * ec - rab * rab * rca * ea;
      dF[2] = (rca - rbc - I) * (rbc * rca) * ec + rca * rca * rab * ea - rbc * rbc * rab * eb;
    }
    return;
  }
  // One pair of phases within 1e-3 rad, the largest gap D >= 1 rad.  The
  // closed form cancels like eps/d^2 here, so evaluate in __float128 with
  // the phases re-based on the median vertex (phases lo <= mid <= hi,
  // P = lo - mid, Q = hi - mid, D = Q - P):
  //   F = e^{-i mid} (phi(P) - phi(Q)) / D,  phi(y) = (e^{-iy} - 1)/y,
  //   dF/dlo = e^{-i mid} (phi'(P) + H)/D,  dF/dhi = -e^{-i mid} (phi'(Q) + H)/D,
  //   dF/dmid = -iF - dF/dlo - dF/dhi,
  // where phi(y) = -(sin(y/2)/(y/2)) (sin(y/2) + i cos(y/2)) has no
  // cancellation and phi'(y) uses its Taylor series for |y| < 1/4.
  // Pinned by the quadrature tests (nearly coincident pairs).
  const __float128 xs[3] = {(__float128)xa, (__float128)xb, (__float128)xc};
  int ord[3] = {0, 1, 2};
  std::sort(ord, ord + 3, [&](int i, int j) { return xs[i] < xs[j] || (xs[i] == xs[j] && i < j); });
  const __float128 mid = xs[ord[1]];
  const __float128 P = xs[ord[0]] - mid, Q = xs[ord[2]] - mid, D = Q - P;
  auto phi = [](__float128 y) -> Quad {
    const __float128 h = y / 2, sh = sinq(h), ch = cosq(h);
    const __float128 k = (h == 0) ? (__float128)1 : sh / h;
    return {-k * sh, -k * ch};
  };
  auto dphi = [&](__float128 y) -> Quad {
    if (fabsq(y) >= (__float128)0.25) {
      const Quad e = qexpmi(y);          // e^{-iy}
      const Quad mie = {e.im, -e.re};    // -i e^{-iy}
      return qscale(qsub(mie, phi(y)), 1 / y);
    }
    // sum_{k>=2} (-i)^k (k-1) y^{k-2} / k!
    Quad sum = {0, 0};
    __float128 yp = 1, fact = 2;
    for (int k = 2; k <= 34; ++k) {
      if (k > 2) { yp *= y; fact *= k; }
      const __float128 c = (k - 1) * yp / fact;
      switch (k & 3) {
        case 0: sum.re += c; break;
        case 1: sum.im -= c; break;
        case 2: sum.re -= c; break;
        case 3: sum.im += c; break;
      }
    }
    return sum;
  };
  const Quad em = qexpmi(mid);
  const Quad H = qscale(qsub(phi(P), phi(Q)), 1 / D);
  const Quad Fq = qmul(em, H);
  *F = q2d(Fq);
  if (dF) {
    const Quad dlo = qmul(em, qscale(qadd(dphi(P), H), 1 / D));
    const Quad dhi = qmul(em, qscale(qadd(dphi(Q), H), -1 / D));
    const Quad miF = {Fq.im, -Fq.re};
    const Quad dmid = qsub(qsub(miF, dlo), dhi);
    dF[ord[0]] = q2d(dlo);
    dF[ord[1]] = q2d(dmid);
    dF[ord[2]] = q2d(dhi);
  }
}

// ---------------------------------------------------------------- reference semantics
// Restatement of fourier_xi_tri_f (.cu:239-278) and grad_fourier_xi_tri_*
// (.cu:74-233) on scalar type T, comparisons <= Seps as in the reference.
template <typename T> inline std::complex<T> fexpmit(T z) {  // f_expmit_f, .cu:61-64
  return std::complex<T>(std::cos(z), -std::sin(z));
}

template <typename T>
inline std::complex<T> ref_F(T xia, T xib, T xic, std::complex<T> ea, std::complex<T> eb,
                             std::complex<T> ec, T Seps) {
  const T xiab = xib - xia, xibc = xic - xib, xica = xia - xic;
  const std::complex<T> I(0, 1);
  const bool sab = std::fabs(xiab) <= Seps, sbc = std::fabs(xibc) <= Seps, sca = std::fabs(xica) <= Seps;
  if ((sab && sbc) || (sbc && sca) || (sca && sab)) {
    T xid = (xia + xib + xic) / (T)3.0;
    return fexpmit(xid) / (T)2.0;
  } else if (sab) {
    T xid = (xia + xib) / (T)2.0;
    std::complex<T> em = fexpmit(xid);
    xid -= xic;
    return (I * em + (em - ec) / xid) / xid;
  } else if (sbc) {
    T xid = (xib + xic) / (T)2.0;
    std::complex<T> em = fexpmit(xid);
    xid -= xia;
    return (I * em + (em - ea) / xid) / xid;
  } else if (sca) {
    T xid = (xic + xia) / (T)2.0;
    std::complex<T> em = fexpmit(xid);
    xid -= xib;
    return (I * em + (em - eb) / xid) / xid;
  }
  return ea / (xiab * xica) + eb / (xibc * xiab) + ec / (xica * xibc);
}

// corner = 0,1,2 -> grad_fourier_xi_tri_{a,b,c}_f
template <typename T>
inline std::complex<T> ref_dF(int corner, T xia, T xib, T xic, std::complex<T> ea,
                              std::complex<T> eb, std::complex<T> ec, T Seps) {
  T xiab = xib - xia, xibc = xic - xib, xica = xia - xic;
  const std::complex<T> I(0, 1), mI(0, -1);
  const bool sab = std::fabs(xiab) <= Seps, sbc = std::fabs(xibc) <= Seps, sca = std::fabs(xica) <= Seps;
  if ((sab && sbc) || (sbc && sca) || (sca && sab)) {
    T xid = (xia + xib + xic) / (T)3.0;
    std::complex<T> ft = fexpmit(xid) / (T)2.0;
    return mI * ft / (T)3.0;
  }
  int odd;
  T xid;
  std::complex<T> eo;
  if (sab) { xid = (xia + xib) / (T)2.0; eo = ec; odd = 2; }
  else if (sbc) { xid = (xib + xic) / (T)2.0; eo = ea; odd = 0; }
  else if (sca) { xid = (xic + xia) / (T)2.0; eo = eb; odd = 1; }
  else {
    xiab = (T)1.0 / xiab; xibc = (T)1.0 / xibc; xica = (T)1.0 / xica;
    if (corner == 0)
      return (xiab - xica - I) * ((xica * xiab) * ea) + xiab * ((xiab * xibc) * eb) - xica * ((xibc * xica) * ec);
    if (corner == 1)
      return (xibc - xiab - I) * ((xiab * xibc) * eb) + xibc * ((xibc * xica) * ec) - xiab * ((xica * xiab) * ea);
    return (xica - xibc - I) * ((xibc * xica) * ec) + xica * ((xica * xiab) * ea) - xibc * ((xiab * xibc) * eb);
  }
  std::complex<T> em = fexpmit(xid);
  const T xo = (odd == 0) ? xia : (odd == 1 ? xib : xic);
  xid -= xo;
  std::complex<T> ft = (I * em + (em - eo) / xid) / xid;
  std::complex<T> g = (em / (T)2.0 - ft) / xid;
  return (corner == odd) ? (mI * ft - (T)2.0 * g) : g;
}

template <typename T> inline T dot3(const T* v, T x, T y, T z) { return v[0] * x + v[1] * y + v[2] * z; }

// trisar and meshar as computed on the host by the reference
// (get_internal_triangle_and_mesh_areas, .cu:38-51; trisar *= 2/meshar, .cu:439)
template <typename T>
void ref_trisar(const T* V, const int64_t* Fc, int64_t nf, const T* coeff, std::vector<T>& trisar) {
  trisar.assign(nf, (T)0);
  for (int64_t k = 0; k < nf; ++k) {
    const T* a = V + 3 * Fc[3 * k]; const T* b = V + 3 * Fc[3 * k + 1]; const T* c = V + 3 * Fc[3 * k + 2];
    T u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, w[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    T n[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
    trisar[k] += (T)0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) * coeff[k];
  }
  T meshar = 0;
  for (int64_t k = 0; k < nf; ++k) meshar += trisar[k];
  const T s = (T)2 / meshar;
  for (int64_t k = 0; k < nf; ++k) trisar[k] *= s;
}

template <typename T>
void ref_forward(const T* V, int64_t nv, const int64_t* Fc, int64_t nf, const T* coeff, const T* xi,
                 int64_t n, T* out, int nthreads) {
  (void)nv;
  std::vector<T> trisar;
  ref_trisar(V, Fc, nf, coeff, trisar);
  const T Seps = std::sqrt(std::numeric_limits<T>::epsilon());  // .cu:454
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t m = 0; m < n; ++m) {
    const T x0 = xi[3 * m], x1 = xi[3 * m + 1], x2 = xi[3 * m + 2];
    std::complex<T> acc(0, 0);
    for (int64_t kt = 0; kt < nf; ++kt) {  // .cu:304-309
      const T xa = dot3(V + 3 * Fc[3 * kt], x0, x1, x2);
      const T xb = dot3(V + 3 * Fc[3 * kt + 1], x0, x1, x2);
      const T xc = dot3(V + 3 * Fc[3 * kt + 2], x0, x1, x2);
      acc += trisar[kt] * ref_F<T>(xa, xb, xc, fexpmit(xa), fexpmit(xb), fexpmit(xc), Seps);
    }
    out[2 * m] = acc.real();
    out[2 * m + 1] = acc.imag();
  }
}

template <typename T>
void ref_backward(const T* V, int64_t nv, const int64_t* Fc, int64_t nf, const T* coeff, const T* xi,
                  int64_t n, const T* grad, const T* ftmesh, T* grad_verts, int nthreads) {
  // host precompute of get_internal_triangle_and_mesh_areas_and_edgenormals (.cu:16-36)
  std::vector<T> trisar(nf, 0), w(9 * nf, 0), gma(3 * nv, 0);
  for (int64_t kt = 0; kt < nf; ++kt) {
    const T* a = V + 3 * Fc[3 * kt]; const T* b = V + 3 * Fc[3 * kt + 1]; const T* c = V + 3 * Fc[3 * kt + 2];
    T u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, q[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    T nr[3] = {u[1] * q[2] - u[2] * q[1], u[2] * q[0] - u[0] * q[2], u[0] * q[1] - u[1] * q[0]};
    T nn = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    trisar[kt] = nn;
    for (int d = 0; d < 3; ++d) nr[d] /= trisar[kt];
    trisar[kt] *= (T)0.5 * coeff[kt];
    const T* P[3] = {a, b, c};
    for (int j = 0; j < 3; ++j) {
      const T* s = P[(j + 2) % 3]; const T* r = P[(j + 1) % 3];
      T e[3] = {s[0] - r[0], s[1] - r[1], s[2] - r[2]};
      T en[3] = {e[1] * nr[2] - e[2] * nr[1], e[2] * nr[0] - e[0] * nr[2], e[0] * nr[1] - e[1] * nr[0]};
      for (int d = 0; d < 3; ++d) {
        w[9 * kt + 3 * j + d] = en[d] * coeff[kt];
        gma[3 * Fc[3 * kt + j] + d] -= w[9 * kt + 3 * j + d] / 2;
      }
    }
  }
  T meshar = 0;
  for (int64_t k = 0; k < nf; ++k) meshar += trisar[k];
  for (int64_t k = 0; k < nf; ++k) trisar[k] *= (T)2 / meshar;   // .cu:483
  for (auto& g : gma) g *= (T)1 / meshar;                          // .cu:484
  for (auto& x : w) x *= (T)1 / meshar;                            // .cu:485
  // vertex -> (face, local index) map (.cu:487-501), without the 30 cap
  std::vector<std::vector<std::pair<int64_t, int>>> vmap(nv);
  for (int64_t i = 0; i < nf; ++i)
    for (int j = 0; j < 3; ++j) vmap[Fc[3 * i + j]].push_back({i, j});
  const int64_t l = 256;  // .cu:470
  const T Seps = std::sqrt(std::numeric_limits<T>::epsilon());
  std::vector<T> big(3 * nv * l, 0);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int64_t idx = 0; idx < nv * l; ++idx) {  // .cu:333-396
    const int64_t kv = idx / l, r = idx % l;
    const int64_t m0 = (int64_t)((T)r * (T)n / (T)l), m1 = (int64_t)(((T)r + (T)1) * (T)n / (T)l);
    T acc[3] = {0, 0, 0};
    for (int64_t m = m0; m < m1; ++m) {
      const T x0 = xi[3 * m], x1 = xi[3 * m + 1], x2 = xi[3 * m + 2];
      std::complex<T> g0(0, 0), g1(0, 0), g2(0, 0);
      for (auto& pr : vmap[kv]) {
        const int64_t t = pr.first; const int j = pr.second;
        const T xa = dot3(V + 3 * Fc[3 * t], x0, x1, x2);
        const T xb = dot3(V + 3 * Fc[3 * t + 1], x0, x1, x2);
        const T xc = dot3(V + 3 * Fc[3 * t + 2], x0, x1, x2);
        const std::complex<T> ea = fexpmit(xa), eb = fexpmit(xb), ec = fexpmit(xc);
        const std::complex<T> gt = ref_dF<T>(j, xa, xb, xc, ea, eb, ec, Seps);
        const std::complex<T> ft = ref_F<T>(xa, xb, xc, ea, eb, ec, Seps);
        g0 += gt * trisar[t] * x0 - ft * w[9 * t + 3 * j + 0];
        g1 += gt * trisar[t] * x1 - ft * w[9 * t + 3 * j + 1];
        g2 += gt * trisar[t] * x2 - ft * w[9 * t + 3 * j + 2];
      }
      const std::complex<T> fm(ftmesh[2 * m], ftmesh[2 * m + 1]);
      g0 -= fm * gma[3 * kv]; g1 -= fm * gma[3 * kv + 1]; g2 -= fm * gma[3 * kv + 2];
      const T gr = grad[2 * m], gi = grad[2 * m + 1];
      acc[0] += g0.real() * gr + g0.imag() * gi;
      acc[1] += g1.real() * gr + g1.imag() * gi;
      acc[2] += g2.real() * gr + g2.imag() * gi;
    }
    for (int d = 0; d < 3; ++d) big[(kv * l + r) * 3 + d] = acc[d];
  }
  for (int64_t v = 0; v < nv; ++v) {  // Additions_kernel, .cu:409-414
    T s[3] = {0, 0, 0};
    for (int64_t r = 0; r < l; ++r)
      for (int d = 0; d < 3; ++d) s[d] += big[(v * l + r) * 3 + d];
    for (int d = 0; d < 3; ++d) grad_verts[3 * v + d] = s[d];
  }
}

// ---------------------------------------------------------------- accurate mesh-level oracle
struct Geo {
  std::vector<double> w;       // 2 c_t A_t / S         (trisar after .cu:439)
  std::vector<double> cA;      // c_t A_t
  std::vector<double> gradA;   // [nf][3][3] dA_t/dv_j
  double S;
};

void geometry(const double* V, const int64_t* Fc, int64_t nf, const double* coeff, Geo& g) {
  g.w.resize(nf); g.cA.resize(nf); g.gradA.assign(9 * nf, 0.0);
  long double S = 0;
  for (int64_t t = 0; t < nf; ++t) {
    const double* P[3] = {V + 3 * Fc[3 * t], V + 3 * Fc[3 * t + 1], V + 3 * Fc[3 * t + 2]};
    double u[3], q[3];
    for (int d = 0; d < 3; ++d) { u[d] = P[1][d] - P[0][d]; q[d] = P[2][d] - P[0][d]; }
    double nr[3] = {u[1] * q[2] - u[2] * q[1], u[2] * q[0] - u[0] * q[2], u[0] * q[1] - u[1] * q[0]};
    double nn = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
    double A = 0.5 * nn;
    g.cA[t] = coeff[t] * A;
    S += g.cA[t];
    if (nn > 0) {
      for (int d = 0; d < 3; ++d) nr[d] /= nn;
      for (int j = 0; j < 3; ++j) {  // dA/dv_j = 1/2 n x (v_{j+2} - v_{j+1})
        const double* s = P[(j + 2) % 3]; const double* r = P[(j + 1) % 3];
        double e[3] = {s[0] - r[0], s[1] - r[1], s[2] - r[2]};
        g.gradA[9 * t + 3 * j + 0] = 0.5 * (nr[1] * e[2] - nr[2] * e[1]);
        g.gradA[9 * t + 3 * j + 1] = 0.5 * (nr[2] * e[0] - nr[0] * e[2]);
        g.gradA[9 * t + 3 * j + 2] = 0.5 * (nr[0] * e[1] - nr[1] * e[0]);
      }
    }
  }
  g.S = (double)S;
  for (int64_t t = 0; t < nf; ++t) g.w[t] = 2.0 * g.cA[t] / g.S;
}

inline double phase(const double* v, const double* x) {
  // products of FP32-representable values are exact in FP64; fma keeps the
  // sum to one rounding per term
  return std::fma(v[2], x[2], std::fma(v[1], x[1], v[0] * x[0]));
}

}  // namespace

extern "C" {

// Single-face values for the pinning tests: out = [F, dFa, dFb, dFc] as (re, im).
void orc_face_accurate(double xa, double xb, double xc, double* out) {
  cd F, dF[3];
  face_accurate(xa, xb, xc, &F, dF);
  out[0] = F.real(); out[1] = F.imag();
  for (int j = 0; j < 3; ++j) { out[2 + 2 * j] = dF[j].real(); out[3 + 2 * j] = dF[j].imag(); }
}

void orc_face_ref_f64(double xa, double xb, double xc, double* out) {
  const double Seps = std::sqrt(std::numeric_limits<double>::epsilon());
  auto ea = fexpmit(xa), eb = fexpmit(xb), ec = fexpmit(xc);
  auto F = ref_F<double>(xa, xb, xc, ea, eb, ec, Seps);
  out[0] = F.real(); out[1] = F.imag();
  for (int j = 0; j < 3; ++j) {
    auto d = ref_dF<double>(j, xa, xb, xc, ea, eb, ec, Seps);
    out[2 + 2 * j] = d.real(); out[3 + 2 * j] = d.imag();
  }
}

void orc_face_ref_f32(float xa, float xb, float xc, float* out) {
  const float Seps = std::sqrt(std::numeric_limits<float>::epsilon());
  auto ea = fexpmit(xa), eb = fexpmit(xb), ec = fexpmit(xc);
  auto F = ref_F<float>(xa, xb, xc, ea, eb, ec, Seps);
  out[0] = F.real(); out[1] = F.imag();
  for (int j = 0; j < 3; ++j) {
    auto d = ref_dF<float>(j, xa, xb, xc, ea, eb, ec, Seps);
    out[2 + 2 * j] = d.real(); out[3 + 2 * j] = d.imag();
  }
}

// Accurate forward: out[m] = sum_t w_t F_t(xi_m), w_t = 2 c_t A_t / sum c A.
// Inputs are the FP32 values the HIP path receives, widened exactly to FP64.
int orc_acc_forward(const double* V, int64_t nv, const int64_t* Fc, int64_t nf, const double* coeff,
                    const double* xi, int64_t n, double* out, int nthreads) {
  (void)nv;
  Geo g;
  geometry(V, Fc, nf, coeff, g);
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads)
  for (int64_t m = 0; m < n; ++m) {
    const double* x = xi + 3 * m;
    long double sr = 0, si = 0;
    for (int64_t t = 0; t < nf; ++t) {
      cd F;
      face_accurate(phase(V + 3 * Fc[3 * t], x), phase(V + 3 * Fc[3 * t + 1], x),
                    phase(V + 3 * Fc[3 * t + 2], x), &F, nullptr);
      sr += g.w[t] * F.real();
      si += g.w[t] * F.imag();
    }
    out[2 * m] = (double)sr;
    out[2 * m + 1] = (double)si;
  }
  return 0;
}

// Accurate autofluorescence variant (fourier_from_mesh_narrowband_autofluorescence,
// deltamic/render.py:228-318): face weight w_t + i (xi . nor2_t) / |xi|^2,
// nor2_t = nor_t regint / regvol, nor_t = a x b + b x c + c x a,
// regvol = -sum_t a . (b x c) / 6, w_t = 2 c_t A_t / S (the reference's
// coefficient-free version is coeff = 1).  At xi = 0 the interior term is its
// limit regint * (sum_t a . (b x c) / 6) / regvol = -regint.  Its adjoint is
// checked against central differences of this function.
int orc_acc_forward_af(const double* V, int64_t nv, const int64_t* Fc, int64_t nf,
                       const double* coeff, double regint, const double* xi, int64_t n,
                       double* out, int nthreads) {
  (void)nv;
  Geo g;
  geometry(V, Fc, nf, coeff, g);
  std::vector<double> nor(3 * nf);
  long double vol = 0;
  for (int64_t t = 0; t < nf; ++t) {
    const double* a = V + 3 * Fc[3 * t];
    const double* b = V + 3 * Fc[3 * t + 1];
    const double* c = V + 3 * Fc[3 * t + 2];
    const double ab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double bc[3] = {b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2], b[0] * c[1] - b[1] * c[0]};
    const double ca[3] = {c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2], c[0] * a[1] - c[1] * a[0]};
    for (int d = 0; d < 3; ++d) nor[3 * t + d] = ab[d] + bc[d] + ca[d];
    vol += (long double)a[0] * bc[0] + (long double)a[1] * bc[1] + (long double)a[2] * bc[2];
  }
  const double regvol = (double)(-vol / 6);
  const double sc = (regvol != 0.0) ? regint / regvol : 0.0;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads)
  for (int64_t m = 0; m < n; ++m) {
    const double* x = xi + 3 * m;
    const long double q = (long double)x[0] * x[0] + (long double)x[1] * x[1] + (long double)x[2] * x[2];
    long double sr = 0, si = 0, tr = 0, ti = 0;
    for (int64_t t = 0; t < nf; ++t) {
      cd F;
      face_accurate(phase(V + 3 * Fc[3 * t], x), phase(V + 3 * Fc[3 * t + 1], x),
                    phase(V + 3 * Fc[3 * t + 2], x), &F, nullptr);
      sr += g.w[t] * F.real();
      si += g.w[t] * F.imag();
      const long double s = sc * ((long double)x[0] * nor[3 * t] + (long double)x[1] * nor[3 * t + 1] +
                                  (long double)x[2] * nor[3 * t + 2]);
      tr += s * F.real();
      ti += s * F.imag();
    }
    if (q > 0) {
      sr -= ti / q;  // i (tr + i ti) / q
      si += tr / q;
    } else if (regvol != 0.0) {
      sr -= regint;
    }
    out[2 * m] = (double)sr;
    out[2 * m + 1] = (double)si;
  }
  return 0;
}

// Accurate adjoint: grad_verts = d/dV sum_m Re(conj(g_m) ftmesh_m), the
// gradient convention of Grad_Mesh_NUFT_kernel (.cu:392-394).  Organised per
// face: P_tj = sum_m Re(conj g dF_j) xi_m, R_t = sum_m Re(conj g F), then
// dL/dv = sum_{(t,j) ni v} [w_t P_tj + (2 c_t / S) R_t dA_t/dv_j]
//         - (K / S) sum_{(t,j) ni v} c_t dA_t/dv_j,  K = sum_t w_t R_t,
// which is the same sum as .cu:365-390 (w[t][j] = -2 c_t dA_t/dv_j / S,
// gradmeshar[v] = sum c_t dA_t/dv / S).  Optional out_face receives the
// per-face partials [nf][10] = (P_a, P_b, T, R) with T = sum Re(conj g (-iF)) xi.
int orc_acc_backward(const double* V, int64_t nv, const int64_t* Fc, int64_t nf, const double* coeff,
                     const double* xi, int64_t n, const double* grad, double* grad_verts,
                     double* out_face, int nthreads) {
  Geo g;
  geometry(V, Fc, nf, coeff, g);
  std::vector<double> part(10 * nf, 0.0);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int64_t t = 0; t < nf; ++t) {
    long double acc[10] = {0};
    const double* P[3] = {V + 3 * Fc[3 * t], V + 3 * Fc[3 * t + 1], V + 3 * Fc[3 * t + 2]};
    for (int64_t m = 0; m < n; ++m) {
      const double* x = xi + 3 * m;
      cd F, dF[3];
      face_accurate(phase(P[0], x), phase(P[1], x), phase(P[2], x), &F, dF);
      const double gr = grad[2 * m], gi = grad[2 * m + 1];
      const double al = gr * dF[0].real() + gi * dF[0].imag();
      const double be = gr * dF[1].real() + gi * dF[1].imag();
      const double ga = gr * F.imag() - gi * F.real();  // Re(conj g (-i F))
      const double rh = gr * F.real() + gi * F.imag();
      for (int d = 0; d < 3; ++d) {
        acc[d] += al * x[d];
        acc[3 + d] += be * x[d];
        acc[6 + d] += ga * x[d];
      }
      acc[9] += rh;
    }
    for (int k = 0; k < 10; ++k) part[10 * t + k] = (double)acc[k];
  }
  if (out_face) std::memcpy(out_face, part.data(), sizeof(double) * 10 * nf);
  long double K = 0;
  for (int64_t t = 0; t < nf; ++t) K += g.w[t] * part[10 * t + 9];
  std::vector<long double> gv(3 * nv, 0.0L);
  for (int64_t t = 0; t < nf; ++t) {
    const double* pt = &part[10 * t];
    double Pj[3][3];
    for (int d = 0; d < 3; ++d) {
      Pj[0][d] = pt[d];
      Pj[1][d] = pt[3 + d];
      Pj[2][d] = pt[6 + d] - pt[d] - pt[3 + d];
    }
    const double cS = (coeff[t] / g.S);
    for (int j = 0; j < 3; ++j) {
      const int64_t v = Fc[3 * t + j];
      for (int d = 0; d < 3; ++d) {
        const double gA = g.gradA[9 * t + 3 * j + d];
        gv[3 * v + d] += g.w[t] * Pj[j][d] + 2.0 * cS * pt[9] * gA - (K / g.S) * coeff[t] * gA;
      }
    }
  }
  for (int64_t i = 0; i < 3 * nv; ++i) grad_verts[i] = (double)gv[i];
  return 0;
}

int orc_ref_forward_f32(const float* V, int64_t nv, const int64_t* Fc, int64_t nf, const float* coeff,
                        const float* xi, int64_t n, float* out, int nthreads) {
  ref_forward<float>(V, nv, Fc, nf, coeff, xi, n, out, nthreads);
  return 0;
}
int orc_ref_forward_f64(const double* V, int64_t nv, const int64_t* Fc, int64_t nf, const double* coeff,
                        const double* xi, int64_t n, double* out, int nthreads) {
  ref_forward<double>(V, nv, Fc, nf, coeff, xi, n, out, nthreads);
  return 0;
}
int orc_ref_backward_f32(const float* V, int64_t nv, const int64_t* Fc, int64_t nf, const float* coeff,
                         const float* xi, int64_t n, const float* grad, const float* ftmesh,
                         float* grad_verts, int nthreads) {
  ref_backward<float>(V, nv, Fc, nf, coeff, xi, n, grad, ftmesh, grad_verts, nthreads);
  return 0;
}
int orc_ref_backward_f64(const double* V, int64_t nv, const int64_t* Fc, int64_t nf, const double* coeff,
                         const double* xi, int64_t n, const double* grad, const double* ftmesh,
                         double* grad_verts, int nthreads) {
  ref_backward<double>(V, nv, Fc, nf, coeff, xi, n, grad, ftmesh, grad_verts, nthreads);
  return 0;
}

}  // extern "C"
