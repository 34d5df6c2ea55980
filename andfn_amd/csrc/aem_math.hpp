// Device math of the analytic-element evaluation (FP64 and FP32 forms).
//
// The element potentials are those of AnDFN's HPC path:
//   line source (intersection, constant-head line, impermeable line), andfn/hpc/hpc_intersection.py:115-145,
//     hpc_const_head_line.py:88-108, hpc_imp_object.py:186-205:
//       Z = (2z - a - b)/(b - a),  chi = Z + sqrt(Z - 1) sqrt(Z + 1)   (hpc_geometry_functions.py:11-31)
//       Omega = sum_j c_j chi^-j + q/(2 pi) log chi                    (hpc_math_functions.py:16-40, 193-213)
//   bounding circle: chi = z/R, Omega = sum_j c_j chi^j                (hpc_bounding_circle.py:52-71)
//   well: chi = (z - c)/r, Omega = q/(2 pi) log chi, NaN inside        (hpc_well.py:38-60)
//   impermeable circle: chi = (z - c)/r, Omega = sum_j c_j chi^-j, NaN inside (hpc_imp_object.py:67-88)
//
// Formulation (same values, fewer operations):
//   * Z = alpha z + beta with alpha = 2/(b - a), beta = -(a + b)/(b - a) precomputed per source;
//   * sqrt(Z - 1) sqrt(Z + 1) = +-sqrt((Z - 1)(Z + 1)), the sign chosen so that Re(conj(Z) s) >= 0,
//     which is the condition |chi| >= 1 of the exterior branch (|Z + s| >= |Z - s|);
//   * one reciprocal w = 1/chi, then Horner in multiply-adds: r <- r w + c_j;
//   * Re log chi = log|chi|^2 / 2, Im log chi = atan2 only where the imaginary part is used.
#pragma once
#include <hip/hip_runtime.h>

namespace aem {

constexpr double kPi = 3.141592653589793;
constexpr double kInv2Pi = 0.15915494309189535;  // 1/(2 pi)

// ---------------------------------------------------------------------------------------
// FP64
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double2 c2(double re, double im) { return make_double2(re, im); }

// 1/x for x > 0 (normal): hardware reciprocal refined by two Newton steps
__device__ __forceinline__ double rcp_pos(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// sqrt(x) for x > 0 (normal): hardware reciprocal square root refined by two Goldschmidt steps
// (g -> sqrt x, h -> 1/(2 sqrt x)) and a final residual correction; within ~1 ulp
__device__ __forceinline__ double sqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  return fma(d, h, g);
}

// log(x) for x > 0 (normal): x = m 2^e with m in [1/sqrt2, sqrt2), log m = 2 atanh(s),
// s = (m - 1)/(m + 1), |s| <= 0.172, by its odd series to s^21 (truncation < 1e-18 relative)
__device__ __forceinline__ double log_pos(double x) {
  int e;
  double m = frexp(x, &e);  // m in [0.5, 1)
  if (m < 0.7071067811865476) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) * rcp_pos(m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, s2, 1.0 / 19.0);
  p = fma(p, s2, 1.0 / 17.0);
  p = fma(p, s2, 1.0 / 15.0);
  p = fma(p, s2, 1.0 / 13.0);
  p = fma(p, s2, 1.0 / 11.0);
  p = fma(p, s2, 1.0 / 9.0);
  p = fma(p, s2, 1.0 / 7.0);
  p = fma(p, s2, 1.0 / 5.0);
  p = fma(p, s2, 1.0 / 3.0);
  const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
  const double de = (double)e;
  const double t = fma(2.0 * s * s2, p, de * ln2_lo);
  return fma(de, ln2_hi, fma(2.0, s, t));
}

// sqrt(x) for x > 0 (normal) as sqrt_pos, with h = 1 / (2 sqrt x) from the same iteration
__device__ __forceinline__ double sqrt_half_rcp_pos(double x, double& h) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  h = 0.5 * y;
  double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, x);
  return fma(d, h, g);
}

// principal square root (branch cut on the negative real axis)
__device__ __forceinline__ double2 csqrt_principal(double x, double y) {
  const double r2 = fma(x, x, y * y);
  if (r2 == 0.0) return c2(0.0, y);
  const double r = sqrt_pos(r2);
  const double t = sqrt_pos(0.5 * (r + fabs(x)));
  const double h = 0.5 * rcp_pos(t);
  if (x >= 0.0) return c2(t, y * h);
  return c2(fabs(y) * h, copysign(t, y));
}

// exterior Joukowsky inverse of Z (|chi| >= 1): chi = Z + s, s^2 = u = Z^2 - 1, the root with
// Re(conj(Z) s) >= 0.  The principal root of u is (t, uy / (2 t)) for ux >= 0 and
// (|uy| / (2 t), sign(uy) t) otherwise, t = sqrt((|u| + |ux|) / 2); 1/(2 t) comes out of the
// square-root iteration itself.
__device__ __forceinline__ double2 joukowsky_exterior(double2 Z) {
  const double ux = fma(-Z.y, Z.y, fma(Z.x, Z.x, -1.0));
  const double uy = 2.0 * Z.x * Z.y;
  const double r2 = fma(ux, ux, uy * uy);
  double2 s = c2(0.0, 0.0);
  if (r2 != 0.0) {
    double h;
    const double t = sqrt_half_rcp_pos(0.5 * (sqrt_pos(r2) + fabs(ux)), h);
    const double v = uy * h;
    s = ux >= 0.0 ? c2(t, v) : c2(fabs(v), copysign(t, uy));
    if (fma(Z.x, s.x, Z.y * s.y) < 0.0) s = c2(-s.x, -s.y);
  }
  return c2(Z.x + s.x, Z.y + s.y);
}

// r <- r * w + c
__device__ __forceinline__ double2 cmadd(double2 r, double2 w, double2 c) {
  return c2(fma(r.x, w.x, fma(-r.y, w.y, c.x)), fma(r.x, w.y, fma(r.y, w.x, c.y)));
}

// P independent Horner chains sharing every coefficient load (P target points per lane)
// r = sum_{j<m} c_j w^j for P target points (P per lane) by Horner; P independent chains share
// every coefficient load (wave-uniform, scalar loads).
template <int P>
__device__ __forceinline__ void horner_p(const double2* __restrict__ c, int m, const double2 (&w)[P],
                                         double2 (&r)[P]) {
#pragma unroll
  for (int k = 0; k < P; ++k) r[k] = c[m - 1];
  int j = m - 2;
  if (j >= 3) {
    // software-pipelined: the next four coefficients are requested before the current four are
    // used, so the scalar-load latency overlaps the multiply-adds (same arithmetic order)
    double2 n0 = c[j], n1 = c[j - 1], n2 = c[j - 2], n3 = c[j - 3];
    for (; j >= 7; j -= 4) {
      const double2 c0 = n0, c1 = n1, c2_ = n2, c3 = n3;
      n0 = c[j - 4];
      n1 = c[j - 5];
      n2 = c[j - 6];
      n3 = c[j - 7];
#pragma unroll
      for (int k = 0; k < P; ++k) {
        r[k] = cmadd(r[k], w[k], c0);
        r[k] = cmadd(r[k], w[k], c1);
        r[k] = cmadd(r[k], w[k], c2_);
        r[k] = cmadd(r[k], w[k], c3);
      }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      r[k] = cmadd(r[k], w[k], n0);
      r[k] = cmadd(r[k], w[k], n1);
      r[k] = cmadd(r[k], w[k], n2);
      r[k] = cmadd(r[k], w[k], n3);
    }
    j -= 4;
  }
  for (; j >= 0; --j) {
    const double2 cj = c[j];
#pragma unroll
    for (int k = 0; k < P; ++k) r[k] = cmadd(r[k], w[k], cj);
  }
}

// r = sum_{j<m} c_j w^j for coefficients with zero imaginary parts (intersections and
// constant-head lines: their fits keep Re only), P points per lane.  The polynomial has real
// coefficients, so it is reduced modulo the real quadratic (x - w)(x - conj w) = x^2 - s x + t,
// s = 2 Re w, t = |w|^2, whose roots are w and conj w (Knuth, TAOCP vol. 2, 4.6.4):
//   b_j = c_j + s b_{j+1} - t b_{j+2}  (j = m-1 .. 0, b_m = b_{m+1} = 0),  p(w) = b_0 - b_1 conj(w).
// Two real multiply-adds per coefficient instead of four, and half the coefficient bytes.
template <int P>
__device__ __forceinline__ void horner_real_p(const double2* __restrict__ c, int m, const double2 (&w)[P],
                                              double2 (&r)[P]) {
  const double* __restrict__ cr = reinterpret_cast<const double*>(c);  // Re c_j at cr[2 j]
  double s[P], t[P], b1[P], b2[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    s[k] = 2.0 * w[k].x;
    t[k] = fma(w[k].x, w[k].x, w[k].y * w[k].y);
    b1[k] = cr[2 * (m - 1)];
    b2[k] = 0.0;
  }
  int j = m - 2;
  if (j >= 3) {
    // the next four coefficients are requested before the current four are used
    double n0 = cr[2 * j], n1 = cr[2 * (j - 1)], n2 = cr[2 * (j - 2)], n3 = cr[2 * (j - 3)];
    for (; j >= 7; j -= 4) {
      const double c0 = n0, c1 = n1, c2_ = n2, c3 = n3;
      n0 = cr[2 * (j - 4)];
      n1 = cr[2 * (j - 5)];
      n2 = cr[2 * (j - 6)];
      n3 = cr[2 * (j - 7)];
#pragma unroll
      for (int k = 0; k < P; ++k) {
        double b = fma(s[k], b1[k], fma(-t[k], b2[k], c0));
        b2[k] = fma(s[k], b, fma(-t[k], b1[k], c1));
        b1[k] = fma(s[k], b2[k], fma(-t[k], b, c2_));
        b2[k] = fma(s[k], b1[k], fma(-t[k], b2[k], c3));
        // (b1, b2) now hold (b_{j-2}, b_{j-3}) in swapped roles: restore the order
        const double u = b1[k];
        b1[k] = b2[k];
        b2[k] = u;
      }
    }
#pragma unroll
    for (int k = 0; k < P; ++k) {
      double b = fma(s[k], b1[k], fma(-t[k], b2[k], n0));
      b2[k] = fma(s[k], b, fma(-t[k], b1[k], n1));
      b1[k] = fma(s[k], b2[k], fma(-t[k], b, n2));
      b2[k] = fma(s[k], b1[k], fma(-t[k], b2[k], n3));
      const double u = b1[k];
      b1[k] = b2[k];
      b2[k] = u;
    }
    j -= 4;
  }
  for (; j >= 0; --j) {
    const double cj = cr[2 * j];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const double b = fma(s[k], b1[k], fma(-t[k], b2[k], cj));
      b2[k] = b1[k];
      b1[k] = b;
    }
  }
  // Re w = s / 2 exactly, so w.x need not stay live through the recurrence
#pragma unroll
  for (int k = 0; k < P; ++k) r[k] = c2(fma(-b2[k], 0.5 * s[k], b1[k]), b2[k] * w[k].y);
}

// sum_{j<m} c_j w^j by Horner; c is wave-uniform (scalar loads).  Coefficients are fetched
// eight at a time ahead of their multiply-adds so the scalar-load latency is paid once per chunk.
__device__ __forceinline__ double2 horner(const double2* __restrict__ c, int m, double2 w) {
  double2 r = c[m - 1];
  int j = m - 2;
  for (; j >= 7; j -= 8) {
    const double2 c0 = c[j], c1 = c[j - 1], c2_ = c[j - 2], c3 = c[j - 3];
    const double2 c4 = c[j - 4], c5 = c[j - 5], c6 = c[j - 6], c7 = c[j - 7];
    r = cmadd(r, w, c0);
    r = cmadd(r, w, c1);
    r = cmadd(r, w, c2_);
    r = cmadd(r, w, c3);
    r = cmadd(r, w, c4);
    r = cmadd(r, w, c5);
    r = cmadd(r, w, c6);
    r = cmadd(r, w, c7);
  }
  for (; j >= 0; --j) r = cmadd(r, w, c[j]);
  return r;
}

// ---------------------------------------------------------------------------------------
// FP32 (mixed-precision far field)
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cmaddf(float2 r, float2 w, float2 c) {
  return make_float2(fmaf(r.x, w.x, fmaf(-r.y, w.y, c.x)), fmaf(r.x, w.y, fmaf(r.y, w.x, c.y)));
}

__device__ __forceinline__ float2 csqrt_principal_f(float x, float y) {
  float r = sqrtf(fmaf(x, x, y * y));
  if (r == 0.0f) return make_float2(0.0f, y);
  float t = sqrtf(0.5f * (r + fabsf(x)));
  if (x >= 0.0f) return make_float2(t, y / (2.0f * t));
  return make_float2(fabsf(y) / (2.0f * t), copysignf(t, y));
}

__device__ __forceinline__ float2 joukowsky_exterior_f(float2 Z) {
  float ux = fmaf(Z.x - 1.0f, Z.x + 1.0f, -Z.y * Z.y);
  float uy = 2.0f * Z.x * Z.y;
  float2 s = csqrt_principal_f(ux, uy);
  if (fmaf(Z.x, s.x, Z.y * s.y) < 0.0f) {
    s.x = -s.x;
    s.y = -s.y;
  }
  return make_float2(Z.x + s.x, Z.y + s.y);
}

// P FP32 Horner chains sharing the coefficient loads
template <int P>
__device__ __forceinline__ void horner_f_p(const float2* __restrict__ c, int m, const float2 (&w)[P],
                                           float2 (&r)[P]) {
#pragma unroll
  for (int k = 0; k < P; ++k) r[k] = c[m - 1];
  for (int j = m - 2; j >= 0; --j) {
    const float2 cj = c[j];
#pragma unroll
    for (int k = 0; k < P; ++k) r[k] = cmaddf(r[k], w[k], cj);
  }
}

// FP32 form of horner_real_p (real coefficients: the second-order recurrence)
template <int P>
__device__ __forceinline__ void horner_real_f_p(const float2* __restrict__ c, int m, const float2 (&w)[P],
                                                float2 (&r)[P]) {
  const float* __restrict__ cr = reinterpret_cast<const float*>(c);  // Re c_j at cr[2 j]
  float s[P], t[P], b1[P], b2[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    s[k] = 2.0f * w[k].x;
    t[k] = fmaf(w[k].x, w[k].x, w[k].y * w[k].y);
    b1[k] = cr[2 * (m - 1)];
    b2[k] = 0.0f;
  }
  for (int j = m - 2; j >= 0; --j) {
    const float cj = cr[2 * j];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const float b = fmaf(s[k], b1[k], fmaf(-t[k], b2[k], cj));
      b2[k] = b1[k];
      b1[k] = b;
    }
  }
#pragma unroll
  for (int k = 0; k < P; ++k) r[k] = make_float2(fmaf(-b2[k], 0.5f * s[k], b1[k]), b2[k] * w[k].y);
}

// FP32 Horner, coefficients fetched eight at a time (wave-uniform scalar loads)
__device__ __forceinline__ float2 horner_f(const float2* __restrict__ c, int m, float2 w) {
  float2 r = c[m - 1];
  int j = m - 2;
  for (; j >= 7; j -= 8) {
    const float2 c0 = c[j], c1 = c[j - 1], c2_ = c[j - 2], c3 = c[j - 3];
    const float2 c4 = c[j - 4], c5 = c[j - 5], c6 = c[j - 6], c7 = c[j - 7];
    r = cmaddf(r, w, c0);
    r = cmaddf(r, w, c1);
    r = cmaddf(r, w, c2_);
    r = cmaddf(r, w, c3);
    r = cmaddf(r, w, c4);
    r = cmaddf(r, w, c5);
    r = cmaddf(r, w, c6);
    r = cmaddf(r, w, c7);
  }
  for (; j >= 0; --j) r = cmaddf(r, w, c[j]);
  return r;
}

}  // namespace aem
