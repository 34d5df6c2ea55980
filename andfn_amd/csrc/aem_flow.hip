// Flow kernels of a solved network (gfx950): the complex discharge vector w = -dOmega/dz at
// points and the tracing of streamlines through one fracture.
//
//   w_points_kernel  w of fracture f at explicit points, one 64-point tile of one fracture per
//                    wavefront (the tiles of the potential sweep): two forms,
//                    jit  = hpc_fracture.calc_w (hpc_fracture.py:139-174) over the jitted element
//                           functions (hpc_intersection.py:179-222, hpc_bounding_circle.py:97-115,
//                           hpc_well.py:89-108, hpc_const_head_line.py:135-164); impermeable
//                           objects are not summed there and the bounding circle enters as
//                           taylor_series_d1(z / r) alone;
//                    !jit = the object model's Fracture.calc_w (fracture.py:453-477) that the
//                           streamline tracking uses: every element type, the bounding circle as
//                           -taylor_series_d1(z / r) / r (bounding.py:91-108), NaN inside wells and
//                           impermeable circles (well.py:129-153, impermeable_object.py:131-224).
//   trace_kernel     one lane per streamline: the inner loop of DFN.streamline_tracking
//                    (dfn.py:2663-2705) inside the line's current fracture, i.e. predictor-
//                    corrector steps (runge_kutta, dfn.py:2779-2834) until the step crosses a
//                    discharge element (check_streamline_exit, dfn.py:2727-2742, with
//                    intersection.py:227-271, const_head.py:208-246, well.py:162-200), w turns NaN
//                    or the point budget is spent.  The points and the speeds |w| / aperture
//                    (fracture.py:479-497) go to a per-line row of `capacity` entries.  The switch
//                    to the neighbouring fracture (get_exit_intersection, dfn.py:2744-2777) is done
//                    by the caller between launches.
//
// Every derivative series is summed by Horner in the reciprocal variable: for a line or an
// impermeable circle, asym_expansion_d1(chi) = -u^2 S(u) with u = 1/chi and
// S(u) = sum_{j>=1} j c_j u^(j-1); for the bounding circle taylor_series_d1(chi) = S(chi)
// (hpc_math_functions.py:75-100, 163-190).  The line map's derivative
// dchi/dz = 2 chi^2 / (chi^2 - 1) * 2 / (b - a) is 2 alpha / (1 - u^2) with alpha = 2 / (b - a).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aem_kernels.hpp"
#include "aem_math.hpp"

namespace aem {

namespace {

constexpr double kQ2Pi = 0.15915494309189535;  // 1 / (2 pi)

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return c2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return c2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return c2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) { return c2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cconj(double2 a) { return c2(a.x, -a.y); }
__device__ __forceinline__ double cabs_(double2 a) { return hypot(a.x, a.y); }
// a / b by Smith's method (the scaling numpy's complex division uses)
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    if (b.x == 0.0 && b.y == 0.0) return c2(a.x / fabs(b.x), a.y / fabs(b.y));
    const double rat = b.y / b.x, scl = 1.0 / (b.x + b.y * rat);
    return c2((a.x + a.y * rat) * scl, (a.y - a.x * rat) * scl);
  }
  const double rat = b.x / b.y, scl = 1.0 / (b.y + b.x * rat);
  return c2((a.x * rat + a.y) * scl, (a.y * rat - a.x) * scl);
}
__device__ __forceinline__ double2 crcp(double2 b) { return cdiv(c2(1.0, 0.0), b); }
__device__ __forceinline__ bool cnan(double2 a) { return isnan(a.x); }

// S(u) = sum_{j=1}^{m-1} j c_j u^(j-1), 0 for m < 2
__device__ __forceinline__ double2 dseries(const double2* __restrict__ c, int m, double2 u) {
  if (m < 2) return c2(0.0, 0.0);
  double2 r = cscale(c[m - 1], (double)(m - 1));
  for (int j = m - 2; j >= 1; --j) r = cmadd(r, u, cscale(c[j], (double)j));
  return r;
}

// w of fracture f at z, all elements but `exclude`
__device__ double2 frac_w(const NetView& net, int f, double2 z, int exclude, bool jit) {
  double2 w = c2(0.0, 0.0);
  const int kbeg = net.mem_ptr[f], kend = net.mem_ptr[f + 1];
  for (int ks = kbeg; ks < kend; ++ks) {
    const Slot sl = net.slots[ks];
    const int src = sl.elem;
    if (src == exclude) continue;
    const int m = net.ncoef[src];
    const double2* __restrict__ cf = net.coef + (size_t)src * net.M;
    if (sl.kind == KIND_LINE || sl.kind == KIND_IMPL) {
      if (jit && sl.kind == KIND_IMPL) continue;
      const double2 Z = cadd(cmul(sl.alpha, z), sl.beta);
      const double2 chi = joukowsky_exterior(Z);
      const double2 u = crcp(chi);
      const double2 u2 = cmul(u, u);
      // asym_expansion_d1 + q / (2 pi chi), sign -1 seen from an intersection's second fracture
      double2 t = cmul(u2, dseries(cf, m, u));
      t = c2(-t.x, -t.y);
      if (sl.kind == KIND_LINE) {
        const double qq = net.q[src] * kQ2Pi;
        t = c2(fma(qq, u.x, t.x), fma(qq, u.y, t.y));
      }
      const double2 jac = cdiv(cscale(sl.alpha, 2.0), c2(1.0 - u2.x, -u2.y));
      w = csub(w, cscale(cmul(t, jac), sl.sign));
    } else if (sl.kind == KIND_BOUND) {
      const double2 chi = cscale(z, sl.alpha.x);
      const double2 s = dseries(cf, m, chi);
      w = jit ? cadd(w, s) : csub(w, cscale(s, sl.alpha.x));
    } else if (sl.kind == KIND_WELL) {
      const double2 chi = c2(fma(sl.alpha.x, z.x, sl.beta.x), fma(sl.alpha.x, z.y, sl.beta.y));
      if (!jit && cabs_(chi) < 1.0 - 1e-10) return c2(__builtin_nan(""), __builtin_nan(""));
      w = csub(w, cscale(crcp(chi), net.q[src] * kQ2Pi * sl.alpha.x));
    } else {  // KIND_IMPC
      if (jit) continue;
      const double2 chi = c2(fma(sl.alpha.x, z.x, sl.beta.x), fma(sl.alpha.x, z.y, sl.beta.y));
      if (cabs_(chi) < 1.0 - 1e-10) return c2(__builtin_nan(""), __builtin_nan(""));
      const double2 u = crcp(chi);
      w = cadd(w, cscale(cmul(cmul(u, u), dseries(cf, m, u)), sl.alpha.x));
    }
  }
  return w;
}

__global__ __launch_bounds__(64) void w_points_kernel(NetView net, TargetView tg, int exclude, int jit,
                                                      double* __restrict__ val) {
  const Tile tile = tg.tiles[blockIdx.x];
  const int p = tile.p0 + (int)threadIdx.x;
  if (p >= tile.pend) return;
  const double2 w = frac_w(net, tile.frac, tg.pt_z[p], exclude, jit != 0);
  val[2 * (int64_t)p] = w.x;
  val[2 * (int64_t)p + 1] = w.y;
}

// ---- streamline tracing ----------------------------------------------------------------------

// z scaled back inside the fracture disk when a step leaves it (dfn.py:2817-2818, 2827-2828)
__device__ __forceinline__ double2 clamp_disk(double2 z, double R) {
  const double a = cabs_(z);
  if (a > R) return cscale(z, R / (a * (1.0 + 1e-5)));
  return z;
}

// predictor-corrector step of signed length ds from z0 (dfn.py:2779-2834); NaN when w(z0) is NaN
__device__ double2 rk_step(const NetView& net, int f, double2 z0, double ds, double R) {
  const double2 w0 = frac_w(net, f, z0, -1, false);
  if (cnan(w0)) return c2(__builtin_nan(""), __builtin_nan(""));
  double2 z1 = cadd(z0, cscale(cdiv(cconj(w0), c2(cabs_(w0), 0.0)), ds));
  z1 = clamp_disk(z1, R);
  double dz = 1e99;
  for (int it = 0; dz > 1e-6 && it < 10; ++it) {
    const double2 w1 = frac_w(net, f, z1, -1, false);
    if (cnan(w1)) break;
    const double2 s = cadd(w0, w1);
    double2 z2 = cadd(z0, cscale(cdiv(cconj(s), c2(cabs_(s), 0.0)), ds));
    z2 = clamp_disk(z2, R);
    dz = cabs_(csub(z2, z1));
    z1 = z2;
  }
  return z1;
}

// crossing of the step z0 -> z1 with the segment e0 -> e1 (line_line_intersection,
// geometry_functions.py:405-438, and the two distance tests of intersection.py:253-271)
__device__ bool cross_line(double2 z0, double2 z1, double2 e0, double2 e1, double2* zc) {
  const double2 d01 = csub(z1, z0), d23 = csub(e1, e0);
  const double2 det = csub(cmul(csub(cconj(z1), cconj(z0)), d23), cmul(d01, csub(cconj(e1), cconj(e0))));
  if (det.x == 0.0 && det.y == 0.0) return false;
  const double2 num = csub(cmul(csub(cmul(cconj(z1), z0), cmul(z1, cconj(z0))), d23),
                           cmul(d01, csub(cmul(cconj(e1), e0), cmul(e1, cconj(e0)))));
  const double2 z = cdiv(num, det);
  // a NaN distance passes these tests, as it does the reference's "> atol" comparisons
  if (fabs(cabs_(csub(z, z0)) + cabs_(csub(z1, z)) - cabs_(csub(z0, z1))) > 1e-12) return false;
  if (fabs(cabs_(csub(z, e0)) + cabs_(csub(z, e1)) - cabs_(csub(e0, e1))) > 1e-12) return false;
  *zc = z;
  return true;
}

// crossing of the step with a well's circle (well.py:162-200 over line_circle_intersection,
// geometry_functions.py:359-402); the point returned is pushed out by 1e-10 in the circle's map
__device__ bool cross_well(double2 z0, double2 z1, double2 c, double r, double2* zc) {
  const double atol = 1e-10;
  const double2 x0c = c2((z0.x - c.x) / r, (z0.y - c.y) / r);  // map_z_circle_to_chi
  const double2 x1c = c2((z1.x - c.x) / r, (z1.y - c.y) / r);
  const double dx = x1c.x - x0c.x, dy = x1c.y - x0c.y;
  double2 chi2, chi3;
  if (dx == 0.0) {
    const double x = x0c.x;
    const double y1 = sqrt(1.0 - x * x);
    chi2 = c2(x, y1);
    chi3 = c2(x, -y1);
  } else {
    const double mm = dy / dx;
    const double b0 = x0c.y - mm * x0c.x;
    const double a = 1.0 + mm * mm, b = 2.0 * b0 * mm, cc = b0 * b0 - 1.0;
    const double disc = b * b - 4.0 * a * cc;
    if (disc < 0.0) return false;
    const double sd = sqrt(disc);
    const double xa = (-b + sd) / (2.0 * a), xb = (-b - sd) / (2.0 * a);
    chi2 = c2(xa, mm * xa + b0);
    chi3 = c2(xb, mm * xb + b0);
  }
  const double l01 = cabs_(csub(x1c, x0c));
  const double diff0 = fabs(cabs_(csub(chi2, x0c)) + cabs_(csub(chi2, x1c)) - l01);
  const double diff1 = fabs(cabs_(csub(chi3, x0c)) + cabs_(csub(chi3, x1c)) - l01);
  if (diff0 > atol && diff1 > atol) return false;
  chi2 = cscale(chi2, 1.0 + atol);
  chi3 = cscale(chi3, 1.0 + atol);
  const double2 z2 = cadd(cscale(chi2, r), c), z3 = cadd(cscale(chi3, r), c);
  *zc = cabs_(csub(z2, z0)) < cabs_(csub(z3, z0)) ? z2 : z3;
  return true;
}

// first discharge element of the fracture (in its element order) the step crosses, -1 if none
__device__ int check_exit(const NetView& net, int f, double2 z0, double2 z1, double2* zc) {
  const int kbeg = net.mem_ptr[f], kend = net.mem_ptr[f + 1];
  for (int ks = kbeg; ks < kend; ++ks) {
    const Slot sl = net.slots[ks];
    const int e = sl.elem;
    if (sl.kind == KIND_LINE) {  // intersections (this fracture's side) and constant-head lines
      const double2* ep = net.ep + 4 * (size_t)e + (sl.sign < 0.0 ? 2 : 0);
      if (cross_line(z0, z1, ep[0], ep[1], zc)) return e;
    } else if (sl.kind == KIND_WELL) {
      if (cross_well(z0, z1, net.center[e], net.radius[e], zc)) return e;
    }
  }
  return -1;
}

__device__ __forceinline__ double speed(const NetView& net, int f, double2 z, double aperture) {
  return cabs_(frac_w(net, f, z, -1, false)) / aperture;
}

__global__ __launch_bounds__(64) void trace_kernel(NetView net, TraceArgs ta) {
  const int i = blockIdx.x * 64 + (int)threadIdx.x;
  if (i >= ta.n) return;
  const int f = ta.frac[i];
  const double R = ta.fradius[f];
  const double ds = (ta.backward ? -ta.ds : ta.ds) * R;
  const double ap = ta.aperture[f];
  // the budget test below keeps n <= max(remaining + 2, 3) <= capacity (checked by the caller);
  // the clamp only guards the row's bounds
  const int64_t remaining = ta.remaining[i] < ta.capacity - 2 ? ta.remaining[i] : ta.capacity - 2;
  double2* pts = ta.pts + (int64_t)i * ta.capacity;
  double* vel = ta.vel + (int64_t)i * ta.capacity;
  double2 z0 = ta.z0[i];
  const double v0 = speed(net, f, z0, ap);
  double2 z1 = rk_step(net, f, z0, ds, R);
  if (cnan(z1)) {  // the reference leaves the tracking loop without a segment (dfn.py:2682-2683)
    ta.npts[i] = -1;
    ta.exit_elem[i] = -1;
    return;
  }
  int64_t n = 0;
  pts[n] = ta.z_start[i];
  vel[n++] = v0;
  double2 z3;
  int elem = check_exit(net, f, z0, z1, &z3);
  while (elem < 0) {
    pts[n] = z1;
    vel[n++] = speed(net, f, z1, ap);
    z0 = z1;
    z1 = rk_step(net, f, z0, ds, R);
    if (cnan(z1)) {
      z3 = z0;
      break;
    }
    elem = check_exit(net, f, z0, z1, &z3);
    if (n > remaining) {  // point budget spent: the step's end point closes the segment
      z3 = z1;
      break;
    }
  }
  pts[n] = z3;
  vel[n++] = speed(net, f, z3, ap);
  ta.npts[i] = n;
  ta.exit_elem[i] = elem;
}

}  // namespace

void launch_w_points(const NetView& net, const TargetView& tg, int ntiles, int exclude, int jit, double* val,
                     hipStream_t st) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(w_points_kernel, dim3(ntiles), dim3(64), 0, st, net, tg, exclude, jit, val);
}

void launch_trace(const NetView& net, const TraceArgs& ta, hipStream_t st) {
  if (ta.n <= 0) return;
  hipLaunchKernelGGL(trace_kernel, dim3((ta.n + 63) / 64), dim3(64), 0, st, net, ta);
}

}  // namespace aem
