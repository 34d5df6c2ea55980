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
//   trace_kernel     one wavefront per streamline, its lanes over the fracture's elements:
//                    DFN.streamline_tracking (dfn.py:2638-2725) segment
//                    by segment.  In a fracture: predictor-corrector steps (runge_kutta,
//                    dfn.py:2779-2834) until the step crosses a discharge element
//                    (check_streamline_exit, dfn.py:2727-2742, with intersection.py:227-271,
//                    const_head.py:208-246, well.py:162-200), w turns NaN or the point budget is
//                    spent; through an intersection into the neighbouring fracture with the
//                    elevation split of get_exit_intersection (dfn.py:2744-2777), else the line
//                    ends.  Points and speeds |w| / aperture (fracture.py:479-497) go to a row of
//                    the launch; a line whose next segment might not fit its row stops with its
//                    state saved and continues in the caller's next launch.
//   trace_pack_kernel  the rows' points and segments packed back to back for the copy out.
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

// w of one source slot of a fracture at z (NaN inside a well or an impermeable circle in the
// object model's form)
__device__ double2 slot_w(const NetView& net, int ks, double2 z, bool jit) {
  const Slot sl = net.slots[ks];
  const int src = sl.elem;
  const int m = net.ncoef[src];
  const double2* __restrict__ cf = net.coef + (size_t)src * net.M;
  if (sl.kind == KIND_LINE || sl.kind == KIND_IMPL) {
    if (jit && sl.kind == KIND_IMPL) return c2(0.0, 0.0);
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
    return cscale(cmul(t, jac), -sl.sign);
  }
  if (sl.kind == KIND_BOUND) {
    const double2 s = dseries(cf, m, cscale(z, sl.alpha.x));
    return jit ? s : cscale(s, -sl.alpha.x);
  }
  const double2 chi = c2(fma(sl.alpha.x, z.x, sl.beta.x), fma(sl.alpha.x, z.y, sl.beta.y));
  if (sl.kind == KIND_WELL) {
    if (!jit && cabs_(chi) < 1.0 - 1e-10) return c2(__builtin_nan(""), __builtin_nan(""));
    return cscale(crcp(chi), -net.q[src] * kQ2Pi * sl.alpha.x);
  }
  // KIND_IMPC
  if (jit) return c2(0.0, 0.0);
  if (cabs_(chi) < 1.0 - 1e-10) return c2(__builtin_nan(""), __builtin_nan(""));
  const double2 u = crcp(chi);
  return cscale(cmul(cmul(u, u), dseries(cf, m, u)), sl.alpha.x);
}

// w of fracture f at z, all elements but `exclude`, summed in the fracture's element order
__device__ double2 frac_w(const NetView& net, int f, double2 z, int exclude, bool jit) {
  double2 w = c2(0.0, 0.0);
  const int kbeg = net.mem_ptr[f], kend = net.mem_ptr[f + 1];
  for (int ks = kbeg; ks < kend; ++ks)
    if (net.slots[ks].elem != exclude) w = cadd(w, slot_w(net, ks, z, jit));
  return w;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// the object model's w of fracture f at z by the whole wavefront: lane l sums slots l, l + 64, ...
// and the butterfly leaves the same total in every lane (each stage adds two values that are equal
// across the lanes it pairs), so the line's control flow stays uniform
__device__ double2 frac_w_wave(const NetView& net, int f, double2 z) {
  const int lane = (int)threadIdx.x & 63;
  double2 w = c2(0.0, 0.0);
  const int kbeg = net.mem_ptr[f], kend = net.mem_ptr[f + 1];
  for (int ks = kbeg + lane; ks < kend; ks += 64) w = cadd(w, slot_w(net, ks, z, false));
  return c2(wave_sum(w.x), wave_sum(w.y));
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
  const double2 w0 = frac_w_wave(net, f, z0);
  if (cnan(w0)) return c2(__builtin_nan(""), __builtin_nan(""));
  double2 z1 = cadd(z0, cscale(cdiv(cconj(w0), c2(cabs_(w0), 0.0)), ds));
  z1 = clamp_disk(z1, R);
  double dz = 1e99;
  for (int it = 0; dz > 1e-6 && it < 10; ++it) {
    const double2 w1 = frac_w_wave(net, f, z1);
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

// first discharge element of the fracture (in its element order) the step crosses, -1 if none:
// lane l tests slots l, l + 64, ... (stopping at its first crossing); the smallest slot wins
__device__ int check_exit(const NetView& net, int f, double2 z0, double2 z1, double2* zc) {
  const int lane = (int)threadIdx.x & 63;
  const int kbeg = net.mem_ptr[f], kend = net.mem_ptr[f + 1];
  int best = kend;
  double2 zb = c2(0.0, 0.0);
  for (int ks = kbeg + lane; ks < kend; ks += 64) {
    const Slot sl = net.slots[ks];
    const int e = sl.elem;
    bool hit = false;
    if (sl.kind == KIND_LINE) {  // intersections (this fracture's side) and constant-head lines
      const double2* ep = net.ep + 4 * (size_t)e + (sl.sign < 0.0 ? 2 : 0);
      hit = cross_line(z0, z1, ep[0], ep[1], &zb);
    } else if (sl.kind == KIND_WELL) {
      hit = cross_well(z0, z1, net.center[e], net.radius[e], &zb);
    }
    if (hit) {
      best = ks;
      break;
    }
  }
  int m = best;
  for (int o = 32; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, 64));
  if (m >= kend) return -1;
  const int src = (m - kbeg) & 63;
  *zc = c2(__shfl(zb.x, src, 64), __shfl(zb.y, src, 64));
  return net.slots[m].elem;
}

__device__ __forceinline__ double speed(const NetView& net, int f, double2 z, double aperture) {
  return cabs_(frac_w_wave(net, f, z)) / aperture;
}

// principal-root line map (hpc_geometry_functions.py:11-31): on the line itself the branch is
// picked by the sign of Im Z, as in the reference
__device__ __forceinline__ double2 line_chi_principal(double2 z, double2 e0, double2 e1) {
  const double2 Z = cdiv(csub(csub(cscale(z, 2.0), e0), e1), csub(e1, e0));
  const double2 a = csqrt_principal(Z.x - 1.0, Z.y), b = csqrt_principal(Z.x + 1.0, Z.y);
  return cadd(Z, cmul(a, b));
}

__device__ __forceinline__ double2 line_z(double2 chi, double2 e0, double2 e1) {
  const double2 Z = cscale(cadd(chi, crcp(chi)), 0.5);
  return cscale(cadd(cmul(Z, csub(e1, e0)), cadd(e0, e1)), 0.5);
}

struct P3 {
  double x, y, z;
};

__device__ __forceinline__ P3 to_3d(const double* fr, double2 z) {  // map_2d_to_3d, geometry_functions.py:161-190
  return P3{z.x * fr[3] + z.y * fr[6] + fr[0], z.x * fr[4] + z.y * fr[7] + fr[1], z.x * fr[5] + z.y * fr[8] + fr[2]};
}

__device__ __forceinline__ double2 to_2d(const double* fr, P3 p) {  // map_3d_to_2d, geometry_functions.py:193-220
  const double dx = p.x - fr[0], dy = p.y - fr[1], dz = p.z - fr[2];
  return c2(dx * fr[3] + dy * fr[4] + dz * fr[5], dx * fr[6] + dy * fr[7] + dz * fr[8]);
}

// one wavefront per line (every lane follows the line; lane 0 writes)
__global__ __launch_bounds__(64) void trace_kernel(NetView net, TraceArgs ta) {
  const int k = blockIdx.x;
  const bool w0 = threadIdx.x == 0;
  const int i = ta.act[k];
  double2* pts = ta.pts + (int64_t)k * ta.capacity;
  double* vel = ta.vel + (int64_t)k * ta.capacity;
  int* sfrac = ta.seg_frac + (int64_t)k * ta.seg_capacity;
  int* slen = ta.seg_len + (int64_t)k * ta.seg_capacity;
  int f = ta.frac[i];
  double2 z0 = ta.z0[i], zs = ta.z_start[i];
  double elev = ta.elev[i];
  int64_t length = ta.length[i];
  int nseg = ta.nseg[i], done = ta.done[i], exit_elem = ta.exit_elem[i];
  int64_t op = 0;
  int os = 0;
  while (!done) {
    // a segment holds at most max(remaining + 2, 3) points; the line waits for the next launch
    // when this row cannot take that many
    const int64_t remaining = ta.max_length - length;
    const int64_t need = remaining + 2 > 3 ? remaining + 2 : 3;
    if (op + need > ta.capacity || os >= ta.seg_capacity) break;
    // ---- one segment (dfn.py:2671-2705) ----
    const double R = ta.fradius[f];
    const double ds = (ta.backward ? -ta.ds : ta.ds) * R;
    const double ap = ta.aperture[f];
    const double v0 = speed(net, f, z0, ap);
    double2 z1 = rk_step(net, f, z0, ds, R);
    if (cnan(z1)) {  // the reference leaves the tracking loop without a segment (dfn.py:2682-2683)
      done = 1;
      break;
    }
    int64_t n = op;
    if (w0) {
      pts[n] = zs;
      vel[n] = v0;
    }
    ++n;
    double2 z3;
    int elem = check_exit(net, f, z0, z1, &z3);
    while (elem < 0) {
      const double v1 = speed(net, f, z1, ap);
      if (w0) {
        pts[n] = z1;
        vel[n] = v1;
      }
      ++n;
      z0 = z1;
      z1 = rk_step(net, f, z0, ds, R);
      if (cnan(z1)) {
        z3 = z0;
        break;
      }
      elem = check_exit(net, f, z0, z1, &z3);
      if (n - op > remaining) {  // point budget spent: the step's end closes the segment
        z3 = z1;
        break;
      }
    }
    const double v3 = speed(net, f, z3, ap);
    if (w0) {
      pts[n] = z3;
      vel[n] = v3;
      sfrac[os] = f;
      slen[os] = (int)(n + 1 - op);
    }
    ++n;
    ++os;
    length += n - op;
    op = n;
    ++nseg;
    exit_elem = elem;
    if (elem < 0 || net.etype[elem] != 0 || nseg >= ta.max_segments) {
      done = 1;
      break;
    }
    // ---- into the neighbouring fracture (get_exit_intersection, dfn.py:2744-2777) ----
    const int f_old = f;
    const int fn = net.efrac0[elem] == f_old ? net.efrac1[elem] : net.efrac0[elem];
    const double* fo = ta.frame + 12 * (int64_t)f_old;
    const double* fnw = ta.frame + 12 * (int64_t)fn;
    const P3 z3d = to_3d(fo, z3);
    const double2* ep = net.ep + 4 * (int64_t)elem + (fn == net.efrac0[elem] ? 0 : 2);
    const double2 chi0 = line_chi_principal(to_2d(fnw, z3d), ep[0], ep[1]);
    const double2 za = line_z(cscale(chi0, 1.0 + 1e-4), ep[0], ep[1]);
    const double2 zb = line_z(cscale(cconj(chi0), 1.0 + 1e-4), ep[0], ep[1]);
    const double abs_a = cabs_(frac_w_wave(net, fn, za));
    const double abs_b = cabs_(frac_w_wave(net, fn, zb));
    double divide = abs_a / (abs_a + abs_b);
    const P3 pa = to_3d(fnw, za);
    const double nz0 = (pa.x - z3d.x) * fo[9] + (pa.y - z3d.y) * fo[10] + (pa.z - z3d.z) * fo[11];
    double2 up, down;
    if (nz0 < 0.0) {
      up = zb;
      down = za;
    } else {
      up = za;
      down = zb;
      divide = 1.0 - divide;
    }
    if (elev < divide) {
      elev /= divide;
      z0 = down;
    } else {
      elev = (elev - divide) / (1.0 - divide);
      z0 = up;
    }
    zs = za;
    f = fn;
  }
  if (!w0) return;
  ta.frac[i] = f;
  ta.z0[i] = z0;
  ta.z_start[i] = zs;
  ta.elev[i] = elev;
  ta.length[i] = length;
  ta.nseg[i] = nseg;
  ta.done[i] = done;
  ta.exit_elem[i] = exit_elem;
  ta.out_pts[k] = op;
  ta.out_segs[k] = os;
}

// one workgroup per row: its points and segments to their packed offsets
__global__ __launch_bounds__(256) void trace_pack_kernel(TracePack tp) {
  const int k = blockIdx.x;
  const int64_t np = tp.out_pts[k];
  const int ns = tp.out_segs[k];
  const double2* src = tp.pts + (int64_t)k * tp.capacity;
  const double* sv = tp.vel + (int64_t)k * tp.capacity;
  for (int64_t j = threadIdx.x; j < np; j += 256) {
    tp.ppts[tp.poff[k] + j] = src[j];
    tp.pvel[tp.poff[k] + j] = sv[j];
  }
  for (int j = threadIdx.x; j < ns; j += 256) {
    tp.pseg_frac[tp.soff[k] + j] = tp.seg_frac[(int64_t)k * tp.seg_capacity + j];
    tp.pseg_len[tp.soff[k] + j] = tp.seg_len[(int64_t)k * tp.seg_capacity + j];
  }
}

}  // namespace

void launch_w_points(const NetView& net, const TargetView& tg, int ntiles, int exclude, int jit, double* val,
                     hipStream_t st) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(w_points_kernel, dim3(ntiles), dim3(64), 0, st, net, tg, exclude, jit, val);
}

void launch_trace(const NetView& net, const TraceArgs& ta, hipStream_t st) {
  if (ta.n <= 0) return;
  hipLaunchKernelGGL(trace_kernel, dim3(ta.n), dim3(64), 0, st, net, ta);
}

void launch_trace_pack(const TracePack& tp, hipStream_t st) {
  if (tp.n <= 0) return;
  hipLaunchKernelGGL(trace_pack_kernel, dim3(tp.n), dim3(256), 0, st, tp);
}

}  // namespace aem
