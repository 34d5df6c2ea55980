// Transposed operators of the analytic-element fixed point, for the implicit adjoint.
//
// At a converged state the solve satisfies (coefficients c, discharges y = [q; C]):
//   c = FIT(EVAL(c, y), t)          collocation sweep + Cauchy/DFT fit (one Jacobi step)
//   A(t) y = RHS(c, t, h)           head equations with the series part on the right
// For an objective J(c, y, t) the Lagrangian L = J + lam.(FIT(EVAL(c, y)) - c) + mu.(RHS(c) - A y)
// is stationary in (c, y) when
//   mu  = A^-T (J_y + EVAL_y^T FIT^T lam)
//   lam = J_c + EVAL_c^T FIT^T lam + RHS_c^T mu,
// which is the transpose of the forward iteration and converges at the same rate when damped the
// same way.  dJ/dp = J_p + lam.FIT_p + mu.(RHS_p - A_p y) then gives every parameter gradient
// from one adjoint solve.
//
// Kernels here:
//   adj_fitT_kernel    FIT^T: dual weights at an element's collocation points from its
//                      coefficient adjoint (the DFT transposed; circle unwrapping transposed into
//                      suffix sums that feed the branch-cut discharge terms)
//   adj_gather_kernel  EVAL^T: for each source element, the sum over every weighted target point of
//                      its fracture(s) of the derivative of the point's value with respect to the
//                      source's coefficients (powers of 1/chi or chi) and discharge (log chi):
//                      one wavefront per source, lanes over targets, wave reductions at the end
//   adj_cross_kernel   branch-cut terms of FIT^T into the discharge adjoint
//   adj_csum_kernel    fracture-constant adjoint (sum of the Re weights of a fracture's points)
//   adj_rhsT_kernel    RHS^T: dual weights at the discharge points from mu
//   adj_update_kernel  damped update of lam with the change / magnitude maxima
//   adj_tgrad_kernel   d c / d t of intersections contracted with lam (uses the forward values)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aem_kernels.hpp"
#include "aem_math.hpp"

namespace aem {

namespace {

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}

// One step of the reduce-scatter below: the lanes with bit O clear keep the first ceil(C/2) of their
// values, the others the rest (a zero stands in for the missing one when C is odd), each adding
// what its partner at distance O sends.  [base, base + real) is the lane's range of real indices.
template <int C, int O>
__device__ __forceinline__ void rs_step(double* a, int lane, int& base, int& real) {
  if constexpr (O > 0) {
    constexpr int H = (C + 1) / 2;
    const bool up = (lane & O) != 0;
#pragma unroll
    for (int k = 0; k < H; ++k) {
      const double lo = a[k];
      const double hi = (k + H < C) ? a[k + H] : 0.0;
      a[k] = (up ? hi : lo) + __shfl_xor(up ? lo : hi, O, 64);
    }
    if (up) {
      base += H;
      real -= H;
    } else {
      real = min(real, H);
    }
    rs_step<H, O / 2>(a, lane, base, real);
  }
}

// Wavefront sums of a[0..C) by recursive halving: afterwards a[0] of lane l holds the total of
// index reduce_scatter(...) when that is >= 0 (C lanes hold one total each, the rest return -1).
// ceil(C/2) + ceil(C/4) + ... exchanges instead of the 6 C of separate butterfly sums.
template <int C>
__device__ __forceinline__ int reduce_scatter(double* a, int lane) {
  int base = 0, real = C;
  rs_step<C, 32>(a, lane, base, real);
  return real >= 1 ? base : -1;
}

__device__ __forceinline__ void atomic_max_pos(unsigned long long* p, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

}  // namespace

constexpr int kAdjChunk = 20;  // coefficient adjoints accumulated per pass over the targets

// FIT^T for one element per wavefront.  Writes the dual weight of every collocation point of the
// element (Re weight for line types, Im weight for circle types) and, for circle types, the suffix
// sums sfx[e][k] = sum_{i > k} w_i that the branch-cut correction kappa_k receives.
__global__ __launch_bounds__(64) void adj_fitT_kernel(NetView net, const FitItem* __restrict__ items, int nitems,
                                                      const double2* __restrict__ lam, const int* __restrict__ pt_off,
                                                      const double2* __restrict__ trig, double2* __restrict__ w,
                                                      double* __restrict__ sfx) {
  __shared__ double sw[kMaxPts];
  const int b = blockIdx.x;
  if (b >= nitems) return;
  const FitItem it = items[b];
  const int e = it.elem;
  const int lane = threadIdx.x;
  const int ty = net.etype[e];
  const int m = net.ncoef[e];
  const int n = net.nint[e];
  const double2* l = lam + (size_t)e * net.M;
  const double two_n = 2.0 / (double)n;
  const double2* tr = trig + 2 * (int64_t)n * (n - 1);  // (cos, sin)(k pi / (2 n)), k < 4 n
  const bool circle = (ty == 1 || ty == 4 || ty == 5);
  double f0 = 1.0, f1 = 0.0;
  if (ty == 0) {
    // c_j = (t0 a1_j - t1 a0_j)/(t0 + t1): side 0 weighs -t1/(t0+t1), side 1 t0/(t0+t1)
    const double t0 = net.t[net.efrac0[e]], t1 = net.t[net.efrac1[e]];
    f0 = -t1 / (t0 + t1);
    f1 = t0 / (t0 + t1);
  } else if (ty == 3) {
    f0 = -1.0;  // c_j = -a_j
  }
  const int p0 = pt_off[it.s0];
  const int p1 = (ty == 0) ? pt_off[it.s1] : 0;
  for (int i = lane; i < n; i += 64) {
    double acc = 0.0;
    if (ty == 1) {
      // c_j = -(2/n) h_j (Y_j + i X_j), h_0 = 1/2, angles 2 pi i j / n: table entry 4 (i j mod n),
      // i j mod n stepped by i (i < n, one subtraction)
      int r = 0;
      for (int j = 0; j < m; ++j) {
        const double2 cs = tr[4 * r];  // angle 2 pi i j / n
        const double h = (j == 0) ? 0.5 : 1.0;
        acc = fma(h, fma(l[j].x, cs.y, l[j].y * cs.x), acc);
        r += i;
        if (r >= n) r -= n;
      }
      acc *= -two_n;
    } else {
      // angle j (2 i + 1) pi / (2 n): table entry j (2 i + 1) mod 4 n, stepped by 2 i + 1 < 4 n
      const int period = 4 * n, step = 2 * i + 1;
      int r = step;
      for (int j = 1; j < m; ++j) {
        const double2 cs = tr[r];
        r += step;
        if (r >= period) r -= period;
        const double s = cs.y, c = cs.x;
        if (ty == 0 || ty == 3) acc = fma(l[j].x, c, acc);
        else if (ty == 4) acc += l[j].x * s - l[j].y * c;  // c_j = ((2/n) Y_j, -(2/n) X_j)
        else acc -= l[j].y * c;                              // c_j = (0, -(2/n) X_j)
      }
      acc *= two_n;
    }
    if (circle) {
      sw[i] = acc;
      w[p0 + i] = make_double2(0.0, acc);
    } else {
      w[p0 + i] = make_double2(f0 * acc, 0.0);
      if (ty == 0) w[p1 + i] = make_double2(f1 * acc, 0.0);
    }
  }
  if (circle) {
    __syncthreads();
    if (lane == 0) {
      // psi~_i = psi_i - sum_{k < i} kappa_k  =>  dJ/dkappa_k = -sum_{i > k} w_i
      double* o = sfx + (size_t)e * 2 * net.M;
      double run = 0.0;
      for (int k = n - 2; k >= 0; --k) {
        run += sw[k + 1];
        o[k] = run;
      }
    }
  }
}

// EVAL^T for one source element per wavefront (items sorted heaviest first).
//   eslot[eslot_ptr[e] .. eslot_ptr[e+1]) = (slot, fracture) of the element's source sides
//   target points of fracture f: [fp_beg[f], fp_end[f]) with coordinates pz and owner pelem
//   w: (Re, Im) dual weight W of every target point
//   excl: skip the source's own points (collocation); use_q: include the log terms
//   accumulate: add into gc / gq instead of overwriting
// A target's weighted value W.(Re, Im)(c_k b^k) (b = 1/chi, z/R) has the derivative
// (Re Q_k, -Im Q_k) with respect to (Re c_k, Im c_k), Q_k = conj(W) b^k.  Intersections and
// constant-head lines have real coefficients (their fit keeps Re only), impermeable lines
// imaginary ones: for them (BOTH = false) one real sequence u_k (Re Q_k or -Im Q_k) is summed,
// generated by u_{k+1} = 2 Re(b) u_k - |b|^2 u_{k-1} (b and conj b are the roots of
// x^2 - 2 Re(b) x + |b|^2, so every real combination of b^k and conj(b)^k follows it): two
// operations per power, with the other component of gc left at zero.  Bounding and impermeable
// circles (BOTH = true) sum the complex Q_k.
template <bool BOTH>
__global__ __launch_bounds__(64) void adj_gather_kernel(NetView net, const int* __restrict__ items, int nitems,
                                                        const int* __restrict__ eslot_ptr,
                                                        const int2* __restrict__ eslot, const int* __restrict__ fp_beg,
                                                        const int* __restrict__ fp_end, const double2* __restrict__ pz,
                                                        const int* __restrict__ pelem, const double2* __restrict__ w,
                                                        int excl, int use_q, int accumulate, double2* __restrict__ gc,
                                                        double* __restrict__ gq) {
  constexpr int NA = BOTH ? kAdjChunk : 1;
  const int b = blockIdx.x;
  if (b >= nitems) return;
  const int src = items[b];
  const int lane = threadIdx.x;
  const int ty = net.etype[src];
  const int m = (ty == 2) ? 0 : net.ncoef[src];
  const bool imc = (ty == 5);  // impermeable line: imaginary coefficients
  // (the complex-coefficient sources are circles: no log term, no line map)
  const bool has_q = !BOTH && use_q && (ty == 0 || ty == 2 || ty == 3);
  const int sb = eslot_ptr[src], se = eslot_ptr[src + 1];
  double aq = 0.0;
  int k0 = 0;
  do {
    const int kc = min(kAdjChunk, m - k0);
    const bool do_q = has_q && k0 == 0;
    double acc[kAdjChunk + NA];  // Re Q_k (or -Im Q_k), then -Im Q_k when BOTH
    double* ar = acc;
    double* ai = acc + kAdjChunk;
#pragma unroll
    for (int k = 0; k < kAdjChunk + NA; ++k) acc[k] = 0.0;
    for (int s = sb; s < se; ++s) {
      const int2 sf = eslot[s];
      const Slot sl = net.slots[sf.x];
      const int pb = fp_beg[sf.y], pe = fp_end[sf.y];
      const bool line = !BOTH && (sl.kind == KIND_LINE || sl.kind == KIND_IMPL);
      const double sg = line ? sl.sign : 1.0;
      const bool slot_q = do_q && (sl.kind == KIND_LINE || sl.kind == KIND_WELL);
      // the next point's owner, weight and position are requested before this point's work
      int pel_n = -1;
      double2 wd_n = make_double2(0.0, 0.0), z_n = make_double2(0.0, 0.0);
      if (pb + lane < pe) {
        pel_n = excl ? pelem[pb + lane] : -1;
        wd_n = w[pb + lane];
        z_n = pz[pb + lane];
      }
      for (int p = pb + lane; p < pe; p += 64) {
        const int pel = pel_n;
        const double2 wd = wd_n, z = z_n;
        if (p + 64 < pe) {
          pel_n = excl ? pelem[p + 64] : -1;
          wd_n = w[p + 64];
          z_n = pz[p + 64];
        }
        if (pel == src) continue;
        if (wd.x == 0.0 && wd.y == 0.0) continue;
        double2 base = make_double2(0.0, 0.0);
        if (line) {
          const double2 Z = make_double2(fma(sl.alpha.x, z.x, fma(-sl.alpha.y, z.y, sl.beta.x)),
                                         fma(sl.alpha.x, z.y, fma(sl.alpha.y, z.x, sl.beta.y)));
          const double2 chi = joukowsky_exterior(Z);
          const double r2 = fma(chi.x, chi.x, chi.y * chi.y);
          const double inv = rcp_pos(r2);
          base = make_double2(chi.x * inv, -chi.y * inv);  // 1/chi
          if (slot_q) aq = fma(sl.sign * kInv2Pi * wd.x, 0.5 * log_pos(r2), aq);
        } else if (sl.kind == KIND_BOUND) {
          base = make_double2(sl.alpha.x * z.x, sl.alpha.x * z.y);  // z / R
        } else {
          const double2 chi = make_double2(fma(sl.alpha.x, z.x, sl.beta.x), fma(sl.alpha.x, z.y, sl.beta.y));
          const double r2 = fma(chi.x, chi.x, chi.y * chi.y);
          if (sl.kind == KIND_WELL) {
            if (slot_q) aq = fma(kInv2Pi * wd.x, 0.5 * log_pos(r2), aq);
          } else {
            const double inv = rcp_pos(r2);
            base = make_double2(chi.x * inv, -chi.y * inv);
          }
        }
        if (kc > 0) {
          // Q_{k0} = conj(W) b^{k0}
          double2 Q = make_double2(sg * wd.x, -sg * wd.y);
          for (int k = 0; k < k0; ++k) Q = cmul(Q, base);
          // every power up to the chunk's end is accumulated (|b| <= 1 at every target, so the
          // sums past ncoef stay finite); only the first kc are written
          if (BOTH) {
#pragma unroll
            for (int k = 0; k < kAdjChunk; ++k) {
              ar[k] += Q.x;
              ai[k] -= Q.y;
              Q = cmul(Q, base);
            }
          } else {
            const double2 Q1 = cmul(Q, base);
            double u0 = imc ? -Q.y : Q.x;
            double u1 = imc ? -Q1.y : Q1.x;
            const double tr = 2.0 * base.x;
            const double nb = -fma(base.x, base.x, base.y * base.y);
#pragma unroll
            for (int k = 0; k < kAdjChunk; ++k) {
              ar[k] += u0;
              const double u2 = fma(tr, u1, nb * u0);
              u0 = u1;
              u1 = u2;
            }
          }
        }
      }
    }
    double2* o = gc + (size_t)src * net.M + k0;
    const int idx = reduce_scatter<kAdjChunk + NA>(acc, lane);
    if (BOTH) {
      // lane of (power, component): one double each
      const int k = idx % kAdjChunk;
      if (idx >= 0 && k < kc) {
        double* d = reinterpret_cast<double*>(o + k) + idx / kAdjChunk;
        *d = accumulate ? *d + acc[0] : acc[0];
      }
    } else if (idx >= 0 && idx < kc) {
      const double2 v = imc ? make_double2(0.0, acc[0]) : make_double2(acc[0], 0.0);
      o[idx] = accumulate ? make_double2(o[idx].x + v.x, o[idx].y + v.y) : v;
    }
    k0 += kAdjChunk;
  } while (k0 < m);
  if (has_q) {
    // Im weights times arg(chi) in a pass of their own, after the sums are written: only the
    // points of circle-type elements carry them, and the atan2 kept out of the series loop leaves
    // its registers to the sums
    for (int s = sb; s < se; ++s) {
      const int2 sf = eslot[s];
      const Slot sl = net.slots[sf.x];
      if (sl.kind != KIND_LINE && sl.kind != KIND_WELL) continue;
      const bool line = (sl.kind == KIND_LINE);
      const double sg = line ? sl.sign : 1.0;
      for (int p = fp_beg[sf.y] + lane; p < fp_end[sf.y]; p += 64) {
        const double wy = w[p].y;
        if (wy == 0.0 || (excl && pelem[p] == src)) continue;
        const double2 z = pz[p];
        double2 chi;
        if (line) {
          const double2 Z = make_double2(fma(sl.alpha.x, z.x, fma(-sl.alpha.y, z.y, sl.beta.x)),
                                         fma(sl.alpha.x, z.y, fma(sl.alpha.y, z.x, sl.beta.y)));
          chi = joukowsky_exterior(Z);
        } else {
          chi = make_double2(fma(sl.alpha.x, z.x, sl.beta.x), fma(sl.alpha.x, z.y, sl.beta.y));
        }
        aq = fma(sg * kInv2Pi * wy, atan2(chi.y, chi.x), aq);
      }
    }
    aq = wsum(aq);
    if (lane == 0) gq[src] = accumulate ? gq[src] + aq : aq;
  }
}

// branch-cut terms: gq[src] -= sign * sfx[e][pos] over the crossings (e, pos, sign) of src
__global__ void adj_cross_kernel(int nsrc, const int* __restrict__ srcs, const int* __restrict__ cptr,
                                 const int4* __restrict__ cent, const double* __restrict__ sfx, int M,
                                 double* __restrict__ gq) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsrc) return;
  const int src = srcs[k];
  double acc = 0.0;
  for (int c = cptr[k]; c < cptr[k + 1]; ++c) {
    const int4 r = cent[c];  // (circle element, position, sign, -)
    acc -= (double)r.z * sfx[(size_t)r.x * 2 * M + r.y];
  }
  gq[src] += acc;
}

// gC[f] (+)= sum of the Re weights of fracture f's points
__global__ __launch_bounds__(64) void adj_csum_kernel(int nf, const int* __restrict__ fp_beg,
                                                      const int* __restrict__ fp_end, const double2* __restrict__ w,
                                                      int accumulate, double* __restrict__ gC) {
  const int f = blockIdx.x;
  if (f >= nf) return;
  double s = 0.0;
  for (int p = fp_beg[f] + (int)threadIdx.x; p < fp_end[f]; p += 64) s += w[p].x;
  s = wsum(s);
  if (threadIdx.x == 0) gC[f] = accumulate ? gC[f] + s : s;
}

// z = J_y + (gq at the discharge unknowns, gC at the fracture-constant unknowns); the launcher
// copies J_y into z first, so padding unknowns keep J_y = 0
__global__ void adj_build_z_kernel(int nde, int nf, const int* __restrict__ de, SysLayout sys,
                                   const double* __restrict__ jy, const double* __restrict__ gq,
                                   const double* __restrict__ gC, double* __restrict__ z) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nde) z[sys.de_pos[de[j]]] = jy[sys.de_pos[de[j]]] + gq[de[j]];
  else if (j < nde + nf) z[sys.fpos[j - nde]] = jy[sys.fpos[j - nde]] + gC[j - nde];
}

// RHS^T: b_j = S1/t1 - S0/t0 (intersection) or phi - S0, S = mean of the series over K points
__global__ __launch_bounds__(64) void adj_rhsT_kernel(NetView net, const RhsItem* __restrict__ items, int nde,
                                                      const int* __restrict__ de_pos, const double* __restrict__ mu,
                                                      const int* __restrict__ pt_off, int K, double2* __restrict__ w) {
  const int j = blockIdx.x;
  if (j >= nde) return;
  const RhsItem it = items[j];
  const int e = it.elem;
  const double m = mu[de_pos[e]] / (double)K;
  double v0, v1 = 0.0;
  if (net.etype[e] == 0) {
    v0 = -m / net.t[net.efrac0[e]];
    v1 = m / net.t[net.efrac1[e]];
  } else {
    v0 = -m;
  }
  const int b0 = pt_off[it.s0];
  for (int i = threadIdx.x; i < K; i += 64) w[b0 + i] = make_double2(v0, 0.0);
  if (it.s1 >= 0) {
    const int b1 = pt_off[it.s1];
    for (int i = threadIdx.x; i < K; i += 64) w[b1 + i] = make_double2(v1, 0.0);
  }
}

// lam += damping * (J_c + gc - lam); stats[0] = max |J_c + gc - lam|, stats[1] = max |J_c + gc|
__global__ void adj_update_kernel(int64_t cnt, const double2* __restrict__ jc, const double2* __restrict__ gc,
                                  double2* __restrict__ lam, double damping, unsigned long long* __restrict__ stats) {
  double dmax = 0.0, vmax = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 t = make_double2(jc[i].x + gc[i].x, jc[i].y + gc[i].y);
    const double2 l = lam[i];
    const double2 d = make_double2(t.x - l.x, t.y - l.y);
    lam[i] = make_double2(fma(damping, d.x, l.x), fma(damping, d.y, l.y));
    dmax = fmax(dmax, fmax(fabs(d.x), fabs(d.y)));
    vmax = fmax(vmax, fmax(fabs(t.x), fabs(t.y)));
  }
  for (int o = 32; o > 0; o >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
    vmax = fmax(vmax, __shfl_xor(vmax, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomic_max_pos(&stats[0], dmax);
    atomic_max_pos(&stats[1], vmax);
  }
}

// s_e = sum_i (v0_i + v1_i) omega_i, omega_i = (2/n) sum_{j >= 1} lam_j cos(j theta_i), for the
// intersections: lam . dc/dt0 = s_e t1/(t0+t1)^2 and lam . dc/dt1 = -s_e t0/(t0+t1)^2
__global__ __launch_bounds__(64) void adj_tgrad_kernel(NetView net, const FitItem* __restrict__ items, int nitems,
                                                       const double2* __restrict__ lam, const double* __restrict__ val,
                                                       const int* __restrict__ pt_off, double* __restrict__ s_out) {
  const int b = blockIdx.x;
  if (b >= nitems) return;
  const FitItem it = items[b];
  const int e = it.elem;
  const int m = net.ncoef[e];
  const int n = net.nint[e];
  const double2* l = lam + (size_t)e * net.M;
  const int64_t period = 4 * (int64_t)n;
  const int p0 = pt_off[it.s0], p1 = pt_off[it.s1];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 64) {
    double om = 0.0;
    for (int j = 1; j < m; ++j) {
      double s, c;
      sincospi((double)(((int64_t)j * (2 * i + 1)) % period) / (double)(2 * n), &s, &c);
      om = fma(l[j].x, c, om);
    }
    acc = fma(val[p0 + i] + val[p1 + i], om, acc);
  }
  acc = wsum(acc) * (2.0 / (double)n);
  if (threadIdx.x == 0) s_out[b] = acc;
}

// B = A^T for an n x n column-major matrix (32 x 32 tiles through LDS)
template <class T>
__global__ __launch_bounds__(256) void transpose_kernel(const T* __restrict__ A, int64_t n, T* __restrict__ B) {
  __shared__ T tile[32][33];
  A += (int64_t)blockIdx.z * n * n;
  B += (int64_t)blockIdx.z * n * n;
  const int64_t bx = (int64_t)blockIdx.x * 32, by = (int64_t)blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = bx + tx, c = by + k;
    if (r < n && c < n) tile[k][tx] = A[c * n + r];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = by + tx, c = bx + k;  // B(r, c) = A(c, r)
    if (r < n && c < n) B[c * n + r] = tile[tx][k];
  }
}

// x = (T)(z - A^T mu)  (the refinement right-hand side of the transposed solve)
template <class T>
__global__ void adj_refine_in_kernel(const double* __restrict__ z, const double* __restrict__ aty, int n,
                                     T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (T)(z[i] - aty[i]);
}

// mu[perm[i]] += rs[perm[i]] * v[i]  (A^-T = R P^T L^-T U^-T)
template <class T>
__global__ void adj_refine_out_kernel(const T* __restrict__ v, const int* __restrict__ perm,
                                      const double* __restrict__ rs, int n, double* __restrict__ mu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int p = perm[i];
    mu[p] += rs[p] * (double)v[i];
  }
}

// packed coefficient vectors (the real components that the fit reads, k < ncoef): pidx maps a
// packed entry to its position in the [E][M] complex array taken as doubles (2 (e M + k) + part)
__global__ void adj_pack_kernel(int64_t np, const int64_t* __restrict__ pidx, const double* __restrict__ full,
                                double* __restrict__ packed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) packed[i] = full[pidx[i]];
}
__global__ void adj_unpack_kernel(int64_t np, const int64_t* __restrict__ pidx, const double* __restrict__ packed,
                                  double* __restrict__ full) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) full[pidx[i]] = packed[i];
}
// out = v - packed(gc)
__global__ void adj_residual_op_kernel(int64_t np, const int64_t* __restrict__ pidx, const double* __restrict__ v,
                                       const double* __restrict__ gc, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) out[i] = v[i] - gc[pidx[i]];
}

// ---- launchers
void launch_adj_fitT(const NetView& net, const FitItem* items, int nitems, const double2* lam, const int* pt_off,
                     const double2* trig, double2* w, double* sfx, hipStream_t st) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(adj_fitT_kernel, dim3(nitems), dim3(64), 0, st, net, items, nitems, lam, pt_off, trig, w, sfx);
}

void launch_adj_gather(const NetView& net, const int* items, int nitems, int nsingle, const int* eslot_ptr,
                       const int2* eslot, const int* fp_beg, const int* fp_end, const double2* pz, const int* pelem,
                       const double2* w, int excl, int use_q, int accumulate, double2* gc, double* gq,
                       hipStream_t st) {
  if (nsingle > 0)
    hipLaunchKernelGGL(adj_gather_kernel<false>, dim3(nsingle), dim3(64), 0, st, net, items, nsingle, eslot_ptr, eslot,
                       fp_beg, fp_end, pz, pelem, w, excl, use_q, accumulate, gc, gq);
  if (nitems > nsingle)
    hipLaunchKernelGGL(adj_gather_kernel<true>, dim3(nitems - nsingle), dim3(64), 0, st, net, items + nsingle,
                       nitems - nsingle, eslot_ptr, eslot, fp_beg, fp_end, pz, pelem, w, excl, use_q, accumulate, gc,
                       gq);
}

void launch_adj_cross(int nsrc, const int* srcs, const int* cptr, const int4* cent, const double* sfx, int M,
                      double* gq, hipStream_t st) {
  if (nsrc <= 0) return;
  hipLaunchKernelGGL(adj_cross_kernel, dim3((nsrc + 255) / 256), dim3(256), 0, st, nsrc, srcs, cptr, cent, sfx, M, gq);
}

void launch_adj_csum(int nf, const int* fp_beg, const int* fp_end, const double2* w, int accumulate, double* gC,
                     hipStream_t st) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(adj_csum_kernel, dim3(nf), dim3(64), 0, st, nf, fp_beg, fp_end, w, accumulate, gC);
}

void launch_adj_build_z(int nde, int nf, const int* de, const SysLayout& sys, int64_t n, const double* jy,
                        const double* gq, const double* gC, double* z, hipStream_t st) {
  // copy first, then the element/fracture entries (two launches: the scatter may cross workgroups)
  if (n <= 0) return;
  (void)hipMemcpyAsync(z, jy, sizeof(double) * n, hipMemcpyDeviceToDevice, st);
  const int m = nde + nf;
  if (m <= 0) return;
  hipLaunchKernelGGL(adj_build_z_kernel, dim3((m + 255) / 256), dim3(256), 0, st, nde, nf, de, sys, jy, gq, gC, z);
}

void launch_adj_rhsT(const NetView& net, const RhsItem* items, int nde, const int* de_pos, const double* mu,
                     const int* pt_off, int K, double2* w, hipStream_t st) {
  if (nde <= 0) return;
  hipLaunchKernelGGL(adj_rhsT_kernel, dim3(nde), dim3(64), 0, st, net, items, nde, de_pos, mu, pt_off, K, w);
}

void launch_adj_update(int64_t cnt, const double2* jc, const double2* gc, double2* lam, double damping,
                       unsigned long long* stats, hipStream_t st) {
  if (cnt <= 0) return;
  int64_t nb = (cnt + 255) / 256;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(adj_update_kernel, dim3((unsigned)nb), dim3(256), 0, st, cnt, jc, gc, lam, damping, stats);
}

void launch_adj_tgrad(const NetView& net, const FitItem* items, int nitems, const double2* lam, const double* val,
                      const int* pt_off, double* s_out, hipStream_t st) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(adj_tgrad_kernel, dim3(nitems), dim3(64), 0, st, net, items, nitems, lam, val, pt_off, s_out);
}

void launch_adj_pack(int64_t np, const int64_t* pidx, const double* full, double* packed, hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(adj_pack_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, np, pidx, full, packed);
}
void launch_adj_unpack(int64_t np, const int64_t* pidx, const double* packed, double* full, hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(adj_unpack_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, np, pidx, packed, full);
}
void launch_adj_residual_op(int64_t np, const int64_t* pidx, const double* v, const double* gc, double* out,
                            hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(adj_residual_op_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, np, pidx, v, gc,
                     out);
}

template <class T>
void launch_transpose(const T* A, int64_t n, T* B, hipStream_t st, int nblk) {
  if (n <= 0 || nblk <= 0) return;
  const unsigned g = (unsigned)((n + 31) / 32);
  hipLaunchKernelGGL(transpose_kernel<T>, dim3(g, g, nblk), dim3(256), 0, st, A, n, B);
}
template void launch_transpose<double>(const double*, int64_t, double*, hipStream_t, int);
template void launch_transpose<float>(const float*, int64_t, float*, hipStream_t, int);

template <class T>
void launch_adj_refine_in(const double* z, const double* aty, int n, T* x, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(adj_refine_in_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, z, aty, n, x);
}
template void launch_adj_refine_in<double>(const double*, const double*, int, double*, hipStream_t);
template void launch_adj_refine_in<float>(const double*, const double*, int, float*, hipStream_t);

template <class T>
void launch_adj_refine_out(const T* v, const int* perm, const double* rs, int n, double* mu, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(adj_refine_out_kernel<T>, dim3((n + 255) / 256), dim3(256), 0, st, v, perm, rs, n, mu);
}
template void launch_adj_refine_out<double>(const double*, const int*, const double*, int, double*, hipStream_t);
template void launch_adj_refine_out<float>(const float*, const int*, const double*, int, double*, hipStream_t);

}  // namespace aem
