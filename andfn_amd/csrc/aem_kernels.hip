// HIP kernels of the analytic-element DFN solve for CDNA4 (gfx950).
//
// One iteration of AnDFN's fixed point (andfn/hpc/hpc_solve.py:152-239) is, on the device:
//   sweep_kernel<RHS>   series-only potential at the K discharge points of every discharge
//                       element side                           (build_head_matrix, hpc_solve.py:778-829)
//   rhs_reduce_kernel   mean over the K points -> right-hand side b
//   [rocSOLVER getrs]   y = A^-1 b with the factors of the constant discharge matrix
//   scatter_kernel      q, fracture constants, relative discharge change (hpc_solve.py:169-174, 720-761)
//   sweep_kernel<COLLOC> potential at every collocation point of every element, all other
//                       elements of the fracture superposed   (calc_omega inside the Cauchy fits,
//                                                              hpc_fracture.py:51-92)
//   unwrap_kernel, dft_kernel (MFMA), fit_kernel
//                       Cauchy/DFT fit, element update, errors, damping (hpc_math_functions.py:243-297,
//                       537-670; element solves in hpc_intersection.py, hpc_bounding_circle.py,
//                       hpc_const_head_line.py, hpc_imp_object.py; hpc_solve.py:383-388)
//   grow_kernel         coefficient-growth test                 (hpc_solve.py:465-559)
//   blend_kernel        second damping step                     (hpc_solve.py:561-575)
//   reduce_kernel       max element error / coefficient change  (hpc_solve.py:196-197, 260-279)
//
// Layout in HBM (structure of arrays):
//   per element: type, frac0, frac1, ncoef, nint (int32); endpoints ep[4] (complex: ep0[0..1], ep1[0..1]);
//     radius, q, phi (f64); center (complex); coef and work coef [E x M] complex, M = coefficient
//     capacity; errors (f64); dpsi_corr [E x 2M] f64.
//   per fracture: t, constant (f64); member slots [mem_ptr[f], mem_ptr[f+1]) of Slot records that
//     carry the source's map (alpha, beta) in that fracture and its sign (-1 for an intersection
//     seen from its second fracture).
//   target sets: per fracture, the collocation (or discharge) points of its element sides are
//     packed contiguously and cut into 64-point tiles; one wavefront per tile walks the
//     fracture's sources, so every source parameter and coefficient is a wave-uniform scalar load.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "aem_kernels.hpp"
#include "aem_math.hpp"

namespace aem {

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// --------------------------------------------------------------------------------------
// Target points
// --------------------------------------------------------------------------------------
struct PointInfo {
  double2 z;
  int elem;     // element the point belongs to (excluded in the collocation sweep)
  int imag;     // 1 when the fit uses Im Omega (circles, impermeable line)
};

__device__ __forceinline__ int find_side(const int* __restrict__ pt_off, int lo, int hi, int p) {
  // largest s in [lo, hi) with pt_off[s] <= p
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pt_off[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// collocation point i of n on element e, side `which` (calc_thetas + map_chi_to_z_*,
// hpc_math_functions.py:673-699, hpc_geometry_functions.py:34-55, 80-99)
__device__ __forceinline__ double2 colloc_point(const NetView& net, int e, int which, int i, int n) {
  int ty = net.etype[e];
  if (ty == 1) {  // full circle: theta_i = 2 pi i / n
    double s, c;
    sincospi(2.0 * (double)i / (double)n, &s, &c);
    double r = net.radius[e];
    double2 ce = net.center[e];
    return c2(fma(r, c, ce.x), fma(r, s, ce.y));
  }
  double s, c;
  sincospi((double)(2 * i + 1) / (double)(2 * n), &s, &c);  // theta_i = (i + 1/2) pi / n
  if (ty == 4) {  // impermeable circle: half-circle angles on the full circle map
    double r = net.radius[e];
    double2 ce = net.center[e];
    return c2(fma(r, c, ce.x), fma(r, s, ce.y));
  }
  const double2* ep = net.ep + 4 * e + 2 * which;
  double2 a = ep[0], b = ep[1];
  // z = ((chi + 1/chi)/2 (b - a) + a + b)/2 with chi on the unit circle: (chi + 1/chi)/2 = cos theta
  return c2(0.5 * fma(c, b.x - a.x, a.x + b.x), 0.5 * fma(c, b.y - a.y, a.y + b.y));
}

// discharge point i of K on element e, side `which` (hpc_intersection.py:13-19,
// hpc_const_head_line.py:167-184, hpc_well.py:111-130)
__device__ __forceinline__ double2 discharge_point(const NetView& net, int e, int which, int i, int K) {
  int ty = net.etype[e];
  if (ty == 2) {
    double th = (K > 1) ? (double)i * ((2.0 * kPi - 2.0 * kPi / K) / (double)(K - 1)) : 0.0;
    double s, c;
    sincos(th, &s, &c);
    double r = net.radius[e];
    double2 ce = net.center[e];
    return c2(r * c + ce.x, r * s + ce.y);
  }
  const double2* ep = net.ep + 4 * e + 2 * which;
  double2 a = ep[0], b = ep[1];
  double2 step = c2((b.x - a.x) / (double)(K + 1), (b.y - a.y) / (double)(K + 1));
  return c2(a.x + (double)(i + 1) * step.x, a.y + (double)(i + 1) * step.y);
}

// --------------------------------------------------------------------------------------
// Sweep: Omega_f(z) = C_f + sum over the fracture's sources (hpc_fracture.py:51-92)
// --------------------------------------------------------------------------------------
// P target points per lane: a tile is 64 P consecutive points of one fracture, lane l owns
// points p0 + l + 64 k.  Every source's parameters and coefficients are loaded once per tile and
// applied to all P points (P independent arithmetic chains); each point's value is summed over
// the sources in the same order for every P, so the result does not depend on P.
template <int MODE, bool MIX, int P>
__device__ __forceinline__ void sweep_tile(const NetView net, const TargetView tg, int K, double* __restrict__ val,
                                           int t) {
  const Tile tile = tg.tiles[t];
  if (tile.p0 >= tile.pend) return;  // padding of the XCD layout (uniform per workgroup)
  const int lane = threadIdx.x;
  const int f = tile.frac;

  constexpr bool colloc_pts = (MODE == MODE_COLLOC || MODE == MODE_COLLOC_HEADS);
  constexpr bool explicit_pts = (MODE == MODE_POINTS || MODE == MODE_POINTS_OMEGA);
  constexpr bool use_q = (MODE != MODE_RHS);
  constexpr bool heads_out = (MODE == MODE_POINTS || MODE == MODE_COLLOC_HEADS || MODE == MODE_DISCH_HEADS);
  double2 z[P];
  int excl[P], want_imag[P];
  bool valid[P];
  int wi_any = 0;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int p = tile.p0 + lane + 64 * k;
    valid[k] = p < tile.pend;
    z[k] = c2(0.0, 0.0);
    excl[k] = -1;
    want_imag[k] = 0;
    if (valid[k]) {
      if (explicit_pts) {
        z[k] = tg.pt_z[p];
        if (MODE == MODE_POINTS_OMEGA) want_imag[k] = 1;  // the stream function too
      } else {
        const int sd = find_side(tg.pt_off, tile.sbeg, tile.send, p);
        const int e = tg.side_elem[sd];
        const int which = tg.side_which[sd];
        const int i = p - tg.pt_off[sd];
        if (colloc_pts) {
          z[k] = colloc_point(net, e, which, i, net.nint[e]);
          if (MODE == MODE_COLLOC) {
            excl[k] = e;
            const int ty = net.etype[e];
            want_imag[k] = (ty == 1 || ty == 4 || ty == 5);
          }
        } else {
          z[k] = discharge_point(net, e, which, i, K);
        }
      }
    }
    wi_any |= want_imag[k];
  }
  // the stream function (atan2 of every log term) only where a point of the same slot k on some
  // lane needs it: a fracture's bounding circle is its first side, so its points fill the first
  // lanes of slot 0 of the fracture's first tile and slot 1 skips the atan2
  bool any_imag[P];
#pragma unroll
  for (int k = 0; k < P; ++k) any_imag[k] = __any(want_imag[k]);
  (void)wi_any;

  double ore[P], oim[P];
  float fre[P], fim[P];  // FP32 accumulators of the far-field terms (mixed precision)
#pragma unroll
  for (int k = 0; k < P; ++k) {
    ore[k] = 0.0;
    oim[k] = 0.0;
    fre[k] = 0.0f;
    fim[k] = 0.0f;
  }
  const int kbeg = net.mem_ptr[f];
  const int kend = net.mem_ptr[f + 1];
  // the next source's slot, coefficient count and discharge are requested before this source's
  // work (a dependent chain of scalar loads otherwise paid at the head of every source)
  Slot sl_n = net.slots[kbeg < kend ? kbeg : 0];
  int m_n = kbeg < kend ? net.ncoef[sl_n.elem] : 0;
  double q_n = kbeg < kend ? net.q[sl_n.elem] : 0.0;
  for (int ks = kbeg; ks < kend; ++ks) {
    const Slot sl = sl_n;
    const int src = sl.elem;
    const int m = m_n;
    const double q_src = q_n;
    if (ks + 1 < kend) {
      sl_n = net.slots[ks + 1];
      m_n = net.ncoef[sl_n.elem];
      q_n = net.q[sl_n.elem];
    }
    const double2* __restrict__ cf = net.coef + (size_t)src * net.M;
    double tre[P], tim[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      tre[k] = 0.0;
      tim[k] = 0.0;
    }
    if (sl.kind == KIND_LINE || sl.kind == KIND_IMPL) {
      double2 Z[P];
      bool far = true;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        Z[k] = c2(fma(sl.alpha.x, z[k].x, fma(-sl.alpha.y, z[k].y, sl.beta.x)),
                  fma(sl.alpha.x, z[k].y, fma(sl.alpha.y, z[k].x, sl.beta.y)));
        far = far && (!valid[k] || fma(Z[k].x, Z[k].x, Z[k].y * Z[k].y) >= net.far2);
      }
      if (MIX && __all(far)) {
        // far field of this source for every point of the tile: FP32 map, series and log
        float2 chif[P], wf[P], hf[P];
        float r2f[P];
#pragma unroll
        for (int k = 0; k < P; ++k) {
          chif[k] = joukowsky_exterior_f(make_float2((float)Z[k].x, (float)Z[k].y));
          r2f[k] = fmaf(chif[k].x, chif[k].x, chif[k].y * chif[k].y);
          const float invf = 1.0f / r2f[k];
          wf[k] = make_float2(chif[k].x * invf, -chif[k].y * invf);
        }
        if (sl.kind == KIND_LINE && net.real_lines) horner_real_f_p<P>(net.coef32 + (size_t)src * net.M, m, wf, hf);
        else horner_f_p<P>(net.coef32 + (size_t)src * net.M, m, wf, hf);
        const float sg = (float)sl.sign;
        const float qq = (sl.kind == KIND_LINE && use_q) ? (float)(sl.sign * q_src * kInv2Pi) : 0.0f;
#pragma unroll
        for (int k = 0; k < P; ++k) {
          float gre = sg * hf[k].x, gim = sg * hf[k].y;
          if (qq != 0.0f) {
            gre = fmaf(qq, 0.5f * logf(r2f[k]), gre);
            if (any_imag[k]) gim = fmaf(qq, atan2f(chif[k].y, chif[k].x), gim);
          }
          if (src != excl[k]) {
            fre[k] += gre;
            fim[k] += gim;
          }
        }
        continue;
      }
      // the log term first, so chi and |chi|^2 are dead during the series
      const double qq = (sl.kind == KIND_LINE && use_q) ? sl.sign * q_src * kInv2Pi : 0.0;
      double2 w[P], h[P];
      double lg[P], at[P];
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 chi = joukowsky_exterior(Z[k]);
        const double r2 = fma(chi.x, chi.x, chi.y * chi.y);
        lg[k] = 0.0;
        at[k] = 0.0;
        if (qq != 0.0) {
          lg[k] = 0.5 * log_pos(r2);
          if (any_imag[k]) at[k] = atan2(chi.y, chi.x);
        }
        const double inv = rcp_pos(r2);
        w[k] = c2(chi.x * inv, -chi.y * inv);
      }
      if (sl.kind == KIND_LINE && net.real_lines) horner_real_p<P>(cf, m, w, h);
      else horner_p<P>(cf, m, w, h);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        tre[k] = sl.sign * h[k].x;
        tim[k] = sl.sign * h[k].y;
        if (qq != 0.0) {
          tre[k] = fma(qq, lg[k], tre[k]);
          if (any_imag[k]) tim[k] = fma(qq, at[k], tim[k]);
        }
      }
    } else if (sl.kind == KIND_BOUND) {
      double2 chi[P], h[P];
#pragma unroll
      for (int k = 0; k < P; ++k) chi[k] = c2(sl.alpha.x * z[k].x, sl.alpha.x * z[k].y);  // z / R
      horner_p<P>(cf, m, chi, h);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        tre[k] = h[k].x;
        tim[k] = h[k].y;
      }
    } else if (sl.kind == KIND_WELL) {
      const double qq = use_q ? q_src * kInv2Pi : 0.0;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 chi = c2(fma(sl.alpha.x, z[k].x, sl.beta.x), fma(sl.alpha.x, z[k].y, sl.beta.y));
        const double r2 = fma(chi.x, chi.x, chi.y * chi.y);
        if (qq != 0.0) {
          tre[k] = qq * (0.5 * log_pos(r2));
          if (any_imag[k]) tim[k] = qq * atan2(chi.y, chi.x);
        }
        if (sqrt(r2) < 1.0 - 1e-12) {
          tre[k] = __builtin_nan("");
          tim[k] = __builtin_nan("");
        }
      }
    } else {  // KIND_IMPC
      double2 w[P], h[P];
      double r2[P];
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const double2 chi = c2(fma(sl.alpha.x, z[k].x, sl.beta.x), fma(sl.alpha.x, z[k].y, sl.beta.y));
        r2[k] = fma(chi.x, chi.x, chi.y * chi.y);
        const double inv = rcp_pos(r2[k]);
        w[k] = c2(chi.x * inv, -chi.y * inv);
      }
      horner_p<P>(cf, m, w, h);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        tre[k] = h[k].x;
        tim[k] = h[k].y;
        if (sqrt(r2[k]) < 1.0 - 1e-10) {
          tre[k] = __builtin_nan("");
          tim[k] = __builtin_nan("");
        }
      }
    }
#pragma unroll
    for (int k = 0; k < P; ++k)
      if (src != excl[k]) {
        ore[k] += tre[k];
        oim[k] += tim[k];
      }
  }
#pragma unroll
  for (int k = 0; k < P; ++k) {
    if (MIX) {
      ore[k] += (double)fre[k];
      oim[k] += (double)fim[k];
    }
    if (valid[k]) {
      const int p = tile.p0 + lane + 64 * k;
      if (MODE == MODE_COLLOC) {
        val[p] = want_imag[k] ? oim[k] : ore[k] + net.constant[f];
      } else if (MODE == MODE_POINTS_OMEGA) {
        val[2 * (int64_t)p] = ore[k] + net.constant[f];
        val[2 * (int64_t)p + 1] = oim[k];
      } else if (heads_out)
        val[p] = (ore[k] + net.constant[f]) / net.t[f];
      else
        val[p] = ore[k];
    }
  }
}

template <int MODE, bool MIX, int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void sweep_kernel(NetView net, TargetView tg, int K,
                                                                                           double* __restrict__ val) {
  if (stopped(net)) return;
  sweep_tile<MODE, MIX, P>(net, tg, K, val, blockIdx.x);
}

// XCD (accelerator complex die) of the executing wavefront: hardware register XCC_ID, bits 3:0
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7; }

// Persistent form of the sweep: a grid sized to the resident wavefronts, each pulling tiles from
// its own XCD's list first (the host's XCD layout puts every fracture's tiles in one list, tile x
// + 8 t of list x, heaviest first), then from the other lists, so a fracture's coefficients stay
// in one L2 and the uneven fracture sizes balance across the dies.  queue[0..7] are the lists'
// heads, queue[8] counts finished workgroups; the last one resets all nine for the next launch.
template <int MODE, bool MIX, int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void sweep_persistent_kernel(
    NetView net, TargetView tg, int K, double* __restrict__ val, int per_xcd, unsigned* __restrict__ queue) {
  if (stopped(net)) return;  // uniform: no workgroup touches the queue
  const int home = xcc_id();
  for (int d = 0; d < 8; ++d) {
    const int x = (home + d) & 7;
    for (;;) {
      unsigned t = 0;
      if (threadIdx.x == 0) t = atomicAdd(&queue[x], 1u);
      t = (unsigned)__builtin_amdgcn_readfirstlane((int)t);
      if (t >= (unsigned)per_xcd) break;
      sweep_tile<MODE, MIX, P>(net, tg, K, val, x + 8 * (int)t);
    }
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[8], 1u) == gridDim.x - 1) {
      for (int x = 0; x < 9; ++x) queue[x] = 0u;
      __threadfence();
    }
  }
}

// coordinates and owner element of every point of a target set (used by the adjoint gathers)
__global__ __launch_bounds__(64) void points_kernel(NetView net, TargetView tg, int K, int colloc,
                                                    double2* __restrict__ pz, int* __restrict__ pelem) {
  const Tile tile = tg.tiles[blockIdx.x];
  const int p = tile.p0 + (int)threadIdx.x;
  if (p >= tile.pend) return;
  const int s = find_side(tg.pt_off, tile.sbeg, tile.send, p);
  const int e = tg.side_elem[s];
  const int which = tg.side_which[s];
  const int i = p - tg.pt_off[s];
  pz[p] = colloc ? colloc_point(net, e, which, i, net.nint[e]) : discharge_point(net, e, which, i, K);
  pelem[p] = e;
}

void launch_points(const NetView& net, const TargetView& tg, int ntiles, int K, int colloc, double2* pz, int* pelem,
                   hipStream_t st) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(points_kernel, dim3(ntiles), dim3(64), 0, st, net, tg, K, colloc, pz, pelem);
}

// --------------------------------------------------------------------------------------
// Fit (hpc_math_functions.py:243-297, 537-670 and the element solves), in three launches:
//   unwrap_kernel  branch-cut correction of circle-type sides, in place on the swept values
//   dft_kernel     the Cauchy/DFT sums of every side as a GEMM on the matrix cores: tiles of 16
//                  sides that share (angle family, n, m); D[16 x 2m] = Phi[16 x n] W[n x 2m],
//                  W = (cos j theta_i | sin j theta_i), v_mfma_f64_16x16x4f64 (FP32 form:
//                  v_mfma_f32_16x16x4f32 with FP32 accumulation)
//   fit_kernel     element solves, errors and damping from the sums, one wavefront per element
// --------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void unwrap_kernel(NetView net, FitView fv, const FitItem* __restrict__ items,
                                                    double* __restrict__ val, const int* __restrict__ pt_off) {
  __shared__ double sdpsi[kMaxPts];
  __shared__ double sd[kMaxPts];
  __shared__ double scv[64];
  __shared__ int scp[64];
  if (stopped(net)) return;
  const FitItem it = items[blockIdx.x];
  const int e = it.elem;
  const int ty = net.etype[e];
  if (!(ty == 1 || ty == 4 || ty == 5)) return;
  const int lane = threadIdx.x;
  const int n = net.nint[e];
  // kappa_i = sum q * sign over the cached crossings, then unwrap (get_dpsi_corr
  // hpc_math_functions.py:406-447; cauchy_integral_domega 583-595).  The lanes fetch the
  // crossings and the values; the two running sums stay sequential in lane 0, in the reference's
  // order, over values already in LDS.
  for (int i = lane; i < n; i += 64) sdpsi[i] = 0.0;
  const int cb = fv.cross_ptr[e];
  const int cn = fv.cross_cnt[e];
  for (int c0 = 0; c0 < cn; c0 += 64) {
    __syncthreads();
    if (c0 + lane < cn) {
      const int2 rec = fv.cross[cb + c0 + lane];  // (position, element); sign in cross_sign
      scp[lane] = rec.x;
      scv[lane] = net.q[rec.y] * (double)fv.cross_sign[cb + c0 + lane];
    }
    __syncthreads();
    if (lane == 0)
      for (int c = 0; c < min(64, cn - c0); ++c) sdpsi[scp[c]] += scv[c];
  }
  __syncthreads();
  double* v = val + pt_off[it.s0];
  double* gd = fv.dpsi_corr + (size_t)e * 2 * net.M;
  // d_i = (v_i - v_{i-1}) - kappa_{i-1}, then v_i = v_0 + d_1 + ... + d_i summed in order
  for (int i = lane; i < n - 1; i += 64) {
    gd[i] = sdpsi[i];
    sd[i + 1] = (v[i + 1] - v[i]) - sdpsi[i];
  }
  __syncthreads();
  if (lane == 0) {
    double psi0 = v[0];
    for (int i = 1; i < n; ++i) {
      psi0 = psi0 + sd[i];
      sd[i] = psi0;
    }
  }
  __syncthreads();
  for (int i = lane + 1; i < n; i += 64) v[i] = sd[i];
}

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// One wavefront per tile (4 per workgroup).  MFMA 16x16x4 operand layout (lane l, measured on
// gfx950): A[l % 16][l / 16] (side l % 16, point k0 + l / 16), B[l / 16][l % 16] (point, column);
// result register r holds D[l / 16 + 4 r][l % 16] for the f64 instruction and
// D[4 (l / 16) + r][l % 16] for the f32 one.
// Column c < m is X_c = sum_i v_i cos(c theta_i), column m + c is Y_c = sum_i v_i sin(c theta_i);
// theta_i = (2 i + 1) pi / (2 n) (family 0) or 2 pi i / n (family 1), the angles of
// calc_thetas (hpc_math_functions.py:673-699), reduced exactly before sincospi.
template <bool F32>
__global__ __launch_bounds__(256) void dft_kernel(const DftTile* __restrict__ tiles, int ntiles,
                                                  const int* __restrict__ sides, const double* __restrict__ val,
                                                  const int* __restrict__ pt_off, int M, double* __restrict__ dft,
                                                  const double2* __restrict__ trig, const float2* __restrict__ trig32) {
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= ntiles) return;  // the whole wavefront
  const DftTile tl = tiles[tile];
  const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
  const int n = tl.n, m = tl.m, ncol = 2 * m;
  const int side_a = r < tl.ns ? sides[tl.s0 + r] : -1;
  const double* va = side_a >= 0 ? val + pt_off[side_a] : nullptr;
  const int64_t tbase = 2 * (int64_t)n * (n - 1);  // this n's rows of the (cos, sin) tables
  for (int c0 = 0; c0 < ncol; c0 += 16) {
    const int col = c0 + r;
    const bool colok = col < ncol;
    const bool sinpart = col >= m;
    const int j = sinpart ? col - m : col;
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    f32x4 accf = {0.0f, 0.0f, 0.0f, 0.0f};
    // angle j theta_i = kk pi / (2 n): kk = 4 (i j mod n) on the full circle (theta_i = 2 pi i / n),
    // kk = j (2 i + 1) mod 4 n at the midpoint angles (theta_i = (2 i + 1) pi / (2 n)); i steps by
    // 4 per pass, so kk steps by 4 (4 j mod n) or 8 j mod 4 n (below the period: one subtraction)
    const int period = 4 * n;
    const int d = tl.family == 1 ? 4 * ((4 * j) % n) : (8 * j) % period;
    int kk = tl.family == 1 ? 4 * ((kq * j) % n) : (j * (2 * kq + 1)) % period;
    for (int k0 = 0; k0 < n; k0 += 4) {
      const int i = k0 + kq;
      const bool iok = i < n;
      const double a = (va != nullptr && iok) ? va[i] : 0.0;
      if (F32) {
        float w = 0.0f;
        if (colok && iok) {
          const float2 t = trig32[tbase + kk];
          w = sinpart ? t.y : t.x;
        }
        accf = __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, w, accf, 0, 0, 0);
      } else {
        double w = 0.0;
        if (colok && iok) {
          const double2 t = trig[tbase + kk];
          w = sinpart ? t.y : t.x;
        }
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, w, acc, 0, 0, 0);
      }
      kk += d;
      if (kk >= period) kk -= period;
    }
    if (colok) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = F32 ? 4 * kq + q : kq + 4 * q;
        if (row < tl.ns) {
          const int side = sides[tl.s0 + row];
          dft[(size_t)side * 2 * M + (sinpart ? M + j : j)] = F32 ? (double)accf[q] : acc[q];
        }
      }
    }
  }
}

// (cos, sin) of k pi / (2 n) for k < 4 n, block row y = n - 1 (the same sincospi arguments the
// DFT used to form per lane, so the weights are unchanged to the bit)
__global__ void trig_table_kernel(int nmax, double2* __restrict__ trig, float2* __restrict__ trig32) {
  const int n = blockIdx.y + 1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (n > nmax || k >= 4 * n) return;
  const int64_t e = 2 * (int64_t)n * (n - 1) + k;
  double s, c;
  sincospi((double)k / (double)(2 * n), &s, &c);
  trig[e] = make_double2(c, s);
  float sf, cf;
  sincospif((float)k / (float)(2 * n), &sf, &cf);
  trig32[e] = make_float2(cf, sf);
}

void launch_trig_tables(int nmax, double2* trig, float2* trig32, hipStream_t st) {
  if (nmax <= 0) return;
  hipLaunchKernelGGL(trig_table_kernel, dim3((unsigned)((4 * nmax + 255) / 256), (unsigned)nmax), dim3(256), 0, st,
                     nmax, trig, trig32);
}

// sum over the G lanes of a group (G a power of two <= 64, groups aligned to G lanes): the last
// log2 G steps of wave_sum, so for a group whose other lanes hold zeros the result is the same
template <int G>
__device__ __forceinline__ double group_sum(double v) {
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Cauchy/DFT fit of one element per group of G lanes (lanes over its coefficients), 256 / G
// elements per workgroup.  blend: 0 = the fit only; 1 = also the first damping step; 2 = both
// damping steps of the iteration (hpc_solve.py:561-575: the one blend_kernel would apply next)
template <int G>
__global__ __launch_bounds__(256) void fit_kernel(NetView net, FitView fv, const FitItem* __restrict__ items,
                                                  int nitems, const double* __restrict__ dft, double damping,
                                                  int blend) {
  if (stopped(net)) return;
  const int b = blockIdx.x * (256 / G) + (int)threadIdx.x / G;
  if (b >= nitems) return;  // the whole group
  const FitItem it = items[b];
  const int e = it.elem;
  const int lane = (int)threadIdx.x % G;
  const int ty = net.etype[e];
  const int m = net.ncoef[e];
  const int n = net.nint[e];
  const double* d0 = dft + (size_t)it.s0 * 2 * net.M;
  const double* d1 = dft + (size_t)(ty == 0 ? it.s1 : it.s0) * 2 * net.M;

  // per-lane coefficients j = lane, lane + G, ...
  double2* wc = net.wcoef + (size_t)e * net.M;
  double2* cf = net.coef + (size_t)e * net.M;
  double s_new_re = 0.0, s_new_im = 0.0, s_tail_re = 0.0, s_tail_im = 0.0, s_old5 = 0.0, s_diff = 0.0;
  const int jtail = (int)(m * 0.9);
  const double t0 = (ty == 0) ? net.t[net.efrac0[e]] : 0.0;
  const double t1 = (ty == 0) ? net.t[net.efrac1[e]] : 0.0;
  for (int j = lane; j < m; j += G) {
    const double x0 = d0[j], y0 = d0[net.M + j], x1 = ty == 0 ? d1[j] : 0.0;
    const double two_n = 2.0 / (double)n;
    double2 cnew;
    if (ty == 0) {
      // c_j = Re[(t0 a1_j - t1 a0_j)/(t0 + t1)], a_j = (2/n) X_j (hpc_intersection.py:95-101)
      double a0 = two_n * x0, a1 = two_n * x1;
      cnew = c2(j == 0 ? 0.0 : (t0 * a1 - t1 * a0) / (t0 + t1), 0.0);
    } else if (ty == 3) {
      cnew = c2(j == 0 ? 0.0 : -(two_n * x0), 0.0);  // hpc_const_head_line.py:71-75
    } else {
      // b_j = (2i/n) sum psi e^{-ij theta} = (2/n)(Y_j + i X_j); b_0 halved
      double bre = two_n * y0, bim = two_n * x0;
      if (j == 0) {
        bre *= 0.5;
        bim *= 0.5;
      }
      if (ty == 1) cnew = c2(-bre, -bim);                           // hpc_bounding_circle.py:151
      else if (ty == 4) cnew = c2(j == 0 ? 0.0 : bre, j == 0 ? 0.0 : -bim);  // conj, c_0 = 0
      else cnew = c2(0.0, j == 0 ? 0.0 : -bim);                     // -i Im b, c_0 = 0
    }
    const double2 old = cf[j];
    wc[j] = cnew;
    s_new_re += cnew.x;
    s_new_im += cnew.y;
    if (j >= jtail) {
      s_tail_re += cnew.x;
      s_tail_im += cnew.y;
    }
    if (j < 5) s_old5 += hypot(old.x, old.y);
    s_diff += hypot(cnew.x - old.x, cnew.y - old.y);
    if (blend) {
      double2 c = c2(damping * cnew.x + (1.0 - damping) * old.x, damping * cnew.y + (1.0 - damping) * old.y);
      if (blend == 2) c = c2(damping * cnew.x + (1.0 - damping) * c.x, damping * cnew.y + (1.0 - damping) * c.y);
      cf[j] = c;
    }
  }
  s_new_re = group_sum<G>(s_new_re);
  s_new_im = group_sum<G>(s_new_im);
  s_tail_re = group_sum<G>(s_tail_re);
  s_tail_im = group_sum<G>(s_tail_im);
  s_old5 = group_sum<G>(s_old5);
  s_diff = group_sum<G>(s_diff);
  if (lane == 0) {
    double err = 0.0, errc = 0.0;
    if (!(s_old5 < 1e-15)) {
      double a = hypot(s_new_re - s_tail_re, s_new_im - s_tail_im);
      double b = hypot(s_new_re + 1e-16, s_new_im);
      err = fabs(a / b - 1.0);
      errc = s_diff / (double)m;
    }
    fv.error_old2[e] = fv.error_old[e];
    fv.error_old[e] = fv.error[e];
    fv.error[e] = err;
    fv.error_coef[e] = errc;
  }
}

// second damping step for every element (hpc_solve.py:561-575); also the only update of wells
__global__ void blend_kernel(NetView net, int ne, double damping) {
  const int e = blockIdx.x;
  if (e >= ne || stopped(net)) return;
  const int m = net.ncoef[e];
  double2* cf = net.coef + (size_t)e * net.M;
  const double2* wc = net.wcoef + (size_t)e * net.M;
  for (int j = threadIdx.x; j < m; j += blockDim.x) {
    double2 a = wc[j], b = cf[j];
    cf[j] = c2(damping * a.x + (1.0 - damping) * b.x, damping * a.y + (1.0 - damping) * b.y);
  }
}

// wells: error = 0 and the first damping step (their work coefficients stay zero); steps = 2 also
// applies the second (the one blend_kernel would apply next)
__global__ void well_kernel(NetView net, FitView fv, const int* __restrict__ wells, int nw, double damping,
                            int steps) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw || stopped(net)) return;
  const int e = wells[k];
  fv.error[e] = 0.0;
  const int m = net.ncoef[e];
  double2* cf = net.coef + (size_t)e * net.M;
  const double2* wc = net.wcoef + (size_t)e * net.M;
  for (int j = 0; j < m; ++j) {
    double2 a = wc[j], b = cf[j];
    for (int r = 0; r < steps; ++r)
      b = c2(damping * a.x + (1.0 - damping) * b.x, damping * a.y + (1.0 - damping) * b.y);
    cf[j] = b;
  }
}

// coefficient-growth test on the damped coefficients (hpc_solve.py:465-559); grows in place
__global__ void grow_kernel(NetView net, FitView fv, int ne, GrowParams gp, const int* __restrict__ sched,
                            int* __restrict__ grow) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  grow[e] = 0;
  if (net.etype[e] == 2) return;
  const int m = net.ncoef[e];
  if (sched) {  // replay of a recorded trajectory: grow where the recorded count rises
    if (sched[e] > m && m < gp.max_coef) {
      net.ncoef_w[e] = m + gp.increase;
      net.nint_w[e] = 2 * (m + gp.increase);
      grow[e] = 1;
    }
    return;
  }
  const double2* c = net.coef + (size_t)e * net.M;
  const double err = fv.error[e];
  double ratio = 0.0;  // the reference's first ratio is multiplied by zero (hpc_solve.py:474-478)
  if (err > fv.error_old[e] && err > fv.error_old2[e] && err > gp.max_error && gp.nit > 5) ratio = gp.max_ratio + 1.0;
  double tre = 0.0, tim = 0.0, hre = 0.0, him = 0.0;
  const int k3 = m > 3 ? m - 3 : 0;
  for (int k = k3; k < m; ++k) {
    tre += c[k].x;
    tim += c[k].y;
  }
  for (int k = 0; k < k3; ++k) {
    hre += c[k].x;
    him += c[k].y;
  }
  hre += 1e-30;
  // |tail / head| = |tail| / |head|
  if (hypot(tre, tim) / hypot(hre, him) > gp.max_ratio && err > gp.max_error) ratio = gp.max_ratio + 1.0;
  if (ratio > gp.max_ratio && m < gp.max_coef && gp.nit > 1) {
    const int nm = m + gp.increase;
    net.ncoef_w[e] = nm;
    net.nint_w[e] = 2 * nm;
    grow[e] = 1;
  }
}

// --------------------------------------------------------------------------------------
// Branch-cut crossings of circle-type elements (find_branch_cuts, hpc_math_functions.py:300-392)
// one wavefront per element; lanes over the collocation edges (i, i+1), i < n - 1
// --------------------------------------------------------------------------------------
__device__ __forceinline__ double2 line_chi(const double2* ep, double2 z) {
  double2 a = ep[0], b = ep[1];
  double2 num = c2(2.0 * z.x - a.x - b.x, 2.0 * z.y - a.y - b.y);
  double2 den = c2(b.x - a.x, b.y - a.y);
  double dd = den.x * den.x + den.y * den.y;
  double2 Z = c2((num.x * den.x + num.y * den.y) / dd, (num.y * den.x - num.x * den.y) / dd);
  return joukowsky_exterior(Z);
}

__device__ __forceinline__ double npsign(double x) { return (double)((x > 0) - (x < 0)); }

__global__ __launch_bounds__(64) void cross_kernel(NetView net, FitView fv, const int* __restrict__ circles, int nc) {
  const int k = blockIdx.x;
  if (k >= nc) return;
  const int e = circles[k];
  const int lane = threadIdx.x;
  const int ty = net.etype[e];
  const int n = net.nint[e];
  const int f = net.efrac0[e];
  const int cb = fv.cross_ptr[e];
  const int cap = fv.cross_cap[e];
  int cnt = 0;
  const int kbeg = net.mem_ptr[f], kend = net.mem_ptr[f + 1];
  for (int kk = kbeg; kk < kend; ++kk) {
    const int src = net.slots[kk].elem;
    const int sty = net.etype[src];
    if (!(sty == 0 || sty == 2 || sty == 3)) continue;
    for (int i0 = 0; i0 < n - 1; i0 += 64) {
      const int i = i0 + lane;
      int hit = 0, sgn = 0;
      if (i < n - 1) {
        double2 za = colloc_point(net, e, 0, i, n);
        double2 zb = colloc_point(net, e, 0, i + 1, n);
        double2 chi0, chi1;
        int sdef = -1;
        if (sty == 0) {
          const bool own0 = net.efrac0[src] == f;
          const double2* ep = net.ep + 4 * src + (own0 ? 0 : 2);
          chi0 = line_chi(ep, za);
          chi1 = line_chi(ep, zb);
          double ln0 = atan2(chi0.y, chi0.x), ln1 = atan2(chi1.y, chi1.x);
          hit = (npsign(ln0) != npsign(ln1)) && (fabs(ln0) + fabs(ln1) > kPi);
          sdef = own0 ? -1 : 1;
        } else {
          if (sty == 2) {
            double r = net.radius[src];
            double2 ce = net.center[src];
            chi0 = c2((za.x - ce.x) / r, (za.y - ce.y) / r);
            chi1 = c2((zb.x - ce.x) / r, (zb.y - ce.y) / r);
          } else {
            chi0 = line_chi(net.ep + 4 * src, za);
            chi1 = line_chi(net.ep + 4 * src, zb);
          }
          hit = (npsign(chi0.y) != npsign(chi1.y)) && (chi0.x < 0.0);
        }
        if (hit) sgn = (ty != 1 && chi0.y < chi1.y) ? -sdef : sdef;
      }
      const unsigned long long bal = __ballot(hit);
      if (bal) {
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        const int slot = cnt + before;
        if (hit && slot < cap) {
          fv.cross[cb + slot] = make_int2(i, src);
          fv.cross_sign[cb + slot] = sgn;
        }
        cnt += __popcll(bal);
      }
    }
  }
  if (lane == 0) {
    fv.cross_cnt[e] = cnt < cap ? cnt : cap;
    if (cnt > cap) atomicOr(fv.status, 1);
  }
}

// --------------------------------------------------------------------------------------
// Right-hand side (build_head_matrix, hpc_solve.py:812-829)
// --------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void rhs_reduce_kernel(NetView net, const RhsItem* __restrict__ items, int nde,
                                                        const double* __restrict__ val, const int* __restrict__ pt_off,
                                                        int K, const int* __restrict__ de_pos,
                                                        double* __restrict__ b) {
  const int j = blockIdx.x;
  if (j >= nde || stopped(net)) return;
  const RhsItem it = items[j];
  const int lane = threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  const int b0 = pt_off[it.s0];
  for (int i = lane; i < K; i += 64) s0 += val[b0 + i];
  if (it.s1 >= 0) {
    const int b1 = pt_off[it.s1];
    for (int i = lane; i < K; i += 64) s1 += val[b1 + i];
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (lane == 0) {
    const int e = it.elem;
    const double om0 = s0 / (double)K;
    if (net.etype[e] == 0) {
      const double om1 = s1 / (double)K;
      b[de_pos[e]] = om1 / net.t[net.efrac1[e]] - om0 / net.t[net.efrac0[e]];
    } else {
      b[de_pos[e]] = net.phi[e] - om0;
    }
  }
}

// --------------------------------------------------------------------------------------
// Right-hand side as a linear map of the coefficients.  With q = C = 0 the potential at a
// discharge point is the sum of the sources' series, so the mean over the K points of side d is
//   om_d = sum over the sources s of the side's fracture, j < ncoef_s, of Re(c_sj W_dsj),
//   W_dsj = sign_s / K sum_k w_s(z_dk)^j
// (w_s the source's series variable at the point, as in the sweep).  W depends on the geometry
// only: rhs_basis_kernel builds it once per handle, rhs_eval_kernel applies it every iteration
// and forms b as rhs_reduce_kernel does from the swept values.  A side with a point inside a
// well or an impermeable circle gets NaN, as the sweep gives there.
// --------------------------------------------------------------------------------------
constexpr int kRhsMaxQ = 8;  // discharge points per lane (K <= 512)

__global__ __launch_bounds__(64) void rhs_basis_kernel(NetView net, TargetView tg, int K,
                                                       const int64_t* __restrict__ wptr, const int* __restrict__ mcap,
                                                       double* __restrict__ W, int* __restrict__ side_nan) {
  const int d = blockIdx.x;
  const int lane = threadIdx.x;
  const int e = tg.side_elem[d];
  const int which = tg.side_which[d];
  const int f = which ? net.efrac1[e] : net.efrac0[e];
  double2 z[kRhsMaxQ];
  bool ok[kRhsMaxQ];
#pragma unroll
  for (int q = 0; q < kRhsMaxQ; ++q) {
    const int i = lane + 64 * q;
    ok[q] = i < K;
    z[q] = ok[q] ? discharge_point(net, e, which, i, K) : c2(0.0, 0.0);
  }
  const double invK = 1.0 / (double)K;
  int inside = 0;
  int64_t o = wptr[d];
  for (int ks = net.mem_ptr[f]; ks < net.mem_ptr[f + 1]; ++ks) {
    const Slot sl = net.slots[ks];
    const int src = sl.elem;
    if (sl.kind == KIND_WELL) {
#pragma unroll
      for (int q = 0; q < kRhsMaxQ; ++q)
        if (ok[q]) {
          const double2 chi = c2(fma(sl.alpha.x, z[q].x, sl.beta.x), fma(sl.alpha.x, z[q].y, sl.beta.y));
          if (sqrt(fma(chi.x, chi.x, chi.y * chi.y)) < 1.0 - 1e-12) inside = 1;
        }
      continue;
    }
    double2 w[kRhsMaxQ], pw[kRhsMaxQ];
#pragma unroll
    for (int q = 0; q < kRhsMaxQ; ++q) {
      w[q] = c2(0.0, 0.0);
      pw[q] = c2(ok[q] ? 1.0 : 0.0, 0.0);
      if (!ok[q]) continue;
      if (sl.kind == KIND_BOUND) {
        w[q] = c2(sl.alpha.x * z[q].x, sl.alpha.x * z[q].y);  // z / R
      } else if (sl.kind == KIND_IMPC) {
        const double2 chi = c2(fma(sl.alpha.x, z[q].x, sl.beta.x), fma(sl.alpha.x, z[q].y, sl.beta.y));
        const double r2 = fma(chi.x, chi.x, chi.y * chi.y);
        if (sqrt(r2) < 1.0 - 1e-10) inside = 1;
        const double inv = rcp_pos(r2);
        w[q] = c2(chi.x * inv, -chi.y * inv);
      } else {  // KIND_LINE, KIND_IMPL
        const double2 Z = c2(fma(sl.alpha.x, z[q].x, fma(-sl.alpha.y, z[q].y, sl.beta.x)),
                             fma(sl.alpha.x, z[q].y, fma(sl.alpha.y, z[q].x, sl.beta.y)));
        const double2 chi = joukowsky_exterior(Z);
        const double inv = rcp_pos(fma(chi.x, chi.x, chi.y * chi.y));
        w[q] = c2(chi.x * inv, -chi.y * inv);
      }
    }
    const double sg = (sl.kind == KIND_LINE || sl.kind == KIND_IMPL) ? sl.sign * invK : invK;
    const int m = mcap[src];
    // Re(c W) = Re c Re W - Im c Im W; real coefficients need the first weight only
    const bool real_c = sl.kind == KIND_LINE && net.real_lines;
    for (int j = 0; j < m; ++j) {
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int q = 0; q < kRhsMaxQ; ++q) {
        ar += pw[q].x;
        ai += pw[q].y;
        const double2 t = pw[q];
        pw[q] = c2(fma(t.x, w[q].x, -t.y * w[q].y), fma(t.x, w[q].y, t.y * w[q].x));
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      if (lane == 0) {
        if (real_c) {
          W[o + j] = sg * ar;
        } else {
          W[o + 2 * j] = sg * ar;
          W[o + 2 * j + 1] = -(sg * ai);
        }
      }
    }
    o += real_c ? m : 2 * m;
  }
  inside = __any(inside) ? 1 : 0;
  if (lane == 0) side_nan[d] = inside;
}

// om_d: the entries of side d are the columns of its fracture's table col[coff[d] ...], the same
// for every side of that fracture, so the table and the gathered coefficients stay in cache while
// W streams from HBM once per iteration
template <class WT>
__device__ __forceinline__ double rhs_side_mean(const NetView& net, int d, const int64_t* __restrict__ wptr,
                                                const WT* __restrict__ W, const unsigned* __restrict__ col,
                                                const int64_t* __restrict__ coff, const int* __restrict__ side_nan) {
  // four entries per lane in flight (column, count and coefficient loads are a dependent chain)
  const double* __restrict__ cre = reinterpret_cast<const double*>(net.coef);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int64_t w0 = wptr[d], t1 = wptr[d + 1], c0 = coff[d] - w0;
  for (int64_t t0 = w0 + threadIdx.x; t0 < t1; t0 += 256) {
    unsigned q[4];
    WT w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t t = t0 + 64 * u;
      q[u] = t < t1 ? col[c0 + t] : 0xffffffffu;
      w[u] = t < t1 ? W[t] : WT(0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int src = (int)(q[u] >> kRhsColShift);
      const int jj = (int)(q[u] & ((1u << kRhsColShift) - 1));
      if (q[u] != 0xffffffffu && (jj >> 1) < net.ncoef[src])
        acc[u] = fma(cre[2 * ((size_t)src * net.M + (jj >> 1)) + (jj & 1)], (double)w[u], acc[u]);
    }
  }
  double a = wave_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
  return side_nan[d] ? __builtin_nan("") : a;
}

template <class WT>
__global__ __launch_bounds__(64) void rhs_eval_kernel(NetView net, const RhsItem* __restrict__ items, int nde,
                                                      const int64_t* __restrict__ wptr, const WT* __restrict__ W,
                                                      const unsigned* __restrict__ col, const int64_t* __restrict__ coff,
                                                      const int* __restrict__ side_nan, const int* __restrict__ de_pos,
                                                      double* __restrict__ b) {
  const int j = blockIdx.x;
  if (j >= nde || stopped(net)) return;
  const RhsItem it = items[j];
  const double om0 = rhs_side_mean(net, it.s0, wptr, W, col, coff, side_nan);
  const double om1 = it.s1 >= 0 ? rhs_side_mean(net, it.s1, wptr, W, col, coff, side_nan) : 0.0;
  if (threadIdx.x == 0) {
    const int e = it.elem;
    if (net.etype[e] == 0) b[de_pos[e]] = om1 / net.t[net.efrac1[e]] - om0 / net.t[net.efrac0[e]];
    else b[de_pos[e]] = net.phi[e] - om0;
  }
}

__global__ void w_to_f32_kernel(int64_t n, const double* __restrict__ W, float* __restrict__ o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (float)W[i];
}

void launch_rhs_basis(const NetView& net, const TargetView& tg, int nsides, int K, const int64_t* wptr,
                      const int* mcap, double* W, int* side_nan, hipStream_t st) {
  if (nsides <= 0) return;
  hipLaunchKernelGGL(rhs_basis_kernel, dim3(nsides), dim3(64), 0, st, net, tg, K, wptr, mcap, W, side_nan);
}

void launch_rhs_to_f32(int64_t n, const double* W, float* W32, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(w_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, W, W32);
}

void launch_rhs_eval(const NetView& net, const RhsItem* items, int nde, const int64_t* wptr, const double* W,
                     const float* W32, const unsigned* col, const int64_t* coff, const int* side_nan,
                     const int* de_pos, double* b, hipStream_t st) {
  if (nde <= 0) return;
  if (W32)
    hipLaunchKernelGGL(rhs_eval_kernel<float>, dim3(nde), dim3(64), 0, st, net, items, nde, wptr, W32, col, coff,
                       side_nan, de_pos, b);
  else
    hipLaunchKernelGGL(rhs_eval_kernel<double>, dim3(nde), dim3(64), 0, st, net, items, nde, wptr, W, col, coff,
                       side_nan, de_pos, b);
}

// --------------------------------------------------------------------------------------
// Discharge matrix (fill_discharge_matrix, hpc_solve.py:1117-1201; discharge terms
// hpc_intersection.py:22-46, hpc_well.py:14-35, hpc_const_head_line.py:14-35)
// one wavefront per (discharge row, side), into the dense column-major blocks or into the values
// of the matrix's CSR pattern (the same pattern the head residual uses)
// --------------------------------------------------------------------------------------
struct DenseTarget {
  double* A;
  SysLayout sys;
  __device__ double* at(int64_t r, int64_t c) const { return A + mat_at(r, c, sys); }
};

struct CsrTarget {
  double* val;
  const int* ptr;
  const int* col;
  // (r, c) is in the pattern by construction: a binary search over the row's sorted columns
  __device__ double* at(int64_t r, int64_t c) const {
    int lo = ptr[r], hi = ptr[r + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (col[mid] < c) lo = mid + 1;
      else hi = mid;
    }
    return val + lo;
  }
};

template <class Tg>
__global__ __launch_bounds__(64) void assemble_kernel(NetView net, TargetView tg, SysLayout sys, int K, Tg A) {
  const int s = blockIdx.x;
  const int e = tg.side_elem[s];
  const int which = tg.side_which[s];
  const int64_t row = sys.de_pos[e];
  const int f = which ? net.efrac1[e] : net.efrac0[e];
  const int lane = threadIdx.x;
  const bool inter = net.etype[e] == 0;
  const double rs = inter ? (which ? -1.0 : 1.0) / net.t[f] : 1.0;
  for (int k = net.mem_ptr[f]; k < net.mem_ptr[f + 1]; ++k) {
    const Slot sl = net.slots[k];
    const int src = sl.elem;
    if (src == e || !(sl.kind == KIND_LINE || sl.kind == KIND_WELL)) continue;
    double acc = 0.0;
    for (int i = lane; i < K; i += 64) {
      double2 z = discharge_point(net, e, which, i, K);
      double r2;
      if (sl.kind == KIND_LINE) {
        double2 Z = c2(fma(sl.alpha.x, z.x, fma(-sl.alpha.y, z.y, sl.beta.x)),
                       fma(sl.alpha.x, z.y, fma(sl.alpha.y, z.x, sl.beta.y)));
        double2 chi = joukowsky_exterior(Z);
        r2 = fma(chi.x, chi.x, chi.y * chi.y);
      } else {
        double2 chi = c2(fma(sl.alpha.x, z.x, sl.beta.x), fma(sl.alpha.x, z.y, sl.beta.y));
        r2 = fma(chi.x, chi.x, chi.y * chi.y);
      }
      acc += (sl.kind == KIND_LINE ? sl.sign : 1.0) * kInv2Pi * 0.5 * log_pos(r2);
    }
    acc = wave_sum(acc);
    // at most two contributions meet in one entry (the two sides of an intersection that shares
    // both fractures with the source), and two-term sums are exact in either order
    if (lane == 0) atomicAdd(A.at(row, sys.de_pos[src]), rs * acc / (double)K);
  }
  if (lane == 0) {
    if (inter) {
      if (which == 0) atomicAdd(A.at(row, sys.fpos[net.efrac0[e]]), 1.0 / net.t[net.efrac0[e]]);
      else atomicAdd(A.at(row, sys.fpos[net.efrac1[e]]), -1.0 / net.t[net.efrac1[e]]);
    } else {
      atomicAdd(A.at(row, sys.fpos[net.efrac0[e]]), 1.0);
    }
  }
}

// fracture continuity rows: sum_e s_e q_e = 0 (hpc_solve.py:1187-1201)
template <class Tg>
__global__ void continuity_kernel(NetView net, SysLayout sys, int nf, Tg A) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int64_t row = sys.fpos[f];
  for (int k = net.mem_ptr[f]; k < net.mem_ptr[f + 1]; ++k) {
    const int e = net.slots[k].elem;
    const int ty = net.etype[e];
    if (!(ty == 0 || ty == 2 || ty == 3)) continue;
    const double v = (ty != 0 || net.efrac0[e] == f) ? 1.0 : -1.0;
    *A.at(row, sys.de_pos[e]) += v;
  }
}

// row equilibration scales of a CSR matrix: rs[r] = 1 / max_k |A[r, k]| (1 for an empty row)
__global__ void csr_row_scale_kernel(int64_t n, const int* __restrict__ ptr, const double* __restrict__ val,
                                     double* __restrict__ rs) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double m = 0.0;
  for (int q = ptr[r]; q < ptr[r + 1]; ++q) m = fmax(m, fabs(val[q]));
  rs[r] = m > 0.0 ? 1.0 / m : 1.0;
}

// identity rows/columns of the padding unknowns of every block
__global__ void pad_identity_kernel(double* __restrict__ A, int64_t npad, int nblk, const int* __restrict__ blk_n) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= npad * nblk) return;
  const int64_t b = g / npad, i = g - b * npad;
  if (i >= blk_n[b]) A[b * npad * npad + i * npad + i] = 1.0;
}

// --------------------------------------------------------------------------------------
// discharge scatter and relative change (hpc_solve.py:169-174, post_matrix_solve 720-761)
// --------------------------------------------------------------------------------------
__global__ void scatter_kernel(NetView net, const int* __restrict__ de, int nde, int nf, SysLayout sys,
                               const double* __restrict__ y) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nde) net.q_w[de[k]] = y[sys.de_pos[de[k]]];
  else if (k < nde + nf) net.constant_w[k - nde] = y[sys.fpos[k - nde]];
}

// accept a discharge solve x: y_old <- y, y <- x, and q and the fracture constants from x (the
// scatter of hpc_solve.py:720-761), one launch
__global__ void accept_kernel(NetView net, const int* __restrict__ de, int nde, int nf, SysLayout sys,
                              const double* __restrict__ x, double* __restrict__ y, double* __restrict__ yold,
                              int64_t n) {
  if (stopped(net)) return;
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    yold[k] = y[k];
    y[k] = x[k];
  }
  if (k < nde) net.q_w[de[k]] = x[sys.de_pos[de[k]]];
  else if (k < nde + nf) net.constant_w[k - nde] = x[sys.fpos[k - nde]];
}

__device__ __forceinline__ void nanmax_merge(double& a, int& ia, double b, int ib) {
  // np.max / np.argmax semantics: NaN wins, first index on ties
  const bool an = a != a, bn = b != b;
  if (an) {
    if (bn && ib < ia) ia = ib;
    return;
  }
  if (bn || b > a || (b == a && ib < ia)) {
    a = b;
    ia = ib;
  }
}

// one block of 1024 threads: max error (with argmax), max error_coef, max |y - y_old|, max |y_old|
// Two-stage reductions: each workgroup reduces a contiguous chunk and writes its partial; the
// last workgroup to finish (atomic counter, reset by that workgroup for the next launch) merges
// the partials, one per thread, by the same tree.  Every merge keeps the larger value, a NaN
// over any number, and the lower index between equal values or two NaNs, so the result does
// not depend on the merge order: ties resolve to the first index as in np.argmax.
constexpr int kRedThreads = 256;
constexpr int kRedMaxGroups = 256;  // <= kRedThreads: the last workgroup merges one partial per thread
struct RedPart {
  double e, c, d, o;
  int ie, pad;
};
struct RedScratch {
  unsigned count_reduce, count_residual, pad[2];
  RedPart parts[kRedMaxGroups];
};

__device__ __forceinline__ void absmax_merge(double& a, double b) { a = (a != a) ? a : ((b != b || b > a) ? b : a); }

__device__ __forceinline__ bool last_group(unsigned* counter) {
  __shared__ bool last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // this group's partial is visible before the count
    last = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (last) __threadfence();  // every other group's partial is visible to this one
  return last;
}

__global__ __launch_bounds__(kRedThreads) void reduce_kernel(FitView fv, int ne, const double* __restrict__ y,
                                                             const double* __restrict__ yold, int n,
                                                             RedScratch* __restrict__ rs,
                                                             IterScalars* __restrict__ out) {
  __shared__ double se[kRedThreads], sc[kRedThreads], sd[kRedThreads], so[kRedThreads];
  __shared__ int si[kRedThreads];
  if (stopped(fv)) return;  // the same for every workgroup of the launch
  const int tid = threadIdx.x, g = blockIdx.x, G = gridDim.x;
  double em = -__builtin_inf(), cm = -__builtin_inf(), dm = 0.0, om = 0.0;
  int ie = 0x7fffffff, dummy = 0;
  {
    const int per = (ne + G - 1) / G, e0 = g * per, e1 = min(ne, e0 + per);
    for (int e = e0 + tid; e < e1; e += kRedThreads) {
      nanmax_merge(em, ie, fv.error[e], e);
      nanmax_merge(cm, dummy, fv.error_coef[e], e);
    }
    const int pn = (n + G - 1) / G, k0 = g * pn, k1 = min(n, k0 + pn);
    for (int k = k0 + tid; k < k1; k += kRedThreads) {
      absmax_merge(dm, fabs(y[k] - yold[k]));
      absmax_merge(om, fabs(yold[k]));
    }
  }
  auto tree = [&]() {
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
      if (tid < s) {
        double a = se[tid];
        int ia = si[tid];
        nanmax_merge(a, ia, se[tid + s], si[tid + s]);
        se[tid] = a;
        si[tid] = ia;
        int d0 = 0;
        double c = sc[tid];
        nanmax_merge(c, d0, sc[tid + s], 0);
        sc[tid] = c;
        double d = sd[tid];
        absmax_merge(d, sd[tid + s]);
        sd[tid] = d;
        double o = so[tid];
        absmax_merge(o, so[tid + s]);
        so[tid] = o;
      }
      __syncthreads();
    }
  };
  se[tid] = em;
  si[tid] = ie;
  sc[tid] = cm;
  sd[tid] = dm;
  so[tid] = om;
  tree();
  if (tid == 0) rs->parts[g] = RedPart{se[0], sc[0], sd[0], so[0], si[0], 0};
  if (!last_group(&rs->count_reduce)) return;
  {
    RedPart q{-__builtin_inf(), -__builtin_inf(), 0.0, 0.0, 0x7fffffff, 0};
    if (tid < G) q = rs->parts[tid];
    se[tid] = q.e;
    si[tid] = q.ie;
    sc[tid] = q.c;
    sd[tid] = q.d;
    so[tid] = q.o;
  }
  tree();
  if (tid == 0) {
    out->error = se[0];
    out->argmax = si[0];
    out->error_coef = sc[0];
    out->dq_max = sd[0];
    out->yold_max = so[0];
    rs->count_reduce = 0;
  }
}

__global__ void init_errors_kernel(FitView fv, int ne) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  fv.error_old[e] = 1e30;
  fv.error[e] = 1e30;
}

// --------------------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------------------
template <int P>
static void launch_sweep_p(int mode, const NetView& net, const TargetView& tg, int ntiles, int K, double* val,
                           hipStream_t st, bool mixed) {
  const dim3 g(ntiles), b(64);
  if (mode == MODE_COLLOC && mixed)
    hipLaunchKernelGGL((sweep_kernel<MODE_COLLOC, true, P>), g, b, 0, st, net, tg, K, val);
  else if (mode == MODE_COLLOC)
    hipLaunchKernelGGL((sweep_kernel<MODE_COLLOC, false, P>), g, b, 0, st, net, tg, K, val);
  else if (mode == MODE_RHS && mixed)
    hipLaunchKernelGGL((sweep_kernel<MODE_RHS, true, P>), g, b, 0, st, net, tg, K, val);
  else if (mode == MODE_RHS)
    hipLaunchKernelGGL((sweep_kernel<MODE_RHS, false, P>), g, b, 0, st, net, tg, K, val);
  else if (mode == MODE_COLLOC_HEADS)
    hipLaunchKernelGGL((sweep_kernel<MODE_COLLOC_HEADS, false, P>), g, b, 0, st, net, tg, K, val);
  else if (mode == MODE_DISCH_HEADS)
    hipLaunchKernelGGL((sweep_kernel<MODE_DISCH_HEADS, false, P>), g, b, 0, st, net, tg, K, val);
  else if (mode == MODE_POINTS_OMEGA)
    hipLaunchKernelGGL((sweep_kernel<MODE_POINTS_OMEGA, false, P>), g, b, 0, st, net, tg, K, val);
  else
    hipLaunchKernelGGL((sweep_kernel<MODE_POINTS, false, P>), g, b, 0, st, net, tg, K, val);
}

template <int P>
static void launch_sweep_persistent_p(int mode, const NetView& net, const TargetView& tg, int ntiles, int K,
                                      double* val, hipStream_t st, bool mixed, int nwg, unsigned* queue) {
  const dim3 g(nwg), b(64);
  const int per = ntiles / 8;
  if (mode == MODE_COLLOC && mixed)
    hipLaunchKernelGGL((sweep_persistent_kernel<MODE_COLLOC, true, P>), g, b, 0, st, net, tg, K, val, per, queue);
  else if (mode == MODE_COLLOC)
    hipLaunchKernelGGL((sweep_persistent_kernel<MODE_COLLOC, false, P>), g, b, 0, st, net, tg, K, val, per, queue);
  else if (mode == MODE_RHS && mixed)
    hipLaunchKernelGGL((sweep_persistent_kernel<MODE_RHS, true, P>), g, b, 0, st, net, tg, K, val, per, queue);
  else
    hipLaunchKernelGGL((sweep_persistent_kernel<MODE_RHS, false, P>), g, b, 0, st, net, tg, K, val, per, queue);
}

template <int MODE, bool MIX, int P>
static int resident_groups(int cus) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, sweep_persistent_kernel<MODE, MIX, P>, 64, 0) != hipSuccess)
    nb = 0;
  return nb * cus;
}

int sweep_persistent_groups(int mode, bool mixed, int points_per_lane, int cus) {
  if (points_per_lane == 2) {
    if (mode == MODE_COLLOC) return mixed ? resident_groups<MODE_COLLOC, true, 2>(cus) : resident_groups<MODE_COLLOC, false, 2>(cus);
    return mixed ? resident_groups<MODE_RHS, true, 2>(cus) : resident_groups<MODE_RHS, false, 2>(cus);
  }
  if (mode == MODE_COLLOC) return mixed ? resident_groups<MODE_COLLOC, true, 1>(cus) : resident_groups<MODE_COLLOC, false, 1>(cus);
  return mixed ? resident_groups<MODE_RHS, true, 1>(cus) : resident_groups<MODE_RHS, false, 1>(cus);
}

void launch_sweep_persistent(int mode, const NetView& net, const TargetView& tg, int ntiles, int K, double* val,
                             hipStream_t st, bool mixed, int points_per_lane, int nwg, unsigned* queue) {
  if (ntiles <= 0 || nwg <= 0) return;
  if (ntiles % 8 != 0) return;  // the XCD layout always pads to whole rows of eight
  if (points_per_lane == 2) launch_sweep_persistent_p<2>(mode, net, tg, ntiles, K, val, st, mixed, nwg, queue);
  else launch_sweep_persistent_p<1>(mode, net, tg, ntiles, K, val, st, mixed, nwg, queue);
}

void launch_sweep(int mode, const NetView& net, const TargetView& tg, int ntiles, int K, double* val, hipStream_t st,
                  bool mixed, int points_per_lane) {
  if (ntiles <= 0) return;
  if (points_per_lane == 2) launch_sweep_p<2>(mode, net, tg, ntiles, K, val, st, mixed);
  else launch_sweep_p<1>(mode, net, tg, ntiles, K, val, st, mixed);
}

// single-precision copy of every element's coefficients (mixed-precision far field)
__global__ void coef_to_f32_kernel(NetView net, int ne, float2* __restrict__ out) {
  const int e = blockIdx.x;
  if (e >= ne || stopped(net)) return;
  const int m = net.ncoef[e];
  const double2* c = net.coef + (size_t)e * net.M;
  float2* o = out + (size_t)e * net.M;
  for (int j = threadIdx.x; j < m; j += blockDim.x) o[j] = make_float2((float)c[j].x, (float)c[j].y);
}

void launch_coef_to_f32(const NetView& net, int ne, float2* out, hipStream_t st) {
  if (ne <= 0) return;
  hipLaunchKernelGGL(coef_to_f32_kernel, dim3(ne), dim3(64), 0, st, net, ne, out);
}

void launch_fit(const NetView& net, const FitView& fv, const FitItem* items, int nitems, const FitItem* citems,
                int ncitems, double* val, const int* pt_off, const DftTile* tiles, int ntiles, const int* dft_sides,
                double* dft, double damping, int blend, hipStream_t st, bool f32) {
  if (nitems <= 0) return;
  if (ncitems > 0) hipLaunchKernelGGL(unwrap_kernel, dim3(ncitems), dim3(64), 0, st, net, fv, citems, val, pt_off);
  if (ntiles > 0) {
    if (f32)
      hipLaunchKernelGGL(dft_kernel<true>, dim3((ntiles + 3) / 4), dim3(256), 0, st, tiles, ntiles, dft_sides, val,
                         pt_off, net.M, dft, fv.trig, fv.trig32);
    else
      hipLaunchKernelGGL(dft_kernel<false>, dim3((ntiles + 3) / 4), dim3(256), 0, st, tiles, ntiles, dft_sides, val,
                         pt_off, net.M, dft, fv.trig, fv.trig32);
  }
  // a group of 32 lanes per element when every count fits (two coefficient passes at most for
  // 64), one wavefront per element otherwise
  if (net.M <= 32)
    hipLaunchKernelGGL(fit_kernel<32>, dim3((nitems + 7) / 8), dim3(256), 0, st, net, fv, items, nitems, dft, damping,
                       blend);
  else
    hipLaunchKernelGGL(fit_kernel<64>, dim3((nitems + 3) / 4), dim3(256), 0, st, net, fv, items, nitems, dft, damping,
                       blend);
}

void launch_blend(const NetView& net, int ne, double damping, hipStream_t st) {
  if (ne <= 0) return;
  hipLaunchKernelGGL(blend_kernel, dim3(ne), dim3(64), 0, st, net, ne, damping);
}

void launch_wells(const NetView& net, const FitView& fv, const int* wells, int nw, double damping, hipStream_t st,
                  int steps) {
  if (nw <= 0) return;
  hipLaunchKernelGGL(well_kernel, dim3((nw + 255) / 256), dim3(256), 0, st, net, fv, wells, nw, damping, steps);
}

void launch_grow(const NetView& net, const FitView& fv, int ne, GrowParams gp, const int* sched, int* grow,
                 hipStream_t st) {
  hipLaunchKernelGGL(grow_kernel, dim3((ne + 255) / 256), dim3(256), 0, st, net, fv, ne, gp, sched, grow);
}

// x[i] = rs[perm[i]] * b[perm[i]]: row equilibration and the LU row permutation in one gather
__global__ void permute_kernel(const double* __restrict__ b, const int* __restrict__ perm,
                               const double* __restrict__ rs, int n, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int p = perm[i];
    x[i] = b[p] * rs[p];
  }
}

// row equilibration of a dense column-major matrix: rs[j] = 1 / max_k |A[j, k]|, A[j, :] *= rs[j].
// Pass 1: each thread takes one row over a chunk of columns and merges its maximum with an
// integer atomic max on the bit pattern (ordered like the value for non-negative doubles).
constexpr int kColChunk = 256;
__global__ void row_absmax_kernel(const double* __restrict__ A, int64_t npad, int nblk,
                                  unsigned long long* __restrict__ mx) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npad * nblk) return;
  const int64_t b = j / npad, r = j - b * npad;
  const double* a = A + b * npad * npad + r;
  const int64_t k0 = (int64_t)blockIdx.y * kColChunk;
  const int64_t k1 = k0 + kColChunk < npad ? k0 + kColChunk : npad;
  double m = 0.0;
  for (int64_t k = k0; k < k1; ++k) m = fmax(m, fabs(a[k * npad]));
  atomicMax(&mx[j], (unsigned long long)__double_as_longlong(m));
}

__global__ void row_apply_kernel(double* __restrict__ A, int64_t npad, int nblk,
                                 const unsigned long long* __restrict__ mx, double* __restrict__ rs) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npad * nblk) return;
  const double m = __longlong_as_double((long long)mx[j]);
  const double s = m > 0.0 ? 1.0 / m : 1.0;
  if (blockIdx.y == 0) rs[j] = s;
  const int64_t b = j / npad, r = j - b * npad;
  double* a = A + b * npad * npad + r;
  const int64_t k0 = (int64_t)blockIdx.y * kColChunk;
  const int64_t k1 = k0 + kColChunk < npad ? k0 + kColChunk : npad;
  for (int64_t k = k0; k < k1; ++k) a[k * npad] *= s;
}

void launch_row_scale(double* A, int64_t npad, int nblk, double* rs, unsigned long long* work, hipStream_t st) {
  const int64_t rows = npad * nblk;
  dim3 grid((unsigned)((rows + 255) / 256), (unsigned)((npad + kColChunk - 1) / kColChunk));
  (void)hipMemsetAsync(work, 0, sizeof(unsigned long long) * rows, st);
  hipLaunchKernelGGL(row_absmax_kernel, grid, dim3(256), 0, st, A, npad, nblk, work);
  hipLaunchKernelGGL(row_apply_kernel, grid, dim3(256), 0, st, A, npad, nblk, work, rs);
}

__global__ void scale_kernel(const double* __restrict__ b, const double* __restrict__ rs, int n, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = b[i] * rs[i];
}

void launch_scale(const double* b, const double* rs, int n, double* x, hipStream_t st) {
  hipLaunchKernelGGL(scale_kernel, dim3((n + 255) / 256), dim3(256), 0, st, b, rs, n, x);
}

// CSR copy of the assembled (unscaled, unfactored) dense matrix: val[k] = A[row, col[k]]
__global__ void csr_gather_kernel(const double* __restrict__ A, SysLayout sys, int64_t nrows,
                                  const int* __restrict__ row_ptr, const int* __restrict__ col,
                                  double* __restrict__ val) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) val[k] = A[mat_at(r, col[k], sys)];
}

void launch_csr_gather(const double* A, const SysLayout& sys, int64_t nrows, const int* row_ptr, const int* col,
                       double* val, hipStream_t st) {
  hipLaunchKernelGGL(csr_gather_kernel, dim3((unsigned)((nrows + 255) / 256)), dim3(256), 0, st, A, sys, nrows,
                     row_ptr, col, val);
}

// y = A x for a CSR matrix (one thread per row)
// y = A x for CSR A: 16 lanes per row stride over its nonzeros (coalesced reads of val and col;
// the rows here hold ~20-60 nonzeros), then a fixed-order reduction within the 16 lanes
constexpr int kSpmvLanes = 16;
__global__ __launch_bounds__(256) void spmv_kernel(int n, const int* __restrict__ row_ptr,
                                                   const int* __restrict__ col, const double* __restrict__ val,
                                                   const double* __restrict__ x, double* __restrict__ y) {
  const int g = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kSpmvLanes);
  const int l = threadIdx.x & (kSpmvLanes - 1);
  double s = 0.0;
  if (g < n) {
    const int k1 = row_ptr[g + 1];
    for (int k = row_ptr[g] + l; k < k1; k += kSpmvLanes) s = fma(val[k], x[col[k]], s);
  }
#pragma unroll
  for (int o = kSpmvLanes / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kSpmvLanes);
  if (g < n && l == 0) y[g] = s;
}

void launch_spmv(int n, const int* row_ptr, const int* col, const double* val, const double* x, double* y,
                 hipStream_t st) {
  const int64_t threads = (int64_t)n * kSpmvLanes;
  hipLaunchKernelGGL(spmv_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, n, row_ptr, col, val, x,
                     y);
}

// y[0:m] = beta * y + alpha * A[0:m, 0:ncols] x for a column-major block (lda) with ncols <= 1024:
// y = beta y + alpha A x for one block of gemv_rows<T> rows of the column-major A (ncols <= 1024
// columns, x staged in LDS).  A 128-byte line holds gemv_rows<T> = 16 doubles (32 floats): lane
// l of a wavefront takes row l % gemv_rows and column group l / gemv_rows, so one load
// instruction reads whole lines (4 column segments in FP64, 2 in FP32).  The columns are split
// over 16 wavefronts, a lane keeps sixteen loads in flight, and the partial sums are combined in
// a fixed order (lane exchanges across the column groups, then the wavefronts through LDS).
// Short panels (the tail of the forward pass, the start of the backward pass) have few rows;
// line-high workgroups give them four (FP64) or two (FP32) times the workgroups, and so the
// streaming rate, of 64-row ones.  ``tri`` skips the zero half of a triangular block (1: lower,
// column c <= row; 2: upper, column c >= row).  Used for the blocked triangular solves.
constexpr int kGemvCols = 1024;
constexpr int kGemvWaves = 16;
template <class T>
constexpr int gemv_rows() {
  return 128 / (int)sizeof(T);
}
template <class T>
__device__ __forceinline__ void gemv_cm_block(const T* __restrict__ A, int64_t lda, int m, int ncols,
                                              const T* __restrict__ x, T* __restrict__ y, T alpha, T beta, int tri,
                                              int rblk) {
  constexpr int RG = gemv_rows<T>(), CG = 64 / RG;
  __shared__ T sx[kGemvCols];
  __shared__ T part[kGemvWaves][RG];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int c = threadIdx.x; c < ncols; c += 64 * kGemvWaves) sx[c] = x[c];
  __syncthreads();
  const int r = lane % RG, cg = lane / RG;
  const int row = rblk * RG + r;
  const int per = (ncols + kGemvWaves - 1) / kGemvWaves;
  int c0 = wv * per, c1 = min(ncols, c0 + per);
  if (tri == 1) c1 = min(c1, rblk * RG + RG);  // lower: columns beyond the block's last row are zero
  if (tri == 2) c0 = max(c0, rblk * RG);       // upper: columns before the block's first row are zero
  T s = 0;
  if (row < m && c0 + cg < c1) {
    int c = c0 + cg;
    const T* a = A + (int64_t)c * lda + row;
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (; c + 15 * CG < c1; c += 16 * CG, a += 16 * CG * lda) {
      T v[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) v[k] = a[CG * k * lda];
#pragma unroll
      for (int k = 0; k < 16; k += 4) {
        s0 = fma(v[k], sx[c + CG * k], s0);
        s1 = fma(v[k + 1], sx[c + CG * (k + 1)], s1);
        s2 = fma(v[k + 2], sx[c + CG * (k + 2)], s2);
        s3 = fma(v[k + 3], sx[c + CG * (k + 3)], s3);
      }
    }
    for (; c < c1; c += CG, a += CG * lda) s0 = fma(*a, sx[c], s0);
    s = (s0 + s1) + (s2 + s3);
  }
#pragma unroll
  for (int o = RG; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);
  if (lane < RG) part[wv][lane] = s;
  __syncthreads();
  if (wv == 0 && lane < RG && row < m) {
    T t = 0;
#pragma unroll
    for (int k = 0; k < kGemvWaves; ++k) t += part[k][lane];
    y[row] = (beta == T(0) ? T(0) : beta * y[row]) + alpha * t;
  }
}

template <class T>
__global__ __launch_bounds__(64 * kGemvWaves) void gemv_cm_kernel(const T* __restrict__ A, int64_t lda, int m,
                                                                  int ncols, const T* __restrict__ x,
                                                                  T* __restrict__ y, T alpha, T beta, int64_t sA,
                                                                  int64_t sx) {
  const int64_t z = blockIdx.z;
  gemv_cm_block<T>(A + z * sA, lda, m, ncols, x + z * sx, y + z * sx, alpha, beta, 0, blockIdx.x);
}

// y_j = B_j x_j for the nb stacked nbk x nbk triangular blocks B_j (block j covers rows
// j*nbk .. min(n, (j+1)*nbk)); blockIdx.y selects the block
template <class T>
__global__ __launch_bounds__(64 * kGemvWaves) void gemv_diag_kernel(const T* __restrict__ B, int nbk, int n, int tri,
                                                                    const T* __restrict__ x, T* __restrict__ y,
                                                                    int64_t sB) {
  const int j = blockIdx.y;
  const int r0 = j * nbk;
  const int bs = min(nbk, n - r0);
  if ((int)blockIdx.x * gemv_rows<T>() >= bs) return;  // uniform per workgroup, before any barrier
  const int64_t z = blockIdx.z;
  gemv_cm_block<T>(B + z * sB + (size_t)j * nbk * nbk, nbk, bs, bs, x + z * n + r0, y + z * n + r0, T(1), T(0), tri,
                   blockIdx.x);
}

template <class T>
void launch_gemv_cm(const T* A, int64_t lda, int m, int ncols, const T* x, T* y, T alpha, T beta, hipStream_t st,
                    int batch, int64_t sA, int64_t sx) {
  if (m <= 0 || ncols <= 0 || batch <= 0) return;
  constexpr int RG = gemv_rows<T>();
  hipLaunchKernelGGL(gemv_cm_kernel<T>, dim3((m + RG - 1) / RG, 1, batch), dim3(64 * kGemvWaves), 0, st, A, lda, m,
                     ncols, x, y, alpha, beta, sA, sx);
}
template void launch_gemv_cm<double>(const double*, int64_t, int, int, const double*, double*, double, double,
                                     hipStream_t, int, int64_t, int64_t);
template void launch_gemv_cm<float>(const float*, int64_t, int, int, const float*, float*, float, float, hipStream_t,
                                    int, int64_t, int64_t);

template <class T>
void launch_gemv_diag(const T* blocks, int nbk, int n, int tri, const T* x, T* y, hipStream_t st, int batch,
                      int64_t sB) {
  if (n <= 0 || batch <= 0) return;
  const int nb = (n + nbk - 1) / nbk;
  constexpr int RG = gemv_rows<T>();
  hipLaunchKernelGGL(gemv_diag_kernel<T>, dim3((nbk + RG - 1) / RG, nb, batch), dim3(64 * kGemvWaves), 0, st, blocks,
                     nbk, n, tri, x, y, sB);
}
template void launch_gemv_diag<double>(const double*, int, int, int, const double*, double*, hipStream_t, int,
                                       int64_t);
template void launch_gemv_diag<float>(const float*, int, int, int, const float*, float*, hipStream_t, int, int64_t);

// diagonal of nb stacked nbk x nbk blocks set to one (unit-triangular inverses)
template <class T>
__global__ void unit_diag_kernel(T* __restrict__ blocks, int nb, int nbk) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nb * nbk) {
    const int b = k / nbk, i = k % nbk;
    blocks[(size_t)b * nbk * nbk + (size_t)i * nbk + i] = T(1);
  }
}

template <class T>
void launch_unit_diag(T* blocks, int nb, int nbk, hipStream_t st) {
  hipLaunchKernelGGL(unit_diag_kernel<T>, dim3((nb * nbk + 255) / 256), dim3(256), 0, st, blocks, nb, nbk);
}
template void launch_unit_diag<double>(double*, int, int, hipStream_t);
template void launch_unit_diag<float>(float*, int, int, hipStream_t);

// zero the triangle opposite to `lower` (strictly upper when lower, strictly lower otherwise) of
// nb stacked nbk x nbk blocks
template <class T>
__global__ void tri_zero_kernel(T* __restrict__ blocks, int64_t nbk, int lower) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbk * nbk) return;
  const int64_t c = i / nbk, r = i - c * nbk;
  if (lower ? c > r : r > c) blocks[(int64_t)blockIdx.y * nbk * nbk + i] = T(0);
}

template <class T>
void launch_tri_zero(T* blocks, int nb, int nbk, int lower, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(tri_zero_kernel<T>, dim3((unsigned)(((int64_t)nbk * nbk + 255) / 256), nb), dim3(256), 0, st,
                     blocks, (int64_t)nbk, lower);
}
template void launch_tri_zero<double>(double*, int, int, int, hipStream_t);
template void launch_tri_zero<float>(float*, int, int, int, hipStream_t);

// B_z = W_z for nb column-major m x nc matrices: W packed (leading dimension m, stride m nc),
// B with leading dimension ldb and stride sB
template <class T>
__global__ void copy_panels_kernel(const T* __restrict__ w, int64_t m, int64_t nc, T* __restrict__ b, int64_t ldb,
                                   int64_t sB) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * nc) return;
  const int64_t z = blockIdx.y, c = i / m, r = i - c * m;
  b[z * sB + c * ldb + r] = w[z * m * nc + i];
}

template <class T>
void launch_copy_panels(const T* w, int64_t m, int64_t nc, int nb, T* b, int64_t ldb, int64_t sB, hipStream_t st) {
  if (m <= 0 || nc <= 0 || nb <= 0) return;
  hipLaunchKernelGGL(copy_panels_kernel<T>, dim3((unsigned)((m * nc + 255) / 256), nb), dim3(256), 0, st, w, m, nc, b,
                     ldb, sB);
}
template void launch_copy_panels<double>(const double*, int64_t, int64_t, int, double*, int64_t, int64_t,
                                         hipStream_t);
template void launch_copy_panels<float>(const float*, int64_t, int64_t, int, float*, int64_t, int64_t, hipStream_t);

// single-precision copy of a matrix
__global__ void to_f32_kernel(const double* __restrict__ a, int64_t n, float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

void launch_to_f32(const double* a, int64_t n, float* b, hipStream_t st) {
  hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, n, b);
}

// refinement step of the mixed-precision discharge solve:
// r = b - A y (FP64, A y given), x32[i] = rs[p] * r[p] with p = perm[i]
__global__ void refine_rhs_kernel(const double* __restrict__ b, const double* __restrict__ ay,
                                  const int* __restrict__ perm, const double* __restrict__ rs, int n,
                                  float* __restrict__ x32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int p = perm[i];
    x32[i] = (float)((b[p] - ay[p]) * rs[p]);
  }
}

// y_new = y + d (d from the single-precision triangular solves)
__global__ void refine_update_kernel(const double* __restrict__ y, const float* __restrict__ d, int n,
                                     double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = y[i] + (double)d[i];
}

void launch_refine_rhs(const double* b, const double* ay, const int* perm, const double* rs, int n, float* x32,
                       hipStream_t st) {
  hipLaunchKernelGGL(refine_rhs_kernel, dim3((n + 255) / 256), dim3(256), 0, st, b, ay, perm, rs, n, x32);
}

void launch_refine_update(const double* y, const float* d, int n, double* x, hipStream_t st) {
  hipLaunchKernelGGL(refine_update_kernel, dim3((n + 255) / 256), dim3(256), 0, st, y, d, n, x);
}

void launch_permute(const double* b, const int* perm, const double* rs, int n, double* x, hipStream_t st) {
  hipLaunchKernelGGL(permute_kernel, dim3((n + 255) / 256), dim3(256), 0, st, b, perm, rs, n, x);
}

void launch_cross(const NetView& net, const FitView& fv, const int* circles, int nc, hipStream_t st) {
  if (nc <= 0) return;
  hipLaunchKernelGGL(cross_kernel, dim3(nc), dim3(64), 0, st, net, fv, circles, nc);
}

void launch_rhs_reduce(const NetView& net, const RhsItem* items, int nde, const double* val, const int* pt_off, int K,
                       const int* de_pos, double* b, hipStream_t st) {
  if (nde <= 0) return;
  hipLaunchKernelGGL(rhs_reduce_kernel, dim3(nde), dim3(64), 0, st, net, items, nde, val, pt_off, K, de_pos, b);
}

void launch_assemble(const NetView& net, const TargetView& tg, int nsides, const SysLayout& sys, int K, int nf,
                     double* A, hipStream_t st) {
  const DenseTarget tgt{A, sys};
  if (nsides > 0) hipLaunchKernelGGL(assemble_kernel<DenseTarget>, dim3(nsides), dim3(64), 0, st, net, tg, sys, K, tgt);
  if (nf > 0)
    hipLaunchKernelGGL(continuity_kernel<DenseTarget>, dim3((nf + 255) / 256), dim3(256), 0, st, net, sys, nf, tgt);
}

void launch_assemble_csr(const NetView& net, const TargetView& tg, int nsides, const SysLayout& sys, int K, int nf,
                         const int* ptr, const int* col, double* val, hipStream_t st) {
  const CsrTarget tgt{val, ptr, col};
  if (nsides > 0) hipLaunchKernelGGL(assemble_kernel<CsrTarget>, dim3(nsides), dim3(64), 0, st, net, tg, sys, K, tgt);
  if (nf > 0)
    hipLaunchKernelGGL(continuity_kernel<CsrTarget>, dim3((nf + 255) / 256), dim3(256), 0, st, net, sys, nf, tgt);
}

void launch_csr_row_scale(int64_t n, const int* ptr, const double* val, double* rs, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(csr_row_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, ptr, val, rs);
}

void launch_pad_identity(double* A, int64_t npad, int nblk, const int* blk_n, hipStream_t st) {
  const int64_t tot = npad * nblk;
  if (tot <= 0) return;
  hipLaunchKernelGGL(pad_identity_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, A, npad, nblk, blk_n);
}

void launch_accept(const NetView& net, const int* de, int nde, int nf, const SysLayout& sys, const double* x, double* y,
                   double* yold, int64_t n, hipStream_t st) {
  const int64_t tot = std::max<int64_t>(n, (int64_t)nde + nf);
  if (tot <= 0) return;
  hipLaunchKernelGGL(accept_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, net, de, nde, nf, sys, x, y,
                     yold, n);
}

void launch_scatter(const NetView& net, const int* de, int nde, int nf, const SysLayout& sys, const double* y,
                    hipStream_t st) {
  int tot = nde + nf;
  if (tot <= 0) return;
  hipLaunchKernelGGL(scatter_kernel, dim3((tot + 255) / 256), dim3(256), 0, st, net, de, nde, nf, sys, y);
}

// run-ahead iteration: the residual of iteration it is recorded in pinned host memory; with
// may_stop (the FP64 phase) a residual <= tol stops the run, and every later state kernel of the
// stream returns at once
__global__ void iter_decide_kernel(const IterScalars* __restrict__ sc, IterCtl* __restrict__ ctl, int may_stop,
                                   double tol, IterHostLog log) {
  if (threadIdx.x != 0 || ctl->stop) return;
  const int it = ctl->it + 1;
  ctl->it = it;
  const double r = sc->residual;
  if (it < log.cap) {
    log.resid[it] = r;
    log.arg[it] = sc->arg_residual;
  }
  if (may_stop && r <= tol) {
    ctl->stop = 1;
    *log.stop_it = it;
  }
  if (log.stamps && it < log.cap) log.stamps[6 * (size_t)it + 1] = __builtin_amdgcn_s_memrealtime();
  __threadfence_system();
}

__global__ void iter_stamp_kernel(const IterCtl* __restrict__ ctl, unsigned long long* __restrict__ stamps, int cap,
                                  int slot, int next) {
  if (threadIdx.x != 0 || ctl->stop) return;
  const int it = ctl->it + next;
  if (it < cap) stamps[6 * (size_t)it + slot] = __builtin_amdgcn_s_memrealtime();
}

void launch_iter_stamp(const IterCtl* ctl, unsigned long long* stamps, int cap, int slot, int next, hipStream_t st) {
  hipLaunchKernelGGL(iter_stamp_kernel, dim3(1), dim3(64), 0, st, ctl, stamps, cap, slot, next);
}

__global__ void iter_reset_kernel(IterCtl* __restrict__ ctl, int it) {
  if (threadIdx.x == 0) {
    ctl->stop = 0;
    ctl->it = it;
  }
}

void launch_iter_decide(const IterScalars* sc, IterCtl* ctl, int may_stop, double tol, IterHostLog log,
                        hipStream_t st) {
  hipLaunchKernelGGL(iter_decide_kernel, dim3(1), dim3(64), 0, st, sc, ctl, may_stop, tol, log);
}

void launch_iter_reset(IterCtl* ctl, int it, hipStream_t st) {
  hipLaunchKernelGGL(iter_reset_kernel, dim3(1), dim3(64), 0, st, ctl, it);
}

void launch_reduce(const FitView& fv, int ne, const double* y, const double* yold, int n, void* scratch,
                   IterScalars* out, hipStream_t st) {
  const int groups = std::max(1, std::min(kRedMaxGroups, (std::max(ne, n) + 1023) / 1024));
  hipLaunchKernelGGL(reduce_kernel, dim3(groups), dim3(kRedThreads), 0, st, fv, ne, y, yold, n,
                     (RedScratch*)scratch, out);
}

void launch_init_errors(const FitView& fv, int ne, hipStream_t st) {
  if (ne <= 0) return;
  hipLaunchKernelGGL(init_errors_kernel, dim3((ne + 255) / 256), dim3(256), 0, st, fv, ne);
}

// max over the head equations of |(A y)_j - b_j| in metres (constant-head/well rows are in
// potential units and are divided by t of their fracture via row_scale, which is 0 on the
// continuity and padding rows)
__global__ __launch_bounds__(kRedThreads) void residual_kernel(const double* __restrict__ ay,
                                                               const double* __restrict__ b,
                                                               const double* __restrict__ row_scale, int nrows,
                                                               RedScratch* __restrict__ rs,
                                                               IterScalars* __restrict__ out) {
  __shared__ double sr[kRedThreads];
  __shared__ int si[kRedThreads];
  const int tid = threadIdx.x, g = blockIdx.x, G = gridDim.x;
  double r = 0.0;
  int ir = 0x7fffffff;
  const int per = (nrows + G - 1) / G, j0 = g * per, j1 = min(nrows, j0 + per);
  for (int j = j0 + tid; j < j1; j += kRedThreads) {
    const double v = fabs(ay[j] - b[j]) * row_scale[j];  // row_scale 0 outside the head equations
    if (v != v || v > r) {
      r = v;
      ir = j;
    }
  }
  auto tree = [&]() {
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
      if (tid < s) {
        double a = sr[tid];
        int ia = si[tid];
        nanmax_merge(a, ia, sr[tid + s], si[tid + s]);
        sr[tid] = a;
        si[tid] = ia;
      }
      __syncthreads();
    }
  };
  sr[tid] = r;
  si[tid] = ir;
  tree();
  if (tid == 0) rs->parts[g] = RedPart{sr[0], 0.0, 0.0, 0.0, si[0], 0};
  if (!last_group(&rs->count_residual)) return;
  sr[tid] = tid < G ? rs->parts[tid].e : 0.0;
  si[tid] = tid < G ? rs->parts[tid].ie : 0x7fffffff;
  tree();
  if (tid == 0) {
    out->residual = sr[0];
    out->arg_residual = si[0];
    rs->count_residual = 0;
  }
}

void launch_residual(const double* ay, const double* b, const double* row_scale, int nde, void* scratch,
                     IterScalars* out, hipStream_t st) {
  const int groups = std::max(1, std::min(kRedMaxGroups, (nde + 1023) / 1024));
  hipLaunchKernelGGL(residual_kernel, dim3(groups), dim3(kRedThreads), 0, st, ay, b, row_scale, nde,
                     (RedScratch*)scratch, out);
}

size_t reduce_scratch_bytes() { return sizeof(RedScratch); }

// pointwise boundary-condition residual in metres: head continuity at every collocation point of
// every intersection, prescribed head at every collocation point of every constant-head line and
// at the discharge points of every well (andfn/intersection.py:197-225, const_head.py:187-206)
__global__ __launch_bounds__(64) void bc_residual_kernel(NetView net, const FitItem* __restrict__ items, int nitems,
                                                         const double* __restrict__ val, const int* __restrict__ pt_off,
                                                         const RhsItem* __restrict__ witems, int nw,
                                                         const double* __restrict__ wval,
                                                         const int* __restrict__ wpt_off, int K,
                                                         double* __restrict__ out_max) {
  const int k = blockIdx.x;
  const int lane = threadIdx.x;
  double r = 0.0;
  if (k < nitems) {
    const FitItem it = items[k];
    const int e = it.elem;
    const int ty = net.etype[e];
    if (ty == 0 || ty == 3) {
      const int n = net.nint[e];
      const int b0 = pt_off[it.s0];
      for (int i = lane; i < n; i += 64) {
        double v = (ty == 0) ? fabs(val[b0 + i] - val[pt_off[it.s1] + i]) : fabs(val[b0 + i] - net.phi[e] / net.t[net.efrac0[e]]);
        r = (v != v || v > r) ? v : r;
      }
    }
  } else if (k - nitems < nw) {
    const RhsItem it = witems[k - nitems];
    const int e = it.elem;
    const int b0 = wpt_off[it.s0];
    for (int i = lane; i < K; i += 64) {
      double v = fabs(wval[b0 + i] - net.phi[e] / net.t[net.efrac0[e]]);
      r = (v != v || v > r) ? v : r;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    double v = __shfl_xor(r, o, 64);
    r = (v != v || v > r) ? v : r;
  }
  if (lane == 0) out_max[k] = r;
}

void launch_bc_residual(const NetView& net, const FitItem* items, int nitems, const double* val, const int* pt_off,
                        const RhsItem* witems, int nw, const double* wval, const int* wpt_off, int K, double* out_max,
                        hipStream_t st) {
  if (nitems + nw <= 0) return;
  hipLaunchKernelGGL(bc_residual_kernel, dim3(nitems + nw), dim3(64), 0, st, net, items, nitems, val, pt_off, witems,
                     nw, wval, wpt_off, K, out_max);
}

}  // namespace aem
