// Right-looking blocked LU of one dense column-major matrix with tournament pivoting, for the
// discharge system (the reference factors the same matrix with SuperLU's partial pivoting,
// hpc_solve.py:121; any stable row pivoting gives the same solution to rounding).
//
// Per 64-column panel [j0, j0 + bw):
//   tslu_kernel   pivot selection by a reduction tree: each group of up to 512 rows (one row per
//                 thread, the panel row in registers) runs Gaussian elimination with partial
//                 pivoting and passes its pivot rows up; 8 candidate sets form the next group.
//                 The last group's elimination is the LU of the winning rows in pivot order, so
//                 it writes the 64 x 64 block L11\U11 directly, and plans the row interchange
//                 (winner k to row j0 + k; displaced top rows to the winners' old rows).
//   apply_kernel  one wavefront per column of the outer panel (below): the interchange; the
//                 panel's top block L11\U11; U12 = L11^-1 A12 in the outer panel's columns.
//   l21_kernel    one thread per row below the panel's top block: L21 = A21 U11^-1.
//   then the update of the outer panel's remaining columns (rocBLAS dgemm).
// Inner panels are grouped into outer panels of kLuOuter columns: after its inner panels,
//   swap_kernel   applies their interchanges to every column outside the outer panel, and the
//                 outer U12 (rocBLAS trsm) and the rank-kLuOuter trailing update (rocBLAS dgemm,
//                 MFMA FP64) follow, so the trailing matrix is read and written once per
//                 kLuOuter columns instead of once per panel.  The host (Solver::calu_factor)
//                 updates the next outer panel's columns first and runs the rest of the
//                 trailing update on a second stream, overlapping the next panel's factorisation.
// The row permutation is kept as perm[i] = original row now at row i.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aem_kernels.hpp"
#include "lu_pivot_key.hpp"

namespace aem {

namespace {

constexpr int kBW = kLuPanel;       // panel width
constexpr int kRows = kLuGroupRows; // rows per tournament group (one per thread)
static_assert(kRows <= (1 << (kPivotIdxBits - 1)), "the pivot key holds the thread index of at most 512 rows");
constexpr int kWaves = kRows / 64;

// y = max(x, x of the lane that DPP control CTRL selects), on 64-bit keys
template <int CTRL>
__device__ __forceinline__ unsigned long long key_max_dpp(unsigned long long x) {
  const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, 0xF, 0xF, false);
  const unsigned long long y = ((unsigned long long)hi << 32) | lo;
  return y > x ? y : x;
}

__device__ __forceinline__ unsigned long long key_readlane(unsigned long long x, int lane) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}

// maximum of a 64-bit key over the wavefront, wave-uniform: quad exchanges (lane ^ 1, lane ^ 2),
// mirrors within 8 and 16 lanes, then the four 16-lane rows combined through readlane
__device__ __forceinline__ unsigned long long wave_key_max(unsigned long long x) {
  x = key_max_dpp<0xB1>(x);   // quad_perm [1, 0, 3, 2]
  x = key_max_dpp<0x4E>(x);   // quad_perm [2, 3, 0, 1]
  x = key_max_dpp<0x141>(x);  // row_half_mirror
  x = key_max_dpp<0x140>(x);  // row_mirror
  unsigned long long m = key_readlane(x, 0);
  const unsigned long long m1 = key_readlane(x, 16), m2 = key_readlane(x, 32), m3 = key_readlane(x, 48);
  m = m1 > m ? m1 : m;
  m = m2 > m ? m2 : m;
  return m3 > m ? m3 : m;
}

__device__ __forceinline__ int lanes_below(unsigned long long mask, int lane) {
  return __popcll(mask & ((1ull << lane) - 1ull));
}

}  // namespace

// One level of the pivot tournament.  Leaf level (cand_in == nullptr): group g holds rows
// j0 + 512 g + t.  Otherwise group g holds the candidates of input sets 8 g .. 8 g + 7
// (thread t: set 8 g + t / 64, entry t % 64).  Outputs the group's pivot rows in pivot order.
__global__ __launch_bounds__(kRows) void tslu_kernel(const double* __restrict__ A, int64_t lda, int64_t n, int64_t j0,
                                                      int bw, const int* __restrict__ cand_in,
                                                      const int* __restrict__ cnt_in, int nsets_in,
                                                      int* __restrict__ cand_out, int* __restrict__ cnt_out,
                                                      int final_level, double* __restrict__ lu11,
                                                      int* __restrict__ pairs, int* __restrict__ perm,
                                                      int64_t* __restrict__ info, int64_t sA, int64_t sc,
                                                      int64_t sp) {
  // block z of a batch: matrix at z sA, candidate sets at z sc (counts at z sc / kBW), pairs at
  // z sp, its own L11\U11, n permutation entries and info word
  {
    const int64_t z = blockIdx.z;
    A += z * sA;
    if (cand_in != nullptr) {
      cand_in += z * sc;
      cnt_in += z * (sc / kBW);
    }
    cand_out += z * sc;
    cnt_out += z * (sc / kBW);
    lu11 += z * kBW * kBW;
    pairs += z * sp;
    perm += z * n;
    info += z;
  }
  __shared__ double prow[kBW];
  __shared__ double pinv;
  __shared__ unsigned long long redk[kWaves];
  __shared__ int swin[kBW];
  __shared__ int sdisp[kBW];
  const int t = threadIdx.x, g = blockIdx.x;
  int64_t row = -1;
  if (cand_in == nullptr) {
    const int64_t r = j0 + (int64_t)g * kRows + t;
    if (r < n) row = r;
  } else {
    const int s = g * kWaves + (t >> 6), i = t & 63;
    if (s < nsets_in && i < cnt_in[s]) row = cand_in[(int64_t)s * kBW + i];
  }
  bool active = row >= 0;
  double a[kBW];
#pragma unroll
  for (int c = 0; c < kBW; ++c) a[c] = (active && c < bw) ? A[row + (j0 + c) * lda] : 0.0;
  int npiv = 0, mypiv = -1;
  // pivot key (lu_pivot_key.hpp): the FP64 bits of |a[k]| with the inverted thread index in
  // the low bits, so one 64-bit max picks the largest magnitude (ties to the lower thread);
  // 0 marks a row that is already a pivot.
  // The 64 steps run as a loop over 8 blocks of 8 columns: the block's columns are copied into
  // cur[] and its 8 steps are unrolled, so every register index is static (the row stays in
  // registers) while the code stays ~2k instructions; fully unrolled, the 64 steps take ~140 KB
  // of code, more than the instruction cache, and every step waits on instruction fetch.  Each
  // element receives the same multiply-adds in the same order either way.
  constexpr int kCB = 8;
#pragma clang loop unroll(disable)
  for (int kb = 0; kb < kBW / kCB; ++kb) {
    if (kb * kCB >= bw) break;
    double cur[kCB];
#pragma unroll
    for (int cb = 0; cb < kBW / kCB; ++cb)
      if (cb == kb) {
#pragma unroll
        for (int q = 0; q < kCB; ++q) cur[q] = a[cb * kCB + q];
      }
#pragma unroll
    for (int q = 0; q < kCB; ++q) {
      const int k = kb * kCB + q;
      if (k < bw) {
        unsigned long long key = 0ull;
        if (active) key = lu_pivot_key(cur[q], t);
        // every row's reciprocal is formed while the pivot search runs, so the winner only
        // stores its own (the division is off the path between the two barriers)
        const double rinv = cur[q] != 0.0 ? 1.0 / cur[q] : 0.0;
        key = wave_key_max(key);
        if ((t & 63) == 0) redk[t >> 6] = key;
        __syncthreads();
        unsigned long long best = redk[0];
#pragma unroll
        for (int w = 1; w < kWaves; ++w) {
          const unsigned long long wk = redk[w];
          best = wk > best ? wk : best;
        }
        if (best != 0ull) {
          npiv = k + 1;
          if (t == lu_pivot_thread(best)) {
#pragma unroll
            for (int q2 = 0; q2 < kCB; ++q2) prow[kb * kCB + q2] = cur[q2];
#pragma unroll
            for (int cb = 0; cb < kBW / kCB; ++cb)
              if (cb > kb) {
#pragma unroll
                for (int q2 = 0; q2 < kCB; ++q2) prow[cb * kCB + q2] = a[cb * kCB + q2];
              }
            pinv = rinv;
            swin[k] = (int)row;
            mypiv = k;
            active = false;
          }
        }
        __syncthreads();
        if (best != 0ull) {
          if (active) {
            const double l = cur[q] * pinv;
            cur[q] = l;
#pragma unroll
            for (int q2 = q + 1; q2 < kCB; ++q2) cur[q2] = fma(-l, prow[kb * kCB + q2], cur[q2]);
#pragma unroll
            for (int cb = 0; cb < kBW / kCB; ++cb)
              if (cb > kb) {
#pragma unroll
                for (int q2 = 0; q2 < kCB; ++q2) a[cb * kCB + q2] = fma(-l, prow[cb * kCB + q2], a[cb * kCB + q2]);
              }
          }
          if (final_level && t == 0 && prow[k] == 0.0 && info[0] == 0) info[0] = j0 + k + 1;
        }
      }
    }
#pragma unroll
    for (int cb = 0; cb < kBW / kCB; ++cb)
      if (cb == kb) {
#pragma unroll
        for (int q = 0; q < kCB; ++q) a[cb * kCB + q] = cur[q];
      }
  }
  __syncthreads();
  if (t < kBW && t < npiv) cand_out[(int64_t)g * kBW + t] = swin[t];
  if (t == 0) cnt_out[g] = npiv;
  if (!final_level) return;
  // the pivot rows in pivot order are the rows of L11\U11
  if (mypiv >= 0) {
#pragma unroll
    for (int c = 0; c < kBW; ++c) lu11[mypiv + c * kBW] = a[c];
  }
  // interchange plan: winner k -> row j0 + k; the top rows that are not winners (in row order)
  // -> the rows the winners outside the top leave (in pivot order)
  const int lane = t & 63;
  int w = -1, d0 = -1, s0 = -1, d1 = -1, s1 = -1;
  if (t < 64) {
    w = lane < npiv ? swin[lane] : -1;
    const bool top = w >= 0 && w < j0 + bw;
    unsigned long long bits = top ? (1ull << (w - j0)) : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bits |= __shfl_xor(bits, o, 64);
    const bool disp = lane < bw && !((bits >> lane) & 1ull);
    const unsigned long long mdisp = __ballot(disp);
    if (disp) sdisp[lanes_below(mdisp, lane)] = (int)(j0 + lane);
    const bool moved = w >= 0 && w != j0 + lane;
    const unsigned long long mmoved = __ballot(moved);
    const bool out = w >= 0 && !top;
    const unsigned long long mout = __ballot(out);
    const int n1 = __popcll(mmoved);
    if (moved) {
      const int i = lanes_below(mmoved, lane);
      d0 = (int)(j0 + lane);
      s0 = w;
      pairs[i] = d0;
      pairs[kLuMaxPairs + i] = s0;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);
    if (out) {
      const int i = lanes_below(mout, lane);
      d1 = w;
      s1 = sdisp[i];
      pairs[n1 + i] = d1;
      pairs[kLuMaxPairs + n1 + i] = s1;
    }
    if (lane == 0) {
      pairs[2 * kLuMaxPairs] = n1 + __popcll(mout);
      pairs[2 * kLuMaxPairs + 1] = npiv;
      if (npiv < bw && info[0] == 0) info[0] = j0 + npiv + 1;
    }
  }
  // perm[dst] = perm[src] for every pair: all reads before any write
  int p0 = 0, p1 = 0;
  if (d0 >= 0) p0 = perm[s0];
  if (d1 >= 0) p1 = perm[s1];
  __syncthreads();
  if (d0 >= 0) perm[d0] = p0;
  if (d1 >= 0) perm[d1] = p1;
}

// One wavefront per column c of the outer panel [c0, c1) (4 per workgroup): the inner panel's
// row interchange, then its top block L11\U11 (panel columns) or U12 = L11^-1 A12 (the outer
// panel's columns right of the inner panel).
__global__ __launch_bounds__(256) void calu_apply_kernel(double* __restrict__ A, int64_t lda, int64_t c0, int64_t c1,
                                                         int64_t j0, int bw, const int* __restrict__ pairs,
                                                         const int* __restrict__ win,
                                                         const double* __restrict__ lu11, int64_t sA, int64_t sc,
                                                         int64_t sp) {
  {
    const int64_t z = blockIdx.z;
    A += z * sA;
    pairs += z * sp;
    win += z * sc;
    lu11 += z * kBW * kBW;
  }
  const int lane = threadIdx.x & 63;
  const int64_t c = c0 + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const bool valid = c < c1;
  double* col = A + (valid ? c : 0) * lda;
  const int np = pairs[2 * kLuMaxPairs];
  const int npiv = pairs[2 * kLuMaxPairs + 1];
  const bool top_written = c >= j0;  // rows j0 .. j0 + bw of these columns are rewritten below
  int d0 = -1, d1 = -1;
  double v0 = 0.0, v1 = 0.0, x = 0.0;
  if (valid) {
    if (lane < np) {
      d0 = pairs[lane];
      v0 = col[pairs[kLuMaxPairs + lane]];
    }
    if (lane + 64 < np) {
      d1 = pairs[lane + 64];
      v1 = col[pairs[kLuMaxPairs + lane + 64]];
    }
    if (c >= j0 + bw && lane < npiv) x = col[win[lane]];
  }
  __syncthreads();  // every read of the column is done before it is rewritten
  if (!valid) return;
  if (d0 >= 0 && !(top_written && d0 < j0 + bw)) col[d0] = v0;
  if (d1 >= 0 && !(top_written && d1 < j0 + bw)) col[d1] = v1;
  if (c < j0) return;
  if (c < j0 + bw) {
    if (lane < bw) col[j0 + lane] = lu11[lane + (c - j0) * kBW];
    return;
  }
  for (int k = 0; k < bw - 1; ++k) {
    const double xk = __shfl(x, k, 64);
    if (lane > k) x = fma(-lu11[lane + k * kBW], xk, x);
  }
  if (lane < bw) col[j0 + lane] = x;
}

// The row interchanges of an outer panel's inner panels (lists 0 .. nlists - 1, in order) on every
// column outside that outer panel [skip0, skip1): one wavefront per column.
__global__ __launch_bounds__(256) void calu_swap_kernel(double* __restrict__ A, int64_t lda, int64_t n, int64_t skip0,
                                                        int64_t skip1, const int* __restrict__ pairs, int nlists,
                                                        int64_t sA, int64_t sp) {
  A += (int64_t)blockIdx.z * sA;
  pairs += (int64_t)blockIdx.z * sp;
  const int lane = threadIdx.x & 63;
  int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= skip0) c += skip1 - skip0;
  const bool valid = c < n;
  double* col = A + (valid ? c : 0) * lda;
  for (int l = 0; l < nlists; ++l) {
    const int* pl = pairs + l * kLuPairStride;
    const int np = pl[2 * kLuMaxPairs];
    int d0 = -1, d1 = -1;
    double v0 = 0.0, v1 = 0.0;
    if (valid) {
      if (lane < np) {
        d0 = pl[lane];
        v0 = col[pl[kLuMaxPairs + lane]];
      }
      if (lane + 64 < np) {
        d1 = pl[lane + 64];
        v1 = col[pl[kLuMaxPairs + lane + 64]];
      }
    }
    __syncthreads();  // reads of this list before its writes
    if (d0 >= 0) col[d0] = v0;
    if (d1 >= 0) col[d1] = v1;
    __syncthreads();  // writes of this list before the next list's reads
  }
}

// One thread per row r >= j0 + bw: L21 row = A21 row U11^-1 by forward substitution over the
// columns, with U11 (stored by rows) and the reciprocals of its diagonal in LDS.
__global__ __launch_bounds__(256) void calu_l21_kernel(double* __restrict__ A, int64_t lda, int64_t n, int64_t j0,
                                                       int bw, const double* __restrict__ lu11, int64_t sA) {
  A += (int64_t)blockIdx.z * sA;
  lu11 += (int64_t)blockIdx.z * kBW * kBW;
  __shared__ double ut[kBW * kBW];  // ut[c * kBW + c2] = U11(c, c2)
  __shared__ double rinv[kBW];
  for (int i = threadIdx.x; i < kBW * kBW; i += 256) {
    const int r = i & (kBW - 1), c = i / kBW;  // lu11 is column-major: lu11[r + c kBW] = (r, c)
    ut[r * kBW + c] = lu11[i];
  }
  __syncthreads();
  if (threadIdx.x < kBW) {
    const double d = ut[threadIdx.x * (kBW + 1)];
    rinv[threadIdx.x] = d != 0.0 ? 1.0 / d : 0.0;  // a zero pivot is reported through info
  }
  __syncthreads();
  const int64_t r = j0 + bw + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double a[kBW];
#pragma unroll
  for (int c = 0; c < kBW; ++c) a[c] = c < bw ? A[r + (j0 + c) * lda] : 0.0;
#pragma clang loop unroll(full)
  for (int c = 0; c < kBW; ++c) {
    if (c < bw) {
      a[c] *= rinv[c];
#pragma unroll
      for (int c2 = c + 1; c2 < kBW; ++c2) a[c2] = fma(-a[c], ut[c * kBW + c2], a[c2]);
    }
  }
#pragma unroll
  for (int c = 0; c < kBW; ++c)
    if (c < bw) A[r + (j0 + c) * lda] = a[c];
}

const int* launch_tslu(const double* A, int64_t lda, int64_t n, int64_t j0, int bw, int* cand[2], int* cnt[2],
                       double* lu11, int* pairs, int* perm, int64_t* info, hipStream_t st, int nblk, int64_t sA,
                       int64_t sp) {
  const int64_t m = n - j0;
  const int64_t sc = (n + kRows - 1) / kRows * kBW;  // candidate sets per block at the leaf level
  int groups = (int)((m + kRows - 1) / kRows);
  int cur = 0;
  hipLaunchKernelGGL(tslu_kernel, dim3(groups, 1, nblk), dim3(kRows), 0, st, A, lda, n, j0, bw, (const int*)nullptr,
                     (const int*)nullptr, 0, cand[cur], cnt[cur], groups == 1 ? 1 : 0, lu11, pairs, perm, info, sA, sc,
                     sp);
  while (groups > 1) {
    const int nsets = groups;
    groups = (nsets + kWaves - 1) / kWaves;
    hipLaunchKernelGGL(tslu_kernel, dim3(groups, 1, nblk), dim3(kRows), 0, st, A, lda, n, j0, bw,
                       (const int*)cand[cur], (const int*)cnt[cur], nsets, cand[cur ^ 1], cnt[cur ^ 1],
                       groups == 1 ? 1 : 0, lu11, pairs, perm, info, sA, sc, sp);
    cur ^= 1;
  }
  return cand[cur];
}

void launch_calu_apply(double* A, int64_t lda, int64_t n, int64_t c0, int64_t c1, int64_t j0, int bw,
                       const int* pairs, const int* win, const double* lu11, hipStream_t st, int nblk, int64_t sA,
                       int64_t sp) {
  if (c1 <= c0) return;
  const int64_t sc = (n + kRows - 1) / kRows * kBW;
  hipLaunchKernelGGL(calu_apply_kernel, dim3((unsigned)((c1 - c0 + 3) / 4), 1, nblk), dim3(256), 0, st, A, lda, c0,
                     c1, j0, bw, pairs, win, lu11, sA, sc, sp);
}

void launch_calu_swap(double* A, int64_t lda, int64_t n, int64_t skip0, int64_t skip1, const int* pairs, int nlists,
                      hipStream_t st, int nblk, int64_t sA, int64_t sp) {
  const int64_t cols = n - (skip1 - skip0);
  if (cols <= 0 || nlists <= 0) return;
  hipLaunchKernelGGL(calu_swap_kernel, dim3((unsigned)((cols + 3) / 4), 1, nblk), dim3(256), 0, st, A, lda, n, skip0,
                     skip1, pairs, nlists, sA, sp);
}

void launch_calu_l21(double* A, int64_t lda, int64_t n, int64_t j0, int bw, const double* lu11, hipStream_t st,
                     int nblk, int64_t sA) {
  const int64_t rows = n - j0 - bw;
  if (rows <= 0) return;
  hipLaunchKernelGGL(calu_l21_kernel, dim3((unsigned)((rows + 255) / 256), 1, nblk), dim3(256), 0, st, A, lda, n, j0,
                     bw, lu11, sA);
}

}  // namespace aem
