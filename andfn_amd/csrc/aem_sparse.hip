// Multifrontal sparse LU of the discharge system: ordering, symbolic analysis, numeric
// factorisation and the per-iteration solves (declarations and overview in aem_sparse.hpp).
#include "aem_sparse.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>

#include "aem_kernels.hpp"
#include "lu_pivot_key.hpp"

namespace aem {

namespace {


#define MF_HIP(x)                                                                                   \
  do {                                                                                              \
    hipError_t _e = (x);                                                                            \
    if (_e != hipSuccess) throw std::runtime_error(std::string(#x " failed: ") + hipGetErrorString(_e)); \
  } while (0)
#define MF_RB(x)                                                                                  \
  do {                                                                                            \
    rocblas_status _s = (x);                                                                      \
    if (_s != rocblas_status_success) throw std::runtime_error(std::string(#x " failed: ") + std::to_string((int)_s)); \
  } while (0)

template <class T>
void dev_upload(T*& d, const std::vector<T>& h, hipStream_t st) {
  MF_HIP(hipMalloc((void**)&d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) MF_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}
template <class T>
void dev_alloc(T*& d, size_t n) {
  MF_HIP(hipMalloc((void**)&d, std::max<size_t>(n, 1) * sizeof(T)));
}
template <class T>
void dev_free(T*& d) {
  if (d) (void)hipFree(d);
  d = nullptr;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(x, o, 64);
    x = y > x ? y : x;
  }
  return x;
}


// ---- factorisation kernels --------------------------------------------------------------------

// row-equilibrated A into the fronts: F[asm[q]] = req[row] * val[q] (one thread per row)
__global__ void mf_scatter_kernel(int nrows, const int* __restrict__ ptr, const double* __restrict__ val,
                                  const double* __restrict__ req, const int64_t* __restrict__ asmap,
                                  double* __restrict__ F) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const double s = req[r];
  for (int q = ptr[r]; q < ptr[r + 1]; ++q) F[asmap[q]] = s * val[q];
}

// extend-add: item (child front, first column): four columns of the child's Schur complement
// (a wavefront each) added into its parent at the parent-local positions rel
__global__ __launch_bounds__(256) void mf_extend_kernel(const MfFront* __restrict__ fr, const int2* __restrict__ items,
                                                        const int* __restrict__ rel, double* __restrict__ F) {
  const int2 it = items[blockIdx.x];
  const MfFront c = fr[it.x];
  const MfFront p = fr[c.parent];
  const int b = c.m - c.k;
  const int t = it.y + (threadIdx.x >> 6);
  if (t >= b) return;
  const int lane = threadIdx.x & 63;
  const int* rl = rel + c.rel;
  const double* src = F + c.off + (int64_t)(c.k + t) * c.m + c.k;
  double* dst = F + p.off + (int64_t)rl[t] * p.m;
  for (int r = lane; r < b; r += 64) dst[rl[r]] += src[r];
}

// In-LDS triangular solves of the small-front kernel on an X block (column-major, leading
// dimension LD, nc columns of k rows): X = L11^-1 X (unit lower, L11 below the diagonal of L) and
// X = U11^-1 X (U11 on and above it, rd its reciprocal diagonal); one barrier per elimination step
template <int LD>
__device__ void lds_lower_unit_solve(const double* L, double* X, int k, int nc) {
  for (int j = 0; j + 1 < k; ++j) {
    const int ni = k - j - 1;
    for (int e = threadIdx.x; e < ni * nc; e += blockDim.x) {
      const int i = j + 1 + e % ni, c = e / ni;
      X[c * LD + i] = fma(-L[j * LD + i], X[c * LD + j], X[c * LD + i]);
    }
    __syncthreads();
  }
}

template <int LD>
__device__ void lds_upper_solve(const double* L, const double* rd, double* X, int k, int nc) {
  for (int j = k - 1; j >= 0; --j) {
    for (int c = threadIdx.x; c < nc; c += blockDim.x) X[c * LD + j] *= rd[j];
    __syncthreads();
    for (int e = threadIdx.x; e < j * nc; e += blockDim.x) {
      const int i = e % j, c = e / j;
      X[c * LD + i] = fma(-L[j * LD + i], X[c * LD + j], X[c * LD + i]);
    }
    __syncthreads();
  }
}

// Partial factorisation of a front with k <= KM pivots, one workgroup: LU of the pivot block with
// partial pivoting over the front's own k rows, then in tiles of XC columns (rows) staged in LDS:
// U12 = L11^-1 P F12, L21 = F21 U11^-1, and L11^-1, U11^-1.  The Schur update F22 -= L21 U12
// follows in mf_schur_kernel.  (KM = 128 holds the pivot block in 129 KiB of LDS and stages
// 24-column tiles in the rest of the 160 KiB.)
template <int KM, int XC>
__global__ __launch_bounds__(256) void mf_small_factor_kernel(const MfFront* __restrict__ fr,
                                                              const int* __restrict__ list, double* __restrict__ F,
                                                              double* __restrict__ INV, int* __restrict__ piv,
                                                              int* __restrict__ status,
                                                              unsigned long long* __restrict__ growth) {
  static_assert(KM <= 128 && XC <= KM, "pivot search covers two rows per lane of one wavefront");
  constexpr int LD = KM + 1;   // leading dimension of L and of X's column tiles (k rows)
  constexpr int LDR = XC + 1;  // leading dimension of X's transposed row tiles (XC rows)
  constexpr int XN = XC * LD > KM * LDR ? XC * LD : KM * LDR;
  __shared__ double L[KM * LD];  // L[c * LD + r] = (r, c) of the pivot block
  __shared__ double X[XN];       // a tile of U12, L21 (transposed) or an inverse
  __shared__ double rdiag[KM];
  __shared__ unsigned long long gmax[KM];  // bits of max |L21(:, j)| (inf: zero pivot)
  __shared__ int prm[KM];
  __shared__ int spiv;
  const int fi = list[blockIdx.x];
  const MfFront f = fr[fi];
  const int m = f.m, k = f.k, b = m - k;
  double* A = F + f.off;
  const int t = threadIdx.x;
  for (int i = t; i < k * k; i += 256) {
    const int r = i % k, c = i / k;
    L[c * LD + r] = A[r + (int64_t)c * m];
  }
  if (t < k) {
    prm[t] = t;
    gmax[t] = 0ull;
  }
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (t < 64) {
      // rows t and t + 64 (the key's index field holds the row)
      unsigned long long key = 0ull;
      if (t >= j && t < k) key = lu_pivot_key(L[j * LD + t], t);
      if (KM > 64 && t + 64 >= j && t + 64 < k) {
        const unsigned long long k2 = lu_pivot_key(L[j * LD + t + 64], t + 64);
        key = k2 > key ? k2 : key;
      }
      key = wave_max_u64(key);
      if (t == 0) spiv = lu_pivot_thread(key);
    }
    __syncthreads();
    const int p = spiv;
    if (p != j) {
      for (int c = t; c < k; c += 256) {
        const double a = L[c * LD + j];
        L[c * LD + j] = L[c * LD + p];
        L[c * LD + p] = a;
      }
      if (t == 0) {
        const int a = prm[j];
        prm[j] = prm[p];
        prm[p] = a;
      }
    }
    __syncthreads();
    const double d = L[j * LD + j];
    if (d == 0.0 && t == 0) {
      atomicMin(status, fi);
      gmax[j] = 0x7ff0000000000000ull;  // +inf
    }
    const double rd = d != 0.0 ? 1.0 / d : 0.0;
    for (int r = j + 1 + t; r < k; r += 256) L[j * LD + r] *= rd;
    __syncthreads();
    const int n2 = k - j - 1;
    for (int i = t; i < n2 * n2; i += 256) {
      const int r = j + 1 + i % n2, c = j + 1 + i / n2;
      L[c * LD + r] = fma(-L[j * LD + r], L[c * LD + j], L[c * LD + r]);
    }
    __syncthreads();
  }
  if (t < k) rdiag[t] = L[t * LD + t] != 0.0 ? 1.0 / L[t * LD + t] : 0.0;
  for (int i = t; i < k * k; i += 256) {
    const int r = i % k, c = i / k;
    A[r + (int64_t)c * m] = L[c * LD + r];
  }
  if (t < k) piv[f.piv + t] = prm[t];
  __syncthreads();
  // U12 = L11^-1 P F12, XC columns at a time
  for (int c0 = 0; c0 < b; c0 += XC) {
    const int nc = min(XC, b - c0);
    for (int e = t; e < k * nc; e += 256) {
      const int r = e % k, c = e / k;
      X[c * LD + r] = A[prm[r] + (int64_t)(k + c0 + c) * m];
    }
    __syncthreads();
    lds_lower_unit_solve<LD>(L, X, k, nc);
    for (int e = t; e < k * nc; e += 256) {
      const int r = e % k, c = e / k;
      A[r + (int64_t)(k + c0 + c) * m] = X[c * LD + r];
    }
    __syncthreads();
  }
  // L21 = F21 U11^-1, XC rows at a time: X holds the rows' transposes, so U11^T X = F21^T is the
  // same upper solve from the left once transposed: solved here as forward substitution over columns
  for (int r0 = 0; r0 < b; r0 += XC) {
    const int nr = min(XC, b - r0);
    for (int e = t; e < k * nr; e += 256) {
      const int r = e % nr, c = e / nr;
      X[c * LDR + r] = A[(k + r0 + r) + (int64_t)c * m];
    }
    __syncthreads();
    for (int j = 0; j < k; ++j) {
      for (int r = t; r < nr; r += 256) X[j * LDR + r] *= rdiag[j];
      __syncthreads();
      const int nc2 = k - j - 1;
      for (int e = t; e < nr * nc2; e += 256) {
        const int r = e % nr, c = j + 1 + e / nr;
        X[c * LDR + r] = fma(-X[j * LDR + r], L[c * LD + j], X[c * LDR + r]);
      }
      __syncthreads();
    }
    for (int e = t; e < k * nr; e += 256) {
      const int r = e % nr, c = e / nr;
      A[(k + r0 + r) + (int64_t)c * m] = X[c * LDR + r];
    }
    __syncthreads();
  }
  // growth of every pivot: max |L21(:, j)| (pivots within the front's own rows keep |L11| <= 1)
  for (int e = t; e < k * b; e += 256) {
    const int r = e % b, j = e / b;
    atomicMax(&gmax[j], (unsigned long long)__double_as_longlong(fabs(A[(k + r) + (int64_t)j * m])));
  }
  __syncthreads();
  if (t < k) growth[f.piv + t] = gmax[t];
  // L11^-1 and U11^-1 against the identity, XC columns at a time
  double* Li = INV + f.inv;
  double* Ui = Li + (int64_t)k * k;
  for (int pass = 0; pass < 2; ++pass) {
    for (int c0 = 0; c0 < k; c0 += XC) {
      const int nc = min(XC, k - c0);
      for (int e = t; e < k * nc; e += 256) {
        const int r = e % k, c = e / k;
        X[c * LD + r] = r == c0 + c ? 1.0 : 0.0;
      }
      __syncthreads();
      if (pass == 0) lds_lower_unit_solve<LD>(L, X, k, nc);
      else lds_upper_solve<LD>(L, rdiag, X, k, nc);
      double* out = pass == 0 ? Li : Ui;
      for (int e = t; e < k * nc; e += 256) {
        const int r = e % k, c = e / k;
        out[r + (int64_t)(c0 + c) * k] = X[c * LD + r];
      }
      __syncthreads();
    }
  }
}

// Schur update of small fronts: tile (front, first boundary row, first boundary column) of 64 x 64
// entries of F22 -= L21 U12 (k <= kMfSmallMax), L21 and U12 panels staged in LDS 64 pivot columns
// at a time, 4 x 4 outputs per thread
__global__ __launch_bounds__(256) void mf_schur_kernel(const MfFront* __restrict__ fr, const int4* __restrict__ tiles,
                                                       double* __restrict__ F) {
  constexpr int KC = 64;
  __shared__ double Ls[KC * 65];
  __shared__ double Us[KC * 65];
  const int4 tl = tiles[blockIdx.x];
  const MfFront f = fr[tl.x];
  const int m = f.m, k = f.k, b = m - k, i0 = tl.y, j0 = tl.z;
  double* A = F + f.off;
  const int t = threadIdx.x;
  const int ti = t & 15, tj = t >> 4;
  double acc[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[q][r] = 0.0;
  for (int c0 = 0; c0 < k; c0 += KC) {
    const int kc = min(KC, k - c0);
    if (c0 > 0) __syncthreads();
    for (int e = t; e < kc * 64; e += 256) {
      const int i = e & 63, c = e >> 6;
      Ls[c * 65 + i] = (i0 + i < b) ? A[(k + i0 + i) + (int64_t)(c0 + c) * m] : 0.0;
    }
    for (int e = t; e < kc * 64; e += 256) {
      const int c = e % kc, j = e / kc;
      Us[c * 65 + j] = (j0 + j < b) ? A[(c0 + c) + (int64_t)(k + j0 + j) * m] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < kc; ++c) {
      double a[4], u[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = Ls[c * 65 + ti + 16 * q];
        u[q] = Us[c * 65 + tj + 16 * q];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[q][r] = fma(a[q], u[r], acc[q][r]);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int j = j0 + tj + 16 * r;
    if (j >= b) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + ti + 16 * q;
      if (i < b) A[(k + i) + (int64_t)(k + j) * m] -= acc[q][r];
    }
  }
}

// growth of a large front's pivots: a wavefront per pivot column j, max |L21(:, j)| over the
// boundary rows (inf for a zero pivot)
__global__ __launch_bounds__(256) void mf_growth_kernel(const double* __restrict__ A, int m, int k,
                                                        unsigned long long* __restrict__ growth) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= k) return;
  const int lane = threadIdx.x & 63;
  const double* col = A + (int64_t)j * m;
  double g = 0.0;
  for (int r = k + lane; r < m; r += 64) g = fmax(g, fabs(col[r]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) g = fmax(g, __shfl_xor(g, o, 64));
  if (lane == 0) growth[j] = col[j] == 0.0 ? 0x7ff0000000000000ull : (unsigned long long)__double_as_longlong(g);
}

__global__ void mf_axpy_kernel(int64_t n, const double* __restrict__ x0, const double* __restrict__ dx,
                               double* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x0[i] + dx[i];
}

// identity in the L11^-1 and U11^-1 blocks of front f (before the triangular solves against it)
__global__ void mf_eye_kernel(double* __restrict__ inv, int k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t kk = (int64_t)k * k;
  if (i >= 2 * kk) return;
  const int64_t e = i % kk;
  inv[i] = (e % k == e / k) ? 1.0 : 0.0;
}

// ---- solve kernels ----------------------------------------------------------------------------
// chunk (front, first row): 64 rows of one front per workgroup.  N-form products (lanes over the
// rows of a column-major block, the four wavefronts over interleaved columns, a fixed-order sum)
// for A x = b; T-form (a wavefront per output, lanes over the column, a shuffle-tree sum) for
// A^T y = c.  Children's update vectors reach a front through its gather lists.

__device__ __forceinline__ double gathered(const int* __restrict__ gptr, const int* __restrict__ gsrc,
                                           const double* __restrict__ vec, int at) {
  double s = 0.0;
  for (int q = gptr[at]; q < gptr[at + 1]; ++q) s += vec[gsrc[q]];
  return s;
}

// The block of a solve item: kind 0 L11^-1 (lower), 1 L21, 2 U12, 3 U11^-1 (upper).
// The block of a solve panel and its element (r, c) = M[r * rs + c * cs].  Kinds 0-3 are the
// products of A x = b: 0 L11^-1 (lower), 1 L21, 2 U12, 3 U11^-1 (upper).  Kinds 4-7 those of
// A^T y = c, with the shapes of kinds 0-3: 4 U11^-T (lower), 5 U12^T, 6 L21^T, 7 L11^-T (upper).
template <int KIND>
__device__ __forceinline__ void seg_block(const MfFront& f, const double* __restrict__ F,
                                          const double* __restrict__ INV, const double*& M, int64_t& rs,
                                          int64_t& cs, int& rows, int& cols) {
  constexpr int S = KIND & 3;                      // shape
  constexpr int B = KIND < 4 ? KIND : 7 - KIND;    // stored block
  const int k = f.k, nb = f.m - f.k;
  rows = S == 1 ? nb : k;
  cols = S == 2 ? nb : k;
  int64_t ld;
  if (B == 0) { M = INV + f.inv; ld = k; }
  else if (B == 1) { M = F + f.off + k; ld = f.m; }
  else if (B == 2) { M = F + f.off + (int64_t)k * f.m; ld = f.m; }
  else { M = INV + f.inv + (int64_t)k * k; ld = k; }
  rs = KIND < 4 ? 1 : ld;
  cs = KIND < 4 ? ld : 1;
}

// After a factorisation: copy each panel's rows of its block into the panel's tile (layout in
// mf_symbolic.hpp); thread t = R p + l of the workgroup holds row r0 + l, columns 2 p and 2 p + 1
// of every column block.
template <int KIND, int R>
__global__ __launch_bounds__(256) void mfs_pack_kernel(const MfFront* __restrict__ fr, const int4* __restrict__ items,
                                                       const int64_t* __restrict__ toff, const double* __restrict__ F,
                                                       const double* __restrict__ INV, double* __restrict__ T) {
  constexpr int C = kMfBlockValues / R;
  static_assert(kMfBlockValues == 512, "a column block is 256 threads x 2 values");
  const int4 it = items[blockIdx.x];
  const MfFront f = fr[it.x];
  const double* M;
  int64_t rs, cs;
  int rows, cols;
  seg_block<KIND>(f, F, INV, M, rs, cs, rows, cols);
  const int l = threadIdx.x % R, p = threadIdx.x / R, r = it.y + l;
  double2* out = reinterpret_cast<double2*>(T + toff[blockIdx.x]) + threadIdx.x;
  for (int cb = 0; cb < it.w; ++cb) {
    double v[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const int c = it.z + cb * C + 2 * p + e;
      bool ok = r < rows && c < cols;
      if ((KIND & 3) == 0) ok = ok && c <= r;
      if ((KIND & 3) == 3) ok = ok && c >= r;
      v[e] = ok ? M[r * rs + c * cs] : 0.0;
    }
    out[cb * 256] = make_double2(v[0], v[1]);
  }
}

// Solve products, one row panel per workgroup, every row computed in full, so no partial leaves
// the workgroup.  For A x = b (the row scales R and the pivot order P inside A = R^-1 P^T L U):
//   kind 0  vec_U  = L11^-1 P (R b + children)_U
//   kind 1  vec_B  = (children)_B - L21 vec_U
//   kind 2  vec_U -= U12 x_B
//   kind 3  x_U    = U11^-1 vec_U   (scattered to the solver's layout)
// and for A^T y' = c (y = R y' after the last level; b is c and x is y' here):
//   kind 4  vec_U  = U11^-T (c + children)_U
//   kind 5  vec_B  = (children)_B - U12^T vec_U
//   kind 6  vec_U -= L21^T y'_B
//   kind 7  y'_{P U} = L11^-T vec_U
// The panel's record carries its front's fields, so one scalar load starts it; the first batch of
// tile loads and the values the epilogue needs are issued before the input is staged, which keeps
// the dependent memory round trips of a small panel few.  The kind's input over the panel's
// columns is staged in LDS kMfStage columns at a time; thread t = 256 q + R p + l accumulates row
// r0 + l over columns 2 p, 2 p + 1 of the column blocks q, q + NQ, ... (eight 16-byte loads in
// flight, the workgroup streaming the panel's tile front to back), and the (256 / R) NQ partials
// of a row are summed in a fixed order (quarters, then a pairwise tree over the column pairs).  NQ = 4 (1,024 threads)
// for launches of few, long panels, so that the loads in flight still cover the memory latency of
// the whole device.
constexpr int kMfStage = 2048;
template <class TT>
struct TilePair;  // two tile values of one thread and column block
template <>
struct TilePair<double> {
  using V = double2;
};
template <>
struct TilePair<float> {
  using V = float2;
};

template <int KIND, int NQ, int R, class TT>
__global__ __launch_bounds__(256 * NQ) __attribute__((amdgpu_waves_per_eu(NQ > 1 ? 8 : 6))) void mfs_panel_kernel(
    const MfPanel* __restrict__ panels, const TT* __restrict__ T, const int* __restrict__ idx,
    const int* __restrict__ piv, const int* __restrict__ gptr, const int* __restrict__ gsrc,
    const double* __restrict__ b, const double* __restrict__ req, double* __restrict__ x, double* __restrict__ vec) {
  using V2 = typename TilePair<TT>::V;
  constexpr int C = kMfBlockValues / R, NP = 256 / R;  // columns per block, column pairs per block
  __shared__ double xs[kMfStage];
  __shared__ double ps[NP * NQ][R + 1];
  constexpr int S = KIND & 3;  // shape (kinds 4-7: the transposed solve)
  const MfPanel pn = panels[blockIdx.x];
  const int k = pn.k, nb = pn.m - pn.k;
  const int rows = S == 1 ? nb : k, cols = S == 2 ? nb : k;
  const int ncol = pn.nblk * C;
  const int t = threadIdx.x, l = t % R, p = (t & 255) / R, q = t >> 8, r = pn.r0 + t;
  // the panel's tile from a uniform base with 32-bit offsets (a tile is far below 4 GiB); FP32
  // tiles (the mixed form's correction solves) hold the same values rounded, in the same layout
  const char* P = reinterpret_cast<const char*>(T + pn.toff);
  constexpr unsigned kPair = sizeof(V2), kBlockBytes = 256u * sizeof(V2);
  const unsigned tb = (unsigned)(t & 255) * kPair;
  auto ld = [&](int cb) { return *reinterpret_cast<const V2*>(P + ((unsigned)cb * kBlockBytes + tb)); };
  V2 m[8];
  {
    const int nblk0 = min(kMfStage, ncol) / C;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int cb = q + u * NQ;
      m[u] = cb < nblk0 ? ld(cb) : V2{0, 0};
    }
  }
  double pre = 0.0;  // epilogue input of row r (threads t < R)
  int xi = 0;
  if (t < R && r < rows) {
    if (S == 1) pre = gathered(gptr, gsrc, vec, pn.gat + k + r);
    if (S == 2) pre = vec[pn.vec + r];
    if (KIND == 3) xi = idx[pn.idx + r];
    if (KIND == 7) xi = idx[pn.idx + piv[pn.piv + r]];
  }
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  bool have = true;  // m holds the first batch of the first stage
  for (int s0 = 0; s0 < ncol; s0 += kMfStage) {
    const int ns = min(kMfStage, ncol - s0), nblk = ns / C;
    __syncthreads();  // the previous stage is consumed
    for (int j = t; j < kMfStage; j += 256 * NQ) {
      const int c = pn.c_lo + s0 + j;
      double v = 0.0;  // beyond the block or the stage: zero (the tile holds zeros there too)
      if (j < ns && c < cols) {
        if (KIND == 0) {
          const int pj = piv[pn.piv + c];
          const int g = idx[pn.idx + pj];
          v = fma(b[g], req[g], gathered(gptr, gsrc, vec, pn.gat + pj));
        } else if (KIND == 4) {
          v = b[idx[pn.idx + c]] + gathered(gptr, gsrc, vec, pn.gat + c);
        } else if (S == 2) {
          v = x[idx[pn.idx + k + c]];
        } else {
          v = vec[pn.vec + c];
        }
      }
      xs[j] = v;
    }
    __syncthreads();
    const int sb = s0 / C;
    for (int cb = q; cb < nblk; cb += 8 * NQ) {
      if (!have) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int c = cb + u * NQ;
          m[u] = c < nblk ? ld(sb + c) : V2{0, 0};
        }
      }
      have = false;
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        // a block beyond the stage holds zeros in m and meets a finite staged value (the index
        // wraps inside the stage buffer, every entry of which the staging wrote)
        const int j0 = ((cb + u * NQ) * C + 2 * p) & (kMfStage - 1);
        const int j1 = ((cb + (u + 1) * NQ) * C + 2 * p) & (kMfStage - 1);
        a0 = fma((double)m[u].x, xs[j0], a0);
        a1 = fma((double)m[u].y, xs[j0 + 1], a1);
        a2 = fma((double)m[u + 1].x, xs[j1], a2);
        a3 = fma((double)m[u + 1].y, xs[j1 + 1], a3);
      }
    }
  }
  ps[q * NP + p][l] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  if (NQ > 1) {
    if (t < 256) {  // thread (p, l) folds the quarters of its own partial
      double v = ps[p][l];
      for (int i = 1; i < NQ; ++i) v += ps[i * NP + p][l];
      ps[p][l] = v;
    }
    __syncthreads();
  }
  // the column pairs of a row, as a pairwise tree in a fixed order
#pragma unroll
  for (int h = NP / 2; h >= 1; h >>= 1) {
    if (t < h * R) ps[p][l] += ps[p + h][l];
    __syncthreads();
  }
  if (t >= R || r >= rows) return;
  const double sum = ps[0][t];
  if (S == 0) vec[pn.vec + r] = sum;
  else if (S == 1) vec[pn.vec + k + r] = pre - sum;
  else if (S == 2) vec[pn.vec + r] = pre - sum;
  else x[xi] = sum;
}

__global__ void mft_scale_kernel(int n, const int* __restrict__ unk, const double* __restrict__ req,
                                 const double* __restrict__ ys, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int g = unk[i];
    y[g] = req ? req[g] * ys[g] : ys[g];
  }
}

}  // namespace

MultifrontalLU::~MultifrontalLU() {
  release();
  for (Lane& l : lanes_) {
    if (l.rb) (void)rocblas_destroy_handle(l.rb);
    if (l.done) (void)hipEventDestroy(l.done);
    if (l.s) (void)hipStreamDestroy(l.s);
  }
  if (fork_) (void)hipEventDestroy(fork_);
}

void MultifrontalLU::ensure_lanes() {
  if (lanes_[0].s) return;
  MF_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
  for (Lane& l : lanes_) {
    MF_HIP(hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking));  // default priority
    MF_HIP(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
    MF_RB(rocblas_create_handle(&l.rb));
    MF_RB(rocblas_set_stream(l.rb, l.s));
  }
}

void MultifrontalLU::release() {
  dev_free(d_fr_);
  dev_free(d_idx_);
  dev_free(d_rel_);
  dev_free(d_piv_);
  dev_free(d_gptr_);
  dev_free(d_gsrc_);
  dev_free(d_small_);
  dev_free(d_tiles_);
  dev_free(d_ext_);
  dev_free(d_asm_);
  dev_free(d_status_);
  dev_free(d_F_);
  dev_free(d_inv_);
  dev_free(d_vec_);
  dev_free(d_ys_);
  dev_free(d_unk_);
  for (Lane& l : lanes_) {
    for (int i = 0; i < 2; ++i) {
      dev_free(l.cand[i]);
      dev_free(l.cnt[i]);
    }
    dev_free(l.lu11);
    dev_free(l.pairs);
  }
  dev_free(d_info_);
  dev_free(d_growth_);
  dev_free(d_gv_);
  dev_free(d_gvoff_);
  dev_free(d_panel_);
  dev_free(d_tile_t_);
  tiles_t_ready_ = false;
  dev_free(d_tile32_);
  tiles32_ready_ = false;
  dev_free(d_tile_);
  dev_free(d_r_);
  dev_free(d_ay_);
  dev_free(d_x0_);
  dev_free(d_dx_);
  analysed_ = false;
}

void MultifrontalLU::analyse(const MfInput& in, hipStream_t st) {
  release();
  ensure_lanes();
  mf_analyse(in, sym_);
  N_ = in.N;
  nrows_csr_ = (int)in.N;
  const MfSymbolic& s = sym_;
  dev_upload(d_fr_, s.fronts, st);
  dev_upload(d_idx_, s.idx, st);
  dev_upload(d_rel_, s.rel, st);
  dev_upload(d_piv_, s.piv0, st);
  dev_upload(d_gptr_, s.gptr, st);
  dev_upload(d_gsrc_, s.gsrc, st);
  dev_upload(d_small_, s.small_list, st);
  {
    static_assert(sizeof(MfI4) == sizeof(int4) && sizeof(MfI2) == sizeof(int2), "device vector layouts");
    int4* t = nullptr;
    dev_alloc(t, s.tiles.size());
    if (!s.tiles.empty()) MF_HIP(hipMemcpyAsync(t, s.tiles.data(), sizeof(int4) * s.tiles.size(), hipMemcpyHostToDevice, st));
    d_tiles_ = t;
    dev_alloc(d_ext_, s.ext.size());
    if (!s.ext.empty())
      MF_HIP(hipMemcpyAsync(d_ext_, s.ext.data(), sizeof(int2) * s.ext.size(), hipMemcpyHostToDevice, st));
  }
  {
    int4* g = nullptr;
    dev_alloc(g, s.gv.size());
    if (!s.gv.empty()) MF_HIP(hipMemcpyAsync(g, s.gv.data(), sizeof(int4) * s.gv.size(), hipMemcpyHostToDevice, st));
    d_gv_ = g;
    dev_upload(d_gvoff_, s.gv_off, st);
    std::vector<MfPanel> pn(s.gv.size());
    for (size_t i = 0; i < s.gv.size(); ++i) {
      const MfFront& f = s.fronts[s.gv[i].x];
      pn[i] = MfPanel{s.gv_off[i], s.gv[i].y, s.gv[i].z, s.gv[i].w, f.k, f.m, f.vec, f.gat, f.piv, f.idx, 0};
    }
    dev_upload(d_panel_, pn, st);
    dev_alloc(d_tile_, (size_t)s.gv_total);
  }
  dev_upload(d_asm_, s.asmap, st);
  dev_upload(d_unk_, s.unk, st);
  dev_alloc(d_status_, 1);
  dev_alloc(d_F_, (size_t)s.front_total);
  dev_alloc(d_inv_, (size_t)s.inv_total);
  dev_alloc(d_vec_, (size_t)s.nvec);
  dev_alloc(d_ys_, (size_t)N_);
  MF_HIP(hipMemsetAsync(d_ys_, 0, sizeof(double) * std::max<int64_t>(N_, 1), st));
  const int64_t groups = (s.big_k + kLuGroupRows - 1) / kLuGroupRows;
  for (Lane& l : lanes_) {
    for (int i = 0; i < 2; ++i) {
      dev_alloc(l.cand[i], (size_t)std::max<int64_t>(groups, 1) * kLuPanel);
      dev_alloc(l.cnt[i], (size_t)std::max<int64_t>(groups, 1));
    }
    dev_alloc(l.lu11, (size_t)kLuPanel * kLuPanel);
    dev_alloc(l.pairs, (size_t)(kLuOuter / kLuPanel) * kLuPairStride);
  }
  dev_alloc(d_info_, std::max<size_t>(s.fronts.size(), 1));
  dev_alloc(d_growth_, s.piv0.size());
  dev_alloc(d_r_, (size_t)N_);
  dev_alloc(d_ay_, (size_t)N_);
  dev_alloc(d_x0_, (size_t)N_);
  dev_alloc(d_dx_, (size_t)N_);
  MF_HIP(hipMemsetAsync(d_x0_, 0, sizeof(double) * std::max<int64_t>(N_, 1), st));
  MF_HIP(hipMemsetAsync(d_dx_, 0, sizeof(double) * std::max<int64_t>(N_, 1), st));
  MF_HIP(hipStreamSynchronize(st));
  analysed_ = true;
}

void MultifrontalLU::factor_big(int fi, const Lane& lane) {
  const MfFront& f = sym_.fronts[fi];
  const int64_t m = f.m, k = f.k;
  double* A = d_F_ + f.off;
  int* perm = d_piv_ + f.piv;
  int64_t* info = d_info_ + fi;
  const double one = 1.0, mone = -1.0;
  hipStream_t st = lane.s;
  rocblas_handle rb = lane.rb;
  int* cand[2] = {lane.cand[0], lane.cand[1]};
  int* cnt[2] = {lane.cnt[0], lane.cnt[1]};
  // the first k steps of the front's LU, pivot search over its first k rows only
  for (int64_t J0 = 0; J0 < k; J0 += kLuOuter) {
    const int64_t J1 = std::min<int64_t>(k, J0 + kLuOuter);
    int nl = 0;
    for (int64_t j0 = J0; j0 < J1; j0 += kLuPanel, ++nl) {
      const int bw = (int)std::min<int64_t>(kLuPanel, J1 - j0);
      int* pairs = lane.pairs + (int64_t)nl * kLuPairStride;
      const int* win = launch_tslu(A, m, k, j0, bw, cand, cnt, lane.lu11, pairs, perm, info, st);
      launch_calu_apply(A, m, k, J0, J1, j0, bw, pairs, win, lane.lu11, st);
      launch_calu_l21(A, m, m, j0, bw, lane.lu11, st);
      const int64_t rows = m - j0 - bw, cols = J1 - j0 - bw;
      if (rows > 0 && cols > 0)
        MF_RB(rocblas_dgemm_64(rb, rocblas_operation_none, rocblas_operation_none, rows, cols, bw, &mone,
                               A + (j0 + bw) + j0 * m, m, A + j0 + (j0 + bw) * m, m, &one,
                               A + (j0 + bw) + (j0 + bw) * m, m));
    }
    launch_calu_swap(A, m, m, J0, J1, lane.pairs, nl, st);
    const int64_t rest = m - J1;
    if (rest > 0) {
      MF_RB(rocblas_dtrsm_64(rb, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit,
                             J1 - J0, rest, &one, A + J0 + J0 * m, m, A + J0 + J1 * m, m));
      MF_RB(rocblas_dgemm_64(rb, rocblas_operation_none, rocblas_operation_none, rest, rest, J1 - J0, &mone,
                             A + J1 + J0 * m, m, A + J0 + J1 * m, m, &one, A + J1 + J1 * m, m));
    }
  }
  // L11^-1 and U11^-1: triangular solves against the identity
  double* Li = d_inv_ + f.inv;
  double* Ui = Li + k * k;
  hipLaunchKernelGGL(mf_eye_kernel, dim3((unsigned)((2 * k * k + 255) / 256)), dim3(256), 0, st, Li, (int)k);
  MF_RB(rocblas_dtrsm_64(rb, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_unit, k,
                         k, &one, A, m, Li, k));
  MF_RB(rocblas_dtrsm_64(rb, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                         rocblas_diagonal_non_unit, k, k, &one, A, m, Ui, k));
  hipLaunchKernelGGL(mf_growth_kernel, dim3((unsigned)((k + 3) / 4)), dim3(256), 0, st, A, (int)m, (int)k,
                     d_growth_ + f.piv);
}

int MultifrontalLU::factor(const double* csr_val, const int* csr_ptr, const double* req, hipStream_t st) {
  if (!analysed_) throw std::runtime_error("multifrontal factor before analysis");
  const int nf = (int)sym_.fronts.size();
  MF_HIP(hipMemsetAsync(d_F_, 0, sizeof(double) * std::max<int64_t>(sym_.front_total, 1), st));
  if (!sym_.piv0.empty())
    MF_HIP(hipMemcpyAsync(d_piv_, sym_.piv0.data(), sizeof(int) * sym_.piv0.size(), hipMemcpyHostToDevice, st));
  MF_HIP(hipMemsetAsync(d_info_, 0, sizeof(int64_t) * std::max(nf, 1), st));
  const int big = 0x7fffffff;
  MF_HIP(hipMemcpyAsync(d_status_, &big, sizeof(int), hipMemcpyHostToDevice, st));
  if (nrows_csr_ > 0)
    hipLaunchKernelGGL(mf_scatter_kernel, dim3((nrows_csr_ + 255) / 256), dim3(256), 0, st, nrows_csr_, csr_ptr,
                       csr_val, req, d_asm_, d_F_);
  for (const MfLevel& L : sym_.levels) {
    for (size_t s = 0; s < L.ext_beg.size(); ++s) {
      const int n = L.ext_end[s] - L.ext_beg[s];
      if (n > 0)
        hipLaunchKernelGGL(mf_extend_kernel, dim3(n), dim3(256), 0, st, d_fr_, d_ext_ + L.ext_beg[s], d_rel_, d_F_);
    }
    // the level's large fronts on the lanes and its small fronts on st run concurrently (they
    // write disjoint fronts); the lanes join st before the next level's extend-add
    bool used[kMfLanes] = {false};
    if (!L.big.empty()) {
      // the level's large fronts are independent: spread over the lanes (largest first, each to
      // the least loaded lane), forked from st after the extend-add
      MF_HIP(hipEventRecord(fork_, st));
      std::vector<int> order(L.big);
      auto work = [&](int i) {
        const double kk = sym_.fronts[i].k, bb = sym_.fronts[i].m - sym_.fronts[i].k;
        return kk * kk * kk / 1.5 + 2.0 * kk * kk * bb + 2.0 * kk * bb * bb;
      };
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
      double load[kMfLanes] = {0.0};
      for (int fi : order) {
        const int li = (int)(std::min_element(load, load + kMfLanes) - load);
        if (!used[li]) {
          MF_HIP(hipStreamWaitEvent(lanes_[li].s, fork_, 0));
          used[li] = true;
        }
        factor_big(fi, lanes_[li]);
        load[li] += work(fi) + 1e7;  // (a launch-latency floor per front)
      }
    }
    const int n0 = L.small_end[0] - L.small_beg[0], n1 = L.small_end[1] - L.small_beg[1],
              n2 = L.small_end[2] - L.small_beg[2], n3 = L.small_end[3] - L.small_beg[3];
    if (n0 > 0)
      hipLaunchKernelGGL((mf_small_factor_kernel<16, 16>), dim3(n0), dim3(256), 0, st, d_fr_,
                         d_small_ + L.small_beg[0], d_F_, d_inv_, d_piv_, d_status_, d_growth_);
    if (n1 > 0)
      hipLaunchKernelGGL((mf_small_factor_kernel<32, 32>), dim3(n1), dim3(256), 0, st, d_fr_,
                         d_small_ + L.small_beg[1], d_F_, d_inv_, d_piv_, d_status_, d_growth_);
    if (n2 > 0)
      hipLaunchKernelGGL((mf_small_factor_kernel<64, 64>), dim3(n2), dim3(256), 0, st, d_fr_,
                         d_small_ + L.small_beg[2], d_F_, d_inv_, d_piv_, d_status_, d_growth_);
    if (n3 > 0)
      hipLaunchKernelGGL((mf_small_factor_kernel<128, 24>), dim3(n3), dim3(256), 0, st, d_fr_,
                         d_small_ + L.small_beg[3], d_F_, d_inv_, d_piv_, d_status_, d_growth_);
    if (L.tile_end > L.tile_beg)
      hipLaunchKernelGGL(mf_schur_kernel, dim3(L.tile_end - L.tile_beg), dim3(256), 0, st, d_fr_,
                         d_tiles_ + L.tile_beg, d_F_);
    for (int li = 0; li < kMfLanes; ++li)
      if (used[li]) {
        MF_HIP(hipEventRecord(lanes_[li].done, lanes_[li].s));
        MF_HIP(hipStreamWaitEvent(st, lanes_[li].done, 0));
      }
  }
  pack_tiles(false, st);
  tiles_t_ready_ = false;
  tiles32_ready_ = false;
  MF_HIP(hipGetLastError());
  int status = 0;
  std::vector<int64_t> info(std::max(nf, 1), 0);
  MF_HIP(hipMemcpyAsync(&status, d_status_, sizeof(int), hipMemcpyDeviceToHost, st));
  MF_HIP(hipMemcpyAsync(info.data(), d_info_, sizeof(int64_t) * std::max(nf, 1), hipMemcpyDeviceToHost, st));
  MF_HIP(hipStreamSynchronize(st));
  int bad = status != big ? status : -1;
  for (int i = 0; i < nf && bad < 0; ++i)
    if (info[i] > 0) bad = i;
  return bad < 0 ? 0 : 1 + bad;
}

template <int K>
static void launch_pack(int R, int n, const MfFront* fr, const int4* items, const int64_t* toff, const double* F,
                        const double* INV, double* T, hipStream_t st) {
  if (R == 16) hipLaunchKernelGGL((mfs_pack_kernel<K, 16>), dim3(n), dim3(256), 0, st, fr, items, toff, F, INV, T);
  else if (R == 8) hipLaunchKernelGGL((mfs_pack_kernel<K, 8>), dim3(n), dim3(256), 0, st, fr, items, toff, F, INV, T);
  else hipLaunchKernelGGL((mfs_pack_kernel<K, 4>), dim3(n), dim3(256), 0, st, fr, items, toff, F, INV, T);
}

void MultifrontalLU::pack_tiles(bool transposed, hipStream_t st) {
  double* T = transposed ? d_tile_t_ : d_tile_;
  for (const MfLevel& L : sym_.levels)
    for (int kind = 0; kind < 4; ++kind) {
      const int n = L.gv_end[kind] - L.gv_beg[kind];
      if (n <= 0) continue;
      const int4* items = d_gv_ + L.gv_beg[kind];
      const int64_t* toff = d_gvoff_ + L.gv_beg[kind];
      const int R = L.prow[kind];
      switch (kind + (transposed ? 4 : 0)) {
        case 0: launch_pack<0>(R, n, d_fr_, items, toff, d_F_, d_inv_, T, st); break;
        case 1: launch_pack<1>(R, n, d_fr_, items, toff, d_F_, d_inv_, T, st); break;
        case 2: launch_pack<2>(R, n, d_fr_, items, toff, d_F_, d_inv_, T, st); break;
        case 3: launch_pack<3>(R, n, d_fr_, items, toff, d_F_, d_inv_, T, st); break;
        case 4: launch_pack<4>(R, n, d_fr_, items, toff, d_F_, d_inv_, T, st); break;
        case 5: launch_pack<5>(R, n, d_fr_, items, toff, d_F_, d_inv_, T, st); break;
        case 6: launch_pack<6>(R, n, d_fr_, items, toff, d_F_, d_inv_, T, st); break;
        default: launch_pack<7>(R, n, d_fr_, items, toff, d_F_, d_inv_, T, st); break;
      }
    }
  MF_HIP(hipGetLastError());
}

void MultifrontalLU::prepare_transposed(hipStream_t st) {
  if (!analysed_) throw std::runtime_error("multifrontal transposed tiles before analysis");
  if (tiles_t_ready_) return;
  if (!d_tile_t_) dev_alloc(d_tile_t_, (size_t)sym_.gv_total);
  pack_tiles(true, st);
  tiles_t_ready_ = true;
}

template <int K, int NQ, int R, class TT>
static void launch_panels(int n, const MfPanel* pn, const TT* T, const int* idx, const int* piv, const int* gptr,
                          const int* gsrc, const double* b, const double* req, double* x, double* vec, hipStream_t st) {
  hipLaunchKernelGGL((mfs_panel_kernel<K, NQ, R, TT>), dim3(n), dim3(256 * NQ), 0, st, pn, T, idx, piv, gptr, gsrc, b,
                     req, x, vec);
}

template <int K, class TT>
static void launch_panels_k(int wide, int R, int n, const MfPanel* pn, const TT* T, const int* idx, const int* piv,
                            const int* gptr, const int* gsrc, const double* b, const double* req, double* x,
                            double* vec, hipStream_t st) {
#define MF_PANELS(NQ, RR) launch_panels<K, NQ, RR, TT>(n, pn, T, idx, piv, gptr, gsrc, b, req, x, vec, st)
  if (R == 16 && wide) MF_PANELS(4, 16);  // the analysis makes only 16-row panels wide
  else if (R == 16) MF_PANELS(1, 16);
  else if (R == 8) MF_PANELS(1, 8);
  else MF_PANELS(1, 4);
#undef MF_PANELS
}

void MultifrontalLU::launch_seg(int kind, const MfLevel& L, const double* b, const double* req, double* x,
                                hipStream_t st, bool f32) {
  const int s = kind & 3;
  const int n = L.gv_end[s] - L.gv_beg[s];
  if (n <= 0) return;
  const MfPanel* pn = d_panel_ + L.gv_beg[s];
  const int w = L.pwide[s], R = L.prow[s];
  const double* T = kind < 4 ? d_tile_ : d_tile_t_;
  if (f32) {  // the FP32 tiles of A x = b
#define MF_K32(K) launch_panels_k<K, float>(w, R, n, pn, d_tile32_, d_idx_, d_piv_, d_gptr_, d_gsrc_, b, req, x, d_vec_, st)
    switch (kind) {
      case 0: MF_K32(0); break;
      case 1: MF_K32(1); break;
      case 2: MF_K32(2); break;
      default: MF_K32(3); break;
    }
#undef MF_K32
    return;
  }
#define MF_K(K) launch_panels_k<K, double>(w, R, n, pn, T, d_idx_, d_piv_, d_gptr_, d_gsrc_, b, req, x, d_vec_, st)
  switch (kind) {
    case 0: MF_K(0); break;
    case 1: MF_K(1); break;
    case 2: MF_K(2); break;
    case 3: MF_K(3); break;
    case 4: MF_K(4); break;
    case 5: MF_K(5); break;
    case 6: MF_K(6); break;
    default: MF_K(7); break;
  }
#undef MF_K
}

void MultifrontalLU::solve(const double* b, const double* req, double* x, hipStream_t st, bool f32) {
  if (f32 && !tiles32_ready_) throw std::runtime_error("multifrontal FP32 solve before prepare_f32");
  const int H = (int)sym_.levels.size();
  for (int h = 0; h < H; ++h) {
    launch_seg(0, sym_.levels[h], b, req, x, st, f32);
    launch_seg(1, sym_.levels[h], b, req, x, st, f32);
  }
  for (int h = H - 1; h >= 0; --h) {
    launch_seg(2, sym_.levels[h], b, req, x, st, f32);
    launch_seg(3, sym_.levels[h], b, req, x, st, f32);
  }
}

__global__ void mf_to_f32_kernel(int64_t n, const double* __restrict__ a, float* __restrict__ o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (float)a[i];
}

void MultifrontalLU::prepare_f32(hipStream_t st) {
  if (!analysed_) throw std::runtime_error("multifrontal FP32 tiles before analysis");
  if (tiles32_ready_) return;
  if (!d_tile32_) dev_alloc(d_tile32_, (size_t)sym_.gv_total);
  if (sym_.gv_total > 0)
    hipLaunchKernelGGL(mf_to_f32_kernel, dim3((unsigned)((sym_.gv_total + 255) / 256)), dim3(256), 0, st,
                       sym_.gv_total, d_tile_, d_tile32_);
  MF_HIP(hipGetLastError());
  tiles32_ready_ = true;
}

void MultifrontalLU::solve_t(const double* c, const double* req, double* y, hipStream_t st) {
  if (!tiles_t_ready_) throw std::runtime_error("multifrontal transposed solve before prepare_transposed");
  const int H = (int)sym_.levels.size();
  for (int h = 0; h < H; ++h) {
    launch_seg(4, sym_.levels[h], c, nullptr, d_ys_, st);
    launch_seg(5, sym_.levels[h], c, nullptr, d_ys_, st);
  }
  for (int h = H - 1; h >= 0; --h) {
    launch_seg(6, sym_.levels[h], c, nullptr, d_ys_, st);
    launch_seg(7, sym_.levels[h], c, nullptr, d_ys_, st);
  }
  if (sym_.nunk > 0)
    hipLaunchKernelGGL(mft_scale_kernel, dim3((unsigned)((sym_.nunk + 255) / 256)), dim3(256), 0, st, (int)sym_.nunk, d_unk_,
                       req, d_ys_, y);
}

bool MultifrontalLU::unstable_pivots(double limit, std::vector<int>& lift, bool& root_zero, hipStream_t st) {
  const size_t np = sym_.piv0.size();
  std::vector<unsigned long long> g(np);
  std::vector<int> pv(np);
  if (np) {
    MF_HIP(hipMemcpyAsync(g.data(), d_growth_, sizeof(unsigned long long) * np, hipMemcpyDeviceToHost, st));
    MF_HIP(hipMemcpyAsync(pv.data(), d_piv_, sizeof(int) * np, hipMemcpyDeviceToHost, st));
  }
  MF_HIP(hipStreamSynchronize(st));
  lift.clear();
  root_zero = false;
  for (const MfFront& f : sym_.fronts)
    for (int j = 0; j < f.k; ++j) {
      double v;
      std::memcpy(&v, &g[f.piv + j], sizeof(v));
      if (v <= limit) continue;  // (a NaN fails the test too)
      if (f.parent < 0) {
        root_zero = root_zero || !(v < HUGE_VAL);
        continue;
      }
      lift.push_back(sym_.idx[f.idx + j]);              // the pivot's column
      lift.push_back(sym_.idx[f.idx + pv[f.piv + j]]);  // and the row it was taken from
    }
  std::sort(lift.begin(), lift.end());
  lift.erase(std::unique(lift.begin(), lift.end()), lift.end());
  return !lift.empty();
}

void MultifrontalLU::solve_correction(const double* b, const double* ay, const double* y, const double* req, double* x,
                                      hipStream_t st, bool f32) {
  launch_adj_refine_in<double>(b, ay, (int)N_, d_r_, st);
  solve(d_r_, req, d_dx_, st, f32);
  hipLaunchKernelGGL(mf_axpy_kernel, dim3((unsigned)((N_ + 255) / 256)), dim3(256), 0, st, N_, y, d_dx_, x);
}

void MultifrontalLU::solve_refined(const double* b, const double* req, double* x, const int* csr_ptr,
                                   const int* csr_col, const double* csr_val, hipStream_t st) {
  solve(b, req, d_x0_, st);
  launch_spmv((int)N_, csr_ptr, csr_col, csr_val, d_x0_, d_ay_, st);
  launch_adj_refine_in<double>(b, d_ay_, (int)N_, d_r_, st);
  solve(d_r_, req, d_dx_, st);
  hipLaunchKernelGGL(mf_axpy_kernel, dim3((unsigned)((N_ + 255) / 256)), dim3(256), 0, st, N_, d_x0_, d_dx_, x);
}

}  // namespace aem
