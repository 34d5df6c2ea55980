// Multifrontal sparse LU of the discharge system (aem_sparse.hip).
//
// The reference factors the sparse discharge matrix once with SuperLU (scipy.sparse.linalg.splu,
// andfn/hpc/hpc_solve.py:121, on the csc_matrix of :885) and solves with it every iteration
// (:651).  Its structure is a union of cliques: the head equation of a discharge element couples
// every discharge element and the constant of each fracture it lies in, a fracture's continuity
// row couples its discharge elements (hpc_solve.py:1117-1201).  Here the unknowns are ordered by
// nested dissection of the fracture graph (fractures as vertices, intersections as edges):
// a fracture set is cut at the median of the fracture centres along its longest extent, and the
// intersections between the two halves form the separator eliminated after both halves.  The
// elimination tree's nodes are dense fronts (pivot unknowns U, boundary unknowns B):
//
//   factor, level by level from the leaves: extend-add of the children's Schur complements,
//   LU of the pivot block with partial pivoting over the front's own rows, L21 = F21 U11^-1,
//   U12 = L11^-1 P F12, F22 -= L21 U12; the inverses of L11 and U11 are kept for the solves;
//   solve (A x = b and A^T y = c), four launches per level, each covering every front of the level.
//
// Fronts with at most 64 pivots, and with 65 to 128 and a boundary of at most 512 rows, are
// factored by one workgroup each in LDS (all fronts of a level and class in one launch, their Schur
// updates in one tiled launch); larger fronts by the tournament-pivoting kernels of aem_lu.hip with
// the pivot search limited to the front's rows, and rocBLAS, on several streams beside them.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <vector>

#include "mf_symbolic.hpp"

namespace aem {

// a solve panel with its front's fields (one record per workgroup of the solve kernels)
struct MfPanel {
  int64_t toff;               // first value of its tile
  int r0, c_lo, nblk;         // first row, first column, column blocks
  int k, m, vec, gat, piv, idx;  // of its front
  int pad;
};
static_assert(sizeof(MfPanel) == 48, "panel record layout");

class MultifrontalLU {
 public:
  MultifrontalLU() = default;
  ~MultifrontalLU();
  MultifrontalLU(const MultifrontalLU&) = delete;
  MultifrontalLU& operator=(const MultifrontalLU&) = delete;

  // ordering, elimination tree, front layout (mf_analyse) and the device tables, once per network;
  // throws std::runtime_error when the network does not fit the kernels
  void analyse(const MfInput& in, hipStream_t st);
  // numeric factorisation of R A (row scales req[N]) from A's CSR values in the analysed pattern;
  // returns 0, or 1 + the first front with an exactly zero pivot
  int factor(const double* csr_val, const int* csr_ptr, const double* req, hipStream_t st);
  // x = A^-1 b (b and x in the solver's layout; x is written on the analysed unknowns only);
  // f32: through the factors rounded to FP32 (prepare_f32), accumulated in FP64
  void solve(const double* b, const double* req, double* x, hipStream_t st, bool f32 = false);
  // the factors rounded to FP32 for solve(..., f32), once per factorisation (outside graph capture)
  void prepare_f32(hipStream_t st);
  // the transposed solve's tiles, once per factorisation (before solve_t; outside graph capture)
  void prepare_transposed(hipStream_t st);
  // y = A^-T c (req == nullptr: y = (R A)^-T c, the row scales left to the caller)
  void solve_t(const double* c, const double* req, double* y, hipStream_t st);
  // x = A^-1 b with one step of iterative refinement against A's CSR values (residual in FP64):
  // the pivot blocks are applied through their explicit inverses, which is accurate to their
  // condition number times the rounding unit; one step brings the solve back to rounding
  void solve_refined(const double* b, const double* req, double* x, const int* csr_ptr, const int* csr_col,
                     const double* csr_val, hipStream_t st);
  // x = y + A^-1 (b - A y) given ay = A y (all in the solver's layout): one solve, the correction
  // of an iterate y, accurate to the condition number times the rounding unit times |b - A y|
  void solve_correction(const double* b, const double* ay, const double* y, const double* req, double* x,
                        hipStream_t st, bool f32 = false);
  // After factor: the unknowns of every pivot whose L21 column exceeds limit in magnitude (or that is
  // zero), its column and the row it was taken from, to be lifted to the parent front by a new
  // analysis (returns whether there are any).  root_zero: a zero pivot in a root front, which has no
  // parent to pass it to (the matrix is singular).
  bool unstable_pivots(double limit, std::vector<int>& lift, bool& root_zero, hipStream_t st);

  bool ready() const { return analysed_; }
  const MfSymbolic& symbolic() const { return sym_; }
  int64_t nnz_factors() const { return sym_.nnz_factors; }  // entries of L and U (diagonals included)

 private:
  bool analysed_ = false;
  MfSymbolic sym_;
  int64_t N_ = 0;
  int nrows_csr_ = 0;

  // device tables
  MfFront* d_fr_ = nullptr;
  int* d_idx_ = nullptr;
  int* d_rel_ = nullptr;
  int* d_piv_ = nullptr;
  int* d_gptr_ = nullptr;
  int* d_gsrc_ = nullptr;
  int* d_small_ = nullptr;
  int4* d_tiles_ = nullptr;
  int2* d_ext_ = nullptr;
  int4* d_gv_ = nullptr;        // solve panels (MfSymbolic::gv)
  int64_t* d_gvoff_ = nullptr;  // per solve panel: first value of its tile
  double* d_tile_ = nullptr;    // the panel tiles (repacked after every factorisation)
  MfPanel* d_panel_ = nullptr;  // per solve panel: its record
  double* d_tile_t_ = nullptr;  // the transposed solve's tiles (packed on demand)
  bool tiles_t_ready_ = false;
  float* d_tile32_ = nullptr;   // the tiles rounded to FP32 (the mixed form's correction solves)
  bool tiles32_ready_ = false;
  int64_t* d_asm_ = nullptr;
  int* d_unk_ = nullptr;     // every analysed unknown (its position in the solver's layout)
  int* d_status_ = nullptr;
  double* d_F_ = nullptr;    // front values
  double* d_inv_ = nullptr;  // L11^-1, U11^-1 of every front
  double* d_vec_ = nullptr;  // per-front solve vectors
  double* d_ys_ = nullptr;   // transposed solve: unscaled result in the solver's layout
  // The large fronts of a level are factored concurrently on kMfLanes streams of default priority,
  // each with its own rocBLAS handle (its own workspace) and tournament work buffers.  HIP shares
  // the process's hardware queues (4 by default) among its streams; 6 lanes beside the handle's
  // stream measured fastest (factorisation of the 64-realisation batch: 163, 94, 120, 96, 73 and
  // 79 ms with 1, 2, 3, 4, 6 and 8 lanes; cfg3: 185, 124, 138, 122, 108 and 111 ms).
  static constexpr int kMfLanes = 6;
  struct Lane {
    hipStream_t s = nullptr;
    rocblas_handle rb = nullptr;
    hipEvent_t done = nullptr;
    int* cand[2] = {nullptr, nullptr};
    int* cnt[2] = {nullptr, nullptr};
    double* lu11 = nullptr;
    int* pairs = nullptr;
  };
  Lane lanes_[kMfLanes];
  hipEvent_t fork_ = nullptr;
  void ensure_lanes();
  int64_t* d_info_ = nullptr;
  unsigned long long* d_growth_ = nullptr;  // per pivot: bits of max |L21 column| (inf: zero pivot)
  double* d_r_ = nullptr;  // refinement: residual, A x0, first solve, correction
  double* d_ay_ = nullptr;
  double* d_x0_ = nullptr;
  double* d_dx_ = nullptr;

  void release();
  void factor_big(int f, const Lane& lane);
  void launch_seg(int kind, const MfLevel& L, const double* b, const double* req, double* x, hipStream_t st,
                  bool f32 = false);
  void pack_tiles(bool transposed, hipStream_t st);
};

}  // namespace aem
