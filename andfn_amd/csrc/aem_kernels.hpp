// Device-side views and kernel launchers of the analytic-element solve.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aem {

constexpr int kMaxPts = 512;  // >= 2 * coefficient capacity handled by the fit kernel

// sweep modes: which points, whether the target's own element is excluded, whether the log
// (discharge) terms and fracture constants are included, and what is written
enum {
  MODE_COLLOC = 0,        // collocation points, own element excluded, full potential; Re (+C) or Im for the fit
  MODE_RHS = 1,           // discharge points, series only (q = C = 0), Re
  MODE_POINTS = 2,        // explicit points, full potential, head = Re Omega / t
  MODE_COLLOC_HEADS = 3,  // collocation points, full potential, head
  MODE_DISCH_HEADS = 4,   // discharge points, full potential, head
  MODE_POINTS_OMEGA = 5   // explicit points, full complex potential Omega (C_f included), two values per point
};
enum { KIND_LINE = 0, KIND_BOUND = 1, KIND_WELL = 2, KIND_IMPC = 3, KIND_IMPL = 4 };

// one (fracture, member element) pair: the element as a source of potential in that fracture
struct Slot {
  int elem;
  int kind;
  double sign;     // -1 for an intersection seen from its second fracture, else +1
  double2 alpha;   // line: 2/(b - a); circle: (1/r, 0)
  double2 beta;    // line: -(a + b)/(b - a); circle: -c/r
};

struct Tile {
  int frac;
  int sbeg, send;  // side range of the fracture in the target set
  int p0, pend;    // first point of the tile, end of the fracture's points
};

struct NetView {
  int M;  // coefficient capacity (stride of coef / wcoef)
  const int* etype;
  const int* efrac0;
  const int* efrac1;
  const int* ncoef;
  const int* nint;
  int* ncoef_w;
  int* nint_w;
  const double2* ep;      // [E][4] = ep0[0], ep0[1], ep1[0], ep1[1]
  const double* radius;
  const double2* center;
  const double* phi;
  const double* q;
  double* q_w;
  double2* coef;          // [E][M]
  double2* wcoef;         // [E][M]
  const float2* coef32;   // [E][M] single-precision copy (mixed-precision far field)
  double far2;            // mixed precision: a line source is evaluated in FP32 where |Z|^2 >= far2 on every lane
  int real_lines;         // 1: every KIND_LINE source (intersection, constant-head line) has real coefficients
  const int* stop;        // device flag of the run-ahead iteration (null outside it): set, the state kernels return
  const double* t;
  const double* constant;
  double* constant_w;
  const int* mem_ptr;     // [F + 1]
  const Slot* slots;      // [S]
};

struct FitView {
  double* error;
  double* error_old;
  double* error_old2;
  double* error_coef;
  double* dpsi_corr;       // [E][2M]
  const int* cross_ptr;    // [E] start of an element's crossing list
  const int* cross_cap;    // [E] capacity of the list
  int* cross_cnt;          // [E]
  int2* cross;             // (edge position, crossing element)
  int* cross_sign;
  int* status;             // bit 0: a crossing list overflowed
  const int* stop;         // as NetView::stop
  // (cos, sin) of k pi / (2 n) for k < 4 n, n = 1 .. nmax, at entry 2 n (n - 1) + k: the DFT
  // weights of every collocation family (launch_trig_tables)
  const double2* trig;
  const float2* trig32;
};
// entries of the tables for n = 1 .. nmax
inline int64_t trig_entries(int nmax) { return 2 * (int64_t)(nmax + 1) * nmax; }
void launch_trig_tables(int nmax, double2* trig, float2* trig32, hipStream_t st);

// a kernel of the run-ahead iteration returns at once when its stop flag is set (the head
// residual met the tolerance earlier in the stream), so iterations enqueued past the stop leave
// the state unchanged
template <class V>
__device__ __forceinline__ bool stopped(const V& v) {
  return v.stop != nullptr && *(volatile const int*)v.stop != 0;
}

// control block of the run-ahead iteration (device memory)
struct IterCtl {
  int stop;  // 1 once an iteration's head residual met the tolerance
  int it;    // iterations whose residual has been decided
};
// per iteration, written by the device into pinned host memory: residual and its row, and the
// iteration that stopped the run (0: none yet)
struct IterHostLog {
  double* resid;
  int* arg;
  int* stop_it;
  int cap;
  unsigned long long* stamps;  // device memory, [cap][6] (launch_iter_stamp)
};

// Unknowns of the discharge system: nblk independent blocks (connected components of the
// network) of npad unknowns each, real unknowns first and identity padding after.  Unknown
// g = b npad + local; block b's matrix is the column-major npad x npad slab at b npad^2.
// Discharge system of one or more independent parts (connected components).  Each part is a
// dense column-major block padded to its size class: the parts are grouped into at most
// kMaxSysClasses classes of similar size, class k holding blocks of npad[k] unknowns stored one
// after the other from matrix entry a0[k] and unknown u0[k] (u0[ncls] = total unknowns).
constexpr int kMaxSysClasses = 8;
struct SysLayout {
  int ncls;
  int64_t u0[kMaxSysClasses + 1];
  int64_t a0[kMaxSysClasses];
  int64_t npad[kMaxSysClasses];
  const int* de_pos;  // [E] unknown of each discharge element (-1 for other elements)
  const int* fpos;    // [F] unknown of each fracture constant
};

// position of (row r, column c) of the same block in the block storage
__host__ __device__ __forceinline__ int64_t mat_at(int64_t r, int64_t c, const SysLayout& s) {
  int k = 0;
  while (k + 1 < s.ncls && r >= s.u0[k + 1]) ++k;
  const int64_t np = s.npad[k];
  const int64_t base = s.u0[k] + (r - s.u0[k]) / np * np;
  return s.a0[k] + (base - s.u0[k]) * np + (c - base) * np + (r - base);
}

struct TargetView {
  const Tile* tiles;
  const int* side_elem;
  const int* side_which;
  const int* pt_off;       // [nsides + 1]
  const double2* pt_z;     // explicit points (MODE_POINTS) or null
};

// One launch of the streamline tracer over lines act[0..n): every line advances from its saved
// state through as many segments as its output rows hold, then saves its state again.
struct TraceArgs {
  int n;
  const int* act;          // [n] lines of this launch
  // per line (indexed by line): state read and written by every launch
  int* frac;               // current fracture
  double2* z0;             // point the tracking continues from
  double2* z_start;        // first point recorded for the next segment
  double* elev;            // elevation (the split at intersection exits)
  int64_t* length;         // points recorded so far
  int* nseg;               // segments recorded so far
  int* done;               // 1 once the line has ended
  int* exit_elem;          // discharge element crossed last, -1 if none
  // network
  const double* fradius;   // [F] fracture radii
  const double* aperture;  // [F] apertures (speed = |w| / aperture)
  const double* frame;     // [F][12] centre, x vector, y vector, normal of each fracture
  double ds;               // step as a fraction of the fracture radius
  int backward;            // 1: against the flow
  int max_segments;        // a line ends after this many segments
  int64_t max_length;      // the point budget of DFN.streamline_tracking
  // this launch's output rows (indexed by position in act)
  int64_t capacity;        // points per row (>= max_length + 2)
  int seg_capacity;        // segments per row (>= 1)
  double2* pts;            // [n][capacity]
  double* vel;             // [n][capacity]
  int* seg_frac;           // [n][seg_capacity]
  int* seg_len;            // [n][seg_capacity]
  int64_t* out_pts;        // [n] points written this launch
  int* out_segs;           // [n] segments written this launch
};

// packs the rows of a trace launch: row k's out_pts[k] points to poff[k], its out_segs[k]
// segments to soff[k]
struct TracePack {
  int n;
  int64_t capacity;
  int seg_capacity;
  const double2* pts;
  const double* vel;
  const int* seg_frac;
  const int* seg_len;
  const int64_t* out_pts;
  const int* out_segs;
  const int64_t* poff;
  const int64_t* soff;
  double2* ppts;
  double* pvel;
  int* pseg_frac;
  int* pseg_len;
};

struct FitItem {
  int elem;
  int s0, s1;  // side indices in the target set (s1 = -1 unless intersection)
};

// The fit's DFT as a GEMM: a tile is up to 16 sides of one target set (sides[s0 .. s0 + ns) of
// the tile list's side array) that share the point count n, the coefficient count m and the
// angle family (0: theta_i = (i + 1/2) pi / n, lines and impermeable objects; 1: theta_i =
// 2 pi i / n, bounding circles).
struct DftTile {
  int s0, ns, n, m, family;
};

struct RhsItem {
  int elem;
  int s0, s1;
};

struct GrowParams {
  double max_error;
  double max_ratio;
  int nit;
  int max_coef;
  int increase;
};

struct IterScalars {
  double error;
  double error_coef;
  double dq_max;
  double yold_max;
  double residual;  // max |A y - b| over the head equations, in metres
  int argmax;
  int arg_residual;
};

// after the head residual of an iteration: record it, stop when may_stop and it is <= tol
void launch_iter_decide(const IterScalars* sc, IterCtl* ctl, int may_stop, double tol, IterHostLog log,
                        hipStream_t st);
void launch_iter_reset(IterCtl* ctl, int it, hipStream_t st);
// device wall clock (constant rate) of the iteration's phase boundaries: stamps[6 it + slot],
// slot 0 iteration start, 1 head residual decided (written by iter_decide_kernel), 2 solve
// accepted, 3 sweep done, 4 fits and wells done; next = 1 before the decision of the iteration
void launch_iter_stamp(const IterCtl* ctl, unsigned long long* stamps, int cap, int slot, int next, hipStream_t st);

// points_per_lane 1 or 2: the target view's tiles must hold 64 x points_per_lane points each
void launch_sweep(int mode, const NetView& net, const TargetView& tg, int ntiles, int K, double* val, hipStream_t st,
                  bool mixed = false, int points_per_lane = 1);
// the collocation or right-hand-side sweep as one persistent launch of nwg workgroups over the
// XCD tile layout (ntiles a multiple of 8); queue: 9 zeroed counters, left zeroed by the launch
void launch_sweep_persistent(int mode, const NetView& net, const TargetView& tg, int ntiles, int K, double* val,
                             hipStream_t st, bool mixed, int points_per_lane, int nwg, unsigned* queue);
// ---- DFN construction (aem_geometry.hip)
struct DiskArrays {
  const double* center;  // [n][3]
  const double* normal;  // [n][3] unit normal
  const double* xvec;    // [n][3] in-plane basis
  const double* yvec;    // [n][3]
  const double* radius;  // [n]
};
struct ChordOptions {
  int shorten_a, shorten_b;  // shrink the chord in disk a's / b's plane about its midpoint
  int clip;                  // 0 off, 1 inside rho r of both disks, 2 inside rho r of disk b
  double se_factor, rho;
};
void geo_chords(int64_t np, const int64_t* ia, const int64_t* ib, const DiskArrays& A, const DiskArrays& B,
                const ChordOptions& opt, double* e0, double* e1, int8_t* ok, hipStream_t st);
void geo_pair_count(int64_t n, const double* c, const double* r, int64_t* count, hipStream_t st);
void geo_pair_write(int64_t n, const double* c, const double* r, const int64_t* offset, int64_t* pi, int64_t* pk,
                    hipStream_t st);
// the message andfn_mi_last_error() returns on this thread
void set_last_error(const char* msg);

// workgroups of the persistent sweep that fit on the device at once (occupancy x compute units)
int sweep_persistent_groups(int mode, bool mixed, int points_per_lane, int cus);
// w = -dOmega/dz at the explicit points of a target view (two values per point); jit != 0 gives
// the jitted form (hpc_fracture.calc_w), 0 the object model's (Fracture.calc_w)
void launch_w_points(const NetView& net, const TargetView& tg, int ntiles, int exclude, int jit, double* val,
                     hipStream_t st);
void launch_trace(const NetView& net, const TraceArgs& ta, hipStream_t st);
void launch_trace_pack(const TracePack& tp, hipStream_t st);
void launch_coef_to_f32(const NetView& net, int ne, float2* out, hipStream_t st);
// unwrap (circle-type sides, in place in val), DFT sums of every side into dft (side s: X_j at
// s 2M + j, Y_j at s 2M + M + j), then the element solves
// citems: the circle-type items among items (types 1, 4, 5), whose sides are unwrapped.
// blend: 0 the fit only, 1 with the first damping step, 2 with both damping steps of the
// iteration (the second is otherwise launch_blend's)
void launch_fit(const NetView& net, const FitView& fv, const FitItem* items, int nitems, const FitItem* citems,
                int ncitems, double* val,
                const int* pt_off, const DftTile* tiles, int ntiles, const int* dft_sides, double* dft, double damping,
                int blend, hipStream_t st, bool f32 = false);
void launch_blend(const NetView& net, int ne, double damping, hipStream_t st);
// steps = 2: also the second damping step of the iteration (the one launch_blend applies)
void launch_wells(const NetView& net, const FitView& fv, const int* wells, int nw, double damping, hipStream_t st,
                  int steps = 1);
void launch_grow(const NetView& net, const FitView& fv, int ne, GrowParams gp, const int* sched, int* grow,
                 hipStream_t st);
void launch_csr_gather(const double* A, const SysLayout& sys, int64_t nrows, const int* row_ptr, const int* col,
                       double* val, hipStream_t st);
void launch_spmv(int n, const int* row_ptr, const int* col, const double* val, const double* x, double* y,
                 hipStream_t st);
// batch > 1: the same product on `batch` slabs, A + z sA, x + z sx, y + z sx
template <class T>
void launch_gemv_cm(const T* A, int64_t lda, int m, int ncols, const T* x, T* y, T alpha, T beta, hipStream_t st,
                    int batch = 1, int64_t sA = 0, int64_t sx = 0);
// y_j = B_j x_j for the stacked nbk x nbk diagonal blocks (tri: 0 full, 1 lower, 2 upper) of
// `batch` systems of n unknowns (blocks + z sB, x and y + z n)
template <class T>
void launch_gemv_diag(const T* blocks, int nbk, int n, int tri, const T* x, T* y, hipStream_t st, int batch = 1,
                      int64_t sB = 0);
template <class T>
void launch_unit_diag(T* blocks, int nb, int nbk, hipStream_t st);
// zero the strictly upper (lower = 1) or strictly lower (lower = 0) triangle of nb stacked blocks
template <class T>
void launch_tri_zero(T* blocks, int nb, int nbk, int lower, hipStream_t st);
// B_z = W_z for nb m x nc matrices (W packed, B with leading dimension ldb and stride sB)
template <class T>
void launch_copy_panels(const T* w, int64_t m, int64_t nc, int nb, T* b, int64_t ldb, int64_t sB, hipStream_t st);
void launch_to_f32(const double* a, int64_t n, float* b, hipStream_t st);

// Blocked LU with tournament pivoting (aem_lu.hip).  Inner panels of kLuPanel columns grouped
// into outer panels of kLuOuter columns; tournament groups of kLuGroupRows rows.  Buffers:
// cand[2] and cnt[2] hold ceil(n / kLuGroupRows) sets of kLuPanel rows and their counts; lu11
// kLuPanel^2 doubles; pairs kLuPairStride ints per inner panel of an outer panel (the row
// interchange: destinations, sources, their count, the pivots found); perm n ints (perm[i] =
// original row now at row i, updated in place); info: first zero pivot (1-based).
constexpr int kLuPanel = 64;
constexpr int kLuOuter = 512;
constexpr int kLuGroupRows = 512;
constexpr int kLuMaxPairs = 2 * kLuPanel;
constexpr int kLuPairStride = 2 * kLuMaxPairs + 2;
// Batches: nblk independent n x n matrices at A + z sA (launch grid z); per block z the
// candidate sets, counts, L11\U11 block, pairs (at z sp), n permutation entries and the info word
// are offset accordingly.
// pivot tournament of panel [j0, j0 + bw); returns the buffer holding the winners in pivot order
const int* launch_tslu(const double* A, int64_t lda, int64_t n, int64_t j0, int bw, int* cand[2], int* cnt[2],
                       double* lu11, int* pairs, int* perm, int64_t* info, hipStream_t st, int nblk = 1,
                       int64_t sA = 0, int64_t sp = 0);
// interchange, L11\U11 and U12 of panel [j0, j0 + bw) on the outer panel's columns [c0, c1)
void launch_calu_apply(double* A, int64_t lda, int64_t n, int64_t c0, int64_t c1, int64_t j0, int bw,
                       const int* pairs, const int* win, const double* lu11, hipStream_t st, int nblk = 1,
                       int64_t sA = 0, int64_t sp = 0);
// L21 = A21 U11^-1 for the rows below panel [j0, j0 + bw)
void launch_calu_l21(double* A, int64_t lda, int64_t n, int64_t j0, int bw, const double* lu11, hipStream_t st,
                     int nblk = 1, int64_t sA = 0);
// the interchanges of nlists inner panels on every column outside [skip0, skip1)
void launch_calu_swap(double* A, int64_t lda, int64_t n, int64_t skip0, int64_t skip1, const int* pairs, int nlists,
                      hipStream_t st, int nblk = 1, int64_t sA = 0, int64_t sp = 0);

void launch_refine_rhs(const double* b, const double* ay, const int* perm, const double* rs, int n, float* x32,
                       hipStream_t st);
void launch_refine_update(const double* y, const float* d, int n, double* x, hipStream_t st);
void launch_permute(const double* b, const int* perm, const double* rs, int n, double* x, hipStream_t st);
void launch_row_scale(double* A, int64_t npad, int nblk, double* rs, unsigned long long* work, hipStream_t st);
void launch_scale(const double* b, const double* rs, int n, double* x, hipStream_t st);
void launch_cross(const NetView& net, const FitView& fv, const int* circles, int nc, hipStream_t st);
// right-hand side as a linear map of the coefficients: basis W (entries [wptr[d], wptr[d + 1]) of
// discharge side d) built once, applied per iteration.  An entry is the real weight of one real
// number of the coefficient state: Re c_j of a source whose coefficients are real (one entry
// per power), Re c_j and Im c_j of the other sources (two entries per power).  A column is
// packed as (source << kRhsColShift) | (2 j + part), part 1 for Im c_j.
constexpr int kRhsMaxK = 512;
constexpr int kRhsColShift = 9;  // 2 j + part < 2 x 256 (coefficient capacity <= kMaxPts / 2)
void launch_rhs_basis(const NetView& net, const TargetView& tg, int nsides, int K, const int64_t* wptr,
                      const int* mcap, double* W, int* side_nan, hipStream_t st);
// the basis rounded to FP32 (the mixed form's FP32 phase)
void launch_rhs_to_f32(int64_t n, const double* W, float* W32, hipStream_t st);
// b from the basis (W32 non-null: its FP32 copy); col / coff: per fracture the packed column of
// every entry and per side the offset of its fracture's columns
void launch_rhs_eval(const NetView& net, const RhsItem* items, int nde, const int64_t* wptr, const double* W,
                     const float* W32, const unsigned* col, const int64_t* coff, const int* side_nan,
                     const int* de_pos, double* b, hipStream_t st);
void launch_rhs_reduce(const NetView& net, const RhsItem* items, int nde, const double* val, const int* pt_off, int K,
                       const int* de_pos, double* b, hipStream_t st);
void launch_assemble(const NetView& net, const TargetView& tg, int nsides, const SysLayout& sys, int K, int nf,
                     double* A, hipStream_t st);
// the same entries into the values of the matrix's CSR pattern (zeroed by the caller)
void launch_assemble_csr(const NetView& net, const TargetView& tg, int nsides, const SysLayout& sys, int K, int nf,
                         const int* ptr, const int* col, double* val, hipStream_t st);
// rs[r] = 1 / max |row r| of a CSR matrix (1 for an empty row)
void launch_csr_row_scale(int64_t n, const int* ptr, const double* val, double* rs, hipStream_t st);
// A[b][i][i] = 1 for the padding unknowns i >= blk_n[b] of every block
void launch_pad_identity(double* A, int64_t npad, int nblk, const int* blk_n, hipStream_t st);
// y_old <- y, y <- x, q and fracture constants <- x (n >= nde + nf unknowns)
void launch_accept(const NetView& net, const int* de, int nde, int nf, const SysLayout& sys, const double* x, double* y,
                   double* yold, int64_t n, hipStream_t st);
void launch_scatter(const NetView& net, const int* de, int nde, int nf, const SysLayout& sys, const double* y,
                    hipStream_t st);
// scratch: reduce_scratch_bytes() of device memory, zeroed once before the first launch (the
// launches leave their counters at zero)
size_t reduce_scratch_bytes();
void launch_reduce(const FitView& fv, int ne, const double* y, const double* yold, int n, void* scratch,
                   IterScalars* out, hipStream_t st);
void launch_init_errors(const FitView& fv, int ne, hipStream_t st);
void launch_residual(const double* ay, const double* b, const double* row_scale, int nde, void* scratch,
                     IterScalars* out, hipStream_t st);
void launch_bc_residual(const NetView& net, const FitItem* items, int nitems, const double* val, const int* pt_off,
                        const RhsItem* witems, int nw, const double* wval, const int* wpt_off, int K, double* out_max,
                        hipStream_t st);


// ---- implicit adjoint (aem_adjoint.hip) and its point tables
void launch_points(const NetView& net, const TargetView& tg, int ntiles, int K, int colloc, double2* pz, int* pelem,
                   hipStream_t st);
// trig: the (cos, sin) tables of launch_trig_tables
void launch_adj_fitT(const NetView& net, const FitItem* items, int nitems, const double2* lam, const int* pt_off,
                     const double2* trig, double2* w, double* sfx, hipStream_t st);
// items[0, nsingle): sources with real or imaginary coefficients (and wells); items[nsingle,
// nitems): sources with complex coefficients (bounding and impermeable circles)
void launch_adj_gather(const NetView& net, const int* items, int nitems, int nsingle, const int* eslot_ptr,
                       const int2* eslot, const int* fp_beg, const int* fp_end, const double2* pz, const int* pelem,
                       const double2* w, int excl, int use_q, int accumulate, double2* gc, double* gq,
                       hipStream_t st);
void launch_adj_cross(int nsrc, const int* srcs, const int* cptr, const int4* cent, const double* sfx, int M,
                      double* gq, hipStream_t st);
void launch_adj_csum(int nf, const int* fp_beg, const int* fp_end, const double2* w, int accumulate, double* gC,
                     hipStream_t st);
void launch_adj_build_z(int nde, int nf, const int* de, const SysLayout& sys, int64_t n, const double* jy,
                        const double* gq, const double* gC, double* z, hipStream_t st);
void launch_adj_rhsT(const NetView& net, const RhsItem* items, int nde, const int* de_pos, const double* mu,
                     const int* pt_off, int K, double2* w, hipStream_t st);
void launch_adj_update(int64_t cnt, const double2* jc, const double2* gc, double2* lam, double damping,
                       unsigned long long* stats, hipStream_t st);
void launch_adj_tgrad(const NetView& net, const FitItem* items, int nitems, const double2* lam, const double* val,
                      const int* pt_off, double* s_out, hipStream_t st);
void launch_adj_pack(int64_t np, const int64_t* pidx, const double* full, double* packed, hipStream_t st);
void launch_adj_unpack(int64_t np, const int64_t* pidx, const double* packed, double* full, hipStream_t st);
void launch_adj_residual_op(int64_t np, const int64_t* pidx, const double* v, const double* gc, double* out,
                            hipStream_t st);
// B_b = A_b^T for the nblk column-major n x n slabs
template <class T>
void launch_transpose(const T* A, int64_t n, T* B, hipStream_t st, int nblk = 1);
template <class T>
void launch_adj_refine_in(const double* z, const double* aty, int n, T* x, hipStream_t st);
template <class T>
void launch_adj_refine_out(const T* v, const int* perm, const double* rs, int n, double* mu, hipStream_t st);

}  // namespace aem
