// Host side of the MI355X analytic-element DFN solve and its C ABI (include/andfn_mi.h).
//
// The call replaces andfn.hpc.hpc_solve.solve (andfn/hpc/hpc_solve.py:55-257): it reads the
// caller's element/fracture records (layouts of andfn/element.py:45-69 and 146-159), runs the
// fixed-point iteration on the GPU and writes the converged state back into the same records.
// The control flow is the reference's: discharge solve, Jacobi sweep with damping, coefficient
// growth with a second damping step, the extra sweep at the iteration cap, and the two-hit
// stop rule on coefficient and discharge changes (hpc_solve.py:152-239).
//
// The discharge system (hpc_solve.py:832-887, 1061-1201) is assembled on the device as a
// dense column-major matrix, factored once with rocSOLVER getrf and solved every iteration
// with getrs (partial pivoting, like SuperLU's splu in the reference).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <functional>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/andfn_mi.h"
#include "aem_kernels.hpp"
#include "aem_sparse.hpp"

namespace {

using aem::FitItem;
using aem::RhsItem;
using aem::Slot;
using aem::Tile;
typedef std::complex<double> cd;

thread_local std::string g_last_error;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

#define HIP_CHECK(x)                                                                     \
  do {                                                                                   \
    hipError_t _e = (x);                                                                 \
    if (_e != hipSuccess) fail(ANDFN_MI_EHIP, "%s failed: %s", #x, hipGetErrorString(_e)); \
  } while (0)

#define RB_CHECK(x)                                                       \
  do {                                                                    \
    rocblas_status _s = (x);                                              \
    if (_s != rocblas_status_success) fail(ANDFN_MI_EHIP, "%s failed: %d", #x, (int)_s); \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    HIP_CHECK(hipMalloc((void**)&p, bytes));
    n = count;
  }
  void upload(const std::vector<T>& v, hipStream_t st) {
    alloc(v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  }
  void zero(size_t count, hipStream_t st) {
    alloc(count);
    HIP_CHECK(hipMemsetAsync(p, 0, std::max<size_t>(count, 1) * sizeof(T), st));
  }
};

// ---- record layouts (andfn/element.py:45-69, 146-159; andfn/constants.py:13-29) ----------
struct Layout {
  int64_t M, ME;
  size_t esz, fsz, e_thetas, e_coef, e_dpsi, e_err, f_elements, f_nel, f_const;
  Layout(int64_t m, int64_t me) : M(m), ME(me) {
    e_thetas = 160;
    e_coef = e_thetas + 16 * M;
    e_dpsi = e_coef + 32 * M;  // coef, old_coef
    e_err = e_dpsi + 16 * M;
    esz = e_err + 32;
    f_elements = 120;
    f_nel = f_elements + 8 * ME;
    f_const = f_nel + 8;
    fsz = f_const + 8;
  }
};

#pragma pack(push, 1)
struct ConstantsRec {
  double RHO, G, SE_FACTOR;
  int32_t MAX_ITERATIONS;
  double MAX_ERROR;
  int32_t MAX_NCOEF;
  double DAMPING;
  int32_t COEF_INCREASE;
  double COEF_RATIO;
  int32_t MAX_ELEMENTS, NCOEF, NINT, NUM_THREADS;
};
#pragma pack(pop)
static_assert(sizeof(ConstantsRec) == 76, "dtype_constants is 76 bytes packed");

template <class T>
inline T rd(const char* p, size_t off) {
  T v;
  std::memcpy(&v, p + off, sizeof(T));
  return v;
}
template <class T>
inline void wr(char* p, size_t off, T v) {
  std::memcpy(p + off, &v, sizeof(T));
}

inline bool is_discharge(int t) { return t == 0 || t == 2 || t == 3; }
inline bool is_circle_like(int t) { return t == 1 || t == 4 || t == 5; }

double2 d2(cd v) { return make_double2(v.real(), v.imag()); }

// algorithmic flops per (point, source) pair: cmul 6, cadd 2, crecip 7, csqrt 12, clog 20
double pair_flops(int kind, int m, bool with_q) {
  switch (kind) {
    case aem::KIND_LINE: return with_q ? 8.0 * m + 61 : 8.0 * m + 41;
    case aem::KIND_IMPL: return 8.0 * m + 41;
    case aem::KIND_BOUND: return 8.0 * m - 6;
    case aem::KIND_WELL: return with_q ? 24.0 : 4.0;
    case aem::KIND_IMPC: return 8.0 * m + 4;
  }
  return 0.0;
}

struct TargetSet {
  std::vector<int> side_elem, side_which, pt_off;
  std::vector<Tile> tiles2;  // 128-point tiles (two points per lane)
  DBuf<Tile> d_tiles2;
  std::vector<int> fsb;  // [F + 1]: sides of fracture f are [fsb[f], fsb[f + 1])
  std::vector<Tile> tiles;
  std::vector<int> frac_nslots;  // sources per tile's fracture, for work accounting
  DBuf<int> d_side_elem, d_side_which, d_pt_off;
  DBuf<Tile> d_tiles;
  // collocation sets: the fit's DFT tiles (sides grouped by angle family, n and m)
  std::vector<aem::DftTile> dft_tiles;
  std::vector<int> dft_sides;
  DBuf<aem::DftTile> d_dft_tiles;
  DBuf<int> d_dft_sides;
  // launch grids: the device copies of the tile lists are laid out for the XCDs (xcd_order) and
  // padded with empty tiles, so they are at least as long as the host lists
  int ndev[2] = {0, 0};
  // the counts the tables were built for (a whole set only): a run that starts from the same
  // counts keeps them
  bool keyed = false;
  std::vector<int> key_nint, key_ncoef;
  int64_t npts() const { return pt_off.empty() ? 0 : pt_off.back(); }
  int ntiles(int ppl) const { return ppl == 2 ? ndev[1] : ndev[0]; }
  aem::TargetView view(int ppl = 1) const {
    aem::TargetView v;
    v.tiles = ppl == 2 ? d_tiles2.p : d_tiles.p;
    v.side_elem = d_side_elem.p;
    v.side_which = d_side_which.p;
    v.pt_off = d_pt_off.p;
    v.pt_z = nullptr;
    return v;
  }
};

class Solver {
 public:
  Solver(void* fr, int64_t nf, int64_t me, void* el, int64_t ne, int64_t M, int64_t K, const void* consts, int dev)
      : L_(M, me), fr_((char*)fr), el_((char*)el), F_((int)nf), E_((int)ne), K_((int)K) {
    if (nf <= 0 || ne <= 0) fail(ANDFN_MI_EINVAL, "empty network (%lld fractures, %lld elements)", (long long)nf, (long long)ne);
    if (M <= 0 || me <= 0) fail(ANDFN_MI_EINVAL, "invalid record capacities");
    if (2 * M > aem::kMaxPts) fail(ANDFN_MI_ECAPACITY, "coefficient capacity %lld exceeds %d", (long long)M, aem::kMaxPts / 2);
    if (consts) std::memcpy(&C_, consts, sizeof(ConstantsRec));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) fail(ANDFN_MI_ENODEV, "no HIP device");
    if (dev < 0 || dev >= ndev) fail(ANDFN_MI_EINVAL, "device %d out of range (%d devices)", dev, ndev);
    HIP_CHECK(hipSetDevice(dev));
    dev_ = dev;
    HIP_CHECK(hipStreamCreate(&st_));
    parse();
    upload();
    RB_CHECK(rocblas_create_handle(&rb_));
    RB_CHECK(rocblas_set_stream(rb_, st_));
    snapshot();
  }

  ~Solver() {
    drop_graph();
    if (rb_) rocblas_destroy_handle(rb_);
    if (rb2_) rocblas_destroy_handle(rb2_);
    for (auto& e : ev_) (void)hipEventDestroy(e);
    for (auto& e : rhs_ev_)
      if (e) (void)hipEventDestroy(e);
    for (auto& p : sol_ev_)
      for (auto& e : p)
        if (e) (void)hipEventDestroy(e);
    if (res_ev_) (void)hipEventDestroy(res_ev_);
    if (h_scal_) (void)hipHostFree(h_scal_);
    drop_iter_graphs();
    for (auto& e : it_end_ev_)
      if (e) (void)hipEventDestroy(e);
    if (h_it_resid_) (void)hipHostFree(h_it_resid_);
    if (h_it_arg_) (void)hipHostFree(h_it_arg_);
    if (h_it_stop_) (void)hipHostFree(h_it_stop_);
    for (auto& e : lu_ev_)
      if (e) (void)hipEventDestroy(e);
    if (st3_) (void)hipStreamDestroy(st3_);
    if (st2_) (void)hipStreamDestroy(st2_);
    if (st_) (void)hipStreamDestroy(st_);
  }

  int solve(const andfn_mi_options* opt, andfn_mi_info* info);
  void write_back();
  // values at points grouped by fracture: what = 0 heads, 1 Omega (two values per point), 2 w in
  // the jitted form, 3 w in the object model's form (two values per point)
  void heads(int64_t npts, const int64_t* frac, const double* z, double* out, int what = 0, int64_t exclude = -1);
  // streamline tracking of n lines; the results stay in the solver until fetched
  void streamlines(int64_t n, const int64_t* frac, const double* z0, const double* elevation, const double* aperture,
                   double ds, int backward, int64_t max_length, int64_t* n_points, int64_t* n_segments,
                   int64_t* exit_element);
  void streamlines_fetch(double* points, double* speed, int64_t* segment_fracture, int64_t* segment_points);
  // heads() with every device buffer in the caller-provided scratch region (no allocation)
  void heads_noalloc(int64_t npts, const int64_t* frac, const double* z, double* out, unsigned char* scratch_dev);
  int device() const { return dev_; }

 private:
  int dev_ = 0;
  Layout L_;
  // coefficients per element held on the device (the stride of every [E][M] array): at least the
  // largest count any element can reach in this network (record counts, 2 M >= nint and the
  // discharge points, MAX_NCOEF + COEF_INCREASE - 1 for growth), rounded up to 32 or 64 or a
  // multiple of 32.  Records are read and written up to Mrec_ = min(M_, L_.M) coefficients; their
  // entries beyond it are left as they are.  (The stride matters to the collocation sweep's
  // coefficient loads: on the 20,000-fracture network, 20 coefficients per element, the sweep
  // takes 926 us per launch at a stride of 20, 924 at 24, 923 at 40, 893 at 32, 892 at 64 and
  // 897 at 200.)
  int64_t M_ = 0, Mrec_ = 0;
  char* fr_;
  char* el_;
  int F_, E_, K_;
  ConstantsRec C_{};
  hipStream_t st_ = nullptr;
  rocblas_handle rb_ = nullptr;
  // LU look-ahead: the trailing update of one outer panel runs on st2_ (default priority, like
  // st_) while st_ factors the next outer panel; created on first use
  hipStream_t st2_ = nullptr;
  rocblas_handle rb2_ = nullptr;
  // st2_ is kept off a few CUs and st3_ has only those CUs (three streams in all: with more,
  // two can share one of the process's hardware queues and its CU mask)
  hipStream_t st3_ = nullptr;
  hipEvent_t lu_ev_[4] = {};
  hipEvent_t rhs_ev_[2] = {};  // start of the right-hand-side phase, alternating by iteration
  hipEvent_t sol_ev_[2][2] = {};  // start / end of the discharge solve, alternating by iteration
  hipEvent_t res_ev_ = nullptr;   // head residual copied to h_scal_
  aem::IterScalars* h_scal_ = nullptr;  // pinned

  // host mirrors (SoA)
  std::vector<int> etype_, efrac0_, efrac1_, ncoef_, nint_;
  std::vector<double2> ep_, center_, coef_;
  std::vector<double> radius_, phi_, q_, t_, constant_, error_, error_old_, error_old2_, error_coef_, dpsi_;
  std::vector<int64_t> frac_id_;
  std::vector<int> mem_ptr_;
  std::vector<Slot> slots_;
  std::vector<int> de_, de_pos_, wells_, circles_, cross_ptr_, cross_cap_;
  int cross_total_ = 0;
  std::vector<int> gcirc_, sched_row_;
  int64_t n_grown_ = 0;

  // device
  DBuf<int> d_etype_, d_efrac0_, d_efrac1_, d_ncoef_, d_nint_, d_mem_ptr_, d_de_, d_de_pos_, d_wells_, d_circles_;
  DBuf<int> d_cross_ptr_, d_cross_cap_, d_cross_cnt_, d_cross_sign_, d_status_, d_grow_, d_grown_circles_;
  DBuf<int2> d_cross_;
  DBuf<double2> d_ep_, d_center_, d_coef_, d_wcoef_;
  DBuf<double> d_radius_, d_phi_, d_q_, d_t_, d_constant_, d_error_, d_error_old_, d_error_old2_, d_error_coef_;
  DBuf<double> d_dpsi_, d_val_, d_valg_, d_valr_, d_dft_, d_A_, d_b_, d_y_, d_yold_, d_x_;
  DBuf<int> d_perm_, d_sched_;
  DBuf<double> d_rowscale_, d_ay_, d_bcmax_, d_csr_val_;
  DBuf<int> d_csr_ptr_, d_csr_col_;
  int64_t nnz_A_ = 0;
  int64_t nnz_factors_ = 0;  // stored entries of the LU factors
  bool csr_ready_ = false;  // the CSR pattern of A (d_csr_ptr_, d_csr_col_) is built
  // discharge-system layout: one dense block per connected component of the network (discharge
  // elements in element order, then fracture constants, then padding), the blocks grouped into
  // size classes of a common padded size (a class's smallest part has at least kClassRatio of
  // its largest part's unknowns), largest class first
  static constexpr double kClassRatio = 0.7;
  // block size rounded up to 32 unknowns: every column of a block (leading dimension npad) then
  // starts on a 128-byte line in FP64 and FP32, so a 64-row segment is whole lines
  static constexpr int64_t kPadAlign = 32;
  struct SysClass {
    int nblk, b0;      // blocks of the class and the global index of its first block
    int64_t npad;      // unknowns per block
    int64_t u0, a0;    // first unknown and first matrix entry of the class
    int64_t inv0;      // first entry of the class's diagonal-block inverses
  };
  std::vector<SysClass> cls_;
  int64_t N_ = 0, Atot_ = 0, invtot_ = 0;
  int nblk_ = 1;  // blocks over all classes
  std::vector<int> fpos_, blk_n_;
  DBuf<int> d_fpos_, d_blk_n_;
  aem::SysLayout sys() const {
    aem::SysLayout l{};
    l.ncls = (int)cls_.size();
    for (int k = 0; k < l.ncls; ++k) {
      l.u0[k] = cls_[k].u0;
      l.a0[k] = cls_[k].a0;
      l.npad[k] = cls_[k].npad;
    }
    l.u0[l.ncls] = N_;
    l.de_pos = d_de_pos_.p;
    l.fpos = d_fpos_.p;
    return l;
  }
  // inverses of the diagonal blocks of L (unit lower) and U, for the blocked triangular solves
  static constexpr int kNB = 1024;
  static int64_t inv_entries(int64_t npad) { return (npad + kNB - 1) / kNB * kNB * kNB; }
  DBuf<double> d_linv_, d_uinv_, d_w_, d_v_, d_req_;
  template <class T>
  void prepare_blocked(T* A, int64_t n, int nblk, T* linv, T* uinv, bool lower_unit, T* ws);
  template <class T>
  void prepare_blocked_all(T* A, DBuf<T>& linv, DBuf<T>& uinv, bool lower_unit = true);
  template <class T>
  void blocked_solve_all(const T* A, const T* linv, const T* uinv, T* x, T* w, T* v);
  template <class T>
  void blocked_solve(const T* A, int n, int nblk, const T* linv, const T* uinv, T* x, T* w, T* v);
  // mixed precision: single-precision factors and blocked-solve buffers
  DBuf<float> d_A32_, d_linv32_, d_uinv32_, d_x32_, d_w32_, d_v32_;
  DBuf<float2> d_coef32_;
  bool mixed_phase_ = false;
  bool solve32_ = false;      // discharge solves by refinement with the FP32-rounded factors
  bool a64_blocked_ = false;  // d_A_ holds the FP64 blocked operators (panel products applied)
  int64_t solves32_ = 0;      // iterations of the current run that used solve32_
  double refine_rate_ = 0.0;  // measured contraction of the FP32-factor refinement
  void switch_to_fp64_solves();
  double refine_probe();
  double far2_ = 16.0;
  // every intersection and constant-head line starts with real coefficients (their fits and the
  // damping keep the imaginary parts at zero), so their series are evaluated as real polynomials
  bool real_lines_ = false;
  int precision_ = ANDFN_MI_FP64;
  // initial state of the mutable arrays, restored at the start of every run
  DBuf<double2> s_coef_;
  DBuf<double> s_q_, s_constant_, s_error_, s_error_old_, s_error_old2_, s_error_coef_, s_dpsi_;
  DBuf<int> s_ncoef_, s_nint_;
  std::vector<int> ncoef0_, nint0_;
  std::vector<double> y0_;
  int lu_mode_ = ANDFN_MI_LU_TRSV;
  int factor_ = ANDFN_MI_FACTOR_SPARSE;
  // multifrontal sparse LU (factor_ == ANDFN_MI_FACTOR_SPARSE): its analysis is built once per
  // handle from the CSR pattern and the fracture centres; a network the analysis cannot take
  // (a pattern that is not the discharge matrix's) is factored densely instead
  aem::MultifrontalLU mf_;
  bool mf_failed_ = false;
  bool sparse_ = false;  // the current factors are the multifrontal ones
  std::string mf_note_;
  std::vector<int> mf_delay_;            // levels each unknown is lifted (pivots found unstable)
  int mf_rounds_ = 0;                    // factorisations of the last run (1 + analyses redone)
  void mf_analyse_now();
  int verbose_ = 0;  // the current run's verbose option (factorisation rounds on stderr)
  std::vector<double> fcenter_;          // [F][3] fracture centres (fracture record offset 24)
  std::vector<double> fradius_;          // [F] fracture radii (fracture record offset 16)
  std::vector<double> fframe_;           // [F][12] centre, x vector, y vector, normal (offsets 24-120)
  // streamline results of the last streamlines() call, per line in input order
  std::vector<std::vector<double2>> sl_pts_;
  std::vector<std::vector<double>> sl_vel_;
  std::vector<std::vector<int>> sl_sfrac_, sl_slen_;
  std::vector<int> h_csr_ptr_, h_csr_col_;  // host copy of the CSR pattern
  void ensure_csr_pattern();
  // blocked LU with tournament pivoting (one connected part): candidate sets, the panel's
  // L11\U11 block and its row interchange
  DBuf<int> lu_cand_[2], lu_cnt_[2], lu_pairs_;
  DBuf<double> lu11_;
  void calu_reserve(int64_t n, int nblk);
  void calu_factor(double* A, int64_t n, int nblk, int* perm, int64_t* info);
  DBuf<Slot> d_slots_;
  DBuf<double2> d_trig_;  // DFT weights (cos, sin) of k pi / (2 n), n <= 2 M (launch_trig_tables)
  DBuf<float2> d_trig32_;
  DBuf<FitItem> d_items_, d_items_g_;
  DBuf<FitItem> d_citems_, d_citems_g_;  // the circle-type items (branch-cut unwrap)
  int ncitems_ = 0, ncitems_g_ = 0;
  int upload_circle_items(const std::vector<FitItem>& items, DBuf<FitItem>& d) {
    std::vector<FitItem> c;
    for (const FitItem& it : items)
      if (is_circle_like(etype_[it.elem])) c.push_back(it);
    d.upload(c, st_);
    return (int)c.size();
  }
  DBuf<RhsItem> d_rhs_items_;
  // right-hand side as a linear map of the coefficients (rhs_basis_kernel), built once per handle
  DBuf<double> d_rbW_;
  DBuf<float> d_rbW32_;     // the basis rounded to FP32 (the mixed form's FP32 phase), made on demand
  bool rb32_ok_ = false;
  DBuf<unsigned> d_rbCol_;  // per fracture: packed (source, coefficient index, part) of every basis column
  DBuf<int64_t> d_rbCoff_;  // per side: first column of its fracture's table
  DBuf<int64_t> d_rbPtr_;
  DBuf<int> d_rbNan_, d_rbMcap_;
  std::vector<int> rb_mcap_;
  bool rb_ok_ = false;
  int64_t rb_entries_ = 0;
  bool ensure_rhs_basis(int max_coef, int coef_increase);
  DBuf<int64_t> d_ipiv_, d_info_;  // 64-bit LAPACK interface: n^2 exceeds 2^31 from ~46k unknowns
  DBuf<aem::IterScalars> d_scal_;
  DBuf<unsigned char> d_red_;  // scratch of the two-stage reductions (counters reset by the kernels)

  TargetSet full_, grown_, rhs_;
  std::vector<FitItem> items_, items_g_;
  std::vector<RhsItem> rhs_items_;
  hipEvent_t ev_[8] = {};

  aem::NetView net() const {
    aem::NetView v;
    v.M = (int)M_;
    v.etype = d_etype_.p;
    v.efrac0 = d_efrac0_.p;
    v.efrac1 = d_efrac1_.p;
    v.ncoef = d_ncoef_.p;
    v.nint = d_nint_.p;
    v.ncoef_w = d_ncoef_.p;
    v.nint_w = d_nint_.p;
    v.ep = d_ep_.p;
    v.radius = d_radius_.p;
    v.center = d_center_.p;
    v.phi = d_phi_.p;
    v.q = d_q_.p;
    v.q_w = d_q_.p;
    v.coef = d_coef_.p;
    v.wcoef = d_wcoef_.p;
    v.t = d_t_.p;
    v.constant = d_constant_.p;
    v.constant_w = d_constant_.p;
    v.mem_ptr = d_mem_ptr_.p;
    v.slots = d_slots_.p;
    v.coef32 = d_coef32_.p;
    v.far2 = far2_;
    v.real_lines = real_lines_ ? 1 : 0;
    v.stop = nullptr;
    return v;
  }
  aem::FitView fit() const {
    aem::FitView v;
    v.error = d_error_.p;
    v.error_old = d_error_old_.p;
    v.error_old2 = d_error_old2_.p;
    v.error_coef = d_error_coef_.p;
    v.dpsi_corr = d_dpsi_.p;
    v.cross_ptr = d_cross_ptr_.p;
    v.cross_cap = d_cross_cap_.p;
    v.cross_cnt = d_cross_cnt_.p;
    v.cross = d_cross_.p;
    v.cross_sign = d_cross_sign_.p;
    v.status = d_status_.p;
    v.stop = nullptr;
    v.trig = d_trig_.p;
    v.trig32 = d_trig32_.p;
    return v;
  }

  void parse();
  void upload();
  void snapshot();
  void restore();
  bool build_targets(TargetSet& ts, const std::vector<char>* only, bool rhs);
  void upload_targets(TargetSet& ts);
  std::vector<Tile> xcd_order(const std::vector<Tile>& tiles, const TargetSet& ts) const;
  void build_items(const TargetSet& ts, std::vector<FitItem>& items, const std::vector<char>* only);
  double set_flops(const TargetSet& ts, bool with_q, int64_t* pairs) const;
  void assemble_and_factor();
  double* discharge_solve(double* b, double* x);
  void solve_rhs();
  void correct_rhs();
  // a discharge solve captured as a HIP graph on its second call (the first runs uncaptured, so
  // any lazy library set-up happens outside the capture)
  struct CapturedSolve {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    bool tried = false;
  };
  CapturedSolve cap_solve_, cap_correct_;
  template <class Fn>
  void run_captured(CapturedSolve& c, Fn&& fn);
  void drop_graph() {
    drop_iter_graphs();
    for (CapturedSolve* c : {&cap_solve_, &cap_correct_}) {
      if (c->e) (void)hipGraphExecDestroy(c->e);
      if (c->g) (void)hipGraphDestroy(c->g);
      *c = CapturedSolve{};
    }
  }
  double head_residual(const double* b, int* row);
  void head_residual_async(const double* b);
  double bc_residual();

  // ---- run-ahead iterations: with the head tolerance as the stop rule and no coefficient growth,
  // the iterations after the first are enqueued one ahead of the host's read of the head
  // residual; the stop is decided on the device (iter_decide_kernel), and the state kernels of an
  // iteration enqueued past the stop return at once.  Each iteration is one HIP graph launch,
  // captured per (FP32 phase, FP32-factor solves, iteration mod 3) on its second use; three
  // copies, so the events of the iteration whose times the host reads are not re-recorded by the
  // one enqueued ahead.
  DBuf<aem::IterCtl> d_ctl_;
  double* h_it_resid_ = nullptr;  // pinned, written by the device: residual of iteration k at [k]
  int* h_it_arg_ = nullptr;
  int* h_it_stop_ = nullptr;      // iteration whose residual stopped the run (0: none)
  aem::IterHostLog it_log_dev_{};  // the same three arrays as the device addresses them
  int it_cap_ = 0;
  struct IterGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    int uses = 0;
    bool failed = false;
  };
  IterGraph it_graphs_[2][2][3];  // [FP32 phase][FP32-factor solves][iteration mod 3]
  // per iteration mod 3: recorded after the iteration's launch, outside any graph (the host waits
  // on it; phase times come from the device's wall-clock stamps, launch_iter_stamp)
  hipEvent_t it_end_ev_[3] = {};
  DBuf<unsigned long long> d_stamps_;
  double wall_khz_ = 0.0;
  int64_t it_graph_launches_ = 0;  // iterations run as graph launches in the last run
  void drop_iter_graphs() {
    for (auto& a : it_graphs_)
      for (auto& b : a)
        for (IterGraph& g : b) {
          if (g.e) (void)hipGraphExecDestroy(g.e);
          if (g.g) (void)hipGraphDestroy(g.g);
          g = IterGraph{};
        }
  }
  void ensure_iter_log(int cap);
  void enqueue_iteration(const aem::NetView& nv, const aem::FitView& fv, bool phase32, bool s32, int par,
                         double head_tol, double damping);
  void launch_iteration(const aem::NetView& nv, const aem::FitView& fv, bool phase32, bool s32, int par,
                        double head_tol, double damping);
  void launch_iteration_body(const aem::NetView& nv, const aem::FitView& fv, bool phase32, bool s32, int par,
                             double head_tol, double damping);

  // tile shape of the collocation and right-hand-side sweeps (points per lane, 1 or 2), picked on
  // the first sweep of each kind by timing both shapes on the same inputs (same results)
  int ppl_[2] = {0, 0};
  // and whether the sweep runs as one persistent launch over the XCD tile lists (the autotuner's
  // other axis); the work queue's nine counters stay zero between launches
  int persist_[2] = {0, 0};
  DBuf<unsigned> d_queue_;
  int persist_groups_[2][2][2] = {};  // [right-hand-side mode][mixed][two points per lane]: resident workgroups
  void launch_sweep_shape(int mode, const aem::NetView& nv, const TargetSet& ts, double* val, bool mixed, int ppl,
                          int persist);
  void sweep(int mode, const TargetSet& ts, double* val, bool mixed);

  // ---- implicit adjoint
  struct AdjTables {
    DBuf<double2> pz;  // point coordinates
    DBuf<int> pelem;   // element owning the point (-1: observation point)
    DBuf<int> fpb, fpe;  // [F] point range of every fracture
  };
  bool have_state_ = false, adjT_valid_ = false;
  bool converged_ = false;  // the last run met its stop rule (head tolerance or AnDFN's two-hit rule)
  DBuf<aem::IterScalars> d_hist_;  // per-iteration scalars for the iteration log
  std::vector<double> head_;  // prescribed head of constant-head elements (phi / t)
  AdjTables tab_full_, tab_rhs_, tab_obs_;
  DBuf<int> d_eslot_ptr_, d_gitems_, d_xsrc_, d_xptr_;
  DBuf<int2> d_eslot_;
  DBuf<int4> d_xent_;
  DBuf<FitItem> d_titems_;
  DBuf<double2> d_lam_, d_jc_, d_gc_, d_wfull_, d_wrhs_, d_wobs_;
  DBuf<double> d_gq_, d_gC_, d_jy0_, d_jy_, d_z_, d_mu_, d_aty_, d_sfx_, d_ts_, d_hw_;
  DBuf<unsigned long long> d_stats_;
  DBuf<double> d_jzero_, d_hv_;
  DBuf<int64_t> d_pidx_;
  DBuf<double> d_gv_, d_xv_, d_wv_;
  DBuf<double> d_V_;
  DBuf<double> d_AT_, d_linvT_, d_uinvT_, d_csrT_val_;
  DBuf<int> d_csrT_ptr_, d_csrT_col_;
  void adj_tables(const TargetSet& ts, bool colloc, AdjTables& t);
  template <class T>
  void make_transposed(const T* A, bool blocked, DBuf<T>& AT, DBuf<T>& linvT, DBuf<T>& uinvT);
  void make_csr_transpose();
  template <class T>
  void adj_mu_step(const T* AT, const T* linvT, const T* uinvT, T* x, T* w, T* v);
  void adj_mu_step_sparse();

 public:
  int adjoint(const andfn_mi_objective* obj, const andfn_mi_adjoint_options* opt, double* grad_t, double* grad_head,
              andfn_mi_adjoint_info* info);
  void set_parameters(const double* t, const double* head);
  void state_discharges(double* q, double* constant);
};

void Solver::parse() {
  etype_.resize(E_);
  efrac0_.resize(E_);
  efrac1_.resize(E_);
  ncoef_.resize(E_);
  nint_.resize(E_);
  ep_.resize(4 * (size_t)E_);
  center_.resize(E_);
  radius_.resize(E_);
  phi_.resize(E_);
  q_.resize(E_);
  error_.resize(E_);
  error_old_.resize(E_);
  error_old2_.resize(E_);
  error_coef_.resize(E_);
  // the device's coefficient stride M_ (see its declaration)
  {
    int64_t mx = 1;
    for (int i = 0; i < E_; ++i) {
      const char* p = el_ + (size_t)i * L_.esz;
      mx = std::max(mx, rd<int64_t>(p, 136));
      mx = std::max(mx, (rd<int64_t>(p, 144) + 1) / 2);
    }
    mx = std::max<int64_t>(mx, (int64_t)C_.MAX_NCOEF + C_.COEF_INCREASE - 1);
    mx = std::max<int64_t>(mx, (K_ + 1) / 2);
    mx = std::min<int64_t>(std::max<int64_t>(mx, 1), L_.M);
    M_ = mx <= 32 ? 32 : mx <= 64 ? 64 : (mx + 31) / 32 * 32;
    Mrec_ = std::min<int64_t>(M_, L_.M);
  }
  coef_.assign((size_t)E_ * M_, make_double2(0, 0));
  dpsi_.assign((size_t)E_ * 2 * M_, 0.0);
  for (int i = 0; i < E_; ++i) {
    const char* p = el_ + (size_t)i * L_.esz;
    int64_t id = rd<int64_t>(p, 0), ty = rd<int64_t>(p, 8), f0 = rd<int64_t>(p, 16), f1 = rd<int64_t>(p, 24);
    int64_t nc = rd<int64_t>(p, 136), ni = rd<int64_t>(p, 144);
    if (id != i) fail(ANDFN_MI_EINVAL, "element %d has _id %lld (ids must equal positions)", i, (long long)id);
    if (ty < 0 || ty > 5) fail(ANDFN_MI_EINVAL, "element %d has invalid type %lld", i, (long long)ty);
    if (f0 < 0 || f0 >= F_ || (ty == 0 && (f1 < 0 || f1 >= F_)))
      fail(ANDFN_MI_EINVAL, "element %d references a fracture out of range", i);
    if (ty != 2 && (nc < 1 || nc > L_.M)) fail(ANDFN_MI_ECAPACITY, "element %d: ncoef %lld outside [1, %lld]", i, (long long)nc, (long long)L_.M);
    if (ty != 2 && (ni < 2 || ni > 2 * L_.M)) fail(ANDFN_MI_ECAPACITY, "element %d: nint %lld outside [2, %lld]", i, (long long)ni, (long long)(2 * L_.M));
    etype_[i] = (int)ty;
    efrac0_[i] = (int)f0;
    efrac1_[i] = (int)f1;
    ncoef_[i] = (int)std::min<int64_t>(std::max<int64_t>(nc, 0), L_.M);
    nint_[i] = (int)std::min<int64_t>(std::max<int64_t>(ni, 0), 2 * L_.M);
    for (int k = 0; k < 4; ++k) ep_[4 * i + k] = d2(rd<cd>(p, 32 + 16 * k));
    radius_[i] = rd<double>(p, 96);
    center_[i] = d2(rd<cd>(p, 104));
    phi_[i] = rd<double>(p, 128);
    q_[i] = rd<double>(p, 152);
    for (int64_t j = 0; j < Mrec_; ++j) coef_[(size_t)i * M_ + j] = d2(rd<cd>(p, L_.e_coef + 16 * j));
    for (int64_t j = 0; j < 2 * Mrec_; ++j) dpsi_[(size_t)i * 2 * M_ + j] = rd<double>(p, L_.e_dpsi + 8 * j);
    error_[i] = rd<double>(p, L_.e_err);
    error_old_[i] = rd<double>(p, L_.e_err + 8);
    error_old2_[i] = rd<double>(p, L_.e_err + 16);
    error_coef_[i] = rd<double>(p, L_.e_err + 24);
  }
  // ANDFN_MI_COMPLEX_LINES set: evaluate every line series with complex coefficients (tests
  // compare the two)
  real_lines_ = std::getenv("ANDFN_MI_COMPLEX_LINES") == nullptr;
  for (int i = 0; i < E_ && real_lines_; ++i)
    if (etype_[i] == 0 || etype_[i] == 3)
      for (int64_t j = 0; j < Mrec_; ++j)
        if (coef_[(size_t)i * M_ + j].y != 0.0) {
          real_lines_ = false;
          break;
        }
  t_.resize(F_);
  constant_.resize(F_);
  frac_id_.resize(F_);
  fcenter_.resize(3 * (size_t)F_);
  fradius_.resize(F_);
  fframe_.resize(12 * (size_t)F_);
  mem_ptr_.assign(F_ + 1, 0);
  slots_.clear();
  for (int f = 0; f < F_; ++f) {
    const char* p = fr_ + (size_t)f * L_.fsz;
    frac_id_[f] = rd<int64_t>(p, 0);
    t_[f] = rd<double>(p, 8);
    for (int a = 0; a < 3; ++a) fcenter_[3 * (size_t)f + a] = rd<double>(p, 24 + 8 * a);
    fradius_[f] = rd<double>(p, 16);
    // centre 24, normal 48, x vector 72, y vector 96 -> frame order centre, x, y, normal
    for (int a = 0; a < 3; ++a) {
      fframe_[12 * (size_t)f + a] = rd<double>(p, 24 + 8 * a);
      fframe_[12 * (size_t)f + 3 + a] = rd<double>(p, 72 + 8 * a);
      fframe_[12 * (size_t)f + 6 + a] = rd<double>(p, 96 + 8 * a);
      fframe_[12 * (size_t)f + 9 + a] = rd<double>(p, 48 + 8 * a);
    }
    constant_[f] = rd<double>(p, L_.f_const);
    int64_t nel = rd<int64_t>(p, L_.f_nel);
    if (nel < 0 || nel > L_.ME) fail(ANDFN_MI_ECAPACITY, "fracture %d: %lld elements exceed capacity %lld", f, (long long)nel, (long long)L_.ME);
    for (int64_t k = 0; k < nel; ++k) {
      int64_t e = rd<int64_t>(p, L_.f_elements + 8 * k);
      if (e < 0 || e >= E_) fail(ANDFN_MI_EINVAL, "fracture %d lists element %lld out of range", f, (long long)e);
      Slot s{};
      s.elem = (int)e;
      s.sign = 1.0;
      int ty = etype_[e];
      if (ty == 0 || ty == 3 || ty == 5) {
        bool second = (ty == 0) && (efrac0_[e] != frac_id_[f]);
        s.kind = (ty == 5) ? aem::KIND_IMPL : aem::KIND_LINE;
        if (second) s.sign = -1.0;
        cd a(ep_[4 * e + (second ? 2 : 0)].x, ep_[4 * e + (second ? 2 : 0)].y);
        cd b(ep_[4 * e + (second ? 3 : 1)].x, ep_[4 * e + (second ? 3 : 1)].y);
        s.alpha = d2(2.0 / (b - a));
        s.beta = d2(-(a + b) / (b - a));
      } else if (ty == 1) {
        s.kind = aem::KIND_BOUND;
        s.alpha = make_double2(1.0 / radius_[e], 0.0);
        s.beta = make_double2(0.0, 0.0);
      } else {
        s.kind = (ty == 2) ? aem::KIND_WELL : aem::KIND_IMPC;
        double r = radius_[e];
        s.alpha = make_double2(1.0 / r, 0.0);
        s.beta = make_double2(-center_[e].x / r, -center_[e].y / r);
      }
      slots_.push_back(s);
    }
    mem_ptr_[f + 1] = (int)slots_.size();
  }
  head_.assign(E_, 0.0);
  for (int e = 0; e < E_; ++e)
    if (etype_[e] == 2 || etype_[e] == 3) head_[e] = phi_[e] / t_[efrac0_[e]];
  de_.clear();
  de_pos_.assign(E_, -1);
  wells_.clear();
  circles_.clear();
  for (int e = 0; e < E_; ++e) {
    if (is_discharge(etype_[e])) {
      de_pos_[e] = (int)de_.size();
      de_.push_back(e);
    }
    if (etype_[e] == 2) wells_.push_back(e);
    if (is_circle_like(etype_[e])) circles_.push_back(e);
  }
  // connected components of the fractures through intersections: independent discharge blocks
  {
    std::vector<int> par(F_);
    for (int f = 0; f < F_; ++f) par[f] = f;
    auto root = [&](int a) {
      while (par[a] != a) a = par[a] = par[par[a]];
      return a;
    };
    for (int e = 0; e < E_; ++e)
      if (etype_[e] == 0) {
        const int a = root(efrac0_[e]), b = root(efrac1_[e]);
        if (a != b) par[std::max(a, b)] = std::min(a, b);
      }
    std::vector<int> bid_of_root(F_, -1), bid(F_);
    nblk_ = 0;
    for (int f = 0; f < F_; ++f) {
      const int r = root(f);
      if (bid_of_root[r] < 0) bid_of_root[r] = nblk_++;
      bid[f] = bid_of_root[r];
    }
    std::vector<int> cnt(nblk_, 0), local_e(E_, -1), local_f(F_);
    for (int e : de_) local_e[e] = cnt[bid[efrac0_[e]]]++;
    for (int f = 0; f < F_; ++f) local_f[f] = cnt[bid[f]]++;
    // blocks by size, largest first (ties in part order); a new class starts below kClassRatio
    // of the current class's size while classes remain
    std::vector<int> order(nblk_);
    for (int b = 0; b < nblk_; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
    std::vector<int> slot(nblk_);  // block index in class order
    cls_.clear();
    blk_n_.assign(nblk_, 0);
    for (int i = 0; i < nblk_; ++i) {
      const int b = order[i];
      if (cls_.empty() || ((double)cnt[b] < kClassRatio * (double)cls_.back().npad &&
                           (int)cls_.size() < aem::kMaxSysClasses))
        cls_.push_back(SysClass{0, i, (cnt[b] + kPadAlign - 1) / kPadAlign * kPadAlign, 0, 0, 0});
      cls_.back().nblk++;
      slot[b] = i;
      blk_n_[i] = cnt[b];
    }
    N_ = Atot_ = invtot_ = 0;
    for (auto& c : cls_) {
      c.u0 = N_;
      c.a0 = Atot_;
      c.inv0 = invtot_;
      N_ += c.nblk * c.npad;
      Atot_ += c.nblk * c.npad * c.npad;
      invtot_ += c.nblk * inv_entries(c.npad);
    }
    if (N_ > INT32_MAX) fail(ANDFN_MI_ECAPACITY, "discharge system of %lld padded unknowns", (long long)N_);
    // first unknown of every block
    std::vector<int64_t> base(nblk_);
    for (const auto& c : cls_)
      for (int k = 0; k < c.nblk; ++k) base[c.b0 + k] = c.u0 + k * c.npad;
    fpos_.assign(F_, 0);
    for (int e : de_) de_pos_[e] = (int)(base[slot[bid[efrac0_[e]]]] + local_e[e]);
    for (int f = 0; f < F_; ++f) fpos_[f] = (int)(base[slot[bid[f]]] + local_f[f]);
  }
  cross_ptr_.assign(E_, 0);
  cross_cap_.assign(E_, 0);
  int tot = 0;
  for (int e : circles_) {
    int f = efrac0_[e];
    int cap = 4 * (mem_ptr_[f + 1] - mem_ptr_[f]) + 16;
    cross_ptr_[e] = tot;
    cross_cap_[e] = cap;
    tot += cap;
  }
  cross_total_ = tot;
}

}  // namespace


namespace {

void Solver::upload() {
  d_etype_.upload(etype_, st_);
  d_efrac0_.upload(efrac0_, st_);
  d_efrac1_.upload(efrac1_, st_);
  d_ncoef_.upload(ncoef_, st_);
  d_nint_.upload(nint_, st_);
  d_ep_.upload(ep_, st_);
  d_center_.upload(center_, st_);
  d_radius_.upload(radius_, st_);
  d_phi_.upload(phi_, st_);
  d_q_.upload(q_, st_);
  d_t_.upload(t_, st_);
  d_constant_.upload(constant_, st_);
  d_coef_.upload(coef_, st_);
  d_wcoef_.zero(coef_.size(), st_);
  d_dpsi_.upload(dpsi_, st_);
  d_error_.upload(error_, st_);
  d_error_old_.upload(error_old_, st_);
  d_error_old2_.upload(error_old2_, st_);
  d_error_coef_.upload(error_coef_, st_);
  d_mem_ptr_.upload(mem_ptr_, st_);
  d_slots_.upload(slots_, st_);
  d_de_.upload(de_, st_);
  d_de_pos_.upload(de_pos_, st_);
  d_fpos_.upload(fpos_, st_);
  d_blk_n_.upload(blk_n_, st_);
  d_wells_.upload(wells_, st_);
  d_circles_.upload(circles_, st_);
  d_cross_ptr_.upload(cross_ptr_, st_);
  d_cross_cap_.upload(cross_cap_, st_);
  d_cross_cnt_.zero(E_, st_);
  d_cross_.alloc(std::max(cross_total_, 1));
  d_cross_sign_.alloc(std::max(cross_total_, 1));
  d_status_.zero(1, st_);
  d_grow_.zero(E_, st_);
  d_scal_.alloc(1);
  d_red_.zero(aem::reduce_scratch_bytes(), st_);
  {
    const int nmax = (int)(2 * M_);  // nint <= 2 M_ (parse)
    d_trig_.alloc((size_t)aem::trig_entries(nmax));
    d_trig32_.alloc((size_t)aem::trig_entries(nmax));
    aem::launch_trig_tables(nmax, d_trig_.p, d_trig32_.p, st_);
    HIP_CHECK(hipGetLastError());
  }
  for (auto& e : ev_) HIP_CHECK(hipEventCreate(&e));
  for (auto& e : rhs_ev_) HIP_CHECK(hipEventCreate(&e));
  for (auto& p : sol_ev_)
    for (auto& e : p) HIP_CHECK(hipEventCreate(&e));
  HIP_CHECK(hipEventCreateWithFlags(&res_ev_, hipEventDisableTiming));
  HIP_CHECK(hipHostMalloc((void**)&h_scal_, sizeof(aem::IterScalars), hipHostMallocDefault));
}

template <class T>
static void dcopy(DBuf<T>& dst, const DBuf<T>& src, size_t n, hipStream_t st) {
  dst.alloc(n);
  HIP_CHECK(hipMemcpyAsync(dst.p, src.p, sizeof(T) * std::max<size_t>(n, 1), hipMemcpyDeviceToDevice, st));
}

void Solver::snapshot() {
  dcopy(s_coef_, d_coef_, coef_.size(), st_);
  dcopy(s_q_, d_q_, E_, st_);
  dcopy(s_constant_, d_constant_, F_, st_);
  dcopy(s_error_, d_error_, E_, st_);
  dcopy(s_error_old_, d_error_old_, E_, st_);
  dcopy(s_error_old2_, d_error_old2_, E_, st_);
  dcopy(s_error_coef_, d_error_coef_, E_, st_);
  dcopy(s_dpsi_, d_dpsi_, dpsi_.size(), st_);
  dcopy(s_ncoef_, d_ncoef_, E_, st_);
  dcopy(s_nint_, d_nint_, E_, st_);
  ncoef0_ = ncoef_;
  nint0_ = nint_;
  y0_.assign(N_, 0.0);
  for (int e : de_) y0_[de_pos_[e]] = q_[e];
  for (int f = 0; f < F_; ++f) y0_[fpos_[f]] = constant_[f];
  HIP_CHECK(hipStreamSynchronize(st_));
}

void Solver::restore() {
  dcopy(d_coef_, s_coef_, coef_.size(), st_);
  dcopy(d_q_, s_q_, E_, st_);
  dcopy(d_constant_, s_constant_, F_, st_);
  dcopy(d_error_, s_error_, E_, st_);
  dcopy(d_error_old_, s_error_old_, E_, st_);
  dcopy(d_error_old2_, s_error_old2_, E_, st_);
  dcopy(d_error_coef_, s_error_coef_, E_, st_);
  dcopy(d_dpsi_, s_dpsi_, dpsi_.size(), st_);
  dcopy(d_ncoef_, s_ncoef_, E_, st_);
  dcopy(d_nint_, s_nint_, E_, st_);
  HIP_CHECK(hipMemsetAsync(d_wcoef_.p, 0, sizeof(double2) * coef_.size(), st_));
  HIP_CHECK(hipMemsetAsync(d_cross_cnt_.p, 0, sizeof(int) * E_, st_));
  HIP_CHECK(hipMemsetAsync(d_status_.p, 0, sizeof(int), st_));
  ncoef_ = ncoef0_;
  nint_ = nint0_;
  n_grown_ = 0;
}

// Pack the target points of every fracture contiguously and cut them into 64-point tiles.
// Collocation set: every non-well element side (an intersection has one side per fracture);
// discharge set (rhs): every discharge element side with K points each.
// Returns false when the set's tables (host and device) were kept: a whole set (only == nullptr)
// already built for the current counts (the discharge-point set depends on the structure alone).
bool Solver::build_targets(TargetSet& ts, const std::vector<char>* only, bool rhs) {
  if (!only && ts.keyed && (rhs || (ts.key_nint == nint_ && ts.key_ncoef == ncoef_))) return false;
  ts.keyed = false;
  ts.side_elem.clear();
  ts.side_which.clear();
  ts.pt_off.assign(1, 0);
  ts.tiles.clear();
  ts.tiles2.clear();
  ts.frac_nslots.clear();
  ts.fsb.assign(F_ + 1, 0);
  for (int f = 0; f < F_; ++f) {
    const int sbeg = (int)ts.side_elem.size();
    ts.fsb[f] = sbeg;
    const int pbeg = ts.pt_off.back();
    for (int k = mem_ptr_[f]; k < mem_ptr_[f + 1]; ++k) {
      const int e = slots_[k].elem;
      const int ty = etype_[e];
      if (rhs ? !is_discharge(ty) : ty == 2) continue;
      if (only && !(*only)[e]) continue;
      ts.side_elem.push_back(e);
      ts.side_which.push_back((ty == 0 && efrac0_[e] != frac_id_[f]) ? 1 : 0);
      ts.pt_off.push_back(ts.pt_off.back() + (rhs ? K_ : nint_[e]));
    }
    const int send = (int)ts.side_elem.size();
    const int pend = ts.pt_off.back();
    for (int p0 = pbeg; p0 < pend; p0 += 64) {
      ts.tiles.push_back(Tile{f, sbeg, send, p0, pend});
      ts.frac_nslots.push_back(mem_ptr_[f + 1] - mem_ptr_[f]);
    }
    for (int p0 = pbeg; p0 < pend; p0 += 128) ts.tiles2.push_back(Tile{f, sbeg, send, p0, pend});
  }
  ts.fsb[F_] = (int)ts.side_elem.size();
  // heaviest fractures first (work of a tile ~ number of sources of its fracture)
  std::vector<int> order(ts.tiles.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ts.frac_nslots[a] > ts.frac_nslots[b]; });
  std::vector<Tile> t2(ts.tiles.size());
  std::vector<int> n2(ts.tiles.size());
  for (size_t i = 0; i < order.size(); ++i) {
    t2[i] = ts.tiles[order[i]];
    n2[i] = ts.frac_nslots[order[i]];
  }
  ts.tiles.swap(t2);
  ts.frac_nslots.swap(n2);
  std::stable_sort(ts.tiles2.begin(), ts.tiles2.end(), [&](const Tile& a, const Tile& b) {
    return mem_ptr_[a.frac + 1] - mem_ptr_[a.frac] > mem_ptr_[b.frac + 1] - mem_ptr_[b.frac];
  });
  ts.dft_tiles.clear();
  ts.dft_sides.clear();
  if (!rhs) {
    // sides grouped by (angle family, n, m), cut into tiles of 16
    const int ns = (int)ts.side_elem.size();
    auto key = [&](int sd) {
      const int e = ts.side_elem[sd];
      return std::make_tuple(etype_[e] == 1 ? 1 : 0, nint_[e], ncoef_[e]);
    };
    ts.dft_sides.resize(ns);
    for (int i = 0; i < ns; ++i) ts.dft_sides[i] = i;
    std::stable_sort(ts.dft_sides.begin(), ts.dft_sides.end(), [&](int a, int b) { return key(a) < key(b); });
    for (int i = 0; i < ns;) {
      int j = i;
      while (j < ns && j - i < 16 && key(ts.dft_sides[j]) == key(ts.dft_sides[i])) ++j;
      const int e = ts.side_elem[ts.dft_sides[i]];
      ts.dft_tiles.push_back(aem::DftTile{i, j - i, nint_[e], ncoef_[e], etype_[e] == 1 ? 1 : 0});
      i = j;
    }
  }
  upload_targets(ts);
  if (!only) {
    ts.keyed = true;
    ts.key_nint = nint_;
    ts.key_ncoef = ncoef_;
  }
  return true;
}

// Device order of a tile list.  Workgroups go to the 8 XCDs of the MI355X round-robin
// (workgroup b to XCD b % 8), and each XCD has its own L2.  Every fracture's tiles are given to
// one XCD (fractures by decreasing work, each to the least loaded XCD), each XCD's tiles keep the
// heaviest-first order, and the device list interleaves the 8 lists so that workgroup b takes
// tile b / 8 of XCD b % 8's list: the tiles that read one fracture's sources share an L2.  Lists
// are padded with empty tiles (p0 == pend), which every tile kernel skips.
std::vector<Tile> Solver::xcd_order(const std::vector<Tile>& tiles, const TargetSet& ts) const {
  constexpr int kXcd = 8;
  std::vector<double> fw(F_, 0.0);
  for (const Tile& t : tiles) fw[t.frac] += (double)std::min(t.pend - t.p0, 128) * (mem_ptr_[t.frac + 1] - mem_ptr_[t.frac]);
  std::vector<int> fr;
  for (int f = 0; f < F_; ++f)
    if (fw[f] > 0.0) fr.push_back(f);
  std::stable_sort(fr.begin(), fr.end(), [&](int a, int b) { return fw[a] > fw[b]; });
  std::vector<int> xcd(F_, 0);
  double load[kXcd] = {0.0};
  for (int f : fr) {
    const int x = (int)(std::min_element(load, load + kXcd) - load);
    xcd[f] = x;
    load[x] += fw[f];
  }
  std::vector<std::vector<Tile>> lists(kXcd);
  for (const Tile& t : tiles) lists[xcd[t.frac]].push_back(t);  // (keeps the heaviest-first order)
  size_t len = 0;
  for (const auto& l : lists) len = std::max(len, l.size());
  std::vector<Tile> out(len * kXcd, Tile{0, 0, 0, 0, 0});
  for (int x = 0; x < kXcd; ++x)
    for (size_t i = 0; i < lists[x].size(); ++i) out[i * kXcd + x] = lists[x][i];
  (void)ts;
  return out;
}

void Solver::upload_targets(TargetSet& ts) {
  ts.d_side_elem.upload(ts.side_elem, st_);
  ts.d_side_which.upload(ts.side_which, st_);
  ts.d_pt_off.upload(ts.pt_off, st_);
  {
    const std::vector<Tile> t1 = xcd_order(ts.tiles, ts), t2 = xcd_order(ts.tiles2, ts);
    ts.d_tiles.upload(t1, st_);
    ts.d_tiles2.upload(t2, st_);
    ts.ndev[0] = (int)t1.size();
    ts.ndev[1] = (int)t2.size();
  }
  ts.d_dft_tiles.upload(ts.dft_tiles, st_);
  ts.d_dft_sides.upload(ts.dft_sides, st_);
}

void Solver::build_items(const TargetSet& ts, std::vector<FitItem>& items, const std::vector<char>* only) {
  std::vector<int> s0(E_, -1), s1(E_, -1);
  for (size_t s = 0; s < ts.side_elem.size(); ++s) {
    if (ts.side_which[s]) s1[ts.side_elem[s]] = (int)s;
    else s0[ts.side_elem[s]] = (int)s;
  }
  items.clear();
  for (int e = 0; e < E_; ++e) {
    if (etype_[e] == 2) continue;
    if (only && !(*only)[e]) continue;
    if (s0[e] < 0 || (etype_[e] == 0 && s1[e] < 0))
      fail(ANDFN_MI_EINVAL, "element %d is not listed by the fracture(s) it belongs to", e);
    items.push_back(FitItem{e, s0[e], s1[e]});
  }
}

double Solver::set_flops(const TargetSet& ts, bool with_q, int64_t* pairs) const {
  // per fracture: points x sum over sources of the per-pair flops at the current ncoef
  double fl = 0.0;
  int64_t pr = 0;
  std::vector<char> seen(F_, 0);
  for (const Tile& t : ts.tiles) {
    if (seen[t.frac]) continue;
    seen[t.frac] = 1;
    const int64_t P = t.pend - ts.pt_off[t.sbeg];
    double per = 0.0;
    for (int k = mem_ptr_[t.frac]; k < mem_ptr_[t.frac + 1]; ++k)
      per += pair_flops(slots_[k].kind, ncoef_[slots_[k].elem], with_q);
    fl += per * (double)P;
    pr += P * (mem_ptr_[t.frac + 1] - mem_ptr_[t.frac]);
  }
  if (pairs) *pairs = pr;
  return fl;
}

// rocBLAS triangular inverse / multiply by element type
static rocblas_status rb_trtri_batched(rocblas_handle h, rocblas_fill f, rocblas_diagonal d, int nbk, const double* A,
                                       int64_t lda, int64_t sa, double* inv, int ldi, int64_t si, int cnt) {
  return rocblas_dtrtri_strided_batched(h, f, d, nbk, A, (rocblas_int)lda, sa, inv, ldi, si, cnt);
}
static rocblas_status rb_trtri_batched(rocblas_handle h, rocblas_fill f, rocblas_diagonal d, int nbk, const float* A,
                                       int64_t lda, int64_t sa, float* inv, int ldi, int64_t si, int cnt) {
  return rocblas_strtri_strided_batched(h, f, d, nbk, A, (rocblas_int)lda, sa, inv, ldi, si, cnt);
}
static rocblas_status rb_trtri(rocblas_handle h, rocblas_fill f, rocblas_diagonal d, int m, const double* A,
                               int64_t lda, double* inv, int ldi) {
  return rocblas_dtrtri(h, f, d, m, A, (rocblas_int)lda, inv, ldi);
}
static rocblas_status rb_trtri(rocblas_handle h, rocblas_fill f, rocblas_diagonal d, int m, const float* A,
                               int64_t lda, float* inv, int ldi) {
  return rocblas_strtri(h, f, d, m, A, (rocblas_int)lda, inv, ldi);
}
static rocblas_status rb_trmm_right(rocblas_handle h, rocblas_fill f, rocblas_diagonal d, int m, int nc,
                                    const double* T, int ldt, double* B, int64_t ldb) {
  const double one = 1.0;
  return rocblas_dtrmm_64(h, rocblas_side_right, f, rocblas_operation_none, d, m, nc, &one, T, ldt, B, ldb, B, ldb);
}

// C_z = A_z B_z (m x nc, inner dimension k) on nb blocks, column-major, strided
static rocblas_status rb_gemm_nn(rocblas_handle h, int64_t m, int64_t nc, int64_t k, const double* A, int64_t lda,
                                 int64_t sA, const double* B, int64_t ldb, int64_t sB, double* C, int64_t ldc,
                                 int64_t sC, int nb) {
  const double one = 1.0, zero = 0.0;
  return rocblas_dgemm_strided_batched_64(h, rocblas_operation_none, rocblas_operation_none, m, nc, k, &one, A, lda, sA,
                                          B, ldb, sB, &zero, C, ldc, sC, nb);
}
static rocblas_status rb_gemm_nn(rocblas_handle h, int64_t m, int64_t nc, int64_t k, const float* A, int64_t lda,
                                 int64_t sA, const float* B, int64_t ldb, int64_t sB, float* C, int64_t ldc, int64_t sC,
                                 int nb) {
  const float one = 1.0f, zero = 0.0f;
  return rocblas_sgemm_strided_batched_64(h, rocblas_operation_none, rocblas_operation_none, m, nc, k, &one, A, lda, sA,
                                          B, ldb, sB, &zero, C, ldc, sC, nb);
}

// the same product on nb blocks: T + z sT, B + z sB
static rocblas_status rb_trmm_right_batched(rocblas_handle h, rocblas_fill f, rocblas_diagonal d, int m, int nc,
                                            const double* T, int ldt, int64_t sT, double* B, int64_t ldb, int64_t sB,
                                            int nb) {
  const double one = 1.0;
  return rocblas_dtrmm_strided_batched_64(h, rocblas_side_right, f, rocblas_operation_none, d, m, nc, &one, T, ldt, sT,
                                          B, ldb, sB, B, ldb, sB, nb);
}

// Blocked triangular-solve operators from the LU factors in A (n x n, column-major, kNB blocks):
// Linv_j / Uinv_j = inverses of the diagonal blocks of L (unit) and U, and, in place in A, the
// panels G_j = L[below j, j] Linv_j and H_j = U[above j, j] Uinv_j.  With them, L z = c becomes
// "c[below j] -= G_j c_j for j ascending, then z_j = Linv_j c_j for all j at once" (and U likewise,
// descending): the sequential chain is one gemv per block and the diagonal blocks are applied by
// a single batched launch.
// lower_unit: L has the unit diagonal (LU factors); false for the transposed factors U^T L^T,
// whose lower triangle carries the diagonal and whose upper triangle is unit.
// For the nblk blocks of one size class (n unknowns each) at A; linv / uinv hold the class's
// inverse blocks (zeroed by the caller).
template <class T>
void Solver::prepare_blocked(T* A, int64_t n, int nblk, T* linv, T* uinv, bool lower_unit, T* ws) {
  const int64_t nb = (n + kNB - 1) / kNB, nfull = n / kNB;
  const rocblas_diagonal dl = lower_unit ? rocblas_diagonal_unit : rocblas_diagonal_non_unit;
  const rocblas_diagonal du = lower_unit ? rocblas_diagonal_non_unit : rocblas_diagonal_unit;
  const int64_t sinv = nb * kNB * kNB;  // inverse blocks of one system
  if (nblk > 1) {
    // several systems of the same (padded) size: per diagonal block j, one call covers them all
    const int64_t sA = n * n;
    for (int64_t j = 0; j < nb; ++j) {
      const int64_t r0 = j * kNB;
      const int bs = (int)std::min<int64_t>(kNB, n - r0);
      RB_CHECK(rb_trtri_batched(rb_, rocblas_fill_lower, dl, bs, A + r0 * (n + 1), n, sA, linv + j * kNB * kNB, kNB,
                                sinv, nblk));
      RB_CHECK(rb_trtri_batched(rb_, rocblas_fill_upper, du, bs, A + r0 * (n + 1), n, sA, uinv + j * kNB * kNB, kNB,
                                sinv, nblk));
    }
  } else {
    for (int b = 0; b < nblk; ++b) {
      T* Ab = A + (int64_t)b * n * n;
      T* li = linv + b * sinv;
      T* ui = uinv + b * sinv;
      if (nfull > 0) {
        RB_CHECK(rb_trtri_batched(rb_, rocblas_fill_lower, dl, kNB, Ab, n, kNB * (n + 1), li, kNB, (int64_t)kNB * kNB,
                                  (int)nfull));
        RB_CHECK(rb_trtri_batched(rb_, rocblas_fill_upper, du, kNB, Ab, n, kNB * (n + 1), ui, kNB, (int64_t)kNB * kNB,
                                  (int)nfull));
      }
      if (nb > nfull) {
        const int64_t r0 = nfull * kNB;
        const int rem = (int)(n - r0);
        const T* blk = Ab + r0 * (n + 1);
        RB_CHECK(rb_trtri(rb_, rocblas_fill_lower, dl, rem, blk, n, li + nfull * kNB * kNB, kNB));
        RB_CHECK(rb_trtri(rb_, rocblas_fill_upper, du, rem, blk, n, ui + nfull * kNB * kNB, kNB));
      }
    }
  }
  // the inverses as dense blocks: the opposite triangles zero, and a unit-triangular inverse
  // (written without its diagonal) given its ones
  aem::launch_tri_zero<T>(linv, (int)(nb * nblk), kNB, 1, st_);
  aem::launch_tri_zero<T>(uinv, (int)(nb * nblk), kNB, 0, st_);
  aem::launch_unit_diag<T>(lower_unit ? linv : uinv, (int)(nb * nblk), kNB, st_);
  // panels: G_j = L[below j, j] Linv_j and H_j = U[above j, j] Uinv_j as dense products (MFMA
  // GEMM) into the workspace ws (at least (n - 1) kNB nblk entries), copied back in place
  const int64_t sA = n * n;
  for (int64_t j = 0; j < nb; ++j) {
    const int64_t r0 = j * kNB;
    const int bs = (int)std::min<int64_t>(kNB, n - r0);
    const int64_t rest = n - r0 - bs;
    if (rest > 0) {
      T* P = A + r0 * n + r0 + bs;
      RB_CHECK(rb_gemm_nn(rb_, rest, bs, bs, P, n, sA, linv + j * kNB * kNB, kNB, sinv, ws, rest, rest * bs, nblk));
      aem::launch_copy_panels<T>(ws, rest, bs, nblk, P, n, sA, st_);
    }
    if (r0 > 0) {
      T* P = A + r0 * n;
      RB_CHECK(rb_gemm_nn(rb_, r0, bs, bs, P, n, sA, uinv + j * kNB * kNB, kNB, sinv, ws, r0, r0 * bs, nblk));
      aem::launch_copy_panels<T>(ws, r0, bs, nblk, P, n, sA, st_);
    }
  }
}

// prepare_blocked for every size class; linv / uinv sized and zeroed here
template <class T>
void Solver::prepare_blocked_all(T* A, DBuf<T>& linv, DBuf<T>& uinv, bool lower_unit) {
  linv.zero((size_t)invtot_, st_);
  uinv.zero((size_t)invtot_, st_);
  size_t wsn = 1;
  for (const auto& c : cls_) wsn = std::max(wsn, (size_t)(c.npad * kNB * c.nblk));
  DBuf<T> ws;
  ws.alloc(wsn);
  for (const auto& c : cls_)
    prepare_blocked<T>(A + c.a0, c.npad, c.nblk, linv.p + c.inv0, uinv.p + c.inv0, lower_unit, ws.p);
  HIP_CHECK(hipStreamSynchronize(st_));  // the workspace is released on return
}

// The CSR pattern of A (rows and columns in the unknown layout), built once per handle on the host
// and uploaded: the head residual's product, the FP64 refinement and the sparse factorisation use it.
void Solver::ensure_csr_pattern() {
  if (csr_ready_) return;
  const int64_t N = N_;
  std::vector<int> row_e(N, -1), row_f(N, -1);
  for (int e : de_) row_e[de_pos_[e]] = e;
  for (int f = 0; f < F_; ++f) row_f[fpos_[f]] = f;
  std::vector<int> rp(1, 0), cols, tmp;
  for (int64_t g = 0; g < N; ++g) {
    tmp.clear();
    if (row_e[g] >= 0) {
      const int e = row_e[g];
      const int sides = etype_[e] == 0 ? 2 : 1;
      for (int sd = 0; sd < sides; ++sd) {
        const int f = sd ? efrac1_[e] : efrac0_[e];
        for (int k = mem_ptr_[f]; k < mem_ptr_[f + 1]; ++k) {
          const int src = slots_[k].elem;
          if (src != e && de_pos_[src] >= 0) tmp.push_back(de_pos_[src]);
        }
        tmp.push_back(fpos_[f]);
      }
    } else if (row_f[g] >= 0) {
      const int f = row_f[g];
      for (int k = mem_ptr_[f]; k < mem_ptr_[f + 1]; ++k)
        if (de_pos_[slots_[k].elem] >= 0) tmp.push_back(de_pos_[slots_[k].elem]);
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    cols.insert(cols.end(), tmp.begin(), tmp.end());
    rp.push_back((int)cols.size());
  }
  d_csr_ptr_.upload(rp, st_);
  d_csr_col_.upload(cols, st_);
  d_csr_val_.alloc(cols.size());
  nnz_A_ = (int64_t)cols.size();
  h_csr_ptr_.swap(rp);
  h_csr_col_.swap(cols);
  csr_ready_ = true;
}

void Solver::mf_analyse_now() {
  aem::MfInput in;
  in.F = F_;
  in.E = E_;
  in.etype = etype_.data();
  in.efrac0 = efrac0_.data();
  in.efrac1 = efrac1_.data();
  in.de_pos = de_pos_.data();
  in.fpos = fpos_.data();
  in.center = fcenter_.data();
  in.N = N_;
  in.csr_ptr = h_csr_ptr_.data();
  in.csr_col = h_csr_col_.data();
  in.delay = mf_delay_.empty() ? nullptr : mf_delay_.data();
  try {
    mf_.analyse(in, st_);
  } catch (const std::exception& e) {
    mf_failed_ = true;
    mf_note_ = e.what();
  }
}

void Solver::assemble_and_factor() {
  drop_graph();
  adjT_valid_ = false;
  const int64_t N = N_;
  const int nblk = nblk_;
  if (nblk > 1 && lu_mode_ != ANDFN_MI_LU_TRSV)
    fail(ANDFN_MI_EINVAL, "a network of %d independent parts is solved with the blocked triangular solves (lu_mode 0)",
         nblk);
  ensure_csr_pattern();
  // the row scale of the head residual follows the transmissivities (which set_parameters may change):
  // 1 on intersection rows, 1/t on constant-head and well rows, 0 elsewhere
  {
    std::vector<double> rs(N, 0.0);
    for (int e : de_) rs[de_pos_[e]] = etype_[e] == 0 ? 1.0 : 1.0 / t_[efrac0_[e]];
    d_rowscale_.upload(rs, st_);
  }
  sparse_ = lu_mode_ == ANDFN_MI_LU_TRSV && factor_ == ANDFN_MI_FACTOR_SPARSE;
  if (sparse_ && !mf_.ready() && !mf_failed_) {
    const auto ta0 = std::chrono::steady_clock::now();
    mf_analyse_now();
    if (verbose_)
      fprintf(stderr, "analysis: %.1f ms\n", 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - ta0).count());
  }
  if (sparse_ && mf_failed_) sparse_ = false;  // the dense tournament LU below instead
  if (sparse_) {
    // A assembled straight into its CSR values, row-equilibrated, factored by the multifrontal LU
    HIP_CHECK(hipMemsetAsync(d_csr_val_.p, 0, sizeof(double) * std::max<int64_t>(nnz_A_, 1), st_));
    aem::TargetView tv = rhs_.view();
    aem::launch_assemble_csr(net(), tv, (int)rhs_.side_elem.size(), sys(), K_, F_, d_csr_ptr_.p, d_csr_col_.p,
                             d_csr_val_.p, st_);
    d_req_.alloc(N);
    aem::launch_csr_row_scale(N, d_csr_ptr_.p, d_csr_val_.p, d_req_.p, st_);
    HIP_CHECK(hipGetLastError());
    d_ay_.alloc(N);
    // Pivots are searched within a front's own rows only.  A pivot whose L21 column grows beyond
    // kGrowthLimit (threshold pivoting at 1e-5 of its column; the solves refine once) or that is
    // zero is delayed: its column and row unknowns are lifted to the root front of their part of
    // the network, where every row is fully summed and partial pivoting is complete, the tree is
    // analysed again and A refactored.  (Lifting one level at a time took three factorisations and
    // two analyses on the 20,000-fracture network, where 5 unknowns were delayed and 1 of them
    // again; a handful of unknowns at the root costs nothing measurable.)  The lifts persist on
    // the handle.  After kMaxRounds the dense tournament LU takes over.
    constexpr double kGrowthLimit = 1e5;
    constexpr int kMaxRounds = 8;
    constexpr int kLiftToRoot = 1 << 20;  // levels: more than any tree has
    bool root_zero = false;
    mf_rounds_ = 0;
    for (;;) {
      int bad = 0;
      std::vector<int> lift;
      const auto tr0 = std::chrono::steady_clock::now();
      try {
        bad = mf_.factor(d_csr_val_.p, d_csr_ptr_.p, d_req_.p, st_);
        ++mf_rounds_;
        mf_.unstable_pivots(kGrowthLimit, lift, root_zero, st_);
      } catch (const std::exception& e) {
        fail(ANDFN_MI_EHIP, "multifrontal factorisation: %s", e.what());
      }
      if (verbose_) {
        size_t again = 0;
        for (int u : lift) again += !mf_delay_.empty() && mf_delay_[u] > 0;
        fprintf(stderr, "factorisation round %d: %.1f ms, %zu unknowns to delay (%zu delayed before)\n", mf_rounds_,
                1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - tr0).count(), lift.size(), again);
      }
      if (lift.empty()) {
        if (bad || root_zero) fail(ANDFN_MI_ESINGULAR, "discharge matrix is exactly singular (zero pivot in front %d)", bad - 1);
        break;
      }
      if (mf_rounds_ >= kMaxRounds) {
        mf_failed_ = true;
        mf_note_ = "pivots still unstable after repeated delays";
        sparse_ = false;
        break;
      }
      if (mf_delay_.empty()) mf_delay_.assign(N, 0);
      for (int u : lift) mf_delay_[u] = kLiftToRoot;
      const auto ta0 = std::chrono::steady_clock::now();
      mf_analyse_now();
      if (verbose_)
        fprintf(stderr, "analysis with the delays: %.1f ms\n",
                1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - ta0).count());
      if (mf_failed_) {
        sparse_ = false;
        break;
      }
    }
  }
  if (sparse_) {
    nnz_factors_ = mf_.nnz_factors();
    solve32_ = false;
    a64_blocked_ = false;
    // the solves write the analysed unknowns only: the padding of x, w and v stays zero
    d_x_.zero(N, st_);
    d_w_.zero(N, st_);
    d_v_.zero(N, st_);
    {
      std::vector<int> id((size_t)N);
      for (int64_t i = 0; i < N; ++i) id[i] = (int)i;
      d_perm_.upload(id, st_);
    }
    if (precision_ == ANDFN_MI_MIXED && lu_mode_ == ANDFN_MI_LU_TRSV) {
      // the correction solves x = y + A^-1 (b - A y) stream the factors rounded to FP32 (half the
      // bytes); the FP64 residual b - A y keeps the iteration's fixed point the FP64 one.  They
      // must contract well to pay off; otherwise the FP64 factors serve throughout.
      mf_.prepare_f32(st_);
      solve32_ = true;
      refine_rate_ = refine_probe();
      if (!(refine_rate_ <= 0.25)) solve32_ = false;
    }
    HIP_CHECK(hipStreamSynchronize(st_));
    return;
  }
  d_A_.zero((size_t)Atot_, st_);
  aem::TargetView tv = rhs_.view();
  aem::launch_assemble(net(), tv, (int)rhs_.side_elem.size(), sys(), K_, F_, d_A_.p, st_);
  for (const auto& c : cls_) aem::launch_pad_identity(d_A_.p + c.a0, c.npad, c.nblk, d_blk_n_.p + c.b0, st_);
  HIP_CHECK(hipGetLastError());
  // a CSR copy of the unfactored matrix for the head residual A y - b
  aem::launch_csr_gather(d_A_.p, sys(), N, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, st_);
  HIP_CHECK(hipStreamSynchronize(st_));
  d_ay_.alloc(N);
  // row equilibration (rows of intersections scale with 1/t), as SuperLU equilibrates by default
  d_req_.alloc(N);
  {
    DBuf<unsigned long long> mx;
    mx.alloc(N);
    for (const auto& c : cls_) aem::launch_row_scale(d_A_.p + c.a0, c.npad, c.nblk, d_req_.p + c.u0, mx.p + c.u0, st_);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st_));
  }
  d_ipiv_.alloc(N);
  d_info_.zero(nblk, st_);
  // the tournament LU also stands in for a multifrontal analysis that could not be built
  const bool tournament = lu_mode_ == ANDFN_MI_LU_TRSV && factor_ != ANDFN_MI_FACTOR_ROCSOLVER;
  if (tournament) {
    std::vector<int> id((size_t)N);
    for (int64_t i = 0; i < N; ++i) id[i] = (int)i;
    d_perm_.upload(id, st_);
    // work buffers sized for the largest class first: calu_factor's allocations are then no-ops
    // (no buffer is released while an earlier class's launches may still use it)
    for (const auto& c : cls_) calu_reserve(c.npad, c.nblk);
    for (const auto& c : cls_) calu_factor(d_A_.p + c.a0, c.npad, c.nblk, d_perm_.p + c.u0, d_info_.p + c.b0);
  } else {
    for (const auto& c : cls_) {
      if (c.nblk == 1)
        RB_CHECK(rocsolver_dgetrf_64(rb_, c.npad, c.npad, d_A_.p + c.a0, c.npad, d_ipiv_.p + c.u0, d_info_.p + c.b0));
      else
        RB_CHECK(rocsolver_dgetrf_strided_batched_64(rb_, c.npad, c.npad, d_A_.p + c.a0, c.npad,
                                                     (rocblas_stride)(c.npad * c.npad), d_ipiv_.p + c.u0,
                                                     (rocblas_stride)c.npad, d_info_.p + c.b0, c.nblk));
    }
  }
  {
    std::vector<int64_t> info(nblk, 0);
    HIP_CHECK(hipMemcpyAsync(info.data(), d_info_.p, sizeof(int64_t) * nblk, hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    for (int b = 0; b < nblk; ++b)
      if (info[b] > 0)
        fail(ANDFN_MI_ESINGULAR, "discharge matrix is exactly singular (block %d, zero pivot at %lld)", b,
             (long long)info[b]);
  }
  nnz_factors_ = Atot_;  // dense blocks: L and U share the n^2 entries of each block
  solve32_ = false;
  a64_blocked_ = false;
  if (precision_ == ANDFN_MI_MIXED) {
    // the per-iteration solves use the factors rounded to FP32 (half the bytes); the FP64 residual
    // of the CSR copy drives the refinement, so the iteration's fixed point is the FP64 one.
    // (rocSOLVER's sgetrf_64 in this ROCm reports spurious zero pivots at n = 65,690, on random
    // matrices too, and is no faster than dgetrf below that size, so the factorisation is FP64.)
    d_A32_.alloc((size_t)Atot_);
    aem::launch_to_f32(d_A_.p, Atot_, d_A32_.p, st_);
    solve32_ = true;
  }
  d_x_.alloc(N);
  d_w_.alloc(N);
  if (lu_mode_ == ANDFN_MI_LU_TRSV && !tournament) {
    // LAPACK row interchanges of every block -> one gather permutation: (P b)[i] = b[perm[i]]
    std::vector<int64_t> ipiv(N);
    HIP_CHECK(hipMemcpyAsync(ipiv.data(), d_ipiv_.p, sizeof(int64_t) * N, hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    std::vector<int> perm(N);
    for (const auto& c : cls_)
      for (int b = 0; b < c.nblk; ++b) {
        const int64_t o = c.u0 + (int64_t)b * c.npad;
        for (int64_t i = 0; i < c.npad; ++i) perm[o + i] = (int)(o + i);
        for (int64_t i = 0; i < c.npad; ++i) std::swap(perm[o + i], perm[o + ipiv[o + i] - 1]);
      }
    d_perm_.upload(perm, st_);
    HIP_CHECK(hipStreamSynchronize(st_));
  }
  if (lu_mode_ == ANDFN_MI_LU_TRSV) {
    if (precision_ == ANDFN_MI_MIXED) {
      prepare_blocked_all<float>(d_A32_.p, d_linv32_, d_uinv32_);
      d_x32_.alloc(N);
      d_w32_.alloc(N);
      d_v32_.alloc(N);
      // the refinement must contract well for the FP32-rounded factors to pay off; otherwise the
      // run solves with the FP64 factors (large or ill-conditioned systems)
      refine_rate_ = refine_probe();
      if (!(refine_rate_ <= 0.25)) switch_to_fp64_solves();
    } else {
      prepare_blocked_all<double>(d_A_.p, d_linv_, d_uinv_);
      a64_blocked_ = true;
    }
    d_w_.alloc(N);
    d_v_.alloc(N);
    HIP_CHECK(hipStreamSynchronize(st_));
  } else if (lu_mode_ == ANDFN_MI_LU_INVERSE) {
    // getri has only the 32-bit interface here (this mode is for comparisons at moderate n)
    const int64_t n = N;  // one block (checked above)
    std::vector<int64_t> p64(n);
    HIP_CHECK(hipMemcpyAsync(p64.data(), d_ipiv_.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    std::vector<rocblas_int> p32(p64.begin(), p64.end());
    DBuf<rocblas_int> ip32, inf32;
    ip32.upload(p32, st_);
    inf32.zero(1, st_);
    RB_CHECK(rocsolver_dgetri(rb_, (rocblas_int)n, d_A_.p, (rocblas_int)n, ip32.p, inf32.p));
    HIP_CHECK(hipStreamSynchronize(st_));
  }
}

// In-place LU of the nblk n x n column-major matrices at A + b n^2 with tournament pivoting
// (aem_lu.hip; one launch per step covers every block).  Outer
// panels of kLuOuter columns; within one, per inner panel of kLuPanel columns: the pivot
// tournament, the interchange with L11\U11 and U12 on the outer panel's columns, L21 = A21 U11^-1
// and the update of the outer panel's remaining columns.  Then the outer panel's interchanges on
// all other columns, its U12 and the trailing update: the next outer panel's columns on st_, the
// rest on st2_, overlapping the next outer panel's factorisation (look-ahead).  Leaves the factors in LAPACK layout (unit
// L below the diagonal, U on and above), the row permutation in perm ((P b)[i] = b[perm[i]]; it
// holds the identity on entry, n entries per block) and the first zero pivot (1-based, 0 if none)
// in info[b] for every block b.
// work buffers of calu_factor for nblk blocks of n unknowns (only ever grown)
void Solver::calu_reserve(int64_t n, int nblk) {
  const int64_t groups = (n + aem::kLuGroupRows - 1) / aem::kLuGroupRows;
  const int64_t sp = (int64_t)(aem::kLuOuter / aem::kLuPanel) * aem::kLuPairStride;
  for (int i = 0; i < 2; ++i) {
    lu_cand_[i].alloc((size_t)(nblk * groups * aem::kLuPanel));
    lu_cnt_[i].alloc((size_t)(nblk * groups));
  }
  lu11_.alloc((size_t)nblk * aem::kLuPanel * aem::kLuPanel);
  lu_pairs_.alloc((size_t)(nblk * sp));
}

void Solver::calu_factor(double* A, int64_t n, int nblk, int* perm, int64_t* info) {
  const int nslots = aem::kLuOuter / aem::kLuPanel;
  const int64_t sA = n * n, sp = (int64_t)nslots * aem::kLuPairStride;
  calu_reserve(n, nblk);
  if (!st2_) {
    // A trailing update fills every CU for a while; the panel kernels of the next outer panel,
    // sharing CUs with it, run at half speed, which matters once the update is shorter than
    // the panel chain.  st2_ is kept off `reserve` CUs of each of the 8 XCDs (mask bit i lies in
    // block g = i / (cu / 8), position p = i % (cu / 8); bits with p % 8 == g and p / 8 < reserve
    // are cleared, which spreads the cleared bits over the XCDs whether the mask enumerates CUs
    // XCD-major or interleaved) and st3_ has only those CUs.  st_ keeps every CU.
    int reserve = 4;
    if (const char* e = std::getenv("ANDFN_MI_LU_RESERVE")) reserve = std::atoi(e);
    int dev = 0, cu = 0;
    HIP_CHECK(hipGetDevice(&dev));
    HIP_CHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
    const int per = cu / 8;
    if (reserve > 0 && cu % 8 == 0 && per % 8 == 0 && reserve <= per / 8) {
      std::vector<uint32_t> mask((cu + 31) / 32, 0u), kept((cu + 31) / 32, 0u);
      for (int i = 0; i < cu; ++i) {
        const int g = i / per, p = i % per;
        if (!(p % 8 == g && p / 8 < reserve))
          mask[i / 32] |= 1u << (i % 32);
        else
          kept[i / 32] |= 1u << (i % 32);
      }
      // a device or runtime that refuses CU masks gets plain streams (and no narrow split)
      if (hipExtStreamCreateWithCUMask(&st2_, (uint32_t)mask.size(), mask.data()) != hipSuccess ||
          hipExtStreamCreateWithCUMask(&st3_, (uint32_t)kept.size(), kept.data()) != hipSuccess) {
        (void)hipGetLastError();
        if (st2_) (void)hipStreamDestroy(st2_);
        if (st3_) (void)hipStreamDestroy(st3_);
        st2_ = st3_ = nullptr;
      }
    }
    if (!st2_) HIP_CHECK(hipStreamCreate(&st2_));
    RB_CHECK(rocblas_create_handle(&rb2_));
    RB_CHECK(rocblas_set_stream(rb2_, st2_));
    for (auto& e : lu_ev_) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  int* cand[2] = {lu_cand_[0].p, lu_cand_[1].p};
  int* cnt[2] = {lu_cnt_[0].p, lu_cnt_[1].p};
  const double one = 1.0, mone = -1.0;
  // C -= Ablk Bblk on every block (rocBLAS, strided-batched for several blocks)
  auto gemm = [&](rocblas_handle h, int64_t m, int64_t nc, int64_t k, const double* a, const double* bm, double* c) {
    if (nblk == 1)
      RB_CHECK(rocblas_dgemm_64(h, rocblas_operation_none, rocblas_operation_none, m, nc, k, &mone, a, n, bm, n, &one,
                                c, n));
    else
      RB_CHECK(rocblas_dgemm_strided_batched_64(h, rocblas_operation_none, rocblas_operation_none, m, nc, k, &mone, a,
                                                n, sA, bm, n, sA, &one, c, n, sA, nblk));
  };
  // Where the work of an outer panel runs, set when the previous outer panel issues its trailing
  // update (which overlaps this outer panel's factorisation):
  //   kSerial  everything on st_ (no trailing update pending: the first outer panel);
  //   kNarrow  a small update (below narrow_flops, one block) on st2_, the panel kernels on st3_
  //            and the inner updates on st_;
  //   kWide    a large update on st_ itself, which keeps every CU for it, and the whole panel
  //            chain on st2_ (the chain waits on the update for CUs either way; there it keeps
  //            the reserved CUs).
  // narrow_flops: 400 GFLOP (measured: on a 19.7k block every outer panel runs faster narrow;
  // on a 65.7k block the updates above it run faster wide).
  enum { kSerial, kNarrow, kWide } mode = kSerial;
  bool pending = false;  // a trailing update on st2_ has not been waited for yet
  static const double narrow_flops = [] {
    const char* e = std::getenv("ANDFN_MI_LU_NARROW_GFLOP");
    return (e ? std::atof(e) : 400.0) * 1e9;
  }();
  for (int64_t J0 = 0; J0 < n; J0 += aem::kLuOuter) {
    const int64_t J1 = std::min<int64_t>(n, J0 + aem::kLuOuter);
    // inner panels: only columns [J0, J1) are read or written (brought up to date by the
    // previous outer panel's update on st_), so they overlap the trailing update on st2_
    int nl = 0;
    for (int64_t j0 = J0; j0 < J1; j0 += aem::kLuPanel, ++nl) {
      const int bw = (int)std::min<int64_t>(aem::kLuPanel, J1 - j0);
      int* pairs = lu_pairs_.p + (int64_t)nl * aem::kLuPairStride;
      hipStream_t pst = mode == kWide ? st2_ : st_;  // panel kernels
      rocblas_handle ib = mode == kWide ? rb2_ : rb_;  // inner update
      if (mode == kNarrow) {
        HIP_CHECK(hipEventRecord(lu_ev_[2], st_));
        HIP_CHECK(hipStreamWaitEvent(st3_, lu_ev_[2], 0));
        pst = st3_;
      }
      const int* win =
          aem::launch_tslu(A, n, n, j0, bw, cand, cnt, lu11_.p, pairs, perm, info, pst, nblk, sA, sp);
      aem::launch_calu_apply(A, n, n, J0, J1, j0, bw, pairs, win, lu11_.p, pst, nblk, sA, sp);
      aem::launch_calu_l21(A, n, n, j0, bw, lu11_.p, pst, nblk, sA);
      if (mode == kNarrow) {
        HIP_CHECK(hipEventRecord(lu_ev_[3], st3_));
        HIP_CHECK(hipStreamWaitEvent(st_, lu_ev_[3], 0));
      }
      const int64_t rows = n - j0 - bw, cols = J1 - j0 - bw;
      if (rows > 0 && cols > 0)
        gemm(ib, rows, cols, bw, A + (j0 + bw) + j0 * n, A + j0 + (j0 + bw) * n, A + (j0 + bw) + (j0 + bw) * n);
    }
    if (mode == kWide) {  // the chain on st2_ ran beside the update on st_
      HIP_CHECK(hipEventRecord(lu_ev_[3], st2_));
      HIP_CHECK(hipStreamWaitEvent(st_, lu_ev_[3], 0));
    }
    // everything below touches columns >= J1 (and < J0): the previous trailing update must be done
    if (pending) {
      HIP_CHECK(hipStreamWaitEvent(st_, lu_ev_[1], 0));
      pending = false;
    }
    aem::launch_calu_swap(A, n, n, J0, J1, lu_pairs_.p, nl, st_, nblk, sA, sp);
    const int64_t rest = n - J1;
    if (rest > 0) {
      if (nblk == 1)
        RB_CHECK(rocblas_dtrsm_64(rb_, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                                  rocblas_diagonal_unit, J1 - J0, rest, &one, A + J0 + J0 * n, n, A + J0 + J1 * n, n));
      else
        RB_CHECK(rocblas_dtrsm_strided_batched_64(rb_, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                                                  rocblas_diagonal_unit, J1 - J0, rest, &one, A + J0 + J0 * n, n, sA,
                                                  A + J0 + J1 * n, n, sA, nblk));
      // the next outer panel's columns first, on st_; the rest of the update on st2_
      const int64_t J2 = std::min<int64_t>(n, J1 + aem::kLuOuter);
      gemm(rb_, rest, J2 - J1, J1 - J0, A + J1 + J0 * n, A + J0 + J1 * n, A + J1 + J1 * n);
      mode = kSerial;
      if (J2 < n) {
        mode = st3_ != nullptr && nblk == 1 && 2.0 * rest * (n - J2) * (J1 - J0) < narrow_flops ? kNarrow : kWide;
        // the next outer panel's columns are up to date: its chain may start
        HIP_CHECK(hipEventRecord(lu_ev_[0], st_));
        HIP_CHECK(hipStreamWaitEvent(st2_, lu_ev_[0], 0));
        if (mode == kNarrow) {
          gemm(rb2_, rest, n - J2, J1 - J0, A + J1 + J0 * n, A + J0 + J2 * n, A + J1 + J2 * n);
          HIP_CHECK(hipEventRecord(lu_ev_[1], st2_));
          pending = true;
        } else {
          gemm(rb_, rest, n - J2, J1 - J0, A + J1 + J0 * n, A + J0 + J2 * n, A + J1 + J2 * n);
        }
      }
    }
  }
  if (pending) HIP_CHECK(hipStreamWaitEvent(st_, lu_ev_[1], 0));
  HIP_CHECK(hipGetLastError());
}

// Mixed form: contraction rate of the refinement y <- y + (LU)_32^-1 R P (b - A y) on the probe
// system A y = A 1: sqrt(r3 / r1) with r_k the head residual after k steps from y = 0.
double Solver::refine_probe() {
  const int n = (int)N_;
  DBuf<double> ones, b0, y, x;
  ones.upload(std::vector<double>(n, 1.0), st_);
  b0.alloc(n);
  y.zero(n, st_);
  x.alloc(n);
  d_ay_.alloc(n);
  aem::launch_spmv(n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, ones.p, b0.p, st_);
  double r[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; ++k) {
    aem::launch_spmv(n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, y.p, d_ay_.p, st_);
    aem::launch_residual(d_ay_.p, b0.p, d_rowscale_.p, n, d_red_.p, d_scal_.p, st_);
    aem::IterScalars sc{};
    HIP_CHECK(hipMemcpyAsync(&sc, d_scal_.p, sizeof(sc), hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    r[k] = sc.residual;
    if (k == 3) break;
    if (sparse_) {
      mf_.solve_correction(b0.p, d_ay_.p, y.p, d_req_.p, x.p, st_, true);
    } else {
      aem::launch_refine_rhs(b0.p, d_ay_.p, d_perm_.p, d_req_.p, n, d_x32_.p, st_);
      blocked_solve_all<float>(d_A32_.p, d_linv32_.p, d_uinv32_.p, d_x32_.p, d_w32_.p, d_v32_.p);
      aem::launch_refine_update(y.p, d_v32_.p, n, x.p, st_);
    }
    HIP_CHECK(hipMemcpyAsync(y.p, x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st_));
  }
  HIP_CHECK(hipStreamSynchronize(st_));
  if (!(r[1] > 0.0)) return 0.0;
  return std::sqrt(r[3] / r[1]);
}

// Mixed form: continue with the FP64 factors (the refinement with the FP32-rounded factors
// stopped contracting).  Builds the FP64 blocked operators on first use.
void Solver::switch_to_fp64_solves() {
  if (!solve32_) return;
  HIP_CHECK(hipStreamSynchronize(st_));  // no graph of this handle is in flight when they are dropped
  if (!sparse_ && !a64_blocked_) {
    prepare_blocked_all<double>(d_A_.p, d_linv_, d_uinv_);
    d_w_.alloc(N_);
    d_v_.alloc(N_);
    a64_blocked_ = true;
  }
  solve32_ = false;
  drop_graph();
  adjT_valid_ = false;
  HIP_CHECK(hipStreamSynchronize(st_));
}

// d_x_ = A^-1 d_b_.  The triangular-solve sequence (one permutation gather and two gemv per
// diagonal block) is launch-bound, so it is captured once into a HIP graph and replayed.
template <class Fn>
void Solver::run_captured(CapturedSolve& c, Fn&& fn) {
  if (!c.tried) {
    c.tried = true;
    fn();  // uncaptured first call: any lazy library set-up happens here
    HIP_CHECK(hipStreamSynchronize(st_));
    // ANDFN_MI_NO_GRAPH set: launch the solve's kernels directly (profiler counter collection
    // instruments every dispatch one by one)
    static const bool no_graph = std::getenv("ANDFN_MI_NO_GRAPH") != nullptr;
    if (!no_graph && hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      fn();
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(st_, &g) == hipSuccess && g) {
        if (hipGraphInstantiate(&c.e, g, nullptr, nullptr, 0) == hipSuccess) {
          c.g = g;
        } else {
          (void)hipGraphDestroy(g);
          c.e = nullptr;
        }
      }
      (void)hipGetLastError();
    }
    return;  // the first call's result is already in place
  }
  if (c.e) {
    HIP_CHECK(hipGraphLaunch(c.e, st_));
  } else {
    fn();
  }
}

void Solver::solve_rhs() {
  if (lu_mode_ != ANDFN_MI_LU_TRSV) {
    (void)discharge_solve(d_b_.p, d_x_.p);
    return;
  }
  run_captured(cap_solve_, [&] { (void)discharge_solve(d_b_.p, d_x_.p); });
}

// x = y + A^-1 (b - A y) with the sparse factors, A y of the current iterate y already in d_ay_
// (the head residual's product): one solve, accurate to the condition number times the rounding
// unit times |b - A y| rather than |b|
void Solver::correct_rhs() {
  run_captured(cap_correct_,
               [&] { mf_.solve_correction(d_b_.p, d_ay_.p, d_y_.p, d_req_.p, d_x_.p, st_, solve32_); });
}

// v = U^-1 L^-1 x with the operators of prepare_blocked (x and w are overwritten), for the nblk
// blocks of n unknowns of one size class at once (every launch covers the same step of every block)
template <class T>
void Solver::blocked_solve(const T* A, int n, int nblk, const T* linv, const T* uinv, T* x, T* w, T* v) {
  const int nb = (n + kNB - 1) / kNB;
  const int64_t sA = (int64_t)n * n, sinv = (int64_t)nb * kNB * kNB;
  for (int j = 0; j < nb; ++j) {
    const int r0 = j * kNB, bs = std::min(kNB, n - r0), rest = n - r0 - bs;
    if (rest > 0)
      aem::launch_gemv_cm<T>(A + (size_t)r0 * n + r0 + bs, n, rest, bs, x + r0, x + r0 + bs, T(-1), T(1), st_, nblk,
                             sA, n);
  }
  aem::launch_gemv_diag<T>(linv, kNB, n, 1, x, w, st_, nblk, sinv);
  for (int j = nb - 1; j > 0; --j) {
    const int r0 = j * kNB, bs = std::min(kNB, n - r0);
    aem::launch_gemv_cm<T>(A + (size_t)r0 * n, n, r0, bs, w + r0, w, T(-1), T(1), st_, nblk, sA, n);
  }
  aem::launch_gemv_diag<T>(uinv, kNB, n, 2, w, v, st_, nblk, sinv);
}

// blocked_solve for every size class (the class's slices of A, the inverses and the vectors)
template <class T>
void Solver::blocked_solve_all(const T* A, const T* linv, const T* uinv, T* x, T* w, T* v) {
  for (const auto& c : cls_)
    blocked_solve<T>(A + c.a0, (int)c.npad, c.nblk, linv + c.inv0, uinv + c.inv0, x + c.u0, w + c.u0, v + c.u0);
}

// x = A^-1 b with the factors (or inverse) of the discharge matrix; returns the buffer holding x
double* Solver::discharge_solve(double* b, double* x) {
  const rocblas_int n = (rocblas_int)N_;  // one block in the getrs / inverse modes
  if (lu_mode_ == ANDFN_MI_LU_GETRS) {
    aem::launch_scale(b, d_req_.p, n, x, st_);
    RB_CHECK(rocsolver_dgetrs_64(rb_, rocblas_operation_none, n, 1, d_A_.p, n, d_ipiv_.p, x, n));
    return x;
  }
  if (lu_mode_ == ANDFN_MI_LU_INVERSE) {
    const double one = 1.0, zero = 0.0;
    aem::launch_scale(b, d_req_.p, n, d_w_.p, st_);
    RB_CHECK(rocblas_dgemv(rb_, rocblas_operation_none, n, n, &one, d_A_.p, n, d_w_.p, 1, &zero, x, 1));
    return x;
  }
  if (sparse_) {
    mf_.solve_refined(b, d_req_.p, x, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, st_);
    return x;
  }
  if (solve32_) {
    // x = y + (L U)^-1 R P (b - A y): one refinement step per iteration, FP32 factors
    aem::launch_spmv(n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, d_y_.p, d_ay_.p, st_);
    aem::launch_refine_rhs(b, d_ay_.p, d_perm_.p, d_req_.p, n, d_x32_.p, st_);
    blocked_solve_all<float>(d_A32_.p, d_linv32_.p, d_uinv32_.p, d_x32_.p, d_w32_.p, d_v32_.p);
    float* v32 = d_v32_.p;
    aem::launch_refine_update(d_y_.p, v32, n, x, st_);
    return x;
  }
  // R P b, then the blocked triangular solves
  aem::launch_permute(b, d_perm_.p, d_req_.p, n, x, st_);
  double* v = d_v_.p;
  blocked_solve_all<double>(d_A_.p, d_linv_.p, d_uinv_.p, x, d_w_.p, v);
  HIP_CHECK(hipMemcpyAsync(x, v, sizeof(double) * n, hipMemcpyDeviceToDevice, st_));
  return x;
}

// Collocation / right-hand-side sweep with the tuned tile shape.  The first call of each kind runs
// both shapes twice on the same inputs (identical values) and keeps the faster.
void Solver::launch_sweep_shape(int mode, const aem::NetView& nv, const TargetSet& ts, double* val, bool mixed,
                                int ppl, int persist) {
  if (persist) {
    if (!d_queue_.p) {
      // first use (the autotuner's trials, never inside a graph capture): the queue and the
      // resident workgroup counts of every variant
      d_queue_.zero(9, st_);
      int cus = 0;
      HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_));
      for (int r = 0; r < 2; ++r)
        for (int m = 0; m < 2; ++m)
          for (int p = 0; p < 2; ++p)
            persist_groups_[r][m][p] = aem::sweep_persistent_groups(r ? aem::MODE_RHS : aem::MODE_COLLOC, m, p + 1, cus);
    }
    const int nwg = persist_groups_[mode == aem::MODE_RHS][mixed ? 1 : 0][ppl == 2];
    if (nwg > 0 && ts.ntiles(ppl) % 8 == 0) {
      aem::launch_sweep_persistent(mode, nv, ts.view(ppl), ts.ntiles(ppl), K_, val, st_, mixed, ppl, nwg, d_queue_.p);
      return;
    }
  }
  aem::launch_sweep(mode, nv, ts.view(ppl), ts.ntiles(ppl), K_, val, st_, mixed, ppl);
}

// Collocation / right-hand-side sweep with the tuned tile shape.  The first call of each kind runs
// every shape twice on the same inputs (identical values): one or two points per lane, one
// workgroup per tile or one persistent launch; the fastest is kept, unless it beats the plain
// one-point grid by less than 3%.
void Solver::sweep(int mode, const TargetSet& ts, double* val, bool mixed) {
  const int slot = mode == aem::MODE_RHS ? 1 : 0;
  const aem::NetView nv = net();
  if (ppl_[slot] == 0 && ts.ntiles(1) > 0) {
    // shape k: points per lane 1 + (k & 1), persistent k >> 1; rounds 0 1 2 3 3 2 1 0
    static const bool no_persist = std::getenv("ANDFN_MI_NO_PERSISTENT") != nullptr;
    const int nshape = no_persist ? 2 : 4;
    float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    for (int r = 0; r < 2 * nshape; ++r) {
      const int k = r < nshape ? r : 2 * nshape - 1 - r;
      float ms = 0.f;
      HIP_CHECK(hipEventRecord(ev_[6], st_));
      launch_sweep_shape(mode, nv, ts, val, mixed, 1 + (k & 1), k >> 1);
      HIP_CHECK(hipEventRecord(ev_[7], st_));
      HIP_CHECK(hipEventSynchronize(ev_[7]));
      HIP_CHECK(hipEventElapsedTime(&ms, ev_[6], ev_[7]));
      best[k] = std::min(best[k], ms);
    }
    int pick = 0;
    for (int k = 1; k < nshape; ++k)
      if (best[k] < 0.97f * best[0] && best[k] < best[pick]) pick = k;
    ppl_[slot] = 1 + (pick & 1);
    persist_[slot] = pick >> 1;
    return;
  }
  const int ppl = ppl_[slot] ? ppl_[slot] : 1;
  launch_sweep_shape(mode, nv, ts, val, mixed, ppl, persist_[slot]);
}

// max_j |(A0 y)_j - b_j| * row_scale_j over the head equations for the current y
// head_residual without waiting: the scalars land in h_scal_ (pinned) when res_ev_ completes
void Solver::head_residual_async(const double* b) {
  const int n = (int)N_;
  aem::launch_spmv(n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, d_y_.p, d_ay_.p, st_);
  aem::launch_residual(d_ay_.p, b, d_rowscale_.p, n, d_red_.p, d_scal_.p, st_);
  HIP_CHECK(hipMemcpyAsync(h_scal_, d_scal_.p, sizeof(aem::IterScalars), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipEventRecord(res_ev_, st_));
}

double Solver::head_residual(const double* b, int* row) {
  const int n = (int)N_;
  aem::launch_spmv(n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, d_y_.p, d_ay_.p, st_);
  aem::launch_residual(d_ay_.p, b, d_rowscale_.p, n, d_red_.p, d_scal_.p, st_);
  aem::IterScalars sc{};
  HIP_CHECK(hipMemcpyAsync(&sc, d_scal_.p, sizeof(sc), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  if (row) *row = sc.arg_residual;
  return sc.residual;
}

// pointwise boundary-condition residual of the current state
double Solver::bc_residual() {
  aem::NetView nv = net();
  DBuf<double> hv, hw;
  hv.alloc(std::max<int64_t>(full_.npts(), 1));
  hw.alloc(std::max<int64_t>(rhs_.npts(), 1));
  aem::launch_sweep(aem::MODE_COLLOC_HEADS, nv, full_.view(), full_.ntiles(1), K_, hv.p, st_);
  std::vector<RhsItem> witems;
  for (const RhsItem& it : rhs_items_)
    if (etype_[it.elem] == 2) witems.push_back(it);
  DBuf<RhsItem> dw;
  dw.upload(witems, st_);
  if (!witems.empty())
    aem::launch_sweep(aem::MODE_DISCH_HEADS, nv, rhs_.view(), rhs_.ntiles(1), K_, hw.p, st_);
  // collocation items of the current counts
  std::vector<FitItem> items;
  build_items(full_, items, nullptr);
  DBuf<FitItem> di;
  di.upload(items, st_);
  const int nb = (int)(items.size() + witems.size());
  d_bcmax_.alloc(std::max(nb, 1));
  aem::launch_bc_residual(nv, di.p, (int)items.size(), hv.p, full_.d_pt_off.p, dw.p, (int)witems.size(), hw.p,
                          rhs_.d_pt_off.p, K_, d_bcmax_.p, st_);
  HIP_CHECK(hipGetLastError());
  std::vector<double> mx(std::max(nb, 1), 0.0);
  HIP_CHECK(hipMemcpyAsync(mx.data(), d_bcmax_.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  double r = 0.0;
  for (int k = 0; k < nb; ++k) r = (mx[k] != mx[k] || mx[k] > r) ? mx[k] : r;
  return r;
}

// Basis of the right-hand side (rhs_basis_kernel) for the current discharge target set: every
// source gets as many powers as its coefficient count can reach (MAX_NCOEF + COEF_INCREASE - 1
// while it may still grow).  Returns false (the iteration then sweeps the discharge points) when
// the discharge points exceed what the kernel takes or the basis would exceed 16 GB.
bool Solver::ensure_rhs_basis(int max_coef, int coef_increase) {
  if (K_ > aem::kRhsMaxK) return false;
  // ANDFN_MI_RHS_SWEEP set: sweep the discharge points every iteration (the fallback path; tests
  // compare the two)
  if (std::getenv("ANDFN_MI_RHS_SWEEP")) return false;
  std::vector<int> mcap(E_, 0);
  for (int e = 0; e < E_; ++e)
    if (etype_[e] != 2)
      mcap[e] = ncoef_[e] < max_coef ? (int)std::min<int64_t>(M_, max_coef + coef_increase - 1) : ncoef_[e];
  if (rb_ok_ && mcap == rb_mcap_) return true;
  // a column packs (source, 2 j + part) into 32 bits
  if ((int64_t)E_ > ((int64_t)1 << (32 - aem::kRhsColShift)) - 1 || 2 * M_ > ((int64_t)1 << aem::kRhsColShift))
    return false;
  const int ns = (int)rhs_.side_elem.size();
  std::vector<int64_t> ptr(ns + 1, 0), coff(std::max(ns, 1), 0);
  std::vector<int64_t> fcol(F_, -1);  // first column of each fracture's table (-1: not built yet)
  std::vector<unsigned> col;
  // entries per power of a source: one for real coefficients (Re c_j), two otherwise
  auto parts = [&](const Slot& sl) { return sl.kind == aem::KIND_LINE && real_lines_ ? 1 : 2; };
  for (int d = 0; d < ns; ++d) {
    const int e = rhs_.side_elem[d];
    const int f = rhs_.side_which[d] ? efrac1_[e] : efrac0_[e];
    if (fcol[f] < 0) {
      // the fracture's columns in the order rhs_basis_kernel writes a side's entries
      fcol[f] = (int64_t)col.size();
      for (int k = mem_ptr_[f]; k < mem_ptr_[f + 1]; ++k)
        if (slots_[k].kind != aem::KIND_WELL) {
          const unsigned src = (unsigned)slots_[k].elem << aem::kRhsColShift;
          const int np = parts(slots_[k]);
          for (int j = 0; j < mcap[slots_[k].elem]; ++j)
            for (int part = 0; part < np; ++part) col.push_back(src | (unsigned)(2 * j + part));
        }
    }
    coff[d] = fcol[f];
    const int64_t fend = (int64_t)col.size();
    int64_t cnt = 0;
    for (int k = mem_ptr_[f]; k < mem_ptr_[f + 1]; ++k)
      if (slots_[k].kind != aem::KIND_WELL) cnt += (int64_t)parts(slots_[k]) * mcap[slots_[k].elem];
    if (fcol[f] + cnt > fend) throw std::logic_error("basis column table shorter than a side's entries");
    ptr[d + 1] = ptr[d] + cnt;
  }
  const int64_t entries = ptr[ns];
  if (entries * (int64_t)sizeof(double) > (int64_t)16 << 30) {
    rb_ok_ = false;
    return false;
  }
  d_rbPtr_.upload(ptr, st_);
  d_rbMcap_.upload(mcap, st_);
  d_rbCol_.upload(col, st_);
  d_rbCoff_.upload(coff, st_);
  d_rbW_.alloc((size_t)std::max<int64_t>(entries, 1));
  d_rbNan_.alloc((size_t)std::max(ns, 1));
  rb32_ok_ = false;
  aem::launch_rhs_basis(net(), rhs_.view(), ns, K_, d_rbPtr_.p, d_rbMcap_.p, d_rbW_.p, d_rbNan_.p, st_);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st_));
  rb_mcap_ = mcap;
  rb_entries_ = entries;
  rb_ok_ = true;
  return true;
}

void Solver::ensure_iter_log(int cap) {
  if (cap <= it_cap_) return;
  if (h_it_resid_) HIP_CHECK(hipHostFree(h_it_resid_));
  if (h_it_arg_) HIP_CHECK(hipHostFree(h_it_arg_));
  h_it_resid_ = nullptr;
  h_it_arg_ = nullptr;
  HIP_CHECK(hipHostMalloc((void**)&h_it_resid_, sizeof(double) * cap, hipHostMallocCoherent | hipHostMallocMapped));
  HIP_CHECK(hipHostMalloc((void**)&h_it_arg_, sizeof(int) * cap, hipHostMallocCoherent | hipHostMallocMapped));
  if (!h_it_stop_)
    HIP_CHECK(hipHostMalloc((void**)&h_it_stop_, sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
  it_cap_ = cap;
  void *dr = nullptr, *da = nullptr, *ds = nullptr;
  HIP_CHECK(hipHostGetDevicePointer(&dr, h_it_resid_, 0));
  HIP_CHECK(hipHostGetDevicePointer(&da, h_it_arg_, 0));
  HIP_CHECK(hipHostGetDevicePointer(&ds, h_it_stop_, 0));
  it_log_dev_ = aem::IterHostLog{(double*)dr, (int*)da, (int*)ds, cap, nullptr};
  drop_iter_graphs();  // captured graphs hold the previous addresses
  if (!it_end_ev_[0]) {
    for (auto& e : it_end_ev_) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int khz = 0;
    HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev_));
    wall_khz_ = khz > 0 ? (double)khz : 100000.0;
  }
  d_ctl_.alloc(1);
  d_stamps_.alloc(6 * (size_t)cap);
  it_log_dev_.stamps = d_stamps_.p;
}

// One iteration after the first (hpc_solve.py:152-239 with the head-residual stop): right-hand side
// of the current coefficients, head residual of the current discharges and the device's stop
// decision, the correction solve and its acceptance, then the sweep, fits, damping and the
// iteration scalars.  The same launches in the same order as the host loop below.
void Solver::enqueue_iteration(const aem::NetView& nv, const aem::FitView& fv, bool phase32, bool s32, int par,
                               double head_tol, double damping) {
  const int n = (int)N_;
  const int nde = (int)de_.size();
  (void)par;
  unsigned long long* stamps = d_stamps_.p;
  aem::launch_iter_stamp(d_ctl_.p, stamps, it_cap_, 0, 1, st_);
  if (phase32) aem::launch_coef_to_f32(nv, E_, d_coef32_.p, st_);
  aem::launch_rhs_eval(nv, d_rhs_items_.p, nde, d_rbPtr_.p, d_rbW_.p, phase32 ? d_rbW32_.p : nullptr, d_rbCol_.p,
                       d_rbCoff_.p, d_rbNan_.p, d_de_pos_.p, d_b_.p, st_);
  aem::launch_spmv(n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, d_y_.p, d_ay_.p, st_);
  aem::launch_residual(d_ay_.p, d_b_.p, d_rowscale_.p, n, d_red_.p, d_scal_.p, st_);
  aem::launch_iter_decide(d_scal_.p, d_ctl_.p, phase32 ? 0 : 1, head_tol, it_log_dev_, st_);
  mf_.solve_correction(d_b_.p, d_ay_.p, d_y_.p, d_req_.p, d_x_.p, st_, s32);
  aem::launch_accept(nv, d_de_.p, nde, F_, sys(), d_x_.p, d_y_.p, d_yold_.p, n, st_);
  aem::launch_iter_stamp(d_ctl_.p, stamps, it_cap_, 2, 0, st_);
  launch_sweep_shape(aem::MODE_COLLOC, nv, full_, d_val_.p, phase32, ppl_[0] ? ppl_[0] : 1, persist_[0]);
  aem::launch_iter_stamp(d_ctl_.p, stamps, it_cap_, 3, 0, st_);
  // both damping steps of the iteration inside the fit and well kernels (the fit items and the
  // wells are every element: no blend launch)
  aem::launch_fit(nv, fv, d_items_.p, (int)items_.size(), d_citems_.p, ncitems_, d_val_.p, full_.d_pt_off.p,
                  full_.d_dft_tiles.p, (int)full_.dft_tiles.size(), full_.d_dft_sides.p, d_dft_.p, damping, 2, st_,
                  phase32);
  aem::launch_wells(nv, fv, d_wells_.p, (int)wells_.size(), damping, st_, 2);
  aem::launch_iter_stamp(d_ctl_.p, stamps, it_cap_, 4, 0, st_);
  aem::launch_reduce(fv, E_, d_y_.p, d_yold_.p, n, d_red_.p, d_scal_.p, st_);
}

// enqueue_iteration as a graph launch: the first use of a configuration runs its launches
// directly (any lazy set-up happens outside a capture), the second captures and instantiates
// them; a configuration whose capture fails keeps running directly
void Solver::launch_iteration(const aem::NetView& nv, const aem::FitView& fv, bool phase32, bool s32, int par,
                              double head_tol, double damping) {
  launch_iteration_body(nv, fv, phase32, s32, par, head_tol, damping);
  HIP_CHECK(hipEventRecord(it_end_ev_[par], st_));
}

void Solver::launch_iteration_body(const aem::NetView& nv, const aem::FitView& fv, bool phase32, bool s32, int par,
                                   double head_tol, double damping) {
  const bool no_graph = std::getenv("ANDFN_MI_NO_GRAPH") != nullptr;
  IterGraph& g = it_graphs_[phase32][s32][par];
  if (g.e) {
    HIP_CHECK(hipGraphLaunch(g.e, st_));
    ++it_graph_launches_;
    return;
  }
  if (no_graph || g.failed || g.uses++ == 0) {
    enqueue_iteration(nv, fv, phase32, s32, par, head_tol, damping);
    return;
  }
  bool ok = hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    try {
      enqueue_iteration(nv, fv, phase32, s32, par, head_tol, damping);
    } catch (...) {
      ok = false;  // nothing captured ran; the launches are made directly below
    }
    hipGraph_t gr = nullptr;
    ok = hipStreamEndCapture(st_, &gr) == hipSuccess && gr && ok;
    if (ok && hipGraphInstantiate(&g.e, gr, nullptr, nullptr, 0) == hipSuccess) {
      g.g = gr;
    } else {
      if (gr) (void)hipGraphDestroy(gr);
      g.e = nullptr;
      ok = false;
    }
  }
  (void)hipGetLastError();
  if (ok) {
    HIP_CHECK(hipGraphLaunch(g.e, st_));
    ++it_graph_launches_;
  } else {
    g.failed = true;
    enqueue_iteration(nv, fv, phase32, s32, par, head_tol, damping);
  }
}

int Solver::solve(const andfn_mi_options* opt, andfn_mi_info* info) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  const double max_error = C_.MAX_ERROR;
  const int max_iterations = C_.MAX_ITERATIONS;
  const double coef_ratio = C_.COEF_RATIO;
  const int max_coef = C_.MAX_NCOEF;
  const int coef_increase = C_.COEF_INCREASE;
  const double damping = C_.DAMPING;
  const int verbose = opt ? opt->verbose : 0;
  verbose_ = verbose;
  lu_mode_ = opt ? opt->lu_mode : ANDFN_MI_LU_TRSV;
  factor_ = opt ? opt->factor : ANDFN_MI_FACTOR_SPARSE;
  if (factor_ != ANDFN_MI_FACTOR_SPARSE && factor_ != ANDFN_MI_FACTOR_TOURNAMENT && factor_ != ANDFN_MI_FACTOR_ROCSOLVER)
    fail(ANDFN_MI_EINVAL, "factor %d unknown", factor_);
  // the multifrontal factors serve the blocked-solve mode only; the other lu_modes are dense
  if (factor_ == ANDFN_MI_FACTOR_SPARSE && lu_mode_ != ANDFN_MI_LU_TRSV) factor_ = ANDFN_MI_FACTOR_ROCSOLVER;
  precision_ = opt ? opt->precision : ANDFN_MI_FP64;
  if (precision_ != ANDFN_MI_FP64 && precision_ != ANDFN_MI_MIXED) fail(ANDFN_MI_EINVAL, "precision %d unknown", precision_);
  if (precision_ == ANDFN_MI_MIXED && lu_mode_ != ANDFN_MI_LU_TRSV)
    fail(ANDFN_MI_EINVAL, "the mixed-precision form uses the blocked triangular solves (lu_mode 0)");
  mixed_phase_ = precision_ == ANDFN_MI_MIXED;
  if (mixed_phase_ && !(opt && opt->head_tol > 0.0))
    fail(ANDFN_MI_EINVAL, "the mixed-precision form stops on the FP64 head residual: set head_tol > 0");
  {
    const double ff = (opt && opt->far_field > 0.0) ? opt->far_field : 4.0;
    far2_ = ff * ff;
  }
  if (mixed_phase_) d_coef32_.zero(coef_.size(), st_);
  if (lu_mode_ < 0 || lu_mode_ > 2) fail(ANDFN_MI_EINVAL, "lu_mode %d unknown", lu_mode_);
  const int32_t* sched = opt ? opt->growth_schedule : nullptr;
  const int64_t sched_n = opt ? opt->schedule_iterations : 0;
  int32_t* record = opt ? opt->growth_record : nullptr;
  const int64_t record_n = opt ? opt->record_iterations : 0;
  double* log_out = opt ? opt->iteration_log : nullptr;
  const int64_t log_n = log_out ? std::min<int64_t>(opt->log_iterations, max_iterations) : 0;
  std::vector<std::pair<int, int>> grow_events;  // (element, iteration) of every growth, for the log
  const double head_tol = opt ? opt->head_tol : 0.0;
  double resid = -1.0;
  int resid_row = -1;
  std::vector<double> rhist;
  int64_t nit_fp32 = 0;
  double best_resid = HUGE_VAL;
  solves32_ = 0;
  if (K_ < 1 || K_ > 2 * L_.M) fail(ANDFN_MI_EINVAL, "discharge_int %d outside [1, %lld]", K_, (long long)(2 * L_.M));
  if (max_coef + coef_increase - 1 > L_.M)
    fail(ANDFN_MI_ECAPACITY, "MAX_NCOEF %d + COEF_INCREASE %d exceeds the coefficient capacity %lld", max_coef,
         coef_increase, (long long)L_.M);
  for (int e = 0; e < E_; ++e)
    if ((etype_[e] == 1 || etype_[e] == 4) && radius_[e] <= 0)
      fail(ANDFN_MI_EINVAL, "element %d: circle radius must be positive", e);

  have_state_ = false;
  converged_ = false;
  restore();
  if (log_n > 0) d_hist_.alloc((size_t)log_n);
  const int nde = (int)de_.size();
  const int n = (int)N_;  // unknowns of the block layout (padding included)
  aem::NetView nv = net();
  aem::FitView fv = fit();

  // setup: tables, initial discharges, errors, branch cuts, matrix
  // (a later run from the same counts keeps the tables, the items and their device copies)
  if (build_targets(full_, nullptr, false)) {
    build_items(full_, items_, nullptr);
    d_items_.upload(items_, st_);
    ncitems_ = upload_circle_items(items_, d_citems_);
  }
  if (build_targets(rhs_, nullptr, true)) {
    std::vector<int> s0(E_, -1), s1(E_, -1);
    for (size_t s = 0; s < rhs_.side_elem.size(); ++s) (rhs_.side_which[s] ? s1 : s0)[rhs_.side_elem[s]] = (int)s;
    rhs_items_.clear();
    for (int e : de_) {
      if (s0[e] < 0 || (etype_[e] == 0 && s1[e] < 0))
        fail(ANDFN_MI_EINVAL, "discharge element %d is not listed by the fracture(s) it belongs to", e);
      rhs_items_.push_back(RhsItem{e, s0[e], etype_[e] == 0 ? s1[e] : -1});
    }
    d_rhs_items_.upload(rhs_items_, st_);
  }
  d_y_.upload(y0_, st_);
  d_yold_.alloc(n);
  d_b_.zero(n, st_);
  d_val_.alloc(std::max<int64_t>(full_.npts(), 1));
  d_valr_.alloc(std::max<int64_t>(rhs_.npts(), 1));
  // DFT sums of every collocation side (a grown subset indexes its own sides from 0)
  d_dft_.alloc(std::max<size_t>(full_.side_elem.size(), 1) * 2 * (size_t)M_);
  aem::launch_init_errors(fv, E_, st_);
  aem::launch_cross(nv, fv, d_circles_.p, (int)circles_.size(), st_);
  HIP_CHECK(hipGetLastError());
  auto t_setup = clk::now();
  assemble_and_factor();
  auto t_fact = clk::now();

  int nit = 0, cnt_error = 0;
  double error_q = 1.0, error = 1.0, error_coef = 1.0;
  int argmax = 0;
  double ms_sweep = 0, ms_rhs = 0, ms_fit = 0, ms_lu = 0, fl_sweep = 0, fl_rhs = 0;
  int64_t n_sweep = 0, n_rhs = 0, pr_sweep = 0, pr_rhs = 0, solve_passes = 0;
  std::vector<int> grow_h(E_);
  std::vector<char> grown(E_);
  aem::IterScalars sc{};
  // head rows of b from the current coefficients: the basis product, or a sweep of the discharge
  // points when no basis could be built
  const bool use_basis = ensure_rhs_basis(max_coef, coef_increase);
  int64_t pr_tmp = 0;
  double rhs_flops = set_flops(rhs_, false, &pr_tmp);
  int64_t rhs_pairs = pr_tmp;
  if (use_basis) {
    rhs_flops = 2.0 * (double)rb_entries_;  // one multiply-add per basis entry
    rhs_pairs = rb_entries_;
  }
  if (use_basis && mixed_phase_ && !rb32_ok_) {
    d_rbW32_.alloc((size_t)std::max<int64_t>(rb_entries_, 1));
    aem::launch_rhs_to_f32(rb_entries_, d_rbW_.p, d_rbW32_.p, st_);
    rb32_ok_ = true;
  }
  auto rhs_head = [&](bool mixed) {
    if (use_basis) {
      // the FP32 phase of the mixed form applies the basis rounded to FP32
      aem::launch_rhs_eval(nv, d_rhs_items_.p, nde, d_rbPtr_.p, d_rbW_.p, mixed ? d_rbW32_.p : nullptr, d_rbCol_.p,
                           d_rbCoff_.p, d_rbNan_.p, d_de_pos_.p, d_b_.p, st_);
    } else {
      sweep(aem::MODE_RHS, rhs_, d_valr_.p, mixed);
      aem::launch_rhs_reduce(nv, d_rhs_items_.p, nde, d_valr_.p, rhs_.d_pt_off.p, K_, d_de_pos_.p, d_b_.p, st_);
    }
  };

  auto sweep_full = [&](bool blend_now) {
    HIP_CHECK(hipEventRecord(ev_[0], st_));
    sweep(aem::MODE_COLLOC, full_, d_val_.p, mixed_phase_);
    HIP_CHECK(hipEventRecord(ev_[1], st_));
    aem::launch_fit(nv, fv, d_items_.p, (int)items_.size(), d_citems_.p, ncitems_, d_val_.p, full_.d_pt_off.p,
                    full_.d_dft_tiles.p,
                    (int)full_.dft_tiles.size(), full_.d_dft_sides.p, d_dft_.p, damping, blend_now ? 1 : 0, st_,
                    mixed_phase_);
    aem::launch_wells(nv, fv, d_wells_.p, (int)wells_.size(), damping, st_);
    HIP_CHECK(hipEventRecord(ev_[2], st_));
    int64_t pr = 0;
    fl_sweep += set_flops(full_, true, &pr);
    pr_sweep += pr;
    ++n_sweep;
  };
  bool grown_timed = false;
  // With a head tolerance (and no per-iteration log) the iteration scalars are not read back
  // every iteration: nothing launched depends on them (the stop test is the head residual, read
  // at the top of the next iteration), so the stream runs on and the scalars and the kernel
  // times of an iteration are collected at the next synchronisation.  The start event of the
  // right-hand side alternates between two events because the next iteration records its own
  // before that synchronisation.
  const bool lazy = head_tol > 0.0 && !verbose;
  int times_pending = 0;  // iteration whose kernel times are still to be collected (0: none)
  auto add_times = [&](int it) {
    hipEvent_t rhs0 = rhs_ev_[it & 1];
    float a = 0, b = 0, c = 0, d = 0;
    if (grown_timed) {
      float g = 0;
      HIP_CHECK(hipEventElapsedTime(&g, ev_[6], ev_[7]));
      ms_sweep += g;
      grown_timed = false;
    }
    HIP_CHECK(hipEventElapsedTime(&a, ev_[0], ev_[1]));
    HIP_CHECK(hipEventElapsedTime(&b, ev_[1], ev_[2]));
    HIP_CHECK(hipEventElapsedTime(&c, rhs0, sol_ev_[it & 1][0]));
    HIP_CHECK(hipEventElapsedTime(&d, sol_ev_[it & 1][0], sol_ev_[it & 1][1]));
    ms_sweep += a;
    ms_fit += b;
    ms_rhs += c;
    ms_lu += d;
  };

  // d_x_ = A^-1 d_b_, then y_old <- y, y <- x and q, constants <- x (one launch).  With the
  // sparse factors and A y of the current iterate at hand (the head residual's product), the
  // solve is the correction x = y + A^-1 (b - A y); otherwise a refined solve of b.
  auto solve_and_accept = [&](int it, bool have_ay) {
    HIP_CHECK(hipEventRecord(sol_ev_[it & 1][0], st_));
    if (have_ay && sparse_ && lu_mode_ == ANDFN_MI_LU_TRSV) {
      correct_rhs();
      ++solve_passes;
    } else {
      solve_rhs();
      solve_passes += sparse_ ? 2 : 1;
    }
    HIP_CHECK(hipEventRecord(sol_ev_[it & 1][1], st_));
    aem::launch_accept(nv, d_de_.p, nde, F_, sys(), d_x_.p, d_y_.p, d_yold_.p, n, st_);
  };
  // drop an accepted solve: y <- y_old, q and constants from it
  auto undo_accept = [&]() {
    HIP_CHECK(hipMemcpyAsync(d_y_.p, d_yold_.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st_));
    aem::launch_scatter(nv, d_de_.p, nde, F_, sys(), d_y_.p, st_);
  };
  // run-ahead iterations (see enqueue_iteration) after the first: head tolerance as the stop rule,
  // no per-iteration log, schedule or record, the basis right-hand side, the multifrontal
  // factors, and no element that can still grow
  bool runahead = lazy && log_n == 0 && !sched && !record && use_basis && sparse_ && lu_mode_ == ANDFN_MI_LU_TRSV &&
                  std::getenv("ANDFN_MI_NO_RUNAHEAD") == nullptr && max_iterations >= 2;
  for (int e = 0; e < E_ && runahead; ++e)
    if (etype_[e] != 2 && ncoef_[e] < max_coef) runahead = false;
  it_graph_launches_ = 0;
  int64_t runahead_iterations = 0;
  while (cnt_error < 2 && nit < max_iterations) {
    if (runahead && nit >= 1) {
      // iterations 2 .. : the host keeps one iteration enqueued ahead of the one whose head
      // residual it waits for; the device stops the run (iter_decide_kernel)
      ensure_iter_log(max_iterations + 2);
      *h_it_stop_ = 0;
      aem::launch_iter_reset(d_ctl_.p, nit, st_);
      aem::NetView nvg = nv;
      aem::FitView fvg = fv;
      nvg.stop = &d_ctl_.p->stop;
      fvg.stop = nvg.stop;
      std::vector<char> ph32(max_iterations + 2, 0), sv32(max_iterations + 2, 0);
      int launched = nit, done = nit;
      auto launch_next = [&]() {
        const int k = ++launched;
        ph32[k] = mixed_phase_;
        sv32[k] = solve32_;
        launch_iteration(nvg, fvg, mixed_phase_, solve32_, k % 3, head_tol, damping);
      };
      launch_next();
      int stopped_at = 0;
      while (done < launched) {
        const int k = done + 1;
        if (launched == k && k < max_iterations) launch_next();
        HIP_CHECK(hipEventSynchronize(it_end_ev_[k % 3]));
        HIP_CHECK(hipGetLastError());
        if (times_pending) {
          add_times(times_pending);  // iteration 1 (the host loop's events)
          times_pending = 0;
        }
        const int st_it = *(volatile int*)h_it_stop_;
        if (st_it > 0) {
          // the device stopped at iteration st_it <= k; later iterations left the state alone
          stopped_at = st_it;
          resid = h_it_resid_[st_it];
          resid_row = h_it_arg_[st_it];
          break;
        }
        resid = h_it_resid_[k];
        resid_row = h_it_arg_[k];
        done = k;
        if (solve32_) {
          // as in the host loop: a residual 4x above the best so far (or not finite) hands the
          // solves to the FP64 factors; the iteration enqueued ahead still used the FP32 ones
          if (!(resid == resid) || !std::isfinite(resid) || (k > 3 && resid > 4.0 * best_resid)) {
            if (verbose)
              fprintf(stderr, "mixed form: head residual %.3e at iteration %d (best %.3e); FP64 factors from here\n",
                      resid, k, best_resid);
            switch_to_fp64_solves();
            mixed_phase_ = false;
          }
          if (resid == resid) best_resid = std::min(best_resid, resid);
        }
        if (mixed_phase_) {
          rhist.push_back(resid);
          const size_t h = rhist.size();
          bool stalled = false;
          if (h >= 20) {
            const double last = *std::min_element(rhist.end() - 10, rhist.end());
            const double prev = *std::min_element(rhist.end() - 20, rhist.end() - 10);
            stalled = last > 0.7 * prev;
          }
          if (resid <= 30.0 * head_tol || stalled) mixed_phase_ = false;
        }
      }
      HIP_CHECK(hipStreamSynchronize(st_));
      HIP_CHECK(hipGetLastError());
      if (times_pending) {
        add_times(times_pending);
        times_pending = 0;
      }
      nit = stopped_at ? stopped_at - 1 : done;
      {
        // work of iterations 2 .. nit (the same right-hand side, solve and sweep every iteration)
        int64_t pr = 0;
        const double fs = set_flops(full_, true, &pr);
        for (int k = 2; k <= nit; ++k) {
          fl_rhs += rhs_flops;
          pr_rhs += rhs_pairs;
          ++n_rhs;
          ++solve_passes;
          fl_sweep += fs;
          pr_sweep += pr;
          ++n_sweep;
          if (sv32[k]) ++solves32_;
        }
      }
      if (nit >= 2) {
        // phase times of iterations 2 .. nit from the device's wall-clock stamps
        std::vector<unsigned long long> stv(6 * (size_t)(nit + 1));
        HIP_CHECK(hipMemcpy(stv.data(), d_stamps_.p, sizeof(unsigned long long) * stv.size(), hipMemcpyDeviceToHost));
        const double ms = 1.0 / wall_khz_;
        for (int k = 2; k <= nit; ++k) {
          const unsigned long long* t = stv.data() + 6 * (size_t)k;
          ms_rhs += (double)(t[1] - t[0]) * ms;
          ms_lu += (double)(t[2] - t[1]) * ms;
          ms_sweep += (double)(t[3] - t[2]) * ms;
          ms_fit += (double)(t[4] - t[3]) * ms;
        }
      }
      nit_fp32 = precision_ == ANDFN_MI_MIXED ? 1 : 0;  // iteration 1 runs in the FP32 phase
      for (int k = 2; k <= nit; ++k) nit_fp32 += ph32[k] ? 1 : 0;
      runahead_iterations = nit - 1;
      HIP_CHECK(hipMemcpyAsync(&sc, d_scal_.p, sizeof(sc), hipMemcpyDeviceToHost, st_));
      HIP_CHECK(hipStreamSynchronize(st_));
      error = sc.error;
      argmax = sc.argmax;
      error_coef = sc.error_coef;
      error_q = sc.dq_max / (sc.yold_max + 1e-16);
      if (!stopped_at && nit == max_iterations) {
        sweep_full(true);
        HIP_CHECK(hipStreamSynchronize(st_));
        float a = 0, b = 0;
        HIP_CHECK(hipEventElapsedTime(&a, ev_[0], ev_[1]));
        HIP_CHECK(hipEventElapsedTime(&b, ev_[1], ev_[2]));
        ms_sweep += a;
        ms_fit += b;
      }
      if (!stopped_at) resid = -1.0;  // not stopped on the head residual: recomputed below
      break;
    }
    ++nit;
    // ---- discharge solve (hpc_solve.py:155-174, 594-663)
    if (mixed_phase_) aem::launch_coef_to_f32(nv, E_, d_coef32_.p, st_);
    HIP_CHECK(hipEventRecord(rhs_ev_[nit & 1], st_));
    // only the head rows are written; the continuity and padding rows of b stay zero
    rhs_head(mixed_phase_);
    bool early = false, early32 = false;
    if (head_tol > 0.0 && nit >= 2) {
      // head residual of the previous iterate against the right-hand side of its coefficients.
      // This iteration's discharge solve is enqueued behind it before the host waits for the
      // value, so the GPU does not idle over the round trip; when the value ends the run or
      // hands the solve to the FP64 factors, the solve is dropped (y restored from y_old).
      head_residual_async(d_b_.p);
      early32 = solve32_;
      if (early32) ++solves32_;
      solve_and_accept(nit, true);
      early = true;
      HIP_CHECK(hipEventSynchronize(res_ev_));
      resid = h_scal_->residual;
      resid_row = h_scal_->arg_residual;
      if (times_pending) {
        add_times(times_pending);
        times_pending = 0;
      }
      if (solve32_) {
        // the FP32-factor refinement must contract: a residual 4x above the best so far (or not
        // finite) hands the solve to the FP64 factors for the rest of the run
        if (!(resid == resid) || !std::isfinite(resid) || (nit > 3 && resid > 4.0 * best_resid)) {
          if (verbose)
            fprintf(stderr, "mixed form: head residual %.3e at iteration %d (best %.3e); FP64 factors from here\n",
                    resid, nit, best_resid);
          if (early) {
            undo_accept();
            if (early32) --solves32_;
            early = false;
          }
          switch_to_fp64_solves();
          if (mixed_phase_) {
            mixed_phase_ = false;
            nit_fp32 = nit - 1;
          }
        }
        if (resid == resid) best_resid = std::min(best_resid, resid);
      }
      if (mixed_phase_) {
        // The FP32 phase ends near the tolerance, or when the best residual of the last 10
        // iterations is no better than 0.7 of the best of the 10 before (the FP32 floor).
        rhist.push_back(resid);
        const size_t h = rhist.size();
        bool stalled = false;
        if (h >= 20) {
          const double last = *std::min_element(rhist.end() - 10, rhist.end());
          const double prev = *std::min_element(rhist.end() - 20, rhist.end() - 10);
          stalled = last > 0.7 * prev;
        }
        if (resid <= 30.0 * head_tol || stalled) {
          mixed_phase_ = false;
          nit_fp32 = nit - 1;
        }
      } else if (resid <= head_tol) {
        if (early) {
          undo_accept();
          if (early32) --solves32_;
        }
        --nit;
        break;
      }
    }
    if (!early) {
      if (solve32_) ++solves32_;
      solve_and_accept(nit, false);
    }
    fl_rhs += rhs_flops;
    pr_rhs += rhs_pairs;
    ++n_rhs;
    // ---- element sweep, fit and first damping (hpc_solve.py:310-390)
    sweep_full(true);
    // ---- coefficient growth and second damping (hpc_solve.py:465-575)
    if (nit >= 2) {
      aem::GrowParams gp{max_error, coef_ratio, nit, max_coef, coef_increase};
      const int* dsched = nullptr;
      if (sched && nit - 1 < sched_n) {
        std::vector<int> row(sched + (size_t)(nit - 1) * E_, sched + (size_t)nit * E_);
        sched_row_.swap(row);
        d_sched_.upload(sched_row_, st_);
        dsched = d_sched_.p;
      }
      // an element grows only below MAX_NCOEF; when none is, the test (and its device-to-host
      // round trip) is skipped, with the same result
      int growable = 0;
      for (int e = 0; e < E_; ++e) growable += etype_[e] != 2 && ncoef_[e] < max_coef;
      if (growable > 0) {
        aem::launch_grow(nv, fv, E_, gp, dsched, d_grow_.p, st_);
        HIP_CHECK(hipMemcpyAsync(grow_h.data(), d_grow_.p, sizeof(int) * E_, hipMemcpyDeviceToHost, st_));
        HIP_CHECK(hipStreamSynchronize(st_));
      } else {
        std::fill(grow_h.begin(), grow_h.end(), 0);
      }
      int ng = 0;
      std::vector<int>& gcirc = gcirc_;
      gcirc.clear();
      for (int e = 0; e < E_; ++e) {
        grown[e] = (char)grow_h[e];
        if (grow_h[e]) {
          ++ng;
          if (log_n > 0) grow_events.emplace_back(e, nit);
          ncoef_[e] += coef_increase;
          nint_[e] = 2 * ncoef_[e];
          if (is_circle_like(etype_[e])) gcirc.push_back(e);
        }
      }
      if (ng) {
        n_grown_ += ng;
        build_targets(grown_, &grown, false);
        build_items(grown_, items_g_, &grown);
        d_items_g_.upload(items_g_, st_);
        ncitems_g_ = upload_circle_items(items_g_, d_citems_g_);
        d_valg_.alloc(std::max<int64_t>(grown_.npts(), 1));
        if (!gcirc.empty()) {
          d_grown_circles_.upload(gcirc, st_);
          aem::launch_cross(nv, fv, d_grown_circles_.p, (int)gcirc.size(), st_);
        }
        HIP_CHECK(hipEventRecord(ev_[6], st_));
        if (mixed_phase_) aem::launch_coef_to_f32(nv, E_, d_coef32_.p, st_);
        sweep(aem::MODE_COLLOC, grown_, d_valg_.p, mixed_phase_);
        HIP_CHECK(hipEventRecord(ev_[7], st_));
        grown_timed = true;
        ++n_sweep;
        aem::launch_fit(nv, fv, d_items_g_.p, (int)items_g_.size(), d_citems_g_.p, ncitems_g_, d_valg_.p,
                        grown_.d_pt_off.p,
                        grown_.d_dft_tiles.p, (int)grown_.dft_tiles.size(), grown_.d_dft_sides.p, d_dft_.p, damping, 0,
                        st_, mixed_phase_);
        int64_t pr = 0;
        fl_sweep += set_flops(grown_, true, &pr);
        pr_sweep += pr;
        // collocation tables for the grown counts, used from the next sweep on
        build_targets(full_, nullptr, false);
        d_val_.alloc(std::max<int64_t>(full_.npts(), 1));
      }
      aem::launch_blend(nv, E_, damping, st_);
    }
    aem::launch_reduce(fv, E_, d_y_.p, d_yold_.p, n, d_red_.p, d_scal_.p, st_);
    if (nit - 1 < log_n)
      HIP_CHECK(hipMemcpyAsync(d_hist_.p + (nit - 1), d_scal_.p, sizeof(aem::IterScalars), hipMemcpyDeviceToDevice, st_));
    if (record && nit - 1 < record_n) std::copy(ncoef_.begin(), ncoef_.end(), record + (size_t)(nit - 1) * E_);
    if (lazy && nit < max_iterations) {
      times_pending = nit;
      continue;
    }
    HIP_CHECK(hipMemcpyAsync(&sc, d_scal_.p, sizeof(sc), hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    HIP_CHECK(hipGetLastError());
    add_times(nit);
    error = sc.error;
    argmax = sc.argmax;
    error_coef = sc.error_coef;
    error_q = sc.dq_max / (sc.yold_max + 1e-16);
    if (verbose)
      fprintf(stderr, "Iteration: %02d, Error E: %.3e, Coef: %.3e, Q: %.3e, Element: %d, N of Coef: %d, Type: %d\n", nit,
              error, error_coef, error_q, argmax, ncoef_[argmax], etype_[argmax]);
    if (nit == max_iterations) {
      sweep_full(true);
      HIP_CHECK(hipStreamSynchronize(st_));
      float a = 0, b = 0;
      HIP_CHECK(hipEventElapsedTime(&a, ev_[0], ev_[1]));
      HIP_CHECK(hipEventElapsedTime(&b, ev_[1], ev_[2]));
      ms_sweep += a;
      ms_fit += b;
    }
    // AnDFN's stop rule (hpc_solve.py:237-239); with a head tolerance the head-residual test
    // at the top of the loop is the stop rule instead
    if (head_tol <= 0.0 && error_coef < max_error && error_q < max_error) {
      ++cnt_error;
      error_q = 1e30;
    }
  }
  HIP_CHECK(hipStreamSynchronize(st_));
  if (times_pending) {
    // the last complete iteration's kernel times and scalars (deferred, see above)
    add_times(times_pending);
    HIP_CHECK(hipMemcpyAsync(&sc, d_scal_.p, sizeof(sc), hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    HIP_CHECK(hipGetLastError());
    error = sc.error;
    argmax = sc.argmax;
    error_coef = sc.error_coef;
    error_q = sc.dq_max / (sc.yold_max + 1e-16);
    times_pending = 0;
  }
  auto t_loop = clk::now();
  const bool stopped_on_head = head_tol > 0.0 && resid >= 0.0 && resid <= head_tol;
  if (!stopped_on_head) {
    // right-hand side of the returned coefficients, then the residual of the returned discharges
    rhs_head(false);
    resid = head_residual(d_b_.p, &resid_row);
  }
  const double bcres = bc_residual();
  int status = 0;
  HIP_CHECK(hipMemcpyAsync(&status, d_status_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  if (status & 1) fail(ANDFN_MI_ECAPACITY, "branch-cut crossing list overflow");
  if (log_n > 0) {
    // AnDFN's log line per iteration: errors, the element of the largest error and its
    // coefficient count at that iteration (initial count plus the growths up to it)
    const int64_t rows = std::min<int64_t>(nit, log_n);
    std::vector<aem::IterScalars> hist((size_t)std::max<int64_t>(rows, 1));
    if (rows > 0)
      HIP_CHECK(hipMemcpyAsync(hist.data(), d_hist_.p, sizeof(aem::IterScalars) * rows, hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    std::sort(grow_events.begin(), grow_events.end());
    for (int64_t k = 0; k < rows; ++k) {
      const aem::IterScalars& h = hist[k];
      const int e = std::min(std::max(h.argmax, 0), E_ - 1);
      const auto lo = std::lower_bound(grow_events.begin(), grow_events.end(), std::make_pair(e, 0));
      const auto hi = std::upper_bound(grow_events.begin(), grow_events.end(), std::make_pair(e, (int)k + 1));
      double* row = log_out + k * ANDFN_MI_LOG_FIELDS;
      row[0] = h.error;
      row[1] = h.error_coef;
      row[2] = h.dq_max / (h.yold_max + 1e-16);
      row[3] = e;
      row[4] = ncoef0_[e] + (double)coef_increase * (double)(hi - lo);
      row[5] = etype_[e];
    }
  }
  auto t_end = clk::now();
  if (info) {
    info->nit = nit;
    info->n_unknowns = nde + F_;
    info->n_grown = n_grown_;
    info->max_error_element = argmax;
    info->error = error;
    info->error_coef = error_coef;
    info->error_q = error_q;
    info->t_setup = std::chrono::duration<double>(t_setup - t0).count();
    info->t_factor = std::chrono::duration<double>(t_fact - t_setup).count();
    info->t_iterate = std::chrono::duration<double>(t_loop - t_fact).count();
    info->t_total = std::chrono::duration<double>(t_end - t0).count();
    info->k_sweep_ms = ms_sweep;
    info->k_rhs_ms = ms_rhs;
    info->k_fit_ms = ms_fit;
    info->k_lu_ms = ms_lu;
    info->n_sweep = n_sweep;
    info->n_rhs = n_rhs;
    info->pairs_sweep = pr_sweep;
    info->pairs_rhs = pr_rhs;
    info->flops_sweep = fl_sweep;
    info->flops_rhs = fl_rhs;
    info->head_residual = resid;
    info->head_residual_row = resid_row;
    info->bc_residual = bcres;
    info->nit_fp32 = precision_ == ANDFN_MI_MIXED ? (mixed_phase_ ? nit : nit_fp32) : 0;
    info->solves_fp32 = solves32_;
    info->refine_rate = precision_ == ANDFN_MI_MIXED ? refine_rate_ : 0.0;
    info->sweep_points_per_lane = ppl_[0];
    info->rhs_points_per_lane = ppl_[1];
    info->sweep_persistent = persist_[0];
    info->rhs_persistent = persist_[1];
    info->nnz_discharge = nnz_A_;
    info->nnz_factors = nnz_factors_;
    info->factor_used = sparse_ ? ANDFN_MI_FACTOR_SPARSE
                                : (lu_mode_ == ANDFN_MI_LU_TRSV && factor_ != ANDFN_MI_FACTOR_ROCSOLVER
                                       ? ANDFN_MI_FACTOR_TOURNAMENT
                                       : ANDFN_MI_FACTOR_ROCSOLVER);
    info->factor_rounds = sparse_ ? mf_rounds_ : 0;
    info->solve_passes = solve_passes;
    info->solve_values = sparse_ ? mf_.symbolic().gv_total : nnz_factors_;
    info->runahead_iterations = runahead_iterations;
    info->graph_iterations = it_graph_launches_;
  }
  mixed_phase_ = false;
  have_state_ = true;
  converged_ = cnt_error >= 2 || stopped_on_head;
  return converged_ ? ANDFN_MI_OK : ANDFN_MI_NOT_CONVERGED;
}

// copy the device state back into the caller's records (fields the reference mutates)
void Solver::write_back() {
  HIP_CHECK(hipMemcpyAsync(coef_.data(), d_coef_.p, sizeof(double2) * coef_.size(), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(dpsi_.data(), d_dpsi_.p, sizeof(double) * dpsi_.size(), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(q_.data(), d_q_.p, sizeof(double) * E_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(constant_.data(), d_constant_.p, sizeof(double) * F_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(ncoef_.data(), d_ncoef_.p, sizeof(int) * E_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(nint_.data(), d_nint_.p, sizeof(int) * E_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(error_.data(), d_error_.p, sizeof(double) * E_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(error_old_.data(), d_error_old_.p, sizeof(double) * E_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(error_old2_.data(), d_error_old2_.p, sizeof(double) * E_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(error_coef_.data(), d_error_coef_.p, sizeof(double) * E_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  const double PI = 3.141592653589793;
  for (int i = 0; i < E_; ++i) {
    char* p = el_ + (size_t)i * L_.esz;
    wr<int64_t>(p, 136, ncoef_[i]);
    wr<int64_t>(p, 144, nint_[i]);
    wr<double>(p, 152, q_[i]);
    for (int64_t j = 0; j < Mrec_; ++j) {
      double2 c = coef_[(size_t)i * M_ + j];
      wr<cd>(p, L_.e_coef + 16 * j, cd(c.x, c.y));
    }
    // thetas[:nint] as calc_thetas writes them (hpc_math_functions.py:673-699)
    const int n = nint_[i];
    double del = PI / n, start = del / 2;
    if (etype_[i] == 1) {
      del = 2 * PI / n;
      start = 0.0;
    }
    for (int k = 0; k < n; ++k) wr<double>(p, L_.e_thetas + 8 * k, start + k * del);
    if (is_circle_like(etype_[i]))
      for (int k = 0; k < n - 1; ++k) wr<double>(p, L_.e_dpsi + 8 * k, dpsi_[(size_t)i * 2 * M_ + k]);
    wr<double>(p, L_.e_err, error_[i]);
    wr<double>(p, L_.e_err + 8, error_old_[i]);
    wr<double>(p, L_.e_err + 16, error_old2_[i]);
    wr<double>(p, L_.e_err + 24, error_coef_[i]);
  }
  for (int f = 0; f < F_; ++f) wr<double>(fr_ + (size_t)f * L_.fsz, L_.f_const, constant_[f]);
}

void Solver::heads(int64_t npts, const int64_t* frac, const double* z, double* out, int what, int64_t exclude) {
  // group the points by fracture into tiles, evaluate Re Omega / t (hpc_fracture.py:204-250), the
  // complex potential Omega itself (calc_omega hpc_fracture.py:51-92) or the discharge vector w
  // (hpc_fracture.py:139-174 / fracture.py:453-477).  Points already grouped by fracture (a flow-net
  // pattern, fracture after fracture) go to the device and back as they are; others are grouped by
  // a counting sort that keeps their order within a fracture.
  if (npts > (int64_t)INT32_MAX) fail(ANDFN_MI_EINVAL, "%lld points exceed the limit of one call", (long long)npts);
  bool grouped = true;
  for (int64_t i = 0; i < npts; ++i) {
    if (frac[i] < 0 || frac[i] >= F_) fail(ANDFN_MI_EINVAL, "point %lld: fracture %lld out of range", (long long)i, (long long)frac[i]);
    if (i > 0 && frac[i] < frac[i - 1]) grouped = false;
  }
  std::vector<int64_t> order;
  if (!grouped) {
    std::vector<int64_t> pos(F_ + 1, 0);
    for (int64_t i = 0; i < npts; ++i) ++pos[frac[i] + 1];
    for (int f = 0; f < F_; ++f) pos[f + 1] += pos[f];
    order.resize(npts);
    for (int64_t i = 0; i < npts; ++i) order[pos[frac[i]]++] = i;
  }
  auto at = [&](int64_t k) { return grouped ? k : order[k]; };
  std::vector<Tile> tiles;
  for (int64_t k = 0; k < npts;) {
    const int64_t f = frac[at(k)];
    const int64_t b = k;
    while (k < npts && frac[at(k)] == f) ++k;
    for (int64_t p0 = b; p0 < k; p0 += 64) tiles.push_back(Tile{(int)f, 0, 0, (int)p0, (int)k});
  }
  DBuf<double2> dz;
  DBuf<Tile> dt;
  DBuf<double> dv;
  dz.alloc(std::max<int64_t>(npts, 1));
  if (grouped) {
    if (npts > 0) HIP_CHECK(hipMemcpyAsync(dz.p, z, sizeof(double2) * npts, hipMemcpyHostToDevice, st_));
  } else {
    std::vector<double2> zs(npts);
    for (int64_t k = 0; k < npts; ++k) zs[k] = make_double2(z[2 * order[k]], z[2 * order[k] + 1]);
    HIP_CHECK(hipMemcpyAsync(dz.p, zs.data(), sizeof(double2) * npts, hipMemcpyHostToDevice, st_));
    HIP_CHECK(hipStreamSynchronize(st_));  // zs leaves scope
  }
  dt.upload(tiles, st_);
  const int w = what == 0 ? 1 : 2;
  dv.alloc(std::max<int64_t>(w * npts, 1));
  aem::TargetView tv{};
  tv.tiles = dt.p;
  tv.pt_z = dz.p;
  if (what >= 2) {
    const int ex = (exclude >= 0 && exclude < E_) ? (int)exclude : -1;
    aem::launch_w_points(net(), tv, (int)tiles.size(), ex, what == 2 ? 1 : 0, dv.p, st_);
  } else {
    aem::launch_sweep(what == 1 ? aem::MODE_POINTS_OMEGA : aem::MODE_POINTS, net(), tv, (int)tiles.size(), 1, dv.p,
                      st_);
  }
  HIP_CHECK(hipGetLastError());
  if (grouped) {
    if (npts > 0) HIP_CHECK(hipMemcpyAsync(out, dv.p, sizeof(double) * w * npts, hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    return;
  }
  std::vector<double> v(w * npts);
  HIP_CHECK(hipMemcpyAsync(v.data(), dv.p, sizeof(double) * w * npts, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  for (int64_t k = 0; k < npts; ++k)
    for (int c = 0; c < w; ++c) out[w * order[k] + c] = v[w * k + c];
}

void Solver::streamlines(int64_t n, const int64_t* frac, const double* z0, const double* elevation,
                         const double* aperture, double ds, int backward, int64_t max_length, int64_t* n_points,
                         int64_t* n_segments, int64_t* exit_element) {
  // DFN.streamline_tracking (dfn.py:2638-2725) for every line at once.  Each launch advances every
  // unfinished line of a chunk as far as its output row allows (kSegRow segments, the point budget
  // plus three points per segment); the rows are packed on the device and appended to the lines'
  // host results, and lines that ran out of row continue in the next launch.
  constexpr int kSegRow = 64;
  constexpr int64_t kChunkBytes = int64_t(1) << 30;  // output rows of one launch
  if (n < 0 || n > (int64_t)INT32_MAX / 2) fail(ANDFN_MI_EINVAL, "%lld lines: outside [0, %d]", (long long)n, INT32_MAX / 2);
  if (max_length < 0 || max_length > (int64_t)1 << 24)
    fail(ANDFN_MI_EINVAL, "max_length %lld outside [0, 2^24]", (long long)max_length);
  if (!(ds > 0.0) || !std::isfinite(ds)) fail(ANDFN_MI_EINVAL, "step %g is not a positive number", ds);
  for (int64_t i = 0; i < n; ++i)
    if (frac[i] < 0 || frac[i] >= F_) fail(ANDFN_MI_EINVAL, "line %lld: fracture %lld out of range", (long long)i, (long long)frac[i]);
  for (int f = 0; f < F_; ++f)
    if (!(aperture[f] > 0.0) || !std::isfinite(aperture[f]))
      fail(ANDFN_MI_EINVAL, "fracture %d: aperture %g is not a positive number", f, aperture[f]);
  sl_pts_.assign(n, {});
  sl_vel_.assign(n, {});
  sl_sfrac_.assign(n, {});
  sl_slen_.assign(n, {});
  if (n == 0) return;
  // a line ends after max(max_length, 1) + 1 segments (the reference has no bound: a line that
  // crossed an intersection at every first step would never end)
  const int max_seg = (int)std::min<int64_t>(std::max<int64_t>(max_length, 1) + 1, INT32_MAX);
  const int64_t cap = std::max<int64_t>(max_length + 2, 3) + 3 * kSegRow;
  const int64_t row_bytes = cap * (int64_t)(sizeof(double2) + sizeof(double)) + kSegRow * 8 + 16;
  const int64_t chunk = std::max<int64_t>(64, std::min<int64_t>(n, kChunkBytes / row_bytes));
  // lines grouped by fracture: a wavefront starts in one fracture's elements
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return frac[a] < frac[b]; });
  std::vector<int> hf(n), hseg(n, 0), hdone(n, 0), hex(n, -1);
  std::vector<double2> hz(n);
  std::vector<double> hel(n);
  std::vector<int64_t> hlen(n, 0);
  for (int64_t k = 0; k < n; ++k) {
    const int64_t i = order[k];
    hf[k] = (int)frac[i];
    hz[k] = make_double2(z0[2 * i], z0[2 * i + 1]);
    hel[k] = elevation[i];
  }
  DBuf<int> d_f, d_seg, d_done, d_ex, d_act, d_sfrac, d_slen, d_osegs, d_psf, d_psl;
  DBuf<double2> d_z0, d_zs, d_pts, d_ppts;
  DBuf<double> d_el, d_vel, d_pvel, d_rad, d_ap, d_frame;
  DBuf<int64_t> d_len, d_opts, d_poff, d_soff;
  d_f.upload(hf, st_);
  d_z0.upload(hz, st_);
  d_zs.upload(hz, st_);
  d_el.upload(hel, st_);
  d_len.upload(hlen, st_);
  d_seg.upload(hseg, st_);
  d_done.upload(hdone, st_);
  d_ex.upload(hex, st_);
  d_rad.upload(fradius_, st_);
  d_ap.upload(std::vector<double>(aperture, aperture + F_), st_);
  d_frame.upload(fframe_, st_);
  d_pts.alloc((size_t)chunk * cap);
  d_vel.alloc((size_t)chunk * cap);
  d_sfrac.alloc((size_t)chunk * kSegRow);
  d_slen.alloc((size_t)chunk * kSegRow);
  d_opts.alloc(chunk);
  d_osegs.alloc(chunk);
  d_poff.alloc(chunk);
  d_soff.alloc(chunk);
  std::vector<int64_t> h_opts(chunk), h_poff(chunk), h_soff(chunk);
  std::vector<int> h_osegs(chunk), act;
  std::vector<double2> h_pts;
  std::vector<double> h_vel;
  std::vector<int> h_sf, h_sl;
  for (int64_t c0 = 0; c0 < n; c0 += chunk) {
    const int64_t c1 = std::min(n, c0 + chunk);
    act.clear();
    for (int64_t k = c0; k < c1; ++k) act.push_back((int)k);
    while (!act.empty()) {
      const int na = (int)act.size();
      d_act.upload(act, st_);
      aem::TraceArgs ta{};
      ta.n = na;
      ta.act = d_act.p;
      ta.frac = d_f.p;
      ta.z0 = d_z0.p;
      ta.z_start = d_zs.p;
      ta.elev = d_el.p;
      ta.length = d_len.p;
      ta.nseg = d_seg.p;
      ta.done = d_done.p;
      ta.exit_elem = d_ex.p;
      ta.fradius = d_rad.p;
      ta.aperture = d_ap.p;
      ta.frame = d_frame.p;
      ta.ds = ds;
      ta.backward = backward ? 1 : 0;
      ta.max_segments = max_seg;
      ta.max_length = max_length;
      ta.capacity = cap;
      ta.seg_capacity = kSegRow;
      ta.pts = d_pts.p;
      ta.vel = d_vel.p;
      ta.seg_frac = d_sfrac.p;
      ta.seg_len = d_slen.p;
      ta.out_pts = d_opts.p;
      ta.out_segs = d_osegs.p;
      aem::launch_trace(net(), ta, st_);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(h_opts.data(), d_opts.p, sizeof(int64_t) * na, hipMemcpyDeviceToHost, st_));
      HIP_CHECK(hipMemcpyAsync(h_osegs.data(), d_osegs.p, sizeof(int) * na, hipMemcpyDeviceToHost, st_));
      HIP_CHECK(hipStreamSynchronize(st_));
      int64_t tp = 0, tsg = 0;
      for (int k = 0; k < na; ++k) {
        h_poff[k] = tp;
        h_soff[k] = tsg;
        tp += h_opts[k];
        tsg += h_osegs[k];
      }
      if (tp > 0) {
        d_ppts.alloc(tp);
        d_pvel.alloc(tp);
        d_psf.alloc(std::max<int64_t>(tsg, 1));
        d_psl.alloc(std::max<int64_t>(tsg, 1));
        HIP_CHECK(hipMemcpyAsync(d_poff.p, h_poff.data(), sizeof(int64_t) * na, hipMemcpyHostToDevice, st_));
        HIP_CHECK(hipMemcpyAsync(d_soff.p, h_soff.data(), sizeof(int64_t) * na, hipMemcpyHostToDevice, st_));
        aem::TracePack pk{};
        pk.n = na;
        pk.capacity = cap;
        pk.seg_capacity = kSegRow;
        pk.pts = d_pts.p;
        pk.vel = d_vel.p;
        pk.seg_frac = d_sfrac.p;
        pk.seg_len = d_slen.p;
        pk.out_pts = d_opts.p;
        pk.out_segs = d_osegs.p;
        pk.poff = d_poff.p;
        pk.soff = d_soff.p;
        pk.ppts = d_ppts.p;
        pk.pvel = d_pvel.p;
        pk.pseg_frac = d_psf.p;
        pk.pseg_len = d_psl.p;
        aem::launch_trace_pack(pk, st_);
        HIP_CHECK(hipGetLastError());
        h_pts.resize(tp);
        h_vel.resize(tp);
        h_sf.resize(tsg);
        h_sl.resize(tsg);
        HIP_CHECK(hipMemcpyAsync(h_pts.data(), d_ppts.p, sizeof(double2) * tp, hipMemcpyDeviceToHost, st_));
        HIP_CHECK(hipMemcpyAsync(h_vel.data(), d_pvel.p, sizeof(double) * tp, hipMemcpyDeviceToHost, st_));
        if (tsg > 0) {
          HIP_CHECK(hipMemcpyAsync(h_sf.data(), d_psf.p, sizeof(int) * tsg, hipMemcpyDeviceToHost, st_));
          HIP_CHECK(hipMemcpyAsync(h_sl.data(), d_psl.p, sizeof(int) * tsg, hipMemcpyDeviceToHost, st_));
        }
      }
      HIP_CHECK(hipMemcpyAsync(hdone.data(), d_done.p, sizeof(int) * n, hipMemcpyDeviceToHost, st_));
      HIP_CHECK(hipStreamSynchronize(st_));
      std::vector<int> next;
      for (int k = 0; k < na; ++k) {
        const int64_t i = order[act[k]];
        sl_pts_[i].insert(sl_pts_[i].end(), h_pts.begin() + h_poff[k], h_pts.begin() + h_poff[k] + h_opts[k]);
        sl_vel_[i].insert(sl_vel_[i].end(), h_vel.begin() + h_poff[k], h_vel.begin() + h_poff[k] + h_opts[k]);
        sl_sfrac_[i].insert(sl_sfrac_[i].end(), h_sf.begin() + h_soff[k], h_sf.begin() + h_soff[k] + h_osegs[k]);
        sl_slen_[i].insert(sl_slen_[i].end(), h_sl.begin() + h_soff[k], h_sl.begin() + h_soff[k] + h_osegs[k]);
        if (!hdone[act[k]]) {
          // a fresh row always takes the next segment, so a line that wrote nothing is stuck
          if (h_opts[k] == 0 && h_osegs[k] == 0) fail(ANDFN_MI_EHIP, "streamline %lld made no progress", (long long)i);
          next.push_back(act[k]);
        }
      }
      act.swap(next);
    }
  }
  HIP_CHECK(hipMemcpyAsync(hex.data(), d_ex.p, sizeof(int) * n, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  for (int64_t k = 0; k < n; ++k) {
    const int64_t i = order[k];
    n_points[i] = (int64_t)sl_pts_[i].size();
    n_segments[i] = (int64_t)sl_sfrac_[i].size();
    exit_element[i] = hex[k];
  }
}

void Solver::streamlines_fetch(double* points, double* speed, int64_t* segment_fracture, int64_t* segment_points) {
  int64_t p = 0, q = 0;
  for (size_t i = 0; i < sl_pts_.size(); ++i) {
    for (size_t j = 0; j < sl_pts_[i].size(); ++j, ++p) {
      points[2 * p] = sl_pts_[i][j].x;
      points[2 * p + 1] = sl_pts_[i][j].y;
      speed[p] = sl_vel_[i][j];
    }
    for (size_t j = 0; j < sl_sfrac_[i].size(); ++j, ++q) {
      segment_fracture[q] = sl_sfrac_[i][j];
      segment_points[q] = sl_slen_[i][j];
    }
  }
}

// ---- heads at points without allocation (andfn_mi_state_heads_noalloc) ----------------------
// Device scratch: the points' coordinates, their heads and the 64-point tiles (at most one tile
// per point), in a region given by the caller of heads_noalloc; the host staging arrays below are
// static (.bss), so nothing is allocated on either side.  One call at a time (a mutex in the entry
// point): every call shares them.
constexpr int64_t kNoAllocMaxPoints = 65536;
constexpr size_t kNoAllocVal = 16 * kNoAllocMaxPoints;                          // double2 z[max] at 0
constexpr size_t kNoAllocTiles = kNoAllocVal + 8 * kNoAllocMaxPoints;           // double val[max]
constexpr size_t kNoAllocBytes = kNoAllocTiles + sizeof(Tile) * kNoAllocMaxPoints;  // Tile tiles[max]
static_assert(kNoAllocVal % 256 == 0 && kNoAllocTiles % 256 == 0, "scratch sections stay 256-byte aligned");
int64_t g_na_order[kNoAllocMaxPoints];
double2 g_na_z[kNoAllocMaxPoints];
Tile g_na_tiles[kNoAllocMaxPoints];
double g_na_val[kNoAllocMaxPoints];

void Solver::heads_noalloc(int64_t npts, const int64_t* frac, const double* z, double* out,
                           unsigned char* scratch_dev) {
  if (npts < 0 || npts > kNoAllocMaxPoints)
    fail(ANDFN_MI_ECAPACITY, "%lld points exceed the allocation-free limit of %lld per call", (long long)npts,
         (long long)kNoAllocMaxPoints);
  if (npts == 0) return;
  for (int64_t i = 0; i < npts; ++i) {
    if (frac[i] < 0 || frac[i] >= F_) fail(ANDFN_MI_EINVAL, "point %lld: fracture %lld out of range", (long long)i, (long long)frac[i]);
    g_na_order[i] = i;
  }
  // points grouped by fracture, in input order within a fracture (the order heads() uses); std::sort
  // on the (fracture, index) pair needs no buffer, unlike a stable sort
  std::sort(g_na_order, g_na_order + npts,
            [&](int64_t a, int64_t b) { return frac[a] != frac[b] ? frac[a] < frac[b] : a < b; });
  int ntiles = 0;
  for (int64_t k = 0; k < npts;) {
    const int64_t f = frac[g_na_order[k]];
    const int64_t b = k;
    while (k < npts && frac[g_na_order[k]] == f) {
      g_na_z[k] = make_double2(z[2 * g_na_order[k]], z[2 * g_na_order[k] + 1]);
      ++k;
    }
    for (int64_t p0 = b; p0 < k; p0 += 64) g_na_tiles[ntiles++] = Tile{(int)f, 0, 0, (int)p0, (int)k};
  }
  double2* dz = (double2*)scratch_dev;
  double* dv = (double*)(scratch_dev + kNoAllocVal);
  Tile* dt = (Tile*)(scratch_dev + kNoAllocTiles);
  HIP_CHECK(hipMemcpyAsync(dz, g_na_z, sizeof(double2) * npts, hipMemcpyHostToDevice, st_));
  HIP_CHECK(hipMemcpyAsync(dt, g_na_tiles, sizeof(Tile) * ntiles, hipMemcpyHostToDevice, st_));
  aem::TargetView tv{};
  tv.tiles = dt;
  tv.pt_z = dz;
  aem::launch_sweep(aem::MODE_POINTS, net(), tv, ntiles, 1, dv, st_);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(g_na_val, dv, sizeof(double) * npts, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  for (int64_t k = 0; k < npts; ++k) out[g_na_order[k]] = g_na_val[k];
}

// ---------------------------------------------------------------------------------------------
// Implicit adjoint
// ---------------------------------------------------------------------------------------------

void Solver::adj_tables(const TargetSet& ts, bool colloc, AdjTables& t) {
  const int64_t np = ts.npts();
  t.pz.alloc(std::max<int64_t>(np, 1));
  t.pelem.alloc(std::max<int64_t>(np, 1));
  aem::launch_points(net(), ts.view(), ts.ntiles(1), K_, colloc ? 1 : 0, t.pz.p, t.pelem.p, st_);
  std::vector<int> b(F_), e(F_);
  for (int f = 0; f < F_; ++f) {
    b[f] = ts.pt_off[ts.fsb[f]];
    e[f] = ts.pt_off[ts.fsb[f + 1]];
  }
  t.fpb.upload(b, st_);
  t.fpe.upload(e, st_);
}

// Transposed factors for A^-T = R P^T L^-T U^-T: the LU factors with the panel products undone
// (L[below j, j] = G_j L_jj, U[above j, j] = H_j U_jj), transposed, and prepared for the same
// blocked solves with a non-unit lower (U^T) and a unit upper (L^T) triangle.
template <class T>
void Solver::make_transposed(const T* A, bool blocked, DBuf<T>& AT, DBuf<T>& linvT, DBuf<T>& uinvT) {
  const size_t tot = (size_t)Atot_;
  {
    DBuf<T> tmp;
    tmp.alloc(tot);
    HIP_CHECK(hipMemcpyAsync(tmp.p, A, sizeof(T) * tot, hipMemcpyDeviceToDevice, st_));
    AT.alloc(tot);
    for (const auto& c : cls_) {
      const int64_t n = c.npad, nb = (n + kNB - 1) / kNB, sA = n * n;
      T* tc = tmp.p + c.a0;
      for (int64_t j = 0; blocked && j < nb; ++j) {
        const int64_t r0 = j * kNB;
        const int bs = (int)std::min<int64_t>(kNB, n - r0);
        const int rest = (int)(n - r0 - bs);
        if (c.nblk > 1) {
          // several systems of the same (padded) size: one call per panel covers them all
          const T* djj = tc + r0 * (n + 1);
          if (rest > 0)
            RB_CHECK(rb_trmm_right_batched(rb_, rocblas_fill_lower, rocblas_diagonal_unit, rest, bs, djj, (int)n, sA,
                                           tc + r0 * n + r0 + bs, n, sA, c.nblk));
          if (r0 > 0)
            RB_CHECK(rb_trmm_right_batched(rb_, rocblas_fill_upper, rocblas_diagonal_non_unit, (int)r0, bs, djj,
                                           (int)n, sA, tc + r0 * n, n, sA, c.nblk));
        } else {
          const T* djj = tc + r0 * (n + 1);
          if (rest > 0)
            RB_CHECK(rb_trmm_right(rb_, rocblas_fill_lower, rocblas_diagonal_unit, rest, bs, djj, (int)n,
                                   tc + r0 * n + r0 + bs, n));
          if (r0 > 0)
            RB_CHECK(rb_trmm_right(rb_, rocblas_fill_upper, rocblas_diagonal_non_unit, (int)r0, bs, djj, (int)n,
                                   tc + r0 * n, n));
        }
      }
      aem::launch_transpose<T>(tc, n, AT.p + c.a0, st_, c.nblk);
    }
    HIP_CHECK(hipStreamSynchronize(st_));
  }
  prepare_blocked_all<T>(AT.p, linvT, uinvT, false);
  HIP_CHECK(hipStreamSynchronize(st_));
}

// CSR of A^T (for the FP64 residual z - A^T mu of the transposed refinement)
void Solver::make_csr_transpose() {
  const int n = (int)N_;
  std::vector<int> rp(n + 1), col(nnz_A_);
  std::vector<double> val(nnz_A_);
  HIP_CHECK(hipMemcpyAsync(rp.data(), d_csr_ptr_.p, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(col.data(), d_csr_col_.p, sizeof(int) * nnz_A_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(val.data(), d_csr_val_.p, sizeof(double) * nnz_A_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  std::vector<int> tp(n + 1, 0), tc(nnz_A_);
  std::vector<double> tv(nnz_A_);
  for (int64_t k = 0; k < nnz_A_; ++k) ++tp[col[k] + 1];
  for (int i = 0; i < n; ++i) tp[i + 1] += tp[i];
  std::vector<int> fill(tp.begin(), tp.end() - 1);
  for (int r = 0; r < n; ++r)
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int d = fill[col[k]]++;
      tc[d] = r;
      tv[d] = val[k];
    }
  d_csrT_ptr_.upload(tp, st_);
  d_csrT_col_.upload(tc, st_);
  d_csrT_val_.upload(tv, st_);
}

// mu += A^-T (z - A^T mu) with the multifrontal factors: A^-T = R (R A)^-T, the solve returning
// (R A)^-T r and the update applying R
void Solver::adj_mu_step_sparse() {
  const int n = (int)N_;
  aem::launch_spmv(n, d_csrT_ptr_.p, d_csrT_col_.p, d_csrT_val_.p, d_mu_.p, d_aty_.p, st_);
  aem::launch_adj_refine_in<double>(d_z_.p, d_aty_.p, n, d_x_.p, st_);
  mf_.solve_t(d_x_.p, nullptr, d_v_.p, st_);
  aem::launch_adj_refine_out<double>(d_v_.p, d_perm_.p, d_req_.p, n, d_mu_.p, st_);
}

// mu += A^-T (z - A^T mu) with the transposed factors in precision T
template <class T>
void Solver::adj_mu_step(const T* AT, const T* linvT, const T* uinvT, T* x, T* w, T* v) {
  const int n = (int)N_;
  aem::launch_spmv(n, d_csrT_ptr_.p, d_csrT_col_.p, d_csrT_val_.p, d_mu_.p, d_aty_.p, st_);
  aem::launch_adj_refine_in<T>(d_z_.p, d_aty_.p, n, x, st_);
  blocked_solve_all<T>(AT, linvT, uinvT, x, w, v);
  aem::launch_adj_refine_out<T>(v, d_perm_.p, d_req_.p, n, d_mu_.p, st_);
}

int Solver::adjoint(const andfn_mi_objective* obj, const andfn_mi_adjoint_options* opt, double* grad_t,
                    double* grad_head, andfn_mi_adjoint_info* info) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  if (!have_state_) fail(ANDFN_MI_EINVAL, "no solved state on this handle: run a solve first");
  if (!converged_ && !(opt && opt->accept_unconverged))
    fail(ANDFN_MI_EINVAL,
         "the last run stopped at MAX_ITERATIONS without meeting its stop rule: the implicit adjoint is the gradient "
         "of the fixed point (set accept_unconverged to use this state anyway)");
  if (lu_mode_ != ANDFN_MI_LU_TRSV) fail(ANDFN_MI_EINVAL, "the adjoint uses the blocked triangular solves (lu_mode 0)");
  if (!grad_t || !grad_head) fail(ANDFN_MI_EINVAL, "null gradient pointer");
  const double tol = (opt && opt->tol > 0) ? opt->tol : 1e-10;
  const int64_t maxit = (opt && opt->max_iterations > 0) ? opt->max_iterations : 5000;
  const double damp = (opt && opt->damping > 0) ? opt->damping : 0.75;
  if (damp > 1.0) fail(ANDFN_MI_EINVAL, "damping %g outside (0, 1]", damp);
  const int verbose = opt ? opt->verbose : 0;
  const int64_t nobs = obj ? obj->n_obs : 0;
  if (nobs < 0 || (nobs > 0 && (!obj->obs_frac || !obj->obs_z || !obj->dhead)))
    fail(ANDFN_MI_EINVAL, "observation arrays missing");
  for (int64_t k = 0; k < nobs; ++k)
    if (obj->obs_frac[k] < 0 || obj->obs_frac[k] >= F_)
      fail(ANDFN_MI_EINVAL, "observation %lld: fracture %lld out of range", (long long)k, (long long)obj->obs_frac[k]);
  const int nde = (int)de_.size(), n = (int)N_;
  const size_t ncf = (size_t)E_ * M_;
  aem::NetView nv = net();

  // point tables of the collocation and discharge point sets at the converged counts
  adj_tables(full_, true, tab_full_);
  adj_tables(rhs_, false, tab_rhs_);
  // source sides of every element as (slot, fracture)
  std::vector<int> eptr(E_ + 1, 0);
  for (const Slot& sl : slots_) ++eptr[sl.elem + 1];
  for (int e = 0; e < E_; ++e) eptr[e + 1] += eptr[e];
  std::vector<int2> esl(std::max<size_t>(slots_.size(), 1));
  {
    std::vector<int> fill(eptr.begin(), eptr.end() - 1);
    for (int f = 0; f < F_; ++f)
      for (int k = mem_ptr_[f]; k < mem_ptr_[f + 1]; ++k) esl[fill[slots_[k].elem]++] = make_int2(k, f);
  }
  d_eslot_ptr_.upload(eptr, st_);
  d_eslot_.upload(esl, st_);
  // gather order: heaviest sources (most target points on their fractures) first
  std::vector<int64_t> work(E_, 0);
  for (int e = 0; e < E_; ++e)
    for (int k = eptr[e]; k < eptr[e + 1]; ++k) {
      const int f = esl[k].y;
      work[e] += full_.pt_off[full_.fsb[f + 1]] - full_.pt_off[full_.fsb[f]];
    }
  // (sources with one coefficient component first, then bounding and impermeable circles)
  std::vector<int> gitems(E_);
  for (int e = 0; e < E_; ++e) gitems[e] = e;
  auto both = [&](int e) { return etype_[e] == 1 || etype_[e] == 4; };
  std::stable_sort(gitems.begin(), gitems.end(), [&](int a, int b) {
    if (both(a) != both(b)) return !both(a);
    return work[a] > work[b];
  });
  const int gsingle = (int)std::count_if(gitems.begin(), gitems.end(), [&](int e) { return !both(e); });
  d_gitems_.upload(gitems, st_);
  // branch-cut crossings regrouped by crossing source: (circle element, position, sign)
  int nxsrc = 0;
  {
    std::vector<int> ccnt(E_);
    std::vector<int2> cr(std::max(cross_total_, 1));
    std::vector<int> cs(std::max(cross_total_, 1));
    HIP_CHECK(hipMemcpyAsync(ccnt.data(), d_cross_cnt_.p, sizeof(int) * E_, hipMemcpyDeviceToHost, st_));
    if (cross_total_ > 0) {
      HIP_CHECK(hipMemcpyAsync(cr.data(), d_cross_.p, sizeof(int2) * cross_total_, hipMemcpyDeviceToHost, st_));
      HIP_CHECK(hipMemcpyAsync(cs.data(), d_cross_sign_.p, sizeof(int) * cross_total_, hipMemcpyDeviceToHost, st_));
    }
    HIP_CHECK(hipStreamSynchronize(st_));
    std::vector<std::vector<int4>> by(E_);
    for (int e : circles_)
      for (int c = 0; c < ccnt[e]; ++c) {
        const int2 r = cr[cross_ptr_[e] + c];
        by[r.y].push_back(make_int4(e, r.x, cs[cross_ptr_[e] + c], 0));
      }
    std::vector<int> xsrc, xptr(1, 0);
    std::vector<int4> xent;
    for (int e = 0; e < E_; ++e)
      if (!by[e].empty()) {
        xsrc.push_back(e);
        xent.insert(xent.end(), by[e].begin(), by[e].end());
        xptr.push_back((int)xent.size());
      }
    d_xsrc_.upload(xsrc, st_);
    d_xptr_.upload(xptr, st_);
    d_xent_.upload(xent, st_);
    nxsrc = (int)xsrc.size();
  }
  // transposed FP64 factors, once per factorisation (also after a mixed-form run: one FP64 step
  // per application is cheaper than three refinement steps with FP32-rounded factors)
  if (!adjT_valid_) {
    if (!sparse_) make_transposed<double>(d_A_.p, a64_blocked_, d_AT_, d_linvT_, d_uinvT_);
    make_csr_transpose();
    adjT_valid_ = true;
  }
  if (sparse_) {
    mf_.prepare_transposed(st_);  // (a no-op until the next factorisation)
    HIP_CHECK(hipGetLastError());
  }
  d_x_.alloc(n);
  d_w_.alloc(n);
  d_v_.alloc(n);
  if (sparse_) HIP_CHECK(hipMemsetAsync(d_v_.p, 0, sizeof(double) * n, st_));  // padding stays zero

  // objective: J_c and J_y (discharge / constant weights plus the observation-head terms)
  std::vector<double> jy0(n, 0.0);
  if (obj && obj->dq)
    for (int e : de_) jy0[de_pos_[e]] = obj->dq[e];
  if (obj && obj->dconstant)
    for (int f = 0; f < F_; ++f) jy0[fpos_[f]] = obj->dconstant[f];
  d_gc_.zero(ncf, st_);
  d_gq_.zero(E_, st_);
  d_gC_.zero(F_, st_);
  std::vector<int64_t> oorder(nobs);
  if (nobs > 0) {
    for (int64_t k = 0; k < nobs; ++k) oorder[k] = k;
    std::stable_sort(oorder.begin(), oorder.end(),
                     [&](int64_t a, int64_t b) { return obj->obs_frac[a] < obj->obs_frac[b]; });
    std::vector<double2> pz(nobs), wv(nobs);
    std::vector<int> pe(nobs, -1), fb(F_, 0), fe(F_, 0);
    for (int64_t i = 0; i < nobs; ++i) {
      const int64_t k = oorder[i];
      const int f = (int)obj->obs_frac[k];
      pz[i] = make_double2(obj->obs_z[2 * k], obj->obs_z[2 * k + 1]);
      wv[i] = make_double2(obj->dhead[k] / t_[f], 0.0);  // h = (Re Omega + C) / t
    }
    for (int64_t i = 0; i < nobs;) {
      const int f = (int)obj->obs_frac[oorder[i]];
      fb[f] = (int)i;
      while (i < nobs && obj->obs_frac[oorder[i]] == f) ++i;
      fe[f] = (int)i;
    }
    tab_obs_.pz.upload(pz, st_);
    tab_obs_.pelem.upload(pe, st_);
    tab_obs_.fpb.upload(fb, st_);
    tab_obs_.fpe.upload(fe, st_);
    d_wobs_.upload(wv, st_);
    aem::launch_adj_gather(nv, d_gitems_.p, E_, gsingle, d_eslot_ptr_.p, d_eslot_.p, tab_obs_.fpb.p, tab_obs_.fpe.p,
                           tab_obs_.pz.p, tab_obs_.pelem.p, d_wobs_.p, 0, 1, 0, d_gc_.p, d_gq_.p, st_);
    aem::launch_adj_csum(F_, tab_obs_.fpb.p, tab_obs_.fpe.p, d_wobs_.p, 0, d_gC_.p, st_);
  }
  d_jc_.alloc(ncf);
  HIP_CHECK(hipMemcpyAsync(d_jc_.p, d_gc_.p, sizeof(double2) * ncf, hipMemcpyDeviceToDevice, st_));
  d_jy0_.upload(jy0, st_);
  d_jy_.alloc(n);
  aem::launch_adj_build_z(nde, F_, d_de_.p, sys(), n, d_jy0_.p, d_gq_.p, d_gC_.p, d_jy_.p, st_);

  // adjoint solve
  const int method = opt ? opt->method : ANDFN_MI_ADJ_GMRES;
  if (method != ANDFN_MI_ADJ_GMRES && method != ANDFN_MI_ADJ_FIXED_POINT) fail(ANDFN_MI_EINVAL, "adjoint method %d unknown", method);
  d_lam_.zero(ncf, st_);
  d_mu_.zero(n, st_);
  d_z_.alloc(n);
  d_aty_.alloc(n);
  d_gc_.zero(ncf, st_);
  d_jzero_.zero(n, st_);
  d_wfull_.zero(std::max<int64_t>(full_.npts(), 1), st_);
  d_wrhs_.zero(std::max<int64_t>(rhs_.npts(), 1), st_);
  d_sfx_.zero((size_t)E_ * 2 * M_, st_);
  d_stats_.zero(4, st_);
  // d_gc_ = B lam (+ RHS_c^T A^-T J_y when with_j) and d_mu_ = A^-T (S_y^T lam (+ J_y)), where B is
  // the fixed-point map's linear part on the coefficients: FIT^T, EVAL^T at the collocation
  // points (coefficients, discharges, constants, branch cuts), A^-T, RHS^T
  // (the collocation gather overwrites every active entry of d_gc_; the entries past an element's
  // ncoef keep the zeros written above)
  auto apply = [&](bool with_j) {
    aem::launch_adj_fitT(nv, d_items_.p, (int)items_.size(), d_lam_.p, full_.d_pt_off.p, d_trig_.p, d_wfull_.p,
                         d_sfx_.p, st_);
    aem::launch_adj_gather(nv, d_gitems_.p, E_, gsingle, d_eslot_ptr_.p, d_eslot_.p, tab_full_.fpb.p, tab_full_.fpe.p,
                           tab_full_.pz.p, tab_full_.pelem.p, d_wfull_.p, 1, 1, 0, d_gc_.p, d_gq_.p, st_);
    aem::launch_adj_cross(nxsrc, d_xsrc_.p, d_xptr_.p, d_xent_.p, d_sfx_.p, (int)M_, d_gq_.p, st_);
    aem::launch_adj_csum(F_, tab_full_.fpb.p, tab_full_.fpe.p, d_wfull_.p, 0, d_gC_.p, st_);
    aem::launch_adj_build_z(nde, F_, d_de_.p, sys(), n, with_j ? d_jy_.p : d_jzero_.p, d_gq_.p, d_gC_.p, d_z_.p, st_);
    HIP_CHECK(hipMemsetAsync(d_mu_.p, 0, sizeof(double) * n, st_));
    if (sparse_) {
      adj_mu_step_sparse();
      adj_mu_step_sparse();  // one step of refinement (the pivot blocks act through their inverses)
    } else {
      adj_mu_step<double>(d_AT_.p, d_linvT_.p, d_uinvT_.p, d_x_.p, d_w_.p, d_v_.p);
    }
    aem::launch_adj_rhsT(nv, d_rhs_items_.p, nde, d_de_pos_.p, d_mu_.p, rhs_.d_pt_off.p, K_, d_wrhs_.p, st_);
    aem::launch_adj_gather(nv, d_gitems_.p, E_, gsingle, d_eslot_ptr_.p, d_eslot_.p, tab_rhs_.fpb.p, tab_rhs_.fpe.p,
                           tab_rhs_.pz.p, tab_rhs_.pelem.p, d_wrhs_.p, 0, 0, 1, d_gc_.p, d_gq_.p, st_);
  };
  auto read_stats = [&](double* v) {
    unsigned long long st4[4];
    HIP_CHECK(hipMemcpyAsync(st4, d_stats_.p, sizeof(st4), hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    for (int k = 0; k < 4; ++k) std::memcpy(&v[k], &st4[k], sizeof(double));
  };
  HIP_CHECK(hipStreamSynchronize(st_));
  HIP_CHECK(hipGetLastError());
  auto t_set = clk::now();
  int64_t nit = 0;
  double delta = 1.0;
  bool conv = false;
  if (method == ANDFN_MI_ADJ_FIXED_POINT) {
    while (nit < maxit) {
      ++nit;
      apply(true);
      HIP_CHECK(hipMemsetAsync(d_stats_.p, 0, sizeof(unsigned long long) * 4, st_));
      aem::launch_adj_update((int64_t)ncf, d_jc_.p, d_gc_.p, d_lam_.p, damp, d_stats_.p, st_);
      double v[4];
      read_stats(v);
      delta = v[1] > 0 ? v[0] / v[1] : 0.0;
      if (verbose && (nit % 50 == 0 || nit == 1))
        fprintf(stderr, "adjoint iteration %lld: relative change %.3e\n", (long long)nit, delta);
      if (!(delta == delta)) fail(ANDFN_MI_EHIP, "adjoint iteration produced NaN at iteration %lld", (long long)nit);
      if (nit >= 3 && delta <= tol) {
        conv = true;
        break;
      }
    }
  } else {
    // restarted GMRES on (I - B) lam = g, g = J_c + RHS_c^T A^-T J_y, in the packed space of the
    // active coefficients (k < ncoef of every element with coefficients), real components only
    // where the fit reads only one: Re for intersections and constant-head lines (real
    // coefficients), Im for impermeable lines.  The other component of such an element is read
    // by nothing and receives nothing (the gather writes it zero), an identity block of I - B
    // that GMRES need not carry.
    std::vector<int64_t> pidx;
    for (int e = 0; e < E_; ++e) {
      const int ty = etype_[e];
      if (ty == 2) continue;
      for (int k = 0; k < ncoef_[e]; ++k) {
        const int64_t q = 2 * ((int64_t)e * M_ + k);
        if (ty != 5) pidx.push_back(q);
        if (ty != 0 && ty != 3) pidx.push_back(q + 1);
      }
    }
    const int64_t np = (int64_t)pidx.size();
    const rocblas_int N = (rocblas_int)np;
    const int restart = 40;
    d_pidx_.upload(pidx, st_);
    d_gv_.alloc(np);
    d_xv_.zero(np, st_);
    d_wv_.alloc(np);
    d_V_.alloc((size_t)N * (restart + 1));
    d_hv_.alloc(restart + 1);
    const double one = 1.0, zero = 0.0, mone = -1.0;
    // g (lam = 0)
    apply(true);
    HIP_CHECK(hipMemsetAsync(d_stats_.p, 0, sizeof(unsigned long long) * 4, st_));
    aem::launch_adj_update((int64_t)ncf, d_jc_.p, d_gc_.p, d_lam_.p, 1.0, d_stats_.p, st_);  // lam = J_c + gc
    aem::launch_adj_pack(np, d_pidx_.p, (const double*)d_lam_.p, d_gv_.p, st_);
    ++nit;
    auto op = [&](const double* v, double* out) {  // out = (I - B) v (reads only the packed entries of lam)
      aem::launch_adj_unpack(np, d_pidx_.p, v, (double*)d_lam_.p, st_);
      apply(false);
      aem::launch_adj_residual_op(np, d_pidx_.p, v, (const double*)d_gc_.p, out, st_);
      ++nit;
    };
    double gnorm = 0.0;
    RB_CHECK(rocblas_dnrm2(rb_, N, d_gv_.p, 1, &gnorm));
    double* V = d_V_.p;
    std::vector<double> H((size_t)(restart + 1) * restart), cs(restart), sn(restart), sv(restart + 1), h(restart + 1),
        h2(restart + 1);
    double* w = d_wv_.p;
    if (gnorm == 0.0) {
      conv = true;
      delta = 0.0;
    }
    bool first = true;
    while (!conv && nit < maxit) {
      // r = g - (I - B) x
      if (first) {
        HIP_CHECK(hipMemcpyAsync(w, d_gv_.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st_));
        first = false;
      } else {
        op(d_xv_.p, d_wv_.p);
        RB_CHECK(rocblas_dscal(rb_, N, &mone, w, 1));
        RB_CHECK(rocblas_daxpy(rb_, N, &one, d_gv_.p, 1, w, 1));
      }
      double beta = 0.0;
      RB_CHECK(rocblas_dnrm2(rb_, N, w, 1, &beta));
      delta = beta / gnorm;
      if (delta <= tol) {
        conv = true;
        break;
      }
      const double ib = 1.0 / beta;
      RB_CHECK(rocblas_dcopy(rb_, N, w, 1, V, 1));
      RB_CHECK(rocblas_dscal(rb_, N, &ib, V, 1));
      std::fill(sv.begin(), sv.end(), 0.0);
      sv[0] = beta;
      int jd = 0;
      for (int j = 0; j < restart && nit < maxit; ++j) {
        op(V + (size_t)j * N, d_wv_.p);
        // classical Gram-Schmidt, repeated when the first pass cancels more than 1 - 1/sqrt(2) of
        // the vector's norm (the criterion of Daniel, Gragg, Kaufman and Stewart, 1976): each
        // pass reads the j + 1 basis vectors twice, the bulk of the iteration's memory traffic
        double w0 = 0.0, hn = 0.0;
        RB_CHECK(rocblas_dnrm2(rb_, N, w, 1, &w0));
        std::fill(h2.begin(), h2.end(), 0.0);
        for (int pass = 0; pass < 2; ++pass) {
          RB_CHECK(rocblas_dgemv(rb_, rocblas_operation_transpose, N, j + 1, &one, V, N, w, 1, &zero, d_hv_.p, 1));
          RB_CHECK(rocblas_dgemv(rb_, rocblas_operation_none, N, j + 1, &mone, V, N, d_hv_.p, 1, &one, w, 1));
          HIP_CHECK(hipMemcpyAsync(pass ? h2.data() : h.data(), d_hv_.p, sizeof(double) * (j + 1),
                                   hipMemcpyDeviceToHost, st_));
          RB_CHECK(rocblas_dnrm2(rb_, N, w, 1, &hn));
          HIP_CHECK(hipStreamSynchronize(st_));
          if (hn >= 0.70710678118654752 * w0) break;
          w0 = hn;
        }
        for (int i = 0; i <= j; ++i) H[(size_t)i * restart + j] = h[i] + h2[i];
        H[(size_t)(j + 1) * restart + j] = hn;
        if (hn > 0.0) {
          const double ih = 1.0 / hn;
          RB_CHECK(rocblas_dcopy(rb_, N, w, 1, V + (size_t)(j + 1) * N, 1));
          RB_CHECK(rocblas_dscal(rb_, N, &ih, V + (size_t)(j + 1) * N, 1));
        }
        // Givens rotations
        for (int i = 0; i < j; ++i) {
          const double a = H[(size_t)i * restart + j], b = H[(size_t)(i + 1) * restart + j];
          H[(size_t)i * restart + j] = cs[i] * a + sn[i] * b;
          H[(size_t)(i + 1) * restart + j] = -sn[i] * a + cs[i] * b;
        }
        const double a = H[(size_t)j * restart + j], b = H[(size_t)(j + 1) * restart + j];
        const double r = std::hypot(a, b);
        cs[j] = r > 0 ? a / r : 1.0;
        sn[j] = r > 0 ? b / r : 0.0;
        H[(size_t)j * restart + j] = r;
        H[(size_t)(j + 1) * restart + j] = 0.0;
        sv[j + 1] = -sn[j] * sv[j];
        sv[j] = cs[j] * sv[j];
        jd = j + 1;
        delta = std::fabs(sv[j + 1]) / gnorm;
        if (verbose && (nit % 20 == 0))
          fprintf(stderr, "adjoint GMRES %lld: relative residual %.3e\n", (long long)nit, delta);
        if (!(delta == delta)) fail(ANDFN_MI_EHIP, "adjoint GMRES produced NaN");
        if (delta <= tol || hn == 0.0) break;
      }
      // x += V y,  H y = s (upper triangular, jd x jd)
      std::vector<double> y(jd);
      for (int i = jd - 1; i >= 0; --i) {
        double acc = sv[i];
        for (int k = i + 1; k < jd; ++k) acc -= H[(size_t)i * restart + k] * y[k];
        y[i] = acc / H[(size_t)i * restart + i];
      }
      HIP_CHECK(hipMemcpyAsync(d_hv_.p, y.data(), sizeof(double) * jd, hipMemcpyHostToDevice, st_));
      RB_CHECK(rocblas_dgemv(rb_, rocblas_operation_none, N, jd, &one, V, N, d_hv_.p, 1, &one, d_xv_.p, 1));
      HIP_CHECK(hipStreamSynchronize(st_));
      if (delta <= tol) conv = true;
    }
    HIP_CHECK(hipMemsetAsync(d_lam_.p, 0, sizeof(double2) * ncf, st_));
    aem::launch_adj_unpack(np, d_pidx_.p, d_xv_.p, (double*)d_lam_.p, st_);
  }
  // mu of the final lam
  apply(true);
  HIP_CHECK(hipGetLastError());
  auto t_it = clk::now();

  // parameter terms at the solved state
  std::vector<FitItem> titems;
  for (const FitItem& it : items_)
    if (etype_[it.elem] == 0) titems.push_back(it);
  d_titems_.upload(titems, st_);
  d_ts_.alloc(std::max<size_t>(titems.size(), 1));
  d_val_.alloc(std::max<int64_t>(full_.npts(), 1));
  sweep(aem::MODE_COLLOC, full_, d_val_.p, false);
  aem::launch_adj_tgrad(nv, d_titems_.p, (int)titems.size(), d_lam_.p, d_val_.p, full_.d_pt_off.p, d_ts_.p, st_);
  d_hw_.alloc(std::max<int64_t>(rhs_.npts(), 1));
  aem::launch_sweep(aem::MODE_DISCH_HEADS, nv, rhs_.view(), rhs_.ntiles(1), K_, d_hw_.p, st_);
  HIP_CHECK(hipGetLastError());
  std::vector<double> ts(titems.size()), hw(rhs_.npts()), mu(n), qv(E_), cv(F_);
  if (!titems.empty())
    HIP_CHECK(hipMemcpyAsync(ts.data(), d_ts_.p, sizeof(double) * titems.size(), hipMemcpyDeviceToHost, st_));
  if (!hw.empty()) HIP_CHECK(hipMemcpyAsync(hw.data(), d_hw_.p, sizeof(double) * hw.size(), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(mu.data(), d_mu_.p, sizeof(double) * n, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(qv.data(), d_q_.p, sizeof(double) * E_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipMemcpyAsync(cv.data(), d_constant_.p, sizeof(double) * F_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  std::vector<double> hk(nobs);
  if (nobs > 0) heads(nobs, obj->obs_frac, obj->obs_z, hk.data());

  for (int f = 0; f < F_; ++f) grad_t[f] = 0.0;
  for (int e = 0; e < E_; ++e) grad_head[e] = 0.0;
  double J = 0.0;
  if (obj && obj->dq)
    for (int e = 0; e < E_; ++e) J += obj->dq[e] * qv[e];
  if (obj && obj->dconstant)
    for (int f = 0; f < F_; ++f) J += obj->dconstant[f] * cv[f];
  // observation heads h = Phi / t:  dh/dt = -h / t
  for (int64_t k = 0; k < nobs; ++k) {
    const int f = (int)obj->obs_frac[k];
    J += obj->dhead[k] * hk[k];
    grad_t[f] -= obj->dhead[k] * hk[k] / t_[f];
  }
  // intersection fits c_j = (t0 a1_j - t1 a0_j)/(t0 + t1)
  for (size_t b = 0; b < titems.size(); ++b) {
    const int e = titems[b].elem;
    const int f0 = efrac0_[e], f1 = efrac1_[e];
    const double t0 = t_[f0], t1 = t_[f1], ss = (t0 + t1) * (t0 + t1);
    grad_t[f0] += ts[b] * t1 / ss;
    grad_t[f1] -= ts[b] * t0 / ss;
  }
  // head equations: intersection rows Phi0/t0 - Phi1/t1 = 0 give mu (h0/t0, -h1/t1); constant-head
  // rows Phi = head t give mu head (d/dt) and mu t (d/dhead)
  for (int j = 0; j < nde; ++j) {
    const RhsItem& it = rhs_items_[j];
    const int e = it.elem;
    const int b0 = rhs_.pt_off[it.s0];
    double h0 = 0.0;
    for (int i = 0; i < K_; ++i) h0 += hw[b0 + i];
    h0 /= (double)K_;
    if (etype_[e] == 0) {
      const int b1 = rhs_.pt_off[it.s1];
      double h1 = 0.0;
      for (int i = 0; i < K_; ++i) h1 += hw[b1 + i];
      h1 /= (double)K_;
      grad_t[efrac0_[e]] += mu[de_pos_[e]] * h0 / t_[efrac0_[e]];
      grad_t[efrac1_[e]] -= mu[de_pos_[e]] * h1 / t_[efrac1_[e]];
    } else {
      const int f = efrac0_[e];
      grad_t[f] += mu[de_pos_[e]] * head_[e];
      grad_head[e] = mu[de_pos_[e]] * t_[f];
    }
  }
  auto t_end = clk::now();
  if (info) {
    info->nit = nit;
    info->delta = delta;
    info->objective = J;
    info->t_setup = std::chrono::duration<double>(t_set - t0).count();
    info->t_iterate = std::chrono::duration<double>(t_it - t_set).count();
    info->t_total = std::chrono::duration<double>(t_end - t0).count();
  }
  return conv ? ANDFN_MI_OK : ANDFN_MI_NOT_CONVERGED;
}

void Solver::set_parameters(const double* t, const double* head) {
  if (t)
    for (int f = 0; f < F_; ++f)
      if (!(t[f] > 0.0) || !std::isfinite(t[f])) fail(ANDFN_MI_EINVAL, "t[%d] = %g must be positive and finite", f, t[f]);
  if (t) std::copy(t, t + F_, t_.begin());
  for (int e = 0; e < E_; ++e)
    if (etype_[e] == 2 || etype_[e] == 3) {
      if (head) head_[e] = head[e];
      phi_[e] = head_[e] * t_[efrac0_[e]];
    }
  d_t_.upload(t_, st_);
  d_phi_.upload(phi_, st_);
  HIP_CHECK(hipStreamSynchronize(st_));
  have_state_ = false;
  adjT_valid_ = false;
}

void Solver::state_discharges(double* q, double* constant) {
  if (q) HIP_CHECK(hipMemcpyAsync(q, d_q_.p, sizeof(double) * E_, hipMemcpyDeviceToHost, st_));
  if (constant) HIP_CHECK(hipMemcpyAsync(constant, d_constant_.p, sizeof(double) * F_, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
}

int guarded(const std::function<int()>& fn) {
  try {
    return fn();
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return ANDFN_MI_EHIP;
  }
}

}  // namespace

struct andfn_mi_handle {
  Solver* s;
};

// The allocation-free entry point's device scratch: one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel or copy (its host-side name is not a device
// address): its address on each device comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that address is used.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
namespace {
constexpr int kNoAllocMaxDevices = 64;
unsigned char* g_scratch_ptr[kNoAllocMaxDevices] = {};
std::mutex g_noalloc_mu;

unsigned char* scratch_on_device(int dev) {
  if (dev < 0 || dev >= kNoAllocMaxDevices) fail(ANDFN_MI_EINVAL, "device %d out of range for the scratch table", dev);
  if (!g_scratch_ptr[dev]) {
    HIP_CHECK(hipSetDevice(dev));
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(ANDFN_MI_EHIP, "no device address for the heads scratch on device %d", dev);
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

namespace aem {
void set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }
}  // namespace aem

extern "C" {

int andfn_mi_create(void* fractures, int64_t n_fractures, int64_t max_elements, void* elements, int64_t n_elements,
                    int64_t max_ncoef, int64_t discharge_int, const void* constants, int32_t device,
                    andfn_mi_handle** out) {
  return guarded([&]() {
    if (!fractures || !elements || !constants || !out) fail(ANDFN_MI_EINVAL, "null pointer");
    *out = nullptr;
    Solver* s = new Solver(fractures, n_fractures, max_elements, elements, n_elements, max_ncoef, discharge_int,
                           constants, device);
    *out = new andfn_mi_handle{s};
    return ANDFN_MI_OK;
  });
}

int andfn_mi_run(andfn_mi_handle* h, const andfn_mi_options* opt, andfn_mi_info* info) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    return h->s->solve(opt, info);
  });
}

int andfn_mi_download(andfn_mi_handle* h) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    h->s->write_back();
    return ANDFN_MI_OK;
  });
}

void andfn_mi_destroy(andfn_mi_handle* h) {
  if (!h) return;
  try {
    delete h->s;
  } catch (...) {
  }
  delete h;
}

int andfn_mi_solve(void* fractures, int64_t n_fractures, int64_t max_elements, void* elements, int64_t n_elements,
                   int64_t max_ncoef, int64_t discharge_int, const void* constants, const int64_t* ntype_elements,
                   const andfn_mi_options* opt, andfn_mi_info* info) {
  (void)ntype_elements;  // unused by the reference as well (hpc_solve.py:55-79)
  return guarded([&]() {
    if (!fractures || !elements || !constants) fail(ANDFN_MI_EINVAL, "null record pointer");
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    int rc;
    clk::time_point t1, t2, t3;
    {
      Solver s(fractures, n_fractures, max_elements, elements, n_elements, max_ncoef, discharge_int, constants,
               opt ? opt->device : 0);
      t1 = clk::now();
      rc = s.solve(opt, info);
      t2 = clk::now();
      s.write_back();
      t3 = clk::now();
    }
    if (opt && opt->verbose) {
      auto ms = [](clk::time_point a, clk::time_point b) { return 1e3 * std::chrono::duration<double>(b - a).count(); };
      fprintf(stderr, "drop-in call: records in %.1f ms, solve %.1f ms, records out %.1f ms, release %.1f ms\n",
              ms(t0, t1), ms(t1, t2), ms(t2, t3), ms(t3, clk::now()));
    }
    return rc;
  });
}

int andfn_mi_heads(const void* fractures, int64_t n_fractures, int64_t max_elements, const void* elements,
                   int64_t n_elements, int64_t max_ncoef, int64_t n_points, const int64_t* frac, const double* z,
                   double* heads, int32_t device) {
  return guarded([&]() {
    if (!fractures || !elements || (n_points > 0 && (!frac || !z || !heads))) fail(ANDFN_MI_EINVAL, "null pointer");
    Solver s((void*)fractures, n_fractures, max_elements, (void*)elements, n_elements, max_ncoef, 1, nullptr, device);
    if (n_points > 0) s.heads(n_points, frac, z, heads);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_omega(const void* fractures, int64_t n_fractures, int64_t max_elements, const void* elements,
                   int64_t n_elements, int64_t max_ncoef, int64_t n_points, const int64_t* frac, const double* z,
                   double* omega, int32_t device) {
  return guarded([&]() {
    if (!fractures || !elements || (n_points > 0 && (!frac || !z || !omega))) fail(ANDFN_MI_EINVAL, "null pointer");
    Solver s((void*)fractures, n_fractures, max_elements, (void*)elements, n_elements, max_ncoef, 1, nullptr, device);
    if (n_points > 0) s.heads(n_points, frac, z, omega, 1);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_state_omega(andfn_mi_handle* h, int64_t n_points, const int64_t* frac, const double* z, double* omega) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    if (n_points > 0 && (!frac || !z || !omega)) fail(ANDFN_MI_EINVAL, "null pointer");
    if (n_points > 0) h->s->heads(n_points, frac, z, omega, 1);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_w(const void* fractures, int64_t n_fractures, int64_t max_elements, const void* elements,
               int64_t n_elements, int64_t max_ncoef, int64_t n_points, const int64_t* frac, const double* z,
               int64_t exclude, int32_t form, double* w, int32_t device) {
  return guarded([&]() {
    if (!fractures || !elements || (n_points > 0 && (!frac || !z || !w))) fail(ANDFN_MI_EINVAL, "null pointer");
    if (form != ANDFN_MI_W_JIT && form != ANDFN_MI_W_OBJECT) fail(ANDFN_MI_EINVAL, "unknown form %d", (int)form);
    Solver s((void*)fractures, n_fractures, max_elements, (void*)elements, n_elements, max_ncoef, 1, nullptr, device);
    if (n_points > 0) s.heads(n_points, frac, z, w, form == ANDFN_MI_W_JIT ? 2 : 3, exclude);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_state_w(andfn_mi_handle* h, int64_t n_points, const int64_t* frac, const double* z, int64_t exclude,
                     int32_t form, double* w) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    if (n_points > 0 && (!frac || !z || !w)) fail(ANDFN_MI_EINVAL, "null pointer");
    if (form != ANDFN_MI_W_JIT && form != ANDFN_MI_W_OBJECT) fail(ANDFN_MI_EINVAL, "unknown form %d", (int)form);
    if (n_points > 0) h->s->heads(n_points, frac, z, w, form == ANDFN_MI_W_JIT ? 2 : 3, exclude);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_state_streamlines(andfn_mi_handle* h, int64_t n_lines, const int64_t* frac, const double* z0,
                               const double* elevation, const double* aperture, double ds, int32_t backward,
                               int64_t max_length, int64_t* n_points, int64_t* n_segments, int64_t* exit_element) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    if (!aperture) fail(ANDFN_MI_EINVAL, "null pointer");
    if (n_lines > 0 && (!frac || !z0 || !elevation || !n_points || !n_segments || !exit_element))
      fail(ANDFN_MI_EINVAL, "null pointer");
    h->s->streamlines(n_lines, frac, z0, elevation, aperture, ds, backward, max_length, n_points, n_segments,
                      exit_element);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_state_streamlines_fetch(andfn_mi_handle* h, double* points, double* speed, int64_t* segment_fracture,
                                     int64_t* segment_points) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    if (!points || !speed || !segment_fracture || !segment_points) fail(ANDFN_MI_EINVAL, "null pointer");
    h->s->streamlines_fetch(points, speed, segment_fracture, segment_points);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_adjoint(andfn_mi_handle* h, const andfn_mi_objective* objective, const andfn_mi_adjoint_options* opt,
                     double* grad_t, double* grad_head, andfn_mi_adjoint_info* info) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    return h->s->adjoint(objective, opt, grad_t, grad_head, info);
  });
}

int andfn_mi_set_parameters(andfn_mi_handle* h, const double* t, const double* head) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    h->s->set_parameters(t, head);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_state_heads(andfn_mi_handle* h, int64_t n_points, const int64_t* frac, const double* z, double* heads) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    if (n_points > 0 && (!frac || !z || !heads)) fail(ANDFN_MI_EINVAL, "null pointer");
    if (n_points > 0) h->s->heads(n_points, frac, z, heads);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_state_heads_noalloc(andfn_mi_handle* h, int64_t n_points, const int64_t* frac, const double* z,
                                 double* heads) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    if (n_points > 0 && (!frac || !z || !heads)) fail(ANDFN_MI_EINVAL, "null pointer");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    unsigned char* ptr = scratch_on_device(h->s->device());
    h->s->heads_noalloc(n_points, frac, z, heads, ptr);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_state_discharges(andfn_mi_handle* h, double* q, double* constant) {
  return guarded([&]() {
    if (!h || !h->s) fail(ANDFN_MI_EINVAL, "null handle");
    h->s->state_discharges(q, constant);
    return ANDFN_MI_OK;
  });
}

int andfn_mi_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* andfn_mi_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
