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
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/andfn_mi.h"
#include "aem_kernels.hpp"

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
  std::vector<Tile> tiles;
  std::vector<int> frac_nslots;  // sources per tile's fracture, for work accounting
  DBuf<int> d_side_elem, d_side_which, d_pt_off;
  DBuf<Tile> d_tiles;
  int64_t npts() const { return pt_off.empty() ? 0 : pt_off.back(); }
  aem::TargetView view() const {
    aem::TargetView v;
    v.tiles = d_tiles.p;
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
    for (auto& e : ev_) (void)hipEventDestroy(e);
    if (st_) (void)hipStreamDestroy(st_);
  }

  int solve(const andfn_mi_options* opt, andfn_mi_info* info);
  void write_back();
  void heads(int64_t npts, const int64_t* frac, const double* z, double* out);

 private:
  Layout L_;
  char* fr_;
  char* el_;
  int F_, E_, K_;
  ConstantsRec C_{};
  hipStream_t st_ = nullptr;
  rocblas_handle rb_ = nullptr;

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
  DBuf<double> d_dpsi_, d_val_, d_valg_, d_valr_, d_A_, d_b_, d_y_, d_yold_, d_x_;
  DBuf<int> d_perm_, d_sched_;
  DBuf<double> d_rowscale_, d_ay_, d_bcmax_, d_csr_val_;
  DBuf<int> d_csr_ptr_, d_csr_col_;
  int64_t nnz_A_ = 0;
  // inverses of the diagonal blocks of L (unit lower) and U, for the blocked triangular solves
  static constexpr int kNB = 1024;
  DBuf<double> d_linv_, d_uinv_, d_w_, d_v_, d_req_;
  template <class T>
  void prepare_blocked(T* A, int64_t n, DBuf<T>& linv, DBuf<T>& uinv);
  template <class T>
  void blocked_solve(const T* A, int n, const T* linv, const T* uinv, T* x, T* w, T* v);
  // mixed precision: single-precision factors and blocked-solve buffers
  DBuf<float> d_A32_, d_linv32_, d_uinv32_, d_x32_, d_w32_, d_v32_;
  DBuf<float2> d_coef32_;
  bool mixed_phase_ = false;
  double far2_ = 16.0;
  int precision_ = ANDFN_MI_FP64;
  // initial state of the mutable arrays, restored at the start of every run
  DBuf<double2> s_coef_;
  DBuf<double> s_q_, s_constant_, s_error_, s_error_old_, s_error_old2_, s_error_coef_, s_dpsi_;
  DBuf<int> s_ncoef_, s_nint_;
  std::vector<int> ncoef0_, nint0_;
  std::vector<double> y0_;
  int lu_mode_ = ANDFN_MI_LU_TRSV;
  DBuf<Slot> d_slots_;
  DBuf<FitItem> d_items_, d_items_g_;
  DBuf<RhsItem> d_rhs_items_;
  DBuf<rocblas_int> d_ipiv_, d_info_;
  DBuf<aem::IterScalars> d_scal_;

  TargetSet full_, grown_, rhs_;
  std::vector<FitItem> items_, items_g_;
  std::vector<RhsItem> rhs_items_;
  hipEvent_t ev_[8] = {};

  aem::NetView net() const {
    aem::NetView v;
    v.M = (int)L_.M;
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
    return v;
  }

  void parse();
  void upload();
  void snapshot();
  void restore();
  void build_targets(TargetSet& ts, const std::vector<char>* only, bool rhs);
  void upload_targets(TargetSet& ts);
  void build_items(const TargetSet& ts, std::vector<FitItem>& items, const std::vector<char>* only);
  double set_flops(const TargetSet& ts, bool with_q, int64_t* pairs) const;
  void assemble_and_factor();
  double* discharge_solve(double* b, double* x);
  void solve_rhs();
  hipGraph_t lu_graph_ = nullptr;
  hipGraphExec_t lu_exec_ = nullptr;
  bool lu_graph_tried_ = false;
  void drop_graph() {
    if (lu_exec_) (void)hipGraphExecDestroy(lu_exec_);
    if (lu_graph_) (void)hipGraphDestroy(lu_graph_);
    lu_exec_ = nullptr;
    lu_graph_ = nullptr;
    lu_graph_tried_ = false;
  }
  double head_residual(const double* b, int* row);
  double bc_residual();
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
  coef_.assign((size_t)E_ * L_.M, make_double2(0, 0));
  dpsi_.assign((size_t)E_ * 2 * L_.M, 0.0);
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
    for (int64_t j = 0; j < L_.M; ++j) coef_[(size_t)i * L_.M + j] = d2(rd<cd>(p, L_.e_coef + 16 * j));
    for (int64_t j = 0; j < 2 * L_.M; ++j) dpsi_[(size_t)i * 2 * L_.M + j] = rd<double>(p, L_.e_dpsi + 8 * j);
    error_[i] = rd<double>(p, L_.e_err);
    error_old_[i] = rd<double>(p, L_.e_err + 8);
    error_old2_[i] = rd<double>(p, L_.e_err + 16);
    error_coef_[i] = rd<double>(p, L_.e_err + 24);
  }
  t_.resize(F_);
  constant_.resize(F_);
  frac_id_.resize(F_);
  mem_ptr_.assign(F_ + 1, 0);
  slots_.clear();
  for (int f = 0; f < F_; ++f) {
    const char* p = fr_ + (size_t)f * L_.fsz;
    frac_id_[f] = rd<int64_t>(p, 0);
    t_[f] = rd<double>(p, 8);
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
  for (auto& e : ev_) HIP_CHECK(hipEventCreate(&e));
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
  const int nde = (int)de_.size();
  y0_.assign(nde + F_, 0.0);
  for (int j = 0; j < nde; ++j) y0_[j] = q_[de_[j]];
  for (int f = 0; f < F_; ++f) y0_[nde + f] = constant_[f];
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
void Solver::build_targets(TargetSet& ts, const std::vector<char>* only, bool rhs) {
  ts.side_elem.clear();
  ts.side_which.clear();
  ts.pt_off.assign(1, 0);
  ts.tiles.clear();
  ts.frac_nslots.clear();
  for (int f = 0; f < F_; ++f) {
    const int sbeg = (int)ts.side_elem.size();
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
  }
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
  upload_targets(ts);
}

void Solver::upload_targets(TargetSet& ts) {
  ts.d_side_elem.upload(ts.side_elem, st_);
  ts.d_side_which.upload(ts.side_which, st_);
  ts.d_pt_off.upload(ts.pt_off, st_);
  ts.d_tiles.upload(ts.tiles, st_);
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
  return rocblas_dtrmm(h, rocblas_side_right, f, rocblas_operation_none, d, m, nc, &one, T, ldt, B, (rocblas_int)ldb,
                       B, (rocblas_int)ldb);
}
static rocblas_status rb_trmm_right(rocblas_handle h, rocblas_fill f, rocblas_diagonal d, int m, int nc,
                                    const float* T, int ldt, float* B, int64_t ldb) {
  const float one = 1.0f;
  return rocblas_strmm(h, rocblas_side_right, f, rocblas_operation_none, d, m, nc, &one, T, ldt, B, (rocblas_int)ldb,
                       B, (rocblas_int)ldb);
}

// Blocked triangular-solve operators from the LU factors in A (n x n, column-major, kNB blocks):
// Linv_j / Uinv_j = inverses of the diagonal blocks of L (unit) and U, and, in place in A, the
// panels G_j = L[below j, j] Linv_j and H_j = U[above j, j] Uinv_j.  With them, L z = c becomes
// "c[below j] -= G_j c_j for j ascending, then z_j = Linv_j c_j for all j at once" (and U likewise,
// descending): the sequential chain is one gemv per block and the diagonal blocks are applied by
// a single batched launch.
template <class T>
void Solver::prepare_blocked(T* A, int64_t n, DBuf<T>& linv, DBuf<T>& uinv) {
  const int64_t nb = (n + kNB - 1) / kNB, nfull = n / kNB;
  linv.zero((size_t)nb * kNB * kNB, st_);
  uinv.zero((size_t)nb * kNB * kNB, st_);
  if (nfull > 0) {
    RB_CHECK(rb_trtri_batched(rb_, rocblas_fill_lower, rocblas_diagonal_unit, kNB, A, n, kNB * (n + 1), linv.p, kNB,
                              (int64_t)kNB * kNB, (int)nfull));
    RB_CHECK(rb_trtri_batched(rb_, rocblas_fill_upper, rocblas_diagonal_non_unit, kNB, A, n, kNB * (n + 1), uinv.p,
                              kNB, (int64_t)kNB * kNB, (int)nfull));
  }
  if (nb > nfull) {
    const int64_t r0 = nfull * kNB;
    const int rem = (int)(n - r0);
    const T* blk = A + r0 * (n + 1);
    RB_CHECK(rb_trtri(rb_, rocblas_fill_lower, rocblas_diagonal_unit, rem, blk, n, linv.p + nfull * kNB * kNB, kNB));
    RB_CHECK(rb_trtri(rb_, rocblas_fill_upper, rocblas_diagonal_non_unit, rem, blk, n, uinv.p + nfull * kNB * kNB,
                      kNB));
  }
  aem::launch_unit_diag<T>(linv.p, (int)nb, kNB, st_);
  for (int64_t j = 0; j < nb; ++j) {
    const int64_t r0 = j * kNB;
    const int bs = (int)std::min<int64_t>(kNB, n - r0);
    const int rest = (int)(n - r0 - bs);
    if (rest > 0)
      RB_CHECK(rb_trmm_right(rb_, rocblas_fill_lower, rocblas_diagonal_unit, rest, bs, linv.p + j * kNB * kNB, kNB,
                             A + r0 * n + r0 + bs, n));
    if (r0 > 0)
      RB_CHECK(rb_trmm_right(rb_, rocblas_fill_upper, rocblas_diagonal_non_unit, (int)r0, bs, uinv.p + j * kNB * kNB,
                             kNB, A + r0 * n, n));
  }
}

void Solver::assemble_and_factor() {
  drop_graph();
  const int nde = (int)de_.size();
  const int64_t n = nde + F_;
  d_A_.zero((size_t)(n * n), st_);
  aem::TargetView tv = rhs_.view();
  aem::launch_assemble(net(), tv, (int)rhs_.side_elem.size(), d_de_pos_.p, K_, nde, F_, n, d_A_.p, st_);
  HIP_CHECK(hipGetLastError());
  // a CSR copy of the unfactored matrix is kept for the head residual A y - b
  {
    std::vector<int> rp(1, 0), cols;
    std::vector<int> tmp;
    for (int j = 0; j < nde; ++j) {
      const int e = de_[j];
      tmp.clear();
      const int sides = etype_[e] == 0 ? 2 : 1;
      for (int sd = 0; sd < sides; ++sd) {
        const int f = sd ? efrac1_[e] : efrac0_[e];
        for (int k = mem_ptr_[f]; k < mem_ptr_[f + 1]; ++k) {
          const int src = slots_[k].elem;
          if (src != e && de_pos_[src] >= 0) tmp.push_back(de_pos_[src]);
        }
        tmp.push_back(nde + f);
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      cols.insert(cols.end(), tmp.begin(), tmp.end());
      rp.push_back((int)cols.size());
    }
    for (int f = 0; f < F_; ++f) {
      tmp.clear();
      for (int k = mem_ptr_[f]; k < mem_ptr_[f + 1]; ++k)
        if (de_pos_[slots_[k].elem] >= 0) tmp.push_back(de_pos_[slots_[k].elem]);
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      cols.insert(cols.end(), tmp.begin(), tmp.end());
      rp.push_back((int)cols.size());
    }
    d_csr_ptr_.upload(rp, st_);
    d_csr_col_.upload(cols, st_);
    d_csr_val_.alloc(cols.size());
    aem::launch_csr_gather(d_A_.p, n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, st_);
    HIP_CHECK(hipStreamSynchronize(st_));
    nnz_A_ = (int64_t)cols.size();
  }
  {
    std::vector<double> rs(nde, 1.0);
    for (int j = 0; j < nde; ++j)
      if (etype_[de_[j]] != 0) rs[j] = 1.0 / t_[efrac0_[de_[j]]];
    d_rowscale_.upload(rs, st_);
    HIP_CHECK(hipStreamSynchronize(st_));
  }
  d_ay_.alloc(n);
  // row equilibration (rows of intersections scale with 1/t), as SuperLU equilibrates by default
  d_req_.alloc(n);
  {
    DBuf<unsigned long long> mx;
    mx.alloc(n);
    aem::launch_row_scale(d_A_.p, n, d_req_.p, mx.p, st_);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st_));
  }
  d_ipiv_.alloc(n);
  d_info_.zero(1, st_);
  if (precision_ == ANDFN_MI_MIXED) {
    // single-precision factors of the equilibrated matrix; the FP64 residual of the CSR copy
    // drives the refinement, so the iteration's fixed point is the FP64 one
    d_A32_.alloc((size_t)(n * n));
    aem::launch_to_f32(d_A_.p, n * n, d_A32_.p, st_);
    RB_CHECK(rocsolver_sgetrf(rb_, (rocblas_int)n, (rocblas_int)n, d_A32_.p, (rocblas_int)n, d_ipiv_.p, d_info_.p));
  } else {
    RB_CHECK(rocsolver_dgetrf(rb_, (rocblas_int)n, (rocblas_int)n, d_A_.p, (rocblas_int)n, d_ipiv_.p, d_info_.p));
  }
  rocblas_int info = 0;
  HIP_CHECK(hipMemcpyAsync(&info, d_info_.p, sizeof(info), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  if (info > 0) fail(ANDFN_MI_ESINGULAR, "discharge matrix is exactly singular (zero pivot at %d)", (int)info);
  d_x_.alloc(n);
  d_w_.alloc(n);
  if (lu_mode_ == ANDFN_MI_LU_TRSV) {
    // LAPACK row interchanges -> one gather permutation: (P b)[i] = b[perm[i]]
    std::vector<rocblas_int> ipiv(n);
    HIP_CHECK(hipMemcpyAsync(ipiv.data(), d_ipiv_.p, sizeof(rocblas_int) * n, hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    std::vector<int> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int)i;
    for (int64_t i = 0; i < n; ++i) std::swap(perm[i], perm[ipiv[i] - 1]);
    d_perm_.upload(perm, st_);
    if (precision_ == ANDFN_MI_MIXED) {
      prepare_blocked<float>(d_A32_.p, n, d_linv32_, d_uinv32_);
      d_x32_.alloc(n);
      d_w32_.alloc(n);
      d_v32_.alloc(n);
    } else {
      prepare_blocked<double>(d_A_.p, n, d_linv_, d_uinv_);
    }
    d_w_.alloc(n);
    d_v_.alloc(n);
    HIP_CHECK(hipStreamSynchronize(st_));
  } else if (lu_mode_ == ANDFN_MI_LU_INVERSE) {
    RB_CHECK(rocsolver_dgetri(rb_, (rocblas_int)n, d_A_.p, (rocblas_int)n, d_ipiv_.p, d_info_.p));
    HIP_CHECK(hipStreamSynchronize(st_));
  }
}

// d_x_ = A^-1 d_b_.  The triangular-solve sequence (one permutation gather and two gemv per
// diagonal block) is launch-bound, so it is captured once into a HIP graph and replayed.
void Solver::solve_rhs() {
  if (lu_mode_ != ANDFN_MI_LU_TRSV) {
    (void)discharge_solve(d_b_.p, d_x_.p);
    return;
  }
  if (!lu_graph_tried_) {
    lu_graph_tried_ = true;
    (void)discharge_solve(d_b_.p, d_x_.p);  // uncaptured first call: any lazy library set-up happens here
    HIP_CHECK(hipStreamSynchronize(st_));
    if (hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      (void)discharge_solve(d_b_.p, d_x_.p);
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(st_, &g) == hipSuccess && g) {
        if (hipGraphInstantiate(&lu_exec_, g, nullptr, nullptr, 0) == hipSuccess) {
          lu_graph_ = g;
        } else {
          (void)hipGraphDestroy(g);
          lu_exec_ = nullptr;
        }
      }
      (void)hipGetLastError();
    }
    return;  // the first call's result is already in d_x_
  }
  if (lu_exec_) {
    HIP_CHECK(hipGraphLaunch(lu_exec_, st_));
  } else {
    (void)discharge_solve(d_b_.p, d_x_.p);
  }
}

// x = A^-1 b with the factors (or inverse) of the discharge matrix; returns the buffer holding x
// v = U^-1 L^-1 x with the operators of prepare_blocked (x and w are overwritten)
template <class T>
void Solver::blocked_solve(const T* A, int n, const T* linv, const T* uinv, T* x, T* w, T* v) {
  const int nb = (n + kNB - 1) / kNB;
  for (int j = 0; j < nb; ++j) {
    const int r0 = j * kNB, bs = std::min(kNB, n - r0), rest = n - r0 - bs;
    if (rest > 0) aem::launch_gemv_cm<T>(A + (size_t)r0 * n + r0 + bs, n, rest, bs, x + r0, x + r0 + bs, T(-1), T(1), st_);
  }
  aem::launch_gemv_diag<T>(linv, kNB, n, 1, x, w, st_);
  for (int j = nb - 1; j > 0; --j) {
    const int r0 = j * kNB, bs = std::min(kNB, n - r0);
    aem::launch_gemv_cm<T>(A + (size_t)r0 * n, n, r0, bs, w + r0, w, T(-1), T(1), st_);
  }
  aem::launch_gemv_diag<T>(uinv, kNB, n, 2, w, v, st_);
}

double* Solver::discharge_solve(double* b, double* x) {
  const rocblas_int n = (rocblas_int)(de_.size() + F_);
  if (lu_mode_ == ANDFN_MI_LU_GETRS) {
    aem::launch_scale(b, d_req_.p, n, x, st_);
    RB_CHECK(rocsolver_dgetrs(rb_, rocblas_operation_none, n, 1, d_A_.p, n, d_ipiv_.p, x, n));
    return x;
  }
  if (lu_mode_ == ANDFN_MI_LU_INVERSE) {
    const double one = 1.0, zero = 0.0;
    aem::launch_scale(b, d_req_.p, n, d_w_.p, st_);
    RB_CHECK(rocblas_dgemv(rb_, rocblas_operation_none, n, n, &one, d_A_.p, n, d_w_.p, 1, &zero, x, 1));
    return x;
  }
  if (precision_ == ANDFN_MI_MIXED) {
    // x = y + (L U)^-1 R P (b - A y): one refinement step per iteration, FP32 factors
    aem::launch_spmv(n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, d_y_.p, d_ay_.p, st_);
    aem::launch_refine_rhs(b, d_ay_.p, d_perm_.p, d_req_.p, n, d_x32_.p, st_);
    blocked_solve<float>(d_A32_.p, n, d_linv32_.p, d_uinv32_.p, d_x32_.p, d_w32_.p, d_v32_.p);
    float* v32 = d_v32_.p;
    aem::launch_refine_update(d_y_.p, v32, n, x, st_);
    return x;
  }
  // R P b, then the blocked triangular solves
  aem::launch_permute(b, d_perm_.p, d_req_.p, n, x, st_);
  double* v = d_v_.p;
  blocked_solve<double>(d_A_.p, n, d_linv_.p, d_uinv_.p, x, d_w_.p, v);
  HIP_CHECK(hipMemcpyAsync(x, v, sizeof(double) * n, hipMemcpyDeviceToDevice, st_));
  return x;
}

// max_j |(A0 y)_j - b_j| * row_scale_j over the head equations for the current y
double Solver::head_residual(const double* b, int* row) {
  const int n = (int)(de_.size() + F_);
  aem::launch_spmv(n, d_csr_ptr_.p, d_csr_col_.p, d_csr_val_.p, d_y_.p, d_ay_.p, st_);
  aem::launch_residual(d_ay_.p, b, d_rowscale_.p, (int)de_.size(), d_scal_.p, st_);
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
  aem::launch_sweep(aem::MODE_COLLOC_HEADS, nv, full_.view(), (int)full_.tiles.size(), K_, hv.p, st_);
  std::vector<RhsItem> witems;
  for (const RhsItem& it : rhs_items_)
    if (etype_[it.elem] == 2) witems.push_back(it);
  DBuf<RhsItem> dw;
  dw.upload(witems, st_);
  if (!witems.empty())
    aem::launch_sweep(aem::MODE_DISCH_HEADS, nv, rhs_.view(), (int)rhs_.tiles.size(), K_, hw.p, st_);
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
  lu_mode_ = opt ? opt->lu_mode : ANDFN_MI_LU_TRSV;
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
  const double head_tol = opt ? opt->head_tol : 0.0;
  double resid = -1.0;
  int resid_row = -1;
  std::vector<double> rhist;
  int64_t nit_fp32 = 0;
  if (K_ < 1 || K_ > 2 * L_.M) fail(ANDFN_MI_EINVAL, "discharge_int %d outside [1, %lld]", K_, (long long)(2 * L_.M));
  if (max_coef + coef_increase - 1 > L_.M)
    fail(ANDFN_MI_ECAPACITY, "MAX_NCOEF %d + COEF_INCREASE %d exceeds the coefficient capacity %lld", max_coef,
         coef_increase, (long long)L_.M);
  for (int e = 0; e < E_; ++e)
    if (etype_[e] != 2 && is_circle_like(etype_[e]) && radius_[e] <= 0)
      fail(ANDFN_MI_EINVAL, "element %d: circle radius must be positive", e);

  restore();
  const int nde = (int)de_.size();
  const int n = nde + F_;
  aem::NetView nv = net();
  aem::FitView fv = fit();

  // setup: tables, initial discharges, errors, branch cuts, matrix
  build_targets(full_, nullptr, false);
  build_items(full_, items_, nullptr);
  d_items_.upload(items_, st_);
  build_targets(rhs_, nullptr, true);
  {
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
  int64_t n_sweep = 0, n_rhs = 0, pr_sweep = 0, pr_rhs = 0;
  std::vector<int> grow_h(E_);
  std::vector<char> grown(E_);
  aem::IterScalars sc{};
  int64_t pr_tmp = 0;
  const double rhs_flops = set_flops(rhs_, false, &pr_tmp);
  const int64_t rhs_pairs = pr_tmp;

  auto sweep_full = [&](bool blend_now) {
    HIP_CHECK(hipEventRecord(ev_[0], st_));
    aem::launch_sweep(aem::MODE_COLLOC, nv, full_.view(), (int)full_.tiles.size(), K_, d_val_.p, st_, mixed_phase_);
    HIP_CHECK(hipEventRecord(ev_[1], st_));
    aem::launch_fit(nv, fv, d_items_.p, (int)items_.size(), d_val_.p, full_.d_pt_off.p, damping, blend_now ? 1 : 0, st_,
                    mixed_phase_);
    aem::launch_wells(nv, fv, d_wells_.p, (int)wells_.size(), damping, st_);
    HIP_CHECK(hipEventRecord(ev_[2], st_));
    int64_t pr = 0;
    fl_sweep += set_flops(full_, true, &pr);
    pr_sweep += pr;
    ++n_sweep;
  };
  bool grown_timed = false;
  auto add_times = [&]() {
    float a = 0, b = 0, c = 0, d = 0;
    if (grown_timed) {
      float g = 0;
      HIP_CHECK(hipEventElapsedTime(&g, ev_[6], ev_[7]));
      ms_sweep += g;
      grown_timed = false;
    }
    HIP_CHECK(hipEventElapsedTime(&a, ev_[0], ev_[1]));
    HIP_CHECK(hipEventElapsedTime(&b, ev_[1], ev_[2]));
    HIP_CHECK(hipEventElapsedTime(&c, ev_[3], ev_[4]));
    HIP_CHECK(hipEventElapsedTime(&d, ev_[4], ev_[5]));
    ms_sweep += a;
    ms_fit += b;
    ms_rhs += c;
    ms_lu += d;
  };

  while (cnt_error < 2 && nit < max_iterations) {
    ++nit;
    // ---- discharge solve (hpc_solve.py:155-174, 594-663)
    if (mixed_phase_) aem::launch_coef_to_f32(nv, E_, d_coef32_.p, st_);
    HIP_CHECK(hipEventRecord(ev_[3], st_));
    aem::launch_sweep(aem::MODE_RHS, nv, rhs_.view(), (int)rhs_.tiles.size(), K_, d_valr_.p, st_, mixed_phase_);
    aem::launch_rhs_reduce(nv, d_rhs_items_.p, nde, d_valr_.p, rhs_.d_pt_off.p, K_, d_b_.p, st_);
    HIP_CHECK(hipMemsetAsync(d_b_.p + nde, 0, sizeof(double) * F_, st_));
    if (head_tol > 0.0 && nit >= 2) {
      // head residual of the previous iterate against the right-hand side of its coefficients
      resid = head_residual(d_b_.p, &resid_row);
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
        --nit;
        HIP_CHECK(hipEventRecord(ev_[4], st_));
        HIP_CHECK(hipEventRecord(ev_[5], st_));
        break;
      }
    }
    HIP_CHECK(hipEventRecord(ev_[4], st_));
    solve_rhs();  // d_x_ = A^-1 d_b_
    HIP_CHECK(hipEventRecord(ev_[5], st_));
    HIP_CHECK(hipMemcpyAsync(d_yold_.p, d_y_.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st_));
    HIP_CHECK(hipMemcpyAsync(d_y_.p, d_x_.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st_));
    aem::launch_scatter(nv, d_de_.p, nde, F_, d_y_.p, st_);
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
      aem::launch_grow(nv, fv, E_, gp, dsched, d_grow_.p, st_);
      HIP_CHECK(hipMemcpyAsync(grow_h.data(), d_grow_.p, sizeof(int) * E_, hipMemcpyDeviceToHost, st_));
      HIP_CHECK(hipStreamSynchronize(st_));
      int ng = 0;
      std::vector<int>& gcirc = gcirc_;
      gcirc.clear();
      for (int e = 0; e < E_; ++e) {
        grown[e] = (char)grow_h[e];
        if (grow_h[e]) {
          ++ng;
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
        d_valg_.alloc(std::max<int64_t>(grown_.npts(), 1));
        if (!gcirc.empty()) {
          d_grown_circles_.upload(gcirc, st_);
          aem::launch_cross(nv, fv, d_grown_circles_.p, (int)gcirc.size(), st_);
        }
        HIP_CHECK(hipEventRecord(ev_[6], st_));
        if (mixed_phase_) aem::launch_coef_to_f32(nv, E_, d_coef32_.p, st_);
        aem::launch_sweep(aem::MODE_COLLOC, nv, grown_.view(), (int)grown_.tiles.size(), K_, d_valg_.p, st_, mixed_phase_);
        HIP_CHECK(hipEventRecord(ev_[7], st_));
        grown_timed = true;
        ++n_sweep;
        aem::launch_fit(nv, fv, d_items_g_.p, (int)items_g_.size(), d_valg_.p, grown_.d_pt_off.p, damping, 0, st_,
                        mixed_phase_);
        int64_t pr = 0;
        fl_sweep += set_flops(grown_, true, &pr);
        pr_sweep += pr;
        // collocation tables for the grown counts, used from the next sweep on
        build_targets(full_, nullptr, false);
        d_val_.alloc(std::max<int64_t>(full_.npts(), 1));
      }
      aem::launch_blend(nv, E_, damping, st_);
    }
    aem::launch_reduce(fv, E_, d_y_.p, d_yold_.p, n, d_scal_.p, st_);
    HIP_CHECK(hipMemcpyAsync(&sc, d_scal_.p, sizeof(sc), hipMemcpyDeviceToHost, st_));
    HIP_CHECK(hipStreamSynchronize(st_));
    HIP_CHECK(hipGetLastError());
    add_times();
    if (record && nit - 1 < record_n) std::copy(ncoef_.begin(), ncoef_.end(), record + (size_t)(nit - 1) * E_);
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
  auto t_loop = clk::now();
  const bool stopped_on_head = head_tol > 0.0 && resid >= 0.0 && resid <= head_tol;
  if (!stopped_on_head) {
    // right-hand side of the returned coefficients, then the residual of the returned discharges
    aem::launch_sweep(aem::MODE_RHS, nv, rhs_.view(), (int)rhs_.tiles.size(), K_, d_valr_.p, st_);
    aem::launch_rhs_reduce(nv, d_rhs_items_.p, nde, d_valr_.p, rhs_.d_pt_off.p, K_, d_b_.p, st_);
    resid = head_residual(d_b_.p, &resid_row);
  }
  const double bcres = bc_residual();
  int status = 0;
  HIP_CHECK(hipMemcpyAsync(&status, d_status_.p, sizeof(int), hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  if (status & 1) fail(ANDFN_MI_ECAPACITY, "branch-cut crossing list overflow");
  auto t_end = clk::now();
  if (info) {
    info->nit = nit;
    info->n_unknowns = n;
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
  }
  mixed_phase_ = false;
  return (cnt_error < 2 && !stopped_on_head) ? ANDFN_MI_NOT_CONVERGED : ANDFN_MI_OK;
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
    for (int64_t j = 0; j < L_.M; ++j) {
      double2 c = coef_[(size_t)i * L_.M + j];
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
      for (int k = 0; k < n - 1; ++k) wr<double>(p, L_.e_dpsi + 8 * k, dpsi_[(size_t)i * 2 * L_.M + k]);
    wr<double>(p, L_.e_err, error_[i]);
    wr<double>(p, L_.e_err + 8, error_old_[i]);
    wr<double>(p, L_.e_err + 16, error_old2_[i]);
    wr<double>(p, L_.e_err + 24, error_coef_[i]);
  }
  for (int f = 0; f < F_; ++f) wr<double>(fr_ + (size_t)f * L_.fsz, L_.f_const, constant_[f]);
}

void Solver::heads(int64_t npts, const int64_t* frac, const double* z, double* out) {
  // group the points by fracture into tiles, evaluate Re Omega / t (hpc_fracture.py:204-250)
  std::vector<int64_t> order(npts);
  for (int64_t i = 0; i < npts; ++i) {
    if (frac[i] < 0 || frac[i] >= F_) fail(ANDFN_MI_EINVAL, "point %lld: fracture %lld out of range", (long long)i, (long long)frac[i]);
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return frac[a] < frac[b]; });
  std::vector<double2> zs(npts);
  std::vector<Tile> tiles;
  for (int64_t k = 0; k < npts;) {
    int64_t f = frac[order[k]];
    int64_t b = k;
    while (k < npts && frac[order[k]] == f) {
      zs[k] = make_double2(z[2 * order[k]], z[2 * order[k] + 1]);
      ++k;
    }
    for (int64_t p0 = b; p0 < k; p0 += 64) tiles.push_back(Tile{(int)f, 0, 0, (int)p0, (int)k});
  }
  DBuf<double2> dz;
  DBuf<Tile> dt;
  DBuf<double> dv;
  dz.upload(zs, st_);
  dt.upload(tiles, st_);
  dv.alloc(std::max<int64_t>(npts, 1));
  aem::TargetView tv{};
  tv.tiles = dt.p;
  tv.pt_z = dz.p;
  aem::launch_sweep(aem::MODE_POINTS, net(), tv, (int)tiles.size(), 1, dv.p, st_);
  HIP_CHECK(hipGetLastError());
  std::vector<double> v(npts);
  HIP_CHECK(hipMemcpyAsync(v.data(), dv.p, sizeof(double) * npts, hipMemcpyDeviceToHost, st_));
  HIP_CHECK(hipStreamSynchronize(st_));
  for (int64_t k = 0; k < npts; ++k) out[order[k]] = v[k];
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
    Solver s(fractures, n_fractures, max_elements, elements, n_elements, max_ncoef, discharge_int, constants,
             opt ? opt->device : 0);
    int rc = s.solve(opt, info);
    s.write_back();
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

int andfn_mi_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* andfn_mi_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
