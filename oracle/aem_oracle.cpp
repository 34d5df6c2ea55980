// CPU FP64 restatement of AnDFN's HPC analytic-element solve (andfn/hpc).
//
// TEST INFRASTRUCTURE ONLY.  This library is the oracle the accelerated solve is checked
// against and the CPU baseline it is timed beside.  Only tests/, the smoke check and the
// benchmark's CPU-baseline leg load it; the product path never does.
//
// It follows the reference operation by operation, in double precision, on the caller's
// record arrays (element_dtype_hpc / fracture_dtype_hpc layouts of andfn/element.py:45-69
// and 146-159), with OpenMP over elements where the reference uses numba.prange:
//   * Joukowsky map with two principal square roots, inverse map   hpc_geometry_functions.py:11-55
//   * Horner asymptotic / Taylor series with a division per term   hpc_math_functions.py:16-40, 103-128
//   * well log term                                                hpc_math_functions.py:193-213
//   * superposition per fracture                                   hpc_fracture.py:51-92
//   * discharge vector w, series derivatives                       hpc_fracture.py:139-174,
//     hpc_math_functions.py:75-100, 163-190, and the object model's calc_w (fracture.py:453-477)
//   * Cauchy fits, powers by repeated multiplication               hpc_math_functions.py:243-297, 450-608
//   * branch-cut bookkeeping                                       hpc_math_functions.py:300-447
//   * error measures                                               hpc_math_functions.py:611-670
//   * collocation angles                                           hpc_math_functions.py:673-723
//   * element solves                                               hpc_intersection.py:49-112,
//     hpc_bounding_circle.py:118-163, hpc_const_head_line.py:38-85, hpc_imp_object.py:17-64, 124-183
//   * Jacobi sweep, damping, coefficient growth                    hpc_solve.py:310-591
//   * head right-hand side                                         hpc_solve.py:778-829
//   * discharge matrix assembly                                    hpc_solve.py:1061-1201, 1426-1459
// The outer loop and the SuperLU factorisation live in oracle.py (as in hpc_solve.py:55-257).
//
// Parity status of each part is recorded with its tests: every function above is pinned by
// the golden runs of AnDFN itself (tests/golden) and, independently, by closed forms.
//
// Build: see oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off; no fast-math, like Numba).

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

typedef std::complex<double> cd;
typedef int64_t i64;

namespace {

const double PI = 3.141592653589793;

// ---- record layouts -------------------------------------------------------------------
struct Layout {
  i64 M;   // MAX_NCOEF capacity
  i64 ME;  // MAX_ELEMENTS capacity
  size_t esz, fsz;
  size_t e_thetas, e_coef, e_old_coef, e_dpsi, e_err;
  size_t f_elements, f_nel, f_const;
  void init(i64 m, i64 me) {
    M = m;
    ME = me;
    e_thetas = 160;
    e_coef = e_thetas + 16 * M;
    e_old_coef = e_coef + 16 * M;
    e_dpsi = e_old_coef + 16 * M;
    e_err = e_dpsi + 16 * M;
    esz = e_err + 32;
    f_elements = 120;
    f_nel = f_elements + 8 * ME;
    f_const = f_nel + 8;
    fsz = f_const + 8;
  }
};

struct El {
  char* p;
  const Layout* L;
  i64& id() const { return *(i64*)(p + 0); }
  i64& type() const { return *(i64*)(p + 8); }
  i64& frac0() const { return *(i64*)(p + 16); }
  i64& frac1() const { return *(i64*)(p + 24); }
  cd* ep0() const { return (cd*)(p + 32); }
  cd* ep1() const { return (cd*)(p + 64); }
  double& radius() const { return *(double*)(p + 96); }
  cd& center() const { return *(cd*)(p + 104); }
  double& head() const { return *(double*)(p + 120); }
  double& phi() const { return *(double*)(p + 128); }
  i64& ncoef() const { return *(i64*)(p + 136); }
  i64& nint() const { return *(i64*)(p + 144); }
  double& q() const { return *(double*)(p + 152); }
  double* thetas() const { return (double*)(p + L->e_thetas); }
  cd* coef() const { return (cd*)(p + L->e_coef); }
  double* dpsi_corr() const { return (double*)(p + L->e_dpsi); }
  double& error() const { return *(double*)(p + L->e_err); }
  double& error_old() const { return *(double*)(p + L->e_err + 8); }
  double& error_old2() const { return *(double*)(p + L->e_err + 16); }
  double& error_coef() const { return *(double*)(p + L->e_err + 24); }
};

struct Fr {
  char* p;
  const Layout* L;
  i64& id() const { return *(i64*)(p + 0); }
  double& t() const { return *(double*)(p + 8); }
  i64* elements() const { return (i64*)(p + L->f_elements); }
  i64& nel() const { return *(i64*)(p + L->f_nel); }
  double& constant() const { return *(double*)(p + L->f_const); }
};

// per-element scratch of the reference's dtype_work (hpc_solve.py:26-46)
struct Work {
  std::vector<cd> coef, coef0, coef1, old_coef, exp_m, exp_p, integral;
  std::vector<double> phi, psi, dpsi;
  std::vector<i64> sign_array, discharge_element, element_pos;
  i64 len_discharge_element = 0;
  bool set_zero = false;
  void init(i64 M) {
    coef.assign(M, 0.0);
    coef0.assign(M, 0.0);
    coef1.assign(M, 0.0);
    old_coef.assign(M, 0.0);
    exp_m.assign(2 * M, 0.0);
    exp_p.assign(2 * M, 0.0);
    integral.assign(2 * M, 0.0);
    phi.assign(2 * M, 0.0);
    psi.assign(2 * M, 0.0);
    dpsi.assign(2 * M, 0.0);
  }
};

struct Oracle {
  Layout L;
  char* fr;
  i64 nf;
  char* el;
  i64 ne;
  i64 K;
  std::vector<Work> work;
  std::vector<i64> de;                 // discharge element ids (types 0, 2, 3) in element order
  std::vector<std::vector<cd>> z0, z1;  // discharge points per discharge element
  El E(i64 i) const { return El{el + (size_t)i * L.esz, &L}; }
  Fr F(i64 i) const { return Fr{fr + (size_t)i * L.fsz, &L}; }
};

// ---- maps and series ------------------------------------------------------------------
inline cd map_z_line_to_chi(cd z, const cd* ep) {
  cd big_z = (2.0 * z - ep[0] - ep[1]) / (ep[1] - ep[0]);
  return big_z + std::sqrt(big_z - 1.0) * std::sqrt(big_z + 1.0);
}

inline cd map_chi_to_z_line(cd chi, const cd* ep) {
  cd big_z = 0.5 * (chi + 1.0 / chi);
  return 0.5 * (big_z * (ep[1] - ep[0]) + ep[0] + ep[1]);
}

inline cd asym_expansion(cd chi, const cd* coef, i64 n, double sgn) {
  cd res(0.0, 0.0);
  for (i64 k = 0; k < n - 1; ++k) {
    res += sgn * coef[n - k - 1];
    res /= chi;
  }
  res += sgn * coef[0];
  return res;
}

inline cd taylor_series(cd chi, const cd* coef, i64 n) {
  cd res(0.0, 0.0);
  for (i64 k = 0; k < n - 1; ++k) {
    res += coef[n - k - 1];
    res *= chi;
  }
  res += coef[0];
  return res;
}

inline cd well_chi(cd chi, double q) { return q / (2 * PI) * std::log(chi); }

const double NaN = std::nan("");

// element potential at z seen from fracture frac_id (hpc_fracture.py:76-91)
inline cd element_omega(const El& e, cd z, i64 frac_id) {
  switch (e.type()) {
    case 0: {
      if (frac_id == e.frac0()) {
        cd chi = map_z_line_to_chi(z, e.ep0());
        return asym_expansion(chi, e.coef(), e.ncoef(), 1.0) + well_chi(chi, e.q());
      }
      cd chi = map_z_line_to_chi(z, e.ep1());
      return asym_expansion(chi, e.coef(), e.ncoef(), -1.0) + well_chi(chi, -e.q());
    }
    case 1: {
      cd chi = z / e.radius();
      return taylor_series(chi, e.coef(), e.ncoef());
    }
    case 2: {
      cd chi = (z - e.center()) / e.radius();
      cd om = well_chi(chi, e.q());
      if (std::abs(chi) < 1 - 1e-12) om = cd(NaN, NaN);
      return om;
    }
    case 3: {
      cd chi = map_z_line_to_chi(z, e.ep0());
      cd om = well_chi(chi, e.q());
      om += asym_expansion(chi, e.coef(), e.ncoef(), 1.0);
      return om;
    }
    case 4: {
      cd chi = (z - e.center()) / e.radius();
      if (std::abs(chi) < 1.0 - 1e-10) return cd(NaN, NaN);
      return asym_expansion(chi, e.coef(), e.ncoef(), 1.0);
    }
    case 5: {
      cd chi = map_z_line_to_chi(z, e.ep0());
      return asym_expansion(chi, e.coef(), e.ncoef(), 1.0);
    }
  }
  return cd(0.0, 0.0);
}

inline cd calc_omega(const Oracle& O, i64 f, cd z, i64 exclude) {
  Fr F = O.F(f);
  cd omega(F.constant(), 0.0);
  i64 n = F.nel();
  const i64* ids = F.elements();
  for (i64 k = 0; k < n; ++k) {
    i64 id = ids[k];
    if (id != exclude) omega += element_omega(O.E(id), z, F.id());
  }
  return omega;
}

// ---- complex discharge vector w ---------------------------------------------------------
// d/dchi of the asymptotic series: -sum_j j c_j chi^(-j-1), a division per term
// (hpc_math_functions.py:75-100)
inline cd asym_expansion_d1(cd chi, const cd* coef, i64 n, double sgn) {
  cd res(0.0, 0.0);
  for (i64 k = 0; k < n; ++k) {
    res -= sgn * coef[n - k - 1] * (double)(n - k - 1);
    res /= chi;
  }
  return res;
}

// d/dchi of the Taylor series: sum_j j c_j chi^(j-1) (hpc_math_functions.py:163-190)
inline cd taylor_series_d1(cd chi, const cd* coef, i64 n) {
  cd res(0.0, 0.0);
  if (n < 2) return res;  // a constant series (the reference would read coef[1] past the slice)
  for (i64 k = 0; k < n - 2; ++k) {
    res += coef[n - k - 1] * (double)(n - k - 1);
    res *= chi;
  }
  res += coef[1];
  return res;
}

// dchi/dz of the line map times the line terms: 2 chi^2 / (chi^2 - 1) * 2 / (ep1 - ep0)
inline cd line_jacobian(cd chi, const cd* ep) { return 2.0 * chi * chi / (chi * chi - 1.0) * 2.0 / (ep[1] - ep[0]); }

// w of one element seen from fracture frac_id.  jit: the jitted functions of the hpc modules,
// summed by hpc_fracture.calc_w (hpc_fracture.py:139-174): intersections, bounding circle, wells
// and constant-head lines only, the bounding circle as taylor_series_d1 alone
// (hpc_bounding_circle.py:97-115), wells without the inside test (hpc_well.py:89-108).
// !jit: the object model's calc_w (fracture.py:453-477), which streamline tracking uses
// (dfn.py:2676-2700): every element type, the bounding circle as -taylor_series_d1 / r
// (bounding.py:91-108), wells and impermeable circles NaN inside (well.py:129-153,
// impermeable_object.py:131-156), impermeable lines (impermeable_object.py:205-224).
inline cd element_w(const El& e, cd z, i64 frac_id, bool jit) {
  switch (e.type()) {
    case 0: {  // hpc_intersection.py:179-222, intersection.py:158-195
      if (frac_id == e.frac0()) {
        cd chi = map_z_line_to_chi(z, e.ep0());
        cd w = -asym_expansion_d1(chi, e.coef(), e.ncoef(), 1.0) - e.q() / (2 * PI * chi);
        return w * line_jacobian(chi, e.ep0());
      }
      cd chi = map_z_line_to_chi(z, e.ep1());
      cd w = -asym_expansion_d1(chi, e.coef(), e.ncoef(), -1.0) + e.q() / (2 * PI * chi);
      return w * line_jacobian(chi, e.ep1());
    }
    case 1: {
      cd chi = z / e.radius();
      if (jit) return taylor_series_d1(chi, e.coef(), e.ncoef());
      return -taylor_series_d1(chi, e.coef(), e.ncoef()) / e.radius();
    }
    case 2: {  // hpc_well.py:106-107, well.py:144-152
      cd chi = (z - e.center()) / e.radius();
      if (!jit && std::abs(chi) < 1.0 - 1e-10) return cd(NaN, NaN);
      return -e.q() / (2 * PI * chi) / e.radius();
    }
    case 3: {  // hpc_const_head_line.py:135-164, const_head.py:166-185
      cd chi = map_z_line_to_chi(z, e.ep0());
      cd w = -asym_expansion_d1(chi, e.coef(), e.ncoef(), 1.0) - e.q() / (2 * PI * chi);
      return w * line_jacobian(chi, e.ep0());
    }
    case 4: {
      if (jit) return cd(0.0, 0.0);
      cd chi = (z - e.center()) / e.radius();
      if (std::abs(chi) < 1.0 - 1e-10) return cd(NaN, NaN);
      return -asym_expansion_d1(chi, e.coef(), e.ncoef(), 1.0) / e.radius();
    }
    case 5: {
      if (jit) return cd(0.0, 0.0);
      cd chi = map_z_line_to_chi(z, e.ep0());
      return -asym_expansion_d1(chi, e.coef(), e.ncoef(), 1.0) * line_jacobian(chi, e.ep0());
    }
  }
  return cd(0.0, 0.0);
}

inline cd calc_w(const Oracle& O, i64 f, cd z, i64 exclude, bool jit) {
  Fr F = O.F(f);
  cd w(0.0, 0.0);
  i64 n = F.nel();
  const i64* ids = F.elements();
  for (i64 k = 0; k < n; ++k) {
    i64 id = ids[k];
    if (id != exclude) w += element_w(O.E(id), z, F.id(), jit);
  }
  return w;
}

// ---- collocation angles and fits ------------------------------------------------------
void calc_thetas(i64 n, i64 type, double* thetas) {
  double del = PI / n;
  double start = del / 2;
  if (type == 1) {
    del = 2 * PI / n;
    start = 0.0;
  }
  for (i64 i = 0; i < n; ++i) thetas[i] = start + i * del;
}

void fill_exp(i64 n, const double* th, cd* out, int sgn) {
  for (i64 i = 0; i < n; ++i) out[i] = cd(std::cos(th[i]), sgn * std::sin(th[i]));
}

void dft_powers(i64 n, i64 m, const double* vals, const Work& w, cd* integral) {
  for (i64 jj = 0; jj < m; ++jj) {
    cd res(0.0, 0.0);
    for (i64 ii = 0; ii < n; ++ii) {
      cd ev(1.0, 0.0);
      for (i64 k = 0; k < jj; ++k) ev *= w.exp_m[ii];
      res += vals[ii] * ev;
    }
    integral[jj] = res;
  }
}

void cauchy_integral_real(const Oracle& O, i64 n, i64 m, i64 f, i64 id, const cd* ep, Work& w, cd* coef) {
  for (i64 ii = 0; ii < n; ++ii) {
    cd z = map_chi_to_z_line(w.exp_p[ii], ep);
    w.phi[ii] = std::real(calc_omega(O, f, z, id));
  }
  dft_powers(n, m, w.phi.data(), w, w.integral.data());
  for (i64 ii = 0; ii < m; ++ii) coef[ii] = 2.0 * w.integral[ii] / (double)n;
  coef[0] = coef[0] / 2.0;
}

// psi unwrap shared by the circle and line stream-function fits (hpc_math_functions.py:503-514, 583-595)
void unwrap_psi(i64 n, const double* dpsi_corr, Work& w) {
  for (i64 i = 1; i < n; ++i) w.dpsi[i] = (w.psi[i] - w.psi[i - 1]) - dpsi_corr[i - 1];
  double psi0 = w.psi[0];
  for (i64 i = 0; i < n; ++i) {
    double psi1 = psi0 + w.dpsi[i];
    w.psi[i] = psi1;
    psi0 = psi1;
  }
}

void cauchy_integral_domega(const Oracle& O, i64 n, i64 m, const double* dpsi_corr, i64 f, i64 id, double radius,
                            cd center, Work& w, cd* coef) {
  for (i64 ii = 0; ii < n; ++ii) {
    cd z = w.exp_p[ii] * radius + center;
    w.psi[ii] = std::imag(calc_omega(O, f, z, id));
  }
  unwrap_psi(n, dpsi_corr, w);
  dft_powers(n, m, w.psi.data(), w, w.integral.data());
  for (i64 ii = 0; ii < m; ++ii) coef[ii] = cd(0.0, 2.0) * w.integral[ii] / (double)n;
  coef[0] = coef[0] / 2.0;
}

void cauchy_integral_domega_line(const Oracle& O, i64 n, i64 m, const double* dpsi_corr, i64 f, i64 id,
                                 const cd* ep, Work& w, cd* coef) {
  for (i64 ii = 0; ii < n; ++ii) {
    cd z = map_chi_to_z_line(w.exp_p[ii], ep);
    w.psi[ii] = std::imag(calc_omega(O, f, z, id));
  }
  unwrap_psi(n, dpsi_corr, w);
  dft_powers(n, m, w.psi.data(), w, w.integral.data());
  for (i64 ii = 0; ii < m; ++ii) coef[ii] = cd(0.0, 2.0) * w.integral[ii] / (double)n;
  coef[0] = coef[0] / 2.0;
}

// ---- error measures (hpc_math_functions.py:611-670) -------------------------------------
double calc_error(const cd* coef, const cd* ref, i64 n) {
  double s = 0.0;
  for (i64 i = 0; i < 5 && i < n; ++i) s += std::abs(ref[i]);
  if (s < 1e-15) return 0.0;
  cd c1(0.0, 0.0), c10(0.0, 0.0);
  for (i64 i = 0; i < n; ++i) c1 += coef[i];
  for (i64 i = (i64)(n * 0.9); i < n; ++i) c10 += coef[i];
  return std::fabs(std::abs(c1 - c10) / std::abs(c1 + 1e-16) - 1);
}

double calc_coef_error(const cd* coef, const cd* ref, i64 n) {
  double s = 0.0;
  for (i64 i = 0; i < 5 && i < n; ++i) s += std::abs(ref[i]);
  if (s < 1e-15) return 0.0;
  double err = 0.0;
  for (i64 i = 0; i < n; ++i) err += std::abs(coef[i] - ref[i]);
  return err / n;
}

// ---- branch cuts (hpc_math_functions.py:300-447) ----------------------------------------
inline double npsign(double x) { return (x > 0) - (x < 0); }

void record_cut(const El& self, Work& w, i64 ii, i64 eid, cd chi0, cd chi1, i64 sign_default) {
  i64 s = sign_default;
  if (self.type() != 1) s = (std::imag(chi0) < std::imag(chi1)) ? -sign_default : sign_default;
  w.element_pos.push_back(ii);
  w.discharge_element.push_back(eid);
  w.sign_array.push_back(s);
  w.len_discharge_element += 1;
}

void find_branch_cuts(const Oracle& O, const El& self, const std::vector<cd>& z_pos, Work& w) {
  w.len_discharge_element = 0;
  w.element_pos.clear();
  w.discharge_element.clear();
  w.sign_array.clear();
  Fr F = O.F(self.frac0());
  i64 nel = F.nel();
  for (i64 ii = 0; ii < self.nint() - 1; ++ii) {
    for (i64 k = 0; k < nel; ++k) {
      El e = O.E(F.elements()[k]);
      i64 t = e.type();
      if (t == 0) {
        bool own0 = e.frac0() == self.frac0();
        const cd* ep = own0 ? e.ep0() : e.ep1();
        cd chi0 = map_z_line_to_chi(z_pos[ii], ep);
        cd chi1 = map_z_line_to_chi(z_pos[ii + 1], ep);
        double ln0 = std::imag(std::log(chi0));
        double ln1 = std::imag(std::log(chi1));
        if (npsign(ln0) != npsign(ln1) && std::fabs(ln0) + std::fabs(ln1) > PI)
          record_cut(self, w, ii, e.id(), chi0, chi1, own0 ? -1 : 1);
      } else if (t == 2 || t == 3) {
        cd chi0, chi1;
        if (t == 2) {
          chi0 = (z_pos[ii] - e.center()) / e.radius();
          chi1 = (z_pos[ii + 1] - e.center()) / e.radius();
        } else {
          chi0 = map_z_line_to_chi(z_pos[ii], e.ep0());
          chi1 = map_z_line_to_chi(z_pos[ii + 1], e.ep0());
        }
        if (npsign(std::imag(chi0)) != npsign(std::imag(chi1)) && std::real(chi0) < 0)
          record_cut(self, w, ii, e.id(), chi0, chi1, -1);
      }
    }
  }
}

void get_dpsi_corr(const Oracle& O, const El& self, Work& w) {
  i64 n = self.nint();
  if (w.len_discharge_element == 0) {
    std::vector<cd> z_pos(n);
    if (self.type() == 1 || self.type() == 4) {
      for (i64 i = 0; i < n; ++i) z_pos[i] = w.exp_p[i] * self.radius() + self.center();
    } else {
      for (i64 i = 0; i < n; ++i) z_pos[i] = map_chi_to_z_line(w.exp_p[i], self.ep0());
    }
    find_branch_cuts(O, self, z_pos, w);
  }
  double* d = self.dpsi_corr();
  for (i64 i = 0; i < n - 1; ++i) d[i] = 0.0;
  for (i64 i = 0; i < w.len_discharge_element; ++i) {
    El e = O.E(w.discharge_element[i]);
    d[w.element_pos[i]] += e.q() * w.sign_array[i];
  }
}

// ---- element solves --------------------------------------------------------------------
void finish_errors(const El& self, Work& w) {
  i64 m = self.ncoef();
  self.error_old2() = self.error_old();
  self.error_old() = self.error();
  self.error() = calc_error(w.coef.data(), w.old_coef.data(), m);
  self.error_coef() = calc_coef_error(w.coef.data(), w.old_coef.data(), m);
}

void save_old(const El& self, Work& w) {
  for (i64 i = 0; i < self.ncoef(); ++i) w.old_coef[i] = self.coef()[i];
}

void solve_intersection(const Oracle& O, const El& self, Work& w) {
  double t0 = O.F(self.frac0()).t(), t1 = O.F(self.frac1()).t();
  i64 n = self.nint(), m = self.ncoef();
  save_old(self, w);
  cauchy_integral_real(O, n, m, self.frac0(), self.id(), self.ep0(), w, w.coef0.data());
  cauchy_integral_real(O, n, m, self.frac1(), self.id(), self.ep1(), w, w.coef1.data());
  for (i64 i = 0; i < m; ++i) w.coef[i] = std::real((t0 * w.coef1[i] - t1 * w.coef0[i]) / (t0 + t1));
  w.coef[0] = 0.0;
  finish_errors(self, w);
}

void solve_bounding(const Oracle& O, const El& self, Work& w) {
  get_dpsi_corr(O, self, w);
  save_old(self, w);
  i64 m = self.ncoef();
  cauchy_integral_domega(O, self.nint(), m, self.dpsi_corr(), self.frac0(), self.id(), self.radius(),
                         self.center(), w, w.coef.data());
  for (i64 i = 0; i < m; ++i) w.coef[i] = -w.coef[i];
  finish_errors(self, w);
}

void solve_chl(const Oracle& O, const El& self, Work& w) {
  save_old(self, w);
  i64 m = self.ncoef();
  cauchy_integral_real(O, self.nint(), m, self.frac0(), self.id(), self.ep0(), w, w.coef.data());
  for (i64 i = 0; i < m; ++i) w.coef[i] = -std::real(w.coef[i]);
  w.coef[0] = 0.0;
  finish_errors(self, w);
}

void solve_imp_circle(const Oracle& O, const El& self, Work& w) {
  get_dpsi_corr(O, self, w);
  save_old(self, w);
  i64 m = self.ncoef();
  cauchy_integral_domega(O, self.nint(), m, self.dpsi_corr(), self.frac0(), self.id(), self.radius(),
                         self.center(), w, w.coef.data());
  for (i64 i = 0; i < m; ++i) w.coef[i] = std::conj(w.coef[i]);
  w.coef[0] = 0.0;
  finish_errors(self, w);
}

void solve_imp_line(const Oracle& O, const El& self, Work& w) {
  get_dpsi_corr(O, self, w);
  save_old(self, w);
  i64 m = self.ncoef();
  cauchy_integral_domega_line(O, self.nint(), m, self.dpsi_corr(), self.frac0(), self.id(), self.ep0(), w,
                              w.coef.data());
  for (i64 i = 0; i < m; ++i) w.coef[i] = cd(0.0, -std::imag(w.coef[i]));
  w.coef[0] = 0.0;
  finish_errors(self, w);
}

void solve_element(const Oracle& O, i64 i, Work& w) {
  El e = O.E(i);
  switch (e.type()) {
    case 0: solve_intersection(O, e, w); break;
    case 1: solve_bounding(O, e, w); break;
    case 2: e.error() = 0.0; break;
    case 3: solve_chl(O, e, w); break;
    case 4: solve_imp_circle(O, e, w); break;
    case 5: solve_imp_line(O, e, w); break;
  }
}

void blend(const El& e, const Work& w, double damping) {
  cd* c = e.coef();
  for (i64 k = 0; k < e.ncoef(); ++k) {
    double re = damping * w.coef[k].real() + (1 - damping) * c[k].real();
    double im = damping * w.coef[k].imag() + (1 - damping) * c[k].imag();
    c[k] = cd(re, im);
  }
}

void prepare_angles(Oracle& O, i64 i) {
  El e = O.E(i);
  i64 n = e.nint();
  calc_thetas(n, e.type(), e.thetas());
  fill_exp(n, e.thetas(), O.work[i].exp_m.data(), -1);
  fill_exp(n, e.thetas(), O.work[i].exp_p.data(), 1);
}

// ---- discharge points and terms (hpc_intersection.py:13-46, hpc_const_head_line.py:14-35, 167-184,
//      hpc_well.py:14-35, 111-130) ----------------------------------------------------------
std::vector<cd> linspace_interior(cd a, cd b, i64 K) {
  std::vector<cd> z(K);
  cd step = (b - a) / (double)(K + 1);
  for (i64 i = 0; i < K; ++i) z[i] = a + (double)(i + 1) * step;
  return z;
}

std::vector<cd> well_points(const El& e, i64 K) {
  std::vector<cd> z(K);
  double stop = 2.0 * PI - 2 * PI / K;
  double step = K > 1 ? stop / (double)(K - 1) : 0.0;
  for (i64 i = 0; i < K; ++i) {
    double th = i * step;
    z[i] = e.radius() * cd(std::cos(th), std::sin(th)) + e.center();
  }
  return z;
}

double discharge_term(const El& e, const std::vector<cd>& z, i64 frac) {
  double phi = 0.0;
  if (e.type() == 0) {
    double sgn = 1.0;
    const cd* ep = e.ep0();
    if (frac != e.frac0()) {
      sgn = -1.0;
      ep = e.ep1();
    }
    for (const cd& zz : z) phi += std::real(well_chi(map_z_line_to_chi(zz, ep), sgn));
  } else if (e.type() == 2) {
    for (const cd& zz : z) phi += std::real(well_chi((zz - e.center()) / e.radius(), 1.0));
  } else if (e.type() == 3) {
    for (const cd& zz : z) phi += std::real(well_chi(map_z_line_to_chi(zz, e.ep0()), 1.0));
  } else {
    return 0.0;
  }
  return phi / (double)z.size();
}

inline bool is_discharge(i64 t) { return t == 0 || t == 2 || t == 3; }

}  // namespace

// =========================================================================================
extern "C" {

void* oracle_create(void* fractures, i64 nf, i64 max_elements, void* elements, i64 ne, i64 max_ncoef, i64 K) {
  Oracle* O = new Oracle();
  O->L.init(max_ncoef, max_elements);
  O->fr = (char*)fractures;
  O->nf = nf;
  O->el = (char*)elements;
  O->ne = ne;
  O->K = K;
  O->work.resize(ne);
  for (i64 i = 0; i < ne; ++i) O->work[i].init(max_ncoef);
  for (i64 i = 0; i < ne; ++i)
    if (is_discharge(O->E(i).type())) O->de.push_back(i);
  O->z0.resize(O->de.size());
  O->z1.resize(O->de.size());
  for (size_t j = 0; j < O->de.size(); ++j) {
    El e = O->E(O->de[j]);
    if (e.type() == 0) {
      O->z0[j] = linspace_interior(e.ep0()[0], e.ep0()[1], K);
      O->z1[j] = linspace_interior(e.ep1()[0], e.ep1()[1], K);
    } else if (e.type() == 2) {
      O->z0[j] = well_points(e, K);
    } else {
      O->z0[j] = linspace_interior(e.ep0()[0], e.ep0()[1], K);
    }
  }
  return O;
}

void oracle_destroy(void* h) { delete (Oracle*)h; }

i64 oracle_num_discharge(void* h) { return (i64)((Oracle*)h)->de.size(); }

// number of COO entries of the discharge matrix (count_discharge_nnz, hpc_solve.py:1061-1104)
i64 oracle_discharge_nnz(void* h) {
  Oracle& O = *(Oracle*)h;
  i64 nnz = 0;
  for (i64 id : O.de) {
    El e = O.E(id);
    i64 fr[2] = {e.frac0(), e.frac1()};
    int nsides = e.type() == 0 ? 2 : 1;
    for (int s = 0; s < nsides; ++s) {
      Fr F = O.F(fr[s]);
      for (i64 k = 0; k < F.nel(); ++k) {
        El ee = O.E(F.elements()[k]);
        if (ee.id() == e.id() || !is_discharge(ee.type())) continue;
        ++nnz;
      }
    }
    nnz += nsides;
  }
  for (i64 f = 0; f < O.nf; ++f) {
    Fr F = O.F(f);
    for (i64 k = 0; k < F.nel(); ++k)
      if (is_discharge(O.E(F.elements()[k]).type())) ++nnz;
  }
  return nnz;
}

// fill_discharge_matrix (hpc_solve.py:1117-1201); rows/cols/data sized by oracle_discharge_nnz
void oracle_discharge_matrix(void* h, i64* rows, i64* cols, double* data) {
  Oracle& O = *(Oracle*)h;
  i64 nde = (i64)O.de.size();
  std::vector<i64> id_to_pos(O.ne, -1);
  for (i64 j = 0; j < nde; ++j) id_to_pos[O.de[j]] = j;
  std::vector<i64> offs(nde + O.nf + 1, 0);
  for (i64 j = 0; j < nde; ++j) {
    El e = O.E(O.de[j]);
    i64 fr[2] = {e.frac0(), e.frac1()};
    int nsides = e.type() == 0 ? 2 : 1;
    i64 c = nsides;
    for (int s = 0; s < nsides; ++s) {
      Fr F = O.F(fr[s]);
      for (i64 k = 0; k < F.nel(); ++k) {
        El ee = O.E(F.elements()[k]);
        if (ee.id() != e.id() && is_discharge(ee.type())) ++c;
      }
    }
    offs[j + 1] = c;
  }
  for (i64 f = 0; f < O.nf; ++f) {
    Fr F = O.F(f);
    i64 c = 0;
    for (i64 k = 0; k < F.nel(); ++k)
      if (is_discharge(O.E(F.elements()[k]).type())) ++c;
    offs[nde + f + 1] = c;
  }
  for (size_t r = 1; r < offs.size(); ++r) offs[r] += offs[r - 1];
#pragma omp parallel for schedule(dynamic, 8)
  for (i64 j = 0; j < nde; ++j) {
    El e = O.E(O.de[j]);
    i64 ptr = offs[j];
    if (e.type() == 0) {
      i64 fr[2] = {e.frac0(), e.frac1()};
      double sg[2] = {1.0, -1.0};
      for (int s = 0; s < 2; ++s) {
        Fr F = O.F(fr[s]);
        const std::vector<cd>& z = s == 0 ? O.z0[j] : O.z1[j];
        for (i64 k = 0; k < F.nel(); ++k) {
          El ee = O.E(F.elements()[k]);
          if (ee.id() == e.id() || !is_discharge(ee.type())) continue;
          rows[ptr] = j;
          cols[ptr] = id_to_pos[ee.id()];
          data[ptr] = (sg[s] * discharge_term(ee, z, fr[s])) / F.t();
          ++ptr;
        }
      }
      rows[ptr] = j;
      cols[ptr] = nde + e.frac0();
      data[ptr] = 1.0 / O.F(e.frac0()).t();
      ++ptr;
      rows[ptr] = j;
      cols[ptr] = nde + e.frac1();
      data[ptr] = -1.0 / O.F(e.frac1()).t();
    } else {
      Fr F = O.F(e.frac0());
      for (i64 k = 0; k < F.nel(); ++k) {
        El ee = O.E(F.elements()[k]);
        if (ee.id() == e.id() || !is_discharge(ee.type())) continue;
        rows[ptr] = j;
        cols[ptr] = id_to_pos[ee.id()];
        data[ptr] = discharge_term(ee, O.z0[j], e.frac0());
        ++ptr;
      }
      rows[ptr] = j;
      cols[ptr] = nde + e.frac0();
      data[ptr] = 1.0;
    }
  }
  for (i64 f = 0; f < O.nf; ++f) {
    Fr F = O.F(f);
    i64 ptr = offs[nde + f];
    for (i64 k = 0; k < F.nel(); ++k) {
      El e = O.E(F.elements()[k]);
      if (!is_discharge(e.type())) continue;
      rows[ptr] = nde + f;
      cols[ptr] = id_to_pos[e.id()];
      data[ptr] = (e.type() != 0 || e.frac0() == F.id()) ? 1.0 : -1.0;
      ++ptr;
    }
  }
}

// error fields to 1e30 and collocation angles (hpc_solve.py:125-137)
void oracle_init(void* h) {
  Oracle& O = *(Oracle*)h;
  for (i64 i = 0; i < O.ne; ++i) {
    El e = O.E(i);
    e.error_old() = 1e30;
    e.error() = 1e30;
    prepare_angles(O, i);
  }
}

// y0 = [q of discharge elements, constants] (get_old_discharges, hpc_solve.py:764-775)
void oracle_old_discharges(void* h, double* y) {
  Oracle& O = *(Oracle*)h;
  i64 nde = (i64)O.de.size();
  for (i64 j = 0; j < nde; ++j) y[j] = O.E(O.de[j]).q();
  for (i64 f = 0; f < O.nf; ++f) y[nde + f] = O.F(f).constant();
}

// pre_matrix_solve + build_head_matrix (hpc_solve.py:666-717, 778-829)
void oracle_head_rhs(void* h, double* b) {
  Oracle& O = *(Oracle*)h;
  for (i64 i = 0; i < O.ne; ++i) O.E(i).q() = 0.0;
  for (i64 f = 0; f < O.nf; ++f) O.F(f).constant() = 0.0;
  i64 nde = (i64)O.de.size();
  i64 K = O.K;
#pragma omp parallel for schedule(dynamic, 8)
  for (i64 j = 0; j < nde; ++j) {
    El e = O.E(O.de[j]);
    cd om(0.0, 0.0);
    for (i64 i = 0; i < K; ++i) om += calc_omega(O, e.frac0(), O.z0[j][i], -1);
    om = om / (double)K;
    if (e.type() == 0) {
      cd om1(0.0, 0.0);
      for (i64 i = 0; i < K; ++i) om1 += calc_omega(O, e.frac1(), O.z1[j][i], -1);
      om1 = om1 / (double)K;
      b[j] = std::real(om1) / O.F(e.frac1()).t() - std::real(om) / O.F(e.frac0()).t();
    } else {
      b[j] = e.phi() - std::real(om);
    }
  }
  for (i64 f = 0; f < O.nf; ++f) b[nde + f] = 0.0;
}

// post_matrix_solve (hpc_solve.py:720-761)
void oracle_set_discharges(void* h, const double* y) {
  Oracle& O = *(Oracle*)h;
  i64 nde = (i64)O.de.size();
  for (i64 j = 0; j < nde; ++j) O.E(O.de[j]).q() = y[j];
  for (i64 f = 0; f < O.nf; ++f) O.F(f).constant() = y[nde + f];
}

// element_solver (hpc_solve.py:310-390)
void oracle_element_solver(void* h, double damping) {
  Oracle& O = *(Oracle*)h;
#pragma omp parallel for schedule(dynamic, 4)
  for (i64 i = 0; i < O.ne; ++i) solve_element(O, i, O.work[i]);
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < O.ne; ++i) blend(O.E(i), O.work[i], damping);
}

// element_solver2 (hpc_solve.py:393-591); returns the number of elements that grew
// force (optional): every element's ncoef after this iteration in a recorded run; when given,
// exactly the elements whose recorded count rises grow (trajectory replay)
i64 oracle_element_solver2(void* h, double max_error, i64 nit, double max_coef_ratio, i64 max_coef,
                           i64 coef_increase, double damping, const int32_t* force) {
  Oracle& O = *(Oracle*)h;
  oracle_element_solver(h, damping);
  if (nit < 2) return 0;
  const i64 ncc = 3;
  i64 grown = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : grown)
  for (i64 i = 0; i < O.ne; ++i) {
    Work& w = O.work[i];
    w.set_zero = false;
    El e = O.E(i);
    if (e.type() == 2) continue;
    i64 m = e.ncoef();
    const cd* c = e.coef();
    // the reference's first ratio is multiplied by zero (hpc_solve.py:474-478)
    double coef_ratio = 0.0;
    if (e.error() > e.error_old() && e.error() > e.error_old2() && e.error() > max_error && nit > 5)
      coef_ratio = max_coef_ratio + 1.0;
    cd tail(0.0, 0.0), head(0.0, 0.0);
    for (i64 k = m - ncc; k < m; ++k) tail += c[k];
    for (i64 k = 0; k < m - ncc; ++k) head += c[k];
    if (std::abs(tail / (head + 1e-30)) > max_coef_ratio && e.error() > max_error) coef_ratio = max_coef_ratio + 1.0;
    if (force) coef_ratio = (force[i] > e.ncoef()) ? max_coef_ratio + 1.0 : 0.0;
    if (coef_ratio > max_coef_ratio && e.ncoef() < max_coef && nit > 1) {
      e.ncoef() = e.ncoef() + coef_increase;
      e.nint() = e.ncoef() * 2;
      w.len_discharge_element = 0;
      prepare_angles(O, i);
      solve_element(O, i, w);
      ++grown;
    }
  }
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < O.ne; ++i) blend(O.E(i), O.work[i], damping);
  return grown;
}

// max element error and its index, max error_coef (get_error, hpc_solve.py:260-279, 197)
void oracle_ncoef(void* h, int32_t* out) {
  Oracle& O = *(Oracle*)h;
  for (i64 i = 0; i < O.ne; ++i) out[i] = (int32_t)O.E(i).ncoef();
}

void oracle_errors(void* h, double* out) {
  Oracle& O = *(Oracle*)h;
  double emax = O.E(0).error(), cmax = O.E(0).error_coef();
  i64 arg = 0;
  for (i64 i = 1; i < O.ne; ++i) {
    El e = O.E(i);
    double v = e.error();
    if (!std::isnan(emax) && (std::isnan(v) || v > emax)) {
      emax = v;
      arg = i;
    }
    double c = e.error_coef();
    if (!std::isnan(cmax) && (std::isnan(c) || c > cmax)) cmax = c;
  }
  out[0] = emax;
  out[1] = (double)arg;
  out[2] = cmax;
}

// Omega of fracture f at points z (all elements, with q and constants): heads and residual checks.
void oracle_omega_at(void* h, i64 n, const i64* frac, const cd* z, i64 exclude, cd* out) {
  Oracle& O = *(Oracle*)h;
#pragma omp parallel for schedule(dynamic, 64)
  for (i64 i = 0; i < n; ++i) out[i] = calc_omega(O, frac[i], z[i], exclude);
}

// w of fracture f at points z; jit != 0: hpc_fracture.calc_w, else Fracture.calc_w
void oracle_w_at(void* h, i64 n, const i64* frac, const cd* z, i64 exclude, i64 jit, cd* out) {
  Oracle& O = *(Oracle*)h;
#pragma omp parallel for schedule(dynamic, 64)
  for (i64 i = 0; i < n; ++i) out[i] = calc_w(O, frac[i], z[i], exclude, jit != 0);
}

// ---- direct access to the building blocks, for closed-form checks ----------------------
void oracle_t_map_z_line_to_chi(i64 n, const cd* z, const cd* ep, cd* out) {
  for (i64 i = 0; i < n; ++i) out[i] = map_z_line_to_chi(z[i], ep);
}
void oracle_t_map_chi_to_z_line(i64 n, const cd* chi, const cd* ep, cd* out) {
  for (i64 i = 0; i < n; ++i) out[i] = map_chi_to_z_line(chi[i], ep);
}
void oracle_t_asym(i64 n, const cd* chi, const cd* coef, i64 m, cd* out) {
  for (i64 i = 0; i < n; ++i) out[i] = asym_expansion(chi[i], coef, m, 1.0);
}
void oracle_t_taylor(i64 n, const cd* chi, const cd* coef, i64 m, cd* out) {
  for (i64 i = 0; i < n; ++i) out[i] = taylor_series(chi[i], coef, m);
}
void oracle_t_asym_d1(i64 n, const cd* chi, const cd* coef, i64 m, cd* out) {
  for (i64 i = 0; i < n; ++i) out[i] = asym_expansion_d1(chi[i], coef, m, 1.0);
}
void oracle_t_taylor_d1(i64 n, const cd* chi, const cd* coef, i64 m, cd* out) {
  for (i64 i = 0; i < n; ++i) out[i] = taylor_series_d1(chi[i], coef, m);
}
// the fit step of cauchy_integral_real (kind 0) / cauchy_integral_domega (kind 1) on given values
void oracle_t_fit(i64 n, i64 m, i64 type, const double* vals, i64 kind, cd* out) {
  Work w;
  w.init(std::max<i64>(n, m) + 1);
  std::vector<double> th(n);
  calc_thetas(n, type, th.data());
  fill_exp(n, th.data(), w.exp_m.data(), -1);
  std::vector<double> v(vals, vals + n);
  dft_powers(n, m, v.data(), w, w.integral.data());
  for (i64 j = 0; j < m; ++j) out[j] = (kind == 0 ? cd(2.0, 0.0) : cd(0.0, 2.0)) * w.integral[j] / (double)n;
  out[0] = out[0] / 2.0;
}
double oracle_t_calc_error(const cd* coef, const cd* ref, i64 n) { return calc_error(coef, ref, n); }
double oracle_t_calc_coef_error(const cd* coef, const cd* ref, i64 n) { return calc_coef_error(coef, ref, n); }

}  // extern "C"
