// DFN construction on the device: candidate pairs of disks, the chord two disks share, the
// constant-head chords of a boundary disk, and the connected components of the fracture graph.
//
// These give the solver's inputs (element endpoints in each fracture's local plane) as AnDFN's
// constructors do:
//   * pairs i < k tested in (i, k) order, the intersection chords created in that order
//     (andfn/dfn.py:1437-1455);
//   * the chord of two disks: line of the two planes, its intersections with both circles, the
//     first two of the four candidates [a0, b0, a1, b1] that lie in both disks
//     (andfn/geometry_functions.py:302-356, 359-402);
//   * shortening about the midpoint by SE_FACTOR (andfn/geometry_functions.py:974-993);
//   * the constant-head chords where a boundary disk cuts the fractures
//     (andfn/geometry_functions.py:878-926);
//   * the fracture graph's connected components (andfn/geometry_functions.py:1121-1294, the
//     connectivity test of andfn/dfn.py:1674-1695).
//
// The arithmetic follows the host builder operation by operation (this file is compiled without
// multiply-add contraction), so on the same inputs the chords agree with it to the last bit,
// except where the rim clipping takes |z| (a hypot, whose last bit the two libraries may round
// differently).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/andfn_mi.h"
#include "aem_kernels.hpp"

namespace aem {
namespace geo {

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 ld3(const double* p, int64_t i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
// Euclidean norm summed left to right
__device__ __forceinline__ double norm(V3 a) { return sqrt((a.x * a.x + a.y * a.y) + a.z * a.z); }

// complex numbers with the component operations of numpy's complex128 arithmetic
struct C {
  double re, im;
};
// x + 1j * y: the real y promoted to complex, multiplied by (0 + 1j), then added
__device__ __forceinline__ C make_xy(double x, double y) {
  const double pr = 0.0 * y - 1.0 * 0.0, pi = 0.0 * 0.0 + 1.0 * y;
  return C{x + pr, 0.0 + pi};
}
__device__ __forceinline__ C cadd(C a, C b) { return C{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ C csub(C a, C b) { return C{a.re - b.re, a.im - b.im}; }
// complex times a real promoted to complex (s + 0j)
__device__ __forceinline__ C cscale(C a, double s) { return C{a.re * s - a.im * 0.0, a.re * 0.0 + a.im * s}; }
// real times complex (lo + 0j)(d)
__device__ __forceinline__ C rscale(double s, C d) { return C{s * d.re - 0.0 * d.im, s * d.im + 0.0 * d.re}; }
// complex divided by 2 (the integer promoted to 2 + 0j; Smith's division with |re| >= |im|)
__device__ __forceinline__ C chalf(C a) {
  const double rat = 0.0 / 2.0, scl = 1.0 / (2.0 + 0.0 * rat);
  return C{(a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl};
}
// NaN-propagating max / min (numpy.maximum / numpy.minimum)
__device__ __forceinline__ double npmax(double a, double b) { return (a >= b || a != a) ? a : b; }
__device__ __forceinline__ double npmin(double a, double b) { return (a <= b || a != a) ? a : b; }

__device__ __forceinline__ C to2d(V3 p, V3 c, V3 xv, V3 yv) {
  const V3 d = sub(p, c);
  return make_xy(dot(d, xv), dot(d, yv));
}
__device__ __forceinline__ V3 to3d(C z, V3 c, V3 xv, V3 yv) {
  return V3{(z.re * xv.x + z.im * yv.x) + c.x, (z.re * xv.y + z.im * yv.y) + c.y, (z.re * xv.z + z.im * yv.z) + c.z};
}

// intersections of the line through z0, z1 with |z| = r (andfn/geometry_functions.py:359-402)
__device__ __forceinline__ bool line_circle(C z0, C z1, double r, C& za, C& zb) {
  const double dx = z1.re - z0.re, dy = z1.im - z0.im;
  const bool vertical = dx == 0.0;
  const double m = dy / (vertical ? 1.0 : dx);
  const double x0 = z0.im - m * z0.re;
  const double a = 1.0 + m * m;
  const double b = (2.0 * x0) * m;
  const double c = x0 * x0 - r * r;
  const double disc = b * b - (4.0 * a) * c;
  const double sq = sqrt(disc < 0.0 ? 0.0 : disc);
  const double x1 = (-b + sq) / (2.0 * a), x2 = (-b - sq) / (2.0 * a);
  const double y1 = m * x1 + x0, y2 = m * x2 + x0;
  const double xv = z0.re;
  const double yv = sqrt(r * r - xv * xv);
  za = vertical ? make_xy(xv, yv) : make_xy(x1, y1);
  zb = vertical ? make_xy(xv, -yv) : make_xy(x2, y2);
  return vertical || disc >= 0.0;
}

struct Disk {
  V3 c, n, x, y;
  double r;
};
__device__ __forceinline__ Disk disk(const DiskArrays& d, int64_t i) {
  return Disk{ld3(d.center, i), ld3(d.normal, i), ld3(d.xvec, i), ld3(d.yvec, i), d.radius[i]};
}

// the chord shared by disks A and B in both planes (andfn/geometry_functions.py:302-356)
__device__ bool chord(const Disk& A, const Disk& B, C e0[2], C e1[2]) {
  const V3 d = cross(A.n, B.n);
  const double d2 = dot(d, d);
  const bool par = d2 < 1e-14;
  const double d2s = par ? 1.0 : d2;
  const double cc0 = dot(A.n, A.c), cc1 = dot(B.n, B.c);
  const V3 u = cross(B.n, d), w = cross(d, A.n);
  const V3 p0{(cc0 * u.x + cc1 * w.x) / d2s, (cc0 * u.y + cc1 * w.y) / d2s, (cc0 * u.z + cc1 * w.z) / d2s};
  const V3 p1{p0.x + d.x, p0.y + d.y, p0.z + d.z};
  C a0, b0, a1, b1;
  const bool ok0 = line_circle(to2d(p0, A.c, A.x, A.y), to2d(p1, A.c, A.x, A.y), A.r, a0, b0);
  const bool ok1 = line_circle(to2d(p0, B.c, B.x, B.y), to2d(p1, B.c, B.x, B.y), B.r, a1, b1);
  const V3 cand[4] = {to3d(a0, A.c, A.x, A.y), to3d(b0, A.c, A.x, A.y), to3d(a1, B.c, B.x, B.y),
                      to3d(b1, B.c, B.x, B.y)};
  bool valid[4];
  int nv = 0;
  for (int j = 0; j < 4; ++j) {
    valid[j] = (norm(sub(cand[j], A.c)) <= A.r + 1e-8) && (norm(sub(cand[j], B.c)) <= B.r + 1e-8);
    nv += valid[j] ? 1 : 0;
  }
  // the first two rows of a stable argsort of ~valid: the valid candidates in index order, then
  // the others (only the first case gives a chord; the second keeps the values defined)
  int pick[2] = {0, 0}, k = 0;
  for (int j = 0; j < 4 && k < 2; ++j)
    if (valid[j]) pick[k++] = j;
  for (int j = 0; j < 4 && k < 2; ++j)
    if (!valid[j]) pick[k++] = j;
  const V3 q0 = cand[pick[0]], q1 = cand[pick[1]];
  e0[0] = to2d(q0, A.c, A.x, A.y);
  e0[1] = to2d(q1, A.c, A.x, A.y);
  e1[0] = to2d(q0, B.c, B.x, B.y);
  e1[1] = to2d(q1, B.c, B.x, B.y);
  return !par && ok0 && ok1 && nv == 2;
}

// shrink a segment about its midpoint (andfn/geometry_functions.py:974-993)
__device__ __forceinline__ void shorten(C e[2], double se) {
  const C m = chalf(cadd(e[0], e[1]));
  const C a = cadd(m, cscale(csub(e[0], m), se)), b = cadd(m, cscale(csub(e[1], m), se));
  e[0] = a;
  e[1] = b;
}

// the parameter interval [lo, hi] of the segment inside |z| <= rho r
__device__ __forceinline__ bool interval(const C e[2], double rr, double& lo, double& hi) {
  const C a = e[0], d = csub(e[1], e[0]);
  const double ad = hypot(d.re, d.im), aa = hypot(a.re, a.im);
  const double A = ad * ad;
  const double B = a.re * d.re - a.im * (-d.im);
  const double Cc = aa * aa - rr * rr;
  const double disc = B * B - A * Cc;
  const double sq = sqrt(npmax(disc, 0.0));
  const double As = A > 0.0 ? A : 1.0;
  lo = (-B - sq) / As;
  hi = (-B + sq) / As;
  return disc >= 0.0 && A > 0.0;
}

__device__ __forceinline__ void cut(C e[2], double lo, double hi) {
  const C d = csub(e[1], e[0]);
  const C a = cadd(e[0], rscale(lo, d)), b = cadd(e[0], rscale(hi, d));
  e[0] = a;
  e[1] = b;
}

// one thread per pair: chords of (A[ia[p]], B[ib[p]]), shortened and clipped as requested
__global__ void chords_kernel(int64_t np, const int64_t* __restrict__ ia, const int64_t* __restrict__ ib,
                              DiskArrays A, DiskArrays B, ChordOptions opt, double* __restrict__ e0,
                              double* __restrict__ e1, int8_t* __restrict__ ok) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const Disk da = disk(A, ia[p]), db = disk(B, ib[p]);
  C s0[2], s1[2];
  bool good = chord(da, db, s0, s1);
  if (opt.shorten_a) shorten(s0, opt.se_factor);
  if (opt.shorten_b) shorten(s1, opt.se_factor);
  if (opt.clip == 1 || opt.clip == 2) {
    double lo = 0.0, hi = 0.0;
    bool okc;
    if (opt.clip == 1) {  // both disks: the same fraction of the chord in both planes
      double lo1 = 0.0, hi1 = 0.0;
      okc = interval(s0, opt.rho * da.r, lo, hi);
      const bool ok1 = interval(s1, opt.rho * db.r, lo1, hi1);
      lo = npmax(lo, lo1);
      hi = npmin(hi, hi1);
      okc = okc && ok1;
    } else {  // disk B only
      okc = interval(s1, opt.rho * db.r, lo, hi);
    }
    lo = npmin(npmax(lo, 0.0), 1.0);
    hi = npmin(npmax(hi, 0.0), 1.0);
    okc = okc && (hi - lo > 1e-9);
    if (opt.clip == 1) cut(s0, lo, hi);
    cut(s1, lo, hi);
    good = good && okc;
  }
  for (int j = 0; j < 2; ++j) {
    e0[4 * p + 2 * j] = s0[j].re;
    e0[4 * p + 2 * j + 1] = s0[j].im;
    e1[4 * p + 2 * j] = s1[j].re;
    e1[4 * p + 2 * j + 1] = s1[j].im;
  }
  ok[p] = good ? 1 : 0;
}

// bounding spheres of disks i < k meet: |c_i - c_k| <= r_i + r_k (the radii carry the tolerance)
__device__ __forceinline__ bool near(const double* c, const double* r, int64_t i, int64_t k) {
  const double dx = c[3 * i] - c[3 * k], dy = c[3 * i + 1] - c[3 * k + 1], dz = c[3 * i + 2] - c[3 * k + 2];
  return sqrt((dx * dx + dy * dy) + dz * dz) <= r[i] + r[k];
}

constexpr int kPairThreads = 256;

// row i: number of k > i whose bounding spheres meet disk i's
__global__ __launch_bounds__(kPairThreads) void pair_count_kernel(int64_t n, const double* __restrict__ c,
                                                                  const double* __restrict__ r,
                                                                  int64_t* __restrict__ count) {
  __shared__ int64_t part[kPairThreads / 64];
  const int64_t i = blockIdx.x;
  int64_t cnt = 0;
  for (int64_t k = i + 1 + threadIdx.x; k < n; k += kPairThreads) cnt += near(c, r, i, k) ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int w = 0; w < kPairThreads / 64; ++w) s += part[w];
    count[i] = s;
  }
}

// row i: the pairs (i, k) in increasing k from offset[i] (a block-wide ordered compaction per
// chunk of kPairThreads columns)
__global__ __launch_bounds__(kPairThreads) void pair_write_kernel(int64_t n, const double* __restrict__ c,
                                                                  const double* __restrict__ r,
                                                                  const int64_t* __restrict__ offset,
                                                                  int64_t* __restrict__ pi, int64_t* __restrict__ pk) {
  __shared__ int wcount[kPairThreads / 64];
  const int64_t i = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int64_t o = offset[i];
  for (int64_t k0 = i + 1; k0 < n; k0 += kPairThreads) {
    const int64_t k = k0 + threadIdx.x;
    const bool hit = k < n && near(c, r, i, k);
    const unsigned long long m = __ballot(hit);
    if (lane == 0) wcount[wv] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < kPairThreads / 64; ++w) {
      if (w < wv) before += wcount[w];
      total += wcount[w];
    }
    if (hit) {
      const int pos = before + __popcll(m & ((1ull << lane) - 1ull));
      pi[o + pos] = i;
      pk[o + pos] = k;
    }
    o += total;
    __syncthreads();
  }
}

}  // namespace geo

void geo_chords(int64_t np, const int64_t* ia, const int64_t* ib, const DiskArrays& A, const DiskArrays& B,
                const ChordOptions& opt, double* e0, double* e1, int8_t* ok, hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(geo::chords_kernel, dim3((unsigned)((np + 127) / 128)), dim3(128), 0, st, np, ia, ib, A, B, opt,
                     e0, e1, ok);
}

void geo_pair_count(int64_t n, const double* c, const double* r, int64_t* count, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(geo::pair_count_kernel, dim3((unsigned)n), dim3(geo::kPairThreads), 0, st, n, c, r, count);
}

void geo_pair_write(int64_t n, const double* c, const double* r, const int64_t* offset, int64_t* pi, int64_t* pk,
                    hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(geo::pair_write_kernel, dim3((unsigned)n), dim3(geo::kPairThreads), 0, st, n, c, r, offset, pi,
                     pk);
}

}  // namespace aem

// ---- C ABI (host) ---------------------------------------------------------------------------
namespace {

#define GEO_HIP(x)                                                                          \
  do {                                                                                      \
    hipError_t _e = (x);                                                                    \
    if (_e != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(_e)); \
  } while (0)

struct GeoError : std::runtime_error {
  int code;
  GeoError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <class Fn>
int geo_guard(Fn&& fn) {
  try {
    return fn();
  } catch (const GeoError& e) {
    aem::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    aem::set_last_error(e.what());
    return ANDFN_MI_EHIP;
  }
}

// device copy of a host array, freed with the holder
template <class T>
struct Dev {
  T* p = nullptr;
  Dev() = default;
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  explicit Dev(size_t n) { GEO_HIP(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T))); }
  Dev(const T* h, size_t n, hipStream_t st) : Dev(n) {
    if (n) GEO_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};

void use_device(int device) {
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) throw GeoError(ANDFN_MI_ENODEV, "no HIP device");
  if (device < 0 || device >= nd) throw GeoError(ANDFN_MI_EINVAL, "device out of range");
  GEO_HIP(hipSetDevice(device));
}

struct DevDisks {
  Dev<double> c, n, x, y, r;
  DevDisks(const andfn_mi_disks* d, hipStream_t st)
      : c(d->center, 3 * (size_t)d->n, st), n(d->normal, 3 * (size_t)d->n, st), x(d->x_vector, 3 * (size_t)d->n, st),
        y(d->y_vector, 3 * (size_t)d->n, st), r(d->radius, (size_t)d->n, st) {}
  aem::DiskArrays view() const { return aem::DiskArrays{c.p, n.p, x.p, y.p, r.p}; }
};

void check_disks(const andfn_mi_disks* d) {
  if (!d || d->n < 0 || (d->n > 0 && (!d->center || !d->normal || !d->x_vector || !d->y_vector || !d->radius)))
    throw GeoError(ANDFN_MI_EINVAL, "invalid disk arrays");
}

}  // namespace

extern "C" {

int andfn_mi_candidate_pairs(int device, const andfn_mi_disks* disks, double tol, int64_t cap, int64_t* pair_i,
                             int64_t* pair_k, int64_t* count) {
  return geo_guard([&]() {
    check_disks(disks);
    if (!count || cap < 0 || (cap > 0 && (!pair_i || !pair_k))) throw GeoError(ANDFN_MI_EINVAL, "invalid output arrays");
    *count = 0;
    const int64_t n = disks->n;
    if (n < 2) return ANDFN_MI_OK;
    use_device(device);
    hipStream_t st = nullptr;
    GEO_HIP(hipStreamCreate(&st));
    std::unique_ptr<ihipStream_t, void (*)(hipStream_t)> hold(st, [](hipStream_t s) { (void)hipStreamDestroy(s); });
    std::vector<double> rt(n);
    for (int64_t i = 0; i < n; ++i) rt[i] = disks->radius[i] + tol;
    Dev<double> dc(disks->center, 3 * (size_t)n, st), dr(rt.data(), (size_t)n, st);
    Dev<int64_t> cnt((size_t)n);
    aem::geo_pair_count(n, dc.p, dr.p, cnt.p, st);
    GEO_HIP(hipGetLastError());
    std::vector<int64_t> hc(n), off(n);
    GEO_HIP(hipMemcpyAsync(hc.data(), cnt.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
    GEO_HIP(hipStreamSynchronize(st));
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) {
      off[i] = total;
      total += hc[i];
    }
    *count = total;
    if (total > cap) throw GeoError(ANDFN_MI_ECAPACITY, "more candidate pairs than the output holds");
    if (total == 0) return ANDFN_MI_OK;
    Dev<int64_t> doff(off.data(), (size_t)n, st), di((size_t)total), dk((size_t)total);
    aem::geo_pair_write(n, dc.p, dr.p, doff.p, di.p, dk.p, st);
    GEO_HIP(hipGetLastError());
    GEO_HIP(hipMemcpyAsync(pair_i, di.p, sizeof(int64_t) * total, hipMemcpyDeviceToHost, st));
    GEO_HIP(hipMemcpyAsync(pair_k, dk.p, sizeof(int64_t) * total, hipMemcpyDeviceToHost, st));
    GEO_HIP(hipStreamSynchronize(st));
    return ANDFN_MI_OK;
  });
}

int andfn_mi_disk_chords(int device, int64_t n, const int64_t* ia, const int64_t* ib, const andfn_mi_disks* a,
                         const andfn_mi_disks* b, int32_t shorten, double se_factor, int32_t clip, double rho,
                         double* e_a, double* e_b, int8_t* ok) {
  return geo_guard([&]() {
    check_disks(a);
    check_disks(b);
    if (n < 0 || (n > 0 && (!ia || !ib || !e_a || !e_b || !ok))) throw GeoError(ANDFN_MI_EINVAL, "invalid pair arrays");
    if (clip < 0 || clip > 2) throw GeoError(ANDFN_MI_EINVAL, "clip must be 0, 1 or 2");
    for (int64_t p = 0; p < n; ++p)
      if (ia[p] < 0 || ia[p] >= a->n || ib[p] < 0 || ib[p] >= b->n)
        throw GeoError(ANDFN_MI_EINVAL, "pair index out of range");
    if (n == 0) return ANDFN_MI_OK;
    use_device(device);
    hipStream_t st = nullptr;
    GEO_HIP(hipStreamCreate(&st));
    std::unique_ptr<ihipStream_t, void (*)(hipStream_t)> hold(st, [](hipStream_t s) { (void)hipStreamDestroy(s); });
    DevDisks da(a, st), db(b, st);
    Dev<int64_t> dia(ia, (size_t)n, st), dib(ib, (size_t)n, st);
    Dev<double> dea(4 * (size_t)n), deb(4 * (size_t)n);
    Dev<int8_t> dok((size_t)n);
    aem::ChordOptions opt{(shorten & 1) ? 1 : 0, (shorten & 2) ? 1 : 0, clip, se_factor, rho};
    aem::geo_chords(n, dia.p, dib.p, da.view(), db.view(), opt, dea.p, deb.p, dok.p, st);
    GEO_HIP(hipGetLastError());
    GEO_HIP(hipMemcpyAsync(e_a, dea.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, st));
    GEO_HIP(hipMemcpyAsync(e_b, deb.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, st));
    GEO_HIP(hipMemcpyAsync(ok, dok.p, (size_t)n, hipMemcpyDeviceToHost, st));
    GEO_HIP(hipStreamSynchronize(st));
    return ANDFN_MI_OK;
  });
}

int andfn_mi_components(int64_t n, int64_t m, const int64_t* a, const int64_t* b, int64_t* label) {
  return geo_guard([&]() {
    if (n < 0 || m < 0 || (m > 0 && (!a || !b)) || (n > 0 && !label)) throw GeoError(ANDFN_MI_EINVAL, "invalid graph arrays");
    // union-find with path halving and union by size
    std::vector<int64_t> parent(n), size(n, 1);
    std::iota(parent.begin(), parent.end(), (int64_t)0);
    auto find = [&](int64_t v) {
      while (parent[v] != v) {
        parent[v] = parent[parent[v]];
        v = parent[v];
      }
      return v;
    };
    for (int64_t j = 0; j < m; ++j) {
      if (a[j] < 0 || a[j] >= n || b[j] < 0 || b[j] >= n) throw GeoError(ANDFN_MI_EINVAL, "edge index out of range");
      int64_t ra = find(a[j]), rb = find(b[j]);
      if (ra == rb) continue;
      if (size[ra] < size[rb]) std::swap(ra, rb);
      parent[rb] = ra;
      size[ra] += size[rb];
    }
    std::vector<int64_t> lab(n, -1);
    int64_t next = 0;
    for (int64_t v = 0; v < n; ++v) {
      const int64_t r = find(v);
      if (lab[r] < 0) lab[r] = next++;
      label[v] = lab[r];
    }
    return ANDFN_MI_OK;
  });
}

}  // extern "C"
