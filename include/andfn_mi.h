/*
 * andfn_mi.h -- C ABI of the MI355X analytic-element DFN solve.
 *
 * The boundary replaces AnDFN's HPC entry point
 *
 *     andfn.hpc.hpc_solve.solve(fracture_struc_array, element_struc_array, discharge_int,
 *                               constants, ntype_elements)          (andfn/hpc/hpc_solve.py:55-257)
 *
 * whose only caller is DFN.solve (andfn/dfn.py:1840-1846).  The record arrays are passed as
 * they are in memory: packed numpy structured arrays with the layouts
 *
 *     element_dtype_hpc   andfn/element.py:45-69    (12992 bytes per record at MAX_NCOEF = 200)
 *     fracture_dtype_hpc  andfn/element.py:146-159  (4136 bytes per record at MAX_ELEMENTS = 500)
 *     dtype_constants     andfn/constants.py:13-29  (0-d record, packed, 76 bytes)
 *
 * and are updated in place exactly as the reference updates them: element coef, ncoef, nint,
 * thetas, q, dpsi_corr, error, error_old, error_old2, error_coef; fracture constant.
 *
 * All entry points return 0 on success, a positive value for a warning (not converged) and a
 * negative ANDFN_MI_E* code on failure; andfn_mi_last_error() gives the message (per thread).
 * Calls are blocking: results are in the caller's arrays when the call returns.
 */
#ifndef ANDFN_MI_H
#define ANDFN_MI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANDFN_MI_OK 0
#define ANDFN_MI_NOT_CONVERGED 1          /* reached MAX_ITERATIONS (the reference returns silently) */
#define ANDFN_MI_EINVAL (-1)              /* invalid argument or inconsistent records */
#define ANDFN_MI_ESINGULAR (-2)           /* discharge matrix is singular (reference: SuperLU RuntimeError) */
#define ANDFN_MI_ECAPACITY (-3)           /* capacity exceeded (reference: IndexError) */
#define ANDFN_MI_EHIP (-4)                /* HIP / rocSOLVER failure */
#define ANDFN_MI_ENODEV (-5)              /* no GPU device */

/* Arithmetic of the solve. */
#define ANDFN_MI_FP64 0   /* every kernel in FP64 (the reference's arithmetic) */
#define ANDFN_MI_MIXED 1  /* FP32 far-field series, FP64 near field, LU, residual and refinement */

/* Per-iteration discharge solve with the LU factors of the constant discharge matrix. */
#define ANDFN_MI_LU_TRSV 0     /* row permutation + blocked triangular solves (the library's own kernels) */
#define ANDFN_MI_LU_GETRS 1    /* rocSOLVER getrs */
#define ANDFN_MI_LU_INVERSE 2  /* explicit inverse formed once (getri), then one gemv per iteration */

/* andfn_mi_options.factor: how the discharge system is factored (lu_mode 0 only; the other
 * lu_modes always use rocSOLVER).  The reference factors the sparse matrix with SuperLU
 * (hpc_solve.py:121, splu of the csc_matrix of :885). */
#define ANDFN_MI_FACTOR_SPARSE 0     /* multifrontal sparse LU on a nested-dissection ordering of the fracture graph
                                      * (the default; falls back to the tournament LU for a network its solve
                                      * kernels cannot take) */
#define ANDFN_MI_FACTOR_ROCSOLVER 1  /* dense: rocSOLVER getrf (partial pivoting; strided-batched for several parts) */
#define ANDFN_MI_FACTOR_TOURNAMENT 2 /* dense: the library's blocked LU with tournament pivoting (all parts of the
                                      * network in one batch) */

typedef struct andfn_mi_options {
  int32_t precision;     /* ANDFN_MI_FP64 or ANDFN_MI_MIXED */
  int32_t device;        /* HIP device ordinal */
  int32_t verbose;       /* >0: per-iteration line on stderr, as hpc_solve.py:200-215 logs */
  int32_t lu_mode;       /* ANDFN_MI_LU_* */
  double far_field;      /* mixed form: |Z| above which a line source is evaluated in FP32 (0: default 4) */
  /* Optional coefficient-count schedule: row k (k = 0, 1, ...) holds every element's ncoef after
   * iteration k + 1.  When given, the growth step of iteration k + 1 grows exactly the elements
   * whose count rises in row k instead of applying the growth test (rows beyond
   * schedule_iterations fall back to the test).  Used to replay another run's trajectory. */
  const int32_t* growth_schedule;
  int64_t schedule_iterations;
  /* Optional output: every element's ncoef after each iteration, record_iterations rows. */
  int32_t* growth_record;
  int64_t record_iterations;
  /* >0: stop as soon as the head residual (below) of the current iterate is <= head_tol metres,
   * instead of AnDFN's coefficient/discharge-change rule; it is evaluated at the start of every
   * iteration from the new right-hand side.  MAX_ITERATIONS still bounds the solve. */
  double head_tol;
  int32_t factor;        /* ANDFN_MI_FACTOR_* */
  int32_t reserved;      /* zero */
  /* Optional output: the fields of AnDFN's per-iteration log line (hpc_solve.py:198-216), one row of
   * ANDFN_MI_LOG_FIELDS doubles per iteration for log_iterations rows: error, error_coef, error_q,
   * the element of the largest error, its ncoef at that iteration and its type. */
  double* iteration_log;
  int64_t log_iterations;
} andfn_mi_options;

#define ANDFN_MI_LOG_FIELDS 6

typedef struct andfn_mi_info {
  int64_t nit;            /* outer iterations taken (hpc_solve.py:152) */
  int64_t n_unknowns;     /* discharge elements + fractures */
  int64_t n_grown;        /* coefficient-growth events over the solve */
  int64_t max_error_element;
  double error;           /* max element error of the last iteration */
  double error_coef;      /* max element coefficient change of the last iteration */
  double error_q;         /* relative discharge change of the last iteration */
  double t_setup;         /* seconds: upload, tables, discharge-matrix assembly */
  double t_factor;        /* seconds: LU factorisation */
  double t_iterate;       /* seconds: the fixed-point iterations */
  double t_total;         /* seconds: whole call, records in to records out */
  double k_sweep_ms;      /* summed device time of the collocation sweep kernel */
  double k_rhs_ms;        /* summed device time of the head right-hand side (basis product or sweep) and residual */
  double k_fit_ms;        /* summed device time of the fit kernel */
  double k_lu_ms;         /* summed device time of the per-iteration discharge solve */
  int64_t n_sweep;        /* launches of the collocation sweep kernel */
  int64_t n_rhs;          /* launches of the right-hand-side sweep kernel */
  int64_t pairs_sweep;    /* target-point x source pairs evaluated by all collocation sweeps */
  int64_t pairs_rhs;      /* right-hand-side basis entries applied (or point x source pairs swept) */
  double flops_sweep;     /* algorithmic flops of all collocation sweeps (line 8m+61, circle 8m-6, well 24) */
  double flops_rhs;       /* algorithmic flops of all right-hand-side evaluations */
  /* head residual of the returned state, metres: max over the discharge (head) equations of
   * |A y - b(c)|, i.e. how far the returned discharges and constants are from satisfying the
   * averaged head conditions of the returned coefficients (intersection rows are head
   * differences; constant-head and well rows are divided by the fracture's t) */
  double head_residual;
  int64_t head_residual_row;
  /* pointwise boundary-condition residual, metres: max over intersection collocation points of
   * |h_f0 - h_f1|, over constant-head-line collocation points and well discharge points of
   * |h - h_prescribed|; bounded below by the series truncation, not by the iteration */
  double bc_residual;
  int64_t nit_fp32;       /* mixed form: iterations run with FP32 far field and fit before the FP64 phase */
  int64_t solves_fp32;    /* mixed form: discharge solves done by refinement with the FP32-rounded factors
                           * (the rest use the FP64 factors: after the refinement stopped contracting) */
  double refine_rate;     /* mixed form: measured contraction per refinement step with the FP32-rounded
                           * factors (above 0.25 the run solves with the FP64 factors throughout) */
  int32_t sweep_points_per_lane;  /* tile shape the autotuner chose for the collocation sweep (1 or 2) */
  int32_t rhs_points_per_lane;    /* and for the right-hand-side sweep (0: the basis product ran instead) */
  int64_t nnz_discharge;  /* structural nonzeros of the discharge matrix (hpc_solve.py:140-141 logs getnnz()) */
  int64_t nnz_factors;    /* entries of the stored LU factors of the discharge matrix (L and U together) */
  int32_t factor_used;    /* ANDFN_MI_FACTOR_* that factored the last run's matrix */
  int32_t factor_rounds;  /* multifrontal LU: factorisations of the last run (more than one when unstable
                           * pivots were delayed to parent fronts and the tree analysed again) */
  int64_t solve_passes;   /* passes through the stored factors by the iterations' discharge solves (a
                           * refined solve makes two, a correction solve x = y + A^-1 (b - A y) one) */
  int64_t solve_values;   /* values one pass streams (multifrontal: the packed panel tiles, FP32 in the
                           * mixed form's correction solves; dense: the n^2 entries of the blocks) */
  int64_t runahead_iterations;  /* iterations enqueued one ahead of the host's read of the head residual,
                                 * stopped on the device (head tolerance, no coefficient growth) */
  int64_t graph_iterations;     /* of those, iterations run as one HIP graph launch each */
  int32_t sweep_persistent;     /* 1: the autotuner chose one persistent launch over the XCD tile lists for
                                 * the collocation sweep (0: one workgroup per tile) */
  int32_t rhs_persistent;       /* and for the right-hand-side sweep */
} andfn_mi_info;

/*
 * Drop-in for andfn.hpc.hpc_solve.solve (andfn/hpc/hpc_solve.py:55-79).
 *   fractures, n_fractures, max_elements : fracture_struc_array, len(), elements[] capacity
 *   elements, n_elements, max_ncoef      : element_struc_array, len(), coef[] capacity
 *   discharge_int                        : number of discharge-integration points K
 *   constants                            : the 0-d dtype_constants record
 *   ntype_elements                       : int64[6] element counts per type (unused, as in the reference)
 *   opt                                  : NULL for defaults (FP64, device 0)
 *   info                                 : NULL or receives iteration count, errors and timings
 */
int andfn_mi_solve(void* fractures, int64_t n_fractures, int64_t max_elements, void* elements,
                   int64_t n_elements, int64_t max_ncoef, int64_t discharge_int, const void* constants,
                   const int64_t* ntype_elements, const andfn_mi_options* opt, andfn_mi_info* info);

/*
 * The same solve in three steps, for callers that solve one network repeatedly (benchmarks,
 * parameter sweeps): create parses the records and uploads them to the device (the records must
 * stay alive until destroy); run solves from the uploaded initial state on the device, and may be
 * called again (every run restarts from that state); download writes the last run's results into
 * the records given to create.  andfn_mi_solve is create + run + download + destroy.
 */
typedef struct andfn_mi_handle andfn_mi_handle;
int andfn_mi_create(void* fractures, int64_t n_fractures, int64_t max_elements, void* elements,
                    int64_t n_elements, int64_t max_ncoef, int64_t discharge_int, const void* constants,
                    int32_t device, andfn_mi_handle** handle);
int andfn_mi_run(andfn_mi_handle* handle, const andfn_mi_options* opt, andfn_mi_info* info);
int andfn_mi_download(andfn_mi_handle* handle);
void andfn_mi_destroy(andfn_mi_handle* handle);

/*
 * Hydraulic heads at points of the solved network: h = Re Omega_f(z) / t_f
 * (andfn/hpc/hpc_fracture.py:204-250, get_heads 302-343).  z is interleaved complex
 * (re, im) in each fracture's local plane; frac gives the fracture index of every point.
 */
int andfn_mi_heads(const void* fractures, int64_t n_fractures, int64_t max_elements, const void* elements,
                   int64_t n_elements, int64_t max_ncoef, int64_t n_points, const int64_t* frac, const double* z,
                   double* heads, int32_t device);

/*
 * Complex potential at points of the solved network: Omega_f(z) = C_f + the sum of every element of
 * fracture f (andfn/hpc/hpc_fracture.py:51-92 calc_omega; the flow nets of calc_flow_net 178-201 and
 * get_flow_nets 254-299 evaluate it on a sunflower pattern per fracture).  z and omega are
 * interleaved complex (re, im); frac gives the fracture index of every point.  Re Omega is the
 * discharge potential Phi (head times t), Im Omega the stream function Psi.
 */
int andfn_mi_omega(const void* fractures, int64_t n_fractures, int64_t max_elements, const void* elements,
                   int64_t n_elements, int64_t max_ncoef, int64_t n_points, const int64_t* frac, const double* z,
                   double* omega, int32_t device);

/* Forms of the discharge vector (andfn_mi_w, andfn_mi_state_w). */
#define ANDFN_MI_W_JIT 0     /* andfn/hpc/hpc_fracture.py:139-174 calc_w: intersections, bounding circle, wells and
                                constant-head lines; the bounding circle as taylor_series_d1(z / r) alone
                                (hpc_bounding_circle.py:97-115), wells without an inside test (hpc_well.py:89-108) */
#define ANDFN_MI_W_OBJECT 1  /* andfn/fracture.py:453-477 Fracture.calc_w, the form streamline tracking uses: every
                                element type, the bounding circle as -taylor_series_d1(z / r) / r (bounding.py:91-108),
                                NaN inside wells and impermeable circles (well.py:129-153,
                                impermeable_object.py:131-224) */

/*
 * Complex discharge vector w = -dOmega/dz at points of the solved network (Q_x - i Q_y), every element
 * of fracture f but element `exclude` (-1: none) summed, in the given form.  z and w are interleaved
 * complex; frac gives the fracture index of every point.  Replaces the per-point Python loops of
 * hpc_fracture.calc_w and Fracture.calc_w.
 */
int andfn_mi_w(const void* fractures, int64_t n_fractures, int64_t max_elements, const void* elements,
               int64_t n_elements, int64_t max_ncoef, int64_t n_points, const int64_t* frac, const double* z,
               int64_t exclude, int32_t form, double* w, int32_t device);

/*
 * Implicit adjoint of the converged fixed point.  No counterpart in the reference (AnDFN has no
 * gradients); it is the gradient companion of andfn_mi_run.  For the objective
 *
 *     J = sum_e dq[e] q_e + sum_f dconstant[f] C_f + sum_k dhead[k] h(obs_z[k] in fracture obs_frac[k])
 *
 * at the state of the last andfn_mi_run on the handle, it returns dJ/dt_f for every fracture
 * transmissivity and dJ/dhead_e for every constant-head element (constant-head lines and wells;
 * 0 for other elements), with the coefficient counts frozen at their converged values.  The
 * adjoint equations are the transpose of the fixed-point iteration (collocation sweep, fit,
 * discharge solve): (I - B^T) lam = g with B the iteration's linear map on the coefficients.
 * They are solved with the transposed operators (A^-T from the run's FP64 LU factors) by
 * restarted GMRES or by the forward's damped iteration.  Returns
 * ANDFN_MI_NOT_CONVERGED if max_iterations is reached first.
 */
typedef struct andfn_mi_objective {
  const double* dq;         /* [n_elements] weights of the element discharges, or NULL */
  const double* dconstant;  /* [n_fractures] weights of the fracture constants, or NULL */
  int64_t n_obs;            /* observation points */
  const int64_t* obs_frac;  /* [n_obs] fracture index of each point */
  const double* obs_z;      /* [2 n_obs] interleaved (re, im) in the fracture's local plane */
  const double* dhead;      /* [n_obs] weights of the heads at the points */
} andfn_mi_objective;

#define ANDFN_MI_ADJ_GMRES 0      /* restarted GMRES on (I - B^T) lam = g */
#define ANDFN_MI_ADJ_FIXED_POINT 1 /* the forward's damped iteration, transposed */

typedef struct andfn_mi_adjoint_options {
  double tol;               /* GMRES: relative residual; fixed point: max |change| / max |adjoint| (0: 1e-10) */
  int64_t max_iterations;   /* operator applications (0: 5000) */
  double damping;           /* fixed point: 0 means 0.75, the forward's effective damping (hpc_solve.py:383-388, 561-575) */
  int32_t verbose;          /* >0: progress lines on stderr */
  int32_t method;           /* ANDFN_MI_ADJ_* */
  /* 0: refuse (ANDFN_MI_EINVAL) a state whose run did not meet its stop rule, since the implicit
   * adjoint is the gradient of the fixed point only; nonzero: accept it (e.g. a run of a fixed
   * iteration count known to be converged) */
  int32_t accept_unconverged;
  int32_t reserved;         /* zero */
} andfn_mi_adjoint_options;

typedef struct andfn_mi_adjoint_info {
  int64_t nit;              /* adjoint iterations */
  double delta;             /* relative change of the adjoint at the last iteration */
  double objective;         /* J at the solved state */
  double t_setup;           /* seconds: point tables, transposed factors, objective terms */
  double t_iterate;         /* seconds: the adjoint iterations */
  double t_total;           /* seconds: whole call */
} andfn_mi_adjoint_info;

int andfn_mi_adjoint(andfn_mi_handle* handle, const andfn_mi_objective* objective,
                     const andfn_mi_adjoint_options* opt, double* grad_t, double* grad_head,
                     andfn_mi_adjoint_info* info);

/*
 * Parameters of a handle's network (no counterpart in the reference, where they are fixed when
 * the DFN is built: const_head.py:45, well.py:38).  t[n_fractures] (> 0) replaces the fracture
 * transmissivities and head[n_elements] the prescribed heads of the constant-head elements (other
 * entries are ignored); phi of a constant-head element becomes head * t of its fracture.  NULL
 * leaves a parameter as it is.  The next run assembles and factors the new discharge matrix; the
 * caller's records are not changed.
 */
int andfn_mi_set_parameters(andfn_mi_handle* handle, const double* t, const double* head);

/* Heads at points of the handle's current state (after a run), as andfn_mi_heads. */
int andfn_mi_state_heads(andfn_mi_handle* handle, int64_t n_points, const int64_t* frac, const double* z,
                         double* heads);

/* Complex potential at points of the handle's current state (after a run), as andfn_mi_omega. */
int andfn_mi_state_omega(andfn_mi_handle* handle, int64_t n_points, const int64_t* frac, const double* z,
                         double* omega);

/* Discharge vector at points of the handle's current state, as andfn_mi_w. */
int andfn_mi_state_w(andfn_mi_handle* handle, int64_t n_points, const int64_t* frac, const double* z,
                     int64_t exclude, int32_t form, double* w);

/*
 * Streamline tracking on the handle's current state: DFN.streamline_tracking (andfn/dfn.py:2638-2725)
 * for n_lines lines at once, line i from z0[i] (interleaved complex, local coordinates) in fracture
 * frac[i] with elevation[i].  In a fracture, predictor-corrector steps of ds times the fracture radius
 * along the object model's w (runge_kutta, dfn.py:2779-2834; backward != 0 against the flow) until a
 * step crosses a discharge element of the fracture (check_streamline_exit, dfn.py:2727-2742:
 * intersections, constant-head lines, wells, in the fracture's element order), w turns NaN or the
 * segment holds more than max_length - (points recorded so far) + 1 points; through an intersection
 * the line continues in the neighbouring fracture, on the side its elevation selects
 * (get_exit_intersection, dfn.py:2744-2777), otherwise it ends.  Speeds are |w| / aperture[fracture]
 * (fracture.py:479-497), one positive aperture per fracture.  A line also ends after
 * max(max_length, 1) + 1 segments (the reference's loop has no such bound).
 * Returns per line the number of points, of segments (one per fracture visited) and the element it
 * left through last (-1: none, the reference's False).  The points themselves stay in the handle
 * until andfn_mi_state_streamlines_fetch (the next call replaces them).
 */
int andfn_mi_state_streamlines(andfn_mi_handle* handle, int64_t n_lines, const int64_t* frac, const double* z0,
                               const double* elevation, const double* aperture, double ds, int32_t backward,
                               int64_t max_length, int64_t* n_points, int64_t* n_segments, int64_t* exit_element);

/*
 * The points of the last andfn_mi_state_streamlines call, line after line: points (interleaved
 * complex, sum of n_points values) and speed (one per point); per segment, line after line, its
 * fracture and its number of points (sum of n_segments values).  A segment's first point is where it
 * starts recording (the reference's z_start), its last the exit point.
 */
int andfn_mi_state_streamlines_fetch(andfn_mi_handle* handle, double* points, double* speed,
                                     int64_t* segment_fracture, int64_t* segment_points);

/* andfn_mi_state_heads without allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The points, their heads and the point tiles live in a fixed device
 * scratch region of the library (2.75 MiB per GPU, held until the process ends; no hipMalloc,
 * hipFree or allocating library call is made), so one call takes at most 65,536 points
 * (ANDFN_MI_ECAPACITY beyond).  Calls are serialised by the library; the scratch keeps the last
 * call's contents. */
int andfn_mi_state_heads_noalloc(andfn_mi_handle* handle, int64_t n_points, const int64_t* frac, const double* z,
                                 double* heads);

/* Discharges q[n_elements] and fracture constants constant[n_fractures] of the handle's current
 * state (either pointer may be NULL), without writing the records. */
int andfn_mi_state_discharges(andfn_mi_handle* handle, double* q, double* constant);

/* Number of visible HIP devices (0 when none). */
int andfn_mi_device_count(void);

/* Message of the last failure on the calling thread. */
/* ---- DFN construction ----------------------------------------------------------------------
 * The input side of the solve: the intersections AnDFN's DFN.get_fracture_intersections creates
 * (andfn/dfn.py:1393-1458: every pair i < k in (i, k) order, the chord of
 * andfn/geometry_functions.py:302-356 with the line/circle roots of :359-402, shortened by
 * shorten_line :974-993), the constant-head lines of set_constant_head_boundary
 * (andfn/dfn.py:1504-1553, geometry_functions.py:878-926), and the connectivity of the fracture
 * graph (andfn/dfn.py:1674-1695, geometry_functions.py:1121-1294). */

/* Disks as structure of arrays: centre, unit normal and in-plane basis vectors as [n][3], radius
 * [n] (the Fracture fields of andfn/fracture.py:55-75). */
typedef struct andfn_mi_disks {
  int64_t n;
  const double* center;
  const double* normal;
  const double* x_vector;
  const double* y_vector;
  const double* radius;
} andfn_mi_disks;

/* Candidate pairs i < k whose bounding spheres meet, |c_i - c_k| <= (r_i + tol) + (r_k + tol), in
 * (i, k) order: every pair AnDFN's pair loop (andfn/dfn.py:1437-1455) can find intersecting.  The
 * pairs go to pair_i / pair_k (capacity cap) and their number to *count; more than cap pairs
 * fail with ANDFN_MI_ECAPACITY after *count is set. */
int andfn_mi_candidate_pairs(int device, const andfn_mi_disks* disks, double tol, int64_t cap, int64_t* pair_i,
                             int64_t* pair_k, int64_t* count);

/* For p < n, the chord shared by disk ia[p] of a and disk ib[p] of b (andfn/geometry_functions.py
 * :302-356): its endpoints in a's plane (e_a[p], two complex numbers as four doubles) and in b's
 * plane (e_b[p]); ok[p] = 1 where the disks cut each other in a chord.  shorten: bit 0 shrinks e_a,
 * bit 1 e_b about the midpoint by se_factor (geometry_functions.py:974-993).  clip (not in AnDFN;
 * 0 = off): 1 keeps the part of the chord inside rho r of both disks (the same fraction in both
 * planes), 2 the part inside rho r of disk b; ok is cleared where nothing remains. */
int andfn_mi_disk_chords(int device, int64_t n, const int64_t* ia, const int64_t* ib, const andfn_mi_disks* a,
                         const andfn_mi_disks* b, int32_t shorten, double se_factor, int32_t clip, double rho,
                         double* e_a, double* e_b, int8_t* ok);

/* Connected components of the graph of n nodes and m edges (a[j], b[j]) -- the fracture graph,
 * whose edges are the intersections: label[v] is the component of node v, the components
 * numbered in the order of their smallest node.  Runs on the host. */
int andfn_mi_components(int64_t n, int64_t m, const int64_t* a, const int64_t* b, int64_t* label);

const char* andfn_mi_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* ANDFN_MI_H */
