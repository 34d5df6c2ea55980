// Symbolic analysis of the multifrontal LU of the discharge system (mf_symbolic.cpp): the
// nested-dissection ordering of the fracture graph, the elimination tree of dense fronts and every
// table the device kernels of aem_sparse.hip index.  Plain C++ (no device code), so the same
// analysis can be built on the host alone and its tables checked against their bounds.
//
// Pivoting is restricted to a front's own rows, so a front's pivot block must be nonsingular and
// well conditioned.  Two mechanisms delay pivots to a parent front: on the structure (a maximum
// matching of the block's own entries) here, and on the values (pivots whose column of L21 grows
// beyond a threshold, reported by the factorisation) through MfInput::delay and a new analysis.
#ifndef ANDFN_MF_SYMBOLIC_HPP
#define ANDFN_MF_SYMBOLIC_HPP

#include <cstdint>
#include <vector>

namespace aem {

constexpr int kMfSmallMax = 128;    // fronts with at most this many pivots: one workgroup in LDS ...
constexpr int kMfSmallMaxB = 512;   // ... if they have more than 64, only with at most this many boundary rows
constexpr int kMfSmall128Min = 16;  // ... and only on a level with at least this many such fronts
// (the one-workgroup kernel stages the boundary of a 65-128 pivot front 24 rows or columns at a
// time, one barrier per pivot and tile, and pays off when a level holds many such fronts, as the
// 64-realisation batch does with 29, 130, 156 and 57 on four levels: 104 against 73 ms per
// factorisation; cfg3's levels hold at most 7, for which the large-front streams are faster:
// 104 against 108 ms)
inline bool mf_is_small(int k, int b) { return k <= 64 || (k <= kMfSmallMax && b <= kMfSmallMaxB); }
constexpr int kMfLeafUnknowns = 48;  // a fracture set is split while its interior has more unknowns

struct MfFront {
  int64_t off;  // first value of the front: m x m, column-major, leading dimension m
  int64_t inv;  // first value of L11^-1 (k x k, column-major) and, k^2 later, U11^-1
  int m, k;     // front size and pivots (the first k of its index list)
  int idx;      // first entry of its index list: m unknowns (pivots, then the boundary)
  int vec;      // first entry of its solve vector (m values)
  int piv;      // first entry of its pivot order (k local rows: position i holds row piv[i])
  int rel;      // first entry of the parent-local positions of its boundary (m - k entries)
  int parent;   // front index of the parent, -1 at a root
  int gat;      // first entry of its gather pointers (m + 1): children's contributions per position
};

// layouts of the device's int2 / int4 (the tables are uploaded as such)
struct MfI2 {
  int x, y;
};
struct MfI4 {
  int x, y, z, w;
};

// symbolic input: the network's discharge unknowns in the solver's layout (positions < N)
struct MfInput {
  int F = 0, E = 0;
  const int* etype = nullptr;
  const int* efrac0 = nullptr;
  const int* efrac1 = nullptr;
  const int* de_pos = nullptr;     // [E] unknown of each discharge element, -1 otherwise
  const int* fpos = nullptr;       // [F] unknown of each fracture constant
  const double* center = nullptr;  // [F][3] fracture centres
  int64_t N = 0;                   // size of the solver's unknown layout (padding included)
  const int* csr_ptr = nullptr;    // [N + 1] pattern of A in that layout (structurally symmetric,
  const int* csr_col = nullptr;    //  columns sorted within a row)
  const int* delay = nullptr;      // [N] or null: levels to lift each unknown above its node (delayed pivots)
  int panel_rows = 0;              // 0: solve panel heights chosen per level; 4, 8 or 16: that height for kinds 1-3
};

struct MfLevel {
  int small_beg[4], small_end[4];     // small-front lists per pivot class (<= 16, 32, 64, 128)
  std::vector<int> big;               // fronts that are not small (mf_is_small, kMfSmall128Min)
  int tile_beg, tile_end;             // Schur tiles of its small fronts
  std::vector<int> ext_beg, ext_end;  // extend-add items per child slot
  int gv_beg[4], gv_end[4];           // solve panels, per kind (mf_panel_shape / mf_panel_columns)
  int prow[4], pwide[4];              // per kind: panel rows (16, 8 or 4); 1 if a panel takes 1,024 threads
};

// The products of the solve A x = b are cut, per level and kind, into row panels of R rows (R =
// 16, 8 or 4, chosen per level and kind so that a launch has enough panels to spread over the
// device; launches of very few panels keep 16 rows and take 1,024 threads per panel): kind 0
// L11^-1 (k x k, lower), 1 L21 (b x k), 2 U12 (k x b), 3 U11^-1 (k x k, upper).
// A panel spans the column blocks (C = 512 / R columns each) its rows need (a triangular kind:
// only those that meet the triangle), so one workgroup computes each of its rows in full.  After
// each factorisation every panel is copied into a tile of its own (512 values per column block,
// within a block the value of row r0 + l and column 2 p + e at (p R + l) 2 + e, zeros outside the
// block and its triangle), so that a solve streams each panel as one contiguous run of memory.
constexpr int kMfBlockValues = 512;
constexpr int kMfWidePanels = 1024;  // launches of fewer panels than this (and long ones) go wide
inline void mf_panel_shape(int kind, int k, int b, int* rows, int* cols) {
  *rows = kind == 1 ? b : k;
  *cols = kind == 2 ? b : k;
}
// first column (a multiple of C = 512 / R) and number of column blocks of the panel at row r0
inline void mf_panel_columns(int kind, int k, int b, int R, int r0, int* c_lo, int* nblk) {
  int rows, cols;
  mf_panel_shape(kind, k, b, &rows, &cols);
  const int C = kMfBlockValues / R;
  const int hi = kind == 0 ? (r0 + R < cols ? r0 + R : cols) : cols;
  *c_lo = kind == 3 ? r0 / C * C : 0;
  *nblk = (hi - *c_lo + C - 1) / C;
}

struct MfSymbolic {
  std::vector<MfFront> fronts;        // in postorder: children before parents
  std::vector<MfLevel> levels;        // by height: leaves first
  std::vector<int> idx;               // index lists of all fronts
  std::vector<int> rel;               // parent-local positions of every front's boundary
  std::vector<int> piv0;              // identity pivot order (reset before each factorisation)
  std::vector<int> gptr, gsrc;        // per front position: the children's solve-vector entries
  std::vector<int> small_list;        // fronts of the small classes, level by level
  std::vector<int> unk;               // every analysed unknown
  std::vector<MfI4> tiles;            // (front, first boundary row, first boundary column, 0)
  std::vector<MfI2> ext;              // extend-add items (child, first boundary column)
  std::vector<MfI4> gv;               // solve panels (front, first row, first column, column blocks)
  std::vector<int64_t> gv_off;        // per solve panel: first value of its tile (kMfBlockValues per column block)
  std::vector<int64_t> asmap;         // CSR entry -> value position in its front
  int64_t nunk = 0, front_total = 0, inv_total = 0, nnz_factors = 0, gv_total = 0;
  int nvec = 0, max_m = 0, big_k = 0;
  double flops = 0.0;                 // factorisation flops (2/3 k^3 + 2 k^2 b + 2 k b^2 per front)
};

// Throws std::runtime_error when a table would not fit the kernels (a front beyond their LDS) or
// the pattern is not the discharge matrix's (not structurally symmetric).
void mf_analyse(const MfInput& in, MfSymbolic& out);

}  // namespace aem

#endif
