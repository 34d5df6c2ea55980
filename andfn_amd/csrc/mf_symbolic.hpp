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

constexpr int kMfSmallMax = 64;     // fronts with at most this many pivots: one workgroup in LDS
constexpr int kMfLeafUnknowns = 48;  // a fracture set is split while its interior has more unknowns
constexpr int kMfMaxLds = 61440;    // bytes of dynamic LDS of the solve kernels (a level's largest k or m - k; their static LDS is 2 KB)

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
};

struct MfLevel {
  int small_beg[3], small_end[3];     // small-front lists per pivot class (<= 16, 32, 64)
  std::vector<int> big;               // fronts with more than kMfSmallMax pivots
  int tile_beg, tile_end;             // Schur tiles of its small fronts
  int ck_beg, ck_end;                 // 64-row chunks over the pivot rows of its fronts
  int cb_beg, cb_end;                 // 64-row chunks over the boundary rows
  int maxk, maxb;                     // largest pivot count and boundary of its fronts
  std::vector<int> ext_beg, ext_end;  // extend-add items per child slot
  int gv_beg[4], gv_end[4];           // solve products, per kind (mf_gemv_shape / mf_gemv_segments)
  int gr_beg[4], gr_end[4];           // chunks of more than one segment, per kind (their partials to sum)
};

// The products of the solve A x = b, per level and kind, are cut into items of 64 rows by
// kMfSeg columns: kind 0 L11^-1 (k x k, lower), 1 L21 (b x k), 2 U12 (k x b), 3 U11^-1 (k x k,
// upper).  A 64-row chunk's items are consecutive; a triangular kind has only the segments that
// meet the triangle.  An item's first column is a multiple of kMfSeg and carries the chunk's
// segment count in its low bits (at most kMfMaxLds / 8 / kMfSeg = 30 segments).  After each
// factorisation every item's block is copied into a tile of its own (64 rows, the segment's columns
// rounded up to even, column pairs interleaved per row, zeros outside the block and its triangle),
// so that a solve streams each tile as one contiguous run of memory.
constexpr int kMfSeg = 256;
constexpr int kMfTileRows = 64;
inline void mf_gemv_shape(int kind, int k, int b, int* rows, int* cols) {
  *rows = kind == 1 ? b : k;
  *cols = kind == 2 ? b : k;
}
// first and one-past-last segment of the chunk starting at row r0
inline void mf_gemv_segments(int kind, int k, int b, int r0, int* s_lo, int* s_hi) {
  int rows, cols;
  mf_gemv_shape(kind, k, b, &rows, &cols);
  const int nseg = (cols + kMfSeg - 1) / kMfSeg;
  const int last_row = r0 + 63 < rows - 1 ? r0 + 63 : rows - 1;
  *s_lo = kind == 3 ? r0 / kMfSeg : 0;
  *s_hi = kind == 0 ? (last_row / kMfSeg + 1 < nseg ? last_row / kMfSeg + 1 : nseg) : nseg;
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
  std::vector<MfI2> ck, cb, ext;      // (front, first row) chunks; (child, first column) items
  std::vector<MfI4> gv;               // solve items (front, first row, first column + the chunk's segment count, its first item)
  std::vector<int64_t> gv_off;        // per solve item: first value of its packed tile (kMfTileRows x even column count)
  std::vector<int> gr;                // first item of every chunk of more than one segment
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
