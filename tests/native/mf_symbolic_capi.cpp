// Host-only C interface to the multifrontal symbolic analysis (andfn_amd/csrc/mf_symbolic.cpp),
// built with g++ by tests/test_multifrontal_symbolic.py to check its tables on the CPU.
#include <cstring>
#include <stdexcept>
#include <string>

#include "mf_symbolic.hpp"

static aem::MfSymbolic g_sym;
static std::string g_err;

extern "C" {

int mf_test_analyse(int F, int E, const int* etype, const int* f0, const int* f1, const int* de_pos, const int* fpos,
                    const double* center, long long N, const int* ptr, const int* col, const int* delay,
                    int panel_rows) {
  aem::MfInput in;
  in.F = F;
  in.E = E;
  in.etype = etype;
  in.efrac0 = f0;
  in.efrac1 = f1;
  in.de_pos = de_pos;
  in.fpos = fpos;
  in.center = center;
  in.N = N;
  in.csr_ptr = ptr;
  in.csr_col = col;
  in.delay = delay;
  in.panel_rows = panel_rows;
  try {
    aem::mf_analyse(in, g_sym);
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  }
  return 0;
}

const char* mf_test_error() { return g_err.c_str(); }

// sizes: 0 fronts, 1 levels, 2 idx, 3 rel, 4 piv0, 5 gptr, 6 gsrc, 7 small_list, 8 unk, 9 tiles,
// 10 (unused), 11 (unused), 12 ext, 13 asmap, 14 front_total, 15 inv_total, 16 nvec, 17 nnz_factors,
// 18 nunk, 19 gv, 20 gv_total
long long mf_test_size(int w) {
  const aem::MfSymbolic& s = g_sym;
  switch (w) {
    case 0: return (long long)s.fronts.size();
    case 1: return (long long)s.levels.size();
    case 2: return (long long)s.idx.size();
    case 3: return (long long)s.rel.size();
    case 4: return (long long)s.piv0.size();
    case 5: return (long long)s.gptr.size();
    case 6: return (long long)s.gsrc.size();
    case 7: return (long long)s.small_list.size();
    case 8: return (long long)s.unk.size();
    case 9: return (long long)s.tiles.size();
    case 12: return (long long)s.ext.size();
    case 13: return (long long)s.asmap.size();
    case 14: return s.front_total;
    case 15: return s.inv_total;
    case 16: return s.nvec;
    case 17: return s.nnz_factors;
    case 18: return s.nunk;
    case 19: return (long long)s.gv.size();
    case 20: return s.gv_total;
  }
  return 0;
}

// fronts as 11 int64 each: off, inv, m, k, idx, vec, piv, rel, parent, gat, 0
void mf_test_fronts(long long* out) {
  for (size_t i = 0; i < g_sym.fronts.size(); ++i) {
    const aem::MfFront& f = g_sym.fronts[i];
    long long* o = out + 11 * i;
    o[0] = f.off; o[1] = f.inv; o[2] = f.m; o[3] = f.k; o[4] = f.idx; o[5] = f.vec; o[6] = f.piv; o[7] = f.rel;
    o[8] = f.parent; o[9] = f.gat; o[10] = 0;
  }
}

// levels as 28 ints each: small_beg[4], small_end[4], tile_beg, tile_end, number of big fronts,
// number of extend slots, gv_beg[4], gv_end[4], prow[4], pwide[4]
void mf_test_levels(int* out) {
  for (size_t h = 0; h < g_sym.levels.size(); ++h) {
    const aem::MfLevel& L = g_sym.levels[h];
    int* o = out + 28 * h;
    for (int c = 0; c < 4; ++c) { o[c] = L.small_beg[c]; o[4 + c] = L.small_end[c]; }
    o[8] = L.tile_beg; o[9] = L.tile_end; o[10] = (int)L.big.size(); o[11] = (int)L.ext_beg.size();
    for (int c = 0; c < 4; ++c) {
      o[12 + c] = L.gv_beg[c]; o[16 + c] = L.gv_end[c]; o[20 + c] = L.prow[c]; o[24 + c] = L.pwide[c];
    }
  }
}

void mf_test_level_big(int h, int* out) { std::memcpy(out, g_sym.levels[h].big.data(), sizeof(int) * g_sym.levels[h].big.size()); }
void mf_test_level_ext(int h, int* beg, int* end) {
  const aem::MfLevel& L = g_sym.levels[h];
  for (size_t s = 0; s < L.ext_beg.size(); ++s) { beg[s] = L.ext_beg[s]; end[s] = L.ext_end[s]; }
}

void mf_test_ints(int w, int* out) {
  const aem::MfSymbolic& s = g_sym;
  const std::vector<int>* v[] = {nullptr, nullptr, &s.idx, &s.rel, &s.piv0, &s.gptr, &s.gsrc, &s.small_list, &s.unk};
  if (w >= 2 && w <= 8) std::memcpy(out, v[w]->data(), sizeof(int) * v[w]->size());
  else if (w == 9) std::memcpy(out, s.tiles.data(), sizeof(int) * 4 * s.tiles.size());
  else if (w == 12) std::memcpy(out, s.ext.data(), sizeof(int) * 2 * s.ext.size());
  else if (w == 13) std::memcpy(out, s.gv.data(), sizeof(int) * 4 * s.gv.size());
}

void mf_test_gvoff(long long* out) {
  for (size_t i = 0; i < g_sym.gv_off.size(); ++i) out[i] = g_sym.gv_off[i];
}

void mf_test_asmap(long long* out) {
  for (size_t i = 0; i < g_sym.asmap.size(); ++i) out[i] = g_sym.asmap[i];
}

}  // extern "C"
