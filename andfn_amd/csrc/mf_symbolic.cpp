// Symbolic analysis of the multifrontal LU of the discharge system (declarations in
// mf_symbolic.hpp).  Nested dissection of the fracture graph: a fracture set is cut at the median
// of the fracture centres along its longest extent until its interior holds at most
// kMfLeafUnknowns unknowns; an intersection whose fractures fall on two sides belongs to the
// separator of the smallest set holding both, a constant-head line or well to its fracture's leaf,
// and a fracture constant to the deepest node holding one of its fracture's discharge elements,
// so that it is eliminated together with a discharge of its fracture (the constant's own
// continuity row has no entry in its column: hpc_solve.py:1187-1201).
#include "mf_symbolic.hpp"

#include <algorithm>
#include <atomic>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>

namespace aem {

namespace {

struct Node {
  int parent = -1, depth = 0;
  std::vector<int> frs;  // fractures (leaves only)
  std::vector<int> U;    // pivot unknowns (positions in the solver's layout)
};

}  // namespace

void mf_analyse(const MfInput& in, MfSymbolic& out) {
  out = MfSymbolic{};
  const int F = in.F, E = in.E;
  const int64_t N = in.N;
  if (in.panel_rows != 0 && in.panel_rows != 4 && in.panel_rows != 8 && in.panel_rows != 16)
    throw std::runtime_error("multifrontal analysis: panel rows must be 0, 4, 8 or 16");
  const int* rp = in.csr_ptr;
  const int* cl = in.csr_col;
  const int64_t nnz = N > 0 ? rp[N] : 0;
  auto is_de = [&](int e) { return in.etype[e] == 0 || in.etype[e] == 2 || in.etype[e] == 3; };
  // per fracture: its intersections and its other discharge elements
  std::vector<std::vector<int>> fx(F), floc(F);
  for (int e = 0; e < E; ++e) {
    if (!is_de(e)) continue;
    if (in.etype[e] == 0) {
      fx[in.efrac0[e]].push_back(e);
      fx[in.efrac1[e]].push_back(e);
    } else {
      floc[in.efrac0[e]].push_back(e);
    }
  }
  // connected components of the fracture graph: independent subtrees
  std::vector<int> par(F);
  std::iota(par.begin(), par.end(), 0);
  auto root = [&](int a) {
    while (par[a] != a) a = par[a] = par[par[a]];
    return a;
  };
  for (int e = 0; e < E; ++e)
    if (in.etype[e] == 0) {
      const int a = root(in.efrac0[e]), b = root(in.efrac1[e]);
      if (a != b) par[std::max(a, b)] = std::min(a, b);
    }
  std::vector<std::vector<int>> comps;
  {
    std::vector<int> cid(F, -1);
    for (int f = 0; f < F; ++f) {
      const int r = root(f);
      if (cid[r] < 0) {
        cid[r] = (int)comps.size();
        comps.emplace_back();
      }
      comps[cid[r]].push_back(f);
    }
  }
  // nested dissection by recursive median bisection of the fracture centres
  std::vector<Node> nodes;
  std::vector<int> leaf_of(F, -1);
  std::vector<char> inset(F, 0);
  std::vector<int> stack;
  for (auto& c : comps) {
    Node nd;
    nd.frs = c;
    nodes.push_back(std::move(nd));
    stack.push_back((int)nodes.size() - 1);
  }
  while (!stack.empty()) {
    const int id = stack.back();
    stack.pop_back();
    std::vector<int> S = std::move(nodes[id].frs);
    int64_t cnt = (int64_t)S.size();
    for (int f : S) inset[f] = 1;
    for (int f : S) {
      cnt += (int64_t)floc[f].size();
      for (int e : fx[f]) {
        const int o = in.efrac0[e] == f ? in.efrac1[e] : in.efrac0[e];
        if (inset[o] && o > f) ++cnt;  // each interior intersection once
      }
    }
    for (int f : S) inset[f] = 0;
    if (cnt <= kMfLeafUnknowns || S.size() < 2) {
      for (int f : S) leaf_of[f] = id;
      nodes[id].frs = std::move(S);
      continue;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int f : S)
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], in.center[3 * f + a]);
        hi[a] = std::max(hi[a], in.center[3 * f + a]);
      }
    int ax = 0;
    for (int a = 1; a < 3; ++a)
      if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
    std::sort(S.begin(), S.end(), [&](int a, int b) {
      const double ca = in.center[3 * a + ax], cb = in.center[3 * b + ax];
      return ca != cb ? ca < cb : a < b;
    });
    const size_t h = S.size() / 2;
    for (int side = 0; side < 2; ++side) {
      Node ch;
      ch.parent = id;
      ch.depth = nodes[id].depth + 1;
      ch.frs.assign(side ? S.begin() + h : S.begin(), side ? S.end() : S.begin() + h);
      nodes.push_back(std::move(ch));
      stack.push_back((int)nodes.size() - 1);
    }
  }
  const int nn = (int)nodes.size();
  auto lca = [&](int a, int b) {
    while (a != b) {
      if (nodes[a].depth > nodes[b].depth) a = nodes[a].parent;
      else if (nodes[b].depth > nodes[a].depth) b = nodes[b].parent;
      else {
        a = nodes[a].parent;
        b = nodes[b].parent;
      }
    }
    return a;
  };
  // unknowns to nodes: an intersection to the separator of its two fractures (their leaves'
  // common ancestor), other discharge elements to their fracture's leaf, a fracture constant to
  // the shallowest node holding one of its fracture's discharge elements, so that its continuity
  // row is pivoted only once every discharge it sums is (a row missing some of them can cancel
  // exactly against another fracture's).  An unknown the numerical factorisation found unstable
  // is moved delay[u] levels further up, at most to its root (the solver passes a delay larger
  // than any depth: the root front, where every row is fully summed).
  auto lift = [&](int nd, int64_t g) {
    for (int d = in.delay ? in.delay[g] : 0; d > 0 && nodes[nd].parent >= 0; --d) nd = nodes[nd].parent;
    return nd;
  };
  std::vector<int> node_of_e(E, -1);
  int64_t nunk = 0;
  for (int e = 0; e < E; ++e) {
    if (!is_de(e)) continue;
    const int nd = in.etype[e] == 0 ? lca(leaf_of[in.efrac0[e]], leaf_of[in.efrac1[e]]) : leaf_of[in.efrac0[e]];
    node_of_e[e] = lift(nd, in.de_pos[e]);
    nodes[node_of_e[e]].U.push_back(in.de_pos[e]);
    ++nunk;
  }
  for (int f = 0; f < F; ++f) {
    int best = -1;
    for (const auto* lst : {&fx[f], &floc[f]})
      for (int e : *lst) {
        const int nd = node_of_e[e];
        if (best < 0 || nodes[nd].depth < nodes[best].depth) best = nd;
      }
    if (best < 0) best = leaf_of[f];
    best = lift(best, in.fpos[f]);
    nodes[best].U.push_back(in.fpos[f]);
    ++nunk;
  }
  // empty nodes are spliced out: their children hang from the nearest non-empty ancestor
  std::vector<int> keep_parent(nn, -1);
  for (int i = 0; i < nn; ++i) {
    int p = nodes[i].parent;
    while (p >= 0 && nodes[p].U.empty()) p = nodes[p].parent;
    keep_parent[i] = p;
  }
  std::vector<std::vector<int>> kids(nn);
  std::vector<int> roots;
  for (int i = 0; i < nn; ++i) {
    if (nodes[i].U.empty()) continue;
    if (keep_parent[i] < 0) roots.push_back(i);
    else kids[keep_parent[i]].push_back(i);
  }
  // postorder (children in creation order, roots in component order)
  std::vector<int> post;
  post.reserve(nn);
  {
    std::vector<std::pair<int, int>> stk;
    for (int r : roots) {
      stk.emplace_back(r, 0);
      while (!stk.empty()) {
        auto& top = stk.back();
        if (top.second < (int)kids[top.first].size()) {
          const int c = kids[top.first][top.second++];
          stk.emplace_back(c, 0);
        } else {
          post.push_back(top.first);
          stk.pop_back();
        }
      }
    }
  }
  // Delayed pivots, decided on the structure.  A front's pivots must be matched, row to column,
  // through A's own entries among them (the fill that the children's Schur complements add is not
  // counted on: its pattern depends on their pivot order and it can vanish, e.g. on a row whose
  // only coupling to a child is one entry).  A maximum matching names the rows and columns left
  // over, and their unknowns are passed up to the parent, which eliminates them after the fill
  // they receive here.
  std::vector<int> parent_of(nn, -1);
  for (int i = 0; i < nn; ++i) parent_of[i] = keep_parent[i];
  std::vector<char> removed(nn, 0), done(N, 0);
  std::vector<std::vector<int>> Bn(nn);
  std::vector<int> bloc(N, -1), stamp(N, -1);
  int stampc = 0;
  std::vector<int> aptr, acol;  // the block's pattern, row by row (block-local columns)
  std::vector<int> matchc, matchr, vis, stk_r, stk_e, la;
  for (int sidx = 0; sidx < (int)post.size(); ++sidx) {
    const int s = post[sidx];
    std::vector<int>& U = nodes[s].U;
    std::sort(U.begin(), U.end());
    for (;;) {
      const int k = (int)U.size();
      if (k == 0) break;
      for (int i = 0; i < k; ++i) bloc[U[i]] = i;
      aptr.assign(k + 1, 0);
      acol.clear();
      for (int i = 0; i < k; ++i) {
        ++stampc;
        for (int q = rp[U[i]]; q < rp[U[i] + 1]; ++q) {
          const int l = bloc[cl[q]];
          if (l >= 0 && stamp[cl[q]] != stampc) {
            stamp[cl[q]] = stampc;
            acol.push_back(l);
          }
        }
        aptr[i + 1] = (int)acol.size();
      }
      // maximum matching of rows to columns: depth-first augmenting paths with look-ahead (each
      // row first takes a free column of its own if it has one; its look-ahead pointer only moves
      // forward, since a matched column stays matched)
      matchc.assign(k, -1);
      matchr.assign(k, -1);
      la.assign(k, 0);
      vis.assign(k, -1);
      auto free_col = [&](int r) {
        for (int& e = la[r]; e < aptr[r + 1] - aptr[r]; ++e)
          if (matchc[acol[aptr[r] + e]] < 0) return acol[aptr[r] + e];
        return -1;
      };
      // augment along the stack: each row on it takes the column it went through, the top row col
      auto augment = [&](int col) {
        for (int d = (int)stk_r.size() - 1; d >= 0; --d) {
          const int rr = stk_r[d];
          const int prev = matchr[rr];
          matchr[rr] = col;
          matchc[col] = rr;
          col = prev;
        }
      };
      for (int i0 = 0; i0 < k; ++i0) {
        // search from row i0: stk_r holds rows, stk_e the next edge index of each
        stk_r.assign(1, i0);
        stk_e.assign(1, 0);
        vis[i0] = i0;
        int jf = free_col(i0);
        if (jf >= 0) {
          augment(jf);
          continue;
        }
        while (!stk_r.empty()) {
          const int r = stk_r.back();
          int& e = stk_e.back();
          if (e >= aptr[r + 1] - aptr[r]) {
            stk_r.pop_back();
            stk_e.pop_back();
            continue;
          }
          const int r2 = matchc[acol[aptr[r] + e++]];  // every column of a row on the stack is matched
          if (vis[r2] == i0) continue;
          vis[r2] = i0;
          stk_r.push_back(r2);
          stk_e.push_back(0);
          jf = free_col(r2);
          if (jf >= 0) {
            augment(jf);
            break;
          }
        }
      }
      std::vector<int> move;
      for (int i = 0; i < k; ++i)
        if (matchr[i] < 0 || matchc[i] < 0) move.push_back(U[i]);
      for (int i = 0; i < k; ++i) bloc[U[i]] = -1;
      const int p = parent_of[s];
      // a root's block is the last Schur complement: nonsingular when A and every earlier block
      // are, whatever the pattern of A's own entries in it
      if (move.empty() || p < 0) break;
      std::vector<int> keep;
      std::sort(move.begin(), move.end());
      for (int g : U)
        if (!std::binary_search(move.begin(), move.end(), g)) keep.push_back(g);
      U.swap(keep);
      nodes[p].U.insert(nodes[p].U.end(), move.begin(), move.end());
    }
    if (U.empty()) {
      // nothing left to eliminate here: the children hang from the parent instead
      removed[s] = 1;
      const int p = parent_of[s];
      if (p >= 0) {
        std::vector<int> nk;
        for (int c : kids[p]) {
          if (c == s) nk.insert(nk.end(), kids[s].begin(), kids[s].end());
          else nk.push_back(c);
        }
        kids[p].swap(nk);
      }
      for (int c : kids[s]) parent_of[c] = p;  // at a root: its children become roots
      continue;
    }
    for (int u : U) done[u] = 1;
    std::vector<int>& bl = Bn[s];
    ++stampc;
    auto add = [&](int v) {
      if (!done[v] && stamp[v] != stampc) {
        stamp[v] = stampc;
        bl.push_back(v);
      }
    };
    for (int u : U)
      for (int q = rp[u]; q < rp[u + 1]; ++q) add(cl[q]);
    for (int c : kids[s])
      for (int v : Bn[c]) add(v);
  }
  {
    std::vector<int> p2;
    for (int s : post)
      if (!removed[s]) p2.push_back(s);
    post.swap(p2);
  }
  for (int s : post)
    if (parent_of[s] >= 0 && removed[parent_of[s]]) throw std::runtime_error("multifrontal analysis: broken tree");
  const int nf = (int)post.size();
  std::vector<int> front_of_node(nn, -1);
  for (int i = 0; i < nf; ++i) front_of_node[post[i]] = i;
  // elimination positions
  std::vector<int> elim(N, -1);
  int ectr = 0;
  for (int i = 0; i < nf; ++i) {
    auto& U = nodes[post[i]].U;
    std::sort(U.begin(), U.end());
    for (int g : U) elim[g] = ectr++;
  }
  std::vector<std::vector<int>> B(nf);
  for (int i = 0; i < nf; ++i) {
    B[i] = std::move(Bn[post[i]]);
    std::sort(B[i].begin(), B[i].end(), [&](int a, int b) { return elim[a] < elim[b]; });
  }
  // fronts, offsets and the index lists
  std::vector<MfFront>& fr = out.fronts;
  fr.assign(nf, MfFront{});
  std::vector<int>& idx = out.idx;
  int64_t off = 0, inv = 0;
  int piv = 0, vec = 0, rel = 0, gat = 0;
  int64_t nnz_factors = 0;
  double flops = 0.0;
  int max_m = 0;
  std::vector<int> height(nf, 0);
  for (int i = 0; i < nf; ++i) {
    const Node& nd = nodes[post[i]];
    MfFront& f = fr[i];
    f.k = (int)nd.U.size();
    f.m = f.k + (int)B[i].size();
    f.off = off;
    f.inv = inv;
    f.idx = (int)idx.size();
    f.vec = vec;
    f.piv = piv;
    f.rel = rel;
    f.gat = gat;
    f.parent = parent_of[post[i]] >= 0 ? front_of_node[parent_of[post[i]]] : -1;
    idx.insert(idx.end(), nd.U.begin(), nd.U.end());
    idx.insert(idx.end(), B[i].begin(), B[i].end());
    off += (int64_t)f.m * f.m;
    inv += 2 * (int64_t)f.k * f.k;
    vec += f.m;
    piv += f.k;
    rel += f.m - f.k;
    gat += f.m + 1;
    const double k = f.k, b = f.m - f.k;
    nnz_factors += (int64_t)(k * k + k + 2 * k * b);
    flops += 2.0 / 3.0 * k * k * k + 2.0 * k * k * b + 2.0 * k * b * b;
    max_m = std::max(max_m, f.m);
    for (int c : kids[post[i]]) height[i] = std::max(height[i], height[front_of_node[c]] + 1);
  }
  out.front_total = off;
  out.inv_total = inv;
  out.nvec = vec;
  out.nunk = nunk;
  out.nnz_factors = nnz_factors;
  out.flops = flops;
  out.max_m = max_m;
  // assembly map, extend-add positions and the solve's gather lists
  std::vector<int64_t>& asmap = out.asmap;
  asmap.assign(nnz, -1);
  std::vector<int>& relv = out.rel;
  relv.assign(rel, -1);
  std::vector<int>& gptr = out.gptr;
  gptr.assign(gat, 0);  // m + 1 pointers per front
  std::vector<std::vector<int>> gl;  // per position of the current parent: its sources
  std::vector<int>& gsrc = out.gsrc;
  std::vector<int> loc(N, -1);
  std::vector<std::vector<int>> fkids(nf);
  for (int i = 0; i < nf; ++i)
    for (int c : kids[post[i]]) fkids[i].push_back(front_of_node[c]);
  // the assembly map, front by front on a few host threads (every matrix entry belongs to one
  // front, so the fronts write disjoint entries); the rest in front order below
  {
    const unsigned hc = std::thread::hardware_concurrency();
    const int nth = (int)std::max(1u, std::min(8u, hc ? hc : 1u));
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    std::string why;
    auto work = [&](std::vector<int>& lc) {
      for (;;) {
        const int i0 = next.fetch_add(16);
        if (i0 >= nf || failed.load()) return;
        for (int i = i0; i < std::min(nf, i0 + 16); ++i) {
          const MfFront& f = fr[i];
          const int* ix = idx.data() + f.idx;
          for (int l = 0; l < f.m; ++l) lc[ix[l]] = l;
          const int first = elim[ix[0]];
          const char* err = nullptr;
          for (int lu = 0; lu < f.k && !err; ++lu) {
            const int u = ix[lu];
            for (int q = rp[u]; q < rp[u + 1]; ++q) {
              const int v = cl[q];
              if (elim[v] < first) continue;  // v eliminated in a descendant, which placed this entry
              const int lv = lc[v];
              if (lv < 0) {
                err = "multifrontal analysis: entry outside its front";
                break;
              }
              asmap[q] = f.off + lu + (int64_t)lv * f.m;
              if (lv >= f.k) {  // (row v, column u) with v in the boundary
                const int* b0 = cl + rp[v];
                const int* b1 = cl + rp[v + 1];
                const int* hit = std::lower_bound(b0, b1, u);
                if (hit == b1 || *hit != u) {
                  err = "multifrontal analysis: pattern is not symmetric";
                  break;
                }
                asmap[hit - cl] = f.off + lv + (int64_t)lu * f.m;
              }
            }
          }
          for (int l = 0; l < f.m; ++l) lc[ix[l]] = -1;
          if (err) {
            if (!failed.exchange(true)) why = err;
            return;
          }
        }
      }
    };
    std::vector<std::vector<int>> locs(nth - 1, std::vector<int>(N, -1));
    std::vector<std::thread> pool;
    for (int t = 0; t + 1 < nth; ++t) pool.emplace_back(work, std::ref(locs[t]));
    work(loc);
    for (auto& th : pool) th.join();
    if (failed.load()) throw std::runtime_error(why);
  }
  for (int i = 0; i < nf; ++i) {
    const MfFront& f = fr[i];
    const int* ix = idx.data() + f.idx;
    for (int l = 0; l < f.m; ++l) loc[ix[l]] = l;
    // children: their boundaries inside this front
    gl.assign(f.m, {});
    for (int c : fkids[i]) {
      const MfFront& fc = fr[c];
      const int* ixc = idx.data() + fc.idx;
      for (int t = 0; t < fc.m - fc.k; ++t) {
        const int l = loc[ixc[fc.k + t]];
        if (l < 0) throw std::runtime_error("multifrontal analysis: a child's boundary is outside its parent");
        relv[fc.rel + t] = l;
        gl[l].push_back(fc.vec + fc.k + t);
      }
    }
    for (int l = 0; l < f.m; ++l) {
      gptr[f.gat + l + 1] = gptr[f.gat + l] + (int)gl[l].size();
      gsrc.insert(gsrc.end(), gl[l].begin(), gl[l].end());
    }
    if (i + 1 < nf) gptr[fr[i + 1].gat] = gptr[f.gat + f.m];
    for (int l = 0; l < f.m; ++l) loc[ix[l]] = -1;
  }
  for (int64_t q = 0; q < nnz; ++q)
    if (asmap[q] < 0) throw std::runtime_error("multifrontal analysis: a matrix entry has no front");
  // levels by height; per level the small-front lists, Schur tiles, solve panels, extend items
  const int H = nf ? *std::max_element(height.begin(), height.end()) + 1 : 0;
  std::vector<MfLevel>& lev = out.levels;
  lev.assign(H, MfLevel{});
  std::vector<std::vector<int>> byh(H);
  for (int i = 0; i < nf; ++i) byh[height[i]].push_back(i);
  std::vector<int>& small_list = out.small_list;
  std::vector<MfI4>& tiles = out.tiles;
  std::vector<MfI2>& ext = out.ext;
  std::vector<char> small(nf, 0);
  for (int h = 0; h < H; ++h) {
    MfLevel& L = lev[h];
    int n128 = 0;  // fronts of the 65-128 class on this level
    for (int i : byh[h]) n128 += fr[i].k > 64 && mf_is_small(fr[i].k, fr[i].m - fr[i].k);
    for (int i : byh[h]) small[i] = fr[i].k <= 64 || (n128 >= kMfSmall128Min && mf_is_small(fr[i].k, fr[i].m - fr[i].k));
    for (int cls = 0; cls < 4; ++cls) {
      const int kmax = 16 << cls, kmin = cls ? (8 << cls) : 0;
      L.small_beg[cls] = (int)small_list.size();
      for (int i : byh[h])
        if (fr[i].k > kmin && fr[i].k <= kmax && small[i]) small_list.push_back(i);
      L.small_end[cls] = (int)small_list.size();
    }
    L.tile_beg = (int)tiles.size();
    for (int i : byh[h]) {
      const MfFront& f = fr[i];
      const int b = f.m - f.k;
      if (!small[i]) {
        L.big.push_back(i);
        continue;
      }
      for (int i0 = 0; i0 < b; i0 += 64)
        for (int j0 = 0; j0 < b; j0 += 64) tiles.push_back(MfI4{i, i0, j0, 0});
    }
    L.tile_end = (int)tiles.size();
    for (int kind = 0; kind < 4; ++kind) {
      // panel rows: the tallest of 16, 8, 4 that still gives the launch kMfWidePanels panels (kind
      // 0 keeps 16: each of its panels gathers its own inputs, so more panels cost more gathers)
      auto panels = [&](int R) {
        int64_t n = 0;
        for (int i : byh[h]) {
          int rows, cols;
          mf_panel_shape(kind, fr[i].k, fr[i].m - fr[i].k, &rows, &cols);
          if (cols > 0) n += (rows + R - 1) / R;
        }
        return n;
      };
      int R = 16;
      if (kind != 0 && in.panel_rows) {
        R = in.panel_rows;
      } else if (kind != 0) {
        while (R > 4 && panels(R) < kMfWidePanels) R /= 2;
        if (panels(R) < kMfWidePanels / 4) R = 16;  // very few panels even at 4 rows: 16 rows, wide
      }
      L.prow[kind] = R;
      L.gv_beg[kind] = (int)out.gv.size();
      int64_t blocks = 0;
      for (int i : byh[h]) {
        const MfFront& f = fr[i];
        const int b = f.m - f.k;
        int rows, cols;
        mf_panel_shape(kind, f.k, b, &rows, &cols);
        if (cols == 0) continue;
        for (int r0 = 0; r0 < rows; r0 += R) {
          int c_lo, nblk;
          mf_panel_columns(kind, f.k, b, R, r0, &c_lo, &nblk);
          out.gv.push_back(MfI4{i, r0, c_lo, nblk});
          out.gv_off.push_back(out.gv_total);
          out.gv_total += (int64_t)kMfBlockValues * nblk;
          blocks += nblk;
        }
      }
      L.gv_end[kind] = (int)out.gv.size();
      // few 16-row panels of many blocks each: 1,024 threads per panel
      const int n = L.gv_end[kind] - L.gv_beg[kind];
      L.pwide[kind] = R == 16 && n > 0 && n < kMfWidePanels && blocks >= 8 * (int64_t)n ? 1 : 0;
    }
    int slots = 0;
    for (int i : byh[h]) slots = std::max(slots, (int)fkids[i].size());
    for (int s = 0; s < slots; ++s) {
      L.ext_beg.push_back((int)ext.size());
      for (int i : byh[h])
        if (s < (int)fkids[i].size()) {
          const int c = fkids[i][s];
          for (int t0 = 0; t0 < fr[c].m - fr[c].k; t0 += 4) ext.push_back(MfI2{c, t0});
        }
      L.ext_end.push_back((int)ext.size());
    }
  }
  std::vector<int>& unk = out.unk;
  for (int i = 0; i < nf; ++i)
    for (int l = 0; l < fr[i].k; ++l) unk.push_back(idx[fr[i].idx + l]);
  out.piv0.assign(piv, 0);
  for (int i = 0; i < nf; ++i)
    for (int l = 0; l < fr[i].k; ++l) out.piv0[fr[i].piv + l] = l;
  out.big_k = 0;
  for (int i = 0; i < nf; ++i)
    if (!small[i]) out.big_k = std::max(out.big_k, fr[i].k);
}

}  // namespace aem
