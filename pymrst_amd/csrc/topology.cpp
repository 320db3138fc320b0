// Host-side construction of the static block pattern of a TPFA operator and
// of the aggregation hierarchy used by the pressure AMG.  Runs once per
// topology; everything that changes per Newton iteration is on the device.
//
// The TPFA pattern is the one the reference's Jacobian has on its cell-cell
// blocks: the diagonal plus (N0,N1) and (N1,N0) for every internal
// connection (pymrst/modules/ad_core/utils/setup_operators_tpfa.py:158-209;
// Grad = -C x, Div = C' x).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "pmx_internal.h"

namespace pmx {

void tpfa_pattern(int n, int nconn, const int* N, HostPattern& pat, std::vector<int>& entry_conn) {
  PMX_CHECK(n >= 0 && nconn >= 0, "negative size");
  std::vector<int> deg(n, 1);
  for (int f = 0; f < nconn; ++f) {
    int a = N[2 * f], b = N[2 * f + 1];
    PMX_CHECK(a >= 0 && a < n && b >= 0 && b < n && a != b, "connection list holds an invalid cell pair");
    deg[a]++;
    deg[b]++;
  }
  pat.n = n;
  pat.rowptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) pat.rowptr[i + 1] = pat.rowptr[i] + deg[i];
  int nnz = pat.rowptr[n];
  std::vector<std::pair<int, int>> ent(nnz);  // (col, conn code)
  std::vector<int> fill(pat.rowptr.begin(), pat.rowptr.end() - 1);
  for (int i = 0; i < n; ++i) ent[fill[i]++] = {i, -1};
  for (int f = 0; f < nconn; ++f) {
    int a = N[2 * f], b = N[2 * f + 1];
    ent[fill[a]++] = {b, 2 * f};
    ent[fill[b]++] = {a, 2 * f + 1};
  }
  pat.col.resize(nnz);
  entry_conn.resize(nnz);
  pat.diag.resize(n);
  for (int i = 0; i < n; ++i) {
    // stable insertion sort by column (rows are short; no temporary buffer)
    for (int k = pat.rowptr[i] + 1; k < pat.rowptr[i + 1]; ++k) {
      const std::pair<int, int> e = ent[k];
      int q = k;
      for (; q > pat.rowptr[i] && ent[q - 1].first > e.first; --q) ent[q] = ent[q - 1];
      ent[q] = e;
    }
    for (int k = pat.rowptr[i]; k < pat.rowptr[i + 1]; ++k) {
      // one connection per block entry: the assembly reads one T and one
      // gravity drop per entry.  Two connections between the same cells
      // carry the same gravity drop and upwind direction, so their sum is
      // one connection with the summed transmissibility.
      PMX_CHECK(k == pat.rowptr[i] || ent[k].first != ent[k - 1].first,
                "connection list holds a cell pair twice (merge them into one connection with the summed T)");
      pat.col[k] = ent[k].first;
      entry_conn[k] = ent[k].second;
      if (ent[k].first == i) pat.diag[i] = k;
    }
  }
}

// Greedy strength-based aggregation (the classic three-phase scheme of
// smoothed-aggregation AMG: seed aggregates from nodes whose strong
// neighbourhood is untouched, attach the rest to their strongest aggregated
// neighbour, then sweep up leftovers).
static int aggregate_greedy(const HostPattern& P, const std::vector<double>& w, double theta,
                            std::vector<int>& agg) {
  const int n = P.n;
  std::vector<double> wmax(n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k)
      if (P.col[k] != i) wmax[i] = std::max(wmax[i], w[k]);
  auto strong = [&](int i, int k) {
    int j = P.col[k];
    if (j == i) return false;
    double t = theta * std::sqrt(wmax[i] * wmax[j]);
    return w[k] > 0.0 && w[k] >= t;
  };
  agg.assign(n, -1);
  int na = 0;
  // phase 1
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nbhd = true;
    int nstrong = 0;
    for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k)
      if (strong(i, k)) {
        ++nstrong;
        if (agg[P.col[k]] >= 0) { free_nbhd = false; break; }
      }
    if (!free_nbhd || nstrong == 0) continue;
    agg[i] = na;
    for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k)
      if (strong(i, k)) agg[P.col[k]] = na;
    ++na;
  }
  // phase 2: attach to the strongest aggregated strong neighbour
  std::vector<int> snap = agg;
  for (int i = 0; i < n; ++i) {
    if (snap[i] >= 0) continue;
    double best = -1.0;
    int ba = -1;
    for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
      int j = P.col[k];
      if (strong(i, k) && snap[j] >= 0 && w[k] > best) { best = w[k]; ba = snap[j]; }
    }
    if (ba >= 0) agg[i] = ba;
  }
  // phase 3: leftovers form aggregates with their unaggregated strong neighbours;
  // a node with no strong neighbour joins its strongest neighbour's aggregate or
  // stays alone.
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    int nstrong = 0;
    for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k)
      if (strong(i, k)) ++nstrong;
    if (nstrong == 0) {
      double best = -1.0;
      int ba = -1;
      for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
        int j = P.col[k];
        if (j != i && agg[j] >= 0 && w[k] > best) { best = w[k]; ba = agg[j]; }
      }
      agg[i] = ba >= 0 ? ba : na++;
      continue;
    }
    agg[i] = na;
    for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k)
      if (strong(i, k) && agg[P.col[k]] < 0) agg[P.col[k]] = na;
    ++na;
  }
  return na;
}

// Pairwise matching by strongest mutual connection (one pass halves the
// number of unknowns at most); repeated passes give aggregates of 4, 8, ...
static int aggregate_pairwise(const HostPattern& P, const std::vector<double>& w, std::vector<int>& agg) {
  const int n = P.n;
  agg.assign(n, -1);
  int na = 0;
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    double best = 0.0;
    int bj = -1;
    for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
      int j = P.col[k];
      if (j != i && agg[j] < 0 && w[k] > best) { best = w[k]; bj = j; }
    }
    agg[i] = na;
    if (bj >= 0) agg[bj] = na;
    ++na;
  }
  return na;
}

// Coarse pattern, Galerkin gather map (for piecewise-constant prolongation)
// and coarse strength weights.  Flat arrays throughout: each coarse row's
// columns are gathered once (a mark per column), sorted in place, and every
// fine entry finds its coarse entry through the row's slot table.
static void coarsen(const HostPattern& P, const std::vector<double>& w, const std::vector<int>& agg, int nc,
                    HostLevel& L, HostPattern& CP, std::vector<double>& cw) {
  const int n = P.n;
  const int nnz = P.rowptr[n];
  L.agg = agg;
  L.ncoarse = nc;
  L.agg_ptr.assign(nc + 1, 0);
  for (int i = 0; i < n; ++i) L.agg_ptr[agg[i] + 1]++;
  for (int I = 0; I < nc; ++I) L.agg_ptr[I + 1] += L.agg_ptr[I];
  L.agg_mem.resize(n);
  {
    std::vector<int> f(L.agg_ptr.begin(), L.agg_ptr.end() - 1);
    for (int i = 0; i < n; ++i) L.agg_mem[f[agg[i]]++] = i;
  }
  CP.n = nc;
  CP.rowptr.assign(nc + 1, 0);
  CP.diag.resize(nc);
  CP.col.clear();
  CP.col.reserve((size_t)nc + (size_t)nnz / 2);
  std::vector<int> mark(nc, -1), slot(nc, 0), dest(nnz);
  for (int I = 0; I < nc; ++I) {
    const size_t r0 = CP.col.size();
    CP.col.push_back(I);
    mark[I] = I;
    for (int m = L.agg_ptr[I]; m < L.agg_ptr[I + 1]; ++m) {
      const int i = L.agg_mem[m];
      for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
        const int J = agg[P.col[k]];
        if (mark[J] != I) {
          mark[J] = I;
          CP.col.push_back(J);
        }
      }
    }
    std::sort(CP.col.begin() + r0, CP.col.end());
    for (size_t q = r0; q < CP.col.size(); ++q) slot[CP.col[q]] = (int)q;
    CP.diag[I] = slot[I];
    for (int m = L.agg_ptr[I]; m < L.agg_ptr[I + 1]; ++m) {
      const int i = L.agg_mem[m];
      for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) dest[k] = slot[agg[P.col[k]]];
    }
    CP.rowptr[I + 1] = (int)CP.col.size();
  }
  const int cnnz = CP.rowptr[nc];
  // gather map: coarse entry K <- the fine entries k with dest[k] == K, in
  // coarse-row, member, entry order
  L.gal_ptr.assign(cnnz + 1, 0);
  for (int k = 0; k < nnz; ++k) L.gal_ptr[dest[k] + 1]++;
  for (int K = 0; K < cnnz; ++K) L.gal_ptr[K + 1] += L.gal_ptr[K];
  L.gal_src.resize(nnz);
  cw.assign(cnnz, 0.0);
  {
    std::vector<int> f(L.gal_ptr.begin(), L.gal_ptr.end() - 1);
    for (int I = 0; I < nc; ++I)
      for (int m = L.agg_ptr[I]; m < L.agg_ptr[I + 1]; ++m) {
        const int i = L.agg_mem[m];
        for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
          const int K = dest[k];
          L.gal_src[f[K]++] = k;
          if (CP.col[K] != I && P.col[k] != i) cw[K] += w[k];
        }
      }
  }
}

void build_aggregation(const HostPattern& fine, const std::vector<double>& w, const AggregationParams& prm,
                       std::vector<HostLevel>& levels) {
  levels.clear();
  levels.emplace_back();
  levels[0].pat = fine;
  levels[0].w = w;
  while (true) {
    HostLevel& L = levels.back();
    const int n = L.pat.n;
    if (n <= prm.coarse_size || (int)levels.size() >= prm.max_levels) break;
    std::vector<int> agg;
    int nc = 0;
    if (prm.pairwise_passes > 0) {
      // repeated pairwise matching on successively coarser graphs
      HostPattern P = L.pat;
      std::vector<double> ww = L.w;
      std::vector<int> total(n);
      std::iota(total.begin(), total.end(), 0);
      for (int pass = 0; pass < prm.pairwise_passes; ++pass) {
        std::vector<int> a;
        int m = aggregate_pairwise(P, ww, a);
        for (int i = 0; i < n; ++i) total[i] = a[total[i]];
        if (pass + 1 < prm.pairwise_passes) {
          HostLevel tmp;
          HostPattern CP;
          std::vector<double> cw;
          coarsen(P, ww, a, m, tmp, CP, cw);
          P = std::move(CP);
          ww = std::move(cw);
        }
        nc = m;
      }
      agg = std::move(total);
    } else {
      nc = aggregate_greedy(L.pat, L.w, prm.theta, agg);
    }
    if (nc <= 0 || (long long)nc * 100 > (long long)prm.min_coarse_reduction_pct * n) break;
    HostPattern CP;
    std::vector<double> cw;
    coarsen(L.pat, L.w, agg, nc, L, CP, cw);
    levels.emplace_back();
    levels.back().pat = std::move(CP);
    levels.back().w = std::move(cw);
  }
  // the last level carries no transfer
  levels.back().agg.clear();
  levels.back().ncoarse = 0;
}

}  // namespace pmx

namespace pmx {

// Position-major tables for a set of lines that covers every cell once.
// Lines are coloured greedily (longest first) so that lines sharing a matrix
// entry get different colours; a Gauss-Seidel sweep over the colours then
// solves all lines of one colour at once.
static void line_tables(const HostPattern& P, std::vector<std::vector<int>>& lines, HostLines& L) {
  const int n = P.n;
  std::stable_sort(lines.begin(), lines.end(),
                   [](const std::vector<int>& a, const std::vector<int>& b) { return a.size() > b.size(); });
  int nl = (int)lines.size();
  std::vector<int> line_of(n, -1);
  for (int l = 0; l < nl; ++l)
    for (int c : lines[l]) {
      PMX_CHECK(line_of[c] < 0, "a cell lies on two lines");
      line_of[c] = l;
    }
  for (int c = 0; c < n; ++c) PMX_CHECK(line_of[c] >= 0, "lines must cover every cell");
  std::vector<int> color(nl, -1), mark;
  int ncolor = 0;
  for (int l = 0; l < nl; ++l) {
    mark.assign(ncolor + 1, 0);
    for (int c : lines[l])
      for (int k = P.rowptr[c]; k < P.rowptr[c + 1]; ++k) {
        const int m = line_of[P.col[k]];
        if (m != l && color[m] >= 0) mark[color[m]] = 1;
      }
    int pick = 0;
    while (mark[pick]) ++pick;
    color[l] = pick;
    ncolor = std::max(ncolor, pick + 1);
  }
  PMX_CHECK(ncolor <= 255, "too many line colours");
  // colour-major order; within a colour the length order above is kept
  std::vector<int> order(nl);
  for (int l = 0; l < nl; ++l) order[l] = l;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return color[a] < color[b]; });
  std::vector<std::vector<int>> sorted;
  L.color_ptr.assign(1, 0);
  L.cell_color.assign(n, 0);
  for (int t = 0; t < nl;) {
    const int c = color[order[t]];
    for (; t < nl && color[order[t]] == c; ++t) {
      for (int cell : lines[order[t]]) L.cell_color[cell] = (unsigned char)c;
      sorted.push_back(std::move(lines[order[t]]));
    }
    while (sorted.size() % LINE_GROUP) sorted.emplace_back();   // empty lines fill the last wave
    L.color_ptr.push_back((int)sorted.size());
  }
  nl = (int)sorted.size();
  int longest = 0;
  for (auto& ln : sorted) longest = std::max(longest, (int)ln.size());
  const int maxlen = (longest + LINE_PAD - 1) / LINE_PAD * LINE_PAD;
  L.nlines = nl;
  L.ncolor = ncolor;
  L.longest = longest;
  L.maxlen = maxlen;
  L.len.resize(nl);
  L.cell.assign((size_t)maxlen * nl, -1);
  L.lo.assign((size_t)maxlen * nl, -1);
  L.up.assign((size_t)maxlen * nl, -1);
  auto entry = [&](int r, int c) {
    auto b0 = P.col.begin() + P.rowptr[r], b1 = P.col.begin() + P.rowptr[r + 1];
    auto it = std::lower_bound(b0, b1, c);
    PMX_CHECK(it != b1 && *it == c, "consecutive line cells are not connected");
    return (int)(it - P.col.begin());
  };
  for (int l = 0; l < nl; ++l) {
    const auto& ln = sorted[l];
    L.len[l] = (int)ln.size();
    for (size_t k = 0; k < ln.size(); ++k) {
      const size_t at = k * nl + l;
      L.cell[at] = ln[k];
      if (k > 0) L.lo[at] = entry(ln[k], ln[k - 1]);
      if (k + 1 < ln.size()) L.up[at] = entry(ln[k], ln[k + 1]);
    }
  }
  // couplings off the line, per position
  const size_t npos = (size_t)maxlen * nl;
  L.pos.assign(n, -1);
  std::vector<int> lof(n, -1);
  for (int l = 0; l < nl; ++l)
    for (size_t k = 0; k < sorted[l].size(); ++k) {
      L.pos[sorted[l][k]] = (int)(k * nl + l);
      lof[sorted[l][k]] = l;
    }
  int noff = 0;
  for (int c = 0; c < n; ++c) {
    int m = 0;
    for (int q = P.rowptr[c]; q < P.rowptr[c + 1]; ++q) m += lof[P.col[q]] != lof[c];
    noff = std::max(noff, m);
  }
  L.noff = noff;
  L.off_q.assign((size_t)noff * npos, -1);
  L.off_nbr.assign((size_t)noff * npos, -1);
  L.off_col.assign((size_t)noff * npos, 255);
  for (int c = 0; c < n; ++c) {
    const size_t at = (size_t)L.pos[c];
    int h = 0;
    for (int q = P.rowptr[c]; q < P.rowptr[c + 1]; ++q) {
      const int j = P.col[q];
      if (lof[j] == lof[c]) continue;
      L.off_q[h * npos + at] = q;
      L.off_nbr[h * npos + at] = L.pos[j];
      L.off_col[h * npos + at] = L.cell_color[j];
      ++h;
    }
  }
}

// Lines of strongly coupled cells: each cell links to its (at most) two
// strongest neighbours; a link is kept when it is mutual, which leaves
// paths and cycles; cycles are cut once.  On layered Cartesian grids with
// thin cells these are the vertical columns.
void build_lines(const HostPattern& P, const std::vector<double>& w, HostLines& L) {
  const int n = P.n;
  std::vector<int> s1(n, -1), s2(n, -1);
  for (int i = 0; i < n; ++i) {
    double b1 = 0.0, b2 = 0.0;
    for (int k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
      const int j = P.col[k];
      if (j == i || !(w[k] > 0.0)) continue;
      if (w[k] > b1) {
        b2 = b1; s2[i] = s1[i];
        b1 = w[k]; s1[i] = j;
      } else if (w[k] > b2) {
        b2 = w[k]; s2[i] = j;
      }
    }
  }
  auto top2 = [&](int i, int j) { return s1[i] == j || s2[i] == j; };
  std::vector<int> nb0(n, -1), nb1(n, -1);
  auto link = [&](int i, int j) {
    if (nb0[i] < 0) nb0[i] = j; else nb1[i] = j;
  };
  for (int i = 0; i < n; ++i)
    for (int j : {s1[i], s2[i]})
      if (j > i && top2(j, i)) { link(i, j); link(j, i); }
  std::vector<char> seen(n, 0);
  std::vector<std::vector<int>> lines;
  auto walk = [&](int start) {
    std::vector<int> line;
    int prev = -1, cur = start;
    while (cur >= 0 && !seen[cur]) {
      seen[cur] = 1;
      line.push_back(cur);
      const int nxt = nb0[cur] != prev ? nb0[cur] : nb1[cur];
      prev = cur;
      cur = nxt;
    }
    return line;
  };
  for (int i = 0; i < n; ++i)   // path ends first
    if (!seen[i] && (nb0[i] < 0 || nb1[i] < 0)) lines.push_back(walk(i));
  for (int i = 0; i < n; ++i)   // what is left lies on cycles: cut at i
    if (!seen[i]) lines.push_back(walk(i));
  std::stable_sort(lines.begin(), lines.end(),
                   [](const std::vector<int>& a, const std::vector<int>& b) { return a[0] < b[0]; });
  line_tables(P, lines, L);
}

// Lines given by the caller (line_ptr[nlines+1], cells[]), e.g. the columns
// of a Cartesian grid; consecutive cells of a line must be connected.  Cells
// on no line become lines of one cell.
void lines_from_list(const HostPattern& P, int nlines, const int* ptr, const int* cells, HostLines& L) {
  std::vector<std::vector<int>> lines(nlines);
  std::vector<char> seen(P.n, 0);
  for (int l = 0; l < nlines; ++l) {
    PMX_CHECK(ptr[l + 1] > ptr[l], "empty line");
    for (int k = ptr[l]; k < ptr[l + 1]; ++k) {
      const int c = cells[k];
      PMX_CHECK(c >= 0 && c < P.n && !seen[c], "line cells must be distinct cells of the grid");
      seen[c] = 1;
      lines[l].push_back(c);
    }
  }
  for (int i = 0; i < P.n; ++i)
    if (!seen[i]) lines.push_back(std::vector<int>{i});
  line_tables(P, lines, L);
}

}  // namespace pmx
