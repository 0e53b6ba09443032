// Host symbolic pass: run once per distinct circuit topology and analysis
// kind. It replaces the per-solve symbolic work that krets asks of faer
// (SparseColMat::try_new_from_triplets + sp_lu at solver/op.rs:48-51,
// dc.rs:81-84, ac.rs:67-70, transient.rs:74-77) with a fixed plan:
//
//  1. the stamp pattern of krets' element stamps (stampable.rs:202-1050) for
//     the analysis kind, with every entry tagged by the element contribution
//     that lands there;
//  2. a structural row matching so every pivot is structurally non-zero
//     (MNA branch rows of V sources, DC inductors and current sources have a
//     zero diagonal);
//  3. a symmetric fill-reducing order of the matched matrix;
//  4. the symbolic LU (fill) of the permuted matrix, the row-oriented factor
//     program, and dependency levels for the factor, both triangular solves
//     and both transposed solves;
//  5. per-slot and per-row gather lists for stamping on the device.
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <cstring>
#include <functional>
#include <queue>
#include <set>
#include <tuple>
#include <string>
#include <vector>

#include "kg_internal.h"

namespace kg {

bool validate_plan(const PatLayout& L, const std::vector<int32_t>& b, std::string& err);

namespace {

struct Entry {
  int r, c;
  int32_t code;
};

// Minimum degree (ties to the lowest index).
std::vector<int> order_min_degree(int n, const std::vector<std::vector<int>>& adj0) {
  std::vector<std::set<int>> adj(n);
  for (int v = 0; v < n; v++) adj[v].insert(adj0[v].begin(), adj0[v].end());
  std::set<std::pair<int, int>> pq;
  for (int v = 0; v < n; v++) pq.insert({(int)adj[v].size(), v});
  std::vector<int> order;
  order.reserve(n);
  while (!pq.empty()) {
    int v = pq.begin()->second;
    pq.erase(pq.begin());
    order.push_back(v);
    std::vector<int> nb(adj[v].begin(), adj[v].end());
    for (int u : nb) {
      pq.erase({(int)adj[u].size(), u});
      adj[u].erase(v);
      for (int w : nb)
        if (w != u) adj[u].insert(w);
      pq.insert({(int)adj[u].size(), u});
    }
    adj[v].clear();
  }
  return order;
}

// Minimum fill (deficiency: neighbour pairs not yet adjacent); ties go to the
// vertex lowest in the elimination tree built so far (depth = 1 + deepest
// eliminated neighbour), then to minimum degree. The depth rule keeps the
// tree balanced: a chain is eliminated from both ends at once (two
// independent dependency chains instead of one, with no fill).
std::vector<int> order_balanced(int n, const std::vector<std::vector<int>>& adj0) {
  std::vector<std::set<int>> adj(n);
  for (int v = 0; v < n; v++) adj[v].insert(adj0[v].begin(), adj0[v].end());
  auto deficiency = [&](int v) {
    int d = 0;
    for (auto a = adj[v].begin(); a != adj[v].end(); ++a)
      for (auto b = std::next(a); b != adj[v].end(); ++b)
        if (!adj[*a].count(*b)) d++;
    return d;
  };
  std::vector<int> depth(n, 0), def(n, 0);
  std::set<std::tuple<int, int, int, int>> pq;  // (fill, depth, degree, vertex)
  auto key = [&](int v) { return std::make_tuple(def[v], depth[v], (int)adj[v].size(), v); };
  for (int v = 0; v < n; v++) {
    def[v] = deficiency(v);
    pq.insert(key(v));
  }
  std::vector<char> done(n, 0);
  std::vector<int> order;
  order.reserve(n);
  while (!pq.empty()) {
    int v = std::get<3>(*pq.begin());
    pq.erase(pq.begin());
    order.push_back(v);
    done[v] = 1;
    std::vector<int> nb(adj[v].begin(), adj[v].end());
    std::set<int> touched(nb.begin(), nb.end());
    for (int u : nb) {
      pq.erase(key(u));
      adj[u].erase(v);
      for (int w : nb)
        if (w != u) adj[u].insert(w);
      depth[u] = std::max(depth[u], depth[v] + 1);
    }
    for (int u : nb) touched.insert(adj[u].begin(), adj[u].end());
    for (int u : touched) {
      if (done[u]) continue;
      pq.erase(key(u));
      def[u] = deficiency(u);
      pq.insert(key(u));
    }
    adj[v].clear();
  }
  return order;
}

// Nested dissection by recursive BFS level-structure bisection: the middle
// level of a level structure rooted at a pseudo-peripheral vertex is the
// separator. On the chain-like graphs of ladder circuits this yields an
// elimination tree of logarithmic depth.
std::vector<int> order_nested(int n, const std::vector<std::vector<int>>& adj) {
  std::vector<int> order;
  order.reserve(n);
  std::vector<int> part(n, 0), dist(n, -1);
  int next_id = 1;
  std::function<void(std::vector<int>&)> rec = [&](std::vector<int>& verts) {
    if (verts.empty()) return;
    if (verts.size() <= 2) {
      for (int v : verts) order.push_back(v);
      return;
    }
    int id = next_id++;
    for (int v : verts) part[v] = id;
    // split into connected components
    std::vector<std::vector<int>> comps;
    for (int v : verts) dist[v] = -1;
    for (int s : verts) {
      if (dist[s] >= 0) continue;
      std::vector<int> comp{s};
      dist[s] = 0;
      for (size_t h = 0; h < comp.size(); h++)
        for (int w : adj[comp[h]])
          if (part[w] == id && dist[w] < 0) {
            dist[w] = 0;
            comp.push_back(w);
          }
      comps.push_back(std::move(comp));
    }
    if (comps.size() > 1) {
      for (auto& c : comps) rec(c);
      return;
    }
    auto bfs = [&](int s, std::vector<int>& seq) {
      for (int v : verts) dist[v] = -1;
      seq.assign(1, s);
      dist[s] = 0;
      for (size_t h = 0; h < seq.size(); h++)
        for (int w : adj[seq[h]])
          if (part[w] == id && dist[w] < 0) {
            dist[w] = dist[seq[h]] + 1;
            seq.push_back(w);
          }
    };
    std::vector<int> seq;
    int start = *std::min_element(verts.begin(), verts.end());
    bfs(start, seq);
    bfs(seq.back(), seq);
    bfs(seq.back(), seq);
    int L = dist[seq.back()];
    if (L < 2) {
      for (int v : verts) order.push_back(v);
      return;
    }
    int mid = L / 2;
    std::vector<int> A, B, S;
    for (int v : seq) {
      if (dist[v] < mid) A.push_back(v);
      else if (dist[v] > mid) B.push_back(v);
      else S.push_back(v);
    }
    rec(A);
    rec(B);
    std::sort(S.begin(), S.end());
    for (int v : S) order.push_back(v);
  };
  std::vector<int> all(n);
  for (int i = 0; i < n; i++) all[i] = i;
  rec(all);
  return order;
}

}  // namespace

// Builds the plan for one topology. Returns the packed int32 buffer; layout
// and status are written to L. err receives a message on failure.
std::vector<int32_t> build_pattern(const kg_element* el, int ne, int n, AnalysisKind ak, Ordering ordk,
                                   const double* rep, double s_rep, PatLayout& L, std::string& err) {
  std::memset(&L, 0, sizeof(L));
  L.n = n;
  L.ne = ne;
  L.status = KG_OK;
  std::vector<Entry> A;
  std::vector<Entry> R;  // rhs: c unused
  auto add = [&](int r, int c, int e, int k) {
    if (r >= 0 && c >= 0) A.push_back({r, c, (e << 3) | k});
  };
  auto rhs = [&](int r, int e, int k) {
    if (r >= 0) R.push_back({r, 0, (e << 3) | k});
  };
  auto four = [&](int p, int m, int e) {
    add(p, p, e, CK_YP);
    add(m, m, e, CK_YP);
    if (p >= 0 && m >= 0) {
      add(p, m, e, CK_YN);
      add(m, p, e, CK_YN);
    }
  };
  auto incidence = [&](int p, int m, int c, int e) {  // stampable.rs:961-983
    add(p, c, e, CK_ONEP);
    add(c, p, e, CK_ONEP);
    add(m, c, e, CK_ONEN);
    add(c, m, e, CK_ONEN);
  };
  bool has_dyn = false, has_diode = false;
  for (int e = 0; e < ne; e++) {
    const kg_element& x = el[e];
    int p = x.plus, m = x.minus, c = x.branch;
    const bool ctl = x.type >= KG_E && x.type <= KG_H;
    const int cp = ctl ? x.ctrl_plus : -1, cm = ctl ? x.ctrl_minus : -1, cb = ctl ? x.ctrl_branch : -1;
    if (p >= n || m >= n || c >= n || cp >= n || cm >= n || cb >= n) {
      err = "element index out of range";
      L.status = KG_INVALID_ARGUMENT;
      return {};
    }
    switch (x.type) {
      case KG_R:  // stampable.rs:202-297
        if (x.g2) {
          incidence(p, m, c, e);
          add(c, c, e, CK_BR);
        } else {
          four(p, m, e);
        }
        break;
      case KG_C:  // stampable.rs:360-506
        if (ak == AK_AC) {
          if (!x.g2) four(p, m, e);
          else {
            add(c, p, e, CK_YN);
            add(c, m, e, CK_YP);
            add(c, c, e, CK_ONEP);
          }
        } else if (ak == AK_TRAN) {
          four(p, m, e);
          rhs(p, e, RK_CP);
          rhs(m, e, RK_CM);
          has_dyn = true;
        }
        break;
      case KG_L:  // stampable.rs:676-817
        incidence(p, m, c, e);
        if (ak != AK_DC) add(c, c, e, CK_BR);
        if (ak == AK_TRAN) {
          rhs(c, e, RK_IND);
          has_dyn = true;
        }
        break;
      case KG_V:  // stampable.rs:960-1050
        incidence(p, m, c, e);
        rhs(c, e, RK_SRC);
        break;
      case KG_I:  // stampable.rs:508-585
        if (ak != AK_AC) add(c, c, e, CK_ONEP);
        add(p, c, e, CK_ONEP);
        add(m, c, e, CK_ONEN);
        rhs(c, e, RK_SRC);
        break;
      case KG_D:  // stampable.rs:587-674
        four(p, m, e);
        if (ak != AK_AC) {
          rhs(p, e, RK_DP);
          rhs(m, e, RK_DM);
        }
        has_diode = true;
        break;
      // Controlled sources (SPICE conventions; no krets counterpart, README.md:56-62).
      // The same stamps in every analysis: they are linear and memoryless.
      case KG_E:  // V(p) - V(m) - k (V(cp) - V(cm)) = 0 on the branch row
        incidence(p, m, c, e);
        add(c, cp, e, CK_KN);
        add(c, cm, e, CK_KP);
        break;
      case KG_G:  // k (V(cp) - V(cm)) flows from p through the source to m
        add(p, cp, e, CK_KP);
        add(p, cm, e, CK_KN);
        add(m, cp, e, CK_KN);
        add(m, cm, e, CK_KP);
        break;
      case KG_F:  // k I(cb) flows from p through the source to m
        add(p, cb, e, CK_KP);
        add(m, cb, e, CK_KN);
        break;
      case KG_H:  // V(p) - V(m) - k I(cb) = 0 on the branch row
        incidence(p, m, c, e);
        add(c, cb, e, CK_KN);
        break;
      default:
        err = "unknown element type";
        L.status = KG_INVALID_ELEMENT_FORMAT;
        return {};
    }
  }
  L.nonlinear = has_diode ? 1 : 0;
  L.has_diode = has_diode ? 1 : 0;
  L.has_dyn = has_dyn ? 1 : 0;

  // unique pattern, row and column adjacency (original numbering)
  std::vector<std::vector<int>> rows(n), cols(n);
  {
    std::vector<std::pair<int, int>> pr;
    for (auto& a : A) pr.push_back({a.r, a.c});
    std::sort(pr.begin(), pr.end());
    pr.erase(std::unique(pr.begin(), pr.end()), pr.end());
    for (auto& q : pr) {
      rows[q.first].push_back(q.second);
      cols[q.second].push_back(q.first);
    }
  }
  // structural matching: diagonal first, then augmenting paths (MC21 style)
  std::vector<int> mrow(n, -1), rowm(n, -1);
  for (int c = 0; c < n; c++)
    if (std::binary_search(rows[c].begin(), rows[c].end(), c)) {
      mrow[c] = c;
      rowm[c] = c;
    }
  std::vector<int> seen(n, -1);
  std::function<bool(int, int)> augment = [&](int c, int stamp) -> bool {
    for (int r : cols[c]) {
      if (seen[r] == stamp) continue;
      seen[r] = stamp;
      if (rowm[r] < 0 || augment(rowm[r], stamp)) {
        rowm[r] = c;
        mrow[c] = r;
        return true;
      }
    }
    return false;
  };
  for (int c = 0; c < n; c++) {
    if (mrow[c] >= 0) continue;
    if (!augment(c, c)) {
      L.status = KG_DECOMPOSITION;
      err = "structurally singular MNA matrix (no pivot for unknown " + std::to_string(c) + ")";
      return {};
    }
  }
  // symmetric order of the matched matrix B(c, c') = A(mrow[c], c')
  std::vector<std::vector<int>> adj(n);
  for (int r = 0; r < n; r++)
    for (int c2 : rows[r]) {
      int u = rowm[r];
      if (u != c2) {
        adj[u].push_back(c2);
        adj[c2].push_back(u);
      }
    }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::vector<int> ord;
  if (ordk == ORD_MINDEG) ord = order_min_degree(n, adj);
  else if (ordk == ORD_BALANCED) ord = order_balanced(n, adj);
  else if (ordk == ORD_NESTED) ord = order_nested(n, adj);
  else {
    ord.resize(n);
    for (int i = 0; i < n; i++) ord[i] = i;
  }
  std::vector<int> cpos(n), rpos(n);
  for (int j = 0; j < n; j++) cpos[ord[j]] = j;

  // Static pivot sequence from one representative numeric factorisation
  // (the approach of KLU's refactorisation): columns in the fill-reducing
  // order, rows chosen by threshold partial pivoting that keeps the matched
  // row when it is within 1e-3 of the column maximum (the
  // threshold of KLU and UMFPACK). This removes the exact
  // cancellations of MNA incidence entries (e.g. an inductor short in DC)
  // that a purely structural matching cannot see.
  std::vector<int> prow(n, -1);
  {
    using cd = std::complex<double>;
    const double vt = 1.380649e-23 * 300.0 / 1.602176634e-19;  // constants.rs:1-4
    auto coef = [&](int32_t code) -> cd {
      int e = code >> 3, k = code & 7;
      const kg_element& x = el[e];
      double v = rep ? rep[e * KG_NPARAM] : 1.0;
      auto y = [&]() -> cd {
        if (x.type == KG_R) return cd(1.0 / v, 0.0);
        if (x.type == KG_C) return ak == AK_AC ? cd(0.0, s_rep * v) : cd(s_rep * v, 0.0);
        if (x.type == KG_D) {
          double nvt = (rep ? rep[e * KG_NPARAM + 1] : 1.0) * vt;
          return cd(v / nvt * std::exp(0.5 / nvt), 0.0);
        }
        return cd(0.0, 0.0);
      };
      switch (k) {
        case CK_YP: return y();
        case CK_YN: return -y();
        case CK_ONEP: return cd(1.0, 0.0);
        case CK_ONEN: return cd(-1.0, 0.0);
        case CK_KP: return cd(v, 0.0);
        case CK_KN: return cd(-v, 0.0);
        default:
          if (x.type == KG_L) return ak == AK_AC ? cd(0.0, -s_rep * v) : cd(-s_rep * v, 0.0);
          if (x.type == KG_R) return cd(-v, 0.0);
          return cd(1.0, 0.0);
      }
    };
    std::vector<std::map<int, cd>> rowv(n);  // original row -> (permuted column -> value)
    for (auto& a : A) rowv[a.r][cpos[a.c]] += coef(a.code);
    std::vector<std::set<int>> colr(n);  // permuted column -> rows that may hold it
    for (int r = 0; r < n; r++)
      for (auto& kv : rowv[r]) colr[kv.first].insert(r);
    std::vector<char> used(n, 0);
    for (int k = 0; k < n; k++) {
      int pref = mrow[ord[k]];
      double best = 0.0;
      int arg = -1;
      for (int r : colr[k]) {
        if (used[r]) continue;
        auto it = rowv[r].find(k);
        double a = it == rowv[r].end() ? 0.0 : std::abs(it->second);
        if (a > best) { best = a; arg = r; }
      }
      int p = -1;
      if (best > 0.0) {
        auto it = used[pref] ? rowv[pref].end() : rowv[pref].find(k);
        double ap = it == rowv[pref].end() ? 0.0 : std::abs(it->second);
        p = ap >= 1e-3 * best ? pref : arg;
      } else {  // numerically singular representative: keep the structural match
        p = !used[pref] ? pref : -1;
        for (int r : colr[k])
          if (p < 0 && !used[r]) p = r;
        if (p < 0) {
          for (int r = 0; r < n && p < 0; r++)
            if (!used[r]) p = r;
        }
      }
      prow[k] = p;
      used[p] = 1;
      auto pit = rowv[p].find(k);
      cd pv = pit == rowv[p].end() ? cd(0.0, 0.0) : pit->second;
      if (std::abs(pv) == 0.0) continue;
      for (int r : colr[k]) {
        if (used[r]) continue;
        auto it = rowv[r].find(k);
        if (it == rowv[r].end()) continue;
        cd l = it->second / pv;
        rowv[r].erase(it);
        for (auto& kv : rowv[p]) {
          if (kv.first <= k) continue;
          rowv[r][kv.first] -= l * kv.second;
          colr[kv.first].insert(r);
        }
      }
    }
  }
  for (int i = 0; i < n; i++) rpos[prow[i]] = i;

  // symbolic LU of M(i, j) = A(prow[i], ord[j])
  std::vector<std::vector<int>> mrows(n);
  for (int r = 0; r < n; r++)
    for (int c2 : rows[r]) mrows[rpos[r]].push_back(cpos[c2]);
  std::vector<std::vector<int>> Lp(n), Up(n);
  std::vector<std::vector<int>> full(n);
  for (int i = 0; i < n; i++) {
    std::set<int> S(mrows[i].begin(), mrows[i].end());
    for (auto it = S.begin(); it != S.end() && *it < i; ++it) {
      int k = *it;
      for (int j : Up[k]) S.insert(j);
    }
    if (!S.count(i)) {
      L.status = KG_DECOMPOSITION;
      err = "internal: missing diagonal after matching";
      return {};
    }
    for (int j : S) {
      if (j < i) Lp[i].push_back(j);
      else if (j > i) Up[i].push_back(j);
    }
    full[i].assign(S.begin(), S.end());
  }
  // CSR of L+U
  std::vector<int32_t> row_ptr(n + 1, 0), col, diag(n);
  for (int i = 0; i < n; i++) {
    row_ptr[i + 1] = row_ptr[i] + (int)full[i].size();
    for (int j : full[i]) {
      if (j == i) diag[i] = (int32_t)col.size();
      col.push_back(j);
    }
  }
  int nnz = (int)col.size();
  L.nnz = nnz;
  auto slot = [&](int i, int j) -> int32_t {
    auto b = col.begin() + row_ptr[i], e = col.begin() + row_ptr[i + 1];
    auto it = std::lower_bound(b, e, j);
    return (it != e && *it == j) ? (int32_t)(it - col.begin()) : -1;
  };
  // factor program (up-looking, row oriented)
  std::vector<int32_t> fptr(n + 1, 0), fprog;
  for (int i = 0; i < n; i++) {
    fprog.push_back((int32_t)Lp[i].size());
    for (int k : Lp[i]) {
      fprog.push_back(slot(i, k));
      fprog.push_back(diag[k]);
      fprog.push_back((int32_t)Up[k].size());
      for (int j : Up[k]) {
        fprog.push_back(slot(i, j));
        fprog.push_back(slot(k, j));
      }
    }
    fptr[i + 1] = (int32_t)fprog.size();
  }
  // levels
  auto group = [&](const std::vector<int>& lev, std::vector<int32_t>& ptr, std::vector<int32_t>& rws) {
    int nl = 0;
    for (int v : lev) nl = std::max(nl, v + 1);
    ptr.assign(nl + 1, 0);
    for (int v : lev) ptr[v + 1]++;
    for (int l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
    rws.assign(n, 0);
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; i++) rws[fill[lev[i]]++] = i;
    return nl;
  };
  std::vector<int> lev(n, 0), blev(n, 0), tlev(n, 0), tblev(n, 0);
  for (int i = 0; i < n; i++)
    for (int k : Lp[i]) lev[i] = std::max(lev[i], lev[k] + 1);
  for (int i = n - 1; i >= 0; i--)
    for (int j : Up[i]) blev[i] = std::max(blev[i], blev[j] + 1);
  // transposed structures: column lists of U (k < i) and L (j > i)
  std::vector<std::vector<std::pair<int, int>>> ucol(n), lcol(n);
  for (int i = 0; i < n; i++) {
    for (int k : Lp[i]) lcol[k].push_back({i, slot(i, k)});
    for (int j : Up[i]) ucol[j].push_back({i, slot(i, j)});
  }
  for (int i = 0; i < n; i++)
    for (auto& kv : ucol[i]) tlev[i] = std::max(tlev[i], tlev[kv.first] + 1);
  for (int i = n - 1; i >= 0; i--)
    for (auto& kv : lcol[i]) tblev[i] = std::max(tblev[i], tblev[kv.first] + 1);
  std::vector<int32_t> lvl_ptr, lvl_rows, blvl_ptr, blvl_rows, tlvl_ptr, tlvl_rows, tblvl_ptr, tblvl_rows;
  L.nlev = group(lev, lvl_ptr, lvl_rows);
  L.nblev = group(blev, blvl_ptr, blvl_rows);
  L.ntlev = group(tlev, tlvl_ptr, tlvl_rows);
  L.ntblev = group(tblev, tblvl_ptr, tblvl_rows);
  std::vector<int32_t> ut_ptr(n + 1, 0), ut, lt_ptr(n + 1, 0), lt;
  for (int i = 0; i < n; i++) {
    for (auto& kv : ucol[i]) { ut.push_back(kv.first); ut.push_back(kv.second); }
    ut_ptr[i + 1] = (int32_t)ut.size();
    for (auto& kv : lcol[i]) { lt.push_back(kv.first); lt.push_back(kv.second); }
    lt_ptr[i + 1] = (int32_t)lt.size();
  }
  // gather lists
  std::vector<std::vector<int32_t>> sc(nnz), rc(n);
  for (auto& a : A) {
    int s = slot(rpos[a.r], cpos[a.c]);
    sc[s].push_back(a.code);
  }
  for (auto& b : R) rc[rpos[b.r]].push_back(b.code);
  std::vector<int32_t> scon_ptr(nnz + 1, 0), scon, rcon_ptr(n + 1, 0), rcon;
  for (int s = 0; s < nnz; s++) {
    scon.insert(scon.end(), sc[s].begin(), sc[s].end());
    scon_ptr[s + 1] = (int32_t)scon.size();
  }
  for (int i = 0; i < n; i++) {
    rcon.insert(rcon.end(), rc[i].begin(), rc[i].end());
    rcon_ptr[i + 1] = (int32_t)rcon.size();
  }
  std::vector<int32_t> col_of(ord.begin(), ord.end()), cposv(cpos.begin(), cpos.end()), isbr(n, 0);
  std::vector<int32_t> et(ne), ep(ne), em(ne), ec(ne);
  for (int e = 0; e < ne; e++) {
    et[e] = el[e].type;
    ep[e] = el[e].plus >= 0 ? cpos[el[e].plus] : -1;
    em[e] = el[e].minus >= 0 ? cpos[el[e].minus] : -1;
    ec[e] = el[e].branch >= 0 ? cpos[el[e].branch] : -1;
    if (el[e].branch >= 0) isbr[cpos[el[e].branch]] = 1;
  }
  // pack
  std::vector<int32_t> buf;
  auto put = [&](const std::vector<int32_t>& v) {
    int32_t off = (int32_t)buf.size();
    buf.insert(buf.end(), v.begin(), v.end());
    while (buf.size() % 4) buf.push_back(0);
    return off;
  };
  // The factor and solve structure first (row pointers, diagonal slots, the
  // factor program, both level lists): a kernel can copy the prefix up to
  // o_col into LDS in one pass.
  L.o_row_ptr = put(row_ptr);
  L.o_diag = put(diag);
  L.o_fprog_ptr = put(fptr);
  L.o_fprog = put(fprog);
  L.o_lvl_ptr = put(lvl_ptr);
  L.o_lvl_rows = put(lvl_rows);
  L.o_blvl_ptr = put(blvl_ptr);
  L.o_blvl_rows = put(blvl_rows);
  L.o_col = put(col);
  L.o_scon_ptr = put(scon_ptr);
  L.o_scon = put(scon);
  L.o_rcon_ptr = put(rcon_ptr);
  L.o_rcon = put(rcon);
  L.o_col_of = put(col_of);
  L.o_cpos = put(cposv);
  L.o_isbr = put(isbr);
  L.o_etype = put(et);
  L.o_ep = put(ep);
  L.o_em = put(em);
  L.o_ec = put(ec);
  L.o_ut_ptr = put(ut_ptr);
  L.o_ut = put(ut);
  L.o_lt_ptr = put(lt_ptr);
  L.o_lt = put(lt);
  L.o_tlvl_ptr = put(tlvl_ptr);
  L.o_tlvl_rows = put(tlvl_rows);
  L.o_tblvl_ptr = put(tblvl_ptr);
  L.o_tblvl_rows = put(tblvl_rows);
  std::vector<int32_t> row_of(n);
  for (int i = 0; i < n; i++) row_of[i] = prow[i];
  L.o_row_of = put(row_of);
  std::vector<int32_t> rposv(rpos.begin(), rpos.end()), erp(ne), erm(ne), erc(ne);
  for (int e = 0; e < ne; e++) {
    erp[e] = el[e].plus >= 0 ? rpos[el[e].plus] : -1;
    erm[e] = el[e].minus >= 0 ? rpos[el[e].minus] : -1;
    erc[e] = el[e].branch >= 0 ? rpos[el[e].branch] : -1;
  }
  L.o_rpos = put(rposv);
  L.o_erp = put(erp);
  L.o_erm = put(erm);
  L.o_erc = put(erc);
  L.total = (int32_t)buf.size();
  if (!validate_plan(L, buf, err)) {
    L.status = KG_UNEXPECTED;
    return {};
  }
  return buf;
}

// Compact plan for the LDS-resident linear transient kernel (see LinHdr).
// Returns false, leaving `out` empty, when the topology does not qualify:
// diodes present, or indices that do not fit 16 bits.
bool build_lin_plan(const PatLayout& L, const std::vector<int32_t>& b, std::vector<uint16_t>& out) {
  out.clear();
  const int n = L.n, nnz = L.nnz, ne = L.ne;
  if (L.status != KG_OK || L.has_diode || L.nonlinear || n >= 0xFFFF || nnz >= 0xFFFF || ne >= (1 << 14))
    return false;
  const int32_t* et = b.data() + L.o_etype;
  const int32_t* ep = b.data() + L.o_ep;
  const int32_t* em = b.data() + L.o_em;
  const int32_t* ec = b.data() + L.o_ec;
  const int32_t* rcp = b.data() + L.o_rcon_ptr;
  const int32_t* rc = b.data() + L.o_rcon;
  // dyn (C, L) and src (V, I) indices per element, in element order
  std::vector<int32_t> dyn_of(ne, -1), src_of(ne, -1);
  std::vector<uint16_t> dyn, src;
  auto u16 = [](int32_t v) { return v < 0 ? LIN_NONE : (uint16_t)v; };
  for (int e = 0; e < ne; e++) {
    if (et[e] == KG_C || et[e] == KG_L) {
      dyn_of[e] = (int32_t)(dyn.size() / LIN_DYN);
      const int kind = et[e] == KG_C ? 1 : 2;
      dyn.insert(dyn.end(), {(uint16_t)((kind << 14) | e), u16(ep[e]), u16(em[e]), u16(ec[e])});
    } else if (et[e] == KG_V || et[e] == KG_I) {
      src_of[e] = (int32_t)(src.size() / LIN_SRC);
      src.insert(src.end(), {(uint16_t)et[e], (uint16_t)e});
    }
  }
  const int ndyn = (int)dyn.size() / LIN_DYN, nsrc = (int)src.size() / LIN_SRC;
  if (ndyn >= (1 << 14) || nsrc >= (1 << 14) || rcp[n] >= 0xFFFF) return false;
  std::vector<uint16_t> rcon;
  for (int q = 0; q < rcp[n]; q++) {
    int e = rc[q] >> 3, k = rc[q] & 7;
    int code;
    if (k == RK_SRC && src_of[e] >= 0) code = (src_of[e] << 2) | LK_SRC;
    else if (k == RK_CP && dyn_of[e] >= 0) code = (dyn_of[e] << 2) | LK_CP;
    else if (k == RK_CM && dyn_of[e] >= 0) code = (dyn_of[e] << 2) | LK_CM;
    else if (k == RK_IND && dyn_of[e] >= 0) code = (dyn_of[e] << 2) | LK_IND;
    else return false;  // diode companions never reach here (has_diode)
    rcon.push_back((uint16_t)code);
  }
  LinHdr h{};
  h.n = n;
  h.nnz = nnz;
  h.nlev = L.nlev;
  h.nblev = L.nblev;
  h.ndyn = ndyn;
  h.ncon = (int32_t)rcon.size();
  h.nsrc = nsrc;
  h.ne = ne;
  const size_t hw = sizeof(LinHdr) / 2;
  std::vector<uint16_t> w(hw, 0);
  auto put = [&](const uint16_t* p, size_t k) {
    while (w.size() % 2) w.push_back(0);  // 4-byte aligned arrays
    int32_t off = (int32_t)w.size();
    w.insert(w.end(), p, p + k);
    return off;
  };
  auto put32 = [&](const int32_t* p, size_t k) {
    std::vector<uint16_t> t(k);
    for (size_t i = 0; i < k; i++) t[i] = (uint16_t)p[i];
    return put(t.data(), k);
  };
  const int32_t* rp = b.data() + L.o_row_ptr;
  const int32_t* dg = b.data() + L.o_diag;
  auto descs = [&](int o_ptr, int o_rows, int nl, bool fwd) {
    std::vector<uint16_t> d;
    const int32_t* lp = b.data() + o_ptr;
    const int32_t* lr = b.data() + o_rows;
    for (int q = 0; q < lp[nl]; q++) {
      const int i = lr[q];
      if (fwd) d.insert(d.end(), {(uint16_t)i, (uint16_t)rp[i], (uint16_t)dg[i], (uint16_t)dg[i]});
      else d.insert(d.end(), {(uint16_t)i, (uint16_t)(dg[i] + 1), (uint16_t)rp[i + 1], (uint16_t)dg[i]});
    }
    return d;
  };
  auto put8 = [&](const std::vector<uint16_t>& d) {  // 8-byte aligned descriptors
    while (w.size() % 4) w.push_back(0);
    int32_t off = (int32_t)w.size();
    w.insert(w.end(), d.begin(), d.end());
    return off;
  };
  h.o_lvl_ptr = put32(b.data() + L.o_lvl_ptr, L.nlev + 1);
  h.o_fdesc = put8(descs(L.o_lvl_ptr, L.o_lvl_rows, L.nlev, true));
  h.o_blvl_ptr = put32(b.data() + L.o_blvl_ptr, L.nblev + 1);
  h.o_bdesc = put8(descs(L.o_blvl_ptr, L.o_blvl_rows, L.nblev, false));
  h.o_col = put32(b.data() + L.o_col, nnz);
  h.o_cpos = put32(b.data() + L.o_cpos, n);
  h.o_rcon_ptr = put32(rcp, n + 1);
  h.o_rcon = put(rcon.data(), rcon.size());
  h.o_dyn = put(dyn.data(), dyn.size());
  h.o_src = put(src.data(), src.size());
  while (w.size() % 8) w.push_back(0);  // 16-byte multiple
  h.words_fwd = (int32_t)w.size();
  // ---- discrete adjoint: transposed solves and R / C / L element records
  {
    const int32_t* utp = b.data() + L.o_ut_ptr;
    const int32_t* ut = b.data() + L.o_ut;
    const int32_t* ltp = b.data() + L.o_lt_ptr;
    const int32_t* lt = b.data() + L.o_lt;
    std::vector<uint16_t> tent, tf, tb;
    auto tdescs = [&](int o_ptr, int o_rows, int nl, const int32_t* pp, const int32_t* pv, bool upper,
                      std::vector<uint16_t>& d) {
      const int32_t* lp = b.data() + o_ptr;
      const int32_t* lr = b.data() + o_rows;
      for (int q = 0; q < lp[nl]; q++) {
        const int i = lr[q];
        const int q0 = (int)tent.size() / 2;
        for (int k = pp[i]; k < pp[i + 1]; k += 2) tent.insert(tent.end(), {(uint16_t)pv[k], (uint16_t)pv[k + 1]});
        const int q1 = (int)tent.size() / 2;
        d.insert(d.end(), {(uint16_t)i, (uint16_t)q0, (uint16_t)q1, (uint16_t)(upper ? dg[i] : 0)});
      }
    };
    tdescs(L.o_tlvl_ptr, L.o_tlvl_rows, L.ntlev, utp, ut, true, tf);
    tdescs(L.o_tblvl_ptr, L.o_tblvl_rows, L.ntblev, ltp, lt, false, tb);
    if (tent.size() / 2 >= 0xFFFF) return false;
    const int32_t* erp = b.data() + L.o_erp;
    const int32_t* erm = b.data() + L.o_erm;
    const int32_t* erc = b.data() + L.o_erc;
    std::vector<uint16_t> adj;
    for (int e = 0; e < ne; e++) {
      if (et[e] != KG_R && et[e] != KG_C && et[e] != KG_L) continue;
      const int kind = et[e] == KG_R ? 0 : et[e] == KG_C ? 1 : 2;
      adj.insert(adj.end(), {(uint16_t)((kind << 14) | e), u16(ep[e]), u16(em[e]), u16(erp[e]), u16(erm[e]),
                             u16(ec[e]), u16(erc[e]), (uint16_t)(dyn_of[e] >= 0 ? dyn_of[e] : LIN_NONE)});
    }
    h.ntlev = L.ntlev;
    h.ntblev = L.ntblev;
    h.o_tlvl_ptr = put32(b.data() + L.o_tlvl_ptr, L.ntlev + 1);
    h.o_tfdesc = put8(tf);
    h.o_tblvl_ptr = put32(b.data() + L.o_tblvl_ptr, L.ntblev + 1);
    h.o_tbdesc = put8(tb);
    h.o_tent = put(tent.data(), tent.size());
    h.nadj = (int32_t)adj.size() / LIN_ADJ;
    h.o_adj = put(adj.data(), adj.size());
    h.tlev0_free = 1;
    const int32_t* tl = b.data() + L.o_tlvl_ptr;
    const int32_t* tr = b.data() + L.o_tlvl_rows;
    for (int q = tl[0]; q < (L.ntlev > 0 ? tl[1] : 0); q++)
      if (utp[tr[q]] != utp[tr[q] + 1]) h.tlev0_free = 0;
    const int32_t* tbl = b.data() + L.o_tblvl_ptr;
    const int32_t* tbr = b.data() + L.o_tblvl_rows;
    for (int q = tbl[0]; q < (L.ntblev > 0 ? tbl[1] : 0); q++)
      if (ltp[tbr[q]] != ltp[tbr[q] + 1]) h.tlev0_free = 0;
  }
  while (w.size() % 8) w.push_back(0);
  h.words = (int32_t)w.size();
  // level-0 forward rows have no strictly lower entries: their forward step
  // is the right-hand side itself
  const int32_t* lp = b.data() + L.o_lvl_ptr;
  const int32_t* lr = b.data() + L.o_lvl_rows;
  h.lev0_free = 1;
  for (int q = lp[0]; q < (L.nlev > 0 ? lp[1] : 0); q++)
    if (dg[lr[q]] != rp[lr[q]]) h.lev0_free = 0;
  std::memcpy(w.data(), &h, sizeof(LinHdr));
  out.swap(w);
  return true;
}

// Checks every index a kernel will dereference before the plan is uploaded.
bool validate_plan(const PatLayout& L, const std::vector<int32_t>& b, std::string& err) {
  const int n = L.n, nnz = L.nnz, ne = L.ne;
  auto in = [](int64_t v, int64_t lo, int64_t hi) { return v >= lo && v < hi; };
  auto fail = [&](const char* what) {
    err = std::string("plan validation failed: ") + what;
    return false;
  };
  const int32_t* rp = b.data() + L.o_row_ptr;
  const int32_t* col = b.data() + L.o_col;
  const int32_t* dg = b.data() + L.o_diag;
  if (rp[0] != 0 || rp[n] != nnz) return fail("row_ptr");
  for (int i = 0; i < n; i++) {
    if (rp[i + 1] < rp[i]) return fail("row_ptr order");
    if (!in(dg[i], rp[i], rp[i + 1]) || col[dg[i]] != i) return fail("diag");
    for (int p = rp[i]; p < rp[i + 1]; p++)
      if (!in(col[p], 0, n) || (p > rp[i] && col[p] <= col[p - 1])) return fail("col");
  }
  const int32_t* fpp = b.data() + L.o_fprog_ptr;
  const int32_t* fp = b.data() + L.o_fprog;
  for (int i = 0; i < n; i++) {
    int q = fpp[i], e = fpp[i + 1];
    int nk = fp[q++];
    for (int k = 0; k < nk; k++) {
      if (q + 3 > e) return fail("fprog length");
      if (!in(fp[q], 0, nnz) || !in(fp[q + 1], 0, nnz)) return fail("fprog slot");
      int nu = fp[q + 2];
      q += 3;
      for (int u = 0; u < nu; u++, q += 2)
        if (q + 2 > e || !in(fp[q], 0, nnz) || !in(fp[q + 1], 0, nnz)) return fail("fprog update");
    }
    if (q != e) return fail("fprog end");
  }
  auto levels = [&](int o_ptr, int o_rows, int nl) {
    const int32_t* lp = b.data() + o_ptr;
    const int32_t* lr = b.data() + o_rows;
    if (lp[0] != 0 || lp[nl] != n) return false;
    std::vector<char> seen(n, 0);
    for (int q = 0; q < n; q++) {
      if (!in(lr[q], 0, n) || seen[lr[q]]) return false;
      seen[lr[q]] = 1;
    }
    return true;
  };
  if (!levels(L.o_lvl_ptr, L.o_lvl_rows, L.nlev) || !levels(L.o_blvl_ptr, L.o_blvl_rows, L.nblev) ||
      !levels(L.o_tlvl_ptr, L.o_tlvl_rows, L.ntlev) || !levels(L.o_tblvl_ptr, L.o_tblvl_rows, L.ntblev))
    return fail("levels");
  const int32_t* scp = b.data() + L.o_scon_ptr;
  const int32_t* sc = b.data() + L.o_scon;
  for (int s = 0; s <= nnz; s++)
    if (s > 0 && scp[s] < scp[s - 1]) return fail("scon_ptr");
  for (int q = 0; q < scp[nnz]; q++)
    if (!in(sc[q] >> 3, 0, ne) || (sc[q] & 7) > CK_KN) return fail("scon code");
  const int32_t* rcp = b.data() + L.o_rcon_ptr;
  const int32_t* rc = b.data() + L.o_rcon;
  for (int q = 0; q < rcp[n]; q++)
    if (!in(rc[q] >> 3, 0, ne) || (rc[q] & 7) > RK_IND) return fail("rcon code");
  for (int j = 0; j < n; j++)
    if (!in(b[L.o_col_of + j], 0, n) || !in(b[L.o_row_of + j], 0, n) || !in(b[L.o_cpos + j], 0, n) ||
        !in(b[L.o_rpos + j], 0, n))
      return fail("permutation");
  for (int e = 0; e < ne; e++) {
    if (!in(b[L.o_etype + e], KG_R, KG_H + 1)) return fail("etype");
    for (int o : {L.o_ep, L.o_em, L.o_ec, L.o_erp, L.o_erm, L.o_erc})
      if (!in(b[o + e], -1, n)) return fail("element index");
    int ty = b[L.o_etype + e];
    if ((ty == KG_V || ty == KG_L || ty == KG_I || ty == KG_E || ty == KG_H) && b[L.o_ec + e] < 0)
      return fail("branch index");
  }
  for (int o_ptr : {L.o_ut_ptr, L.o_lt_ptr}) {
    const int32_t* pp = b.data() + o_ptr;
    const int32_t* v = b.data() + (o_ptr == L.o_ut_ptr ? L.o_ut : L.o_lt);
    for (int i = 0; i < n; i++)
      for (int q = pp[i]; q < pp[i + 1]; q += 2)
        if (!in(v[q], 0, n) || !in(v[q + 1], 0, nnz)) return fail("transposed lists");
  }
  return true;
}

}  // namespace kg
