// kh_symbolic.cpp — host symbolic phase (see kh_symbolic.h).
#include "kh_symbolic.h"

#include <algorithm>
#include <map>
#include <set>

#include "kh_devices.h"

namespace kh {

std::vector<int32_t> minimum_degree(int n, const std::vector<int32_t>& ri, const std::vector<int32_t>& ci) {
  std::vector<std::vector<int32_t>> adj(n);
  for (size_t k = 0; k < ri.size(); k++) {
    int i = ri[k], j = ci[k];
    if (i != j) {
      adj[i].push_back(j);
      adj[j].push_back(i);
    }
  }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  // (degree, index): the first element is the reference's choice, the
  // lowest-index vertex of least current degree (solver/sparse.go:428-433).
  std::set<std::pair<int32_t, int32_t>> pq;
  for (int i = 0; i < n; i++) pq.insert({(int32_t)adj[i].size(), i});
  std::vector<int32_t> order;
  order.reserve(n);
  std::vector<int32_t> nbr, tmp;
  while (!pq.empty()) {
    auto it = pq.begin();
    int v = it->second;
    pq.erase(it);
    order.push_back(v);
    nbr.swap(adj[v]);
    adj[v].clear();
    for (int32_t u : nbr) {
      // adj[u] <- (adj[u] \ {v}) U (nbr \ {u})   (solver/sparse.go:440-447)
      pq.erase({(int32_t)adj[u].size(), u});
      auto& au = adj[u];
      tmp.clear();
      size_t a = 0, b = 0;
      while (a < au.size() || b < nbr.size()) {
        int32_t x;
        if (b >= nbr.size() || (a < au.size() && au[a] < nbr[b])) x = au[a++];
        else if (a >= au.size() || nbr[b] < au[a]) x = nbr[b++];
        else { x = au[a]; a++; b++; }
        if (x == v || x == u) continue;
        tmp.push_back(x);
      }
      au.swap(tmp);
      pq.insert({(int32_t)au.size(), u});
    }
    nbr.clear();
  }
  return order;
}

Pattern analyze_pattern(int n, const std::vector<int32_t>& slot_i, const std::vector<int32_t>& slot_j,
                        const std::vector<int32_t>& slots) {
  Pattern P;
  std::vector<int32_t> ri, cj;
  ri.reserve(slots.size());
  cj.reserve(slots.size());
  for (int32_t s : slots) {
    ri.push_back(slot_i[s]);
    cj.push_back(slot_j[s]);
  }
  std::vector<int32_t> order = minimum_degree(n, ri, cj);
  P.q.assign(n, 0);
  P.qinv.assign(n, 0);
  for (int k = 0; k < n; k++) {
    P.q[k] = order[k];
    P.qinv[order[k]] = k;
  }
  P.ap.assign(n + 1, 0);
  for (size_t k = 0; k < slots.size(); k++) P.ap[P.qinv[cj[k]] + 1]++;
  for (int k = 0; k < n; k++) P.ap[k + 1] += P.ap[k];
  P.arow.assign(slots.size(), 0);
  P.aslot.assign(slots.size(), 0);
  std::vector<int32_t> next(P.ap.begin(), P.ap.begin() + n);
  for (size_t k = 0; k < slots.size(); k++) {
    int c = P.qinv[cj[k]];
    int p = next[c]++;
    P.arow[p] = P.qinv[ri[k]];
    P.aslot[p] = slots[k];
  }
  std::vector<std::pair<int32_t, int32_t>> col;
  for (int c = 0; c < n; c++) {
    col.clear();
    for (int p = P.ap[c]; p < P.ap[c + 1]; p++) col.push_back({P.arow[p], P.aslot[p]});
    std::sort(col.begin(), col.end());
    for (int p = P.ap[c]; p < P.ap[c + 1]; p++) {
      P.arow[p] = col[p - P.ap[c]].first;
      P.aslot[p] = col[p - P.ap[c]].second;
    }
  }
  // Structural fill with the diagonal as pivot where it is reachable (the
  // common outcome of the diagonal-preference rule), else the lowest row.
  std::vector<int32_t> pinv(n, -1), lp(n + 1, 0), li, mark(n, -1), stack, pst, out;
  int64_t nu = 0;
  for (int k = 0; k < n; k++) {
    lp[k] = (int32_t)li.size();
    out.clear();
    for (int p = P.ap[k]; p < P.ap[k + 1]; p++) {
      int r = P.arow[p];
      if (mark[r] == k) continue;
      stack.assign(1, r);
      pst.assign(1, 0);
      mark[r] = k;
      while (!stack.empty()) {
        int v = stack.back();
        int j = pinv[v];
        bool pushed = false;
        if (j >= 0) {
          for (int q = lp[j] + pst.back(); q < lp[j + 1]; q++) {
            pst.back()++;
            int w = li[q];
            if (mark[w] == k) continue;
            mark[w] = k;
            stack.push_back(w);
            pst.push_back(0);
            pushed = true;
            break;
          }
        }
        if (!pushed) {
          out.push_back(v);
          stack.pop_back();
          pst.pop_back();
        }
      }
    }
    int piv = -1;
    for (int r : out) {
      if (pinv[r] >= 0) { nu++; continue; }
      if (piv < 0 || r < piv) piv = r;
    }
    if (pinv[k] < 0 && mark[k] == k) piv = k;
    if (piv < 0) {  // structurally singular column: keep counting without a pivot
      lp[k + 1] = (int32_t)li.size();
      continue;
    }
    pinv[piv] = k;
    for (int r : out)
      if (pinv[r] < 0) li.push_back(r);
    lp[k + 1] = (int32_t)li.size();
  }
  P.fill_estimate = (int64_t)li.size() + nu + n;
  return P;
}

int32_t PatternPool::add(const Pattern& p, const std::string& key) {
  PatDesc d = {};
  auto put = [&](const std::vector<int32_t>& v) {
    int32_t off = (int32_t)pool.size();
    pool.insert(pool.end(), v.begin(), v.end());
    return off;
  };
  d.q = put(p.q);
  d.qinv = put(p.qinv);
  d.ap = put(p.ap);
  d.arow = put(p.arow);
  d.aslot = put(p.aslot);
  d.nnz = (int32_t)p.arow.size();
  int32_t id = (int32_t)desc.size();
  desc.push_back(d);
  by_key[key] = id;
  if (p.fill_estimate > max_fill) max_fill = p.fill_estimate;
  return id;
}

// Add pairs (terminal index pairs) each device's Load can stamp, in the order
// of its code over all modes (device/passive.go:72-83, device/reactive.go:60-106
// and :230-259, device/source.go:102-112, device/diode.go:262-273,
// device/mosfet.go:591-611).
static const int kPairsR[][2] = {{0, 0}, {1, 1}, {0, 1}, {1, 0}};
static const int kPairsL[][2] = {{0, 2}, {1, 2}, {2, 0}, {2, 1}, {2, 2}};
static const int kPairsV[][2] = {{0, 2}, {1, 2}, {2, 0}, {2, 1}};
static const int kPairsD[][2] = {{2, 2}, {1, 1}, {2, 1}, {1, 2}, {0, 0}, {0, 2}, {2, 0}};
static const int kPairsM[][2] = {
    {4, 2}, {4, 5}, {5, 2}, {5, 5}, {4, 4}, {4, 5}, {5, 4}, {5, 5}, {4, 3}, {4, 5}, {5, 3}, {5, 5},  // channel
    {3, 3}, {3, 5}, {5, 3}, {5, 5},                                                                  // bulk-source
    {3, 3}, {3, 4}, {4, 3}, {4, 4},                                                                  // bulk-drain
    {2, 2}, {2, 5}, {5, 2}, {5, 5}, {2, 2}, {2, 4}, {4, 2}, {4, 4}, {2, 2}, {2, 3}, {3, 2}, {3, 3},  // gate charges
    {3, 3}, {3, 5}, {5, 3}, {5, 5}, {3, 3}, {3, 4}, {4, 3}, {4, 4},                                  // junction charges
    {0, 0}, {4, 4}, {0, 4}, {4, 0}, {1, 1}, {5, 5}, {1, 5}, {5, 1}};                                 // RD, RS

// BJT terms: 0 c, 1 e, 2 b, 3 s, 4 cp, 5 bp, 6 ep (device/bjt.go:492-515, in Load's Add order).
static const int kPairsQ[][2] = {
    {4, 5}, {4, 6}, {6, 5}, {6, 6}, {4, 5}, {4, 4}, {6, 5}, {6, 4},  // transport (cp -> ep)
    {5, 5}, {5, 6}, {6, 5}, {6, 6},                                  // base-emitter
    {5, 5}, {5, 4}, {4, 5}, {4, 4},                                  // base-collector
    {3, 3}, {3, 4}, {4, 3}, {4, 4},                                  // substrate, vertical (s -> cp)
    {5, 5}, {5, 3}, {3, 5}, {3, 3},                                  // substrate, lateral (bp -> s)
    {5, 5}, {5, 6}, {6, 5}, {6, 6}, {5, 5}, {5, 4}, {6, 5}, {6, 4},  // base-emitter charge
    {2, 2}, {2, 4}, {4, 2}, {4, 4},                                  // external base-collector charge
    {2, 2}, {5, 5}, {2, 5}, {5, 2},                                  // RB
    {0, 0}, {4, 4}, {0, 4}, {4, 0},                                  // RC
    {1, 1}, {6, 6}, {1, 6}, {6, 1}};                                 // RE

// Controlled sources (device/controlled.go:76-107): terminals 0 n+, 1 n-, 2
// own branch, 3 nc+ or the controlling branch, 4 nc-; pairs in stamp order.
static const int kPairsE[6][2] = {{0, 2}, {1, 2}, {2, 0}, {2, 1}, {2, 3}, {2, 4}};
static const int kPairsH[5][2] = {{0, 2}, {1, 2}, {2, 0}, {2, 1}, {2, 3}};
static const int kPairsG[4][2] = {{0, 3}, {0, 4}, {1, 3}, {1, 4}};
static const int kPairsF[2][2] = {{0, 3}, {1, 3}};
// Mutual inductance (device/mutual.go:104-108): terminals 0 and 1 are the
// coupled inductors' branches.
static const int kPairsK[2][2] = {{0, 1}, {1, 0}};

bool Topology::build(int32_t numNodes, const kh_element* elems, int32_t n_elems, int32_t np, const double* params_c0) {
  num_nodes = numNodes;
  nel = n_elems;
  nparams = np;
  et.assign((size_t)nel * ET_W, -1);
  int branches = 0;
  std::vector<int32_t> internal;
  auto alloc = [&]() { return numNodes - 1 + branches++; };
  int soff = 0;
  nonlin = false;
  kinds = 0;
  nl_idx.clear();
  rx_idx.clear();
  src_idx.clear();
  level_rows.clear();
  level_vals.clear();
  for (int e = 0; e < nel; e++) {
    const kh_element& E = elems[e];
    int32_t* t = &et[(size_t)e * ET_W];
    t[ET_KIND] = E.kind;
    t[ET_POFF] = E.poff;
    for (int k = 0; k < KH_NT; k++) t[ET_TERM + k] = -1;
    auto unk = [&](int node) { return node - 1; };
    int ns = 0, npar = 0;
    switch (E.kind) {
      case K_R: t[ET_TERM] = unk(E.node[0]); t[ET_TERM + 1] = unk(E.node[1]); ns = NS_R; npar = KH_NPAR_R; break;
      case K_C: t[ET_TERM] = unk(E.node[0]); t[ET_TERM + 1] = unk(E.node[1]); ns = NS_C; npar = KH_NPAR_C; break;
      case K_I: t[ET_TERM] = unk(E.node[0]); t[ET_TERM + 1] = unk(E.node[1]); ns = NS_I; npar = KH_NPAR_SRC; break;
      case K_V:
      case K_L:
        t[ET_TERM] = unk(E.node[0]);
        t[ET_TERM + 1] = unk(E.node[1]);
        t[ET_TERM + 2] = alloc();
        ns = E.kind == K_V ? NS_V : NS_L;
        npar = E.kind == K_V ? KH_NPAR_SRC : KH_NPAR_L;
        break;
      case K_D: {
        t[ET_TERM] = unk(E.node[0]);
        t[ET_TERM + 1] = unk(E.node[1]);
        t[ET_TERM + 2] = t[ET_TERM];
        layout_rows.push_back(E.poff + 14);
        layout_vals.push_back(params_c0[E.poff + 14] > 0);
        if (params_c0[E.poff + 14] > 0) {
          int j = alloc();
          internal.push_back(j);
          t[ET_TERM + 2] = j;
        }
        ns = NS_D;
        npar = KH_NPAR_D;
        nonlin = true;
        break;
      }
      case K_M: {
        int d = unk(E.node[0]), s = unk(E.node[1]);
        t[ET_TERM] = d; t[ET_TERM + 1] = s; t[ET_TERM + 2] = unk(E.node[2]); t[ET_TERM + 3] = unk(E.node[3]);
        int dp = d, sp = s;
        for (int r : {7, 8}) {
          layout_rows.push_back(E.poff + r);
          layout_vals.push_back(params_c0[E.poff + r] > 0);
        }
        if (params_c0[E.poff + 7] > 0) { dp = alloc(); internal.push_back(dp); }
        if (params_c0[E.poff + 8] > 0) { sp = alloc(); internal.push_back(sp); }
        t[ET_TERM + 4] = dp; t[ET_TERM + 5] = sp;
        ns = NS_M;
        npar = KH_NPAR_M;
        nonlin = true;
        break;
      }
      case K_Q: {  // cp, bp, ep behind RC, RB, RE (device/bjt.go:250-268)
        int c = unk(E.node[0]), em = unk(E.node[1]), b = unk(E.node[2]);
        t[ET_TERM] = c; t[ET_TERM + 1] = em; t[ET_TERM + 2] = b; t[ET_TERM + 3] = unk(E.node[3]);
        int cp = c, bp = b, ep = em;
        for (int r : {18, 14, 17}) {
          layout_rows.push_back(E.poff + r);
          layout_vals.push_back(params_c0[E.poff + r] > 0);
        }
        if (params_c0[E.poff + 18] > 0) { cp = alloc(); internal.push_back(cp); }
        if (params_c0[E.poff + 14] > 0) { bp = alloc(); internal.push_back(bp); }
        if (params_c0[E.poff + 17] > 0) { ep = alloc(); internal.push_back(ep); }
        t[ET_TERM + 4] = cp; t[ET_TERM + 5] = bp; t[ET_TERM + 6] = ep;
        ns = NS_Q;
        npar = KH_NPAR_Q;
        nonlin = true;
        break;
      }
      case K_E:
      case K_F:
      case K_G:
      case K_H:
        t[ET_TERM] = unk(E.node[0]);
        t[ET_TERM + 1] = unk(E.node[1]);
        if (E.kind == K_E || E.kind == K_H) t[ET_TERM + 2] = alloc();  // Setup: the branch of a voltage output
        if (E.kind == K_E || E.kind == K_G) {
          t[ET_TERM + 3] = unk(E.node[2]);
          t[ET_TERM + 4] = unk(E.node[3]);
        }
        ns = NS_CTL;
        npar = KH_NPAR_CTL;
        break;
      case K_K:  // Setup allocates nothing; Link below finds the inductors' branches
        ns = NS_K;
        npar = KH_NPAR_K;
        break;
      default:
        error = "element " + std::to_string(e) + ": kind " + std::to_string(E.kind) + " is not supported";
        return false;
    }
    if (E.poff < 0 || E.poff + npar > np) {
      error = "element " + std::to_string(e) + ": parameter block out of range";
      return false;
    }
    t[ET_SOFF] = soff;
    soff += ns;
    kinds |= 1u << E.kind;
    if (E.kind == K_D || E.kind == K_M || E.kind == K_Q) nl_idx.push_back(e);
    if (E.kind == K_C || E.kind == K_L || E.kind == K_D || E.kind == K_M || E.kind == K_Q || E.kind == K_K)
      rx_idx.push_back(e);
    if (E.kind == K_V || E.kind == K_I) src_idx.push_back(e);
    if (E.kind == K_M) {
      if (params_c0[E.poff + 34] >= 2) kinds |= 1u << K_M23;  // level 2 or 3 currents
      level_rows.push_back(E.poff + 34);
      level_vals.push_back(params_c0[E.poff + 34]);
    }
  }
  nstate = soff;
  // Link (device/controlled.go:67-77): F and H sense the branch current of the
  // element node[2] names, which must own a registered branch (V, L, E, H).
  for (int e = 0; e < nel; e++) {
    const kh_element& E = elems[e];
    if (E.kind != K_F && E.kind != K_H) continue;
    const int c = E.node[2];
    const int ck = (c >= 0 && c < nel) ? elems[c].kind : -1;
    if (ck != K_V && ck != K_L && ck != K_E && ck != K_H) {
      error = "element " + std::to_string(e) + ": controlling source not found";
      return false;
    }
    et[(size_t)e * ET_W + ET_TERM + 3] = et[(size_t)c * ET_W + ET_TERM + 2];
  }
  // Mutual.Link (device/mutual.go:57-72): the two coupled inductors' branches.
  for (int e = 0; e < nel; e++) {
    const kh_element& E = elems[e];
    if (E.kind != K_K) continue;
    for (int t = 0; t < 2; t++) {
      const int c = E.node[t];
      if (c < 0 || c >= nel || elems[c].kind != K_L || (t == 1 && c == E.node[0])) {
        error = "element " + std::to_string(e) + ": coupled inductor not found";
        return false;
      }
      et[(size_t)e * ET_W + ET_TERM + t] = et[(size_t)c * ET_W + ET_TERM + 2];
    }
  }
  n = numNodes - 1 + branches;
  if (n <= 0) {
    error = "circuit has no unknowns";
    return false;
  }
  isV.assign(n, 0);
  for (int i = 0; i < numNodes - 1; i++) isV[i] = 1;
  for (int j : internal) isV[j] = 1;
  // Union of value slots, numbered by first appearance in element order.
  std::map<int64_t, int32_t> slot;
  slot_i.clear();
  slot_j.clear();
  auto get = [&](int i, int j) {
    int64_t key = (int64_t)i * n + j;
    auto it = slot.find(key);
    if (it != slot.end()) return it->second;
    int32_t s = (int32_t)slot_i.size();
    slot[key] = s;
    slot_i.push_back(i);
    slot_j.push_back(j);
    return s;
  };
  for (int e = 0; e < nel; e++) {
    int32_t* t = &et[(size_t)e * ET_W];
    const int(*pairs)[2] = nullptr;
    int np2 = 0;
    switch (t[ET_KIND]) {
      case K_R: case K_C: pairs = kPairsR; np2 = 4; break;
      case K_L: pairs = kPairsL; np2 = 5; break;
      case K_V: pairs = kPairsV; np2 = 4; break;
      case K_D: pairs = kPairsD; np2 = 7; break;
      case K_M: pairs = kPairsM; np2 = (int)(sizeof(kPairsM) / sizeof(kPairsM[0])); break;
      case K_Q: pairs = kPairsQ; np2 = (int)(sizeof(kPairsQ) / sizeof(kPairsQ[0])); break;
      case K_E: pairs = kPairsE; np2 = 6; break;
      case K_F: pairs = kPairsF; np2 = 2; break;
      case K_G: pairs = kPairsG; np2 = 4; break;
      case K_H: pairs = kPairsH; np2 = 5; break;
      case K_K: pairs = kPairsK; np2 = 2; break;
      default: break;
    }
    for (int k = 0; k < KH_NT * KH_NT; k++) t[ET_SLOT + k] = -1;
    for (int p = 0; p < np2; p++) {
      int ti = pairs[p][0], tj = pairs[p][1];
      int ui = t[ET_TERM + ti], uj = t[ET_TERM + tj];
      if (ui < 0 || uj < 0 || ui >= n || uj >= n) continue;
      t[ET_SLOT + ti * KH_NT + tj] = get(ui, uj);
    }
  }
  diag_slot.assign(n, -1);
  for (int i = 0; i < n; i++) diag_slot[i] = get(i, i);
  nslots = (int32_t)slot_i.size();
  return true;
}

WavePlan build_wave_plan(const Topology& T, int rec_limit) {
  WavePlan w;
  const int nel = T.nel, ns = T.nslots;
  w.rec_off.assign(nel, 0);
  w.ch_el.push_back(0);
  int cur = 0, cnt = 0;
  for (int k = 0; k < nel; k++) {
    const int cap = record_capacity(T.et[(size_t)k * ET_W + ET_KIND]);
    if (cnt == 64 || (cnt > 0 && cur + cap > rec_limit)) {
      w.ch_el.push_back(k);
      w.rec_cap = std::max(w.rec_cap, cur);
      cur = 0;
      cnt = 0;
    }
    w.rec_off[k] = cur;
    cur += cap;
    cnt++;
  }
  w.ch_el.push_back(nel);
  w.rec_cap = std::max(w.rec_cap, cur);
  const int nch = (int)w.ch_el.size() - 1;
  w.ch_t.push_back(0);
  w.tel_ptr.push_back(0);
  for (int c = 0; c < nch; c++) {
    std::map<int32_t, std::vector<int32_t>> touch;  // target -> elements, in element order
    for (int k = w.ch_el[c]; k < w.ch_el[c + 1]; k++) {
      const int32_t* e = &T.et[(size_t)k * ET_W];
      std::set<int32_t> mine;
      for (int q = 0; q < KH_NT * KH_NT; q++)
        if (e[ET_SLOT + q] >= 0) mine.insert(e[ET_SLOT + q]);
      for (int t = 0; t < KH_NT; t++)
        if (e[ET_TERM + t] >= 0 && e[ET_TERM + t] < T.n) mine.insert(ns + e[ET_TERM + t]);
      for (int32_t tg : mine) touch[tg].push_back(k);
    }
    for (auto& kv : touch) {
      w.tgt.push_back(kv.first);
      for (int32_t k : kv.second) w.tel.push_back(k);
      w.tel_ptr.push_back((int32_t)w.tel.size());
    }
    w.ch_t.push_back((int32_t)w.tgt.size());
  }
  // AbsProduct rows: slot order, and row-major (column order within a row).
  std::vector<std::vector<int32_t>> rows(T.n);
  for (int s2 = 0; s2 < ns; s2++) rows[T.slot_i[s2]].push_back(s2);
  w.ro_ptr_sparse.push_back(0);
  w.ro_ptr_dense.push_back(0);
  for (int i = 0; i < T.n; i++) {
    for (int32_t s2 : rows[i]) w.ro_slot_sparse.push_back(s2);
    w.ro_ptr_sparse.push_back((int32_t)w.ro_slot_sparse.size());
    std::vector<int32_t> r = rows[i];
    std::sort(r.begin(), r.end(), [&](int32_t a, int32_t b) { return T.slot_j[a] < T.slot_j[b]; });
    for (int32_t s2 : r) w.ro_slot_dense.push_back(s2);
    w.ro_ptr_dense.push_back((int32_t)w.ro_slot_dense.size());
  }
  return w;
}

}  // namespace kh
