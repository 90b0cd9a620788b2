// pzg_renumber.cpp -- bandwidth renumbering of the connects, on the host.
//
// TPZAnalysis::OptimizeBandwidth (Analysis/TPZAnalysis.cpp:225-300) with the
// default renumbering TPZSloanRenumbering (RENUMBER, :74-78):
//   * the node graph of TPZRenumbering::ConvertGraph (External/
//     TPZRenumbering.cpp:76-110): per connect the sorted set of the other
//     connects of its elements;
//   * pseudo-peripheral start and end nodes (TPZCutHillMcKee::SGraph::
//     PseudoPeripheralNodes, External/TPZCutHillMcKee.cpp:165-240, with
//     SmallestDegree, RootedLevelStructure, SortNodes, ShrinkLastLevel);
//   * Sloan's numbering (TPZSloanRenumbering::Resequence, External/
//     TPZSloanRenumbering.cpp:23-161): priority W1 * distance-to-end -
//     W2 * (degree + 1), W1 = 1, W2 = 2, raised by W2 as nodes become
//     active; the node of highest priority is numbered next.
// The permutation is the reference's node for node.  The reference scans
// its whole queue for the highest priority at every step (O(n |Q|)); here a
// max segment tree over the queue slots returns the same node (the first
// slot holding the highest priority) in O(log n).  The queue keeps the
// reference's slot order: a node enters at most once, at the end, and a
// removed slot takes the last entry.
//
// No GPU work: the numbering is sequential by construction.
#include <algorithm>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "pzg_internal.h"

namespace pzg {

namespace {

struct Graph {
  int64_t n = 0;
  std::vector<int64_t> xadj, adj;
  int64_t degree(int64_t v) const { return xadj[v + 1] - xadj[v]; }
};

// ConvertGraph: node -> elements, then the sorted union of their nodes.
int convert_graph(int64_t nnodes, int64_t nel, const int64_t *eptr, const int64_t *elist, Graph &g) {
  std::vector<int64_t> cnt(nnodes + 1, 0);
  for (int64_t e = 0; e < nel; e++)
    for (int64_t k = eptr[e]; k < eptr[e + 1]; k++) {
      const int64_t v = elist[k];
      if (v < 0 || v >= nnodes) {
        set_error("pzg_sloan_renumber: element graph entry out of range");
        return PZG_EINVAL;
      }
      cnt[v + 1]++;
    }
  for (int64_t v = 0; v < nnodes; v++) cnt[v + 1] += cnt[v];
  std::vector<int64_t> n2e(cnt[nnodes]);
  {
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < nel; e++)
      for (int64_t k = eptr[e]; k < eptr[e + 1]; k++) n2e[fill[elist[k]]++] = e;
  }
  g.n = nnodes;
  g.xadj.assign(nnodes + 1, 0);
  const int T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<int64_t>> part(T);
  std::vector<int64_t> deg(nnodes, 0);
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++)
    th.emplace_back([&, t] {
      const int64_t v0 = nnodes * t / T, v1 = nnodes * (t + 1) / T;
      std::vector<int64_t> buf;
      for (int64_t v = v0; v < v1; v++) {
        buf.clear();
        for (int64_t k = cnt[v]; k < cnt[v + 1]; k++) {
          const int64_t e = n2e[k];
          for (int64_t j = eptr[e]; j < eptr[e + 1]; j++)
            if (elist[j] != v) buf.push_back(elist[j]);
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        deg[v] = (int64_t)buf.size();
        part[t].insert(part[t].end(), buf.begin(), buf.end());
      }
    });
  for (auto &x : th) x.join();
  for (int64_t v = 0; v < nnodes; v++) g.xadj[v + 1] = g.xadj[v] + deg[v];
  g.adj.resize(g.xadj[nnodes]);
  for (int t = 0; t < T; t++)  // thread t's nodes start at xadj[v0]
    std::copy(part[t].begin(), part[t].end(), g.adj.begin() + g.xadj[nnodes * t / T]);
  return PZG_OK;
}

// RootedLevelStructure: breadth-first levels from root (a node joins the
// first level that reaches it, in the order of its parents and of their
// sorted adjacency).  lv: nodes level by level; lptr: level starts.
int64_t levels(const Graph &g, int64_t root, std::vector<int64_t> &lv, std::vector<int64_t> &lptr,
               std::vector<char> &seen) {
  std::fill(seen.begin(), seen.end(), 0);
  lv.resize(g.n);
  lptr.clear();
  int64_t head = 0, tail = 0;
  lv[tail++] = root;
  seen[root] = 1;
  while (head < tail) {
    const int64_t end = tail;
    lptr.push_back(head);
    for (; head < end; head++) {
      const int64_t v = lv[head];
      for (int64_t k = g.xadj[v]; k < g.xadj[v + 1]; k++) {
        const int64_t a = g.adj[k];
        if (!seen[a]) {
          seen[a] = 1;
          lv[tail++] = a;
        }
      }
    }
  }
  lptr.push_back(tail);
  return (int64_t)lptr.size() - 1;
}

void peripheral(const Graph &g, int64_t &start, int64_t &endn) {
  std::vector<int64_t> lv, lptr, lv2, lptr2, last;
  std::vector<char> seen(g.n);
  int64_t mindeg = 1000000, s = -1;  // SmallestDegree: first node of least degree
  for (int64_t i = 0; i < g.n; i++)
    if (g.degree(i) < mindeg) {
      mindeg = g.degree(i);
      s = i;
    }
  int64_t e = -1;
  int64_t nlev = levels(g, s, lv, lptr, seen);
  int count = 0;
  while (count < 10 || e == -1) {
    count++;
    const int64_t cps = s, cpe = e;
    if (!nlev) break;
    last.assign(lv.begin() + lptr[nlev - 1], lv.begin() + lptr[nlev]);
    const int64_t nlast = (int64_t)last.size();
    std::stable_sort(last.begin(), last.end(),
                     [&](int64_t a, int64_t b) { return g.degree(a) < g.degree(b); });
    // ShrinkLastLevel: the last node of each degree, at most (nlast+2)/2
    int64_t nk = 0;
    for (int64_t i = 0; i < nlast; i++) {
      if (i + 1 < nlast && g.degree(last[i + 1]) == g.degree(last[i])) continue;
      last[nk++] = last[i];
    }
    if ((nlast + 2) / 2 < nk) nk = (nlast + 2) / 2;
    int64_t we = 1000000000L;
    const int64_t hs = nlev;
    for (int64_t iq = 0; iq < nk; iq++) {
      const int64_t q = last[iq];
      const int64_t h = levels(g, q, lv2, lptr2, seen);
      int64_t w = 0;
      for (int64_t j = 0; j < h; j++) w = std::max(w, lptr2[j + 1] - lptr2[j]);
      if (h > hs && w < we) {
        s = q;
        lv.swap(lv2);
        lptr.swap(lptr2);
        nlev = h;
        break;
      } else if (w < we) {
        e = q;
        we = w;
      }
    }
    if (cps == s && cpe == e) break;
  }
  start = s;
  endn = e;
}

// Max segment tree over the queue slots; leftmost slot of the maximum.
struct SlotTree {
  int64_t size = 1;
  std::vector<int64_t> t;
  static constexpr int64_t NONE = std::numeric_limits<int64_t>::min();
  explicit SlotTree(int64_t n) {
    while (size < n) size <<= 1;
    t.assign(2 * size, NONE);
  }
  void set(int64_t slot, int64_t v) {
    int64_t i = slot + size;
    t[i] = v;
    for (i >>= 1; i >= 1; i >>= 1) t[i] = std::max(t[2 * i], t[2 * i + 1]);
  }
  int64_t leftmost_max() const {
    int64_t i = 1;
    while (i < size) i = t[2 * i] == t[i] ? 2 * i : 2 * i + 1;
    return i - size;
  }
};

}  // namespace

int sloan_renumber(int64_t nnodes, int64_t nel, const int64_t *eptr, const int64_t *elist,
                   int64_t *perm) {
  if (nnodes <= 0 || nel < 0 || !eptr || !perm || (nel > 0 && !elist)) {
    set_error("pzg_sloan_renumber: invalid arguments");
    return PZG_EINVAL;
  }
  Graph g;
  PZG_TRY(convert_graph(nnodes, nel, eptr, elist, g));
  int64_t s, e;
  peripheral(g, s, e);
  std::vector<int64_t> lv, lptr;
  std::vector<char> seen(g.n);
  const int64_t nlev = levels(g, e, lv, lptr, seen);
  if (lptr[nlev] != g.n) {
    set_error("pzg_sloan_renumber: the connect graph is not connected");
    return PZG_EINVAL;
  }
  constexpr int64_t W1 = 1, W2 = 2;
  enum : char { INACTIVE = 0, PREACTIVE = 1, ACTIVE = 2, POSTACTIVE = 3 };
  std::vector<int64_t> prio(g.n), Q(g.n), pos(g.n, -1);
  for (int64_t l = 0; l < nlev; l++)
    for (int64_t k = lptr[l]; k < lptr[l + 1]; k++) prio[lv[k]] = W1 * l;
  for (int64_t i = 0; i < g.n; i++) prio[i] -= W2 * (g.degree(i) + 1);
  std::vector<char> hist(g.n, 0), st(g.n, INACTIVE);
  SlotTree tree(g.n);
  int64_t qn = 0, nr = 0;
  auto push = [&](int64_t v) {
    if (hist[v]) return false;
    hist[v] = 1;
    Q[qn] = v;
    pos[v] = qn;
    tree.set(qn, prio[v]);
    qn++;
    return true;
  };
  auto raise = [&](int64_t v) {
    prio[v] += W2;
    if (hist[v] && st[v] != POSTACTIVE) tree.set(pos[v], prio[v]);
  };
  st[s] = PREACTIVE;
  push(s);
  while (qn) {
    const int64_t qi = tree.leftmost_max(), cur = Q[qi];
    // remove(qi): the last entry moves into slot qi
    const int64_t lastv = Q[qn - 1];
    Q[qi] = lastv;
    pos[lastv] = qi;
    tree.set(qi, prio[lastv]);
    tree.set(qn - 1, SlotTree::NONE);
    qn--;
    if (st[cur] == PREACTIVE)
      for (int64_t k = g.xadj[cur]; k < g.xadj[cur + 1]; k++) {
        const int64_t a = g.adj[k];
        raise(a);
        if (push(a) && st[a] == INACTIVE) st[a] = PREACTIVE;
      }
    perm[cur] = nr++;
    st[cur] = POSTACTIVE;
    for (int64_t k = g.xadj[cur]; k < g.xadj[cur + 1]; k++) {
      const int64_t a = g.adj[k];
      if (st[a] != PREACTIVE) continue;
      st[a] = ACTIVE;
      raise(a);
      for (int64_t j = g.xadj[a]; j < g.xadj[a + 1]; j++) {
        const int64_t b = g.adj[j];
        if (st[b] != POSTACTIVE) raise(b);
        if (push(b) && st[b] == INACTIVE) st[b] = PREACTIVE;
      }
    }
  }
  if (nr != g.n) {
    set_error("pzg_sloan_renumber: not every node was numbered");
    return PZG_EINVAL;
  }
  return PZG_OK;
}

}  // namespace pzg
