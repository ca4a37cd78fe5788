// Octree + interaction lists (see let3d.hpp).  Citations are to the reference's
// fmm3d_mpi/let3d_mpi.cpp unless stated otherwise.
#include "let3d.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>

namespace kifmm {

using u128 = unsigned __int128;

double Tree::radius(int i) const { return rad0 / double(1 << depth[i]); }

void Tree::center(int i, double* c) const {
  const int tmp = 1 << depth[i];
  for (int d = 0; d < 3; d++) c[d] = (ctr[d] - rad0) + (2 * rad0) * (path[3 * i + d] + 0.5) / double(tmp);
}

namespace {

struct Builder {
  Tree& t;
  const std::vector<u128>& cells;  // sorted distinct depth-L cell keys
  int L, pts_max;

  int push(int par, int d, int px, int py, int pz) {
    int i = int(t.parent.size());
    t.parent.push_back(par); t.child.push_back(-1); t.depth.push_back(d);
    t.path.push_back(px); t.path.push_back(py); t.path.push_back(pz);
    return i;
  }
  void split(int p) {  // 8 children, index 4a+2b+c (CreateAncestors :641-655)
    t.child[p] = int(t.parent.size());
    const int d = t.depth[p] + 1, x = t.path[3 * p], y = t.path[3 * p + 1], z = t.path[3 * p + 2];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int c = 0; c < 2; c++) push(p, d, 2 * x + a, 2 * y + b, 2 * z + c);
  }
  // Pre-order walk in Dendro Morton order; a child block is appended on entry.
  void walk(int p, u128 base, size_t lo, size_t hi) {
    const int d = t.depth[p] + 1;
    const u128 span = u128(1) << (3 * (L - d));
    for (int m = 0; m < 8; m++) {  // Morton digit (z<<2)|(y<<1)|x
      const int x = m & 1, y = (m >> 1) & 1, z = (m >> 2) & 1;
      const int c = t.child[p] + 4 * x + 2 * y + z;
      const u128 cb = base + u128(m) * span;
      size_t clo = std::lower_bound(cells.begin() + lo, cells.begin() + hi, cb) - cells.begin();
      size_t chi = std::lower_bound(cells.begin() + clo, cells.begin() + hi, cb + span) - cells.begin();
      if (d < L && chi - clo > size_t(pts_max)) {
        split(c);
        walk(c, cb, clo, chi);
      }
    }
  }
};

inline u128 interleave(uint32_t x, uint32_t y, uint32_t z, int L) {
  u128 k = 0;
  for (int b = L - 1; b >= 0; b--)
    k = (k << 3) | u128(((z >> b) & 1u) << 2 | ((y >> b) & 1u) << 1 | ((x >> b) & 1u));
  return k;
}

// Descend by x >= centre (srcData :1255-1267).
inline int locate(const Tree& t, const double* p) {
  int node = 0;
  while (t.child[node] >= 0) {
    const int tmp = 1 << t.depth[node];
    int c = t.child[node];
    for (int d = 0; d < 3; d++) {
      double ctr = (t.ctr[d] - t.rad0) + (2 * t.rad0) * (t.path[3 * node + d] + 0.5) / double(tmp);
      c += (p[d] >= ctr) << (2 - d);
    }
    node = c;
  }
  return node;
}

// findGlbNode :1833-1853
inline int find_node(const Tree& t, int want, const int* wpath) {
  int cur = 0, tp[3] = {wpath[0], wpath[1], wpath[2]};
  while (t.depth[cur] < want && t.child[cur] >= 0) {
    const int step = 1 << (want - t.depth[cur] - 1);
    int ch = t.child[cur];
    for (int d = 0; d < 3; d++) {
      int c = tp[d] / step;
      tp[d] -= c * step;
      ch += c << (2 - d);
    }
    cur = ch;
  }
  return cur;
}

// adjacent :1856-1876 (64-bit to stay exact at depth 30)
inline bool adjacent(const Tree& t, int a, int b) {
  const int da = t.depth[a], db = t.depth[b], md = std::max(da, db);
  const int64_t ra = int64_t(1) << (md - da), rb = int64_t(1) << (md - db), rs = ra + rb;
  int64_t mx = 0;
  bool within = true;
  for (int d = 0; d < 3; d++) {
    int64_t ca = (2 * int64_t(t.path[3 * a + d]) + 1) * ra;
    int64_t cb = (2 * int64_t(t.path[3 * b + d]) + 1) * rb;
    int64_t df = ca > cb ? ca - cb : cb - ca;
    within = within && df <= rs;
    mx = std::max(mx, df);
  }
  return within && mx == rs;
}

void sort_unique(std::vector<int>& v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

// calGlbNodeLists :1637-1768
void node_lists(const Tree& t, int B, std::vector<int>* out, std::vector<int>& queue) {
  for (int k = 0; k < 4; k++) out[k].clear();
  const bool bleaf = t.child[B] < 0;
  if (B != 0) {
    const int P = t.parent[B], D = t.depth[B], lim = 1 << D;
    for (int i = -2; i < 4; i++)
      for (int j = -2; j < 4; j++)
        for (int k = -2; k < 4; k++) {
          int tp[3] = {2 * t.path[3 * P] + i, 2 * t.path[3 * P + 1] + j, 2 * t.path[3 * P + 2] + k};
          bool inside = true, same = true;
          for (int d = 0; d < 3; d++) {
            inside = inside && tp[d] >= 0 && tp[d] < lim;
            same = same && tp[d] == t.path[3 * B + d];
          }
          if (!inside || same) continue;
          const int R = find_node(t, D, tp);
          const bool adj = adjacent(t, R, B);
          if (t.depth[R] < D) {
            if (adj) { if (bleaf) out[LIST_U].push_back(R); }
            else out[LIST_X].push_back(R);
          } else if (t.depth[R] == D) {
            if (!adj) out[LIST_V].push_back(R);
            else if (bleaf) {  // breadth-first over R's subtree
              queue.assign(1, R);
              for (size_t q = 0; q < queue.size(); q++) {
                const int F = queue[q];
                if (!adjacent(t, F, B)) out[LIST_W].push_back(F);
                else if (t.child[F] < 0) out[LIST_U].push_back(F);
                else for (int c = 0; c < 8; c++) queue.push_back(t.child[F] + c);
              }
            }
          }
        }
  }
  if (bleaf) out[LIST_U].push_back(B);
  for (int k = 0; k < 4; k++) sort_unique(out[k]);
  // filters :1748-1765
  auto keep = [&](std::vector<int>& v, bool by_count) {
    size_t m = 0;
    for (int a : v)
      if (by_count ? t.src_num[a] > 0 : (t.tag[a] & TAG_SRCE) != 0) v[m++] = a;
    v.resize(m);
  };
  keep(out[LIST_U], true); keep(out[LIST_V], false); keep(out[LIST_W], false); keep(out[LIST_X], true);
}

}  // namespace

int build_tree(const double* spos, int64_t ns, const double* tpos, int64_t nt, const double ctr[3],
               int root_level, int pts_max, int max_level, Tree& t, std::string* err) {
  if (max_level < 1 || max_level > 30) { if (err) *err = "maxLevel must lie in [1,30]"; return -2; }
  if (pts_max < 1) { if (err) *err = "ptsMax must be >= 1"; return -2; }
  if (ns > INT32_MAX || nt > INT32_MAX) { if (err) *err = "too many points for 32-bit indices"; return -2; }
  t = Tree();
  t.max_level = max_level; t.root_level = root_level;
  t.rad0 = std::pow(2.0, -root_level);
  for (int d = 0; d < 3; d++) t.ctr[d] = ctr[d];
  const int L = max_level;
  const double lo[3] = {ctr[0] - t.rad0, ctr[1] - t.rad0, ctr[2] - t.rad0};
  for (int64_t i = 0; i < ns; i++)
    for (int d = 0; d < 3; d++) {
      double v = spos[3 * i + d];
      if (!(v >= lo[d] && v <= lo[d] + 2 * t.rad0)) {
        if (err) *err = "source point " + std::to_string(i) + " lies outside the root box";
        return -2;
      }
    }
  for (int64_t i = 0; i < nt; i++)
    for (int d = 0; d < 3; d++) {
      double v = tpos[3 * i + d];
      if (!(v >= lo[d] && v <= lo[d] + 2 * t.rad0)) {
        if (err) *err = "target point " + std::to_string(i) + " lies outside the root box";
        return -2;
      }
    }

  // distinct depth-L cells of the sources
  std::vector<u128> cells(ns);
  const double scale = std::ldexp(1.0, L - 1 + root_level);
  const uint32_t cmax = (1u << L) - 1u;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ns; i++) {
    uint32_t c[3];
    for (int d = 0; d < 3; d++) {
      double v = (spos[3 * i + d] - lo[d]) * scale;
      c[d] = v <= 0.0 ? 0u : (v >= double(cmax) ? cmax : uint32_t(v));
    }
    cells[i] = interleave(c[0], c[1], c[2], L);
  }
  std::sort(cells.begin(), cells.end());
  cells.erase(std::unique(cells.begin(), cells.end()), cells.end());

  Builder b{t, cells, L, pts_max};
  b.push(-1, 0, 0, 0, 0);
  b.split(0);
  b.walk(0, u128(0), 0, cells.size());
  const int N = int(t.parent.size());
  t.n_nodes = N;
  for (int i = 0; i < N; i++) t.max_depth = std::max(t.max_depth, t.depth[i]);
  t.tag.assign(N, 0);

  // sources (srcData :1209-1465 on one rank)
  std::vector<int> leaf_s(ns), leaf_t(nt);
  std::vector<std::atomic<int>> cnt(N);
  for (auto& c : cnt) c.store(0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ns; i++) { leaf_s[i] = locate(t, spos + 3 * i); cnt[leaf_s[i]].fetch_add(1); }
  t.src_cnt.assign(N, 0);
  for (int i = 0; i < N; i++) t.src_cnt[i] = cnt[i].load();
  for (int i = N - 1; i >= 1; i--) t.src_cnt[t.parent[i]] += t.src_cnt[i];  // children follow parents
  t.src_beg.assign(N, -1); t.src_num.assign(N, -1);
  int off = 0;
  for (int i = 1; i < N; i++) {
    if (!t.leaf(i) || t.src_cnt[i] > 0) {
      t.tag[i] |= TAG_SRCE | TAG_OWNR;
      if (t.leaf(i)) { t.src_beg[i] = off; t.src_num[i] = t.src_cnt[i]; off += t.src_cnt[i]; }
    }
    if (t.src_cnt[i] > 0) t.tag[i] |= TAG_CBTR;
  }
  t.src_perm.assign(ns, 0);
  {
    std::vector<int> fill(N, 0);
    for (int64_t i = 0; i < ns; i++) { int l = leaf_s[i]; t.src_perm[t.src_beg[l] + fill[l]++] = int(i); }
  }
  // targets (trgData :1470-1631)
  for (auto& c : cnt) c.store(0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nt; i++) { leaf_t[i] = locate(t, tpos + 3 * i); cnt[leaf_t[i]].fetch_add(1); }
  t.trg_cnt.assign(N, 0);
  for (int i = 0; i < N; i++) t.trg_cnt[i] = cnt[i].load();
  for (int i = N - 1; i >= 1; i--) t.trg_cnt[t.parent[i]] += t.trg_cnt[i];
  t.trg_beg.assign(N, -1); t.trg_num.assign(N, -1);
  off = 0;
  for (int i = 0; i < N; i++)
    if (t.trg_cnt[i] > 0) {
      t.tag[i] |= TAG_EVTR;
      if (t.leaf(i)) { t.trg_beg[i] = off; t.trg_num[i] = t.trg_cnt[i]; off += t.trg_cnt[i]; }
    }
  t.trg_perm.assign(nt, 0);
  {
    std::vector<int> fill(N, 0);
    for (int64_t i = 0; i < nt; i++) { int l = leaf_t[i]; t.trg_perm[t.trg_beg[l] + fill[l]++] = int(i); }
  }

  // lists for every node that holds targets
  std::vector<std::vector<int>> per[4];
  for (int k = 0; k < 4; k++) per[k].resize(N);
#pragma omp parallel
  {
    std::vector<int> out[4], queue;
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < N; i++) {
      if (t.trg_cnt[i] <= 0) continue;
      node_lists(t, i, out, queue);
      for (int k = 0; k < 4; k++) per[k][i] = out[k];
    }
  }
  for (int k = 0; k < 4; k++) {
    t.off[k].assign(N + 1, 0);
    for (int i = 0; i < N; i++) t.off[k][i + 1] = t.off[k][i] + int64_t(per[k][i].size());
    t.lst[k].resize(t.off[k][N]);
    for (int i = 0; i < N; i++) {
      std::copy(per[k][i].begin(), per[k][i].end(), t.lst[k].begin() + t.off[k][i]);
      for (int a : per[k][i]) t.tag[a] |= TAG_USER;
    }
  }
  return 0;
}

void merge_trees(std::vector<Tree>& parts, const std::vector<int64_t>& src_off, const std::vector<int64_t>& trg_off,
                 Tree& out) {
  out = Tree();
  if (parts.empty()) return;
  out.max_level = parts[0].max_level;
  out.root_level = parts[0].root_level;
  out.rad0 = parts[0].rad0;
  for (int d = 0; d < 3; d++) out.ctr[d] = parts[0].ctr[d];
  int64_t nn = 0, nl[4] = {0, 0, 0, 0};
  for (const Tree& t : parts) {
    nn += t.n_nodes;
    for (int k = 0; k < 4; k++) nl[k] += int64_t(t.lst[k].size());
  }
  out.n_nodes = int(nn);
  for (int k = 0; k < 4; k++) {
    out.off[k].reserve(size_t(nn) + 1);
    out.off[k].push_back(0);
    out.lst[k].reserve(size_t(nl[k]));
  }
  int node0 = 0;
  for (size_t p = 0; p < parts.size(); p++) {
    Tree& t = parts[p];
    const int so = int(src_off[p]), to = int(trg_off[p]);
    auto shift = [](std::vector<int>& dst, const std::vector<int>& src, int by) {
      for (int v : src) dst.push_back(v >= 0 ? v + by : v);
    };
    shift(out.parent, t.parent, node0);
    shift(out.child, t.child, node0);
    out.depth.insert(out.depth.end(), t.depth.begin(), t.depth.end());
    out.path.insert(out.path.end(), t.path.begin(), t.path.end());
    out.tag.insert(out.tag.end(), t.tag.begin(), t.tag.end());
    out.src_cnt.insert(out.src_cnt.end(), t.src_cnt.begin(), t.src_cnt.end());
    out.trg_cnt.insert(out.trg_cnt.end(), t.trg_cnt.begin(), t.trg_cnt.end());
    shift(out.src_beg, t.src_beg, so);
    shift(out.trg_beg, t.trg_beg, to);
    out.src_num.insert(out.src_num.end(), t.src_num.begin(), t.src_num.end());
    out.trg_num.insert(out.trg_num.end(), t.trg_num.begin(), t.trg_num.end());
    shift(out.src_perm, t.src_perm, so);
    shift(out.trg_perm, t.trg_perm, to);
    for (int k = 0; k < 4; k++) {
      const int64_t base = out.off[k].back();
      for (int i = 0; i < t.n_nodes; i++) out.off[k].push_back(base + t.off[k][size_t(i) + 1]);
      shift(out.lst[k], t.lst[k], node0);
    }
    out.max_depth = std::max(out.max_depth, t.max_depth);
    node0 += t.n_nodes;
    t = Tree();  // release the part
  }
}

}  // namespace kifmm
