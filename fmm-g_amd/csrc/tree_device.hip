// Octree, tags, packing and U/V/W/X lists built on the GPU: the same Tree as
// build_tree (one problem) or merge_trees over per-problem build_tree calls (a
// forest), node for node and list entry for list entry.  Citations are to the
// reference's fmm3d_mpi/let3d_mpi.cpp.
//
// The leaf rule (a node splits iff it holds more than ptsMax distinct
// depth-maxLevel cells and lies above maxLevel; the root always splits) and the
// CreateAncestors numbering (:611-726: child blocks appended in a pre-order walk
// that visits children in Morton order) are evaluated data-parallel over the
// sorted distinct cells:
//
//   * cells are keys (problem << 3L) | morton, morton with z as the most
//     significant bit of each octal digit; sorted and made unique;
//   * lcp(a, b) = number of leading octal digits two cells of one problem share
//     (0 across problems).  The node of depth d that starts at cell i splits iff
//     i starts a depth-d group (lcp(i-1, i) < d) and the group holds more than
//     ptsMax cells (lcp(i, i+ptsMax) >= d), 1 <= d <= L-1; so cell i starts the
//     split nodes of depths lcp(i-1, i)+1 .. min(L-1, lcp(i, i+ptsMax));
//   * split nodes in (cell, depth) order are the pre-order walk, so with every
//     problem's root counted first in its problem, the global rank of split
//     node (i, d) of problem p is g = p + 1 + base[i] + (d - lcp(i-1, i) - 1)
//     (base: exclusive scan of the per-cell counts), the root of problem p has
//     rank p + base[first cell of p], the child block of split node g starts at
//     node p + 1 + 8 g and the root of problem p is node p + 8 g_root;
//   * the descendants of split node g are the nodes [p+1+8g, p+1+8 g_end),
//     g_end the rank of the first split node past its cell range, so subtree
//     point counts are differences of one prefix sum over the leaves.
//
// Points descend by x >= centre with the host's FP64 centre expression (this
// file is compiled without FP contraction), leaves are packed in node order by
// a stable radix sort of (leaf, caller index), and the lists are emitted per
// target node (one wave per node, one lane per probe of calGlbNodeLists
// :1637-1768, a stackless walk for the breadth-first W/U search), filtered, then
// sorted and made unique as (kind, node, entry) keys.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_binary_search.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "tree_device.hpp"

namespace kifmm {

namespace {

using u128 = unsigned __int128;
using u64 = unsigned long long;

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

struct TreeConst {
  double lo[3];     // root box corner ctr - rad0
  double hi[3];     // lo + 2 rad0
  double scale;     // 2^(L - 1 + rootLevel): depth-L cell of a coordinate
  double ctr[3], rad0;
  int L;            // maxLevel
  int pts_max;
  int nprob;
};

__device__ inline int clz128(u128 x) {
  const u64 hi = u64(x >> 64), lo = u64(x);
  return hi ? __builtin_clzll(hi) : 64 + (lo ? __builtin_clzll(lo) : 64);
}

// leading octal digits shared by two cell keys; 0 when their problems differ
__device__ inline int lcp_oct(u128 a, u128 b, int L) {
  if ((a >> (3 * L)) != (b >> (3 * L))) return 0;
  const u128 x = a ^ b;
  if (x == 0) return L;
  return (clz128(x) - (128 - 3 * L)) / 3;
}

__device__ inline u128 interleave(uint32_t x, uint32_t y, uint32_t z, int L) {
  u128 k = 0;
  for (int b = L - 1; b >= 0; b--)
    k = (k << 3) | u128(((z >> b) & 1u) << 2 | ((y >> b) & 1u) << 1 | ((x >> b) & 1u));
  return k;
}

template <typename T>
__device__ inline int64_t lower_bound_dev(const T* a, int64_t n, T v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// problem of a point: the last p with off[p] <= i (off: nprob + 1 offsets)
__device__ inline int problem_of(const int64_t* off, int nprob, int64_t i) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Points outside the root box: the smallest offending index (host error text).
__global__ void k_bounds(const double* pos, int64_t n, TreeConst c, unsigned long long* first_bad) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool ok = true;
  for (int d = 0; d < 3; d++) {
    const double v = pos[3 * i + d];
    ok = ok && (v >= c.lo[d] && v <= c.hi[d]);
  }
  if (!ok) atomicMin(first_bad, (unsigned long long)i);
}

// Depth-L cell key of every source (build_tree's cell loop).
__global__ void k_cells(const double* spos, int64_t n, const int64_t* soff, TreeConst c, u128* keys) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t cmax = (1u << c.L) - 1u;
  uint32_t q[3];
  for (int d = 0; d < 3; d++) {
    const double v = (spos[3 * i + d] - c.lo[d]) * c.scale;
    q[d] = v <= 0.0 ? 0u : (v >= double(cmax) ? cmax : uint32_t(v));
  }
  const int p = c.nprob > 1 ? problem_of(soff, c.nprob, i) : 0;
  keys[i] = (u128(p) << (3 * c.L)) | interleave(q[0], q[1], q[2], c.L);
}

// Per distinct cell: lcp with its predecessor and the number of split nodes it starts.
__global__ void k_split_count(const u128* cells, int64_t m, TreeConst c, unsigned char* lcp_prev, int* cnt) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int lp = i == 0 ? 0 : lcp_oct(cells[i - 1], cells[i], c.L);
  const int ln = i + c.pts_max < m ? lcp_oct(cells[i], cells[i + c.pts_max], c.L) : 0;
  lcp_prev[i] = (unsigned char)lp;
  cnt[i] = max(0, min(c.L - 1, ln) - lp);
}

// Split node descriptors by global rank: the cell it starts at (its key prefix
// is that cell's key truncated at its depth), its depth, its problem.
struct SplitDesc {
  u128 pkey;  // (problem << 3L) | morton prefix, low 3(L-d) bits zero
  int depth, prob;
};

__global__ void k_split_roots(const u128* cells, int64_t m, const int64_t* base, TreeConst c, SplitDesc* desc,
                              int* groot) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= c.nprob) return;
  const u128 key = u128(p) << (3 * c.L);
  const int64_t f = lower_bound_dev(cells, m, key);
  const int g = int(p + base[f]);
  groot[p] = g;
  desc[g].pkey = key;
  desc[g].depth = 0;
  desc[g].prob = p;
}

__global__ void k_split_expand(const u128* cells, int64_t m, const unsigned char* lcp_prev, const int* cnt,
                               const int64_t* base, TreeConst c, SplitDesc* desc) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m || cnt[i] == 0) return;
  const u128 key = cells[i];
  const int p = int(key >> (3 * c.L));
  for (int j = 0; j < cnt[i]; j++) {
    const int d = lcp_prev[i] + 1 + j;
    const int g = int(p + 1 + base[i] + j);
    const u128 low = (u128(1) << (3 * (c.L - d))) - 1;
    desc[g].pkey = key & ~low;
    desc[g].depth = d;
    desc[g].prob = p;
  }
}

// global rank of the split node that starts at cell j with depth d (d >= 1)
__device__ inline int rank_at(int p, int64_t j, int d, const int64_t* base, const unsigned char* lcp_prev) {
  return int(p + 1 + base[j] + (d - lcp_prev[j] - 1));
}

// Node records of every split node and of its 8 children (CreateAncestors:
// child block at p + 1 + 8g, child 4x+2y+z at path 2 path + (x, y, z)).
__global__ void k_split_nodes(const SplitDesc* desc, int G, const u128* cells, int64_t m, const int64_t* base,
                              const unsigned char* lcp_prev, const int* groot, TreeConst c, int* snode,
                              int* parent, int* child, int* depth, int* path, int* proot) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const SplitDesc s = desc[g];
  const int L = c.L, d = s.depth, p = s.prob;
  const int rootnode = p + 8 * groot[p];
  const u128 pre = (s.pkey >> (3 * (L - d))) & ((u128(1) << (3 * d)) - 1);  // morton prefix, d digits
  int px = 0, py = 0, pz = 0;
  for (int b = d - 1; b >= 0; b--) {
    const int dig = int(pre >> (3 * b)) & 7;
    px = (px << 1) | (dig & 1);
    py = (py << 1) | ((dig >> 1) & 1);
    pz = (pz << 1) | ((dig >> 2) & 1);
  }
  int node;
  if (d == 0) {
    node = rootnode;
    parent[node] = -1;
    depth[node] = 0;
    path[3 * node] = 0; path[3 * node + 1] = 0; path[3 * node + 2] = 0;
    proot[node] = rootnode;
  } else {
    // the parent: the split node of depth d-1 whose range starts at the first
    // cell with this node's depth-(d-1) prefix
    const u128 plow = (u128(1) << (3 * (L - d + 1))) - 1;
    const int64_t j = lower_bound_dev(cells, m, s.pkey & ~plow);
    const int gp = d - 1 == 0 ? groot[p] : rank_at(p, j, d - 1, base, lcp_prev);
    const int dig = int(pre) & 7;
    node = p + 1 + 8 * gp + 4 * (dig & 1) + 2 * ((dig >> 1) & 1) + ((dig >> 2) & 1);
  }
  snode[g] = node;
  const int cb = p + 1 + 8 * g;
  for (int x = 0; x < 2; x++)
    for (int y = 0; y < 2; y++)
      for (int z = 0; z < 2; z++) {
        const int k = cb + 4 * x + 2 * y + z;
        parent[k] = node;
        child[k] = -1;
        depth[k] = d + 1;
        path[3 * k] = 2 * px + x; path[3 * k + 1] = 2 * py + y; path[3 * k + 2] = 2 * pz + z;
        proot[k] = rootnode;
      }
}

__global__ void k_split_child(const SplitDesc* desc, int G, const int* snode, int* child) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  child[snode[g]] = desc[g].prob + 1 + 8 * g;
}

// Point -> leaf by x >= centre (srcData :1255-1267), from the point's problem root.
__global__ void k_locate(const double* pos, int64_t n, const int64_t* off, const int* groot, const int* child,
                         const int* depth, const int* path, TreeConst c, int* leaf_of, int* leaf_cnt) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = c.nprob > 1 ? problem_of(off, c.nprob, i) : 0;
  int node = p + 8 * groot[p];
  const double x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  while (child[node] >= 0) {
    const int tmp = 1 << depth[node];
    int ch = child[node];
    for (int d = 0; d < 3; d++) {
      const double ctr = (c.ctr[d] - c.rad0) + (2 * c.rad0) * (path[3 * node + d] + 0.5) / double(tmp);
      ch += (x[d] >= ctr) << (2 - d);
    }
    node = ch;
  }
  leaf_of[i] = node;
  atomicAdd(&leaf_cnt[node], 1);
}

// First packed slot of every occupied leaf (the stable sort keeps the caller's
// order inside a leaf); leaves hold their count, other nodes -1.
__global__ void k_leaf_ranges(const int* sorted_leaf, int64_t n, int* beg) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  if (q == 0 || sorted_leaf[q] != sorted_leaf[q - 1]) beg[sorted_leaf[q]] = int(q);
}

// Subtree point counts: leaves their own; split node g the sum over the node
// range of its descendants [p+1+8g, p+1+8 g_end).
__global__ void k_subtree_counts(const SplitDesc* desc, int G, const u128* cells, int64_t m, const int64_t* base,
                                 const int* snode, TreeConst c, const int64_t* spre, const int64_t* tpre,
                                 int* scnt, int* tcnt) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const SplitDesc s = desc[g];
  const u128 span = u128(1) << (3 * (c.L - s.depth));
  const int64_t e = lower_bound_dev(cells, m, s.pkey + span);
  const int gend = int(s.prob + 1 + base[e]);
  const int64_t a = s.prob + 1 + 8 * int64_t(g), b = s.prob + 1 + 8 * int64_t(gend);
  scnt[snode[g]] = int(spre[b] - spre[a]);
  tcnt[snode[g]] = int(tpre[b] - tpre[a]);
}

// Tags (srcData :1301-1326, trgData :1547-1558) and the packed ranges of the
// leaves: src_beg/src_num only for occupied leaves, trg_beg/trg_num likewise.
__global__ void k_tags(int N, const int* parent, const int* child, const int* scnt, const int* tcnt, int* sbeg,
                       int* snum, int* tbeg, int* tnum, int* tag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const bool leaf = child[k] < 0, root = parent[k] < 0;
  int t = 0;
  if (!root) {
    if (!leaf || scnt[k] > 0) t |= TAG_SRCE | TAG_OWNR;
    if (scnt[k] > 0) t |= TAG_CBTR;
  }
  if (tcnt[k] > 0) t |= TAG_EVTR;
  tag[k] = t;
  if (!(leaf && scnt[k] > 0)) { sbeg[k] = -1; snum[k] = -1; } else snum[k] = scnt[k];
  if (!(leaf && tcnt[k] > 0)) { tbeg[k] = -1; tnum[k] = -1; } else tnum[k] = tcnt[k];
}

struct ListCtx {
  const int* parent; const int* child; const int* depth; const int* path; const int* proot;
  const int* tcnt; const int* snum; const int* tag;
  int N, bits;
};

// findGlbNode :1833-1853, from the problem root
__device__ inline int find_node(const ListCtx& t, int root, int want, int tx, int ty, int tz) {
  int cur = root, tp[3] = {tx, ty, tz};
  while (t.depth[cur] < want && t.child[cur] >= 0) {
    const int sh = want - t.depth[cur] - 1;
    int ch = t.child[cur];
    for (int d = 0; d < 3; d++) {
      const int cc = tp[d] >> sh;
      tp[d] -= cc << sh;
      ch += cc << (2 - d);
    }
    cur = ch;
  }
  return cur;
}

// adjacent :1856-1876 (64-bit, exact at depth 30)
__device__ inline bool adjacent(const ListCtx& t, int a, int b) {
  const int da = t.depth[a], db = t.depth[b], md = max(da, db);
  const int64_t ra = int64_t(1) << (md - da), rb = int64_t(1) << (md - db), rs = ra + rb;
  int64_t mx = 0;
  bool within = true;
  for (int d = 0; d < 3; d++) {
    const int64_t ca = (2 * int64_t(t.path[3 * a + d]) + 1) * ra;
    const int64_t cb = (2 * int64_t(t.path[3 * b + d]) + 1) * rb;
    const int64_t df = ca > cb ? ca - cb : cb - ca;
    within = within && df <= rs;
    mx = max(mx, df);
  }
  return within && mx == rs;
}

// filters :1748-1765: U and X need sources (srcNum > 0), V and W the source tag
__device__ inline bool keep(const ListCtx& t, int kind, int r) {
  return (kind == LIST_U || kind == LIST_X) ? t.snum[r] > 0 : (t.tag[r] & TAG_SRCE) != 0;
}

struct Emit {
  u64* out;     // null: count only
  int64_t pos;  // next slot
  int64_t n;    // entries so far
  __device__ inline void put(const ListCtx& t, int kind, int b, int r) {
    if (!keep(t, kind, r)) return;
    if (out) out[pos++] = (u64(kind) << (2 * t.bits)) | (u64(b) << t.bits) | u64(r);
    n++;
  }
};

// The entries one probe of node B contributes (calGlbNodeLists :1650-1746).
__device__ void probe(const ListCtx& t, int B, int q, Emit& em) {
  const int P = t.parent[B], D = t.depth[B], lim = 1 << D;
  const bool bleaf = t.child[B] < 0;
  const int i = q / 36 - 2, j = (q / 6) % 6 - 2, k = q % 6 - 2;
  const int tp[3] = {2 * t.path[3 * P] + i, 2 * t.path[3 * P + 1] + j, 2 * t.path[3 * P + 2] + k};
  bool inside = true, same = true;
  for (int d = 0; d < 3; d++) {
    inside = inside && tp[d] >= 0 && tp[d] < lim;
    same = same && tp[d] == t.path[3 * B + d];
  }
  if (!inside || same) return;
  const int R = find_node(t, t.proot[B], D, tp[0], tp[1], tp[2]);
  const bool adj = adjacent(t, R, B);
  if (t.depth[R] < D) {
    if (adj) { if (bleaf) em.put(t, LIST_U, B, R); }
    else em.put(t, LIST_X, B, R);
    return;
  }
  if (!adj) { em.put(t, LIST_V, B, R); return; }
  if (!bleaf) return;
  // R's subtree (the reference's breadth-first queue), walked depth first
  // without a stack: children are contiguous blocks, so the next sibling is
  // F + 1 while F is not the block's last
  int F = R;
  for (;;) {
    bool down = false;
    if (!adjacent(t, F, B)) em.put(t, LIST_W, B, F);
    else if (t.child[F] < 0) em.put(t, LIST_U, B, F);
    else down = true;
    if (down) { F = t.child[F]; continue; }
    bool done = true;
    while (F != R) {
      const int par = t.parent[F];
      if (F - t.child[par] < 7) { F++; done = false; break; }
      F = par;
    }
    if (done) break;
  }
}

// One wave per node, one lane per probe (216 probes in 4 rounds) plus B itself
// for a leaf.  Count pass (out == null): per-node entry counts; fill pass: the
// node's entries at its exclusive-scan offset.
__global__ __launch_bounds__(256) void k_lists(ListCtx t, int64_t* cnt, const int64_t* off, u64* out) {
  const int lane = threadIdx.x & 63;
  const int B = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (B >= t.N) return;  // whole wave exits together
  if (t.tcnt[B] <= 0) {
    if (!out && lane == 0) cnt[B] = 0;
    return;
  }
  const bool root = t.parent[B] < 0;
  Emit em{nullptr, 0, 0};
  if (!root)
    for (int q = lane; q < 216; q += 64) probe(t, B, q, em);
  if (lane == 0 && t.child[B] < 0) em.put(t, LIST_U, B, B);
  // wave prefix sum of the lane counts
  int64_t incl = em.n;
  for (int s = 1; s < 64; s <<= 1) {
    const int64_t v = __shfl_up(incl, s, 64);
    if (lane >= s) incl += v;
  }
  if (!out) {
    if (lane == 63) cnt[B] = incl;
    return;
  }
  Emit fe{out, off[B] + incl - em.n, 0};
  if (!root)
    for (int q = lane; q < 216; q += 64) probe(t, B, q, fe);
  if (lane == 0 && t.child[B] < 0) fe.put(t, LIST_U, B, B);
}

// CSR offsets of every kind from the sorted unique keys: off[k][b] = first
// entry of (k, b) minus the first entry of kind k.
__global__ void k_list_offsets(const u64* keys, int64_t m, int N, int bits, int64_t* off) {
  const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= 4 * int64_t(N + 1)) return;
  const int k = int(idx / (N + 1)), b = int(idx % (N + 1));
  const u64 kb = u64(k) << (2 * bits);
  off[idx] = lower_bound_dev(keys, m, kb | (u64(b) << bits)) - lower_bound_dev(keys, m, kb);
}

__global__ void k_list_entries(const u64* keys, int64_t m, int bits, int* lst, int* tag) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const int r = int(keys[q] & ((u64(1) << bits) - 1));
  lst[q] = r;
  atomicOr(&tag[r], int(TAG_USER));
}

__global__ void k_iota(int* v, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = int(i);
}

__global__ void k_widen(const int* in, int64_t* out, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

inline unsigned blocks(int64_t n, int b = 256) { return unsigned((n + b - 1) / b); }

// Device allocations of one build, released together.
struct Pool {
  std::vector<void*> ptrs;
  ~Pool() { for (void* p : ptrs) (void)hipFree(p); }
  template <typename T> hipError_t get(T** p, int64_t n) {
    *p = nullptr;
    const size_t bytes = size_t(std::max<int64_t>(n, 1)) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};

#define TRY(x)                                                       \
  do {                                                               \
    hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) {                                          \
      if (err) *err = std::string(#x) + ": " + hipGetErrorString(e_); \
      return -5;                                                     \
    }                                                                \
  } while (0)

template <typename T> hipError_t download(std::vector<T>& v, const T* d, int64_t n, hipStream_t s) {
  v.resize(size_t(n));
  if (n == 0) return hipSuccess;
  hipError_t e = hipMemcpyAsync(v.data(), d, size_t(n) * sizeof(T), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

int nbits(int64_t v) {  // smallest b with v < 2^b
  int b = 0;
  while ((int64_t(1) << b) <= v) b++;
  return b;
}

}  // namespace

int build_tree_device(int nprob, const int64_t* ns, const double* spos, const int64_t* nt, const double* tpos,
                      const double ctr[3], int root_level, int pts_max, int max_level, Tree& t, std::string* err,
                      hipStream_t s, double* ms) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  if (max_level < 1 || max_level > 30) { if (err) *err = "maxLevel must lie in [1,30]"; return -2; }
  if (pts_max < 1) { if (err) *err = "ptsMax must be >= 1"; return -2; }
  if (nprob < 1) { if (err) *err = "no problems"; return -2; }
  std::vector<int64_t> soff(nprob + 1, 0), toff(nprob + 1, 0);
  for (int p = 0; p < nprob; p++) {
    if (ns[p] < 0 || nt[p] < 0) { if (err) *err = "negative point count"; return -2; }
    soff[p + 1] = soff[p] + ns[p];
    toff[p + 1] = toff[p] + nt[p];
  }
  const int64_t NS = soff[nprob], NT = toff[nprob];
  for (int p = 0; p < nprob; p++)
    if (ns[p] > INT32_MAX || nt[p] > INT32_MAX) { if (err) *err = "too many points for 32-bit indices"; return -2; }
  if (NS > INT32_MAX || NT > INT32_MAX) { if (err) *err = "too many points for 32-bit indices"; return -2; }
  const int L = max_level;
  if (3 * L + nbits(nprob) > 127) { if (err) *err = "too many problems for the cell keys"; return -2; }

  t = Tree();
  t.max_level = max_level;
  t.root_level = root_level;
  t.rad0 = std::pow(2.0, -root_level);
  for (int d = 0; d < 3; d++) t.ctr[d] = ctr[d];
  TreeConst c{};
  for (int d = 0; d < 3; d++) {
    c.lo[d] = ctr[d] - t.rad0;
    c.hi[d] = c.lo[d] + 2 * t.rad0;
    c.ctr[d] = ctr[d];
  }
  c.rad0 = t.rad0;
  c.scale = std::ldexp(1.0, L - 1 + root_level);
  c.L = L;
  c.pts_max = pts_max;
  c.nprob = nprob;

  Pool pool;
  double *d_spos, *d_tpos;
  int64_t *d_soff, *d_toff;
  unsigned long long* d_bad;
  TRY(pool.get(&d_spos, 3 * NS));
  TRY(pool.get(&d_tpos, 3 * NT));
  TRY(pool.get(&d_soff, nprob + 1));
  TRY(pool.get(&d_toff, nprob + 1));
  TRY(pool.get(&d_bad, 2));
  if (NS) TRY(hipMemcpyAsync(d_spos, spos, size_t(3 * NS) * 8, hipMemcpyHostToDevice, s));
  if (NT) TRY(hipMemcpyAsync(d_tpos, tpos, size_t(3 * NT) * 8, hipMemcpyHostToDevice, s));
  TRY(hipMemcpyAsync(d_soff, soff.data(), size_t(nprob + 1) * 8, hipMemcpyHostToDevice, s));
  TRY(hipMemcpyAsync(d_toff, toff.data(), size_t(nprob + 1) * 8, hipMemcpyHostToDevice, s));
  const unsigned long long none[2] = {~0ull, ~0ull};
  TRY(hipMemcpyAsync(d_bad, none, sizeof(none), hipMemcpyHostToDevice, s));
  if (NS) k_bounds<<<blocks(NS), 256, 0, s>>>(d_spos, NS, c, d_bad);
  if (NT) k_bounds<<<blocks(NT), 256, 0, s>>>(d_tpos, NT, c, d_bad + 1);
  TRY(hipGetLastError());
  unsigned long long bad[2];
  TRY(hipMemcpyAsync(bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  // the host builder's report: the first problem with a bad point, its sources
  // checked before its targets
  int bw = -1, bp = nprob;
  int64_t bi = 0;
  for (int w = 0; w < 2; w++) {
    if (bad[w] == ~0ull) continue;
    const std::vector<int64_t>& off = w == 0 ? soff : toff;
    const int64_t gi = int64_t(bad[w]);
    const int p = int(std::upper_bound(off.begin(), off.end(), gi) - off.begin()) - 1;
    if (p < bp) { bp = p; bw = w; bi = gi - off[p]; }
  }
  if (bw >= 0) {
    std::string msg = std::string(bw == 0 ? "source" : "target") + " point " + std::to_string(bi) +
                      " lies outside the root box";
    if (nprob > 1) msg = "problem " + std::to_string(bp) + ": " + msg;
    if (err) *err = msg;
    return -2;
  }

  // ---- distinct cells, sorted ----
  u128 *d_keys, *d_cells;
  int64_t* d_m;
  TRY(pool.get(&d_keys, NS));
  TRY(pool.get(&d_cells, NS));
  TRY(pool.get(&d_m, 1));
  const unsigned key_bits = unsigned(3 * L + nbits(nprob - 1));
  int64_t M = 0;
  if (NS) {
    k_cells<<<blocks(NS), 256, 0, s>>>(d_spos, NS, d_soff, c, d_keys);
    TRY(hipGetLastError());
    size_t tb = 0, tb2 = 0;
    TRY(rocprim::radix_sort_keys(nullptr, tb, d_keys, d_cells, size_t(NS), 0, key_bits, s));
    TRY(rocprim::unique(nullptr, tb2, d_cells, d_keys, d_m, size_t(NS), rocprim::equal_to<u128>(), s));
    void* tmp;
    TRY(pool.get(reinterpret_cast<char**>(&tmp), int64_t(std::max(tb, tb2))));
    tb = std::max(tb, tb2);
    TRY(rocprim::radix_sort_keys(tmp, tb, d_keys, d_cells, size_t(NS), 0, key_bits, s));
    TRY(rocprim::unique(tmp, tb, d_cells, d_keys, d_m, size_t(NS), rocprim::equal_to<u128>(), s));
    TRY(hipMemcpyAsync(&M, d_m, 8, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    std::swap(d_keys, d_cells);  // d_cells: the M distinct cells
  }

  // ---- split nodes ----
  unsigned char* d_lcp;
  int* d_cnt;
  int64_t* d_base;  // split nodes before each cell: 64-bit, checked before the node arrays
  TRY(pool.get(&d_lcp, M));
  TRY(pool.get(&d_cnt, M + 1));
  TRY(pool.get(&d_base, M + 1));
  TRY(hipMemsetAsync(d_cnt, 0, size_t(M + 1) * 4, s));
  if (M) k_split_count<<<blocks(M), 256, 0, s>>>(d_cells, M, c, d_lcp, d_cnt);
  TRY(hipGetLastError());
  {
    size_t tb = 0;
    TRY(rocprim::exclusive_scan(nullptr, tb, d_cnt, d_base, int64_t(0), size_t(M + 1), rocprim::plus<int64_t>(), s));
    char* tmp;
    TRY(pool.get(&tmp, int64_t(tb)));
    TRY(rocprim::exclusive_scan(tmp, tb, d_cnt, d_base, int64_t(0), size_t(M + 1), rocprim::plus<int64_t>(), s));
  }
  int64_t S = 0;
  TRY(hipMemcpyAsync(&S, d_base + M, 8, hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  const int64_t G = int64_t(S) + nprob, N64 = nprob + 8 * G;
  if (N64 > INT32_MAX / 2) { if (err) *err = "tree too large for 32-bit node indices"; return -2; }
  const int N = int(N64);
  SplitDesc* d_desc;
  int *d_groot, *d_snode, *d_parent, *d_child, *d_depth, *d_path, *d_proot;
  TRY(pool.get(&d_desc, G));
  TRY(pool.get(&d_groot, nprob));
  TRY(pool.get(&d_snode, G));
  TRY(pool.get(&d_parent, N));
  TRY(pool.get(&d_child, N));
  TRY(pool.get(&d_depth, N));
  TRY(pool.get(&d_path, 3 * int64_t(N)));
  TRY(pool.get(&d_proot, N));
  k_split_roots<<<blocks(nprob), 256, 0, s>>>(d_cells, M, d_base, c, d_desc, d_groot);
  if (M) k_split_expand<<<blocks(M), 256, 0, s>>>(d_cells, M, d_lcp, d_cnt, d_base, c, d_desc);
  k_split_nodes<<<blocks(G), 256, 0, s>>>(d_desc, int(G), d_cells, M, d_base, d_lcp, d_groot, c, d_snode, d_parent,
                                          d_child, d_depth, d_path, d_proot);
  k_split_child<<<blocks(G), 256, 0, s>>>(d_desc, int(G), d_snode, d_child);
  TRY(hipGetLastError());

  // ---- points -> leaves, packing ----
  int *d_sleaf, *d_tleaf, *d_slc, *d_tlc, *d_sperm, *d_tperm, *d_iota, *d_sorted;
  TRY(pool.get(&d_sleaf, NS));
  TRY(pool.get(&d_tleaf, NT));
  TRY(pool.get(&d_slc, N));
  TRY(pool.get(&d_tlc, N));
  TRY(pool.get(&d_sperm, NS));
  TRY(pool.get(&d_tperm, NT));
  TRY(pool.get(&d_iota, std::max(NS, NT)));
  TRY(pool.get(&d_sorted, std::max(NS, NT)));
  TRY(hipMemsetAsync(d_slc, 0, size_t(N) * 4, s));
  TRY(hipMemsetAsync(d_tlc, 0, size_t(N) * 4, s));
  if (NS) k_locate<<<blocks(NS), 256, 0, s>>>(d_spos, NS, d_soff, d_groot, d_child, d_depth, d_path, c, d_sleaf, d_slc);
  if (NT) k_locate<<<blocks(NT), 256, 0, s>>>(d_tpos, NT, d_toff, d_groot, d_child, d_depth, d_path, c, d_tleaf, d_tlc);
  if (std::max(NS, NT)) k_iota<<<blocks(std::max(NS, NT)), 256, 0, s>>>(d_iota, std::max(NS, NT));
  TRY(hipGetLastError());
  int *d_sbeg, *d_snum, *d_tbeg, *d_tnum, *d_scnt, *d_tcnt, *d_tag;
  TRY(pool.get(&d_sbeg, N));
  TRY(pool.get(&d_snum, N));
  TRY(pool.get(&d_tbeg, N));
  TRY(pool.get(&d_tnum, N));
  TRY(pool.get(&d_scnt, N));
  TRY(pool.get(&d_tcnt, N));
  TRY(pool.get(&d_tag, N));
  const unsigned node_bits = unsigned(nbits(N));
  {
    size_t tb = 0, tbs = 0;
    const size_t nmax = size_t(std::max(NS, NT));
    TRY(rocprim::radix_sort_pairs(nullptr, tb, d_sleaf, d_sorted, d_iota, d_sperm, nmax, 0, node_bits, s));
    char* tmp;
    TRY(pool.get(&tmp, int64_t(tb)));
    if (NS) {
      tbs = tb;
      TRY(rocprim::radix_sort_pairs(tmp, tbs, d_sleaf, d_sorted, d_iota, d_sperm, size_t(NS), 0, node_bits, s));
      k_leaf_ranges<<<blocks(NS), 256, 0, s>>>(d_sorted, NS, d_sbeg);
    }
    if (NT) {
      tbs = tb;
      TRY(rocprim::radix_sort_pairs(tmp, tbs, d_tleaf, d_sorted, d_iota, d_tperm, size_t(NT), 0, node_bits, s));
      k_leaf_ranges<<<blocks(NT), 256, 0, s>>>(d_sorted, NT, d_tbeg);
    }
    TRY(hipGetLastError());
  }
  // subtree counts: prefix sums of the leaf counts in node order
  int64_t *d_spre, *d_tpre, *d_wide;
  TRY(pool.get(&d_spre, N + 1));
  TRY(pool.get(&d_tpre, N + 1));
  TRY(pool.get(&d_wide, N + 1));
  {
    size_t tb = 0;
    TRY(rocprim::exclusive_scan(nullptr, tb, d_wide, d_spre, int64_t(0), size_t(N + 1), rocprim::plus<int64_t>(), s));
    char* tmp;
    TRY(pool.get(&tmp, int64_t(tb)));
    TRY(hipMemsetAsync(d_wide + N, 0, 8, s));
    k_widen<<<blocks(N), 256, 0, s>>>(d_slc, d_wide, N);
    TRY(rocprim::exclusive_scan(tmp, tb, d_wide, d_spre, int64_t(0), size_t(N + 1), rocprim::plus<int64_t>(), s));
    k_widen<<<blocks(N), 256, 0, s>>>(d_tlc, d_wide, N);
    TRY(rocprim::exclusive_scan(tmp, tb, d_wide, d_tpre, int64_t(0), size_t(N + 1), rocprim::plus<int64_t>(), s));
  }
  TRY(hipMemcpyAsync(d_scnt, d_slc, size_t(N) * 4, hipMemcpyDeviceToDevice, s));
  TRY(hipMemcpyAsync(d_tcnt, d_tlc, size_t(N) * 4, hipMemcpyDeviceToDevice, s));
  k_subtree_counts<<<blocks(G), 256, 0, s>>>(d_desc, int(G), d_cells, M, d_base, d_snode, c, d_spre, d_tpre, d_scnt,
                                             d_tcnt);
  k_tags<<<blocks(N), 256, 0, s>>>(N, d_parent, d_child, d_scnt, d_tcnt, d_sbeg, d_snum, d_tbeg, d_tnum, d_tag);
  TRY(hipGetLastError());

  // ---- lists ----
  ListCtx lc{d_parent, d_child, d_depth, d_path, d_proot, d_tcnt, d_snum, d_tag, N, int(node_bits)};
  int64_t *d_lcnt, *d_loff;
  TRY(pool.get(&d_lcnt, N + 1));
  TRY(pool.get(&d_loff, N + 1));
  TRY(hipMemsetAsync(d_lcnt + N, 0, 8, s));
  k_lists<<<blocks(N, 4), 256, 0, s>>>(lc, d_lcnt, nullptr, nullptr);
  TRY(hipGetLastError());
  {
    size_t tb = 0;
    TRY(rocprim::exclusive_scan(nullptr, tb, d_lcnt, d_loff, int64_t(0), size_t(N + 1), rocprim::plus<int64_t>(), s));
    char* tmp;
    TRY(pool.get(&tmp, int64_t(tb)));
    TRY(rocprim::exclusive_scan(tmp, tb, d_lcnt, d_loff, int64_t(0), size_t(N + 1), rocprim::plus<int64_t>(), s));
  }
  int64_t R = 0;
  TRY(hipMemcpyAsync(&R, d_loff + N, 8, hipMemcpyDeviceToHost, s));
  TRY(hipStreamSynchronize(s));
  u64 *d_raw, *d_sorted_keys;
  TRY(pool.get(&d_raw, R));
  TRY(pool.get(&d_sorted_keys, R));
  k_lists<<<blocks(N, 4), 256, 0, s>>>(lc, nullptr, d_loff, d_raw);
  TRY(hipGetLastError());
  int64_t U = 0;
  if (R) {
    size_t tb = 0, tb2 = 0;
    const unsigned lbits = 2 * node_bits + 2;
    TRY(rocprim::radix_sort_keys(nullptr, tb, d_raw, d_sorted_keys, size_t(R), 0, lbits, s));
    TRY(rocprim::unique(nullptr, tb2, d_sorted_keys, d_raw, d_m, size_t(R), rocprim::equal_to<u64>(), s));
    tb = std::max(tb, tb2);
    char* tmp;
    TRY(pool.get(&tmp, int64_t(tb)));
    TRY(rocprim::radix_sort_keys(tmp, tb, d_raw, d_sorted_keys, size_t(R), 0, lbits, s));
    TRY(rocprim::unique(tmp, tb, d_sorted_keys, d_raw, d_m, size_t(R), rocprim::equal_to<u64>(), s));
    TRY(hipMemcpyAsync(&U, d_m, 8, hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
  }
  // d_raw: the U sorted unique (kind, node, entry) keys
  int64_t* d_koff;
  int* d_lst;
  TRY(pool.get(&d_koff, 4 * int64_t(N + 1)));
  TRY(pool.get(&d_lst, U));
  k_list_offsets<<<blocks(4 * int64_t(N + 1)), 256, 0, s>>>(d_raw, U, N, int(node_bits), d_koff);
  if (U) k_list_entries<<<blocks(U), 256, 0, s>>>(d_raw, U, int(node_bits), d_lst, d_tag);
  TRY(hipGetLastError());
  TRY(hipStreamSynchronize(s));
  const auto t1 = clk::now();

  // ---- download into the host Tree ----
  t.n_nodes = N;
  TRY(download(t.parent, d_parent, N, s));
  TRY(download(t.child, d_child, N, s));
  TRY(download(t.depth, d_depth, N, s));
  TRY(download(t.path, d_path, 3 * int64_t(N), s));
  TRY(download(t.tag, d_tag, N, s));
  TRY(download(t.src_cnt, d_scnt, N, s));
  TRY(download(t.trg_cnt, d_tcnt, N, s));
  TRY(download(t.src_beg, d_sbeg, N, s));
  TRY(download(t.src_num, d_snum, N, s));
  TRY(download(t.trg_beg, d_tbeg, N, s));
  TRY(download(t.trg_num, d_tnum, N, s));
  TRY(download(t.src_perm, d_sperm, NS, s));
  TRY(download(t.trg_perm, d_tperm, NT, s));
  std::vector<int64_t> koff;
  TRY(download(koff, d_koff, 4 * int64_t(N + 1), s));
  std::vector<int> all;
  TRY(download(all, d_lst, U, s));
  int64_t base = 0;
  for (int k = 0; k < 4; k++) {
    t.off[k].assign(koff.begin() + int64_t(k) * (N + 1), koff.begin() + int64_t(k + 1) * (N + 1));
    const int64_t len = t.off[k][N];
    t.lst[k].assign(all.begin() + base, all.begin() + base + len);
    base += len;
  }
  for (int i = 0; i < N; i++) t.max_depth = std::max(t.max_depth, t.depth[i]);
  const auto t2 = clk::now();
  if (ms) {
    ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    ms[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
  }
  return 0;
}

}  // namespace kifmm
