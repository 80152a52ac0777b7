// tree_build.cpp -- host construction of the bisection tree, bit-exact with the
// reference's treemake (tree3d.cpp:1345-1551), treedivide (1163-1304),
// treeshrink (1099-1127) and treeinit (1045-1073).
//
// The reference keeps one heap array of member indices per node and copies it
// on every split.  Here every node owns a contiguous range of one permutation
// array, split in place by a stable partition that also carries the
// coordinates, so each level streams memory once per split direction.  Nodes
// are processed a generation at a time (all children of the previous
// generation): a node's split depends only on its own members, and its
// children's numbers are the running node count when the reference reaches it,
// i.e. a prefix sum over the generation.  Big nodes use a parallel partition,
// generations with many nodes run their nodes in parallel.
//
// Arithmetic that decides the tree (midpoints, aspect ratios, squared radii,
// `<=` comparisons) is written exactly as the reference evaluates it and this
// file is compiled with -ffp-contract=off.
#include "tree_build.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace bit {
namespace {

struct Arrays {
  int* P;
  double* X;
  double* Y;
  double* Z;
  int* tP;
  double* tX;
  double* tY;
  double* tZ;
};

// treeinit (tree3d.cpp:1045-1073).
void box_init(double* f, double x0, double x1, double y0, double y1, double z0, double z1) {
  double x = 0.5 * (x1 + x0);
  double y = 0.5 * (y1 + y0);
  double z = 0.5 * (z1 + z0);
  f[BX0] = x0; f[BX1] = x1; f[BY0] = y0; f[BY1] = y1; f[BZ0] = z0; f[BZ1] = z1;
  f[BXMID] = x; f[BYMID] = y; f[BZMID] = z;
  f[BAXY] = std::fabs((x1 - x0) / (y1 - y0));
  f[BAXZ] = std::fabs((x1 - x0) / (z1 - z0));
  f[BAYZ] = std::fabs((y1 - y0) / (z1 - z0));
  x = x1 - x;
  y = y1 - y;
  z = z1 - z;
  f[BSQR] = x * x + y * y + z * z;
}

struct MinMax {
  double mn[3], mx[3];
  bool any;
  void reset() {
    any = false;
    for (int d = 0; d < 3; d++) { mn[d] = 0; mx[d] = 0; }
  }
  void add(double x, double y, double z) {
    double v[3] = {x, y, z};
    if (!any) {
      for (int d = 0; d < 3; d++) mn[d] = mx[d] = v[d];
      any = true;
      return;
    }
    for (int d = 0; d < 3; d++) {
      if (v[d] > mx[d]) mx[d] = v[d];
      if (v[d] < mn[d]) mn[d] = v[d];
    }
  }
  void merge(const MinMax& o) {
    if (!o.any) return;
    if (!any) { *this = o; return; }
    for (int d = 0; d < 3; d++) {
      if (o.mx[d] > mx[d]) mx[d] = o.mx[d];
      if (o.mn[d] < mn[d]) mn[d] = o.mn[d];
    }
  }
};

// Stable partition of slots [b,e) by coord(dir) <= mid, carrying P,X,Y,Z.
// Returns the count on the `<=` side; mm[0], mm[1] receive each side's
// bounding box (treeshrink).  `threads` > 1 splits the range over threads.
int stable_partition(Arrays& a, int b, int e, int dir, double mid, MinMax mm[2], int threads) {
  const double* C = dir == 0 ? a.X : (dir == 1 ? a.Y : a.Z);
  const int n = e - b;
  mm[0].reset();
  mm[1].reset();
  if (threads <= 1 || n < (1 << 16)) {
    int nl = 0;
    for (int i = b; i < e; i++) nl += (C[i] <= mid);
    int l = b, r = b + nl;
    for (int i = b; i < e; i++) {
      int s = (C[i] <= mid) ? l++ : r++;
      a.tP[s] = a.P[i];
      a.tX[s] = a.X[i];
      a.tY[s] = a.Y[i];
      a.tZ[s] = a.Z[i];
      mm[s < b + nl ? 0 : 1].add(a.X[i], a.Y[i], a.Z[i]);
    }
    std::memcpy(a.P + b, a.tP + b, sizeof(int) * n);
    std::memcpy(a.X + b, a.tX + b, sizeof(double) * n);
    std::memcpy(a.Y + b, a.tY + b, sizeof(double) * n);
    std::memcpy(a.Z + b, a.tZ + b, sizeof(double) * n);
    return nl;
  }
  const int T = threads;
  std::vector<int> cl(T + 1, 0), cr(T + 1, 0);
  std::vector<MinMax> lm(T), rm(T);
  auto lo = [&](int t) { return b + (int)((int64_t)n * t / T); };
#pragma omp parallel num_threads(T)
  {
    int t = omp_get_thread_num();
    int s0 = lo(t), s1 = lo(t + 1), c = 0;
    for (int i = s0; i < s1; i++) c += (C[i] <= mid);
    cl[t + 1] = c;
    cr[t + 1] = (s1 - s0) - c;
#pragma omp barrier
#pragma omp single
    {
      for (int k = 0; k < T; k++) {
        cl[k + 1] += cl[k];
        cr[k + 1] += cr[k];
      }
    }
    int nl = cl[T];
    int l = b + cl[t], r = b + nl + cr[t];
    lm[t].reset();
    rm[t].reset();
    for (int i = s0; i < s1; i++) {
      int s;
      if (C[i] <= mid) {
        s = l++;
        lm[t].add(a.X[i], a.Y[i], a.Z[i]);
      } else {
        s = r++;
        rm[t].add(a.X[i], a.Y[i], a.Z[i]);
      }
      a.tP[s] = a.P[i];
      a.tX[s] = a.X[i];
      a.tY[s] = a.Y[i];
      a.tZ[s] = a.Z[i];
    }
#pragma omp barrier
    std::memcpy(a.P + s0, a.tP + s0, sizeof(int) * (s1 - s0));
    std::memcpy(a.X + s0, a.tX + s0, sizeof(double) * (s1 - s0));
    std::memcpy(a.Y + s0, a.tY + s0, sizeof(double) * (s1 - s0));
    std::memcpy(a.Z + s0, a.tZ + s0, sizeof(double) * (s1 - s0));
  }
  for (int t = 0; t < T; t++) {
    mm[0].merge(lm[t]);
    mm[1].merge(rm[t]);
  }
  return cl[T];
}

struct Slot {
  bool used = false;
  int b = 0, e = 0;
  double f[BOX_FIELDS];
  int nummemb() const { return used ? e - b : 0; }
};

// The children block one node creates: up to 8 slots (16 leave room for the
// z-split index below to point past the created ones).
struct Block {
  Slot s[16];
  int nchild = 0;
  int kids[8];
};

struct Splitter {
  Arrays& a;
  int maxmemb;
  double lim;
  int threads;

  // treedivide (tree3d.cpp:1163-1304): the first non-empty side goes to c1,
  // the second to c2; c1 may be the split node itself.
  int divide(Block& B, int dir, Slot src, int c1, int c2) {
    MinMax mm[2];
    int nl = stable_partition(a, src.b, src.e, dir, src.f[BXMID + dir], mm, threads);
    int lo[2] = {src.b, src.b + nl}, hi[2] = {src.b + nl, src.e};
    int target = c1, made = 0;
    for (int sd = 0; sd < 2; sd++) {
      if (hi[sd] > lo[sd]) {
        Slot& t = B.s[target];
        t.used = true;
        t.b = lo[sd];
        t.e = hi[sd];
        box_init(t.f, mm[sd].mn[0], mm[sd].mx[0], mm[sd].mn[1], mm[sd].mx[1], mm[sd].mn[2],
                 mm[sd].mx[2]);
        made++;
        target = c2;
      }
    }
    return made;
  }
  bool want_y(const Slot& c) const {
    return c.f[BAXY] < lim && c.f[BAYZ] > lim && c.nummemb() > maxmemb;
  }
  bool want_z(const Slot& c) const {
    return c.used && c.f[BAXZ] < lim && c.f[BAYZ] < lim && c.nummemb() > maxmemb;
  }

  // One pass of treemake's loop body for a node with too many members
  // (tree3d.cpp:1418-1545), on slots numbered from the block start.  After a
  // y-split, the z-split check addresses slot j + k*(nchild2+1) exactly as the
  // reference indexes tree[numtree+j+k*(nchild2+1)] (tree3d.cpp:1458-1477).
  bool split(const Slot& parent, Block& B) {
    int& nchild = B.nchild;
    int nchild2 = 0;
    auto add = [&](int slot) {
      if (nchild >= 8) return false;
      B.kids[nchild++] = slot;
      return true;
    };
    if (parent.f[BAXY] > lim && parent.f[BAXZ] > lim) {
      int xc = divide(B, 0, parent, 0, 1);
      int ind2 = xc;
      for (int j = 0; j < xc; j++) {
        if (!add(j)) return false;
        if (want_y(B.s[j])) {
          int yc = divide(B, 1, B.s[j], j, ind2);
          if (yc == 2) {
            if (!add(ind2)) return false;
            ind2++;
          }
          for (int k = 0; k < yc; k++) {
            int t = j + k * (nchild2 + 1);
            if (t < 16 && want_z(B.s[t])) {
              int zc = divide(B, 2, B.s[t], t, ind2);
              if (zc == 2) {
                if (!add(ind2)) return false;
                nchild2++;
                ind2++;
              }
            }
          }
        } else if (want_z(B.s[j])) {
          int zc = divide(B, 2, B.s[j], j, ind2);
          if (zc == 2) {
            if (!add(ind2)) return false;
            ind2++;
          }
        }
        if (ind2 >= 15) return false;
      }
    } else if (parent.f[BAXY] < lim && parent.f[BAYZ] > lim) {
      int yc = divide(B, 1, parent, 0, 1);
      int ind2 = yc;
      for (int j = 0; j < yc; j++) {
        if (!add(j)) return false;
        if (want_z(B.s[j])) {
          int zc = divide(B, 2, B.s[j], j, ind2);
          if (zc == 2) {
            if (!add(ind2)) return false;
            ind2++;
          }
        }
      }
    } else {
      int zc = divide(B, 2, parent, 0, 1);
      for (int j = 0; j < zc; j++)
        if (!add(j)) return false;
    }
    return true;
  }
};

// Spreads the low 10 bits of v so that two zero bits separate each bit.
inline uint32_t spread3(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

inline uint32_t quant(double v, double lo, double hi) {
  if (!(hi > lo)) return 0;
  double t = (v - lo) / (hi - lo) * 1023.0;
  if (t < 0) t = 0;
  if (t > 1023) t = 1023;
  return (uint32_t)t;
}

// Orders the slots of every leaf along a Morton curve over the leaf's box, so
// that 64 consecutive slots (one target batch) are close in space and take
// the same traversal decisions.  Leaf ranges, and therefore every node's
// member set, are unchanged.
void order_leaves(HostTree& t, Arrays& a, int threads) {
  std::vector<int> leaves;
  for (int i = 0; i < t.numtree; i++)
    if (t.children[(size_t)i * 8] < 0 && t.end[i] - t.begin[i] > 1) leaves.push_back(i);
#pragma omp parallel num_threads(threads)
  {
    std::vector<std::pair<uint32_t, int>> key;
#pragma omp for schedule(dynamic, 4)
    for (size_t li = 0; li < leaves.size(); li++) {
      const int lf = leaves[li];
      const int b = t.begin[lf], e = t.end[lf];
      const double* f = &t.box[(size_t)lf * BOX_FIELDS];
      key.resize(e - b);
      for (int s = b; s < e; s++) {
        uint32_t k = spread3(quant(a.X[s], f[BX0], f[BX1])) |
                     (spread3(quant(a.Y[s], f[BY0], f[BY1])) << 1) |
                     (spread3(quant(a.Z[s], f[BZ0], f[BZ1])) << 2);
        key[s - b] = {k, s};
      }
      std::stable_sort(key.begin(), key.end(),
                       [](const std::pair<uint32_t, int>& u, const std::pair<uint32_t, int>& v) {
                         return u.first < v.first;
                       });
      for (int s = b; s < e; s++) {
        int src = key[s - b].second;
        a.tP[s] = a.P[src];
        a.tX[s] = a.X[src];
        a.tY[s] = a.Y[src];
        a.tZ[s] = a.Z[src];
      }
      std::memcpy(a.P + b, a.tP + b, sizeof(int) * (e - b));
      std::memcpy(a.X + b, a.tX + b, sizeof(double) * (e - b));
      std::memcpy(a.Y + b, a.tY + b, sizeof(double) * (e - b));
      std::memcpy(a.Z + b, a.tZ + b, sizeof(double) * (e - b));
    }
  }
}

}  // namespace

bool build_tree(int n, const double* x, const double* y, const double* z, int maxmemb,
                HostTree& out, std::string& err, int max_threads) {
  if (n <= 0) {
    err = "tree needs at least one particle";
    return false;
  }
  if (maxmemb < 1) {
    err = "maxmemb must be >= 1";
    return false;
  }
  const double lim = 1.0 / std::sqrt(2.0);  // aspectlimit, tree3d.cpp:1371
  const int threads =
      std::max(1, max_threads > 0 ? std::min(max_threads, omp_get_max_threads())
                                  : omp_get_max_threads());
  const int node_cap = 4 * n + 64;

  std::vector<int> P(n), tP(n);
  std::vector<double> X(x, x + n), Y(y, y + n), Z(z, z + n), tX(n), tY(n), tZ(n);
  for (int i = 0; i < n; i++) P[i] = i;
  Arrays a{P.data(), X.data(), Y.data(), Z.data(), tP.data(), tX.data(), tY.data(), tZ.data()};

  out = HostTree();
  out.n = n;
  // Root box over all particles (tree3d.cpp:1385-1403).
  double mnx = x[0], mxx = x[0], mny = y[0], mxy = y[0], mnz = z[0], mxz = z[0];
  for (int i = 0; i < n; i++) {
    if (mnx > x[i]) mnx = x[i];
    if (mxx < x[i]) mxx = x[i];
    if (mny > y[i]) mny = y[i];
    if (mxy < y[i]) mxy = y[i];
    if (mnz > z[i]) mnz = z[i];
    if (mxz < z[i]) mxz = z[i];
  }
  out.box.resize(BOX_FIELDS);
  box_init(out.box.data(), mnx, mxx, mny, mxy, mnz, mxz);
  out.children.assign(8, -1);
  out.begin.push_back(0);
  out.end.push_back(n);
  out.depth.push_back(0);

  int gen_lo = 0, gen_hi = 1;
  while (gen_lo < gen_hi) {
    const int cnt = gen_hi - gen_lo;
    std::vector<Block> blocks(cnt);
    std::vector<char> ok(cnt, 1);
    auto run = [&](int idx, int part_threads) {
      int i = gen_lo + idx;
      if (out.end[i] - out.begin[i] <= maxmemb) return;
      Slot parent;
      parent.used = true;
      parent.b = out.begin[i];
      parent.e = out.end[i];
      std::memcpy(parent.f, &out.box[(size_t)i * BOX_FIELDS], sizeof(parent.f));
      Splitter sp{a, maxmemb, lim, part_threads};
      ok[idx] = sp.split(parent, blocks[idx]) ? 1 : 0;
    };
    if (cnt < 2 * threads) {
      for (int idx = 0; idx < cnt; idx++) run(idx, threads);
    } else {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
      for (int idx = 0; idx < cnt; idx++) run(idx, 1);
    }
    for (int idx = 0; idx < cnt; idx++)
      if (!ok[idx]) {
        err = "tree split overflow (more than 8 children)";
        return false;
      }
    std::vector<int> start(cnt);
    int next = gen_hi;
    for (int idx = 0; idx < cnt; idx++) {
      start[idx] = next;
      next += blocks[idx].nchild;
    }
    if (next > node_cap) {
      err = "tree exceeds 4n+64 nodes: particles cannot be separated into leaves of maxmemb";
      return false;
    }
    out.box.resize((size_t)next * BOX_FIELDS);
    out.children.resize((size_t)next * 8, -1);
    out.begin.resize(next);
    out.end.resize(next);
    out.depth.resize(next);
    for (int idx = 0; idx < cnt; idx++) {
      const Block& B = blocks[idx];
      int parent = gen_lo + idx;
      for (int s = 0; s < B.nchild; s++) {
        int g = start[idx] + s;
        std::memcpy(&out.box[(size_t)g * BOX_FIELDS], B.s[s].f, sizeof(B.s[s].f));
        out.begin[g] = B.s[s].b;
        out.end[g] = B.s[s].e;
        out.depth[g] = out.depth[parent] + 1;
        out.maxdepth = std::max(out.maxdepth, out.depth[g]);
      }
      for (int c = 0; c < B.nchild; c++) out.children[(size_t)parent * 8 + c] = start[idx] + B.kids[c];
    }
    gen_lo = gen_hi;
    gen_hi = next;
  }
  out.numtree = gen_hi;
  order_leaves(out, a, threads);
  out.perm = std::move(P);
  return true;
}

}  // namespace bit
