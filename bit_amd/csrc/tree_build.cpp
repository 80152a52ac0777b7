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
  int par = 0;  // splits that moved these slots, mod 2
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

// One pass of treemake's loop body for a node with too many members
// (tree3d.cpp:1418-1545) as a resumable program: it requests one stable split
// (treedivide, tree3d.cpp:1163-1304) at a time and continues when the split's
// result (side counts and bounding boxes) is applied, so the splits of all
// nodes of a generation can be executed together by a partition engine.  Slots
// are numbered from the node's block start; after a y-split the z-split check
// addresses slot j + k*(nchild2+1) exactly as the reference indexes
// tree[numtree+j+k*(nchild2+1)] (tree3d.cpp:1458-1477).
class SplitProgram {
 public:
  Block B;
  bool failed = false;

  void start(const Slot& parent, int maxmemb, double lim) {
    parent_ = parent;
    maxmemb_ = maxmemb;
    lim_ = lim;
    pc_ = 0;
    advance();
  }
  bool done() const { return pc_ == kDone; }
  const PartOp& op() const { return op_; }

  // treedivide's outcome: the first non-empty side goes to c1, the second to c2.
  void apply(const PartResult& r) {
    int lo[2] = {src_.b, src_.b + r.nl}, hi[2] = {src_.b + r.nl, src_.e};
    int target = c1_;
    made_ = 0;
    for (int sd = 0; sd < 2; sd++) {
      if (hi[sd] > lo[sd]) {
        Slot& t = B.s[target];
        t.used = true;
        t.b = lo[sd];
        t.e = hi[sd];
        t.par = src_.par ^ 1;
        box_init(t.f, r.mn[sd][0], r.mx[sd][0], r.mn[sd][1], r.mx[sd][1], r.mn[sd][2],
                 r.mx[sd][2]);
        made_++;
        target = c2_;
      }
    }
    advance();
  }

 private:
  static constexpr int kDone = -1;
  Slot parent_, src_;
  int maxmemb_ = 0;
  double lim_ = 0;
  int pc_ = 0, xc_ = 0, yc_ = 0, j_ = 0, k_ = 0, ind2_ = 0, nchild2_ = 0, made_ = 0;
  int c1_ = 0, c2_ = 0;
  PartOp op_{};

  bool want_y(const Slot& c) const {
    return c.used && c.f[BAXY] < lim_ && c.f[BAYZ] > lim_ && c.nummemb() > maxmemb_;
  }
  bool want_z(const Slot& c) const {
    return c.used && c.f[BAXZ] < lim_ && c.f[BAYZ] < lim_ && c.nummemb() > maxmemb_;
  }
  void request(int dir, const Slot& src, int c1, int c2) {
    src_ = src;
    c1_ = c1;
    c2_ = c2;
    op_ = PartOp{src.b, src.e, dir, src.f[BXMID + dir], src.par};
  }
  bool add(int slot) {
    if (B.nchild >= 8) {
      failed = true;
      pc_ = kDone;
      return false;
    }
    B.kids[B.nchild++] = slot;
    return true;
  }

  // Runs the program until it requests a split (returns) or ends (pc_ = kDone).
  void advance() {
    for (;;) {
      switch (pc_) {
        case 0:  // first split direction from the parent's aspect ratios
          if (parent_.f[BAXY] > lim_ && parent_.f[BAXZ] > lim_) {
            request(0, parent_, 0, 1);
            pc_ = 1;
          } else if (parent_.f[BAXY] < lim_ && parent_.f[BAYZ] > lim_) {
            request(1, parent_, 0, 1);
            pc_ = 20;
          } else {
            request(2, parent_, 0, 1);
            pc_ = 30;
          }
          return;
        case 1:  // after the x-split
          xc_ = made_;
          ind2_ = xc_;
          j_ = 0;
          pc_ = 2;
          break;
        case 2:  // loop over x-children j
          if (j_ >= xc_) {
            pc_ = kDone;
            return;
          }
          if (!add(j_)) return;
          if (want_y(B.s[j_])) {
            request(1, B.s[j_], j_, ind2_);
            pc_ = 3;
            return;
          }
          if (want_z(B.s[j_])) {
            request(2, B.s[j_], j_, ind2_);
            pc_ = 6;
            return;
          }
          j_++;
          break;
        case 3:  // after the y-split of x-child j
          yc_ = made_;
          if (yc_ == 2) {
            if (!add(ind2_)) return;
            ind2_++;
          }
          k_ = 0;
          pc_ = 4;
          break;
        case 4: {  // z-split checks after a y-split
          if (k_ >= yc_) {
            j_++;
            pc_ = 2;
            break;
          }
          int t = j_ + k_ * (nchild2_ + 1);
          if (t < 16 && want_z(B.s[t])) {
            request(2, B.s[t], t, ind2_);
            pc_ = 5;
            return;
          }
          k_++;
          break;
        }
        case 5:  // after a z-split in the k loop
          if (made_ == 2) {
            if (!add(ind2_)) return;
            nchild2_++;
            ind2_++;
          }
          k_++;
          pc_ = 4;
          break;
        case 6:  // after a z-split of an x-child that was not y-split
          if (made_ == 2) {
            if (!add(ind2_)) return;
            ind2_++;
          }
          j_++;
          pc_ = 2;
          break;
        case 20:  // after a y-first split
          yc_ = made_;
          ind2_ = yc_;
          j_ = 0;
          pc_ = 21;
          break;
        case 21:
          if (j_ >= yc_) {
            pc_ = kDone;
            return;
          }
          if (!add(j_)) return;
          if (want_z(B.s[j_])) {
            request(2, B.s[j_], j_, ind2_);
            pc_ = 22;
            return;
          }
          j_++;
          break;
        case 22:
          if (made_ == 2) {
            if (!add(ind2_)) return;
            ind2_++;
          }
          j_++;
          pc_ = 21;
          break;
        case 30:  // after a z-only split
          for (int jj = 0; jj < made_; jj++)
            if (!add(jj)) return;
          pc_ = kDone;
          return;
        default:
          return;
      }
      if (ind2_ >= 15) {
        failed = true;
        pc_ = kDone;
        return;
      }
    }
  }
};

// Partitions on the host arrays: ops of one round in parallel, or one big op
// with a parallel partition.
class HostEngine : public PartitionEngine {
 public:
  HostEngine(Arrays& a, int threads) : a_(a), threads_(threads) {}
  bool run(const std::vector<PartOp>& ops, std::vector<PartResult>& res,
           std::string&) override {
    res.resize(ops.size());
    auto one = [&](size_t i, int th) {
      MinMax mm[2];
      const PartOp& o = ops[i];
      int nl = stable_partition(a_, o.b, o.e, o.dir, o.mid, mm, th);
      PartResult& r = res[i];
      r.nl = nl;
      for (int sd = 0; sd < 2; sd++)
        for (int d = 0; d < 3; d++) {
          r.mn[sd][d] = mm[sd].mn[d];
          r.mx[sd][d] = mm[sd].mx[d];
        }
    };
    if ((int)ops.size() < 2 * threads_) {
      for (size_t i = 0; i < ops.size(); i++) one(i, threads_);
    } else {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads_)
      for (size_t i = 0; i < ops.size(); i++) one(i, 1);
    }
    return true;
  }

 private:
  Arrays& a_;
  int threads_;
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
        uint32_t k = leaf_morton_key(a.X[s], a.Y[s], a.Z[s], f);
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

uint32_t leaf_morton_key(double x, double y, double z, const double* f) {
  return spread3(quant(x, f[BX0], f[BX1])) | (spread3(quant(y, f[BY0], f[BY1])) << 1) |
         (spread3(quant(z, f[BZ0], f[BZ1])) << 2);
}

bool build_tree_generic(PartitionEngine& eng, int n, const double mn[3], const double mx[3],
                        int maxmemb, HostTree& out, std::string& err) {
  if (n <= 0) {
    err = "tree needs at least one particle";
    return false;
  }
  if (maxmemb < 1) {
    err = "maxmemb must be >= 1";
    return false;
  }
  const double lim = 1.0 / std::sqrt(2.0);  // aspectlimit, tree3d.cpp:1371
  const int node_cap = 4 * n + 64;
  out = HostTree();
  out.n = n;
  out.box.resize(BOX_FIELDS);
  box_init(out.box.data(), mn[0], mx[0], mn[1], mx[1], mn[2], mx[2]);  // tree3d.cpp:1385-1403
  out.children.assign(8, -1);
  out.begin.push_back(0);
  out.end.push_back(n);
  out.depth.push_back(0);
  out.par.push_back(0);

  int gen_lo = 0, gen_hi = 1;
  std::vector<PartOp> ops;
  std::vector<PartResult> res;
  std::vector<int> who;
  while (gen_lo < gen_hi) {
    const int cnt = gen_hi - gen_lo;
    std::vector<SplitProgram> prog(cnt);
    std::vector<int> active;
    for (int idx = 0; idx < cnt; idx++) {
      int i = gen_lo + idx;
      if (out.end[i] - out.begin[i] <= maxmemb) continue;
      Slot parent;
      parent.used = true;
      parent.b = out.begin[i];
      parent.e = out.end[i];
      parent.par = out.par[i];
      std::memcpy(parent.f, &out.box[(size_t)i * BOX_FIELDS], sizeof(parent.f));
      prog[idx].start(parent, maxmemb, lim);
      active.push_back(idx);
    }
    // rounds: every node with a pending split advances by one split
    while (!active.empty()) {
      ops.clear();
      who.clear();
      for (int idx : active)
        if (!prog[idx].done()) {
          ops.push_back(prog[idx].op());
          who.push_back(idx);
        }
      if (ops.empty()) break;
      if (!eng.run(ops, res, err)) return false;
      for (size_t k = 0; k < ops.size(); k++) prog[who[k]].apply(res[k]);
      std::vector<int> still;
      for (int idx : active)
        if (!prog[idx].done()) still.push_back(idx);
      active.swap(still);
    }
    std::vector<int> start(cnt);
    int next = gen_hi;
    for (int idx = 0; idx < cnt; idx++) {
      if (prog[idx].failed) {
        err = "tree split overflow (more than 8 children)";
        return false;
      }
      start[idx] = next;
      next += prog[idx].B.nchild;
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
    out.par.resize(next);
    for (int idx = 0; idx < cnt; idx++) {
      const Block& B = prog[idx].B;
      int parent = gen_lo + idx;
      for (int sl = 0; sl < B.nchild; sl++) {
        int g = start[idx] + sl;
        std::memcpy(&out.box[(size_t)g * BOX_FIELDS], B.s[sl].f, sizeof(B.s[sl].f));
        out.begin[g] = B.s[sl].b;
        out.end[g] = B.s[sl].e;
        out.depth[g] = out.depth[parent] + 1;
        out.par[g] = (unsigned char)B.s[sl].par;
        out.maxdepth = std::max(out.maxdepth, out.depth[g]);
      }
      for (int c = 0; c < B.nchild; c++)
        out.children[(size_t)parent * 8 + c] = start[idx] + B.kids[c];
    }
    gen_lo = gen_hi;
    gen_hi = next;
  }
  out.numtree = gen_hi;
  return true;
}

bool build_tree(int n, const double* x, const double* y, const double* z, int maxmemb,
                HostTree& out, std::string& err, int max_threads) {
  if (n <= 0) {
    err = "tree needs at least one particle";
    return false;
  }
  const int threads =
      std::max(1, max_threads > 0 ? std::min(max_threads, omp_get_max_threads())
                                  : omp_get_max_threads());
  std::vector<int> P(n), tP(n);
  std::vector<double> X(x, x + n), Y(y, y + n), Z(z, z + n), tX(n), tY(n), tZ(n);
  for (int i = 0; i < n; i++) P[i] = i;
  Arrays a{P.data(), X.data(), Y.data(), Z.data(), tP.data(), tX.data(), tY.data(), tZ.data()};
  // root bounding box over all particles (tree3d.cpp:1385-1403)
  double mn[3] = {x[0], y[0], z[0]}, mx[3] = {x[0], y[0], z[0]};
  for (int i = 0; i < n; i++) {
    if (mn[0] > x[i]) mn[0] = x[i];
    if (mx[0] < x[i]) mx[0] = x[i];
    if (mn[1] > y[i]) mn[1] = y[i];
    if (mx[1] < y[i]) mx[1] = y[i];
    if (mn[2] > z[i]) mn[2] = z[i];
    if (mx[2] < z[i]) mx[2] = z[i];
  }
  HostEngine eng(a, threads);
  if (!build_tree_generic(eng, n, mn, mx, maxmemb, out, err)) return false;
  order_leaves(out, a, threads);
  out.perm = std::move(P);
  return true;
}

}  // namespace bit
