// tree3d.cpp -- the reference library interface (tree3d.h) on top of the
// MI355X C-ABI (include/bit_amd.h).  Built as bit_amd/libtree3d.so.
//
// A tree made by treemake is bound to its PARTICLE array; the first evaluation
// that needs the device creates a bit_system for it (same tree: both are built
// by the same bit-exact host code) and every batched result is cached per
// (tree, start node, quantity, nterms, acc, alpha) until treeempty releases the
// tree.
#include "tree3d.h"

#include <algorithm>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <tuple>
#include <vector>

#include "bit_amd.h"
#include "problem.h"

namespace {

std::mutex g_mu;
bit_ctx* g_ctx = nullptr;

// Self term shielded_treepot adds for the target particle itself, per unit
// tot_charge/weight (tree3d.cpp:3575-3596 of the reference).
const double kSelfIntG = -0.0000483811229725;

enum Quantity { kField = 0, kPot = 1, kShielded = 2 };

struct Bound {
  PARTICLE* part = nullptr;
  int npart = 0;
  int numtree = 0;
  int maxmemb = 0;
  bit_system* sys = nullptr;
  std::map<std::tuple<int, int, int, double, double>, std::vector<double>> cache;
};
std::map<TREE*, Bound> g_bound;
std::set<double***> g_own_moments;  // moment blocks allocated by treemoments

bit_ctx* context() {
  if (!g_ctx && bit_ctx_create(0, 0, &g_ctx) != BIT_OK) {
    cout << "[tree3d] no GPU context: " << bit_last_error() << endl;
    g_ctx = nullptr;
  }
  return g_ctx;
}

void coords(const PARTICLE* part, int n, std::vector<double>& x, std::vector<double>& y,
            std::vector<double>& z, std::vector<double>& q) {
  x.resize(n);
  y.resize(n);
  z.resize(n);
  q.resize(n);
  for (int i = 0; i < n; i++) {
    x[i] = part[i].x;
    y[i] = part[i].y;
    z[i] = part[i].z;
    q[i] = part[i].tot_charge;
  }
}

bit_system* make_system(PARTICLE* part, int n, int maxmemb) {
  bit_ctx* c = context();
  if (!c) return nullptr;
  std::vector<double> x, y, z, q;
  coords(part, n, x, y, z, q);
  bit_system* s = nullptr;
  if (bit_system_create(c, n, x.data(), y.data(), z.data(), q.data(), maxmemb, &s) != BIT_OK) {
    cout << "[tree3d] " << bit_last_error() << endl;
    return nullptr;
  }
  return s;
}

void release(Bound& b) {
  if (b.sys) bit_system_destroy(b.sys);
  b.sys = nullptr;
  b.cache.clear();
}

bit_params params(int nterms, double acc, double kappa, double delta2, double mac_sub,
                  int root) {
  bit_params p;
  bit_params_default(&p);
  p.root = root;
  p.nterms = nterms;
  p.acc = acc;
  p.kappa = kappa;
  p.delta2 = delta2;
  p.mac_sub = mac_sub;
  p.precision = BIT_FP64;
  return p;
}

// Field (3 per particle) or potential of every particle, in the reference
// semantics of the per-target function the quantity stands for, from the
// traversal that starts at node `root` (the reference's `ind`).  `node` is that
// node (nullptr: the whole particle set): shielded_treepot adds its self term
// when the target's own leaf is summed directly, that is for the node's members.
bool evaluate(bit_system* s, int n, const PARTICLE* part, int what, int nterms, double acc,
              double alpha, int root, const TREE* node, std::vector<double>& out) {
  if (what == kField) {
    out.resize(3 * (size_t)n);
    bit_params p = params(nterms, acc, 0.0, DEL, DEL, root);  // treeforce: rs = r^2 + DEL
    p.kernel_order = KERNEL_ORDER;  // regularized family (problem.h)
    if (bit_eval_field(s, &p, out.data(), 0) != BIT_OK) return false;
    return true;
  }
  out.resize(n);
  if (what == kPot) {
    bit_params p = params(nterms, acc, 0.0, DEL, DEL, root);  // treepotential: rs = r^2 + DEL
    return bit_eval_potential(s, &p, out.data(), 0) == BIT_OK;
  }
  // shielded_treepot: unregularized kernel, MAC shifted by DEL, result -phi plus
  // the target's self term.
  bit_params p = params(nterms, acc, alpha, 0.0, DEL, root);
  if (bit_eval_potential(s, &p, out.data(), 0) != BIT_OK) return false;
  for (int i = 0; i < n; i++) out[i] = -out[i];
  auto self = [&](int i) { out[i] += part[i].tot_charge * kSelfIntG / part[i].weight; };
  if (!node)
    for (int i = 0; i < n; i++) self(i);
  else
    for (int m = 0; m < node->nummemb; m++) self(node->members[m]);
  return true;
}

// The cached batched result for a per-target call, or nullptr (message printed).
const double* lookup(const char* fn, TREE* tree, PARTICLE* part, int ind, int p, double x,
                     double y, double z, int what, int nterms, double acc, double alpha) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_bound.find(tree);
  if (it == g_bound.end() || it->second.part != part) {
    cout << "[" << fn << "] tree was not built by treemake for this particle array" << endl;
    return nullptr;
  }
  Bound& b = it->second;
  if (ind < 0 || ind >= b.numtree) {
    cout << "[" << fn << "] node " << ind << " is not in the tree (numtree " << b.numtree
         << ")" << endl;
    return nullptr;
  }
  if (p < 0 || p >= b.npart || part[p].x != x || part[p].y != y || part[p].z != z) {
    cout << "[" << fn << "] served at particle targets (0 <= p < npart, x,y,z = particle p) "
         << "or at points (p == -1) only" << endl;
    return nullptr;
  }
  auto key = std::make_tuple(ind, what, nterms, acc, alpha);
  auto c = b.cache.find(key);
  if (c == b.cache.end()) {
    if (!b.sys) b.sys = make_system(part, b.npart, b.maxmemb);
    if (!b.sys) return nullptr;
    std::vector<double> v;
    if (!evaluate(b.sys, b.npart, part, what, nterms, acc, alpha, ind, &tree[ind], v)) {
      cout << "[" << fn << "] " << bit_last_error() << endl;
      return nullptr;
    }
    c = b.cache.emplace(key, std::move(v)).first;
  }
  return c->second.data();
}

// One arbitrary point (the reference's p = -1 calls): every source of the
// subtree at node ind contributes, pairs closer than eps skipped.  Returns
// false (message printed) when the tree is not bound or the evaluation fails.
bool point_eval(const char* fn, TREE* tree, PARTICLE* part, int ind, int what, int nterms,
                double acc, double alpha, double x, double y, double z, double* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_bound.find(tree);
  if (it == g_bound.end() || it->second.part != part) {
    cout << "[" << fn << "] tree was not built by treemake for this particle array" << endl;
    return false;
  }
  Bound& b = it->second;
  if (ind < 0 || ind >= b.numtree) {
    cout << "[" << fn << "] node " << ind << " is not in the tree (numtree " << b.numtree
         << ")" << endl;
    return false;
  }
  if (!b.sys) b.sys = make_system(part, b.npart, b.maxmemb);
  if (!b.sys) return false;
  int rc;
  if (what == kField) {
    bit_params p = params(nterms, acc, 0.0, DEL, DEL, ind);
    p.kernel_order = KERNEL_ORDER;
    rc = bit_eval_field_points(b.sys, &p, 1, &x, &y, &z, out);
  } else if (what == kPot) {
    bit_params p = params(nterms, acc, 0.0, DEL, DEL, ind);
    rc = bit_eval_potential_points(b.sys, &p, 1, &x, &y, &z, out);
  } else {
    bit_params p = params(nterms, acc, alpha, 0.0, DEL, ind);
    rc = bit_eval_potential_points(b.sys, &p, 1, &x, &y, &z, out);
    out[0] = -out[0];  // shielded_treepot's sign, near and far alike
  }
  if (rc != BIT_OK) {
    cout << "[" << fn << "] " << bit_last_error() << endl;
    return false;
  }
  return true;
}

void fill_tree_node(TREE& t, const double* f) {
  t.x0 = f[0]; t.x1 = f[1]; t.y0 = f[2]; t.y1 = f[3]; t.z0 = f[4]; t.z1 = f[5];
  t.xmid = f[6]; t.ymid = f[7]; t.zmid = f[8];
  t.aspectxy = f[9]; t.aspectxz = f[10]; t.aspectyz = f[11];
  t.sqradius = f[12];
}

// Whole-system driver: part[i].{x,y,z}force from the batched result.
void drive(const char* fn, int npart, PARTICLE* part, int what, int nterms, double acc,
           double alpha, int maxmemb, bool print_numtree) {
  std::lock_guard<std::mutex> lk(g_mu);
  bit_system* s = make_system(part, npart, maxmemb);
  if (!s) return;
  if (print_numtree) cout << "numtree = " << bit_system_numtree(s) << endl;
  std::vector<double> v;
  if (!evaluate(s, npart, part, what, nterms, acc, alpha, 0, nullptr, v)) {
    cout << "[" << fn << "] " << bit_last_error() << endl;
    bit_system_destroy(s);
    return;
  }
  for (int i = 0; i < npart; i++) {  // the reference's product order
    if (what == kField) {
      part[i].xforce = v[3 * (size_t)i] * charge_constant * part[i].tot_charge;
      part[i].yforce = v[3 * (size_t)i + 1] * charge_constant * part[i].tot_charge;
      part[i].zforce = v[3 * (size_t)i + 2] * charge_constant * part[i].tot_charge;
    } else {
      part[i].xforce = v[i] * charge_constant * part[i].tot_charge;
    }
  }
  bit_system_destroy(s);
}

}  // namespace

// ---------------------------------------------------------------------------
// tree
// ---------------------------------------------------------------------------
void treeinit(TREE* t, double x0, double x1, double y0, double y1, double z0, double z1) {
  double x = 0.5 * (x1 + x0), y = 0.5 * (y1 + y0), z = 0.5 * (z1 + z0);
  double f[13] = {x0, x1, y0, y1, z0, z1, x, y, z, fabs((x1 - x0) / (y1 - y0)),
                  fabs((x1 - x0) / (z1 - z0)), fabs((y1 - y0) / (z1 - z0)), 0.0};
  double hx = x1 - x, hy = y1 - y, hz = z1 - z;
  f[12] = hx * hx + hy * hy + hz * hz;
  fill_tree_node(*t, f);
  for (int i = 0; i < NUMCHILD; i++) t->children[i] = -1;
}

void treeshrink(TREE* t, PARTICLE* part) {
  int j = t->members[0];
  double mx = part[j].x, Mx = mx, my = part[j].y, My = my, mz = part[j].z, Mz = mz;
  for (int i = 0; i < t->nummemb; i++) {
    j = t->members[i];
    Mx = std::max(Mx, part[j].x);
    mx = std::min(mx, part[j].x);
    My = std::max(My, part[j].y);
    my = std::min(my, part[j].y);
    Mz = std::max(Mz, part[j].z);
    mz = std::min(mz, part[j].z);
  }
  treeinit(t, mx, Mx, my, My, mz, Mz);
}

int treedivide(int dir, TREE* split, TREE* c1, TREE* c2, PARTICLE* part) {
  double mid = dir == 0 ? split->xmid : (dir == 1 ? split->ymid : split->zmid);
  std::vector<int> side[2];
  for (int j = 0; j < split->nummemb; j++) {
    int i = split->members[j];
    double v = dir == 0 ? part[i].x : (dir == 1 ? part[i].y : part[i].z);
    side[v <= mid ? 0 : 1].push_back(i);
  }
  TREE* dst = c1;
  int made = 0;
  for (int s = 0; s < 2; s++) {
    if (side[s].empty()) continue;
    delete[] dst->members;
    dst->nummemb = (int)side[s].size();
    dst->members = new int[side[s].size()];
    std::copy(side[s].begin(), side[s].end(), dst->members);
    treeshrink(dst, part);
    made++;
    dst = c2;
  }
  return made;
}

int treemake(int npart, int nterms, int maxmemb, PARTICLE* part, TREE* tree) {
  (void)nterms;
  std::vector<double> x, y, z, q;
  coords(part, npart, x, y, z, q);
  bit_tree* t = nullptr;
  if (bit_tree_build(npart, x.data(), y.data(), z.data(), maxmemb, &t) != BIT_OK) {
    cout << "[treemake] " << bit_last_error() << endl;
    return 0;
  }
  const int nt = bit_tree_numtree(t);
  std::vector<double> boxes(13 * (size_t)nt);
  std::vector<int> ch(8 * (size_t)nt), b(nt), e(nt), perm(npart);
  bit_tree_export(t, boxes.data(), ch.data(), b.data(), e.data(), perm.data());
  bit_tree_destroy(t);
  for (int i = 0; i < nt; i++) {
    TREE& d = tree[i];
    fill_tree_node(d, &boxes[13 * (size_t)i]);
    for (int c = 0; c < NUMCHILD; c++) d.children[c] = ch[8 * (size_t)i + c];
    d.nummemb = e[i] - b[i];
    delete[] d.members;
    d.members = new int[d.nummemb];
    std::copy(perm.begin() + b[i], perm.begin() + e[i], d.members);
    std::sort(d.members, d.members + d.nummemb);  // reference order: ascending indices
  }
  std::lock_guard<std::mutex> lk(g_mu);
  Bound& bd = g_bound[tree];
  release(bd);
  bd.part = part;
  bd.npart = npart;
  bd.numtree = nt;
  bd.maxmemb = maxmemb;
  return nt;
}

void treeempty(int numtree, TREE* tree) {
  if (!tree) return;
  for (int i = 0; i < numtree; i++) {
    delete[] tree[i].members;
    tree[i].members = nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  for (int i = 0; i < numtree; i++) {
    if (!tree[i].moments) continue;
    if (g_own_moments.erase(tree[i].moments)) {  // one block made by treemoments
      delete[] tree[i].moments[0][0];
      delete[] tree[i].moments[0];
    }
    delete[] tree[i].moments;  // as the reference: the outer array only otherwise
    tree[i].moments = nullptr;
  }
  auto it = g_bound.find(tree);
  if (it != g_bound.end()) {
    release(it->second);
    g_bound.erase(it);
  }
}

void treemoments(int numtree, int nterms, TREE* tree, PARTICLE* part) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_bound.find(tree);
  if (it == g_bound.end() || it->second.part != part) {
    cout << "[treemoments] tree was not built by treemake for this particle array" << endl;
    return;
  }
  Bound& b = it->second;
  if (!b.sys) {
    b.sys = make_system(part, b.npart, b.maxmemb);
    if (!b.sys) return;
  } else {
    // the reference recomputes the moments from part[].tot_charge on every
    // call (tree3d.cpp:2685-2707): upload the current charges, and drop the
    // results computed with the previous ones
    std::vector<double> q(b.npart);
    for (int i = 0; i < b.npart; i++) q[i] = part[i].tot_charge;
    if (bit_system_set_charges(b.sys, q.data(), 0) != BIT_OK) {
      cout << "[treemoments] " << bit_last_error() << endl;
      return;
    }
    b.cache.clear();
  }
  if (bit_moments(b.sys, nterms) != BIT_OK) {
    cout << "[treemoments] " << bit_last_error() << endl;
    return;
  }
  const size_t cube = (size_t)nterms * nterms * nterms;
  std::vector<double> m(cube * numtree);
  if (bit_moments_download(b.sys, m.data()) != BIT_OK) {
    cout << "[treemoments] " << bit_last_error() << endl;
    return;
  }
  for (int t = 0; t < numtree; t++) {
    if (!tree[t].moments) {
      // one contiguous block, indexable as moments[i][j][k]
      double* data = new double[cube];
      double** rows = new double*[(size_t)nterms * nterms];
      tree[t].moments = new double**[nterms];
      g_own_moments.insert(tree[t].moments);
      for (int i = 0; i < nterms; i++) {
        tree[t].moments[i] = rows + (size_t)i * nterms;
        for (int j = 0; j < nterms; j++) rows[(size_t)i * nterms + j] = data + ((size_t)i * nterms + j) * nterms;
      }
    }
    for (int i = 0; i < nterms; i++)
      for (int j = 0; j < nterms; j++)
        for (int k = 0; k < nterms; k++)
          tree[t].moments[i][j][k] = m[cube * t + ((size_t)i * nterms + j) * nterms + k];
  }
}

// ---------------------------------------------------------------------------
// per-target evaluation
// ---------------------------------------------------------------------------
double treeforce(int ind, int p, int nterms, double x, double y, double z, double acc,
                 double* xfor, double* yfor, double* zfor, TREE* tree, PARTICLE* part) {
  if (p == -1) {
    double E[3];
    if (point_eval("treeforce", tree, part, ind, kField, nterms, acc, 0.0, x, y, z, E)) {
      *xfor += E[0];
      *yfor += E[1];
      *zfor += E[2];
    }
    return 0;
  }
  const double* E = lookup("treeforce", tree, part, ind, p, x, y, z, kField, nterms, acc, 0.0);
  if (E) {
    *xfor += E[3 * (size_t)p];
    *yfor += E[3 * (size_t)p + 1];
    *zfor += E[3 * (size_t)p + 2];
  }
  return 0;
}

double treepotential(int ind, int p, int nterms, double x, double y, double z, double acc,
                     double* pot, TREE* tree, PARTICLE* part) {
  if (p == -1) {
    double v;
    if (point_eval("treepotential", tree, part, ind, kPot, nterms, acc, 0.0, x, y, z, &v))
      *pot += v;
    return 0;
  }
  const double* v = lookup("treepotential", tree, part, ind, p, x, y, z, kPot, nterms, acc, 0.0);
  if (v) *pot += v[p];
  return 0;
}

double shielded_treepot(int ind, int p, int nterms, double x, double y, double z, double alpha,
                        double acc, double* pot, TREE* tree, PARTICLE* part, int, int, int,
                        double, double, double) {
  if (p == -1) {
    double v;
    if (point_eval("shielded_treepot", tree, part, ind, kShielded, nterms, acc, alpha, x, y, z,
                   &v))
      *pot += v;
    return 0;
  }
  const double* v =
      lookup("shielded_treepot", tree, part, ind, p, x, y, z, kShielded, nterms, acc, alpha);
  if (v) *pot += v[p];
  return 0;
}

// ---------------------------------------------------------------------------
// whole-system drivers
// ---------------------------------------------------------------------------
void freespace_particles_tree(int npart, PARTICLE* part, int nterms, double acc, TREE*, int,
                              int maxmemb) {
  drive("freespace_particles_tree", npart, part, kField, nterms, acc, 0.0, maxmemb, true);
}

void freespace_pot_tree(int npart, PARTICLE* part, int nterms, double acc, TREE*, int,
                        int maxmemb) {
  drive("freespace_pot_tree", npart, part, kPot, nterms, acc, 0.0, maxmemb, false);
}

void shielded_freespace_pot_tree(int npart, PARTICLE* part, double alpha, int nterms,
                                 double acc, TREE*, int, int maxmemb) {
  drive("shielded_freespace_pot_tree", npart, part, kShielded, nterms, acc, alpha, maxmemb,
        false);
}

// ---------------------------------------------------------------------------
// direct sums (no tree): GPU FP64 sums over every particle
// ---------------------------------------------------------------------------
namespace {

enum DirectKind { kDsNoreg = 0, kDsReg = 1, kDsShieldedPot = 2 };

// The particle array as uploaded for the direct sums: a one-leaf system and a
// snapshot of the fields the sums read.  Results per (kind, alpha), computed
// for all particles at once, are served while the array still holds the
// snapshot's values, so per-particle calls from many threads cost one GPU
// evaluation and a comparison each.
struct DirectState {
  const PARTICLE* part = nullptr;
  int npart = 0;
  std::vector<uint64_t> snap;  // bit patterns of x, y, z, tot_charge per particle
  bit_system* sys = nullptr;
  std::map<std::pair<int, double>, std::shared_ptr<const std::vector<double>>> results;
  ~DirectState() {
    if (sys) bit_system_destroy(sys);
  }
};
// guarded by g_mu; never destroyed, so no device call runs at process exit
std::shared_ptr<DirectState>& g_direct = *new std::shared_ptr<DirectState>();

uint64_t bits(double v) {
  uint64_t u;
  std::memcpy(&u, &v, sizeof u);
  return u;
}

bool same_particles(const DirectState& st, const PARTICLE* part, int npart) {
  if (st.part != part || st.npart != npart) return false;
  const uint64_t* s = st.snap.data();
  for (int i = 0; i < npart; i++, s += 4)
    if (s[0] != bits(part[i].x) || s[1] != bits(part[i].y) || s[2] != bits(part[i].z) ||
        s[3] != bits(part[i].tot_charge))
      return false;
  return true;
}

// The current direct-sum state of `part` (made anew when the array changed),
// with g_mu held by the caller's lock `lk` on return; nullptr on failure
// (message printed).
std::shared_ptr<DirectState> direct_state(const char* fn, const PARTICLE* part, int npart,
                                          std::unique_lock<std::mutex>& lk) {
  for (;;) {
    std::shared_ptr<DirectState> st = g_direct;
    lk.unlock();
    const bool same = st && same_particles(*st, part, npart);  // O(n), outside the lock
    lk.lock();
    if (g_direct != st) continue;  // replaced meanwhile: compare again
    if (same) return st;
    auto ns = std::make_shared<DirectState>();
    ns->part = part;
    ns->npart = npart;
    ns->snap.resize(4 * (size_t)npart);
    for (int i = 0; i < npart; i++) {
      ns->snap[4 * (size_t)i] = bits(part[i].x);
      ns->snap[4 * (size_t)i + 1] = bits(part[i].y);
      ns->snap[4 * (size_t)i + 2] = bits(part[i].z);
      ns->snap[4 * (size_t)i + 3] = bits(part[i].tot_charge);
    }
    ns->sys = make_system(const_cast<PARTICLE*>(part), npart, std::max(1, npart));
    if (!ns->sys) {
      cout << "[" << fn << "] direct sum: no device copy of the particles" << endl;
      return nullptr;
    }
    g_direct = ns;
    return ns;
  }
}

// All-particle result of a direct-sum kind: 3 per particle (fields) or 1
// (shielded potential, the reference's -phi).
std::shared_ptr<const std::vector<double>> direct_all(const char* fn, const PARTICLE* part,
                                                      int npart, int kind, double alpha) {
  std::unique_lock<std::mutex> lk(g_mu);
  std::shared_ptr<DirectState> st = direct_state(fn, part, npart, lk);
  if (!st) return nullptr;
  auto key = std::make_pair(kind, kind == kDsShieldedPot ? alpha : 0.0);
  auto it = st->results.find(key);
  if (it != st->results.end()) return it->second;
  std::vector<int> idx(npart);
  for (int i = 0; i < npart; i++) idx[i] = i;
  auto v = std::make_shared<std::vector<double>>();
  int rc;
  if (kind == kDsShieldedPot) {
    v->resize(npart);
    rc = bit_direct_potential(st->sys, alpha, 0.0, npart, idx.data(), v->data());
    for (double& x : *v) x = -x;  // shielded_pot_particle_ds returns -phi
  } else {
    v->resize(3 * (size_t)npart);
    rc = kind == kDsReg ? bit_direct_field(st->sys, DEL, KERNEL_ORDER, npart, idx.data(), v->data())
                        : bit_direct_field(st->sys, 0.0, 0, npart, idx.data(), v->data());
  }
  if (rc != BIT_OK) {
    cout << "[" << fn << "] " << bit_last_error() << endl;
    return nullptr;
  }
  st->results.emplace(key, v);
  return v;
}

void particle_field_ds(const char* fn, int kind, int ind, int npart, PARTICLE* part,
                       double* xfor, double* yfor, double* zfor) {
  if (ind < 0 || ind >= npart) {
    cout << "[" << fn << "] particle index " << ind << " outside [0, " << npart << ")" << endl;
    return;
  }
  auto v = direct_all(fn, part, npart, kind, 0.0);
  if (!v) return;
  *xfor = (*v)[3 * (size_t)ind];
  *yfor = (*v)[3 * (size_t)ind + 1];
  *zfor = (*v)[3 * (size_t)ind + 2];
}

void point_field_ds(const char* fn, bool reg, double x, double y, double z, int npart,
                    PARTICLE* part, double* xfor, double* yfor, double* zfor) {
  if (npart < 1) {
    *xfor = *yfor = *zfor = 0.0;
    return;
  }
  std::unique_lock<std::mutex> lk(g_mu);
  std::shared_ptr<DirectState> st = direct_state(fn, part, npart, lk);
  if (!st) return;
  double E[3];
  int rc = reg ? bit_direct_field_points(st->sys, DEL, KERNEL_ORDER, 1, &x, &y, &z, E)
               : bit_direct_field_points(st->sys, 0.0, 0, 1, &x, &y, &z, E);
  if (rc != BIT_OK) {
    cout << "[" << fn << "] " << bit_last_error() << endl;
    return;
  }
  *xfor = E[0];
  *yfor = E[1];
  *zfor = E[2];
}

void all_fields_ds(const char* fn, int kind, int npart, PARTICLE* part) {
  if (npart < 1) return;
  auto v = direct_all(fn, part, npart, kind, 0.0);
  if (!v) return;
  for (int i = 0; i < npart; i++) {  // the reference's product order
    part[i].xforce = (*v)[3 * (size_t)i] * charge_constant * part[i].tot_charge;
    part[i].yforce = (*v)[3 * (size_t)i + 1] * charge_constant * part[i].tot_charge;
    part[i].zforce = (*v)[3 * (size_t)i + 2] * charge_constant * part[i].tot_charge;
  }
}

}  // namespace

void eval_grad_kernel_noreg(double dx, double dy, double dz, double rs, double* xfor,
                            double* yfor, double* zfor) {
  const double t = 1 / (rs * sqrt(rs));
  *xfor = -dx * t;
  *yfor = -dy * t;
  *zfor = -dz * t;
}

void eval_grad_kernel_reg(double dx, double dy, double dz, double rs, double* xfor,
                          double* yfor, double* zfor) {
  // the regularized family's factor chain, in the reference's operation order
  // (tree3d.cpp:80-91)
  const double r = sqrt(rs);
  double t = 1.0 / (r * rs), f = t;
  for (int l = 1; l <= KERNEL_ORDER; l++) {
    f = f * (l + 0.5) / l * DEL / rs;
    t += f;
  }
  *xfor = -dx * t;
  *yfor = -dy * t;
  *zfor = -dz * t;
}

void force_point_ds_noreg(double x, double y, double z, int npart, PARTICLE* part,
                          double* xfor, double* yfor, double* zfor) {
  point_field_ds("force_point_ds_noreg", false, x, y, z, npart, part, xfor, yfor, zfor);
}

void force_point_ds_reg(double x, double y, double z, int npart, PARTICLE* part,
                        double* xfor, double* yfor, double* zfor) {
  point_field_ds("force_point_ds_reg", true, x, y, z, npart, part, xfor, yfor, zfor);
}

void force_particle_ds_noreg(int ind, int npart, PARTICLE* part, double* xfor, double* yfor,
                             double* zfor) {
  particle_field_ds("force_particle_ds_noreg", kDsNoreg, ind, npart, part, xfor, yfor, zfor);
}

void force_particle_ds_reg(int ind, int npart, PARTICLE* part, double* xfor, double* yfor,
                           double* zfor) {
  particle_field_ds("force_particle_ds_reg", kDsReg, ind, npart, part, xfor, yfor, zfor);
}

void freespace_particles_ds_noreg(int npart, PARTICLE* part) {
  all_fields_ds("freespace_particles_ds_noreg", kDsNoreg, npart, part);
}

void freespace_particles_ds_reg(int npart, PARTICLE* part) {
  all_fields_ds("freespace_particles_ds_reg", kDsReg, npart, part);
}

void shielded_pot_particle_ds(int ind, int npart, PARTICLE* part, double* pot, double alpha) {
  if (ind < 0 || ind >= npart) {
    cout << "[shielded_pot_particle_ds] particle index " << ind << " outside [0, " << npart
         << ")" << endl;
    return;
  }
  auto v = direct_all("shielded_pot_particle_ds", part, npart, kDsShieldedPot, alpha);
  if (v) *pot = (*v)[ind];
}

void shielded_freespace_pot_ds(int npart, PARTICLE* part, double alpha) {
  if (npart < 1) return;
  auto v = direct_all("shielded_freespace_pot_ds", part, npart, kDsShieldedPot, alpha);
  if (!v) return;
  for (int i = 0; i < npart; i++) part[i].xforce = (*v)[i] * charge_constant * part[i].tot_charge;
}

// ---------------------------------------------------------------------------
// far-field primitives: one cluster's Taylor sum at one offset
// ---------------------------------------------------------------------------
namespace {

// sum_{|k| < nterms} M_k a_k(d) for G(R) = exp(-kappa R)/R, R^2 = dist, with
// a_k = (-1)^|k| / k! d^k G the Taylor coefficients about the cluster centre
// (moment[i][j][k] the reference's raw moments).  Radial-derivative form:
// with g_m = (R^-1 d/dR)^m G,
//   a_k = (-1)^|k| sum_{2l <= k} d^(k-2l) g_(|k|-|l|) / (l! (k-2l)! 2^|l|)
// (chain rule for G(sqrt(2S)), S = dist/2, one axis at a time), and
//   R^2 g_(m+1) = kappa^2 g_(m-1) - (2m+1) g_m,  g_-1 = -exp(-kappa R)/kappa
// (for kappa = 0 the first term vanishes).
double cluster_sum(int nterms, double dx, double dy, double dz, double dist, double*** moment,
                   double kappa) {
  if (nterms < 1) return 0.0;
  const int p = nterms - 1;
  const double R = sqrt(dist), e = exp(-kappa * R);
  std::vector<double> g(nterms + 1);
  g[0] = e / R;
  if (p >= 1) g[1] = -e * (kappa * R + 1.0) / (R * dist);
  for (int m = 1; m + 1 <= p; m++) g[m + 1] = (kappa * kappa * g[m - 1] - (2 * m + 1) * g[m]) / dist;
  // powers d^a / a! and 1 / (b! 2^b) per axis
  std::vector<double> px(nterms), py(nterms), pz(nterms), h(nterms);
  px[0] = py[0] = pz[0] = h[0] = 1.0;
  for (int a = 1; a < nterms; a++) {
    px[a] = px[a - 1] * dx / a;
    py[a] = py[a - 1] * dy / a;
    pz[a] = pz[a - 1] * dz / a;
    h[a] = h[a - 1] / (2.0 * a);
  }
  double sum = 0.0;
  for (int n = 0; n <= p; n++)
    for (int i = n; i >= 0; i--)
      for (int j = n - i; j >= 0; j--) {
        const int k = n - i - j;
        double a = 0.0;
        for (int l0 = 0; 2 * l0 <= i; l0++)
          for (int l1 = 0; 2 * l1 <= j; l1++)
            for (int l2 = 0; 2 * l2 <= k; l2++)
              a += px[i - 2 * l0] * py[j - 2 * l1] * pz[k - 2 * l2] * h[l0] * h[l1] * h[l2] *
                   g[n - l0 - l1 - l2];
        sum += moment[i][j][k] * ((n & 1) ? -a : a);
      }
  return sum;
}

// Taylor coefficients T_k(d) = (1/k!) d^k G of the regularized family
// G_L(R) = sum_{l <= L} w_l DEL^l R^-(2l+1), R^2 = rs, w_l = prod_{j<=l} (j-1/2)/j
// (L = KERNEL_ORDER; tree3d.cpp:614-791), for |k| <= deg, in t[i][j][k]
// (i + j + k <= deg).  Radial derivatives of each term:
//   (R^-1 d/dR)^m R^-(2l+1) = (-1)^m prod_{i<m} (2l+2i+1) R^-(2l+2m+1).
void family_coeffs(int deg, double dx, double dy, double dz, double rs,
                   std::vector<double>& t) {
  const double R = sqrt(rs);
  std::vector<double> g(deg + 1, 0.0);
  double w = 1.0;  // w_l DEL^l
  for (int l = 0; l <= KERNEL_ORDER; l++) {
    if (l > 0) w *= (l - 0.5) / l * DEL;
    double term = w / pow(R, 2 * l + 1);  // (R^-1 d/dR)^m R^-(2l+1) for m = 0, 1, ...
    for (int m = 0; m <= deg; m++) {
      g[m] += term;
      term *= -(2.0 * l + 2.0 * m + 1.0) / rs;
    }
  }
  std::vector<double> px(deg + 1), py(deg + 1), pz(deg + 1), h(deg + 1);
  px[0] = py[0] = pz[0] = h[0] = 1.0;
  for (int a = 1; a <= deg; a++) {
    px[a] = px[a - 1] * dx / a;
    py[a] = py[a - 1] * dy / a;
    pz[a] = pz[a - 1] * dz / a;
    h[a] = h[a - 1] / (2.0 * a);
  }
  const int n1 = deg + 1;
  t.assign((size_t)n1 * n1 * n1, 0.0);
  for (int i = 0; i <= deg; i++)
    for (int j = 0; i + j <= deg; j++)
      for (int k = 0; i + j + k <= deg; k++) {
        double a = 0.0;
        for (int l0 = 0; 2 * l0 <= i; l0++)
          for (int l1 = 0; 2 * l1 <= j; l1++)
            for (int l2 = 0; 2 * l2 <= k; l2++)
              a += px[i - 2 * l0] * py[j - 2 * l1] * pz[k - 2 * l2] * h[l0] * h[l1] * h[l2] *
                   g[i + j + k - l0 - l1 - l2];
        t[((size_t)i * n1 + j) * n1 + k] = a;
      }
}

}  // namespace

// One cluster's field at offset (dx, dy, dz) = target - centre, rs = |d|^2 + DEL
// (tree3d.cpp:456-810): E_e += sum_{|k| < nterms} (-1)^(|k|+1) (k_e + 1) M_k
// T_{k+e}(d), the same sum as the reference's, with its coefficients from the
// family's radial derivatives instead of the reference's recurrences.
void treetaylorforce(int nterms, double dx, double dy, double dz, double rs, double* xfor,
                     double* yfor, double* zfor, double*** moment) {
  if (nterms < 1) return;
  std::vector<double> t;
  family_coeffs(nterms, dx, dy, dz, rs, t);
  const int n1 = nterms + 1;
  auto T = [&](int i, int j, int k) { return t[((size_t)i * n1 + j) * n1 + k]; };
  for (int m = 0; m < nterms; m++)
    for (int k = 0; k < nterms - m; k++)
      for (int i = 0; i < nterms - m - k; i++) {
        const double s = ((i + k + m + 1) % 2 == 0) ? 1.0 : -1.0;
        const double M = moment[i][k][m];
        *xfor += s * M * T(i + 1, k, m) * (i + 1);
        *yfor += s * M * T(i, k + 1, m) * (k + 1);
        *zfor += s * M * T(i, k, m + 1) * (m + 1);
      }
}

void shielded_treetaylorpot(int nterms, double dx, double dy, double dz, double dist,
                            double* pot, double*** moment, double kappa) {
  *pot -= cluster_sum(nterms, dx, dy, dz, dist, moment, kappa);
}

void treetaylorpotential(int nterms, double dx, double dy, double dz, double rs, double* pot,
                         double*** moment) {
  *pot += cluster_sum(nterms, dx, dy, dz, rs, moment, 0.0);
}

// ---------------------------------------------------------------------------
// particle set-up (the reference's test-data generators)
// ---------------------------------------------------------------------------
double RandL(void) { return rand() / ((double)RAND_MAX); }

void boxparticles(int npart, double dlength, double dheight, double ddepth, PARTICLE* part,
                  double density) {
  const double weight = density * dlength * dheight * ddepth / npart;
  srand(100);
  const double hx = dlength / 2, hy = dheight / 2, hz = ddepth / 2;
  for (int i = 0; i < npart; i++) {
    part[i].charge = -1.0;
    part[i].weight = weight;
    part[i].x = hx / 2 + RandL() * hx;
    part[i].y = hy / 2 + RandL() * hy;
    part[i].z = hz / 2 + RandL() * hz;
    part[i].u = (RandL() - 0.5) * 1e6;
    part[i].v = (RandL() - 0.5) * 1e6;
    part[i].w = (RandL() - 0.5) * 1e6;
    part[i].tot_charge = part[i].charge * part[i].weight * elec;
  }
}

int bounding_box(int npart, PARTICLE* part, double& Mx, double& mx, double& My, double& my,
                 double& Mz, double& mz) {
  mx = Mx = part[0].x;
  my = My = part[0].y;
  mz = Mz = part[0].z;
  for (int i = 0; i < npart; i++) {
    mx = std::min(mx, part[i].x);
    Mx = std::max(Mx, part[i].x);
    my = std::min(my, part[i].y);
    My = std::max(My, part[i].y);
    mz = std::min(mz, part[i].z);
    Mz = std::max(Mz, part[i].z);
  }
  return 0;
}

double sq(double x) { return x * x; }
