// bit_api.cpp -- implementation of the C-ABI in include/bit_amd.h.
//
// Host side of one system: the tree is built on the host (bit-exact with the
// reference treemake), the sources are sorted into the tree's slot order and
// uploaded once, and every evaluation runs on the context's stream:
//   moments (cached per nterms) -> batch lists (cached per MAC parameters)
//   -> far field -> near field + scatter to the caller's order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <functional>
#include <map>
#include <atomic>
#include <chrono>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bit_amd.h"
#include "gmres.h"
#include "kernels.h"
#include "tree_build.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(BIT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

#define BIT_TRY(expr)         \
  do {                        \
    int r_ = (expr);          \
    if (r_ != BIT_OK) return r_; \
  } while (0)

constexpr int kMomChunk = 4096;

int ncoef(int nt) { return nt * (nt + 1) * (nt + 2) / 6; }

}  // namespace

struct bit_tree {
  bit::HostTree tree;
};

struct bit_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  size_t budget = 0;
  std::atomic<size_t> in_use{0};  // systems of one context may be used from several threads
  std::atomic<int> inject{0};
  int host_threads = 0;           // tree-build threads, 0 = all
  std::vector<std::pair<void*, size_t>> user_bufs;  // bit_device_alloc
  std::atomic<int> live_systems{0};
  // Device blocks released by the library's own buffers, kept for reuse (a
  // system's buffers are freed and requested again at every rebuild).  Every
  // API call synchronises its stream before returning, so a released block has
  // no work in flight.  Guarded by pool_mu.
  std::mutex pool_mu;
  std::multimap<size_t, void*> pool;
  size_t pooled = 0;
  // Per-batch list staging capacities from the context's last list build,
  // [0] particles as targets, [1] external points; far then near.  A hint
  // only: a batch that outgrows them is rebuilt by count-then-fill.
  std::atomic<int> list_cap[2][2] = {{{0}, {0}}, {{0}, {0}}};
  // this device's scratch of bit_eval_potential_noalloc (a static device
  // array shared by every context on the device)
  void* scratch = nullptr;
  // device tree builds: each generation's result, written by the device into
  // pinned coherent host memory and waited for by polling (guarded by gen_mu)
  std::mutex gen_mu;
  bit::dev::GenResult* gen_host = nullptr;
  // ... and the node table's download: pinned staging, a side stream (default
  // priority) and the event it waits for
  void* pin = nullptr;
  size_t pin_bytes = 0;
  hipStream_t side = nullptr;
  hipEvent_t built = nullptr;
  // an evaluation's moments run on the side stream, concurrently with its list
  // build (fork: the tree is done; join: the moments are)
  std::mutex fork_mu;
  hipEvent_t fork = nullptr, join = nullptr;
  hipEvent_t lists_done = nullptr;  // near lists and counts ready (far ordering may follow)
  // persistent launches of mixed evaluations: their work counters (two, used
  // in turn; serve_mu orders the launches on the context's stream)
  std::mutex serve_mu;
  void* serve_ctr = nullptr;
  int serve_parity = 0;
};

namespace {

constexpr size_t kPoolGranule = 1 << 16;       // block sizes are multiples of 64 KiB
constexpr size_t kPoolMax = (size_t)8 << 30;    // bytes kept cached at most

size_t pool_round(size_t n) { return (n + kPoolGranule - 1) / kPoolGranule * kPoolGranule; }

// Frees cached blocks, largest first, until at most `keep` bytes stay cached.
void pool_trim(bit_ctx* c, size_t keep) {
  while (c->pooled > keep && !c->pool.empty()) {
    auto it = std::prev(c->pool.end());
    (void)hipFree(it->second);
    c->pooled -= it->first;
    c->pool.erase(it);
  }
}

// A block of at least `need` bytes (size returned in `got`): a cached block of
// the same rounded size class (up to 2x) if there is one, else a new one.
hipError_t pool_get(bit_ctx* c, size_t need, void** p, size_t* got) {
  const size_t want = pool_round(need);
  {
    std::lock_guard<std::mutex> lk(c->pool_mu);
    auto it = c->pool.lower_bound(want);
    if (it != c->pool.end() && it->first <= 2 * want) {
      *p = it->second;
      *got = it->first;
      c->pooled -= it->first;
      c->pool.erase(it);
      return hipSuccess;
    }
    if (c->budget && c->in_use + c->pooled + want > c->budget) pool_trim(c, 0);
  }
  hipError_t e = hipMalloc(p, want);
  if (e != hipSuccess) {  // cached blocks may be what fills the device
    (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> lk(c->pool_mu);
      pool_trim(c, 0);
    }
    e = hipMalloc(p, want);
  }
  if (e == hipSuccess) *got = want;
  return e;
}

void pool_put(bit_ctx* c, void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(c->pool_mu);
  c->pool.emplace(bytes, p);
  c->pooled += bytes;
  pool_trim(c, kPoolMax);
}

}  // namespace

namespace {

// A device allocation accounted against the context's HBM budget.
struct DevBuf {
  bit_ctx* ctx = nullptr;
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      ctx->in_use -= bytes;
      pool_put(ctx, p, bytes);
    }
    p = nullptr;
    bytes = 0;
  }
  // Grows (never shrinks) to at least `need` bytes; contents are not kept.
  int ensure(bit_ctx* c, size_t need) {
    ctx = c;
    if (need <= bytes && p) return BIT_OK;
    release();
    if (need == 0) need = 8;
    if (c->inject.load() > 0 && c->inject.fetch_sub(1) > 0)
      return fail(BIT_E_OOM, "injected device allocation failure");
    if (c->budget && c->in_use + need > c->budget)
      return fail(BIT_E_OOM, "HBM budget exceeded: need " + std::to_string(need) +
                                 " bytes, in use " + std::to_string(c->in_use) + " of " +
                                 std::to_string(c->budget));
    size_t got = 0;
    hipError_t e = pool_get(c, need, &p, &got);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(BIT_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bytes = got;
    c->in_use += got;
    return BIT_OK;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  void start(hipStream_t s) {
    if (!a) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
    }
    (void)hipEventRecord(a, s);
  }
  void stop(hipStream_t s) { (void)hipEventRecord(b, s); }
  double ms() {
    float v = 0.f;
    if (a && b && hipEventElapsedTime(&v, a, b) == hipSuccess) return v;
    return 0.0;
  }
};

}  // namespace

// Interaction lists of one target set.
struct ListSet {
  bool valid = false;
  double acc = 0, delta2 = 0, sub = 0;
  int root = 0;
  int nfar = 0, nnear = 0;
  int passes = 0;  // traversals of the last build
  DevBuf far_cnt, near_cnt, far_off, near_off, far_items, near_items;
};

struct bit_system {
  bit_ctx* ctx = nullptr;
  bit::HostTree tree;
  int stack_cap = 0;
  int nleaves = 0;
  bool view_ready = false;  // node arrays and sloc written by the device build
  // tree + sources
  DevBuf center, child, range, scale, perm, leaves, sx, sy, sz, sq, sloc;
  // moments
  int mom_nterms = 0;
  int nchunks = 0;
  DevBuf chunk_node, chunk_rng, node_chunk0, node_nchunk, partial, mom, mom_hat64, qc32, qc64;
  int qc64_nterms = 0;  // nterms of qc64, 0 = not computed for the current moments
  // field coefficients Q^(e) (FP32, FP64) and the nterms they were made for
  DevBuf qf32, qf64;
  int qf_nterms[2] = {0, 0};
  // internal nodes grouped by height (leaves 0): up_off[h-1]..up_off[h] in up_nodes
  DevBuf up_nodes;
  std::vector<int> up_off;
  // lists: pl for the particles as targets (cached per MAC parameters), ql for
  // the last external point set
  ListSet pl, ql;
  DevBuf scan_tmp, lstats, pstat, err_flag, stage_far, stage_near;
  bit::dev::ListStats hstats{};
  // the last list build's pair statistics (4 totals on the device), copied
  // into stats when they are asked for
  DevBuf pair_dev;
  mutable bool pair_pending = false;
  // external point targets (batch order) and their results
  DevBuf qx, qy, qz, qperm, qphi, qout, qdesc;
  bit::dev::EvalDesc qhdesc{};
  // outputs
  // per-target partial results (potential: the near field's; field: the far field's)
  DevBuf part, out;
  // a deferred prepare_eval left the moments on the side stream (join_moments)
  bool join_pending = false;
  // direct sum (particle targets tslot; point targets dpx, dpy, dpz)
  DevBuf tslot, dpart, dout, dpx, dpy, dpz;
  // charge updates
  DevBuf qin;
  // evaluation descriptor (host copy kept alive for the async upload)
  bit::dev::EvalDesc hdesc{};
  DevBuf desc;
  bit_stats stats{};
  Timer t_mom, t_lists, t_far, t_near;

  bit::dev::TreeView view() const {
    bit::dev::TreeView v;
    v.n = tree.n;
    v.numtree = tree.numtree;
    v.center = center.as<double4>();
    v.child = child.as<int>();
    v.range = range.as<int2>();
    v.scale = scale.as<double>();
    v.sx = sx.as<double>();
    v.sy = sy.as<double>();
    v.sz = sz.as<double>();
    v.sq = sq.as<double>();
    v.sloc = sloc.as<float4>();
    v.perm = perm.as<int>();
    v.nt = tree.n;
    v.self = 1;
    v.tx = v.sx;
    v.ty = v.sy;
    v.tz = v.sz;
    v.tperm = v.perm;
    return v;
  }
};

namespace {

int upload(bit_ctx* c, DevBuf& b, const void* host, size_t bytes) {
  BIT_TRY(b.ensure(c, bytes));
  HIP_TRY(hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, c->stream));
  return BIT_OK;
}

bool finite_all(int n, const double* a) {
  for (int i = 0; i < n; i++)
    if (!std::isfinite(a[i])) return false;
  return true;
}

// The moments' work layout of a tree: leaves in chunks of kMomChunk slots
// (chunk node, slot range; per node its first chunk and count) and the internal
// nodes grouped by height (leaves 0): up_off[h-1]..up_off[h] in up.  A child's
// index exceeds its parent's, so one backward sweep gives every height.
struct MomLayout {
  std::vector<int> cn, cr, c0, ck, up, off;
};

void mom_layout(const bit::HostTree& T, MomLayout& M) {
  const int nt = T.numtree;
  std::vector<int> height(nt, 0);
  M.c0.assign(nt, 0);
  M.ck.assign(nt, 0);
  int hmax = 0;
  for (int i = nt - 1; i >= 0; i--)
    for (int q = 0; q < 8; q++) {
      const int ch = T.children[(size_t)i * 8 + q];
      if (ch >= 0) height[i] = std::max(height[i], height[ch] + 1);
    }
  for (int i = 0; i < nt; i++) {
    M.c0[i] = (int)M.cn.size();
    int k = 0;
    if (height[i] == 0) {
      const int b = T.begin[i], e = T.end[i];
      for (int x = b; x < e; x += kMomChunk, k++) {
        M.cn.push_back(i);
        M.cr.push_back(x);
        M.cr.push_back(std::min(e, x + kMomChunk));
      }
    }
    M.ck[i] = k;
    hmax = std::max(hmax, height[i]);
  }
  M.off.assign(1, 0);
  std::vector<int> count(hmax + 1, 0);
  for (int i = 0; i < nt; i++) count[height[i]]++;
  std::vector<int> at(hmax + 1, 0);
  for (int h = 1; h <= hmax; h++) {
    at[h] = M.off.back();
    M.off.push_back(M.off.back() + count[h]);
  }
  M.up.assign(M.off.back(), 0);
  for (int i = 0; i < nt; i++)
    if (height[i] > 0) M.up[at[height[i]]++] = i;
}

// Generations of the device build: particles that can never be separated
// (coincident, or more than maxmemb in one FP64 point) add a generation each.
// Each generation halves the largest extent of every node it splits, and an
// FP64 extent halves at most about 2,100 times, so separable inputs need far
// fewer than this many generations.
constexpr int kMaxGenerations = 8192;

// Builds the tree of n particles whose coordinates are in HBM (original order)
// and leaves the system's sorted sources, permutation and node table on the
// device.  Every split decision is taken on the device (split programs, one
// thread per node, kernels.hip); the host launches a generation's rounds and
// reads back only the generation's size, then the finished node table.
int build_on_device(bit_ctx* ctx, bit_system* s, int n, const double* dx, const double* dy,
                    const double* dz, const double* dq, int maxmemb) {
  hipStream_t st = ctx->stream;
  const size_t nb = (size_t)n * sizeof(double);
  DevBuf tP, tX, tY, tZ;
  BIT_TRY(s->perm.ensure(ctx, (size_t)n * sizeof(int)));
  BIT_TRY(s->sx.ensure(ctx, nb));
  BIT_TRY(s->sy.ensure(ctx, nb));
  BIT_TRY(s->sz.ensure(ctx, nb));
  BIT_TRY(tP.ensure(ctx, (size_t)n * sizeof(int)));
  BIT_TRY(tX.ensure(ctx, nb));
  BIT_TRY(tY.ensure(ctx, nb));
  BIT_TRY(tZ.ensure(ctx, nb));
  HIP_TRY(hipMemcpyAsync(s->sx.p, dx, nb, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->sy.p, dy, nb, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->sz.p, dz, nb, hipMemcpyDeviceToDevice, st));
  HIP_TRY(bit::dev::iota(n, s->perm.as<int>(), st));
  bit::dev::SlotArrays a{s->perm.as<int>(), s->sx.as<double>(), s->sy.as<double>(),
                         s->sz.as<double>(), tP.as<int>(), tX.as<double>(), tY.as<double>(),
                         tZ.as<double>()};
  // node table, grown between generations (a generation of cnt nodes makes at
  // most 8 cnt children)
  int cap = 0;
  DevBuf nbox, nchild, nbegin, nend, ndepth, npar;
  auto node_view = [&]() {
    return bit::dev::BuildNodes{nbox.as<double>(), nchild.as<int>(), nbegin.as<int>(),
                                nend.as<int>(), ndepth.as<int>(), npar.as<int>()};
  };
  auto grow = [&](int need, int used) -> int {
    if (need <= cap) return BIT_OK;
    const int ncap = std::max(need, 2 * cap);
    DevBuf b2, c2, g2, e2, d2, p2;
    BIT_TRY(b2.ensure(ctx, (size_t)ncap * 13 * sizeof(double)));
    BIT_TRY(c2.ensure(ctx, (size_t)ncap * 8 * sizeof(int)));
    BIT_TRY(g2.ensure(ctx, (size_t)ncap * sizeof(int)));
    BIT_TRY(e2.ensure(ctx, (size_t)ncap * sizeof(int)));
    BIT_TRY(d2.ensure(ctx, (size_t)ncap * sizeof(int)));
    BIT_TRY(p2.ensure(ctx, (size_t)ncap * sizeof(int)));
    if (used > 0) {
      HIP_TRY(hipMemcpyAsync(b2.p, nbox.p, (size_t)used * 13 * sizeof(double), hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(c2.p, nchild.p, (size_t)used * 8 * sizeof(int), hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(g2.p, nbegin.p, (size_t)used * sizeof(int), hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(e2.p, nend.p, (size_t)used * sizeof(int), hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(d2.p, ndepth.p, (size_t)used * sizeof(int), hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipMemcpyAsync(p2.p, npar.p, (size_t)used * sizeof(int), hipMemcpyDeviceToDevice, st));
      HIP_TRY(hipStreamSynchronize(st));  // the old blocks go back to the cache
    }
    std::swap(nbox.p, b2.p); std::swap(nbox.bytes, b2.bytes);
    std::swap(nchild.p, c2.p); std::swap(nchild.bytes, c2.bytes);
    std::swap(nbegin.p, g2.p); std::swap(nbegin.bytes, g2.bytes);
    std::swap(nend.p, e2.p); std::swap(nend.bytes, e2.bytes);
    std::swap(ndepth.p, d2.p); std::swap(ndepth.bytes, d2.bytes);
    std::swap(npar.p, p2.p); std::swap(npar.bytes, p2.bytes);
    nbox.ctx = nchild.ctx = nbegin.ctx = nend.ctx = ndepth.ctx = npar.ctx = ctx;
    cap = ncap;
    return BIT_OK;
  };
  BIT_TRY(grow(std::max(1024, 8 * (n / std::max(1, maxmemb)) + 64), 0));
  // root bounding box (tree3d.cpp:1385-1403) and root node, on the device
  const int mblocks = std::max(1, std::min(256, (n + 255) / 256));
  DevBuf part;
  BIT_TRY(part.ensure(ctx, (size_t)mblocks * 6 * sizeof(double)));
  HIP_TRY(bit::dev::minmax(n, s->sx.as<double>(), s->sy.as<double>(), s->sz.as<double>(),
                           mblocks, part.as<double>(), st));
  // generations: the programs of one, and those of the next (started as its
  // children are written)
  const double lim = 1.0 / std::sqrt(2.0);  // aspectlimit, tree3d.cpp:1371
  const long long node_cap = 4LL * n + 64;
  DevBuf progs, next, cres, nkid, start;
  // round layouts, alternating between generations: generation g runs from
  // lay[g & 1], written while generation g - 1 ends
  struct LayoutBufs {
    DevBuf ops, nch, cbeg, cprog, ntot;
    // sized for at most cnt programs
    int ensure(bit_ctx* c, int n, int cnt) {
      BIT_TRY(ops.ensure(c, (size_t)2 * cnt * sizeof(bit::dev::PartOpDev)));
      BIT_TRY(nch.ensure(c, (size_t)2 * cnt * sizeof(int)));
      BIT_TRY(cbeg.ensure(c, (size_t)(cnt + 1) * sizeof(int)));
      BIT_TRY(cprog.ensure(
          c, ((size_t)(n + bit::dev::kPartChunk - 1) / bit::dev::kPartChunk + cnt) * sizeof(int)));
      BIT_TRY(ntot.ensure(c, 2 * sizeof(int)));
      return BIT_OK;
    }
    bit::dev::RoundLayout view() {
      return bit::dev::RoundLayout{ops.as<bit::dev::PartOpDev>(), nch.as<int>(), cbeg.as<int>(),
                                   cprog.as<int>(), ntot.as<int>()};
    }
  } lay[2];
  BIT_TRY(progs.ensure(ctx, sizeof(bit::dev::DevProg)));
  BIT_TRY(lay[0].ensure(ctx, n, 1));
  HIP_TRY(bit::dev::build_root(n, mblocks, part.as<double>(), node_view(),
                               progs.as<bit::dev::DevProg>(), maxmemb, lim, lay[0].view(), st));
  std::lock_guard<std::mutex> gen_lock(ctx->gen_mu);
  if (!ctx->gen_host)
    HIP_TRY(hipHostMalloc((void**)&ctx->gen_host, sizeof(bit::dev::GenResult),
                          hipHostMallocCoherent));
  bit::dev::GenResult* hr = ctx->gen_host;
  // Waits for generation `token`'s result: polls the pinned word, and the
  // stream now and then (a finished stream without the token is an error).
  auto wait_gen = [&](int token) -> int {
    for (long it = 0;; it++) {
      if (__atomic_load_n(&hr->seq, __ATOMIC_ACQUIRE) == token) return BIT_OK;
      if ((it & 1023) == 1023) {
        const hipError_t q = hipStreamQuery(st);
        if (q == hipSuccess) {
          if (__atomic_load_n(&hr->seq, __ATOMIC_ACQUIRE) == token) return BIT_OK;
          return fail(BIT_E_HIP, "tree generation finished without its result");
        }
        if (q != hipErrorNotReady) return fail(BIT_E_HIP, hipGetErrorString(q));
      }
    }
  };
  bit::dev::GenResult gr{};
  int gen_lo = 0, gen_hi = 1, depth = 0, maxdepth = 0, maxsize = n;
  bool work = n > maxmemb;
  while (work) {
    const int cnt = gen_hi - gen_lo;
    if (++depth > kMaxGenerations)
      return fail(BIT_E_TREE, "tree deeper than " + std::to_string(kMaxGenerations) +
                                  " levels: particles cannot be separated into leaves of maxmemb");
    BIT_TRY(grow(gen_hi + 8 * cnt, gen_hi));
    const int maxch = (n + bit::dev::kPartChunk - 1) / bit::dev::kPartChunk + cnt;
    BIT_TRY(next.ensure(ctx, (size_t)8 * cnt * sizeof(bit::dev::DevProg)));
    LayoutBufs& cur = lay[(depth - 1) & 1];
    LayoutBufs& nxt = lay[depth & 1];
    BIT_TRY(nxt.ensure(ctx, n, 8 * cnt));  // free: generation g - 1 has ended
    BIT_TRY(nkid.ensure(ctx, (size_t)cnt * sizeof(int)));
    BIT_TRY(start.ensure(ctx, (size_t)cnt * sizeof(int)));
    BIT_TRY(cres.ensure(ctx, (size_t)maxch * sizeof(bit::dev::ChunkResult)));
    bit::dev::BuildScratch w{progs.as<bit::dev::DevProg>(), next.as<bit::dev::DevProg>(),
                             cur.view(),
                             nxt.view(),
                             cres.as<bit::dev::ChunkResult>(),
                             nkid.as<int>(),
                             start.as<int>(),
                             hr};
    __atomic_store_n(&hr->seq, 0, __ATOMIC_RELEASE);
    HIP_TRY(bit::dev::build_generation(a, n, gen_lo, gen_hi, depth, maxmemb, lim, maxsize,
                                       node_view(), w, st));
    // The generation's last kernel (k_prog_write) may still run: buffers
    // resized below are only rewritten by kernels queued after it.
    BIT_TRY(wait_gen(depth));
    gr.next = hr->next;
    gr.failed = hr->failed;
    gr.maxdepth = hr->maxdepth;
    gr.splits = hr->splits;
    gr.maxsize = hr->maxsize;
    std::swap(progs.p, next.p);
    std::swap(progs.bytes, next.bytes);
    if (gr.failed) return fail(BIT_E_TREE, "tree split overflow (more than 8 children)");
    if (gr.next > node_cap)
      return fail(BIT_E_TREE,
                  "tree exceeds 4n+64 nodes: particles cannot be separated into leaves of maxmemb");
    maxdepth = std::max(maxdepth, gr.maxdepth);
    gen_lo = gen_hi;
    gen_hi = gr.next;
    work = gr.splits > 0;
    maxsize = gr.maxsize;
  }
  // The node table goes to the host on a side stream (pinned staging) while
  // the device sorts the leaves (Morton order, same keys as the host build,
  // straight from the node table), gathers the charges and writes the
  // evaluation's node arrays.
  const int nt = gen_hi;
  const size_t off_child = (size_t)nt * bit::BOX_FIELDS * sizeof(double);
  const size_t off_int = off_child + (size_t)nt * 8 * sizeof(int);  // begin, end, depth, par
  const size_t pin_need = off_int + (size_t)4 * nt * sizeof(int);
  if (!ctx->side) HIP_TRY(hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
  if (!ctx->built) HIP_TRY(hipEventCreateWithFlags(&ctx->built, hipEventDisableTiming));
  if (ctx->pin_bytes < pin_need) {
    if (ctx->pin) HIP_TRY(hipHostFree(ctx->pin));
    ctx->pin = nullptr;
    ctx->pin_bytes = 0;
    HIP_TRY(hipHostMalloc(&ctx->pin, pin_need, hipHostMallocDefault));
    ctx->pin_bytes = pin_need;
  }
  char* pin = static_cast<char*>(ctx->pin);
  int* pint = reinterpret_cast<int*>(pin + off_int);
  HIP_TRY(hipEventRecord(ctx->built, st));
  HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->built, 0));
  HIP_TRY(hipMemcpyAsync(pin, nbox.p, off_child, hipMemcpyDeviceToHost, ctx->side));
  HIP_TRY(hipMemcpyAsync(pin + off_child, nchild.p, (size_t)nt * 8 * sizeof(int),
                         hipMemcpyDeviceToHost, ctx->side));
  HIP_TRY(hipMemcpyAsync(pint, nbegin.p, nt * sizeof(int), hipMemcpyDeviceToHost, ctx->side));
  HIP_TRY(hipMemcpyAsync(pint + nt, nend.p, nt * sizeof(int), hipMemcpyDeviceToHost, ctx->side));
  HIP_TRY(hipMemcpyAsync(pint + 2 * nt, ndepth.p, nt * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->side));
  HIP_TRY(hipMemcpyAsync(pint + 3 * nt, npar.p, nt * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->side));
  HIP_TRY(bit::dev::leaf_morton(a, nt, node_view(), st));
  // charges in slot order, the node arrays and the FP32 leaf-local copy
  BIT_TRY(s->sq.ensure(ctx, nb));
  BIT_TRY(s->sloc.ensure(ctx, (size_t)n * sizeof(float4)));
  BIT_TRY(s->center.ensure(ctx, (size_t)nt * 4 * sizeof(double)));
  BIT_TRY(s->scale.ensure(ctx, (size_t)nt * sizeof(double)));
  BIT_TRY(s->range.ensure(ctx, (size_t)nt * 2 * sizeof(int)));
  BIT_TRY(s->child.ensure(ctx, (size_t)nt * 8 * sizeof(int)));
  s->tree.n = n;
  s->tree.numtree = nt;
  HIP_TRY(bit::dev::set_charges(s->view(), dq, s->sq.as<double>(), s->sloc.as<float4>(), st));
  HIP_TRY(bit::dev::tree_view(nt, node_view(), s->view(), s->sloc.as<float4>(), st));
  // the host's copy of the node table, from the staging, behind the device work
  HIP_TRY(hipStreamSynchronize(ctx->side));
  bit::HostTree& T = s->tree;
  T = bit::HostTree();
  T.n = n;
  T.numtree = nt;
  T.maxdepth = maxdepth;
  T.box.assign(reinterpret_cast<const double*>(pin),
               reinterpret_cast<const double*>(pin) + (size_t)nt * bit::BOX_FIELDS);
  const int* pch = reinterpret_cast<const int*>(pin + off_child);
  T.children.assign(pch, pch + (size_t)nt * 8);
  T.begin.assign(pint, pint + nt);
  T.end.assign(pint + nt, pint + 2 * nt);
  T.depth.assign(pint + 2 * nt, pint + 3 * nt);
  T.par.assign(pint + 3 * nt, pint + 4 * nt);
  int nleaves = 0;
  for (int i = 0; i < nt; i++) nleaves += T.children[(size_t)i * 8] < 0;
  s->nleaves = nleaves;
  s->stack_cap = 8 * (T.maxdepth + 2) + 8;
  s->view_ready = true;
  // the moments' work layout, also behind the device work: the host tables
  // staged in the pinned buffer (free again) and copied on the stream
  MomLayout M;
  mom_layout(T, M);
  const std::vector<int>* tab[5] = {&M.cn, &M.cr, &M.c0, &M.ck, &M.up};
  DevBuf* dst[5] = {&s->chunk_node, &s->chunk_rng, &s->node_chunk0, &s->node_nchunk,
                    &s->up_nodes};
  size_t need = 0;
  for (const auto* v : tab) need += std::max<size_t>(1, v->size()) * sizeof(int);
  if (ctx->pin_bytes < need) {
    if (ctx->pin) HIP_TRY(hipHostFree(ctx->pin));
    ctx->pin = nullptr;
    ctx->pin_bytes = 0;
    HIP_TRY(hipHostMalloc(&ctx->pin, need, hipHostMallocDefault));
    ctx->pin_bytes = need;
  }
  char* at = static_cast<char*>(ctx->pin);
  for (int k = 0; k < 5; k++) {
    const size_t bytes = tab[k]->size() * sizeof(int);
    BIT_TRY(dst[k]->ensure(ctx, std::max<size_t>(1, tab[k]->size()) * sizeof(int)));
    if (bytes) {
      std::memcpy(at, tab[k]->data(), bytes);
      HIP_TRY(hipMemcpyAsync(dst[k]->p, at, bytes, hipMemcpyHostToDevice, st));
    }
    at += std::max<size_t>(1, tab[k]->size()) * sizeof(int);
  }
  s->up_off = M.off;
  s->nchunks = (int)M.cn.size();
  HIP_TRY(hipStreamSynchronize(st));  // the node table, staging and scratch die here
  return BIT_OK;
}

// Host inputs of up to this many particles are built on the host (the same
// split program and Morton leaf order, so the tree is identical) and uploaded
// sorted; larger ones are uploaded and built on the device.  Measured on 16-blob
// sets (host build on two threads / upload + device build): 5,000 particles
// 0.48 / 0.57 ms, 20,000 1.7 / 0.95 ms, 200,000 15.1 / 2.0 ms, 500,000 41 / 3.1 ms.
constexpr int kHostBuildMax = 8192;

int build_on_host(bit_ctx* ctx, bit_system* s, int n, const double* x, const double* y,
                  const double* z, const double* dq, int maxmemb) {
  std::string err;
  if (!bit::build_tree(n, x, y, z, maxmemb, s->tree, err, ctx->host_threads))
    return fail(BIT_E_TREE, err);
  const std::vector<int>& P = s->tree.perm;
  std::vector<double> a(n), b(n), c(n);
  for (int i = 0; i < n; i++) {
    a[i] = x[P[i]];
    b[i] = y[P[i]];
    c[i] = z[P[i]];
  }
  const size_t nb = (size_t)n * sizeof(double);
  BIT_TRY(upload(ctx, s->perm, P.data(), (size_t)n * sizeof(int)));
  BIT_TRY(upload(ctx, s->sx, a.data(), nb));
  BIT_TRY(upload(ctx, s->sy, b.data(), nb));
  BIT_TRY(upload(ctx, s->sz, c.data(), nb));
  BIT_TRY(s->sq.ensure(ctx, nb));
  BIT_TRY(s->sloc.ensure(ctx, (size_t)n * sizeof(float4)));
  HIP_TRY(bit::dev::set_charges(s->view(), dq, s->sq.as<double>(), s->sloc.as<float4>(),
                                ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));  // the sorted host copies die here
  return BIT_OK;
}

// Host copy of the permutation (downloaded on first use).
int ensure_host_perm(const bit_system* cs) {
  bit_system* s = const_cast<bit_system*>(cs);
  if (!s->tree.perm.empty()) return BIT_OK;
  s->tree.perm.resize(s->tree.n);
  HIP_TRY(hipMemcpyAsync(s->tree.perm.data(), s->perm.p, (size_t)s->tree.n * sizeof(int),
                         hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return BIT_OK;
}

// Upload the node table and create the system; sorted sources come from
// `sorted` (host, 4 arrays) or are gathered on the device from d_src.
int finish_system(bit_ctx* ctx, bit_system* s) {
  if (s->view_ready) return BIT_OK;  // written by the device build
  const bit::HostTree& T = s->tree;
  const int nt = T.numtree;
  std::vector<double> cen(4 * (size_t)nt), scl(nt);
  std::vector<int> rng(2 * (size_t)nt), lv;
  for (int i = 0; i < nt; i++) {
    const double* f = &T.box[(size_t)i * bit::BOX_FIELDS];
    cen[4 * i + 0] = f[bit::BXMID];
    cen[4 * i + 1] = f[bit::BYMID];
    cen[4 * i + 2] = f[bit::BZMID];
    cen[4 * i + 3] = f[bit::BSQR];
    scl[i] = f[bit::BSQR] > 0 ? std::sqrt(f[bit::BSQR]) : 1.0;
    rng[2 * i] = T.begin[i];
    rng[2 * i + 1] = T.end[i];
    if (T.children[(size_t)i * 8] < 0) lv.push_back(i);
  }
  s->nleaves = (int)lv.size();
  s->stack_cap = 8 * (T.maxdepth + 2) + 8;
  BIT_TRY(upload(ctx, s->center, cen.data(), cen.size() * sizeof(double)));
  BIT_TRY(upload(ctx, s->scale, scl.data(), scl.size() * sizeof(double)));
  BIT_TRY(upload(ctx, s->range, rng.data(), rng.size() * sizeof(int)));
  BIT_TRY(upload(ctx, s->child, T.children.data(), T.children.size() * sizeof(int)));
  BIT_TRY(upload(ctx, s->leaves, lv.data(), std::max<size_t>(1, lv.size()) * sizeof(int)));
  HIP_TRY(bit::dev::leaf_local(s->view(), s->leaves.as<int>(), s->nleaves, s->sloc.as<float4>(),
                               ctx->stream));
  // the node tables above are host temporaries of this function
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return BIT_OK;
}

// field: the regularized family (kernel_order >= 1) applies to the field only,
// as KERNEL_ORDER does in the reference (tree3d.cpp:88, 614).
int check_params(const bit_system* s, const bit_params* p, bool field = false) {
  if (!s || !p) return fail(BIT_E_INVALID, "null system or params");
  if (p->kernel_order < 0 || p->kernel_order > BIT_MAX_KERNEL_ORDER)
    return fail(BIT_E_INVALID,
                "kernel_order must be in [0, " + std::to_string(BIT_MAX_KERNEL_ORDER) + "]");
  if (!field && p->kernel_order != 0)
    return fail(BIT_E_INVALID, "kernel_order >= 1 (regularized family) applies to the field only");
  if (p->nterms < 1 || p->nterms > BIT_MAX_NTERMS)
    return fail(BIT_E_INVALID, "nterms must be in [1, " + std::to_string(BIT_MAX_NTERMS) + "]");
  if (!std::isfinite(p->acc) || !std::isfinite(p->kappa) || !std::isfinite(p->delta2) ||
      !std::isfinite(p->mac_sub) || p->delta2 < 0 || p->kappa < 0)
    return fail(BIT_E_INVALID, "acc/kappa/delta2/mac_sub must be finite, kappa,delta2 >= 0");
  if (p->precision != BIT_FP64 && p->precision != BIT_FP32)
    return fail(BIT_E_INVALID, "precision must be BIT_FP64 or BIT_FP32");
  if (p->root < 0 || p->root >= s->tree.numtree)
    return fail(BIT_E_INVALID, "root must be a node of the tree: [0, " +
                                   std::to_string(s->tree.numtree) + ")");
  return BIT_OK;
}

// Lists built for these MAC parameters and start node.
bool lists_match(const ListSet& L, const bit_params* p) {
  return L.valid && L.acc == p->acc && L.delta2 == p->delta2 && L.sub == p->mac_sub &&
         L.root == p->root;
}

// Staging capacity for a batch's next build: the largest count seen plus an
// eighth, or 0 (two traversals) when the staging regions would exceed four
// times the entries the lists actually hold.
static int stage_cap(unsigned long long mx, int nb, long long total) {
  const long long cap = (long long)mx + (long long)(mx / 8) + 4;
  if (cap > INT_MAX / 2 || cap * nb > 4 * total + 64LL * nb) return 0;
  return (int)cap;
}

// One traversal when the capacities from the context's last build of this
// kind hold every batch (count and stage, exclusive scan, compaction); two
// otherwise (count, exclusive scan, fill).  Both give the same lists.  Each
// batch's far items are then put in the far kernel's pairing order, with the
// pair statistics: on far_stream when given (after an event that marks the
// rest of the build; the caller joins far_stream before the far field), else
// on the context's stream.
int build_lists(bit_system* s, ListSet& L, const bit::dev::TreeView& tv, const bit_params* p,
                hipStream_t far_stream = nullptr) {
  L.valid = false;  // set again only when both passes were issued
  bit_ctx* c = s->ctx;
  const int nb = (tv.nt + bit::dev::kBatch - 1) / bit::dev::kBatch;
  bit::dev::MacParams mp{p->acc, p->delta2, p->mac_sub, p->root};
  BIT_TRY(L.far_cnt.ensure(c, (size_t)(nb + 1) * sizeof(int)));
  BIT_TRY(L.near_cnt.ensure(c, (size_t)(nb + 1) * sizeof(int)));
  BIT_TRY(L.far_off.ensure(c, (size_t)(nb + 1) * sizeof(int)));
  BIT_TRY(L.near_off.ensure(c, (size_t)(nb + 1) * sizeof(int)));
  BIT_TRY(s->lstats.ensure(c, sizeof(bit::dev::ListStats)));
  BIT_TRY(s->pstat.ensure(c, (size_t)nb * 4 * sizeof(unsigned long long)));
  BIT_TRY(s->err_flag.ensure(c, sizeof(int)));
  HIP_TRY(hipMemsetAsync(L.far_cnt.p, 0, (size_t)(nb + 1) * sizeof(int), c->stream));
  HIP_TRY(hipMemsetAsync(L.near_cnt.p, 0, (size_t)(nb + 1) * sizeof(int), c->stream));
  HIP_TRY(hipMemsetAsync(s->err_flag.p, 0, sizeof(int), c->stream));
  const int kind = &L == &s->ql ? 1 : 0;  // the context's staging hint for these lists
  const int cap_far = c->list_cap[kind][0].load();
  const int cap_near = c->list_cap[kind][1].load();
  bool staged = cap_far > 0 && cap_near > 0;
  if (staged && c->budget) {  // staging is optional: never the allocation that breaks a budget
    const size_t nf = (size_t)nb * cap_far * sizeof(bit::dev::Item);
    const size_t nn = (size_t)nb * cap_near * sizeof(bit::dev::Item);
    const size_t extra = (nf > s->stage_far.bytes ? nf : 0) + (nn > s->stage_near.bytes ? nn : 0);
    staged = c->in_use + extra <= c->budget;
  }
  if (staged) {
    BIT_TRY(s->stage_far.ensure(c, (size_t)nb * cap_far * sizeof(bit::dev::Item)));
    BIT_TRY(s->stage_near.ensure(c, (size_t)nb * cap_near * sizeof(bit::dev::Item)));
  }
  s->t_lists.start(c->stream);
  if (staged)
    HIP_TRY(bit::dev::lists_stage(tv, mp, s->stack_cap, L.far_cnt.as<int>(),
                                  L.near_cnt.as<int>(), cap_far, cap_near,
                                  s->stage_far.as<bit::dev::Item>(),
                                  s->stage_near.as<bit::dev::Item>(), s->err_flag.as<int>(),
                                  c->stream));
  else
    HIP_TRY(bit::dev::lists_count(tv, mp, s->stack_cap, L.far_cnt.as<int>(),
                                  L.near_cnt.as<int>(), s->err_flag.as<int>(), c->stream));
  HIP_TRY(bit::dev::list_stats(nb, L.far_cnt.as<int>(), L.near_cnt.as<int>(),
                               s->lstats.as<bit::dev::ListStats>(), c->stream));
  size_t tmp = 0;
  HIP_TRY(bit::dev::exclusive_scan(L.far_cnt.as<int>(), L.far_off.as<int>(), nb + 1, nullptr, tmp,
                                   c->stream));
  BIT_TRY(s->scan_tmp.ensure(c, tmp));
  tmp = s->scan_tmp.bytes;
  HIP_TRY(bit::dev::exclusive_scan(L.far_cnt.as<int>(), L.far_off.as<int>(), nb + 1, s->scan_tmp.p,
                                   tmp, c->stream));
  tmp = s->scan_tmp.bytes;
  HIP_TRY(bit::dev::exclusive_scan(L.near_cnt.as<int>(), L.near_off.as<int>(), nb + 1,
                                   s->scan_tmp.p, tmp, c->stream));
  int tot[2] = {0, 0};
  int eflag = 0;
  HIP_TRY(hipMemcpyAsync(&tot[0], L.far_off.as<int>() + nb, sizeof(int), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(&tot[1], L.near_off.as<int>() + nb, sizeof(int), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(&s->hstats, s->lstats.p, sizeof(bit::dev::ListStats),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(&eflag, s->err_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (eflag) return fail(BIT_E_STATE, "traversal stack overflow");
  L.nfar = tot[0];
  L.nnear = tot[1];
  BIT_TRY(L.far_items.ensure(c, (size_t)std::max(1, tot[0]) * sizeof(bit::dev::Item)));
  BIT_TRY(L.near_items.ensure(c, (size_t)std::max(1, tot[1]) * sizeof(bit::dev::Item)));
  const bool fits = staged && s->hstats.far_max <= (unsigned long long)cap_far &&
                    s->hstats.near_max <= (unsigned long long)cap_near;
  // near items from the staging regions (or in place after the second
  // traversal) with their pair statistics; far items into the pairing order,
  // from the staging regions or in place
  if (!fits)
    HIP_TRY(bit::dev::lists_fill(tv, mp, s->stack_cap, L.far_off.as<int>(),
                                 L.near_off.as<int>(), L.far_items.as<bit::dev::Item>(),
                                 L.near_items.as<bit::dev::Item>(), s->err_flag.as<int>(),
                                 c->stream));
  HIP_TRY(bit::dev::near_compact(nb, L.near_off.as<int>(), fits ? cap_near : 0,
                                 fits ? s->stage_near.as<bit::dev::Item>() : nullptr,
                                 L.near_items.as<bit::dev::Item>(), tv.range,
                                 s->pstat.as<unsigned long long>(), c->stream));
  BIT_TRY(s->pair_dev.ensure(c, 4 * sizeof(unsigned long long)));
  hipStream_t fs = c->stream;
  if (far_stream) {
    s->t_lists.stop(c->stream);
    if (!c->lists_done) HIP_TRY(hipEventCreateWithFlags(&c->lists_done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(c->lists_done, c->stream));
    HIP_TRY(hipStreamWaitEvent(far_stream, c->lists_done, 0));
    fs = far_stream;
  }
  HIP_TRY(bit::dev::far_order(nb, L.far_off.as<int>(), fits ? cap_far : 0,
                              fits ? s->stage_far.as<bit::dev::Item>()
                                   : L.far_items.as<bit::dev::Item>(),
                              L.far_items.as<bit::dev::Item>(),
                              s->pstat.as<unsigned long long>(), fs));
  if (!far_stream) s->t_lists.stop(c->stream);
  HIP_TRY(bit::dev::pair_stats(nb, s->pstat.as<unsigned long long>(),
                               s->pair_dev.as<unsigned long long>(), fs));
  s->pair_pending = true;
  L.passes = fits ? 1 : 2;
  const long long total = (long long)tot[0] + tot[1];
  c->list_cap[kind][0].store(stage_cap(s->hstats.far_max, nb, total));
  c->list_cap[kind][1].store(stage_cap(s->hstats.near_max, nb, total));
  L.valid = true;
  L.acc = p->acc;
  L.delta2 = p->delta2;
  L.sub = p->mac_sub;
  L.root = p->root;
  s->stats.far_items = tot[0];
  s->stats.near_items = tot[1];
  s->stats.ms_lists = -1.0;  // filled after the next synchronisation
  s->stats.list_passes = L.passes;
  return BIT_OK;
}

// Lists of the particles as targets, cached per MAC parameters and start
// node; far_stream as build_lists.
int ensure_lists(bit_system* s, const bit_params* p, hipStream_t far_stream = nullptr) {
  ListSet& L = s->pl;
  if (lists_match(L, p)) return BIT_OK;
  return build_lists(s, L, s->view(), p, far_stream);
}

}  // namespace

extern "C" {

void bit_params_default(bit_params* p) {
  p->nterms = 6;
  p->acc = 0.5;
  p->kappa = 0.0;
  p->delta2 = 0.0;
  p->mac_sub = 0.0;
  p->precision = BIT_FP64;
  p->kernel_order = 0;
  p->root = 0;
}

const char* bit_last_error(void) { return g_err.c_str(); }

int bit_ctx_create(int device, size_t hbm_budget_bytes, bit_ctx** out) {
  if (!out) return fail(BIT_E_INVALID, "null out");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev)
    return fail(BIT_E_INVALID, "device " + std::to_string(device) + " not present (" +
                                   std::to_string(ndev) + " visible)");
  HIP_TRY(hipSetDevice(device));
  bit_ctx* c = new bit_ctx();
  c->device = device;
  c->budget = hbm_budget_bytes;
  hipError_t e = bit::dev::noalloc_scratch(device, &c->scratch);
  if (e != hipSuccess) {
    delete c;
    return fail(BIT_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(BIT_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  c->stream = c->own;
  *out = c;
  return BIT_OK;
}

int bit_ctx_destroy(bit_ctx* c) {
  if (!c) return BIT_OK;
  if (c->live_systems > 0)
    return fail(BIT_E_STATE, std::to_string(c->live_systems) +
                                 " system(s) still alive: destroy them before their context");
  (void)hipStreamSynchronize(c->stream);
  for (auto& b : c->user_bufs) (void)hipFree(b.first);
  {
    std::lock_guard<std::mutex> lk(c->pool_mu);
    pool_trim(c, 0);
  }
  if (c->own) (void)hipStreamDestroy(c->own);
  if (c->gen_host) (void)hipHostFree(c->gen_host);
  if (c->pin) (void)hipHostFree(c->pin);
  if (c->built) (void)hipEventDestroy(c->built);
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  if (c->lists_done) (void)hipEventDestroy(c->lists_done);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->serve_ctr) (void)hipFree(c->serve_ctr);
  delete c;
  return BIT_OK;
}

int bit_ctx_set_stream(bit_ctx* c, void* stream) {
  if (!c) return fail(BIT_E_INVALID, "null context");
  // the persistent launches' counters are handed from launch to launch in
  // stream order: the old stream's launches finish before the new one's (every
  // call already ends synchronised; this covers calls from other threads, and a
  // caller's stream destroyed before the switch is not an error)
  std::lock_guard<std::mutex> lk(c->serve_mu);
  if (c->serve_ctr) (void)hipStreamSynchronize(c->stream);
  c->stream = stream ? static_cast<hipStream_t>(stream) : c->own;
  return BIT_OK;
}

int bit_ctx_sync(bit_ctx* c) {
  if (!c) return fail(BIT_E_INVALID, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

int bit_valu_ceiling(bit_ctx* c, double* fp32_tflops, double* fp64_tflops) {
  if (!c || !fp32_tflops || !fp64_tflops) return fail(BIT_E_INVALID, "null argument");
  HIP_TRY(hipSetDevice(c->device));
  int ncu = 0;
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
  const int blocks = 8 * ncu, iters = 20000;  // eight waves per SIMD
  DevBuf out;
  BIT_TRY(out.ensure(c, (size_t)blocks * 256 * sizeof(double)));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  double res[2] = {0.0, 0.0};
  hipError_t err = hipSuccess;
  for (int fp64 = 0; fp64 < 2 && err == hipSuccess; fp64++)
    for (int rep = 0; rep < 2 && err == hipSuccess; rep++) {  // the first run warms up
      err = hipEventRecord(e0, c->stream);
      if (err == hipSuccess) err = bit::dev::fma_ceiling(blocks, iters, fp64, out.p, c->stream);
      if (err == hipSuccess) err = hipEventRecord(e1, c->stream);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (err == hipSuccess && ms > 0)  // 8 FMAs = 16 flops per thread and step
        res[fp64] = 16.0 * iters * (double)blocks * 256 / (ms * 1e-3) / 1e12;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != hipSuccess) return fail(BIT_E_HIP, hipGetErrorString(err));
  *fp32_tflops = res[0];
  *fp64_tflops = res[1];
  return BIT_OK;
}

size_t bit_ctx_bytes_in_use(const bit_ctx* c) { return c ? c->in_use.load() : 0; }

int bit_ctx_set_host_threads(bit_ctx* c, int threads) {
  if (!c || threads < 0) return fail(BIT_E_INVALID, "bad context or thread count");
  c->host_threads = threads;
  return BIT_OK;
}

int bit_ctx_inject_alloc_failures(bit_ctx* c, int k) {
  if (!c) return fail(BIT_E_INVALID, "null context");
  c->inject.store(k < 0 ? 0 : k);
  return BIT_OK;
}

int bit_device_alloc(bit_ctx* c, size_t bytes, void** out) {
  if (!c || !out) return fail(BIT_E_INVALID, "null context or out");
  *out = nullptr;
  HIP_TRY(hipSetDevice(c->device));
  if (bytes == 0) bytes = 8;
  if (c->budget && c->in_use + bytes > c->budget)
    return fail(BIT_E_OOM, "HBM budget exceeded by a device buffer of " + std::to_string(bytes));
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(BIT_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  c->in_use += bytes;
  c->user_bufs.emplace_back(p, bytes);
  *out = p;
  return BIT_OK;
}

int bit_device_free(bit_ctx* c, void* p) {
  if (!c) return fail(BIT_E_INVALID, "null context");
  if (!p) return BIT_OK;
  for (size_t i = 0; i < c->user_bufs.size(); i++) {
    if (c->user_bufs[i].first == p) {
      HIP_TRY(hipStreamSynchronize(c->stream));
      HIP_TRY(hipFree(p));
      c->in_use -= c->user_bufs[i].second;
      c->user_bufs.erase(c->user_bufs.begin() + i);
      return BIT_OK;
    }
  }
  return fail(BIT_E_INVALID, "pointer was not allocated by bit_device_alloc on this context");
}

int bit_copy_to_device(bit_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return fail(BIT_E_INVALID, "null argument");
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

int bit_copy_to_host(bit_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return fail(BIT_E_INVALID, "null argument");
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

int bit_copy_on_device(bit_ctx* c, void* dst, const void* src, size_t bytes) {
  if (!c || (!dst && bytes) || (!src && bytes)) return fail(BIT_E_INVALID, "null argument");
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

static int create_common(bit_ctx* ctx, int n, const double* x, const double* y, const double* z,
                         const double* q, int maxmemb, bit_system** out, bool device_src) {
  if (!ctx || !out) return fail(BIT_E_INVALID, "null context or out");
  *out = nullptr;
  if (n < 1) return fail(BIT_E_INVALID, "n must be >= 1");
  if (maxmemb < 1) return fail(BIT_E_INVALID, "maxmemb must be >= 1");
  if (!x || !y || !z || !q) return fail(BIT_E_INVALID, "null coordinate or charge array");
  HIP_TRY(hipSetDevice(ctx->device));
  bit_system* s = new bit_system();
  s->ctx = ctx;
  auto t0 = std::chrono::steady_clock::now();
  const double *dx = x, *dy = y, *dz = z, *dq = q;
  DevBuf ux, uy, uz, uq;  // host inputs are staged in HBM first
  int rc = BIT_OK;
  if (!device_src) {
    if (!finite_all(n, x) || !finite_all(n, y) || !finite_all(n, z) || !finite_all(n, q)) {
      delete s;
      return fail(BIT_E_INVALID, "non-finite particle coordinate or charge");
    }
    const size_t nb = (size_t)n * sizeof(double);
    if ((rc = upload(ctx, uq, q, nb)) == BIT_OK) dq = uq.as<double>();
    if (rc == BIT_OK && n > kHostBuildMax &&  // the device build reads the coordinates in HBM
        (rc = upload(ctx, ux, x, nb)) == BIT_OK && (rc = upload(ctx, uy, y, nb)) == BIT_OK &&
        (rc = upload(ctx, uz, z, nb)) == BIT_OK) {
      dx = ux.as<double>();
      dy = uy.as<double>();
      dz = uz.as<double>();
    }
  }
  if (rc == BIT_OK && device_src) {  // inputs from HBM are checked on the device
    DevBuf flag;
    int h = 0;
    if ((rc = flag.ensure(ctx, sizeof(int))) == BIT_OK) {
      hipError_t e = hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream);
      if (e == hipSuccess) e = bit::dev::nonfinite(n, dx, dy, dz, dq, flag.as<int>(), ctx->stream);
      if (e == hipSuccess) e = hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) rc = fail(BIT_E_HIP, hipGetErrorString(e));
      else if (h) rc = fail(BIT_E_INVALID, "non-finite particle coordinate or charge");
    }
  }
  if (rc == BIT_OK) {
    if (!device_src && n <= kHostBuildMax)
      rc = build_on_host(ctx, s, n, x, y, z, dq, maxmemb);
    else
      rc = build_on_device(ctx, s, n, dx, dy, dz, dq, maxmemb);
  }
  auto t1 = std::chrono::steady_clock::now();
  s->stats.ms_tree = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (rc == BIT_OK) rc = finish_system(ctx, s);
  if (rc != BIT_OK) {
    (void)hipStreamSynchronize(ctx->stream);  // no copy may still read host staging
    delete s;
    return rc;
  }
  auto t2 = std::chrono::steady_clock::now();
  s->stats.ms_upload = std::chrono::duration<double, std::milli>(t2 - t1).count();
  ctx->live_systems++;
  *out = s;
  return BIT_OK;
}

int bit_system_create(bit_ctx* ctx, int n, const double* x, const double* y, const double* z,
                      const double* q, int maxmemb, bit_system** out) {
  return create_common(ctx, n, x, y, z, q, maxmemb, out, false);
}

int bit_system_create_device(bit_ctx* ctx, int n, const double* x, const double* y,
                             const double* z, const double* q, int maxmemb, bit_system** out) {
  return create_common(ctx, n, x, y, z, q, maxmemb, out, true);
}

int bit_system_destroy(bit_system* s) {
  if (!s) return BIT_OK;
  (void)hipStreamSynchronize(s->ctx->stream);
  s->ctx->live_systems--;
  delete s;
  return BIT_OK;
}

int bit_system_numtree(const bit_system* s) { return s ? s->tree.numtree : 0; }
int bit_system_maxdepth(const bit_system* s) { return s ? s->tree.maxdepth : 0; }

static void export_host_tree(const bit::HostTree& T, double* boxes, int* children, int* begin,
                             int* end, int* perm) {
  if (boxes) std::memcpy(boxes, T.box.data(), T.box.size() * sizeof(double));
  if (children) std::memcpy(children, T.children.data(), T.children.size() * sizeof(int));
  if (begin) std::memcpy(begin, T.begin.data(), T.begin.size() * sizeof(int));
  if (end) std::memcpy(end, T.end.data(), T.end.size() * sizeof(int));
  if (perm) std::memcpy(perm, T.perm.data(), T.perm.size() * sizeof(int));
}

int bit_tree_build(int n, const double* x, const double* y, const double* z, int maxmemb,
                   bit_tree** out) {
  if (!out) return fail(BIT_E_INVALID, "null out");
  *out = nullptr;
  if (n < 1 || !x || !y || !z) return fail(BIT_E_INVALID, "need n >= 1 and coordinates");
  if (!finite_all(n, x) || !finite_all(n, y) || !finite_all(n, z))
    return fail(BIT_E_INVALID, "non-finite particle coordinate");
  bit_tree* t = new bit_tree();
  std::string err;
  if (!bit::build_tree(n, x, y, z, maxmemb, t->tree, err)) {
    delete t;
    return fail(BIT_E_TREE, err);
  }
  *out = t;
  return BIT_OK;
}

int bit_tree_destroy(bit_tree* t) {
  delete t;
  return BIT_OK;
}

int bit_tree_numtree(const bit_tree* t) { return t ? t->tree.numtree : 0; }
int bit_tree_maxdepth(const bit_tree* t) { return t ? t->tree.maxdepth : 0; }

int bit_tree_export(const bit_tree* t, double* boxes, int* children, int* begin, int* end,
                    int* perm) {
  if (!t) return fail(BIT_E_INVALID, "null tree");
  export_host_tree(t->tree, boxes, children, begin, end, perm);
  return BIT_OK;
}

int bit_system_export_tree(const bit_system* s, double* boxes, int* children, int* begin,
                           int* end, int* perm) {
  if (!s) return fail(BIT_E_INVALID, "null system");
  if (perm) BIT_TRY(ensure_host_perm(s));
  export_host_tree(s->tree, boxes, children, begin, end, perm);
  return BIT_OK;
}

}  // extern "C"

// Moments of the system's current charges, launched on `st` (the chunk tables
// are made on the first call, on the context's stream).
static int moments_on(bit_system* s, int nterms, hipStream_t st) {
  if (!s) return fail(BIT_E_INVALID, "null system");
  if (nterms < 1 || nterms > BIT_MAX_NTERMS)
    return fail(BIT_E_INVALID, "nterms must be in [1, " + std::to_string(BIT_MAX_NTERMS) + "]");
  if (s->mom_nterms == nterms) return BIT_OK;
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const bit::HostTree& T = s->tree;
  const int nt = T.numtree, nc = ncoef(nterms);
  // State changes only after every allocation and upload succeeded: a failed
  // call (e.g. BIT_E_OOM) leaves the system as it was, never half-initialised.
  s->mom_nterms = 0;
  s->qc64_nterms = 0;
  s->qf_nterms[0] = s->qf_nterms[1] = 0;
  if (s->nchunks == 0) {
    MomLayout M;
    mom_layout(T, M);
    BIT_TRY(upload(c, s->chunk_node, M.cn.data(), M.cn.size() * sizeof(int)));
    BIT_TRY(upload(c, s->chunk_rng, M.cr.data(), M.cr.size() * sizeof(int)));
    BIT_TRY(upload(c, s->node_chunk0, M.c0.data(), M.c0.size() * sizeof(int)));
    BIT_TRY(upload(c, s->node_nchunk, M.ck.data(), M.ck.size() * sizeof(int)));
    if (!M.up.empty()) BIT_TRY(upload(c, s->up_nodes, M.up.data(), M.up.size() * sizeof(int)));
    HIP_TRY(hipStreamSynchronize(c->stream));  // host tables die at scope end
    s->up_off = M.off;
    s->nchunks = (int)M.cn.size();
  }
  BIT_TRY(s->partial.ensure(c, (size_t)s->nchunks * nc * sizeof(double)));
  BIT_TRY(s->mom.ensure(c, (size_t)nt * nc * sizeof(double)));
  BIT_TRY(s->mom_hat64.ensure(c, (size_t)nt * nc * sizeof(double)));
  BIT_TRY(s->qc32.ensure(c, (size_t)nt * bit::dev::nqpad(nterms) * sizeof(float)));
  s->t_mom.start(st);
  HIP_TRY(bit::dev::moments_partial(s->view(), nterms, s->nchunks, s->chunk_node.as<int>(),
                                    s->chunk_rng.as<int2>(), s->partial.as<double>(), st));
  HIP_TRY(bit::dev::moments_reduce(s->view(), nterms, s->node_chunk0.as<int>(),
                                   s->node_nchunk.as<int>(), s->partial.as<double>(),
                                   s->mom.as<double>(), s->mom_hat64.as<double>(), st));
  for (size_t h = 1; h < s->up_off.size(); h++)
    HIP_TRY(bit::dev::moments_upward(s->view(), nterms, s->up_off[h] - s->up_off[h - 1],
                                     s->up_nodes.as<int>() + s->up_off[h - 1],
                                     s->mom.as<double>(), s->mom_hat64.as<double>(), st));
  HIP_TRY(bit::dev::far_coeffs32(nt, nterms, s->mom_hat64.as<double>(), s->qc32.as<float>(), st));
  s->t_mom.stop(st);
  s->mom_nterms = nterms;
  s->stats.ms_moments = -1.0;
  return BIT_OK;
}

extern "C" {

int bit_moments(bit_system* s, int nterms) {
  if (!s) return fail(BIT_E_INVALID, "null system");
  return moments_on(s, nterms, s->ctx->stream);
}

}  // extern "C"

// FP64 cluster coefficients for the current moments, computed on first use.
static int ensure_qc64(bit_system* s, int nterms, hipStream_t st = nullptr) {
  if (s->qc64_nterms == nterms) return BIT_OK;
  bit_ctx* c = s->ctx;
  const int nt = s->tree.numtree;
  BIT_TRY(s->qc64.ensure(c, (size_t)nt * bit::dev::nqpad(nterms) * sizeof(double)));
  HIP_TRY(bit::dev::far_coeffs64(nt, nterms, s->mom_hat64.as<double>(), s->qc64.as<double>(),
                                 st ? st : c->stream));
  s->qc64_nterms = nterms;
  return BIT_OK;
}

// Field coefficients Q^(e) for the current moments, computed on first use
// (index 0: FP32, 1: FP64).
static int ensure_qf(bit_system* s, int nterms, bool fp32, hipStream_t st = nullptr) {
  const int k = fp32 ? 0 : 1;
  if (s->qf_nterms[k] == nterms) return BIT_OK;
  bit_ctx* c = s->ctx;
  if (!st) st = c->stream;
  const int nt = s->tree.numtree;
  const size_t per = (size_t)3 * bit::dev::nqpad(nterms + 1);
  if (fp32) {
    BIT_TRY(s->qf32.ensure(c, (size_t)nt * per * sizeof(float)));
    HIP_TRY(bit::dev::field_coeffs32(nt, nterms, s->mom_hat64.as<double>(), s->qf32.as<float>(),
                                     st));
  } else {
    BIT_TRY(s->qf64.ensure(c, (size_t)nt * per * sizeof(double)));
    HIP_TRY(bit::dev::field_coeffs64(nt, nterms, s->mom_hat64.as<double>(), s->qf64.as<double>(),
                                     st));
  }
  s->qf_nterms[k] = nterms;
  return BIT_OK;
}

extern "C" {

int bit_moments_download(const bit_system* s, double* out) {
  if (!s || !out) return fail(BIT_E_INVALID, "null system or output");
  if (s->mom_nterms == 0) return fail(BIT_E_STATE, "moments not computed");
  const int nt = s->tree.numtree, p = s->mom_nterms, nc = ncoef(p);
  std::vector<double> packed((size_t)nt * nc);
  HIP_TRY(hipMemcpyAsync(packed.data(), s->mom.p, packed.size() * sizeof(double),
                         hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  const size_t cube = (size_t)p * p * p;
  for (int t = 0; t < nt; t++) {
    double* M = out + cube * t;
    std::fill(M, M + cube, 0.0);
    for (int n = 0; n < p; n++)
      for (int m = 0; m <= n; m++)
        for (int k = 0; k <= m; k++) {
          int i = n - m, j = m - k;
          int c = n * (n + 1) * (n + 2) / 6 + m * (m + 1) / 2 + k;
          M[((size_t)i * p + j) * p + k] = packed[(size_t)t * nc + c];
        }
  }
  return BIT_OK;
}

static void collect_times(bit_system* s) {
  if (s->stats.ms_moments < 0) s->stats.ms_moments = s->t_mom.ms();
  if (s->stats.ms_lists < 0) s->stats.ms_lists = s->t_lists.ms();
  s->stats.ms_far = s->t_far.ms();
  s->stats.ms_near = s->t_near.ms();
}

// The moments for params (and, after them on the same stream, the cluster
// coefficients `coeffs` launches) and the lists.  When both have to be made,
// the moments run on the side stream while the lists (bound by scalar issue)
// build on the context's stream, and the lists' far ordering follows the
// moments on the side stream; the context's stream then waits for the side
// stream, or, with defer_join, leaves that to join_moments (the caller runs
// work that needs only the near lists in between).
static int moments_and_lists(bit_system* s, const bit_params* p, bool defer_join,
                             const std::function<int(hipStream_t)>& coeffs) {
  bit_ctx* c = s->ctx;
  const ListSet& L = s->pl;
  const bool lists_ok = lists_match(L, p);
  if (s->mom_nterms != p->nterms && !lists_ok && s->nchunks > 0) {
    std::lock_guard<std::mutex> lk(c->fork_mu);
    if (!c->side) HIP_TRY(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
    if (!c->fork) HIP_TRY(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
    if (!c->join) HIP_TRY(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(c->fork, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->side, c->fork, 0));
    int rc = moments_on(s, p->nterms, c->side);
    if (rc == BIT_OK) rc = coeffs(c->side);
    // the lists' far ordering follows on the side stream (it overlaps the near
    // field, which needs only the near lists)
    if (rc == BIT_OK) rc = ensure_lists(s, p, c->side);
    HIP_TRY(hipEventRecord(c->join, c->side));
    if (defer_join && rc == BIT_OK) s->join_pending = true;
    else HIP_TRY(hipStreamWaitEvent(c->stream, c->join, 0));
    return rc;
  }
  BIT_TRY(bit_moments(s, p->nterms));
  BIT_TRY(coeffs(c->stream));
  return ensure_lists(s, p);
}

// Moments and lists for params, the near field's per-target buffer, and the
// evaluation descriptor (result to `dst`, a device pointer); nothing is
// launched for the near/far kernels.  defer_join: the moments may still run on
// the side stream when this returns (join_moments before the far field).
static int prepare_eval(bit_system* s, const bit_params* p, double* dst,
                        bool defer_join = false) {
  BIT_TRY(check_params(s, p));
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int n = s->tree.n;
  // FP64's cluster coefficients follow the moments on their stream (FP32's are
  // made with the moments)
  BIT_TRY(moments_and_lists(s, p, defer_join, [&](hipStream_t st) {
    return p->precision == BIT_FP64 ? ensure_qc64(s, p->nterms, st) : BIT_OK;
  }));
  BIT_TRY(s->part.ensure(c, (size_t)n * sizeof(double)));
  if (!dst) {
    BIT_TRY(s->out.ensure(c, (size_t)n * sizeof(double)));
    dst = s->out.as<double>();
  }
  bit::dev::EvalDesc& d = s->hdesc;
  d.tv = s->view();
  d.kappa = p->kappa;
  d.delta2 = p->delta2;
  d.qc32 = s->qc32.as<float>();
  d.qc64 = s->qc64.as<double>();
  d.far_off = s->pl.far_off.as<int>();
  d.far_items = s->pl.far_items.as<bit::dev::Item>();
  d.near_off = s->pl.near_off.as<int>();
  d.near_items = s->pl.near_items.as<bit::dev::Item>();
  d.phi_near = s->part.as<double>();
  d.out = dst;
  d.nterms = p->nterms;
  return BIT_OK;
}

// The context's stream waits for the moments a deferred prepare_eval left on
// the side stream.  The join event is only ever recorded on the side stream,
// so waiting on a later record of it covers this system's moments too.
static int join_moments(bit_system* s) {
  if (!s->join_pending) return BIT_OK;
  s->join_pending = false;
  HIP_TRY(hipStreamWaitEvent(s->ctx->stream, s->ctx->join, 0));
  return BIT_OK;
}

int bit_eval_potential(bit_system* s, const bit_params* p, double* out, int out_on_device) {
  if (!out) return fail(BIT_E_INVALID, "null output");
  const int rc = prepare_eval(s, p, out_on_device ? out : nullptr, true);
  struct Join {  // every exit leaves the stream joined with the moments
    bit_system* s;
    ~Join() { (void)join_moments(s); }
  } join{s};
  BIT_TRY(rc);
  bit_ctx* c = s->ctx;
  const int n = s->tree.n;
  const int nb = (n + bit::dev::kBatch - 1) / bit::dev::kBatch;
  BIT_TRY(upload(c, s->desc, &s->hdesc, sizeof(bit::dev::EvalDesc)));
  const bool fp32 = p->precision == BIT_FP32, scr = p->kappa != 0.0;
  const bit::dev::EvalDesc* dd = s->desc.as<bit::dev::EvalDesc>();
  s->t_near.start(c->stream);
  HIP_TRY(bit::dev::near_field(dd, nullptr, nb, fp32, scr, c->stream));
  s->t_near.stop(c->stream);
  BIT_TRY(join_moments(s));  // the far field needs the moments from the side stream
  s->t_far.start(c->stream);
  HIP_TRY(bit::dev::far_field(dd, nullptr, nb, p->nterms, fp32, scr, c->stream));
  s->t_far.stop(c->stream);
  if (!out_on_device)
    HIP_TRY(hipMemcpyAsync(out, s->hdesc.out, (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  collect_times(s);
  return BIT_OK;
}

int bit_eval_potential_noalloc(bit_system* s, const bit_params* p, double* d_out) {
  BIT_TRY(check_params(s, p));
  if (!d_out) return fail(BIT_E_INVALID, "null output");
  const int n = s->tree.n;
  if (n > bit::dev::kNoAllocMaxN)
    return fail(BIT_E_INVALID, "the allocation-free evaluation takes at most " +
                                   std::to_string(bit::dev::kNoAllocMaxN) + " particles (system has " +
                                   std::to_string(n) + ")");
  const ListSet& L = s->pl;
  if (s->mom_nterms != p->nterms || !lists_match(L, p) ||
      (p->precision == BIT_FP64 && s->qc64_nterms != p->nterms))
    return fail(BIT_E_STATE, "system not prepared for these parameters: call bit_system_prepare");
  bit_ctx* c = s->ctx;
  if (!c->scratch) return fail(BIT_E_STATE, "context has no allocation-free scratch");
  HIP_TRY(hipSetDevice(c->device));
  char* base = static_cast<char*>(c->scratch);  // this device's copy, never the symbol
  auto* dd = reinterpret_cast<bit::dev::EvalDesc*>(base);
  bit::dev::EvalDesc d = s->hdesc;
  d.tv = s->view();
  d.kappa = p->kappa;
  d.delta2 = p->delta2;
  d.qc32 = s->qc32.as<float>();
  d.qc64 = s->qc64.as<double>();
  d.far_off = L.far_off.as<int>();
  d.far_items = L.far_items.as<bit::dev::Item>();
  d.near_off = L.near_off.as<int>();
  d.near_items = L.near_items.as<bit::dev::Item>();
  d.phi_near = reinterpret_cast<double*>(base + bit::dev::kNoAllocDescBytes);
  d.out = d_out;
  d.nterms = p->nterms;
  const int nb = (n + bit::dev::kBatch - 1) / bit::dev::kBatch;
  const bool fp32 = p->precision == BIT_FP32, scr = p->kappa != 0.0;
  HIP_TRY(bit::dev::put_desc(d, dd, c->stream));
  HIP_TRY(bit::dev::near_field(dd, nullptr, nb, fp32, scr, c->stream));
  HIP_TRY(bit::dev::far_field(dd, nullptr, nb, p->nterms, fp32, scr, c->stream));
  return BIT_OK;
}

int bit_system_prepare(bit_system* s, const bit_params* p) {
  BIT_TRY(prepare_eval(s, p, nullptr));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  collect_times(s);
  return BIT_OK;
}

namespace {
int mixed_launch(bit_ctx* ctx, const std::vector<bit_system*>& sys,
                 const std::vector<bit_params>& prm, std::vector<bit::dev::EvalDesc>& descs,
                 std::vector<int2>& map, DevBuf& d_desc, DevBuf& d_map, hipStream_t st);
}  // namespace

int bit_eval_potential_multi(bit_ctx* ctx, int count, bit_system* const* systems,
                             const bit_params* params, double* const* outs, int out_on_device,
                             double* kernel_ms) {
  if (!ctx || count < 0 || (count > 0 && (!systems || !params || !outs)))
    return fail(BIT_E_INVALID, "bad multi-evaluation arguments");
  if (count == 0) return BIT_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  {
    // a system's moments, lists and result buffers serve one entry of a launch
    std::vector<const bit_system*> seen(systems, systems + count);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
      return fail(BIT_E_INVALID, "a system appears more than once in a mixed launch");
  }
  for (int i = 0; i < count; i++) {
    if (!systems[i] || !outs[i]) return fail(BIT_E_INVALID, "null system or output");
    if (systems[i]->ctx->device != ctx->device)
      return fail(BIT_E_INVALID, "systems of a mixed launch must share the device");
    BIT_TRY(prepare_eval(systems[i], &params[i], out_on_device ? outs[i] : nullptr));
    HIP_TRY(hipStreamSynchronize(systems[i]->ctx->stream));  // lists/moments ready
  }
  std::vector<bit_system*> sys(systems, systems + count);
  std::vector<bit_params> prm(params, params + count);
  std::vector<bit::dev::EvalDesc> descs;
  std::vector<int2> map;
  DevBuf d_desc, d_map;
  Timer tm;
  tm.start(ctx->stream);
  BIT_TRY(mixed_launch(ctx, sys, prm, descs, map, d_desc, d_map, ctx->stream));
  tm.stop(ctx->stream);
  if (!out_on_device)
    for (int i = 0; i < count; i++)
      HIP_TRY(hipMemcpyAsync(outs[i], systems[i]->hdesc.out,
                             (size_t)systems[i]->tree.n * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (kernel_ms) *kernel_ms = tm.ms();
  return BIT_OK;
}

int bit_eval_field(bit_system* s, const bit_params* p, double* out, int out_on_device) {
  BIT_TRY(check_params(s, p, true));
  if (!out) return fail(BIT_E_INVALID, "null output");
  if (p->kappa != 0.0) return fail(BIT_E_INVALID, "field evaluation supports kappa = 0 only");
  if (p->nterms > BIT_MAX_NTERMS - 1)
    return fail(BIT_E_INVALID, "field needs nterms <= " + std::to_string(BIT_MAX_NTERMS - 1));
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int n = s->tree.n;
  const bool fp32 = p->precision == BIT_FP32;
  // the field's coefficients follow the moments on their stream, beside the lists
  BIT_TRY(moments_and_lists(s, p, false,
                            [&](hipStream_t st) { return ensure_qf(s, p->nterms, fp32, st); }));
  BIT_TRY(s->part.ensure(c, (size_t)3 * n * sizeof(double)));
  double* dst = out;
  if (!out_on_device) {
    BIT_TRY(s->out.ensure(c, (size_t)3 * n * sizeof(double)));
    dst = s->out.as<double>();
  }
  BIT_TRY(ensure_qf(s, p->nterms, fp32));
  s->t_far.start(c->stream);
  HIP_TRY(bit::dev::far_field_grad(s->view(), p->nterms, p->delta2, p->kernel_order, fp32,
                                   fp32 ? (const void*)s->qf32.p : (const void*)s->qf64.p,
                                   s->pl.far_off.as<int>(), s->pl.far_items.as<bit::dev::Item>(),
                                   s->part.as<double>(), c->stream));
  s->t_far.stop(c->stream);
  s->t_near.start(c->stream);
  HIP_TRY(bit::dev::near_field_grad(s->view(), p->delta2, p->kernel_order, fp32,
                                    s->pl.near_off.as<int>(),
                                    s->pl.near_items.as<bit::dev::Item>(), s->part.as<double>(),
                                    dst, c->stream));
  s->t_near.stop(c->stream);
  if (!out_on_device)
    HIP_TRY(hipMemcpyAsync(out, dst, (size_t)3 * n * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  collect_times(s);
  return BIT_OK;
}

// External targets: Morton order over the root box so that each batch of 64
// points is compact, uploaded with the permutation back to the caller's order.
static int setup_points(bit_system* s, int m, const double* x, const double* y, const double* z,
                        bit::dev::TreeView& tv) {
  if (m < 1 || !x || !y || !z) return fail(BIT_E_INVALID, "need m >= 1 points and coordinates");
  if (!finite_all(m, x) || !finite_all(m, y) || !finite_all(m, z))
    return fail(BIT_E_INVALID, "non-finite point coordinate");
  bit_ctx* c = s->ctx;
  const double* root = &s->tree.box[0];
  std::vector<std::pair<uint32_t, int>> key(m);
  for (int i = 0; i < m; i++) key[i] = {bit::leaf_morton_key(x[i], y[i], z[i], root), i};
  std::stable_sort(key.begin(), key.end(),
                   [](const std::pair<uint32_t, int>& u, const std::pair<uint32_t, int>& v) {
                     return u.first < v.first;
                   });
  std::vector<double> hx(m), hy(m), hz(m);
  std::vector<int> hp(m);
  for (int i = 0; i < m; i++) {
    int k = key[i].second;
    hx[i] = x[k];
    hy[i] = y[k];
    hz[i] = z[k];
    hp[i] = k;
  }
  BIT_TRY(upload(c, s->qx, hx.data(), (size_t)m * sizeof(double)));
  BIT_TRY(upload(c, s->qy, hy.data(), (size_t)m * sizeof(double)));
  BIT_TRY(upload(c, s->qz, hz.data(), (size_t)m * sizeof(double)));
  BIT_TRY(upload(c, s->qperm, hp.data(), (size_t)m * sizeof(int)));
  HIP_TRY(hipStreamSynchronize(c->stream));  // host staging dies here
  tv = s->view();
  tv.nt = m;
  tv.self = 0;
  tv.tx = s->qx.as<double>();
  tv.ty = s->qy.as<double>();
  tv.tz = s->qz.as<double>();
  tv.tperm = s->qperm.as<int>();
  return BIT_OK;
}

int bit_eval_potential_points(bit_system* s, const bit_params* p, int m, const double* x,
                              const double* y, const double* z, double* out) {
  BIT_TRY(check_params(s, p));
  if (m < 0) return fail(BIT_E_INVALID, "negative point count");
  if (m == 0) return BIT_OK;  // no points, nothing to evaluate
  if (!out) return fail(BIT_E_INVALID, "null output");
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  BIT_TRY(bit_moments(s, p->nterms));
  bit::dev::TreeView tv;
  BIT_TRY(setup_points(s, m, x, y, z, tv));
  BIT_TRY(build_lists(s, s->ql, tv, p));
  BIT_TRY(s->qphi.ensure(c, (size_t)m * sizeof(double)));
  BIT_TRY(s->qout.ensure(c, (size_t)m * sizeof(double)));
  bit::dev::EvalDesc& d = s->qhdesc;
  d.tv = tv;
  d.kappa = p->kappa;
  d.delta2 = p->delta2;
  d.qc32 = s->qc32.as<float>();
  if (p->precision == BIT_FP64) BIT_TRY(ensure_qc64(s, p->nterms));
  d.qc64 = s->qc64.as<double>();
  d.far_off = s->ql.far_off.as<int>();
  d.far_items = s->ql.far_items.as<bit::dev::Item>();
  d.near_off = s->ql.near_off.as<int>();
  d.near_items = s->ql.near_items.as<bit::dev::Item>();
  d.phi_near = s->qphi.as<double>();
  d.out = s->qout.as<double>();
  d.nterms = p->nterms;
  BIT_TRY(upload(c, s->qdesc, &d, sizeof(bit::dev::EvalDesc)));
  const int nb = (m + bit::dev::kBatch - 1) / bit::dev::kBatch;
  const bool fp32 = p->precision == BIT_FP32, scr = p->kappa != 0.0;
  const bit::dev::EvalDesc* dd = s->qdesc.as<bit::dev::EvalDesc>();
  HIP_TRY(bit::dev::near_field(dd, nullptr, nb, fp32, scr, c->stream));
  HIP_TRY(bit::dev::far_field(dd, nullptr, nb, p->nterms, fp32, scr, c->stream));
  HIP_TRY(hipMemcpyAsync(out, s->qout.p, (size_t)m * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

int bit_eval_field_points(bit_system* s, const bit_params* p, int m, const double* x,
                          const double* y, const double* z, double* out) {
  BIT_TRY(check_params(s, p, true));
  if (m < 0) return fail(BIT_E_INVALID, "negative point count");
  if (m == 0) return BIT_OK;  // no points, nothing to evaluate
  if (!out) return fail(BIT_E_INVALID, "null output");
  if (p->kappa != 0.0) return fail(BIT_E_INVALID, "field evaluation supports kappa = 0 only");
  if (p->nterms > BIT_MAX_NTERMS - 1)
    return fail(BIT_E_INVALID, "field needs nterms <= " + std::to_string(BIT_MAX_NTERMS - 1));
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  BIT_TRY(bit_moments(s, p->nterms));
  bit::dev::TreeView tv;
  BIT_TRY(setup_points(s, m, x, y, z, tv));
  BIT_TRY(build_lists(s, s->ql, tv, p));
  BIT_TRY(s->qphi.ensure(c, (size_t)3 * m * sizeof(double)));
  BIT_TRY(s->qout.ensure(c, (size_t)3 * m * sizeof(double)));
  const bool fp32 = p->precision == BIT_FP32;
  BIT_TRY(ensure_qf(s, p->nterms, fp32));
  HIP_TRY(bit::dev::far_field_grad(tv, p->nterms, p->delta2, p->kernel_order, fp32,
                                   fp32 ? (const void*)s->qf32.p : (const void*)s->qf64.p,
                                   s->ql.far_off.as<int>(), s->ql.far_items.as<bit::dev::Item>(),
                                   s->qphi.as<double>(), c->stream));
  HIP_TRY(bit::dev::near_field_grad(tv, p->delta2, p->kernel_order, fp32,
                                    s->ql.near_off.as<int>(),
                                    s->ql.near_items.as<bit::dev::Item>(), s->qphi.as<double>(),
                                    s->qout.as<double>(), c->stream));
  HIP_TRY(hipMemcpyAsync(out, s->qout.p, (size_t)3 * m * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

int bit_system_stats(const bit_system* s, bit_stats* out) {
  if (!s || !out) return fail(BIT_E_INVALID, "null system or output");
  // the last list build's pair totals (made on the side stream when its join
  // is still pending)
  BIT_TRY(join_moments(const_cast<bit_system*>(s)));
  if (s->pair_pending) {
    unsigned long long v[4];
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipMemcpyAsync(v, s->pair_dev.p, sizeof(v), hipMemcpyDeviceToHost, s->ctx->stream));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    bit_stats& st = const_cast<bit_system*>(s)->stats;
    st.far_pairs = (int64_t)v[0];
    st.near_pairs = (int64_t)v[1];
    st.lane_near_pairs = 64 * (int64_t)v[2];
    st.lane_far = 64 * (int64_t)v[3];
    s->pair_pending = false;
  }
  *out = s->stats;
  return BIT_OK;
}

int bit_system_export_lists(bit_system* s, const bit_params* p, int* nbatch, int* nfar,
                            int* nnear, int* far_off, uint64_t* far_mask, int* far_node,
                            int* near_off, uint64_t* near_mask, int* near_node) {
  BIT_TRY(check_params(s, p, true));
  if (!nbatch || !nfar || !nnear) return fail(BIT_E_INVALID, "null size outputs");
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  BIT_TRY(ensure_lists(s, p));
  const int nb = (s->tree.n + bit::dev::kBatch - 1) / bit::dev::kBatch;
  *nbatch = nb;
  *nfar = s->pl.nfar;
  *nnear = s->pl.nnear;
  if (!far_off) return BIT_OK;
  std::vector<bit::dev::Item> f(std::max(1, s->pl.nfar)), n(std::max(1, s->pl.nnear));
  HIP_TRY(hipMemcpyAsync(far_off, s->pl.far_off.p, (nb + 1) * sizeof(int), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipMemcpyAsync(near_off, s->pl.near_off.p, (nb + 1) * sizeof(int),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(f.data(), s->pl.far_items.p, s->pl.nfar * sizeof(bit::dev::Item),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(n.data(), s->pl.near_items.p, s->pl.nnear * sizeof(bit::dev::Item),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int i = 0; i < s->pl.nfar; i++) {
    far_mask[i] = f[i].mask;
    far_node[i] = f[i].node;
  }
  for (int i = 0; i < s->pl.nnear; i++) {
    near_mask[i] = n[i].mask;
    near_node[i] = n[i].node;
  }
  return BIT_OK;
}

int bit_system_set_charges(bit_system* s, const double* q, int on_device) {
  if (!s || !q) return fail(BIT_E_INVALID, "null system or charges");
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int n = s->tree.n;
  const double* src = q;
  if (!on_device) {
    if (!finite_all(n, q)) return fail(BIT_E_INVALID, "non-finite charge");
    BIT_TRY(upload(c, s->qin, q, (size_t)n * sizeof(double)));
    src = s->qin.as<double>();
  }
  s->mom_nterms = 0;  // moments depend on the charges
  HIP_TRY(bit::dev::set_charges(s->view(), src, s->sq.as<double>(), s->sloc.as<float4>(),
                                c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));  // the host array may be reused at once
  return BIT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Second-kind solves with their Krylov state resident in HBM, stepped in
// serving rounds.
//
// GMRES(m) with every n-vector in HBM: the basis V, the residual/work vector,
// the iterate, the weights and the right-hand side.  Arnoldi orthogonalises by
// classical Gram-Schmidt applied twice; the Hessenberg columns, the Givens
// rotations and the small least-squares solve stay on the host, all in FP64.
// The stopping rule is restarted_gmres's: ||b - A x|| <= tol ||b||, checked on
// the Givens estimate inside a cycle and on the true residual at each restart.
//
// A round advances every active solve by one matrix-vector product (a treecode
// evaluation of w * u, u = x or V_j) and the Arnoldi step that follows it.  The
// products of all solves go into the round's mixed far/near launches together
// with the round's plain evaluations; each Gram-Schmidt pass, the norm and the
// vector updates are one batched launch for all solves, with one read-back per
// pass.  Every solve's arithmetic touches only its own vectors, with the same
// partition of its n, so a solve gives the same iterates alone or in any mix.
// ---------------------------------------------------------------------------
struct bit_krylov {
  bit_system* sys = nullptr;
  bit_params p{};
  double diag = 0, self = 0, tol = 0, bnorm = 0, relres = 0;
  int n = 0, restart = 0, maxit = 0;
  int phase = 0;  // 0: the next product is A x (true residual); 1: A V_j (Arnoldi)
  int j = 0, iters = 0, restarts = 0;
  bool done = false, converged = false, breakdown = false;
  DevBuf w, b, x, V, r, q, phi, partial;
  std::vector<double> H, cs, sn, g, y, h1, h2;
};

namespace {

template <typename T>
int upload_vec(bit_ctx* c, DevBuf& d, const std::vector<T>& v, hipStream_t st) {
  BIT_TRY(d.ensure(c, std::max<size_t>(1, v.size()) * sizeof(T)));
  if (!v.empty()) HIP_TRY(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return BIT_OK;
}

// One persistent launch per precision on `st` (the context's stream) over
// systems whose descriptors (hdesc) were set by prepare_eval: every batch of
// every system is one work item, systems taken in order of decreasing nterms,
// then size, so the costliest batches start first and the cheapest fill the
// tail.  The host tables stay alive in `descs`/`map` until the caller
// synchronises.
int mixed_launch(bit_ctx* ctx, const std::vector<bit_system*>& sys,
                 const std::vector<bit_params>& prm, std::vector<bit::dev::EvalDesc>& descs,
                 std::vector<int2>& map, DevBuf& d_desc, DevBuf& d_map, hipStream_t st) {
  const int count = (int)sys.size();
  if (count == 0) return BIT_OK;
  std::vector<int> order(count);
  for (int i = 0; i < count; i++) order[i] = i;
  auto key = [&](int i) {
    return std::make_tuple(prm[i].precision, -prm[i].nterms, -sys[i]->tree.n);
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
  descs.assign(count, bit::dev::EvalDesc{});
  map.clear();
  std::vector<std::tuple<int, int, bool>> groups;  // (first item, items, FP32) per precision
  for (int k = 0; k < count;) {
    int e = k;
    while (e < count && prm[order[e]].precision == prm[order[k]].precision) e++;
    const int first = (int)map.size();
    for (int j = k; j < e; j++) {
      bit_system* s = sys[order[j]];
      descs[j] = s->hdesc;
      const int nb = (s->tree.n + bit::dev::kBatch - 1) / bit::dev::kBatch;
      for (int b = 0; b < nb; b++) map.push_back(make_int2(j, b));
    }
    groups.emplace_back(first, (int)map.size() - first, prm[order[k]].precision == BIT_FP32);
    k = e;
  }
  BIT_TRY(upload_vec(ctx, d_desc, descs, st));
  BIT_TRY(upload_vec(ctx, d_map, map, st));
  std::lock_guard<std::mutex> lk(ctx->serve_mu);
  if (!ctx->serve_ctr) {
    HIP_TRY(hipMalloc(&ctx->serve_ctr, 2 * sizeof(int)));
    HIP_TRY(hipMemsetAsync(ctx->serve_ctr, 0, 2 * sizeof(int), st));  // once, at creation
  }
  for (const auto& g : groups) {
    HIP_TRY(bit::dev::serve_field(d_desc.as<bit::dev::EvalDesc>(),
                                  d_map.as<int2>() + std::get<0>(g), std::get<1>(g),
                                  std::get<2>(g), static_cast<int*>(ctx->serve_ctr),
                                  ctx->serve_parity, st));
    ctx->serve_parity ^= 1;
  }
  return BIT_OK;
}

// One dot pass over the solves in `ks`: out[i][0..ncols) = V_j . r (basis
// columns, ncols = j + 1) or r . r (norm), read back to the host.
int dot_pass(bit_ctx* ctx, const std::vector<bit_krylov*>& ks, bool norm, DevBuf& d_jobs,
             DevBuf& d_out, std::vector<double>& h_out, int ld, hipStream_t st) {
  const int cnt = (int)ks.size();
  if (cnt == 0) return BIT_OK;
  std::vector<bit::dev::DotJob> jobs(cnt);
  BIT_TRY(d_out.ensure(ctx, (size_t)cnt * ld * sizeof(double)));
  int maxcols = 1;
  for (int i = 0; i < cnt; i++) {
    bit_krylov* k = ks[i];
    const int nc = norm ? 1 : k->j + 1;
    maxcols = std::max(maxcols, nc);
    jobs[i] = bit::dev::DotJob{k->n, nc, norm ? k->r.as<double>() : k->V.as<double>(),
                               k->r.as<double>(), k->partial.as<double>(),
                               d_out.as<double>() + (size_t)i * ld};
  }
  BIT_TRY(upload_vec(ctx, d_jobs, jobs, st));
  HIP_TRY(bit::dev::dots_multi(cnt, maxcols, d_jobs.as<bit::dev::DotJob>(), st));
  h_out.resize((size_t)cnt * ld);
  HIP_TRY(hipMemcpyAsync(h_out.data(), d_out.p, h_out.size() * sizeof(double),
                         hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));  // also ends the job upload's use of `jobs`
  return BIT_OK;
}

// r -= V h (h per solve in coef[i*ld ..]) for the solves in `ks`.
int project_pass(bit_ctx* ctx, const std::vector<bit_krylov*>& ks, const std::vector<double>& coef,
                 int ld, DevBuf& d_jobs, DevBuf& d_coef, std::vector<bit::dev::CombJob>& jobs,
                 hipStream_t st) {
  const int cnt = (int)ks.size();
  if (cnt == 0) return BIT_OK;
  BIT_TRY(upload_vec(ctx, d_coef, coef, st));
  jobs.resize(cnt);
  int maxn = 1;
  for (int i = 0; i < cnt; i++) {
    bit_krylov* k = ks[i];
    maxn = std::max(maxn, k->n);
    jobs[i] = bit::dev::CombJob{k->n, k->j + 1, k->V.as<double>(),
                                d_coef.as<double>() + (size_t)i * ld, -1.0, k->r.as<double>()};
  }
  BIT_TRY(upload_vec(ctx, d_jobs, jobs, st));
  HIP_TRY(bit::dev::combine_multi(cnt, maxn, d_jobs.as<bit::dev::CombJob>(), st));
  return BIT_OK;
}

int serve_round(bit_ctx* ctx, int ne, bit_system* const* es, const bit_params* ep,
                double* const* eouts, int nk, bit_krylov* const* ks, double* kernel_ms) {
  if (!ctx || ne < 0 || nk < 0 || (ne > 0 && (!es || !ep || !eouts)) || (nk > 0 && !ks))
    return fail(BIT_E_INVALID, "bad serving-round arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  std::vector<bit_krylov*> act;
  std::vector<const bit_system*> seen;
  for (int i = 0; i < ne; i++) {
    if (!es[i] || !eouts[i]) return fail(BIT_E_INVALID, "null system or output");
    if (es[i]->ctx->device != ctx->device)
      return fail(BIT_E_INVALID, "systems of a round must share the device");
    seen.push_back(es[i]);
  }
  for (int i = 0; i < nk; i++) {
    if (!ks[i]) return fail(BIT_E_INVALID, "null solve");
    if (ks[i]->sys->ctx->device != ctx->device)
      return fail(BIT_E_INVALID, "solves of a round must share the device");
    seen.push_back(ks[i]->sys);
    if (!ks[i]->done) act.push_back(ks[i]);
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
    return fail(BIT_E_INVALID, "a system appears more than once in a round");
  const int na = (int)act.size();
  std::vector<bit_system*> msys;
  std::vector<bit_params> mprm;
  for (int i = 0; i < ne; i++) {
    BIT_TRY(prepare_eval(es[i], &ep[i], nullptr));
    HIP_TRY(hipStreamSynchronize(es[i]->ctx->stream));  // moments/lists ready
    msys.push_back(es[i]);
    mprm.push_back(ep[i]);
  }
  Timer tm;
  tm.start(st);
  DevBuf d_vjobs, d_desc, d_map, d_djobs, d_dout, d_cjobs, d_coef, d_sjobs;
  std::vector<bit::dev::VecJob> vjobs;
  int maxn = 1;
  // charges w * u of each active solve's product, u = x (phase 0) or V_j
  for (bit_krylov* k : act) {
    const double* u = k->phase == 0 ? k->x.as<double>() : k->V.as<double>() + (size_t)k->j * k->n;
    vjobs.push_back(bit::dev::VecJob{k->n, 0, 0.0, 0.0, k->w.as<double>(), u, nullptr, nullptr,
                                     k->q.as<double>()});
    maxn = std::max(maxn, k->n);
  }
  BIT_TRY(upload_vec(ctx, d_vjobs, vjobs, st));
  HIP_TRY(bit::dev::vec_multi(na, maxn, d_vjobs.as<bit::dev::VecJob>(), st));
  for (bit_krylov* k : act) {
    bit_system* s = k->sys;
    HIP_TRY(bit::dev::set_charges(s->view(), k->q.as<double>(), s->sq.as<double>(),
                                  s->sloc.as<float4>(), st));
    s->mom_nterms = 0;
    BIT_TRY(moments_on(s, k->p.nterms, st));
    if (k->p.precision == BIT_FP64) BIT_TRY(ensure_qc64(s, k->p.nterms, st));
    BIT_TRY(prepare_eval(s, &k->p, k->phi.as<double>()));  // cached lists; sets hdesc
    msys.push_back(s);
    mprm.push_back(k->p);
  }
  std::vector<bit::dev::EvalDesc> descs;
  std::vector<int2> map;
  BIT_TRY(mixed_launch(ctx, msys, mprm, descs, map, d_desc, d_map, st));
  // y = (diag + self w) u + phi: into r (Arnoldi), or r = b - y (residual)
  std::vector<bit::dev::VecJob> pjobs;
  std::vector<bit_krylov*> arn;
  for (bit_krylov* k : act) {
    const double* u = k->phase == 0 ? k->x.as<double>() : k->V.as<double>() + (size_t)k->j * k->n;
    pjobs.push_back(bit::dev::VecJob{k->n, k->phase == 0 ? 4 : 1, k->diag, k->self,
                                     k->w.as<double>(), u, k->phi.as<double>(),
                                     k->b.as<double>(), k->r.as<double>()});
    if (k->phase == 1) arn.push_back(k);
  }
  DevBuf d_pjobs;
  BIT_TRY(upload_vec(ctx, d_pjobs, pjobs, st));
  HIP_TRY(bit::dev::vec_multi(na, maxn, d_pjobs.as<bit::dev::VecJob>(), st));
  // classical Gram-Schmidt, twice
  int ld = 1;
  for (bit_krylov* k : arn) ld = std::max(ld, k->restart + 1);
  std::vector<double> hd, coef;
  std::vector<bit::dev::CombJob> cjobs;
  for (int pass = 0; pass < 2 && !arn.empty(); pass++) {
    BIT_TRY(dot_pass(ctx, arn, false, d_djobs, d_dout, hd, ld, st));
    for (size_t i = 0; i < arn.size(); i++) {
      std::vector<double>& h = pass == 0 ? arn[i]->h1 : arn[i]->h2;
      h.assign(hd.begin() + i * ld, hd.begin() + i * ld + arn[i]->j + 1);
    }
    BIT_TRY(project_pass(ctx, arn, hd, ld, d_cjobs, d_coef, cjobs, st));
  }
  // norms of every active solve's r
  std::vector<double> nrm;
  BIT_TRY(dot_pass(ctx, act, true, d_djobs, d_dout, nrm, 1, st));
  std::vector<bit::dev::VecJob> sjobs;   // V_0 = r / beta, V_{j+1} = r / hn
  std::vector<bit::dev::CombJob> xjobs;  // x += V y at the end of a cycle
  std::vector<std::vector<double>> ys;
  int maxx = 1;
  for (int i = 0; i < na; i++) {
    bit_krylov* k = act[i];
    const double rr = nrm[i];
    if (k->phase == 0) {
      const double beta = std::sqrt(rr);
      k->relres = beta / k->bnorm;
      if (k->relres <= k->tol) {
        k->converged = true;
        k->done = true;
        continue;
      }
      if (k->iters >= k->maxit) {
        k->done = true;
        continue;
      }
      sjobs.push_back(bit::dev::VecJob{k->n, 3, 1.0 / beta, 0.0, nullptr, k->r.as<double>(),
                                       nullptr, nullptr, k->V.as<double>()});
      std::fill(k->H.begin(), k->H.end(), 0.0);
      std::fill(k->g.begin(), k->g.end(), 0.0);
      k->g[0] = beta;
      k->j = 0;
      k->phase = 1;
      continue;
    }
    const int j = k->j, ldh = k->restart;
    for (int t = 0; t <= j; t++) k->H[(size_t)t * ldh + j] = k->h1[t] + k->h2[t];
    const double hn = std::sqrt(rr);
    k->H[(size_t)(j + 1) * ldh + j] = hn;
    if (hn > 0)
      sjobs.push_back(bit::dev::VecJob{k->n, 3, 1.0 / hn, 0.0, nullptr, k->r.as<double>(),
                                       nullptr, nullptr, k->V.as<double>() + (size_t)(j + 1) * k->n});
    bit::detail::givens_column(k->H, ldh, j, k->cs, k->sn, k->g);
    k->iters++;
    k->relres = std::fabs(k->g[j + 1]) / k->bnorm;
    if (hn == 0.0) k->breakdown = true;  // the Krylov space is invariant: exact
    k->j = j + 1;
    if (k->relres <= k->tol || hn == 0.0 || k->j == k->restart || k->iters >= k->maxit) {
      bit::detail::back_solve(k->H, ldh, k->g, k->j, k->y);
      ys.push_back(k->y);
      xjobs.push_back(bit::dev::CombJob{k->n, k->j, k->V.as<double>(), nullptr, 1.0,
                                        k->x.as<double>()});
      maxx = std::max(maxx, k->n);
      k->restarts++;
      k->phase = 0;
    }
  }
  DevBuf d_y;
  std::vector<double> yflat;
  for (size_t i = 0; i < xjobs.size(); i++) yflat.insert(yflat.end(), ys[i].begin(), ys[i].end());
  BIT_TRY(upload_vec(ctx, d_y, yflat, st));
  for (size_t i = 0, off = 0; i < xjobs.size(); off += ys[i].size(), i++)
    xjobs[i].c = d_y.as<double>() + off;
  BIT_TRY(upload_vec(ctx, d_sjobs, sjobs, st));
  HIP_TRY(bit::dev::vec_multi((int)sjobs.size(), maxn, d_sjobs.as<bit::dev::VecJob>(), st));
  BIT_TRY(upload_vec(ctx, d_cjobs, xjobs, st));
  HIP_TRY(bit::dev::combine_multi((int)xjobs.size(), maxx, d_cjobs.as<bit::dev::CombJob>(), st));
  tm.stop(st);
  for (int i = 0; i < ne; i++)
    HIP_TRY(hipMemcpyAsync(eouts[i], es[i]->hdesc.out, (size_t)es[i]->tree.n * sizeof(double),
                           hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (kernel_ms) *kernel_ms = tm.ms();
  return BIT_OK;
}

}  // namespace

extern "C" {

int bit_krylov_create(bit_system* s, const bit_params* p, double diag, double self,
                      const double* w, const double* f, const double* sigma0, int restart,
                      double tol, int maxit, bit_krylov** out) {
  if (!out) return fail(BIT_E_INVALID, "null out");
  *out = nullptr;
  BIT_TRY(check_params(s, p));
  if (!w || !f) return fail(BIT_E_INVALID, "null w or f");
  if (restart < 1 || maxit < 1 || !(tol > 0)) return fail(BIT_E_INVALID, "bad GMRES settings");
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t st = c->stream;
  const int n = s->tree.n;
  if (!finite_all(n, w) || !finite_all(n, f) || (sigma0 && !finite_all(n, sigma0)) ||
      !std::isfinite(diag) || !std::isfinite(self))
    return fail(BIT_E_INVALID, "non-finite weights, right-hand side, initial guess or diagonal");
  BIT_TRY(ensure_lists(s, p));  // the MAC's lists once; every product reuses them
  auto* k = new bit_krylov();
  auto cleanup = [&](int rc) {
    (void)hipStreamSynchronize(st);
    delete k;
    return rc;
  };
  k->sys = s;
  k->p = *p;
  k->diag = diag;
  k->self = self;
  k->tol = tol;
  k->n = n;
  k->restart = restart;
  k->maxit = maxit;
  const size_t nb = (size_t)n * sizeof(double);
  int rc;
  if ((rc = upload(c, k->w, w, nb)) != BIT_OK || (rc = upload(c, k->b, f, nb)) != BIT_OK ||
      (rc = k->x.ensure(c, nb)) != BIT_OK || (rc = k->V.ensure(c, (size_t)(restart + 1) * nb)) != BIT_OK ||
      (rc = k->r.ensure(c, nb)) != BIT_OK || (rc = k->q.ensure(c, nb)) != BIT_OK ||
      (rc = k->phi.ensure(c, nb)) != BIT_OK ||
      (rc = k->partial.ensure(c, (size_t)bit::dev::kDotBlocks * (restart + 1) * sizeof(double))) != BIT_OK)
    return cleanup(rc);
  hipError_t e = sigma0 ? hipMemcpyAsync(k->x.p, sigma0, nb, hipMemcpyHostToDevice, st)
                        : hipMemsetAsync(k->x.p, 0, nb, st);
  if (e != hipSuccess) return cleanup(fail(BIT_E_HIP, hipGetErrorString(e)));
  // ||f||
  DevBuf d_job, d_out;
  bit::dev::DotJob job{n, 1, k->b.as<double>(), k->b.as<double>(), k->partial.as<double>(), nullptr};
  if ((rc = d_out.ensure(c, sizeof(double))) != BIT_OK) return cleanup(rc);
  job.out = d_out.as<double>();
  double bb = 0.0;
  if ((rc = upload(c, d_job, &job, sizeof(job))) != BIT_OK) return cleanup(rc);
  e = bit::dev::dots_multi(1, 1, d_job.as<bit::dev::DotJob>(), st);
  if (e == hipSuccess) e = hipMemcpyAsync(&bb, d_out.p, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return cleanup(fail(BIT_E_HIP, hipGetErrorString(e)));
  collect_times(s);  // the list build's time
  k->bnorm = std::sqrt(bb);
  k->H.assign((size_t)(restart + 1) * restart, 0.0);
  k->cs.assign(restart, 0.0);
  k->sn.assign(restart, 0.0);
  k->g.assign(restart + 1, 0.0);
  if (k->bnorm == 0.0) {  // x = 0 solves it exactly
    e = hipMemsetAsync(k->x.p, 0, nb, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return cleanup(fail(BIT_E_HIP, hipGetErrorString(e)));
    k->done = k->converged = true;
    k->relres = 0.0;
  }
  *out = k;
  return BIT_OK;
}

int bit_krylov_destroy(bit_krylov* k) {
  if (!k) return BIT_OK;
  (void)hipStreamSynchronize(k->sys->ctx->stream);
  delete k;
  return BIT_OK;
}

int bit_krylov_status(const bit_krylov* k, int* done, int* iters, double* relres) {
  if (!k) return fail(BIT_E_INVALID, "null solve");
  if (done) *done = k->done ? (k->converged ? 1 : 2) : 0;
  if (iters) *iters = k->iters;
  if (relres) *relres = k->relres;
  return BIT_OK;
}

int bit_krylov_solution(const bit_krylov* k, double* sigma) {
  if (!k || !sigma) return fail(BIT_E_INVALID, "null solve or output");
  bit_ctx* c = k->sys->ctx;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipMemcpyAsync(sigma, k->x.p, (size_t)k->n * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

int bit_serve_round(bit_ctx* ctx, int n_eval, bit_system* const* systems,
                    const bit_params* params, double* const* outs, int n_solve,
                    bit_krylov* const* solves, double* kernel_ms) {
  return serve_round(ctx, n_eval, systems, params, outs, n_solve, solves, kernel_ms);
}

int bit_solve_second_kind(bit_system* s, const bit_params* p, double diag, double self,
                          const double* w, const double* f, double* sigma, int restart,
                          double tol, int maxit, int* iters, double* relres) {
  if (!sigma) return fail(BIT_E_INVALID, "null sigma");
  bit_krylov* k = nullptr;
  BIT_TRY(bit_krylov_create(s, p, diag, self, w, f, sigma, restart, tol, maxit, &k));
  int rc = BIT_OK;
  while (rc == BIT_OK && !k->done) rc = serve_round(s->ctx, 0, nullptr, nullptr, nullptr, 1, &k, nullptr);
  if (rc == BIT_OK) rc = bit_krylov_solution(k, sigma);
  if (rc == BIT_OK) {
    if (iters) *iters = k->iters;
    if (relres) *relres = k->relres;
  }
  bit_krylov_destroy(k);
  return rc;
}

int bit_gmres_dense(int n, const double* A, const double* b, double* x, int restart, double tol,
                    int maxit, int* iters, double* relres) {
  if (n < 1 || !A || !b || !x) return fail(BIT_E_INVALID, "bad dense GMRES arguments");
  auto matvec = [&](const double* in, double* out) {
    for (int i = 0; i < n; i++) {
      double acc = 0.0;
      for (int j = 0; j < n; j++) acc += A[(size_t)i * n + j] * in[j];
      out[i] = acc;
    }
  };
  bit::GmresResult r = bit::restarted_gmres(matvec, b, x, n, restart, tol, maxit);
  if (iters) *iters = r.iters;
  if (relres) *relres = r.relres;
  return BIT_OK;
}

int bit_gmres(int n, bit_matvec_fn matvec, void* user, const double* b, double* x, int restart,
              double tol, int maxit, int* iters, double* relres) {
  if (n < 1 || !matvec || !b || !x) return fail(BIT_E_INVALID, "bad GMRES arguments");
  auto mv = [&](const double* in, double* out) { matvec(user, in, out); };
  bit::GmresResult r = bit::restarted_gmres(mv, b, x, n, restart, tol, maxit);
  if (iters) *iters = r.iters;
  if (relres) *relres = r.relres;
  return BIT_OK;
}

}  // extern "C"

// Sorted slots of the particle targets tidx (original indices) in s->tslot.
static int upload_target_slots(bit_system* s, int ntargets, const int* tidx) {
  bit_ctx* c = s->ctx;
  const int n = s->tree.n;
  BIT_TRY(ensure_host_perm(s));
  std::vector<int> inv(n), slot(ntargets);
  for (int i = 0; i < n; i++) inv[s->tree.perm[i]] = i;
  for (int t = 0; t < ntargets; t++) {
    if (tidx[t] < 0 || tidx[t] >= n) return fail(BIT_E_INVALID, "target index out of range");
    slot[t] = inv[tidx[t]];
  }
  BIT_TRY(upload(c, s->tslot, slot.data(), slot.size() * sizeof(int)));
  HIP_TRY(hipStreamSynchronize(c->stream));  // `slot` is a host temporary
  return BIT_OK;
}

// Source slices of a direct sum: enough work-groups to fill the device, each
// slice at least 1024 sources.
static int direct_split(int n, int ntargets) {
  const int tblocks = (ntargets + 255) / 256;
  const int nsplit = std::max(1, std::min(1024, 4096 / tblocks));
  return std::min(nsplit, std::max(1, n / 1024));
}

extern "C" {

int bit_direct_potential(bit_system* s, double kappa, double delta2, int ntargets,
                         const int* tidx, double* out) {
  if (!s || (!tidx && ntargets > 0) || (!out && ntargets > 0))
    return fail(BIT_E_INVALID, "null argument");
  if (ntargets < 0) return fail(BIT_E_INVALID, "ntargets < 0");
  if (ntargets == 0) return BIT_OK;
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int n = s->tree.n;
  BIT_TRY(upload_target_slots(s, ntargets, tidx));
  const int nsplit = direct_split(n, ntargets);
  BIT_TRY(s->dpart.ensure(c, (size_t)nsplit * ntargets * sizeof(double)));
  BIT_TRY(s->dout.ensure(c, (size_t)ntargets * sizeof(double)));
  HIP_TRY(bit::dev::direct_potential(s->view(), kappa, delta2, ntargets, s->tslot.as<int>(),
                                     nsplit, s->dpart.as<double>(), s->dout.as<double>(),
                                     c->stream));
  HIP_TRY(hipMemcpyAsync(out, s->dout.p, (size_t)ntargets * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}


int bit_direct_field(bit_system* s, double delta2, int kernel_order, int ntargets,
                     const int* tidx, double* out) {
  if (!s || (!tidx && ntargets > 0) || (!out && ntargets > 0))
    return fail(BIT_E_INVALID, "null argument");
  if (ntargets < 0) return fail(BIT_E_INVALID, "ntargets < 0");
  if (!std::isfinite(delta2) || delta2 < 0 || kernel_order < 0 ||
      kernel_order > BIT_MAX_KERNEL_ORDER)
    return fail(BIT_E_INVALID, "delta2 must be finite and >= 0, kernel_order in [0, " +
                                   std::to_string(BIT_MAX_KERNEL_ORDER) + "]");
  if (ntargets == 0) return BIT_OK;
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int n = s->tree.n;
  BIT_TRY(upload_target_slots(s, ntargets, tidx));
  const int nsplit = direct_split(n, ntargets);
  BIT_TRY(s->dpart.ensure(c, (size_t)3 * nsplit * ntargets * sizeof(double)));
  BIT_TRY(s->dout.ensure(c, (size_t)3 * ntargets * sizeof(double)));
  HIP_TRY(bit::dev::direct_field(s->view(), delta2, kernel_order, ntargets, s->tslot.as<int>(),
                                 nullptr, nullptr, nullptr, nsplit, s->dpart.as<double>(),
                                 s->dout.as<double>(), c->stream));
  HIP_TRY(hipMemcpyAsync(out, s->dout.p, (size_t)3 * ntargets * sizeof(double),
                         hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

int bit_direct_field_points(bit_system* s, double delta2, int kernel_order, int m,
                            const double* x, const double* y, const double* z, double* out) {
  if (!s) return fail(BIT_E_INVALID, "null system");
  if (m < 0) return fail(BIT_E_INVALID, "negative point count");
  if (!std::isfinite(delta2) || delta2 < 0 || kernel_order < 0 ||
      kernel_order > BIT_MAX_KERNEL_ORDER)
    return fail(BIT_E_INVALID, "delta2 must be finite and >= 0, kernel_order in [0, " +
                                   std::to_string(BIT_MAX_KERNEL_ORDER) + "]");
  if (m == 0) return BIT_OK;
  if (!x || !y || !z || !out) return fail(BIT_E_INVALID, "null point or output array");
  if (!finite_all(m, x) || !finite_all(m, y) || !finite_all(m, z))
    return fail(BIT_E_INVALID, "non-finite point coordinate");
  bit_ctx* c = s->ctx;
  HIP_TRY(hipSetDevice(c->device));
  const int n = s->tree.n;
  const size_t mb = (size_t)m * sizeof(double);
  BIT_TRY(upload(c, s->dpx, x, mb));
  BIT_TRY(upload(c, s->dpy, y, mb));
  BIT_TRY(upload(c, s->dpz, z, mb));
  const int nsplit = direct_split(n, m);
  BIT_TRY(s->dpart.ensure(c, (size_t)3 * nsplit * m * sizeof(double)));
  BIT_TRY(s->dout.ensure(c, 3 * mb));
  HIP_TRY(bit::dev::direct_field(s->view(), delta2, kernel_order, m, nullptr,
                                 s->dpx.as<double>(), s->dpy.as<double>(), s->dpz.as<double>(),
                                 nsplit, s->dpart.as<double>(), s->dout.as<double>(), c->stream));
  HIP_TRY(hipMemcpyAsync(out, s->dout.p, 3 * mb, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return BIT_OK;
}

}  // extern "C"
