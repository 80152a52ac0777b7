// kernels.h -- launchers of the CDNA4 kernels of the treecode (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bit {
namespace dev {

constexpr int kBatch = 64;     // targets per batch = one wavefront
constexpr int kMaxNterms = 13; // BIT_MAX_NTERMS

// One entry of a batch interaction list: the cluster/leaf `node` applies to
// the batch's lanes set in `mask` (exactly those targets for which the
// reference traversal reaches that node and takes that branch).
struct Item {
  unsigned long long mask;
  int node;
  int pad;
};

// Per-system device view of the tree and the sorted sources.
struct TreeView {
  int n;        // particles
  int numtree;  // nodes
  const double4* center;  // (xmid, ymid, zmid, sqradius) per node
  const int* child;       // 8 per node, -1 = none
  const int2* range;      // [begin, end) into the sorted arrays
  const double* scale;    // s = sqrt(sqradius) (1 when 0) per node
  const double* sx;       // sorted sources, FP64
  const double* sy;
  const double* sz;
  const double* sq;
  const float4* sloc;     // sorted sources in leaf-centre coordinates, FP32 (x,y,z,q)
  const int* perm;        // sorted slot -> original index
  // targets, in batch order (the sources themselves, or external points)
  int nt;
  int self;               // 1: targets are the sources (self pair excluded)
  const double* tx;
  const double* ty;
  const double* tz;
  const int* tperm;       // target slot -> caller's index
};

struct MacParams {
  double acc;
  double delta2;
  double mac_sub;
  int root;  // node the traversal starts at (0: the whole tree)
};

// counters accumulated by the list-count pass
struct ListStats {
  unsigned long long far_items, near_items;
  unsigned long long far_max, near_max;  // largest per-batch entry counts
};

// ---- tree build on the device ----
// Slot order: P (original index), coordinates.  Two buffers: a split moves a
// range from the buffer its node's data is in to the other one (the node's
// parity), so no copy back is needed; the leaf pass writes the first buffer.
struct SlotArrays {
  int* P;
  double* X;
  double* Y;
  double* Z;
  int* tP;
  double* tX;
  double* tY;
  double* tZ;
};
struct PartOpDev {
  int b, e, dir, src;  // src: buffer holding the range (0: P..Z, 1: tP..tZ)
  double mid;
};
struct ChunkResult {
  int nl, pad;
  double v[12];  // min side0 xyz, min side1 xyz, max side0 xyz, max side1 xyz
};
struct LeafBox {
  int b, e, src, pad;  // src: buffer holding the leaf's slots
  double x0, x1, y0, y1, z0, z1;
};
#ifndef BIT_PART_CHUNK
#define BIT_PART_CHUNK 4096
#endif
constexpr int kPartChunk = BIT_PART_CHUNK;  // slots per work-group of a split
constexpr int kMortonMax = 4096;  // largest leaf sorted along the Morton curve

hipError_t iota(int n, int* p, hipStream_t s);
// flag |= 1 if any of x, y, z, q[0..n) is NaN or infinite
hipError_t nonfinite(int n, const double* x, const double* y, const double* z, const double* q,
                     int* flag, hipStream_t s);
hipError_t minmax(int n, const double* X, const double* Y, const double* Z, int blocks,
                  double* part, hipStream_t s);
// ---- tree build with the split decisions on the device ----
// One slot of a node's children block and the node's split program (the
// host's SplitProgram state, run by one device thread per node).
struct DevSlot {
  int b, e, par, used;
  double f[13];  // TREE box fields
};
struct DevProg {
  DevSlot s[16];
  DevSlot parent, src;
  int kids[8];
  int nchild, pc, xc, yc, j, k, ind2, nchild2, made, c1, c2, failed, pending, maxmemb;
  double lim;
  PartOpDev op;
};
// Node table being built (capacity >= every node written).
struct BuildNodes {
  double* box;  // 13 per node
  int* child;   // 8 per node
  int* begin;
  int* end;
  int* depth;
  int* par;
};
struct GenResult {
  int next;      // end of the next generation (node count so far)
  int failed;    // a program overflowed its children block or did not finish
  int maxdepth;  // depth of the generation's children
  int splits;    // children with more than maxmemb members (next generation's work)
  int maxsize;   // the largest of them
  int seq;       // written last (system scope, release): the generation's token
};
// Generations whose nodes all hold at most this many slots run one work-group
// per node (k_node_program) instead of rounds across the generation (4M-blob
// tree, whole build: 4.20 / 4.18 / 4.80 / 5.30 ms at 32K / 64K / 128K / 256K;
// 256 threads up to 8K / 16K / 32K: 4.25 / 4.18 / 4.24 ms).
#ifndef BIT_NODE_PROGRAM_MAX
#define BIT_NODE_PROGRAM_MAX 65536
#endif
#ifndef BIT_NODE_PROGRAM_WIDE
#define BIT_NODE_PROGRAM_WIDE 16384
#endif
constexpr int kNodeProgramMax = BIT_NODE_PROGRAM_MAX;
// ... with 256 threads per node up to this size, 1024 above it.
constexpr int kNodeProgramWide = BIT_NODE_PROGRAM_WIDE;
// A generation's rounds, for cnt programs: ops, nch 2 cnt (two buffers);
// cbeg cnt + 1; cprog ceil(n / kPartChunk) + cnt; ntot 2.
struct RoundLayout {
  PartOpDev* ops;  // a round's request per program
  int* nch;        // its chunks, tagged with the round
  int* cbeg;       // first chunk of each program's budget
  int* cprog;      // program of each budget chunk
  int* ntot;       // [0] budget chunks in all, [1] programs
};
// Per-generation work arrays, sized for cnt programs (next: 8 cnt) and
// ceil(n / kPartChunk) + cnt chunks (cres).
struct BuildScratch {
  DevProg* progs;     // the generation's programs
  DevProg* next;      // the next generation's, started by k_prog_write
  RoundLayout cur;    // this generation's rounds (written by the last generation)
  RoundLayout nxt;    // the next generation's, written by k_prog_number / k_prog_write
  ChunkResult* cres;  // per chunk: count results
  int* nkid;
  int* start;
  GenResult* result;
};
constexpr int kMaxSplitsPerNode = 7;  // x, then y and two z per x-child
// Root node from k_minmax's mblocks partial boxes, its program and the layout
// of its rounds.
hipError_t build_root(int n, int mblocks, const double* part, const BuildNodes& nd,
                      DevProg* progs, int maxmemb, double lim, const RoundLayout& L,
                      hipStream_t s);
// Nodes [gen_lo, gen_hi) at depth depth - 1, the largest that splits of
// maxsize slots: split programs, run in kMaxSplitsPerNode partition rounds
// across the generation or (maxsize <= kNodeProgramMax) one work-group per
// node; children written from gen_hi on with their programs started in
// w.next; *w.result receives the next generation's end.
hipError_t build_generation(const SlotArrays& a, int n, int gen_lo, int gen_hi, int depth,
                            int maxmemb, double lim, int maxsize, const BuildNodes& nd,
                            const BuildScratch& w, hipStream_t s);

// Morton order inside the leaves of a node table of numtree nodes, written to
// the first buffer (P..Z) from the buffer each leaf's slots are in (every
// slot covered).
hipError_t leaf_morton(const SlotArrays& a, int numtree, const BuildNodes& nd, hipStream_t s);
// blocks x 256 threads of eight independent FMA chains of iters steps each
// (FP64 or FP32), one result per thread into out.
hipError_t fma_ceiling(int blocks, int iters, bool fp64, void* out, hipStream_t s);
// tv's center / scale / range / child (writable device arrays of numtree
// nodes) from a device-built node table, then sloc's positions relative to
// each leaf's midpoint (tv's sorted sources and charges in place).
hipError_t tree_view(int numtree, const BuildNodes& nd, const TreeView& tv, float4* sloc,
                     hipStream_t s);

hipError_t gather_sources(const TreeView& tv, const double* x, const double* y,
                          const double* z, const double* q, double* sx, double* sy,
                          double* sz, double* sq, hipStream_t s);
hipError_t set_charges(const TreeView& tv, const double* q, double* sq, float4* sloc,
                       hipStream_t s);
hipError_t leaf_local(const TreeView& tv, const int* leaves, int nleaves, float4* sloc,
                      hipStream_t s);

hipError_t moments_partial(const TreeView& tv, int nterms, int nchunks, const int* chunk_node,
                           const int2* chunk_rng, double* partial, hipStream_t s);
// Upward pass: moments of `nodes` (internal nodes whose children are done)
// from their children's raw moments; writes raw and scaled copies.
hipError_t moments_upward(const TreeView& tv, int nterms, int nnodes, const int* nodes,
                          double* mom, double* mom_hat64, hipStream_t s);
hipError_t moments_reduce(const TreeView& tv, int nterms, const int* node_chunk0,
                          const int* node_nchunk, const double* partial, double* mom,
                          double* mom_hat64, hipStream_t s);

// Far-field coefficients of the FP32 far kernel per node (radial-derivative
// form of the Taylor sum): nqcoef(nterms) values per node from the scaled
// FP64 moments, stored in tiles of nqpad(nterms) floats.
constexpr int nqcoef(int nt) {
  int n = 0;
  for (int d = 0; d < nt; d++) n += (d + 1) * (d + 2) / 2 * ((nt - 1 - d) / 2 + 1);
  return n;
}
// Floats per node in the coefficient buffer (tiles padded to whole wavefronts).
constexpr int nqpad(int nt) { return (nqcoef(nt) + 63) / 64 * 64; }
hipError_t far_coeffs32(int numtree, int nterms, const double* mom_hat64, float* qc32,
                        hipStream_t s);
hipError_t far_coeffs64(int numtree, int nterms, const double* mom_hat64, double* qc64,
                        hipStream_t s);

hipError_t lists_count(const TreeView& tv, MacParams mp, int stack_cap, int* far_cnt,
                       int* near_cnt, int* err_flag, hipStream_t s);
// Totals and maxima over nb batches of the counts.
hipError_t list_stats(int nb, const int* far_cnt, const int* near_cnt, ListStats* out,
                      hipStream_t s);
// Totals over nb batches of pstat (4 per batch: far pairs, near pairs,
// k_near32's pair slots per lane, k_farq's sweeps), written by far_order and
// near_compact, into out[4].
hipError_t pair_stats(int nb, const unsigned long long* pstat, unsigned long long* out,
                      hipStream_t s);
// Writes the items at the scanned offsets (far_order then orders the far
// items in place).
hipError_t lists_fill(const TreeView& tv, MacParams mp, int stack_cap, const int* far_off,
                      const int* near_off, Item* far_items, Item* near_items, int* err_flag,
                      hipStream_t s);
// One traversal that counts like lists_count and writes batch b's first
// cap_far / cap_near entries to far_stage[b*cap_far ..] / near_stage[b*cap_near ..].
hipError_t lists_stage(const TreeView& tv, MacParams mp, int stack_cap, int* far_cnt,
                       int* near_cnt, int cap_far, int cap_near, Item* far_stage,
                       Item* near_stage, int* err_flag, hipStream_t s);
// Far items of batch b from src + b * cap (cap > 0: staging regions) or
// src + far_off[b] (cap == 0; src may be far_items itself) to far_items at
// far_off[b], each batch's in the far kernel's pairing order (k_far_order in
// kernels.hip), with the batch's far pair statistics (pstat[4b], [4b + 3]).
hipError_t far_order(int nb, const int* far_off, int cap, const Item* src, Item* far_items,
                     unsigned long long* pstat, hipStream_t s);
// Near items copied from stage + b * cap when cap > 0 (else read in place), with
// the batch's near pair statistics (pstat[4b + 1], [4b + 2]; range: the nodes'
// slot ranges).
hipError_t near_compact(int nb, const int* near_off, int cap, const Item* stage,
                        Item* near_items, const int2* range, unsigned long long* pstat,
                        hipStream_t s);
hipError_t exclusive_scan(const int* in, int* out, int n, void* tmp, size_t& tmp_bytes,
                          hipStream_t s);

// One system of a (mixed) evaluation launch, read from HBM by every
// work-group that serves one of its batches.
struct EvalDesc {
  TreeView tv;
  double kappa;
  double delta2;
  const float* qc32;     // far coefficients, FP32 (far_coeffs32)
  const double* qc64;    // far coefficients, FP64 (far_coeffs64)
  const int* far_off;    // [nbatch+1]
  const Item* far_items;
  const int* near_off;
  const Item* near_items;
  double* phi_near;      // near-field result, sorted order (near kernel -> far kernel)
  double* out;           // potential, caller's original order
  int nterms;            // Taylor terms (read per item by the persistent launch)
};

// near field -> d.phi_near; map == nullptr: one system, work-group b = batch b;
// else map[g] = (system, batch).  Runs first: it needs the lists, not the moments.
hipError_t near_field(const EvalDesc* descs, const int2* map, int nblocks, bool fp32,
                      bool screened, hipStream_t s);
// far field + combine: d.out[perm[t]] = far(t) + d.phi_near[t], after near_field.
// One launch serves one (nterms, precision, screened) group.
hipError_t far_field(const EvalDesc* descs, const int2* map, int nblocks, int nterms, bool fp32,
                     bool screened, hipStream_t s);

// Persistent evaluation of many systems in one launch (all nterms and both
// kernels; one precision): work-groups sized to what the device holds at once
// take (system, batch) items from a device counter, in the order of `items`,
// and run each batch's near field, then its far field and the combine, until
// the counter passes nitems.  Results equal those of near_field + far_field.
// ctr: two ints with ctr[parity] zero at the launch; the launch zeroes
// ctr[parity ^ 1] for the next one (launches that share ctr share a stream).
hipError_t serve_field(const EvalDesc* descs, const int2* items, int nitems, bool fp32,
                       int* ctr, int parity, hipStream_t s);

// Krylov vectors (FP64): out[j] = V_j . w for j < ncols (deterministic order;
// partial holds kDotBlocks * ncols doubles), y += sign * sum_j c[j] V_j, and the
// elementwise steps of vec_op: 0 y = u v; 1 y = (a + b u) v + z; 2 y = u - v;
// 3 y = a v.
constexpr int kDotBlocks = 256;
hipError_t dots(int n, const double* V, size_t ld, int ncols, const double* w, double* partial,
                double* out, hipStream_t s);
hipError_t combine_cols(int n, const double* V, size_t ld, int ncols, const double* c,
                        double sign, double* y, hipStream_t s);
// Batched Krylov steps of many solves in one launch each: job i of a launch
// works on solve i's own vectors only, with the same per-element arithmetic
// (and the same kDotBlocks partition of its n) as the single-solve kernels.
struct DotJob {      // out[j] = V_j . w, j < ncols (partial: kDotBlocks * ncols)
  int n, ncols;
  const double* V;   // ld = n
  const double* w;
  double* partial;
  double* out;
};
struct CombJob {     // y += sign * sum_{j < ncols} c[j] V_j
  int n, ncols;
  const double* V;   // ld = n
  const double* c;
  double sign;
  double* y;
};
struct VecJob {      // vec_op's elementwise steps; op 4: y = c - ((a + b u) v + z)
  int n, op;
  double a, b;
  const double *u, *v, *z, *c;
  double* y;
};
hipError_t dots_multi(int count, int maxcols, const DotJob* jobs, hipStream_t s);
hipError_t combine_multi(int count, int maxn, const CombJob* jobs, hipStream_t s);
hipError_t vec_multi(int count, int maxn, const VecJob* jobs, hipStream_t s);
hipError_t vec_op(int n, int op, double a, double b, const double* u, const double* v,
                  const double* z, double* y, hipStream_t s);

// Field (Coulomb, regularized; korder >= 1 adds the terms l = 1..korder of the
// regularized family): per-node coefficients Q^(e) of the radial-derivative form
// at degree nterms (3 x nqpad(nterms + 1) values per node, FP32 or FP64), E_far
// sorted [3n], then near + scatter [3n].
hipError_t field_coeffs32(int numtree, int nterms, const double* mom_hat64, float* qf,
                          hipStream_t s);
hipError_t field_coeffs64(int numtree, int nterms, const double* mom_hat64, double* qf,
                          hipStream_t s);
hipError_t far_field_grad(const TreeView& tv, int nterms, double delta2, int korder, bool fp32,
                          const void* qf, const int* far_off, const Item* far_items,
                          double* e_far, hipStream_t s);
hipError_t near_field_grad(const TreeView& tv, double delta2, int korder, bool fp32,
                           const int* near_off, const Item* near_items, const double* e_far,
                           double* out, hipStream_t s);

// FP64 direct sum at ntargets sorted target slots; partial sums over
// nsplit source chunks, then reduced in order.
hipError_t direct_potential(const TreeView& tv, double kappa, double delta2, int ntargets,
                            const int* tslot, int nsplit, double* partial, double* out,
                            hipStream_t s);
// FP64 direct-sum field E_t = sum_j q_j (x_t - x_j) G'(rs), rs = r^2 + delta2, the
// regularized family up to korder, at sorted particle slots tslot (own slot
// skipped) or, with tslot == nullptr, at points (qx, qy, qz) (every source).
// partial holds 3 * nsplit * ntargets doubles; out[3 * ntargets].
hipError_t direct_field(const TreeView& tv, double delta2, int korder, int ntargets,
                        const int* tslot, const double* qx, const double* qy, const double* qz,
                        int nsplit, double* partial, double* out, hipStream_t s);

// ---- allocation-free evaluation ----
// Scratch of bit_eval_potential_noalloc: the evaluation descriptor, then the
// near-field results of up to kNoAllocMaxN targets.
constexpr size_t kNoAllocDescBytes = 1024;
constexpr int kNoAllocMaxN = 1 << 22;
constexpr size_t kNoAllocScratchBytes = kNoAllocDescBytes + (size_t)kNoAllocMaxN * sizeof(double);
static_assert(sizeof(EvalDesc) <= kNoAllocDescBytes, "descriptor region too small");
// Device address of the calling device's scratch (resolved once per device,
// with `device` current).
hipError_t noalloc_scratch(int device, void** out);
// *dst = d, stored by a kernel that takes d by value (stream-ordered, no copy
// from host memory).
hipError_t put_desc(const EvalDesc& d, EvalDesc* dst, hipStream_t s);

}  // namespace dev
}  // namespace bit
