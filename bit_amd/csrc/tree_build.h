// tree_build.h -- host construction of the bisection tree, bit-exact with the
// reference's treemake (tree3d.cpp:1345-1551).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace bit {

// Box fields of one node, in the order of TREE (tree3d.h:51-74).
enum BoxField {
  BX0 = 0, BX1, BY0, BY1, BZ0, BZ1, BXMID, BYMID, BZMID, BAXY, BAXZ, BAYZ, BSQR, BOX_FIELDS
};

struct HostTree {
  int n = 0;
  int numtree = 0;
  int maxdepth = 0;
  std::vector<double> box;    // BOX_FIELDS per node
  std::vector<int> children;  // 8 per node, -1 = none, reference order
  std::vector<int> begin;     // node i's members are the set perm[begin[i] .. end[i]-1];
                              // inside a leaf, slots follow a Morton curve over the box
  std::vector<int> end;
  std::vector<int> depth;
  std::vector<int> perm;      // sorted slot -> original particle index
  // Number of splits that moved node i's slots, mod 2 (an engine that moves
  // slots between two buffers finds node i's data in buffer par[i]).
  std::vector<unsigned char> par;
};

// Builds the tree over n particles (coordinates in original order).  Returns
// false with a message when the particles cannot be separated into leaves of
// at most maxmemb (coincident points) or the input is invalid.
// One stable split request: slots [b, e) by coordinate `dir` <= mid.
struct PartOp {
  int b, e, dir;
  double mid;
  int par;  // splits that moved these slots so far, mod 2
};
// Its outcome: nl slots on the `<=` side, then each side's bounding box
// (mn/mx[side][axis], meaningful for a non-empty side).
struct PartResult {
  int nl;
  double mn[2][3];
  double mx[2][3];
};

// Executes the split requests of one round (disjoint ranges) on whatever
// holds the slot arrays (host memory or HBM).
class PartitionEngine {
 public:
  virtual ~PartitionEngine() {}
  virtual bool run(const std::vector<PartOp>& ops, std::vector<PartResult>& res,
                   std::string& err) = 0;
};

// The node structure of treemake for n particles whose bounding box is
// [mn, mx]: generations of nodes, each node's split program advanced one
// split per round, splits executed by `eng`.  Fills everything but perm.
bool build_tree_generic(PartitionEngine& eng, int n, const double mn[3], const double mx[3],
                        int maxmemb, HostTree& out, std::string& err);

// Morton key of a position inside a leaf box (10 bits per axis, interleaved).
uint32_t leaf_morton_key(double x, double y, double z, const double* box);

// max_threads > 0 caps the OpenMP threads used (0 = all available).
bool build_tree(int n, const double* x, const double* y, const double* z, int maxmemb,
                HostTree& out, std::string& err, int max_threads = 0);

}  // namespace bit
