// kernels.hip -- CDNA4 (gfx950) kernels of the particle-cluster treecode.
//
// Layout in HBM (built by bit_api.cpp):
//   * sources sorted so every tree node is a contiguous slot range; FP64 SoA
//     (sx, sy, sz, sq) and an FP32 float4 copy in leaf-centre coordinates;
//   * node table SoA: double4 centre+sqradius, int[8] children, int2 range,
//     double scale s = sqrt(sqradius);
//   * moments packed by total-degree shell, T(p) = C(p+3,3) per node, FP64,
//     plus a copy scaled by s^-|k| (FP64) and the far kernels' cluster
//     coefficients (radial-derivative form, FP32 and FP64; see qstep);
//   * interaction lists per batch of 64 consecutive sorted targets: (mask,
//     node) items, far and near, exactly the reference traversal's choices.
// Work mapping: one wavefront per batch, one lane per target.  Every lane owns
// its target's sums, so a target's result does not depend on what else runs.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <mutex>
#include <type_traits>
#include <utility>

#include "kernels.h"

namespace bit {
namespace dev {

// ---------------------------------------------------------------------------
// Packed moments.  Index of multi-index (i,j,k), n = i+j+k:
//   n(n+1)(n+2)/6 + (j+k)(j+k+1)/2 + k
// i.e. shells of increasing degree, inside a shell i descending, then k
// ascending.
// ---------------------------------------------------------------------------
constexpr int cidx(int i, int j, int k) {
  int n = i + j + k, m = j + k;
  return n * (n + 1) * (n + 2) / 6 + m * (m + 1) / 2 + k;
}
constexpr int ncoef(int nt) { return nt * (nt + 1) * (nt + 2) / 6; }

// Pointers read from a descriptor are generic (flat) to the compiler; loads
// through a global-address-space pointer with a wave-uniform address can be
// scalar loads (the values then live in SGPRs, not VGPRs).
template <typename T>
using gptr = const __attribute__((address_space(1))) T*;
// Read-only data: loads with wave-uniform addresses through a constant
// address-space pointer are scalar loads.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> as_global(const T* p) {
  return (gptr<T>)(p);
}


// ---------------------------------------------------------------------------
// Radial-derivative form of the far field (FP32 kernel).  The same truncated
// Taylor sum sum_{|k|<=p} M_k a_k, evaluated through the radial derivatives of
// the kernel instead of the coefficient recurrence.  With G(R) = exp(-kappa R)/R,
// g_m = (R^-1 d/dR)^m G and d = target - centre, every Taylor coefficient is
//   a_k = (-1)^|k| sum_{2j<=k} d^(k-2j) g_{|k|-|j|} / (j! (k-2j)! 2^|j|)
// (the chain rule for G(sqrt(2 S)), S = (|d|^2 + delta^2)/2, one axis at a
// time).  In the variables u = d/R, rho = s/R, w = rho u, z = rho^2 and
// ghat_m = g_m R^(2m+1) this regroups as
//   sum_k M_k a_k = (1/R) sum_t z^t sum_{|l|<=p-2t} ghat_{|l|+t} Q_{|l|+t,l} w^l,
//   Q_{m,l} = (-1)^|l| / l! * sum_{|j|=m-|l|} Mhat_{l+2j} / (j! 2^|j|),
// with Mhat_k = M_k / s^|k|.  Q depends on the cluster only and is computed
// once per node (k_qcoef).  With the ratios r_m = ghat_{m+1}/ghat_m the inner
// sum is ghat_t times a nested Horner sweep (l0, l1, l2 descending) whose step
// in direction e at degree n multiplies by r_{n+t} w_e, so every Q costs one
// FMA per (target, cluster) pair:
//   B = fma(B, r_{l0+l1+l2+t} w_2, Q),  A = fma(A, r_{l0+l1+t} w_1, B),
//   Phi = fma(Phi, r_{l0+t} w_0, A),    S = fma(S, z, ghat_t Phi).
// The ratios obey r_0 = -(1 + lam), r_m = lam^2 / r_{m-1} - (2m+1), lam =
// kappa R (from ghat_{m+1} = lam^2 ghat_{m-1} - (2m+1) ghat_m, whose terms
// share one sign); Coulomb: r_m = -(2m+1).
// Q order = consumption order: t descending, then l0, l1, l2 descending.
// ---------------------------------------------------------------------------
template <int NT>
struct QTab {
  static constexpr int P = NT - 1;
  static constexpr int NQ = nqcoef(NT);
  int l[NQ][3];
  int t[NQ];
  constexpr QTab() : l(), t() {
    int q = 0;
    for (int tt = P / 2; tt >= 0; tt--) {
      const int L = P - 2 * tt;
      for (int l0 = L; l0 >= 0; l0--)
        for (int l1 = L - l0; l1 >= 0; l1--)
          for (int l2 = L - l0 - l1; l2 >= 0; l2--) {
            l[q][0] = l0;
            l[q][1] = l1;
            l[q][2] = l2;
            t[q] = tt;
            q++;
          }
    }
  }
};
template <int NT>
inline constexpr QTab<NT> kQTab{};
template <int NT>
__constant__ QTab<NT> kQTabDev = QTab<NT>();

__constant__ double kInvFact[kMaxNterms + 1] = {
    1.0,           1.0,           1.0 / 2,        1.0 / 6,         1.0 / 24,
    1.0 / 120,     1.0 / 720,     1.0 / 5040,     1.0 / 40320,     1.0 / 362880,
    1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600, 1.0 / 6227020800.0};

// One thread per (node, Q entry), from the FP64 scaled moments.  Each node's
// tile is padded with zeros to nqpad(NT) values.
template <int NT, typename T>
__global__ __launch_bounds__(256) void k_qcoef(int numtree, const double* __restrict__ mom_hat,
                                               T* __restrict__ qc) {
  constexpr int NQ = nqcoef(NT), NQP = nqpad(NT), NC = ncoef(NT);
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)numtree * NQP) return;
  const int node = (int)(g / NQP), q = (int)(g - (long long)node * NQP);
  if (q >= NQ) {
    qc[g] = T(0);
    return;
  }
  const int l0 = kQTabDev<NT>.l[q][0], l1 = kQTabDev<NT>.l[q][1], l2 = kQTabDev<NT>.l[q][2];
  const int t = kQTabDev<NT>.t[q];
  const double* M = mom_hat + (size_t)node * NC;
  double sum = 0.0;
  for (int j0 = 0; j0 <= t; j0++)
    for (int j1 = 0; j0 + j1 <= t; j1++) {
      const int j2 = t - j0 - j1;
      sum = fma(M[cidx(l0 + 2 * j0, l1 + 2 * j1, l2 + 2 * j2)],
                kInvFact[j0] * kInvFact[j1] * kInvFact[j2], sum);
    }
  const double w = ldexp(kInvFact[l0] * kInvFact[l1] * kInvFact[l2], -t);
  qc[g] = (T)(((l0 + l1 + l2) & 1 ? -w : w) * sum);
}

template <int NT, typename T>
static void launch_qcoef(int numtree, const double* mom_hat, T* qc, hipStream_t s) {
  const long long total = (long long)numtree * nqpad(NT);
  hipLaunchKernelGGL((k_qcoef<NT, T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                     numtree, mom_hat, qc);
}

template <typename T, int... NTs>
static bool dispatch_qcoef(int nterms, std::integer_sequence<int, NTs...>, int numtree,
                           const double* mom_hat, T* qc, hipStream_t s) {
  bool done = false;
  ((nterms == NTs + 1 ? (launch_qcoef<NTs + 1, T>(numtree, mom_hat, qc, s), done = true)
                      : false),
   ...);
  return done;
}

template <typename T>
static hipError_t far_coeffs(int numtree, int nterms, const double* mom_hat64, T* qc,
                             hipStream_t s) {
  if (numtree == 0) return hipSuccess;
  if (!dispatch_qcoef(nterms, std::make_integer_sequence<int, kMaxNterms>{}, numtree, mom_hat64,
                      qc, s))
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t far_coeffs32(int numtree, int nterms, const double* mom_hat64, float* qc32,
                        hipStream_t s) {
  return far_coeffs(numtree, nterms, mom_hat64, qc32, s);
}

hipError_t far_coeffs64(int numtree, int nterms, const double* mom_hat64, double* qc64,
                        hipStream_t s) {
  return far_coeffs(numtree, nterms, mom_hat64, qc64, s);
}

__device__ __forceinline__ float fmaT(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return fma(a, b, c); }

// Per-pair values of the sweep: mult[e][m] = r_m w_e, ghat_t, z.
template <int NT, typename T>
struct QPair {
  T mult[3][NT > 1 ? NT - 1 : 1];
  T g[NT / 2 + 1];
  T z;
};

// w = rho u, z = rho^2, c0 = exp(-lam) (screened), lam = kappa R.  FP32 uses
// the hardware reciprocal for the ratio recurrence, FP64 the division.
template <int NT, bool SCR, typename T>
__device__ __forceinline__ void qpair_init(QPair<NT, T>& v, T w0, T w1, T w2, T z, T c0, T lam) {
  constexpr int P = NT - 1;
  T r[P > 0 ? P : 1];
  if constexpr (SCR) {
    if constexpr (P > 0) r[0] = -(T(1) + lam);
    const T lam2 = lam * lam;
#pragma unroll
    for (int m = 1; m < P; m++) {
      if constexpr (std::is_same_v<T, float>)
        r[m] = fmaf(lam2, __builtin_amdgcn_rcpf(r[m - 1]), -float(2 * m + 1));
      else
        r[m] = lam2 / r[m - 1] - double(2 * m + 1);
    }
    v.g[0] = c0;
  } else {
#pragma unroll
    for (int m = 0; m < P; m++) r[m] = -T(2 * m + 1);
    v.g[0] = T(1);
  }
#pragma unroll
  for (int t = 1; t <= P / 2; t++) v.g[t] = v.g[t - 1] * r[t - 1];
#pragma unroll
  for (int m = 0; m < P; m++) {
    v.mult[0][m] = r[m] * w0;
    v.mult[1][m] = r[m] * w1;
    v.mult[2][m] = r[m] * w2;
  }
  v.z = z;
}

template <typename T>
struct QState {
  T b, a, phi, s;  // Horner sums over l2, l1, l0 and t
};

// Entry C of the sweep for one target (q = the cluster's Q_C).
template <int NT, int C, typename T>
__device__ __forceinline__ void qstep(const QPair<NT, T>& v, T q, QState<T>& z) {
  constexpr const QTab<NT>& tb = kQTab<NT>;
  constexpr int l0 = tb.l[C][0], l1 = tb.l[C][1], l2 = tb.l[C][2], t = tb.t[C];
  constexpr int L = NT - 1 - 2 * t;
  if constexpr (l2 == L - l0 - l1) z.b = q;
  else z.b = fmaT(z.b, v.mult[2][l0 + l1 + l2 + t], q);
  if constexpr (l2 == 0) {
    if constexpr (l1 == L - l0) z.a = z.b;
    else z.a = fmaT(z.a, v.mult[1][l0 + l1 + t], z.b);
    if constexpr (l1 == 0) {
      if constexpr (l0 == L) z.phi = z.a;
      else z.phi = fmaT(z.phi, v.mult[0][l0 + t], z.a);
      if constexpr (l0 == 0) {
        if constexpr (t == (NT - 1) / 2) z.s = v.g[t] * z.phi;
        else z.s = fmaT(z.s, v.z, v.g[t] * z.phi);
      }
    }
  }
}

// The sweep over the cluster's Q (in LDS) for one target.
template <int NT, typename T, int... C>
__device__ __forceinline__ T qall(std::integer_sequence<int, C...>, const QPair<NT, T>& v,
                                  const T* M) {
  QState<T> z{T(0), T(0), T(0), T(0)};
  (qstep<NT, C, T>(v, M[C], z), ...);
  return z.s;
}

// The FP32 sweep as generated inline asm (gen_far_sweep.py): the same operations
// as qall<NT, float>, each accumulation in the 4-byte v_fmac form.
#include "far_sweep_asm.h"

// Byte address in LDS of a pointer into __shared__ memory.
__device__ __forceinline__ unsigned lds_addr(const float* p) {
  return (unsigned)(size_t)(const __attribute__((address_space(3))) float*)p;
}

// The FP32 sweep of one target over a Q tile in LDS: the generated asm where
// it exists, else the compiler's (identical operations either way).
template <int NT>
__device__ __forceinline__ float sweep32(const QPair<NT, float>& v, const float* M) {
  if constexpr (NT <= kAsmSweepMaxNt) return qsweep32<NT>(v, lds_addr(M));
  else return qall<NT, float>(std::make_integer_sequence<int, nqcoef(NT)>{}, v, M);
}

__device__ __forceinline__ unsigned long long ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long uniform64(unsigned long long v) {
  unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// ---------------------------------------------------------------------------
// Sorting the sources
// ---------------------------------------------------------------------------
__global__ void k_gather(int n, const int* __restrict__ perm, const double* __restrict__ x,
                         const double* __restrict__ y, const double* __restrict__ z,
                         const double* __restrict__ q, double* sx, double* sy, double* sz,
                         double* sq) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int p = perm[t];
  sx[t] = x[p];
  sy[t] = y[p];
  sz[t] = z[p];
  sq[t] = q[p];
}

hipError_t gather_sources(const TreeView& tv, const double* x, const double* y,
                          const double* z, const double* q, double* sx, double* sy,
                          double* sz, double* sq, hipStream_t s) {
  int blocks = (tv.n + 255) / 256;
  hipLaunchKernelGGL(k_gather, dim3(blocks), dim3(256), 0, s, tv.n, tv.perm, x, y, z, q, sx,
                     sy, sz, sq);
  return hipGetLastError();
}

// New charges for the same positions (original order in q).
__global__ void k_set_charges(int n, const int* __restrict__ perm, const double* __restrict__ q,
                              double* sq, float4* sloc) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double v = q[perm[t]];
  sq[t] = v;
  sloc[t].w = (float)v;
}

hipError_t set_charges(const TreeView& tv, const double* q, double* sq, float4* sloc,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_set_charges, dim3((tv.n + 255) / 256), dim3(256), 0, s, tv.n, tv.perm, q,
                     sq, sloc);
  return hipGetLastError();
}

// One block per leaf: FP32 source copy relative to the leaf's box midpoint.
__global__ void k_leaf_local(const int* __restrict__ leaves, const double4* __restrict__ center,
                             const int2* __restrict__ range, const double* __restrict__ sx,
                             const double* __restrict__ sy, const double* __restrict__ sz,
                             const double* __restrict__ sq, float4* sloc) {
  int leaf = leaves[blockIdx.x];
  double4 c = center[leaf];
  int2 r = range[leaf];
  for (int t = r.x + threadIdx.x; t < r.y; t += blockDim.x) {
    sloc[t] = make_float4((float)(sx[t] - c.x), (float)(sy[t] - c.y), (float)(sz[t] - c.z),
                          (float)sq[t]);
  }
}

hipError_t leaf_local(const TreeView& tv, const int* leaves, int nleaves, float4* sloc,
                      hipStream_t s) {
  if (nleaves == 0) return hipSuccess;
  hipLaunchKernelGGL(k_leaf_local, dim3(nleaves), dim3(256), 0, s, leaves, tv.center, tv.range,
                     tv.sx, tv.sy, tv.sz, tv.sq, sloc);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Tree build on the device: stable splits of slot ranges (treedivide,
// tree3d.cpp:1163-1304) for all nodes of a round, and the Morton order inside
// leaves.  Only comparisons and min/max: the results are exact, so the tree is
// the one the host build (and the reference) produces.
// ---------------------------------------------------------------------------
// One of the two slot buffers.
struct SlotBuf {
  int* P;
  double* X;
  double* Y;
  double* Z;
};
__device__ __forceinline__ SlotBuf slot_buf(const SlotArrays& a, int which) {
  return which ? SlotBuf{a.tP, a.tX, a.tY, a.tZ} : SlotBuf{a.P, a.X, a.Y, a.Z};
}
__device__ __forceinline__ double coord_of(const SlotBuf& a, int dir, int i) {
  return dir == 0 ? a.X[i] : (dir == 1 ? a.Y[i] : a.Z[i]);
}

__global__ __launch_bounds__(256) void k_minmax(int n, const double* __restrict__ X,
                                                const double* __restrict__ Y,
                                                const double* __restrict__ Z, double* part) {
  __shared__ double red[6][256];
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double c[3] = {X[i], Y[i], Z[i]};
    for (int d = 0; d < 3; d++) {
      v[d] = fmin(v[d], c[d]);
      v[3 + d] = fmax(v[3 + d], c[d]);
    }
  }
  for (int k = 0; k < 6; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 6; k++)
        red[k][threadIdx.x] = k < 3 ? fmin(red[k][threadIdx.x], red[k][threadIdx.x + w])
                                    : fmax(red[k][threadIdx.x], red[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_nonfinite(int n, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, const double* __restrict__ q, int* flag) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n && !(isfinite(x[t]) && isfinite(y[t]) && isfinite(z[t]) && isfinite(q[t])))
    atomicOr(flag, 1);
}

hipError_t nonfinite(int n, const double* x, const double* y, const double* z, const double* q,
                     int* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_nonfinite, dim3((n + 255) / 256), dim3(256), 0, s, n, x, y, z, q, flag);
  return hipGetLastError();
}

__global__ void k_iota(int n, int* p) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = t;
}

hipError_t iota(int n, int* p, hipStream_t s) {
  hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, n, p);
  return hipGetLastError();
}

hipError_t minmax(int n, const double* X, const double* Y, const double* Z, int blocks,
                  double* part, hipStream_t s) {
  hipLaunchKernelGGL(k_minmax, dim3(blocks), dim3(256), 0, s, n, X, Y, Z, part);
  return hipGetLastError();
}

// Slots [s0, s1) of one split by an NB-thread work-group: how many go to the
// `<=` side, and both sides' boxes (into *out, written by thread 0; the block
// is synchronised on return).  Four slots per thread and pass (their loads in
// flight together); the side decides by select which box a slot widens.
// red / cnt: one entry per wave.
template <int NB>
__device__ __forceinline__ void count_range(const double* __restrict__ X,
                                            const double* __restrict__ Y,
                                            const double* __restrict__ Z, int s0, int s1, int dir,
                                            double mid, double (*red)[NB / 64], int* cnt,
                                            ChunkResult* out) {
  int c = 0;
  // v[k]: min side 0 xyz, min side 1 xyz, max side 0 xyz, max side 1 xyz
  double v[12];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    v[k] = INFINITY;
    v[6 + k] = -INFINITY;
  }
  for (int i0 = s0 + threadIdx.x; i0 < s1; i0 += 4 * NB) {
    double px[4], py[4], pz[4];
    bool in[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = i0 + u * NB;
      in[u] = i < s1;
      const int j = in[u] ? i : s0;
      px[u] = X[j];
      py[u] = Y[j];
      pz[u] = Z[j];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (!in[u]) continue;
      const double p[3] = {px[u], py[u], pz[u]};
      const double cv = dir == 0 ? p[0] : (dir == 1 ? p[1] : p[2]);
      const bool left = cv <= mid;
      c += left;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        v[d] = left ? fmin(v[d], p[d]) : v[d];
        v[3 + d] = left ? v[3 + d] : fmin(v[3 + d], p[d]);
        v[6 + d] = left ? fmax(v[6 + d], p[d]) : v[6 + d];
        v[9 + d] = left ? v[9 + d] : fmax(v[9 + d], p[d]);
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    c += __shfl_xor(c, o);
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const double w = __shfl_xor(v[k], o);
      v[k] = k < 6 ? fmin(v[k], w) : fmax(v[k], w);
    }
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    cnt[wave] = c;
#pragma unroll
    for (int k = 0; k < 12; k++) red[k][wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ChunkResult r;
    r.nl = 0;
    r.pad = 0;
    for (int w = 0; w < NB / 64; w++) r.nl += cnt[w];
    for (int k = 0; k < 12; k++) {
      double x = red[k][0];
      for (int w = 1; w < NB / 64; w++) x = k < 6 ? fmin(x, red[k][w]) : fmax(x, red[k][w]);
      r.v[k] = x;
    }
    *out = r;
  }
  __syncthreads();
}

// A round's chunks: the programs' fixed budgets of kPartChunk-slot chunks,
// laid out once per generation (cbeg[idx] .. cbeg[idx + 1], chunk c's program
// cprog[c]).  Round r's request of program idx is ops[idx] over the first
// nch[idx] chunks of its budget, where nch holds (r << kTagShift) | chunks and
// an entry tagged with another round counts as no request: rounds alternate
// between two (ops, nch) buffers, and a program writes the next round's entry
// only while it has requests.  Returns false for a chunk with no work.
constexpr int kTagShift = 20;
__device__ __forceinline__ bool round_chunk(const int* __restrict__ cprog,
                                            const int* __restrict__ cbeg,
                                            const int* __restrict__ nch,
                                            const int* __restrict__ ntot, int r, int c, int& idx,
                                            int& k, int& nc) {
  if (c >= *ntot) return false;  // grid sized for the most chunks
  idx = cprog[c];
  k = c - cbeg[idx];
  const int v = nch[idx];
  nc = (v >> kTagShift) == r ? v & ((1 << kTagShift) - 1) : 0;
  return k < nc;
}

// Per chunk of a round: count_range of the chunk's slots.
__global__ __launch_bounds__(256) void k_part_count(SlotArrays sa, const PartOpDev* __restrict__ ops,
                                                    const int* __restrict__ cprog,
                                                    const int* __restrict__ cbeg,
                                                    const int* __restrict__ nch,
                                                    const int* __restrict__ ntot, int r,
                                                    ChunkResult* __restrict__ cres) {
  __shared__ double red[12][4];
  __shared__ int cnt[4];
  int idx, k, nc;
  if (!round_chunk(cprog, cbeg, nch, ntot, r, blockIdx.x, idx, k, nc)) return;
  const PartOpDev op = ops[idx];
  const int s0 = op.b + k * kPartChunk, s1 = min(op.e, s0 + kPartChunk);
  count_range<256>(op.src ? sa.tX : sa.X, op.src ? sa.tY : sa.Y, op.src ? sa.tZ : sa.Z, s0, s1,
              op.dir, op.mid, red, cnt, cres + blockIdx.x);
}

// Stable scatter of slots [s0, s1) of one split into the other buffer by an
// NB-thread work-group: the `<=` slots to op_b + lrun + rank, the others to
// op_b + opnl + rrun + rank.  Tiles of 4 NB slots, four rows of NB (slot
// t0 + NB u + thread); a slot's rank among its tile's `<=` slots comes from
// the wave ballots of all four rows, one barrier per tile, and every load of
// the tile is in flight before it.  The block is synchronised on return.
template <int NB>
__device__ __forceinline__ void scatter_range(const SlotBuf& a, const SlotBuf& t, int s0, int s1,
                                              int dir, double mid, int op_b, int opnl, int lrun,
                                              int rrun, int (*wsum)[NB / 64]) {
  const double* __restrict__ C = dir == 0 ? a.X : (dir == 1 ? a.Y : a.Z);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int t0 = s0; t0 < s1; t0 += 4 * NB) {
    bool in[4], left[4];
    int before[4], pv[4];
    double xv[4], yv[4], zv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = t0 + u * NB + threadIdx.x;
      in[u] = i < s1;
      const int j = in[u] ? i : s0;
      pv[u] = a.P[j];
      xv[u] = a.X[j];
      yv[u] = a.Y[j];
      zv[u] = a.Z[j];
      left[u] = in[u] && C[j] <= mid;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const unsigned long long bl = __ballot(left[u]);
      before[u] = __popcll(bl & below);
      if (lane == 0) wsum[u][wave] = __popcll(bl);
    }
    __syncthreads();
    int rowoff = 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      int woff = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < NB / 64; w++) {
        const int c = wsum[u][w];
        woff += w < wave ? c : 0;
        tot += c;
      }
      if (in[u]) {
        const int lrank = rowoff + woff + before[u];
        const int pos = u * NB + threadIdx.x;  // slot's place in the tile
        const int dst = left[u] ? op_b + lrun + lrank : op_b + opnl + rrun + (pos - lrank);
        t.P[dst] = pv[u];
        t.X[dst] = xv[u];
        t.Y[dst] = yv[u];
        t.Z[dst] = zv[u];
      }
      rowoff += tot;
    }
    const int ntile = min(4 * NB, s1 - t0);
    lrun += rowoff;
    rrun += ntile - rowoff;
    __syncthreads();  // wsum is rewritten by the next tile
  }
}

// ---------------------------------------------------------------------------
// Tree build with the split decisions on the device.  Every node of a
// generation runs treemake's loop body (tree3d.cpp:1418-1545) as the same
// resumable split program the host build uses, one thread per node: a round
// applies each program's last split result, advances it to its next split
// request (at most seven per node), and lays the requests out as chunks for
// the partition kernels; after the last round the children are numbered in
// node order (a prefix sum over the generation, as the reference numbers
// them) and written to the node table.  The host only reads the generation's
// size between generations.  The arithmetic that decides the tree (treeinit's
// midpoints, aspect ratios and squared radius, the `<=` tests) rounds as the
// reference's: no contraction.
// ---------------------------------------------------------------------------
namespace {

constexpr int kDone = -1;
constexpr int kF = 13;  // box fields per node
enum { FX0 = 0, FX1, FY0, FY1, FZ0, FZ1, FXMID, FYMID, FZMID, FAXY, FAXZ, FAYZ, FSQR };

// treeinit (tree3d.cpp:1045-1073).
__device__ void box_init_dev(double* f, double x0, double x1, double y0, double y1, double z0,
                             double z1) {
#pragma clang fp contract(off)
  double x = 0.5 * (x1 + x0);
  double y = 0.5 * (y1 + y0);
  double z = 0.5 * (z1 + z0);
  f[FX0] = x0; f[FX1] = x1; f[FY0] = y0; f[FY1] = y1; f[FZ0] = z0; f[FZ1] = z1;
  f[FXMID] = x; f[FYMID] = y; f[FZMID] = z;
  f[FAXY] = fabs((x1 - x0) / (y1 - y0));
  f[FAXZ] = fabs((x1 - x0) / (z1 - z0));
  f[FAYZ] = fabs((y1 - y0) / (z1 - z0));
  x = x1 - x;
  y = y1 - y;
  z = z1 - z;
  f[FSQR] = x * x + y * y + z * z;
}

__device__ bool want_y(const DevSlot& c, double lim, int maxmemb) {
  return c.used && c.f[FAXY] < lim && c.f[FAYZ] > lim && (c.e - c.b) > maxmemb;
}
__device__ bool want_z(const DevSlot& c, double lim, int maxmemb) {
  return c.used && c.f[FAXZ] < lim && c.f[FAYZ] < lim && (c.e - c.b) > maxmemb;
}

__device__ void request(DevProg& P, int dir, const DevSlot& src, int c1, int c2) {
  P.src = src;
  P.c1 = c1;
  P.c2 = c2;
  P.op = PartOpDev{src.b, src.e, dir, src.par, src.f[FXMID + dir]};
}

__device__ bool add_kid(DevProg& P, int slot) {
  if (P.nchild >= 8) {
    P.failed = 1;
    P.pc = kDone;
    return false;
  }
  P.kids[P.nchild++] = slot;
  return true;
}

// Runs the program until it requests a split (returns) or ends (pc = kDone):
// the same states and slot arithmetic as the host's SplitProgram::advance.
__device__ void advance(DevProg& P) {
  const double lim = P.lim;
  const int mm = P.maxmemb;
  for (;;) {
    switch (P.pc) {
      case 0:
        if (P.parent.f[FAXY] > lim && P.parent.f[FAXZ] > lim) {
          request(P, 0, P.parent, 0, 1);
          P.pc = 1;
        } else if (P.parent.f[FAXY] < lim && P.parent.f[FAYZ] > lim) {
          request(P, 1, P.parent, 0, 1);
          P.pc = 20;
        } else {
          request(P, 2, P.parent, 0, 1);
          P.pc = 30;
        }
        return;
      case 1:
        P.xc = P.made;
        P.ind2 = P.xc;
        P.j = 0;
        P.pc = 2;
        break;
      case 2:
        if (P.j >= P.xc) {
          P.pc = kDone;
          return;
        }
        if (!add_kid(P, P.j)) return;
        if (want_y(P.s[P.j], lim, mm)) {
          request(P, 1, P.s[P.j], P.j, P.ind2);
          P.pc = 3;
          return;
        }
        if (want_z(P.s[P.j], lim, mm)) {
          request(P, 2, P.s[P.j], P.j, P.ind2);
          P.pc = 6;
          return;
        }
        P.j++;
        break;
      case 3:
        P.yc = P.made;
        if (P.yc == 2) {
          if (!add_kid(P, P.ind2)) return;
          P.ind2++;
        }
        P.k = 0;
        P.pc = 4;
        break;
      case 4: {
        if (P.k >= P.yc) {
          P.j++;
          P.pc = 2;
          break;
        }
        const int t = P.j + P.k * (P.nchild2 + 1);
        if (t < 16 && want_z(P.s[t], lim, mm)) {
          request(P, 2, P.s[t], t, P.ind2);
          P.pc = 5;
          return;
        }
        P.k++;
        break;
      }
      case 5:
        if (P.made == 2) {
          if (!add_kid(P, P.ind2)) return;
          P.nchild2++;
          P.ind2++;
        }
        P.k++;
        P.pc = 4;
        break;
      case 6:
        if (P.made == 2) {
          if (!add_kid(P, P.ind2)) return;
          P.ind2++;
        }
        P.j++;
        P.pc = 2;
        break;
      case 20:
        P.yc = P.made;
        P.ind2 = P.yc;
        P.j = 0;
        P.pc = 21;
        break;
      case 21:
        if (P.j >= P.yc) {
          P.pc = kDone;
          return;
        }
        if (!add_kid(P, P.j)) return;
        if (want_z(P.s[P.j], lim, mm)) {
          request(P, 2, P.s[P.j], P.j, P.ind2);
          P.pc = 22;
          return;
        }
        P.j++;
        break;
      case 22:
        if (P.made == 2) {
          if (!add_kid(P, P.ind2)) return;
          P.ind2++;
        }
        P.j++;
        P.pc = 21;
        break;
      case 30:
        for (int jj = 0; jj < P.made; jj++)
          if (!add_kid(P, jj)) return;
        P.pc = kDone;
        return;
      default:
        return;
    }
    if (P.ind2 >= 15) {
      P.failed = 1;
      P.pc = kDone;
      return;
    }
  }
}

// treedivide's outcome: the first non-empty side goes to c1, the second to c2.
__device__ void apply(DevProg& P, const ChunkResult& r) {
  const int lo[2] = {P.src.b, P.src.b + r.nl}, hi[2] = {P.src.b + r.nl, P.src.e};
  int target = P.c1;
  P.made = 0;
  for (int sd = 0; sd < 2; sd++) {
    if (hi[sd] > lo[sd]) {
      DevSlot& t = P.s[target];
      t.used = 1;
      t.b = lo[sd];
      t.e = hi[sd];
      t.par = P.src.par ^ 1;
      // r.v: min side0 xyz, min side1 xyz, max side0 xyz, max side1 xyz
      box_init_dev(t.f, r.v[3 * sd], r.v[6 + 3 * sd], r.v[3 * sd + 1], r.v[6 + 3 * sd + 1],
                   r.v[3 * sd + 2], r.v[6 + 3 * sd + 2]);
      P.made++;
      target = P.c2;
    }
  }
  advance(P);
}

}  // namespace

__global__ void k_root_init(int n, int mblocks, const double* __restrict__ part, BuildNodes nd) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double mn[3] = {part[0], part[1], part[2]}, mx[3] = {part[3], part[4], part[5]};
  for (int b = 1; b < mblocks; b++)
    for (int d = 0; d < 3; d++) {
      mn[d] = fmin(mn[d], part[6 * b + d]);
      mx[d] = fmax(mx[d], part[6 * b + 3 + d]);
    }
  box_init_dev(nd.box, mn[0], mx[0], mn[1], mx[1], mn[2], mx[2]);  // tree3d.cpp:1385-1403
  for (int c = 0; c < 8; c++) nd.child[c] = -1;
  nd.begin[0] = 0;
  nd.end[0] = n;
  nd.depth[0] = 0;
  nd.par[0] = 0;
}

// A node's program: treemake's loop body run to its first split request
// (pc = kDone at once for a node of at most maxmemb members).
__device__ void prog_start(DevProg& P, int b, int e, int par, const double* box, int maxmemb,
                           double lim) {
  P.nchild = 0;
  P.failed = 0;
  P.pending = 0;
  P.made = 0;
  P.nchild2 = 0;
  P.xc = P.yc = P.j = P.k = P.ind2 = 0;
  P.maxmemb = maxmemb;
  P.lim = lim;
  for (int sl = 0; sl < 16; sl++) P.s[sl].used = 0;
  if (e - b <= maxmemb) {
    P.pc = kDone;
    return;
  }
  P.parent.used = 1;
  P.parent.b = b;
  P.parent.e = e;
  P.parent.par = par;
  for (int k = 0; k < kF; k++) P.parent.f[k] = box[k];
  P.pc = 0;
  advance(P);
}

// Chunks of kPartChunk slots a node of this size is given for its requests
// (none for a node of at most maxmemb members, which does not split).
__device__ __forceinline__ int chunk_budget(int size, int maxmemb) {
  return size > maxmemb ? (size + kPartChunk - 1) / kPartChunk : 0;
}

// Round 0's entry of a started program: its first request, or none.
__device__ __forceinline__ void first_request(const DevProg& P, int j, int cnt,
                                              const RoundLayout& L) {
  if (P.pc != kDone) {
    L.ops[j] = P.op;
    L.nch[j] = max(1, (P.op.e - P.op.b + kPartChunk - 1) / kPartChunk);  // tag 0
  } else {
    L.nch[j] = 0;
  }
  L.nch[cnt + j] = -1;  // no round's tag
}

// The root's program (generation 0) and its rounds' layout.
__global__ __launch_bounds__(256) void k_prog_root(BuildNodes nd, DevProg* __restrict__ progs,
                                                   int maxmemb, double lim, RoundLayout L) {
  __shared__ int bud;
  if (threadIdx.x == 0) {
    prog_start(progs[0], nd.begin[0], nd.end[0], nd.par[0], nd.box, maxmemb, lim);
    bud = chunk_budget(nd.end[0] - nd.begin[0], maxmemb);
    L.cbeg[0] = 0;
    L.cbeg[1] = bud;
    L.ntot[0] = bud;
    L.ntot[1] = 1;
    first_request(progs[0], 0, 1, L);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < bud; k += 256) L.cprog[k] = 0;
}

// Exclusive scan by one 1024-thread work-group of val(i), i in [0, cnt), in
// index order; emit(i, x, offset) is called for every i.  Returns the total.
template <typename Val, typename Emit>
__device__ int block_scan_f(int cnt, int* red, Val val, Emit emit) {
  int carry = 0;
  for (int base = 0; base < cnt; base += 1024) {
    const int i = base + threadIdx.x;
    const int x = i < cnt ? val(i) : 0;
    red[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int y = threadIdx.x >= (unsigned)off ? red[threadIdx.x - off] : 0;
      __syncthreads();
      red[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < cnt) emit(i, x, carry + red[threadIdx.x] - x);
    const int tot = red[1023];
    __syncthreads();
    carry += tot;
  }
  return carry;
}

// Per chunk of round r: the chunk's bases in its request's two runs (the
// `<=` counts of the request's earlier chunks, all full) and the request's
// `<=` total from the round's count results, then scatter_range.  The
// request's first chunk also forms the request's result (total and both
// sides' boxes), applies it to the program and writes the program's next
// request into the other buffers (ops_n / nch_n, tagged r + 1); a request
// larger than the program's budget fails the program.
__global__ __launch_bounds__(256) void k_part_scatter(SlotArrays sa, const PartOpDev* __restrict__ ops,
                                                      const int* __restrict__ cprog,
                                                      const int* __restrict__ cbeg,
                                                      const int* __restrict__ nch,
                                                      const int* __restrict__ ntot, int r,
                                                      const ChunkResult* __restrict__ cres,
                                                      DevProg* __restrict__ progs,
                                                      PartOpDev* __restrict__ ops_n,
                                                      int* __restrict__ nch_n) {
  __shared__ int wsum[4][4];  // [row][wave] counts of `<=` slots
  __shared__ int part[2][4];
  __shared__ double vred[12][4];
  int idx, k, nc;
  if (!round_chunk(cprog, cbeg, nch, ntot, r, blockIdx.x, idx, k, nc)) return;
  const PartOpDev op = ops[idx];
  const int c0 = cbeg[idx];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int before = 0, total = 0;
  for (int i = threadIdx.x; i < nc; i += 256) {
    const int v = cres[c0 + i].nl;
    total += v;
    before += i < k ? v : 0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    before += __shfl_xor(before, o);
    total += __shfl_xor(total, o);
  }
  if (lane == 0) {
    part[0][wave] = before;
    part[1][wave] = total;
  }
  if (k == 0) {  // the request's boxes: min / max over its chunks
    double v[12];
#pragma unroll
    for (int q = 0; q < 6; q++) {
      v[q] = INFINITY;
      v[6 + q] = -INFINITY;
    }
    for (int i = threadIdx.x; i < nc; i += 256) {
      const ChunkResult& x = cres[c0 + i];
#pragma unroll
      for (int q = 0; q < 6; q++) {
        v[q] = fmin(v[q], x.v[q]);
        v[6 + q] = fmax(v[6 + q], x.v[6 + q]);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
      for (int q = 0; q < 12; q++) {
        const double w = __shfl_xor(v[q], o);
        v[q] = q < 6 ? fmin(v[q], w) : fmax(v[q], w);
      }
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < 12; q++) vred[q][wave] = v[q];
  }
  __syncthreads();
  const int lrun = part[0][0] + part[0][1] + part[0][2] + part[0][3];
  const int opnl = part[1][0] + part[1][1] + part[1][2] + part[1][3];
  const int s0 = op.b + k * kPartChunk, s1 = min(op.e, s0 + kPartChunk);
  scatter_range<256>(slot_buf(sa, op.src), slot_buf(sa, op.src ^ 1), s0, s1, op.dir, op.mid,
                     op.b, opnl, lrun, k * kPartChunk - lrun, wsum);
  if (k == 0 && threadIdx.x == 0) {
    ChunkResult res;
    res.nl = opnl;
    res.pad = 0;
    for (int q = 0; q < 12; q++) {
      double x = vred[q][0];
      for (int w = 1; w < 4; w++) x = q < 6 ? fmin(x, vred[q][w]) : fmax(x, vred[q][w]);
      res.v[q] = x;
    }
    DevProg& P = progs[idx];
    apply(P, res);
    if (P.pc != kDone) {
      const int c = max(1, (P.op.e - P.op.b + kPartChunk - 1) / kPartChunk);
      if (c > cbeg[idx + 1] - c0) {
        P.failed = 1;
        P.pc = kDone;
      } else {
        ops_n[idx] = P.op;
        nch_n[idx] = ((r + 1) << kTagShift) | c;
      }
    }
  }
}

// After a generation: each program's children (none, and the failure flag,
// for a program that overflowed its block or did not finish), numbered in
// node order by a prefix sum over the generation, as the reference numbers
// them; the next generation's end and work; when the next generation runs in
// rounds, its chunk budgets laid out in program (child) order in L.
// res may be host memory: its fields, then seq = depth (system scope).
__global__ __launch_bounds__(1024) void k_prog_number(int gen_hi, int depth, int cnt, int maxmemb,
                                                      const DevProg* __restrict__ progs,
                                                      int* __restrict__ nkid,
                                                      int* __restrict__ start,
                                                      GenResult* __restrict__ res,
                                                      RoundLayout L) {
  __shared__ int red[1024];
  __shared__ int big, bigsize, failed;
  if (threadIdx.x == 0) {
    big = 0;
    bigsize = 0;
    failed = 0;
  }
  __syncthreads();
  const int total = block_scan_f(
      cnt, red,
      [&](int i) {
        const DevProg& P = progs[i];
        if (P.failed || P.pc != kDone) {
          atomicOr(&failed, 1);
          return 0;
        }
        for (int sl = 0; sl < P.nchild; sl++) {
          const int sz = P.s[sl].e - P.s[sl].b;
          if (sz > maxmemb) {
            atomicAdd(&big, 1);
            atomicMax(&bigsize, sz);
          }
        }
        return P.nchild;
      },
      [&](int i, int x, int off) {
        nkid[i] = x;
        start[i] = off;
      });
  __syncthreads();
  if (bigsize > kNodeProgramMax) {  // the next generation runs in rounds
    // (thread t handles the same programs as in the scan above: it reads
    // the nkid / start it wrote)
    const int chunks = block_scan_f(
        cnt, red,
        [&](int i) {
          const DevProg& P = progs[i];
          int b = 0;
          for (int sl = 0; sl < nkid[i]; sl++) b += chunk_budget(P.s[sl].e - P.s[sl].b, maxmemb);
          return b;
        },
        [&](int i, int x, int off) {
          const DevProg& P = progs[i];
          const int j0 = start[i];
          for (int sl = 0; sl < nkid[i]; sl++) {
            const int b = chunk_budget(P.s[sl].e - P.s[sl].b, maxmemb);
            L.cbeg[j0 + sl] = off;
            for (int k = 0; k < b; k++) L.cprog[off + k] = j0 + sl;
            off += b;
          }
        });
    if (threadIdx.x == 0) {
      L.cbeg[total] = chunks;
      L.ntot[0] = chunks;
    }
  }
  if (threadIdx.x == 0) {
    L.ntot[1] = total;
    res->next = gen_hi + total;
    res->failed = failed;
    res->maxdepth = total > 0 ? depth : 0;
    res->splits = big;
    res->maxsize = bigsize;
    __hip_atomic_store(&res->seq, depth, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The generation's children: thread (program, slot) writes node gen_hi +
// start + slot of the node table and the parent's child `slot`, and starts
// the node's program for the next generation (with its round-0 entry in L).
__global__ __launch_bounds__(256) void k_prog_write(int gen_lo, int gen_hi, int depth, int cnt,
                                                    int maxmemb, double lim,
                                                    const DevProg* __restrict__ progs,
                                                    const int* __restrict__ nkid,
                                                    const int* __restrict__ start, BuildNodes nd,
                                                    DevProg* __restrict__ next, RoundLayout L) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int idx = t >> 3, sl = t & 7;
  if (idx >= cnt || sl >= nkid[idx]) return;
  const DevProg& P = progs[idx];
  const int st = start[idx];
  const int g = gen_hi + st + sl;
  const DevSlot& S = P.s[sl];
  for (int k = 0; k < kF; k++) nd.box[(size_t)g * kF + k] = S.f[k];
  nd.begin[g] = S.b;
  nd.end[g] = S.e;
  nd.depth[g] = depth;
  nd.par[g] = S.par;
  for (int c = 0; c < 8; c++) nd.child[(size_t)g * 8 + c] = -1;
  nd.child[(size_t)(gen_lo + idx) * 8 + sl] = gen_hi + st + P.kids[sl];
  DevProg& Q = next[st + sl];
  prog_start(Q, S.b, S.e, S.par, S.f, maxmemb, lim);
  first_request(Q, st + sl, L.ntot[1], L);
}

// A generation of small nodes: one work-group runs one node's whole split
// program, each requested split a count pass and a stable scatter over the
// node's own range (no rounds across the generation).  The generation's
// time is its largest node's chain of splits: NB = 1024 threads for
// generations with large nodes, 256 otherwise.
template <int NB>
__global__ __launch_bounds__(NB) void k_node_program(SlotArrays sa, int cnt,
                                                     DevProg* __restrict__ progs) {
  constexpr int NW = NB / 64;
  __shared__ double red[12][NW];
  __shared__ int cnts[NW];
  __shared__ int wsum[4][NW];
  __shared__ ChunkResult res;
  __shared__ PartOpDev op;
  __shared__ int done;
  const int idx = blockIdx.x;
  if (idx >= cnt) return;
  DevProg& P = progs[idx];
  for (;;) {
    if (threadIdx.x == 0) {
      done = P.pc == kDone;
      op = P.op;
    }
    __syncthreads();
    if (done) break;
    const PartOpDev o = op;
    const SlotBuf a = slot_buf(sa, o.src);
    count_range<NB>(a.X, a.Y, a.Z, o.b, o.e, o.dir, o.mid, red, cnts, &res);
    scatter_range<NB>(a, slot_buf(sa, o.src ^ 1), o.b, o.e, o.dir, o.mid, o.b, res.nl, 0, 0,
                      wsum);
    if (threadIdx.x == 0) apply(P, res);
    __syncthreads();
  }
}

hipError_t build_root(int n, int mblocks, const double* part, const BuildNodes& nd,
                      DevProg* progs, int maxmemb, double lim, const RoundLayout& L,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_root_init, dim3(1), dim3(64), 0, s, n, mblocks, part, nd);
  hipLaunchKernelGGL(k_prog_root, dim3(1), dim3(256), 0, s, nd, progs, maxmemb, lim, L);
  return hipGetLastError();
}

hipError_t build_generation(const SlotArrays& a, int n, int gen_lo, int gen_hi, int depth,
                            int maxmemb, double lim, int maxsize, const BuildNodes& nd,
                            const BuildScratch& w, hipStream_t s) {
  const int cnt = gen_hi - gen_lo;
  const int maxch = (n + kPartChunk - 1) / kPartChunk + cnt;
  if (maxsize <= kNodeProgramWide) {
    hipLaunchKernelGGL(k_node_program<256>, dim3(cnt), dim3(256), 0, s, a, cnt, w.progs);
  } else if (maxsize <= kNodeProgramMax) {
    hipLaunchKernelGGL(k_node_program<1024>, dim3(cnt), dim3(1024), 0, s, a, cnt, w.progs);
  } else {
    const RoundLayout& L = w.cur;
    for (int r = 0; r < kMaxSplitsPerNode; r++) {
      const int cur = (r & 1) * cnt, nxt = ((r + 1) & 1) * cnt;
      hipLaunchKernelGGL(k_part_count, dim3(maxch), dim3(256), 0, s, a,
                         (const PartOpDev*)L.ops + cur, (const int*)L.cprog, (const int*)L.cbeg,
                         (const int*)L.nch + cur, (const int*)L.ntot, r, w.cres);
      hipLaunchKernelGGL(k_part_scatter, dim3(maxch), dim3(256), 0, s, a,
                         (const PartOpDev*)L.ops + cur, (const int*)L.cprog, (const int*)L.cbeg,
                         (const int*)L.nch + cur, (const int*)L.ntot, r,
                         (const ChunkResult*)w.cres, w.progs, L.ops + nxt, L.nch + nxt);
    }
  }
  hipLaunchKernelGGL(k_prog_number, dim3(1), dim3(1024), 0, s, gen_hi, depth, cnt, maxmemb,
                     (const DevProg*)w.progs, w.nkid, w.start, w.result, w.nxt);
  hipLaunchKernelGGL(k_prog_write, dim3((cnt * 8 + 255) / 256), dim3(256), 0, s, gen_lo, gen_hi,
                     depth, cnt, maxmemb, lim, (const DevProg*)w.progs, (const int*)w.nkid,
                     (const int*)w.start, nd, w.next, w.nxt);
  return hipGetLastError();
}

// Morton order inside each leaf (same key as the host build): bitonic sort of
// (key, position) pairs in LDS, then a gather into the first buffer (through
// the second for leaves held in the first).  Leaves larger than kMortonMax keep
// their order.
__device__ __forceinline__ unsigned spread3_dev(unsigned v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ unsigned quant_dev(double v, double lo, double hi) {
#pragma clang fp contract(off)
  if (!(hi > lo)) return 0;
  double t = (v - lo) / (hi - lo) * 1023.0;
  if (t < 0) t = 0;
  if (t > 1023) t = 1023;
  return (unsigned)t;
}

// One block per node of the table; internal nodes return at once.  A leaf's
// slots are in the buffer par names; its box is the node's (treeinit fields).
__global__ __launch_bounds__(256) void k_leaf_morton(SlotArrays sa, BuildNodes nd) {
  __shared__ unsigned long long key[kMortonMax];
  const int node = blockIdx.x;
  if (nd.child[(size_t)node * 8] >= 0) return;
  const double* f = nd.box + (size_t)node * kF;
  const LeafBox L{nd.begin[node], nd.end[node], nd.par[node], 0,
                  f[FX0],         f[FX1],       f[FY0],       f[FY1], f[FZ0], f[FZ1]};
  const int n = L.e - L.b;
  const SlotBuf a = slot_buf(sa, L.src), out = slot_buf(sa, 0);
  if (n <= 1 || n > kMortonMax) {
    if (L.src)
      for (int i = threadIdx.x; i < n; i += 256) {
        out.P[L.b + i] = a.P[L.b + i];
        out.X[L.b + i] = a.X[L.b + i];
        out.Y[L.b + i] = a.Y[L.b + i];
        out.Z[L.b + i] = a.Z[L.b + i];
      }
    return;
  }
  int m = 1;
  while (m < n) m <<= 1;
  for (int i = threadIdx.x; i < m; i += 256) {
    unsigned long long k = ~0ull;
    if (i < n) {
      const int s = L.b + i;
      unsigned mk = spread3_dev(quant_dev(a.X[s], L.x0, L.x1)) |
                    (spread3_dev(quant_dev(a.Y[s], L.y0, L.y1)) << 1) |
                    (spread3_dev(quant_dev(a.Z[s], L.z0, L.z1)) << 2);
      k = ((unsigned long long)mk << 32) | (unsigned)i;
    }
    key[i] = k;
  }
  __syncthreads();
  for (int size = 2; size <= m; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < m; i += 256) {
        int j = i ^ stride;
        if (j > i) {
          bool up = (i & size) == 0;
          unsigned long long u = key[i], v = key[j];
          if ((u > v) == up) {
            key[i] = v;
            key[j] = u;
          }
        }
      }
      __syncthreads();
    }
  }
  // gather into the other buffer, then (leaves held in the first buffer) back
  const SlotBuf g = slot_buf(sa, L.src ^ 1);
  for (int i = threadIdx.x; i < n; i += 256) {
    const int src = L.b + (int)(key[i] & 0xffffffffu);
    g.P[L.b + i] = a.P[src];
    g.X[L.b + i] = a.X[src];
    g.Y[L.b + i] = a.Y[src];
    g.Z[L.b + i] = a.Z[src];
  }
  if (L.src == 0) {
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += 256) {
      out.P[L.b + i] = g.P[L.b + i];
      out.X[L.b + i] = g.X[L.b + i];
      out.Y[L.b + i] = g.Y[L.b + i];
      out.Z[L.b + i] = g.Z[L.b + i];
    }
  }
}

// The evaluation's node arrays from a device-built node table: one thread
// per node.
__global__ void k_tree_view(int numtree, BuildNodes nd, double4* __restrict__ center,
                            double* __restrict__ scale, int2* __restrict__ range,
                            int* __restrict__ child) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numtree) return;
  const double* f = nd.box + (size_t)i * kF;
  center[i] = make_double4(f[FXMID], f[FYMID], f[FZMID], f[FSQR]);
  scale[i] = f[FSQR] > 0 ? sqrt(f[FSQR]) : 1.0;
  range[i] = make_int2(nd.begin[i], nd.end[i]);
  for (int c = 0; c < 8; c++) child[(size_t)i * 8 + c] = nd.child[(size_t)i * 8 + c];
}

// One block per node, leaves only: FP32 source copy relative to the leaf's
// box midpoint (as k_leaf_local).
__global__ void k_leaf_local_nodes(const int* __restrict__ child,
                                   const double4* __restrict__ center,
                                   const int2* __restrict__ range, const double* __restrict__ sx,
                                   const double* __restrict__ sy, const double* __restrict__ sz,
                                   const double* __restrict__ sq, float4* sloc) {
  const int leaf = blockIdx.x;
  if (child[(size_t)leaf * 8] >= 0) return;
  const double4 c = center[leaf];
  const int2 r = range[leaf];
  for (int t = r.x + threadIdx.x; t < r.y; t += blockDim.x)
    sloc[t] = make_float4((float)(sx[t] - c.x), (float)(sy[t] - c.y), (float)(sz[t] - c.z),
                          (float)sq[t]);
}

hipError_t tree_view(int numtree, const BuildNodes& nd, const TreeView& tv, float4* sloc,
                     hipStream_t s) {
  if (numtree == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tree_view, dim3((numtree + 255) / 256), dim3(256), 0, s, numtree, nd,
                     const_cast<double4*>(tv.center), const_cast<double*>(tv.scale),
                     const_cast<int2*>(tv.range), const_cast<int*>(tv.child));
  hipLaunchKernelGGL(k_leaf_local_nodes, dim3(numtree), dim3(256), 0, s, tv.child, tv.center,
                     tv.range, tv.sx, tv.sy, tv.sz, tv.sq, sloc);
  return hipGetLastError();
}

hipError_t leaf_morton(const SlotArrays& a, int numtree, const BuildNodes& nd, hipStream_t s) {
  if (numtree == 0) return hipSuccess;
  hipLaunchKernelGGL(k_leaf_morton, dim3(numtree), dim3(256), 0, s, a, nd);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Moments (treemoments, tree3d.cpp:2631-2711), FP64.  A chunk is a slice of one
// node's slot range; 256 threads, one coefficient per thread (two above 256).
// Powers q*dx^i, dy^j, dz^k of 64 particles at a time are staged in LDS.
// Partials are summed per node in chunk order (deterministic).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_moments_partial(
    int nterms, const int* __restrict__ chunk_node, const int2* __restrict__ chunk_rng,
    const double4* __restrict__ center, const double* __restrict__ sx,
    const double* __restrict__ sy, const double* __restrict__ sz,
    const double* __restrict__ sq, double* __restrict__ partial) {
  __shared__ double PX[64][kMaxNterms + 1];
  __shared__ double PY[64][kMaxNterms + 1];
  __shared__ double PZ[64][kMaxNterms + 1];
  const int ch = blockIdx.x;
  const int node = chunk_node[ch];
  const int2 r = chunk_rng[ch];
  const double4 c = center[node];
  const int nc = ncoef(nterms);
  const int tid = threadIdx.x;
  // coefficient(s) of this thread
  int ci[2][3];
  int cnt = 0;
  for (int cc = tid; cc < nc && cnt < 2; cc += 256, cnt++) {
    // invert the packed index
    int n = 0;
    while (ncoef(n + 1) <= cc) n++;
    int rem = cc - ncoef(n);
    int m = 0;
    while ((m + 1) * (m + 2) / 2 <= rem) m++;
    int k = rem - m * (m + 1) / 2;
    ci[cnt][0] = n - m;
    ci[cnt][1] = m - k;
    ci[cnt][2] = k;
  }
  double acc0 = 0.0, acc1 = 0.0;
  for (int base = r.x; base < r.y; base += 64) {
    if (tid < 64) {
      int j = base + tid;
      double px = 0.0, py = 0.0, pz = 0.0;
      if (j < r.y) {
        px = sq[j];
        py = 1.0;
        pz = 1.0;
      }
      double dx = j < r.y ? sx[j] - c.x : 0.0;
      double dy = j < r.y ? sy[j] - c.y : 0.0;
      double dz = j < r.y ? sz[j] - c.z : 0.0;
      for (int k = 0; k < nterms; k++) {
        PX[tid][k] = px;
        PY[tid][k] = py;
        PZ[tid][k] = pz;
        px *= dx;
        py *= dy;
        pz *= dz;
      }
    }
    __syncthreads();
    if (cnt > 0) {
      const int i0 = ci[0][0], j0 = ci[0][1], k0 = ci[0][2];
#pragma unroll 8
      for (int m = 0; m < 64; m++) acc0 = fma(PX[m][i0] * PY[m][j0], PZ[m][k0], acc0);
    }
    if (cnt > 1) {
      const int i1 = ci[1][0], j1 = ci[1][1], k1 = ci[1][2];
#pragma unroll 8
      for (int m = 0; m < 64; m++) acc1 = fma(PX[m][i1] * PY[m][j1], PZ[m][k1], acc1);
    }
    __syncthreads();
  }
  if (cnt > 0) partial[(size_t)ch * nc + tid] = acc0;
  if (cnt > 1) partial[(size_t)ch * nc + tid + 256] = acc1;
}

hipError_t moments_partial(const TreeView& tv, int nterms, int nchunks, const int* chunk_node,
                           const int2* chunk_rng, double* partial, hipStream_t s) {
  if (nchunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_moments_partial, dim3(nchunks), dim3(256), 0, s, nterms, chunk_node,
                     chunk_rng, tv.center, tv.sx, tv.sy, tv.sz, tv.sq, partial);
  return hipGetLastError();
}

__global__ void k_moments_reduce(int numtree, int nterms, const int* __restrict__ node_chunk0,
                                 const int* __restrict__ node_nchunk,
                                 const double* __restrict__ scale,
                                 const double* __restrict__ partial, double* mom,
                                 double* mom_hat64) {
  const int nc = ncoef(nterms);
  long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)numtree * nc) return;
  int node = (int)(g / nc), c = (int)(g % nc);
  int c0 = node_chunk0[node], k = node_nchunk[node];
  double s = 0.0;
  for (int i = 0; i < k; i++) s += partial[(size_t)(c0 + i) * nc + c];
  int n = 0;
  while (ncoef(n + 1) <= c) n++;
  double inv = 1.0 / scale[node], f = 1.0;
  for (int i = 0; i < n; i++) f *= inv;
  mom[g] = s;
  mom_hat64[g] = s * f;
}

hipError_t moments_reduce(const TreeView& tv, int nterms, const int* node_chunk0,
                          const int* node_nchunk, const double* partial, double* mom,
                          double* mom_hat64, hipStream_t s) {
  long long total = (long long)tv.numtree * ncoef(nterms);
  int blocks = (int)((total + 255) / 256);
  hipLaunchKernelGGL(k_moments_reduce, dim3(blocks), dim3(256), 0, s, tv.numtree, nterms,
                     node_chunk0, node_nchunk, tv.scale, partial, mom, mom_hat64);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Upward pass for the internal nodes of one height: node P's moments about its
// centre are the sum over its children C (stored order) of C's moments
// shifted by h = centre_C - centre_P, from the binomial expansion
//   (t + h)^i = sum_{a<=i} C(i,a) h^{i-a} t^a,
// applied one axis at a time in LDS.  Mathematically the same sums as the
// per-particle moments (tree3d.cpp:2631-2711), at a cost per child instead of
// per particle and level.
// ---------------------------------------------------------------------------
constexpr int kMaxCoef = ncoef(kMaxNterms);

// C(n, a) for n < kMaxNterms (exact in FP64) and the (i, j, k) of every
// packed coefficient index.
struct M2MTab {
  double binom[kMaxNterms][kMaxNterms];
  unsigned char ijk[kMaxCoef][3];
  constexpr M2MTab() : binom(), ijk() {
    for (int n = 0; n < kMaxNterms; n++) {
      binom[n][n] = 1.0;
      for (int a = n; a > 0; a--) binom[n][a - 1] = binom[n][a] * a / (n - a + 1);
    }
    for (int n = 0; n < kMaxNterms; n++)
      for (int m = 0; m <= n; m++)
        for (int kk = 0; kk <= m; kk++) {
          const int c = cidx(n - m, m - kk, kk);
          ijk[c][0] = (unsigned char)(n - m);
          ijk[c][1] = (unsigned char)(m - kk);
          ijk[c][2] = (unsigned char)kk;
        }
  }
};
__constant__ M2MTab kM2M = M2MTab();

// sum over a <= idx_e of C(idx_e, a) h^{idx_e - a} src[idx with idx_e = a],
// a descending
__device__ __forceinline__ double shift_1d(const double* src, int e, int i, int j, int k,
                                           double h) {
  int v[3] = {i, j, k};
  const int n = v[e];
  double hp = 1.0, sum = 0.0;
  for (int a = n; a >= 0; a--) {
    v[e] = a;
    sum = fma(kM2M.binom[n][a] * hp, src[cidx(v[0], v[1], v[2])], sum);
    hp *= h;
  }
  return sum;
}

// One work-group per parent; the eight children are shifted together, one
// axis per pass over all (child, coefficient) pairs, then summed per
// coefficient in child order.  NB threads: 1024 for the upper heights, whose
// few parents leave the GPU idle, so each parent's passes finish sooner; the
// arithmetic per coefficient is the same either way.
template <int NB>
__global__ __launch_bounds__(NB) void k_m2m(int nterms, const int* __restrict__ nodes,
                                             const int* __restrict__ child,
                                             const double4* __restrict__ center,
                                             const double* __restrict__ scale,
                                             double* __restrict__ mom,
                                             double* __restrict__ mom_hat64) {
  __shared__ double A[8 * kMaxCoef], B[8 * kMaxCoef];
  __shared__ double H[8][3];
  __shared__ int K[8];
  const int P = nodes[blockIdx.x];
  const int nc = ncoef(nterms);
  const int tid = threadIdx.x;
  if (tid < 8) {
    const int C = child[8 * P + tid];
    K[tid] = C;
    if (C >= 0) {
      const double4 cp = center[P], cc = center[C];
      H[tid][0] = cc.x - cp.x;
      H[tid][1] = cc.y - cp.y;
      H[tid][2] = cc.z - cp.z;
    }
  }
  __syncthreads();
  const int tot = 8 * nc;
  for (int x = tid; x < tot; x += NB) {
    const int q = x / nc, c = x - q * nc;
    A[x] = K[q] >= 0 ? mom[(size_t)K[q] * nc + c] : 0.0;
  }
  __syncthreads();
  for (int x = tid; x < tot; x += NB) {
    const int q = x / nc, c = x - q * nc;
    if (K[q] >= 0)
      B[x] = shift_1d(A + q * nc, 0, kM2M.ijk[c][0], kM2M.ijk[c][1], kM2M.ijk[c][2], H[q][0]);
  }
  __syncthreads();
  for (int x = tid; x < tot; x += NB) {
    const int q = x / nc, c = x - q * nc;
    if (K[q] >= 0)
      A[x] = shift_1d(B + q * nc, 1, kM2M.ijk[c][0], kM2M.ijk[c][1], kM2M.ijk[c][2], H[q][1]);
  }
  __syncthreads();
  for (int x = tid; x < tot; x += NB) {
    const int q = x / nc, c = x - q * nc;
    if (K[q] >= 0)
      B[x] = shift_1d(A + q * nc, 2, kM2M.ijk[c][0], kM2M.ijk[c][1], kM2M.ijk[c][2], H[q][2]);
  }
  __syncthreads();
  const double inv = 1.0 / scale[P];
  for (int c = tid; c < nc; c += NB) {
    double acc = 0.0;
    for (int q = 0; q < 8; q++)
      if (K[q] >= 0) acc += B[q * nc + c];
    const int n = kM2M.ijk[c][0] + kM2M.ijk[c][1] + kM2M.ijk[c][2];
    double f = 1.0;
    for (int i = 0; i < n; i++) f *= inv;
    const size_t g = (size_t)P * nc + c;
    mom[g] = acc;
    mom_hat64[g] = acc * f;
  }
}

// Heights with fewer parents than this run 1024 threads per parent.
constexpr int kM2MWide = 512;

hipError_t moments_upward(const TreeView& tv, int nterms, int nnodes, const int* nodes,
                          double* mom, double* mom_hat64, hipStream_t s) {
  if (nnodes == 0) return hipSuccess;
  if (nterms < 1 || nterms > kMaxNterms) return hipErrorInvalidValue;
  if (nnodes < kM2MWide)
    hipLaunchKernelGGL(k_m2m<1024>, dim3(nnodes), dim3(1024), 0, s, nterms, nodes, tv.child,
                       tv.center, tv.scale, mom, mom_hat64);
  else
    hipLaunchKernelGGL(k_m2m<256>, dim3(nnodes), dim3(256), 0, s, nterms, nodes, tv.child,
                       tv.center, tv.scale, mom, mom_hat64);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batch interaction lists.  One wavefront walks the tree for its 64 targets
// with a shared depth-first stack of (node, mask) entries, children pushed in
// reverse so they are visited in children[0..7] order, from the node mp.root
// (0, or a subtree as the reference's `ind`).  At each node every
// lane evaluates the reference MAC (tree3d.cpp:3469-3473 / 4687-4691) in FP64
// without contraction, so the decisions are bit-identical to the reference:
//   rs = dx*dx + dy*dy + dz*dz + delta2;  accept = sqradius < acc*(rs - mac_sub)
// accepted lanes -> far item; rejected lanes at a childless node -> near item.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool mac_accept(double px, double py, double pz, double4 c,
                                        MacParams mp) {
#pragma clang fp contract(off)
  double dx = px - c.x;
  double dy = py - c.y;
  double dz = pz - c.z;
  double rs = dx * dx + dy * dy + dz * dz + mp.delta2;
  return c.w < mp.acc * (rs - mp.mac_sub);
}

// Modes of k_lists: count only; fill at the scanned offsets; count and stage
// the items into fixed per-batch regions of cap_far / cap_near entries
// (entries beyond a region's capacity are counted, not written).
enum { kListCount = 0, kListFill = 1, kListStage = 2 };

// Near items with at most this many targets are evaluated source-parallel
// (FP32: measured best among 24, 32, 40 and 48 at the 4M-blob point; FP64: 32).
#ifndef BIT_NEAR_SPARSE32
#define BIT_NEAR_SPARSE32 40
#endif
constexpr int kNearSparse32 = BIT_NEAR_SPARSE32;
constexpr int kNearSparse = 32;

// Pair slots per lane of k_near32 for a near item of pc targets over a leaf of
// len sources: the leaf's length for a target-parallel item, four per 64
// sources per four targets for a source-parallel one.
__device__ __forceinline__ unsigned long long near_item_slots(int pc, unsigned long long len) {
  return pc <= kNearSparse32 ? (unsigned long long)((pc + 3) / 4) * ((len + 63) / 64) * 4 : len;
}

template <int MODE>
__global__ __launch_bounds__(64) void k_lists(TreeView tv, MacParams mp, int stack_cap,
                                              int* __restrict__ far_cnt,
                                              int* __restrict__ near_cnt,
                                              const int* __restrict__ far_off,
                                              const int* __restrict__ near_off, int cap_far,
                                              int cap_near, Item* __restrict__ far_items,
                                              Item* __restrict__ near_items, int* err_flag) {
  extern __shared__ unsigned char smem[];
  unsigned long long* st_mask = reinterpret_cast<unsigned long long*>(smem);
  int* st_node = reinterpret_cast<int*>(st_mask + stack_cap);
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  const int t = b * kBatch + lane;
  const bool valid = t < tv.nt;
  const int tt = valid ? t : tv.nt - 1;
  const double px = tv.tx[tt], py = tv.ty[tt], pz = tv.tz[tt];
  const unsigned long long active = ballot(valid);
  // node data through constant-address-space pointers (the tree is read-only
  // here): the wave-uniform addresses become scalar loads, issued together
  // once per visit
  const cptr<double> gcenter = (cptr<double>)(tv.center);
  const cptr<int> gchild = (cptr<int>)(tv.child);
  if (lane == 0) {
    st_node[0] = mp.root;
    st_mask[0] = active;
  }
  int sp = 1;
  int nf = 0, nn = 0;
  constexpr bool kWrite = MODE != kListCount;
  const size_t fbase = MODE == kListFill    ? (size_t)far_off[b]
                       : MODE == kListStage ? (size_t)b * (size_t)cap_far
                                            : 0;
  const size_t nbase = MODE == kListFill    ? (size_t)near_off[b]
                       : MODE == kListStage ? (size_t)b * (size_t)cap_near
                                            : 0;
  // entries a batch may write: all at the scanned offsets, cap when staging
  const int flim = MODE == kListStage ? cap_far : INT_MAX;
  const int nlim = MODE == kListStage ? cap_near : INT_MAX;
  // Items are gathered in registers, entry k of each run of 64 in lane k, and
  // written a run at a time (one coalesced store instead of 64 single-lane ones).
  unsigned long long fbm = 0ull, nbm = 0ull;
  int fbn = 0, nbn = 0;
  __syncthreads();
  while (sp > 0) {
    sp--;
    const int node = uniform(st_node[sp]);
    const unsigned long long m = uniform64(st_mask[sp]);
    __syncthreads();
    const unsigned un = (unsigned)node;  // 32-bit offsets: no sign extension
    const double4 c = make_double4(gcenter[4u * un], gcenter[4u * un + 1u],
                                   gcenter[4u * un + 2u], gcenter[4u * un + 3u]);
    int kids[8];
#pragma unroll
    for (int q = 0; q < 8; q++) kids[q] = gchild[8u * un + (unsigned)q];
    const bool acc = mac_accept(px, py, pz, c, mp);
    const unsigned long long am = ballot(acc) & m;
    if (am) {
      if (kWrite) {
        if (lane == (nf & 63)) {
          fbm = am;
          fbn = node;
        }
        if ((nf & 63) == 63 && nf - 63 + lane < flim)
          far_items[fbase + (nf - 63 + lane)] = Item{fbm, fbn, 0};
      }
      nf++;
    }
    const unsigned long long rm = m & ~am;
    if (rm) {
      if (kids[0] < 0) {
        if (kWrite) {
          if (lane == (nn & 63)) {
            nbm = rm;
            nbn = node;
          }
          if ((nn & 63) == 63 && nn - 63 + lane < nlim)
            near_items[nbase + (nn - 63 + lane)] = Item{nbm, nbn, 0};
        }
        nn++;
      } else {
        int kid = kids[0];  // child `lane` (lanes 0..7)
#pragma unroll
        for (int q = 1; q < 8; q++) kid = (lane == q) ? kids[q] : kid;
        // children fill the first slots, -1 after: the leading valid ones
        const int nch = __builtin_ctzll(~ballot(lane < 8 && kid >= 0));
        if (sp + nch > stack_cap) {
          if (lane == 0) atomicOr(err_flag, 1);
          break;
        }
        if (lane < nch) {  // child `lane` goes to slot sp + nch - 1 - lane
          st_node[sp + (nch - 1 - lane)] = kid;
          st_mask[sp + (nch - 1 - lane)] = rm;
        }
        sp += nch;
      }
    }
    __syncthreads();
  }
  if (kWrite) {  // the last, partial runs
    const int rf = nf & 63, rn = nn & 63;
    if (lane < rf && nf - rf + lane < flim) far_items[fbase + (nf - rf + lane)] = Item{fbm, fbn, 0};
    if (lane < rn && nn - rn + lane < nlim)
      near_items[nbase + (nn - rn + lane)] = Item{nbm, nbn, 0};
  }
  if (MODE != kListFill && lane == 0) {
    far_cnt[b] = nf;
    near_cnt[b] = nn;
  }
}

// Totals and maxima of the per-batch list counts (one work-group, fixed order).
__global__ __launch_bounds__(1024) void k_list_stats(int nb, const int* __restrict__ far_cnt,
                                                     const int* __restrict__ near_cnt,
                                                     ListStats* __restrict__ out) {
  __shared__ unsigned long long red[4][1024];
  unsigned long long v[4] = {0, 0, 0, 0};  // two sums, two maxima
  for (int b = threadIdx.x; b < nb; b += 1024) {
    const unsigned long long f = (unsigned long long)far_cnt[b];
    const unsigned long long n = (unsigned long long)near_cnt[b];
    v[0] += f;
    v[1] += n;
    v[2] = f > v[2] ? f : v[2];
    v[3] = n > v[3] ? n : v[3];
  }
  for (int k = 0; k < 4; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int k = 0; k < 2; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
      for (int k = 2; k < 4; k++) {
        const unsigned long long o = red[k][threadIdx.x + w];
        red[k][threadIdx.x] = o > red[k][threadIdx.x] ? o : red[k][threadIdx.x];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = ListStats{red[0][0], red[1][0], red[2][0], red[3][0]};
}

hipError_t list_stats(int nb, const int* far_cnt, const int* near_cnt, ListStats* out,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_list_stats, dim3(1), dim3(1024), 0, s, nb, far_cnt, near_cnt, out);
  return hipGetLastError();
}

hipError_t lists_count(const TreeView& tv, MacParams mp, int stack_cap, int* far_cnt,
                       int* near_cnt, int* err_flag, hipStream_t s) {
  int nb = (tv.nt + kBatch - 1) / kBatch;
  size_t sh = (size_t)stack_cap * (sizeof(unsigned long long) + sizeof(int));
  hipLaunchKernelGGL(k_lists<kListCount>, dim3(nb), dim3(kBatch), sh, s, tv, mp, stack_cap,
                     far_cnt, near_cnt, (const int*)nullptr, (const int*)nullptr, 0, 0,
                     (Item*)nullptr, (Item*)nullptr, err_flag);
  return hipGetLastError();
}

hipError_t lists_stage(const TreeView& tv, MacParams mp, int stack_cap, int* far_cnt,
                       int* near_cnt, int cap_far, int cap_near, Item* far_stage,
                       Item* near_stage, int* err_flag, hipStream_t s) {
  if (cap_far <= 0 || cap_near <= 0) return hipErrorInvalidValue;
  int nb = (tv.nt + kBatch - 1) / kBatch;
  size_t sh = (size_t)stack_cap * (sizeof(unsigned long long) + sizeof(int));
  hipLaunchKernelGGL(k_lists<kListStage>, dim3(nb), dim3(kBatch), sh, s, tv, mp, stack_cap,
                     far_cnt, near_cnt, (const int*)nullptr, (const int*)nullptr, cap_far,
                     cap_near, far_stage, near_stage, err_flag);
  return hipGetLastError();
}

hipError_t lists_fill(const TreeView& tv, MacParams mp, int stack_cap, const int* far_off,
                      const int* near_off, Item* far_items, Item* near_items, int* err_flag,
                      hipStream_t s) {
  int nb = (tv.nt + kBatch - 1) / kBatch;
  size_t sh = (size_t)stack_cap * (sizeof(unsigned long long) + sizeof(int));
  hipLaunchKernelGGL(k_lists<kListFill>, dim3(nb), dim3(kBatch), sh, s, tv, mp, stack_cap,
                     (int*)nullptr, (int*)nullptr, far_off, near_off, 0, 0, far_items,
                     near_items, err_flag);
  return hipGetLastError();
}

// Far items of a batch are ordered for the pairing in windows of this many
// consecutive items (traversal order between windows; each lane holds
// kFarOrderWindow / 64 of a window's items in registers).  At 4M blobs, windows
// of 64 / 128 / 256 / a whole batch: 93.6 / 94.7 / 95.6 / 96.2% of the far
// kernel's lane slots used, 3.7 / 4.9 / 7.4 / 10.3 GB of its L2-to-memory
// traffic per launch (2.9 GB in traversal order), far kernel time within
// 0.2 ms of each other, 128 the fastest.
#ifndef BIT_FAR_ORDER_WINDOW
#define BIT_FAR_ORDER_WINDOW 128
#endif
constexpr int kFarOrderWindow = BIT_FAR_ORDER_WINDOW;
constexpr int kFarOrderPer = kFarOrderWindow / kBatch;
static_assert(kFarOrderWindow % kBatch == 0, "a window is whole 64-item chunks");

__device__ __forceinline__ int item_pc(uint4 v) {
  return __popcll(((unsigned long long)v.y << 32) | v.x);
}
__device__ __forceinline__ int lane_of(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Far items to their scanned offsets, in k_farq's pairing order, and each
// batch's far pair statistics (pstat[4b]: pairs, pstat[4b + 3]: k_farq's
// sweeps); one wavefront per batch.  Batch b's items are read from
// src + b * cap (cap > 0: staging regions) or src + far_off[b] (cap == 0; src
// may be the output itself: each window is read whole before it is written).
//
// k_farq sweeps two consecutive items together when their target counts sum to
// at most 64, else one.  Within each window of kFarOrderWindow consecutive
// items (traversal order between windows, so that batches processed together
// still visit clusters in about the same order and share them in the L2), the
// order below makes that rule form as many pairs as any pairing of the
// window's items can (the two-pointer matching of items sorted by target count):
//   bucket c = the items of c targets (1..64), each in traversal order;
//   repeat, with j the largest non-empty bucket and i the smallest:
//     i == j: emit bucket j (pairs of two when 2j <= 64) and stop;
//     i + j <= 64: emit k = min(|i|, |j|) pairs, each the next item of j then
//                  the next of i;
//     else: emit bucket j, each item alone (no remaining item fits with it).
// An item emitted alone is never followed by one it fits with (j + any
// remaining count > 64), a pair's second item starts no unit, and a window's
// last unit ends at the window's end, so the kernel's rule forms exactly these
// units.  The order depends only on the batch's items in traversal order.
//
// Work per window: an item's rank in its bucket comes from a match of equal
// counts (seven ballots per 64 items); the bucket walk above runs on bucket
// counts (one lane per bucket) and records, per bucket, segments of
// consecutive ranks with their output position and stride (1, or 2 for one
// side of a run of pairs); each item then finds its segment and is stored once.
__global__ __launch_bounds__(64) void k_far_order(int nb, const int* __restrict__ far_off,
                                                  int cap, const uint4* src_all,
                                                  uint4* far_items,
                                                  unsigned long long* __restrict__ pstat) {
  // segments: at most two per walk step, and every step but the last empties
  // a bucket, so at most 2 * 64 + 1
  __shared__ int fill[kBatch];
  __shared__ int seg_first[kBatch];
  __shared__ int4 seg[4 * kBatch];  // (first rank, count, output position, stride)
  __shared__ int seg_next[4 * kBatch];
  const int b = blockIdx.x;
  const int lane = threadIdx.x;
  if (b >= nb) return;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int n = far_off[b + 1] - far_off[b];
  uint4* dst = far_items + far_off[b];
  const uint4* src = cap > 0 ? src_all + (size_t)b * (size_t)cap : src_all + far_off[b];
  unsigned long long pairs = 0, units = 0;  // per lane / wave-uniform
  for (int w0 = 0; w0 < n; w0 += kFarOrderWindow) {
    const int wn = min(kFarOrderWindow, n - w0);
    fill[lane] = 0;
    seg_first[lane] = -1;
    uint4 v[kFarOrderPer];
    int rk[kFarOrderPer];
#pragma unroll
    for (int u = 0; u < kFarOrderPer; u++) {
      const int i = u * kBatch + lane;
      v[u] = i < wn ? src[w0 + i] : make_uint4(0, 0, 0, 0);
    }
    __syncthreads();
    // rank of each item among the window's earlier items of its count
#pragma unroll
    for (int u = 0; u < kFarOrderPer; u++) {
      if (u * kBatch >= wn) break;
      const bool in = u * kBatch + lane < wn;
      const int pc = in ? item_pc(v[u]) : 0;
      unsigned long long same = ballot(in);
#pragma unroll
      for (int bit = 0; bit < 7; bit++) {
        const bool on = (pc >> bit) & 1;
        const unsigned long long m = ballot(on);
        same &= on ? m : ~m;
      }
      const int f = in ? fill[pc - 1] : 0;
      rk[u] = f + __popcll(same & below);
      if (in && (same & below) == 0ull) fill[pc - 1] = f + __popcll(same);
      __syncthreads();
    }
    // the bucket walk (lane l: bucket l + 1)
    int cnt = fill[lane];
    pairs += (unsigned long long)cnt * (unsigned long long)(lane + 1);
    int take = 0, last = -1, p = 0, ns = 0;
    auto add_seg = [&](int bl, int r0, int k, int p0, int stride) {
      const int lb = lane_of(last, bl);
      if (lane == 0) {
        seg[ns] = make_int4(r0, k, p0, stride);
        seg_next[ns] = -1;
        if (lb < 0) seg_first[bl] = ns;
        else seg_next[lb] = ns;
      }
      last = lane == bl ? ns : last;
      ns++;
    };
    for (;;) {
      const unsigned long long nz = ballot(cnt > 0);
      if (!nz) break;
      const int jl = 63 - __builtin_clzll(nz), il = __builtin_ctzll(nz);  // bucket - 1
      const int cj = lane_of(cnt, jl), tj = lane_of(take, jl);
      if (il == jl) {
        add_seg(jl, tj, cj, p, 1);
        units += 2 * (jl + 1) <= kBatch ? (unsigned long long)(cj + 1) / 2 : cj;
        break;
      }
      if ((il + 1) + (jl + 1) <= kBatch) {
        const int k = min(cj, lane_of(cnt, il));
        add_seg(jl, tj, k, p, 2);
        add_seg(il, lane_of(take, il), k, p + 1, 2);
        p += 2 * k;
        units += k;
        cnt -= (lane == jl || lane == il) ? k : 0;
        take += (lane == jl || lane == il) ? k : 0;
      } else {
        add_seg(jl, tj, cj, p, 1);
        p += cj;
        units += cj;
        cnt -= lane == jl ? cj : 0;
        take += lane == jl ? cj : 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kFarOrderPer; u++) {
      if (u * kBatch >= wn) break;
      if (u * kBatch + lane < wn) {
        int s = seg_first[item_pc(v[u]) - 1];
        int4 g = seg[s];
        while (rk[u] >= g.x + g.y) {
          s = seg_next[s];
          g = seg[s];
        }
        dst[w0 + g.z + (rk[u] - g.x) * g.w] = v[u];
      }
    }
    __syncthreads();  // the window's LDS tables are rewritten by the next
  }
  for (int o = 32; o > 0; o >>= 1) pairs += __shfl_xor(pairs, o);
  if (lane == 0) {
    pstat[4 * b] = pairs;
    pstat[4 * b + 3] = units;
  }
}

// Near items to their scanned offsets (from staging regions of cap entries when
// stage is non-null, else read in place) and each batch's near pair statistics
// (pstat[4b + 1]: pairs, pstat[4b + 2]: k_near32's pair slots); one wavefront
// per batch.
__global__ __launch_bounds__(256) void k_near_compact(int nb, const int* __restrict__ near_off,
                                                      int cap, const uint4* __restrict__ stage,
                                                      uint4* __restrict__ near_items,
                                                      const int2* __restrict__ range,
                                                      unsigned long long* __restrict__ pstat) {
  const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (b >= nb) return;
  const int n = near_off[b + 1] - near_off[b];
  uint4* dst = near_items + near_off[b];
  const uint4* src = stage ? stage + (size_t)b * (size_t)cap : dst;
  unsigned long long s0 = 0, s1 = 0;
  for (int i = lane; i < n; i += kBatch) {
    const uint4 v = src[i];
    if (src != dst) dst[i] = v;
    const int2 r = range[(int)v.z];
    const unsigned long long len = (unsigned long long)(r.y - r.x);
    const int pc = item_pc(v);
    s0 += (unsigned long long)pc * len;
    s1 += near_item_slots(pc, len);
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o);
    s1 += __shfl_xor(s1, o);
  }
  if (lane == 0) {
    pstat[4 * b + 1] = s0;
    pstat[4 * b + 2] = s1;
  }
}

// Totals of the per-batch pair statistics (one work-group, fixed order).
__global__ __launch_bounds__(1024) void k_pair_stats(int nb,
                                                     const unsigned long long* __restrict__ pstat,
                                                     unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[4][1024];
  unsigned long long v[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += 1024)
    for (int k = 0; k < 4; k++) v[k] += pstat[4 * b + k];
  for (int k = 0; k < 4; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

hipError_t pair_stats(int nb, const unsigned long long* pstat, unsigned long long* out,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_pair_stats, dim3(1), dim3(1024), 0, s, nb, pstat, out);
  return hipGetLastError();
}

hipError_t far_order(int nb, const int* far_off, int cap, const Item* src, Item* far_items,
                     unsigned long long* pstat, hipStream_t s) {
  static_assert(sizeof(Item) == sizeof(uint4), "items are copied as 16-byte words");
  if (nb == 0) return hipSuccess;
  if (!src) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_far_order, dim3(nb), dim3(kBatch), 0, s, nb, far_off, cap,
                     reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(far_items),
                     pstat);
  return hipGetLastError();
}

hipError_t near_compact(int nb, const int* near_off, int cap, const Item* stage,
                        Item* near_items, const int2* range, unsigned long long* pstat,
                        hipStream_t s) {
  if (nb == 0) return hipSuccess;
  if ((cap > 0) != (stage != nullptr)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_near_compact, dim3((nb + 3) / 4), dim3(256), 0, s, nb, near_off, cap,
                     reinterpret_cast<const uint4*>(stage), reinterpret_cast<uint4*>(near_items),
                     range, pstat);
  return hipGetLastError();
}

hipError_t exclusive_scan(const int* in, int* out, int n, void* tmp, size_t& tmp_bytes,
                          hipStream_t s) {
  return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, n, s);
}

// ---------------------------------------------------------------------------
// Far field.  Per item the wave's lanes in the mask evaluate the cluster's
// Taylor sum in the radial-derivative form (qall): FP32 (k_farq, per-cluster
// values added with Kahan compensation) and FP64 (k_farq64).  Both stage the
// cluster's coefficients in LDS.  The field kernels (k_farq_field) use the
// same form with three coefficient sets per node.
// ---------------------------------------------------------------------------
template <typename T>
struct Acc;
template <>
struct Acc<double> {
  double s = 0.0;
  __device__ void add(double v) { s += v; }
  __device__ double value() const { return s; }
};
template <>
struct Acc<float> {
  float s = 0.0f, c = 0.0f;
  __device__ void add(float v) {
    float y = v - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ double value() const { return (double)s - (double)c; }
};

// Work-groups are dealt to the eight XCDs in turn (work-group i to XCD i mod 8),
// and each XCD has its own L2.  Consecutive batches share most of their
// clusters and leaves, so work-group i takes unit xcd_unit(i, n): runs of
// kXcdRun consecutive units per XCD (for i below the last whole round of
// 8 kXcdRun, the identity after it).  A bijection on [0, n).  At 4M blobs,
// against the identity, runs of 16 / 64 / 128 cut the far kernel's L2 misses
// by about 19 / 47 / 59% and the near kernel's by 34 / 67 / 80% at unchanged times
// (both kernels are bound by issue, not memory); runs of 256 and more slow
// the step (the XCDs' shares of dense regions grow uneven).
constexpr int kXcdRun = 64;
__device__ __forceinline__ int xcd_unit(int i, int n) {
  const int full = n / (8 * kXcdRun) * (8 * kXcdRun);
  if (i >= full) return i;
  const int x = i & 7, s = i >> 3;
  return (s / kXcdRun) * (8 * kXcdRun) + x * kXcdRun + s % kXcdRun;
}

// Work-group -> (system, batch): unit u = xcd_unit(work-group) is batch u of a
// single system (map == nullptr), else entry u of the mixed launch's table.
__device__ __forceinline__ void which(const int2* __restrict__ map, int& sys, int& b) {
  const int u = xcd_unit(blockIdx.x, gridDim.x);
  if (map) {
    int2 m = map[u];
    sys = uniform(m.x);
    b = uniform(m.y);
  } else {
    sys = 0;
    b = u;
  }
}


// Per-target set-up of the sweep for one cluster: the offset from the centre
// is formed in FP64 and rounded, the rest is FP32 (hardware rsqrt and exp2).
// Returns 1/R.
template <int NT, bool SCR>
__device__ __forceinline__ float qsetup(QPair<NT, float>& v, double px, double py, double pz,
                                        double cx, double cy, double cz, float s, float kappa,
                                        float delta2) {
  const float dx = (float)(px - cx), dy = (float)(py - cy), dz = (float)(pz - cz);
  const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, delta2)));
  const float ri = __builtin_amdgcn_rsqf(r2);
  const float R = r2 * ri;
  const float rho = s * ri, f = rho * ri;
  const float lam = kappa * R;
  const float c0 = SCR ? __builtin_amdgcn_exp2f(-1.4426950408889634f * lam) : 1.0f;
  qpair_init<NT, SCR, float>(v, dx * f, dy * f, dz * f, rho * rho, c0, lam);
  return ri;
}

// The same in FP64 throughout.
template <int NT, bool SCR>
__device__ __forceinline__ double qsetup64(QPair<NT, double>& v, double px, double py, double pz,
                                           double cx, double cy, double cz, double s,
                                           double kappa, double delta2) {
  const double dx = px - cx, dy = py - cy, dz = pz - cz;
  const double R = sqrt(dx * dx + dy * dy + dz * dz + delta2);
  const double rinv = 1.0 / R;
  const double rho = s * rinv, f = rho * rinv;
  const double lam = kappa * R;
  const double c0 = SCR ? exp(-lam) : 1.0;
  qpair_init<NT, SCR, double>(v, dx * f, dy * f, dz * f, rho * rho, c0, lam);
  return rinv;
}

// LDS of the FP32 far sweep: two tiles (the second offset by 16 floats onto
// other banks), the packed unit's lane map and per-item results.
constexpr int far32_tiles(int nt) { return 2 * nqpad(nt) + 16; }
struct FarLds32 {
  int tslot[kBatch];      // packed unit: lane j's target lane
  float res[2][kBatch];   // packed unit: per-item results by target lane
};

// One batch's far field (FP32) in the radial-derivative form: per lane, the
// Kahan sum over the batch's cluster items of its target (unspecified on lanes
// past the last target).  The cluster's Q is staged in LDS (the next tiles are
// loaded into registers while the current ones are evaluated) and read as
// broadcasts.  Two consecutive items with at most 64 targets together share
// one sweep: lanes are remapped to the first item's targets, then the
// second's, each reading its item's tile, and the results return to the
// targets' lanes, which add them in list order, so the results are those of
// one item at a time.
template <int NT, bool SCR>
__device__ __forceinline__ double far32_batch(const EvalDesc& d, int b, float* mbuf,
                                              FarLds32& fl) {
  constexpr int NQP = nqpad(NT);
  constexpr int PF = NQP / kBatch;   // Q values per lane per tile
  constexpr int T2 = NQP + 16;       // second tile's offset: other LDS banks
  int* tslot = fl.tslot;
  float (*res)[kBatch] = fl.res;
  const TreeView& tv = d.tv;
  const float kappa32 = (float)d.kappa, delta2_32 = (float)d.delta2;
  const gptr<float> qc = as_global(d.qc32);
  const gptr<Item> items = as_global(d.far_items);
  // read-only node data at wave-uniform addresses: scalar loads
  const cptr<double> center = (cptr<double>)(tv.center);  // xmid, ymid, zmid, sqradius
  const cptr<double> scale = (cptr<double>)(tv.scale);
  const gptr<int> far_off = as_global(d.far_off);
  const int lane = threadIdx.x;
  const int t = b * kBatch + lane;
  const bool valid = t < tv.nt;
  const int tt = valid ? t : tv.nt - 1;
  const double px = tv.tx[tt], py = tv.ty[tt], pz = tv.tz[tt];
  Acc<float> acc;
  const int i0 = uniform(far_off[b]), i1 = uniform(far_off[b + 1]);
  // items [u, u + n): n = 2 when item u+1 exists and the two items have at
  // most 64 targets together (one sweep, lanes remapped to their targets)
  auto unit_len = [&](int u) -> int {
    // the list build orders the items for this rule window by window
    // (k_far_order): a unit never spans two windows
    if (u + 1 >= i1 || (u + 1 - i0) % kFarOrderWindow == 0) return 1;
    const unsigned long long m0 = uniform64(items[u].mask), m1 = uniform64(items[u + 1].mask);
    return __popcll(m0) + __popcll(m1) <= kBatch ? 2 : 1;
  };
  const unsigned long long below = (1ull << lane) - 1ull;
  float pa[PF], pb[PF];
  auto fetch = [&](int u, int n) {  // node tiles are padded to NQP floats
    const int na = uniform(items[u].node);
#pragma unroll
    for (int j = 0; j < PF; j++) pa[j] = qc[(size_t)na * NQP + j * kBatch + lane];
    if (n == 2) {
      const int nb = uniform(items[u + 1].node);
#pragma unroll
      for (int j = 0; j < PF; j++) pb[j] = qc[(size_t)nb * NQP + j * kBatch + lane];
    }
  };
  int u = i0, n = 0;
  if (u < i1) {
    n = unit_len(u);
    fetch(u, n);
  }
  while (u < i1) {
    const unsigned long long m0 = uniform64(items[u].mask);
    const unsigned long long m1 = n == 2 ? uniform64(items[u + 1].mask) : 0ull;
    const int n0 = uniform(items[u].node);
    const int n1 = n == 2 ? uniform(items[u + 1].node) : n0;
#pragma unroll
    for (int j = 0; j < PF; j++) mbuf[j * kBatch + lane] = pa[j];
    if (n == 2) {
#pragma unroll
      for (int j = 0; j < PF; j++) mbuf[T2 + j * kBatch + lane] = pb[j];
    }
    __syncthreads();
    const int un = u + n;
    int nn = 0;
    if (un < i1) {
      nn = unit_len(un);
      fetch(un, nn);
    }
    const bool in0 = (m0 >> lane) & 1ull, in1 = (m1 >> lane) & 1ull;
    if (n == 1) {
      if (in0) {
        const double cx = center[4 * n0], cy = center[4 * n0 + 1], cz = center[4 * n0 + 2];
        QPair<NT, float> v;
        const float rinv = qsetup<NT, SCR>(v, px, py, pz, cx, cy, cz, (float)scale[n0],
                                           kappa32, delta2_32);
        acc.add(sweep32<NT>(v, mbuf) * rinv);
      }
    } else {
      // Packed unit: lane j takes the j-th target of the first item, then the
      // targets of the second; each result returns to its target's lane, which
      // adds the first item's before the second's (list order, as one item at
      // a time).
      const int k0 = __popcll(m0), k1 = __popcll(m1);
      if (in0) tslot[__popcll(m0 & below)] = lane;
      if (in1) tslot[k0 + __popcll(m1 & below)] = lane;
      // both clusters' centres and scales by uniform (scalar) loads, chosen per
      // lane below, instead of per-lane loads on the unit's critical path
      const double c0x = center[4 * n0], c0y = center[4 * n0 + 1], c0z = center[4 * n0 + 2];
      const double c1x = center[4 * n1], c1y = center[4 * n1 + 1], c1z = center[4 * n1 + 2];
      const float s0 = (float)scale[n0], s1 = (float)scale[n1];
      __syncthreads();
      const bool act = lane < k0 + k1, second = lane >= k0;
      const int tl = act ? tslot[lane] : lane;
      const double qx = __shfl(px, tl), qy = __shfl(py, tl), qz = __shfl(pz, tl);
      if (act) {
        QPair<NT, float> v;
        const float rinv = qsetup<NT, SCR>(v, qx, qy, qz, second ? c1x : c0x,
                                           second ? c1y : c0y, second ? c1z : c0z,
                                           second ? s1 : s0, kappa32, delta2_32);
        res[second ? 1 : 0][tl] = sweep32<NT>(v, mbuf + (second ? T2 : 0)) * rinv;
      }
      __syncthreads();
      if (in0) acc.add(res[0][lane]);
      if (in1) acc.add(res[1][lane]);
    }
    __syncthreads();  // the tiles and slots are overwritten next unit
    u = un;
    n = nn;
  }
  return acc.value();
}

// FP32 far field of one batch per work-group, plus the combine with the near
// field's results: out[perm[t]] = far(t) + phi_near[t].  Four waves per SIMD:
// the sweep is chains of dependent FMAs, and one wave's issue rate is well
// below the SIMD's.
template <int NT, bool SCR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_farq(
    const EvalDesc* __restrict__ descs, const int2* __restrict__ map) {
  __shared__ __align__(16) float mbuf[far32_tiles(NT)];
  __shared__ FarLds32 fl;
  int sys, b;
  which(map, sys, b);
  const EvalDesc& d = descs[sys];
  const double f = far32_batch<NT, SCR>(d, b, mbuf, fl);
  const int t = b * kBatch + threadIdx.x;
  if (t < d.tv.nt) d.out[d.tv.tperm[t]] = f + d.phi_near[t];
}

// FP64 far field: the same radial-derivative sweep in FP64 over FP64 cluster
// coefficients, one item at a time (a tile is 6.5 KB at p = 11).
template <int NT, bool SCR>
__device__ __forceinline__ double far64_batch(const EvalDesc& d, int b, double* mbuf) {
  constexpr int NQ = nqcoef(NT);
  constexpr int NQP = nqpad(NT);
  constexpr int PF = NQP / kBatch;  // Q values per lane per tile
  const TreeView& tv = d.tv;
  const double kappa = d.kappa, delta2 = d.delta2;
  const gptr<double> qc = as_global(d.qc64);
  const gptr<Item> items = as_global(d.far_items);
  const gptr<double> center = (gptr<double>)(tv.center);  // xmid, ymid, zmid, sqradius
  const gptr<double> scale = as_global(tv.scale);
  const gptr<int> far_off = as_global(d.far_off);
  const int lane = threadIdx.x;
  const int t = b * kBatch + lane;
  const bool valid = t < tv.nt;
  const int tt = valid ? t : tv.nt - 1;
  const double px = tv.tx[tt], py = tv.ty[tt], pz = tv.tz[tt];
  Acc<double> acc;
  const int i0 = uniform(far_off[b]), i1 = uniform(far_off[b + 1]);
  double pf[PF];
  auto fetch = [&](int node) {  // node tiles are padded to NQP values
#pragma unroll
    for (int j = 0; j < PF; j++) pf[j] = qc[(size_t)node * NQP + j * kBatch + lane];
  };
  if (i0 < i1) fetch(uniform(items[i0].node));
  for (int it = i0; it < i1; it++) {
    const unsigned long long mask = uniform64(items[it].mask);
    const int node = uniform(items[it].node);
#pragma unroll
    for (int j = 0; j < PF; j++) mbuf[j * kBatch + lane] = pf[j];
    __syncthreads();
    if (it + 1 < i1) fetch(uniform(items[it + 1].node));
    if ((mask >> lane) & 1ull) {
      const double cx = center[4 * node], cy = center[4 * node + 1], cz = center[4 * node + 2];
      QPair<NT, double> v;
      const double rinv =
          qsetup64<NT, SCR>(v, px, py, pz, cx, cy, cz, scale[node], kappa, delta2);
      acc.add(qall<NT, double>(std::make_integer_sequence<int, NQ>{}, v, mbuf) * rinv);
    }
    __syncthreads();  // the tile is overwritten next item
  }
  return acc.value();
}

template <int NT, bool SCR>
__global__ __launch_bounds__(64) void k_farq64(const EvalDesc* __restrict__ descs,
                                               const int2* __restrict__ map) {
  __shared__ __align__(16) double mbuf[nqpad(NT)];
  int sys, b;
  which(map, sys, b);
  const EvalDesc& d = descs[sys];
  const double f = far64_batch<NT, SCR>(d, b, mbuf);
  const int t = b * kBatch + threadIdx.x;
  if (t < d.tv.nt) d.out[d.tv.tperm[t]] = f + d.phi_near[t];
}

template <typename T, bool SCR, int NT>
static void launch_far(const EvalDesc* descs, const int2* map, int nblocks, hipStream_t s) {
  if constexpr (std::is_same_v<T, float>)
    hipLaunchKernelGGL((k_farq<NT, SCR>), dim3(nblocks), dim3(kBatch), 0, s, descs, map);
  else
    hipLaunchKernelGGL((k_farq64<NT, SCR>), dim3(nblocks), dim3(kBatch), 0, s, descs, map);
}

template <typename T, bool SCR, int... NTs>
static bool dispatch_far(int nterms, std::integer_sequence<int, NTs...>, const EvalDesc* descs,
                         const int2* map, int nblocks, hipStream_t s) {
  bool done = false;
  ((nterms == NTs + 1 ? (launch_far<T, SCR, NTs + 1>(descs, map, nblocks, s), done = true)
                      : false),
   ...);
  return done;
}

hipError_t far_field(const EvalDesc* descs, const int2* map, int nblocks, int nterms, bool fp32,
                     bool screened, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  using Seq = std::make_integer_sequence<int, kMaxNterms>;
  bool ok;
  if (fp32)
    ok = screened ? dispatch_far<float, true>(nterms, Seq{}, descs, map, nblocks, s)
                  : dispatch_far<float, false>(nterms, Seq{}, descs, map, nblocks, s);
  else
    ok = screened ? dispatch_far<double, true>(nterms, Seq{}, descs, map, nblocks, s)
                  : dispatch_far<double, false>(nterms, Seq{}, descs, map, nblocks, s);
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Near field (leaf direct sum, tree3d.cpp:3573-3621): sources of each listed
// leaf are staged 64 at a time in LDS and read by all lanes as broadcasts.
// FP32: positions in leaf-centre coordinates (the target's offset from the
// leaf centre is formed once in FP64), hardware rsqrt/exp2, groups of 16
// terms summed plainly and added with Kahan compensation.  The self pair is
// excluded only in leaves that hold some of the batch's targets.
// ---------------------------------------------------------------------------
// g + q G(r) for one pair, the sum folded into the last FMA.
template <bool SCR>
__device__ __forceinline__ float pair32(float g, float ox, float oy, float oz, float4 s, float d2,
                                        float kneg2) {
  float dx = ox - s.x, dy = oy - s.y, dz = oz - s.z;
  float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, d2)));
  float ri = __builtin_amdgcn_rsqf(r2);
  if constexpr (SCR) {
    float e = __builtin_amdgcn_exp2f(kneg2 * (r2 * ri));
    return fmaf(s.w * e, ri, g);
  } else {
    return fmaf(s.w, ri, g);
  }
}

template <bool SCR>
__device__ __forceinline__ double pair64(double px, double py, double pz, double4 s, double d2,
                                         double kappa) {
  double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
  double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, d2)));
  double ri = rsqrt(r2);  // one reciprocal square root instead of sqrt and a division
  if constexpr (SCR) return s.w * exp(-kappa * (r2 * ri)) * ri;
  else return s.w * ri;
}


__device__ __forceinline__ float bcast_lane(float x, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}
__device__ __forceinline__ double bcast_lane64(double x, int l) {
  const unsigned long long v = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// A source in the near kernel's coordinates: positions times csf when the
// coordinates are scaled (charge unchanged).
template <bool SCALED>
__device__ __forceinline__ float4 scale_src(float4 q, float csf) {
  if constexpr (SCALED) return make_float4(q.x * csf, q.y * csf, q.z * csf, q.w);
  else return q;
}

// A float4 through a global-address-space pointer (one 16-byte global load;
// HIP's vector types do not bind references in other address spaces).
__device__ __forceinline__ float4 load4(gptr<float4> p, int i) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  const v4 v = ((const __attribute__((address_space(1))) v4*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}

// One source-parallel sweep of the sparse near mode: lane l takes sources
// j0 + l, j0 + l + 64, ... below j1, each against four broadcast targets
// (sx, sy, sz), and sums per target into g.  MODE 0: every pair; 1: the
// target's own slot (own[u]) skipped; 2: sources with R^2 <= eps skipped
// (point targets).  The mode is fixed per item, outside the source loop.
template <bool SCR, bool SCALED, int MODE>
__device__ __forceinline__ void near_sparse_sweep(gptr<float4> sloc, int j0, int j1,
                                                  const float (&sx)[4], const float (&sy)[4],
                                                  const float (&sz)[4], const int (&own)[4],
                                                  float d2, float kneg2, float csf, float eps2,
                                                  float (&g)[4]) {
  for (int j = j0; j < j1; j += 64) {
    const float4 q = scale_src<SCALED>(load4(sloc, j), csf);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if constexpr (MODE == 0) {
        g[u] = pair32<SCR>(g[u], sx[u], sy[u], sz[u], q, d2, kneg2);
      } else if constexpr (MODE == 1) {
        const float v = pair32<SCR>(0.0f, sx[u], sy[u], sz[u], q, d2, kneg2);
        g[u] += j == own[u] ? 0.0f : v;
      } else {
        const float dx = sx[u] - q.x, dy = sy[u] - q.y, dz = sz[u] - q.z;
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, d2)));
        const float v = pair32<SCR>(0.0f, sx[u], sy[u], sz[u], q, d2, kneg2);
        g[u] += r2 > eps2 ? v : 0.0f;
      }
    }
  }
}

// One batch of the FP32 near field.  SCALED (screened kernel): positions are
// multiplied by cs = kappa log2(e), so that exp(-kappa R) = exp2(-R') with
// R' = cs R needs no multiplication per pair; 1/R' = 1/(cs R) is corrected by
// one multiplication by cs per target.  Otherwise cs = 1 and the exponent is
// kneg2 R.
template <bool SCR, bool SCALED>
__device__ __forceinline__ double near32_batch(const EvalDesc& d, int b, float4* tile, double cs) {
  const TreeView& tv = d.tv;
  const float csf = (float)cs;
  const float d2 = (float)(d.delta2 * cs * cs);
  const float eps2 = SCALED ? (float)(1e-14 * cs * cs) : 1e-14f;  // point targets' R^2 cut
  const float kneg2 = SCALED ? -1.0f : (float)(-d.kappa * 1.4426950408889634);
  const int lane = threadIdx.x;
  const int t = b * kBatch + lane;
  const bool valid = t < tv.nt;
  const int tt = valid ? t : tv.nt - 1;
  const double px = tv.tx[tt], py = tv.ty[tt], pz = tv.tz[tt];
  const int t0 = b * kBatch, t1 = min(tv.nt, t0 + kBatch);
  const gptr<float4> sloc = as_global(tv.sloc);
  Acc<float> acc;
  const int i0 = uniform(d.near_off[b]), i1 = uniform(d.near_off[b + 1]);
  for (int it = i0; it < i1; it++) {
    const unsigned long long mask = uniform64(d.near_items[it].mask);
    const int node = uniform(d.near_items[it].node);
    const int2 r = tv.range[node];
    const double4 c = tv.center[node];
    const bool act = (mask >> lane) & 1ull;
    // the sources' operation sequence (offset rounded to FP32, then scaled in
    // FP32), so that a target on a source gives a zero offset exactly
    const float ox = (float)(px - c.x) * csf, oy = (float)(py - c.y) * csf,
                oz = (float)(pz - c.z) * csf;
    // particle targets: the self pair lies in leaves holding some of the
    // batch's targets; point targets: pairs closer than eps are skipped instead
    const bool selfp = tv.self ? (r.x < t1 && r.y > t0) : true;
    if (__popcll(mask) <= kNearSparse32) {
      // Few targets: lanes take the leaf's sources 64 at a time and the wave
      // walks the item's targets in lane order, four at a time (one source
      // load feeds four pairs); a target's sum over the leaf is one wave
      // reduction, added to its lane's accumulator.
      unsigned long long m = mask;
      while (m) {
        int tl[4];
        bool on[4];
        float sx[4], sy[4], sz[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          on[u] = m != 0ull;
          tl[u] = on[u] ? __builtin_ctzll(m) : 0;
          m &= m - 1;
          sx[u] = bcast_lane(ox, tl[u]);
          sy[u] = bcast_lane(oy, tl[u]);
          sz[u] = bcast_lane(oz, tl[u]);
        }
        float g[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const int own[4] = {t0 + tl[0], t0 + tl[1], t0 + tl[2], t0 + tl[3]};
        if (!selfp)
          near_sparse_sweep<SCR, SCALED, 0>(sloc, r.x + lane, r.y, sx, sy, sz, own, d2,
                                                 kneg2, csf, eps2, g);
        else if (tv.self)
          near_sparse_sweep<SCR, SCALED, 1>(sloc, r.x + lane, r.y, sx, sy, sz, own, d2,
                                                 kneg2, csf, eps2, g);
        else
          near_sparse_sweep<SCR, SCALED, 2>(sloc, r.x + lane, r.y, sx, sy, sz, own, d2,
                                                 kneg2, csf, eps2, g);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
          for (int u = 0; u < 4; u++) g[u] += __shfl_xor(g[u], o);
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (on[u] && lane == tl[u]) acc.add(g[u]);
      }
      continue;
    }
    for (int base = r.x; base < r.y; base += 64) {
      const int cnt = min(64, r.y - base);
      if (lane < cnt) {
        tile[lane] = scale_src<SCALED>(load4(sloc, base + lane), csf);
      }
      __syncthreads();
      if (act) {
        if (!selfp) {
          int j = 0;
          for (; j + 16 <= cnt; j += 16) {
            float g = 0.0f;
#pragma unroll
            for (int u = 0; u < 16; u++) g = pair32<SCR>(g, ox, oy, oz, tile[j + u], d2, kneg2);
            acc.add(g);
          }
          float g = 0.0f;
          for (; j < cnt; j++) g = pair32<SCR>(g, ox, oy, oz, tile[j], d2, kneg2);
          acc.add(g);
        } else if (tv.self) {
          const int me = t - base;
          int j = 0;
          for (; j + 16 <= cnt; j += 16) {
            float g = 0.0f;
#pragma unroll
            for (int u = 0; u < 16; u++) {
              float v = pair32<SCR>(0.0f, ox, oy, oz, tile[j + u], d2, kneg2);
              g += (j + u == me) ? 0.0f : v;
            }
            acc.add(g);
          }
          float g = 0.0f;
          for (; j < cnt; j++) {
            float v = pair32<SCR>(0.0f, ox, oy, oz, tile[j], d2, kneg2);
            g += (j == me) ? 0.0f : v;
          }
          acc.add(g);
        } else {  // point targets: skip coincident sources (R^2 <= eps)
          float g = 0.0f;
          for (int j = 0; j < cnt; j++) {
            float4 q = tile[j];
            float dx = ox - q.x, dy = oy - q.y, dz = oz - q.z;
            float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, d2)));
            float v = pair32<SCR>(0.0f, ox, oy, oz, q, d2, kneg2);
            g += r2 > eps2 ? v : 0.0f;
          }
          acc.add(g);
        }
      }
      __syncthreads();
    }
  }
  return SCALED ? (double)acc.value() * cs : acc.value();
}

// One batch's near field (FP32), per lane: the screened kernel in coordinates
// scaled by cs = kappa log2(e) while cs keeps leaf-size offsets and their
// squares far from FP32's range limits.
template <bool SCR>
__device__ __forceinline__ double near32_value(const EvalDesc& d, int b, float4* tile) {
  if constexpr (SCR) {
    const double cs = d.kappa * 1.4426950408889634;
    if (cs >= 0x1p-20 && cs <= 0x1p20) return near32_batch<true, true>(d, b, tile, cs);
  }
  return near32_batch<SCR, false>(d, b, tile, 1.0);
}

// At least six waves per SIMD: left alone the compiler keeps a whole group of
// 16 sources in flight (114 VGPRs, 4 waves); with the bound it keeps fewer
// (66 VGPRs, 7 waves), which hides more of the LDS and transcendental latency.
template <bool SCR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_near32(
    const EvalDesc* __restrict__ descs, const int2* __restrict__ map) {
  __shared__ float4 tile[64];
  int sys, b;
  which(map, sys, b);
  const EvalDesc& d = descs[sys];
  const double v = near32_value<SCR>(d, b, tile);
  const int t = b * kBatch + threadIdx.x;
  if (t < d.tv.nt) d.phi_near[t] = v;
}

template <bool SCR>
__device__ __forceinline__ double near64_batch(const EvalDesc& d, int b, double4* tile) {
  const TreeView& tv = d.tv;
  const double d2 = d.delta2, kappa = d.kappa;
  const int lane = threadIdx.x;
  const int t = b * kBatch + lane;
  const bool valid = t < tv.nt;
  const int tt = valid ? t : tv.nt - 1;
  const double px = tv.tx[tt], py = tv.ty[tt], pz = tv.tz[tt];
  const int t0 = b * kBatch, t1 = min(tv.nt, t0 + kBatch);
  double acc = 0.0;
  const int i0 = uniform(d.near_off[b]), i1 = uniform(d.near_off[b + 1]);
  for (int it = i0; it < i1; it++) {
    const unsigned long long mask = uniform64(d.near_items[it].mask);
    const int node = uniform(d.near_items[it].node);
    const int2 r = tv.range[node];
    const bool act = (mask >> lane) & 1ull;
    // particle targets: the self pair lies in leaves holding some of the
    // batch's targets; point targets: pairs closer than eps are skipped instead
    const bool selfp = tv.self ? (r.x < t1 && r.y > t0) : true;
    if (__popcll(mask) <= kNearSparse) {
      // Few targets: source-parallel as in k_near32, four targets per pass.
      unsigned long long m = mask;
      while (m) {
        int tl[4];
        bool on[4];
        double sx[4], sy[4], sz[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          on[u] = m != 0ull;
          tl[u] = on[u] ? __builtin_ctzll(m) : 0;
          m &= m - 1;
          sx[u] = bcast_lane64(px, tl[u]);
          sy[u] = bcast_lane64(py, tl[u]);
          sz[u] = bcast_lane64(pz, tl[u]);
        }
        double g[4] = {0.0, 0.0, 0.0, 0.0};
        for (int j = r.x + lane; j < r.y; j += 64) {
          const double4 q = make_double4(tv.sx[j], tv.sy[j], tv.sz[j], tv.sq[j]);
#pragma unroll
          for (int u = 0; u < 4; u++) {
            if (tv.self) {
              if (selfp && j == t0 + tl[u]) continue;  // the target's own slot
            } else {  // point targets: skip coincident sources (R^2 <= eps)
              const double dx = sx[u] - q.x, dy = sy[u] - q.y, dz = sz[u] - q.z;
              if (dx * dx + dy * dy + dz * dz + d2 <= 1e-14) continue;
            }
            g[u] += pair64<SCR>(sx[u], sy[u], sz[u], q, d2, kappa);
          }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
          for (int u = 0; u < 4; u++) g[u] += __shfl_xor(g[u], o);
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (on[u] && lane == tl[u]) acc += g[u];
      }
      continue;
    }
    for (int base = r.x; base < r.y; base += 64) {
      const int cnt = min(64, r.y - base);
      if (lane < cnt) {
        int j = base + lane;
        tile[lane] = make_double4(tv.sx[j], tv.sy[j], tv.sz[j], tv.sq[j]);
      }
      __syncthreads();
      if (act) {
        const int me = tv.self && selfp ? t - base : -1;
        for (int j = 0; j < cnt; j++) {
          if (j == me) continue;
          if (!tv.self) {  // point targets: skip coincident sources (R^2 <= eps)
            double4 q = tile[j];
            double dx = px - q.x, dy = py - q.y, dz = pz - q.z;
            if (dx * dx + dy * dy + dz * dz + d2 <= 1e-14) continue;
          }
          acc += pair64<SCR>(px, py, pz, tile[j], d2, kappa);
        }
      }
      __syncthreads();
    }
  }
  return acc;
}

template <bool SCR>
__global__ __launch_bounds__(64) void k_near64(const EvalDesc* __restrict__ descs,
                                               const int2* __restrict__ map) {
  __shared__ double4 tile[64];
  int sys, b;
  which(map, sys, b);
  const EvalDesc& d = descs[sys];
  const double v = near64_batch<SCR>(d, b, tile);
  const int t = b * kBatch + threadIdx.x;
  if (t < d.tv.nt) d.phi_near[t] = v;
}

hipError_t near_field(const EvalDesc* descs, const int2* map, int nblocks, bool fp32,
                      bool screened, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  if (fp32) {
    if (screened)
      hipLaunchKernelGGL(k_near32<true>, dim3(nblocks), dim3(kBatch), 0, s, descs, map);
    else
      hipLaunchKernelGGL(k_near32<false>, dim3(nblocks), dim3(kBatch), 0, s, descs, map);
  } else {
    if (screened)
      hipLaunchKernelGGL(k_near64<true>, dim3(nblocks), dim3(kBatch), 0, s, descs, map);
    else
      hipLaunchKernelGGL(k_near64<false>, dim3(nblocks), dim3(kBatch), 0, s, descs, map);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Persistent launch of a serving round or a mixed evaluation: one wave per
// work-group, as many work-groups as the device holds at once, each taking
// (system, batch) items from a device counter until the counter passes
// nitems (every wave reaches that exit).  An item runs the batch's near field
// and then its far field in the same work-group (their LDS is shared, so a
// barrier separates them) and writes far + near, the combine of the two-kernel
// path, so a target's result is the same either way.  The far sweep is chosen
// per item from the system's nterms and kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int next_item(int* ctr) {
  int v = 0;
  if (threadIdx.x == 0) v = atomicAdd(ctr, 1);
  return uniform(__shfl(v, 0));
}

template <bool SCR, int... NTs>
__device__ __forceinline__ double far32_any(std::integer_sequence<int, NTs...>, int nt,
                                            const EvalDesc& d, int b, float* mbuf,
                                            FarLds32& fl) {
  double r = 0.0;
  (void)((nt == NTs + 1 ? (r = far32_batch<NTs + 1, SCR>(d, b, mbuf, fl), true) : false) || ...);
  return r;
}

template <bool SCR, int... NTs>
__device__ __forceinline__ double far64_any(std::integer_sequence<int, NTs...>, int nt,
                                            const EvalDesc& d, int b, double* mbuf) {
  double r = 0.0;
  (void)((nt == NTs + 1 ? (r = far64_batch<NTs + 1, SCR>(d, b, mbuf), true) : false) || ...);
  return r;
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_serve32(
    const EvalDesc* __restrict__ descs, const int2* __restrict__ items, int nitems,
    int* __restrict__ ctr, int parity) {
  __shared__ __align__(16) float mbuf[far32_tiles(kMaxNterms)];  // near tile aliases it
  __shared__ FarLds32 fl;
  using Seq = std::make_integer_sequence<int, kMaxNterms>;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr[parity ^ 1] = 0;
  for (;;) {
    const int i = next_item(ctr + parity);
    if (i >= nitems) break;
    const int2 m = items[i];
    const int sys = uniform(m.x), b = uniform(m.y);
    const EvalDesc& d = descs[sys];
    const bool scr = d.kappa != 0.0;
    float4* tile = reinterpret_cast<float4*>(mbuf);
    const double nv = scr ? near32_value<true>(d, b, tile) : near32_value<false>(d, b, tile);
    __syncthreads();
    const double fv = scr ? far32_any<true>(Seq{}, d.nterms, d, b, mbuf, fl)
                          : far32_any<false>(Seq{}, d.nterms, d, b, mbuf, fl);
    const int t = b * kBatch + threadIdx.x;
    if (t < d.tv.nt) d.out[d.tv.tperm[t]] = fv + nv;
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void k_serve64(const EvalDesc* __restrict__ descs,
                                                const int2* __restrict__ items, int nitems,
                                                int* __restrict__ ctr, int parity) {
  __shared__ __align__(16) double mbuf[nqpad(kMaxNterms)];  // near tile aliases it
  using Seq = std::make_integer_sequence<int, kMaxNterms>;
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr[parity ^ 1] = 0;
  for (;;) {
    const int i = next_item(ctr + parity);
    if (i >= nitems) break;
    const int2 m = items[i];
    const int sys = uniform(m.x), b = uniform(m.y);
    const EvalDesc& d = descs[sys];
    const bool scr = d.kappa != 0.0;
    double4* tile = reinterpret_cast<double4*>(mbuf);
    const double nv = scr ? near64_batch<true>(d, b, tile) : near64_batch<false>(d, b, tile);
    __syncthreads();
    const double fv = scr ? far64_any<true>(Seq{}, d.nterms, d, b, mbuf)
                          : far64_any<false>(Seq{}, d.nterms, d, b, mbuf);
    const int t = b * kBatch + threadIdx.x;
    if (t < d.tv.nt) d.out[d.tv.tperm[t]] = fv + nv;
    __syncthreads();
  }
}
static_assert(nqpad(kMaxNterms) * sizeof(double) >= 64 * sizeof(double4), "near tile fits");

// Work-groups of one persistent launch on the current device (what its CUs
// hold at once), computed once per device.
static int serve_groups(bool fp32) {
  static std::mutex mu;
  static int cached[2][64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  std::lock_guard<std::mutex> lk(mu);
  int& g = cached[fp32 ? 0 : 1][dev];
  if (g == 0) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    const hipError_t e =
        fp32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_serve32, kBatch, 0)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_serve64, kBatch, 0);
    if (e != hipSuccess || per <= 0 || cus <= 0) return 0;
    g = cus * per;
  }
  return g;
}

hipError_t serve_field(const EvalDesc* descs, const int2* items, int nitems, bool fp32,
                       int* ctr, int parity, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const int groups = serve_groups(fp32);
  if (groups <= 0) return hipErrorInvalidDevice;
  const int nb = std::min(groups, nitems);
  parity &= 1;
  if (fp32)
    hipLaunchKernelGGL(k_serve32, dim3(nb), dim3(kBatch), 0, s, descs, items, nitems, ctr, parity);
  else
    hipLaunchKernelGGL(k_serve64, dim3(nb), dim3(kBatch), 0, s, descs, items, nitems, ctr, parity);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Field of the regularized Coulomb kernel (treeforce, tree3d.cpp:851-971, with
// treetaylorforce 456-810 and the KERNEL_ORDER family 614-791), FP64 or FP32.
// The far term of one cluster, E_e = sum_{|k|<=p} (k_e+1) M_k a_{k+e}, is a
// Taylor sum of degree p+1 with moments M^(e)_k' = k'_e M_{k'-e}: the same
// radial-derivative form as the potential, over three per-node coefficient
// sets Q^(e) (k_qcoef_field) and with the radial derivatives of the
// regularized family G_L = sum_{l<=L} w_l delta^2l R^-(2l+1):
//   ghat_m = R^(2m+1) (R^-1 d/dR)^m G_L = (-1)^m sum_l w_l (delta^2/R^2)^l
//            prod_{i<m} (2l + 2i + 1),
// the sweep's ratios r_m = ghat_{m+1} / ghat_m (Coulomb: -(2m+1)).  Since
// Mhat^(e)_k' = k'_e Mhat_{k'-e} / s, the Q^(e) omit the 1/s and each pair's
// sums are multiplied by 1/(R s).
// ---------------------------------------------------------------------------
// One thread per (node, component e, Q entry); NT = nterms + 1 (degree p + 1).
template <int NT, typename T>
__global__ __launch_bounds__(256) void k_qcoef_field(int numtree, const double* __restrict__ mom_hat,
                                                     T* __restrict__ qf) {
  constexpr int NQ = nqcoef(NT), NQP = nqpad(NT), NCM = ncoef(NT - 1);
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)numtree * 3 * NQP) return;
  const int node = (int)(g / (3 * NQP));
  const int rem = (int)(g - (long long)node * 3 * NQP);
  const int e = rem / NQP, q = rem - e * NQP;
  if (q >= NQ) {
    qf[g] = T(0);
    return;
  }
  const int l[3] = {kQTabDev<NT>.l[q][0], kQTabDev<NT>.l[q][1], kQTabDev<NT>.l[q][2]};
  const int t = kQTabDev<NT>.t[q];
  const double* M = mom_hat + (size_t)node * NCM;
  double sum = 0.0;
  for (int j0 = 0; j0 <= t; j0++)
    for (int j1 = 0; j0 + j1 <= t; j1++) {
      const int j2 = t - j0 - j1;
      int k[3] = {l[0] + 2 * j0, l[1] + 2 * j1, l[2] + 2 * j2};
      const int ke = k[e];
      if (ke == 0) continue;  // M^(e)_k' = k'_e M_{k'-e}
      k[e] -= 1;
      sum = fma((double)ke * M[cidx(k[0], k[1], k[2])],
                kInvFact[j0] * kInvFact[j1] * kInvFact[j2], sum);
    }
  const double w = ldexp(kInvFact[l[0]] * kInvFact[l[1]] * kInvFact[l[2]], -t);
  qf[g] = (T)(((l[0] + l[1] + l[2]) & 1 ? -w : w) * sum);
}

template <int NT, typename T>
static void launch_qcoef_field(int numtree, const double* mom_hat, T* qf, hipStream_t s) {
  const long long total = (long long)numtree * 3 * nqpad(NT);
  hipLaunchKernelGGL((k_qcoef_field<NT, T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     s, numtree, mom_hat, qf);
}

template <typename T, int... NTs>
static bool dispatch_qcoef_field(int nterms, std::integer_sequence<int, NTs...>, int numtree,
                                 const double* mom_hat, T* qf, hipStream_t s) {
  bool done = false;
  ((nterms + 1 == NTs + 2
        ? (launch_qcoef_field<NTs + 2, T>(numtree, mom_hat, qf, s), done = true)
        : false),
   ...);
  return done;
}

template <typename T>
static hipError_t field_coeffs(int numtree, int nterms, const double* mom_hat64, T* qf,
                               hipStream_t s) {
  if (numtree == 0) return hipSuccess;
  if (!dispatch_qcoef_field(nterms, std::make_integer_sequence<int, kMaxNterms - 1>{}, numtree,
                            mom_hat64, qf, s))
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t field_coeffs32(int numtree, int nterms, const double* mom_hat64, float* qf,
                          hipStream_t s) {
  return field_coeffs(numtree, nterms, mom_hat64, qf, s);
}
hipError_t field_coeffs64(int numtree, int nterms, const double* mom_hat64, double* qf,
                          hipStream_t s) {
  return field_coeffs(numtree, nterms, mom_hat64, qf, s);
}

__device__ __forceinline__ float rsqrtT(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrtT(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float rcpT(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpT(double x) { return 1.0 / x; }

// Per-pair values of the field sweep from the target's offset d from the
// centre (rounded to T), the cluster scale s and the family; returns 1/(R s).
template <int NT, bool FAM, typename T>
__device__ __forceinline__ T field_setup(QPair<NT, T>& v, T dx, T dy, T dz, T s, T delta2,
                                         int korder) {
  constexpr int P = NT - 1;
  const T r2 = fmaT(dx, dx, fmaT(dy, dy, fmaT(dz, dz, delta2)));
  const T ri = rsqrtT(r2);
  const T rho = s * ri, f = rho * ri;
  T r[P > 0 ? P : 1];
  T g0 = T(1);
  if constexpr (FAM) {
    // ghat_m = (-1)^m sum_l t_l prod_{i<m} (2l + 2i + 1), t_l = w_l (delta^2/R^2)^l
    constexpr int LM = 8;  // BIT_MAX_KERNEL_ORDER
    T tl[LM + 1];
    const T x = delta2 * ri * ri;
    tl[0] = T(1);
#pragma unroll
    for (int l = 1; l <= LM; l++) tl[l] = l <= korder ? tl[l - 1] * (T(l - 0.5) / T(l)) * x : T(0);
    T gh[P + 1];
#pragma unroll
    for (int m = 0; m <= P; m++) {
      T sum = T(0);
#pragma unroll
      for (int l = 0; l <= LM; l++) {
        sum += tl[l];
        tl[l] *= T(2 * l + 2 * m + 1);
      }
      gh[m] = (m & 1) ? -sum : sum;
    }
#pragma unroll
    for (int m = 0; m < P; m++) r[m] = gh[m + 1] * rcpT(gh[m]);
    g0 = gh[0];
  } else {
#pragma unroll
    for (int m = 0; m < P; m++) r[m] = -T(2 * m + 1);
  }
  v.g[0] = g0;
#pragma unroll
  for (int t = 1; t <= P / 2; t++) v.g[t] = v.g[t - 1] * r[t - 1];
  const T w0 = dx * f, w1 = dy * f, w2 = dz * f;
#pragma unroll
  for (int m = 0; m < P; m++) {
    v.mult[0][m] = r[m] * w0;
    v.mult[1][m] = r[m] * w1;
    v.mult[2][m] = r[m] * w2;
  }
  v.z = rho * rho;
  return ri * rcpT(s);
}

// One wavefront per batch, one lane per target; a far item's three Q^(e) tiles
// are staged in LDS and read as broadcasts.  FP32: the offset from the centre
// is formed in FP64 and rounded, per-cluster sums Kahan-compensated.
template <int NT, bool FAM, typename T>
__global__ __launch_bounds__(64) void k_farq_field(TreeView tv, double delta2, int korder,
                                                   const T* __restrict__ qf,
                                                   const int* __restrict__ far_off,
                                                   const Item* __restrict__ items,
                                                   double* __restrict__ e_far) {
  constexpr int NQ = nqcoef(NT), NQP = nqpad(NT), PF = 3 * NQP / kBatch;
  __shared__ __align__(16) T tile[3 * NQP];
  const int b = xcd_unit(blockIdx.x, gridDim.x), lane = threadIdx.x;
  const int t = b * kBatch + lane;
  const bool valid = t < tv.nt;
  const int tt = valid ? t : tv.nt - 1;
  const double px = tv.tx[tt], py = tv.ty[tt], pz = tv.tz[tt];
  const T d2 = (T)delta2;
  Acc<T> ax, ay, az;
  const int i0 = uniform(far_off[b]), i1 = uniform(far_off[b + 1]);
  for (int it = i0; it < i1; it++) {
    const unsigned long long mask = uniform64(items[it].mask);
    const int node = uniform(items[it].node);
    const T* src = qf + (size_t)node * 3 * NQP;
#pragma unroll
    for (int j = 0; j < PF; j++) tile[j * kBatch + lane] = src[j * kBatch + lane];
    __syncthreads();
    if ((mask >> lane) & 1ull) {
      const double4 c = tv.center[node];
      QPair<NT, T> v;
      const T sc = field_setup<NT, FAM, T>(v, (T)(px - c.x), (T)(py - c.y), (T)(pz - c.z),
                                           (T)tv.scale[node], d2, korder);
      ax.add(qall<NT, T>(std::make_integer_sequence<int, NQ>{}, v, tile) * sc);
      ay.add(qall<NT, T>(std::make_integer_sequence<int, NQ>{}, v, tile + NQP) * sc);
      az.add(qall<NT, T>(std::make_integer_sequence<int, NQ>{}, v, tile + 2 * NQP) * sc);
    }
    __syncthreads();  // the tile is overwritten by the next item
  }
  if (valid) {
    e_far[3 * (size_t)t + 0] = (double)ax.value();
    e_far[3 * (size_t)t + 1] = (double)ay.value();
    e_far[3 * (size_t)t + 2] = (double)az.value();
  }
}

template <int NT, typename T>
static void launch_farq_field(const TreeView& tv, double delta2, int korder, const T* qf,
                              const int* far_off, const Item* items, double* e_far,
                              hipStream_t s) {
  const int nb = (tv.nt + kBatch - 1) / kBatch;
  if (korder > 0)
    hipLaunchKernelGGL((k_farq_field<NT, true, T>), dim3(nb), dim3(kBatch), 0, s, tv, delta2,
                       korder, qf, far_off, items, e_far);
  else
    hipLaunchKernelGGL((k_farq_field<NT, false, T>), dim3(nb), dim3(kBatch), 0, s, tv, delta2, 0,
                       qf, far_off, items, e_far);
}

template <typename T, int... NTs>
static bool dispatch_farq_field(int nterms, std::integer_sequence<int, NTs...>,
                                const TreeView& tv, double delta2, int korder, const T* qf,
                                const int* far_off, const Item* items, double* e_far,
                                hipStream_t s) {
  bool done = false;
  ((nterms + 1 == NTs + 2
        ? (launch_farq_field<NTs + 2, T>(tv, delta2, korder, qf, far_off, items, e_far, s),
           done = true)
        : false),
   ...);
  return done;
}

hipError_t far_field_grad(const TreeView& tv, int nterms, double delta2, int korder, bool fp32,
                          const void* qf, const int* far_off, const Item* far_items,
                          double* e_far, hipStream_t s) {
  if (korder < 0 || korder > 8) return hipErrorInvalidValue;
  if (tv.nt == 0) return hipSuccess;
  const bool ok =
      fp32 ? dispatch_farq_field(nterms, std::make_integer_sequence<int, kMaxNterms - 1>{}, tv,
                                 delta2, korder, (const float*)qf, far_off, far_items, e_far, s)
           : dispatch_farq_field(nterms, std::make_integer_sequence<int, kMaxNterms - 1>{}, tv,
                                 delta2, korder, (const double*)qf, far_off, far_items, e_far, s);
  if (!ok) return hipErrorInvalidValue;
  return hipGetLastError();
}

// FP32 near field of the field: sources in leaf-centre coordinates (sloc), the
// target's offset from the leaf centre formed in FP64 and rounded, groups of 16
// pairs summed plainly and added with Kahan compensation; the family's factor
// chain as eval_grad_kernel_reg (tree3d.cpp:80-91).
__global__ __launch_bounds__(64) void k_near_grad32(TreeView tv, double d2, int korder,
                                                    const int* __restrict__ near_off,
                                                    const Item* __restrict__ items,
                                                    const double* __restrict__ e_far,
                                                    double* __restrict__ out) {
  __shared__ float4 tile[64];
  const int b = xcd_unit(blockIdx.x, gridDim.x), lane = threadIdx.x;
  const int t = b * kBatch + lane;
  const bool valid = t < tv.nt;
  const int tt = valid ? t : tv.nt - 1;
  const double px = tv.tx[tt], py = tv.ty[tt], pz = tv.tz[tt];
  const int t0 = b * kBatch, t1 = min(tv.nt, t0 + kBatch);
  const float dd = (float)d2;
  Acc<float> ax, ay, az;
  const int i0 = uniform(near_off[b]), i1 = uniform(near_off[b + 1]);
  for (int it = i0; it < i1; it++) {
    const unsigned long long mask = uniform64(items[it].mask);
    const int node = uniform(items[it].node);
    const int2 r = tv.range[node];
    const double4 c = tv.center[node];
    const bool act = (mask >> lane) & 1ull;
    const float ox = (float)(px - c.x), oy = (float)(py - c.y), oz = (float)(pz - c.z);
    const bool selfp = tv.self ? (r.x < t1 && r.y > t0) : true;
    for (int base = r.x; base < r.y; base += 64) {
      const int cnt = min(64, r.y - base);
      if (lane < cnt) tile[lane] = tv.sloc[base + lane];
      __syncthreads();
      if (act) {
        const int me = tv.self && selfp ? t - base : -1;
        for (int j0 = 0; j0 < cnt; j0 += 16) {
          float gx = 0.0f, gy = 0.0f, gz = 0.0f;
          const int j1 = min(cnt, j0 + 16);
          for (int j = j0; j < j1; j++) {
            if (j == me) continue;
            const float4 q = tile[j];
            const float dx = ox - q.x, dy = oy - q.y, dz = oz - q.z;
            const float rs = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, dd)));
            const float ri = __builtin_amdgcn_rsqf(rs);
            float temp = ri * ri * ri, f = temp;
            for (int l = 1; l <= korder; l++) {
              f = f * ((l + 0.5f) / l) * (dd * ri * ri);
              temp += f;
            }
            const float wq = q.w * temp;
            gx = fmaf(dx, wq, gx);
            gy = fmaf(dy, wq, gy);
            gz = fmaf(dz, wq, gz);
          }
          ax.add(gx);
          ay.add(gy);
          az.add(gz);
        }
      }
      __syncthreads();
    }
  }
  if (valid) {
    const size_t o = 3 * (size_t)tv.tperm[t];
    out[o + 0] = e_far[3 * (size_t)t + 0] + ax.value();
    out[o + 1] = e_far[3 * (size_t)t + 1] + ay.value();
    out[o + 2] = e_far[3 * (size_t)t + 2] + az.value();
  }
}

__global__ __launch_bounds__(64) void k_near_grad(TreeView tv, double d2, int korder,
                                                  const int* __restrict__ near_off,
                                                  const Item* __restrict__ items,
                                                  const double* __restrict__ e_far,
                                                  double* __restrict__ out) {
  __shared__ double4 tile[64];
  const int b = xcd_unit(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const int t = b * kBatch + lane;
  const bool valid = t < tv.nt;
  const int tt = valid ? t : tv.nt - 1;
  const double px = tv.tx[tt], py = tv.ty[tt], pz = tv.tz[tt];
  const int t0 = b * kBatch, t1 = min(tv.nt, t0 + kBatch);
  double ex = 0, ey = 0, ez = 0;
  const int i0 = uniform(near_off[b]), i1 = uniform(near_off[b + 1]);
  for (int it = i0; it < i1; it++) {
    const unsigned long long mask = uniform64(items[it].mask);
    const int node = uniform(items[it].node);
    const int2 r = tv.range[node];
    const bool act = (mask >> lane) & 1ull;
    // particle targets: the self pair lies in leaves holding some of the
    // batch's targets; point targets: pairs closer than eps are skipped instead
    const bool selfp = tv.self ? (r.x < t1 && r.y > t0) : true;
    for (int base = r.x; base < r.y; base += 64) {
      const int cnt = min(64, r.y - base);
      if (lane < cnt) {
        int j = base + lane;
        tile[lane] = make_double4(tv.sx[j], tv.sy[j], tv.sz[j], tv.sq[j]);
      }
      __syncthreads();
      if (act) {
        const int me = tv.self && selfp ? t - base : -1;
        for (int j = 0; j < cnt; j++) {
          if (j == me) continue;
          double4 s = tile[j];
          double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
          double r2 = dx * dx + dy * dy + dz * dz + d2;
          double w;
          if (korder == 0) {
            w = s.w / (r2 * sqrt(r2));
          } else {  // eval_grad_kernel_reg, tree3d.cpp:80-91
            const double r = sqrt(r2);
            double temp = 1.0 / (r * r2), f = temp;
            for (int l = 1; l <= korder; l++) {
              f = f * (l + 0.5) / l * d2 / r2;
              temp += f;
            }
            w = s.w * temp;
          }
          ex += dx * w;
          ey += dy * w;
          ez += dz * w;
        }
      }
      __syncthreads();
    }
  }
  if (valid) {
    size_t o = 3 * (size_t)tv.tperm[t];
    out[o + 0] = e_far[3 * (size_t)t + 0] + ex;
    out[o + 1] = e_far[3 * (size_t)t + 1] + ey;
    out[o + 2] = e_far[3 * (size_t)t + 2] + ez;
  }
}

hipError_t near_field_grad(const TreeView& tv, double delta2, int korder, bool fp32,
                           const int* near_off, const Item* near_items, const double* e_far,
                           double* out, hipStream_t s) {
  if (tv.nt == 0) return hipSuccess;
  int nb = (tv.nt + kBatch - 1) / kBatch;
  if (fp32)
    hipLaunchKernelGGL(k_near_grad32, dim3(nb), dim3(kBatch), 0, s, tv, delta2, korder,
                       near_off, near_items, e_far, out);
  else
    hipLaunchKernelGGL(k_near_grad, dim3(nb), dim3(kBatch), 0, s, tv, delta2, korder, near_off,
                       near_items, e_far, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Krylov vectors of the GMRES solve (FP64, resident in HBM).  Dot products are
// deterministic: kDotBlocks blocks each sum one fixed contiguous slice in a
// fixed tree order, then one fixed-order sum over the blocks.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dots(int n, const double* __restrict__ V, size_t ld,
                                              const double* __restrict__ w,
                                              double* __restrict__ partial) {
  __shared__ double red[256];
  const int b = blockIdx.x, j = blockIdx.y;
  const int per = (n + kDotBlocks - 1) / kDotBlocks;
  const int i0 = min(n, b * per), i1 = min(n, i0 + per);
  const double* v = V + (size_t)j * ld;
  double s = 0.0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) s = fma(v[i], w[i], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(size_t)j * kDotBlocks + b] = red[0];
}

__global__ __launch_bounds__(kDotBlocks) void k_dots_final(const double* __restrict__ partial,
                                                           double* __restrict__ out) {
  __shared__ double red[kDotBlocks];
  const int j = blockIdx.x;
  red[threadIdx.x] = partial[(size_t)j * kDotBlocks + threadIdx.x];
  __syncthreads();
  for (int k = kDotBlocks / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

hipError_t dots(int n, const double* V, size_t ld, int ncols, const double* w, double* partial,
                double* out, hipStream_t s) {
  if (ncols < 1) return hipSuccess;
  hipLaunchKernelGGL(k_dots, dim3(kDotBlocks, ncols), dim3(256), 0, s, n, V, ld, w, partial);
  hipLaunchKernelGGL(k_dots_final, dim3(ncols), dim3(kDotBlocks), 0, s, partial, out);
  return hipGetLastError();
}

// y += sign * sum_{j < ncols} c[j] V_j, j ascending.
__global__ void k_combine_cols(int n, const double* __restrict__ V, size_t ld, int ncols,
                               const double* __restrict__ c, double sign, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = y[i];
  for (int j = 0; j < ncols; j++) acc = fma(sign * c[j], V[(size_t)j * ld + i], acc);
  y[i] = acc;
}

hipError_t combine_cols(int n, const double* V, size_t ld, int ncols, const double* c,
                        double sign, double* y, hipStream_t s) {
  if (ncols < 1) return hipSuccess;
  hipLaunchKernelGGL(k_combine_cols, dim3((n + 255) / 256), dim3(256), 0, s, n, V, ld, ncols, c,
                     sign, y);
  return hipGetLastError();
}

// Elementwise steps of the second-kind operator y = (diag + self w) v + phi(w v)
// and of the Arnoldi process.
__global__ void k_vec(int n, int op, double a, double b, const double* __restrict__ u,
                      const double* __restrict__ v, const double* __restrict__ z,
                      double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (op) {
    case 0: y[i] = u[i] * v[i]; break;                   // charges w v
    case 1: y[i] = (a + b * u[i]) * v[i] + z[i]; break;  // (diag + self w) v + phi
    case 2: y[i] = u[i] - v[i]; break;                   // residual b - A x
    default: y[i] = v[i] * a; break;                     // scaled copy
  }
}

hipError_t vec_op(int n, int op, double a, double b, const double* u, const double* v,
                  const double* z, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_vec, dim3((n + 255) / 256), dim3(256), 0, s, n, op, a, b, u, v, z, y);
  return hipGetLastError();
}

// Batched forms: blockIdx.z (dots) or blockIdx.y picks the solve; blocks past a
// solve's own extent return at once.
__global__ __launch_bounds__(256) void k_dots_multi(const DotJob* __restrict__ jobs) {
  __shared__ double red[256];
  const DotJob J = jobs[blockIdx.z];
  const int b = blockIdx.x, j = blockIdx.y;
  if (j >= J.ncols) return;
  const int per = (J.n + kDotBlocks - 1) / kDotBlocks;
  const int i0 = min(J.n, b * per), i1 = min(J.n, i0 + per);
  const double* v = J.V + (size_t)j * J.n;
  double s = 0.0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) s = fma(v[i], J.w[i], s);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) J.partial[(size_t)j * kDotBlocks + b] = red[0];
}

__global__ __launch_bounds__(kDotBlocks) void k_dots_final_multi(const DotJob* __restrict__ jobs) {
  __shared__ double red[kDotBlocks];
  const DotJob J = jobs[blockIdx.y];
  const int j = blockIdx.x;
  if (j >= J.ncols) return;
  red[threadIdx.x] = J.partial[(size_t)j * kDotBlocks + threadIdx.x];
  __syncthreads();
  for (int k = kDotBlocks / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) J.out[j] = red[0];
}

hipError_t dots_multi(int count, int maxcols, const DotJob* jobs, hipStream_t s) {
  if (count < 1 || maxcols < 1) return hipSuccess;
  hipLaunchKernelGGL(k_dots_multi, dim3(kDotBlocks, maxcols, count), dim3(256), 0, s, jobs);
  hipLaunchKernelGGL(k_dots_final_multi, dim3(maxcols, count), dim3(kDotBlocks), 0, s, jobs);
  return hipGetLastError();
}

__global__ void k_combine_multi(const CombJob* __restrict__ jobs) {
  const CombJob J = jobs[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= J.n) return;
  double acc = J.y[i];
  for (int j = 0; j < J.ncols; j++) acc = fma(J.sign * J.c[j], J.V[(size_t)j * J.n + i], acc);
  J.y[i] = acc;
}

hipError_t combine_multi(int count, int maxn, const CombJob* jobs, hipStream_t s) {
  if (count < 1 || maxn < 1) return hipSuccess;
  hipLaunchKernelGGL(k_combine_multi, dim3((maxn + 255) / 256, count), dim3(256), 0, s, jobs);
  return hipGetLastError();
}

__global__ void k_vec_multi(const VecJob* __restrict__ jobs) {
  const VecJob J = jobs[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= J.n) return;
  switch (J.op) {
    case 0: J.y[i] = J.u[i] * J.v[i]; break;
    case 1: J.y[i] = (J.a + J.b * J.u[i]) * J.v[i] + J.z[i]; break;
    case 2: J.y[i] = J.u[i] - J.v[i]; break;
    case 4: J.y[i] = J.c[i] - ((J.a + J.b * J.u[i]) * J.v[i] + J.z[i]); break;
    default: J.y[i] = J.v[i] * J.a; break;
  }
}

hipError_t vec_multi(int count, int maxn, const VecJob* jobs, hipStream_t s) {
  if (count < 1 || maxn < 1) return hipSuccess;
  hipLaunchKernelGGL(k_vec_multi, dim3((maxn + 255) / 256, count), dim3(256), 0, s, jobs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 direct sum (shielded_pot_particle_ds, tree3d.cpp:1966-2016): 256 targets
// per block, sources split into nsplit slices whose partial sums are added in
// slice order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_direct(TreeView tv, double kappa, double d2,
                                                int ntargets, const int* __restrict__ tslot,
                                                int nsplit, double* __restrict__ partial) {
  __shared__ double4 tile[256];
  const int tb = blockIdx.x, sl = blockIdx.y;
  const int ti = tb * 256 + threadIdx.x;
  const bool valid = ti < ntargets;
  const int me = valid ? tslot[ti] : 0;
  const double px = tv.sx[me], py = tv.sy[me], pz = tv.sz[me];
  const long long per = ((long long)tv.n + nsplit - 1) / nsplit;
  const int s0 = (int)min((long long)tv.n, per * sl), s1 = (int)min((long long)tv.n, per * (sl + 1));
  double acc = 0.0;
  for (int base = s0; base < s1; base += 256) {
    const int cnt = min(256, s1 - base);
    if ((int)threadIdx.x < cnt) {
      int j = base + threadIdx.x;
      tile[threadIdx.x] = make_double4(tv.sx[j], tv.sy[j], tv.sz[j], tv.sq[j]);
    }
    __syncthreads();
    for (int j = 0; j < cnt; j++) {
      if (base + j == me) continue;
      double4 s = tile[j];
      double dx = px - s.x, dy = py - s.y, dz = pz - s.z;
      double R = sqrt(dx * dx + dy * dy + dz * dz + d2);
      acc += s.w * exp(-kappa * R) / R;
    }
    __syncthreads();
  }
  if (valid) partial[(size_t)sl * ntargets + ti] = acc;
}

__global__ void k_direct_reduce(int ntargets, int nsplit, const double* __restrict__ partial,
                                double* out) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= ntargets) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; k++) s += partial[(size_t)k * ntargets + ti];
  out[ti] = s;
}

hipError_t direct_potential(const TreeView& tv, double kappa, double delta2, int ntargets,
                            const int* tslot, int nsplit, double* partial, double* out,
                            hipStream_t s) {
  if (ntargets == 0) return hipSuccess;
  dim3 grid((ntargets + 255) / 256, nsplit);
  hipLaunchKernelGGL(k_direct, grid, dim3(256), 0, s, tv, kappa, delta2, ntargets, tslot,
                     nsplit, partial);
  hipLaunchKernelGGL(k_direct_reduce, dim3((ntargets + 255) / 256), dim3(256), 0, s, ntargets,
                     nsplit, partial, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 direct-sum field (force_particle_ds_noreg/_reg, force_point_ds_noreg/_reg,
// tree3d.cpp:125-330).  Per pair, in the reference's operation order:
//   d = x_source - x_target, rs = |d|^2 + d2, temp = 1/(sqrt(rs) rs) (plus the
//   regularized family's terms, eval_grad_kernel_reg tree3d.cpp:80-91),
//   E += q (-d temp).
// Particle targets (tslot != nullptr) skip their own slot only; point targets
// take every source, so a point on a source gives what the reference gives
// there (0/0 with d2 = 0).  Partial sums over nsplit source slices, added in
// slice order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_direct_field(TreeView tv, double d2, int korder,
                                                      int ntargets, const int* __restrict__ tslot,
                                                      const double* __restrict__ qx,
                                                      const double* __restrict__ qy,
                                                      const double* __restrict__ qz, int nsplit,
                                                      double* __restrict__ partial) {
  __shared__ double4 tile[256];
  const int tb = blockIdx.x, sl = blockIdx.y;
  const int ti = tb * 256 + threadIdx.x;
  const bool valid = ti < ntargets;
  const int tc = valid ? ti : 0;
  int me = -1;
  double px, py, pz;
  if (tslot) {
    me = tslot[tc];
    px = tv.sx[me];
    py = tv.sy[me];
    pz = tv.sz[me];
  } else {
    px = qx[tc];
    py = qy[tc];
    pz = qz[tc];
  }
  const long long per = ((long long)tv.n + nsplit - 1) / nsplit;
  const int s0 = (int)min((long long)tv.n, per * sl), s1 = (int)min((long long)tv.n, per * (sl + 1));
  double ex = 0.0, ey = 0.0, ez = 0.0;
  for (int base = s0; base < s1; base += 256) {
    const int cnt = min(256, s1 - base);
    if ((int)threadIdx.x < cnt) {
      int j = base + threadIdx.x;
      tile[threadIdx.x] = make_double4(tv.sx[j], tv.sy[j], tv.sz[j], tv.sq[j]);
    }
    __syncthreads();
    for (int j = 0; j < cnt; j++) {
      if (base + j == me) continue;
      const double4 s = tile[j];
      const double dx = s.x - px, dy = s.y - py, dz = s.z - pz;
      const double rs = dx * dx + dy * dy + dz * dz + d2;
      const double r = sqrt(rs);
      double temp = 1.0 / (r * rs), f = temp;
      for (int l = 1; l <= korder; l++) {
        f = f * (l + 0.5) / l * d2 / rs;
        temp += f;
      }
      ex += s.w * (-dx * temp);
      ey += s.w * (-dy * temp);
      ez += s.w * (-dz * temp);
    }
    __syncthreads();
  }
  if (valid) {
    double* p = partial + 3 * ((size_t)sl * ntargets + ti);
    p[0] = ex;
    p[1] = ey;
    p[2] = ez;
  }
}

__global__ void k_direct_field_reduce(int ntargets, int nsplit, const double* __restrict__ partial,
                                      double* out) {
  int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= ntargets) return;
  double e[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nsplit; k++)
    for (int c = 0; c < 3; c++) e[c] += partial[3 * ((size_t)k * ntargets + ti) + c];
  for (int c = 0; c < 3; c++) out[3 * (size_t)ti + c] = e[c];
}

hipError_t direct_field(const TreeView& tv, double delta2, int korder, int ntargets,
                        const int* tslot, const double* qx, const double* qy, const double* qz,
                        int nsplit, double* partial, double* out, hipStream_t s) {
  if (ntargets == 0) return hipSuccess;
  dim3 grid((ntargets + 255) / 256, nsplit);
  hipLaunchKernelGGL(k_direct_field, grid, dim3(256), 0, s, tv, delta2, korder, ntargets, tslot,
                     qx, qy, qz, nsplit, partial);
  hipLaunchKernelGGL(k_direct_field_reduce, dim3((ntargets + 255) / 256), dim3(256), 0, s,
                     ntargets, nsplit, partial, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free evaluation: its scratch (the evaluation descriptor and the
// near-field results of up to kNoAllocMaxN targets) lives in this file-scope
// device array, one copy per GPU, held until the process ends.  Host code never
// uses the array's name as an address: noalloc_scratch() resolves each device's
// copy once with hipGetSymbolAddress and hands out that pointer.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // device is current
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

// The descriptor arrives by value as a kernel argument and is stored to the
// scratch by one lane: no host-to-device copy, nothing allocated.
__global__ void k_put_desc(EvalDesc d, EvalDesc* dst) {
  if (threadIdx.x == 0) *dst = d;
}

hipError_t put_desc(const EvalDesc& d, EvalDesc* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_put_desc, dim3(1), dim3(64), 0, s, d, dst);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The FMA issue ceiling of this build's instruction mix (no packed FP32): eight
// independent FMA chains per thread, eight waves per SIMD, in the fastest FMA
// form, v_fmac (the accumulator tied to the destination, both multiplicands in
// VGPRs).  The kernels' VALU throughput is compared against it.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fma_ceiling(T* __restrict__ out, int iters, T b, T c) {
  T a[8], q[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    a[i] = T(threadIdx.x) * T(1e-3) + T(i);
    q[i] = c * T(i + 1) + T(threadIdx.x) * T(1e-6);
  }
  const T m = b + T(threadIdx.x) * T(1e-9);  // per-lane values: VGPR operands
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] = fmaT(q[i], m, a[i]);
  }
  T sum = T(0);
#pragma unroll
  for (int i = 0; i < 8; i++) sum += a[i];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = sum;
}

hipError_t fma_ceiling(int blocks, int iters, bool fp64, void* out, hipStream_t s) {
  if (fp64)
    hipLaunchKernelGGL(k_fma_ceiling<double>, dim3(blocks), dim3(256), 0, s, (double*)out, iters,
                       0.999999, 1e-6);
  else
    hipLaunchKernelGGL(k_fma_ceiling<float>, dim3(blocks), dim3(256), 0, s, (float*)out, iters,
                       0.999999f, 1e-6f);
  return hipGetLastError();
}

}  // namespace dev
}  // namespace bit
