// Device SAH kd-tree builder.  It builds the host builder's tree
// (geom.cpp::build_kdtree, the intent of RefineNode / CalculateCostsOfSplit,
// raytracer/environment.go:181-387) node for node and reference for
// reference, level by level on the GPU.
//
// A level holds the nodes still to be refined, their triangle lists back to
// back.  Per level:
//   1. candidate splits per node and axis: the midpoint, then the vertices of
//      every (1 + n/10)-th listed triangle that lie inside the node;
//   2. left/right counts of every candidate over the node's list, in chunks
//      of 2048 references (integer atomics, so chunk order is immaterial),
//      and the list's extent per axis;
//   3. the cheapest candidate in candidate order (ties to the first, as the
//      sequential scan) against the cost of a leaf;
//   4. a stable partition of each split node's list into its children's
//      lists in the host builder's order: left = negatives then straddlers,
//      right = straddlers then positives in reverse.
// After the last level the tree is renumbered in the host builder's
// depth-first allocation order: the k-th split node in preorder gets
// children 2k + 1 and 2k + 2, and leaves append their lists in preorder.
//
// Float min/max follow std::min / std::max of the host code, including
// which of two equal values (+0 and -0) is kept: the first in list order.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"
#include "vr_internal.h"

namespace vr {
namespace {

constexpr int kKdMaxDepth = 21;         // raytracer/kdtree/constants.go:28
constexpr float kKdCostTrav = 75.0f;    // :25
constexpr float kKdCostIsect = 167.0f;  // :26
constexpr int kAxisCands = 32;          // candidates per axis: at most 1 + 3 * 9
constexpr int kCands = 3 * kAxisCands;
constexpr int kChunk = 2048;            // references per partition chunk
constexpr int kEvalChunk = 512;         // references per candidate-evaluation chunk
constexpr uint32_t kLeafType = 3u;

struct KNode {
  int64_t off;    // first reference of the list in the level's array
  int32_t n;      // references
  int32_t depth;  // RefineNode depth (+100 below 20 references with an empty side)
  float mn[3], mx[3];
};
struct KDec {
  int32_t kind;   // split axis, or kLeafType
  float trial;    // the chosen candidate (the classification value)
  float moved;    // the split value stored in the node
  int32_t nl, nb, nr;
  int32_t cdepth; // the children's depth
  int32_t child;  // left child's index in the next level (right: child + 1)
  int64_t loff, roff;  // the children's lists in the next level's array
};

// std::min(x, y) / std::max(x, y)
__device__ __forceinline__ float smin(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float smax(float x, float y) { return (x < y) ? y : x; }

// Keys for 64-bit atomicMin that pick the first (in list order) of the least
// (min_key) or greatest (max_key) values under float comparison: the value's
// order key (-0 merged into +0: they compare equal; inverted for the
// greatest) above the list position.  kNone: no value.
__device__ __forceinline__ uint32_t order_key(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) == 0u) u = 0u;  // -0 == +0
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ unsigned long long min_key(float f, uint32_t pos) {
  return ((unsigned long long)order_key(f) << 32) | pos;
}
__device__ __forceinline__ unsigned long long max_key(float f, uint32_t pos) {
  return ((unsigned long long)~order_key(f) << 32) | pos;
}
constexpr unsigned long long kNone = ~0ull;
__device__ __forceinline__ uint32_t key_pos(unsigned long long k) { return (uint32_t)k; }
__device__ __forceinline__ unsigned long long kmin(unsigned long long a, unsigned long long b) {
  return b < a ? b : a;
}

// classify (geom.cpp Builder::classify): 1 positive, -1 negative, 0 straddling
__device__ __forceinline__ int kd_classify(float mn, float mx, float sv) {
  if (mn >= sv) return 1;
  if (mx <= sv) return -1;
  if (mn == mx) return 1;
  return 0;
}

__device__ __forceinline__ float box_sa(const float mn[3], const float mx[3]) {
  const float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
  return 2.0f * ((d0 * d2 + d0 * d1) + d1 * d2);
}

// Exclusive scan of one value per thread over a block of B threads; *total
// receives the sum.  Every thread of the block must call it.
template <int B = 1024>
__device__ int64_t block_scan(int64_t v, int64_t *sh, int64_t *total) {
  const int t = (int)threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < B; d <<= 1) {
    const int64_t x = t >= d ? sh[t - d] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  const int64_t incl = sh[t];
  if (total) *total = sh[B - 1];
  __syncthreads();
  return incl - v;
}

// largest i < n with a[i] <= x (a non-decreasing, a[0] <= x)
template <class T>
__device__ __forceinline__ int32_t last_le(const T *a, int32_t n, T x) {
  int32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (a[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_kd_extents(const float *__restrict__ v9, int64_t n, float *__restrict__ lo,
                             float *__restrict__ hi, int32_t *__restrict__ refs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float a = v9[9 * t + c], b = v9[9 * t + 3 + c], d = v9[9 * t + 6 + c];
    lo[3 * t + c] = smin(a, smin(b, d));
    hi[3 * t + c] = smax(a, smax(b, d));
  }
  refs[t] = (int32_t)t;
}

// Root bounds: mn = std::min over the triangles in order from 1e23 (the
// first of equal values kept), likewise mx from -1e23; one block.  ext (6
// keys) is kNone on entry.
__global__ __launch_bounds__(1024) void k_kd_root(const float *__restrict__ lo,
                                                  const float *__restrict__ hi, int64_t n,
                                                  KNode *__restrict__ root,
                                                  unsigned long long *__restrict__ ext) {
  const int t = (int)threadIdx.x;
  for (int c = 0; c < 3; ++c) {
    unsigned long long bmin = kNone, bmax = kNone;
    for (int64_t i = t; i < n; i += 1024) {
      const float l = lo[3 * i + c], h = hi[3 * i + c];
      if (l < 1.0e23f) bmin = kmin(bmin, min_key(l, (uint32_t)i));
      if (h > -1.0e23f) bmax = kmin(bmax, max_key(h, (uint32_t)i));
    }
    if (bmin != kNone) atomicMin(&ext[c], bmin);
    if (bmax != kNone) atomicMin(&ext[3 + c], bmax);
  }
  __syncthreads();
  if (t == 0) {
    KNode r;
    r.off = 0;
    r.n = (int32_t)n;
    r.depth = 0;
    for (int c = 0; c < 3; ++c) {
      // read through the atomic unit, where the other threads' minima landed
      const unsigned long long kn = atomicAdd(&ext[c], 0ull), kx = atomicAdd(&ext[3 + c], 0ull);
      r.mn[c] = kn == kNone ? 1.0e23f : lo[3 * (int64_t)key_pos(kn) + c];
      r.mx[c] = kx == kNone ? -1.0e23f : hi[3 * (int64_t)key_pos(kx) + c];
      if (n == 0) r.mn[c] = r.mx[c] = 0.0f;
    }
    *root = r;
  }
}

__device__ __forceinline__ bool refinable(const KNode &nd) {
  return nd.n >= 3 && nd.depth <= kKdMaxDepth;
}

__device__ __forceinline__ int64_t part_chunks(const KNode &nd) {
  return ((int64_t)nd.n + kChunk - 1) / kChunk;
}
__device__ __forceinline__ int64_t eval_chunks(const KNode &nd) {
  return refinable(nd) ? ((int64_t)nd.n + kEvalChunk - 1) / kEvalChunk : 0;
}

// chunk_start[i] = partition chunks of the nodes before i (ceil(n / kChunk)
// each), eval_start[i] = evaluation chunks (ceil(n / kEvalChunk) per node to
// refine); entry N = totals.  One block.
__global__ __launch_bounds__(1024) void k_kd_chunks(const KNode *__restrict__ nodes, int32_t N,
                                                    int64_t *__restrict__ chunk_start,
                                                    int64_t *__restrict__ eval_start) {
  __shared__ int64_t sh[1024];
  const int t = (int)threadIdx.x;
  const int32_t per = (N + 1023) / 1024;
  const int32_t a = min(N, t * per), b = min(N, a + per);
  int64_t s = 0, se = 0;
  for (int32_t i = a; i < b; ++i) {
    const KNode nd = nodes[i];
    s += part_chunks(nd);
    se += eval_chunks(nd);
  }
  int64_t total, etotal;
  int64_t run = block_scan(s, sh, &total);
  int64_t erun = block_scan(se, sh, &etotal);
  for (int32_t i = a; i < b; ++i) {
    const KNode nd = nodes[i];
    chunk_start[i] = run;
    eval_start[i] = erun;
    run += part_chunks(nd);
    erun += eval_chunks(nd);
  }
  if (t == 0) {
    chunk_start[N] = total;
    eval_start[N] = etotal;
  }
}

// Candidate splits of node i on axis a, in the host builder's order.
// Also clears the node's counters for the evaluation: cl, cr of the axis's
// candidates and the axis's two extent keys.
__global__ void k_kd_cands(const KNode *__restrict__ nodes, int32_t N,
                           const int32_t *__restrict__ refs, const float *__restrict__ v9,
                           float *__restrict__ cv, int32_t *__restrict__ cc,
                           int32_t *__restrict__ cl, int32_t *__restrict__ cr,
                           unsigned long long *__restrict__ ext) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= 3 * (int64_t)N) return;
  const int32_t i = (int32_t)(g / 3);
  const int a = (int)(g % 3);
  const KNode nd = nodes[i];
  int cnt = 0;
  for (int j = 0; j < kAxisCands; ++j) {
    cl[(int64_t)i * kCands + a * kAxisCands + j] = 0;
    cr[(int64_t)i * kCands + a * kAxisCands + j] = 0;
  }
  ext[6 * (int64_t)i + a] = kNone;
  ext[6 * (int64_t)i + 3 + a] = kNone;
  if (refinable(nd)) {
    float *o = cv + (int64_t)i * kCands + a * kAxisCands;
    const float mn = nd.mn[a], mx = nd.mx[a];
    o[cnt++] = 0.5f * (mn + mx);
    const int32_t skip = 1 + nd.n / 10;
    for (int32_t ts = skip - 1; ts < nd.n && cnt < kAxisCands; ts += skip) {
      const float *v = v9 + 9 * (int64_t)refs[nd.off + ts];
      for (int tv = 0; tv < 3 && cnt < kAxisCands; ++tv) {
        const float trial = v[3 * tv + a];
        if (trial > mx || trial < mn) continue;
        o[cnt++] = trial;
      }
    }
  }
  cc[g] = cnt;
}

// Left/right counts of every candidate of a node over one evaluation chunk
// of its list (atomically added; two threads per candidate, one per half of
// the chunk), and the chunk's per-axis extents (first of the least lo /
// greatest hi values, as list positions).
__global__ __launch_bounds__(256) void k_kd_eval(
    const KNode *__restrict__ nodes, int32_t N, const int64_t *__restrict__ eval_start,
    const int32_t *__restrict__ refs, const float *__restrict__ lo, const float *__restrict__ hi,
    const float *__restrict__ cv, const int32_t *__restrict__ cc, int32_t *__restrict__ cl,
    int32_t *__restrict__ cr, unsigned long long *__restrict__ ext) {
  __shared__ float s_lo[3][kEvalChunk], s_hi[3][kEvalChunk];
  const int64_t ch = blockIdx.x;
  if (ch >= eval_start[N]) return;
  const int32_t i = last_le(eval_start, N, ch);
  const KNode nd = nodes[i];
  if (!refinable(nd)) return;  // (such nodes have no evaluation chunks)
  const int64_t k0 = (ch - eval_start[i]) * kEvalChunk;
  const int32_t m = (int32_t)min((int64_t)kEvalChunk, (int64_t)nd.n - k0);
  for (int32_t j = (int)threadIdx.x; j < m; j += 256) {
    const int64_t t = refs[nd.off + k0 + j];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s_lo[c][j] = lo[3 * t + c];
      s_hi[c][j] = hi[3 * t + c];
    }
  }
  __syncthreads();
  const int tid = (int)threadIdx.x;
  if (tid < 2 * kCands) {
    const int c = tid % kCands, half = tid / kCands;
    const int a = c / kAxisCands, j = c % kAxisCands;
    if (j < cc[3 * (int64_t)i + a]) {
      const float sv = cv[(int64_t)i * kCands + c];
      const int32_t mh = (m + 1) >> 1;
      const int32_t q0 = half ? mh : 0, q1 = half ? m : mh;
      int32_t l = 0, r = 0;
      for (int32_t q = q0; q < q1; ++q) {
        const int k = kd_classify(s_lo[a][q], s_hi[a][q], sv);
        r += k > 0;
        l += k < 0;
      }
      if (l) atomicAdd(&cl[(int64_t)i * kCands + c], l);
      if (r) atomicAdd(&cr[(int64_t)i * kCands + c], r);
    }
  } else {
    // 64 threads: per axis, the first least lo and first greatest hi (NaN
    // never taken, as `v < min` / `v > max` in the host code)
    const int u = tid - 2 * kCands;
    for (int a = 0; a < 3; ++a) {
      unsigned long long bmin = kNone, bmax = kNone;
      for (int32_t q = u; q < m; q += 64) {
        const uint32_t pos = (uint32_t)(k0 + q);
        const float l = s_lo[a][q], h = s_hi[a][q];
        if (l == l) bmin = kmin(bmin, min_key(l, pos));
        if (h == h) bmax = kmin(bmax, max_key(h, pos));
      }
      if (bmin != kNone) atomicMin(&ext[6 * (int64_t)i + a], bmin);
      if (bmax != kNone) atomicMin(&ext[6 * (int64_t)i + 3 + a], bmax);
    }
  }
}

// Cost of candidate c of node nd (geom.cpp Builder::evaluate).
struct CandEval {
  float cost, moved;
  int32_t l, b, r;
};
__device__ __forceinline__ CandEval cand_eval(const KNode &nd, int a, float sv, int32_t l,
                                              int32_t r, double minc, double maxc) {
  CandEval e;
  e.l = l;
  e.r = r;
  e.b = nd.n - l - r;
  float moved = sv;
  if (l != 0 && e.b == 0 && r == 0) moved = (float)maxc;
  if (r != 0 && e.b == 0 && l == 0) moved = (float)minc;
  float lmx[3] = {nd.mx[0], nd.mx[1], nd.mx[2]}, rmn[3] = {nd.mn[0], nd.mn[1], nd.mn[2]};
  lmx[a] = moved;
  rmn[a] = moved;
  const float sal = box_sa(nd.mn, lmx), sar = box_sa(rmn, nd.mx);
  const float isa = 1.0f / box_sa(nd.mn, nd.mx);
  e.cost = kKdCostTrav +
           kKdCostIsect * (((float)e.b + (sal * isa) * (float)l) + (sar * isa) * (float)r);
  e.moved = moved;
  return e;
}

// One wave per node: the first cheapest candidate, then split or leaf.
__global__ __launch_bounds__(256) void k_kd_select(
    const KNode *__restrict__ nodes, int32_t N, const int32_t *__restrict__ refs,
    const float *__restrict__ lo, const float *__restrict__ hi, const float *__restrict__ cv,
    const int32_t *__restrict__ cc, const int32_t *__restrict__ cl, const int32_t *__restrict__ cr,
    const unsigned long long *__restrict__ ext, KDec *__restrict__ dec) {
  const int32_t i = (int32_t)blockIdx.x * 4 + (int32_t)(threadIdx.x >> 6);
  const int lane = (int)(threadIdx.x & 63);
  if (i >= N) return;
  const KNode nd = nodes[i];
  KDec d;
  d.kind = (int32_t)kLeafType;
  d.trial = d.moved = 0.0f;
  d.nl = d.nb = d.nr = 0;
  d.cdepth = 0;
  d.child = -1;
  d.loff = d.roff = 0;
  if (refinable(nd)) {
    double minc[3], maxc[3];
    for (int a = 0; a < 3; ++a) {
      const unsigned long long kn = ext[6 * (int64_t)i + a], kx = ext[6 * (int64_t)i + 3 + a];
      // host: minc = 1e23 (double), replaced by each listed value below it
      minc[a] = 1.0e23;
      maxc[a] = -1.0e23;
      if (kn != kNone) {
        const float v = lo[3 * (int64_t)refs[nd.off + key_pos(kn)] + a];
        if ((double)v < minc[a]) minc[a] = v;
      }
      if (kx != kNone) {
        const float v = hi[3 * (int64_t)refs[nd.off + key_pos(kx)] + a];
        if ((double)v > maxc[a]) maxc[a] = v;
      }
    }
    float best = 1.0e23f;
    int bidx = -1;
    for (int c = lane; c < kCands; c += 64) {
      const int a = c / kAxisCands, j = c % kAxisCands;
      if (j >= cc[3 * (int64_t)i + a]) continue;
      const int64_t q = (int64_t)i * kCands + c;
      const CandEval e = cand_eval(nd, a, cv[q], cl[q], cr[q], minc[a], maxc[a]);
      if (e.cost < best) {
        best = e.cost;
        bidx = c;
      }
    }
    // across lanes: least cost, then least candidate index
    for (int s = 1; s < 64; s <<= 1) {
      const float ob = __shfl_xor(best, s);
      const int oi = __shfl_xor(bidx, s);
      if (oi >= 0 && (bidx < 0 || ob < best || (ob == best && oi < bidx))) {
        best = ob;
        bidx = oi;
      }
    }
    const float no_split = (float)(double)((int64_t)167 * nd.n);
    if (bidx >= 0 && !(no_split <= best)) {
      const int a = bidx / kAxisCands;
      const int64_t q = (int64_t)i * kCands + bidx;
      const CandEval e = cand_eval(nd, a, cv[q], cl[q], cr[q], minc[a], maxc[a]);
      d.kind = a;
      d.trial = cv[q];
      d.moved = e.moved;
      d.nl = e.l;
      d.nb = e.b;
      d.nr = e.r;
      d.cdepth = nd.depth + ((nd.n < 20 && (e.l == 0 || e.r == 0)) ? 100 : 0) + 1;
    }
  }
  if (lane == 0) dec[i] = d;
}

// Children slots and list offsets in the next level (one block); tot[0] =
// next level's nodes, tot[1] = its references.
__global__ __launch_bounds__(1024) void k_kd_nodescan(int32_t N, KDec *__restrict__ dec,
                                                      int64_t *__restrict__ tot) {
  __shared__ int64_t sh[1024];
  const int t = (int)threadIdx.x;
  const int32_t per = (N + 1023) / 1024;
  const int32_t a = min(N, t * per), b = min(N, a + per);
  int64_t sn = 0, sr = 0;
  for (int32_t i = a; i < b; ++i) {
    const KDec &d = dec[i];
    if (d.kind != (int32_t)kLeafType) {
      sn += 2;
      sr += (int64_t)d.nl + 2 * (int64_t)d.nb + d.nr;
    }
  }
  int64_t tn, tr;
  int64_t rn = block_scan(sn, sh, &tn);
  int64_t rr = block_scan(sr, sh, &tr);
  for (int32_t i = a; i < b; ++i) {
    KDec &d = dec[i];
    if (d.kind != (int32_t)kLeafType) {
      d.child = (int32_t)rn;
      d.loff = rr;
      d.roff = rr + d.nl + d.nb;
      rn += 2;
      rr += (int64_t)d.nl + 2 * (int64_t)d.nb + d.nr;
    }
  }
  if (t == 0) {
    tot[0] = tn;
    tot[1] = tr;
  }
}

__global__ void k_kd_children(const KNode *__restrict__ nodes, const KDec *__restrict__ dec,
                              int32_t N, KNode *__restrict__ next) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const KDec d = dec[i];
  if (d.kind == (int32_t)kLeafType) return;
  const KNode nd = nodes[i];
  KNode L = nd, R = nd;
  L.off = d.loff;
  L.n = d.nl + d.nb;
  L.depth = d.cdepth;
  L.mx[d.kind] = d.moved;
  R.off = d.roff;
  R.n = d.nb + d.nr;
  R.depth = d.cdepth;
  R.mn[d.kind] = d.moved;
  next[d.child] = L;
  next[d.child + 1] = R;
}

// Per chunk of a split node whose list spans several chunks: counts of
// negatives, straddlers and positives (packed 16 bits each; a chunk holds at
// most 2048 references).  A single-chunk list needs none: its ranks start
// at 0.
__global__ __launch_bounds__(256) void k_kd_partcount(
    const KNode *__restrict__ nodes, const KDec *__restrict__ dec, int32_t N,
    const int64_t *__restrict__ chunk_start, const int32_t *__restrict__ refs,
    const float *__restrict__ lo, const float *__restrict__ hi, int64_t *__restrict__ ccnt) {
  __shared__ int64_t sh[256];
  const int64_t ch = blockIdx.x;
  if (ch >= chunk_start[N]) return;
  const int32_t i = last_le(chunk_start, N, ch);
  const KDec d = dec[i];
  if (d.kind == (int32_t)kLeafType || nodes[i].n <= kChunk) return;
  int64_t v = 0;
  {
    const KNode nd = nodes[i];
    const int64_t k0 = (ch - chunk_start[i]) * kChunk;
    const int32_t m = (int32_t)min((int64_t)kChunk, (int64_t)nd.n - k0);
    for (int32_t j = (int)threadIdx.x; j < m; j += 256) {
      const int64_t t = refs[nd.off + k0 + j];
      const int k = kd_classify(lo[3 * t + d.kind], hi[3 * t + d.kind], d.trial);
      v += k < 0 ? 1 : (k == 0 ? (1ll << 16) : (1ll << 32));
    }
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) ccnt[ch] = sh[0];
}

// Per split node spanning several chunks (one thread each): exclusive prefix
// of its chunks' class counts, pfx[3 ch + class] = references of that class
// in the node's earlier chunks.
__global__ void k_kd_chunkbase(const KNode *__restrict__ nodes, const KDec *__restrict__ dec,
                               int32_t N, const int64_t *__restrict__ chunk_start,
                               const int64_t *__restrict__ ccnt, int64_t *__restrict__ pfx) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  if (dec[i].kind == (int32_t)kLeafType || nodes[i].n <= kChunk) return;
  int64_t run[3] = {0, 0, 0};
  for (int64_t c = chunk_start[i]; c < chunk_start[i + 1]; ++c) {
    const int64_t v = ccnt[c];
    for (int cls = 0; cls < 3; ++cls) {
      pfx[3 * c + cls] = run[cls];
      run[cls] += (v >> (16 * cls)) & 0xffff;
    }
  }
}

// Writes reference tri of class cls (0 negative, 1 straddling, 2 positive)
// with rank k within its class into the children's lists.
__device__ __forceinline__ void kd_place_ref(const KDec &d, int cls, int64_t k, int32_t tri,
                                             int32_t *__restrict__ next_refs, int64_t next_total,
                                             int32_t *__restrict__ err) {
  const int64_t nl = d.nl, nb = d.nb, nr = d.nr;
  if (cls == 0) {  // negative: left, in order
    if (k >= nl || d.loff + k >= next_total) {
      atomicOr(err, 1);
      return;
    }
    next_refs[d.loff + k] = tri;
  } else if (cls == 1) {  // straddling: both, in order
    if (k >= nb || d.roff + k >= next_total) {
      atomicOr(err, 1);
      return;
    }
    next_refs[d.loff + nl + k] = tri;
    next_refs[d.roff + k] = tri;
  } else {  // positive: right, reversed
    if (k >= nr || d.roff + nb + nr - 1 - k >= next_total) {
      atomicOr(err, 1);
      return;
    }
    next_refs[d.roff + nb + (nr - 1 - k)] = tri;
  }
}

// Stable partition of one chunk of a split node's list.  Up to 64
// references: one wave, ranks from ballots.  Otherwise thread t takes
// references 8t .. 8t + 7 and ranks come from a block scan, offset by the
// node's earlier chunks (pfx).
__global__ __launch_bounds__(256) void k_kd_scatter(
    const KNode *__restrict__ nodes, const KDec *__restrict__ dec, int32_t N,
    const int64_t *__restrict__ chunk_start, const int32_t *__restrict__ refs,
    const float *__restrict__ lo, const float *__restrict__ hi, const int64_t *__restrict__ pfx,
    int32_t *__restrict__ next_refs, int64_t next_total, int32_t *__restrict__ err) {
  __shared__ int64_t sh[256];
  const int64_t ch = blockIdx.x;
  if (ch >= chunk_start[N]) return;
  const int32_t i = last_le(chunk_start, N, ch);
  const KDec d = dec[i];
  if (d.kind == (int32_t)kLeafType) return;
  const KNode nd = nodes[i];
  const int64_t c0 = chunk_start[i];
  const int64_t k0 = (ch - c0) * kChunk;
  const int32_t m = (int32_t)min((int64_t)kChunk, (int64_t)nd.n - k0);
  const int t = (int)threadIdx.x;
  auto classify_ref = [&](int32_t tri) {
    const int k = kd_classify(lo[3 * (int64_t)tri + d.kind], hi[3 * (int64_t)tri + d.kind], d.trial);
    return k < 0 ? 0 : (k == 0 ? 1 : 2);
  };
  if (m <= 64) {  // block-uniform: one wave, no barrier
    if (t >= 64) return;
    const bool live = t < m;
    const int32_t tri = live ? refs[nd.off + k0 + t] : 0;
    const int cls = live ? classify_ref(tri) : 3;
    const uint64_t below = (1ull << t) - 1ull;
    const uint64_t b0 = __ballot(cls == 0), b1 = __ballot(cls == 1), b2 = __ballot(cls == 2);
    if (!live) return;
    const uint64_t mine = cls == 0 ? b0 : (cls == 1 ? b1 : b2);
    const int64_t base = nd.n > kChunk ? pfx[3 * ch + cls] : 0;
    kd_place_ref(d, cls, base + __popcll(mine & below), tri, next_refs, next_total, err);
    return;
  }
  constexpr int PER = 8;  // 256 threads x 8 = one chunk
  int cls[PER];
  int32_t tri[PER];
  int64_t v = 0;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int32_t j = t * PER + u;
    cls[u] = 3;  // none
    tri[u] = 0;
    if (j < m) {
      tri[u] = refs[nd.off + k0 + j];
      cls[u] = classify_ref(tri[u]);
      v += 1ll << (16 * cls[u]);
    }
  }
  const int64_t ex = block_scan<256>(v, sh, nullptr);
  int64_t rk[3];
  for (int c = 0; c < 3; ++c)
    rk[c] = (nd.n > kChunk ? pfx[3 * ch + c] : 0) + ((ex >> (16 * c)) & 0xffff);
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    if (cls[u] == 3) continue;
    kd_place_ref(d, cls[u], rk[cls[u]]++, tri[u], next_refs, next_total, err);
  }
}

// Subtree totals, deepest level first: split nodes and leaf references.
__global__ void k_kd_subtree(const KNode *__restrict__ nodes, const KDec *__restrict__ dec,
                             int32_t N, const int64_t *__restrict__ ns_next,
                             const int64_t *__restrict__ nr_next, int64_t *__restrict__ ns,
                             int64_t *__restrict__ nr) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const KDec d = dec[i];
  if (d.kind == (int32_t)kLeafType) {
    ns[i] = 0;
    nr[i] = nodes[i].n;
  } else {
    ns[i] = 1 + ns_next[d.child] + ns_next[d.child + 1];
    nr[i] = nr_next[d.child] + nr_next[d.child + 1];
  }
}

// Depth-first numbering, root level first: a node's index, its preorder rank
// among split nodes and its leaves' first reference.
__global__ void k_kd_place(const KNode *__restrict__ nodes, const KDec *__restrict__ dec,
                           int32_t N, const int64_t *__restrict__ gi, const int64_t *__restrict__ rk,
                           const int64_t *__restrict__ ro, const int64_t *__restrict__ ns_next,
                           const int64_t *__restrict__ nr_next, int64_t *__restrict__ gi_next,
                           int64_t *__restrict__ rk_next, int64_t *__restrict__ ro_next,
                           uint32_t *__restrict__ children, float *__restrict__ split,
                           int64_t n_nodes, int32_t *__restrict__ err) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const KDec d = dec[i];
  const int64_t g = gi[i];
  if (g < 0 || g >= n_nodes) {
    atomicOr(err, 2);
    return;
  }
  if (d.kind == (int32_t)kLeafType) {
    children[g] = kLeafType + ((uint32_t)ro[i] << 2);
    split[g] = (float)nodes[i].n;
  } else {
    const int64_t k = rk[i];
    children[g] = (uint32_t)d.kind + ((uint32_t)(1 + 2 * k) << 2);
    split[g] = d.moved;
    const int32_t c = d.child;
    gi_next[c] = 1 + 2 * k;
    rk_next[c] = k + 1;
    ro_next[c] = ro[i];
    gi_next[c + 1] = 2 + 2 * k;
    rk_next[c + 1] = k + 1 + ns_next[c];
    ro_next[c + 1] = ro[i] + nr_next[c];
  }
}

// Leaf lists into the triangle index list: one thread per reference of the
// level.
__global__ void k_kd_leafcopy(const KNode *__restrict__ nodes, const KDec *__restrict__ dec,
                              int32_t N, const int32_t *__restrict__ refs, int64_t R,
                              const int64_t *__restrict__ ro, int32_t *__restrict__ idx,
                              int64_t n_idx, int32_t *__restrict__ err) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= R) return;
  int32_t lo = 0, hi = N - 1;  // last node whose list starts at or before p
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (nodes[mid].off <= p) lo = mid;
    else hi = mid - 1;
  }
  if (dec[lo].kind != (int32_t)kLeafType) return;
  const int64_t q = ro[lo] + (p - nodes[lo].off);
  if (p - nodes[lo].off >= nodes[lo].n || q >= n_idx) {
    atomicOr(err, 4);
    return;
  }
  idx[q] = refs[p];
}

// Device memory for one build: bump allocation from a few large blocks,
// released after the stream has drained.
struct Arena {
  std::vector<void *> blocks;
  std::vector<size_t> sizes;
  size_t cur = 0, used = 0;
  ~Arena() {
    for (void *p : blocks) (void)hipFree(p);
  }
  void reset() {
    cur = 0;
    used = 0;
  }
  template <class T>
  int get(T **out, int64_t n) {
    size_t bytes = ((size_t)std::max<int64_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
    while (cur < blocks.size() && used + bytes > sizes[cur]) {
      ++cur;
      used = 0;
    }
    if (cur == blocks.size()) {
      const size_t sz = std::max(bytes, (size_t)64 << 20);
      void *p = nullptr;
      if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        return fail(VR_E_NOMEM, "kd-tree build: hipMalloc(" + std::to_string(sz) + ")");
      }
      blocks.push_back(p);
      sizes.push_back(sz);
      used = 0;
    }
    *out = reinterpret_cast<T *>(static_cast<char *>(blocks[cur]) + used);
    used += bytes;
    return VR_OK;
  }
};

struct Level {
  KNode *nodes = nullptr;
  KDec *dec = nullptr;
  int32_t *refs = nullptr;
  int32_t N = 0;
  int64_t R = 0;
  int64_t *ns = nullptr, *nr = nullptr;  // subtree totals
  int64_t *gi = nullptr, *rk = nullptr, *ro = nullptr;
};

#define KD_HIP(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      (void)hipStreamSynchronize(st);                                                   \
      return fail(VR_E_HIP, std::string("kd-tree build: ") + #call + ": " +            \
                                hipGetErrorString(e_));                                 \
    }                                                                                   \
  } while (0)
#define KD_TRY(x)                      \
  do {                                 \
    int rc_ = (x);                     \
    if (rc_) {                         \
      (void)hipStreamSynchronize(st);  \
      return rc_;                      \
    }                                  \
  } while (0)

unsigned grid(int64_t n, int b) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

int build(const float *v9_host, int64_t n, hipStream_t st, KdTree *out) {
  Arena keep, tmp;
  float *v9, *lo, *hi;
  int32_t *err;
  unsigned long long *rext;
  int64_t *tot;
  KD_TRY(keep.get(&v9, 9 * n));
  KD_TRY(keep.get(&lo, 3 * n));
  KD_TRY(keep.get(&hi, 3 * n));
  KD_TRY(keep.get(&err, 1));
  KD_TRY(keep.get(&rext, 6));
  KD_TRY(keep.get(&tot, 2));
  std::vector<Level> lv(1);
  KD_TRY(keep.get(&lv[0].nodes, 1));
  KD_TRY(keep.get(&lv[0].refs, n));
  lv[0].N = 1;
  lv[0].R = n;
  if (n) KD_HIP(hipMemcpyAsync(v9, v9_host, 9 * n * sizeof(float), hipMemcpyHostToDevice, st));
  KD_HIP(hipMemsetAsync(err, 0, sizeof(int32_t), st));
  KD_HIP(hipMemsetAsync(rext, 0xff, 6 * sizeof(unsigned long long), st));  // kNone
  if (n) hipLaunchKernelGGL(k_kd_extents, dim3(grid(n, 256)), dim3(256), 0, st, v9, n, lo, hi,
                            lv[0].refs);
  hipLaunchKernelGGL(k_kd_root, dim3(1), dim3(1024), 0, st, lo, hi, n, lv[0].nodes, rext);
  KD_HIP(hipGetLastError());
  for (size_t L = 0;; ++L) {
    Level &c = lv[L];
    const int32_t N = c.N;
    tmp.reset();
    float *cv;
    int32_t *cc, *cl, *cr;
    unsigned long long *ext;
    int64_t *cs, *es;
    KD_TRY(keep.get(&c.dec, N));
    KD_TRY(tmp.get(&cv, (int64_t)N * kCands));
    KD_TRY(tmp.get(&cc, 3 * (int64_t)N));
    KD_TRY(tmp.get(&cl, (int64_t)N * kCands));
    KD_TRY(tmp.get(&cr, (int64_t)N * kCands));
    KD_TRY(tmp.get(&ext, 6 * (int64_t)N));
    KD_TRY(tmp.get(&cs, (int64_t)N + 1));
    KD_TRY(tmp.get(&es, (int64_t)N + 1));
    const int64_t max_chunks = c.R / kChunk + N + 1;
    const int64_t max_echunks = c.R / kEvalChunk + N + 1;
    hipLaunchKernelGGL(k_kd_chunks, dim3(1), dim3(1024), 0, st, c.nodes, N, cs, es);
    hipLaunchKernelGGL(k_kd_cands, dim3(grid(3 * (int64_t)N, 256)), dim3(256), 0, st, c.nodes, N,
                       c.refs, v9, cv, cc, cl, cr, ext);
    hipLaunchKernelGGL(k_kd_eval, dim3(grid(max_echunks, 1)), dim3(256), 0, st, c.nodes, N, es,
                       c.refs, lo, hi, cv, cc, cl, cr, ext);
    hipLaunchKernelGGL(k_kd_select, dim3(grid(N, 4)), dim3(256), 0, st, c.nodes, N, c.refs, lo,
                       hi, cv, cc, cl, cr, ext, c.dec);
    hipLaunchKernelGGL(k_kd_nodescan, dim3(1), dim3(1024), 0, st, N, c.dec, tot);
    KD_HIP(hipGetLastError());
    int64_t th[2] = {0, 0};
    int32_t lerr = 0;
    KD_HIP(hipMemcpyAsync(th, tot, sizeof th, hipMemcpyDeviceToHost, st));
    KD_HIP(hipMemcpyAsync(&lerr, err, sizeof lerr, hipMemcpyDeviceToHost, st));
    KD_HIP(hipStreamSynchronize(st));
    if (lerr)
      return fail(VR_E_STATE, "kd-tree build: inconsistent partition (" + std::to_string(lerr) + ")");
    if (th[0] == 0) break;
    if (th[0] >= ((int64_t)1 << 30) || th[1] >= ((int64_t)1 << 31) - 1)
      KD_TRY(fail(VR_E_CAPACITY, "kd-tree level exceeds 2^30 nodes or 2^31 references"));
    if (L + 1 > (size_t)kKdMaxDepth + 2)
      KD_TRY(fail(VR_E_STATE, "kd-tree build: more levels than the depth limit allows"));
    lv.emplace_back();
    Level &c2 = lv[L];  // lv may have moved
    Level &nx = lv[L + 1];
    nx.N = (int32_t)th[0];
    nx.R = th[1];
    KD_TRY(keep.get(&nx.nodes, nx.N));
    KD_TRY(keep.get(&nx.refs, nx.R));
    // every slot is written by the partition; zero first so that even an
    // inconsistent partition (reported through err) leaves valid indices
    KD_HIP(hipMemsetAsync(nx.refs, 0, sizeof(int32_t) * (size_t)std::max<int64_t>(nx.R, 1), st));
    int64_t *ccnt, *pfx;
    KD_TRY(tmp.get(&ccnt, max_chunks));
    KD_TRY(tmp.get(&pfx, 3 * max_chunks));
    hipLaunchKernelGGL(k_kd_children, dim3(grid(N, 256)), dim3(256), 0, st, c2.nodes, c2.dec, N,
                       nx.nodes);
    hipLaunchKernelGGL(k_kd_partcount, dim3(grid(max_chunks, 1)), dim3(256), 0, st, c2.nodes,
                       c2.dec, N, cs, c2.refs, lo, hi, ccnt);
    hipLaunchKernelGGL(k_kd_chunkbase, dim3(grid(N, 256)), dim3(256), 0, st, c2.nodes, c2.dec, N,
                       cs, ccnt, pfx);
    hipLaunchKernelGGL(k_kd_scatter, dim3(grid(max_chunks, 1)), dim3(256), 0, st, c2.nodes,
                       c2.dec, N, cs, c2.refs, lo, hi, pfx, nx.refs, nx.R, err);
    KD_HIP(hipGetLastError());
  }
  // bottom-up subtree totals
  const int nL = (int)lv.size();
  for (int L = nL - 1; L >= 0; --L) {
    Level &c = lv[L];
    KD_TRY(keep.get(&c.ns, c.N));
    KD_TRY(keep.get(&c.nr, c.N));
    KD_TRY(keep.get(&c.gi, c.N));
    KD_TRY(keep.get(&c.rk, c.N));
    KD_TRY(keep.get(&c.ro, c.N));
    const int64_t *nsn = L + 1 < nL ? lv[L + 1].ns : c.ns;  // the deepest level has no splits
    const int64_t *nrn = L + 1 < nL ? lv[L + 1].nr : c.nr;
    hipLaunchKernelGGL(k_kd_subtree, dim3(grid(c.N, 256)), dim3(256), 0, st, c.nodes, c.dec, c.N,
                       nsn, nrn, c.ns, c.nr);
  }
  KD_HIP(hipGetLastError());
  int64_t rt[2] = {0, 0};
  KD_HIP(hipMemcpyAsync(&rt[0], lv[0].ns, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  KD_HIP(hipMemcpyAsync(&rt[1], lv[0].nr, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  KD_HIP(hipStreamSynchronize(st));
  const int64_t n_nodes = 1 + 2 * rt[0], n_idx = rt[1];
  if (n_nodes >= ((int64_t)1 << 30) || n_idx >= ((int64_t)1 << 30))
    KD_TRY(fail(VR_E_CAPACITY, "kd-tree exceeds 2^30 nodes or leaf references"));
  uint32_t *children;
  float *split;
  int32_t *idx;
  KD_TRY(keep.get(&children, n_nodes));
  KD_TRY(keep.get(&split, n_nodes));
  KD_TRY(keep.get(&idx, n_idx));
  const int64_t zero = 0;
  KD_HIP(hipMemcpyAsync(lv[0].gi, &zero, sizeof zero, hipMemcpyHostToDevice, st));
  KD_HIP(hipMemcpyAsync(lv[0].rk, &zero, sizeof zero, hipMemcpyHostToDevice, st));
  KD_HIP(hipMemcpyAsync(lv[0].ro, &zero, sizeof zero, hipMemcpyHostToDevice, st));
  for (int L = 0; L < nL; ++L) {
    Level &c = lv[L];
    const bool deeper = L + 1 < nL;
    hipLaunchKernelGGL(k_kd_place, dim3(grid(c.N, 256)), dim3(256), 0, st, c.nodes, c.dec, c.N,
                       c.gi, c.rk, c.ro, deeper ? lv[L + 1].ns : c.ns,
                       deeper ? lv[L + 1].nr : c.nr, deeper ? lv[L + 1].gi : c.gi,
                       deeper ? lv[L + 1].rk : c.rk, deeper ? lv[L + 1].ro : c.ro, children,
                       split, n_nodes, err);
    if (c.R > 0)
      hipLaunchKernelGGL(k_kd_leafcopy, dim3(grid(c.R, 256)), dim3(256), 0, st, c.nodes, c.dec,
                         c.N, c.refs, c.R, c.ro, idx, n_idx, err);
  }
  KD_HIP(hipGetLastError());
  out->children.resize((size_t)n_nodes);
  out->split.resize((size_t)n_nodes);
  out->idx.resize((size_t)n_idx);
  KNode root;
  int32_t herr = 0;
  KD_HIP(hipMemcpyAsync(out->children.data(), children, 4 * (size_t)n_nodes,
                        hipMemcpyDeviceToHost, st));
  KD_HIP(hipMemcpyAsync(out->split.data(), split, 4 * (size_t)n_nodes, hipMemcpyDeviceToHost, st));
  if (n_idx)
    KD_HIP(hipMemcpyAsync(out->idx.data(), idx, 4 * (size_t)n_idx, hipMemcpyDeviceToHost, st));
  KD_HIP(hipMemcpyAsync(&root, lv[0].nodes, sizeof root, hipMemcpyDeviceToHost, st));
  KD_HIP(hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, st));
  KD_HIP(hipStreamSynchronize(st));
  if (herr)
    return fail(VR_E_STATE, "kd-tree build: inconsistent partition (" + std::to_string(herr) + ")");
  std::memcpy(out->bmin, root.mn, sizeof root.mn);
  std::memcpy(out->bmax, root.mx, sizeof root.mx);
  out->max_depth = nL - 1;
  return VR_OK;
}

}  // namespace

int build_kdtree_device(const float *v9, int64_t n, hipStream_t st, KdTree *out) {
  if (n < 0 || n >= ((int64_t)1 << 31)) return fail(VR_E_INVALID, "kd-tree build: bad triangle count");
  *out = KdTree();
  return build(v9, n, st, out);
}

}  // namespace vr
