// BSP clusters and PVS on the GPU.
//
//  k_point_cluster   ClusterFromPoint / PointInLeaf (rad/clustertable/
//                    point.go:10-36), one thread per point.
//  k_patch_cluster   Patch.ClusterNumber (rad/patches/subdivide.go:100-115):
//                    origin first, then the winding points.
//  k_decompress_vis  GetVisCache / DecompressVis (rad/lightmap/vis.go:9-94),
//                    one thread per cluster row.
//  k_pair_cull_pvs   the candidate pairs of MakeTransfers restricted to
//                    cluster pairs the PVS lets see each other.
#include "kernels.h"

namespace vr {

constexpr float kTestEpsilon = 0.03125f;  // TEST_EPSILON, vmath/constants.go:7

// PointInLeaf from the root.  The recursion returns children[0]'s leaf for a
// point in front of the plane, children[1]'s behind it, and for a point
// within the epsilon children[0]'s result unless its cluster is -1, then
// children[1]'s.  That is: the first leaf, in a depth-first walk taking
// children[0] before children[1] at every within-epsilon node, whose cluster
// is not -1; -1 when there is none.  The pending children[1] of
// within-epsilon nodes go on a stack of at most one entry per tree level
// (the host rejects deeper trees).
__device__ int bsp_cluster(const DevBsp &b, float x, float y, float z) {
  int stk[VR_BSP_MAX_DEPTH];
  int sp = 0;
  int node = 0;
  for (;;) {
    if (node < 0) {
      const int c = b.leaf_cluster[-1 - node];
      if (c != -1 || sp == 0) return c;
      node = stk[--sp];
      continue;
    }
    const int *nd = b.nodes + 3 * node;
    const float4 pl = b.planes[nd[0]];
    // mgl32 Vec3.Dot order, then minus the plane distance
    const float d = ((x * pl.x + y * pl.y) + z * pl.z) - pl.w;
    if (d > kTestEpsilon) {
      node = nd[1];
    } else if (d < -kTestEpsilon) {
      node = nd[2];
    } else {
      stk[sp++] = nd[2];
      node = nd[1];
    }
  }
}

__global__ __launch_bounds__(256) void k_point_cluster(DevBsp b, const float *__restrict__ xyz,
                                                       int64_t n, int16_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  out[i] = (int16_t)bsp_cluster(b, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

__global__ __launch_bounds__(256) void k_patch_cluster(DevBsp b, const vr_patch *__restrict__ p,
                                                       int64_t n, int32_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const vr_patch &q = p[i];
  int c = bsp_cluster(b, q.origin[0], q.origin[1], q.origin[2]);
  if (c == -1) {
    const int np = min(max(q.n_points, 0), VR_FACE_MAX_POINTS);
    for (int k = 0; k < np; ++k) {
      const int ck = bsp_cluster(b, q.points[k][0], q.points[k][1], q.points[k][2]);
      if (ck != -1) {
        c = ck;
        break;
      }
    }
  }
  out[i] = c;
}

void launch_point_cluster(const DevBsp &b, const float *xyz, int64_t n, int16_t *out,
                          hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_point_cluster, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, b, xyz,
                     n, out);
}

void launch_patch_cluster(const DevBsp &b, const vr_patch *p, int64_t n, int32_t *out,
                          hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_patch_cluster, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, b, p, n,
                     out);
}

// DecompressVis as the Source run-length format defines it (the loop in
// rad/lightmap/vis.go:53-94 copies its input instead; this takes the intent
// its comments and GetVisCache's use state): until the row is full, a
// nonzero byte is copied, a zero byte is followed by a count of zero bytes
// (count 0 is an error, a run past the row is clipped and counted).  Every
// input read is checked against the buffer; err bit 1 = zero count, bit 2 =
// read past the end.
__global__ __launch_bounds__(256) void k_decompress_vis(const uint8_t *__restrict__ vis,
                                                        int64_t vis_len,
                                                        const int32_t *__restrict__ ofs, int32_t nc,
                                                        uint8_t *__restrict__ rows,
                                                        unsigned long long *__restrict__ overrun,
                                                        uint32_t *__restrict__ err) {
  const int c = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (c >= nc) return;
  const int64_t row = (nc + 7) >> 3;
  uint8_t *out = rows + (int64_t)c * row;
  int64_t in = ofs[c];
  int64_t o = 0;
  while (o < row) {
    if (in >= vis_len) {
      atomicOr(err, 2u);
      break;
    }
    const uint8_t v = vis[in];
    if (v != 0) {
      out[o++] = v;
      ++in;
      continue;
    }
    if (in + 1 >= vis_len) {
      atomicOr(err, 2u);
      break;
    }
    int64_t cnt = vis[in + 1];
    if (cnt == 0) {
      atomicOr(err, 1u);
      break;
    }
    in += 2;
    if (o + cnt > row) {
      cnt = row - o;
      atomicAdd(overrun, 1ull);
    }
    for (int64_t k = 0; k < cnt; ++k) out[o + k] = 0;
    o += cnt;
  }
  for (; o < row; ++o) out[o] = 0;  // only after an error: leave no stale bytes
}

void launch_decompress_vis(const uint8_t *vis, int64_t vis_len, const int32_t *ofs, int32_t nc,
                           uint8_t *rows, unsigned long long *overrun, uint32_t *err,
                           hipStream_t st) {
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_decompress_vis, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, vis,
                     vis_len, ofs, nc, rows, overrun, err);
}

// Candidate pairs with a PVS.  Patches are grouped by cluster (ascending
// patch index inside a group; the patches with a negative cluster form one
// more group that sees every group).  Each table entry is a pair of groups
// (a <= b) that see each other; its work is |a| virtual rows of |b| columns,
// cut into blocks of 256 * kPvsIt columns.  A block finds its entry by binary
// search over the entries' block prefix, tests the facing cull on the pair
// in ascending patch order (the same expression as the unculled enumeration,
// so the kept set is the unculled set restricted to visible groups), and
// compacts the survivors with the wave ballots and one atomic per block.
constexpr int kPvsIt = 16;

__global__ __launch_bounds__(256) void k_pair_cull_pvs(DevPatches p, const int32_t *__restrict__ mem,
                                                       const PvsEntry *__restrict__ tab,
                                                       const int64_t *__restrict__ blk_pre,
                                                       int32_t n_tab, int64_t blk0,
                                                       int2 *__restrict__ pairs,
                                                       uint32_t *__restrict__ count, uint32_t cap) {
  __shared__ uint32_t s_cnt[4];
  __shared__ uint32_t s_base;
  const int64_t gb = blk0 + (int64_t)blockIdx.x;
  int lo = 0, hi = n_tab - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk_pre[mid] <= gb) lo = mid;
    else hi = mid - 1;
  }
  const PvsEntry e = tab[lo];
  const int64_t local = gb - blk_pre[lo];
  const int ia = (int)(local / e.cpr);
  const int jb = (int)(local - (int64_t)ia * e.cpr) * (256 * kPvsIt);
  const int i = mem[e.a_off + ia];
  const float ix = p.ox[i], iy = p.oy[i], iz = p.oz[i];
  const float inx = p.nx[i], iny = p.ny[i], inz = p.nz[i];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t m[kPvsIt];
  uint32_t wtot = 0;
#pragma unroll
  for (int it = 0; it < kPvsIt; ++it) {
    const int jl = jb + it * 256 + (int)threadIdx.x;
    bool keep = false;
    if (jl < e.b_n && (!e.diag || jl > ia)) {
      const int j = mem[e.b_off + jl];
      const float jx = p.ox[j], jy = p.oy[j], jz = p.oz[j];
      const float jnx = p.nx[j], jny = p.ny[j], jnz = p.nz[j];
      const bool ij = i < j;  // row = the smaller patch index
      const float rx = ij ? ix : jx, ry = ij ? iy : jy, rz = ij ? iz : jz;
      const float qx = ij ? inx : jnx, qy = ij ? iny : jny, qz = ij ? inz : jnz;
      const float sx = ij ? jx : ix, sy = ij ? jy : iy, sz = ij ? jz : iz;
      const float tx = ij ? jnx : inx, ty = ij ? jny : iny, tz = ij ? jnz : inz;
      const float dx = sx - rx, dy = sy - ry, dz = sz - rz;
      const float di = (qx * dx + qy * dy) + qz * dz;
      const float dj = -((tx * dx + ty * dy) + tz * dz);
      keep = di > 0.0f && dj > 0.0f;
    }
    m[it] = __ballot(keep);
    wtot += (uint32_t)__popcll(m[it]);
  }
  if (lane == 0) s_cnt[wave] = wtot;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    s_base = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  uint32_t base = s_base;
  for (int w = 0; w < wave; ++w) base += s_cnt[w];
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (int it = 0; it < kPvsIt; ++it) {
    if ((m[it] >> lane) & 1ull) {
      const uint32_t slot = base + (uint32_t)__popcll(m[it] & below);
      const int j = mem[e.b_off + jb + it * 256 + (int)threadIdx.x];
      if (slot < cap) pairs[slot] = make_int2(min(i, j), max(i, j));
    }
    base += (uint32_t)__popcll(m[it]);
  }
}

int pvs_block_cols() { return 256 * kPvsIt; }

void launch_pair_cull_pvs(const DevPatches &p, const int32_t *mem, const PvsEntry *tab,
                          const int64_t *blk_pre, int32_t n_tab, int64_t blk0, int64_t n_blk,
                          int2 *pairs, uint32_t *count, uint32_t cap, hipStream_t st) {
  if (n_blk <= 0 || n_tab <= 0) return;
  hipLaunchKernelGGL(k_pair_cull_pvs, dim3((unsigned)n_blk), dim3(256), 0, st, p, mem, tab, blk_pre,
                     n_tab, blk0, pairs, count, cap);
}

}  // namespace vr
