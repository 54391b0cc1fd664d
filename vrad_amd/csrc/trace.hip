// Visibility kernels: batched Trace4Rays (closest hit) and TestLine
// (segment visibility) over the resident kd-tree.
//
// Optimised path: one wave per block, XCD-aware ray ranges, per-lane
// traversal stacks with the top S entries in LDS (S x 8 B x 64 lanes per
// wave) and the rest in scratch.  Plain path: one thread per ray, 256-thread
// blocks, the whole stack in private scratch memory.
#include "kernels.h"

namespace vr {

// R rays per lane: a wave takes 64 R consecutive rays in R rounds of 64
// (round q: rays base + 64 q + lane), reusing its LDS stack.
template <int S, int D, int R>
__global__ __launch_bounds__(64) void k_trace_closest_lds(DevScene sc, DevRays r, int64_t n,
                                                          int32_t skip, int32_t *__restrict__ hit,
                                                          float *__restrict__ hit_t,
                                                          float *__restrict__ hit_n, uint32_t xg) {
  __shared__ uint2 s_stk[S * 64];
  const int64_t base =
      (int64_t)xcd_group_chunk(blockIdx.x, gridDim.x, xg) * 64 * R + threadIdx.x;
  StackSpill<S, D> spill;
  HybridStack<S, D> st(s_stk, (int)threadIdx.x, spill);
  for (int q = 0; q < R; ++q) {
    const int64_t i = base + 64 * q;
    if (i >= n) return;
    float t;
    int32_t best, id;
    trace_ray<D, false>(sc, st, r.ox[i], r.oy[i], r.oz[i], r.dx[i], r.dy[i], r.dz[i], r.tmin[i],
                        r.tmax[i], skip, t, best, id);
    hit[i] = best;
    hit_t[i] = t;
    if (hit_n) {
      float4 pl = best >= 0 ? sc.planes[best] : make_float4(0.f, 0.f, 0.f, 0.f);
      hit_n[3 * i + 0] = pl.x;
      hit_n[3 * i + 1] = pl.y;
      hit_n[3 * i + 2] = pl.z;
    }
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_trace_closest_plain(DevScene sc, DevRays r, int64_t n,
                                                             int32_t skip, int32_t *hit,
                                                             float *hit_t, float *hit_n) {
  const int64_t i = (int64_t)xcd_chunk(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (i >= n) return;
  ScratchStack<D> st;
  float t;
  int32_t best, id;
  trace_ray<D, false>(sc, st, r.ox[i], r.oy[i], r.oz[i], r.dx[i], r.dy[i], r.dz[i], r.tmin[i],
                      r.tmax[i], skip, t, best, id);
  hit[i] = best;
  hit_t[i] = t;
  if (hit_n) {
    float4 pl = best >= 0 ? sc.planes[best] : make_float4(0.f, 0.f, 0.f, 0.f);
    hit_n[3 * i + 0] = pl.x;
    hit_n[3 * i + 1] = pl.y;
    hit_n[3 * i + 2] = pl.z;
  }
}

// TestLine batches, R segments per lane.  R = 1: one segment per lane.
// R > 1: lane l owns segments base + R l + q (q < R), consecutive ones (in a
// MakeTransfers batch: one origin patch and neighbouring targets, mostly
// blocked by the same triangle).  A wave walks each lane's first segment,
// then tests each lane's following segments only against the triangle that
// blocked the segment before it (occluded_by: exact, the walk would accept
// that hit); the segments it does not decide are queued in LDS and walked
// 64 at a time, so lanes stay busy whatever the hit pattern.
template <int S, int D, int R>
__global__ __launch_bounds__(64) void k_test_line_lds(DevScene sc, DevSegs s, int64_t n,
                                                      int32_t skip, uint8_t *__restrict__ vis,
                                                      uint8_t *__restrict__ sky, uint32_t xg) {
  __shared__ uint2 s_stk[S * 64];
  __shared__ int32_t s_q[R > 1 ? 64 * (R - 1) : 1];  // queued segments (offsets from base)
  const int lane = (int)threadIdx.x;
  const int64_t base = (int64_t)xcd_group_chunk(blockIdx.x, gridDim.x, xg) * 64 * R;
  StackSpill<S, D> spill;
  HybridStack<S, D> st(s_stk, lane, spill);
  auto walk = [&](int64_t i, uint32_t &occ) {
    bool hs;
    const bool v = test_line_cached<D>(sc, st, s.sx[i], s.sy[i], s.sz[i], s.ex[i], s.ey[i],
                                       s.ez[i], skip, hs, occ);
    vis[i] = v ? 1 : 0;
    if (sky) sky[i] = hs ? 1 : 0;
  };
  if constexpr (R == 1) {
    const int64_t i = base + lane;
    if (i < n) {
      bool hs;
      const bool v = test_line<D>(sc, st, s.sx[i], s.sy[i], s.sz[i], s.ex[i], s.ey[i], s.ez[i],
                                  skip, hs);
      vis[i] = v ? 1 : 0;
      if (sky) sky[i] = hs ? 1 : 0;
    }
  } else {
    const int64_t first = base + (int64_t)lane * R;
    uint32_t occ = kNoRef;
    if (first < n) walk(first, occ);
    uint32_t qmask = 0;
    for (int q = 1; q < R; ++q) {
      const int64_t i = first + q;
      if (i >= n) break;
      if (occ != kNoRef &&
          occluded_by<D>(sc, st, s.sx[i], s.sy[i], s.sz[i], s.ex[i], s.ey[i], s.ez[i], skip, occ)) {
        vis[i] = 0;
        if (sky) sky[i] = 0;
      } else {
        qmask |= 1u << q;
      }
    }
    // exclusive prefix of the lanes' queue counts
    const int cnt = __builtin_popcount(qmask);
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int o = __shfl_up(incl, d);
      if (lane >= d) incl += o;
    }
    int at = incl - cnt;
    const int total = __shfl(incl, 63);
    for (uint32_t m = qmask; m; m &= m - 1) s_q[at++] = lane * R + __builtin_ctz(m);
    __syncthreads();  // one wave: the queue is written before it is read
    for (int r = 0; r < total; r += 64) {
      const int k = r + lane;
      uint32_t none = kNoRef;
      if (k < total) walk(base + s_q[k], none);
    }
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_test_line_plain(DevScene sc, DevSegs s, int64_t n,
                                                         int32_t skip, uint8_t *vis,
                                                         uint8_t *sky) {
  const int64_t i = (int64_t)xcd_chunk(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (i >= n) return;
  ScratchStack<D> st;
  bool hs;
  bool v = test_line<D>(sc, st, s.sx[i], s.sy[i], s.sz[i], s.ex[i], s.ey[i], s.ez[i], skip, hs);
  vis[i] = v ? 1 : 0;
  if (sky) sky[i] = hs ? 1 : 0;
}

// (traversal-stack entries in LDS, rays per lane)
#define VR_LDS_VARIANTS(X)                                                                     \
  X(2, 1) X(2, 2) X(2, 4) X(2, 8) X(4, 1) X(4, 2) X(4, 4) X(4, 8) X(6, 1) X(6, 2) X(6, 4)       \
  X(6, 8) X(8, 1) X(8, 2) X(8, 4) X(8, 8) X(10, 1) X(10, 2) X(10, 4) X(10, 8) X(12, 1)          \
  X(12, 2) X(12, 4) X(12, 8) X(16, 1) X(16, 2) X(16, 4) X(16, 8) X(24, 1) X(24, 2) X(24, 4)    \
  X(24, 8)

void launch_trace_closest(const DevScene &sc, const DevRays &r, int64_t n, int32_t skip,
                          int32_t *hit, float *hit_t, float *hit_n, int path, hipStream_t st) {
  if (n <= 0) return;
  if (path == VR_PATH_PLAIN) {
    hipLaunchKernelGGL(k_trace_closest_plain<kStackDepth>, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, st, sc, r, n, skip, hit, hit_t, hit_n);
    return;
  }
  const int R = tuning().trace_rays_per_lane;
  const dim3 g((unsigned)((n + 64 * R - 1) / (64 * R)));
  switch (tuning().trace_lds_entries * 8 + R) {
#define X(S, RR)                                                                              \
  case S * 8 + RR:                                                                            \
    hipLaunchKernelGGL((k_trace_closest_lds<S, kStackDepth, RR>), g, dim3(64), 0, st, sc, r, n, \
                       skip, hit, hit_t, hit_n, (uint32_t)tuning().trace_xcd_group);          \
    break;
    VR_LDS_VARIANTS(X)
#undef X
  }
}

void launch_test_line(const DevScene &sc, const DevSegs &s, int64_t n, int32_t skip,
                      uint8_t *vis, uint8_t *sky, int path, hipStream_t st) {
  if (n <= 0) return;
  if (path == VR_PATH_PLAIN) {
    hipLaunchKernelGGL(k_test_line_plain<kStackDepth>, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, st, sc, s, n, skip, vis, sky);
    return;
  }
  const int R = tuning().trace_rays_per_lane;
  const dim3 g((unsigned)((n + 64 * R - 1) / (64 * R)));
  switch (tuning().trace_lds_entries * 8 + R) {
#define X(S, RR)                                                                               \
  case S * 8 + RR:                                                                             \
    hipLaunchKernelGGL((k_test_line_lds<S, kStackDepth, RR>), g, dim3(64), 0, st, sc, s, n, skip, \
                       vis, sky, (uint32_t)tuning().trace_xcd_group);                          \
    break;
    VR_LDS_VARIANTS(X)
#undef X
  }
}

}  // namespace vr
