// GatherSampleLight: direct light at sample points from point, spot, surface,
// sun and sky-ambient lights, each term gated by TestLine visibility.  The
// per-light formulas and their operation order are those of the CPU
// restatement; lights are summed in list order.
#include "kernels.h"

namespace vr {

#define VR_MAX_TRACE_LENGTH_F ((float)(1.732050807569 * 32768.0))

template <int D, class Stack>
__device__ void gather_one(const DevScene &sc, Stack &st, float px, float py, float pz, float nx,
                           float ny, float nz, const vr_light *__restrict__ lights,
                           int32_t n_lights, const float *__restrict__ dirs, float eps,
                           int32_t skip, float &r, float &g, float &b) {
  const float sx = px + eps * nx, sy = py + eps * ny, sz = pz + eps * nz;
  const float MTL = VR_MAX_TRACE_LENGTH_F;
  float ar = 0.0f, ag = 0.0f, ab = 0.0f;
  for (int32_t li = 0; li < n_lights; ++li) {
    const vr_light &L = lights[li];
    float scale = 0.0f;
    if (L.type == VR_EMIT_SURFACE) {
      const float lx = L.origin[0] - px, ly = L.origin[1] - py, lz = L.origin[2] - pz;
      const float dist2 = (lx * lx + ly * ly) + lz * lz;
      if (!(dist2 > 0.0f)) continue;
      const float dist = __builtin_sqrtf(dist2);
      const float inv = 1.0f / dist;
      const float ux = lx * inv, uy = ly * inv, uz = lz * inv;
      const float dot = (nx * ux + ny * uy) + nz * uz;
      if (!(dot > 0.0f)) continue;
      const float dot2 = -((L.normal[0] * ux + L.normal[1] * uy) + L.normal[2] * uz);
      if (!(dot2 > 0.0f)) continue;
      if (!segment_visible<D>(sc, st, sx, sy, sz, L.origin[0], L.origin[1], L.origin[2], skip))
        continue;
      scale = (dot * dot2) / dist2;
    } else if (L.type == VR_EMIT_POINT || L.type == VR_EMIT_SPOTLIGHT) {
      const float lx = L.origin[0] - px, ly = L.origin[1] - py, lz = L.origin[2] - pz;
      const float dist2 = (lx * lx + ly * ly) + lz * lz;
      if (!(dist2 > 0.0f)) continue;
      const float dist = __builtin_sqrtf(dist2);
      const float inv = 1.0f / dist;
      const float ux = lx * inv, uy = ly * inv, uz = lz * inv;
      const float dot = (nx * ux + ny * uy) + nz * uz;
      if (!(dot > 0.0f)) continue;
      const float fd = dist < L.cap_distance ? dist : L.cap_distance;
      const float den = (L.constant_attn + L.linear_attn * fd) + L.quadratic_attn * (fd * fd);
      if (!(den > 0.0f)) continue;
      float fall = 1.0f / den;
      if (L.type == VR_EMIT_SPOTLIGHT) {
        const float dot2 = -((L.normal[0] * ux + L.normal[1] * uy) + L.normal[2] * uz);
        if (!(dot2 > L.stopdot2)) continue;
        if (dot2 <= L.stopdot) fall = fall * ((dot2 - L.stopdot2) / (L.stopdot - L.stopdot2));
        // the cone exponent in float64 then rounded: one rounding of a
        // near-exact power, so the value does not depend on a float powf
        if (L.exponent != 0.0f && L.exponent != 1.0f)
          fall = fall * (float)pow((double)dot2, (double)L.exponent);
      }
      if (L.end_fade > L.start_fade) {
        if (dist >= L.end_fade) continue;
        if (dist > L.start_fade) {
          const float x = (dist - L.start_fade) / (L.end_fade - L.start_fade);
          const float sm = (x * x) * (3.0f - 2.0f * x);
          fall = fall * (1.0f - sm);
        }
      }
      if (!segment_visible<D>(sc, st, sx, sy, sz, L.origin[0], L.origin[1], L.origin[2], skip))
        continue;
      scale = dot * fall;
    } else if (L.type == VR_EMIT_SKYLIGHT) {
      const float dot = -((nx * L.normal[0] + ny * L.normal[1]) + nz * L.normal[2]);
      if (!(dot > 0.0f)) continue;
      const float ex = sx - MTL * L.normal[0], ey = sy - MTL * L.normal[1],
                  ez = sz - MTL * L.normal[2];
      if (!segment_visible<D>(sc, st, sx, sy, sz, ex, ey, ez, skip)) continue;
      scale = dot;
    } else if (L.type == VR_EMIT_SKYAMBIENT) {
      float a = 0.0f, w = 0.0f;
      for (int k = 0; k < 162; ++k) {
        const float dx = dirs[3 * k], dy = dirs[3 * k + 1], dz = dirs[3 * k + 2];
        const float ck = (nx * dx + ny * dy) + nz * dz;
        if (!(ck > 0.0f)) continue;
        w = w + ck;
        const float ex = sx + MTL * dx, ey = sy + MTL * dy, ez = sz + MTL * dz;
        if (segment_visible<D>(sc, st, sx, sy, sz, ex, ey, ez, skip)) a = a + ck;
      }
      if (!(w > 0.0f)) continue;
      scale = a / w;
    } else {
      continue;
    }
    ar = ar + L.intensity[0] * scale;
    ag = ag + L.intensity[1] * scale;
    ab = ab + L.intensity[2] * scale;
  }
  r = ar;
  g = ag;
  b = ab;
}

template <int S, int D>
__global__ __launch_bounds__(64) void k_gather_lds(DevScene sc, DevSamples s, int64_t n,
                                                   const vr_light *__restrict__ lights,
                                                   int32_t n_lights, const float *__restrict__ dirs,
                                                   float eps, int32_t skip,
                                                   float *__restrict__ rgb) {
  __shared__ uint2 s_stk[S * 64];
  const int64_t i =
      (int64_t)xcd_group_chunk(blockIdx.x, gridDim.x, kXcdGroup) * 64 + threadIdx.x;
  if (i >= n) return;
  StackSpill<S, D> spill;
  HybridStack<S, D> st(s_stk, (int)threadIdx.x, spill);
  float r, g, b;
  gather_one<D>(sc, st, s.px[i], s.py[i], s.pz[i], s.nx[i], s.ny[i], s.nz[i], lights, n_lights,
                dirs, eps, skip, r, g, b);
  rgb[3 * i + 0] = r;
  rgb[3 * i + 1] = g;
  rgb[3 * i + 2] = b;
}

template <int D>
__global__ __launch_bounds__(256) void k_gather_plain(DevScene sc, DevSamples s, int64_t n,
                                                      const vr_light *__restrict__ lights,
                                                      int32_t n_lights,
                                                      const float *__restrict__ dirs, float eps,
                                                      int32_t skip, float *__restrict__ rgb) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  ScratchStack<D> st;
  float r, g, b;
  gather_one<D>(sc, st, s.px[i], s.py[i], s.pz[i], s.nx[i], s.ny[i], s.nz[i], lights, n_lights,
                dirs, eps, skip, r, g, b);
  rgb[3 * i + 0] = r;
  rgb[3 * i + 1] = g;
  rgb[3 * i + 2] = b;
}

// Many scenes' samples in one launch: block b (after the XCD remap) belongs
// to the job whose block range holds it (binary search; the job index is
// uniform per wave), and lane l gathers that job's sample
// (b - first block) * 64 + l from the job's scene and lights.
template <int S, int D>
__global__ __launch_bounds__(64) void k_gather_jobs(const GatherJob *__restrict__ jobs, int n_jobs,
                                                    const float *__restrict__ dirs, float eps,
                                                    int32_t skip) {
  __shared__ uint2 s_stk[S * 64];
  const int64_t b = (int64_t)xcd_group_chunk(blockIdx.x, gridDim.x, kXcdGroup);
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].blk_begin <= b) lo = mid;
    else hi = mid - 1;
  }
  const int jx = __builtin_amdgcn_readfirstlane(lo);
  const GatherJob &J = jobs[jx];
  const int64_t i = (b - J.blk_begin) * 64 + threadIdx.x;
  if (i >= J.n) return;
  StackSpill<S, D> spill;
  HybridStack<S, D> st(s_stk, (int)threadIdx.x, spill);
  const DevSamples &s = J.s;
  float r, g, bb;
  gather_one<D>(J.sc, st, s.px[i], s.py[i], s.pz[i], s.nx[i], s.ny[i], s.nz[i], J.lights,
                J.n_lights, dirs, eps, skip, r, g, bb);
  J.rgb[3 * i + 0] = r;
  J.rgb[3 * i + 1] = g;
  J.rgb[3 * i + 2] = bb;
}

void launch_gather_jobs(const GatherJob *jobs, int n_jobs, int64_t total_blocks,
                        const float *sky_dirs, float eps, int32_t skip, int lds_entries,
                        hipStream_t st) {
  if (n_jobs <= 0 || total_blocks <= 0) return;
#define VR_G_JOBS(S)                                                                          \
  hipLaunchKernelGGL((k_gather_jobs<S, kStackDepth>), dim3((unsigned)total_blocks), dim3(64), 0, \
                     st, jobs, n_jobs, sky_dirs, eps, skip)
  switch (lds_entries) {
    case 2: VR_G_JOBS(2); break;
    case 4: VR_G_JOBS(4); break;
    case 6: VR_G_JOBS(6); break;
    case 10: VR_G_JOBS(10); break;
    case 12: VR_G_JOBS(12); break;
    case 16: VR_G_JOBS(16); break;
    case 24: VR_G_JOBS(24); break;
    default: VR_G_JOBS(8); break;
  }
#undef VR_G_JOBS
}

// CanLeafTraceToSky (rad/lightmap/lightmap.go:425-451): one wave per leaf.
// The leaf's centre is float32(mins + maxs) * 0.5 per axis; lane l traces
// the segments from it to (a * -MAX_TRACE_LENGTH) + centre for the
// directions a = dirs[l], dirs[l + 64], dirs[l + 128] (< 162) with skip id
// -1, and the wave stops after the first round in which some lane's
// segment is visible (a miss, or a sky hit before any other).  "Any" does
// not depend on the order the directions are taken in.
template <int S, int D>
__global__ __launch_bounds__(64) void k_leaf_sky(DevScene sc, const int16_t *__restrict__ mins,
                                                 const int16_t *__restrict__ maxs, int64_t n,
                                                 const float *__restrict__ dirs,
                                                 uint8_t *__restrict__ out) {
  __shared__ uint2 s_stk[S * 64];
  const int64_t leaf = (int64_t)xcd_group_chunk(blockIdx.x, gridDim.x, kXcdGroup);
  if (leaf >= n) return;
  StackSpill<S, D> spill;
  HybridStack<S, D> st(s_stk, (int)threadIdx.x, spill);
  const float cx = (float)((int32_t)mins[3 * leaf + 0] + (int32_t)maxs[3 * leaf + 0]) * 0.5f;
  const float cy = (float)((int32_t)mins[3 * leaf + 1] + (int32_t)maxs[3 * leaf + 1]) * 0.5f;
  const float cz = (float)((int32_t)mins[3 * leaf + 2] + (int32_t)maxs[3 * leaf + 2]) * 0.5f;
  const float nm = -VR_MAX_TRACE_LENGTH_F;
  uint8_t any = 0;
  for (int k0 = 0; k0 < 162; k0 += 64) {
    const int k = k0 + (int)threadIdx.x;
    bool vis = false;
    if (k < 162) {
      vis = segment_visible<D>(sc, st, cx, cy, cz, dirs[3 * k] * nm + cx,
                               dirs[3 * k + 1] * nm + cy, dirs[3 * k + 2] * nm + cz, -1);
    }
    if (__any(vis)) {
      any = 1;
      break;
    }
  }
  if (threadIdx.x == 0) out[leaf] = any;
}

void launch_leaf_sky(const DevScene &sc, const int16_t *mins, const int16_t *maxs, int64_t n,
                     const float *sky_dirs, uint8_t *out, hipStream_t st) {
  if (n <= 0) return;
#define VR_LEAFSKY(S)                                                                      \
  hipLaunchKernelGGL((k_leaf_sky<S, kStackDepth>), dim3((unsigned)n), dim3(64), 0, st, sc, \
                     mins, maxs, n, sky_dirs, out)
  switch (tuning().trace_lds_entries) {
    case 2: VR_LEAFSKY(2); break;
    case 4: VR_LEAFSKY(4); break;
    case 6: VR_LEAFSKY(6); break;
    case 10: VR_LEAFSKY(10); break;
    case 12: VR_LEAFSKY(12); break;
    case 16: VR_LEAFSKY(16); break;
    case 24: VR_LEAFSKY(24); break;
    default: VR_LEAFSKY(8); break;
  }
#undef VR_LEAFSKY
}

void launch_gather_light(const DevScene &sc, const DevSamples &s, int64_t n,
                         const vr_light *lights, int32_t n_lights, const float *sky_dirs,
                         float eps, int32_t skip, float *rgb, int path, int lds_entries,
                         hipStream_t st) {
  if (n <= 0) return;
  if (path == VR_PATH_PLAIN)
    hipLaunchKernelGGL(k_gather_plain<kStackDepth>, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, st, sc, s, n, lights, n_lights, sky_dirs, eps, skip, rgb);
  else {
    // the traversal-stack entries kept in LDS: the caller's (autotuned) knob
#define VR_GATHER(S)                                                                           \
  hipLaunchKernelGGL((k_gather_lds<S, kStackDepth>), dim3((unsigned)((n + 63) / 64)), dim3(64), 0, \
                     st, sc, s, n, lights, n_lights, sky_dirs, eps, skip, rgb)
    switch (lds_entries) {
      case 2: VR_GATHER(2); break;
      case 4: VR_GATHER(4); break;
      case 6: VR_GATHER(6); break;
      case 10: VR_GATHER(10); break;
      case 12: VR_GATHER(12); break;
      case 16: VR_GATHER(16); break;
      case 24: VR_GATHER(24); break;
      default: VR_GATHER(8); break;
    }
#undef VR_GATHER
  }
}

}  // namespace vr
