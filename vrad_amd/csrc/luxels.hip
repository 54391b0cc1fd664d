// The luxel stage on the GPU.
//
//  k_face_extents  CalcFaceExtents (rad/world/face.go:14-90), one thread per
//                  face.
//  k_face_luxels   luxel positions on the face plane and the leaf patch each
//                  luxel takes its value from; one 256-thread workgroup per
//                  256 luxels of a face, the face's leaf patches staged in LDS
//                  64 at a time (their windings mapped to lightmap
//                  coordinates once per workgroup).
#include "kernels.h"

namespace vr {

// val = p . lightmapVecs[j].xyz + lightmapVecs[j][3], left to right in FP32
// (the expression at rad/world/face.go:39-42)
__device__ __forceinline__ float lux_coord(const float *p, const float *l, float o) {
  return ((p[0] * l[0] + p[1] * l[1]) + p[2] * l[2]) + o;
}

__global__ __launch_bounds__(256) void k_face_extents(const vr_face *__restrict__ faces, int64_t n,
                                                      const float *__restrict__ lux_ofs,
                                                      vr_face_lightmap *__restrict__ out) {
  const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (f >= n) return;
  const vr_face &fc = faces[f];
  vr_face_lightmap r;
  r.mins[0] = r.mins[1] = r.size[0] = r.size[1] = 0;
  r.luxel_ofs = 0;
  r.n_luxels = 0;
  r.flags = 0;
  if (fc.flags & (VR_FACE_SKY | VR_FACE_NOLIGHT)) {
    r.flags = VR_LM_UNLIT;
    out[f] = r;
    return;
  }
  const int np = fc.n_points;
  if (np < 1 || np > VR_FACE_MAX_POINTS) {
    r.flags = VR_LM_NO_POINTS;
    out[f] = r;
    return;
  }
  const float o[2] = {lux_ofs ? lux_ofs[2 * f] : 0.0f, lux_ofs ? lux_ofs[2 * f + 1] : 0.0f};
  float mn[2] = {1e24f, 1e24f}, mx[2] = {-1e24f, -1e24f};
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < 2; ++j) {
      const float v = lux_coord(fc.points[i], j ? fc.lux_t : fc.lux_s, o[j]);
      if (v < mn[j]) mn[j] = v;
      if (v > mx[j]) mx[j] = v;
    }
  bool big = false;
  for (int i = 0; i < 2; ++i) {
    const float lo = (float)floor((double)mn[i]);
    const float hi = (float)ceil((double)mx[i]);
    // a side beyond the luxel limit (the reference logs "Bad surface
    // extents") or beyond int32 gets no luxels
    if (!(lo >= -2147483648.0f && hi <= 2147483520.0f)) {
      big = true;
      continue;
    }
    r.mins[i] = (int32_t)lo;
    const float sz = hi - lo;
    if (sz > (float)(VR_LIGHTMAP_MAX_DIM + 1)) big = true;
    r.size[i] = sz < 2147483520.0f ? (int32_t)sz : 0;
  }
  if (big) {
    r.flags = VR_LM_TOO_BIG;
  } else {
    r.n_luxels = (r.size[0] + 1) * (r.size[1] + 1);
  }
  out[f] = r;
}

void launch_face_extents(const vr_face *faces, int64_t n, const float *lux_ofs,
                         vr_face_lightmap *out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_face_extents, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, faces, n,
                     lux_ofs, out);
}

// LuxelSpaceToWorld: the point P with P . lux_s + o_s = us, P . lux_t + o_t
// = ut and P . n = n . points[0], by Cramer's rule in FP64 in this order.
// Returns false when the three vectors are dependent.
__device__ bool luxel_to_world(const vr_face &fc, const float *o, float us, float ut, float *P) {
  const double a0 = fc.lux_s[0], a1 = fc.lux_s[1], a2 = fc.lux_s[2];
  const double b0 = fc.lux_t[0], b1 = fc.lux_t[1], b2 = fc.lux_t[2];
  const double n0 = fc.normal[0], n1 = fc.normal[1], n2 = fc.normal[2];
  const double bn0 = b1 * n2 - b2 * n1, bn1 = b2 * n0 - b0 * n2, bn2 = b0 * n1 - b1 * n0;
  const double na0 = n1 * a2 - n2 * a1, na1 = n2 * a0 - n0 * a2, na2 = n0 * a1 - n1 * a0;
  const double ab0 = a1 * b2 - a2 * b1, ab1 = a2 * b0 - a0 * b2, ab2 = a0 * b1 - a1 * b0;
  const double det = (a0 * bn0 + a1 * bn1) + a2 * bn2;
  if (det == 0.0) {
    P[0] = P[1] = P[2] = 0.0f;
    return false;
  }
  const double d = ((n0 * (double)fc.points[0][0] + n1 * (double)fc.points[0][1]) +
                    n2 * (double)fc.points[0][2]);
  const double r0 = (double)us - (double)o[0], r1 = (double)ut - (double)o[1];
  P[0] = (float)(((r0 * bn0 + r1 * na0) + d * ab0) / det);
  P[1] = (float)(((r0 * bn1 + r1 * na1) + d * ab1) / det);
  P[2] = (float)(((r0 * bn2 + r1 * na2) + d * ab2) / det);
  return true;
}

constexpr int kLeafTile = 64;

__global__ __launch_bounds__(256) void k_face_luxels(
    const vr_face *__restrict__ faces, const float *__restrict__ lux_ofs,
    const vr_face_lightmap *__restrict__ lm, const int2 *__restrict__ blocks,
    const int32_t *__restrict__ leaf_ofs, const int32_t *__restrict__ leaf_idx,
    const vr_patch *__restrict__ patches, const float *__restrict__ patch_value,
    float *__restrict__ pos, float *__restrict__ nrm, int32_t *__restrict__ lpatch,
    float *__restrict__ value) {
  __shared__ float s_u[kLeafTile][VR_FACE_MAX_POINTS], s_v[kLeafTile][VR_FACE_MAX_POINTS];
  __shared__ float s_ou[kLeafTile], s_ov[kLeafTile];
  __shared__ int s_np[kLeafTile];
  const int2 bk = blocks[blockIdx.x];  // (face, first luxel of the block)
  const int f = bk.x;
  const vr_face &fc = faces[f];
  const vr_face_lightmap L = lm[f];
  const float o[2] = {lux_ofs ? lux_ofs[2 * (int64_t)f] : 0.0f,
                      lux_ofs ? lux_ofs[2 * (int64_t)f + 1] : 0.0f};
  const int l = bk.y + (int)threadIdx.x;
  const bool live = l < L.n_luxels;
  const int w = L.size[0] + 1;
  const float us = (float)(L.mins[0] + (live ? l % w : 0));
  const float ut = (float)(L.mins[1] + (live ? l / w : 0));
  int found = -1, best = -1;
  float bestd = 0.0f;
  const int l0 = leaf_ofs[f], l1 = leaf_ofs[f + 1];
  for (int t0 = l0; t0 < l1; t0 += kLeafTile) {
    const int nt = min(kLeafTile, l1 - t0);
    __syncthreads();
    if ((int)threadIdx.x < nt) {
      const vr_patch &q = patches[leaf_idx[t0 + threadIdx.x]];
      const int np = min(max(q.n_points, 0), VR_FACE_MAX_POINTS);
      for (int k = 0; k < np; ++k) {
        s_u[threadIdx.x][k] = lux_coord(q.points[k], fc.lux_s, o[0]);
        s_v[threadIdx.x][k] = lux_coord(q.points[k], fc.lux_t, o[1]);
      }
      s_np[threadIdx.x] = np;
      s_ou[threadIdx.x] = lux_coord(q.origin, fc.lux_s, o[0]);
      s_ov[threadIdx.x] = lux_coord(q.origin, fc.lux_t, o[1]);
    }
    __syncthreads();
    if (!live || found >= 0) continue;
    for (int k = 0; k < nt; ++k) {
      const int np = s_np[k];
      if (np >= 3) {
        bool pos_side = false, neg_side = false;
        for (int e = 0; e < np; ++e) {
          const int e1 = e + 1 == np ? 0 : e + 1;
          const float cr = (s_u[k][e1] - s_u[k][e]) * (ut - s_v[k][e]) -
                           (s_v[k][e1] - s_v[k][e]) * (us - s_u[k][e]);
          pos_side |= cr > 0.0f;
          neg_side |= cr < 0.0f;
        }
        if ((pos_side || neg_side) && !(pos_side && neg_side)) {
          found = t0 + k;
          break;
        }
      }
      const float du = us - s_ou[k], dv = ut - s_ov[k];
      const float d2 = du * du + dv * dv;
      if (best < 0 || d2 < bestd) {
        best = t0 + k;
        bestd = d2;
      }
    }
  }
  if (!live) return;
  const int64_t at = L.luxel_ofs + l;
  if (pos || nrm) {
    float P[3];
    luxel_to_world(fc, o, us, ut, P);
    if (pos) {
      pos[3 * at] = P[0];
      pos[3 * at + 1] = P[1];
      pos[3 * at + 2] = P[2];
    }
    if (nrm) {
      nrm[3 * at] = fc.normal[0];
      nrm[3 * at + 1] = fc.normal[1];
      nrm[3 * at + 2] = fc.normal[2];
    }
  }
  const int pick = found >= 0 ? found : best;
  const int p = pick >= 0 ? leaf_idx[pick] : -1;
  if (lpatch) lpatch[at] = p;
  if (value) {
    for (int c = 0; c < 3; ++c)
      value[3 * at + c] = (p >= 0 && patch_value) ? patch_value[3 * (int64_t)p + c] : 0.0f;
  }
}

void launch_face_luxels(const vr_face *faces, const float *lux_ofs, const vr_face_lightmap *lm,
                        const int2 *blocks, int64_t n_blocks, const int32_t *leaf_ofs,
                        const int32_t *leaf_idx, const vr_patch *patches,
                        const float *patch_value, float *pos, float *nrm, int32_t *lpatch,
                        float *value, hipStream_t st) {
  if (n_blocks <= 0) return;
  hipLaunchKernelGGL(k_face_luxels, dim3((unsigned)n_blocks), dim3(256), 0, st, faces, lux_ofs, lm,
                     blocks, leaf_ofs, leaf_idx, patches, patch_value, pos, nrm, lpatch, value);
}

// ---- the lighting lump: ColorRGBExp32 luxels -------------------------------
// VectorToColorRGBExp32, the Source engine's luxel encoding (named in the
// reference's compute/leafambient/lighting.go:67; the reference's lump writer
// is a stub, cmd/tasks/finish/main.go:15-18): e = floor(log2(max channel)) - 7
// puts the largest channel in [128, 256) after scaling by 2^-e, and each
// channel is truncated to a byte; bytes r, g, b, e (signed) little endian.
// Zero for a zero, negative or non-finite maximum, or e below -127.
__device__ __forceinline__ uint32_t rgbexp32(float r, float g, float b) {
  float m = r > g ? r : g;
  m = m > b ? m : b;
  if (!(m > 0.0f) || !(m <= 3.402823466e38f)) return 0u;
  int ex;
  (void)frexpf(m, &ex);
  const int e = (ex - 1) - 7;
  if (e < -127 || e > 127) return 0u;
  const float scale = ldexpf(1.0f, -e);
  const float c[3] = {r, g, b};
  uint32_t out = 0u;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float v = c[k] > 0.0f ? c[k] * scale : 0.0f;
    int q = (int)v;
    q = q > 255 ? 255 : q;
    out |= (uint32_t)q << (8 * k);
  }
  return out | ((uint32_t)(uint8_t)(int8_t)e << 24);
}

// One wave per face with luxels: its average colour (float64 lane partials of
// luxels k = lane mod 64, an xor butterfly 32 .. 1, the mean rounded to
// float32) at entry at[f], then its luxels.
__global__ __launch_bounds__(256) void k_lighting_lump(const float *__restrict__ lux,
                                                       const vr_face_lightmap *__restrict__ lm,
                                                       const int64_t *__restrict__ at,
                                                       int64_t n_faces,
                                                       uint32_t *__restrict__ lump) {
  const int lane = threadIdx.x & 63;
  const int64_t f = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (f >= n_faces) return;
  const int64_t n = lm[f].n_luxels;
  if (n <= 0) return;
  const float *v = lux + 3 * lm[f].luxel_ofs;
  uint32_t *out = lump + at[f];
  double sr = 0.0, sg = 0.0, sb = 0.0;
  for (int64_t k = lane; k < n; k += 64) {
    const float r = v[3 * k], g = v[3 * k + 1], b = v[3 * k + 2];
    sr = sr + (double)r;
    sg = sg + (double)g;
    sb = sb + (double)b;
    out[1 + k] = rgbexp32(r, g, b);
  }
  for (int off = 32; off >= 1; off >>= 1) {
    sr = sr + __shfl_xor(sr, off);
    sg = sg + __shfl_xor(sg, off);
    sb = sb + __shfl_xor(sb, off);
  }
  if (lane == 0)
    out[0] = rgbexp32((float)(sr / (double)n), (float)(sg / (double)n), (float)(sb / (double)n));
}

void launch_lighting_lump(const float *lux, const vr_face_lightmap *lm, const int64_t *at,
                          int64_t n_faces, uint32_t *lump, hipStream_t st) {
  if (n_faces <= 0) return;
  hipLaunchKernelGGL(k_lighting_lump, dim3((unsigned)((n_faces + 3) / 4)), dim3(256), 0, st, lux,
                     lm, at, n_faces, lump);
}

}  // namespace vr
