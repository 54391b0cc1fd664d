// Patch generation on the GPU: MakePatchForFace (rad/patches/face.go:30-197)
// and SubdividePatches (rad/patches/subdivide.go:22-438).
//
// Two builders give the same patches in the same order (the reference's:
// base patches in face order, then each base patch's children in the order
// SubdividePatch creates them, child 1's subtree before child 2's):
//  - level-parallel (the default): every patch of one subdivision level is
//    split by its own thread, the children form the next level, and the
//    nodes are then renumbered (descendant counts bottom-up, positions
//    top-down) and scattered;
//  - a thread per face running the reference's depth-first recursion with
//    the pending second children on a per-face stack (count pass, scan,
//    fill pass).
// Subdivision is control flow over a few points per patch, so a thread per
// patch (not a wave) is the unit of work.
//
// Float operations follow mathgl's float32 vector ops, one rounding each
// (built with -ffp-contract=off), as the CPU restatement does.  Where the Go
// code does not do what it intends, the intent: WindingCenter
// (vmath/polygon/winding.go:231-239) averages the points (scale 1 / n), and
// WindingAreaAndBalancePoint (subdivide.go:412-438) returns its area-weighted
// centre.  A child's normal is GetPhongNormal (rad/lightmap/
// normallist.go:51-144) at its centre when per-vertex normals (PairEdges)
// are given, else its face's normal.
#include <cstring>
#include <vector>

#include "kernels.h"
#include "vr_internal.h"

namespace vr {
namespace {

constexpr int kMaxPts = VR_FACE_MAX_POINTS;
constexpr float kOnEpsilon = 0.1f;  // vmath.ON_EPSILON
constexpr int kMaxDepth = 40;       // pending second children per face (tree depth)

struct Wind {
  int n;
  float p[kMaxPts][3];
};

__device__ __forceinline__ float len3(const float v[3]) {
  return __builtin_sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}
__device__ __forceinline__ void cross3(const float a[3], const float b[3], float o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
// WindingArea (winding.go:217-229)
__device__ float wind_area(const Wind &w) {
  float total = 0.0f;
  for (int i = 2; i < w.n; ++i) {
    float d1[3], d2[3], c[3];
    for (int k = 0; k < 3; ++k) {
      d1[k] = w.p[i - 1][k] - w.p[0][k];
      d2[k] = w.p[i][k] - w.p[0][k];
    }
    cross3(d1, d2, c);
    total += len3(c);
  }
  return total * 0.5f;
}
// WindingBounds (winding.go:241-260)
__device__ void wind_bounds(const Wind &w, float mn[3], float mx[3]) {
  for (int k = 0; k < 3; ++k) {
    mn[k] = 99999.0f;
    mx[k] = -99999.0f;
  }
  for (int i = 0; i < w.n; ++i)
    for (int k = 0; k < 3; ++k) {
      const float v = w.p[i][k];
      if (v < mn[k]) mn[k] = v;
      if (v > mx[k]) mx[k] = v;
    }
}
// WindingAreaAndBalancePoint (subdivide.go:412-438); has = false: no winding
__device__ float wind_area_balance(const Wind &w, bool has, float c[3]) {
  float total = 0.0f;
  c[0] = c[1] = c[2] = 0.0f;
  if (!has) return 0.0f;
  for (int i = 2; i < w.n; ++i) {
    float d1[3], d2[3], x[3];
    for (int k = 0; k < 3; ++k) {
      d1[k] = w.p[i - 1][k] - w.p[0][k];
      d2[k] = w.p[i][k] - w.p[0][k];
    }
    cross3(d1, d2, x);
    const float area = len3(x);
    total += area;
    const float s = area / 3.0f;
    for (int k = 0; k < 3; ++k) c[k] = c[k] + s * w.p[i - 1][k];
    for (int k = 0; k < 3; ++k) c[k] = c[k] + s * w.p[i][k];
    for (int k = 0; k < 3; ++k) c[k] = c[k] + s * w.p[0][k];
  }
  if (total != 0.0f) {
    const float inv = 1.0f / total;
    for (int k = 0; k < 3; ++k) c[k] = c[k] * inv;
  }
  return total * 0.5f;
}
// ClipWindingEpsilon (subdivide.go:269-347) against x[axis] = dist.  Returns
// false past the point capacity.
__device__ bool clip_axis(const Wind &in, int axis, float dist, Wind &f, bool &has_f, Wind &b,
                          bool &has_b) {
  float dists[kMaxPts + 1];
  int sides[kMaxPts + 1], counts[3] = {0, 0, 0};
  const float nrm[3] = {axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
  int i;
  for (i = 0; i < in.n; ++i) {
    float dot = (in.p[i][0] * nrm[0] + in.p[i][1] * nrm[1]) + in.p[i][2] * nrm[2];
    dot = dot - dist;
    dists[i] = dot;
    sides[i] = dot > kOnEpsilon ? 0 : (dot < -kOnEpsilon ? 1 : 2);
    counts[sides[i]]++;
  }
  sides[i] = sides[0];
  dists[i] = dists[0];
  has_f = has_b = false;
  if (counts[0] == 0) {
    b = in;
    has_b = true;
    return true;
  }
  if (counts[1] == 0) {
    f = in;
    has_f = true;
    return true;
  }
  f.n = b.n = 0;
  has_f = has_b = true;
  for (i = 0; i < in.n; ++i) {
    if (f.n + 2 > kMaxPts || b.n + 2 > kMaxPts) return false;
    const float *p1 = in.p[i];
    if (sides[i] == 2) {
      for (int k = 0; k < 3; ++k) f.p[f.n][k] = b.p[b.n][k] = p1[k];
      ++f.n;
      ++b.n;
      continue;
    }
    if (sides[i] == 0) {
      for (int k = 0; k < 3; ++k) f.p[f.n][k] = p1[k];
      ++f.n;
    }
    if (sides[i] == 1) {
      for (int k = 0; k < 3; ++k) b.p[b.n][k] = p1[k];
      ++b.n;
    }
    if (sides[i + 1] == 2 || sides[i + 1] == sides[i]) continue;
    const float *p2 = in.p[(i + 1) % in.n];
    const float dot = dists[i] / (dists[i] - dists[i + 1]);
    for (int j = 0; j < 3; ++j) {
      float m;
      if (nrm[j] == 1.0f) m = dist;
      else if (nrm[j] == -1.0f) m = -dist;
      else m = p1[j] + dot * (p2[j] - p1[j]);
      f.p[f.n][j] = m;
      b.p[b.n][j] = m;
    }
    ++f.n;
    ++b.n;
  }
  return true;
}

__device__ __forceinline__ bool zero_light(const float v[3]) {
  return fabsf(v[0]) < 1e-20f && fabsf(v[1]) < 1e-20f && fabsf(v[2]) < 1e-20f;
}

// MakePatchForFace (face.go:30-197): false when the face has no patch
// (area <= 0).
__device__ bool base_patch(const vr_face &fc, int32_t face, float max_chop, vr_patch &P) {
  Wind w;
  w.n = fc.n_points;
  for (int i = 0; i < w.n; ++i)
    for (int k = 0; k < 3; ++k) w.p[i][k] = fc.points[i][k];
  const float area = wind_area(w);
  if (area <= 0.0f) return false;
  P.face = face;
  P.parent = P.child1 = P.child2 = -1;
  float cs0 = 0.0f, cs1 = 0.0f;
  for (int j = 0; j < 3; ++j) cs0 += fc.lux_s[j] * fc.lux_s[j];
  for (int j = 0; j < 3; ++j) cs1 += fc.lux_t[j] * fc.lux_t[j];
  cs0 = __builtin_sqrtf(cs0);
  cs1 = __builtin_sqrtf(cs1);
  P.lux_scale = (cs0 + cs1) / 2.0f;
  P.chop = max_chop;
  P.area = area;
  float c[3] = {0.0f, 0.0f, 0.0f};
  for (int i = 0; i < w.n; ++i)
    for (int k = 0; k < 3; ++k) c[k] = w.p[i][k] + c[k];
  const float sc = 1.0f / (float)w.n;
  for (int k = 0; k < 3; ++k) P.origin[k] = c[k] * sc;
  for (int k = 0; k < 3; ++k) P.normal[k] = fc.normal[k];
  wind_bounds(w, P.mins, P.maxs);
  P.n_points = w.n;
  for (int i = 0; i < kMaxPts; ++i)
    for (int k = 0; k < 3; ++k) P.points[i][k] = i < w.n ? w.p[i][k] : 0.0f;
  return true;
}

// GetPhongNormal (rad/lightmap/normallist.go:51-144) at spot on face fc:
// c is the face centroid (the base patch's centre, GetFaceCentroids), vn the
// face's per-vertex normals (PairEdges).  For the first edge (j, j + 1)
// whose wedge from the centroid holds the spot (a1 >= 0 and a2 >= 0 in the
// wedge's coordinates), the face normal, vn[j] and vn[j + 1] are blended
// with weights 1 - a1 - a2, a1, a2 and normalised (mgl32 ops, one rounding
// each); otherwise the face normal.  A blend of length below 1e-20 (where
// the Go code stops with log.Fatalf) keeps the face normal.
__device__ void phong_normal(const vr_face &fc, const float c[3], const float *__restrict__ vn,
                             const float spot[3], float out[3]) {
  auto dot = [](const float *a, const float *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; };
  for (int k = 0; k < 3; ++k) out[k] = fc.normal[k];
  const int E = fc.n_points;
  for (int j = 0; j < E; ++j) {
    const int j2 = (j + 1) % E;
    float v1[3], v2[3], vs[3];
    for (int k = 0; k < 3; ++k) {
      v1[k] = fc.points[j][k] - c[k];
      v2[k] = fc.points[j2][k] - c[k];
      vs[k] = spot[k] - c[k];
    }
    const float aa = dot(v1, v1), bb = dot(v2, v2), ab = dot(v1, v2);
    const float a1 = (bb * dot(v1, vs) - ab * dot(vs, v2)) / (aa * bb - ab * ab);
    const float a2 = (dot(vs, v2) - a1 * ab) / bb;
    if (a1 >= 0.0f && a2 >= 0.0f) {
      const float sc = (1.0f - a1) - a2;
      float ph[3];
      for (int k = 0; k < 3; ++k) ph[k] = fc.normal[k] * sc;
      for (int k = 0; k < 3; ++k) ph[k] = ph[k] + vn[3 * j + k] * a1;
      for (int k = 0; k < 3; ++k) ph[k] = ph[k] + vn[3 * j2 + k] * a2;
      const float len = __builtin_sqrtf(dot(ph, ph));
      if (len < 1.0e-20f) return;  // the reference stops the bake here
      const float inv = 1.0f / len;
      for (int k = 0; k < 3; ++k) out[k] = ph[k] * inv;
      return;
    }
  }
}

// SubdividePatch's step for one patch P (subdivide.go:166-262 with
// CreateChildPatch, :353-410): decides the split (updating P.chop as the
// reference does) and fills the two children (parent / child indices left
// to the caller).  face_mn / face_mx are the face's base-patch bounds,
// centroid its base patch's centre, vn its per-vertex normals (nullable).
// Returns 1 split, 0 leaf, -1 a winding past the point capacity.
__device__ int split_step(const vr_face &fc, vr_patch &P, const float face_mn[3],
                          const float face_mx[3], const float centroid[3],
                          const float *__restrict__ vn, float min_chop, vr_patch ch[2]) {
  bool split = false;
  Wind o1, o2;
  float c1[3], c2[3], a1 = 0.0f, a2 = 0.0f;
  if (!(fc.flags & VR_FACE_SKY)) {
    float total[3];
    float widest = -1.0f;
    int wa = -1;
    bool should = false;
    for (int i = 0; i < 3; ++i) total[i] = (P.maxs[i] - P.mins[i]) * P.lux_scale;
    for (int i = 0; i < 3; ++i) {
      if (total[i] > widest) {
        wa = i;
        widest = total[i];
      }
      if (total[i] >= P.chop && total[i] >= min_chop) should = true;
    }
    if (!should && wa != -1) {
      if (total[wa] > total[(wa + 1) % 3] * 2.0f && total[wa] > total[(wa + 2) % 3] * 2.0f) {
        if (P.chop > min_chop) {
          should = true;
          const float h = P.chop / 2.0f;
          P.chop = min_chop > h ? min_chop : h;
        }
      }
    }
    if (should) {
      const float dist = (P.mins[wa] + P.maxs[wa]) * 0.5f;
      Wind w;
      w.n = P.n_points;
      for (int i = 0; i < w.n; ++i)
        for (int q = 0; q < 3; ++q) w.p[i][q] = P.points[i][q];
      bool h1, h2;
      if (!clip_axis(w, wa, dist, o1, h1, o2, h2)) return -1;
      a1 = wind_area_balance(o1, h1, c1);
      a2 = wind_area_balance(o2, h2, c2);
      split = !(a1 == 0.0f || a2 == 0.0f);
    }
  }
  if (!split) return 0;
  const Wind *ow[2] = {&o1, &o2};
  const float *ca[2] = {c1, c2};
  const float ar[2] = {a1, a2};
  for (int q = 0; q < 2; ++q) {
    vr_patch &C = ch[q];
    C = P;
    C.child1 = C.child2 = -1;
    C.n_points = ow[q]->n;
    for (int i = 0; i < kMaxPts; ++i)
      for (int z = 0; z < 3; ++z) C.points[i][z] = i < ow[q]->n ? ow[q]->p[i][z] : 0.0f;
    C.area = ar[q];
    for (int z = 0; z < 3; ++z) C.origin[z] = ca[q][z];
    if (vn)
      phong_normal(fc, centroid, vn, C.origin, C.normal);
    else
      for (int z = 0; z < 3; ++z) C.normal[z] = fc.normal[z];
    wind_bounds(*ow[q], C.mins, C.maxs);
    if (zero_light(fc.base_light)) {
      float t[3];
      for (int i = 0; i < 3; ++i) t[i] = (C.maxs[i] - C.mins[i]) * C.lux_scale;
      if (C.chop > min_chop && t[0] < C.chop && t[1] < C.chop && t[2] < C.chop) {
        for (int i = 0; i < 3; ++i) {
          if ((face_mx[i] == C.maxs[i] || face_mn[i] == C.mins[i]) && t[i] > min_chop) {
            const float h = C.chop / 2.0f;
            C.chop = min_chop > h ? min_chop : h;
            break;
          }
        }
      }
    }
  }
  return 1;
}

struct Pending {  // a second child waiting for its subdivision
  int64_t idx;
  vr_patch p;
};

// The subdivision of one base patch (index bidx), children written (WRITE)
// from out[cbase] on; returns the number of children, or -1 when a winding
// or the stack overflows.
template <bool WRITE>
__device__ int64_t subdivide_face(const vr_face &fc, vr_patch P, int64_t bidx, int64_t cbase,
                                  float min_chop, const float *vn, vr_patch *out, Pending *stk) {
  const float face_mn[3] = {P.mins[0], P.mins[1], P.mins[2]};
  const float face_mx[3] = {P.maxs[0], P.maxs[1], P.maxs[2]};
  const float centroid[3] = {P.origin[0], P.origin[1], P.origin[2]};
  int64_t made = 0, k = bidx;
  int sp = 0;
  for (;;) {
    vr_patch ch[2];
    const int r = split_step(fc, P, face_mn, face_mx, centroid, vn, min_chop, ch);
    if (r < 0) return -1;
    if (r) {
      const int64_t i1 = cbase + made, i2 = cbase + made + 1;
      made += 2;
      ch[0].parent = ch[1].parent = (int32_t)k;
      P.child1 = (int32_t)i1;
      P.child2 = (int32_t)i2;
      if (WRITE) {
        out[k] = P;
        out[i1] = ch[0];
        out[i2] = ch[1];
      }
      if (sp >= kMaxDepth) return -1;
      stk[sp].idx = i2;
      stk[sp].p = ch[1];
      ++sp;
      k = i1;  // child 1 next, child 2 after child 1's subtree
      P = ch[0];
      continue;
    }
    if (WRITE) out[k] = P;  // final state (its chop may have changed)
    if (sp == 0) break;
    --sp;
    k = stk[sp].idx;
    P = stk[sp].p;
  }
  return made;
}

// Pass 1: base[f] = 1 if the face has a patch; cnt[f] = its children
// (-1: overflow).  Pass 2 (WRITE): the base patch at bofs[f], children from
// cofs[f].
template <bool WRITE>
__global__ __launch_bounds__(64) void k_make_patches(const vr_face *__restrict__ faces,
                                                     int64_t n_faces, float max_chop,
                                                     float min_chop, int32_t *__restrict__ base,
                                                     int64_t *__restrict__ cnt,
                                                     const int64_t *__restrict__ bofs,
                                                     const int64_t *__restrict__ cofs,
                                                     vr_patch *__restrict__ out,
                                                     Pending *__restrict__ stacks,
                                                     const float *__restrict__ vnorm) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  const vr_face fc = faces[f];
  if (fc.n_points < 0 || fc.n_points > kMaxPts) {
    if (!WRITE) {
      base[f] = 0;
      cnt[f] = -1;
    }
    return;
  }
  vr_patch P;
  const bool has = base_patch(fc, (int32_t)f, max_chop, P);
  if (!WRITE) base[f] = has ? 1 : 0;
  if (!has) {
    if (!WRITE) cnt[f] = 0;
    return;
  }
  Pending *stk = stacks + f * kMaxDepth;
  if (fc.flags & VR_FACE_NOCHOP) {  // PreventSubdivision
    if (WRITE) out[bofs[f]] = P;
    else cnt[f] = 0;
    return;
  }
  const int64_t m = subdivide_face<WRITE>(fc, P, WRITE ? bofs[f] : 0, WRITE ? cofs[f] : 0,
                                          min_chop,
                                          vnorm ? vnorm + f * (kMaxPts * 3) : nullptr, out, stk);
  if (!WRITE) cnt[f] = m;
}

// ---- level-parallel subdivision -------------------------------------------
// Every patch of one subdivision level is split by its own thread; the
// children form the next level.  After the last level the nodes are
// renumbered into the reference's array order: base patches in face order,
// then, for each split, its two children back to back followed by child 1's
// descendants and then child 2's (desc = descendants of a node, bottom-up;
// positions top-down), and scattered with final parent / child indices.

// Base patches: one thread per face; has[f] = 1 when the face has a patch
// (tmp[f] the patch, fb[f] its bounds).
__global__ __launch_bounds__(64) void k_lv_base(const vr_face *__restrict__ faces, int64_t n,
                                                float max_chop, vr_patch *__restrict__ tmp,
                                                uint8_t *__restrict__ has) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const vr_face &fc = faces[f];
  has[f] = 0;
  if (fc.n_points < 0 || fc.n_points > kMaxPts) return;
  vr_patch P;
  if (!base_patch(fc, (int32_t)f, max_chop, P)) return;
  tmp[f] = P;
  has[f] = 1;
}

// Level 0 from the faces with a patch, in face order (rank = base index).
__global__ __launch_bounds__(256) void k_lv_gather(const vr_patch *__restrict__ tmp,
                                                   const int32_t *__restrict__ face_of, int64_t n,
                                                   vr_patch *__restrict__ lv,
                                                   float *__restrict__ fbounds) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const vr_patch P = tmp[face_of[i]];
  lv[i] = P;
  for (int k = 0; k < 3; ++k) {  // base bounds and centre (the face centroid)
    fbounds[9 * (int64_t)P.face + k] = P.mins[k];
    fbounds[9 * (int64_t)P.face + 3 + k] = P.maxs[k];
    fbounds[9 * (int64_t)P.face + 6 + k] = P.origin[k];
  }
}

// One level: split[i] = 1 when node i splits (its record updated in place,
// its children in kids[2 i], kids[2 i + 1]); err = a winding overflow.
__global__ __launch_bounds__(64) void k_lv_split(const vr_face *__restrict__ faces,
                                                 const float *__restrict__ fbounds,
                                                 vr_patch *__restrict__ lv, int64_t n,
                                                 float min_chop, vr_patch *__restrict__ kids,
                                                 uint8_t *__restrict__ split,
                                                 uint32_t *__restrict__ err,
                                                 const float *__restrict__ vnorm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vr_patch P = lv[i];
  const vr_face &fc = faces[P.face];
  split[i] = 0;
  if (fc.flags & VR_FACE_NOCHOP) return;  // PreventSubdivision: never subdivided
  const float *b = fbounds + 9 * (int64_t)P.face;
  vr_patch ch[2];
  const int r = split_step(fc, P, b, b + 3, b + 6,
                           vnorm ? vnorm + (int64_t)P.face * (kMaxPts * 3) : nullptr, min_chop, ch);
  if (r < 0) {
    atomicOr(err, 1u);
    return;
  }
  lv[i] = P;
  if (r) {
    kids[2 * i] = ch[0];
    kids[2 * i + 1] = ch[1];
    split[i] = 1;
  }
}

// The next level: the children of split node i at 2 cbase[i], 2 cbase[i] + 1
// (par: their parent's index in this level).
__global__ __launch_bounds__(256) void k_lv_emit(const vr_patch *__restrict__ kids,
                                                 const uint8_t *__restrict__ split,
                                                 const int32_t *__restrict__ cbase, int64_t n,
                                                 vr_patch *__restrict__ next,
                                                 int32_t *__restrict__ par) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n || !split[i]) return;
  const int64_t c = 2 * (int64_t)cbase[i];
  next[c] = kids[2 * i];
  next[c + 1] = kids[2 * i + 1];
  par[c] = par[c + 1] = (int32_t)i;
}

// desc[i] = descendants of node i (next-level desc known).
__global__ __launch_bounds__(256) void k_lv_desc(const uint8_t *__restrict__ split,
                                                 const int32_t *__restrict__ cbase, int64_t n,
                                                 const int64_t *__restrict__ desc_next,
                                                 int64_t *__restrict__ desc) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  desc[i] = split[i] ? 2 + desc_next[2 * (int64_t)cbase[i]] + desc_next[2 * (int64_t)cbase[i] + 1]
                     : 0;
}

// Final indices of the next level from this level's (fin) and the position
// of each split node's child pair (pair).
__global__ __launch_bounds__(256) void k_lv_place(const uint8_t *__restrict__ split,
                                                  const int32_t *__restrict__ cbase, int64_t n,
                                                  const int64_t *__restrict__ pair,
                                                  const int64_t *__restrict__ desc_next,
                                                  int64_t *__restrict__ fin_next,
                                                  int64_t *__restrict__ pair_next) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n || !split[i]) return;
  const int64_t c1 = 2 * (int64_t)cbase[i], c2 = c1 + 1, q = pair[i];
  fin_next[c1] = q;
  fin_next[c2] = q + 1;
  pair_next[c1] = q + 2;
  pair_next[c2] = q + 2 + desc_next[c1];
}

// out[fin[i]] = node i with final parent and child indices.
__global__ __launch_bounds__(256) void k_lv_scatter(const vr_patch *__restrict__ lv, int64_t n,
                                                    const int64_t *__restrict__ fin,
                                                    const int32_t *__restrict__ par,
                                                    const int64_t *__restrict__ fin_prev,
                                                    const uint8_t *__restrict__ split,
                                                    const int32_t *__restrict__ cbase,
                                                    const int64_t *__restrict__ fin_next,
                                                    vr_patch *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  vr_patch P = lv[i];
  P.parent = fin_prev ? (int32_t)fin_prev[par[i]] : -1;
  if (split[i]) {
    P.child1 = (int32_t)fin_next[2 * (int64_t)cbase[i]];
    P.child2 = (int32_t)fin_next[2 * (int64_t)cbase[i] + 1];
  } else {
    P.child1 = P.child2 = -1;
  }
  out[fin[i]] = P;
}

}  // namespace

// Level-parallel MakePatches + SubdividePatches; the same patches, in the
// same order, as the thread-per-face kernel.
static int make_patches_levels(const vr_face *faces, int64_t n_faces, float max_chop,
                               float min_chop, const float *vnorm, vr_patch *out, int64_t cap,
                               int64_t *n_out, hipStream_t st) {
  struct Buf {  // owns one device allocation; moves, never copies
    void *p = nullptr;
    Buf() = default;
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    Buf(Buf &&o) noexcept : p(o.p) { o.p = nullptr; }
    ~Buf() {
      if (p) (void)hipFree(p);
    }
  };
  struct Level {
    Buf lv, par, split, cbase, desc, fin, pair;  // par: parent in the level above
    int64_t n = 0;
  };
  constexpr int kMaxLevels = 64;
  std::vector<Level> L(1);
  L.reserve(kMaxLevels + 1);
  Buf d_faces, d_tmp, d_has, d_fof, d_fb, d_kids, d_err, d_out, d_vn;
  int rc = VR_OK;
  auto H = [&](hipError_t e, const char *what) {
    if (e == hipSuccess) return VR_OK;
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? VR_E_NOMEM : VR_E_HIP,
                std::string("make patches: ") + what + ": " + hipGetErrorString(e));
  };
  auto M = [&](Buf &b, size_t bytes) { return H(hipMalloc(&b.p, bytes ? bytes : 16), "hipMalloc"); };
  const unsigned gf = (unsigned)((n_faces + 63) / 64);
  if ((rc = M(d_faces, sizeof(vr_face) * n_faces)) || (rc = M(d_tmp, sizeof(vr_patch) * n_faces)) ||
      (rc = M(d_has, n_faces)) || (rc = M(d_fb, 36 * n_faces)) || (rc = M(d_err, 4)) ||
      (rc = H(hipMemsetAsync(d_err.p, 0, 4, st), "memset")) ||
      (rc = H(hipMemcpyAsync(d_faces.p, faces, sizeof(vr_face) * n_faces, hipMemcpyHostToDevice, st),
              "upload")))
    return rc;
  const size_t vnb = sizeof(float) * kMaxPts * 3 * (size_t)n_faces;
  if (vnorm && ((rc = M(d_vn, vnb)) ||
                (rc = H(hipMemcpyAsync(d_vn.p, vnorm, vnb, hipMemcpyHostToDevice, st), "upload"))))
    return rc;
  hipLaunchKernelGGL(k_lv_base, dim3(gf), dim3(64), 0, st, (const vr_face *)d_faces.p, n_faces,
                     max_chop, (vr_patch *)d_tmp.p, (uint8_t *)d_has.p);
  std::vector<uint8_t> has((size_t)n_faces);
  if ((rc = H(hipGetLastError(), "launch")) ||
      (rc = H(hipMemcpyAsync(has.data(), d_has.p, n_faces, hipMemcpyDeviceToHost, st), "download")) ||
      (rc = H(hipStreamSynchronize(st), "sync")))
    return rc;
  std::vector<int32_t> face_of;
  for (int64_t f = 0; f < n_faces; ++f)
    if (has[(size_t)f]) face_of.push_back((int32_t)f);
  const int64_t nb = (int64_t)face_of.size();
  L[0].n = nb;
  if (nb > 0) {
    if ((rc = M(d_fof, 4 * nb)) || (rc = M(L[0].lv, sizeof(vr_patch) * nb)) ||
        (rc = H(hipMemcpyAsync(d_fof.p, face_of.data(), 4 * nb, hipMemcpyHostToDevice, st), "upload")))
      return rc;
    hipLaunchKernelGGL(k_lv_gather, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st,
                       (const vr_patch *)d_tmp.p, (const int32_t *)d_fof.p, nb,
                       (vr_patch *)L[0].lv.p, (float *)d_fb.p);
  }
  // levels down to the leaves
  std::vector<uint8_t> sp;
  std::vector<int32_t> cb;
  for (size_t l = 0; L[l].n > 0; ++l) {
    if (l >= (size_t)kMaxLevels)
      return fail(VR_E_CAPACITY, "make patches: more than 64 subdivision levels");
    Level &C = L[l];
    const int64_t n = C.n;
    Buf kids;
    if ((rc = M(kids, sizeof(vr_patch) * 2 * n)) || (rc = M(C.split, n)) || (rc = M(C.cbase, 4 * n)))
      return rc;
    hipLaunchKernelGGL(k_lv_split, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st,
                       (const vr_face *)d_faces.p, (const float *)d_fb.p, (vr_patch *)C.lv.p, n,
                       min_chop, (vr_patch *)kids.p, (uint8_t *)C.split.p, (uint32_t *)d_err.p,
                       (const float *)d_vn.p);
    sp.resize((size_t)n);
    uint32_t err = 0;
    if ((rc = H(hipGetLastError(), "launch")) ||
        (rc = H(hipMemcpyAsync(sp.data(), C.split.p, n, hipMemcpyDeviceToHost, st), "download")) ||
        (rc = H(hipMemcpyAsync(&err, d_err.p, 4, hipMemcpyDeviceToHost, st), "download")) ||
        (rc = H(hipStreamSynchronize(st), "sync")))
      return rc;
    if (err) return fail(VR_E_CAPACITY, "make patches: a winding exceeds 16 points");
    cb.resize((size_t)n);
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i) {
      cb[(size_t)i] = (int32_t)k;
      k += sp[(size_t)i];
    }
    if (2 * k >= (1ll << 31)) return fail(VR_E_CAPACITY, "make patches: more than 2^31 patches");
    if ((rc = H(hipMemcpyAsync(C.cbase.p, cb.data(), 4 * n, hipMemcpyHostToDevice, st), "upload")))
      return rc;
    L.emplace_back();
    Level &N = L.back();
    Level &Cc = L[l];  // emplace_back may move the vector (reserved: no reallocation)
    N.n = 2 * k;
    if (N.n > 0) {
      if ((rc = M(N.lv, sizeof(vr_patch) * N.n)) || (rc = M(N.par, 4 * N.n))) return rc;
      hipLaunchKernelGGL(k_lv_emit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                         (const vr_patch *)kids.p, (const uint8_t *)Cc.split.p,
                         (const int32_t *)Cc.cbase.p, n, (vr_patch *)N.lv.p, (int32_t *)N.par.p);
    }
    if ((rc = H(hipGetLastError(), "launch")) || (rc = H(hipStreamSynchronize(st), "sync")))
      return rc;  // kids is freed at the end of this iteration
  }
  const size_t nl = L.size() - 1;  // L[nl] is the empty level
  int64_t total = 0;
  for (size_t l = 0; l < nl; ++l) total += L[l].n;
  *n_out = total;
  if (!out) return VR_OK;
  if (cap < total)
    return fail(VR_E_CAPACITY, "make patches: output holds " + std::to_string(cap) + " of " +
                                   std::to_string(total) + " patches");
  if (total == 0) return VR_OK;
  // descendants bottom-up
  for (size_t l = nl; l-- > 0;) {
    Level &C = L[l];
    if ((rc = M(C.desc, 8 * C.n))) return rc;
    const int64_t *dn = l + 1 < nl ? (const int64_t *)L[l + 1].desc.p : nullptr;
    if (!dn) {
      if ((rc = H(hipMemsetAsync(C.desc.p, 0, 8 * C.n, st), "memset"))) return rc;
    } else {
      hipLaunchKernelGGL(k_lv_desc, dim3((unsigned)((C.n + 255) / 256)), dim3(256), 0, st,
                         (const uint8_t *)C.split.p, (const int32_t *)C.cbase.p, C.n, dn,
                         (int64_t *)C.desc.p);
    }
  }
  // final indices top-down: base patches 0..nb-1, their child pairs after all
  // base patches in base order
  {
    std::vector<int64_t> d0((size_t)nb), fin0((size_t)nb), pair0((size_t)nb);
    if ((rc = H(hipMemcpyAsync(d0.data(), L[0].desc.p, 8 * nb, hipMemcpyDeviceToHost, st), "download")) ||
        (rc = H(hipStreamSynchronize(st), "sync")))
      return rc;
    int64_t at = nb;
    for (int64_t i = 0; i < nb; ++i) {
      fin0[(size_t)i] = i;
      pair0[(size_t)i] = at;
      at += d0[(size_t)i];
    }
    if ((rc = M(L[0].fin, 8 * nb)) || (rc = M(L[0].pair, 8 * nb)) ||
        (rc = H(hipMemcpyAsync(L[0].fin.p, fin0.data(), 8 * nb, hipMemcpyHostToDevice, st), "upload")) ||
        (rc = H(hipMemcpyAsync(L[0].pair.p, pair0.data(), 8 * nb, hipMemcpyHostToDevice, st), "upload")) ||
        (rc = H(hipStreamSynchronize(st), "sync")))
      return rc;
  }
  for (size_t l = 0; l + 1 < nl; ++l) {
    Level &C = L[l], &N = L[l + 1];
    if ((rc = M(N.fin, 8 * N.n)) || (rc = M(N.pair, 8 * N.n))) return rc;
    hipLaunchKernelGGL(k_lv_place, dim3((unsigned)((C.n + 255) / 256)), dim3(256), 0, st,
                       (const uint8_t *)C.split.p, (const int32_t *)C.cbase.p, C.n,
                       (const int64_t *)C.pair.p, (const int64_t *)N.desc.p, (int64_t *)N.fin.p,
                       (int64_t *)N.pair.p);
  }
  if ((rc = M(d_out, sizeof(vr_patch) * total))) return rc;
  for (size_t l = 0; l < nl; ++l) {
    Level &C = L[l];
    hipLaunchKernelGGL(k_lv_scatter, dim3((unsigned)((C.n + 255) / 256)), dim3(256), 0, st,
                       (const vr_patch *)C.lv.p, C.n, (const int64_t *)C.fin.p,
                       (const int32_t *)C.par.p,
                       l > 0 ? (const int64_t *)L[l - 1].fin.p : (const int64_t *)nullptr,
                       (const uint8_t *)C.split.p, (const int32_t *)C.cbase.p,
                       l + 1 < nl ? (const int64_t *)L[l + 1].fin.p : (const int64_t *)nullptr,
                       (vr_patch *)d_out.p);
  }
  if ((rc = H(hipGetLastError(), "launch")) ||
      (rc = H(hipMemcpyAsync(out, d_out.p, sizeof(vr_patch) * total, hipMemcpyDeviceToHost, st),
              "download")) ||
      (rc = H(hipStreamSynchronize(st), "sync")))
    return rc;
  return VR_OK;
}

int make_patches_device(const vr_face *faces, int64_t n_faces, float max_chop, float min_chop,
                        const float *vnorm, vr_patch *out, int64_t cap, int64_t *n_out,
                        hipStream_t st) {
  *n_out = 0;
  if (n_faces == 0) return VR_OK;
  if (tuning().patch_build == 1)
    return make_patches_levels(faces, n_faces, max_chop, min_chop, vnorm, out, cap, n_out, st);
  void *d_faces = nullptr, *d_base = nullptr, *d_cnt = nullptr, *d_ofs = nullptr,
       *d_out = nullptr, *d_stk = nullptr, *d_vn = nullptr;
  int rc = VR_OK;
  auto release = [&]() {
    (void)hipStreamSynchronize(st);
    for (void *p : {d_faces, d_base, d_cnt, d_ofs, d_out, d_stk, d_vn})
      if (p) (void)hipFree(p);
  };
  auto H = [&](hipError_t e, const char *what) {
    if (e == hipSuccess) return VR_OK;
    (void)hipGetLastError();
    return fail(VR_E_HIP, std::string("make patches: ") + what + ": " + hipGetErrorString(e));
  };
  const size_t fb = sizeof(vr_face) * (size_t)n_faces;
  const size_t sb = sizeof(Pending) * kMaxDepth * (size_t)n_faces;
  if ((rc = H(hipMalloc(&d_faces, fb), "hipMalloc")) || (rc = H(hipMalloc(&d_base, 4 * n_faces), "hipMalloc")) ||
      (rc = H(hipMalloc(&d_cnt, 8 * n_faces), "hipMalloc")) ||
      (rc = H(hipMalloc(&d_ofs, 16 * n_faces), "hipMalloc")) ||
      (rc = H(hipMalloc(&d_stk, sb), "hipMalloc")) ||
      (rc = H(hipMemcpyAsync(d_faces, faces, fb, hipMemcpyHostToDevice, st), "upload"))) {
    release();
    return rc == VR_E_HIP ? VR_E_NOMEM : rc;
  }
  const size_t vnb = sizeof(float) * kMaxPts * 3 * (size_t)n_faces;
  if (vnorm && ((rc = H(hipMalloc(&d_vn, vnb), "hipMalloc")) ||
                (rc = H(hipMemcpyAsync(d_vn, vnorm, vnb, hipMemcpyHostToDevice, st), "upload")))) {
    release();
    return rc;
  }
  const dim3 g((unsigned)((n_faces + 63) / 64));
  hipLaunchKernelGGL(k_make_patches<false>, g, dim3(64), 0, st, (const vr_face *)d_faces, n_faces,
                     max_chop, min_chop, (int32_t *)d_base, (int64_t *)d_cnt,
                     (const int64_t *)nullptr, (const int64_t *)nullptr, (vr_patch *)nullptr,
                     (Pending *)d_stk, (const float *)d_vn);
  std::vector<int32_t> hb((size_t)n_faces);
  std::vector<int64_t> hc((size_t)n_faces);
  if ((rc = H(hipGetLastError(), "count launch")) ||
      (rc = H(hipMemcpyAsync(hb.data(), d_base, 4 * n_faces, hipMemcpyDeviceToHost, st), "download")) ||
      (rc = H(hipMemcpyAsync(hc.data(), d_cnt, 8 * n_faces, hipMemcpyDeviceToHost, st), "download")) ||
      (rc = H(hipStreamSynchronize(st), "sync"))) {
    release();
    return rc;
  }
  // offsets: base patches in face order, then children in face order
  std::vector<int64_t> ofs(2 * (size_t)n_faces);
  int64_t nb = 0, nc = 0;
  for (int64_t f = 0; f < n_faces; ++f) {
    if (hc[(size_t)f] < 0) {
      release();
      return fail(VR_E_CAPACITY, "make patches: a winding or the subdivision stack overflows");
    }
    ofs[(size_t)f] = nb;
    nb += hb[(size_t)f];
  }
  for (int64_t f = 0; f < n_faces; ++f) {
    ofs[(size_t)(n_faces + f)] = nb + nc;
    nc += hc[(size_t)f];
  }
  const int64_t total = nb + nc;
  *n_out = total;
  if (!out) {
    release();
    return VR_OK;
  }
  if (cap < total) {
    release();
    return fail(VR_E_CAPACITY, "make patches: output holds " + std::to_string(cap) + " of " +
                                   std::to_string(total) + " patches");
  }
  const size_t ob = sizeof(vr_patch) * (size_t)std::max<int64_t>(total, 1);
  if ((rc = H(hipMalloc(&d_out, ob), "hipMalloc")) ||
      (rc = H(hipMemcpyAsync(d_ofs, ofs.data(), 16 * n_faces, hipMemcpyHostToDevice, st), "upload"))) {
    release();
    return rc;
  }
  hipLaunchKernelGGL(k_make_patches<true>, g, dim3(64), 0, st, (const vr_face *)d_faces, n_faces,
                     max_chop, min_chop, (int32_t *)nullptr, (int64_t *)nullptr,
                     (const int64_t *)d_ofs, (const int64_t *)d_ofs + n_faces, (vr_patch *)d_out,
                     (Pending *)d_stk, (const float *)d_vn);
  if ((rc = H(hipGetLastError(), "fill launch")) ||
      (rc = H(hipMemcpyAsync(out, d_out, sizeof(vr_patch) * (size_t)total, hipMemcpyDeviceToHost, st),
              "download")) ||
      (rc = H(hipStreamSynchronize(st), "sync"))) {
    release();
    return rc;
  }
  release();
  return VR_OK;
}

}  // namespace vr
