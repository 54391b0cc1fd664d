// Device-side kd-tree traversal and triangle test shared by the visibility,
// transfer and direct-light kernels.  gfx950, wave64.
//
// Built with -ffp-contract=off: the triangle test below rounds once per
// operation in the same association as the CPU restatement, so hit index and
// hit t are bit-identical to it.  Traversal arithmetic only culls, with a
// conservative slab margin, so its rounding never changes a result.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace vr {

struct DevScene {
  const uint2 *nodes;     // x = children ((index << 2) | type), y = split bits
  const float4 *refs;     // 3 float4 per leaf reference (leaf order)
  const float4 *planes;   // per triangle (nx, ny, nz, d), original order
  float bmin[3], bmax[3];
  float M;                // max(1, max |scene coordinate|)
  int has_sky;
  int empty;              // no triangles
  // a second tree over the TRACE_ID_SKY triangles alone (when has_sky)
  const uint2 *sky_nodes;
  const float4 *sky_refs;
  float sky_bmin[3], sky_bmax[3];
};

constexpr float kTiny = 1e-20f;

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Remap workgroup b of nb so that XCD x takes one
// contiguous range of work chunks: neighbouring rays (same origin patch in a
// MakeTransfers batch) then share one L2.  A bijection on [0, nb).
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, k = b >> 3, per = nb >> 3, rem = nb & 7u;
  return x * per + min(x, rem) + k;
}
// The same with the work dealt to the XCDs in groups of g consecutive
// chunks, round-robin (group q to XCD q mod 8), so that an XCD whose
// contiguous range would hold more work than the others' does not set the
// launch's end; chunks of a group stay on one XCD.  g = 0 is xcd_chunk.  The
// blocks past the last whole round of 8 g take xcd_chunk over the rest.
// A bijection on [0, nb).  Groups of 256 chunks (16 K rays) took the
// 8x8-room level's TestLine batch from 0.766 to 0.638 ms and its PVS
// MakeTransfers from 0.615 to 0.530 s against one range per XCD.
constexpr uint32_t kXcdGroup = 256;
__device__ __forceinline__ uint32_t xcd_group_chunk(uint32_t b, uint32_t nb, uint32_t g) {
  if (g == 0) return xcd_chunk(b, nb);
  const uint32_t full = nb / (8u * g) * (8u * g);
  if (b >= full) return full + xcd_chunk(b - full, nb - full);
  const uint32_t x = b & 7u, k = b >> 3;
  return ((k / g) * 8u + x) * g + k % g;
}
constexpr uint32_t kTriMask = (1u << 28) - 1u;

__device__ __forceinline__ float sel3(int c, float a, float b, float d) {
  return c == 0 ? a : (c == 1 ? b : d);
}

// Projected-barycentric hit test (frozen semantics of Trace4Rays, which is a
// stub at raytracer/environment.go:140-145): accept t in [tmin, tbest),
// equal t resolved to the lower triangle index.  SKIPSKY ignores
// TRACE_ID_SKY triangles.
template <bool SKIPSKY = false>
__device__ __forceinline__ void tri_test(const float4 *__restrict__ refs, uint32_t k, float ox,
                                         float oy, float oz, float dx, float dy, float dz,
                                         float tmin, int32_t skip, float &tbest, int32_t &best,
                                         int32_t &best_id, uint32_t &best_ref) {
  const float4 q0 = refs[3 * (size_t)k + 0];
  const float4 q1 = refs[3 * (size_t)k + 1];
  const float4 q2 = refs[3 * (size_t)k + 2];
  const int32_t id = __float_as_int(q2.z);
  if (id == skip) return;
  if (SKIPSKY && (id & 0x01000000)) return;
  const uint32_t meta = __float_as_uint(q2.w);
  const int32_t tri = (int32_t)(meta & kTriMask);
  const float den = (q0.x * dx + q0.y * dy) + q0.z * dz;
  if (!(den != 0.0f)) return;
  const float num = q0.w - ((q0.x * ox + q0.y * oy) + q0.z * oz);
  const float t = num / den;
  if (!(t >= tmin && (t < tbest || (t == tbest && tri < best)))) return;
  const int c0 = (int)((meta >> 28) & 3u), c1 = (int)(meta >> 30);
  const float h0 = sel3(c0, ox, oy, oz) + t * sel3(c0, dx, dy, dz);
  const float h1 = sel3(c1, ox, oy, oz) + t * sel3(c1, dx, dy, dz);
  const float b0 = (q1.x * h0 + q1.y * h1) + q1.z;
  const float b1 = (q1.w * h0 + q2.x * h1) + q2.y;
  if (b0 >= 0.0f && b1 >= 0.0f && (b0 + b1) <= 1.0f) {
    tbest = t;
    best = tri;
    best_id = id;
    best_ref = k;
  }
}
constexpr uint32_t kNoRef = 0xffffffffu;

// Per-lane traversal stack entries are 8-byte records: the node index in
// the low 24 bits, then the entry's interval [a, b] as two 20-bit floats
// (sign, exponent, the top 11 mantissa bits), a rounded toward -inf and b
// toward +inf.  A popped interval can only be wider than the pushed one, so
// a walk may visit a node more, never fewer, and its result is unchanged.
// Trees hold fewer than 2^24 nodes (the scene build checks).
constexpr uint32_t kStackNodeBits = 24;
__device__ __forceinline__ uint32_t f20_down(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u >> 12) + ((u >> 31) & ((u & 0xfffu) != 0u));  // negative: magnitude up
}
__device__ __forceinline__ uint32_t f20_up(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u >> 12) + ((~u >> 31) & ((u & 0xfffu) != 0u));  // positive: magnitude up
}
__device__ __forceinline__ uint2 stack_pack(uint32_t n, float a, float b) {
  const uint32_t ka = f20_down(a), kb = f20_up(b);
  return make_uint2(n | (ka << kStackNodeBits), (ka >> 8) | (kb << 12));
}
__device__ __forceinline__ void stack_unpack(uint2 w, uint32_t &n, float &a, float &b) {
  n = w.x & ((1u << kStackNodeBits) - 1u);
  a = __uint_as_float(((w.x >> kStackNodeBits) | ((w.y & 0xfffu) << 8)) << 12);
  b = __uint_as_float((w.y >> 12) << 12);
}

// The first S entries of a lane's stack sit in LDS (entry e of lane l at
// [e * 64 + l]: one ds_write_b64 / ds_read_b64 per push / pop, conflict
// free), deeper entries in private scratch.  S = D keeps the whole stack in
// LDS.  The deeper entries live in a separate object (StackSpill) the stack
// points to: an array indexed at run time sits in private memory, and keeping
// it out of this struct lets the LDS pointer stay in a register.
template <int S, int D>
struct StackSpill {
  uint64_t v[D - S > 0 ? D - S : 1];
};
#define VR_PRIVATE __attribute__((address_space(5)))
template <int S, int D>
struct HybridStack {
  uint2 *e;  // this lane's entry 0 in LDS; entry k at e + 64 k
  VR_PRIVATE uint64_t *pv;  // private-memory pointer: scratch (not flat) accesses
  __device__ HybridStack(uint2 *lds, int lane, StackSpill<S, D> &sp)
      : e(lds + lane), pv((VR_PRIVATE uint64_t *)sp.v) {}
  __device__ void push(int sp, uint32_t n, float a, float b) {
    const uint2 w = stack_pack(n, a, b);
    if (S == D || sp < S)
      e[64 * sp] = w;
    else
      pv[sp - S] = (uint64_t)w.x | ((uint64_t)w.y << 32);
  }
  __device__ void get(int sp, uint32_t &n, float &a, float &b) const {
    uint2 w;
    if (S == D || sp < S) {
      w = e[64 * sp];
    } else {
      const uint64_t v = pv[sp - S];
      w = make_uint2((uint32_t)v, (uint32_t)(v >> 32));
    }
    stack_unpack(w, n, a, b);
  }
};

template <int D>
struct ScratchStack {
  uint32_t node[D];
  float a[D], b[D];
  __device__ void push(int sp, uint32_t n, float x, float y) {
    node[sp] = n;
    a[sp] = x;
    b[sp] = y;
  }
  __device__ void get(int sp, uint32_t &n, float &x, float &y) const {
    n = node[sp];
    x = a[sp];
    y = b[sp];
  }
};

// Interval bounds in the walk are compared as int32 bit patterns: one
// v_min_i32 / v_max_i32, with no NaN quieting of loop-carried values.  For
// two non-negative floats, or one of each sign, this is the float min / max;
// for two negatives (only when a caller's tmin < 0) "min" returns the larger
// and "max" the smaller, so an upper bound (min) only grows and a lower
// bound (max) only shrinks: intervals widen, the walk visits more, and the
// result is unchanged.
__device__ __forceinline__ float tmin_wide(float a, float b) {
  return __int_as_float(min(__float_as_int(a), __float_as_int(b)));
}
__device__ __forceinline__ float tmax_wide(float a, float b) {
  return __int_as_float(max(__float_as_int(a), __float_as_int(b)));
}

// Walks one kd-tree (nodes/refs/bounds) for hits with tmin <= t < tmax that
// come before (tbest, best) in (t, triangle index) order, updating them; the
// result is the first of the hits in that order, bit for bit the brute-force
// result.  ANY: return at the first accepted hit.  SKIPSKY: ignore
// TRACE_ID_SKY triangles.  POSTPONE: lanes hold their next leaf and walk on
// until every lane of the wave holds one (below).  Returns whether a hit was accepted (best_ref =
// its leaf-reference index).  D = stack capacity; the host guarantees tree
// depth < D.
// What a walk needs of its ray besides origin and direction, the same for
// every tree of a scene (the margin dl depends on the scene's extent M):
// the reciprocal direction, and the slab terms of a split plane s on axis a,
//   t = ((s +- dl) - o_a) i_a = s i_a + (+-dl - o_a) i_a,
// the near child's exit with +dl toward the ray (copysign(dl, i_a)), the far
// child's entry with -dl; the terms after s i_a are fixed per ray, and their
// rounding (and the fma's) stays far inside the dl margin.  A segment that
// walks the sky tree and then the full tree computes this once.
struct RayPre {
  float ix, iy, iz, dl, ex, ey, ez, fx, fy, fz;
};
__device__ __forceinline__ RayPre ray_pre(float M, float ox, float oy, float oz, float dx,
                                          float dy, float dz) {
  RayPre r;
  const float om = fmaxf(fabsf(ox), fmaxf(fabsf(oy), fabsf(oz)));
  r.dl = (M + om) * 0x1p-15f;
  r.ix = 1.0f / (fabsf(dx) < kTiny ? copysignf(kTiny, dx) : dx);
  r.iy = 1.0f / (fabsf(dy) < kTiny ? copysignf(kTiny, dy) : dy);
  r.iz = 1.0f / (fabsf(dz) < kTiny ? copysignf(kTiny, dz) : dz);
  r.ex = (copysignf(r.dl, r.ix) - ox) * r.ix;
  r.fx = (-copysignf(r.dl, r.ix) - ox) * r.ix;
  r.ey = (copysignf(r.dl, r.iy) - oy) * r.iy;
  r.fy = (-copysignf(r.dl, r.iy) - oy) * r.iy;
  r.ez = (copysignf(r.dl, r.iz) - oz) * r.iz;
  r.fz = (-copysignf(r.dl, r.iz) - oz) * r.iz;
  return r;
}

template <int D, bool ANY, bool SKIPSKY, bool POSTPONE, class Stack>
__device__ bool walk_tree(const uint2 *__restrict__ nodes, const float4 *__restrict__ refs,
                          const float *bmin, const float *bmax, const RayPre &rp, Stack &st,
                          float ox, float oy, float oz, float dx, float dy, float dz,
                          float tmin, float tmax, int32_t skip, float &tbest, int32_t &best,
                          int32_t &best_id, uint32_t &best_ref) {
  const int32_t best0 = best;
  const float dl = rp.dl, ix = rp.ix, iy = rp.iy, iz = rp.iz;
  float ta = tmin, tb = tmax;
  {
    float t1 = ((bmin[0] - dl) - ox) * ix, t2 = ((bmax[0] + dl) - ox) * ix;
    ta = fmaxf(ta, fminf(t1, t2));
    tb = fminf(tb, fmaxf(t1, t2));
    t1 = ((bmin[1] - dl) - oy) * iy;
    t2 = ((bmax[1] + dl) - oy) * iy;
    ta = fmaxf(ta, fminf(t1, t2));
    tb = fminf(tb, fmaxf(t1, t2));
    t1 = ((bmin[2] - dl) - oz) * iz;
    t2 = ((bmax[2] + dl) - oz) * iz;
    ta = fmaxf(ta, fminf(t1, t2));
    tb = fminf(tb, fmaxf(t1, t2));
  }
  if (!(ta <= tb)) return false;
  // An any-hit walk returns at its first accepted hit, so tbest is fixed
  // while it runs: clamping tb to it here leaves every min(tb, tbest) of
  // the walk unchanged (an integer-order min, associative), and a stacked
  // entry, pushed only when its start is at most that bound, never needs
  // the pop's prune.
  if constexpr (ANY) tb = tmin_wide(tb, tbest);
  const float ex = rp.ex, ey = rp.ey, ez = rp.ez, fx = rp.fx, fy = rp.fy, fz = rp.fz;
  int sp = 0;
  uint32_t node = 0;
  // next stacked node whose interval can still hold a hit before tbest
  auto pop = [&]() -> bool {
    while (sp > 0) {
      --sp;
      float a, b;
      uint32_t n;
      st.get(sp, n, a, b);
      if (!ANY && a > tbest) continue;
      node = n;
      ta = a;
      tb = b;
      return true;
    }
    return false;
  };
  if constexpr (POSTPONE) {
    // "while-while" with postponed leaves: a lane descends through inner
    // nodes and empty leaves to its next leaf holding triangles and keeps it;
    // until every lane of the wave holds one, a lane that already does walks
    // on to its next such leaf and stops there.  Then the wave tests the held
    // leaves together.  Visits differ from a depth-first walk only in order
    // and, for a closest hit, in leaves tested before a hit found later would
    // have pruned them; the result is the least hit in (t, triangle) order
    // either way.  A step assigns each loop value once, by selects, and the
    // push is a predicated store.  Closest hits take this walk: their long
    // walks through many leaves keep more lanes busy (4 M hemisphere rays on
    // an 8x8-room level: 1.59 against 1.87 ms).  Any-hit walks, which mostly
    // end at their first leaf, do not (0.73 against 0.64 ms on 4 M segments).
    uint32_t lstart = 0, lcnt = 0;  // the held leaf's triangles (none: lcnt 0)
    bool live = true;               // node, ta, tb: the next node to visit
    for (;;) {
      uint2 nd = nodes[node];
      for (;;) {
        if ((nd.x & 3u) == 3u) {
          const uint32_t c = (uint32_t)__uint_as_float(nd.y);
          if (c != 0u) {
            if (lcnt != 0u) break;  // a second leaf: visit it after the held one
            lstart = nd.x >> 2;
            lcnt = c;
          }
          if (!pop()) {
            live = false;
            break;
          }
        } else {
          const uint32_t type = nd.x & 3u;
          const float sv = __uint_as_float(nd.y);
          const float i_a = sel3((int)type, ix, iy, iz);
          const float near_end = __builtin_fmaf(sv, i_a, sel3((int)type, ex, ey, ez));
          const float far_start = __builtin_fmaf(sv, i_a, sel3((int)type, fx, fy, fz));
          const uint32_t left = nd.x >> 2;
          const bool fwd = i_a >= 0.0f;
          const uint32_t near_n = fwd ? left : left + 1, far_n = fwd ? left + 1 : left;
          const float tbe = ANY ? tb : tmin_wide(tb, tbest);
          const bool near_ok = ta <= tmin_wide(near_end, tbe);
          const float fs = tmax_wide(far_start, ta);
          const bool far_ok = fs <= tbe;
          // the stack holds at most one entry per level of the path to the
          // current node, and the host checks tree depth < D
          if (near_ok && far_ok) {
            st.push(sp, far_n, fs, tb);
            ++sp;
          }
          node = near_ok ? near_n : far_n;
          const float tb_near = tmin_wide(tb, near_end);
          ta = near_ok ? ta : fs;
          tb = near_ok ? tb_near : tb;
          if (!near_ok && !far_ok && !pop()) {
            live = false;
            break;
          }
        }
        nd = nodes[node];
        if (__ballot(lcnt == 0u) == 0) break;  // every walking lane holds a leaf
      }
      for (uint32_t k = lstart; k < lstart + lcnt; ++k)
        tri_test<SKIPSKY>(refs, k, ox, oy, oz, dx, dy, dz, tmin, skip, tbest, best, best_id,
                          best_ref);
      if (ANY && best != best0) return true;
      lcnt = 0u;
      if (!live) break;
    }
  } else {
    // "while-while": a lane descends through inner nodes to its next leaf
    // before the wave tests leaves, so lanes test triangles together.  Each
    // lane visits the same nodes in the same order as one step per iteration.
    // A step assigns each loop value once, by selects, and the push is a
    // predicated store.  A step where neither child can hold a hit leaves the
    // descent as an empty leaf: the pop that follows every leaf continues the
    // walk, as a pop right there would.
    constexpr uint2 kEmptyLeaf = {3u, 0u};  // leaf, no triangles
    for (;;) {
      uint2 nd = nodes[node];
      while ((nd.x & 3u) != 3u) {
        const uint32_t type = nd.x & 3u;
        const float sv = __uint_as_float(nd.y);
        const float i_a = sel3((int)type, ix, iy, iz);
        const float near_end = __builtin_fmaf(sv, i_a, sel3((int)type, ex, ey, ez));
        const float far_start = __builtin_fmaf(sv, i_a, sel3((int)type, fx, fy, fz));
        const uint32_t left = nd.x >> 2;
        const bool fwd = i_a >= 0.0f;
        const uint32_t near_n = fwd ? left : left + 1, far_n = fwd ? left + 1 : left;
        const float tbe = ANY ? tb : tmin_wide(tb, tbest);
        const bool near_ok = ta <= tmin_wide(near_end, tbe);
        const float fs = tmax_wide(far_start, ta);
        const bool far_ok = fs <= tbe;
        // the stack holds at most one entry per level of the path to the
        // current node, and the host checks tree depth < D
        if (near_ok && far_ok) {
          st.push(sp, far_n, fs, tb);
          ++sp;
        }
        node = near_ok ? near_n : far_n;
        const float tb_near = tmin_wide(tb, near_end);
        ta = near_ok ? ta : fs;
        tb = near_ok ? tb_near : tb;
        if (!near_ok && !far_ok) {
          nd = kEmptyLeaf;
          break;
        }
        nd = nodes[node];
      }
      const uint32_t start = nd.x >> 2;
      const uint32_t cnt = (uint32_t)__uint_as_float(nd.y);
      for (uint32_t k = start; k < start + cnt; ++k)
        tri_test<SKIPSKY>(refs, k, ox, oy, oz, dx, dy, dz, tmin, skip, tbest, best, best_id,
                          best_ref);
      if (ANY && best != best0) return true;
      if (!pop()) break;
    }
  }
  return best != best0;
}

// Closest hit in [tmin, tmax) (ANY: stop at the first hit).
template <int D, bool ANY, class Stack>
__device__ void trace_ray(const DevScene &sc, Stack &st, float ox, float oy, float oz, float dx,
                          float dy, float dz, float tmin, float tmax, int32_t skip, float &t_out,
                          int32_t &best_out, int32_t &id_out) {
  float tbest = tmax;
  int32_t best = -1, best_id = 0;
  t_out = 0.0f;
  best_out = -1;
  id_out = 0;
  if (sc.empty) return;
  uint32_t ref = kNoRef;
  const RayPre rp = ray_pre(sc.M, ox, oy, oz, dx, dy, dz);
  if (walk_tree<D, ANY, false, !ANY>(sc.nodes, sc.refs, sc.bmin, sc.bmax, rp, st, ox, oy, oz, dx,
                                     dy, dz, tmin, tmax, skip, tbest, best, best_id, ref)) {
    t_out = tbest;
    best_out = best;
    id_out = best_id;
  }
}

// TestLineDoesHitSky for one segment (raytracer/trace/testline.go:22-51):
// dir = delta * (1/len), trace [0, len), occluded iff the closest hit is not
// a TRACE_ID_SKY triangle.  Returns visible; *sky = closest hit is sky.
// With sky triangles present this is evaluated in two walks with the same
// result: the closest sky hit (ts, is) on the sky-only tree, then the full
// tree stops at the first non-sky hit before (ts, is) in (t, index) order;
// the segment is occluded iff there is one.
//
// occluder (in/out): a leaf reference of the full tree to test first (kNoRef
// for none), typically the triangle that blocked the caller's previous,
// nearby segment.  It is tested with the same arithmetic and bound as the
// walk, so a hit there is exactly a hit the walk would accept and the result
// does not depend on it.  On return it holds the blocking triangle found, if
// any, else it is unchanged.
template <int D, class Stack>
__device__ __forceinline__ bool test_line_cached(const DevScene &sc, Stack &st, float sx,
                                                 float sy, float sz, float ex, float ey, float ez,
                                                 int32_t skip, bool &sky, uint32_t &occluder) {
  const float vx = ex - sx, vy = ey - sy, vz = ez - sz;
  const float len2 = (vx * vx + vy * vy) + vz * vz;
  const float len = __builtin_sqrtf(len2);
  const float inv = 1.0f / len;
  const float dx = vx * inv, dy = vy * inv, dz = vz * inv;
  sky = false;
  if (sc.empty) return true;
  float t = len;
  int32_t best = -1, id = 0;
  uint32_t ref = kNoRef;
  const RayPre rp = ray_pre(sc.M, sx, sy, sz, dx, dy, dz);
  if (sc.has_sky)
    walk_tree<D, false, false, false>(sc.sky_nodes, sc.sky_refs, sc.sky_bmin, sc.sky_bmax, rp, st,
                                      sx, sy, sz, dx, dy, dz, 0.0f, len, skip, t, best, id, ref);
  const int32_t sky_best = best;
  bool occluded = false;
  if (occluder != kNoRef) {
    tri_test<true>(sc.refs, occluder, sx, sy, sz, dx, dy, dz, 0.0f, skip, t, best, id, ref);
    occluded = best != sky_best;
  }
  if (!occluded) {
    ref = kNoRef;
    occluded = walk_tree<D, true, true, false>(sc.nodes, sc.refs, sc.bmin, sc.bmax, rp, st, sx,
                                               sy, sz, dx, dy, dz, 0.0f, len, skip, t, best, id,
                                               ref);
    if (occluded) occluder = ref;
  }
  sky = !occluded && sky_best >= 0;
  return !occluded;
}
// TestLineDoesHitSky without an occluder hint, with the walks in the order
// that suits a visible segment: any opaque hit on [0, len) first.  Without
// one the segment is visible, and its closest hit is sky iff the sky tree
// has any hit on [0, len).  With one, the closest sky hit before it in
// (t, index) order (none: occluded), then any opaque hit before that sky
// hit (test_line_cached's second walk) decide; a sky hit found that way is
// the closest of all, since every sky hit it passed over lies after an
// opaque one.  Same results as test_line_cached.
template <int D, class Stack>
__device__ __forceinline__ bool test_line(const DevScene &sc, Stack &st, float sx, float sy,
                                          float sz, float ex, float ey, float ez, int32_t skip,
                                          bool &sky) {
  const float vx = ex - sx, vy = ey - sy, vz = ez - sz;
  const float len2 = (vx * vx + vy * vy) + vz * vz;
  const float len = __builtin_sqrtf(len2);
  const float inv = 1.0f / len;
  const float dx = vx * inv, dy = vy * inv, dz = vz * inv;
  sky = false;
  if (sc.empty) return true;
  const RayPre rp = ray_pre(sc.M, sx, sy, sz, dx, dy, dz);
  float t = len;
  int32_t best = -1, id = 0;
  uint32_t ref = kNoRef;
  if (!walk_tree<D, true, true, false>(sc.nodes, sc.refs, sc.bmin, sc.bmax, rp, st, sx, sy, sz,
                                       dx, dy, dz, 0.0f, len, skip, t, best, id, ref)) {
    if (sc.has_sky)
      sky = walk_tree<D, true, false, false>(sc.sky_nodes, sc.sky_refs, sc.sky_bmin, sc.sky_bmax,
                                             rp, st, sx, sy, sz, dx, dy, dz, 0.0f, len, skip, t,
                                             best, id, ref);
    return true;
  }
  if (!sc.has_sky) return false;
  const int32_t first = best;
  walk_tree<D, false, false, false>(sc.sky_nodes, sc.sky_refs, sc.sky_bmin, sc.sky_bmax, rp, st,
                                    sx, sy, sz, dx, dy, dz, 0.0f, len, skip, t, best, id, ref);
  if (best == first) return false;
  sky = !walk_tree<D, true, true, false>(sc.nodes, sc.refs, sc.bmin, sc.bmax, rp, st, sx, sy, sz,
                                         dx, dy, dz, 0.0f, len, skip, t, best, id, ref);
  return sky;
}

// test_line's visibility without its sky output, with the walks in the
// order that suits a visible segment: first any non-sky hit on [0, len)
// (none: visible, and the sky tree is never walked); then the closest sky
// hit before that one in (t, index) order (none: the closest hit is not
// sky, occluded); then, as test_line does, any non-sky hit before that sky
// hit.  The three steps decide what test_line decides: a sky hit found in
// the second is the closest of all, since every sky hit it passed over lies
// after a non-sky hit.
template <int D, class Stack>
__device__ __forceinline__ bool segment_visible(const DevScene &sc, Stack &st, float sx, float sy,
                                                float sz, float ex, float ey, float ez,
                                                int32_t skip) {
  const float vx = ex - sx, vy = ey - sy, vz = ez - sz;
  const float len2 = (vx * vx + vy * vy) + vz * vz;
  const float len = __builtin_sqrtf(len2);
  const float inv = 1.0f / len;
  const float dx = vx * inv, dy = vy * inv, dz = vz * inv;
  if (sc.empty) return true;
  const RayPre rp = ray_pre(sc.M, sx, sy, sz, dx, dy, dz);
  float t = len;
  int32_t best = -1, id = 0;
  uint32_t ref = kNoRef;
  if (!walk_tree<D, true, true, false>(sc.nodes, sc.refs, sc.bmin, sc.bmax, rp, st, sx, sy, sz,
                                       dx, dy, dz, 0.0f, len, skip, t, best, id, ref))
    return true;
  if (!sc.has_sky) return false;
  const int32_t first = best;
  walk_tree<D, false, false, false>(sc.sky_nodes, sc.sky_refs, sc.sky_bmin, sc.sky_bmax, rp, st,
                                    sx, sy, sz, dx, dy, dz, 0.0f, len, skip, t, best, id, ref);
  if (best == first) return false;
  return !walk_tree<D, true, true, false>(sc.nodes, sc.refs, sc.bmin, sc.bmax, rp, st, sx, sy, sz,
                                          dx, dy, dz, 0.0f, len, skip, t, best, id, ref);
}

// The first half of test_line_cached alone: whether the segment is occluded
// by `occluder` (a full-tree leaf reference) as test_line would decide it
// (the closest sky hit first, then the same hit test and bound).  False
// means only that this triangle does not decide it.
template <int D, class Stack>
__device__ __forceinline__ bool occluded_by(const DevScene &sc, Stack &st, float sx, float sy,
                                            float sz, float ex, float ey, float ez, int32_t skip,
                                            uint32_t occluder) {
  const float vx = ex - sx, vy = ey - sy, vz = ez - sz;
  const float len2 = (vx * vx + vy * vy) + vz * vz;
  const float len = __builtin_sqrtf(len2);
  const float inv = 1.0f / len;
  const float dx = vx * inv, dy = vy * inv, dz = vz * inv;
  if (sc.empty) return false;
  float t = len;
  int32_t best = -1, id = 0;
  uint32_t ref = kNoRef;
  if (sc.has_sky)
    walk_tree<D, false, false, false>(sc.sky_nodes, sc.sky_refs, sc.sky_bmin, sc.sky_bmax,
                                      ray_pre(sc.M, sx, sy, sz, dx, dy, dz), st, sx, sy, sz, dx,
                                      dy, dz, 0.0f, len, skip, t, best, id, ref);
  const int32_t sky_best = best;
  tri_test<true>(sc.refs, occluder, sx, sy, sz, dx, dy, dz, 0.0f, skip, t, best, id, ref);
  return best != sky_best;
}

}  // namespace vr
