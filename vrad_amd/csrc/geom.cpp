// Host geometry of the radiosity-bake path: triangle intersection format and
// the SAH kd-tree, both computed once per scene before upload to HBM.
//
// Compiled with -ffp-contract=off and SSE scalar FP: every FP32 expression is
// written left-associated and rounds once per operation.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <future>
#include <thread>
#include <unordered_map>

#include "vr_internal.h"

namespace vr {

int host_threads(int requested) {
  if (requested > 0) return requested;
  unsigned hc = std::thread::hardware_concurrency();
  return hc ? (int)hc : 1;
}

template <class F>
static void parallel_ranges(int64_t n, int n_threads, F fn) {
  if (n_threads <= 1 || n < 4096) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + n_threads - 1) / n_threads;
  for (int t = 0; t < n_threads; ++t) {
    int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { fn(b, e); });
  }
  for (auto &x : th) x.join();
}

void parallel_for(int64_t n, int n_threads, const std::function<void(int64_t, int64_t)> &fn) {
  parallel_ranges(n, n_threads, [&](int64_t b, int64_t e) { fn(b, e); });
}

// GetEdgeEquation, vmath/polygon/edge.go:5-25
static void edge_equation(const float p1[3], const float p2[3], int c0, int c1,
                          const float in[3], float *out) {
  float nx = p1[c1] - p2[c1];
  float ny = p2[c0] - p1[c0];
  float d = -(nx * p1[c0] + ny * p1[c1]);
  float tr = (in[c0] * nx + in[c1] * ny) + d;
  if (tr < 0) {
    nx = -nx;
    ny = -ny;
    d = -d;
    tr = -tr;
  }
  out[0] = nx / tr;
  out[1] = ny / tr;
  out[2] = d / tr;
}

// ChangeIntoIntersectionFormat, raytracer/cache/optimisedtriangle.go:30-78.
// Vector ops follow mgl32 (Sub, Cross, Normalize = v * (1/Len), Dot).
static void precompute_one(const float *v, int32_t id, uint8_t flags, TriIntersect *t) {
  float p1[3] = {v[0], v[1], v[2]}, p2[3] = {v[3], v[4], v[5]}, p3[3] = {v[6], v[7], v[8]};
  float e1[3], e2[3], N[3];
  for (int c = 0; c < 3; ++c) {
    e1[c] = p2[c] - p1[c];
    e2[c] = p3[c] - p1[c];
  }
  N[0] = e1[1] * e2[2] - e1[2] * e2[1];
  N[1] = e1[2] * e2[0] - e1[0] * e2[2];
  N[2] = e1[0] * e2[1] - e1[1] * e2[0];
  float len2 = (N[0] * N[0] + N[1] * N[1]) + N[2] * N[2];
  float len = (float)std::sqrt((double)len2);
  float l = 1.0f / len;
  for (int c = 0; c < 3; ++c) N[c] = N[c] * l;
  int drop = 0;
  for (int c = 1; c < 3; ++c)
    if (std::fabs((double)N[c]) > std::fabs((double)N[drop])) drop = c;
  std::memset(t, 0, sizeof *t);
  t->nx = N[0];
  t->ny = N[1];
  t->nz = N[2];
  t->d = (N[0] * p1[0] + N[1] * p1[1]) + N[2] * p1[2];
  int c0 = (drop + 1) % 3, c1 = (drop + 2) % 3;
  t->c0 = (uint8_t)c0;
  t->c1 = (uint8_t)c1;
  edge_equation(p1, p2, c0, c1, p3, t->e);
  edge_equation(p2, p3, c0, c1, p1, t->e + 3);
  t->id = id;
  t->flags = flags;
}

void precompute_triangles(const float *v9, const int32_t *ids, const uint8_t *flags,
                          int64_t n, TriIntersect *out, int n_threads) {
  parallel_ranges(n, n_threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i)
      precompute_one(v9 + 9 * i, ids ? ids[i] : 0, flags ? flags[i] : 0, out + i);
  });
}

// ---------------------------------------------------------------------------
// SAH kd-tree.  Follows the intent of RefineNode / CalculateCostsOfSplit
// (raytracer/environment.go:181-387): candidate splits at the node midpoint
// and at the vertices of every (1 + n/10)-th triangle on all three axes;
// cost 75 + 167 * (n_both + SA_L/SA * n_left + SA_R/SA * n_right); a
// one-sided split moves to the extreme vertex coordinate; leaf when n < 3,
// depth > 21 or 167 * n <= best cost; depth += 100 below 20 triangles with an
// empty side.  Defects of the Go code (bounds never written, Vertex() reading
// one coordinate, midpoint precedence, counts passed by value, partition
// off-by-one) are repaired as the comments there intend.
// ---------------------------------------------------------------------------
namespace {

enum { kPos = 1, kNeg = -1, kStrad = 0 };
constexpr int kMaxTreeDepth = 21;              // raytracer/kdtree/constants.go:28
constexpr float kCostTraversal = 75.0f;        // :25
constexpr float kCostIntersection = 167.0f;    // :26
constexpr int64_t kParallelNode = 32768;       // spawn / parallel-search threshold

struct Sub {
  std::vector<uint32_t> ch;
  std::vector<float> sp;
  std::vector<int32_t> idx;
  int max_depth = 0;
};

struct Builder {
  const float *v9;
  std::vector<float> lo, hi;  // per triangle per axis vertex extents
  int n_threads;
  std::atomic<int> busy{1};

  int classify(int32_t t, int axis, float sv) const {
    float mn = lo[3 * (size_t)t + axis], mx = hi[3 * (size_t)t + axis];
    if (mn >= sv) return kPos;
    if (mx <= sv) return kNeg;
    if (mn == mx) return kPos;
    return kStrad;
  }

  static float box_sa(const float mn[3], const float mx[3]) {
    float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return 2.0f * ((d0 * d2 + d0 * d1) + d1 * d2);
  }

  struct Cand {
    int axis;
    float trial;
  };
  struct Eval {
    float cost, moved;
    int64_t nl, nr, nb;
  };

  Eval evaluate(const Cand &c, const int32_t *list, int64_t n, const float mn[3],
                const float mx[3]) const {
    int64_t l = 0, r = 0, b = 0;
    double minc = 1.0e23, maxc = -1.0e23;
    const int a = c.axis;
    const float sv = c.trial;
    for (int64_t i = 0; i < n; ++i) {
      size_t t = (size_t)list[i];
      float mnv = lo[3 * t + a], mxv = hi[3 * t + a];
      if ((double)mnv < minc) minc = mnv;
      if ((double)mxv > maxc) maxc = mxv;
      if (mnv >= sv) ++r;
      else if (mxv <= sv) ++l;
      else if (mnv == mxv) ++r;
      else ++b;
    }
    float moved = sv;
    if (l != 0 && b == 0 && r == 0) moved = (float)maxc;
    if (r != 0 && b == 0 && l == 0) moved = (float)minc;
    float lmx[3] = {mx[0], mx[1], mx[2]}, rmn[3] = {mn[0], mn[1], mn[2]};
    lmx[a] = moved;
    rmn[a] = moved;
    float sal = box_sa(mn, lmx), sar = box_sa(rmn, mx);
    float isa = 1.0f / box_sa(mn, mx);
    float cost = kCostTraversal +
                 kCostIntersection * (((float)b + (sal * isa) * (float)l) + (sar * isa) * (float)r);
    return Eval{cost, moved, l, r, b};
  }

  static void make_leaf(Sub &S, int64_t node, const int32_t *list, int64_t n, int level) {
    S.ch[node] = 3u + ((uint32_t)S.idx.size() << 2);
    S.sp[node] = (float)n;
    S.idx.insert(S.idx.end(), list, list + n);
    S.max_depth = std::max(S.max_depth, level);
  }

  // Append subtree C (root at C.ch[0]) under S's node `slot`.
  static void merge(Sub &S, int64_t slot, const Sub &C) {
    const uint32_t base = (uint32_t)S.ch.size(), ibase = (uint32_t)S.idx.size();
    auto remap = [&](uint32_t x) -> uint32_t {
      uint32_t type = x & 3u, v = x >> 2;
      if (type == 3u) return 3u + ((v + ibase) << 2);
      return type + ((base + v - 1u) << 2);
    };
    S.ch[slot] = remap(C.ch[0]);
    S.sp[slot] = C.sp[0];
    for (size_t k = 1; k < C.ch.size(); ++k) {
      S.ch.push_back(remap(C.ch[k]));
      S.sp.push_back(C.sp[k]);
    }
    S.idx.insert(S.idx.end(), C.idx.begin(), C.idx.end());
    S.max_depth = std::max(S.max_depth, C.max_depth);
  }

  Sub build_sub(std::vector<int32_t> list, const float mn[3], const float mx[3], int depth,
                int level) {
    Sub S;
    S.ch.push_back(0);
    S.sp.push_back(0.0f);
    refine(S, 0, list, mn, mx, depth, level);
    return S;
  }

  void refine(Sub &S, int64_t node, std::vector<int32_t> &list, const float mn[3],
              const float mx[3], int depth, int level) {
    const int64_t n = (int64_t)list.size();
    if (n < 3 || depth > kMaxTreeDepth) {
      make_leaf(S, node, list.data(), n, level);
      return;
    }
    std::vector<Cand> cands;
    const int64_t skip = 1 + n / 10;
    for (int axis = 0; axis < 3; ++axis) {
      for (int64_t ts = -1; ts < n; ts += skip) {
        for (int tv = 0; tv < 3; ++tv) {
          float trial;
          if (ts == -1) {
            trial = 0.5f * (mn[axis] + mx[axis]);
          } else {
            trial = v9[9 * (size_t)list[ts] + 3 * tv + axis];
            if (trial > mx[axis] || trial < mn[axis]) continue;
          }
          cands.push_back(Cand{axis, trial});
          if (ts == -1) break;
        }
      }
    }
    std::vector<Eval> ev(cands.size());
    if (n >= kParallelNode && n_threads > 1) {
      int nt = std::min<int>(n_threads, (int)cands.size());
      std::vector<std::thread> th;
      for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
          for (size_t c = (size_t)t; c < cands.size(); c += (size_t)nt)
            ev[c] = evaluate(cands[c], list.data(), n, mn, mx);
        });
      for (auto &x : th) x.join();
    } else {
      for (size_t c = 0; c < cands.size(); ++c) ev[c] = evaluate(cands[c], list.data(), n, mn, mx);
    }
    // first strict improvement in candidate order, as the sequential scan
    float best_cost = 1.0e23f;
    int64_t bi = -1;
    for (size_t c = 0; c < cands.size(); ++c)
      if (ev[c].cost < best_cost) {
        best_cost = ev[c].cost;
        bi = (int64_t)c;
      }
    float no_split = (float)((int64_t)167 * n);
    if (bi < 0 || no_split <= best_cost) {
      make_leaf(S, node, list.data(), n, level);
      return;
    }
    const Cand bc = cands[(size_t)bi];
    const Eval be = ev[(size_t)bi];
    std::vector<int32_t> nl((size_t)n);
    int64_t lo_o = 0, ro = 0, bo = 0;
    for (int64_t i = 0; i < n; ++i) {
      int32_t t = list[i];
      switch (classify(t, bc.axis, bc.trial)) {
        case kNeg: nl[(size_t)lo_o++] = t; break;
        case kPos: ++ro; nl[(size_t)(n - ro)] = t; break;
        default: nl[(size_t)(be.nl + bo)] = t; ++bo; break;
      }
    }
    std::vector<int32_t>().swap(list);
    const uint32_t left = (uint32_t)S.ch.size();
    S.ch.push_back(0);
    S.ch.push_back(0);
    S.sp.push_back(0.0f);
    S.sp.push_back(0.0f);
    S.ch[node] = (uint32_t)bc.axis + (left << 2);
    S.sp[node] = be.moved;
    if (n < 20 && (be.nl == 0 || be.nr == 0)) depth += 100;
    float lmx[3] = {mx[0], mx[1], mx[2]}, rmn[3] = {mn[0], mn[1], mn[2]};
    lmx[bc.axis] = be.moved;
    rmn[bc.axis] = be.moved;
    std::vector<int32_t> L(nl.begin(), nl.begin() + (be.nl + be.nb));
    std::vector<int32_t> R(nl.begin() + be.nl, nl.end());
    std::vector<int32_t>().swap(nl);
    const bool spawn = (int64_t)L.size() >= kParallelNode && (int64_t)R.size() >= kParallelNode &&
                       busy.load() < n_threads;
    if (spawn) {
      busy.fetch_add(1);
      float mn_c[3] = {mn[0], mn[1], mn[2]}, lmx_c[3] = {lmx[0], lmx[1], lmx[2]};
      auto fut = std::async(std::launch::async, [this, &L, mn_c, lmx_c, depth, level]() {
        Sub r = build_sub(std::move(L), mn_c, lmx_c, depth + 1, level + 1);
        busy.fetch_sub(1);
        return r;
      });
      Sub Rs = build_sub(std::move(R), rmn, mx, depth + 1, level + 1);
      Sub Ls = fut.get();
      merge(S, left, Ls);
      merge(S, left + 1, Rs);
    } else {
      refine(S, left, L, mn, lmx, depth + 1, level + 1);
      refine(S, left + 1, R, rmn, mx, depth + 1, level + 1);
    }
  }
};

}  // namespace

int build_kdtree(const float *v9, int64_t n, int n_threads, KdTree *out) {
  Builder b;
  b.v9 = v9;
  b.n_threads = host_threads(n_threads);
  b.lo.resize(3 * (size_t)std::max<int64_t>(n, 1));
  b.hi.resize(3 * (size_t)std::max<int64_t>(n, 1));
  float mn[3] = {1.0e23f, 1.0e23f, 1.0e23f}, mx[3] = {-1.0e23f, -1.0e23f, -1.0e23f};
  for (int64_t t = 0; t < n; ++t)
    for (int c = 0; c < 3; ++c) {
      float a = v9[9 * t + c], bb = v9[9 * t + 3 + c], cc = v9[9 * t + 6 + c];
      float l = std::min(a, std::min(bb, cc)), h = std::max(a, std::max(bb, cc));
      b.lo[3 * t + c] = l;
      b.hi[3 * t + c] = h;
      mn[c] = std::min(mn[c], l);
      mx[c] = std::max(mx[c], h);
    }
  if (n == 0)
    for (int c = 0; c < 3; ++c) mn[c] = mx[c] = 0.0f;
  std::vector<int32_t> root((size_t)n);
  for (int64_t t = 0; t < n; ++t) root[t] = (int32_t)t;
  Sub S = b.build_sub(std::move(root), mn, mx, 0, 0);
  if (S.idx.size() >= ((size_t)1 << 30) || S.ch.size() >= ((size_t)1 << 30))
    return fail(VR_E_CAPACITY, "kd-tree exceeds 2^30 nodes or leaf references");
  out->children = std::move(S.ch);
  out->split = std::move(S.sp);
  out->idx = std::move(S.idx);
  std::memcpy(out->bmin, mn, sizeof mn);
  std::memcpy(out->bmax, mx, sizeof mx);
  out->max_depth = S.max_depth;
  return VR_OK;
}

}  // namespace vr

namespace vr {

// PairEdges (rad/lightmap/lightmap.go:37-216) without displacements or
// smoothing groups (every face's SmoothingGroups is 0): for vertex j of face
// i, the face normals of the other faces that use the same vertex, when
// their cosine with face i's normal is at least the threshold, are added in
// face order to a zero vector, then face i's normal, and the sum is
// normalised (mgl32 Add, Dot, Normalize).  A vertex is identified by its
// position, bit for bit (the BSP's shared vertex index).  The loop over
// faces per vertex takes j below the face's own edge count, as the Go code's
// comment intends (it iterates over the map's edge count).
void pair_edges(const vr_face *faces, int64_t n, float threshold, float *out) {
  struct Key {
    uint32_t x, y, z;
    bool operator==(const Key &o) const { return x == o.x && y == o.y && z == o.z; }
  };
  struct Hash {
    size_t operator()(const Key &k) const {
      uint64_t h = (uint64_t)k.x * 0x9E3779B97F4A7C15ull;
      h ^= (uint64_t)k.y * 0xC2B2AE3D27D4EB4Full + (h >> 29);
      h ^= (uint64_t)k.z * 0x165667B19E3779F9ull + (h >> 31);
      return (size_t)h;
    }
  };
  auto key = [](const float p[3]) {
    Key k;
    std::memcpy(&k.x, &p[0], 4);
    std::memcpy(&k.y, &p[1], 4);
    std::memcpy(&k.z, &p[2], 4);
    return k;
  };
  std::unordered_map<Key, std::vector<int64_t>, Hash> users;
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < faces[i].n_points; ++j) {
      std::vector<int64_t> &u = users[key(faces[i].points[j])];
      if (u.empty() || u.back() != i) u.push_back(i);
    }
  for (int64_t i = 0; i < n; ++i) {
    const float *fn = faces[i].normal;
    float *o = out + (size_t)i * VR_FACE_MAX_POINTS * 3;
    std::fill(o, o + VR_FACE_MAX_POINTS * 3, 0.0f);
    for (int j = 0; j < faces[i].n_points; ++j) {
      float s[3] = {0.0f, 0.0f, 0.0f};
      for (int64_t k : users[key(faces[i].points[j])]) {
        if (k == i) continue;
        const float *nn = faces[k].normal;
        const float c = (nn[0] * fn[0] + nn[1] * fn[1]) + nn[2] * fn[2];
        if (!(c >= threshold)) continue;
        for (int q = 0; q < 3; ++q) s[q] = s[q] + nn[q];
      }
      for (int q = 0; q < 3; ++q) s[q] = s[q] + fn[q];
      const float len = (float)std::sqrt((double)((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]));
      const float inv = 1.0f / len;
      for (int q = 0; q < 3; ++q) o[3 * j + q] = s[q] * inv;
    }
  }
}

}  // namespace vr
