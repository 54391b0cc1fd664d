// Internal declarations shared by the host (.cpp) and device (.hip) parts of
// libvrad_hip.so.  Not part of the C-ABI.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <functional>
#include <vector>

#include "../../include/vrad_hip.h"

namespace vr {

// ---- error handling (thread-local last error) ----------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// ---- host geometry ---------------------------------------------------------
// TriIntersectData, raytracer/cache/triangle/triintersectdata.go:3-22 (48 B)
struct TriIntersect {
  float nx, ny, nz, d;
  int32_t id;
  float e[6];
  uint8_t c0, c1, flags, unused;
};
static_assert(sizeof(TriIntersect) == 48, "TriIntersectData is 48 bytes");

// ChangeIntoIntersectionFormat for n triangles (optimisedtriangle.go:30-78)
void precompute_triangles(const float *v9, const int32_t *ids, const uint8_t *flags,
                          int64_t n, TriIntersect *out, int n_threads);

// kd-tree in the reference's flattened encoding
struct KdTree {
  std::vector<uint32_t> children;  // (index << 2) | type
  std::vector<float> split;        // split value, or float(count) in leaves
  std::vector<int32_t> idx;        // TriangleIndexList
  float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
  int max_depth = 0;
};
// RefineNode intent (raytracer/environment.go:238-387); subtrees are built on
// n_threads host threads and stitched in the sequential allocation order.
int build_kdtree(const float *v9, int64_t n, int n_threads, KdTree *out);

// PairEdges' per-vertex normals (rad/lightmap/lightmap.go:37-216) of n faces:
// out holds VR_FACE_MAX_POINTS x 3 floats per face.
void pair_edges(const vr_face *faces, int64_t n, float threshold, float *out);

// ---- procedural scenes -----------------------------------------------------
struct GenScene {
  std::vector<float> v9;
  std::vector<int32_t> ids;
  std::vector<uint8_t> flags;
  std::vector<float> po, pn, parea, prefl;  // patches
  std::vector<vr_light> lights;
  std::vector<double> faces;  // per patch rectangle P + a U + b V: P, U, V, normal (12 each)
  std::vector<double> grid_x, grid_y;  // room cell boundaries (wall centre lines)
};

// ---- device-side records ---------------------------------------------------
// One leaf reference = 3 float4:
//   q0 = (nx, ny, nz, d)
//   q1 = (e0, e1, e2, e3)
//   q2 = (e4, e5, bits(id), bits(tri | c0 << 28 | c1 << 30))
constexpr int kRefQuads = 3;
constexpr uint32_t kTriIndexMask = (1u << 28) - 1u;

int host_threads(int requested);
// fn(begin, end) over [0, n) split into n_threads contiguous ranges on host
// threads (in the calling thread when n is small or n_threads <= 1)
void parallel_for(int64_t n, int n_threads, const std::function<void(int64_t, int64_t)> &fn);

}  // namespace vr
