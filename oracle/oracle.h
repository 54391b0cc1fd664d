/*
 * oracle.h -- CPU restatement of the VRAD radiosity-bake hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker.
 *
 * Every function below restates a reference function (Galaco/VRAD, a Go port
 * of Source SDK VRAD) or, where the reference has only a stub or nothing, the
 * semantics frozen for this path.  Citations are file:line in the reference.
 *
 * Arithmetic: IEEE FP32, no contraction (built with -ffp-contract=off, SSE
 * scalar), correctly rounded division and sqrt; FP64 where stated.  The
 * "kernel order" functions reproduce the exact summation order of the HIP
 * kernels so that FP32 results are bit-identical.
 */
#ifndef VRAD_ORACLE_H
#define VRAD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* raytracer/constants.go:9-11 */
#define OR_TRACE_ID_SKY 0x01000000
#define OR_TRACE_ID_OPAQUE 0x02000000
#define OR_TRACE_ID_STATICPROP 0x04000000
/* raytracer/kdtree/constants.go:25-28 */
#define OR_COST_OF_TRAVERSAL 75
#define OR_COST_OF_INTERSECTION 167
#define OR_MAX_TREE_DEPTH 21
/* common/constants/constants.go:15-19 */
#define OR_MAX_TRACE_LENGTH (1.732050807569 * 32768.0)

/* raytracer/cache/triangle/triintersectdata.go:3-22 -- 48 bytes */
typedef struct {
  float nx, ny, nz, d;   /* FlNx, FlNy, FlNz, FlD */
  int32_t id;            /* NTriangleID */
  float e[6];            /* ProjectedEdgeEquations */
  uint8_t c0, c1;        /* NCoordSelect0, NCoordSelect1 */
  uint8_t flags, unused; /* NFlags, Unused0 */
} or_tri_t;

typedef struct {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
} or_rays_t;

typedef struct {
  const float *sx, *sy, *sz, *ex, *ey, *ez;
} or_segs_t;

typedef struct {
  int64_t n_nodes;
  uint32_t *children; /* OptimisedKDNode.Children: (index << 2) | type */
  float *split;       /* SplittingPlaneValue (leaf: float(count)) */
  int64_t n_idx;
  int32_t *idx;       /* TriangleIndexList */
  float bmin[3], bmax[3];
  int max_depth;      /* deepest leaf (root = 0) */
} or_kdtree_t;

/* ---- triangle precompute: ChangeIntoIntersectionFormat ---- */
void or_precompute(const float *v9, const int32_t *ids, const uint8_t *flags,
                   int64_t n, or_tri_t *out);

/* ---- closest hit, brute force over every triangle (the reference's
 *      effective acceleration structure is one leaf holding all triangles) */
void or_trace_brute(const or_tri_t *tris, int64_t ntri, const or_rays_t *r,
                    int64_t n, int32_t skip_id, int32_t *hit, float *hit_t);
/* FP64 brute force on the same vertices (stated-error reference) */
void or_trace_brute_f64(const float *v9, const int32_t *ids, int64_t ntri,
                        const or_rays_t *r, int64_t n, int32_t skip_id,
                        int32_t *hit, double *hit_t);

/* ---- SAH kd-tree (RefineNode intent) and traversal ---- */
int or_kd_build(const float *v9, int64_t ntri, or_kdtree_t *out);
void or_kd_free(or_kdtree_t *t);
void or_trace_kd(const or_kdtree_t *kd, const or_tri_t *tris, const or_rays_t *r,
                 int64_t n, int32_t skip_id, int32_t *hit, float *hit_t,
                 int nthreads);

/* ---- TestLineDoesHitSky intent ---- */
void or_testline(const or_kdtree_t *kd /* NULL = brute force */,
                 const or_tri_t *tris, int64_t ntri, const or_segs_t *s,
                 int64_t n, int32_t skip_id, uint8_t *visible,
                 uint8_t *hit_sky, int nthreads);
/* the same closest-hit query on a kd-tree, also returning the total node
 * visits and triangle tests over the n segments (the traversal's work) */
void or_testline_stats(const or_kdtree_t *kd, const or_tri_t *tris, const or_segs_t *s,
                       int64_t n, int32_t skip_id, uint8_t *visible, int nthreads,
                       int64_t *visits, int64_t *tests);

/* ---- MakeTransfers (frozen semantics) ---- */
typedef struct {
  const float *ox, *oy, *oz, *nx, *ny, *nz, *area;
} or_patches_t;
typedef struct {
  int64_t n_rows, nnz;
  int64_t *row_ptr; /* n_rows + 1, unpadded */
  int32_t *col;
  float *coef;
  double *row_sum;  /* FP64 row sums before normalisation */
} or_csr_t;
int or_make_transfers(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                      const or_patches_t *p, int64_t np, float normal_offset,
                      int32_t skip_id, int normalize, int nthreads,
                      or_csr_t *out);
void or_csr_free(or_csr_t *c);
/* FP64 row sum in the wave order of the kernels (see or_row_sum_f64) */
double or_row_sum_f64(const float *coef, int64_t len);

/* ---- direct light (GatherSampleLight, frozen semantics) ---- */
typedef struct {
  int32_t type;        /* emittype_t: 0 surface, 1 point, 2 spot, 3 skylight (sun), 5 sky ambient */
  float origin[3];
  float intensity[3];
  float normal[3];     /* spot direction / direction the sun's light travels */
  float constant_attn, linear_attn, quadratic_attn;
  float stopdot, stopdot2, exponent;
  float start_fade, end_fade, cap_distance;
} or_light_t;
typedef struct {
  const float *px, *py, *pz, *nx, *ny, *nz;
} or_samples_t;
void or_gather_light(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                     const or_samples_t *s, int64_t n, const or_light_t *lights,
                     int32_t n_lights, float normal_offset, int32_t skip_id,
                     float *rgb, int nthreads);
void or_sky_dirs(float *dirs /* 162*3 */);
/* CanLeafTraceToSky per leaf: mins/maxs are n*3 int16 (dleaf_t bounds) */
void or_can_leaf_trace_to_sky(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                              const int16_t *mins, const int16_t *maxs, int64_t n,
                              uint8_t *can_see_sky, int nthreads);

/* ---- BounceLight + FP64 energy ---- */
/* FP32 in the exact order of a bounce kernel (coef_kind 0=f32, 1=f16, 2=bf16
 * with per-row scale; order 0 = wave per row, 1 = sequential, 2 = LDS-tiled
 * with tile_cols / group, 3 = stream, 4 = stream with rows dealt to chunks in
 * runs of consecutive rows, 5 = stream with each chunk's rows ranked by row
 * index); energy reduced in the kernels' fixed tree order.
 * Order 2 needs each row's columns in ascending order. */
int or_bounce_f32(int64_t n, const int64_t *row_ptr, const int32_t *col,
                  const void *coef, int coef_kind, const float *row_scale,
                  const float *emit0, const float *refl, const float *area,
                  float *total /* in/out, 3n */, int max_bounces,
                  double stop_frac, double *energy, int order, int tile_cols,
                  int group, int nthreads);
/* FP64 recurrence, sequential order, FP64 coefficients as given (f32 input) */
int or_bounce_f64(int64_t n, const int64_t *row_ptr, const int32_t *col,
                  const float *coef, const float *emit0, const float *refl,
                  const float *area, double *total /* in/out, 3n */,
                  int max_bounces, double stop_frac, double *energy);
/* fixed-tree FP64 reduction used by the energy kernels */
double or_energy_reduce(const double *e, int64_t n);

/* Patch generation: MakePatchForFace (rad/patches/face.go:30-197) and
 * SubdividePatches (rad/patches/subdivide.go:22-438).  Layouts equal the
 * product's vr_face / vr_patch. */
#define OR_FACE_MAX_POINTS 16
typedef struct {
  int32_t n_points;
  float points[OR_FACE_MAX_POINTS][3];
  float normal[3];
  float lux_s[3], lux_t[3];
  float base_light[3];
  uint32_t flags; /* 1 sky, 2 no subdivision */
} or_face_t;
typedef struct {
  int32_t face, parent, child1, child2;
  float origin[3], normal[3];
  float area, chop, lux_scale;
  float mins[3], maxs[3];
  int32_t n_points;
  float points[OR_FACE_MAX_POINTS][3];
} or_patch_t;
/* Writes up to cap patches, returns the count in *n_out (0 ok, -1 a winding
 * past OR_FACE_MAX_POINTS). */
int or_make_patches(const or_face_t *faces, int64_t n_faces, float max_chop, float min_chop,
                    or_patch_t *out, int64_t cap, int64_t *n_out);
/* with GetPhongNormal child normals from per-face vertex normals
 * (OR_FACE_MAX_POINTS x 3 floats per face; NULL = or_make_patches) */
int or_make_patches_phong(const or_face_t *faces, int64_t n_faces, float max_chop,
                          float min_chop, const float *vnorm, or_patch_t *out, int64_t cap,
                          int64_t *n_out);
/* GetPhongNormal at spot for a face with centroid c and vertex normals vn */
void or_phong_normal(const or_face_t *fc, const float c[3], const float *vn, const float spot[3],
                     float out[3]);
/* PairEdges' vertex normals, OR_FACE_MAX_POINTS x 3 floats per face */
void or_pair_edges(const or_face_t *faces, int64_t n, float threshold, float *out);

/* ---- luxel stage ---- */
typedef struct {
  int32_t mins[2], size[2];
  int64_t luxel_ofs;
  int32_t n_luxels;
  uint32_t flags; /* 1 unlit, 2 too big, 4 no points, 8 degenerate lightmap vectors */
} or_face_lightmap_t;
/* CalcFaceExtents (rad/world/face.go:14-90) per face; face flags 1 (sky)
 * and 4 (nolight) are skipped as UpdateAllFaceLightmapExtents does
 * (rad/start.go:102-112).  Returns the luxel total. */
int64_t or_calc_face_extents(const or_face_t *faces, int64_t n, const float *lux_ofs,
                             or_face_lightmap_t *out);
/* Luxel positions and the leaf patch each luxel takes (first containing
 * winding in lightmap coordinates, else nearest centre); outputs may be
 * NULL. */
void or_face_luxels(const or_face_t *faces, int64_t n_faces, const float *lux_ofs,
                    const or_face_lightmap_t *lm, const or_patch_t *patches, int64_t n_patches,
                    float *pos, int32_t *patch);
/* ColorRGBExp32 (the Source lighting lump's luxel: r, g, b bytes and a
 * signed exponent byte, packed little endian into a uint32): e =
 * floor(log2(max channel)) - 7, each channel trunc(c * 2^-e); 0 for a zero,
 * negative or non-finite maximum or e below -127. */
uint32_t or_rgbexp32(float r, float g, float b);
/* The lighting lump of a set of faces: for each face (in order) with
 * n_luxels > 0, its average colour (ColorRGBExp32) and then its luxels;
 * lightofs[f] = byte offset of face f's first luxel, -1 without luxels.
 * Average: float64 channel sums, 64 lane partials (luxel k to lane k % 64),
 * combined by xor butterfly (32 .. 1), divided by n in float64, rounded to
 * float32.  Returns the lump size in bytes; lump may be NULL to size it. */
int64_t or_lighting_lump(const float *luxel_rgb, const or_face_lightmap_t *lm, int64_t n_faces,
                         uint32_t *lump, int32_t *lightofs);

/* ---- BSP clusters and PVS (rad/clustertable/point.go, rad/lightmap/vis.go) */
typedef struct {
  float normal[3];
  float dist;
} or_plane_t;
typedef struct {
  int32_t planenum;
  int32_t children[2]; /* < 0: leaf -1-child */
} or_bspnode_t;
/* ClusterFromPoint / PointInLeaf from node 0 (recursive, as written). */
int16_t or_cluster_from_point(const or_plane_t *planes, const or_bspnode_t *nodes,
                              const int16_t *leaf_cluster, const float p[3]);
/* Patch.ClusterNumber (rad/patches/subdivide.go:100-115). */
void or_patch_clusters(const or_plane_t *planes, const or_bspnode_t *nodes,
                       const int16_t *leaf_cluster, const or_patch_t *patches, int64_t n,
                       int32_t *out);
/* DecompressVis into n_clusters rows of (n_clusters + 7) >> 3 bytes; returns
 * 0, -1 on a zero count, -2 on a read past vis_len; *n_overrun counts
 * clipped runs. */
int or_decompress_vis(const uint8_t *vis, int64_t vis_len, const int32_t *ofs, int32_t n_clusters,
                      uint8_t *rows, int64_t *n_overrun);
/* MakeTransfers restricted to pairs whose clusters pass PVSCheck in either
 * direction (cluster < 0 passes); n_clusters == 0: every pair. */
int or_make_transfers_pvs(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                          const or_patches_t *p, int64_t np, float normal_offset,
                          int32_t skip_id, int normalize, int nthreads, const int32_t *cluster,
                          const uint8_t *pvs_rows, int32_t n_clusters, or_csr_t *out);

/* FP16 / BF16 encode-decode (round to nearest even), used for the precision
 * form's coefficient storage. */
uint16_t or_f32_to_f16(float f);
float or_f16_to_f32(uint16_t h);
uint16_t or_f32_to_bf16(float f);
float or_bf16_to_f32(uint16_t h);

#ifdef __cplusplus
}
#endif
#endif
