/*
 * vrad_hip.h -- C-ABI of the MI355X radiosity-bake path.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (VR_OK = 0, negative VR_E_* on failure) and never aborts; vr_last_error()
 * returns the calling thread's last message.  Host pointers passed in are
 * read or written only for the duration of the call (cgo pointer rules).
 *
 * Each declaration names the reference (Galaco/VRAD, Go) interface it
 * replaces as file:line; where the reference has a stub or nothing, the
 * comment says so.  A Go caller binds these with cgo, one thin wrapper per
 * reference function that turns the status into a Go error.
 *
 * Threading: calls are serialised by one library lock; the host-pointer
 * forms are synchronous, the *_dev forms are asynchronous on the stream
 * given (NULL = the library stream).
 */
#ifndef VRAD_HIP_H
#define VRAD_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define VR_OK 0
#define VR_E_INVALID -1     /* bad argument */
#define VR_E_NOMEM -2       /* host or device allocation failed */
#define VR_E_HIP -3         /* HIP runtime error */
#define VR_E_UNSUPPORTED -4 /* e.g. a transparent-triangle callback */
#define VR_E_STATE -5       /* wrong object state (scene not built, ...) */
#define VR_E_NODEVICE -6    /* no usable GPU */
#define VR_E_CAPACITY -7    /* request exceeds a configured capacity */

/* raytracer/constants.go:9-11 */
#define VR_TRACE_ID_SKY 0x01000000
#define VR_TRACE_ID_OPAQUE 0x02000000
#define VR_TRACE_ID_STATICPROP 0x04000000
/* raytracer/constants.go:5-7 */
#define VR_RTE_FLAGS_FAST_TREE_GENERATION 1
#define VR_RTE_FLAGS_DONT_STORE_TRIANGLE_COLORS 2
#define VR_RTE_FLAGS_DONT_STORE_TRIANGLE_MATERIALS 4
/* raytracer/cache/optimisedkdnode.go:3-6 */
#define VR_KDNODE_STATE_XSPLIT 0
#define VR_KDNODE_STATE_YSPLIT 1
#define VR_KDNODE_STATE_ZSPLIT 2
#define VR_KDNODE_STATE_LEAF 3
/* emittype_t as used by rad/lightmap/lights.go:71,216,235,388,392 */
#define VR_EMIT_SURFACE 0
#define VR_EMIT_POINT 1
#define VR_EMIT_SPOTLIGHT 2
#define VR_EMIT_SKYLIGHT 3
#define VR_EMIT_SKYAMBIENT 5

/* kernel paths: the optimised path (LDS traversal stacks, wave-per-row
 * gathers, device-side stop test) and the plain path it is compared with
 * (scratch-memory stacks, one thread per row, host energy read per bounce) */
#define VR_PATH_OPT 0
#define VR_PATH_PLAIN 1

/* transfer-coefficient storage */
#define VR_COEF_F32 0
#define VR_COEF_F16 1
#define VR_COEF_BF16 2

/* ---- library ----------------------------------------------------------- */
/* Thread-local text of the last failure of this thread ("" if none). */
const char *vr_last_error(void);
/* Select the HIP device and create the library stream.  Called implicitly
 * by the first entry point that needs the GPU.  No reference counterpart
 * (the reference is single-process CPU code). */
int vr_init(int device);
/* Select VR_PATH_OPT / VR_PATH_PLAIN for the calls that follow. */
int vr_set_path(int path);
/* Kernel-shape knobs (the autotuner's interface; no reference counterpart).
 * Results do not depend on:
 *   "trace_lds_entries"  traversal-stack entries per lane held in LDS
 *                        (2, 4, 6, 8, 10, 12, 16, 24, default 10; the rest spill
 *                        to scratch)
 *   "trace_rays_per_lane" rays each lane of the visibility kernels takes
 *                        (1, 2, 4, 8): rays per wavefront = 64 x this; for
 *                        TestLine above 1 a lane's consecutive segments are
 *                        first tested against the triangle that blocked the
 *                        one before (results unchanged)
 *   "trace_xcd_group"    TestLine, closest hit and the MakeTransfers pair
 *                        trace deal their workgroups to the 8 XCDs in
 *                        groups of this many consecutive 64-ray chunks,
 *                        round-robin (default 256; 0 = one contiguous range
 *                        per XCD)
 *   "bounce_waves"       persistent waves of the wave-per-row gather
 *   "kd_build"           kd-tree construction in vr_scene_build: 0 = host
 *                        threads, 1 = GPU, 2 (default) = GPU for a tree
 *                        (main or sky) of "kd_build_device_min" (default
 *                        2048) triangles or more; the tree is the same bit
 *                        for bit
 *   "patch_build"        vr_make_patches: 1 (default) = level-parallel (every
 *                        patch of a subdivision level split by its own
 *                        thread, then renumbered), 0 = a thread per face;
 *                        the same patches in the same order
 * These select the FP32 summation order of BounceLight (the CPU
 * restatement takes the same values):
 *   "bounce_kernel"      2 = LDS stream gather (default), 0 = LDS-tiled
 *                        gather, 1 = wave per row
 *   "bounce_tile_cols"   2048 / 4096 (default) patches per LDS tile
 *   "bounce_group"       tiled gather: 8 / 16 lanes per row segment;
 *                        0 (default) = 16 for FP32 coefficients, 8 for 16-bit
 *   "bounce_passes"      tiled gather: 1 / 2 passes of a row segment loaded
 *                        ahead of use; 0 (default) = 1
 *   "bounce_prefetch"    passes (stream: 1, 2) or steps (tiled: 1, 3, 7) the
 *                        gather's loads run ahead; 0 (default) = 1 for the
 *                        stream gather, 3 for the tiled one
 *   "bounce_placement"   stream gather: how rows are dealt to its 1024-row
 *                        chunks and their waves; 0 (default) = a stride
 *                        through the longest-first order of all rows, 1 =
 *                        runs of consecutive rows, 2 = placement 0's chunks
 *                        with their rows dealt in row-index order
 * The kernel, tile width, group and placement fix the summation order; the
 * other bounce knobs change only the schedule.  vr_bounce_light fails with
 * VR_E_UNSUPPORTED for a combination that has no compiled kernel. */
int vr_set_tuning(const char *name, int value);
int vr_get_tuning(const char *name, int *value);
/* Kernel timing with HIP events on the launching stream.  name = kernel
 * family ("trace_closest", "trace_occluded", "bounce_gather", "energy",
 * "make_transfers", "gather_light"). */
int vr_profile_enable(int on);
int vr_profile_reset(void);
int vr_profile_get(const char *name, double *total_ms, int64_t *launches);
/* Read bandwidth (GB/s) a plain streaming kernel reaches on this device: the
 * best of `reps` timed passes of 16-byte loads over `bytes` of device memory
 * (no reference counterpart; the practical ceiling beside the HBM peak). */
int vr_measure_read_bandwidth(int64_t bytes, int32_t reps, double *gbps);

/* ---- scene: raytracer.Environment ------------------------------------ */
typedef struct vr_scene vr_scene;

/* Bulk form of Environment.AddTriangleWithMaterial / AddTriangle
 * (raytracer/environment.go:41-69): verts9 holds n_tri x VertexCoordData[9]
 * (TriGeometryData, raytracer/cache/triangle/trigeometrydata.go:3-12),
 * ids the NTriangleID values, flags the NFlags (may be NULL). */
int vr_scene_create(vr_scene **out, const float *verts9, const int32_t *ids,
                    const uint8_t *flags, int64_t n_tri);

/* Environment.SetupAccelerationStructure (raytracer/environment.go:119-138):
 * SAH kd-tree after RefineNode/CalculateCostsOfSplit (:181-387) with the
 * costs of raytracer/kdtree/constants.go:25-28, then
 * ChangeIntoIntersectionFormat per triangle
 * (raytracer/cache/optimisedtriangle.go:30-78), upload to HBM.  The tree is
 * built on the GPU or on n_threads host threads (tuning knob "kd_build"),
 * the same tree either way.  n_threads <= 0 uses every host core. */
int vr_scene_build(vr_scene *s, int32_t n_threads);
/* vr_scene_build's kd-tree and intersection format on the host alone (no
 * GPU): lets vr_scene_get_kdtree / vr_scene_get_triangles run on a host
 * without a device.  A later vr_scene_build keeps its result. */
int vr_scene_build_host(vr_scene *s, int32_t n_threads);

typedef struct {
  int64_t n_tri, n_nodes, n_leaf_refs;
  int32_t max_depth, has_sky;
  float bmin[3], bmax[3]; /* Environment.MinBound / MaxBound */
  int64_t device_bytes;
} vr_scene_info;
int vr_scene_get_info(const vr_scene *s, vr_scene_info *out);

/* Environment.OptimizedKDTree + TriangleIndexList
 * (raytracer/environment.go:33-35; node encoding
 * raytracer/cache/optimisedkdnode.go:15-54). */
int vr_scene_get_kdtree(const vr_scene *s, uint32_t *children, float *split,
                        int32_t *triangle_index_list);
/* Environment.OptimizedTriangleList in intersection format: n_tri records of
 * 48 bytes laid out as TriIntersectData
 * (raytracer/cache/triangle/triintersectdata.go:3-22); GetTriangle (:422). */
int vr_scene_get_triangles(const vr_scene *s, void *tri_intersect_data);
int vr_scene_destroy(vr_scene *s);

/* ---- visibility ------------------------------------------------------ */
typedef struct {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
} vr_rays;
typedef struct {
  const float *sx, *sy, *sz, *ex, *ey, *ez;
} vr_segs;

/* Environment.Trace4Rays (raytracer/environment.go:140-145, a stub in the
 * reference), batched: for each ray the closest triangle with
 * tmin <= t < tmax (ties to the lower index), skipping triangles whose
 * NTriangleID == skip_id.  hit_idx = -1 for no hit, else the index into
 * OptimizedTriangleList (raytracer/types/result.go:8-12);
 * hit_normal (3n, nullable) = the hit triangle's plane normal.
 * A non-NULL transparency callback has no C form (VR_E_UNSUPPORTED is
 * returned by the Go wrapper before calling here). */
int vr_trace_closest(vr_scene *s, const vr_rays *rays, int64_t n,
                     int32_t skip_id, int32_t *hit_idx, float *hit_t,
                     float *hit_normal);

/* trace.TestLineDoesHitSky (raytracer/trace/testline.go:18-94; panics at
 * :21 in the reference), batched: dir = (end-start)*(1/len), trace
 * [0, len); occluded iff the closest hit is not a TRACE_ID_SKY triangle.
 * visible[i] = 1 - occluded (fractionVisible, :91-93); hit_sky nullable. */
int vr_test_line(vr_scene *s, const vr_segs *segs, int64_t n, int32_t skip_id,
                 uint8_t *visible, uint8_t *hit_sky);

/* Device-pointer forms of the two calls above: every pointer (including the
 * members of *rays / *segs, read on the host) is device memory; the call
 * returns after enqueueing on `stream` (hipStream_t, NULL = library stream).
 * No reference counterpart: they let callers keep ray batches resident. */
int vr_trace_closest_dev(vr_scene *s, const vr_rays *rays, int64_t n,
                         int32_t skip_id, int32_t *hit_idx, float *hit_t,
                         float *hit_normal, void *stream);
int vr_test_line_dev(vr_scene *s, const vr_segs *segs, int64_t n,
                     int32_t skip_id, uint8_t *visible, uint8_t *hit_sky,
                     void *stream);

/* ---- transfers: rad/patches MakeTransfers ----------------------------- */
typedef struct {
  const float *ox, *oy, *oz, *nx, *ny, *nz, *area;
} vr_patches;
typedef struct {
  float normal_offset;  /* segment end offset along each normal (0.1) */
  int32_t skip_id;
  int32_t normalize;    /* scale rows with FP64 sum > 1 by 1/sum */
  int64_t batch_pairs;  /* candidate pairs per visibility batch (<=0: 32M) */
} vr_xfer_params;
typedef struct vr_xfers vr_xfers;
typedef struct {
  int64_t n_rows, nnz, stored_nnz, stored_bytes;
  int32_t precision;
  int64_t pairs_considered, segments_traced;
} vr_xfers_info;

/* MakeTransfers (absent in the reference; result records
 * common/types/transfer.go:3-6 and Patch.NumTransfers/Transfers,
 * common/types/patch.go:60-61): leaf patches i<j passing the facing cull
 * are tested with TestLine between the normal-offset origins, visible pairs
 * get F_ij = cos_i cos_j A_j / (pi r^2) and the CSR stays resident. */
int vr_make_transfers(vr_scene *s, const vr_patches *p, int64_t n_patches,
                      const vr_xfer_params *params, vr_xfers **out);
/* MakeTransfers for many scenes at once (no reference counterpart): scene
 * k's patches[k] (n_patches[k] of them) give out[k],
 * equal bit for bit to vr_make_transfers(scenes[k], ...).  Every round, each
 * unfinished scene culls its next candidate rows into its own region of one
 * pair buffer (params->batch_pairs pairs per round, default 32M, shared
 * among the scenes still culling, at least a row each), and one visibility
 * launch walks all regions through a table of scene descriptors.  A scene
 * with params->batch_pairs candidate pairs or more fills a round alone and
 * runs as vr_make_transfers would.  The plain path runs the scenes one by
 * one.
 * On failure no out[k] is set. */
int vr_make_transfers_batch(vr_scene *const *scenes, const vr_patches *patches,
                            const int64_t *n_patches, int32_t n_scenes,
                            const vr_xfer_params *params, vr_xfers **out);
/* Upload a host CSR (row_ptr n_rows+1, col/coef nnz) as transfers; each row
 * is stored with its columns in ascending order (stable). */
int vr_xfers_from_csr(vr_xfers **out, int64_t n_rows, const int64_t *row_ptr,
                      const int32_t *col, const float *coef, int32_t precision);
/* Synthetic transfers on the device (no reference counterpart; the
 * precision workload's CSR): row lengths lognormal (mu = ln(mean) -
 * sigma^2/2), clipped to [1, max_len] and rescaled to the mean; columns
 * VR_SYNTH_UNIFORM: a uniformly random subset of [0,n)\{i} drawn without
 * replacement (the first distinct values of a seeded stream of uniform
 * draws), or VR_SYNTH_STRATIFIED: one per equal stratum; sorted per row;
 * coef = u/sum(u) * s_i with s_i ~ U[0.3,0.9]. */
#define VR_SYNTH_UNIFORM 0
#define VR_SYNTH_STRATIFIED 1
int vr_xfers_synthetic(vr_xfers **out, uint64_t seed, int64_t n_rows,
                       double mean_len, double sigma, int64_t max_len,
                       int32_t precision, int32_t columns);
/* Re-encode the coefficients: FP16/BF16 store coef/row_max with the row
 * max as an FP32 per-row scale.  VR_E_STATE while a server holds x. */
int vr_xfers_set_precision(vr_xfers *x, int32_t precision);
int vr_xfers_get_info(const vr_xfers *x, vr_xfers_info *out);
/* Copy out the logical CSR (coef decoded to FP32 without the row scale for
 * FP16/BF16; raw_coef16 nullable gets the stored 16-bit words; row_scale
 * nullable). */
int vr_xfers_download(const vr_xfers *x, int64_t *row_ptr, int32_t *col,
                      float *coef, uint16_t *raw_coef16, float *row_scale);
/* VR_E_STATE (and x stays valid) while a server holds x. */
int vr_xfers_destroy(vr_xfers *x);

/* ---- direct light: rad/lightmap GatherSampleLight --------------------- */
/* worldlight.WorldLight fields used by rad/lightmap/lights.go:168-341 plus
 * DirectLight's fade/cap (common/types/light.go:31-36). */
typedef struct {
  int32_t type;
  float origin[3];
  float intensity[3];
  float normal[3];
  float constant_attn, linear_attn, quadratic_attn;
  float stopdot, stopdot2, exponent;
  float start_fade, end_fade, cap_distance;
} vr_light;
typedef struct {
  const float *px, *py, *pz, *nx, *ny, *nz;
} vr_samples;

/* GatherSampleLight (absent in the reference): for every sample, the sum
 * over lights (list order) of intensity x falloff x TestLine visibility. */
int vr_gather_sample_light(vr_scene *s, const vr_samples *smp, int64_t n,
                           const vr_light *lights, int32_t n_lights,
                           float normal_offset, int32_t skip_id, float *rgb);

/* GatherSampleLight for many scenes in one launch (no reference
 * counterpart): scene k's n[k] samples with its lights[k]
 * (n_lights[k]) write rgb[k] (3 n[k]); per scene equal bit for bit to
 * vr_gather_sample_light.  The plain path runs the scenes one by one. */
int vr_gather_sample_light_batch(vr_scene *const *scenes, const vr_samples *samples,
                                 const int64_t *n, const vr_light *const *lights,
                                 const int32_t *n_lights, int32_t n_scenes,
                                 float normal_offset, int32_t skip_id, float *const *rgb);

/* CanLeafTraceToSky (rad/lightmap/lightmap.go:425-451), batched over leaves.
 * Leaf i's bounds are mins[3i..3i+2], maxs[3i..3i+2] (dleaf_t int16); its
 * centre is float32(mins + maxs) * 0.5 per axis (32-bit sum).  can_see_sky[i]
 * = 1 iff, for one of the 162 Anorms directions a (vmath/constants.go:15-184),
 * TestLineDoesHitSky from the centre to (a * -MAX_TRACE_LENGTH) + centre with
 * skip id -1 finds the segment visible (a miss, or a TRACE_ID_SKY hit before
 * any other triangle); else 0.  n < 2^32. */
int vr_can_leaf_trace_to_sky(vr_scene *s, const int16_t *mins, const int16_t *maxs,
                             int64_t n, uint8_t *can_see_sky);

/* ---- bounce: rad/lightmap BounceLight --------------------------------- */
typedef struct {
  int32_t max_bounces;  /* -bounce (cmd/args.go:66); numBounce default 2 */
  double stop_frac;     /* stop when E_k <= stop_frac * sum_m E_m */
} vr_bounce_params;

/* BounceLight (absent in the reference; state Patch.TotalLight,
 * Reflectivity, Area, common/types/patch.go:33-42): Jacobi gather
 * add = F.emit, total += add, emit = add*refl, FP64 energy per bounce.
 * emit0/refl/total are 3n RGB, area n.  energy gets one FP64 per bounce. */
int vr_bounce_light(vr_xfers *x, const float *emit0, const float *refl,
                    const float *area, float *total,
                    const vr_bounce_params *params, double *energy,
                    int32_t *n_bounces);

/* vr_bounce_light without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Arguments and results are those of vr_bounce_light, bit for
 * bit.  The device working set (emit ping-pong, total, reflectivity, area,
 * per-row energy terms, energy partials, bounce state) lives in one fixed
 * array per GPU, held until the process ends: at most 524288 rows
 * (VR_E_CAPACITY above) and 1024 bounces.  Runs the LDS-tiled or stream
 * gather of the optimised path (else VR_E_UNSUPPORTED) on a transfer set
 * whose gather layout for the current tuning is already built (else
 * VR_E_STATE: run vr_bounce_light on it once).  Every call on a GPU shares
 * that GPU's array, which keeps the last call's contents; the library lock
 * serialises calls within a process, and a call that overlapped another on
 * the same array would race. */
int vr_bounce_light_noalloc(vr_xfers *x, const float *emit0, const float *refl,
                            const float *area, float *total,
                            const vr_bounce_params *params, double *energy,
                            int32_t *n_bounces);

/* Batched BounceLight (no reference counterpart): every job is an
 * independent transfer set with its own host arrays; one launch per bounce
 * covers all jobs' rows (the stream or the LDS-tiled gather), each
 * job keeps its own FP64 energy and stop test on the device.  Jobs share the
 * coefficient precision.  use_graph != 0 captures 8 bounces as one hipGraph
 * and replays it until every job has stopped.  Per job the results equal
 * vr_bounce_light's bit for bit. */
typedef struct {
  vr_xfers *xfers;
  const float *emit0, *refl, *area; /* 3n, 3n, n */
  float *total;                     /* 3n, in/out */
  double *energy;                   /* max_bounces, out */
  int32_t n_bounces;                /* out */
} vr_bounce_job;
int vr_bounce_light_batch(vr_bounce_job *jobs, int32_t n_jobs,
                          const vr_bounce_params *params, int32_t use_graph);

/* ---- serving: re-bake requests against resident scenes ----------------- */
/* No reference counterpart.  A server keeps
 * each scene's tree, patches, transfers and the transfers' tile-blocked copy
 * resident in HBM and runs re-bake requests: the scene's lights with 1..k of
 * them moved, direct light re-gathered, then `bounces` bounces from it.
 *
 * VR_SERVE_BATCHED: continuous batching.  A scheduler thread admits queued
 * requests in arrival order while a slot (max_active) and the request's
 * working memory (carved from an hbm_budget arena) are free; every iteration
 * runs up to eight bounces of every admitted request, one batched launch per
 * bounce, so requests that arrive while others run join at the next
 * iteration.  The
 * scenes share one coefficient precision (FP16 in the serving bench).
 * VR_SERVE_PLAIN: the comparison path, one request at a time in arrival
 * order with the plain kernels (scratch-stack gather, one thread per CSR row
 * on the FP32 coefficients, a host read of the energy after every bounce).
 * A request's result does not depend on what else runs with it.
 * Scenes are added before the first request; the scene and transfer handles
 * must outlive the server.  The launch knobs (vr_set_tuning) are read once,
 * at vr_server_create; the batched scenes share one gather shape.  While a
 * server holds a transfer set, vr_xfers_set_precision, vr_xfers_destroy and
 * a vr_bounce_light / vr_bounce_light_batch needing another gather layout of
 * it return VR_E_STATE before any work. */
#define VR_SERVE_BATCHED 0
#define VR_SERVE_PLAIN 1
typedef struct vr_server vr_server;
typedef struct {
  int32_t mode;        /* VR_SERVE_BATCHED / VR_SERVE_PLAIN */
  int32_t max_active;  /* requests on the device at once (1..256; plain: 1) */
  int64_t hbm_budget;  /* bytes of request working memory (<= 0: 8 GiB) */
  int32_t bounces;     /* bounces per request */
  float normal_offset; /* sample offset along the normal for visibility */
} vr_server_params;
typedef struct {
  int64_t submitted, completed, iterations, max_active_seen, pending, active;
} vr_server_stats;
int vr_server_create(vr_server **out, const vr_server_params *params);
/* patches: sample points and normals (n = the transfers' rows), reflectivity
 * 3n, area n; lights: the scene's base light list.  *scene_id is its index. */
int vr_server_add_scene(vr_server *sv, vr_scene *scene, vr_xfers *xfers,
                        const vr_samples *patches, const float *refl,
                        const float *area, const vr_light *lights,
                        int32_t n_lights, int32_t *scene_id);
/* Queue a re-bake of scene_id with lights light_index[i] moved by
 * offset[3i..3i+2]; the request's clock starts here. */
int vr_server_submit(vr_server *sv, int32_t scene_id, const int32_t *light_index,
                     const float *offset, int32_t n_moved, int64_t *ticket);
/* Wait for a request (timeout_ms < 0: no limit; VR_E_STATE on timeout) and
 * take its result: total 3n, energy per bounce, bounces done, latency
 * (completion - submit) in ms; any output may be NULL.  A ticket is taken
 * once. */
int vr_server_wait(vr_server *sv, int64_t ticket, int32_t timeout_ms, float *total,
                   double *energy, int32_t *n_bounces, double *latency_ms);
int vr_server_get_stats(vr_server *sv, vr_server_stats *out);
/* Finishes the queued requests, then stops the scheduler. */
int vr_server_destroy(vr_server *sv);

/* ---- patch generation ---------------------------------------------------- */
#define VR_FACE_MAX_POINTS 16
#define VR_FACE_SKY 1u    /* SURF_SKY: the face's patch is never subdivided */
#define VR_FACE_NOCHOP 2u /* PreventSubdivision (SURF_NOCHOP, or NOLIGHT without LIGHT) */
/* One face: its winding, plane normal and lightmap vectors (texinfo
 * lightmapVecsLuxelsPerWorldUnits, xyz parts). */
typedef struct {
  int32_t n_points;
  float points[VR_FACE_MAX_POINTS][3];
  float normal[3];
  float lux_s[3], lux_t[3];
  float base_light[3]; /* BaseLightForFace: emitted light of a light surface */
  uint32_t flags;      /* VR_FACE_* */
} vr_face;
/* One patch (common/types/patch.go): face, parent and children indices (-1
 * none), centre, normal, area, chop (luxel widths), luxels per world unit,
 * bounds and winding. */
typedef struct {
  int32_t face, parent, child1, child2;
  float origin[3], normal[3];
  float area, chop, lux_scale;
  float mins[3], maxs[3];
  int32_t n_points;
  float points[VR_FACE_MAX_POINTS][3];
} vr_patch;
/* MakePatches / MakePatchForFace (rad/patches/build.go:21-68,
 * rad/patches/face.go:30-197) then SubdividePatches (rad/patches/
 * subdivide.go:22-128): a base patch per face of positive area (in face
 * order), then each base patch's subdivision (SubdividePatch, ClipWinding-
 * Epsilon, CreateChildPatch, WindingAreaAndBalancePoint, :166-438) appended
 * in the order it creates children; leaves have child1 = -1.  max_chop /
 * min_chop are maxChop / minChop (face.go:24-25, 4 and 4).  Runs on the GPU.
 * Returns the patch count in *n_out; with out == NULL or cap too small only
 * counts (VR_E_CAPACITY when out != NULL). */
int vr_make_patches(const vr_face *faces, int64_t n_faces, float max_chop, float min_chop,
                    vr_patch *out, int64_t cap, int64_t *n_out);

/* PairEdges' vertex normals (rad/lightmap/lightmap.go:37-228): for vertex j
 * of face i, the normalised sum of the normals of the other faces that share
 * that vertex (same position, bit for bit) and whose normal's dot with face
 * i's is >= smoothing_threshold (smoothingThreshold, lightmap.go:29,
 * 0.7071067), taken in face order, plus face i's own normal last.  Writes
 * VR_FACE_MAX_POINTS x 3 floats per face (unused vertices 0).  Faces carry no
 * smoothing groups or displacements (both 0 / absent in the reference's
 * test maps).  Runs on the host: it is a hash join over vertices, O(points). */
int vr_pair_edges(const vr_face *faces, int64_t n_faces, float smoothing_threshold,
                  float *vertex_normals);
/* vr_make_patches with GetPhongNormal (rad/lightmap/normallist.go:51-144,
 * called at rad/patches/subdivide.go:387): each child patch's normal is the
 * face normal bent toward vertex_normals (vr_pair_edges' output, or any
 * VR_FACE_MAX_POINTS x 3 per face) at the child's centre, interpolated in
 * the fan triangle (face centroid, vertex j, vertex j+1) that holds it; the
 * face centroid is the base patch's centre (face.go:153).  vertex_normals
 * NULL is vr_make_patches. */
int vr_make_patches_phong(const vr_face *faces, int64_t n_faces, float max_chop, float min_chop,
                          const float *vertex_normals, vr_patch *out, int64_t cap,
                          int64_t *n_out);

/* ---- luxel stage: lightmap extents and luxels ------------------------- */
#define VR_FACE_NOLIGHT 4u /* SURF_NOLIGHT: no lightmap, like SURF_SKY */
/* MAX_LIGHTMAP_DIM_WITHOUT_BORDER = MAX_DISP_LIGHTMAP_DIM_WITHOUT_BORDER
 * (common/constants/constants.go:22,27): one limit for faces and
 * displacements. */
#define VR_LIGHTMAP_MAX_DIM 125
#define VR_LM_UNLIT 1u      /* VR_FACE_SKY or VR_FACE_NOLIGHT: skipped, no luxels */
#define VR_LM_TOO_BIG 2u    /* a side above VR_LIGHTMAP_MAX_DIM + 1 luxels: no luxels */
#define VR_LM_NO_POINTS 4u  /* n_points < 1: no luxels */
#define VR_LM_DEGENERATE 8u /* lightmap vectors and normal dependent: luxel positions 0 */
/* A face's lightmap: LightmapTextureMinsInLuxels / SizeInLuxels and where
 * its (size[0] + 1) x (size[1] + 1) luxels start (luxel (s, t) at index
 * luxel_ofs + t * (size[0] + 1) + s). */
typedef struct {
  int32_t mins[2], size[2];
  int64_t luxel_ofs;
  int32_t n_luxels;
  uint32_t flags; /* VR_LM_* */
} vr_face_lightmap;
/* CalcFaceExtents (rad/world/face.go:14-90) for every face, skipping sky and
 * nolight faces as UpdateAllFaceLightmapExtents does (rad/start.go:102-
 * 112): per point, val = p . lightmapVecs[j].xyz + lightmapVecs[j][3] in
 * FP32; mins = floor, size = ceil(max) - floor(min).  lux_ofs holds the two
 * offsets lightmapVecs[0][3], [1][3] per face (NULL: 0).  Luxel offsets are
 * assigned in face order; *n_luxels is the total.  Runs on the GPU. */
int vr_calc_face_extents(const vr_face *faces, int64_t n_faces, const float *lux_ofs,
                         vr_face_lightmap *out, int64_t *n_luxels);
/* The luxels of every face (patch-to-luxel interpolation, FinalLightFace,
 * is absent in the reference): luxel (s, t) lies on the face plane at
 * lightmap coordinates (mins[0] + s, mins[1] + t), solved in FP64; it takes
 * the value of the first leaf patch of its face (patch order) whose
 * winding, mapped to lightmap coordinates, contains it, else of the leaf
 * patch whose centre is nearest in lightmap coordinates.  lm is
 * vr_calc_face_extents's output for the same faces; patches as
 * vr_make_patches writes them (face indices into faces).  Outputs per
 * luxel, each may be NULL: pos 3, normal 3 (the face normal), patch 1 (-1:
 * the face has no leaf patch), value 3 (patch_value[patch], 3 per patch;
 * 0 without a patch).  Runs on the GPU. */
int vr_face_luxels(const vr_face *faces, int64_t n_faces, const float *lux_ofs,
                   const vr_face_lightmap *lm, const vr_patch *patches, int64_t n_patches,
                   const float *patch_value, float *pos, float *normal, int32_t *patch,
                   float *value);

/* The BSP lighting lump of the luxel light (the on-disk output; the
 * reference's writer is a stub, cmd/tasks/finish/main.go:15-18).  For each
 * face (in order) with n_luxels > 0: its average colour, then its luxels,
 * each a ColorRGBExp32 (bytes r, g, b, signed exponent e: colour = bytes x
 * 2^e; e = floor(log2(max channel)) - 7, channels truncated, as Source's
 * VectorToColorRGBExp32 named at compute/leafambient/lighting.go:67).
 * lightofs[f] (nullable) = byte offset of face f's first luxel, -1 without
 * luxels (dface_t.lightofs; the average sits in the 4 bytes before it).
 * luxel_rgb: 3 floats per luxel (n_luxels), indexed by lm's luxel_ofs.  The
 * average is the mean of the face's luxels, summed in float64 (64 lane
 * partials, luxel k to lane k mod 64, xor butterfly 32 .. 1) and rounded to
 * float32.  *lump_bytes gets the size; lump NULL only sizes it (lump_cap:
 * bytes available).  Runs on the GPU. */
int vr_lighting_lump(const float *luxel_rgb, int64_t n_luxels, const vr_face_lightmap *lm,
                     int64_t n_faces, uint8_t *lump, int64_t lump_cap, int32_t *lightofs,
                     int64_t *lump_bytes);

/* ---- BSP clusters and PVS: rad/clustertable, rad/lightmap/vis.go -------- */
/* The BSP lumps PointInLeaf reads (cache.GetLumpCache().Planes / Nodes /
 * Leafs, rad/clustertable/point.go:14-36): a plane's normal and distance, a
 * node's plane and children (child < 0: leaf -1-child), a leaf's cluster.
 * Node 0 is the root; the nodes reachable from it must form a tree of depth
 * <= VR_BSP_MAX_DEPTH. */
#define VR_BSP_MAX_DEPTH 64
typedef struct {
  float normal[3];
  float dist;
} vr_bsp_plane;
typedef struct {
  int32_t planenum;
  int32_t children[2];
} vr_bsp_node;
typedef struct {
  const vr_bsp_plane *planes;
  int64_t n_planes;
  const vr_bsp_node *nodes;
  int64_t n_nodes;
  const int16_t *leaf_cluster;
  int64_t n_leafs;
} vr_bsp;
/* ClusterFromPoint (rad/clustertable/point.go:10-36) per point: descend by
 * dist = p.normal - plane.dist against TEST_EPSILON (vmath/constants.go:7);
 * a point within the epsilon tries children[0] first and takes children[1]
 * when that gives cluster -1.  xyz is 3n floats.  Runs on the GPU. */
int vr_cluster_from_points(const vr_bsp *bsp, const float *xyz, int64_t n, int16_t *cluster);
/* Patch.ClusterNumber as SubdividePatches assigns it (rad/patches/
 * subdivide.go:100-115): the cluster of the patch origin, or when that is -1
 * of the first winding point whose cluster is not -1.  Runs on the GPU. */
int vr_patch_clusters(const vr_bsp *bsp, const vr_patch *patches, int64_t n, int32_t *cluster);
/* GetVisCache / DecompressVis (rad/lightmap/vis.go:9-94) for every cluster:
 * row c (row = (n_clusters + 7) >> 3 bytes at rows + c * row) is the
 * run-length-decoded PVS at visdata[pvs_ofs[c]] (a nonzero byte is copied,
 * a zero byte is followed by a count of zero bytes).  A run past the row is
 * clipped (counted in *n_overrun, may be NULL); a zero count, an offset of -1
 * or a read past vis_len is VR_E_INVALID.  Runs on the GPU. */
int vr_decompress_vis(const uint8_t *visdata, int64_t vis_len, const int32_t *pvs_ofs,
                      int32_t n_clusters, uint8_t *rows, int64_t *n_overrun);
/* MakeTransfers with PVS culling of candidate pairs: a pair (i, j) is
 * considered only when PVSCheck (rad/lightmap/lightmap.go:413-422) passes
 * for it in either direction, patch_cluster[i] < 0 passing everything (the
 * -1 rule of PVSCheck and GetVisCache).  pvs_rows is n_clusters rows as
 * vr_decompress_vis writes them; n_clusters == 0 considers every pair.
 * Otherwise as vr_make_transfers, whose result it equals whenever the PVS
 * holds every visible pair. */
int vr_make_transfers_pvs(vr_scene *s, const vr_patches *p, int64_t n_patches,
                          const vr_xfer_params *params, const int32_t *patch_cluster,
                          const uint8_t *pvs_rows, int32_t n_clusters, vr_xfers **out);

/* ---- procedural scenes (inputs for tests and benchmarks) --------------- */
typedef struct vr_gen vr_gen;
/* Closed room 1024x1024x256 with n_pillars seeded axis-aligned pillars,
 * shell tessellated at `tess` units, patches chopped at `chop`, one point
 * light (intensity 10000, quadratic attenuation 1). */
int vr_gen_room(vr_gen **out, uint64_t seed, int32_t n_pillars, float chop,
                float tess);
/* Grid of rooms_x x rooms_y connected rooms (sizes 256-2048, doorways
 * 128x192, pillars, ramps, skylights), n_point_lights point lights and one
 * sun when sun != 0. */
int vr_gen_level(vr_gen **out, uint64_t seed, int32_t rooms_x, int32_t rooms_y,
                 float chop, float tess, int32_t n_point_lights, int32_t sun);
int vr_gen_counts(const vr_gen *g, int64_t *n_tri, int64_t *n_patches,
                  int32_t *n_lights);
/* verts9 9n, ids n, flags n, patch arrays (o/n/refl 3 per patch, area 1),
 * lights n_lights; any pointer may be NULL. */
int vr_gen_copy(const vr_gen *g, float *verts9, int32_t *ids, uint8_t *flags,
                float *patch_origin, float *patch_normal, float *patch_area,
                float *patch_refl, vr_light *lights);
/* The faces the generator's patches were chopped from (rectangles, in
 * emission order), as vr_make_patches input with lightmap vectors of
 * 1/luxel_size luxels per unit along the rectangle's edges; out may be NULL
 * (count only). */
int vr_gen_faces(const vr_gen *g, float luxel_size, vr_face *out, int64_t *n_faces);
/* The generator's room grid: rooms span [xs[c], xs[c+1]] x [ys[r], ys[r+1]]
 * (wall centre lines; a room scene is one cell 0..1024); room r * nx + c.
 * xs / ys may be NULL (counts only: *nx and *ny are the cell counts, the
 * arrays hold one more). */
int vr_gen_grid(const vr_gen *g, double *xs, int32_t *nx, double *ys, int32_t *ny);
int vr_gen_destroy(vr_gen *g);

#ifdef __cplusplus
}
#endif
#endif
