// Kernel launchers (defined in the .hip files), called by capi.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/vrad_hip.h"
#include "trace_dev.h"

namespace vr {

constexpr int kStackDepth = 24;  // > max kd depth (21 + 1 leaf level)

// Launch-shape knobs (vr_set_tuning); every value gives identical results.
struct Tuning {
  int trace_lds_entries = 10;  // traversal-stack entries per lane kept in LDS: 2/4/6/8/10/12/16/24
  int trace_rays_per_lane = 1; // rays per lane of the visibility kernels (1, 2, 4, 8): a wave
                               // covers 64 x this many rays
  int trace_xcd_group = 256;   // TestLine, closest hit, MakeTransfers: workgroups dealt to the
                               // XCDs in groups of this many (0 = one contiguous range each)
  int bounce_waves = 8192;     // persistent waves of the wave-per-row gather
  // Bounce gather kernel and its shape.  Unlike the knobs above these fix the
  // FP32 summation order (the CPU restatement follows the same settings).
  int bounce_kernel = 2;       // 0 = LDS-tiled, 1 = wave per row, 2 = LDS stream
  int bounce_tile_cols = 4096; // 2048 / 4096 patches per LDS tile
  int bounce_group = 0;        // 8 / 16 lanes per row segment, 0 = by precision
  int bounce_passes = 0;       // passes of a segment loaded ahead of use, 0 = by precision
  int bounce_prefetch = 0;     // steps/passes the loads run ahead, 0 = by kernel (3 / 1)
  // rows dealt to the gather's 1024-row chunks: 0 = a stride through the
  // longest-first order of all rows (every chunk the same mix), 1 = chunk c
  // takes the c-th run of consecutive rows, longest-first inside the run (a
  // chunk's rows share their column tiles when neighbouring patches see the
  // same patches, as MakeTransfers rows do), 2 = placement 0's chunks with
  // their rows dealt to the waves in row-index order (each wave's rows spread
  // over the scene, so the waves' work per column tile evens out)
  int bounce_placement = 0;
  // kd-tree construction in vr_scene_build (the same tree either way)
  int kd_build = 2;                 // 0 = host threads, 1 = GPU, 2 = GPU from a size on
  int kd_build_device_min = 2048;   // triangles at which kd_build = 2 uses the GPU
  // patch generation in vr_make_patches (the same patches either way)
  int patch_build = 1;              // 0 = a thread per face, 1 = level-parallel
};
Tuning &tuning();

// ---- patch generation (patches.hip) ----
// MakePatchForFace + SubdividePatches on the GPU; *n_out = patch count,
// written to out when cap allows (out == NULL: count only).
// vnorm (nullable): PairEdges' per-vertex normals, VR_FACE_MAX_POINTS x 3 per
// face; children then take GetPhongNormal at their centre
int make_patches_device(const vr_face *faces, int64_t n_faces, float max_chop, float min_chop,
                        const float *vnorm, vr_patch *out, int64_t cap, int64_t *n_out,
                        hipStream_t st);

// ---- kd-tree construction on the device (kdbuild.hip) ----
struct KdTree;
// The host builder's tree (build_kdtree) built on the GPU on stream st and
// returned in host memory.
int build_kdtree_device(const float *v9, int64_t n, hipStream_t st, KdTree *out);

struct DevRays {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
};
struct DevSegs {
  const float *sx, *sy, *sz, *ex, *ey, *ez;
};
struct DevPatches {
  const float *ox, *oy, *oz, *nx, *ny, *nz, *area;
};

// ---- visibility (trace.hip) ----
void launch_trace_closest(const DevScene &sc, const DevRays &r, int64_t n, int32_t skip,
                          int32_t *hit, float *hit_t, float *hit_n, int path, hipStream_t st);
void launch_test_line(const DevScene &sc, const DevSegs &s, int64_t n, int32_t skip,
                      uint8_t *vis, uint8_t *sky, int path, hipStream_t st);

// ---- transfers (transfers.hip) ----
void launch_pair_cull(const DevPatches &p, int64_t np, int64_t i0, int64_t i1, int2 *pairs,
                      uint32_t *count, uint32_t cap, hipStream_t st);
void launch_pair_trace(const DevScene &sc, const DevPatches &p, const int2 *pairs,
                       const uint32_t *count, uint32_t cap, float eps, int32_t skip,
                       uint32_t *bits, int64_t words, int path, hipStream_t st);
// One scene's candidate pairs in a many-scene visibility launch: blocks
// [blk_begin, blk_begin + ceil(count / 64)) of the launch trace pairs[0, count)
// on sc with patches p, setting bits of the scene's N x N matrix (words per row).
struct TraceJob {
  DevScene sc;
  DevPatches p;
  const int2 *pairs;
  uint32_t *bits;
  int64_t words;
  int64_t count;
  int64_t blk_begin;
};
void launch_pair_trace_jobs(const TraceJob *jobs, int n_jobs, int64_t total_blocks, float eps,
                            int32_t skip, int lds_entries, hipStream_t st);
void launch_row_count(const uint32_t *bits, int64_t np, int64_t words, int32_t *cnt,
                      hipStream_t st);
// exclusive scan of padded lengths (multiple of 8) into rp[0..n]
void launch_row_scan(const int32_t *cnt, int64_t n, int64_t *rp, hipStream_t st);
void launch_row_fill(const uint32_t *bits, int64_t np, int64_t words, const DevPatches &p,
                     const int64_t *rp, const int32_t *cnt, int32_t *col, float *coef,
                     hipStream_t st);
void launch_row_normalize(int64_t n, const int64_t *rp, float *coef, double *row_sum,
                          int normalize, hipStream_t st);

// ---- luxel stage (luxels.hip) ----
void launch_face_extents(const vr_face *faces, int64_t n, const float *lux_ofs,
                         vr_face_lightmap *out, hipStream_t st);
// blocks: (face, first luxel) per 256-thread workgroup; leaf_ofs / leaf_idx:
// each face's leaf patches in patch order
void launch_face_luxels(const vr_face *faces, const float *lux_ofs, const vr_face_lightmap *lm,
                        const int2 *blocks, int64_t n_blocks, const int32_t *leaf_ofs,
                        const int32_t *leaf_idx, const vr_patch *patches,
                        const float *patch_value, float *pos, float *nrm, int32_t *lpatch,
                        float *value, hipStream_t st);

// the lighting lump: per face with luxels, entry at[f] = its average colour,
// then its luxels as ColorRGBExp32
void launch_lighting_lump(const float *lux, const vr_face_lightmap *lm, const int64_t *at,
                          int64_t n_faces, uint32_t *lump, hipStream_t st);

// ---- BSP clusters and PVS (pvs.hip) ----
struct DevBsp {
  const float4 *planes;         // normal xyz, dist
  const int *nodes;             // planenum, children[0], children[1]
  const int16_t *leaf_cluster;
};
// A pair of patch groups that see each other: |a| virtual rows of |b|
// columns, cpr blocks per row; diag = the same group (columns > row only).
struct PvsEntry {
  int32_t a_off, a_n, b_off, b_n, cpr, diag;
};
void launch_point_cluster(const DevBsp &b, const float *xyz, int64_t n, int16_t *out,
                          hipStream_t st);
void launch_patch_cluster(const DevBsp &b, const vr_patch *p, int64_t n, int32_t *out,
                          hipStream_t st);
void launch_decompress_vis(const uint8_t *vis, int64_t vis_len, const int32_t *ofs, int32_t nc,
                           uint8_t *rows, unsigned long long *overrun, uint32_t *err,
                           hipStream_t st);
int pvs_block_cols();  // columns one block of the PVS pair cull covers
void launch_pair_cull_pvs(const DevPatches &p, const int32_t *mem, const PvsEntry *tab,
                          const int64_t *blk_pre, int32_t n_tab, int64_t blk0, int64_t n_blk,
                          int2 *pairs, uint32_t *count, uint32_t cap, hipStream_t st);

// ---- bounce (bounce.hip) ----
struct DevXfers {
  int64_t n;
  const int64_t *rp;       // padded row starts (multiple of 8), n + 1
  const int32_t *col;
  const void *coef;        // float or uint16 (f16 / bf16)
  const float *scale;      // per-row scale for 16-bit storage
  const int32_t *order;    // rows sorted by length, longest first
  int kind;                // VR_COEF_*
};
struct DevTiles {
  const int32_t *tp;       // n x (nt + 1) row positions of the tile boundaries
  int nt, cols, group;
  int ostride;             // bytes per patch in the LDS emit tile: 12 (tiled), 16 (stream)
  int64_t n_chunks;        // workgroups of 1024 rows
  // tile-blocked copy: block (c, t) holds the segments of chunk c's rows for
  // column tile t back to back in owner order u = 64 wave + lane, each padded
  // to a multiple of G with zero coefficients; an entry's column is stored as
  // its 16-bit byte offset in the LDS emit tile, ostride (col - t C), whose
  // bit 0 (always 0 in an offset) is set on the first entry of each segment
  const uint16_t *bidx;
  const void *bcoef;
  const int64_t *bstart;   // n_chunks * nt block starts
  const int32_t *boff;     // n_chunks * nt blocks of 1025 segment offsets, stride 1028
};
struct BounceState {       // device-resident bounce bookkeeping
  double cum;
  int32_t converged;
  int32_t k_done;
};
// One transfer set's bounce in a batched launch of the LDS-tiled gather and
// the FP64 energy.  The bounce index is state->k_done: the gather reads
// emit[k & 1] and writes emit[(k + 1) & 1]; every kernel returns at once for
// a job that has converged or done max_bounces.
struct BounceJob {
  DevXfers x;
  DevTiles tl;
  float *emit[2];
  float *total;
  const float *refl, *area;
  double *e_row, *scratch, *energy;
  BounceState *state;
  int64_t chunk_begin;     // first gather workgroup of the job
  int64_t eblk_begin;      // first stage-1 energy block of the job
  int64_t e_blocks;        // ceil(n / 2048) stage-1 blocks (n <= 2048^2)
  int32_t max_bounces;
  double stop_frac;
};
constexpr int kMaxTilesPerJob = 1024;  // block starts of a chunk cached in LDS
// Chunks of a transfer set, at most 1024 rows each (one gather workgroup,
// whose LDS takes a whole CU, so kChunkTarget = 256 of them run at once):
// up to kChunkTarget full chunks, min(kChunkTarget, ceil(n / 256)) chunks
// (at least 256 rows each while that stays under 1024); beyond, a whole
// multiple of kChunkTarget, so the launch runs in full rounds.  A constant
// (not the device's CU count) because the stream gather's summation order
// depends on the chunking.
constexpr int64_t kChunkTarget = 256;
__host__ __device__ inline int64_t gather_chunks(int64_t n) {
  if (n <= 0) return 0;
  constexpr int64_t round_rows = kChunkTarget * 1024;
  if (n > round_rows) return kChunkTarget * ((n + round_rows - 1) / round_rows);
  const int64_t b = (n + 255) / 256;
  return b < kChunkTarget ? b : kChunkTarget;
}
// Rank (in the longest-first row order, within chunk c: row order[c + rank *
// n_chunks]) of owner slot (wave w, lane j): serpentine over the 16 waves, so
// each wave's 64 rows sum to about the same length.
__host__ __device__ inline int64_t slot_rank(int w, int j) {
  return (int64_t)j * 16 + ((j & 1) ? 15 - w : w);
}
// Gather kernel and shape: kernel 0 = LDS-tiled (group, passes, prefetch),
// 2 = stream (prefetch only); cols = tile width.
struct GatherShape {
  int kernel, cols, group, passes, prefetch, placement;
};
void launch_bounce(const DevXfers &x, const float *emit_in, float *emit_out, float *total,
                   const float *refl, const float *area, double *e_row, uint32_t *row_counter,
                   const BounceState *state, int path, hipStream_t st);
bool bounce_gather_variant(const GatherShape &sh);
// one bounce of every job: gather (jobs share precision and shape), then the
// energy tree and stop test of each job
void launch_bounce_jobs(int kind, const GatherShape &sh, const BounceJob *jobs, int n_jobs,
                        int64_t total_chunks, int64_t total_eblocks, hipStream_t st);
void launch_gather_jobs(int kind, const GatherShape &sh, const BounceJob *jobs, int n_jobs,
                        int64_t total_chunks, hipStream_t st);
void launch_energy_jobs(const BounceJob *jobs, int n_jobs, int64_t total_eblocks, hipStream_t st);
// total = direct; emit = direct * refl (3n floats, no aliasing)
void launch_emit_init(int64_t n, const float *direct, const float *refl, float *emit,
                      float *total, hipStream_t st);
void launch_tile_ptrs(int64_t n, const int64_t *rp, const int32_t *cnt, const int32_t *col,
                      int cols, int nt, int32_t *tp, hipStream_t st);
// blocked layout: offsets (boff) and per-block totals, then the copy
void launch_block_offsets(const DevXfers &x, const DevTiles &tl, int G, int64_t *btot,
                          int32_t *boff, hipStream_t st);
void launch_block_fill(const DevXfers &x, const DevTiles &tl, int G, int coef_bytes,
                       uint16_t *bidx, void *bcoef, hipStream_t st);
// energy of the bounce just gathered; updates state / energy[k]
void launch_energy(const double *e_row, int64_t n, double *scratch, double *energy, int k,
                   double stop_frac, BounceState *state, hipStream_t st);
// columns: 0 = uniform without replacement (bitmaps: synthetic_waves(n) x
// synthetic_bitmap_words(n) zeroed words, left zeroed), 1 = one per stratum
int64_t synthetic_bitmap_words(int64_t n);
int64_t synthetic_waves(int64_t n);
void launch_synthetic_fill(int64_t n, const int64_t *rp, const int32_t *len, uint64_t seed,
                           int columns, uint32_t *bitmaps, int32_t *col, float *coef,
                           hipStream_t st);
void launch_encode16(int64_t n, const int64_t *rp, const float *coef32, uint16_t *coef16,
                     float *scale, int kind, hipStream_t st);
void launch_decode16(int64_t total, const uint16_t *coef16, float *out, int kind,
                     hipStream_t st);
// grid-stride 16-byte loads over `bytes` (a multiple of 16); one float per
// thread into out (blocks x 256)
void launch_read_probe(const void *in, int64_t bytes, float *out, int blocks, hipStream_t st);

// ---- direct light (light.hip) ----
struct DevSamples {
  const float *px, *py, *pz, *nx, *ny, *nz;
};
// lds_entries: traversal-stack entries per lane in LDS (2/4/8/12/16/24)
void launch_gather_light(const DevScene &sc, const DevSamples &s, int64_t n,
                         const vr_light *lights, int32_t n_lights, const float *sky_dirs,
                         float eps, int32_t skip, float *rgb, int path, int lds_entries,
                         hipStream_t st);
// One scene's samples in a many-scene direct-light launch: blocks
// [blk_begin, blk_begin + ceil(n / 64)) gather samples [0, n) of s on sc with
// lights[0, n_lights), writing rgb (3 n).
struct GatherJob {
  DevScene sc;
  DevSamples s;
  const vr_light *lights;
  int32_t n_lights;
  float *rgb;
  int64_t n;
  int64_t blk_begin;
};
void launch_gather_jobs(const GatherJob *jobs, int n_jobs, int64_t total_blocks,
                        const float *sky_dirs, float eps, int32_t skip, int lds_entries,
                        hipStream_t st);
// CanLeafTraceToSky per leaf (n waves; n < 2^32): out[i] = 0 / 1
void launch_leaf_sky(const DevScene &sc, const int16_t *mins, const int16_t *maxs, int64_t n,
                     const float *sky_dirs, uint8_t *out, hipStream_t st);

}  // namespace vr
