// BounceLight: sparse CSR gather add = F . emit, total += add,
// emit' = add * refl, and the FP64 per-bounce energy with its stop test.
//
// Optimised gather: one wave per row, rows taken longest-first from an
// atomic counter; lane l owns row positions p with (p / G) % 64 == l
// (G = 4 for FP32, 8 for FP16/BF16: one 16-byte coefficient load per lane),
// FMA-accumulates them in increasing p, then xor-butterflies 32..1.  That
// order is part of the result and is reproduced by the CPU restatement.
// Rows are padded to a multiple of 8 with (col = row, coef = 0), which adds
// exact zeros.  Plain gather: one thread per row, sequential order.
#include <hip/hip_fp16.h>

#include "kernels.h"

namespace vr {

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
__device__ __forceinline__ float f16_to_f32(uint16_t h) {
  return __half2float(__ushort_as_half(h));
}

template <int KIND>
__device__ __forceinline__ void gather_chunk(const int32_t *__restrict__ col, const void *coefp,
                                             int64_t k, const float *__restrict__ emit, float &r,
                                             float &g, float &b) {
  if constexpr (KIND == VR_COEF_F32) {
    const int4 c = *reinterpret_cast<const int4 *>(col + k);
    const float4 w = *reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(coefp) + k);
    const int cc[4] = {c.x, c.y, c.z, c.w};
    const float ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float *e = emit + 3 * (int64_t)cc[j];
      r = __builtin_fmaf(ww[j], e[0], r);
      g = __builtin_fmaf(ww[j], e[1], g);
      b = __builtin_fmaf(ww[j], e[2], b);
    }
  } else {
    const int4 c0 = *reinterpret_cast<const int4 *>(col + k);
    const int4 c1 = *reinterpret_cast<const int4 *>(col + k + 4);
    const uint4 w = *reinterpret_cast<const uint4 *>(reinterpret_cast<const uint16_t *>(coefp) + k);
    const int cc[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint16_t h = (uint16_t)(ww[j >> 1] >> (16 * (j & 1)));
      const float wf = KIND == VR_COEF_F16 ? f16_to_f32(h) : bf16_to_f32(h);
      const float *e = emit + 3 * (int64_t)cc[j];
      r = __builtin_fmaf(wf, e[0], r);
      g = __builtin_fmaf(wf, e[1], g);
      b = __builtin_fmaf(wf, e[2], b);
    }
  }
}

// One counter increment per wave, returned to every lane as a scalar.  All
// lanes issue the atomic (lane 0 adds 1, the others 0): a lane-0-only branch
// in the loop header lets the optimiser send the other lanes straight back to
// the broadcast, where they would read a stale value and never leave the loop.
__device__ __forceinline__ uint32_t wave_fetch_add(uint32_t *ctr, int lane) {
  const uint32_t v = atomicAdd(ctr, lane == 0 ? 1u : 0u);
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

template <int KIND>
__global__ __launch_bounds__(256) void k_bounce_wave(DevXfers x, const float *__restrict__ emit_in,
                                                     float *__restrict__ emit_out,
                                                     float *__restrict__ total,
                                                     const float *__restrict__ refl,
                                                     const float *__restrict__ area,
                                                     double *__restrict__ e_row,
                                                     uint32_t *__restrict__ row_counter,
                                                     const BounceState *__restrict__ state) {
  if (state->converged) return;
  constexpr int G = KIND == VR_COEF_F32 ? 4 : 8;
  const int lane = threadIdx.x & 63;
  for (;;) {
    const uint32_t slot = wave_fetch_add(row_counter, lane);
    if ((int64_t)slot >= x.n) break;
    const int64_t i = x.order[slot];
    const int64_t s = x.rp[i], e = x.rp[i + 1];
    float r = 0.0f, g = 0.0f, b = 0.0f;
    for (int64_t k = s + (int64_t)lane * G; k < e; k += 64 * G)
      gather_chunk<KIND>(x.col, x.coef, k, emit_in, r, g, b);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      r = r + __shfl_xor(r, off);
      g = g + __shfl_xor(g, off);
      b = b + __shfl_xor(b, off);
    }
    if (lane == 0) {
      if constexpr (KIND != VR_COEF_F32) {
        const float sc = x.scale[i];
        r = r * sc;
        g = g * sc;
        b = b * sc;
      }
      total[3 * i + 0] = total[3 * i + 0] + r;
      total[3 * i + 1] = total[3 * i + 1] + g;
      total[3 * i + 2] = total[3 * i + 2] + b;
      emit_out[3 * i + 0] = r * refl[3 * i + 0];
      emit_out[3 * i + 1] = g * refl[3 * i + 1];
      emit_out[3 * i + 2] = b * refl[3 * i + 2];
      e_row[i] = (double)area[i] * (((double)r + (double)g) + (double)b);
    }
  }
}

__global__ __launch_bounds__(256) void k_bounce_plain(DevXfers x, const float *__restrict__ emit_in,
                                                      float *__restrict__ emit_out,
                                                      float *__restrict__ total,
                                                      const float *__restrict__ refl,
                                                      const float *__restrict__ area,
                                                      double *__restrict__ e_row,
                                                      const BounceState *__restrict__ state) {
  if (state->converged) return;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= x.n) return;
  const float *coef = reinterpret_cast<const float *>(x.coef);
  float r = 0.0f, g = 0.0f, b = 0.0f;
  for (int64_t k = x.rp[i]; k < x.rp[i + 1]; ++k) {
    const float w = coef[k];
    const float *e = emit_in + 3 * (int64_t)x.col[k];
    r = __builtin_fmaf(w, e[0], r);
    g = __builtin_fmaf(w, e[1], g);
    b = __builtin_fmaf(w, e[2], b);
  }
  total[3 * i + 0] = total[3 * i + 0] + r;
  total[3 * i + 1] = total[3 * i + 1] + g;
  total[3 * i + 2] = total[3 * i + 2] + b;
  emit_out[3 * i + 0] = r * refl[3 * i + 0];
  emit_out[3 * i + 1] = g * refl[3 * i + 1];
  emit_out[3 * i + 2] = b * refl[3 * i + 2];
  e_row[i] = (double)area[i] * (((double)r + (double)g) + (double)b);
}

void launch_bounce(const DevXfers &x, const float *emit_in, float *emit_out, float *total,
                   const float *refl, const float *area, double *e_row, uint32_t *row_counter,
                   const BounceState *state, int path, hipStream_t st) {
  if (x.n <= 0) return;
  if (path == VR_PATH_PLAIN) {
    hipLaunchKernelGGL(k_bounce_plain, dim3((unsigned)((x.n + 255) / 256)), dim3(256), 0, st, x,
                       emit_in, emit_out, total, refl, area, e_row, state);
    return;
  }
  // persistent: enough waves to fill 256 CUs several times over
  const int64_t tw = tuning().bounce_waves;
  int64_t waves = x.n < tw ? x.n : tw;
  unsigned blocks = (unsigned)((waves + 3) / 4);
  if (x.kind == VR_COEF_F32)
    hipLaunchKernelGGL(k_bounce_wave<VR_COEF_F32>, dim3(blocks), dim3(256), 0, st, x, emit_in,
                       emit_out, total, refl, area, e_row, row_counter, state);
  else if (x.kind == VR_COEF_F16)
    hipLaunchKernelGGL(k_bounce_wave<VR_COEF_F16>, dim3(blocks), dim3(256), 0, st, x, emit_in,
                       emit_out, total, refl, area, e_row, row_counter, state);
  else
    hipLaunchKernelGGL(k_bounce_wave<VR_COEF_BF16>, dim3(blocks), dim3(256), 0, st, x, emit_in,
                       emit_out, total, refl, area, e_row, row_counter, state);
}

// ---- LDS-tiled gather ------------------------------------------------------
// The emit vector is cut into column tiles of C patches that are staged in
// LDS.  A workgroup of 16 waves owns up to 1024 rows: chunk c takes rows
// order[c + k * n_chunks], a stride through the longest-first order, so every
// chunk and every wave gets the same mix of lengths; slot j of wave w is
// chunk row k = 16 j + w.  For each tile the wave's 64 lanes form 64/GS
// groups of GS lanes; in step s group gi takes slot j = s (64/GS) + gi, i.e.
// that row's segment (its entries with columns in the tile, a contiguous
// range because columns are sorted).  Lane l of the group owns the G-entry
// chunks at segment positions G l + G GS i, FMA-accumulates them in order,
// the group butterflies (GS/2 .. 1), and lane s of the group, which owns slot
// j, adds the segment sum to its row total; tiles in order.
//
// Coefficients come from a tile-blocked copy: block (c, t) holds the segments
// of chunk c for tile t back to back in slot order (u = 64 w + j), each
// padded to a multiple of G with (offset 0, coef 0), so every load is a whole
// vector and the loads of one step cover a few contiguous lines.  Columns are
// stored as 16-bit byte offsets into the LDS emit tile.
//
// Pipeline (per wave): the loads of step s + D are issued before step s is
// gathered, into a ring of D + 1 register slots, and every load is issued
// unconditionally (a lane with nothing to load reads the block's first chunk
// and ignores it), so the in-order memory counter waits only for the step it
// needs.  The wave's segment offsets of tile t + 1 are loaded into registers
// during tile t, so tile t + 1's first loads leave before the tile barrier.
// The emit tile t + 1 is copied into the second LDS buffer by DMA at the start
// of tile t.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x3 __attribute__((ext_vector_type(3)));  // 16-byte aligned
// vector load through a global-address-space pointer, so it is issued as a
// global (not flat) load: flat loads also count against lgkmcnt, and the LDS
// gather's waits would then wait for them
template <class V>
__device__ __forceinline__ V load_global(const char *p) {
  return *reinterpret_cast<const __attribute__((address_space(1))) V *>(
      (const __attribute__((address_space(1))) char *)p);
}
// the same with the non-temporal hint: data read once, not to displace the
// emit tiles that every workgroup re-reads from L2
template <class V>
__device__ __forceinline__ V load_global_nt(const char *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const __attribute__((address_space(1))) V *>(
      (const __attribute__((address_space(1))) char *)p));
}

// One G-entry chunk of the tile-blocked copy: 16-bit LDS offsets and the
// coefficients, loaded with 32-bit offsets from wave-uniform block bases.
template <int KIND>
struct BChunk;
template <>
struct BChunk<VR_COEF_F32> {
  uint2 o;
  float4 w;
  __device__ void load(const char *ib, const char *cb, uint32_t q) {
    const u32x2 a = load_global<u32x2>(ib + 2u * q);
    const f32x4 b = load_global<f32x4>(cb + 4u * q);
    o = make_uint2(a.x, a.y);
    w = make_float4(b.x, b.y, b.z, b.w);
  }
  __device__ void load_nt(const char *ib, const char *cb, uint32_t q) {
    const u32x2 a = load_global_nt<u32x2>(ib + 2u * q);
    const f32x4 b = load_global_nt<f32x4>(cb + 4u * q);
    o = make_uint2(a.x, a.y);
    w = make_float4(b.x, b.y, b.z, b.w);
  }
  // bit 0 of an entry's offset marks a segment's first entry (not an address bit)
  __device__ uint32_t off(int m) const {
    const uint32_t d = m < 2 ? o.x : o.y;
    return (m & 1) ? (d >> 16) : (d & 0xfffeu);
  }
  __device__ bool head() const { return (o.x & 1u) != 0; }
  __device__ float coef(int m) const { return m == 0 ? w.x : m == 1 ? w.y : m == 2 ? w.z : w.w; }
};
template <int KIND>
struct BChunk {
  uint4 o, h;
  __device__ void load(const char *ib, const char *cb, uint32_t q) {
    const u32x4 a = load_global<u32x4>(ib + 2u * q);
    const u32x4 b = load_global<u32x4>(cb + 2u * q);
    o = make_uint4(a.x, a.y, a.z, a.w);
    h = make_uint4(b.x, b.y, b.z, b.w);
  }
  __device__ void load_nt(const char *ib, const char *cb, uint32_t q) {
    const u32x4 a = load_global_nt<u32x4>(ib + 2u * q);
    const u32x4 b = load_global_nt<u32x4>(cb + 2u * q);
    o = make_uint4(a.x, a.y, a.z, a.w);
    h = make_uint4(b.x, b.y, b.z, b.w);
  }
  __device__ uint32_t off(int m) const {
    const uint32_t d = (m >> 1) == 0 ? o.x : (m >> 1) == 1 ? o.y : (m >> 1) == 2 ? o.z : o.w;
    return (m & 1) ? (d >> 16) : (d & 0xfffeu);
  }
  __device__ bool head() const { return (o.x & 1u) != 0; }
  __device__ float coef(int m) const {
    const uint32_t wd = (m >> 1) == 0 ? h.x : (m >> 1) == 1 ? h.y : (m >> 1) == 2 ? h.z : h.w;
    const uint16_t v = (uint16_t)(wd >> (16 * (m & 1)));
    return KIND == VR_COEF_F16 ? f16_to_f32(v) : bf16_to_f32(v);
  }
};

// A 4-entry chunk of 16-bit coefficients (the stream layout's chunk for every
// precision): 16-bit offsets and coefficients interleaved as one 16-byte
// record per chunk (4 offsets, then 4 coefficients) at ib + 4 q, q the
// chunk's first entry: one 128-bit load per chunk.
template <int KIND>
struct QChunk16 {
  uint2 o, h;
  __device__ void load_nt(const char *ib, const char *, uint32_t q) {
    const u32x4 v = load_global_nt<u32x4>(ib + 4u * q);
    o = make_uint2(v.x, v.y);
    h = make_uint2(v.z, v.w);
  }
  __device__ float coef(int m) const {
    const uint32_t wd = m < 2 ? h.x : h.y;
    const uint16_t v = (uint16_t)(wd >> (16 * (m & 1)));
    return KIND == VR_COEF_F16 ? f16_to_f32(v) : bf16_to_f32(v);
  }
};
template <int KIND>
struct StreamChunkT {
  using type = QChunk16<KIND>;
};
template <>
struct StreamChunkT<VR_COEF_F32> {
  using type = BChunk<VR_COEF_F32>;
};
template <int KIND>
using StreamChunk = typename StreamChunkT<KIND>::type;

// Exact xor-lane exchange inside 16-lane rows with DPP (no LDS traffic):
// quad_perm for 1 and 2, bank-masked row shifts for 4, row rotate for 8.
template <int OFF>
__device__ __forceinline__ float dpp_xor(float v) {
  const int x = __float_as_int(v);
  int y;
  if constexpr (OFF == 1) {
    y = __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  } else if constexpr (OFF == 2) {
    y = __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  } else if constexpr (OFF == 4) {
    y = __builtin_amdgcn_update_dpp(0, x, 0x104, 0xF, 0x5, false);  // row_shl:4, banks 0,2
    y = __builtin_amdgcn_update_dpp(y, x, 0x114, 0xF, 0xA, false);  // row_shr:4, banks 1,3
  } else {
    static_assert(OFF == 8, "xor offset");
    y = __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, true);  // row_ror:8
  }
  return __int_as_float(y);
}

// Row rotate by 4 (row_ror:4).  After an xor-8 stage every 16-lane row has
// period 8, so lane l's rotate partner holds the value of lane l ^ 4.
__device__ __forceinline__ float dpp_ror4(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x124, 0xF, 0xF, true));
}

// Butterfly GS/2 .. 1 over groups of GS lanes (each lane ends with the sum).
template <int GS>
__device__ __forceinline__ float group_sum(float v) {
  if constexpr (GS >= 16) {
    v = v + dpp_xor<8>(v);
    v = v + dpp_ror4(v);
  } else if constexpr (GS >= 8) {
    v = v + dpp_xor<4>(v);
  }
  v = v + dpp_xor<2>(v);
  v = v + dpp_xor<1>(v);
  return v;
}

// Asynchronous global -> LDS copy of 16-byte pieces (gfx950 LDS DMA): lane l
// of the wave writes its piece at lds_base + 16 l; the wave's pieces are
// consecutive in LDS.  Completion is tracked by vmcnt like any global load.
__device__ __forceinline__ void dma16(const void *src, void *lds_base) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)lds_base, 16, 0,
                                   0);
}

template <int KIND>
__device__ __forceinline__ void gather_bchunk(const BChunk<KIND> &ch, const char *te, float &r,
                                              float &g, float &b) {
  constexpr int G = KIND == VR_COEF_F32 ? 4 : 8;
#pragma unroll
  for (int m = 0; m < G; ++m) {
    const float *e = reinterpret_cast<const float *>(te + ch.off(m));
    const float w = ch.coef(m);
    r = __builtin_fmaf(w, e[0], r);
    g = __builtin_fmaf(w, e[1], g);
    b = __builtin_fmaf(w, e[2], b);
  }
}

// The same over 16-byte LDS records (offset = 16 patch): one 12-byte LDS
// read per entry.
template <int KIND>
__device__ __forceinline__ void gather_bchunk16(const BChunk<KIND> &ch, const char *te, float &r,
                                                float &g, float &b) {
  constexpr int G = KIND == VR_COEF_F32 ? 4 : 8;
#pragma unroll
  for (int m = 0; m < G; ++m) {
    const f32x3 e = *reinterpret_cast<const f32x3 *>(te + ch.off(m));
    const float w = ch.coef(m);
    r = __builtin_fmaf(w, e.x, r);
    g = __builtin_fmaf(w, e.y, g);
    b = __builtin_fmaf(w, e.z, b);
  }
}

template <int KIND, int P>
struct StepLoad {
  BChunk<KIND> ch[P];
  int q, sb;
};

template <int V>
struct IntC {
  static constexpr int value = V;
};

// Index of the job whose range [begin_j, begin_{j+1}) holds b (begin_0 = 0,
// ascending); `begin` is the member of BounceJob named by the pointer-to-member.
template <int64_t BounceJob::*begin>
__device__ __forceinline__ int find_job(const BounceJob *__restrict__ jobs, int n_jobs,
                                        int64_t b) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].*begin <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ bool job_done(const BounceJob &jb) {
  return jb.state->converged || jb.state->k_done >= jb.max_bounces;
}

// A row's gathered sum (ar, ag, ab; times the row scale for 16-bit
// coefficients): total += add, emit = add * refl, e_row = area * (r + g + b).
template <int KIND>
__device__ __forceinline__ void finish_row(const BounceJob &jb, int64_t row, float ar, float ag,
                                           float ab, float *__restrict__ emit_out) {
  if constexpr (KIND != VR_COEF_F32) {
    const float sc = jb.x.scale[row];
    ar = ar * sc;
    ag = ag * sc;
    ab = ab * sc;
  }
  float *__restrict__ total = jb.total;
  const float *__restrict__ refl = jb.refl;
  total[3 * row + 0] = total[3 * row + 0] + ar;
  total[3 * row + 1] = total[3 * row + 1] + ag;
  total[3 * row + 2] = total[3 * row + 2] + ab;
  emit_out[3 * row + 0] = ar * refl[3 * row + 0];
  emit_out[3 * row + 1] = ag * refl[3 * row + 1];
  emit_out[3 * row + 2] = ab * refl[3 * row + 2];
  jb.e_row[row] = (double)jb.area[row] * (((double)ar + (double)ag) + (double)ab);
}

// One workgroup of 16 waves per chunk of a job; emit tiles (r, g, b per
// patch, as in global memory) are copied into two LDS buffers by DMA, tile
// t + 1 while tile t is gathered.  The copy reads emit in whole 16-byte
// pieces, so emit must stay readable up to the next 16-byte boundary past its
// 12 n bytes (the bounce buffers are padded to 256 bytes).
template <int KIND, int C, int GS, int P, int D>
__global__ __launch_bounds__(1024) void k_bounce_tiled(const BounceJob *__restrict__ jobs,
                                                       int n_jobs) {
  const int jid = find_job<&BounceJob::chunk_begin>(jobs, n_jobs, blockIdx.x);
  const BounceJob &jb = jobs[jid];
  if (job_done(jb)) return;
  const int kb = jb.state->k_done;
  const DevXfers x = jb.x;
  const DevTiles tl = jb.tl;
  const float *__restrict__ emit_in = jb.emit[kb & 1];
  float *__restrict__ emit_out = jb.emit[(kb + 1) & 1];
  constexpr int G = KIND == VR_COEF_F32 ? 4 : 8;
  constexpr int T = 1024;              // threads
  constexpr int RS = 1028;             // stride of a block's segment offsets
  constexpr int NG = 64 / GS;          // rows per step (one per lane group)
  constexpr int NS = GS;               // steps per tile: 64 slots / NG
  constexpr int PASS = GS * G;         // entries per pass of a group
  constexpr int EP = 3 * C / 4;        // 16-byte pieces of a full emit tile
  constexpr int ER = (EP + T - 1) / T; // copy rounds per tile
  static_assert(NS % (D + 1) == 0, "steps rotate through D + 1 load slots");
  __shared__ __attribute__((aligned(16))) float s_e[2][4 * T * ER];
  __shared__ int64_t s_bs[kMaxTilesPerJob];  // this chunk's block starts
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int gi = lane / GS, gl = lane % GS;
  const int64_t c = (int64_t)blockIdx.x - jb.chunk_begin;
  const int nt = tl.nt;
  for (int t = (int)threadIdx.x; t < nt; t += T) s_bs[t] = tl.bstart[c * nt + t];
  __syncthreads();
  const int j_own = gl * NG + gi;  // the slot whose row total this lane keeps
  const int64_t oi = c + slot_rank(wave, j_own) * tl.n_chunks;
  const int64_t row = oi < x.n ? (int64_t)x.order[oi] : -1;
  const char *bidx = reinterpret_cast<const char *>(tl.bidx);
  const char *bcoef = reinterpret_cast<const char *>(tl.bcoef);
  const int cbytes = KIND == VR_COEF_F32 ? 4 : 2;
  float ar = 0.0f, ag = 0.0f, ab = 0.0f;

  // emit tile tn -> LDS buffer buf, every lane issuing ER pieces (lanes past
  // the tile's end copy its last piece again into their own slots)
  auto fetch_emit = [&](int tn, int buf) {
    const int64_t p1 = (int64_t)tn * C;
    const int np = (int)min((int64_t)C, x.n - p1);
    const int last = (12 * np + 15) / 16 - 1;
    const char *src = reinterpret_cast<const char *>(emit_in + 3 * p1);
#pragma unroll
    for (int m = 0; m < ER; ++m) {
      const int k = min((int)threadIdx.x + T * m, last);
      dma16(src + 16 * k, &s_e[buf][4 * (T * m + 64 * wave)]);
    }
  };
  // this lane's slot offsets (start, end) in block (c, tn)
  auto slot_offsets = [&](int tn, int &sa, int &sb) {
    const int32_t *bo = tl.boff + (c * nt + tn) * (int64_t)RS + 64 * wave + lane;
    sa = bo[0];
    sb = bo[1];
  };
  // loads of step s of a tile with block start bs, slot offsets (oa, ob)
  auto issue = [&](int s, int64_t bs, int oa, int ob, StepLoad<KIND, P> &L) {
    const int j = s * NG + gi;
    const int sa = __shfl(oa, j);
    L.sb = __shfl(ob, j);
    L.q = sa + gl * G;
    const char *ib = bidx + 2 * bs;
    const char *cb = bcoef + cbytes * bs;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const int qk = L.q + k * PASS;
      L.ch[k].load(ib, cb, qk < L.sb ? (uint32_t)qk : 0u);
    }
  };

  int oa, ob;
  slot_offsets(0, oa, ob);
  int64_t bs = s_bs[0];
  StepLoad<KIND, P> ring[D + 1];  // step s of a tile lives in slot s % (D + 1)
  fetch_emit(0, 0);
#pragma unroll
  for (int k = 0; k < D; ++k) issue(k, bs, oa, ob, ring[k]);
  __builtin_amdgcn_s_waitcnt(0);  // the first tile's copy (fewer loads follow it)

  // Before a tile's barrier each wave needs only its own copy of the tile to
  // have landed (and its LDS reads of the previous tile done).  The copy was
  // issued before the 2 P NS chunk loads of the previous tile's steps, and the
  // memory counter retires in order, so vmcnt <= min(63, 2 P NS) implies it.
  constexpr int WAIT_N = 2 * P * NS < 63 ? 2 * P * NS : 63;
  constexpr int WAIT_IMM = (WAIT_N & 15) | ((WAIT_N >> 4) << 14) | (7 << 4);  // lgkmcnt 0
  auto tile = [&](auto bufc, int t) {
    constexpr int buf = decltype(bufc)::value;
    __builtin_amdgcn_s_waitcnt(WAIT_IMM);
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();  // every wave's copy; tile t - 1 is done
    __asm__ volatile("" ::: "memory");
    const int tn = min(t + 1, nt - 1);
    fetch_emit(tn, buf ^ 1);       // on the last tile a harmless re-copy
    __asm__ volatile("" ::: "memory");  // the step loads below stay after the copy
    int oa_n, ob_n;
    slot_offsets(tn, oa_n, ob_n);
    const int64_t bs_n = s_bs[tn];
    const char *te = reinterpret_cast<const char *>(s_e[buf]);

    // straight-line steps: every wait on the in-order memory counter is exact
#pragma unroll
    for (int st = 0; st < NS; ++st) {
      StepLoad<KIND, P> &cur = ring[st % (D + 1)];
      StepLoad<KIND, P> &ahead = ring[(st + D) % (D + 1)];
      if (st + D < NS) issue(st + D, bs, oa, ob, ahead);
      else issue(st + D - NS, bs_n, oa_n, ob_n, ahead);  // next tile's first steps
      float r = 0.0f, g = 0.0f, b = 0.0f;
#pragma unroll
      for (int k = 0; k < P; ++k)
        if (cur.q + k * PASS < cur.sb) gather_bchunk<KIND>(cur.ch[k], te, r, g, b);
      for (int q = cur.q + P * PASS; q < cur.sb; q += PASS) {
        BChunk<KIND> more;
        more.load(bidx + 2 * bs, bcoef + cbytes * bs, (uint32_t)q);
        gather_bchunk<KIND>(more, te, r, g, b);
      }
      r = group_sum<GS>(r);
      g = group_sum<GS>(g);
      b = group_sum<GS>(b);
      if (gl == st) {
        ar = ar + r;
        ag = ag + g;
        ab = ab + b;
      }
    }
    oa = oa_n;
    ob = ob_n;
    bs = bs_n;
  };
  for (int t = 0; t < nt; t += 2) {
    tile(IntC<0>{}, t);
    if (t + 1 < nt) tile(IntC<1>{}, t + 1);
  }
  __builtin_amdgcn_s_waitcnt(0);  // the last (unused) copies and loads
  if (row >= 0) finish_row<KIND>(jb, row, ar, ag, ab, emit_out);
}

// ---- stream gather ------------------------------------------------------------
// Record slot of patch p (tile-relative) in the stream gather's LDS emit
// tile: p with its low 4 bits xor'ed with bits 3-6.  A permutation inside
// each aligned group of 16 records.  A pass's lane l reads its chunk entry m
// at stream entry 8 l + m, so a run of consecutive columns (MakeTransfers
// rows are mostly such runs) would put lanes l and l + 2 on the same bank
// group; after the xor any 16 consecutive lanes of such a run read 16
// different record slots (8 lanes: 8 different slots mod 8).
__host__ __device__ inline int emit_slot(int p) { return p ^ ((p >> 3) & 15); }

// Lane source through DPP for the segmented scan (0 where the source lane is
// outside the row; the scan's conditions never use such a lane).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_src(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, true));
}

// Per-stage conditions of the segmented scan for a lane whose segment starts
// at lane s: the stage's source lane exists and lies in the segment.
struct ScanMasks {
  bool m1, m2, m4, m8, m15, m31;
};
__device__ __forceinline__ ScanMasks scan_masks(int lane, int s) {
  const int row0 = lane & ~15;
  const int sr = s > row0 ? s : row0;  // the segment's start within this 16-lane row
  ScanMasks k;
  k.m1 = lane - 1 >= sr;
  k.m2 = lane - 2 >= sr;
  k.m4 = lane - 4 >= sr;
  k.m8 = lane - 8 >= sr;
  k.m15 = (lane & 16) != 0 && s < row0;
  k.m31 = lane >= 32 && s < 32;
  return k;
}

// Segmented inclusive scan over the 64 lanes in a fixed stage order; a stage
// whose condition holds sets v = v + (source lane's v):
//   row shifts right by 1, 2, 4, 8 (source lane - d, same 16-lane row);
//   rows 1 and 3 take lane 15 of rows 0 and 2; rows 2 and 3 take lane 31.
// A segment's last lane ends with the sum of the segment's lanes.
__device__ __forceinline__ float seg_scan(float v, const ScanMasks &k) {
  float w;
  w = v + dpp_src<0x111, 0xF>(v);  // row_shr:1
  v = k.m1 ? w : v;
  w = v + dpp_src<0x112, 0xF>(v);  // row_shr:2
  v = k.m2 ? w : v;
  w = v + dpp_src<0x114, 0xF>(v);  // row_shr:4
  v = k.m4 ? w : v;
  w = v + dpp_src<0x118, 0xF>(v);  // row_shr:8
  v = k.m8 ? w : v;
  w = v + dpp_src<0x142, 0xF>(v);  // row_bcast:15 (used in rows 1 and 3)
  v = k.m15 ? w : v;
  w = v + dpp_src<0x143, 0xF>(v);  // row_bcast:31 (used in rows 2 and 3)
  v = k.m31 ? w : v;
  return v;
}

__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// One wave's share of one tile: its 64 owner segments back to back from entry
// a to e of the block (padded to G), read in passes of 128 G entries.
struct StreamMeta {
  int a, e, np;
  const char *ib, *cb;
};

// Stream-layout metadata of a G-entry chunk, kept in the low 4 bits of its
// entries' LDS offsets (offsets are multiples of 16):
//   entry 0: bit 0 start (first chunk of its segment), bit 1 last (last
//            chunk of its segment), bits 2-3 = s bits 0-1
//   entry 2: s bits 2-5, s = the lane of the segment's first chunk in this
//            pass (0 when the segment began in an earlier pass)
//   entry 1: slot bits 0-3, entry 3: slot bits 4-5 (the segment's owner
//            slot, 0..63)
__host__ __device__ inline uint32_t stream_chunk_meta(int m, bool start, bool last, int s,
                                                      int slot) {
  if (m == 0) return (start ? 1u : 0u) | (last ? 2u : 0u) | (uint32_t)((s & 3) << 2);
  if (m == 2) return (uint32_t)((s >> 2) & 15);
  if (m == 1) return (uint32_t)(slot & 15);
  if (m == 3) return (uint32_t)((slot >> 4) & 3);
  return 0u;
}

struct ChunkMeta {
  bool start, last;
  int s, slot;
};
template <class CH>
__device__ __forceinline__ ChunkMeta chunk_meta(const CH &ch) {
  ChunkMeta r;
  const uint32_t x = ch.o.x, y = ch.o.y;
  r.start = (x & 1u) != 0;
  r.last = (x & 2u) != 0;
  r.s = (int)(((x >> 2) & 3u) | ((y & 15u) << 2));
  r.slot = (int)(((x >> 16) & 15u) | (((y >> 16) & 3u) << 4));
  return r;
}

// Gather over 16-byte LDS records of a 4-entry stream-layout chunk
// (metadata bits masked off the offsets), continuing the sums r, g, b.
template <class CH>
__device__ __forceinline__ void gather_stream(const CH &ch, const char *te, float &r, float &g,
                                              float &b) {
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const uint32_t d = m < 2 ? ch.o.x : ch.o.y;
    const uint32_t off = (m & 1) ? (d >> 16) & 0xfff0u : d & 0xfff0u;
    const f32x4 e = *reinterpret_cast<const f32x4 *>(te + off);
    const float w = ch.coef(m);
    r = __builtin_fmaf(w, e.x, r);
    g = __builtin_fmaf(w, e.y, g);
    b = __builtin_fmaf(w, e.z, b);
  }
}

// The two chunks a lane reads in one pass (A = stream chunk 2 lane, B =
// chunk 2 lane + 1).
template <int KIND>
struct PassLoad {
  StreamChunk<KIND> a, b;
};

// Memory slot (in chunks from the pass start) of stream chunk c of a pass of
// nc chunks: the A chunks (even c) first, then the B chunks (odd c).
__host__ __device__ inline int stream_pass_slot(int c, int nc) {
  return (c & 1) ? ((nc + 1) >> 1) + (c >> 1) : (c >> 1);
}

// Stream gather: the same workgroups, tile-blocked copy and LDS emit tiles as
// the LDS-tiled kernel, but each wave reads its 64 segments of a tile as one
// stream of G-entry chunks, in passes of 128 chunks: lane l takes chunks
// 2 l (A) and 2 l + 1 (B), so every lane of a pass carries entries whatever
// the segment lengths.  Per pass and lane:
//   pA = fma chain over A's entries from 0;
//   v  = fma chain over B's entries from (B starts a segment ? 0 : pA), or pA
//        when B lies past the wave's last chunk;
//   S  = segmented inclusive scan of v over the lanes (the lane's last
//        chunk's segment, from lane s), carry = S of the lane before (0 for
//        lane 0);
//   when B starts a segment, A's segment ends here: it adds
//        (A starts a segment ? pA : carry + pA) to its slot;
//   when the lane's last chunk ends its segment, or it is lane 63, S is added
//        to that chunk's slot.
// Slot sums live in LDS (per wave, 64 x 16 bytes) and are added to in pass
// order; each slot gets at most one addition per pass.  Loads run D passes
// ahead; the first D passes of tile t + 1 are issued when tile t starts, into
// the other half of a two-tile register ring.  Emit tiles sit in LDS as
// 16-byte records (offsets 16 emit_slot(col - t C)); the next emit tile is
// loaded into registers when a tile starts and written to the other LDS
// buffer when it ends (no LDS DMA: the compiler waits for every pending load
// before an LDS read while a DMA to LDS is in flight).
template <int KIND, int C, int D>
__global__ __launch_bounds__(1024) void k_bounce_stream(const BounceJob *__restrict__ jobs,
                                                        int n_jobs) {
  const int jid = find_job<&BounceJob::chunk_begin>(jobs, n_jobs, blockIdx.x);
  const BounceJob &jb = jobs[jid];
  if (job_done(jb)) return;
  const int kb = jb.state->k_done;
  const DevXfers x = jb.x;
  const DevTiles tl = jb.tl;
  const float *__restrict__ emit_in = jb.emit[kb & 1];
  float *__restrict__ emit_out = jb.emit[(kb + 1) & 1];
  constexpr int G = 4;                 // entries per chunk, every precision
  constexpr int PW = 128 * G;          // entries per pass of a wave
  constexpr int T = 1024;              // threads
  constexpr int RS = 1028;             // stride of a block's segment offsets
  constexpr int EPT = C / T;           // emit records copied per thread and tile
  constexpr int cbytes = KIND == VR_COEF_F32 ? 4 : 2;
  static_assert(C % T == 0 && 16 * C <= 65536, "16-byte LDS records with 16-bit offsets");
  __shared__ __attribute__((aligned(16))) float s_e[2][4 * C];  // (r, g, b, -) per patch
  __shared__ __attribute__((aligned(16))) float s_acc[16 * 64 * 4];  // slot sums per wave
  __shared__ int64_t s_bs[kMaxTilesPerJob];  // this chunk's block starts
  __shared__ int s_tn[kMaxTilesPerJob];      // this chunk's tiles holding entries, in order
  __shared__ int s_wn[16];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t c = (int64_t)blockIdx.x - jb.chunk_begin;
  const int nt = tl.nt;
  static_assert(kMaxTilesPerJob <= T, "one thread per tile lists the tiles");
  // Tiles where no row of the chunk has entries add nothing to any sum: the
  // chunk walks only the others (block totals are boff[block][1024]), in
  // tile order, so the sums do not change.
  const int tt = (int)threadIdx.x;
  const bool tile_live = tt < nt && tl.boff[(c * nt + tt) * (int64_t)RS + 1024] > 0;
  const uint64_t live_bits = __ballot(tile_live);
  if (tt < nt) s_bs[tt] = tl.bstart[c * nt + tt];
  if (lane == 0) s_wn[wave] = __builtin_popcountll(live_bits);
  f32x3 *const acc = reinterpret_cast<f32x3 *>(s_acc) + 64 * wave;
  acc[lane] = f32x3{0.0f, 0.0f, 0.0f};
  __syncthreads();
  int live_before = 0, n_live = 0;
  for (int w2 = 0; w2 < 16; ++w2) {
    live_before += w2 < wave ? s_wn[w2] : 0;
    n_live += s_wn[w2];
  }
  if (tile_live)
    s_tn[live_before + __builtin_popcountll(live_bits & ((1ull << lane) - 1ull))] = tt;
  __syncthreads();
  // k-th tile walked (a chunk with no entries at all walks tile 0, empty)
  const int ntl = __builtin_amdgcn_readfirstlane(n_live > 0 ? n_live : 1);
  auto tile_at = [&](int k) {
    return n_live > 0 ? __builtin_amdgcn_readfirstlane(s_tn[k]) : 0;
  };
  const int64_t oi = c + slot_rank(wave, lane) * tl.n_chunks;  // this lane's owner slot
  const int64_t row = oi < x.n ? (int64_t)x.order[oi] : -1;
  const char *bidx = reinterpret_cast<const char *>(tl.bidx);
  const char *bcoef = reinterpret_cast<const char *>(tl.bcoef);

  // emit tile tn, patch threadIdx.x + T m per thread (threads past the
  // tile's end load its last patch again), stored as 16-byte LDS records
  f32x3 epc[EPT];
  auto load_emit = [&](int tn) {
    const int64_t p1 = (int64_t)tn * C;
    const int last = (int)min((int64_t)C, x.n - p1) - 1;
    const char *src = reinterpret_cast<const char *>(emit_in + 3 * p1);
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
      const int k = min((int)threadIdx.x + T * m, last);
      const char *q = src + 12 * k;
      epc[m] = f32x3{load_global<float>(q), load_global<float>(q + 4), load_global<float>(q + 8)};
    }
  };
  auto store_emit = [&](int buf) {
#pragma unroll
    for (int m = 0; m < EPT; ++m)
      *reinterpret_cast<f32x3 *>(&s_e[buf][4 * emit_slot(T * m + (int)threadIdx.x)]) = epc[m];
  };
  // the wave's stream [a, e) of block (c, tn): its first and past-the-last
  // entries, read through the scalar cache (wave-uniform, read-only here), so
  // they do not count on the vector memory counter
  auto stream_bounds = [&](int tn, int &sa, int &sb) {
    const __attribute__((address_space(4))) int32_t *bo =
        (const __attribute__((address_space(4))) int32_t *)(tl.boff +
                                                            (c * nt + tn) * (int64_t)RS + 64 * wave);
    sa = bo[0];
    sb = bo[64];
  };
  auto meta = [&](int tn, int sa, int sb, bool live) {
    StreamMeta m;
    m.a = sa;
    m.e = sb;
    m.np = live ? (m.e - m.a + PW - 1) / PW : 0;
    const int64_t bs = uniform64(s_bs[tn]);
    // 16-bit kinds: offsets and coefficients interleaved, 4 bytes per entry
    m.ib = bidx + (KIND == VR_COEF_F32 ? 2 : 4) * bs;
    m.cb = bcoef + cbytes * bs;
    return m;
  };
  // loads of pass i: in memory a pass holds its A chunks, then its B chunks
  // (stream_pass_slot), so each load reads whole lines; lanes without a
  // chunk, and passes past the tile's last, load the pass's first chunk
  auto issue = [&](int i, const StreamMeta &m, PassLoad<KIND> &L) {
    const int p0 = m.a + i * PW;
    const int nc = i < m.np ? min(128, (m.e - p0) / G) : 0;
    const int ha = (nc + 1) >> 1;
    const int qa = lane < ha ? p0 + G * lane : m.a;
    const int qb = lane < nc - ha ? p0 + G * (ha + lane) : m.a;
    if constexpr (KIND == VR_COEF_F32) {
      // FP32 offsets are stored in stream order: the lane's A and B chunks
      // (2 lane, 2 lane + 1) are adjacent, one 16-byte load; coefficients
      // keep the A-then-B slots
      const int qn = lane < ha ? p0 + 2 * G * lane : m.a;
      const u32x4 o = load_global_nt<u32x4>(m.ib + 2u * (uint32_t)qn);
      const f32x4 wa = load_global_nt<f32x4>(m.cb + 4u * (uint32_t)qa);
      const f32x4 wb = load_global_nt<f32x4>(m.cb + 4u * (uint32_t)qb);
      L.a.o = make_uint2(o.x, o.y);
      L.b.o = make_uint2(o.z, o.w);
      L.a.w = make_float4(wa.x, wa.y, wa.z, wa.w);
      L.b.w = make_float4(wb.x, wb.y, wb.z, wb.w);
    } else {
      L.a.load_nt(m.ib, m.cb, (uint32_t)qa);
      L.b.load_nt(m.ib, m.cb, (uint32_t)qb);
    }
  };
  auto deliver = [&](int slot, float r, float g, float b) {
    f32x3 v = acc[slot];
    v.x = v.x + r;
    v.y = v.y + g;
    v.z = v.z + b;
    acc[slot] = v;
  };
  auto consume = [&](int i, const StreamMeta &m, const PassLoad<KIND> &L, const char *te) {
    const int qa = m.a + i * PW + 2 * G * lane;  // A's stream position
    const bool va = qa < m.e, vb = qa + G < m.e;
    const ChunkMeta ma = chunk_meta(L.a), mb = chunk_meta(L.b);
    const bool startb = vb && mb.start;
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    gather_stream(L.a, te, ar, ag, ab);
    float r = startb ? 0.0f : ar, g = startb ? 0.0f : ag, b = startb ? 0.0f : ab;
    if (vb) gather_stream(L.b, te, r, g, b);
    const int s = vb ? mb.s : ma.s;
    const ScanMasks k = scan_masks(lane, s);
    r = seg_scan(r, k);
    g = seg_scan(g, k);
    b = seg_scan(b, k);
    // S of the lane before (whole-wave shift right by one, 0 into lane 0)
    const float cr = dpp_src<0x138, 0xF>(r), cg = dpp_src<0x138, 0xF>(g),
                cb = dpp_src<0x138, 0xF>(b);
    if (va && startb) {
      const bool own = ma.start;  // A's segment began at A: no carry
      deliver(ma.slot, own ? ar : cr + ar, own ? ag : cg + ag, own ? ab : cb + ab);
    }
    const bool endl = vb ? (mb.last || lane == 63) : (ma.last || lane == 63);
    if (va && endl) deliver(vb ? mb.slot : ma.slot, r, g, b);
  };
  // one pass: its prefetch, then its gather; the scheduling barriers keep
  // every pass's loads in program order, so each load is consumed D passes
  // after it was issued and its wait leaves the later ones open
  auto pass = [&](int i, const StreamMeta &m, PassLoad<KIND> &ahead, bool has,
                  const PassLoad<KIND> &cur, const char *te) {
    issue(i + D, m, ahead);
    __builtin_amdgcn_sched_barrier(0);
    if (has) consume(i, m, cur, te);
    __builtin_amdgcn_sched_barrier(0);
  };

  int sa0, sb0, sa1, sb1;
  stream_bounds(tile_at(0), sa0, sb0);
  stream_bounds(tile_at(min(1, ntl - 1)), sa1, sb1);
  StreamMeta m0 = meta(tile_at(0), sa0, sb0, true);
  PassLoad<KIND> ring[2][D + 1];  // [tile parity][pass % (D + 1)]
  load_emit(tile_at(0));
#pragma unroll
  for (int k = 0; k < D; ++k) issue(k, m0, ring[0][k]);
  store_emit(0);

  // Before a tile's barrier each wave needs its LDS writes of the tile's emit
  // done (and its LDS reads of the previous tile): lgkmcnt 0, vmcnt left open.
  constexpr int WAIT_LGKM0 = 15 | (3 << 14) | (7 << 4);
  // the k-th walked tile
  auto tile = [&](auto parc, int k) {
    constexpr int p = decltype(parc)::value;
    __builtin_amdgcn_s_waitcnt(WAIT_LGKM0);
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();  // every wave's copy; the tile before is done
    __asm__ volatile("" ::: "memory");
    const int tn1 = tile_at(min(k + 1, ntl - 1));
    load_emit(tn1);  // on the last tile a harmless re-load
    int sa2, sb2;
    stream_bounds(tile_at(min(k + 2, ntl - 1)), sa2, sb2);
    const StreamMeta m1 = meta(tn1, sa1, sb1, k + 1 < ntl);
#pragma unroll
    for (int k = 0; k < D; ++k) issue(k, m1, ring[p ^ 1][k]);
    const char *te = reinterpret_cast<const char *>(s_e[p]);
    // every pass issues its loads whether or not it has entries, so the
    // compiler's waits on the ring are exact on every path
#pragma unroll
    for (int k = 0; k <= D; ++k)
      pass(k, m0, ring[p][(k + D) % (D + 1)], k < m0.np, ring[p][k], te);
    int i = D + 1;
    for (; m0.np - i >= D + 1; i += D + 1) {
#pragma unroll
      for (int k = 0; k <= D; ++k)
        pass(i + k, m0, ring[p][(k + D) % (D + 1)], true, ring[p][k], te);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      if (i + k < m0.np) pass(i + k, m0, ring[p][(k + D) % (D + 1)], true, ring[p][k], te);
    }
    store_emit(p ^ 1);
    m0 = m1;
    sa1 = sa2;
    sb1 = sb2;
  };
  for (int k = 0; k < ntl; k += 2) {
    tile(IntC<0>{}, k);
    if (k + 1 < ntl) tile(IntC<1>{}, k + 1);
  }
  __builtin_amdgcn_s_waitcnt(0);  // the last (unused) loads
  if (row >= 0) {
    const f32x3 v = acc[lane];
    finish_row<KIND>(jb, row, v.x, v.y, v.z, emit_out);
  }
}

// Blocked layout, pass 1: one workgroup per block (c, t): padded segment
// lengths of the 1024 owner slots, exclusive scan -> boff (stride 1028),
// total -> btot.
__global__ __launch_bounds__(1024) void k_block_offsets(DevXfers x, DevTiles tl, int G,
                                                        int64_t *__restrict__ btot,
                                                        int32_t *__restrict__ boff) {
  __shared__ int64_t sc[1024];
  const int R = 1024, RS = R + 4;
  const int64_t blk = blockIdx.x;  // c * nt + t
  const int64_t c = blk / tl.nt;
  const int t = (int)(blk % tl.nt);
  const int u = threadIdx.x;
  int64_t len = 0;
  if (u < R) {
    const int w = u / 64, j = u % 64;
    const int64_t oi = c + slot_rank(w, j) * tl.n_chunks;
    if (oi < x.n) {
      const int64_t row = x.order[oi];
      const int32_t *tp = tl.tp + row * (int64_t)(tl.nt + 1);
      len = tp[t + 1] - tp[t];
      len = (len + G - 1) / G * G;
    }
  }
  sc[u] = len;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
    const int64_t v = u >= off ? sc[u - off] : 0;
    __syncthreads();
    sc[u] += v;
    __syncthreads();
  }
  int32_t *bo = boff + blk * RS;
  if (u < R) bo[u + 1] = (int32_t)sc[u];
  if (u == 0) {
    bo[0] = 0;
    btot[blk] = sc[R - 1];
  }
}

// Blocked layout, pass 2: copy each owner's segment with its columns turned
// into LDS byte offsets ostride (col - t C) (stream layout: 16 emit_slot(col
// - t C)), bit 0 set on the segment's first entry; pad with (offset 0,
// coef 0).  One wave per (block, gather wave w): it copies the segments of
// w's 64 owner slots one after another, each with all 64 lanes.
__global__ __launch_bounds__(256) void k_block_fill(DevXfers x, DevTiles tl, int G,
                                                    int coef_bytes, uint16_t *__restrict__ bidx,
                                                    void *__restrict__ bcoef) {
  const int lane = threadIdx.x & 63;
  const int R = 1024, RS = R + 4;
  const int64_t gw = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nblk = tl.n_chunks * (int64_t)tl.nt;
  if (gw >= nblk * 16) return;
  const int64_t blk = gw / 16;
  const int w = (int)(gw % 16);
  const int64_t c = blk / tl.nt;
  const int t = (int)(blk % tl.nt);
  for (int j = 0; j < 64; ++j) {
    const int u = 64 * w + j;
    const int64_t oi = c + slot_rank(w, j) * tl.n_chunks;
    if (oi >= x.n) continue;
    const int64_t row = x.order[oi];
    const int32_t *tp = tl.tp + row * (int64_t)(tl.nt + 1);
    const int64_t src = x.rp[row] + tp[t];
    const int len = tp[t + 1] - tp[t];
    const int32_t *bo = tl.boff + blk * RS;
    const int64_t dst = tl.bstart[blk] + bo[u];
    const int plen = bo[u + 1] - bo[u];
    const int p0 = t * tl.cols;
    const int ss = bo[u] - bo[64 * w];  // the segment's start in the wave's stream
    const int wlen = bo[64 * w + 64] - bo[64 * w];
    const int nch = plen / G;
    for (int q = lane; q < plen; q += 64) {
      const bool real = q < len;
      uint32_t bits;
      int64_t at = dst + q;
      if (tl.ostride == 16) {  // stream layout: chunk metadata, A/B chunk order per pass
        const int pw = 128 * G, pos = ss + q, ps = pos / pw * pw, k = q / G;
        const int sl = ss >= ps ? ((ss - ps) / G) >> 1 : 0;
        bits = stream_chunk_meta(q % G, k == 0, k == nch - 1, sl, j);
        const int nc = min(128, (wlen - ps) / G);
        at = tl.bstart[blk] + bo[64 * w] + ps + G * stream_pass_slot((pos - ps) / G, nc) + q % G;
      } else {  // tiled layout: bit 0 marks the segment's first entry
        bits = q == 0 ? 1u : 0u;
      }
      const int cp = real ? x.col[src + q] - p0 : 0;
      const uint16_t off =
          (uint16_t)((tl.ostride == 16 ? 16 * emit_slot(cp) : tl.ostride * cp) | (int)bits);
      if (tl.ostride == 16 && coef_bytes == 2) {
        // stream layout, 16-bit: chunk record of 4 offsets then 4 coefficients
        // (chunks start at multiples of G = 4 entries)
        const int64_t rec = (at >> 2) * 8, m4 = at & 3;
        bidx[rec + m4] = off;
        bidx[rec + 4 + m4] = real ? reinterpret_cast<const uint16_t *>(x.coef)[src + q] : (uint16_t)0;
        continue;
      }
      // stream layout, FP32: offsets in stream order (the A and B chunks of a
      // lane adjacent), coefficients in the A-then-B slots
      bidx[tl.ostride == 16 ? dst + q : at] = off;
      if (coef_bytes == 4)
        reinterpret_cast<float *>(bcoef)[at] =
            real ? reinterpret_cast<const float *>(x.coef)[src + q] : 0.0f;
      else
        reinterpret_cast<uint16_t *>(bcoef)[at] =
            real ? reinterpret_cast<const uint16_t *>(x.coef)[src + q] : (uint16_t)0;
    }
  }
}

void launch_block_offsets(const DevXfers &x, const DevTiles &tl, int G, int64_t *btot,
                          int32_t *boff, hipStream_t st) {
  const int64_t nblk = tl.n_chunks * (int64_t)tl.nt;
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_block_offsets, dim3((unsigned)nblk), dim3(1024), 0, st, x, tl, G, btot,
                     boff);
}

void launch_block_fill(const DevXfers &x, const DevTiles &tl, int G, int coef_bytes,
                       uint16_t *bidx, void *bcoef, hipStream_t st) {
  const int64_t waves = tl.n_chunks * (int64_t)tl.nt * 16;
  if (waves <= 0) return;
  hipLaunchKernelGGL(k_block_fill, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, x, tl, G,
                     coef_bytes, bidx, bcoef);
}

// tile boundaries per row: tp[r][t] = first row position with column >= t*C
__global__ __launch_bounds__(256) void k_tile_ptrs(int64_t n, const int64_t *__restrict__ rp,
                                                   const int32_t *__restrict__ cnt,
                                                   const int32_t *__restrict__ col, int cols,
                                                   int nt, int32_t *__restrict__ tp) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (row >= n) return;
  const int32_t *c = col + rp[row];
  const int L = cnt[row];
  for (int t = lane; t <= nt; t += 64) {
    const int64_t key = (int64_t)t * cols;
    int lo = 0, hi = L;  // first index with c[idx] >= key
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((int64_t)c[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    tp[row * (int64_t)(nt + 1) + t] = t == nt ? L : lo;
  }
}

void launch_tile_ptrs(int64_t n, const int64_t *rp, const int32_t *cnt, const int32_t *col,
                      int cols, int nt, int32_t *tp, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_tile_ptrs, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, n, rp, cnt, col,
                     cols, nt, tp);
}

// Compiled shapes (C, GS, P, D) of the LDS-tiled gather.
#define VR_TILED_SHAPES(X) \
  X(4096, 16, 1, 1)        \
  X(4096, 16, 1, 3)        \
  X(4096, 16, 1, 7)        \
  X(4096, 16, 2, 1)        \
  X(4096, 16, 2, 3)        \
  X(4096, 8, 2, 1)         \
  X(4096, 8, 2, 3)         \
  X(4096, 8, 1, 3)         \
  X(4096, 8, 1, 7)         \
  X(2048, 8, 2, 1)         \
  X(2048, 16, 2, 1)

// Compiled shapes (C, D) of the stream gather.
#define VR_STREAM_SHAPES(X) \
  X(4096, 1)                \
  X(4096, 2)                \
  X(2048, 1)

bool bounce_gather_variant(const GatherShape &sh) {
#define VR_HAS(CC, GG, PP, DD) \
  if (sh.cols == CC && sh.group == GG && sh.passes == PP && sh.prefetch == DD) return true;
#define VR_HAS_S(CC, DD) \
  if (sh.cols == CC && sh.prefetch == DD) return true;
  if (sh.kernel == 0) {
    VR_TILED_SHAPES(VR_HAS)
  } else if (sh.kernel == 2) {
    VR_STREAM_SHAPES(VR_HAS_S)
  }
#undef VR_HAS_S
#undef VR_HAS
  return false;
}

void launch_gather_jobs(int kind, const GatherShape &sh, const BounceJob *jobs, int n_jobs,
                        int64_t total_chunks, hipStream_t st) {
  if (n_jobs <= 0 || total_chunks <= 0) return;
  const dim3 g((unsigned)total_chunks);
#define VR_GO(KER, K, ...) \
  hipLaunchKernelGGL((KER<K, __VA_ARGS__>), g, dim3(1024), 0, st, jobs, n_jobs)
#define VR_KINDS(KER, ...)                                   \
  {                                                          \
    if (kind == VR_COEF_F32) VR_GO(KER, VR_COEF_F32, __VA_ARGS__);  \
    else if (kind == VR_COEF_F16) VR_GO(KER, VR_COEF_F16, __VA_ARGS__); \
    else VR_GO(KER, VR_COEF_BF16, __VA_ARGS__);              \
    return;                                                  \
  }
#define VR_SHAPE(CC, GG, PP, DD)                                                          \
  if (sh.cols == CC && sh.group == GG && sh.passes == PP && sh.prefetch == DD)            \
    VR_KINDS(k_bounce_tiled, CC, GG, PP, DD)
#define VR_SHAPE_S(CC, DD) \
  if (sh.cols == CC && sh.prefetch == DD) VR_KINDS(k_bounce_stream, CC, DD)
  if (sh.kernel == 0) {
    VR_TILED_SHAPES(VR_SHAPE)
  } else if (sh.kernel == 2) {
    VR_STREAM_SHAPES(VR_SHAPE_S)
  }
#undef VR_SHAPE_S
#undef VR_SHAPE
#undef VR_KINDS
#undef VR_GO
}

// ---- FP64 energy: fixed tree.  Each block of 256 threads reduces 2048
// values: thread t sums [base + t + 256 j] for j = 0..7 from 0.0, then an
// LDS tree halves 128..1.  Stages repeat until one value is left.
__global__ __launch_bounds__(256) void k_energy_stage(const double *__restrict__ in, int64_t m,
                                                      double *__restrict__ out,
                                                      const BounceState *__restrict__ state) {
  if (state->converged) return;
  __shared__ double v[256];
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * 2048;
  double a = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t k = base + t + 256 * j;
    if (k < m) a = a + in[k];
  }
  v[t] = a;
  __syncthreads();
  for (int s = 128; s >= 1; s >>= 1) {
    if (t < s) v[t] = v[t] + v[t + s];
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = v[0];
}

__global__ void k_energy_final(const double *__restrict__ e, double *__restrict__ energy, int k,
                               double stop_frac, BounceState *__restrict__ state) {
  if (state->converged) return;
  const double ek = e[0];
  energy[k] = ek;
  const double cum = state->cum + ek;
  state->cum = cum;
  state->k_done = k + 1;
  if (ek <= stop_frac * cum) state->converged = 1;
}

void launch_energy(const double *e_row, int64_t n, double *scratch, double *energy, int k,
                   double stop_frac, BounceState *state, hipStream_t st) {
  // scratch holds ceil(n/2048) + ceil(n/2048^2) + ... + 1 doubles
  const double *in = e_row;
  double *out = scratch;
  int64_t m = n;
  if (m <= 0) {
    hipLaunchKernelGGL(k_energy_stage, dim3(1), dim3(256), 0, st, in, (int64_t)0, out, state);
    in = out;
  }
  while (m > 1) {
    int64_t nb = (m + 2047) / 2048;
    hipLaunchKernelGGL(k_energy_stage, dim3((unsigned)nb), dim3(256), 0, st, in, m, out, state);
    in = out;
    out += nb;
    m = nb;
  }
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(1), 0, st, in, energy, k, stop_frac, state);
}

// ---- FP64 energy of batched jobs: the same tree as launch_energy for
// n <= 2048^2 (stage 1: blocks of 2048 over e_row; stage 2: one block over
// the stage-1 partials when there is more than one), then the stop test.
__device__ __forceinline__ double reduce_block2048(const double *__restrict__ in, int64_t m,
                                                   int64_t blk, double *v) {
  const int t = threadIdx.x;
  const int64_t base = blk * 2048;
  double a = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t k = base + t + 256 * j;
    if (k < m) a = a + in[k];
  }
  v[t] = a;
  __syncthreads();
  for (int s = 128; s >= 1; s >>= 1) {
    if (t < s) v[t] = v[t] + v[t + s];
    __syncthreads();
  }
  return v[0];
}

__global__ __launch_bounds__(256) void k_energy_jobs1(const BounceJob *__restrict__ jobs,
                                                      int n_jobs) {
  __shared__ double v[256];
  const int jid = find_job<&BounceJob::eblk_begin>(jobs, n_jobs, blockIdx.x);
  const BounceJob &jb = jobs[jid];
  if (job_done(jb)) return;
  const int64_t lb = (int64_t)blockIdx.x - jb.eblk_begin;
  const double r = reduce_block2048(jb.e_row, jb.x.n, lb, v);
  if (threadIdx.x == 0) jb.scratch[lb] = r;
}

__global__ __launch_bounds__(256) void k_energy_jobs2(const BounceJob *__restrict__ jobs) {
  __shared__ double v[256];
  const BounceJob &jb = jobs[blockIdx.x];
  if (job_done(jb) || jb.e_blocks <= 1) return;
  const double r = reduce_block2048(jb.scratch, jb.e_blocks, 0, v);
  if (threadIdx.x == 0) jb.scratch[jb.e_blocks] = r;
}

__global__ __launch_bounds__(64) void k_energy_jobs_final(const BounceJob *__restrict__ jobs,
                                                          int n_jobs) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j >= n_jobs) return;
  const BounceJob &jb = jobs[j];
  if (job_done(jb)) return;
  const double ek = jb.e_blocks > 1 ? jb.scratch[jb.e_blocks]
                    : jb.e_blocks == 1 ? jb.scratch[0] : 0.0;
  BounceState *stt = jb.state;
  const int k = stt->k_done;
  jb.energy[k] = ek;
  const double cum = stt->cum + ek;
  stt->cum = cum;
  stt->k_done = k + 1;
  if (ek <= jb.stop_frac * cum) stt->converged = 1;
}

void launch_energy_jobs(const BounceJob *jobs, int n_jobs, int64_t total_eblocks, hipStream_t st) {
  if (n_jobs <= 0) return;
  if (total_eblocks > 0)
    hipLaunchKernelGGL(k_energy_jobs1, dim3((unsigned)total_eblocks), dim3(256), 0, st, jobs,
                       n_jobs);
  hipLaunchKernelGGL(k_energy_jobs2, dim3((unsigned)n_jobs), dim3(256), 0, st, jobs);
  hipLaunchKernelGGL(k_energy_jobs_final, dim3((unsigned)((n_jobs + 63) / 64)), dim3(64), 0, st,
                     jobs, n_jobs);
}

void launch_bounce_jobs(int kind, const GatherShape &sh, const BounceJob *jobs, int n_jobs,
                        int64_t total_chunks, int64_t total_eblocks, hipStream_t st) {
  launch_gather_jobs(kind, sh, jobs, n_jobs, total_chunks, st);
  launch_energy_jobs(jobs, n_jobs, total_eblocks, st);
}

// ---- first bounce input of a re-bake: total = direct, emit = direct * refl
__global__ __launch_bounds__(256) void k_emit_init(int64_t n3, const float *__restrict__ direct,
                                                   const float *__restrict__ refl,
                                                   float *__restrict__ emit,
                                                   float *__restrict__ total) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n3) return;
  const float d = direct[i];
  emit[i] = d * refl[i];
  total[i] = d;
}

void launch_emit_init(int64_t n, const float *direct, const float *refl, float *emit,
                      float *total, hipStream_t st) {
  const int64_t n3 = 3 * n;
  if (n3 <= 0) return;
  hipLaunchKernelGGL(k_emit_init, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, n3, direct,
                     refl, emit, total);
}

// ---- synthetic transfers (lognormal row lengths, uniform columns) ---------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t z = mix64(seed + 0x9E3779B97F4A7C15ull * (a * 0x100000001B3ull + b + 1));
  return ((double)(z >> 11) + 1.0) * 0x1p-53;  // (0, 1]
}

// Coefficients of row i: coef_p = u_p / sum(u) * s_i, u_p = u01(seed, i, p),
// the sum in FP64 in wave order (lane partials, then an xor butterfly).
__device__ __forceinline__ void synthetic_coefs(uint64_t seed, int64_t i, int64_t L, int lane,
                                                double &sum, double &si) {
  sum = 0.0;
  for (int64_t p = lane; p < L; p += 64) sum += u01(seed, (uint64_t)i, (uint64_t)p);
  for (int off = 32; off >= 1; off >>= 1) sum = sum + __shfl_xor(sum, off);
  si = 0.3 + 0.6 * u01(seed ^ 0xA5A5A5A5ull, (uint64_t)i, 0xFFFFFFFFull);
}

// Stratified columns (the named option): one wave per row, column p drawn in
// stratum [p m / L, (p + 1) m / L) of [0, m), m = n - 1, mapped past i.
__global__ __launch_bounds__(256) void k_synthetic_fill_strata(
    int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ len, uint64_t seed,
    int32_t *__restrict__ col, float *__restrict__ coef) {
  const int lane = threadIdx.x & 63;
  const int64_t wave =
      (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  for (int64_t i = wave; i < n; i += nwaves) {
    const int64_t L = len[i], s = rp[i], e = rp[i + 1];
    double sum, si;
    synthetic_coefs(seed, i, L, lane, sum, si);
    const int64_t m = n - 1;
    for (int64_t p = lane; p < e - s; p += 64) {
      if (p < L) {
        const int64_t a = (p * m) / L, b = ((p + 1) * m) / L;  // stratum [a, b)
        const double u = u01(seed ^ 0x5bd1e995ull, (uint64_t)i, (uint64_t)p);
        int64_t c = a + (int64_t)(u * (double)(b - a));
        if (c >= b) c = b - 1;
        if (c >= i) ++c;
        col[s + p] = (int32_t)c;
        coef[s + p] = (float)(u01(seed, (uint64_t)i, (uint64_t)p) / sum * si);
      } else {
        col[s + p] = (int32_t)i;
        coef[s + p] = 0.0f;
      }
    }
  }
}

// Uniform columns without replacement (the default): one wave per row.  The
// draws d = 0, 1, 2, ... give candidates c_d = floor(u01(seed ^ K, i, d) m)
// in [0, m), m = n - 1; the row takes the first L distinct candidates, which
// is a uniformly random L-subset.  A round takes draws 64 r + lane: a draw is
// new when no lower lane drew the same value this round and its bit in the
// wave's bitmap of [0, m) was clear (the atomic OR that sets it returns the
// old word); only the first L - taken new draws in lane order are kept (the
// others clear their bit again), so the result equals the sequential
// process.  The bitmap is then read in column order to write the row sorted
// (mapped past i) and cleared on the way.  Bitmap words are accessed only
// with device-scope atomics, loads and stores, which all act at L2.
__global__ __launch_bounds__(256) void k_synthetic_fill_uniform(
    int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ len, uint64_t seed,
    uint32_t *__restrict__ bitmaps, int64_t words, int32_t *__restrict__ col,
    float *__restrict__ coef) {
  const int lane = threadIdx.x & 63;
  const int64_t wave =
      (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  uint32_t *bm = bitmaps + wave * words;
  const int64_t m = n - 1;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int64_t i = wave; i < n; i += nwaves) {
    const int64_t L = len[i], s = rp[i], e = rp[i + 1];
    int64_t taken = 0;
    for (int64_t r = 0; taken < L; ++r) {
      const double u = u01(seed ^ 0x5bd1e995ull, (uint64_t)i, (uint64_t)(64 * r + lane));
      int c = (int)(u * (double)m);
      if (c >= m) c = (int)m - 1;
      bool dup = false;
      for (int j = 0; j < 63; ++j) {
        const int v = __shfl(c, j);
        if (j < lane && v == c) dup = true;
      }
      const uint32_t bit = 1u << (c & 31);
      uint32_t old = bit;
      if (!dup)
        old = __hip_atomic_fetch_or(&bm[c >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const bool fresh = (old & bit) == 0u;
      const int rank = __popcll(__ballot(fresh) & lt);
      const bool keep = fresh && (int64_t)rank < L - taken;
      if (fresh && !keep)
        __hip_atomic_fetch_and(&bm[c >> 5], ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      taken += __popcll(__ballot(keep));
    }
    double sum, si;
    synthetic_coefs(seed, i, L, lane, sum, si);
    // read the bitmap in column order: 64 words per step, a wave prefix of
    // their bit counts gives each set bit its position in the row
    int64_t base = 0;
    for (int64_t w0 = 0; w0 < words; w0 += 64) {
      const int64_t w = w0 + lane;
      uint32_t bits =
          w < words ? __hip_atomic_load(&bm[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
      const int cnt = __popc(bits);
      int incl = cnt;
      for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
      }
      int64_t pos = base + incl - cnt;
      if (bits) __hip_atomic_store(&bm[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      while (bits) {
        const int b = __ffs(bits) - 1;
        bits &= bits - 1u;
        int64_t c = (w << 5) + b;
        if (c >= i) ++c;
        col[s + pos] = (int32_t)c;
        coef[s + pos] = (float)(u01(seed, (uint64_t)i, (uint64_t)pos) / sum * si);
        ++pos;
      }
      base += __shfl(incl, 63);
    }
    for (int64_t p = L + lane; p < e - s; p += 64) {
      col[s + p] = (int32_t)i;
      coef[s + p] = 0.0f;
    }
    __threadfence();  // the clears land before the next row's atomics
  }
}

int64_t synthetic_bitmap_words(int64_t n) { return (n - 1 + 31) / 32; }
int64_t synthetic_waves(int64_t n) {
  // up to 8192 waves, the bitmaps held to 512 MiB
  const int64_t words = synthetic_bitmap_words(n);
  int64_t w = ((int64_t)512 << 20) / (4 * std::max<int64_t>(words, 1));
  w = std::min<int64_t>(8192, std::max<int64_t>(4, w)) & ~(int64_t)3;
  return std::min<int64_t>(w, ((n + 3) / 4) * 4);
}

void launch_synthetic_fill(int64_t n, const int64_t *rp, const int32_t *len, uint64_t seed,
                           int columns, uint32_t *bitmaps, int32_t *col, float *coef,
                           hipStream_t st) {
  if (n <= 0) return;
  if (columns == 1) {
    int64_t blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_synthetic_fill_strata, dim3((unsigned)blocks), dim3(256), 0, st, n, rp,
                       len, seed, col, coef);
    return;
  }
  const int64_t blocks = synthetic_waves(n) / 4;
  hipLaunchKernelGGL(k_synthetic_fill_uniform, dim3((unsigned)blocks), dim3(256), 0, st, n, rp,
                     len, seed, bitmaps, synthetic_bitmap_words(n), col, coef);
}

// ---- 16-bit coefficient encoding: scale = row max, q = rne(coef / scale) --
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(256) void k_encode16(int64_t n, const int64_t *__restrict__ rp,
                                                  const float *__restrict__ c32,
                                                  uint16_t *__restrict__ c16,
                                                  float *__restrict__ scale, int kind) {
  const int lane = threadIdx.x & 63;
  const int64_t wave =
      (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  for (int64_t i = wave; i < n; i += nwaves) {
    const int64_t s = rp[i], e = rp[i + 1];
    float mx = 0.0f;
    for (int64_t k = s + lane; k < e; k += 64) mx = fmaxf(mx, c32[k]);
    for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    const float sc = mx > 0.0f ? mx : 1.0f;
    if (lane == 0) scale[i] = sc;
    for (int64_t k = s + lane; k < e; k += 64) {
      const float q = c32[k] / sc;
      c16[k] = kind == VR_COEF_F16 ? __half_as_ushort(__float2half_rn(q)) : f32_to_bf16_rne(q);
    }
  }
}

void launch_encode16(int64_t n, const int64_t *rp, const float *coef32, uint16_t *coef16,
                     float *scale, int kind, hipStream_t st) {
  if (n <= 0) return;
  int64_t blocks = (n + 3) / 4;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_encode16, dim3((unsigned)blocks), dim3(256), 0, st, n, rp, coef32, coef16,
                     scale, kind);
}

__global__ void k_decode16(int64_t total, const uint16_t *__restrict__ c16, float *__restrict__ out,
                           int kind) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= total) return;
  out[k] = kind == VR_COEF_F16 ? f16_to_f32(c16[k]) : bf16_to_f32(c16[k]);
}

void launch_decode16(int64_t total, const uint16_t *coef16, float *out, int kind,
                     hipStream_t st) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_decode16, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, total,
                     coef16, out, kind);
}

// ---- streaming-read probe: the read bandwidth a plain kernel reaches ----
__global__ __launch_bounds__(256) void k_read_probe(const f32x4 *__restrict__ in, int64_t n4,
                                                    float *__restrict__ out) {
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nth = (int64_t)gridDim.x * 256;
  f32x4 a = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t i = tid; i < n4; i += nth) a += in[i];
  out[tid] = (a.x + a.y) + (a.z + a.w);
}

void launch_read_probe(const void *in, int64_t bytes, float *out, int blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_read_probe, dim3((unsigned)blocks), dim3(256), 0, st,
                     reinterpret_cast<const f32x4 *>(in), bytes / 16, out);
}

}  // namespace vr
