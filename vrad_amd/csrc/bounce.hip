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

// ---- synthetic transfers (config of the precision form) -------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t z = mix64(seed + 0x9E3779B97F4A7C15ull * (a * 0x100000001B3ull + b + 1));
  return ((double)(z >> 11) + 1.0) * 0x1p-53;  // (0, 1]
}

// one wave per row: columns one per stratum of [0, n-1) mapped past the row
// index; coef = u / sum(u) * s_i, sum in FP64 in the wave order
__global__ __launch_bounds__(256) void k_synthetic_fill(int64_t n, const int64_t *__restrict__ rp,
                                                        const int32_t *__restrict__ len,
                                                        uint64_t seed, int32_t *__restrict__ col,
                                                        float *__restrict__ coef) {
  const int lane = threadIdx.x & 63;
  const int64_t wave =
      (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nwaves = (int64_t)gridDim.x * 4;
  for (int64_t i = wave; i < n; i += nwaves) {
    const int64_t L = len[i], s = rp[i], e = rp[i + 1];
    double sum = 0.0;
    for (int64_t p = lane; p < L; p += 64) sum += u01(seed, (uint64_t)i, (uint64_t)p);
    for (int off = 32; off >= 1; off >>= 1) sum = sum + __shfl_xor(sum, off);
    const double si = 0.3 + 0.6 * u01(seed ^ 0xA5A5A5A5ull, (uint64_t)i, 0xFFFFFFFFull);
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

void launch_synthetic_fill(int64_t n, const int64_t *rp, const int32_t *len, uint64_t seed,
                           int32_t *col, float *coef, hipStream_t st) {
  if (n <= 0) return;
  int64_t blocks = (n + 3) / 4;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_synthetic_fill, dim3((unsigned)blocks), dim3(256), 0, st, n, rp, len, seed,
                     col, coef);
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

}  // namespace vr
