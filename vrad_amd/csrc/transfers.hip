// MakeTransfers: patch-pair visibility batches and CSR construction.
//
//  1. k_pair_cull   candidate pairs (i, j > i) of a row range that pass the
//                   facing cull, compacted with a wave ballot + one atomic.
//  2. k_pair_trace  TestLine(o_i + eps n_i -> o_j + eps n_j) per candidate;
//                   visible pairs set bits (i,j) and (j,i) of an N x N
//                   bit matrix (N^2/8 bytes: 10 GB at 280k patches, which an
//                   MI355X holds with room to spare).
//  3. k_row_count   popcount per row; k_row_scan pads lengths to 8 and scans.
//  4. k_row_fill    columns in ascending order, F_ij in FP32.
//  5. k_row_norm    FP64 row sum in the bounce-kernel wave order (G = 4);
//                   rows with sum > 1 are scaled by 1/sum.
#include "kernels.h"

namespace vr {

#define VR_PI_F 3.14159265358979323846f

// A block of 256 threads tests 256 * kCullIt candidates (row, j) of one row;
// the waves keep their ballots, the block reserves its survivors' slots with
// one atomic, and every wave writes its pairs in j order.
constexpr int kCullIt = 16;

__global__ __launch_bounds__(256) void k_pair_cull(DevPatches p, int64_t np, int64_t i0,
                                                   int64_t bpr, int2 *__restrict__ pairs,
                                                   uint32_t *__restrict__ count, uint32_t cap) {
  __shared__ uint32_t s_cnt[4];
  __shared__ uint32_t s_base;
  const int64_t row = i0 + (int64_t)blockIdx.x / bpr;
  const int64_t jb = row + 1 + ((int64_t)blockIdx.x % bpr) * (256 * kCullIt);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float rx = p.ox[row], ry = p.oy[row], rz = p.oz[row];
  const float qx = p.nx[row], qy = p.ny[row], qz = p.nz[row];
  uint64_t m[kCullIt];
  uint32_t wtot = 0;
#pragma unroll
  for (int it = 0; it < kCullIt; ++it) {
    const int64_t j = jb + it * 256 + threadIdx.x;
    bool keep = false;
    if (j < np) {
      const float dx = p.ox[j] - rx, dy = p.oy[j] - ry, dz = p.oz[j] - rz;
      const float di = (qx * dx + qy * dy) + qz * dz;
      const float dj = -((p.nx[j] * dx + p.ny[j] * dy) + p.nz[j] * dz);
      keep = di > 0.0f && dj > 0.0f;
    }
    m[it] = __ballot(keep);
    wtot += (uint32_t)__popcll(m[it]);
  }
  if (lane == 0) s_cnt[wave] = wtot;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t tot = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
    s_base = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  uint32_t base = s_base;
  for (int w = 0; w < wave; ++w) base += s_cnt[w];
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (int it = 0; it < kCullIt; ++it) {
    if ((m[it] >> lane) & 1ull) {
      const uint32_t slot = base + (uint32_t)__popcll(m[it] & below);
      if (slot < cap) pairs[slot] = make_int2((int)row, (int)(jb + it * 256 + threadIdx.x));
    }
    base += (uint32_t)__popcll(m[it]);
  }
}

void launch_pair_cull(const DevPatches &p, int64_t np, int64_t i0, int64_t i1, int2 *pairs,
                      uint32_t *count, uint32_t cap, hipStream_t st) {
  if (i1 <= i0) return;
  const int64_t maxlen = np - 1 - i0;
  if (maxlen <= 0) return;
  const int64_t bpr = (maxlen + 256 * kCullIt - 1) / (256 * kCullIt);
  hipLaunchKernelGGL(k_pair_cull, dim3((unsigned)((i1 - i0) * bpr)), dim3(256), 0, st, p, np, i0,
                     bpr, pairs, count, cap);
}

__device__ __forceinline__ void set_pair_bits(uint32_t *bits, int64_t words, int i, int j) {
  atomicOr(bits + (int64_t)i * words + (j >> 5), 1u << (j & 31));
  atomicOr(bits + (int64_t)j * words + (i >> 5), 1u << (i & 31));
}

template <int S, int D>
__global__ __launch_bounds__(64) void k_pair_trace_lds(DevScene sc, DevPatches p,
                                                       const int2 *__restrict__ pairs,
                                                       const uint32_t *__restrict__ count,
                                                       uint32_t cap, float eps, int32_t skip,
                                                       uint32_t *__restrict__ bits,
                                                       int64_t words, uint32_t xg) {
  __shared__ uint2 s_stk[S * 64];
  // The grid is sized for cap pairs; only the first ceil(n / 64) blocks have
  // work.  Blocks are dealt to the XCDs round-robin, so those blocks spread
  // over all of them, and the XCD-aware remap runs over just those.
  const uint32_t n = min(*count, cap);
  const uint32_t nb = (n + 63u) / 64u;
  if (blockIdx.x >= nb) return;
  const uint32_t k = xcd_group_chunk(blockIdx.x, nb, xg) * 64 + threadIdx.x;
  if (k >= n) return;
  const int2 ij = pairs[k];
  const int i = ij.x, j = ij.y;
  StackSpill<S, D> spill;
  HybridStack<S, D> st(s_stk, (int)threadIdx.x, spill);
  const bool vis = segment_visible<D>(sc, st, p.ox[i] + eps * p.nx[i],
                                      p.oy[i] + eps * p.ny[i], p.oz[i] + eps * p.nz[i],
                                      p.ox[j] + eps * p.nx[j], p.oy[j] + eps * p.ny[j],
                                      p.oz[j] + eps * p.nz[j], skip);
  if (vis) set_pair_bits(bits, words, i, j);
}

template <int D>
__global__ __launch_bounds__(256) void k_pair_trace_plain(DevScene sc, DevPatches p,
                                                          const int2 *__restrict__ pairs,
                                                          const uint32_t *__restrict__ count,
                                                          uint32_t cap, float eps, int32_t skip,
                                                          uint32_t *__restrict__ bits,
                                                          int64_t words) {
  const uint32_t n = min(*count, cap);
  const uint32_t nb = (n + 255u) / 256u;  // blocks with work (see k_pair_trace_lds)
  if (blockIdx.x >= nb) return;
  const uint32_t k = xcd_chunk(blockIdx.x, nb) * 256 + threadIdx.x;
  if (k >= n) return;
  const int2 ij = pairs[k];
  const int i = ij.x, j = ij.y;
  ScratchStack<D> st;
  const bool vis = segment_visible<D>(sc, st, p.ox[i] + eps * p.nx[i],
                                      p.oy[i] + eps * p.ny[i], p.oz[i] + eps * p.nz[i],
                                      p.ox[j] + eps * p.nx[j], p.oy[j] + eps * p.ny[j],
                                      p.oz[j] + eps * p.nz[j], skip);
  if (vis) set_pair_bits(bits, words, i, j);
}

void launch_pair_trace(const DevScene &sc, const DevPatches &p, const int2 *pairs,
                       const uint32_t *count, uint32_t cap, float eps, int32_t skip,
                       uint32_t *bits, int64_t words, int path, hipStream_t st) {
  if (cap == 0) return;
  if (path == VR_PATH_PLAIN)
    hipLaunchKernelGGL(k_pair_trace_plain<kStackDepth>, dim3((cap + 255) / 256), dim3(256), 0, st,
                       sc, p, pairs, count, cap, eps, skip, bits, words);
  else {
    // the traversal-stack entries kept in LDS follow the autotuned knob
#define VR_PAIR_TRACE(S)                                                                    \
  hipLaunchKernelGGL((k_pair_trace_lds<S, kStackDepth>), dim3((cap + 63) / 64), dim3(64), 0, st, \
                     sc, p, pairs, count, cap, eps, skip, bits, words,                          \
                     (uint32_t)tuning().trace_xcd_group)
    switch (tuning().trace_lds_entries) {
      case 2: VR_PAIR_TRACE(2); break;
      case 4: VR_PAIR_TRACE(4); break;
      case 6: VR_PAIR_TRACE(6); break;
      case 10: VR_PAIR_TRACE(10); break;
      case 12: VR_PAIR_TRACE(12); break;
      case 16: VR_PAIR_TRACE(16); break;
      case 24: VR_PAIR_TRACE(24); break;
      default: VR_PAIR_TRACE(8); break;
    }
#undef VR_PAIR_TRACE
  }
}

// Many scenes' candidate pairs in one launch: block b (after the XCD remap)
// belongs to the job whose block range holds it (binary search over the
// jobs' first blocks; the job index is the same for the whole wave, so its
// descriptor comes through scalar loads), and lane l traces that job's pair
// (b - first block) * 64 + l on the job's scene.
template <int S, int D>
__global__ __launch_bounds__(64) void k_pair_trace_jobs(const TraceJob *__restrict__ jobs,
                                                        int n_jobs, float eps, int32_t skip) {
  __shared__ uint2 s_stk[S * 64];
  const int64_t b = (int64_t)xcd_group_chunk(blockIdx.x, gridDim.x, kXcdGroup);
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {  // last job with blk_begin <= b
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].blk_begin <= b) lo = mid;
    else hi = mid - 1;
  }
  const int jx = __builtin_amdgcn_readfirstlane(lo);
  const TraceJob &J = jobs[jx];
  const int64_t k = (b - J.blk_begin) * 64 + threadIdx.x;
  if (k >= (int64_t)J.count) return;
  const int2 ij = J.pairs[k];
  const int i = ij.x, j = ij.y;
  const DevPatches &p = J.p;
  StackSpill<S, D> spill;
  HybridStack<S, D> st(s_stk, (int)threadIdx.x, spill);
  const bool vis = segment_visible<D>(J.sc, st, p.ox[i] + eps * p.nx[i],
                                      p.oy[i] + eps * p.ny[i], p.oz[i] + eps * p.nz[i],
                                      p.ox[j] + eps * p.nx[j], p.oy[j] + eps * p.ny[j],
                                      p.oz[j] + eps * p.nz[j], skip);
  if (vis) set_pair_bits(J.bits, J.words, i, j);
}

void launch_pair_trace_jobs(const TraceJob *jobs, int n_jobs, int64_t total_blocks, float eps,
                            int32_t skip, int lds_entries, hipStream_t st) {
  if (n_jobs <= 0 || total_blocks <= 0) return;
#define VR_PT_JOBS(S)                                                                          \
  hipLaunchKernelGGL((k_pair_trace_jobs<S, kStackDepth>), dim3((unsigned)total_blocks), dim3(64), \
                     0, st, jobs, n_jobs, eps, skip)
  switch (lds_entries) {
    case 2: VR_PT_JOBS(2); break;
    case 4: VR_PT_JOBS(4); break;
    case 6: VR_PT_JOBS(6); break;
    case 10: VR_PT_JOBS(10); break;
    case 12: VR_PT_JOBS(12); break;
    case 16: VR_PT_JOBS(16); break;
    case 24: VR_PT_JOBS(24); break;
    default: VR_PT_JOBS(8); break;
  }
#undef VR_PT_JOBS
}

// one wave per row
__global__ __launch_bounds__(256) void k_row_count(const uint32_t *__restrict__ bits, int64_t np,
                                                   int64_t words, int32_t *__restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  if (row >= np) return;
  const uint32_t *rb = bits + row * words;
  int c = 0;
  for (int64_t w = lane; w < words; w += 64) c += __popc(rb[w]);
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
  if (lane == 0) cnt[row] = c;
}

void launch_row_count(const uint32_t *bits, int64_t np, int64_t words, int32_t *cnt,
                      hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_row_count, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, st, bits, np,
                     words, cnt);
}

// single-block exclusive scan of padded lengths: rp[i] = sum_{k<i} pad8(cnt[k])
__global__ __launch_bounds__(1024) void k_row_scan(const int32_t *__restrict__ cnt, int64_t n,
                                                   int64_t *__restrict__ rp) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t b = t * per, e = min(n, b + per);
  int64_t s = 0;
  for (int64_t k = b; k < e; ++k) s += ((int64_t)cnt[k] + 7) & ~(int64_t)7;
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t run = 0;
    for (int k = 0; k < 1024; ++k) {
      int64_t v = part[k];
      part[k] = run;
      run += v;
    }
    rp[n] = run;
  }
  __syncthreads();
  int64_t run = part[t];
  for (int64_t k = b; k < e; ++k) {
    rp[k] = run;
    run += ((int64_t)cnt[k] + 7) & ~(int64_t)7;
  }
}

void launch_row_scan(const int32_t *cnt, int64_t n, int64_t *rp, hipStream_t st) {
  hipLaunchKernelGGL(k_row_scan, dim3(1), dim3(1024), 0, st, cnt, n, rp);
}

__device__ __forceinline__ float form_factor(const DevPatches &p, int64_t i, int64_t j) {
  const float dx = p.ox[j] - p.ox[i], dy = p.oy[j] - p.oy[i], dz = p.oz[j] - p.oz[i];
  const float di = (p.nx[i] * dx + p.ny[i] * dy) + p.nz[i] * dz;
  const float dj = -((p.nx[j] * dx + p.ny[j] * dy) + p.nz[j] * dz);
  const float r2 = (dx * dx + dy * dy) + dz * dz;
  return ((di * dj) * p.area[j]) / ((VR_PI_F * r2) * r2);
}

// One wave per row, over the bit matrix row in chunks of 64 words (2048
// columns): each lane takes a word, the set bits' columns go to LDS at their
// rank in the chunk (ascending), and consecutive lanes then write consecutive
// entries (col and F_ij) to the CSR, so the stores are contiguous.
__global__ __launch_bounds__(64) void k_row_fill(const uint32_t *__restrict__ bits, int64_t np,
                                                 int64_t words, DevPatches p,
                                                 const int64_t *__restrict__ rp,
                                                 const int32_t *__restrict__ cnt,
                                                 int32_t *__restrict__ col,
                                                 float *__restrict__ coef) {
  __shared__ int32_t s_col[64 * 32];
  const int lane = (int)threadIdx.x;
  const int64_t row = blockIdx.x;
  if (row >= np) return;
  const uint32_t *rb = bits + row * words;
  const int64_t s = rp[row];
  int64_t run = 0;
  for (int64_t w0 = 0; w0 < words; w0 += 64) {
    const int64_t w = w0 + lane;
    uint32_t m = w < words ? rb[w] : 0u;
    const int pc = __popc(m);
    // inclusive wave scan of pc
    int inc = pc;
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(inc, off);
      if (lane >= off) inc += v;
    }
    const int tot = __shfl(inc, 63);
    if (tot == 0) continue;
    int at = inc - pc;
    while (m) {
      const int bit = __ffs(m) - 1;
      m &= m - 1u;
      s_col[at++] = (int32_t)(w * 32 + bit);
    }
    __syncthreads();  // one wave: the chunk's columns are in LDS before they are read
    for (int k = lane; k < tot; k += 64) {
      const int32_t j = s_col[k];
      col[s + run + k] = j;
      coef[s + run + k] = form_factor(p, row, j);
    }
    __syncthreads();  // and read before the next chunk overwrites them
    run += tot;
  }
  const int64_t c = cnt[row], e = rp[row + 1];
  for (int64_t k = s + c + lane; k < e; k += 64) {
    col[k] = (int32_t)row;
    coef[k] = 0.0f;
  }
}

void launch_row_fill(const uint32_t *bits, int64_t np, int64_t words, const DevPatches &p,
                     const int64_t *rp, const int32_t *cnt, int32_t *col, float *coef,
                     hipStream_t st) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_row_fill, dim3((unsigned)np), dim3(64), 0, st, bits, np, words, p, rp, cnt,
                     col, coef);
}

// FP64 row sum in wave order (lane owns p with (p/4)%64 == lane), butterfly
__global__ __launch_bounds__(256) void k_row_norm(int64_t n, const int64_t *__restrict__ rp,
                                                  float *__restrict__ coef,
                                                  double *__restrict__ row_sum, int normalize) {
  const int lane = threadIdx.x & 63;
  const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  if (row >= n) return;
  const int64_t s = rp[row], e = rp[row + 1];
  double a = 0.0;
  for (int64_t k = s + 4 * lane; k < e; k += 256) {
    const float4 w = *reinterpret_cast<const float4 *>(coef + k);
    a = a + (double)w.x;
    a = a + (double)w.y;
    a = a + (double)w.z;
    a = a + (double)w.w;
  }
  for (int off = 32; off >= 1; off >>= 1) a = a + __shfl_xor(a, off);
  if (lane == 0) row_sum[row] = a;
  if (normalize && a > 1.0)
    for (int64_t k = s + lane; k < e; k += 64) coef[k] = (float)((double)coef[k] / a);
}

void launch_row_normalize(int64_t n, const int64_t *rp, float *coef, double *row_sum,
                          int normalize, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_row_norm, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, n, rp, coef,
                     row_sum, normalize);
}

}  // namespace vr
