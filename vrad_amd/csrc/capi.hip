// C-ABI of libvrad_hip.so (declared in include/vrad_hip.h): scene residency,
// visibility batches, transfer construction, direct light, bounce driver,
// procedural inputs and kernel timing.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kernels.h"
#include "vr_internal.h"

namespace vr {
int gen_room(uint64_t seed, int n_pillars, double chop, double tess, GenScene *out);
int gen_level(uint64_t seed, int rx, int ry, double chop, double tess, int n_lights, int sun,
              GenScene *out);

Tuning &tuning() {
  static Tuning t;
  return t;
}

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

namespace {

std::recursive_mutex g_mu;
bool g_init = false;
int g_device = 0;
hipStream_t g_stream = nullptr;
// a second stream, on which MakeTransfers runs every other round of its
// visibility batches (default priority, like g_stream)
hipStream_t g_stream2 = nullptr;
int g_path = VR_PATH_OPT;
bool g_prof = false;

struct ProfEntry {
  std::string name;
  hipEvent_t a, b;
};
std::vector<ProfEntry> g_pending;
std::map<std::string, std::pair<double, int64_t>> g_prof_tot;

void prof_drain() {
  for (auto &p : g_pending) {
    float ms = 0.0f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto &t = g_prof_tot[p.name];
      t.first += ms;
      t.second += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  g_pending.clear();
}

// Brackets the launches in its scope with HIP events on `st` when enabled.
struct Prof {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t st;
  const char *name;
  Prof(const char *nm, hipStream_t s) : st(s), name(nm) {
    if (!g_prof) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      a = b = nullptr;
      return;
    }
    (void)hipEventRecord(a, st);
  }
  ~Prof() {
    if (!a) return;
    (void)hipEventRecord(b, st);
    g_pending.push_back(ProfEntry{name, a, b});
  }
};

#define VR_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(VR_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

// ---- the device working set of vr_bounce_light_noalloc ----------------------
// Callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc) get a bounce whose
// device working set lives here: one file-scope array per GPU, loaded with
// the code object and held until the process ends.  Layout (each part
// 256-byte aligned): emit ping-pong and total (12 B per row + 16 each),
// reflectivity (12 B per row), area (4 B), per-row energy terms (8 B), the
// energy tree's partials, the per-bounce energies, the bounce state and the
// one-job table.  In host code `scratch` names a host-side stand-in, never
// the device memory: host code uses only the address hipGetSymbolAddress
// returns for the current device (g_scratch).
constexpr int64_t kNoAllocMaxRows = (int64_t)1 << 19;
constexpr int kNoAllocMaxBounces = 1024;
constexpr size_t na_pad(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaVec = na_pad(12 * (size_t)kNoAllocMaxRows + 16);
constexpr size_t kNaEmit0 = 0, kNaEmit1 = kNaEmit0 + kNaVec, kNaTotal = kNaEmit1 + kNaVec;
constexpr size_t kNaRefl = kNaTotal + kNaVec;
constexpr size_t kNaArea = kNaRefl + na_pad(12 * (size_t)kNoAllocMaxRows);
constexpr size_t kNaERow = kNaArea + na_pad(4 * (size_t)kNoAllocMaxRows);
constexpr size_t kNaEScr = kNaERow + na_pad(8 * (size_t)kNoAllocMaxRows);
constexpr size_t kNaEnergy = kNaEScr + na_pad(8 * (size_t)(kNoAllocMaxRows / 2048 + 2));
constexpr size_t kNaState = kNaEnergy + na_pad(8 * (size_t)kNoAllocMaxBounces);
constexpr size_t kNaJob = kNaState + 256;
constexpr size_t kNaBytes = kNaJob + 1024;
static_assert(sizeof(BounceState) <= 256 && sizeof(BounceJob) <= 1024, "scratch layout");
static __device__ __align__(256) unsigned char scratch[kNaBytes];
constexpr int kMaxDevices = 64;
char *g_scratch[kMaxDevices] = {};

int ensure_init() {
  if (g_init) return VR_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(VR_E_NODEVICE, "no HIP device available");
  if (g_device < 0 || g_device >= n || g_device >= kMaxDevices)
    return fail(VR_E_INVALID, "no such device");
  VR_HIP(hipSetDevice(g_device));
  VR_HIP(hipStreamCreate(&g_stream));
  VR_HIP(hipStreamCreate(&g_stream2));
  // this device's copy of the allocation-free bounce's working set
  void *sp = nullptr;
  VR_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  g_scratch[g_device] = static_cast<char *>(sp);
  g_init = true;
  return VR_OK;
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(VR_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return VR_OK;
}

// device buffer
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t n) {
    release();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(VR_E_NOMEM, "hipMalloc(" + std::to_string(n) + "): " + hipGetErrorString(e));
    }
    bytes = n;
    return VR_OK;
  }
  int upload(const void *h, size_t n, hipStream_t st) {
    int rc = alloc(n);
    if (rc) return rc;
    if (n && h) VR_HIP(hipMemcpyAsync(p, h, n, hipMemcpyHostToDevice, st));
    return VR_OK;
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

template <class T>
int upload_vec(DBuf &b, const std::vector<T> &v, hipStream_t st) {
  return b.upload(v.data(), v.size() * sizeof(T), st);
}

}  // namespace
}  // namespace vr

using namespace vr;

// ---------------------------------------------------------------------------
struct vr_scene {
  int64_t n_tri = 0;
  std::vector<float> v9;
  std::vector<int32_t> ids;
  std::vector<uint8_t> flags;
  bool built = false, host_built = false;
  std::vector<TriIntersect> tris;
  KdTree kd;
  KdTree kd_sky;                   // over the TRACE_ID_SKY triangles only
  std::vector<int32_t> sky_index;  // kd_sky triangle -> triangle index
  DBuf nodes, refs, planes, sky_nodes, sky_refs;
  DevScene dev{};
  int has_sky = 0;
  int64_t device_bytes = 0;
};

struct vr_xfers {
  int64_t n = 0, nnz = 0, stored = 0;
  int32_t precision = VR_COEF_F32;
  int64_t pairs = 0, segments = 0;
  std::vector<int64_t> rp_host;  // padded starts
  std::vector<int32_t> cnt_host; // logical lengths
  DBuf rp, col, coef32, coef16, scale, order, cnt;
  int placement = 0;  // the row placement `order` was built for (bounce_placement)
  DBuf tp;            // tile boundaries for the LDS-tiled gather
  int tp_cols = 0;
  // tile-blocked copy of the coefficients for the LDS-tiled gather
  DBuf bidx, bcoef, bstart, boff;
  int blk_cols = 0, blk_prec = -1, blk_stride = 0;
  int64_t blk_chunks = 0;
  // servers holding this set: while > 0 its buffers are read by a serving
  // thread, so nothing may reallocate or free them
  int served = 0;
  DevXfers dev() const {
    DevXfers x;
    x.n = n;
    x.rp = rp.as<int64_t>();
    x.col = col.as<int32_t>();
    x.coef = precision == VR_COEF_F32 ? coef32.p : coef16.p;
    x.scale = scale.as<float>();
    x.order = order.as<int32_t>();
    x.kind = precision;
    return x;
  }
};

struct vr_gen {
  GenScene g;
};

namespace vr {
namespace {

// The gather kernels read chunk c's k-th row as order[c + k n_chunks] (and
// deal a chunk's k-th row to wave slot_rank^-1(k)).
// Placement 0: order = every row longest-first (ties by index), so chunk c
// takes a stride through it.  Placement 1: chunk c's rows are the c-th run of
// consecutive rows (runs of ceil((n - c) / n_chunks) rows, the chunk sizes of
// placement 0), longest-first inside the run.  Placement 2: chunk c's rows
// are placement 0's, ranked by row index, so every wave holds rows from all
// over the index range (and so, for MakeTransfers rows, all over the scene).
void row_order(const std::vector<int32_t> &cnt, int placement, std::vector<int32_t> &order) {
  const int64_t n = (int64_t)cnt.size();
  order.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) order[i] = (int32_t)i;
  auto longer = [&](int32_t a, int32_t b) { return cnt[a] > cnt[b]; };
  if (placement != 1) {
    std::stable_sort(order.begin(), order.end(), longer);
    if (placement == 2) {
      const int64_t nc = gather_chunks(n);
      std::vector<int32_t> mine;
      for (int64_t c = 0; c < nc; ++c) {
        mine.clear();
        for (int64_t k = c; k < n; k += nc) mine.push_back(order[(size_t)k]);
        std::sort(mine.begin(), mine.end());
        for (size_t k = 0; k < mine.size(); ++k) order[(size_t)(c + (int64_t)k * nc)] = mine[k];
      }
    }
    return;
  }
  // Run r (consecutive rows, r-th in index order) has the size chunk r would
  // have.  A chunk's rows alone fix each row's sums, so the runs of one size
  // may go to any chunk of that size: the heaviest first, since workgroups
  // start in chunk order and the heaviest then finish soonest.
  const int64_t nc = gather_chunks(n);
  std::vector<int64_t> run_first((size_t)nc, 0), run_size((size_t)nc, 0), weight((size_t)nc, 0);
  int64_t first = 0;
  for (int64_t r = 0; r < nc; ++r) {
    run_first[(size_t)r] = first;
    run_size[(size_t)r] = (n - r + nc - 1) / nc;
    for (int64_t k = 0; k < run_size[(size_t)r]; ++k) weight[(size_t)r] += cnt[(size_t)(first + k)];
    first += run_size[(size_t)r];
  }
  std::vector<int64_t> slot((size_t)nc);
  for (int64_t r = 0; r < nc; ++r) slot[(size_t)r] = r;
  // chunks 0 .. big-1 hold one row more than the rest
  const int64_t big = n % nc ? n % nc : nc;
  auto heavier = [&](int64_t a, int64_t b) { return weight[(size_t)a] > weight[(size_t)b]; };
  std::stable_sort(slot.begin(), slot.begin() + big, heavier);
  std::stable_sort(slot.begin() + big, slot.end(), heavier);
  std::vector<int32_t> run;
  for (int64_t c = 0; c < nc; ++c) {
    const int64_t r = slot[(size_t)c], size = run_size[(size_t)r];
    run.resize((size_t)size);
    for (int64_t k = 0; k < size; ++k) run[(size_t)k] = (int32_t)(run_first[(size_t)r] + k);
    std::stable_sort(run.begin(), run.end(), longer);
    for (int64_t k = 0; k < size; ++k) order[(size_t)(c + k * nc)] = run[(size_t)k];
  }
}

// padded row starts (rows padded to a multiple of 8 entries)
void layout_rows(const std::vector<int32_t> &cnt, std::vector<int64_t> &rp) {
  const int64_t n = (int64_t)cnt.size();
  rp.assign((size_t)n + 1, 0);
  for (int64_t i = 0; i < n; ++i) rp[i + 1] = rp[i] + (((int64_t)cnt[i] + 7) & ~(int64_t)7);
}

int finish_layout(vr_xfers *x, hipStream_t st) {
  std::vector<int32_t> order;
  layout_rows(x->cnt_host, x->rp_host);
  row_order(x->cnt_host, 0, order);
  x->placement = 0;
  int rc = upload_vec(x->rp, x->rp_host, st);
  if (!rc) rc = upload_vec(x->order, order, st);
  if (!rc) rc = upload_vec(x->cnt, x->cnt_host, st);
  x->stored = x->rp_host.back();
  return rc;
}

// The 162 fixed sky directions: the reference's Anorms table
// (vmath/constants.go:15-184), the vertices of a frequency-4 geodesic
// icosphere rounded to 6 decimals.  CanLeafTraceToSky walks them
// (rad/lightmap/lightmap.go:435-448); sky ambient weights them by cosine.
// Each literal converts to the nearest float32, as Go's mgl32.Vec3 literal does.
static const float kAnorms[162][3] = {
    {-0.525731f, 0.000000f, 0.850651f}, {-0.442863f, 0.238856f, 0.864188f}, {-0.295242f, 0.000000f, 0.955423f},
    {-0.309017f, 0.500000f, 0.809017f}, {-0.162460f, 0.262866f, 0.951056f}, {0.000000f, 0.000000f, 1.000000f},
    {0.000000f, 0.850651f, 0.525731f}, {-0.147621f, 0.716567f, 0.681718f}, {0.147621f, 0.716567f, 0.681718f},
    {0.000000f, 0.525731f, 0.850651f}, {0.309017f, 0.500000f, 0.809017f}, {0.525731f, 0.000000f, 0.850651f},
    {0.295242f, 0.000000f, 0.955423f}, {0.442863f, 0.238856f, 0.864188f}, {0.162460f, 0.262866f, 0.951056f},
    {-0.681718f, 0.147621f, 0.716567f}, {-0.809017f, 0.309017f, 0.500000f}, {-0.587785f, 0.425325f, 0.688191f},
    {-0.850651f, 0.525731f, 0.000000f}, {-0.864188f, 0.442863f, 0.238856f}, {-0.716567f, 0.681718f, 0.147621f},
    {-0.688191f, 0.587785f, 0.425325f}, {-0.500000f, 0.809017f, 0.309017f}, {-0.238856f, 0.864188f, 0.442863f},
    {-0.425325f, 0.688191f, 0.587785f}, {-0.716567f, 0.681718f, -0.147621f}, {-0.500000f, 0.809017f, -0.309017f},
    {-0.525731f, 0.850651f, 0.000000f}, {0.000000f, 0.850651f, -0.525731f}, {-0.238856f, 0.864188f, -0.442863f},
    {0.000000f, 0.955423f, -0.295242f}, {-0.262866f, 0.951056f, -0.162460f}, {0.000000f, 1.000000f, 0.000000f},
    {0.000000f, 0.955423f, 0.295242f}, {-0.262866f, 0.951056f, 0.162460f}, {0.238856f, 0.864188f, 0.442863f},
    {0.262866f, 0.951056f, 0.162460f}, {0.500000f, 0.809017f, 0.309017f}, {0.238856f, 0.864188f, -0.442863f},
    {0.262866f, 0.951056f, -0.162460f}, {0.500000f, 0.809017f, -0.309017f}, {0.850651f, 0.525731f, 0.000000f},
    {0.716567f, 0.681718f, 0.147621f}, {0.716567f, 0.681718f, -0.147621f}, {0.525731f, 0.850651f, 0.000000f},
    {0.425325f, 0.688191f, 0.587785f}, {0.864188f, 0.442863f, 0.238856f}, {0.688191f, 0.587785f, 0.425325f},
    {0.809017f, 0.309017f, 0.500000f}, {0.681718f, 0.147621f, 0.716567f}, {0.587785f, 0.425325f, 0.688191f},
    {0.955423f, 0.295242f, 0.000000f}, {1.000000f, 0.000000f, 0.000000f}, {0.951056f, 0.162460f, 0.262866f},
    {0.850651f, -0.525731f, 0.000000f}, {0.955423f, -0.295242f, 0.000000f}, {0.864188f, -0.442863f, 0.238856f},
    {0.951056f, -0.162460f, 0.262866f}, {0.809017f, -0.309017f, 0.500000f}, {0.681718f, -0.147621f, 0.716567f},
    {0.850651f, 0.000000f, 0.525731f}, {0.864188f, 0.442863f, -0.238856f}, {0.809017f, 0.309017f, -0.500000f},
    {0.951056f, 0.162460f, -0.262866f}, {0.525731f, 0.000000f, -0.850651f}, {0.681718f, 0.147621f, -0.716567f},
    {0.681718f, -0.147621f, -0.716567f}, {0.850651f, 0.000000f, -0.525731f}, {0.809017f, -0.309017f, -0.500000f},
    {0.864188f, -0.442863f, -0.238856f}, {0.951056f, -0.162460f, -0.262866f}, {0.147621f, 0.716567f, -0.681718f},
    {0.309017f, 0.500000f, -0.809017f}, {0.425325f, 0.688191f, -0.587785f}, {0.442863f, 0.238856f, -0.864188f},
    {0.587785f, 0.425325f, -0.688191f}, {0.688191f, 0.587785f, -0.425325f}, {-0.147621f, 0.716567f, -0.681718f},
    {-0.309017f, 0.500000f, -0.809017f}, {0.000000f, 0.525731f, -0.850651f}, {-0.525731f, 0.000000f, -0.850651f},
    {-0.442863f, 0.238856f, -0.864188f}, {-0.295242f, 0.000000f, -0.955423f}, {-0.162460f, 0.262866f, -0.951056f},
    {0.000000f, 0.000000f, -1.000000f}, {0.295242f, 0.000000f, -0.955423f}, {0.162460f, 0.262866f, -0.951056f},
    {-0.442863f, -0.238856f, -0.864188f}, {-0.309017f, -0.500000f, -0.809017f}, {-0.162460f, -0.262866f, -0.951056f},
    {0.000000f, -0.850651f, -0.525731f}, {-0.147621f, -0.716567f, -0.681718f}, {0.147621f, -0.716567f, -0.681718f},
    {0.000000f, -0.525731f, -0.850651f}, {0.309017f, -0.500000f, -0.809017f}, {0.442863f, -0.238856f, -0.864188f},
    {0.162460f, -0.262866f, -0.951056f}, {0.238856f, -0.864188f, -0.442863f}, {0.500000f, -0.809017f, -0.309017f},
    {0.425325f, -0.688191f, -0.587785f}, {0.716567f, -0.681718f, -0.147621f}, {0.688191f, -0.587785f, -0.425325f},
    {0.587785f, -0.425325f, -0.688191f}, {0.000000f, -0.955423f, -0.295242f}, {0.000000f, -1.000000f, 0.000000f},
    {0.262866f, -0.951056f, -0.162460f}, {0.000000f, -0.850651f, 0.525731f}, {0.000000f, -0.955423f, 0.295242f},
    {0.238856f, -0.864188f, 0.442863f}, {0.262866f, -0.951056f, 0.162460f}, {0.500000f, -0.809017f, 0.309017f},
    {0.716567f, -0.681718f, 0.147621f}, {0.525731f, -0.850651f, 0.000000f}, {-0.238856f, -0.864188f, -0.442863f},
    {-0.500000f, -0.809017f, -0.309017f}, {-0.262866f, -0.951056f, -0.162460f}, {-0.850651f, -0.525731f, 0.000000f},
    {-0.716567f, -0.681718f, -0.147621f}, {-0.716567f, -0.681718f, 0.147621f}, {-0.525731f, -0.850651f, 0.000000f},
    {-0.500000f, -0.809017f, 0.309017f}, {-0.238856f, -0.864188f, 0.442863f}, {-0.262866f, -0.951056f, 0.162460f},
    {-0.864188f, -0.442863f, 0.238856f}, {-0.809017f, -0.309017f, 0.500000f}, {-0.688191f, -0.587785f, 0.425325f},
    {-0.681718f, -0.147621f, 0.716567f}, {-0.442863f, -0.238856f, 0.864188f}, {-0.587785f, -0.425325f, 0.688191f},
    {-0.309017f, -0.500000f, 0.809017f}, {-0.147621f, -0.716567f, 0.681718f}, {-0.425325f, -0.688191f, 0.587785f},
    {-0.162460f, -0.262866f, 0.951056f}, {0.442863f, -0.238856f, 0.864188f}, {0.162460f, -0.262866f, 0.951056f},
    {0.309017f, -0.500000f, 0.809017f}, {0.147621f, -0.716567f, 0.681718f}, {0.000000f, -0.525731f, 0.850651f},
    {0.425325f, -0.688191f, 0.587785f}, {0.587785f, -0.425325f, 0.688191f}, {0.688191f, -0.587785f, 0.425325f},
    {-0.955423f, 0.295242f, 0.000000f}, {-0.951056f, 0.162460f, 0.262866f}, {-1.000000f, 0.000000f, 0.000000f},
    {-0.850651f, 0.000000f, 0.525731f}, {-0.955423f, -0.295242f, 0.000000f}, {-0.951056f, -0.162460f, 0.262866f},
    {-0.864188f, 0.442863f, -0.238856f}, {-0.951056f, 0.162460f, -0.262866f}, {-0.809017f, 0.309017f, -0.500000f},
    {-0.864188f, -0.442863f, -0.238856f}, {-0.951056f, -0.162460f, -0.262866f}, {-0.809017f, -0.309017f, -0.500000f},
    {-0.681718f, 0.147621f, -0.716567f}, {-0.681718f, -0.147621f, -0.716567f}, {-0.850651f, 0.000000f, -0.525731f},
    {-0.688191f, 0.587785f, -0.425325f}, {-0.587785f, 0.425325f, -0.688191f}, {-0.425325f, 0.688191f, -0.587785f},
    {-0.425325f, -0.688191f, -0.587785f}, {-0.587785f, -0.425325f, -0.688191f}, {-0.688191f, -0.587785f, -0.425325f},
};

std::vector<float> sky_dirs() {
  return std::vector<float>(&kAnorms[0][0], &kAnorms[0][0] + 162 * 3);
}

// staging for host-pointer calls: one device block carved into arrays
struct Staging {
  DBuf buf;
  std::vector<std::pair<size_t, size_t>> parts;
  int make(const std::vector<size_t> &sizes) {
    size_t off = 0;
    parts.clear();
    for (size_t s : sizes) {
      parts.push_back({off, s});
      off += (s + 255) & ~(size_t)255;
    }
    return buf.alloc(off);
  }
  template <class T>
  T *ptr(int k) const {
    return reinterpret_cast<T *>(buf.as<char>() + parts[(size_t)k].first);
  }
};

}  // namespace
}  // namespace vr

#define LOCK std::lock_guard<std::recursive_mutex> lock_(g_mu)
#define TRY(x)            \
  do {                    \
    int rc_ = (x);        \
    if (rc_) return rc_;  \
  } while (0)

extern "C" {

const char *vr_last_error(void) { return g_err.c_str(); }

int vr_init(int device) {
  LOCK;
  if (g_init) {
    if (device == g_device) return VR_OK;
    return fail(VR_E_STATE, "library already initialised on another device");
  }
  g_device = device;
  return ensure_init();
}

int vr_set_path(int path) {
  LOCK;
  if (path != VR_PATH_OPT && path != VR_PATH_PLAIN) return fail(VR_E_INVALID, "unknown path");
  g_path = path;
  return VR_OK;
}

int vr_set_tuning(const char *name, int value) {
  LOCK;
  if (!name) return fail(VR_E_INVALID, "name is NULL");
  const std::string k(name);
  if (k == "trace_lds_entries") {
    if (value != 2 && value != 4 && value != 6 && value != 8 && value != 10 && value != 12 &&
        value != 16 && value != 24)
      return fail(VR_E_INVALID, "trace_lds_entries must be 2, 4, 6, 8, 10, 12, 16 or 24");
    tuning().trace_lds_entries = value;
  } else if (k == "trace_rays_per_lane") {
    if (value != 1 && value != 2 && value != 4 && value != 8)
      return fail(VR_E_INVALID, "trace_rays_per_lane must be 1, 2, 4 or 8");
    tuning().trace_rays_per_lane = value;
  } else if (k == "trace_xcd_group") {
    if (value < 0 || value > 65536) return fail(VR_E_INVALID, "trace_xcd_group must be 0..65536");
    tuning().trace_xcd_group = value;
  } else if (k == "bounce_waves") {
    if (value < 1 || value > (1 << 20)) return fail(VR_E_INVALID, "bounce_waves out of range");
    tuning().bounce_waves = value;
  } else if (k == "bounce_kernel") {
    if (value < 0 || value > 2) return fail(VR_E_INVALID, "bounce_kernel must be 0, 1 or 2");
    tuning().bounce_kernel = value;
  } else if (k == "bounce_tile_cols") {
    if (value != 2048 && value != 4096)
      return fail(VR_E_INVALID, "bounce_tile_cols must be 2048 or 4096");
    tuning().bounce_tile_cols = value;
  } else if (k == "bounce_group") {
    if (value != 0 && value != 8 && value != 16)
      return fail(VR_E_INVALID, "bounce_group must be 0 (by precision), 8 or 16");
    tuning().bounce_group = value;
  } else if (k == "bounce_prefetch") {
    if (value < 0 || value > 7)
      return fail(VR_E_INVALID, "bounce_prefetch must be 0 (by kernel) to 7");
    tuning().bounce_prefetch = value;
  } else if (k == "bounce_placement") {
    if (value < 0 || value > 2)
      return fail(VR_E_INVALID,
                  "bounce_placement must be 0 (strided), 1 (row runs) or 2 (strided, by index)");
    tuning().bounce_placement = value;
  } else if (k == "bounce_passes") {
    if (value < 0 || value > 2)
      return fail(VR_E_INVALID, "bounce_passes must be 0 (by precision), 1 or 2");
    tuning().bounce_passes = value;
  } else if (k == "kd_build") {
    if (value < 0 || value > 2)
      return fail(VR_E_INVALID, "kd_build must be 0 (host), 1 (device) or 2 (by size)");
    tuning().kd_build = value;
  } else if (k == "kd_build_device_min") {
    if (value < 0) return fail(VR_E_INVALID, "kd_build_device_min must be >= 0");
    tuning().kd_build_device_min = value;
  } else if (k == "patch_build") {
    if (value < 0 || value > 1)
      return fail(VR_E_INVALID, "patch_build must be 0 (a thread per face) or 1 (level-parallel)");
    tuning().patch_build = value;
  } else {
    return fail(VR_E_INVALID, "unknown tuning knob " + k);
  }
  return VR_OK;
}

int vr_get_tuning(const char *name, int *value) {
  LOCK;
  if (!name || !value) return fail(VR_E_INVALID, "NULL argument");
  const std::string k(name);
  if (k == "trace_lds_entries") *value = tuning().trace_lds_entries;
  else if (k == "trace_rays_per_lane") *value = tuning().trace_rays_per_lane;
  else if (k == "trace_xcd_group") *value = tuning().trace_xcd_group;
  else if (k == "bounce_waves") *value = tuning().bounce_waves;
  else if (k == "bounce_kernel") *value = tuning().bounce_kernel;
  else if (k == "bounce_tile_cols") *value = tuning().bounce_tile_cols;
  else if (k == "bounce_group") *value = tuning().bounce_group;
  else if (k == "bounce_passes") *value = tuning().bounce_passes;
  else if (k == "bounce_prefetch") *value = tuning().bounce_prefetch;
  else if (k == "bounce_placement") *value = tuning().bounce_placement;
  else if (k == "kd_build") *value = tuning().kd_build;
  else if (k == "kd_build_device_min") *value = tuning().kd_build_device_min;
  else if (k == "patch_build") *value = tuning().patch_build;
  else return fail(VR_E_INVALID, "unknown tuning knob " + k);
  return VR_OK;
}

int vr_profile_enable(int on) {
  LOCK;
  g_prof = on != 0;
  return VR_OK;
}

int vr_profile_reset(void) {
  LOCK;
  prof_drain();
  g_prof_tot.clear();
  return VR_OK;
}

int vr_profile_get(const char *name, double *total_ms, int64_t *launches) {
  LOCK;
  if (!name) return fail(VR_E_INVALID, "name is NULL");
  prof_drain();
  auto it = g_prof_tot.find(name);
  if (total_ms) *total_ms = it == g_prof_tot.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == g_prof_tot.end() ? 0 : it->second.second;
  return VR_OK;
}

int vr_measure_read_bandwidth(int64_t bytes, int32_t reps, double *gbps) {
  LOCK;
  if (!gbps || bytes < (1 << 20) || reps < 1) return fail(VR_E_INVALID, "bad read-probe arguments");
  TRY(ensure_init());
  bytes &= ~(int64_t)15;
  const int blocks = 2048;
  DBuf in, out;
  TRY(in.alloc((size_t)bytes));
  TRY(out.alloc(4 * (size_t)blocks * 256));
  hipStream_t st = g_stream;
  VR_HIP(hipMemsetAsync(in.p, 0, (size_t)bytes, st));
  hipEvent_t e0, e1;
  VR_HIP(hipEventCreate(&e0));
  VR_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int rc = VR_OK;
  for (int r = 0; r <= reps && rc == VR_OK; ++r) {  // the first launch is untimed
    hipError_t e = hipEventRecord(e0, st);
    launch_read_probe(in.p, bytes, out.as<float>(), blocks, st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) rc = fail(VR_E_HIP, std::string("read probe: ") + hipGetErrorString(e));
    if (r > 0 && ms < best) best = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == VR_OK) *gbps = (double)bytes / (best * 1e-3) / 1e9;
  return rc;
}

// ---- scene ----------------------------------------------------------------
int vr_scene_create(vr_scene **out, const float *verts9, const int32_t *ids, const uint8_t *flags,
                    int64_t n_tri) {
  if (!out) return fail(VR_E_INVALID, "out is NULL");
  *out = nullptr;
  if (n_tri < 0) return fail(VR_E_INVALID, "n_tri < 0");
  if (n_tri > 0 && (!verts9 || !ids)) return fail(VR_E_INVALID, "verts9/ids are NULL");
  if (n_tri >= (int64_t)kTriMask) return fail(VR_E_CAPACITY, "more than 2^28 triangles");
  std::unique_ptr<vr_scene> s(new (std::nothrow) vr_scene);
  if (!s) return fail(VR_E_NOMEM, "scene");
  s->n_tri = n_tri;
  s->v9.assign(verts9, verts9 + 9 * n_tri);
  s->ids.assign(ids, ids + n_tri);
  if (flags) s->flags.assign(flags, flags + n_tri);
  else s->flags.assign((size_t)n_tri, 0);
  *out = s.release();
  return VR_OK;
}

// Triangle intersection records and both kd-trees of a scene; the trees on
// the host threads or, with `device`, on the GPU (the same trees).
static int build_scene_trees(vr_scene *s, int32_t n_threads, bool device) {
  if (s->host_built) return VR_OK;
  const int nt = host_threads(n_threads);
  s->tris.resize((size_t)s->n_tri);
  precompute_triangles(s->v9.data(), s->ids.data(), s->flags.data(), s->n_tri, s->tris.data(), nt);
  // kd_build 2: each tree by its own size (the sky tree is usually a few
  // dozen triangles, where the host is faster than ~20 levels of launches)
  auto build = [&](const float *v9, int64_t n, KdTree *out) {
    const bool dev = device && (tuning().kd_build != 2 || n >= tuning().kd_build_device_min);
    if (dev) return build_kdtree_device(v9, n, g_stream, out);
    return build_kdtree(v9, n, nt, out);
  };
  s->kd = KdTree();
  TRY(build(s->v9.data(), s->n_tri, &s->kd));
  if (s->kd.max_depth >= kStackDepth)
    return fail(VR_E_CAPACITY, "kd-tree deeper than the traversal stack");
  if ((int64_t)s->kd.children.size() >= ((int64_t)1 << kStackNodeBits))
    return fail(VR_E_CAPACITY, "kd-tree of 2^24 nodes or more (traversal-stack records)");
  s->has_sky = 0;
  for (const TriIntersect &t : s->tris)
    if (t.id & VR_TRACE_ID_SKY) s->has_sky = 1;
  // a second tree over the sky triangles alone: TestLine finds the closest
  // sky hit there, then stops the full tree at the first opaque hit before it
  s->kd_sky = KdTree();
  s->sky_index.clear();
  if (s->has_sky) {
    std::vector<float> v9s;
    for (int64_t i = 0; i < s->n_tri; ++i)
      if (s->tris[(size_t)i].id & VR_TRACE_ID_SKY) {
        s->sky_index.push_back((int32_t)i);
        v9s.insert(v9s.end(), s->v9.begin() + 9 * i, s->v9.begin() + 9 * i + 9);
      }
    TRY(build(v9s.data(), (int64_t)s->sky_index.size(), &s->kd_sky));
    if (s->kd_sky.max_depth >= kStackDepth)
      return fail(VR_E_CAPACITY, "sky kd-tree deeper than the traversal stack");
    if ((int64_t)s->kd_sky.children.size() >= ((int64_t)1 << kStackNodeBits))
      return fail(VR_E_CAPACITY, "sky kd-tree of 2^24 nodes or more");
  }
  s->host_built = true;
  return VR_OK;
}

int vr_scene_build_host(vr_scene *s, int32_t n_threads) {
  LOCK;
  if (!s) return fail(VR_E_INVALID, "scene is NULL");
  return build_scene_trees(s, n_threads, false);
}

int vr_scene_build(vr_scene *s, int32_t n_threads) {
  LOCK;
  if (!s) return fail(VR_E_INVALID, "scene is NULL");
  TRY(ensure_init());
  const int kb = tuning().kd_build;
  const bool device = kb == 1 || (kb == 2 && s->n_tri >= tuning().kd_build_device_min);
  TRY(build_scene_trees(s, n_threads, device));
  const int nt = host_threads(n_threads);
  const KdTree &kd = s->kd;
  // device records of a tree: nodes, and 3 float4 per leaf reference whose
  // meta holds the triangle's index in OptimizedTriangleList (`map` turns a
  // tree's own triangle numbering into that index)
  auto flatten = [&](const KdTree &t, const std::vector<int32_t> *map, std::vector<uint2> &nodes,
                     std::vector<float4> &refs) {
    nodes.resize(t.children.size());
    parallel_for((int64_t)nodes.size(), nt, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; ++k) {
        uint32_t sb;
        std::memcpy(&sb, &t.split[(size_t)k], 4);
        nodes[(size_t)k] = make_uint2(t.children[(size_t)k], sb);
      }
    });
    refs.resize(t.idx.size() * 3);
    parallel_for((int64_t)t.idx.size(), nt, [&](int64_t b, int64_t e) {
      for (int64_t k = b; k < e; ++k) {
        const int32_t ti = map ? (*map)[(size_t)t.idx[(size_t)k]] : t.idx[(size_t)k];
        const TriIntersect &tr = s->tris[(size_t)ti];
        uint32_t meta = (uint32_t)ti | ((uint32_t)tr.c0 << 28) | ((uint32_t)tr.c1 << 30);
        float idf, metaf;
        std::memcpy(&idf, &tr.id, 4);
        std::memcpy(&metaf, &meta, 4);
        refs[(size_t)(3 * k + 0)] = make_float4(tr.nx, tr.ny, tr.nz, tr.d);
        refs[(size_t)(3 * k + 1)] = make_float4(tr.e[0], tr.e[1], tr.e[2], tr.e[3]);
        refs[(size_t)(3 * k + 2)] = make_float4(tr.e[4], tr.e[5], idf, metaf);
      }
    });
  };
  std::vector<uint2> nodes, sky_nodes;
  std::vector<float4> refs, sky_refs;
  flatten(kd, nullptr, nodes, refs);
  if (s->has_sky) {
    flatten(s->kd_sky, &s->sky_index, sky_nodes, sky_refs);
    TRY(upload_vec(s->sky_nodes, sky_nodes, g_stream));
    TRY(upload_vec(s->sky_refs, sky_refs, g_stream));
  }
  std::vector<float4> planes((size_t)s->n_tri);
  parallel_for(s->n_tri, nt, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const TriIntersect &t = s->tris[(size_t)i];
      planes[(size_t)i] = make_float4(t.nx, t.ny, t.nz, t.d);
    }
  });
  TRY(upload_vec(s->nodes, nodes, g_stream));
  TRY(upload_vec(s->refs, refs, g_stream));
  TRY(upload_vec(s->planes, planes, g_stream));
  VR_HIP(hipStreamSynchronize(g_stream));
  DevScene d{};
  d.nodes = s->nodes.as<uint2>();
  d.refs = s->refs.as<float4>();
  d.planes = s->planes.as<float4>();
  float M = 1.0f;
  for (int c = 0; c < 3; ++c) {
    d.bmin[c] = kd.bmin[c];
    d.bmax[c] = kd.bmax[c];
    M = std::max(M, std::max(std::fabs(kd.bmin[c]), std::fabs(kd.bmax[c])));
  }
  d.M = M;
  d.has_sky = s->has_sky;
  d.empty = s->n_tri == 0 ? 1 : 0;
  if (s->has_sky) {
    d.sky_nodes = s->sky_nodes.as<uint2>();
    d.sky_refs = s->sky_refs.as<float4>();
    for (int c = 0; c < 3; ++c) {
      d.sky_bmin[c] = s->kd_sky.bmin[c];
      d.sky_bmax[c] = s->kd_sky.bmax[c];
    }
  }
  s->dev = d;
  s->device_bytes = (int64_t)(s->nodes.bytes + s->refs.bytes + s->planes.bytes +
                              s->sky_nodes.bytes + s->sky_refs.bytes);
  s->built = true;
  return VR_OK;
}

int vr_scene_get_info(const vr_scene *s, vr_scene_info *out) {
  if (!s || !out) return fail(VR_E_INVALID, "NULL argument");
  std::memset(out, 0, sizeof *out);
  out->n_tri = s->n_tri;
  out->n_nodes = (int64_t)s->kd.children.size();
  out->n_leaf_refs = (int64_t)s->kd.idx.size();
  out->max_depth = s->kd.max_depth;
  out->has_sky = s->has_sky;
  for (int c = 0; c < 3; ++c) {
    out->bmin[c] = s->kd.bmin[c];
    out->bmax[c] = s->kd.bmax[c];
  }
  out->device_bytes = s->device_bytes;
  return VR_OK;
}

int vr_scene_get_kdtree(const vr_scene *s, uint32_t *children, float *split,
                        int32_t *triangle_index_list) {
  if (!s) return fail(VR_E_INVALID, "scene is NULL");
  if (!s->host_built) return fail(VR_E_STATE, "scene not built");
  if (children) std::memcpy(children, s->kd.children.data(), 4 * s->kd.children.size());
  if (split) std::memcpy(split, s->kd.split.data(), 4 * s->kd.split.size());
  if (triangle_index_list)
    std::memcpy(triangle_index_list, s->kd.idx.data(), 4 * s->kd.idx.size());
  return VR_OK;
}

int vr_scene_get_triangles(const vr_scene *s, void *out) {
  if (!s || !out) return fail(VR_E_INVALID, "NULL argument");
  if (!s->host_built) return fail(VR_E_STATE, "scene not built");
  std::memcpy(out, s->tris.data(), sizeof(TriIntersect) * s->tris.size());
  return VR_OK;
}

int vr_scene_destroy(vr_scene *s) {
  LOCK;
  delete s;
  return VR_OK;
}

// ---- visibility ---------------------------------------------------------
int vr_trace_closest_dev(vr_scene *s, const vr_rays *r, int64_t n, int32_t skip_id,
                         int32_t *hit, float *hit_t, float *hit_n, void *stream) {
  LOCK;
  if (!s || !r) return fail(VR_E_INVALID, "NULL argument");
  if (!s->built) return fail(VR_E_STATE, "scene not built");
  if (n < 0) return fail(VR_E_INVALID, "n < 0");
  if (n > 0 && (!hit || !hit_t)) return fail(VR_E_INVALID, "outputs are NULL");
  TRY(ensure_init());
  hipStream_t st = stream ? (hipStream_t)stream : g_stream;
  DevRays dr{r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->tmin, r->tmax};
  {
    Prof p("trace_closest", st);
    launch_trace_closest(s->dev, dr, n, skip_id, hit, hit_t, hit_n, g_path, st);
  }
  return check_launch();
}

int vr_test_line_dev(vr_scene *s, const vr_segs *g, int64_t n, int32_t skip_id, uint8_t *vis,
                     uint8_t *sky, void *stream) {
  LOCK;
  if (!s || !g) return fail(VR_E_INVALID, "NULL argument");
  if (!s->built) return fail(VR_E_STATE, "scene not built");
  if (n < 0) return fail(VR_E_INVALID, "n < 0");
  if (n > 0 && !vis) return fail(VR_E_INVALID, "visible is NULL");
  TRY(ensure_init());
  hipStream_t st = stream ? (hipStream_t)stream : g_stream;
  DevSegs ds{g->sx, g->sy, g->sz, g->ex, g->ey, g->ez};
  {
    Prof p("test_line", st);
    launch_test_line(s->dev, ds, n, skip_id, vis, sky, g_path, st);
  }
  return check_launch();
}

int vr_trace_closest(vr_scene *s, const vr_rays *r, int64_t n, int32_t skip_id, int32_t *hit,
                     float *hit_t, float *hit_n) {
  LOCK;
  if (!s || !r) return fail(VR_E_INVALID, "NULL argument");
  if (!s->built) return fail(VR_E_STATE, "scene not built");
  if (n < 0) return fail(VR_E_INVALID, "n < 0");
  if (n == 0) return VR_OK;
  if (!hit || !hit_t || !r->ox || !r->oy || !r->oz || !r->dx || !r->dy || !r->dz || !r->tmin ||
      !r->tmax)
    return fail(VR_E_INVALID, "NULL ray array or output");
  TRY(ensure_init());
  const size_t fb = 4 * (size_t)n;
  Staging S;
  TRY(S.make({fb, fb, fb, fb, fb, fb, fb, fb, fb, fb, hit_n ? 3 * fb : 0}));
  const float *src[8] = {r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->tmin, r->tmax};
  for (int k = 0; k < 8; ++k)
    VR_HIP(hipMemcpyAsync(S.ptr<float>(k), src[k], fb, hipMemcpyHostToDevice, g_stream));
  vr_rays dr{S.ptr<float>(0), S.ptr<float>(1), S.ptr<float>(2), S.ptr<float>(3),
             S.ptr<float>(4), S.ptr<float>(5), S.ptr<float>(6), S.ptr<float>(7)};
  TRY(vr_trace_closest_dev(s, &dr, n, skip_id, S.ptr<int32_t>(8), S.ptr<float>(9),
                           hit_n ? S.ptr<float>(10) : nullptr, nullptr));
  VR_HIP(hipMemcpyAsync(hit, S.ptr<int32_t>(8), fb, hipMemcpyDeviceToHost, g_stream));
  VR_HIP(hipMemcpyAsync(hit_t, S.ptr<float>(9), fb, hipMemcpyDeviceToHost, g_stream));
  if (hit_n) VR_HIP(hipMemcpyAsync(hit_n, S.ptr<float>(10), 3 * fb, hipMemcpyDeviceToHost, g_stream));
  VR_HIP(hipStreamSynchronize(g_stream));
  return VR_OK;
}

int vr_test_line(vr_scene *s, const vr_segs *g, int64_t n, int32_t skip_id, uint8_t *vis,
                 uint8_t *sky) {
  LOCK;
  if (!s || !g) return fail(VR_E_INVALID, "NULL argument");
  if (!s->built) return fail(VR_E_STATE, "scene not built");
  if (n < 0) return fail(VR_E_INVALID, "n < 0");
  if (n == 0) return VR_OK;
  if (!vis || !g->sx || !g->sy || !g->sz || !g->ex || !g->ey || !g->ez)
    return fail(VR_E_INVALID, "NULL segment array or output");
  TRY(ensure_init());
  const size_t fb = 4 * (size_t)n;
  Staging S;
  TRY(S.make({fb, fb, fb, fb, fb, fb, (size_t)n, (size_t)n}));
  const float *src[6] = {g->sx, g->sy, g->sz, g->ex, g->ey, g->ez};
  for (int k = 0; k < 6; ++k)
    VR_HIP(hipMemcpyAsync(S.ptr<float>(k), src[k], fb, hipMemcpyHostToDevice, g_stream));
  vr_segs ds{S.ptr<float>(0), S.ptr<float>(1), S.ptr<float>(2),
             S.ptr<float>(3), S.ptr<float>(4), S.ptr<float>(5)};
  TRY(vr_test_line_dev(s, &ds, n, skip_id, S.ptr<uint8_t>(6), sky ? S.ptr<uint8_t>(7) : nullptr,
                       nullptr));
  VR_HIP(hipMemcpyAsync(vis, S.ptr<uint8_t>(6), (size_t)n, hipMemcpyDeviceToHost, g_stream));
  if (sky) VR_HIP(hipMemcpyAsync(sky, S.ptr<uint8_t>(7), (size_t)n, hipMemcpyDeviceToHost, g_stream));
  VR_HIP(hipStreamSynchronize(g_stream));
  return VR_OK;
}

// ---- transfers ------------------------------------------------------------
__global__ void k_add_count(const uint32_t *c, uint32_t cap, unsigned long long *tot) {
  atomicAdd(tot, (unsigned long long)min(*c, cap));  // rounds on two streams add here
}

// The candidate-pair plan of a PVS: patches grouped by cluster, the group
// pairs that see each other, and their block prefix (k_pair_cull_pvs).
struct PvsPlan {
  std::vector<int32_t> mem;
  std::vector<vr::PvsEntry> tab;
  std::vector<int64_t> blk_pre;
  int64_t pairs = 0;
};

static int pvs_plan(const int32_t *cl, int64_t np, const uint8_t *rows, int32_t nc, PvsPlan *P) {
  const int32_t ng = nc + 1;  // group nc: patches with a negative cluster
  std::vector<int64_t> goff((size_t)ng + 1, 0);
  for (int64_t i = 0; i < np; ++i) {
    if (cl[i] >= nc) return fail(VR_E_INVALID, "patch " + std::to_string(i) + ": cluster >= n_clusters");
    ++goff[(size_t)(cl[i] < 0 ? nc : cl[i]) + 1];
  }
  for (int32_t g = 0; g < ng; ++g) goff[(size_t)g + 1] += goff[(size_t)g];
  P->mem.assign((size_t)np, 0);
  {
    std::vector<int64_t> at(goff.begin(), goff.end() - 1);
    for (int64_t i = 0; i < np; ++i) P->mem[(size_t)at[(size_t)(cl[i] < 0 ? nc : cl[i])]++] = (int32_t)i;
  }
  const int64_t row = ((int64_t)nc + 7) >> 3;
  auto bit = [&](int32_t a, int32_t b) { return (rows[(int64_t)a * row + (b >> 3)] >> (b & 7)) & 1; };
  const int64_t bc = vr::pvs_block_cols();
  P->tab.clear();
  P->blk_pre.assign(1, 0);
  P->pairs = 0;
  for (int32_t a = 0; a < ng; ++a) {
    const int64_t na = goff[(size_t)a + 1] - goff[(size_t)a];
    if (na == 0) continue;
    for (int32_t b = a; b < ng; ++b) {
      const int64_t nb = goff[(size_t)b + 1] - goff[(size_t)b];
      if (nb == 0) continue;
      if (a < nc && b < nc && !bit(a, b) && !bit(b, a)) continue;
      vr::PvsEntry e;
      e.a_off = (int32_t)goff[(size_t)a];
      e.a_n = (int32_t)na;
      e.b_off = (int32_t)goff[(size_t)b];
      e.b_n = (int32_t)nb;
      e.cpr = (int32_t)((nb + bc - 1) / bc);
      e.diag = a == b;
      P->tab.push_back(e);
      P->blk_pre.push_back(P->blk_pre.back() + na * e.cpr);
      P->pairs += a == b ? na * (na - 1) / 2 : na * nb;
    }
  }
  return VR_OK;
}

// The CSR of a finished visibility bit matrix: row counts, padded layout,
// columns and form factors, then the optional row normalisation.
static int transfers_from_bits(vr_xfers *x, const DBuf &bits, int64_t words,
                               const DevPatches &dp, int32_t normalize, hipStream_t st) {
  const int64_t np = x->n;
  TRY(x->cnt.alloc(4 * (size_t)std::max<int64_t>(np, 1)));
  launch_row_count(bits.as<uint32_t>(), np, words, x->cnt.as<int32_t>(), st);
  TRY(check_launch());
  x->cnt_host.assign((size_t)np, 0);
  if (np > 0)
    VR_HIP(hipMemcpyAsync(x->cnt_host.data(), x->cnt.p, 4 * (size_t)np, hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  x->nnz = 0;
  for (int32_t c : x->cnt_host) x->nnz += c;
  TRY(finish_layout(x, st));
  TRY(x->col.alloc(4 * (size_t)x->stored));
  TRY(x->coef32.alloc(4 * (size_t)x->stored));
  launch_row_fill(bits.as<uint32_t>(), np, words, dp, x->rp.as<int64_t>(), x->cnt.as<int32_t>(),
                  x->col.as<int32_t>(), x->coef32.as<float>(), st);
  DBuf rowsum;
  TRY(rowsum.alloc(8 * (size_t)std::max<int64_t>(np, 1)));
  launch_row_normalize(np, x->rp.as<int64_t>(), x->coef32.as<float>(), rowsum.as<double>(),
                       normalize, st);
  TRY(check_launch());
  VR_HIP(hipStreamSynchronize(st));
  return VR_OK;
}

static int make_transfers_impl(vr_scene *s, const vr_patches *p, int64_t np,
                               const vr_xfer_params *prm, const int32_t *cl,
                               const uint8_t *pvs, int32_t nc, vr_xfers **out) {
  if (!s || !p || !out) return fail(VR_E_INVALID, "NULL argument");
  *out = nullptr;
  if (!s->built) return fail(VR_E_STATE, "scene not built");
  if (np < 0 || np >= (1ll << 31)) return fail(VR_E_INVALID, "bad patch count");
  if (np > 0 && (!p->ox || !p->oy || !p->oz || !p->nx || !p->ny || !p->nz || !p->area))
    return fail(VR_E_INVALID, "NULL patch array");
  TRY(ensure_init());
  vr_xfer_params P{0.1f, -1, 1, 0};
  if (prm) P = *prm;
  // a batch holds at most every pair of the scene, and at least one row of
  // the unculled enumeration or one block of the PVS one
  const int64_t all_pairs = np * (np - 1) / 2;
  const int64_t cap = std::max<int64_t>(
      std::min<int64_t>(P.batch_pairs > 0 ? P.batch_pairs : (int64_t)32 << 20, all_pairs),
      std::max<int64_t>(np - 1, nc > 0 ? (int64_t)vr::pvs_block_cols() : 1));
  if (cap >= (1ll << 31)) return fail(VR_E_INVALID, "batch_pairs too large");
  hipStream_t st = g_stream;
  std::unique_ptr<vr_xfers> x(new vr_xfers);
  x->n = np;
  const size_t fb = 4 * (size_t)std::max<int64_t>(np, 1);
  Staging S;
  TRY(S.make({fb, fb, fb, fb, fb, fb, fb}));
  const float *src[7] = {p->ox, p->oy, p->oz, p->nx, p->ny, p->nz, p->area};
  for (int k = 0; k < 7 && np > 0; ++k)
    VR_HIP(hipMemcpyAsync(S.ptr<float>(k), src[k], 4 * (size_t)np, hipMemcpyHostToDevice, st));
  DevPatches dp{S.ptr<float>(0), S.ptr<float>(1), S.ptr<float>(2), S.ptr<float>(3),
                S.ptr<float>(4), S.ptr<float>(5), S.ptr<float>(6)};
  const int64_t words = (np + 31) / 32;
  // The optimised path alternates rounds between two streams, each with its
  // own pair buffer and count: a round's cull and the start of its trace
  // run beside the end of the previous round's trace (the launch's tail).
  // Both set bits of one matrix (atomicOr) and add to one segment total.
  const int n_sides = g_path == VR_PATH_PLAIN ? 1 : 2;
  DBuf bits, pairs[2], count[2], segtot;
  TRY(bits.alloc((size_t)np * (size_t)words * 4));
  VR_HIP(hipMemsetAsync(bits.p, 0, bits.bytes, st));
  for (int q = 0; q < n_sides; ++q) {
    TRY(pairs[q].alloc((size_t)cap * sizeof(int2)));
    TRY(count[q].alloc(4));
  }
  TRY(segtot.alloc(8));
  VR_HIP(hipMemsetAsync(segtot.p, 0, 8, st));
  hipStream_t side_st[2] = {st, g_stream2};
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  if (n_sides == 2) {
    // the second stream starts after the clears on the first, and the first
    // goes on (CSR build, segment total) after everything on the second
    VR_HIP(hipEventCreateWithFlags(&ev_start, hipEventDisableTiming));
    if (hipEventCreateWithFlags(&ev_end, hipEventDisableTiming) != hipSuccess) {
      (void)hipEventDestroy(ev_start);
      return fail(VR_E_HIP, "hipEventCreate failed");
    }
  }
  struct Events {
    hipEvent_t a, b;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } events{ev_start, ev_end};
  // the second stream waits for everything queued on the first so far (the
  // clears, the patch and PVS uploads); called before the first round
  auto start_second = [&]() -> int {
    if (n_sides == 2) {
      VR_HIP(hipEventRecord(ev_start, st));
      VR_HIP(hipStreamWaitEvent(g_stream2, ev_start, 0));
    }
    return VR_OK;
  };
  int round = 0;
  {
    Prof pr(nc > 0 ? "make_transfers_pvs" : "make_transfers", st);
    // the buffers and stream of the current round
    auto side = [&]() { return round % n_sides; };
    auto trace_batch = [&]() {
      const int q = side();
      launch_pair_trace(s->dev, dp, pairs[q].as<int2>(), count[q].as<uint32_t>(),
                        (uint32_t)cap, P.normal_offset, P.skip_id, bits.as<uint32_t>(), words,
                        g_path, side_st[q]);
      hipLaunchKernelGGL(k_add_count, dim3(1), dim3(1), 0, side_st[q], count[q].as<uint32_t>(),
                         (uint32_t)cap, segtot.as<unsigned long long>());
      ++round;
      return check_launch();
    };
    if (nc > 0) {
      PvsPlan plan;
      TRY(pvs_plan(cl, np, pvs, nc, &plan));
      x->pairs = plan.pairs;
      const int64_t n_blk = plan.blk_pre.back();
      if (n_blk > 0) {
        DBuf mem, tab, pre;
        TRY(mem.upload(plan.mem.data(), 4 * plan.mem.size(), st));
        TRY(tab.upload(plan.tab.data(), sizeof(vr::PvsEntry) * plan.tab.size(), st));
        TRY(pre.upload(plan.blk_pre.data(), 8 * plan.blk_pre.size(), st));
        // a block keeps at most pvs_block_cols() pairs
        const int64_t per = std::max<int64_t>(1, cap / vr::pvs_block_cols());
        TRY(start_second());
        for (int64_t b0 = 0; b0 < n_blk; b0 += per) {
          const int q = side();
          VR_HIP(hipMemsetAsync(count[q].p, 0, 4, side_st[q]));
          launch_pair_cull_pvs(dp, mem.as<int32_t>(), tab.as<vr::PvsEntry>(), pre.as<int64_t>(),
                               (int32_t)plan.tab.size(), b0, std::min(per, n_blk - b0),
                               pairs[q].as<int2>(), count[q].as<uint32_t>(), (uint32_t)cap,
                               side_st[q]);
          TRY(trace_batch());
        }
        // mem / tab / pre are freed on return
        for (int q = 0; q < n_sides; ++q) VR_HIP(hipStreamSynchronize(side_st[q]));
      }
    } else {
      TRY(start_second());
      int64_t i0 = 0;
      while (i0 + 1 < np) {
        const int64_t rowlen = np - 1 - i0;
        int64_t rows = std::max<int64_t>(1, cap / rowlen);
        const int64_t i1 = std::min<int64_t>(np - 1, i0 + rows);
        x->pairs += (2 * np - i0 - i1 - 1) * (i1 - i0) / 2;
        const int q = side();
        VR_HIP(hipMemsetAsync(count[q].p, 0, 4, side_st[q]));
        launch_pair_cull(dp, np, i0, i1, pairs[q].as<int2>(), count[q].as<uint32_t>(),
                         (uint32_t)cap, side_st[q]);
        TRY(trace_batch());
        i0 = i1;
      }
    }
    if (n_sides == 2) {
      VR_HIP(hipEventRecord(ev_end, g_stream2));
      VR_HIP(hipStreamWaitEvent(st, ev_end, 0));
    }
    unsigned long long segs = 0;
    VR_HIP(hipMemcpyAsync(&segs, segtot.p, 8, hipMemcpyDeviceToHost, st));
    VR_HIP(hipStreamSynchronize(st));
    x->segments = (int64_t)segs;
    TRY(transfers_from_bits(x.get(), bits, words, dp, P.normalize, st));
  }
  *out = x.release();
  return VR_OK;
}

int vr_make_transfers(vr_scene *s, const vr_patches *p, int64_t np, const vr_xfer_params *prm,
                      vr_xfers **out) {
  LOCK;
  return make_transfers_impl(s, p, np, prm, nullptr, nullptr, 0, out);
}

static int make_transfers_shared(vr_scene *const *scenes, const vr_patches *patches,
                                 const int64_t *n_patches, const std::vector<int32_t> &which,
                                 const vr_xfer_params &P, int64_t budget, vr_xfers **out);

int vr_make_transfers_batch(vr_scene *const *scenes, const vr_patches *patches,
                            const int64_t *n_patches, int32_t n_scenes,
                            const vr_xfer_params *prm, vr_xfers **out) {
  LOCK;
  if (n_scenes < 0) return fail(VR_E_INVALID, "n_scenes < 0");
  if (n_scenes == 0) return VR_OK;
  if (!scenes || !patches || !n_patches || !out) return fail(VR_E_INVALID, "NULL argument");
  for (int32_t k = 0; k < n_scenes; ++k) out[k] = nullptr;
  if (g_path == VR_PATH_PLAIN) {  // the comparison path: one scene at a time
    for (int32_t k = 0; k < n_scenes; ++k) {
      const int rc = make_transfers_impl(scenes[k], &patches[k], n_patches[k], prm, nullptr,
                                         nullptr, 0, &out[k]);
      if (rc) {
        for (int32_t q = 0; q < k; ++q) {
          delete out[q];
          out[q] = nullptr;
        }
        return rc;
      }
    }
    return VR_OK;
  }
  for (int32_t k = 0; k < n_scenes; ++k) {
    const vr_scene *sc = scenes[k];
    const vr_patches *p = &patches[k];
    const int64_t np = n_patches[k];
    const std::string who = "scene " + std::to_string(k) + ": ";
    if (!sc) return fail(VR_E_INVALID, who + "NULL scene");
    if (!sc->built) return fail(VR_E_STATE, who + "scene not built");
    if (np < 0 || np >= (1ll << 31)) return fail(VR_E_INVALID, who + "bad patch count");
    if (np > 0 && (!p->ox || !p->oy || !p->oz || !p->nx || !p->ny || !p->nz || !p->area))
      return fail(VR_E_INVALID, who + "NULL patch array");
  }
  TRY(ensure_init());
  vr_xfer_params P{0.1f, -1, 1, 0};
  if (prm) P = *prm;
  const int64_t budget = P.batch_pairs > 0 ? P.batch_pairs : (int64_t)32 << 20;
  // A scene with a whole round's budget of candidate pairs fills the GPU on
  // its own: it takes the single-scene path (the same result), one launch
  // per round of its own, and the shared launches serve the smaller scenes.
  std::vector<int32_t> shared;
  int rc = VR_OK;
  for (int32_t k = 0; k < n_scenes && rc == VR_OK; ++k) {
    const int64_t np = n_patches[k];
    if (np * (np - 1) / 2 >= budget)
      rc = make_transfers_impl(scenes[k], &patches[k], np, &P, nullptr, nullptr, 0, &out[k]);
    else
      shared.push_back(k);
  }
  if (rc == VR_OK) rc = make_transfers_shared(scenes, patches, n_patches, shared, P, budget, out);
  if (rc != VR_OK) {
    for (int32_t k = 0; k < n_scenes; ++k) {
      delete out[k];
      out[k] = nullptr;
    }
  }
  return rc;
}

// vr_make_transfers_batch's shared launches for the scenes in `which`.
static int make_transfers_shared(vr_scene *const *scenes, const vr_patches *patches,
                                 const int64_t *n_patches, const std::vector<int32_t> &which,
                                 const vr_xfer_params &P, int64_t budget, vr_xfers **out) {
  hipStream_t st = g_stream;
  const int32_t n_scenes = (int32_t)which.size();
  if (n_scenes == 0) return VR_OK;
  // every round, each unfinished scene culls its next rows into its own
  // region of one pair buffer, and the regions' pairs are traced in one
  // launch.  The round's budget is shared among the scenes still culling
  // (at least one whole row each), so the rounds do not outlast the small
  // scenes with small shares.
  struct One {
    std::unique_ptr<vr_xfers> x;
    Staging S;
    DevPatches dp{};
    DBuf bits;
    int64_t np = 0, words = 0, cap = 0, off = 0, i0 = 0;
  };
  std::vector<One> sc((size_t)n_scenes);
  int64_t total_cap = budget;  // a round's shares plus one row per scene
  for (int32_t k = 0; k < n_scenes; ++k) {
    One &o = sc[(size_t)k];
    const vr_patches *p = &patches[which[(size_t)k]];
    o.np = n_patches[which[(size_t)k]];
    o.x.reset(new vr_xfers);
    o.x->n = o.np;
    const size_t fb = 4 * (size_t)std::max<int64_t>(o.np, 1);
    TRY(o.S.make({fb, fb, fb, fb, fb, fb, fb}));
    const float *src[7] = {p->ox, p->oy, p->oz, p->nx, p->ny, p->nz, p->area};
    for (int q = 0; q < 7 && o.np > 0; ++q)
      VR_HIP(hipMemcpyAsync(o.S.ptr<float>(q), src[q], 4 * (size_t)o.np, hipMemcpyHostToDevice, st));
    o.dp = DevPatches{o.S.ptr<float>(0), o.S.ptr<float>(1), o.S.ptr<float>(2), o.S.ptr<float>(3),
                      o.S.ptr<float>(4), o.S.ptr<float>(5), o.S.ptr<float>(6)};
    o.words = (o.np + 31) / 32;
    TRY(o.bits.alloc((size_t)std::max<int64_t>(o.np, 1) * (size_t)std::max<int64_t>(o.words, 1) * 4));
    VR_HIP(hipMemsetAsync(o.bits.p, 0, o.bits.bytes, st));
    if (budget >= (1ll << 31) || o.np - 1 >= (1ll << 31))
      return fail(VR_E_INVALID, "batch_pairs too large");
    total_cap += std::max<int64_t>(o.np - 1, 0);
  }
  DBuf pairs, counts, d_jobs;
  TRY(pairs.alloc((size_t)std::max<int64_t>(total_cap, 1) * sizeof(int2)));
  TRY(counts.alloc(4 * (size_t)n_scenes));
  TRY(d_jobs.alloc(sizeof(TraceJob) * (size_t)n_scenes));
  std::vector<uint32_t> hc((size_t)n_scenes);
  std::vector<TraceJob> jobs;
  const int lds = tuning().trace_lds_entries;
  {
    Prof pr("make_transfers", st);
    for (;;) {
      bool any = false;
      int64_t active = 0, off = 0;
      for (int32_t k = 0; k < n_scenes; ++k) active += sc[(size_t)k].i0 + 1 < sc[(size_t)k].np;
      const int64_t share = std::max<int64_t>(1, budget / std::max<int64_t>(active, 1));
      VR_HIP(hipMemsetAsync(counts.p, 0, 4 * (size_t)n_scenes, st));
      for (int32_t k = 0; k < n_scenes; ++k) {
        One &o = sc[(size_t)k];
        if (o.i0 + 1 >= o.np) continue;
        any = true;
        const int64_t rowlen = o.np - 1 - o.i0;
        // share <= budget / active and rowlen <= np - 1, so the regions fit
        o.cap = std::max<int64_t>(rowlen, share);
        o.off = off;
        off += o.cap;
        const int64_t rows = std::max<int64_t>(1, o.cap / rowlen);
        const int64_t i1 = std::min<int64_t>(o.np - 1, o.i0 + rows);
        o.x->pairs += (2 * o.np - o.i0 - i1 - 1) * (i1 - o.i0) / 2;
        launch_pair_cull(o.dp, o.np, o.i0, i1, pairs.as<int2>() + o.off,
                         counts.as<uint32_t>() + k, (uint32_t)o.cap, st);
        o.i0 = i1;
      }
      if (!any) break;
      TRY(check_launch());
      VR_HIP(hipMemcpyAsync(hc.data(), counts.p, 4 * (size_t)n_scenes, hipMemcpyDeviceToHost, st));
      VR_HIP(hipStreamSynchronize(st));
      jobs.clear();
      int64_t blocks = 0;
      for (int32_t k = 0; k < n_scenes; ++k) {
        One &o = sc[(size_t)k];
        const int64_t c = std::min<int64_t>(hc[(size_t)k], o.cap);
        if (c <= 0) continue;
        o.x->segments += c;
        TraceJob j;
        j.sc = scenes[which[(size_t)k]]->dev;
        j.p = o.dp;
        j.pairs = pairs.as<int2>() + o.off;
        j.bits = o.bits.as<uint32_t>();
        j.words = o.words;
        j.count = c;
        j.blk_begin = blocks;
        blocks += (c + 63) / 64;
        jobs.push_back(j);
      }
      if (jobs.empty()) continue;
      if (blocks >= ((int64_t)1 << 32)) return fail(VR_E_CAPACITY, "too many pairs per launch");
      VR_HIP(hipMemcpyAsync(d_jobs.p, jobs.data(), sizeof(TraceJob) * jobs.size(),
                            hipMemcpyHostToDevice, st));
      launch_pair_trace_jobs(d_jobs.as<TraceJob>(), (int)jobs.size(), blocks, P.normal_offset,
                             P.skip_id, lds, st);
      TRY(check_launch());
      // the job table is rewritten next round: wait for this launch
      VR_HIP(hipStreamSynchronize(st));
    }
    for (int32_t k = 0; k < n_scenes; ++k) {
      One &o = sc[(size_t)k];
      TRY(transfers_from_bits(o.x.get(), o.bits, o.words, o.dp, P.normalize, st));
      o.bits.release();
    }
  }
  for (int32_t k = 0; k < n_scenes; ++k) out[which[(size_t)k]] = sc[(size_t)k].x.release();
  return VR_OK;
}

int vr_gather_sample_light_batch(vr_scene *const *scenes, const vr_samples *samples,
                                 const int64_t *n, const vr_light *const *lights,
                                 const int32_t *n_lights, int32_t n_scenes, float normal_offset,
                                 int32_t skip_id, float *const *rgb) {
  LOCK;
  if (n_scenes < 0) return fail(VR_E_INVALID, "n_scenes < 0");
  if (n_scenes == 0) return VR_OK;
  if (!scenes || !samples || !n || !n_lights || !rgb) return fail(VR_E_INVALID, "NULL argument");
  int64_t ns_tot = 0, nl_tot = 0;
  for (int32_t k = 0; k < n_scenes; ++k) {
    const std::string who = "scene " + std::to_string(k) + ": ";
    if (!scenes[k]) return fail(VR_E_INVALID, who + "NULL scene");
    if (!scenes[k]->built) return fail(VR_E_STATE, who + "scene not built");
    if (n[k] < 0 || n_lights[k] < 0) return fail(VR_E_INVALID, who + "negative count");
    if (n[k] > 0 && (!rgb[k] || !samples[k].px || !samples[k].py || !samples[k].pz ||
                     !samples[k].nx || !samples[k].ny || !samples[k].nz))
      return fail(VR_E_INVALID, who + "NULL sample array or output");
    if (n_lights[k] > 0 && (!lights || !lights[k])) return fail(VR_E_INVALID, who + "NULL lights");
    ns_tot += n[k];
    nl_tot += n_lights[k];
  }
  if (ns_tot == 0) return VR_OK;
  if (g_path == VR_PATH_PLAIN) {  // the comparison path: one scene at a time
    for (int32_t k = 0; k < n_scenes; ++k)
      TRY(vr_gather_sample_light(scenes[k], &samples[k], n[k], lights ? lights[k] : nullptr,
                                 n_lights[k], normal_offset, skip_id, rgb[k]));
    return VR_OK;
  }
  TRY(ensure_init());
  hipStream_t st = g_stream;
  Staging S;
  const size_t fb = 4 * (size_t)ns_tot;
  TRY(S.make({fb, fb, fb, fb, fb, fb, 3 * fb, sizeof(vr_light) * (size_t)std::max<int64_t>(nl_tot, 1),
              162 * 3 * 4, sizeof(GatherJob) * (size_t)n_scenes}));
  std::vector<GatherJob> jobs;
  int64_t at = 0, lat = 0, blocks = 0;
  for (int32_t k = 0; k < n_scenes; ++k) {
    const int64_t m = n[k];
    const float *src[6] = {samples[k].px, samples[k].py, samples[k].pz,
                           samples[k].nx, samples[k].ny, samples[k].nz};
    for (int q = 0; q < 6 && m > 0; ++q)
      VR_HIP(hipMemcpyAsync(S.ptr<float>(q) + at, src[q], 4 * (size_t)m, hipMemcpyHostToDevice, st));
    if (n_lights[k] > 0)
      VR_HIP(hipMemcpyAsync(S.ptr<vr_light>(7) + lat, lights[k], sizeof(vr_light) * (size_t)n_lights[k],
                            hipMemcpyHostToDevice, st));
    if (m > 0) {
      GatherJob j;
      j.sc = scenes[k]->dev;
      j.s = DevSamples{S.ptr<float>(0) + at, S.ptr<float>(1) + at, S.ptr<float>(2) + at,
                       S.ptr<float>(3) + at, S.ptr<float>(4) + at, S.ptr<float>(5) + at};
      j.lights = S.ptr<vr_light>(7) + lat;
      j.n_lights = n_lights[k];
      j.rgb = S.ptr<float>(6) + 3 * at;
      j.n = m;
      j.blk_begin = blocks;
      blocks += (m + 63) / 64;
      jobs.push_back(j);
    }
    at += m;
    lat += n_lights[k];
  }
  if (blocks >= ((int64_t)1 << 32)) return fail(VR_E_CAPACITY, "too many samples per launch");
  std::vector<float> dirs = sky_dirs();
  VR_HIP(hipMemcpyAsync(S.ptr<float>(8), dirs.data(), dirs.size() * 4, hipMemcpyHostToDevice, st));
  VR_HIP(hipMemcpyAsync(S.ptr<GatherJob>(9), jobs.data(), sizeof(GatherJob) * jobs.size(),
                        hipMemcpyHostToDevice, st));
  {
    Prof p("gather_light", st);
    launch_gather_jobs(S.ptr<GatherJob>(9), (int)jobs.size(), blocks, S.ptr<float>(8),
                       normal_offset, skip_id, tuning().trace_lds_entries, st);
  }
  TRY(check_launch());
  at = 0;
  for (int32_t k = 0; k < n_scenes; ++k) {
    if (n[k] > 0)
      VR_HIP(hipMemcpyAsync(rgb[k], S.ptr<float>(6) + 3 * at, 12 * (size_t)n[k],
                            hipMemcpyDeviceToHost, st));
    at += n[k];
  }
  VR_HIP(hipStreamSynchronize(st));
  return VR_OK;
}

int vr_make_transfers_pvs(vr_scene *s, const vr_patches *p, int64_t np,
                          const vr_xfer_params *prm, const int32_t *patch_cluster,
                          const uint8_t *pvs_rows, int32_t n_clusters, vr_xfers **out) {
  LOCK;
  if (n_clusters < 0) return fail(VR_E_INVALID, "n_clusters < 0");
  if (n_clusters > 0 && np > 0 && (!patch_cluster || !pvs_rows))
    return fail(VR_E_INVALID, "patch_cluster / pvs_rows is NULL");
  return make_transfers_impl(s, p, np, prm, patch_cluster, pvs_rows, np > 0 ? n_clusters : 0, out);
}

int vr_xfers_from_csr(vr_xfers **out, int64_t n, const int64_t *row_ptr, const int32_t *col,
                      const float *coef, int32_t precision) {
  LOCK;
  if (!out) return fail(VR_E_INVALID, "out is NULL");
  *out = nullptr;
  if (n < 0 || (n > 0 && !row_ptr)) return fail(VR_E_INVALID, "bad rows");
  TRY(ensure_init());
  std::unique_ptr<vr_xfers> x(new vr_xfers);
  x->n = n;
  x->cnt_host.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    int64_t l = row_ptr[i + 1] - row_ptr[i];
    if (l < 0 || l >= (1ll << 31)) return fail(VR_E_INVALID, "row_ptr not monotone");
    x->cnt_host[(size_t)i] = (int32_t)l;
  }
  x->nnz = n > 0 ? row_ptr[n] - row_ptr[0] : 0;
  if (x->nnz > 0 && (!col || !coef)) return fail(VR_E_INVALID, "col/coef NULL");
  for (int64_t k = 0; k < x->nnz; ++k)
    if (col[row_ptr[0] + k] < 0 || col[row_ptr[0] + k] >= n)
      return fail(VR_E_INVALID, "column index out of range");
  TRY(finish_layout(x.get(), g_stream));
  std::vector<int32_t> pc((size_t)x->stored);
  std::vector<float> pf((size_t)x->stored);
  std::vector<int64_t> perm;
  for (int64_t i = 0; i < n; ++i) {
    int64_t s = x->rp_host[i], l = x->cnt_host[i], e = x->rp_host[i + 1];
    // rows are stored with ascending columns (the tiled gather reads each
    // column tile of a row as one contiguous segment)
    perm.resize((size_t)l);
    for (int64_t k = 0; k < l; ++k) perm[(size_t)k] = row_ptr[i] + k;
    std::stable_sort(perm.begin(), perm.end(),
                     [&](int64_t a, int64_t b) { return col[a] < col[b]; });
    for (int64_t k = 0; k < l; ++k) {
      pc[s + k] = col[perm[(size_t)k]];
      pf[s + k] = coef[perm[(size_t)k]];
    }
    for (int64_t k = s + l; k < e; ++k) {
      pc[k] = (int32_t)i;
      pf[k] = 0.0f;
    }
  }
  TRY(upload_vec(x->col, pc, g_stream));
  TRY(upload_vec(x->coef32, pf, g_stream));
  VR_HIP(hipStreamSynchronize(g_stream));
  vr_xfers *raw = x.release();
  if (precision != VR_COEF_F32) {
    int rc = vr_xfers_set_precision(raw, precision);
    if (rc) {
      delete raw;
      return rc;
    }
  }
  *out = raw;
  return VR_OK;
}

int vr_xfers_synthetic(vr_xfers **out, uint64_t seed, int64_t n, double mean_len, double sigma,
                       int64_t max_len, int32_t precision, int32_t columns) {
  LOCK;
  if (!out) return fail(VR_E_INVALID, "out is NULL");
  *out = nullptr;
  if (n < 2 || mean_len < 1 || sigma < 0 || max_len < 1)
    return fail(VR_E_INVALID, "bad synthetic parameters");
  if (columns != VR_SYNTH_UNIFORM && columns != VR_SYNTH_STRATIFIED)
    return fail(VR_E_INVALID, "unknown column layout");
  if (n > INT32_MAX) return fail(VR_E_CAPACITY, "columns are int32");
  TRY(ensure_init());
  std::unique_ptr<vr_xfers> x(new vr_xfers);
  x->n = n;
  const int64_t cap = std::min<int64_t>(max_len, n - 1);
  uint64_t st = seed ^ 0x243F6A8885A308D3ull;
  auto next = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const double mu = std::log(mean_len) - 0.5 * sigma * sigma;
  std::vector<double> raw((size_t)n);
  double sum = 0;
  for (int64_t i = 0; i < n; ++i) {
    double u1 = ((double)(next() >> 11) + 1.0) * 0x1p-53, u2 = (double)(next() >> 11) * 0x1p-53;
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
    double L = std::min<double>((double)cap, std::max(1.0, std::exp(mu + sigma * z)));
    raw[(size_t)i] = L;
    sum += L;
  }
  const double f = mean_len * (double)n / sum;
  x->cnt_host.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i)
    x->cnt_host[(size_t)i] =
        (int32_t)std::min<double>((double)cap, std::max(1.0, std::floor(raw[(size_t)i] * f + 0.5)));
  x->nnz = 0;
  for (int32_t c : x->cnt_host) x->nnz += c;
  TRY(finish_layout(x.get(), g_stream));
  TRY(x->col.alloc(4 * (size_t)x->stored));
  TRY(x->coef32.alloc(4 * (size_t)x->stored));
  DBuf bitmaps;
  if (columns == VR_SYNTH_UNIFORM) {
    const size_t bb = 4 * (size_t)synthetic_waves(n) * (size_t)synthetic_bitmap_words(n);
    TRY(bitmaps.alloc(bb));
    VR_HIP(hipMemsetAsync(bitmaps.p, 0, bb, g_stream));
  }
  launch_synthetic_fill(n, x->rp.as<int64_t>(), x->cnt.as<int32_t>(), seed, columns,
                        bitmaps.as<uint32_t>(), x->col.as<int32_t>(), x->coef32.as<float>(),
                        g_stream);
  TRY(check_launch());
  VR_HIP(hipStreamSynchronize(g_stream));
  vr_xfers *rx = x.release();
  if (precision != VR_COEF_F32) {
    int rc = vr_xfers_set_precision(rx, precision);
    if (rc) {
      delete rx;
      return rc;
    }
  }
  *out = rx;
  return VR_OK;
}

int vr_xfers_set_precision(vr_xfers *x, int32_t precision) {
  LOCK;
  if (!x) return fail(VR_E_INVALID, "xfers is NULL");
  if (x->served) return fail(VR_E_STATE, "transfers are held by a server");
  if (precision != VR_COEF_F32 && precision != VR_COEF_F16 && precision != VR_COEF_BF16)
    return fail(VR_E_INVALID, "unknown precision");
  TRY(ensure_init());
  if (precision != VR_COEF_F32) {
    TRY(x->coef16.alloc(2 * (size_t)x->stored));
    TRY(x->scale.alloc(4 * (size_t)std::max<int64_t>(x->n, 1)));
    launch_encode16(x->n, x->rp.as<int64_t>(), x->coef32.as<float>(), x->coef16.as<uint16_t>(),
                    x->scale.as<float>(), precision, g_stream);
    TRY(check_launch());
    VR_HIP(hipStreamSynchronize(g_stream));
  }
  x->precision = precision;
  return VR_OK;
}

int vr_xfers_get_info(const vr_xfers *x, vr_xfers_info *out) {
  if (!x || !out) return fail(VR_E_INVALID, "NULL argument");
  std::memset(out, 0, sizeof *out);
  out->n_rows = x->n;
  out->nnz = x->nnz;
  out->stored_nnz = x->stored;
  out->precision = x->precision;
  const int64_t cb = x->precision == VR_COEF_F32 ? 4 : 2;
  out->stored_bytes = x->stored * (4 + cb) + 8 * (x->n + 1) + (x->precision ? 4 * x->n : 0);
  out->pairs_considered = x->pairs;
  out->segments_traced = x->segments;
  return VR_OK;
}

int vr_xfers_download(const vr_xfers *x, int64_t *row_ptr, int32_t *col, float *coef,
                      uint16_t *raw16, float *row_scale) {
  LOCK;
  if (!x) return fail(VR_E_INVALID, "xfers is NULL");
  TRY(ensure_init());
  const int64_t n = x->n, T = x->stored;
  if (row_ptr) {
    row_ptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) row_ptr[i + 1] = row_ptr[i] + x->cnt_host[(size_t)i];
  }
  std::vector<int32_t> pc;
  std::vector<float> pf;
  std::vector<uint16_t> ph;
  if (col) {
    pc.resize((size_t)T);
    VR_HIP(hipMemcpyAsync(pc.data(), x->col.p, 4 * (size_t)T, hipMemcpyDeviceToHost, g_stream));
  }
  if (x->precision == VR_COEF_F32) {
    if (coef) {
      pf.resize((size_t)T);
      VR_HIP(hipMemcpyAsync(pf.data(), x->coef32.p, 4 * (size_t)T, hipMemcpyDeviceToHost, g_stream));
    }
    if (raw16) return fail(VR_E_STATE, "raw 16-bit coefficients requested from FP32 transfers");
  } else {
    if (coef || raw16) {
      ph.resize((size_t)T);
      VR_HIP(hipMemcpyAsync(ph.data(), x->coef16.p, 2 * (size_t)T, hipMemcpyDeviceToHost, g_stream));
    }
    if (coef) {
      DBuf tmp;
      TRY(tmp.alloc(4 * (size_t)T));
      launch_decode16(T, x->coef16.as<uint16_t>(), tmp.as<float>(), x->precision, g_stream);
      TRY(check_launch());
      pf.resize((size_t)T);
      VR_HIP(hipMemcpyAsync(pf.data(), tmp.p, 4 * (size_t)T, hipMemcpyDeviceToHost, g_stream));
      VR_HIP(hipStreamSynchronize(g_stream));
    }
    if (row_scale && n > 0)
      VR_HIP(hipMemcpyAsync(row_scale, x->scale.p, 4 * (size_t)n, hipMemcpyDeviceToHost, g_stream));
  }
  VR_HIP(hipStreamSynchronize(g_stream));
  int64_t k = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t s = x->rp_host[(size_t)i];
    for (int64_t q = 0; q < x->cnt_host[(size_t)i]; ++q, ++k) {
      if (col) col[k] = pc[(size_t)(s + q)];
      if (coef) coef[k] = pf[(size_t)(s + q)];
      if (raw16) raw16[k] = ph[(size_t)(s + q)];
    }
  }
  return VR_OK;
}

int vr_xfers_destroy(vr_xfers *x) {
  LOCK;
  if (x && x->served) return fail(VR_E_STATE, "transfers are held by a server");
  delete x;
  return VR_OK;
}

// ---- direct light -----------------------------------------------------------
int vr_gather_sample_light(vr_scene *s, const vr_samples *smp, int64_t n, const vr_light *lights,
                           int32_t n_lights, float normal_offset, int32_t skip_id, float *rgb) {
  LOCK;
  if (!s || !smp) return fail(VR_E_INVALID, "NULL argument");
  if (!s->built) return fail(VR_E_STATE, "scene not built");
  if (n < 0 || n_lights < 0) return fail(VR_E_INVALID, "negative count");
  if (n == 0) return VR_OK;
  if (!rgb || (n_lights > 0 && !lights)) return fail(VR_E_INVALID, "NULL lights/output");
  TRY(ensure_init());
  const size_t fb = 4 * (size_t)n;
  Staging S;
  TRY(S.make({fb, fb, fb, fb, fb, fb, 3 * fb, sizeof(vr_light) * (size_t)std::max(n_lights, 1),
              162 * 3 * 4}));
  const float *src[6] = {smp->px, smp->py, smp->pz, smp->nx, smp->ny, smp->nz};
  for (int k = 0; k < 6; ++k) {
    if (!src[k]) return fail(VR_E_INVALID, "NULL sample array");
    VR_HIP(hipMemcpyAsync(S.ptr<float>(k), src[k], fb, hipMemcpyHostToDevice, g_stream));
  }
  if (n_lights > 0)
    VR_HIP(hipMemcpyAsync(S.ptr<vr_light>(7), lights, sizeof(vr_light) * (size_t)n_lights,
                          hipMemcpyHostToDevice, g_stream));
  std::vector<float> dirs = sky_dirs();
  VR_HIP(hipMemcpyAsync(S.ptr<float>(8), dirs.data(), dirs.size() * 4, hipMemcpyHostToDevice,
                        g_stream));
  DevSamples ds{S.ptr<float>(0), S.ptr<float>(1), S.ptr<float>(2),
                S.ptr<float>(3), S.ptr<float>(4), S.ptr<float>(5)};
  {
    Prof p("gather_light", g_stream);
    launch_gather_light(s->dev, ds, n, S.ptr<vr_light>(7), n_lights, S.ptr<float>(8),
                        normal_offset, skip_id, S.ptr<float>(6), g_path,
                        tuning().trace_lds_entries, g_stream);
  }
  TRY(check_launch());
  VR_HIP(hipMemcpyAsync(rgb, S.ptr<float>(6), 3 * fb, hipMemcpyDeviceToHost, g_stream));
  VR_HIP(hipStreamSynchronize(g_stream));
  return VR_OK;
}

int vr_can_leaf_trace_to_sky(vr_scene *s, const int16_t *mins, const int16_t *maxs, int64_t n,
                             uint8_t *can_see_sky) {
  LOCK;
  if (!s) return fail(VR_E_INVALID, "NULL scene");
  if (!s->built) return fail(VR_E_STATE, "scene not built");
  if (n < 0) return fail(VR_E_INVALID, "negative count");
  if (n == 0) return VR_OK;
  if (!mins || !maxs || !can_see_sky) return fail(VR_E_INVALID, "NULL bounds/output");
  if (n > (int64_t)UINT32_MAX) return fail(VR_E_CAPACITY, "more than 2^32 - 1 leaves per call");
  TRY(ensure_init());
  const size_t bb = 6 * (size_t)n;
  Staging S;
  TRY(S.make({bb, bb, (size_t)n, 162 * 3 * 4}));
  VR_HIP(hipMemcpyAsync(S.ptr<int16_t>(0), mins, bb, hipMemcpyHostToDevice, g_stream));
  VR_HIP(hipMemcpyAsync(S.ptr<int16_t>(1), maxs, bb, hipMemcpyHostToDevice, g_stream));
  std::vector<float> dirs = sky_dirs();
  VR_HIP(hipMemcpyAsync(S.ptr<float>(3), dirs.data(), dirs.size() * 4, hipMemcpyHostToDevice,
                        g_stream));
  {
    Prof p("leaf_sky", g_stream);
    launch_leaf_sky(s->dev, S.ptr<int16_t>(0), S.ptr<int16_t>(1), n, S.ptr<float>(3),
                    S.ptr<uint8_t>(2), g_stream);
  }
  TRY(check_launch());
  VR_HIP(hipMemcpyAsync(can_see_sky, S.ptr<uint8_t>(2), (size_t)n, hipMemcpyDeviceToHost,
                        g_stream));
  VR_HIP(hipStreamSynchronize(g_stream));
  return VR_OK;
}

// ---- bounce ------------------------------------------------------------------
namespace vr {
namespace {

// Gather shape of the LDS-tiled or stream kernel for x at the current tuning.
using TiledShape = GatherShape;
int tiled_shape(const vr_xfers *x, const Tuning &tu, TiledShape &sh) {
  sh.kernel = tu.bounce_kernel;
  sh.cols = tu.bounce_tile_cols;
  if (sh.kernel == 2) {  // the stream gather has no lane groups or passes
    sh.group = 0;
    sh.passes = 0;
  } else {
    // defaults by precision: 16-lane groups for FP32, 8 for FP16/BF16 chunks
    sh.group = tu.bounce_group ? tu.bounce_group : (x->precision == VR_COEF_F32 ? 16 : 8);
    sh.passes = tu.bounce_passes ? tu.bounce_passes : 1;
  }
  sh.prefetch = tu.bounce_prefetch ? tu.bounce_prefetch : (sh.kernel == 2 ? 1 : 3);
  sh.placement = tu.bounce_placement;
  if (!bounce_gather_variant(sh))
    return fail(VR_E_UNSUPPORTED, "no compiled LDS gather for this tuning shape");
  return VR_OK;
}

// Whether x's tile pointers and tile-blocked copy already fit shape sh (so
// prepare_tiles allocates and launches nothing).
bool tiles_ready(const vr_xfers *x, const TiledShape &sh) {
  if (x->n == 0) return true;
  const int ostride = sh.kernel == 2 ? 16 : 12;
  return x->tp_cols == sh.cols && x->tp.p && x->blk_cols == sh.cols &&
         x->blk_prec == x->precision && x->blk_chunks == gather_chunks(x->n) &&
         x->blk_stride == ostride && x->placement == sh.placement;
}

// Tile pointers and the tile-blocked copy of x for the tile width, built once
// and kept with the transfers; fills tl.
int prepare_tiles(vr_xfers *x, const TiledShape &sh, DevTiles &tl, hipStream_t st) {
  const int64_t n = x->n;
  const int cols = sh.cols;
  const int nt = (int)((n + cols - 1) / cols);
  if (nt > kMaxTilesPerJob)
    return fail(VR_E_CAPACITY, "more than " + std::to_string(kMaxTilesPerJob) +
                                   " column tiles; use a wider tile");
  tl = DevTiles{};
  tl.nt = nt;
  tl.cols = cols;
  tl.group = sh.group;
  tl.n_chunks = gather_chunks(n);
  if (n == 0) return VR_OK;
  if (x->served && !tiles_ready(x, sh))
    return fail(VR_E_STATE, "transfers are held by a server with another gather layout");
  if (x->placement != sh.placement) {  // deal the rows again; the blocked copy follows them
    std::vector<int32_t> order;
    row_order(x->cnt_host, sh.placement, order);
    // into the same buffer (n rows either way), so the order pointer in every
    // DevXfers already taken of x stays valid
    if (x->order.bytes < sizeof(int32_t) * order.size())
      return fail(VR_E_STATE, "row order buffer smaller than the rows");
    VR_HIP(hipMemcpyAsync(x->order.p, order.data(), sizeof(int32_t) * order.size(),
                          hipMemcpyHostToDevice, st));
    VR_HIP(hipStreamSynchronize(st));
    x->placement = sh.placement;
    x->blk_cols = 0;
  }
  const DevXfers dx = x->dev();
  if (x->tp_cols != cols || !x->tp.p) {
    TRY(x->tp.alloc(4 * (size_t)n * (size_t)(nt + 1)));
    launch_tile_ptrs(n, x->rp.as<int64_t>(), x->cnt.as<int32_t>(), x->col.as<int32_t>(), cols, nt,
                     x->tp.as<int32_t>(), st);
    TRY(check_launch());
    x->tp_cols = cols;
    x->blk_cols = 0;
  }
  tl.tp = x->tp.as<int32_t>();
  tl.ostride = sh.kernel == 2 ? 16 : 12;
  if (x->blk_cols != cols || x->blk_prec != x->precision || x->blk_chunks != tl.n_chunks ||
      x->blk_stride != tl.ostride) {
    // entries per chunk: 4 in the stream layout, else 4 (FP32) / 8 (16-bit)
    const int G = (tl.ostride == 16 || x->precision == VR_COEF_F32) ? 4 : 8;
    const int cb = x->precision == VR_COEF_F32 ? 4 : 2;
    const int64_t nblk = tl.n_chunks * nt;
    DBuf btot;
    TRY(btot.alloc(8 * (size_t)nblk));
    TRY(x->boff.alloc(4 * (size_t)nblk * 1028));
    tl.boff = x->boff.as<int32_t>();
    launch_block_offsets(dx, tl, G, btot.as<int64_t>(), x->boff.as<int32_t>(), st);
    TRY(check_launch());
    std::vector<int64_t> tot((size_t)nblk), start((size_t)nblk);
    VR_HIP(hipMemcpyAsync(tot.data(), btot.p, 8 * (size_t)nblk, hipMemcpyDeviceToHost, st));
    VR_HIP(hipStreamSynchronize(st));
    int64_t run = 0;
    for (int64_t b = 0; b < nblk; ++b) {
      start[(size_t)b] = run;
      run += tot[(size_t)b];
    }
    TRY(upload_vec(x->bstart, start, st));
    // the stream layout interleaves 16-bit offsets and coefficients in bidx
    const bool inter = tl.ostride == 16 && cb == 2;
    TRY(x->bidx.alloc((inter ? 4 : 2) * (size_t)std::max<int64_t>(run, 1) + 64));
    TRY(x->bcoef.alloc(inter ? 64 : (size_t)cb * (size_t)std::max<int64_t>(run, 1) + 64));
    tl.bstart = x->bstart.as<int64_t>();
    launch_block_fill(dx, tl, G, cb, x->bidx.as<uint16_t>(), x->bcoef.p, st);
    TRY(check_launch());
    VR_HIP(hipStreamSynchronize(st));
    x->blk_cols = cols;
    x->blk_prec = x->precision;
    x->blk_chunks = tl.n_chunks;
    x->blk_stride = tl.ostride;
  }
  tl.bidx = x->bidx.as<uint16_t>();
  tl.bcoef = x->bcoef.p;
  tl.bstart = x->bstart.as<int64_t>();
  tl.boff = x->boff.as<int32_t>();
  return VR_OK;
}

// Device buffers of one job of a tiled bounce (emit x2, total, refl, area,
// e_row, energy scratch, energy); the state lives in a shared array.
struct JobBuffers {
  Staging S;
  int make(int64_t n, int K) {
    const size_t vb = 12 * (size_t)std::max<int64_t>(n, 1);
    const int64_t eb = (n + 2047) / 2048;
    return S.make({vb, vb, vb, vb, 4 * (size_t)std::max<int64_t>(n, 1),
                   8 * (size_t)std::max<int64_t>(n, 1), 8 * (size_t)(eb + 2),
                   8 * (size_t)std::max(K, 1)});
  }
};

void fill_job(BounceJob &jb, const vr_xfers *x, const DevTiles &tl, JobBuffers &b,
              BounceState *state, int K, double stop_frac) {
  jb.x = x->dev();
  jb.tl = tl;
  jb.emit[0] = b.S.ptr<float>(0);
  jb.emit[1] = b.S.ptr<float>(1);
  jb.total = b.S.ptr<float>(2);
  jb.refl = b.S.ptr<float>(3);
  jb.area = b.S.ptr<float>(4);
  jb.e_row = b.S.ptr<double>(5);
  jb.scratch = b.S.ptr<double>(6);
  jb.energy = b.S.ptr<double>(7);
  jb.state = state;
  jb.e_blocks = (x->n + 2047) / 2048;
  jb.max_bounces = K;
  jb.stop_frac = stop_frac;
}

// Runs the bounces of a set of prepared jobs (device table `d_jobs`, states
// `d_states`) in chunks of 8, optionally as a captured graph of one chunk
// replayed until every job has stopped.
int run_jobs(const TiledShape &sh, int kind, const BounceJob *d_jobs, int n_jobs,
             int64_t total_chunks, int64_t total_eblocks, BounceState *d_states, int K,
             bool use_graph, hipStream_t st) {
  std::vector<BounceState> hs((size_t)n_jobs);
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int rc = VR_OK;
  if (use_graph && K > 0) {
    VR_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < 8; ++i)
      launch_bounce_jobs(kind, sh, d_jobs, n_jobs,
                         total_chunks, total_eblocks, st);
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess) return fail(VR_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(VR_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
  }
  for (int done = 0; done < K;) {
    if (exec) {
      hipError_t e = hipGraphLaunch(exec, st);
      if (e != hipSuccess) {
        rc = fail(VR_E_HIP, std::string("graph launch: ") + hipGetErrorString(e));
        break;
      }
    } else {
      for (int i = 0; i < std::min(8, K - done); ++i) {
        {
          Prof p("bounce_gather", st);
          launch_gather_jobs(kind, sh, d_jobs, n_jobs,
                             total_chunks, st);
        }
        {
          Prof p("energy", st);
          launch_energy_jobs(d_jobs, n_jobs, total_eblocks, st);
        }
      }
      if ((rc = check_launch())) break;
    }
    done += 8;
    hipError_t e = hipMemcpyAsync(hs.data(), d_states, sizeof(BounceState) * (size_t)n_jobs,
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = fail(VR_E_HIP, std::string("bounce: ") + hipGetErrorString(e));
      break;
    }
    bool all = true;
    for (const BounceState &b : hs) all = all && (b.converged || b.k_done >= K);
    if (all) break;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return rc;
}

}  // namespace
}  // namespace vr

int vr_bounce_light(vr_xfers *x, const float *emit0, const float *refl, const float *area,
                    float *total, const vr_bounce_params *prm, double *energy,
                    int32_t *n_bounces) {
  LOCK;
  if (!x || !prm) return fail(VR_E_INVALID, "NULL argument");
  const int64_t n = x->n;
  const int K = prm->max_bounces;
  if (K < 0) return fail(VR_E_INVALID, "max_bounces < 0");
  if (n > 0 && (!emit0 || !refl || !area || !total)) return fail(VR_E_INVALID, "NULL array");
  if (K > 0 && !energy) return fail(VR_E_INVALID, "energy is NULL");
  if (g_path == VR_PATH_PLAIN && x->precision != VR_COEF_F32)
    return fail(VR_E_UNSUPPORTED, "plain path gathers FP32 transfers only");
  TRY(ensure_init());
  hipStream_t st = g_stream;
  const Tuning tu = tuning();
  if (g_path == VR_PATH_OPT && tu.bounce_kernel != 1) {
    // LDS-tiled or stream gather: a batch of one job
    vr_bounce_job jb{x, emit0, refl, area, total, energy, 0};
    const int rc = vr_bounce_light_batch(&jb, 1, prm, 0);
    if (n_bounces) *n_bounces = jb.n_bounces;
    return rc;
  }
  const size_t vb = 12 * (size_t)std::max<int64_t>(n, 1);
  int64_t scratch_n = 2;
  for (int64_t m = n; m > 1; m = (m + 2047) / 2048) scratch_n += (m + 2047) / 2048;
  Staging S;
  TRY(S.make({vb, vb, vb, vb, 4 * (size_t)std::max<int64_t>(n, 1),
              8 * (size_t)std::max<int64_t>(n, 1), 8 * (size_t)scratch_n,
              8 * (size_t)std::max(K, 1), 4 * (size_t)std::max(K, 1), sizeof(BounceState)}));
  float *em[2] = {S.ptr<float>(0), S.ptr<float>(1)};
  float *d_total = S.ptr<float>(2), *d_refl = S.ptr<float>(3), *d_area = S.ptr<float>(4);
  double *e_row = S.ptr<double>(5), *scr = S.ptr<double>(6), *d_energy = S.ptr<double>(7);
  uint32_t *counters = S.ptr<uint32_t>(8);
  BounceState *state = S.ptr<BounceState>(9);
  if (n > 0) {
    VR_HIP(hipMemcpyAsync(em[0], emit0, 12 * (size_t)n, hipMemcpyHostToDevice, st));
    VR_HIP(hipMemcpyAsync(d_total, total, 12 * (size_t)n, hipMemcpyHostToDevice, st));
    VR_HIP(hipMemcpyAsync(d_refl, refl, 12 * (size_t)n, hipMemcpyHostToDevice, st));
    VR_HIP(hipMemcpyAsync(d_area, area, 4 * (size_t)n, hipMemcpyHostToDevice, st));
  }
  VR_HIP(hipMemsetAsync(counters, 0, 4 * (size_t)std::max(K, 1), st));
  VR_HIP(hipMemsetAsync(state, 0, sizeof(BounceState), st));
  const DevXfers dx = x->dev();
  BounceState hs{0.0, 0, 0};
  const int chunk = g_path == VR_PATH_PLAIN ? 1 : 8;
  int k = 0;
  while (k < K) {
    const int kend = std::min(K, k + chunk);
    for (; k < kend; ++k) {
      {
        Prof p("bounce_gather", st);
        launch_bounce(dx, em[k & 1], em[(k + 1) & 1], d_total, d_refl, d_area, e_row,
                      counters + k, state, g_path, st);
      }
      {
        Prof p("energy", st);
        launch_energy(e_row, n, scr, d_energy, k, prm->stop_frac, state, st);
      }
      TRY(check_launch());
    }
    VR_HIP(hipMemcpyAsync(&hs, state, sizeof hs, hipMemcpyDeviceToHost, st));
    VR_HIP(hipStreamSynchronize(st));
    if (hs.converged) break;
  }
  const int done = hs.k_done;
  if (done > 0) VR_HIP(hipMemcpyAsync(energy, d_energy, 8 * (size_t)done, hipMemcpyDeviceToHost, st));
  if (n > 0) VR_HIP(hipMemcpyAsync(total, d_total, 12 * (size_t)n, hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  if (n_bounces) *n_bounces = done;
  return VR_OK;
}

int vr_bounce_light_batch(vr_bounce_job *jobs, int32_t n_jobs, const vr_bounce_params *prm,
                          int32_t use_graph) {
  LOCK;
  if (!jobs || !prm || n_jobs < 0) return fail(VR_E_INVALID, "NULL argument or n_jobs < 0");
  const int K = prm->max_bounces;
  if (K < 0) return fail(VR_E_INVALID, "max_bounces < 0");
  if (n_jobs == 0) return VR_OK;
  const Tuning tu = tuning();
  if (g_path != VR_PATH_OPT || tu.bounce_kernel == 1)
    return fail(VR_E_UNSUPPORTED,
                "batched bounces run on the LDS-tiled or stream gather (optimised path)");
  const int kind = jobs[0].xfers ? jobs[0].xfers->precision : -1;
  for (int j = 0; j < n_jobs; ++j) {
    const vr_bounce_job &b = jobs[j];
    if (!b.xfers) return fail(VR_E_INVALID, "job " + std::to_string(j) + ": transfers are NULL");
    if (b.xfers->precision != kind)
      return fail(VR_E_INVALID, "jobs of one batch must share the coefficient precision");
    if (b.xfers->n > 0 && (!b.emit0 || !b.refl || !b.area || !b.total))
      return fail(VR_E_INVALID, "job " + std::to_string(j) + ": NULL array");
    if (K > 0 && !b.energy) return fail(VR_E_INVALID, "job " + std::to_string(j) + ": energy is NULL");
    if (b.xfers->n > (int64_t)2048 * 2048)
      return fail(VR_E_CAPACITY, "a batched job holds at most 2048^2 rows");
  }
  TRY(ensure_init());
  hipStream_t st = g_stream;
  TiledShape sh;
  TRY(tiled_shape(jobs[0].xfers, tu, sh));
  for (int j = 0; j < n_jobs; ++j)  // refuse before any work
    if (jobs[j].xfers->served && !tiles_ready(jobs[j].xfers, sh))
      return fail(VR_E_STATE, "job " + std::to_string(j) +
                                  ": transfers are held by a server with another gather layout");
  std::vector<BounceJob> hj((size_t)n_jobs);
  std::vector<JobBuffers> bufs((size_t)n_jobs);
  DBuf d_states, d_jobs;
  TRY(d_states.alloc(sizeof(BounceState) * (size_t)n_jobs));
  VR_HIP(hipMemsetAsync(d_states.p, 0, sizeof(BounceState) * (size_t)n_jobs, st));
  int64_t chunks = 0, eblocks = 0;
  for (int j = 0; j < n_jobs; ++j) {
    vr_xfers *x = jobs[j].xfers;
    const int64_t n = x->n;
    DevTiles tl;
    TRY(prepare_tiles(x, sh, tl, st));
    TRY(bufs[(size_t)j].make(n, K));
    BounceJob &jb = hj[(size_t)j];
    fill_job(jb, x, tl, bufs[(size_t)j], d_states.as<BounceState>() + j, K, prm->stop_frac);
    jb.chunk_begin = chunks;
    jb.eblk_begin = eblocks;
    chunks += tl.n_chunks;
    eblocks += jb.e_blocks;
    if (n > 0) {
      VR_HIP(hipMemcpyAsync(jb.emit[0], jobs[j].emit0, 12 * (size_t)n, hipMemcpyHostToDevice, st));
      VR_HIP(hipMemcpyAsync(jb.total, jobs[j].total, 12 * (size_t)n, hipMemcpyHostToDevice, st));
      VR_HIP(hipMemcpyAsync((void *)jb.refl, jobs[j].refl, 12 * (size_t)n, hipMemcpyHostToDevice, st));
      VR_HIP(hipMemcpyAsync((void *)jb.area, jobs[j].area, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    }
  }
  TRY(d_jobs.upload(hj.data(), sizeof(BounceJob) * (size_t)n_jobs, st));
  TRY(run_jobs(sh, kind, d_jobs.as<BounceJob>(), n_jobs, chunks, eblocks,
               d_states.as<BounceState>(), K, use_graph != 0, st));
  std::vector<BounceState> hs((size_t)n_jobs);
  VR_HIP(hipMemcpyAsync(hs.data(), d_states.p, sizeof(BounceState) * (size_t)n_jobs,
                        hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  for (int j = 0; j < n_jobs; ++j) {
    const BounceJob &jb = hj[(size_t)j];
    const int done = hs[(size_t)j].k_done;
    const int64_t n = jobs[j].xfers->n;
    if (done > 0)
      VR_HIP(hipMemcpyAsync(jobs[j].energy, jb.energy, 8 * (size_t)done, hipMemcpyDeviceToHost, st));
    if (n > 0) VR_HIP(hipMemcpyAsync(jobs[j].total, jb.total, 12 * (size_t)n, hipMemcpyDeviceToHost, st));
    jobs[j].n_bounces = done;
  }
  VR_HIP(hipStreamSynchronize(st));
  return VR_OK;
}

int vr_bounce_light_noalloc(vr_xfers *x, const float *emit0, const float *refl,
                            const float *area, float *total, const vr_bounce_params *prm,
                            double *energy, int32_t *n_bounces) {
  LOCK;
  if (!x || !prm) return fail(VR_E_INVALID, "NULL argument");
  const int64_t n = x->n;
  const int K = prm->max_bounces;
  if (K < 0) return fail(VR_E_INVALID, "max_bounces < 0");
  if (n > 0 && (!emit0 || !refl || !area || !total)) return fail(VR_E_INVALID, "NULL array");
  if (K > 0 && !energy) return fail(VR_E_INVALID, "energy is NULL");
  if (n > kNoAllocMaxRows)
    return fail(VR_E_CAPACITY, "the allocation-free bounce holds at most 524288 rows");
  if (K > kNoAllocMaxBounces)
    return fail(VR_E_CAPACITY, "the allocation-free bounce runs at most 1024 bounces");
  const Tuning tu = tuning();
  if (g_path != VR_PATH_OPT || tu.bounce_kernel == 1)
    return fail(VR_E_UNSUPPORTED,
                "the allocation-free bounce runs the LDS-tiled or stream gather (optimised path)");
  TRY(ensure_init());
  char *base = g_scratch[g_device];
  if (!base) return fail(VR_E_STATE, "no working set on this device");
  TiledShape sh;
  TRY(tiled_shape(x, tu, sh));
  if (!tiles_ready(x, sh))
    return fail(VR_E_STATE,
                "the gather layout of these transfers is not built for the current tuning "
                "(run vr_bounce_light on them once)");
  hipStream_t st = g_stream;
  DevTiles tl;
  TRY(prepare_tiles(x, sh, tl, st));  // the layout is ready: fills tl only
  BounceJob jb;
  jb.x = x->dev();
  jb.tl = tl;
  jb.emit[0] = reinterpret_cast<float *>(base + kNaEmit0);
  jb.emit[1] = reinterpret_cast<float *>(base + kNaEmit1);
  jb.total = reinterpret_cast<float *>(base + kNaTotal);
  jb.refl = reinterpret_cast<float *>(base + kNaRefl);
  jb.area = reinterpret_cast<float *>(base + kNaArea);
  jb.e_row = reinterpret_cast<double *>(base + kNaERow);
  jb.scratch = reinterpret_cast<double *>(base + kNaEScr);
  jb.energy = reinterpret_cast<double *>(base + kNaEnergy);
  jb.state = reinterpret_cast<BounceState *>(base + kNaState);
  jb.e_blocks = (n + 2047) / 2048;
  jb.max_bounces = K;
  jb.stop_frac = prm->stop_frac;
  jb.chunk_begin = 0;
  jb.eblk_begin = 0;
  BounceJob *d_job = reinterpret_cast<BounceJob *>(base + kNaJob);
  if (n > 0) {
    VR_HIP(hipMemcpyAsync(jb.emit[0], emit0, 12 * (size_t)n, hipMemcpyHostToDevice, st));
    VR_HIP(hipMemcpyAsync(jb.total, total, 12 * (size_t)n, hipMemcpyHostToDevice, st));
    VR_HIP(hipMemcpyAsync((void *)jb.refl, refl, 12 * (size_t)n, hipMemcpyHostToDevice, st));
    VR_HIP(hipMemcpyAsync((void *)jb.area, area, 4 * (size_t)n, hipMemcpyHostToDevice, st));
  }
  // the array keeps the last call's contents: the bounce state starts at zero
  VR_HIP(hipMemsetAsync(jb.state, 0, sizeof(BounceState), st));
  VR_HIP(hipMemcpyAsync(d_job, &jb, sizeof jb, hipMemcpyHostToDevice, st));
  const int kind = x->precision;
  BounceState hs{0.0, 0, 0};
  for (int done = 0; done < K; done += 8) {
    for (int i = 0; i < std::min(8, K - done); ++i) {
      launch_gather_jobs(kind, sh, d_job, 1, tl.n_chunks, st);
      launch_energy_jobs(d_job, 1, jb.e_blocks, st);
    }
    TRY(check_launch());
    VR_HIP(hipMemcpyAsync(&hs, jb.state, sizeof hs, hipMemcpyDeviceToHost, st));
    VR_HIP(hipStreamSynchronize(st));
    if (hs.converged || hs.k_done >= K) break;
  }
  const int nb = hs.k_done;
  if (nb > 0)
    VR_HIP(hipMemcpyAsync(energy, jb.energy, 8 * (size_t)nb, hipMemcpyDeviceToHost, st));
  if (n > 0) VR_HIP(hipMemcpyAsync(total, jb.total, 12 * (size_t)n, hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  if (n_bounces) *n_bounces = nb;
  return VR_OK;
}

// ---- serving: resident scenes, re-bake requests, continuous batching ------------
namespace vr {
namespace {

// First-fit allocator over one device arena (offsets in bytes, 256-aligned).
struct Arena {
  DBuf mem;
  size_t cap = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  int init(size_t bytes) {
    cap = (bytes + 255) & ~(size_t)255;
    TRY(mem.alloc(cap));
    free_.clear();
    free_[0] = cap;
    return VR_OK;
  }
  bool alloc(size_t bytes, size_t *off) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  char *ptr(size_t off) const { return mem.as<char>() + off; }
};

struct ServeScene {
  vr_scene *scene = nullptr;
  vr_xfers *xfers = nullptr;
  int64_t n = 0;
  DBuf samples;  // px py pz nx ny nz (6 n floats)
  DBuf refl, area;
  std::vector<vr_light> lights;
  DevTiles tl{};
};

// Parts of a request's block in the arena, each padded to 256 bytes.
struct ReqLayout {
  size_t emit0, emit1, total, e_row, scratch, energy, lights, counters, bytes;
  ReqLayout(int64_t n, int bounces, int n_lights) {
    auto pad = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t vb = pad(12 * (size_t)std::max<int64_t>(n, 1) + 16);
    const int64_t eb = (n + 2047) / 2048;
    size_t o = 0;
    emit0 = o; o += vb;
    emit1 = o; o += vb;
    total = o; o += vb;
    e_row = o; o += pad(8 * (size_t)std::max<int64_t>(n, 1));
    scratch = o; o += pad(8 * (size_t)(eb + 2));
    energy = o; o += pad(8 * (size_t)std::max(bounces, 1));
    lights = o; o += pad(sizeof(vr_light) * (size_t)std::max(n_lights, 1));
    counters = o; o += pad(4 * (size_t)std::max(bounces, 1));
    bytes = o;
  }
};

struct ServeReq {
  int64_t ticket = 0;
  int scene = 0;
  std::vector<vr_light> lights;
  std::chrono::steady_clock::time_point t_submit, t_done;
  size_t off = 0, bytes = 0;
  int slot = -1;
  int status = VR_OK;
  std::string err;
  bool done = false;
  std::vector<float> total;
  std::vector<double> energy;
  int n_bounces = 0;
};

}  // namespace
}  // namespace vr

struct vr_server {
  vr_server_params prm{};
  std::vector<std::unique_ptr<vr::ServeScene>> scenes;
  vr::Arena arena;
  DBuf sky, states, jobs_dev;
  vr::BounceJob *jobs_host = nullptr;  // pinned mirror of the job table
  hipStream_t stream = nullptr;
  // launch knobs fixed at creation: the serving thread reads these, never the
  // library-wide tuning a concurrent vr_set_tuning may change
  vr::Tuning tune{};
  vr::TiledShape shape{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<vr::ServeReq *> pending;
  std::vector<vr::ServeReq *> active;
  std::map<int64_t, std::unique_ptr<vr::ServeReq>> reqs;
  std::vector<int> free_slots;
  int64_t next_ticket = 1;
  bool stop = false, started = false;
  std::thread worker;
  vr_server_stats st{};
};

namespace vr {
namespace {

int serve_request_setup(vr_server *sv, ServeReq *r) {
  // re-gather the direct light with the moved lights, then emit = direct * refl
  const ServeScene &sc = *sv->scenes[(size_t)r->scene];
  const ReqLayout L(sc.n, sv->prm.bounces, (int)r->lights.size());
  char *base = sv->arena.ptr(r->off);
  hipStream_t st = sv->stream;
  if (!r->lights.empty())
    VR_HIP(hipMemcpyAsync(base + L.lights, r->lights.data(), sizeof(vr_light) * r->lights.size(),
                          hipMemcpyHostToDevice, st));
  const float *p = sc.samples.as<float>();
  const int64_t n = sc.n;
  DevSamples ds{p, p + n, p + 2 * n, p + 3 * n, p + 4 * n, p + 5 * n};
  float *direct = reinterpret_cast<float *>(base + L.emit1);  // free until the first bounce
  const int path = sv->prm.mode == VR_SERVE_PLAIN ? VR_PATH_PLAIN : VR_PATH_OPT;
  launch_gather_light(sc.scene->dev, ds, n, reinterpret_cast<vr_light *>(base + L.lights),
                      (int32_t)r->lights.size(), sv->sky.as<float>(), sv->prm.normal_offset, -1,
                      direct, path, sv->tune.trace_lds_entries, st);
  launch_emit_init(n, direct, sc.refl.as<float>(), reinterpret_cast<float *>(base + L.emit0),
                   reinterpret_cast<float *>(base + L.total), st);
  VR_HIP(hipMemsetAsync(sv->states.as<BounceState>() + r->slot, 0, sizeof(BounceState), st));
  return check_launch();
}

void serve_job(vr_server *sv, ServeReq *r, BounceJob &jb) {
  const ServeScene &sc = *sv->scenes[(size_t)r->scene];
  const ReqLayout L(sc.n, sv->prm.bounces, (int)r->lights.size());
  char *base = sv->arena.ptr(r->off);
  jb.x = sc.xfers->dev();
  jb.tl = sc.tl;
  jb.emit[0] = reinterpret_cast<float *>(base + L.emit0);
  jb.emit[1] = reinterpret_cast<float *>(base + L.emit1);
  jb.total = reinterpret_cast<float *>(base + L.total);
  jb.refl = sc.refl.as<float>();
  jb.area = sc.area.as<float>();
  jb.e_row = reinterpret_cast<double *>(base + L.e_row);
  jb.scratch = reinterpret_cast<double *>(base + L.scratch);
  jb.energy = reinterpret_cast<double *>(base + L.energy);
  jb.state = sv->states.as<BounceState>() + r->slot;
  jb.e_blocks = (sc.n + 2047) / 2048;
  jb.max_bounces = sv->prm.bounces;
  jb.stop_frac = 0.0;
}

int serve_finish(vr_server *sv, ServeReq *r, int n_bounces) {
  const ServeScene &sc = *sv->scenes[(size_t)r->scene];
  const ReqLayout L(sc.n, sv->prm.bounces, (int)r->lights.size());
  char *base = sv->arena.ptr(r->off);
  r->total.resize(3 * (size_t)sc.n);
  r->energy.resize((size_t)std::max(n_bounces, 0));
  if (sc.n > 0)
    VR_HIP(hipMemcpyAsync(r->total.data(), base + L.total, 12 * (size_t)sc.n,
                          hipMemcpyDeviceToHost, sv->stream));
  if (n_bounces > 0)
    VR_HIP(hipMemcpyAsync(r->energy.data(), base + L.energy, 8 * (size_t)n_bounces,
                          hipMemcpyDeviceToHost, sv->stream));
  r->n_bounces = n_bounces;
  return VR_OK;
}

// Plain sequential re-bake of one request: scratch-stack gather, one thread
// per CSR row in FP32, a host read of the energy after every bounce.
int serve_plain(vr_server *sv, ServeReq *r) {
  TRY(serve_request_setup(sv, r));
  const ServeScene &sc = *sv->scenes[(size_t)r->scene];
  const ReqLayout L(sc.n, sv->prm.bounces, (int)r->lights.size());
  char *base = sv->arena.ptr(r->off);
  hipStream_t st = sv->stream;
  DevXfers dx = sc.xfers->dev();
  dx.coef = sc.xfers->coef32.p;
  dx.kind = VR_COEF_F32;
  float *em[2] = {reinterpret_cast<float *>(base + L.emit0), reinterpret_cast<float *>(base + L.emit1)};
  BounceState *state = sv->states.as<BounceState>() + r->slot;
  BounceState hs{};
  int k = 0;
  for (; k < sv->prm.bounces; ++k) {
    launch_bounce(dx, em[k & 1], em[(k + 1) & 1], reinterpret_cast<float *>(base + L.total),
                  sc.refl.as<float>(), sc.area.as<float>(), reinterpret_cast<double *>(base + L.e_row),
                  reinterpret_cast<uint32_t *>(base + L.counters) + k, state, VR_PATH_PLAIN, st);
    launch_energy(reinterpret_cast<double *>(base + L.e_row), sc.n,
                  reinterpret_cast<double *>(base + L.scratch),
                  reinterpret_cast<double *>(base + L.energy), k, 0.0, state, st);
    TRY(check_launch());
    VR_HIP(hipMemcpyAsync(&hs, state, sizeof hs, hipMemcpyDeviceToHost, st));
    VR_HIP(hipStreamSynchronize(st));
    if (hs.converged) {
      ++k;
      break;
    }
  }
  TRY(serve_finish(sv, r, hs.k_done));
  VR_HIP(hipStreamSynchronize(st));
  return VR_OK;
}

// bounce launches of the batched server between two reads of the states
constexpr int kServeBurst = 8;

void serve_loop(vr_server *sv) {
  (void)hipSetDevice(g_device);
  std::vector<BounceState> hs((size_t)sv->prm.max_active);
  for (;;) {
    std::vector<ServeReq *> admitted;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty() || !sv->active.empty(); });
      if (sv->stop && sv->pending.empty() && sv->active.empty()) return;
      // admission: FIFO, while a slot and the request's footprint are free
      while (!sv->pending.empty() && !sv->free_slots.empty()) {
        ServeReq *r = sv->pending.front();
        size_t off = 0;
        if (!sv->arena.alloc(r->bytes, &off)) break;
        r->off = off;
        r->slot = sv->free_slots.back();
        sv->free_slots.pop_back();
        sv->pending.pop_front();
        admitted.push_back(r);
      }
    }
    std::vector<ServeReq *> finished;
    int rc = VR_OK;
    if (sv->prm.mode == VR_SERVE_PLAIN) {
      for (ServeReq *r : admitted) {
        rc = serve_plain(sv, r);
        r->status = rc;
        if (rc) r->err = vr_last_error();
        finished.push_back(r);
      }
    } else {
      for (ServeReq *r : admitted) {
        rc = serve_request_setup(sv, r);
        if (rc) {
          r->status = rc;
          r->err = vr_last_error();
          finished.push_back(r);
        } else {
          sv->active.push_back(r);
        }
      }
      const int na = (int)sv->active.size();
      if (na > 0) {
        // one bounce of every active request per batched launch, up to
        // kServeBurst launches (no more than the most bounces any active
        // request still needs) before the host reads the requests' states:
        // each launch skips a request that has stopped, so a burst only
        // saves host round trips; requests that arrive meanwhile join at
        // the next burst
        int most = 0;
        for (ServeReq *r : sv->active) {
          const bool fresh =
              std::find(admitted.begin(), admitted.end(), r) != admitted.end();
          const int done = fresh ? 0 : hs[(size_t)r->slot].k_done;
          most = std::max(most, sv->prm.bounces - done);
        }
        const int burst = std::max(1, std::min(kServeBurst, most));
        int64_t chunks = 0, eblocks = 0;
        for (int j = 0; j < na; ++j) {
          BounceJob &jb = sv->jobs_host[j];
          serve_job(sv, sv->active[(size_t)j], jb);
          jb.chunk_begin = chunks;
          jb.eblk_begin = eblocks;
          chunks += jb.tl.n_chunks;
          eblocks += jb.e_blocks;
        }
        hipStream_t st = sv->stream;
        hipError_t e = hipMemcpyAsync(sv->jobs_dev.p, sv->jobs_host, sizeof(BounceJob) * (size_t)na,
                                      hipMemcpyHostToDevice, st);
        for (int b = 0; b < burst && e == hipSuccess; ++b) {
          launch_bounce_jobs(sv->scenes[0]->xfers->precision, sv->shape,
                             sv->jobs_dev.as<BounceJob>(), na, chunks, eblocks, st);
          e = hipGetLastError();
        }
        if (e == hipSuccess)
          e = hipMemcpyAsync(hs.data(), sv->states.p, sizeof(BounceState) * hs.size(),
                             hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        std::vector<ServeReq *> still;
        for (ServeReq *r : sv->active) {
          const BounceState &b = hs[(size_t)r->slot];
          if (e != hipSuccess) {
            r->status = fail(VR_E_HIP, std::string("serving: ") + hipGetErrorString(e));
            r->err = vr_last_error();
            finished.push_back(r);
          } else if (b.converged || b.k_done >= sv->prm.bounces) {
            r->status = serve_finish(sv, r, b.k_done);
            if (r->status) r->err = vr_last_error();
            finished.push_back(r);
          } else {
            still.push_back(r);
          }
        }
        sv->active.swap(still);
        (void)hipStreamSynchronize(st);
        sv->st.iterations += burst;  // batched bounce launches
        sv->st.max_active_seen = std::max<int64_t>(sv->st.max_active_seen, na);
      }
    }
    const auto now = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(sv->mu);
    for (ServeReq *r : finished) {
      r->t_done = now;
      r->done = true;
      sv->arena.release(r->off, r->bytes);
      sv->free_slots.push_back(r->slot);
      sv->st.completed += 1;
    }
    sv->st.active = (int64_t)sv->active.size();
    if (!finished.empty()) sv->cv_done.notify_all();
  }
}

}  // namespace
}  // namespace vr

int vr_server_create(vr_server **out, const vr_server_params *p) {
  LOCK;
  if (!out || !p) return fail(VR_E_INVALID, "NULL argument");
  *out = nullptr;
  if (p->mode != VR_SERVE_BATCHED && p->mode != VR_SERVE_PLAIN)
    return fail(VR_E_INVALID, "unknown serving mode");
  if (p->max_active < 1 || p->max_active > 256) return fail(VR_E_INVALID, "max_active not in [1, 256]");
  if (p->bounces < 1 || p->bounces > 1000) return fail(VR_E_INVALID, "bounces not in [1, 1000]");
  TRY(ensure_init());
  std::unique_ptr<vr_server> sv(new vr_server);
  sv->prm = *p;
  sv->tune = tuning();
  if (sv->prm.mode == VR_SERVE_PLAIN) sv->prm.max_active = 1;
  if (sv->prm.hbm_budget <= 0) sv->prm.hbm_budget = (int64_t)8 << 30;
  if (!(sv->prm.normal_offset >= 0)) sv->prm.normal_offset = 0.1f;
  TRY(sv->arena.init((size_t)sv->prm.hbm_budget));
  TRY(sv->states.alloc(sizeof(BounceState) * (size_t)sv->prm.max_active));
  TRY(sv->jobs_dev.alloc(sizeof(BounceJob) * (size_t)sv->prm.max_active));
  VR_HIP(hipHostMalloc((void **)&sv->jobs_host, sizeof(BounceJob) * (size_t)sv->prm.max_active,
                       hipHostMallocDefault));
  VR_HIP(hipStreamCreate(&sv->stream));
  std::vector<float> dirs = sky_dirs();
  TRY(sv->sky.upload(dirs.data(), dirs.size() * 4, sv->stream));
  for (int i = sv->prm.max_active - 1; i >= 0; --i) sv->free_slots.push_back(i);
  VR_HIP(hipStreamSynchronize(sv->stream));
  *out = sv.release();
  return VR_OK;
}

int vr_server_add_scene(vr_server *sv, vr_scene *scene, vr_xfers *x, const vr_samples *smp,
                        const float *refl, const float *area, const vr_light *lights,
                        int32_t n_lights, int32_t *scene_id) {
  LOCK;
  if (!sv || !scene || !x || !smp || !scene_id) return fail(VR_E_INVALID, "NULL argument");
  if (!scene->built) return fail(VR_E_STATE, "scene not built");
  if (n_lights < 0 || (n_lights > 0 && !lights)) return fail(VR_E_INVALID, "bad lights");
  const int64_t n = x->n;
  if (n > 0 && (!refl || !area || !smp->px || !smp->py || !smp->pz || !smp->nx || !smp->ny || !smp->nz))
    return fail(VR_E_INVALID, "NULL patch array");
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->started) return fail(VR_E_STATE, "scenes are added before the first request");
  }
  if (sv->prm.mode == VR_SERVE_BATCHED) {
    if (!sv->scenes.empty() && sv->scenes[0]->xfers->precision != x->precision)
      return fail(VR_E_INVALID, "the scenes of a server share the coefficient precision");
    if (n > (int64_t)2048 * 2048) return fail(VR_E_CAPACITY, "a served scene holds at most 2048^2 patches");
  }
  std::unique_ptr<vr::ServeScene> sc(new vr::ServeScene);
  sc->scene = scene;
  sc->xfers = x;
  sc->n = n;
  sc->lights.assign(lights, lights + n_lights);
  const size_t fb = 4 * (size_t)std::max<int64_t>(n, 1);
  TRY(sc->samples.alloc(6 * fb));
  TRY(sc->refl.alloc(3 * fb));
  TRY(sc->area.alloc(fb));
  const float *src[6] = {smp->px, smp->py, smp->pz, smp->nx, smp->ny, smp->nz};
  if (n > 0) {
    for (int k = 0; k < 6; ++k)
      VR_HIP(hipMemcpyAsync(sc->samples.as<float>() + k * n, src[k], 4 * (size_t)n,
                            hipMemcpyHostToDevice, g_stream));
    VR_HIP(hipMemcpyAsync(sc->refl.p, refl, 12 * (size_t)n, hipMemcpyHostToDevice, g_stream));
    VR_HIP(hipMemcpyAsync(sc->area.p, area, 4 * (size_t)n, hipMemcpyHostToDevice, g_stream));
  }
  if (sv->prm.mode == VR_SERVE_BATCHED) {
    vr::TiledShape sh{};
    TRY(vr::tiled_shape(x, sv->tune, sh));
    if (!sv->scenes.empty() &&
        (sh.kernel != sv->shape.kernel || sh.cols != sv->shape.cols ||
         sh.group != sv->shape.group || sh.passes != sv->shape.passes ||
         sh.prefetch != sv->shape.prefetch || sh.placement != sv->shape.placement))
      return fail(VR_E_INVALID, "the scenes of a server share one gather shape");
    TRY(vr::prepare_tiles(x, sh, sc->tl, g_stream));
    sv->shape = sh;
  } else if (!x->coef32.p) {
    return fail(VR_E_STATE, "plain serving needs the FP32 coefficients");
  }
  VR_HIP(hipStreamSynchronize(g_stream));
  x->served += 1;
  *scene_id = (int32_t)sv->scenes.size();
  sv->scenes.push_back(std::move(sc));
  return VR_OK;
}

int vr_server_submit(vr_server *sv, int32_t scene_id, const int32_t *light_index,
                     const float *offset, int32_t n_moved, int64_t *ticket) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!sv || !ticket) return fail(VR_E_INVALID, "NULL argument");
  if (scene_id < 0 || scene_id >= (int32_t)sv->scenes.size()) return fail(VR_E_INVALID, "unknown scene");
  if (n_moved < 0 || (n_moved > 0 && (!light_index || !offset))) return fail(VR_E_INVALID, "bad moves");
  const vr::ServeScene &sc = *sv->scenes[(size_t)scene_id];
  std::unique_ptr<vr::ServeReq> r(new vr::ServeReq);
  r->scene = scene_id;
  r->lights = sc.lights;
  for (int i = 0; i < n_moved; ++i) {
    const int li = light_index[i];
    if (li < 0 || li >= (int)r->lights.size()) return fail(VR_E_INVALID, "light index out of range");
    for (int c = 0; c < 3; ++c) r->lights[(size_t)li].origin[c] += offset[3 * i + c];
  }
  r->bytes = vr::ReqLayout(sc.n, sv->prm.bounces, (int)r->lights.size()).bytes;
  if (r->bytes > sv->arena.cap)
    return fail(VR_E_CAPACITY, "request needs more device memory than the serving budget");
  r->t_submit = t0;
  std::lock_guard<std::mutex> lk(sv->mu);
  if (!sv->started) {
    sv->started = true;
    sv->worker = std::thread(vr::serve_loop, sv);
  }
  r->ticket = sv->next_ticket++;
  *ticket = r->ticket;
  sv->pending.push_back(r.get());
  sv->reqs[r->ticket] = std::move(r);
  sv->st.submitted += 1;
  sv->cv_work.notify_one();
  return VR_OK;
}

int vr_server_wait(vr_server *sv, int64_t ticket, int32_t timeout_ms, float *total, double *energy,
                   int32_t *n_bounces, double *latency_ms) {
  if (!sv) return fail(VR_E_INVALID, "NULL server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return fail(VR_E_INVALID, "unknown ticket");
  vr::ServeReq *r = it->second.get();
  auto ready = [&] { return r->done; };
  if (timeout_ms < 0) {
    sv->cv_done.wait(lk, ready);
  } else if (!sv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    return fail(VR_E_STATE, "request not finished");
  }
  if (r->status) {
    const int rc = fail(r->status, r->err);
    sv->reqs.erase(it);
    return rc;
  }
  if (total && !r->total.empty()) std::memcpy(total, r->total.data(), 4 * r->total.size());
  if (energy && !r->energy.empty()) std::memcpy(energy, r->energy.data(), 8 * r->energy.size());
  if (n_bounces) *n_bounces = r->n_bounces;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  sv->reqs.erase(it);
  return VR_OK;
}

int vr_server_get_stats(vr_server *sv, vr_server_stats *out) {
  if (!sv || !out) return fail(VR_E_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  *out = sv->st;
  out->pending = (int64_t)sv->pending.size();
  return VR_OK;
}

int vr_server_destroy(vr_server *sv) {
  if (!sv) return VR_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
    sv->cv_work.notify_all();
  }
  if (sv->worker.joinable()) sv->worker.join();
  if (sv->stream) {
    (void)hipStreamSynchronize(sv->stream);
    (void)hipStreamDestroy(sv->stream);
  }
  if (sv->jobs_host) (void)hipHostFree(sv->jobs_host);
  {
    LOCK;
    for (auto &sc : sv->scenes) sc->xfers->served -= 1;
  }
  delete sv;
  return VR_OK;
}

// ---- procedural inputs ---------------------------------------------------------
int vr_gen_room(vr_gen **out, uint64_t seed, int32_t n_pillars, float chop, float tess) {
  if (!out) return fail(VR_E_INVALID, "out is NULL");
  *out = nullptr;
  if (!(chop > 0) || !(tess > 0) || n_pillars < 0) return fail(VR_E_INVALID, "bad parameters");
  std::unique_ptr<vr_gen> g(new vr_gen);
  TRY(gen_room(seed, n_pillars, chop, tess, &g->g));
  *out = g.release();
  return VR_OK;
}

int vr_gen_level(vr_gen **out, uint64_t seed, int32_t rx, int32_t ry, float chop, float tess,
                 int32_t n_lights, int32_t sun) {
  if (!out) return fail(VR_E_INVALID, "out is NULL");
  *out = nullptr;
  if (rx < 1 || ry < 1 || !(chop > 0) || !(tess > 0) || n_lights < 0)
    return fail(VR_E_INVALID, "bad parameters");
  std::unique_ptr<vr_gen> g(new vr_gen);
  TRY(gen_level(seed, rx, ry, chop, tess, n_lights, sun, &g->g));
  *out = g.release();
  return VR_OK;
}

int vr_gen_counts(const vr_gen *g, int64_t *n_tri, int64_t *n_patches, int32_t *n_lights) {
  if (!g) return fail(VR_E_INVALID, "gen is NULL");
  if (n_tri) *n_tri = (int64_t)g->g.ids.size();
  if (n_patches) *n_patches = (int64_t)g->g.parea.size();
  if (n_lights) *n_lights = (int32_t)g->g.lights.size();
  return VR_OK;
}

int vr_gen_copy(const vr_gen *g, float *verts9, int32_t *ids, uint8_t *flags, float *po, float *pn,
                float *pa, float *pr, vr_light *lights) {
  if (!g) return fail(VR_E_INVALID, "gen is NULL");
  const GenScene &s = g->g;
  if (verts9) std::memcpy(verts9, s.v9.data(), 4 * s.v9.size());
  if (ids) std::memcpy(ids, s.ids.data(), 4 * s.ids.size());
  if (flags) std::memcpy(flags, s.flags.data(), s.flags.size());
  if (po) std::memcpy(po, s.po.data(), 4 * s.po.size());
  if (pn) std::memcpy(pn, s.pn.data(), 4 * s.pn.size());
  if (pa) std::memcpy(pa, s.parea.data(), 4 * s.parea.size());
  if (pr) std::memcpy(pr, s.prefl.data(), 4 * s.prefl.size());
  if (lights) std::memcpy(lights, s.lights.data(), sizeof(vr_light) * s.lights.size());
  return VR_OK;
}

int vr_make_patches(const vr_face *faces, int64_t n_faces, float max_chop, float min_chop,
                    vr_patch *out, int64_t cap, int64_t *n_out) {
  return vr_make_patches_phong(faces, n_faces, max_chop, min_chop, nullptr, out, cap, n_out);
}

int vr_pair_edges(const vr_face *faces, int64_t n_faces, float smoothing_threshold,
                  float *vertex_normals) {
  LOCK;
  if (n_faces < 0) return fail(VR_E_INVALID, "n_faces < 0");
  if (n_faces > 0 && (!faces || !vertex_normals)) return fail(VR_E_INVALID, "NULL argument");
  for (int64_t f = 0; f < n_faces; ++f)
    if (faces[f].n_points < 0 || faces[f].n_points > VR_FACE_MAX_POINTS)
      return fail(VR_E_INVALID, "face " + std::to_string(f) + ": n_points outside 0.." +
                                    std::to_string(VR_FACE_MAX_POINTS));
  pair_edges(faces, n_faces, smoothing_threshold, vertex_normals);
  return VR_OK;
}

int vr_make_patches_phong(const vr_face *faces, int64_t n_faces, float max_chop, float min_chop,
                          const float *vertex_normals, vr_patch *out, int64_t cap,
                          int64_t *n_out) {
  LOCK;
  if (!n_out) return fail(VR_E_INVALID, "n_out is NULL");
  *n_out = 0;
  if (n_faces < 0) return fail(VR_E_INVALID, "n_faces < 0");
  if (n_faces > 0 && !faces) return fail(VR_E_INVALID, "faces is NULL");
  if (n_faces >= ((int64_t)1 << 31)) return fail(VR_E_CAPACITY, "more than 2^31 faces");
  if (!(max_chop > 0.0f) || !(min_chop > 0.0f))
    return fail(VR_E_INVALID, "max_chop and min_chop must be > 0");
  for (int64_t f = 0; f < n_faces; ++f)
    if (faces[f].n_points < 0 || faces[f].n_points > VR_FACE_MAX_POINTS)
      return fail(VR_E_INVALID, "face " + std::to_string(f) + ": n_points outside 0.." +
                                    std::to_string(VR_FACE_MAX_POINTS));
  TRY(ensure_init());
  Prof p("make_patches", g_stream);
  return make_patches_device(faces, n_faces, max_chop, min_chop, vertex_normals, out, cap, n_out,
                             g_stream);
}

// Checks the BSP the kernels walk: every plane and child index reachable
// from node 0 is in range, no node is reached twice (a tree, so every walk
// ends), and no root-to-leaf path holds more than VR_BSP_MAX_DEPTH nodes (the
// bound of the walk's stack).  Then uploads the three lumps.
struct BspDev {
  DBuf planes, nodes, leaves;
  DevBsp dev{};
};
static int bsp_check(const vr_bsp *b) {
  if (!b) return fail(VR_E_INVALID, "bsp is NULL");
  if (b->n_nodes < 1 || b->n_planes < 1 || b->n_leafs < 0)
    return fail(VR_E_INVALID, "BSP needs at least one node and one plane");
  if (b->n_nodes >= (1ll << 31) || b->n_planes >= (1ll << 31) || b->n_leafs > (1ll << 31))
    return fail(VR_E_INVALID, "BSP lump too large");
  if (!b->planes || !b->nodes || (b->n_leafs > 0 && !b->leaf_cluster))
    return fail(VR_E_INVALID, "BSP lump is NULL");
  std::vector<uint8_t> seen((size_t)b->n_nodes, 0);
  std::vector<std::pair<int32_t, int32_t>> stk{{0, 1}};
  seen[0] = 1;
  while (!stk.empty()) {
    const auto [n, depth] = stk.back();
    stk.pop_back();
    const vr_bsp_node &nd = b->nodes[n];
    if (nd.planenum < 0 || nd.planenum >= b->n_planes)
      return fail(VR_E_INVALID, "node " + std::to_string(n) + ": plane index out of range");
    for (int k = 0; k < 2; ++k) {
      const int32_t c = nd.children[k];
      if (c < 0) {
        if ((int64_t)(-1 - (int64_t)c) >= b->n_leafs)
          return fail(VR_E_INVALID, "node " + std::to_string(n) + ": leaf index out of range");
        continue;
      }
      if (c >= b->n_nodes)
        return fail(VR_E_INVALID, "node " + std::to_string(n) + ": child index out of range");
      if (seen[(size_t)c]) return fail(VR_E_INVALID, "BSP nodes do not form a tree");
      if (depth + 1 > VR_BSP_MAX_DEPTH)
        return fail(VR_E_INVALID, "BSP deeper than " + std::to_string(VR_BSP_MAX_DEPTH));
      seen[(size_t)c] = 1;
      stk.push_back({c, depth + 1});
    }
  }
  return VR_OK;
}
static int bsp_upload(const vr_bsp *b, BspDev *d, hipStream_t st) {
  static_assert(sizeof(vr_bsp_plane) == 16 && sizeof(vr_bsp_node) == 12, "lump records");
  TRY(d->planes.upload(b->planes, 16 * (size_t)b->n_planes, st));
  TRY(d->nodes.upload(b->nodes, 12 * (size_t)b->n_nodes, st));
  TRY(d->leaves.upload(b->leaf_cluster, 2 * (size_t)b->n_leafs, st));
  d->dev.planes = d->planes.as<float4>();
  d->dev.nodes = d->nodes.as<int>();
  d->dev.leaf_cluster = d->leaves.as<int16_t>();
  return VR_OK;
}

int vr_calc_face_extents(const vr_face *faces, int64_t n, const float *lux_ofs,
                         vr_face_lightmap *out, int64_t *n_luxels) {
  LOCK;
  if (!n_luxels) return fail(VR_E_INVALID, "n_luxels is NULL");
  *n_luxels = 0;
  if (n < 0 || n >= (1ll << 31)) return fail(VR_E_INVALID, "bad face count");
  if (n == 0) return VR_OK;
  if (!faces || !out) return fail(VR_E_INVALID, "faces / out is NULL");
  TRY(ensure_init());
  static_assert(sizeof(vr_face_lightmap) == 32, "vr_face_lightmap layout");
  hipStream_t st = g_stream;
  DBuf df, dofs, dout;
  TRY(df.upload(faces, sizeof(vr_face) * (size_t)n, st));
  if (lux_ofs) TRY(dofs.upload(lux_ofs, 8 * (size_t)n, st));
  TRY(dout.alloc(sizeof(vr_face_lightmap) * (size_t)n));
  {
    Prof pr("face_extents", st);
    launch_face_extents(df.as<vr_face>(), n, lux_ofs ? dofs.as<float>() : nullptr,
                        dout.as<vr_face_lightmap>(), st);
    TRY(check_launch());
  }
  VR_HIP(hipMemcpyAsync(out, dout.p, sizeof(vr_face_lightmap) * (size_t)n, hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  int64_t at = 0;
  for (int64_t f = 0; f < n; ++f) {
    out[f].luxel_ofs = at;
    at += out[f].n_luxels;
  }
  *n_luxels = at;
  return VR_OK;
}

int vr_lighting_lump(const float *luxel_rgb, int64_t n_luxels, const vr_face_lightmap *lm,
                     int64_t n_faces, uint8_t *lump, int64_t lump_cap, int32_t *lightofs,
                     int64_t *lump_bytes) {
  LOCK;
  if (!lm || !lump_bytes) return fail(VR_E_INVALID, "NULL argument");
  if (n_faces < 0 || n_luxels < 0) return fail(VR_E_INVALID, "negative count");
  std::vector<int64_t> at((size_t)n_faces);
  int64_t e = 0;  // 4-byte entries
  for (int64_t f = 0; f < n_faces; ++f) {
    const int64_t n = lm[f].n_luxels;
    if (n <= 0) {
      at[(size_t)f] = -1;
      continue;
    }
    if (lm[f].luxel_ofs < 0 || lm[f].luxel_ofs > n_luxels - n)
      return fail(VR_E_INVALID, "face " + std::to_string(f) + ": luxels outside luxel_rgb");
    at[(size_t)f] = e;
    e += 1 + n;
  }
  if (4 * e > (int64_t)INT32_MAX)
    return fail(VR_E_CAPACITY, "lighting lump above 2 GiB (int32 face offsets)");
  *lump_bytes = 4 * e;
  if (lightofs)
    for (int64_t f = 0; f < n_faces; ++f)
      lightofs[f] = at[(size_t)f] < 0 ? -1 : (int32_t)(4 * (at[(size_t)f] + 1));
  if (!lump) return VR_OK;
  if (lump_cap < 4 * e) return fail(VR_E_CAPACITY, "lump buffer smaller than lump_bytes");
  if (e == 0) return VR_OK;
  if (!luxel_rgb) return fail(VR_E_INVALID, "luxel_rgb is NULL");
  TRY(ensure_init());
  Staging S;
  TRY(S.make({12 * (size_t)std::max<int64_t>(n_luxels, 1),
              sizeof(vr_face_lightmap) * (size_t)std::max<int64_t>(n_faces, 1),
              8 * (size_t)std::max<int64_t>(n_faces, 1), 4 * (size_t)e}));
  if (n_luxels > 0)
    VR_HIP(hipMemcpyAsync(S.ptr<float>(0), luxel_rgb, 12 * (size_t)n_luxels,
                          hipMemcpyHostToDevice, g_stream));
  VR_HIP(hipMemcpyAsync(S.ptr<vr_face_lightmap>(1), lm, sizeof(vr_face_lightmap) * (size_t)n_faces,
                        hipMemcpyHostToDevice, g_stream));
  VR_HIP(hipMemcpyAsync(S.ptr<int64_t>(2), at.data(), 8 * (size_t)n_faces, hipMemcpyHostToDevice,
                        g_stream));
  launch_lighting_lump(S.ptr<float>(0), S.ptr<vr_face_lightmap>(1), S.ptr<int64_t>(2), n_faces,
                       S.ptr<uint32_t>(3), g_stream);
  TRY(check_launch());
  VR_HIP(hipMemcpyAsync(lump, S.ptr<uint32_t>(3), 4 * (size_t)e, hipMemcpyDeviceToHost, g_stream));
  VR_HIP(hipStreamSynchronize(g_stream));
  return VR_OK;
}

int vr_face_luxels(const vr_face *faces, int64_t n_faces, const float *lux_ofs,
                   const vr_face_lightmap *lm, const vr_patch *patches, int64_t n_patches,
                   const float *patch_value, float *pos, float *normal, int32_t *patch,
                   float *value) {
  LOCK;
  if (n_faces < 0 || n_faces >= (1ll << 31) || n_patches < 0 || n_patches >= (1ll << 31))
    return fail(VR_E_INVALID, "bad face or patch count");
  if (n_faces == 0) return VR_OK;
  if (!faces || !lm) return fail(VR_E_INVALID, "faces / lm is NULL");
  if (n_patches > 0 && !patches) return fail(VR_E_INVALID, "patches is NULL");
  // the luxel layout must be the one vr_calc_face_extents assigns
  int64_t total = 0;
  for (int64_t f = 0; f < n_faces; ++f) {
    const vr_face_lightmap &L = lm[f];
    const bool ok = L.n_luxels == 0 ||
                    (L.mins[0] > -(1 << 30) && L.mins[0] < (1 << 30) && L.mins[1] > -(1 << 30) &&
                     L.mins[1] < (1 << 30) && L.size[0] >= 0 && L.size[1] >= 0 &&
                     L.size[0] <= VR_LIGHTMAP_MAX_DIM + 1 &&
                     L.size[1] <= VR_LIGHTMAP_MAX_DIM + 1 &&
                     L.n_luxels == (L.size[0] + 1) * (L.size[1] + 1));
    if (!ok || L.n_luxels < 0 || L.luxel_ofs != total)
      return fail(VR_E_INVALID, "face " + std::to_string(f) + ": lightmap record is not "
                                "vr_calc_face_extents's layout");
    total += L.n_luxels;
  }
  // each face's leaf patches, in patch order
  std::vector<int32_t> leaf_ofs((size_t)n_faces + 1, 0), leaf_idx;
  for (int64_t p = 0; p < n_patches; ++p) {
    if (patches[p].child1 != -1) continue;
    if (patches[p].face < 0 || patches[p].face >= n_faces)
      return fail(VR_E_INVALID, "patch " + std::to_string(p) + ": face index out of range");
    ++leaf_ofs[(size_t)patches[p].face + 1];
  }
  for (int64_t f = 0; f < n_faces; ++f) leaf_ofs[(size_t)f + 1] += leaf_ofs[(size_t)f];
  leaf_idx.assign((size_t)leaf_ofs.back(), 0);
  {
    std::vector<int32_t> at(leaf_ofs.begin(), leaf_ofs.end() - 1);
    for (int64_t p = 0; p < n_patches; ++p)
      if (patches[p].child1 == -1) leaf_idx[(size_t)at[(size_t)patches[p].face]++] = (int32_t)p;
  }
  std::vector<int2> blocks;
  for (int64_t f = 0; f < n_faces; ++f)
    for (int32_t l = 0; l < lm[f].n_luxels; l += 256) blocks.push_back(make_int2((int)f, l));
  if (total == 0) return VR_OK;
  TRY(ensure_init());
  hipStream_t st = g_stream;
  DBuf df, dofs, dlm, dblk, dlo, dli, dp, dv, opos, onrm, opat, oval;
  TRY(df.upload(faces, sizeof(vr_face) * (size_t)n_faces, st));
  if (lux_ofs) TRY(dofs.upload(lux_ofs, 8 * (size_t)n_faces, st));
  TRY(dlm.upload(lm, sizeof(vr_face_lightmap) * (size_t)n_faces, st));
  TRY(dblk.upload(blocks.data(), sizeof(int2) * blocks.size(), st));
  TRY(dlo.upload(leaf_ofs.data(), 4 * leaf_ofs.size(), st));
  TRY(dli.upload(leaf_idx.data(), 4 * leaf_idx.size(), st));
  TRY(dp.upload(patches, sizeof(vr_patch) * (size_t)n_patches, st));
  if (patch_value && value) TRY(dv.upload(patch_value, 12 * (size_t)n_patches, st));
  if (pos) TRY(opos.alloc(12 * (size_t)total));
  if (normal) TRY(onrm.alloc(12 * (size_t)total));
  if (patch) TRY(opat.alloc(4 * (size_t)total));
  if (value) TRY(oval.alloc(12 * (size_t)total));
  {
    Prof pr("face_luxels", st);
    launch_face_luxels(df.as<vr_face>(), lux_ofs ? dofs.as<float>() : nullptr,
                       dlm.as<vr_face_lightmap>(), dblk.as<int2>(), (int64_t)blocks.size(),
                       dlo.as<int32_t>(), dli.as<int32_t>(), dp.as<vr_patch>(),
                       (patch_value && value) ? dv.as<float>() : nullptr, opos.as<float>(),
                       onrm.as<float>(), opat.as<int32_t>(), oval.as<float>(), st);
    TRY(check_launch());
  }
  if (pos) VR_HIP(hipMemcpyAsync(pos, opos.p, 12 * (size_t)total, hipMemcpyDeviceToHost, st));
  if (normal) VR_HIP(hipMemcpyAsync(normal, onrm.p, 12 * (size_t)total, hipMemcpyDeviceToHost, st));
  if (patch) VR_HIP(hipMemcpyAsync(patch, opat.p, 4 * (size_t)total, hipMemcpyDeviceToHost, st));
  if (value) VR_HIP(hipMemcpyAsync(value, oval.p, 12 * (size_t)total, hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  return VR_OK;
}

int vr_cluster_from_points(const vr_bsp *bsp, const float *xyz, int64_t n, int16_t *cluster) {
  LOCK;
  if (n < 0) return fail(VR_E_INVALID, "n < 0");
  if (n > 0 && (!xyz || !cluster)) return fail(VR_E_INVALID, "xyz / cluster is NULL");
  TRY(bsp_check(bsp));
  TRY(ensure_init());
  hipStream_t st = g_stream;
  BspDev d;
  TRY(bsp_upload(bsp, &d, st));
  if (n == 0) return VR_OK;
  DBuf pts, out;
  TRY(pts.upload(xyz, 12 * (size_t)n, st));
  TRY(out.alloc(2 * (size_t)n));
  {
    Prof pr("cluster_from_points", st);
    launch_point_cluster(d.dev, pts.as<float>(), n, out.as<int16_t>(), st);
    TRY(check_launch());
  }
  VR_HIP(hipMemcpyAsync(cluster, out.p, 2 * (size_t)n, hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  return VR_OK;
}

int vr_patch_clusters(const vr_bsp *bsp, const vr_patch *patches, int64_t n, int32_t *cluster) {
  LOCK;
  if (n < 0) return fail(VR_E_INVALID, "n < 0");
  if (n > 0 && (!patches || !cluster)) return fail(VR_E_INVALID, "patches / cluster is NULL");
  TRY(bsp_check(bsp));
  TRY(ensure_init());
  hipStream_t st = g_stream;
  BspDev d;
  TRY(bsp_upload(bsp, &d, st));
  if (n == 0) return VR_OK;
  DBuf pp, out;
  TRY(pp.upload(patches, sizeof(vr_patch) * (size_t)n, st));
  TRY(out.alloc(4 * (size_t)n));
  {
    Prof pr("patch_clusters", st);
    launch_patch_cluster(d.dev, pp.as<vr_patch>(), n, out.as<int32_t>(), st);
    TRY(check_launch());
  }
  VR_HIP(hipMemcpyAsync(cluster, out.p, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  return VR_OK;
}

int vr_decompress_vis(const uint8_t *visdata, int64_t vis_len, const int32_t *pvs_ofs,
                      int32_t n_clusters, uint8_t *rows, int64_t *n_overrun) {
  LOCK;
  if (n_overrun) *n_overrun = 0;
  if (n_clusters < 0 || vis_len < 0) return fail(VR_E_INVALID, "negative size");
  if (n_clusters == 0) return VR_OK;
  if (!pvs_ofs || !rows || (vis_len > 0 && !visdata))
    return fail(VR_E_INVALID, "visdata / pvs_ofs / rows is NULL");
  for (int32_t c = 0; c < n_clusters; ++c) {
    if (pvs_ofs[c] == -1) return fail(VR_E_INVALID, "cluster " + std::to_string(c) + ": visofs == -1");
    if (pvs_ofs[c] < 0 || pvs_ofs[c] >= vis_len)
      return fail(VR_E_INVALID, "cluster " + std::to_string(c) + ": PVS offset outside the vis data");
  }
  TRY(ensure_init());
  hipStream_t st = g_stream;
  const size_t row = ((size_t)n_clusters + 7) >> 3;
  DBuf vis, ofs, out, flags;
  TRY(vis.upload(visdata, (size_t)vis_len, st));
  TRY(ofs.upload(pvs_ofs, 4 * (size_t)n_clusters, st));
  TRY(out.alloc(row * (size_t)n_clusters));
  TRY(flags.alloc(16));
  VR_HIP(hipMemsetAsync(flags.p, 0, 16, st));
  {
    Prof pr("decompress_vis", st);
    launch_decompress_vis(vis.as<uint8_t>(), vis_len, ofs.as<int32_t>(), n_clusters,
                          out.as<uint8_t>(), flags.as<unsigned long long>(),
                          flags.as<uint32_t>() + 2, st);
    TRY(check_launch());
  }
  unsigned long long f[2] = {0, 0};
  VR_HIP(hipMemcpyAsync(f, flags.p, 16, hipMemcpyDeviceToHost, st));
  VR_HIP(hipMemcpyAsync(rows, out.p, row * (size_t)n_clusters, hipMemcpyDeviceToHost, st));
  VR_HIP(hipStreamSynchronize(st));
  const uint32_t err = (uint32_t)f[1];
  if (err & 1u) return fail(VR_E_INVALID, "DecompressVis: 0 repeat");
  if (err & 2u) return fail(VR_E_INVALID, "DecompressVis: run-length data ends inside a row");
  if (n_overrun) *n_overrun = (int64_t)f[0];
  return VR_OK;
}

int vr_gen_faces(const vr_gen *g, float luxel_size, vr_face *out, int64_t *n_faces) {
  if (!g || !n_faces) return fail(VR_E_INVALID, "NULL argument");
  if (!(luxel_size > 0.0f)) return fail(VR_E_INVALID, "luxel_size must be > 0");
  const std::vector<double> &F = g->g.faces;
  const int64_t n = (int64_t)(F.size() / 12);
  *n_faces = n;
  if (!out) return VR_OK;
  for (int64_t k = 0; k < n; ++k) {
    const double *f = &F[(size_t)(12 * k)];
    vr_face &o = out[k];
    std::memset(&o, 0, sizeof o);
    o.n_points = 4;
    for (int c = 0; c < 3; ++c) {
      const double P = f[c], U = f[3 + c], V = f[6 + c];
      o.points[0][c] = (float)P;
      o.points[1][c] = (float)(P + U);
      o.points[2][c] = (float)(P + U + V);
      o.points[3][c] = (float)(P + V);
      o.normal[c] = (float)f[9 + c];
    }
    const double lu = std::sqrt(f[3] * f[3] + f[4] * f[4] + f[5] * f[5]);
    const double lv = std::sqrt(f[6] * f[6] + f[7] * f[7] + f[8] * f[8]);
    for (int c = 0; c < 3; ++c) {
      o.lux_s[c] = (float)(f[3 + c] / (lu * luxel_size));
      o.lux_t[c] = (float)(f[6 + c] / (lv * luxel_size));
    }
  }
  return VR_OK;
}

int vr_gen_grid(const vr_gen *g, double *xs, int32_t *nx, double *ys, int32_t *ny) {
  if (!g || !nx || !ny) return fail(VR_E_INVALID, "NULL argument");
  const std::vector<double> &X = g->g.grid_x, &Y = g->g.grid_y;
  *nx = X.empty() ? 0 : (int32_t)X.size() - 1;
  *ny = Y.empty() ? 0 : (int32_t)Y.size() - 1;
  if (xs) std::copy(X.begin(), X.end(), xs);
  if (ys) std::copy(Y.begin(), Y.end(), ys);
  return VR_OK;
}

int vr_gen_destroy(vr_gen *g) {
  delete g;
  return VR_OK;
}

}  // extern "C"
