// extern "C" boundary of libtlbm.so (declared in include/tlbm.h).
//
// Host side of the step library: argument checks, the specialisation table, launch shapes,
// the report-step FP64 reduction, managed/device memory, streams and events.  Everything a
// caller needs to drive the step from C, ctypes, cgo or JNI without C++ or torch types.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "adjoint.hpp"
#include "registry.hpp"
#include "tlbm.h"

namespace {

thread_local std::string g_error;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(TLBM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace

// used by the other translation units of the library (derived.hip)
namespace tlbm_detail {
__attribute__((visibility("hidden"))) int fail_message(const char* msg) {
  return fail(TLBM_ERR_INVALID, "%s", msg);
}
__attribute__((visibility("hidden"))) int fail_hip(hipError_t e, const char* what) {
  return hip_fail(e, what);
}
}  // namespace tlbm_detail

namespace {

#define TLBM_HIP(call)                                      \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return hip_fail(e_, #call);       \
  } while (0)

// ---- specialisation table ------------------------------------------------------------------
constexpr int kHandles = 2 * 3 * 3 * tlbm::kObstacles * 3 * 2;  // heat, storage, collision,
                                                                  // obstacle, side, diag

tlbm::StepTable g_tables[3][3];
tlbm::BatchTable g_batch_tables[3][3];
tlbm::HeatTable g_heat_tables[3];
// launch override per handle: threads per block along x (0 = default)
int g_block_x[kHandles];
std::once_flag g_once;

void build_tables() {
  tlbm_register_0_0(g_tables[0][0], g_batch_tables[0][0]);
  tlbm_register_0_1(g_tables[0][1], g_batch_tables[0][1]);
  tlbm_register_0_2(g_tables[0][2], g_batch_tables[0][2]);
  tlbm_register_1_0(g_tables[1][0], g_batch_tables[1][0]);
  tlbm_register_1_1(g_tables[1][1], g_batch_tables[1][1]);
  tlbm_register_1_2(g_tables[1][2], g_batch_tables[1][2]);
  tlbm_register_2_0(g_tables[2][0], g_batch_tables[2][0]);
  tlbm_register_2_1(g_tables[2][1], g_batch_tables[2][1]);
  tlbm_register_2_2(g_tables[2][2], g_batch_tables[2][2]);
  tlbm_register_heat_0(g_heat_tables[0]);
  tlbm_register_heat_1(g_heat_tables[1]);
  tlbm_register_heat_2(g_heat_tables[2]);
  for (int& b : g_block_x) b = 0;
}

struct Decoded {
  int storage, collision, obstacle, side, diag, heat;
};

bool decode(int h, Decoded& d) {
  if (h < 0 || h >= kHandles) return false;
  d.diag = h % 2;
  h /= 2;
  d.side = h % 3;
  h /= 3;
  d.obstacle = h % tlbm::kObstacles;
  h /= tlbm::kObstacles;
  d.collision = h % 3;
  h /= 3;
  d.storage = h % 3;
  d.heat = h / 3;
  return !(d.heat && d.storage != TLBM_F32);
}

// ---- report workspace, one per (device, stream), grow-only ----------------------------------
// A report step's block partials live here until its reduction (same stream) has read them, so
// report steps queued on different streams each need their own buffer; steps on one stream are
// ordered by the stream.  A graph captured from report steps holds the buffer's address in its
// kernel nodes, so each buffer counts the live graphs that use it: a buffer the stream outgrows
// (or whose stream is destroyed) is freed only once no graph refers to it any more.
struct WsBuffer {
  double* ptr = nullptr;
  size_t bytes = 0;
  int device = 0;
  int graphs = 0;        // live tlbm_graph objects whose kernels use this buffer
  bool retired = false;  // no longer a stream's current buffer: free when graphs reaches 0
};
std::mutex g_ws_mutex;
std::map<std::pair<int, hipStream_t>, WsBuffer*> g_ws;
// buffers handed to report steps recorded while a stream was capturing, per stream
std::map<hipStream_t, std::vector<WsBuffer*>> g_ws_captured;

// Frees a retired buffer that no graph uses (caller holds g_ws_mutex).
void ws_free_if_unused(WsBuffer* b) {
  if (!b->retired || b->graphs > 0) return;
  int current = 0;
  const bool restore = hipGetDevice(&current) == hipSuccess;
  if (b->ptr && hipSetDevice(b->device) == hipSuccess) {
    (void)hipDeviceSynchronize();  // queued work on any stream may still use it
    (void)hipFree(b->ptr);
  }
  if (restore) (void)hipSetDevice(current);
  delete b;
}

int workspace(size_t bytes, hipStream_t stream, double** out) {
  int dev = 0;
  TLBM_HIP(hipGetDevice(&dev));
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  TLBM_HIP(hipStreamIsCapturing(stream, &cs));
  const bool capturing = cs != hipStreamCaptureStatusNone;
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  WsBuffer*& w = g_ws[{dev, stream}];
  if (!w || w->bytes < bytes) {
    if (capturing)
      return fail(TLBM_ERR_INVALID,
                  "report workspace of this stream must grow to %zu bytes, which cannot happen "
                  "under graph capture: run one report step eagerly first", bytes);
    if (w) {
      w->retired = true;
      ws_free_if_unused(w);
      w = nullptr;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc (report workspace)");
    w = new WsBuffer();
    w->ptr = static_cast<double*>(p);
    w->bytes = bytes;
    w->device = dev;
  }
  if (capturing) {
    std::vector<WsBuffer*>& used = g_ws_captured[stream];
    if (std::find(used.begin(), used.end(), w) == used.end()) used.push_back(w);
  }
  *out = w->ptr;
  return TLBM_OK;
}

// The buffers a capture on `stream` used, now referenced by a graph (or dropped with it).
std::vector<WsBuffer*> take_captured(hipStream_t stream, bool keep) {
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  std::vector<WsBuffer*> used;
  auto it = g_ws_captured.find(stream);
  if (it != g_ws_captured.end()) {
    used.swap(it->second);
    g_ws_captured.erase(it);
  }
  if (keep)
    for (WsBuffer* b : used) ++b->graphs;
  return keep ? used : std::vector<WsBuffer*>();
}

void release_graph_buffers(std::vector<WsBuffer*>& used) {
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  for (WsBuffer* b : used) {
    --b->graphs;
    ws_free_if_unused(b);
  }
  used.clear();
}

// Retires the report workspaces of a stream (on every device) once its work has finished;
// a buffer that graphs still use is freed with the last of them.
void release_workspaces(hipStream_t stream) {
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  for (auto it = g_ws.begin(); it != g_ws.end();) {
    if (it->first.second == stream) {
      if (it->second) {
        it->second->retired = true;
        ws_free_if_unused(it->second);
      }
      it = g_ws.erase(it);
    } else {
      ++it;
    }
  }
  g_ws_captured.erase(stream);
}

// ---- partials of tlbm_step_noalloc: a fixed device array, no allocation ---------------------
// Each device holds its own copy; host code must use the address hipGetSymbolAddress gives on
// that device (the array's name is only a host-side handle), looked up once per device.
static __device__ __align__(256) double scratch[TLBM_NSUMS * TLBM_NOALLOC_MAX_BLOCKS];
std::mutex g_scratch_mutex;
std::map<int, double*> g_scratch_ptr;

int scratch_on_current_device(double** out) {
  int dev = 0;
  TLBM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    TLBM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(TLBM_ERR_HIP, "hipGetSymbolAddress gave NULL for the partials array");
    it = g_scratch_ptr.emplace(dev, static_cast<double*>(p)).first;
  }
  *out = it->second;
  return TLBM_OK;
}

// Sums partials[k][0..nblocks) in a fixed order (block k sums row k): deterministic FP64 totals.
__global__ void __launch_bounds__(1024) reduce_partials(const double* __restrict__ partials,
                                                        uint32_t nblocks, double* __restrict__ out) {
  __shared__ double red[1024];
  const int k = blockIdx.x;
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < nblocks; i += blockDim.x)
    s += partials[static_cast<size_t>(k) * nblocks + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

__global__ void coarsen_kernel(const float* __restrict__ src, float* __restrict__ dst, int nx,
                               int ny, int s, int mx, int my, int mz) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t m = static_cast<uint32_t>(mx) * my * mz;
  if (i >= m) return;
  const uint32_t a = i % mx, b = (i / mx) % my, c = i / (static_cast<uint32_t>(mx) * my);
  dst[i] = src[(static_cast<size_t>(c) * s * ny + static_cast<size_t>(b) * s) * nx + a * s];
}

// Plain device copy, 16 bytes per lane (the bandwidth calibration kernel of tlbm_copy).
__global__ void __launch_bounds__(256) copy16_kernel(const uint4* __restrict__ src,
                                                     uint4* __restrict__ dst, uint64_t n16) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

// Elements of one population buffer: 19 planes of qstride, or 19 planes of k nx per tile of k rows.
uint64_t population_elems(const tlbm_params& p) {
  if (p.tile_rows <= 0) return 19ull * static_cast<uint64_t>(p.qstride);
  const uint64_t k = static_cast<uint64_t>(p.tile_rows);
  const uint64_t rows = static_cast<uint64_t>(p.ny) * static_cast<uint64_t>(p.nz);
  return 19ull * k * static_cast<uint64_t>(p.nx) * ((rows + k - 1) / k);
}

// Elements of one temperature buffer: 7 planes of qstride, or 7 planes of k nx per tile of k rows.
uint64_t heat_population_elems(const tlbm_params& p) {
  return population_elems(p) / 19 * 7;
}

int check_params(const tlbm_params* p, int obstacle = -1) {
  if (!p) return fail(TLBM_ERR_INVALID, "params is NULL");
  if (obstacle == TLBM_FIELD && !p->phi)
    return fail(TLBM_ERR_INVALID, "a TLBM_FIELD obstacle needs params->phi");
  if (obstacle == TLBM_HEATSINK && (!(p->obs[7] >= 1.0f) || !(p->obs[6] >= 0.0f)))
    return fail(TLBM_ERR_INVALID, "heat sink fin pitch must be >= 1 and thickness >= 0");
  if (p->nx <= 0 || p->ny <= 0 || p->nz <= 0)
    return fail(TLBM_ERR_INVALID, "grid %dx%dx%d must be positive", p->nx, p->ny, p->nz);
  const uint64_t n = static_cast<uint64_t>(p->nx) * p->ny * p->nz;
  if (n > 0xFFFFFFFFull)
    return fail(TLBM_ERR_INVALID, "%llu cells exceed the 32-bit cell index",
                static_cast<unsigned long long>(n));
  if (p->qstride < static_cast<int64_t>(n))
    return fail(TLBM_ERR_INVALID, "qstride %lld < cells %llu", static_cast<long long>(p->qstride),
                static_cast<unsigned long long>(n));
  if (static_cast<uint64_t>(p->qstride) * 18 + n > (1ull << 40))
    return fail(TLBM_ERR_INVALID, "population buffer too large");
  if (p->tile_rows < 0 || p->tile_rows > (1 << 20) || (p->tile_rows & (p->tile_rows - 1)))
    return fail(TLBM_ERR_INVALID, "tile_rows %d is not 0 or a power of two up to 2^20",
                p->tile_rows);
  if (population_elems(*p) > (1ull << 40))
    return fail(TLBM_ERR_INVALID, "population buffer too large");
  if (p->xcd_group < 0 || p->xcd_group > 64 || (p->xcd_group & (p->xcd_group - 1)))
    return fail(TLBM_ERR_INVALID, "xcd_group %d is not 0 or a power of two up to 64",
                p->xcd_group);
  return TLBM_OK;
}

// Chunk order of a quiet launch (lbm_step_kernel and pull_copy_kernel, xcd_chunk): runs of
// `group` consecutive blocks per XCD.  0 picks 2 for 16-bit storage -- a pair of 256-cell blocks
// is one 512-cell row, whose x-shifted pulls then share their edge lines in one L2 (measured on
// the access pattern at 512x256x256: 5.81 -> 6.10 TB/s) -- and plain order (1) for FP32 and the
// temperature variant (FP32 pairs measured 6.08 -> 5.92 TB/s).
uint32_t quiet_group(const tlbm_params& p, int storage, bool heat) {
  if (p.xcd_group) return static_cast<uint32_t>(p.xcd_group);
  return heat || storage == TLBM_F32 ? 1u : 2u;
}

// Grid slots for `nblocks` chunks: whole runs of 8 groups (the extra slots exit at once).
uint64_t group_slots(uint64_t nblocks, uint32_t group) {
  if (group <= 1) return nblocks;
  const uint64_t run = 8ull * group;
  return (nblocks + run - 1) / run * run;
}

// Non-temporal population stores pay off once the two population buffers (`elements` each)
// are well beyond the MI355X's 256 MB Infinity Cache; measured on rows of 128-512 cells, plain
// stores win up to about 300 MB (the next step hits what they leave in the cache) and
// non-temporal ones from about 600 MB.
constexpr uint64_t kStreamStoreBytes = 512ull << 20;

bool stream_stores(uint64_t elements, int storage) {
  return 2 * elements * (storage == TLBM_F32 ? 4 : 2) > kStreamStoreBytes;
}

// elem_bytes: bytes per stored population (4 for FP32, 2 for FP16/BF16).
tlbm::StepArgs make_args(const tlbm_params& p, int obstacle, uint32_t elem_bytes) {
  tlbm::StepArgs a{};
  const bool tiled = p.tile_rows > 0;
  a.ns = tiled ? static_cast<uint64_t>(p.tile_rows) * static_cast<uint64_t>(p.nx)
               : static_cast<uint64_t>(p.qstride);
  // the temperature planes: tiled as the populations (7 planes per tile), or q-major with qstride
  a.gns = tiled ? a.ns : static_cast<uint64_t>(p.qstride);
  a.g_tile_stride = tiled ? 7ull * a.ns : 0ull;
  a.g_off32 = heat_population_elems(p) * 4 <= 0xFFFFFFFFull ? 1u : 0u;
  a.tile_stride = tiled ? 19ull * a.ns : 0ull;
  a.tile_log2 = 31u;
  a.tile_mask = 0xFFFFFFFFu;
  if (tiled) {
    uint32_t l = 0;
    while ((1 << l) < p.tile_rows) ++l;
    a.tile_log2 = l;
    a.tile_mask = static_cast<uint32_t>(p.tile_rows - 1);
  }
  a.row_waves = p.nx % 64 == 0 ? 1u : 0u;
  a.nx = p.nx;
  a.ny = p.ny;
  a.nz = p.nz;
  // Buffer-resource addressing (lbm_step.hpp) needs every plane offset plus every lane offset
  // below 2^32 bytes: the buffer's length plus a cell index and one row.
  const uint64_t cells = static_cast<uint64_t>(p.nx) * p.ny * p.nz;
  const uint64_t span = population_elems(p) + cells + static_cast<uint64_t>(p.nx);
  a.off32 = span * elem_bytes < (1ull << 32) ? 1u : 0u;
  if (a.off32) {
    const int64_t plane = static_cast<int64_t>(p.nx) * p.ny;
    const int64_t ns = static_cast<int64_t>(a.ns);
    for (int q = 0; q < tlbm::kQ; ++q) {
      const int64_t shift = tlbm::kCX[q] + tlbm::kCY[q] * static_cast<int64_t>(p.nx) +
                            tlbm::kCZ[q] * plane;
      a.plane_off[q] = static_cast<uint32_t>(q * ns * elem_bytes);
      // q-major planes: plane q's start less its pull shift, >= 0 for q >= 1 since
      // qstride >= cells >= the shift; q = 0 has no shift.  (Tiles add the source row's offset
      // to plane_off in the kernel instead.)
      a.pull_off[q] = tiled ? 0u : static_cast<uint32_t>((q * ns - shift) * elem_bytes);
    }
  }
  a.div_nx = tlbm::make_divisor(static_cast<uint32_t>(p.nx));
  a.div_ny = tlbm::make_divisor(static_cast<uint32_t>(p.ny));
  a.omega = p.omega;
  a.tau0 = p.tau0;
  a.smag = 25.456f * p.cs * p.cs;
  a.magic = p.magic;
  for (int i = 0; i < 8; ++i) a.o[i] = p.obs[i];
  float r = 0.0f;
  if (obstacle == TLBM_CYLINDER) {
    r = p.obs[2];
    a.reach[0] = a.reach[1] = std::fabs(r) + 2.0f;
    a.reach[2] = 0.0f;
  } else if (obstacle == TLBM_SPHERE) {
    r = p.obs[3];
    a.reach[0] = a.reach[1] = a.reach[2] = std::fabs(r) + 2.0f;
  } else if (obstacle == TLBM_BOX) {
    a.reach[0] = std::fabs(p.obs[3]) + 2.0f;
    a.reach[1] = std::fabs(p.obs[4]) + 2.0f;
    a.reach[2] = std::fabs(p.obs[5]) + 2.0f;
  }
  a.phi = p.phi;
  a.r2 = r * r;
  // periodic faces the obstacle reaches: cells on those faces test their wrapped sources
  const float ctr[3] = {p.obs[0], p.obs[1], obstacle == TLBM_CYLINDER ? 0.0f : p.obs[2]};
  const int ext[3] = {p.nx, p.ny, p.nz};
  a.wrap = obstacle == TLBM_FIELD ? 7 : 0;  // a field may be solid anywhere
  if (obstacle == TLBM_HEATSINK) {
    // the duct walls (y = 0, ny - 1) are within near()'s bound; the sink box may touch x or z
    if (p.obs[0] < 2.0f || p.obs[1] > static_cast<float>(p.nx) - 2.0f) a.wrap |= 1;
    if (p.obs[2] < 2.0f || p.obs[3] > static_cast<float>(p.nz) - 2.0f) a.wrap |= 4;
  }
  for (int d = 0; d < 3 && obstacle != TLBM_FIELD && obstacle != TLBM_HEATSINK; ++d) {
    if (d == 2 && obstacle == TLBM_CYLINDER) continue;  // infinite along z: z wraps change nothing
    const float lo = std::fabs(0.0f - ctr[d]), hi = std::fabs(static_cast<float>(ext[d] - 1) - ctr[d]);
    if (lo < a.reach[d] || hi < a.reach[d] || !std::isfinite(lo) || !std::isfinite(hi))
      a.wrap |= 1 << d;
  }
  a.u0 = p.u0;
  a.s0 = p.sponge_start;
  a.s_inv = 1.0f / (static_cast<float>(p.nx) - p.sponge_start);
  a.s_k = p.sponge_strength;
  for (int i = 0; i < 3; ++i) a.lid[i] = p.lid[i];
  a.omega_rot = p.omega_rot;
  return a;
}

template <int O, int S>
hipError_t launch_init(const tlbm::StepArgs& a, float seed_amp, float noise_amp, uint64_t seed,
                       dim3 g, dim3 b, hipStream_t st) {
  hipLaunchKernelGGL((tlbm::lbm_init_kernel<O, S>), g, b, 0, st, a, seed_amp, noise_amp, seed);
  return hipGetLastError();
}

using InitLaunch = hipError_t (*)(const tlbm::StepArgs&, float, float, uint64_t, dim3, dim3,
                                  hipStream_t);
const InitLaunch kInit[3][tlbm::kObstacles] = {
    {&launch_init<0, 0>, &launch_init<1, 0>, &launch_init<2, 0>, &launch_init<3, 0>,
     &launch_init<4, 0>},
    {&launch_init<0, 1>, &launch_init<1, 1>, &launch_init<2, 1>, &launch_init<3, 1>,
     &launch_init<4, 1>},
    {&launch_init<0, 2>, &launch_init<1, 2>, &launch_init<2, 2>, &launch_init<3, 2>,
     &launch_init<4, 2>}};
// [storage][obstacle]

// Per-domain totals of a batched report step: block (d, k) sums partials[k][chunk] over the
// chunks of domain d in a fixed order.
__global__ void __launch_bounds__(256) reduce_domains(const double* __restrict__ partials,
                                                      uint32_t nchunks,
                                                      const uint32_t* __restrict__ chunk_off,
                                                      double* __restrict__ out) {
  __shared__ double red[256];
  const int d = blockIdx.x, k = blockIdx.y;
  const uint32_t c0 = chunk_off[d], c1 = chunk_off[d + 1];
  double s = 0.0;
  for (uint32_t c = c0 + threadIdx.x; c < c1; c += blockDim.x)
    s += partials[static_cast<size_t>(k) * nchunks + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[static_cast<size_t>(d) * TLBM_NSUMS + k] = red[0];
}

}  // namespace

// A batch of independent lattices stepped by one launch (see include/tlbm.h).
struct tlbm_batch {
  int handle = -1;  // quiet handle of the specialisation (report = handle + 1)
  Decoded spec{};
  int device = 0;
  std::vector<tlbm_params> params;
  std::vector<tlbm::BatchDomain> dom;
  std::vector<uint32_t> chunk_off;  // ndom + 1
  uint64_t elements = 0, cells = 0;
  uint32_t nchunks = 0;
  bool nt = false;  // quiet steps use non-temporal population stores (large batches)
  int block = 256;
  int grid = 0;  // persistent grid (blocks); 0 = one block per chunk
  uint32_t group = 1;  // consecutive chunks per XCD run (batch_chunk)
  tlbm::BatchDomain* d_dom = nullptr;
  uint16_t* d_chunk_dom = nullptr;
  uint32_t* d_chunk_off = nullptr;
  double* d_partials = nullptr;
};

struct tlbm_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<WsBuffer*> workspaces;  // report workspaces the graph's kernels write
};

namespace {

void batch_free_tables(tlbm_batch* b) {
  if (b->d_dom) (void)hipFree(b->d_dom);
  if (b->d_chunk_dom) (void)hipFree(b->d_chunk_dom);
  if (b->d_chunk_off) (void)hipFree(b->d_chunk_off);
  if (b->d_partials) (void)hipFree(b->d_partials);
  b->d_dom = nullptr;
  b->d_chunk_dom = nullptr;
  b->d_chunk_off = nullptr;
  b->d_partials = nullptr;
}

// (Re)build the chunk tables for the current block size and upload them.
int batch_build(tlbm_batch* b) {
  const int ndom = static_cast<int>(b->params.size());
  b->dom.assign(ndom, tlbm::BatchDomain{});
  b->chunk_off.assign(ndom + 1, 0);
  uint64_t elem = 0, cell = 0, chunk = 0;
  for (int d = 0; d < ndom; ++d) {
    const tlbm_params& p = b->params[d];
    const uint64_t n = static_cast<uint64_t>(p.nx) * p.ny * p.nz;
    tlbm::BatchDomain& D = b->dom[d];
    // the domain's layout as a single lattice's: q-major planes of qstride, or tiles of rows
    D.a = make_args(p, b->spec.obstacle, b->spec.storage == TLBM_F32 ? 4u : 2u);
    D.elem_off = elem;
    D.cell_off = static_cast<uint32_t>(cell);
    D.chunk_off = static_cast<uint32_t>(chunk);
    b->chunk_off[d] = static_cast<uint32_t>(chunk);
    elem += population_elems(p);
    cell += n;
    chunk += (n + b->block - 1) / b->block;
  }
  if (cell > 0xFFFFFFFFull || chunk > 0xFFFFFFFFull)
    return fail(TLBM_ERR_INVALID, "batch of %llu cells is too large",
                static_cast<unsigned long long>(cell));
  b->chunk_off[ndom] = static_cast<uint32_t>(chunk);
  b->nt = stream_stores(elem, b->spec.storage);
  b->elements = elem;
  b->cells = cell;
  b->nchunks = static_cast<uint32_t>(chunk);
  std::vector<uint16_t> chunk_dom(b->nchunks);
  for (int d = 0; d < ndom; ++d)
    for (uint32_t c = b->chunk_off[d]; c < b->chunk_off[d + 1]; ++c)
      chunk_dom[c] = static_cast<uint16_t>(d);
  batch_free_tables(b);
  TLBM_HIP(hipMalloc(&b->d_dom, sizeof(tlbm::BatchDomain) * ndom));
  TLBM_HIP(hipMalloc(&b->d_chunk_dom, sizeof(uint16_t) * b->nchunks));
  TLBM_HIP(hipMalloc(&b->d_chunk_off, sizeof(uint32_t) * (ndom + 1)));
  TLBM_HIP(hipMalloc(&b->d_partials, sizeof(double) * TLBM_NSUMS * b->nchunks));
  TLBM_HIP(hipMemcpy(b->d_dom, b->dom.data(), sizeof(tlbm::BatchDomain) * ndom,
                     hipMemcpyHostToDevice));
  TLBM_HIP(hipMemcpy(b->d_chunk_dom, chunk_dom.data(), sizeof(uint16_t) * b->nchunks,
                     hipMemcpyHostToDevice));
  TLBM_HIP(hipMemcpy(b->d_chunk_off, b->chunk_off.data(), sizeof(uint32_t) * (ndom + 1),
                     hipMemcpyHostToDevice));
  return TLBM_OK;
}

}  // namespace

extern "C" {

const char* tlbm_last_error(void) { return g_error.c_str(); }

int tlbm_version(void) { return 1; }

int tlbm_device_count(int* count) {
  if (!count) return fail(TLBM_ERR_INVALID, "count is NULL");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    *count = 0;
    return fail(TLBM_ERR_NO_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  *count = n;
  return TLBM_OK;
}

int tlbm_set_device(int device) {
  TLBM_HIP(hipSetDevice(device));
  return TLBM_OK;
}

int tlbm_alloc_managed(size_t bytes, void** out) {
  if (!out) return fail(TLBM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (bytes == 0) bytes = 4;
  int dev = 0;
  TLBM_HIP(hipGetDevice(&dev));
  void* p = nullptr;
  hipError_t e = hipMallocManaged(&p, bytes, hipMemAttachGlobal);
  if (e != hipSuccess) return hip_fail(e, "hipMallocManaged");
  // Keep the pages on the GPU: the kernels write these fields, host readers (in-situ I/O)
  // fault them over on demand.  Advice is a hint; a runtime that declines it is not an error.
  if (hipMemAdvise(p, bytes, hipMemAdviseSetPreferredLocation, dev) != hipSuccess)
    (void)hipGetLastError();
  if (hipMemPrefetchAsync(p, bytes, dev, nullptr) != hipSuccess) (void)hipGetLastError();
  e = hipMemset(p, 0, bytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(p);
    return hip_fail(e, "zero-fill of managed allocation");
  }
  *out = p;
  return TLBM_OK;
}

int tlbm_alloc_device(size_t bytes, void** out) {
  if (!out) return fail(TLBM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (bytes == 0) bytes = 4;
  TLBM_HIP(hipMalloc(out, bytes));
  return TLBM_OK;
}

int tlbm_alloc_host(size_t bytes, void** out) {
  if (!out) return fail(TLBM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (bytes == 0) bytes = 4;
  TLBM_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
  return TLBM_OK;
}

int tlbm_free(void* ptr) {
  if (ptr) TLBM_HIP(hipFree(ptr));
  return TLBM_OK;
}

int tlbm_free_host(void* ptr) {
  if (ptr) TLBM_HIP(hipHostFree(ptr));
  return TLBM_OK;
}

int tlbm_memcpy(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return TLBM_OK;
  if (!dst || !src) return fail(TLBM_ERR_INVALID, "memcpy with NULL pointer");
  TLBM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
  return TLBM_OK;
}

int tlbm_memcpy_async(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return TLBM_OK;
  if (!dst || !src) return fail(TLBM_ERR_INVALID, "memcpy with NULL pointer");
  TLBM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, static_cast<hipStream_t>(stream)));
  return TLBM_OK;
}

int tlbm_memset(void* ptr, int value, size_t bytes, void* stream) {
  if (bytes == 0) return TLBM_OK;
  if (!ptr) return fail(TLBM_ERR_INVALID, "memset of NULL");
  // memset nodes in captured graphs have given wrong results on replays after the first
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  TLBM_HIP(hipStreamIsCapturing(static_cast<hipStream_t>(stream), &cs));
  if (cs != hipStreamCaptureStatusNone)
    return fail(TLBM_ERR_INVALID, "tlbm_memset is not allowed under graph capture");
  TLBM_HIP(hipMemsetAsync(ptr, value, bytes, static_cast<hipStream_t>(stream)));
  return TLBM_OK;
}

int tlbm_stream_create(void** stream) {
  if (!stream) return fail(TLBM_ERR_INVALID, "stream is NULL");
  hipStream_t s;
  TLBM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return TLBM_OK;
}

int tlbm_stream_destroy(void* stream) {
  if (!stream) return TLBM_OK;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  TLBM_HIP(hipStreamSynchronize(st));
  release_workspaces(st);  // the stream's report workspaces, now idle
  TLBM_HIP(hipStreamDestroy(st));
  return TLBM_OK;
}

int tlbm_stream_sync(void* stream) {
  TLBM_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return TLBM_OK;
}

int tlbm_device_sync(void) {
  TLBM_HIP(hipDeviceSynchronize());
  return TLBM_OK;
}

int tlbm_event_create(void** event) {
  if (!event) return fail(TLBM_ERR_INVALID, "event is NULL");
  hipEvent_t e;
  TLBM_HIP(hipEventCreate(&e));
  *event = e;
  return TLBM_OK;
}

int tlbm_event_destroy(void* event) {
  if (event) TLBM_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
  return TLBM_OK;
}

int tlbm_event_record(void* event, void* stream) {
  if (!event) return fail(TLBM_ERR_INVALID, "event is NULL");
  TLBM_HIP(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
  return TLBM_OK;
}

int tlbm_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!start || !stop || !ms) return fail(TLBM_ERR_INVALID, "NULL event or output");
  TLBM_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  TLBM_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
  return TLBM_OK;
}

int tlbm_kernel_lookup(const tlbm_spec* spec, int* handle) {
  if (!spec || !handle) return fail(TLBM_ERR_INVALID, "spec or handle is NULL");
  std::call_once(g_once, build_tables);
  const tlbm_spec& s = *spec;
  if (s.collision < 0 || s.collision > 2) return fail(TLBM_ERR_INVALID, "collision %d", s.collision);
  if (s.obstacle < 0 || s.obstacle >= tlbm::kObstacles)
    return fail(TLBM_ERR_INVALID, "obstacle %d", s.obstacle);
  if (s.diag < 0 || s.diag > 1) return fail(TLBM_ERR_INVALID, "diag %d", s.diag);
  if (s.side < 0 || s.side > 2) return fail(TLBM_ERR_INVALID, "side %d", s.side);
  if (s.storage < 0 || s.storage > 2) return fail(TLBM_ERR_INVALID, "storage %d", s.storage);
  if (s.heat < 0 || s.heat > 1) return fail(TLBM_ERR_INVALID, "heat %d", s.heat);
  if (s.heat && s.storage != TLBM_F32)
    return fail(TLBM_ERR_INVALID, "the temperature variant needs TLBM_F32 storage");
  *handle = (((((s.heat * 3 + s.storage) * 3 + s.collision) * tlbm::kObstacles + s.obstacle) * 3 +
              s.side) * 2) + s.diag;
  return TLBM_OK;
}

int tlbm_set_launch(int handle, int block_x) {
  std::call_once(g_once, build_tables);
  Decoded d;
  if (!decode(handle, d)) return fail(TLBM_ERR_INVALID, "bad kernel handle %d", handle);
  if (block_x < 0 || block_x > 256 || block_x % 64 != 0)
    return fail(TLBM_ERR_INVALID, "block_x must be 0, 64, 128, 192 or 256 (got %d)", block_x);
  g_block_x[handle] = block_x;
  return TLBM_OK;
}

int tlbm_population_elems(const tlbm_params* params, uint64_t* elems) {
  if (!elems) return fail(TLBM_ERR_INVALID, "elems is NULL");
  *elems = 0;
  if (int rc = check_params(params)) return rc;
  *elems = population_elems(*params);
  return TLBM_OK;
}

int tlbm_init(int handle, void* f, const tlbm_params* params, float seed_amp, float noise_amp,
              uint64_t noise_seed, void* stream) {
  std::call_once(g_once, build_tables);
  Decoded d;
  if (!decode(handle, d)) return fail(TLBM_ERR_INVALID, "bad kernel handle %d", handle);
  if (!f) return fail(TLBM_ERR_INVALID, "f is NULL");
  if (int rc = check_params(params, d.obstacle)) return rc;
  tlbm::StepArgs a = make_args(*params, d.obstacle, d.storage == TLBM_F32 ? 4u : 2u);
  a.fout = f;
  const uint64_t n = static_cast<uint64_t>(params->nx) * params->ny * params->nz;
  const uint64_t nblocks = (n + 255) / 256;
  if (nblocks * 256 > 0xFFFFFFFFull) return fail(TLBM_ERR_INVALID, "grid exceeds the launch limit");
  const dim3 block(256);
  const dim3 grid(static_cast<uint32_t>(nblocks));
  hipError_t e = kInit[d.storage][d.obstacle](a, seed_amp, noise_amp, noise_seed, grid, block,
                                              static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "lbm_init launch");
  return TLBM_OK;
}

}  // extern "C"

namespace {

// tlbm_step and tlbm_step_heat: `heat` is NULL for the flow step alone.
// scratch: a caller-provided partials buffer of scratch_blocks blocks (the no-allocation entry
// point), or NULL for the stream's workspace.
int step_impl(int handle, const Decoded& d, const void* f_in, void* f_out, const float* g_in,
              float* g_out, const tlbm_params* params, const tlbm_heat* heat,
              const tlbm_fields* fields, double* sums, void* stream, double* scratch = nullptr,
              uint32_t scratch_blocks = 0) {
  if (!f_in || !f_out) return fail(TLBM_ERR_INVALID, "f_in or f_out is NULL");
  if (f_in == f_out) return fail(TLBM_ERR_INVALID, "f_in and f_out must be distinct buffers");
  if (int rc = check_params(params, d.obstacle)) return rc;
  tlbm::StepArgs a = make_args(*params, d.obstacle, d.storage == TLBM_F32 ? 4u : 2u);
  a.fin = f_in;
  a.fout = f_out;
  // the temperature step adds 7 FP32 planes (g) to the 19 population planes
  a.stream_stores = stream_stores(population_elems(*params) +
                                      (heat ? heat_population_elems(*params) : 0),
                                  heat ? TLBM_F32 : d.storage) ? 1u : 0u;
  if (heat) {
    if (!g_in || !g_out) return fail(TLBM_ERR_INVALID, "g_in or g_out is NULL");
    if (g_in == g_out) return fail(TLBM_ERR_INVALID, "g_in and g_out must be distinct buffers");
    if (!heat->t_solid) return fail(TLBM_ERR_INVALID, "heat->t_solid is NULL");
    if (d.diag == TLBM_REPORT && (!heat->temp || !heat->q_solid))
      return fail(TLBM_ERR_INVALID, "a report step with heat needs heat->temp and heat->q_solid");
    if (!(heat->alpha >= 0.0f) || !(heat->pr_t > 0.0f))
      return fail(TLBM_ERR_INVALID, "heat alpha must be >= 0 and pr_t > 0");
    a.g_in = g_in;
    a.g_out = g_out;
    a.t_solid = heat->t_solid;
    a.q_solid = heat->q_solid;
    a.temp = heat->temp;
    a.alpha = heat->alpha;
    a.nu_mol = heat->nu;
    a.pr_t = heat->pr_t;
  }
  // one thread per cell, blocks of consecutive cells in index order (lbm_step_kernel); the
  // block size of this launch, else the handle's (tlbm_set_launch), else 256
  if (params->block_x < 0 || params->block_x > 256 || params->block_x % 64 != 0)
    return fail(TLBM_ERR_INVALID, "params->block_x must be 0, 64, 128, 192 or 256 (got %d)",
                params->block_x);
  const int bx = params->block_x ? params->block_x : (g_block_x[handle] ? g_block_x[handle] : 256);
  const uint64_t n = static_cast<uint64_t>(params->nx) * params->ny * params->nz;
  const uint64_t nblocks64 = (n + bx - 1) / bx;
  if (nblocks64 * bx > 0xFFFFFFFFull)
    return fail(TLBM_ERR_INVALID, "grid of %llu threads exceeds the launch limit",
                static_cast<unsigned long long>(nblocks64 * bx));
  const dim3 block(bx);
  dim3 grid(static_cast<uint32_t>(nblocks64));
  const uint32_t nblocks = grid.x;
  const uint32_t group = d.diag == TLBM_QUIET ? quiet_group(*params, d.storage, heat != nullptr) : 1u;
  a.xcd_group = group;
  a.nchunks = nblocks;
  const uint64_t slots = group_slots(nblocks64, group);
  if (slots * bx > 0xFFFFFFFFull) return fail(TLBM_ERR_INVALID, "grid exceeds the launch limit");
  grid = dim3(static_cast<uint32_t>(slots));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (d.diag == TLBM_REPORT) {
    if (!fields && !sums)
      return fail(TLBM_ERR_INVALID, "a report step needs the field arrays, the sums or both");
    if (fields) {
      const float* const* ptrs = reinterpret_cast<const float* const*>(fields);
      for (int k = 0; k < 13; ++k)
        if (!ptrs[k]) return fail(TLBM_ERR_INVALID, "report field %d is NULL", k);
      a.fld = *fields;
    }
    if (sums) {
      double* ws = scratch;
      if (scratch) {
        if (nblocks > scratch_blocks)
          return fail(TLBM_ERR_INVALID,
                      "report step of %u blocks exceeds the %u blocks of the no-allocation "
                      "partials array", nblocks, scratch_blocks);
      } else if (int rc = workspace(sizeof(double) * TLBM_NSUMS * nblocks, st, &ws)) {
        return rc;
      }
      a.partials = ws;
      a.nblocks = nblocks;
    }
  }
  hipError_t e = heat ? g_heat_tables[d.collision][d.obstacle][d.side][d.diag](a, grid, block, st)
                      : g_tables[d.storage][d.collision][d.obstacle][d.side][d.diag](a, grid, block,
                                                                                     st);
  if (e != hipSuccess) return hip_fail(e, "lbm_step launch");
  if (d.diag == TLBM_REPORT && sums) {
    hipLaunchKernelGGL(reduce_partials, dim3(TLBM_NSUMS), dim3(1024), 0, st, a.partials, nblocks,
                       sums);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "report reduction launch");
  }
  return TLBM_OK;
}

}  // namespace

extern "C" {

int tlbm_step(int handle, const void* f_in, void* f_out, const tlbm_params* params,
              const tlbm_fields* fields, double* sums, void* stream) {
  std::call_once(g_once, build_tables);
  Decoded d;
  if (!decode(handle, d)) return fail(TLBM_ERR_INVALID, "bad kernel handle %d", handle);
  if (d.heat) return fail(TLBM_ERR_INVALID, "handle %d has heat: step it with tlbm_step_heat", handle);
  return step_impl(handle, d, f_in, f_out, nullptr, nullptr, params, nullptr, fields, sums, stream);
}

int tlbm_step_noalloc(int handle, const void* f_in, void* f_out, const tlbm_params* params,
                      const tlbm_fields* fields, double* sums, void* stream) {
  std::call_once(g_once, build_tables);
  Decoded d;
  if (!decode(handle, d)) return fail(TLBM_ERR_INVALID, "bad kernel handle %d", handle);
  if (d.heat) return fail(TLBM_ERR_INVALID, "handle %d has heat: step it with tlbm_step_heat", handle);
  double* partials = nullptr;
  if (int rc = scratch_on_current_device(&partials)) return rc;
  // every block of a report step writes all TLBM_NSUMS of its own partials before the
  // reduction reads them, so the array needs no clearing between calls
  return step_impl(handle, d, f_in, f_out, nullptr, nullptr, params, nullptr, fields, sums, stream,
                   partials, TLBM_NOALLOC_MAX_BLOCKS);
}

int tlbm_step_heat(int handle, const void* f_in, void* f_out, const float* g_in, float* g_out,
                   const tlbm_params* params, const tlbm_heat* heat, const tlbm_fields* fields,
                   double* sums, void* stream) {
  std::call_once(g_once, build_tables);
  Decoded d;
  if (!decode(handle, d)) return fail(TLBM_ERR_INVALID, "bad kernel handle %d", handle);
  if (!d.heat) return fail(TLBM_ERR_INVALID, "handle %d has no heat: look it up with heat = 1", handle);
  if (!heat) return fail(TLBM_ERR_INVALID, "heat is NULL");
  return step_impl(handle, d, f_in, f_out, g_in, g_out, params, heat, fields, sums, stream);
}

int tlbm_copy(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return TLBM_OK;
  if (!dst || !src) return fail(TLBM_ERR_INVALID, "copy with NULL pointer");
  if (bytes % 16 || reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16)
    return fail(TLBM_ERR_INVALID, "copy needs 16-byte aligned pointers and size");
  const uint64_t n16 = bytes / 16;
  const uint64_t blocks = (n16 + 255) / 256;
  if (blocks > 0xFFFFFFFFull) return fail(TLBM_ERR_INVALID, "copy too large");
  hipLaunchKernelGGL(copy16_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const uint4*>(src),
                     static_cast<uint4*>(dst), n16);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "copy launch");
  return TLBM_OK;
}

int tlbm_pull_copy(const void* f_in, void* f_out, const tlbm_params* params, int storage,
                   void* stream) {
  if (!f_in || !f_out) return fail(TLBM_ERR_INVALID, "f_in or f_out is NULL");
  if (f_in == f_out) return fail(TLBM_ERR_INVALID, "f_in and f_out must be distinct buffers");
  if (storage < TLBM_F32 || storage > TLBM_BF16) return fail(TLBM_ERR_INVALID, "storage %d", storage);
  if (int rc = check_params(params)) return rc;
  tlbm::StepArgs a = make_args(*params, TLBM_CYLINDER, storage == TLBM_F32 ? 4u : 2u);
  a.fin = f_in;
  a.fout = f_out;
  const bool nt = stream_stores(population_elems(*params), storage);
  const uint64_t n = static_cast<uint64_t>(params->nx) * params->ny * params->nz;
  const uint64_t nblocks = (n + 255) / 256;
  a.xcd_group = quiet_group(*params, storage, false);  // the quiet step's chunk order
  a.nchunks = static_cast<uint32_t>(nblocks);
  const uint64_t slots = group_slots(nblocks, a.xcd_group);
  if (slots * 256 > 0xFFFFFFFFull) return fail(TLBM_ERR_INVALID, "grid exceeds the launch limit");
  const dim3 grid(static_cast<uint32_t>(slots)), block(256);
  hipStream_t st = static_cast<hipStream_t>(stream);
  // the quiet step's builds (registry.hpp launch_step): non-temporal stores on large lattices,
  // per-wave row addressing where rows are whole waves
  const bool rows = a.row_waves != 0;
  if (storage == TLBM_F32) {
    if (nt && rows) hipLaunchKernelGGL((tlbm::pull_copy_kernel<float, true, true>), grid, block, 0, st, a);
    else if (nt) hipLaunchKernelGGL((tlbm::pull_copy_kernel<float, true, false>), grid, block, 0, st, a);
    else if (rows) hipLaunchKernelGGL((tlbm::pull_copy_kernel<float, false, true>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((tlbm::pull_copy_kernel<float, false, false>), grid, block, 0, st, a);
  } else {
    if (nt && rows) hipLaunchKernelGGL((tlbm::pull_copy_kernel<uint16_t, true, true>), grid, block, 0, st, a);
    else if (nt) hipLaunchKernelGGL((tlbm::pull_copy_kernel<uint16_t, true, false>), grid, block, 0, st, a);
    else if (rows) hipLaunchKernelGGL((tlbm::pull_copy_kernel<uint16_t, false, true>), grid, block, 0, st, a);
    else hipLaunchKernelGGL((tlbm::pull_copy_kernel<uint16_t, false, false>), grid, block, 0, st, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "pull copy launch");
  return TLBM_OK;
}

int tlbm_coarsen(const float* src, float* dst, int nx, int ny, int nz, int s, int mx, int my,
                 int mz, void* stream) {
  if (!src || !dst) return fail(TLBM_ERR_INVALID, "src or dst is NULL");
  if (nx <= 0 || ny <= 0 || nz <= 0 || s <= 0 || mx <= 0 || my <= 0 || mz <= 0)
    return fail(TLBM_ERR_INVALID, "coarsen sizes must be positive");
  if (static_cast<int64_t>(mx - 1) * s >= nx || static_cast<int64_t>(my - 1) * s >= ny ||
      static_cast<int64_t>(mz - 1) * s >= nz)
    return fail(TLBM_ERR_INVALID, "coarse grid %dx%dx%d stride %d exceeds %dx%dx%d", mx, my, mz,
                s, nx, ny, nz);
  const uint64_t m = static_cast<uint64_t>(mx) * my * mz;
  if (m > 0xFFFFFFFFull) return fail(TLBM_ERR_INVALID, "coarse grid too large");
  const uint32_t blocks = static_cast<uint32_t>((m + 255) / 256);
  hipLaunchKernelGGL(coarsen_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     src, dst, nx, ny, s, mx, my, mz);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "coarsen launch");
  return TLBM_OK;
}

int tlbm_batch_create(int handle, const tlbm_params* params, int ndomains, tlbm_batch** out) {
  std::call_once(g_once, build_tables);
  if (!out) return fail(TLBM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  Decoded d;
  if (!decode(handle, d)) return fail(TLBM_ERR_INVALID, "bad kernel handle %d", handle);
  if (d.heat) return fail(TLBM_ERR_INVALID, "batched lattices have no temperature variant");
  if (!params || ndomains <= 0 || ndomains > 65535)
    return fail(TLBM_ERR_INVALID, "need 1..65535 domains (got %d)", ndomains);
  for (int i = 0; i < ndomains; ++i) {
    tlbm_params p = params[i];
    const int64_t n = static_cast<int64_t>(p.nx) * p.ny * p.nz;
    if (p.qstride < n) p.qstride = n;  // a plane stride below the cell count means "packed"
    if (int rc = check_params(&p, d.obstacle)) return rc;
  }
  tlbm_batch* b = new tlbm_batch();
  b->handle = handle - d.diag;
  b->spec = d;
  (void)hipGetDevice(&b->device);
  b->params.assign(params, params + ndomains);
  for (auto& p : b->params)
    p.qstride = std::max<int64_t>(p.qstride, static_cast<int64_t>(p.nx) * p.ny * p.nz);
  if (int rc = batch_build(b)) {
    batch_free_tables(b);
    delete b;
    return rc;
  }
  *out = b;
  return TLBM_OK;
}

int tlbm_batch_destroy(tlbm_batch* b) {
  if (!b) return TLBM_OK;
  batch_free_tables(b);
  delete b;
  return TLBM_OK;
}

int tlbm_batch_layout(const tlbm_batch* b, int domain, uint64_t* elem_offset,
                      uint64_t* cell_offset, uint64_t* elements, uint64_t* cells) {
  if (!b) return fail(TLBM_ERR_INVALID, "batch is NULL");
  const int ndom = static_cast<int>(b->params.size());
  if (domain < 0 || domain > ndom)
    return fail(TLBM_ERR_INVALID, "domain %d outside 0..%d", domain, ndom);
  if (elem_offset) *elem_offset = domain < ndom ? b->dom[domain].elem_off : b->elements;
  if (cell_offset) *cell_offset = domain < ndom ? b->dom[domain].cell_off : b->cells;
  if (elements) *elements = b->elements;
  if (cells) *cells = b->cells;
  return TLBM_OK;
}

int tlbm_batch_set_launch(tlbm_batch* b, int block_x, int grid_blocks) {
  if (!b) return fail(TLBM_ERR_INVALID, "batch is NULL");
  if (block_x == 0) block_x = 256;
  if (block_x < 64 || block_x > 256 || block_x % 64 != 0)
    return fail(TLBM_ERR_INVALID, "block_x must be 0, 64, 128, 192 or 256 (got %d)", block_x);
  if (grid_blocks < 0) return fail(TLBM_ERR_INVALID, "grid_blocks must be >= 0");
  b->grid = grid_blocks;
  if (block_x != b->block) {
    TLBM_HIP(hipDeviceSynchronize());  // tables may be in use by queued work
    b->block = block_x;
    return batch_build(b);
  }
  return TLBM_OK;
}

int tlbm_batch_set_order(tlbm_batch* b, int group) {
  if (!b) return fail(TLBM_ERR_INVALID, "batch is NULL");
  if (group < 1 || group > 64) return fail(TLBM_ERR_INVALID, "group must be 1..64 (got %d)", group);
  b->group = static_cast<uint32_t>(group);
  return TLBM_OK;
}

int tlbm_batch_init(tlbm_batch* b, void* f, float seed_amp, float noise_amp, uint64_t noise_seed,
                    void* stream) {
  if (!b || !f) return fail(TLBM_ERR_INVALID, "batch or f is NULL");
  for (size_t i = 0; i < b->params.size(); ++i) {
    const tlbm_params& p = b->params[i];
    tlbm::StepArgs a = b->dom[i].a;
    a.fout = static_cast<char*>(f) + b->dom[i].elem_off * (b->spec.storage == TLBM_F32 ? 4 : 2);
    const uint64_t n = static_cast<uint64_t>(p.nx) * p.ny * p.nz;
    const dim3 grid(static_cast<uint32_t>((n + 255) / 256));
    hipError_t e = kInit[b->spec.storage][b->spec.obstacle](a, seed_amp, noise_amp, noise_seed,
                                                           grid, dim3(256),
                                                           static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "batch init launch");
  }
  return TLBM_OK;
}

int tlbm_batch_step(tlbm_batch* b, const void* f_in, void* f_out, int report,
                    const tlbm_fields* fields, double* sums, void* stream) {
  if (!b || !f_in || !f_out) return fail(TLBM_ERR_INVALID, "batch, f_in or f_out is NULL");
  if (f_in == f_out) return fail(TLBM_ERR_INVALID, "f_in and f_out must be distinct buffers");
  tlbm::BatchArgs a{};
  a.fin = f_in;
  a.fout = f_out;
  a.dom = b->d_dom;
  a.chunk_dom = b->d_chunk_dom;
  a.ndom = static_cast<int32_t>(b->params.size());
  a.nchunks = b->nchunks;
  a.group = b->group;
  const uint64_t run = 8ull * b->group;
  a.nslots = b->group > 1 ? static_cast<uint32_t>((b->nchunks + run - 1) / run * run) : b->nchunks;
  if (report) {
    if (fields) {
      const float* const* ptrs = reinterpret_cast<const float* const*>(fields);
      for (int k = 0; k < 13; ++k)
        if (!ptrs[k]) return fail(TLBM_ERR_INVALID, "report field %d is NULL", k);
      a.fld = *fields;
    }
    a.partials = sums ? b->d_partials : nullptr;
  }
  // one block per chunk slot, or a persistent grid that strides over the slots (a multiple of
  // 8 blocks when chunks run in groups, so that a slot keeps its XCD)
  uint32_t grid = b->grid > 0 ? std::min<uint32_t>(b->grid, a.nslots) : a.nslots;
  if (b->group > 1 && b->grid > 0) grid = std::max<uint32_t>(8u, grid / 8u * 8u);
  const Decoded& d = b->spec;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = g_batch_tables[d.storage][d.collision][d.obstacle][d.side][report ? 1 : 0](
      a, b->nt, dim3(grid), dim3(b->block), st);
  if (e != hipSuccess) return hip_fail(e, "batch step launch");
  if (report && sums) {
    hipLaunchKernelGGL(reduce_domains, dim3(static_cast<uint32_t>(b->params.size()), TLBM_NSUMS),
                       dim3(256), 0, st, b->d_partials, b->nchunks, b->d_chunk_off, sums);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch reduction launch");
  }
  return TLBM_OK;
}

int tlbm_batch_graph(tlbm_batch* b, void* f_a, void* f_b, int steps, int report_every,
                     const tlbm_fields* fields, double* sums, void* stream, tlbm_graph** out) {
  if (!out) return fail(TLBM_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (!b || !f_a || !f_b) return fail(TLBM_ERR_INVALID, "batch or buffers are NULL");
  if (!stream) return fail(TLBM_ERR_INVALID, "graph capture needs a created (non-null) stream");
  if (steps <= 0 || steps % 2) return fail(TLBM_ERR_INVALID, "steps must be even and > 0");
  if (report_every < 0) return fail(TLBM_ERR_INVALID, "report_every must be >= 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  TLBM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = TLBM_OK;
  for (int i = 0; i < steps && rc == TLBM_OK; ++i) {
    const bool report = report_every > 0 && (i + 1) % report_every == 0;
    rc = tlbm_batch_step(b, i % 2 ? f_b : f_a, i % 2 ? f_a : f_b, report ? 1 : 0, fields,
                         report ? sums : nullptr, stream);
  }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc != TLBM_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_fail(e, "hipGraphInstantiate");
  }
  tlbm_graph* g = new tlbm_graph();
  g->graph = graph;
  g->exec = exec;
  *out = g;
  return TLBM_OK;
}

int tlbm_capture_begin(void* stream) {
  if (!stream) return fail(TLBM_ERR_INVALID, "graph capture needs a created (non-null) stream");
  take_captured(static_cast<hipStream_t>(stream), false);  // forget an abandoned capture's list
  TLBM_HIP(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                 hipStreamCaptureModeThreadLocal));
  return TLBM_OK;
}

int tlbm_capture_end(void* stream, tlbm_graph** out) {
  if (!stream) return fail(TLBM_ERR_INVALID, "graph capture needs a created (non-null) stream");
  if (out) *out = nullptr;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (e != hipSuccess) {
    take_captured(st, false);
    if (graph) (void)hipGraphDestroy(graph);
    return hip_fail(e, "hipStreamEndCapture");
  }
  if (!out) {
    take_captured(st, false);
    (void)hipGraphDestroy(graph);
    return fail(TLBM_ERR_INVALID, "out is NULL (the capture was discarded)");
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    take_captured(st, false);
    (void)hipGraphDestroy(graph);
    return hip_fail(e, "hipGraphInstantiate");
  }
  tlbm_graph* g = new tlbm_graph();
  g->graph = graph;
  g->exec = exec;
  g->workspaces = take_captured(st, true);  // kept alive until tlbm_graph_destroy
  *out = g;
  return TLBM_OK;
}

int tlbm_graph_launch(tlbm_graph* g, void* stream) {
  if (!g) return fail(TLBM_ERR_INVALID, "graph is NULL");
  TLBM_HIP(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  return TLBM_OK;
}

int tlbm_graph_destroy(tlbm_graph* g) {
  if (!g) return TLBM_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  release_graph_buffers(g->workspaces);
  delete g;
  return TLBM_OK;
}

int tlbm_adjoint_step(const tlbm_params* params, const float* f_t, const float* f_t1,
                      const float* mu_in, float* mu_out, float lam_t, float lam_t1,
                      const int32_t* block_origin, const int32_t* block_size, double* gphi,
                      int final_links, void* stream) {
  if (int rc = check_params(params, TLBM_FIELD)) return rc;
  if (params->tile_rows != 0)
    return fail(TLBM_ERR_INVALID, "the adjoint step takes q-major population planes (tile_rows 0)");
  if (!f_t1 || !mu_in || (!final_links && (!f_t || !mu_out)))
    return fail(TLBM_ERR_INVALID, "adjoint buffers are NULL");
  if (!final_links && mu_in == mu_out)
    return fail(TLBM_ERR_INVALID, "mu_in and mu_out must be distinct buffers");
  if (!block_origin || !block_size || !gphi)
    return fail(TLBM_ERR_INVALID, "parameter block and its gradient buffer are required");
  tlbm::AdjArgs a{};
  a.f_t = f_t;
  a.f_t1 = f_t1;
  a.mu_in = mu_in;
  a.mu_out = mu_out;
  a.phi = params->phi;
  a.ns = static_cast<uint64_t>(params->qstride);  // >= n (check_params)
  a.nx = params->nx;
  a.ny = params->ny;
  a.nz = params->nz;
  a.omega = params->omega;
  a.u0 = params->u0;
  a.s0 = params->sponge_start;
  a.s_inv = 1.0f / (static_cast<float>(params->nx) - params->sponge_start);
  a.s_k = params->sponge_strength;
  a.lam_t = lam_t;
  a.lam_t1 = lam_t1;
  const int32_t dims[3] = {params->nx, params->ny, params->nz};
  for (int i = 0; i < 3; ++i) {
    a.blk[i] = block_origin[i];
    a.bn[i] = block_size[i];
    if (block_size[i] <= 0 || block_origin[i] < 0 ||
        static_cast<int64_t>(block_origin[i]) + block_size[i] > dims[i])
      return fail(TLBM_ERR_INVALID, "parameter block [%d, %d) is not inside [0, %d) on axis %d",
                  block_origin[i], block_origin[i] + block_size[i], dims[i], i);
  }
  a.gphi = gphi;
  // one thread per cell, blocks of 256 consecutive cells (lbm_adjoint_kernel)
  const uint64_t n = static_cast<uint64_t>(params->nx) * params->ny * params->nz;
  const uint64_t nblocks = (n + 255) / 256;
  a.div_nx = tlbm::make_divisor(static_cast<uint32_t>(params->nx));
  a.div_ny = tlbm::make_divisor(static_cast<uint32_t>(params->ny));
  if (nblocks * 256 > 0xFFFFFFFFull) return fail(TLBM_ERR_INVALID, "grid exceeds the launch limit");
  const dim3 grid(static_cast<uint32_t>(nblocks)), block(256);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (final_links)
    hipLaunchKernelGGL(tlbm::lbm_adjoint_kernel<true>, grid, block, 0, st, a);
  else
    hipLaunchKernelGGL(tlbm::lbm_adjoint_kernel<false>, grid, block, 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "adjoint launch");
  return TLBM_OK;
}

int tlbm_device_name(char* buf, int len) {
  if (!buf || len <= 0) return fail(TLBM_ERR_INVALID, "name buffer is NULL or empty");
  int dev = 0;
  TLBM_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  TLBM_HIP(hipGetDeviceProperties(&prop, dev));
  snprintf(buf, static_cast<size_t>(len), "%s", prop.gcnArchName);
  return TLBM_OK;
}

int tlbm_device_cus(int* cus) {
  if (!cus) return fail(TLBM_ERR_INVALID, "cus is NULL");
  int dev = 0;
  TLBM_HIP(hipGetDevice(&dev));
  TLBM_HIP(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
  return TLBM_OK;
}

}  // extern "C"
