// Derived fields of the output grid: the velocity gradient at the points, and the vorticity
// magnitude and Q-criterion at the cell centres (derived.py:48-93 and the default expressions
// "mag(curl(velocity))@cells", "qcrit(velocity)@cells", derived.py:26-29).
//
// Grid: mx x my x mz points, index p = (k * my + j) * mx + i, spacing h (inv_h = 1 / h); cells
// are (mx - 1)(my - 1)(mz - 1), indexed the same way over their lower corner.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "tlbm.h"

namespace tlbm {

// d u_c / d x_d at every point: central differences, one-sided at the grid's faces (a face
// point differences with its one neighbour over one spacing), zero along an axis of one point
// (derived.py:52-89).  jac holds 9 planes of m floats: jac[(3 c + d) m + p].
namespace {
__global__ void __launch_bounds__(256)
    velocity_gradient_kernel(const float* __restrict__ ux, const float* __restrict__ uy,
                             const float* __restrict__ uz, float* __restrict__ jac, int mx, int my,
                             int mz, float inv_h) {
  const uint32_t m = static_cast<uint32_t>(mx) * my * mz;
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  const int i = static_cast<int>(p % static_cast<uint32_t>(mx));
  const int j = static_cast<int>((p / static_cast<uint32_t>(mx)) % static_cast<uint32_t>(my));
  const int k = static_cast<int>(p / (static_cast<uint32_t>(mx) * my));
  const int xl = i > 0 ? i - 1 : i, xh = i < mx - 1 ? i + 1 : i;
  const int yl = j > 0 ? j - 1 : j, yh = j < my - 1 ? j + 1 : j;
  const int zl = k > 0 ? k - 1 : k, zh = k < mz - 1 ? k + 1 : k;
  const float sx = xh > xl ? inv_h / static_cast<float>(xh - xl) : 0.0f;
  const float sy = yh > yl ? inv_h / static_cast<float>(yh - yl) : 0.0f;
  const float sz = zh > zl ? inv_h / static_cast<float>(zh - zl) : 0.0f;
  const uint32_t row = static_cast<uint32_t>(mx) * (static_cast<uint32_t>(j) + my * k);
  const uint32_t x0 = row + xl, x1 = row + xh;
  const uint32_t y0 = i + mx * (static_cast<uint32_t>(yl) + my * k);
  const uint32_t y1 = i + mx * (static_cast<uint32_t>(yh) + my * k);
  const uint32_t z0 = i + mx * (static_cast<uint32_t>(j) + my * zl);
  const uint32_t z1 = i + mx * (static_cast<uint32_t>(j) + my * zh);
  const float* const u[3] = {ux, uy, uz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    jac[(3 * c + 0) * m + p] = (u[c][x1] - u[c][x0]) * sx;
    jac[(3 * c + 1) * m + p] = (u[c][y1] - u[c][y0]) * sy;
    jac[(3 * c + 2) * m + p] = (u[c][z1] - u[c][z0]) * sz;
  }
}
}  // namespace

// |curl u| and Q = -1/2 (G_00^2 + G_11^2 + G_22^2) - (G_01 G_10 + G_02 G_20 + G_12 G_21) at each
// cell centre, with G the cell's velocity gradient: BASIS = false, the mean of its eight
// corners' point gradients (the solver's gradient, the points' as velocity_gradient_kernel forms
// them); BASIS = true, the trilinear cell's gradient at its centre, the mean of the four edge
// differences along each axis.  Both read the velocity only: no point gradient goes through HBM.
//
// Mapping.  A block is 64 x 4 cell columns (one wave per y row) and a chunk of 32 cells along
// z; it walks the point planes of its chunk upwards.  Each plane's tile (the block's corner
// points plus one point before and two after along x and y, 67 x 7, three components) is
// staged in LDS (double-buffered, one barrier per plane); the global loads of the next plane
// are issued before the current one is used, so their latency overlaps the arithmetic.  From
// LDS each thread forms, for each component, three sums over its 2 x 2 corner points (x in
// {i, i+1}, y in {j, j+1}):
//   px(z): the x-derivatives (lattice) or the x edge differences (basis),
//   py(z): the same along y,
//   s(z):  the four values, whose differences across planes give the z-derivatives;
// cell k takes planes k and k + 1 (lattice z: s at planes k - 1 .. k + 2, one-sided at the
// grid's faces).  Blocks go to the XCDs in contiguous runs of tiles, so the halo rows that
// neighbouring blocks share meet in one L2.
constexpr int kCellTileX = 64;
constexpr int kCellTileY = 4;
constexpr int kCellChunkZ = 32;
constexpr int kHaloX = kCellTileX + 3;   // points x0 - 1 .. x0 + 65 of a tile at x0
constexpr int kHaloY = kCellTileY + 3;
constexpr int kPitch = kHaloX + 1;
constexpr int kTileFloats = 3 * kHaloY * kPitch;
constexpr int kCellThreads = kCellTileX * kCellTileY;
constexpr int kSlots = (kHaloY * kHaloX + kCellThreads - 1) / kCellThreads;  // 2

// A thread's corner column in tile coordinates (point (x, y) of the tile is at x - x0 + 1,
// y - y0 + 1 for a tile of cells starting at x0, y0).
struct CellColumn {
  int lx, lxl, lxh1;   // corner column i, its left neighbour, i + 1's right neighbour
  int ly, lyl, lyh1;   // corner row j, its lower neighbour, j + 1's upper neighbour
  float sx0, sx1, sy0, sy1;
};

__device__ __forceinline__ float at(const float* t, int c, int ly, int lx) {
  return t[(c * kHaloY + ly) * kPitch + lx];
}

template <bool BASIS>
__device__ __forceinline__ void plane_sums(const float* t, int c, const CellColumn& k, float& px,
                                           float& py, float& s) {
  const float b0 = at(t, c, k.ly, k.lx), c0 = at(t, c, k.ly, k.lx + 1);
  const float b1 = at(t, c, k.ly + 1, k.lx), c1 = at(t, c, k.ly + 1, k.lx + 1);
  s = (b0 + c0) + (b1 + c1);
  if (BASIS) {
    px = (c0 - b0) + (c1 - b1);
    py = (b1 - b0) + (c1 - c0);
  } else {
    const float a0 = at(t, c, k.ly, k.lxl), d0 = at(t, c, k.ly, k.lxh1);
    const float a1 = at(t, c, k.ly + 1, k.lxl), d1 = at(t, c, k.ly + 1, k.lxh1);
    const float e0 = at(t, c, k.lyl, k.lx), e1 = at(t, c, k.lyl, k.lx + 1);
    const float h0 = at(t, c, k.lyh1, k.lx), h1 = at(t, c, k.lyh1, k.lx + 1);
    // d/dx at corners (i, y) and (i + 1, y), y = j, j + 1; d/dy at corners (x, j), (x, j + 1)
    px = ((c0 - a0) + (c1 - a1)) * k.sx0 + ((d0 - b0) + (d1 - b1)) * k.sx1;
    py = ((b1 - e0) + (c1 - e1)) * k.sy0 + ((h0 - b0) + (h1 - c0)) * k.sy1;
  }
}

template <bool BASIS>
__global__ void __launch_bounds__(kCellThreads)
    vortex_fields_kernel(const float* __restrict__ ux, const float* __restrict__ uy,
                         const float* __restrict__ uz, float* __restrict__ vorticity,
                         float* __restrict__ qcrit, int mx, int my, int mz, float inv_h,
                         uint32_t xtiles, uint32_t ytiles, uint32_t tiles, uint32_t per_xcd) {
  __shared__ float lds[2][kTileFloats];
  const uint32_t b = blockIdx.x;
  const uint32_t tile = (b % 8u) * per_xcd + b / 8u;   // XCD b % 8 takes a contiguous run
  if (tile >= tiles) return;                           // uniform over the block
  const int cx = mx - 1, cy = my - 1, cz = mz - 1;
  const uint32_t tx = tile % xtiles, ty = (tile / xtiles) % ytiles, tz = tile / (xtiles * ytiles);
  const int x0 = static_cast<int>(tx) * kCellTileX, y0 = static_cast<int>(ty) * kCellTileY;
  const int k0 = static_cast<int>(tz) * kCellChunkZ;
  const int k1 = min(k0 + kCellChunkZ, cz);           // cells k0 .. k1 - 1, corner planes k0 .. k1
  // planes zlo .. zhi, walked upwards (lattice: one more on each side for the z-derivatives)
  const int zlo = (!BASIS && k0 > 0) ? k0 - 1 : k0;
  const int zhi = BASIS ? k1 : min(k1 + 1, mz - 1);
  // every thread stages planes and meets the barriers; threads past the grid's last cell
  // column or row compute the last one's values and store nothing
  const int ir = x0 + static_cast<int>(threadIdx.x % kCellTileX);
  const int jr = y0 + static_cast<int>(threadIdx.x / kCellTileX);
  const bool active = ir < cx && jr < cy;
  const int i = min(ir, cx - 1), j = min(jr, cy - 1);
  CellColumn k;
  const int xl = i > 0 ? i - 1 : 0, xh1 = i + 2 < mx ? i + 2 : i + 1;
  const int yl = j > 0 ? j - 1 : 0, yh1 = j + 2 < my ? j + 2 : j + 1;
  k.lx = i - x0 + 1;
  k.lxl = xl - x0 + 1;
  k.lxh1 = xh1 - x0 + 1;
  k.ly = j - y0 + 1;
  k.lyl = yl - y0 + 1;
  k.lyh1 = yh1 - y0 + 1;
  k.sx0 = inv_h / static_cast<float>(i + 1 - xl);
  k.sx1 = inv_h / static_cast<float>(xh1 - i);
  k.sy0 = inv_h / static_cast<float>(j + 1 - yl);
  k.sy1 = inv_h / static_cast<float>(yh1 - j);
  const uint32_t umx = static_cast<uint32_t>(mx);
  const uint32_t plane = umx * static_cast<uint32_t>(my);
  const float* const u[3] = {ux, uy, uz};

  // Staging: the 67 x 7 tile points of one component are r = threadIdx.x and r + 256 (< 469);
  // their tile and plane offsets are the same for every component and plane.  Points outside
  // the grid are never read and not staged (loff < 0).
  int loff[kSlots];
  uint32_t goff[kSlots];
#pragma unroll
  for (int sl = 0; sl < kSlots; ++sl) {
    const int r = static_cast<int>(threadIdx.x) + sl * kCellThreads;
    const int hy = r / kHaloX, hx = r % kHaloX;
    const int gx = x0 - 1 + hx, gy = y0 - 1 + hy;
    const bool in = r < kHaloY * kHaloX && gx >= 0 && gx < mx && gy >= 0 && gy < my;
    loff[sl] = in ? hy * kPitch + hx : -1;
    goff[sl] = in ? static_cast<uint32_t>(gy) * umx + static_cast<uint32_t>(gx) : 0u;
  }
  float pre[3][kSlots];   // the next plane's tile points, loaded ahead
  auto fetch = [&](int z) {
    const uint32_t zoff = static_cast<uint32_t>(z) * plane;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int sl = 0; sl < kSlots; ++sl)
        pre[c][sl] = loff[sl] >= 0 ? u[c][zoff + goff[sl]] : 0.0f;
  };

  const float q = BASIS ? 0.25f * inv_h : 0.125f;
  float s1[3], s2[3], px1[3], px2[3], py1[3], py2[3], jz1[3];   // planes z - 1, z - 2
  auto emit = [&](int kc, const float (&gx)[3], const float (&gy)[3], const float (&gz)[3]) {
    if (!active) return;
    float G[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      G[c][0] = q * gx[c];
      G[c][1] = q * gy[c];
      G[c][2] = q * gz[c];
    }
    const uint32_t e = static_cast<uint32_t>(i) +
                       static_cast<uint32_t>(cx) * (static_cast<uint32_t>(j) +
                                                    static_cast<uint32_t>(cy) *
                                                        static_cast<uint32_t>(kc));
    if (vorticity) {
      const float wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
      vorticity[e] = sqrtf(wx * wx + wy * wy + wz * wz);
    }
    if (qcrit) {
      qcrit[e] = -0.5f * (G[0][0] * G[0][0] + G[1][1] * G[1][1] + G[2][2] * G[2][2]) -
                 (G[0][1] * G[1][0] + G[0][2] * G[2][0] + G[1][2] * G[2][1]);
    }
  };

  fetch(zlo);
  int buf = 0;
  for (int z = zlo; z <= zhi; ++z) {
    float* t = lds[buf];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int sl = 0; sl < kSlots; ++sl)
        if (loff[sl] >= 0) t[c * kHaloY * kPitch + loff[sl]] = pre[c][sl];
    __syncthreads();
    buf ^= 1;
    if (z < zhi) fetch(z + 1);
    float px[3], py[3], s[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) plane_sums<BASIS>(t, c, k, px[c], py[c], s[c]);
    if (BASIS) {
      if (z > k0) {   // cell z - 1 from planes z - 1 and z
        float gx[3], gy[3], gz[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          gx[c] = px1[c] + px[c];
          gy[c] = py1[c] + py[c];
          gz[c] = s[c] - s1[c];
        }
        emit(z - 1, gx, gy, gz);
      }
    } else {
      const int cp = z - 1;   // corner plane whose z-derivative plane z completes
      if (cp >= k0) {
        const float sz = inv_h / static_cast<float>(z - max(cp - 1, 0));
        float jz[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) jz[c] = (s[c] - (cp > 0 ? s2[c] : s1[c])) * sz;
        if (cp > k0) {   // cell cp - 1 from corner planes cp - 1 and cp
          float gx[3], gy[3], gz[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            gx[c] = px2[c] + px1[c];
            gy[c] = py2[c] + py1[c];
            gz[c] = jz1[c] + jz[c];
          }
          emit(cp - 1, gx, gy, gz);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) jz1[c] = jz[c];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      s2[c] = s1[c];
      s1[c] = s[c];
      px2[c] = px1[c];
      px1[c] = px[c];
      py2[c] = py1[c];
      py1[c] = py[c];
    }
  }
  if (!BASIS && zhi == k1) {
    // the grid's last plane is corner plane k1: one-sided z-derivative, then cell k1 - 1
    float gx[3], gy[3], gz[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      gx[c] = px2[c] + px1[c];
      gy[c] = py2[c] + py1[c];
      gz[c] = jz1[c] + (s1[c] - s2[c]) * inv_h;
    }
    emit(k1 - 1, gx, gy, gz);
  }
}

}  // namespace tlbm

// error reporting (tlbm_last_error) lives in capi.hip
namespace tlbm_detail {
__attribute__((visibility("hidden"))) int fail_message(const char* msg);
__attribute__((visibility("hidden"))) int fail_hip(hipError_t e, const char* what);
}  // namespace tlbm_detail
using tlbm_detail::fail_hip;
using tlbm_detail::fail_message;

extern "C" {

int tlbm_velocity_gradient(const float* ux, const float* uy, const float* uz, float* jac, int mx,
                           int my, int mz, float inv_h, void* stream) {
  if (!ux || !uy || !uz || !jac) return fail_message("velocity gradient: NULL pointer");
  if (mx <= 0 || my <= 0 || mz <= 0) return fail_message("grid sizes must be positive");
  const uint64_t m = static_cast<uint64_t>(mx) * my * mz;
  if (9 * m > 0xFFFFFFFFull) return fail_message("grid too large for the gradient");
  hipLaunchKernelGGL(tlbm::velocity_gradient_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)),
                     dim3(256), 0, static_cast<hipStream_t>(stream), ux, uy, uz, jac, mx, my,
                     mz, inv_h);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "gradient launch");
  return TLBM_OK;
}

int tlbm_vortex_fields(const float* ux, const float* uy, const float* uz, float* vorticity,
                       float* qcrit, int mx, int my, int mz, float inv_h, int basis, void* stream) {
  if (mx < 2 || my < 2 || mz < 2) return fail_message("cell fields need >= 2 points per axis");
  if (!ux || !uy || !uz) return fail_message("cell fields: NULL velocity array");
  if (!vorticity && !qcrit) return TLBM_OK;
  const uint64_t m = static_cast<uint64_t>(mx) * my * mz;
  if (m > 0xFFFFFFFFull) return fail_message("grid too large for the cell fields");
  const uint64_t xtiles = (mx - 1 + tlbm::kCellTileX - 1) / tlbm::kCellTileX;
  const uint64_t ytiles = (my - 1 + tlbm::kCellTileY - 1) / tlbm::kCellTileY;
  const uint64_t ztiles = (mz - 1 + tlbm::kCellChunkZ - 1) / tlbm::kCellChunkZ;
  const uint64_t tiles = xtiles * ytiles * ztiles;
  const uint64_t per_xcd = (tiles + 7) / 8;
  if (8 * per_xcd > 0xFFFFFFFFull) return fail_message("grid too large for the cell fields");
  const dim3 grid(static_cast<uint32_t>(8 * per_xcd)), block(tlbm::kCellTileX * tlbm::kCellTileY);
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t a = static_cast<uint32_t>(xtiles), c = static_cast<uint32_t>(ytiles),
                 t = static_cast<uint32_t>(tiles), p = static_cast<uint32_t>(per_xcd);
  if (basis)
    hipLaunchKernelGGL(tlbm::vortex_fields_kernel<true>, grid, block, 0, st, ux, uy, uz,
                       vorticity, qcrit, mx, my, mz, inv_h, a, c, t, p);
  else
    hipLaunchKernelGGL(tlbm::vortex_fields_kernel<false>, grid, block, 0, st, ux, uy, uz,
                       vorticity, qcrit, mx, my, mz, inv_h, a, c, t, p);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "cell fields launch");
  return TLBM_OK;
}

}  // extern "C"
