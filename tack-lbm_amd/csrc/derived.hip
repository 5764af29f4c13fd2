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

// |curl u| and Q = -1/2 (G_00^2 + G_11^2 + G_22^2) - (G_01 G_10 + G_02 G_20 + G_12 G_21) at each
// cell centre, with G the cell's velocity gradient: BASIS = false, the mean of its eight
// corners' point gradients (the solver's gradient, the points' as velocity_gradient_kernel forms
// them); BASIS = true, the trilinear cell's gradient at its centre, the mean of the four edge
// differences along each axis.  Both read the velocity only: no point gradient goes through HBM.
//
// Mapping.  A block is 64 x 4 cell columns (one wave per y row) and a chunk of kz cells along z;
// each thread marches up its column plane by plane.  Per point plane z it forms, for each
// component, three sums over its 2 x 2 corner points (x in {i, i+1}, y in {j, j+1}):
//   px(z): the x-derivatives (lattice) or the x edge differences (basis),
//   py(z): the same along y,
//   s(z):  the four values, whose differences across planes give the z-derivatives;
// a cell k then takes planes k and k + 1 (lattice z: s at planes k - 1 .. k + 2, one-sided at
// the grid's faces).  The x and y neighbours come from the cache lines the wave's own loads and
// the neighbouring rows' loads bring in; blocks go to the XCDs in contiguous runs of tiles so
// rows that share lines share an L2.
constexpr int kCellTileX = 64;
constexpr int kCellTileY = 4;
constexpr int kCellChunkZ = 32;

struct CellColumn {
  // offsets within a plane of the corner rows' points and of the y neighbour rows
  uint32_t rj, rj1, ryl, ryh;
  uint32_t i, xl, xh1;   // x of the corner column i, its left neighbour, i + 1's right neighbour
  float sx0, sx1, sy0, sy1;
};

template <bool BASIS>
__device__ __forceinline__ void plane_sums(const float* __restrict__ v, const CellColumn& t,
                                           float& px, float& py, float& s) {
  const float b0 = v[t.rj + t.i], c0 = v[t.rj + t.i + 1];
  const float b1 = v[t.rj1 + t.i], c1 = v[t.rj1 + t.i + 1];
  s = (b0 + c0) + (b1 + c1);
  if (BASIS) {
    px = (c0 - b0) + (c1 - b1);
    py = (b1 - b0) + (c1 - c0);
  } else {
    const float a0 = v[t.rj + t.xl], d0 = v[t.rj + t.xh1];
    const float a1 = v[t.rj1 + t.xl], d1 = v[t.rj1 + t.xh1];
    const float e0 = v[t.ryl + t.i], e1 = v[t.ryl + t.i + 1];
    const float h0 = v[t.ryh + t.i], h1 = v[t.ryh + t.i + 1];
    // d/dx at corners (i, y) and (i + 1, y), y = j, j + 1; d/dy at corners (x, j), (x, j + 1)
    px = ((c0 - a0) + (c1 - a1)) * t.sx0 + ((d0 - b0) + (d1 - b1)) * t.sx1;
    py = ((b1 - e0) + (c1 - e1)) * t.sy0 + ((h0 - b0) + (h1 - c0)) * t.sy1;
  }
}

__device__ __forceinline__ float corner_sum(const float* __restrict__ v, const CellColumn& t) {
  return (v[t.rj + t.i] + v[t.rj + t.i + 1]) + (v[t.rj1 + t.i] + v[t.rj1 + t.i + 1]);
}

template <bool BASIS>
__global__ void __launch_bounds__(kCellTileX * kCellTileY)
    vortex_fields_kernel(const float* __restrict__ ux, const float* __restrict__ uy,
                         const float* __restrict__ uz, float* __restrict__ vorticity,
                         float* __restrict__ qcrit, int mx, int my, int mz, float inv_h,
                         uint32_t xtiles, uint32_t ytiles, uint32_t tiles, uint32_t per_xcd) {
  const uint32_t b = blockIdx.x;
  const uint32_t tile = (b % 8u) * per_xcd + b / 8u;   // XCD b % 8 takes a contiguous run
  if (tile >= tiles) return;
  const int cx = mx - 1, cy = my - 1, cz = mz - 1;
  const uint32_t tx = tile % xtiles, ty = (tile / xtiles) % ytiles, tz = tile / (xtiles * ytiles);
  const int i = static_cast<int>(tx) * kCellTileX + static_cast<int>(threadIdx.x % kCellTileX);
  const int j = static_cast<int>(ty) * kCellTileY + static_cast<int>(threadIdx.x / kCellTileX);
  if (i >= cx || j >= cy) return;
  const int k0 = static_cast<int>(tz) * kCellChunkZ;
  const int k1 = min(k0 + kCellChunkZ, cz);           // cells k0 .. k1 - 1, corner planes k0 .. k1
  CellColumn t;
  const uint32_t umx = static_cast<uint32_t>(mx);
  const int yl = j > 0 ? j - 1 : 0, yh1 = j + 2 < my ? j + 2 : j + 1;
  t.i = static_cast<uint32_t>(i);
  t.xl = static_cast<uint32_t>(i > 0 ? i - 1 : 0);
  t.xh1 = static_cast<uint32_t>(i + 2 < mx ? i + 2 : i + 1);
  t.rj = static_cast<uint32_t>(j) * umx;
  t.rj1 = t.rj + umx;
  t.ryl = static_cast<uint32_t>(yl) * umx;
  t.ryh = static_cast<uint32_t>(yh1) * umx;
  t.sx0 = inv_h / static_cast<float>(i + 1 - static_cast<int>(t.xl));
  t.sx1 = inv_h / static_cast<float>(static_cast<int>(t.xh1) - i);
  t.sy0 = inv_h / static_cast<float>(j + 1 - yl);
  t.sy1 = inv_h / static_cast<float>(yh1 - j);
  const uint32_t plane = umx * static_cast<uint32_t>(my);
  const float* const u[3] = {ux, uy, uz};
  const float q = BASIS ? 0.25f * inv_h : 0.125f;

  float px[3], py[3], s[3];        // plane kk
  float ppx[3], ppy[3], pz[3];     // plane kk - 1 (and its z-derivative sum, lattice)
  float sm[3];                     // s at plane max(kk - 1, 0) (lattice)
#pragma unroll
  for (int c = 0; c < 3; ++c) plane_sums<BASIS>(u[c] + k0 * plane, t, px[c], py[c], s[c]);
  if (!BASIS) {
#pragma unroll
    for (int c = 0; c < 3; ++c) sm[c] = k0 > 0 ? corner_sum(u[c] + (k0 - 1) * plane, t) : s[c];
  }
  uint32_t e = static_cast<uint32_t>(i) +
               static_cast<uint32_t>(cx) * (static_cast<uint32_t>(j) + static_cast<uint32_t>(cy) *
                                                                        static_cast<uint32_t>(k0));
  const uint32_t cplane = static_cast<uint32_t>(cx) * static_cast<uint32_t>(cy);
  for (int kk = k0; kk <= k1; ++kk) {
    float nx[3], ny[3], ns[3], zs[3];
    if (kk < k1) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
        plane_sums<BASIS>(u[c] + (kk + 1) * plane, t, nx[c], ny[c], ns[c]);
    } else if (!BASIS) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        nx[c] = ny[c] = 0.0f;
        ns[c] = kk + 1 < mz ? corner_sum(u[c] + (kk + 1) * plane, t) : s[c];
      }
    }
    if (!BASIS) {
      const float sz = inv_h / static_cast<float>(min(kk + 1, mz - 1) - max(kk - 1, 0));
#pragma unroll
      for (int c = 0; c < 3; ++c) zs[c] = (ns[c] - sm[c]) * sz;
    }
    if (kk > k0) {
      float G[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        G[c][0] = q * (ppx[c] + px[c]);
        G[c][1] = q * (ppy[c] + py[c]);
        G[c][2] = BASIS ? q * (s[c] - pz[c]) : q * (pz[c] + zs[c]);
      }
      if (vorticity) {
        const float wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
        vorticity[e] = sqrtf(wx * wx + wy * wy + wz * wz);
      }
      if (qcrit) {
        qcrit[e] = -0.5f * (G[0][0] * G[0][0] + G[1][1] * G[1][1] + G[2][2] * G[2][2]) -
                   (G[0][1] * G[1][0] + G[0][2] * G[2][0] + G[1][2] * G[2][1]);
      }
      e += cplane;
    }
    if (kk == k1) break;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ppx[c] = px[c];
      ppy[c] = py[c];
      pz[c] = BASIS ? s[c] : zs[c];
      if (!BASIS) sm[c] = s[c];
      px[c] = nx[c];
      py[c] = ny[c];
      s[c] = ns[c];
    }
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
