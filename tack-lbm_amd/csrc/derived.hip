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
// cell centre, with G the cell's velocity gradient: basis = 0, the mean of its eight corners'
// point gradients (the solver's gradient); basis = 1, the trilinear cell's gradient at its
// centre, the mean of the four edge differences along each axis.
__global__ void __launch_bounds__(256)
    vortex_fields_kernel(const float* __restrict__ ux, const float* __restrict__ uy,
                         const float* __restrict__ uz, const float* __restrict__ jac,
                         float* __restrict__ vorticity, float* __restrict__ qcrit, int mx, int my,
                         int mz, float inv_h, int basis) {
  const int cx = mx - 1, cy = my - 1, cz = mz - 1;
  const uint32_t nc = static_cast<uint32_t>(cx) * cy * cz;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  const int i = static_cast<int>(e % static_cast<uint32_t>(cx));
  const int j = static_cast<int>((e / static_cast<uint32_t>(cx)) % static_cast<uint32_t>(cy));
  const int k = static_cast<int>(e / (static_cast<uint32_t>(cx) * cy));
  const uint32_t m = static_cast<uint32_t>(mx) * my * mz;
  auto pt = [&](int di, int dj, int dk) {
    return static_cast<uint32_t>(i + di) +
           mx * (static_cast<uint32_t>(j + dj) + my * static_cast<uint32_t>(k + dk));
  };
  float G[3][3];
  if (basis) {
    const float* const u[3] = {ux, uy, uz};
    const float q = 0.25f * inv_h;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float* v = u[c];
      float gx = 0.0f, gy = 0.0f, gz = 0.0f;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          gx += v[pt(1, a, b)] - v[pt(0, a, b)];
          gy += v[pt(a, 1, b)] - v[pt(a, 0, b)];
          gz += v[pt(a, b, 1)] - v[pt(a, b, 0)];
        }
      G[c][0] = gx * q;
      G[c][1] = gy * q;
      G[c][2] = gz * q;
    }
  } else {
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const float* plane = jac + static_cast<uint64_t>(r) * m;
      float s = 0.0f;
#pragma unroll
      for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
          for (int di = 0; di < 2; ++di) s += plane[pt(di, dj, dk)];
      G[r / 3][r % 3] = 0.125f * s;
    }
  }
  if (vorticity) {
    const float wx = G[2][1] - G[1][2], wy = G[0][2] - G[2][0], wz = G[1][0] - G[0][1];
    vorticity[e] = sqrtf(wx * wx + wy * wy + wz * wz);
  }
  if (qcrit) {
    qcrit[e] = -0.5f * (G[0][0] * G[0][0] + G[1][1] * G[1][1] + G[2][2] * G[2][2]) -
               (G[0][1] * G[1][0] + G[0][2] * G[2][0] + G[1][2] * G[2][1]);
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

int tlbm_vortex_fields(const float* ux, const float* uy, const float* uz, const float* jac,
                       float* vorticity, float* qcrit, int mx, int my, int mz, float inv_h,
                       int basis, void* stream) {
  if (mx < 2 || my < 2 || mz < 2) return fail_message("cell fields need >= 2 points per axis");
  if (basis ? (!ux || !uy || !uz) : !jac)
    return fail_message(basis ? "basis gradient needs the velocity arrays"
                                    : "lattice gradient needs the point gradient (jac)");
  if (!vorticity && !qcrit) return TLBM_OK;
  const uint64_t m = static_cast<uint64_t>(mx) * my * mz;
  if (9 * m > 0xFFFFFFFFull) return fail_message("grid too large for the cell fields");
  const uint64_t nc = static_cast<uint64_t>(mx - 1) * (my - 1) * (mz - 1);
  hipLaunchKernelGGL(tlbm::vortex_fields_kernel, dim3(static_cast<uint32_t>((nc + 255) / 256)),
                     dim3(256), 0, static_cast<hipStream_t>(stream), ux, uy, uz, jac, vorticity,
                     qcrit, mx, my, mz, inv_h, basis);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "cell fields launch");
  return TLBM_OK;
}

}  // extern "C"
