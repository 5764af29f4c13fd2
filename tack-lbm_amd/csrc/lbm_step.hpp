// Device code of the fused D3Q19 step for gfx950.
//
// One thread per lattice cell, blocks laid along x so that each population plane is read and
// written in coalesced wave-wide runs; y and z come from blockIdx, so every y/z boundary
// decision (periodic wrap, slip mirror, lid) is uniform across the wave.  The kernel is a
// template over collision x obstacle x side x report x storage: each combination the
// reference's Tack compiler would specialise (lbm.py:1-7) is its own instantiation, with no
// runtime branch on the configuration.
//
// Arithmetic follows lbm.py:397-513 with these re-arrangements, all exact or within FP32
// rounding of the reference's expression:
//   * walls: the solid test of a neighbour is skipped for cells whose 3x3x3 neighbourhood
//     cannot reach the obstacle (an exact bound, see Obstacle::near); the bounced population
//     (lbm.py:427) is loaded only on wall links, and the moving-wall term only there;
//   * moments: summed in the reference's q order, skipping the c = 0 terms (adding +0 is exact);
//   * equilibria: computed once per pair (q, opp q), feq = A +- B with A = w rho (1 - 1.5 u.u
//     + cu^2/2), B = w rho cu;
//   * collision: post - feq written directly (BGK: (1-omega) neq; TRT: per pair
//     (1-omega+) n+ +- (1-omega-) n-), then the sponge blend of lbm.py:446-456.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "tlbm.h"

namespace tlbm {

constexpr int kQ = 19;
// D3Q19 ordering of lbm.py:15-18: rest, 6 faces, 12 edges; pairs (2k+1, 2k+2) are opposite.
constexpr int kCX[kQ] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int kCY[kQ] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int kCZ[kQ] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
// FP32 weights, as the reference's device table (lbm.py:18, 48).
constexpr float kW[kQ] = {1.0f / 3, 1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18, 1.0f / 18,
                          1.0f / 18, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36,
                          1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36, 1.0f / 36,
                          1.0f / 36};

// lbm.py:28-30 and 33-35, evaluated at compile time.
constexpr int opp(int q) {
  for (int p = 0; p < kQ; ++p)
    if (kCX[p] == -kCX[q] && kCY[p] == -kCY[q] && kCZ[p] == -kCZ[q]) return p;
  return -1;
}
constexpr int miry(int q) {
  for (int p = 0; p < kQ; ++p)
    if (kCX[p] == kCX[q] && kCY[p] == -kCY[q] && kCZ[p] == kCZ[q]) return p;
  return -1;
}
static_assert(opp(1) == 2 && opp(8) == 7 && opp(18) == 17, "pair layout");
static_assert(miry(3) == 4 && miry(7) == 9 && miry(15) == 18, "y mirror");

template <int I, int N, class F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// ---- population storage -------------------------------------------------------------------
struct StoreF32 {
  using T = float;
  static constexpr int kind = TLBM_F32;
  __device__ static __forceinline__ float load(const T* p, uint32_t i, float) { return p[i]; }
  __device__ static __forceinline__ void store(T* p, uint32_t i, float v, float) { p[i] = v; }
  __device__ static __forceinline__ T encode(float v, float) { return v; }
};

// Deviation from the rest equilibrium w_q, binary16, round to nearest even.
struct StoreF16 {
  using T = _Float16;
  static constexpr int kind = TLBM_F16;
  __device__ static __forceinline__ float load(const T* p, uint32_t i, float w) {
    return w + static_cast<float>(p[i]);
  }
  __device__ static __forceinline__ T encode(float v, float w) { return static_cast<T>(v - w); }
  __device__ static __forceinline__ void store(T* p, uint32_t i, float v, float w) {
    p[i] = encode(v, w);
  }
};

// Deviation from w_q, bfloat16 (upper half of the binary32 pattern), round to nearest even.
struct StoreBF16 {
  using T = uint16_t;
  static constexpr int kind = TLBM_BF16;
  __device__ static __forceinline__ float load(const T* p, uint32_t i, float w) {
    return w + __uint_as_float(static_cast<uint32_t>(p[i]) << 16);
  }
  __device__ static __forceinline__ T encode(float v, float w) {
    const uint32_t u = __float_as_uint(v - w);
    return static_cast<T>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  }
  __device__ static __forceinline__ void store(T* p, uint32_t i, float v, float w) {
    p[i] = encode(v, w);
  }
};

// ---- kernel arguments (host fills them from tlbm_params) ----------------------------------
struct StepArgs {
  const void* fin;
  void* fout;
  uint64_t ns;               // plane stride (elements)
  int32_t nx, ny, nz, pad;
  float omega;               // BGK
  float tau0, smag, magic;   // smag = 25.456 cs^2 (lbm.py:93)
  float o[6];                // obstacle parameters
  float r2;                  // r*r (cylinder, sphere)
  float reach[3];            // half extents + 2 (cells) for the neighbourhood bound
  float u0, s0, s_len, s_k;  // sponge: start, nx - start, strength
  float lid[3];
  const float* omega_rot;
  tlbm_fields fld;
  double* partials;          // report: per-block FP64 partial sums [TLBM_NSUMS][nblocks]
  uint32_t nblocks;
};

// ---- obstacles: lbm.py:144-227 ---------------------------------------------------------------
// solid() uses one rounding per subtract, multiply and add (no contraction), as a float32
// evaluation of the reference's expression.  near() is a conservative bound: when it is false
// for a cell that is not on a periodic wrap, none of its 19 pull sources is solid.
struct Cylinder {
  static constexpr bool kMoving = true;
  __device__ static __forceinline__ float solid(const StepArgs& a, int x, int y, int) {
#pragma clang fp contract(off)
    const float dx = static_cast<float>(x) - a.o[0];
    const float dy = static_cast<float>(y) - a.o[1];
    return (dx * dx + dy * dy < a.r2) ? 1.0f : 0.0f;
  }
  __device__ static __forceinline__ bool near(const StepArgs& a, int x, int y, int) {
    return fabsf(static_cast<float>(x) - a.o[0]) < a.reach[0] &&
           fabsf(static_cast<float>(y) - a.o[1]) < a.reach[1];
  }
  // unit tangent (lbm.py:159-169)
  __device__ static __forceinline__ void wall_vel(const StepArgs& a, float xm, float ym, float,
                                                  float& u, float& v, float& w) {
    u = -(ym - a.o[1]) / a.o[2];
    v = (xm - a.o[0]) / a.o[2];
    w = 0.0f;
  }
};

struct Sphere {
  static constexpr bool kMoving = false;
  __device__ static __forceinline__ float solid(const StepArgs& a, int x, int y, int z) {
#pragma clang fp contract(off)
    const float dx = static_cast<float>(x) - a.o[0];
    const float dy = static_cast<float>(y) - a.o[1];
    const float dz = static_cast<float>(z) - a.o[2];
    return (dx * dx + dy * dy + dz * dz < a.r2) ? 1.0f : 0.0f;
  }
  __device__ static __forceinline__ bool near(const StepArgs& a, int x, int y, int z) {
    return fabsf(static_cast<float>(x) - a.o[0]) < a.reach[0] &&
           fabsf(static_cast<float>(y) - a.o[1]) < a.reach[1] &&
           fabsf(static_cast<float>(z) - a.o[2]) < a.reach[2];
  }
  __device__ static __forceinline__ void wall_vel(const StepArgs&, float, float, float, float& u,
                                                  float& v, float& w) {
    u = v = w = 0.0f;
  }
};

struct Box {
  static constexpr bool kMoving = false;
  __device__ static __forceinline__ float solid(const StepArgs& a, int x, int y, int z) {
    const float d = fmaxf(fmaxf(fabsf(static_cast<float>(x) - a.o[0]) - a.o[3],
                                fabsf(static_cast<float>(y) - a.o[1]) - a.o[4]),
                          fabsf(static_cast<float>(z) - a.o[2]) - a.o[5]);
    return (d < 0.0f) ? 1.0f : 0.0f;
  }
  __device__ static __forceinline__ bool near(const StepArgs& a, int x, int y, int z) {
    return fabsf(static_cast<float>(x) - a.o[0]) < a.reach[0] &&
           fabsf(static_cast<float>(y) - a.o[1]) < a.reach[1] &&
           fabsf(static_cast<float>(z) - a.o[2]) < a.reach[2];
  }
  __device__ static __forceinline__ void wall_vel(const StepArgs&, float, float, float, float& u,
                                                  float& v, float& w) {
    u = v = w = 0.0f;
  }
};

template <int OBS>
using ObstacleT = std::conditional_t<OBS == TLBM_CYLINDER, Cylinder,
                                     std::conditional_t<OBS == TLBM_SPHERE, Sphere, Box>>;

template <int STORAGE>
using StorageT = std::conditional_t<STORAGE == TLBM_F32, StoreF32,
                                    std::conditional_t<STORAGE == TLBM_F16, StoreF16, StoreBF16>>;

// ---- equilibria ------------------------------------------------------------------------------
// feq_q = w rho (1 + cu + cu^2/2 - 1.5 u.u), cu = 3 c.u  (lbm.py:52-55), for all q at once.
__device__ __forceinline__ void equilibria(float rho, float ux, float uy, float uz, float (&e)[kQ]) {
  const float base = 1.0f - 1.5f * (ux * ux + uy * uy + uz * uz);
  e[0] = kW[0] * rho * base;
  static_for<0, 9>([&](auto pc) {
    constexpr int q = 2 * decltype(pc)::value + 1;
    constexpr int o = q + 1;
    float cu = 0.0f;
    if constexpr (kCX[q] == 1) cu += ux;
    if constexpr (kCX[q] == -1) cu -= ux;
    if constexpr (kCY[q] == 1) cu += uy;
    if constexpr (kCY[q] == -1) cu -= uy;
    if constexpr (kCZ[q] == 1) cu += uz;
    if constexpr (kCZ[q] == -1) cu -= uz;
    cu *= 3.0f;
    const float wr = kW[q] * rho;
    const float A = wr * (base + 0.5f * cu * cu);
    const float B = wr * cu;
    e[q] = A + B;
    e[o] = A - B;
  });
}

// tau with the Smagorinsky eddy viscosity, lbm.py:72-93.  s[p] = neq_q + neq_opp(q) of pair p
// (q = 2p + 1); the stress components only need these symmetric parts.
__device__ __forceinline__ float smagorinsky_tau(const float (&neq)[kQ], float rho, float tau0,
                                                 float smag) {
  float pxx = 0.0f, pyy = 0.0f, pzz = 0.0f, pxy = 0.0f, pxz = 0.0f, pyz = 0.0f;
  static_for<0, 9>([&](auto pc) {
    constexpr int q = 2 * decltype(pc)::value + 1;
    const float s = neq[q] + neq[q + 1];
    constexpr int ex = kCX[q], ey = kCY[q], ez = kCZ[q];
    if constexpr (ex != 0) pxx += s;
    if constexpr (ey != 0) pyy += s;
    if constexpr (ez != 0) pzz += s;
    if constexpr (ex * ey == 1) pxy += s;
    if constexpr (ex * ey == -1) pxy -= s;
    if constexpr (ex * ez == 1) pxz += s;
    if constexpr (ex * ez == -1) pxz -= s;
    if constexpr (ey * ez == 1) pyz += s;
    if constexpr (ey * ez == -1) pyz -= s;
  });
  const float pi = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz +
                         2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
  return 0.5f * (tau0 + sqrtf(tau0 * tau0 + smag * pi / rho));
}

// ---- FP64 block reduction for the report variant -------------------------------------------
template <int BLOCK>
__device__ __forceinline__ void block_sums(double (&v)[TLBM_NSUMS], double* partials,
                                           uint32_t nblocks, uint32_t bid) {
  __shared__ double red[BLOCK / 64][TLBM_NSUMS];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < TLBM_NSUMS; ++k) {
    double s = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) red[wave][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < TLBM_NSUMS) {
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) s += red[w][threadIdx.x];
    partials[static_cast<size_t>(threadIdx.x) * nblocks + bid] = s;
  }
}

// ---- the fused step --------------------------------------------------------------------------
template <int COLL, int OBS, int SIDE, bool REPORT, int STORAGE>
__global__ void __launch_bounds__(256) lbm_step_kernel(const StepArgs a) {
  using Obs = ObstacleT<OBS>;
  using St = StorageT<STORAGE>;
  using T = typename St::T;
  const T* __restrict__ fin = static_cast<const T*>(a.fin);
  T* __restrict__ fout = static_cast<T*>(a.fout);

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int tx = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int z = blockIdx.z;
  const bool active = tx < nx;
  if constexpr (!REPORT) {
    if (!active) return;
  }
  const int x = active ? tx : nx - 1;
  const uint32_t snx = static_cast<uint32_t>(nx);
  const uint32_t plane = snx * static_cast<uint32_t>(ny);
  const uint32_t idx = (static_cast<uint32_t>(z) * ny + y) * snx + x;

  // pull sources along each axis (x, z periodic: lbm.py:412, 415; y by side: lbm.py:342-369)
  const int xsrc[3] = {x == nx - 1 ? 0 : x + 1, x, x == 0 ? nx - 1 : x - 1};  // cx = -1, 0, +1
  const int zsrc[3] = {z == nz - 1 ? 0 : z + 1, z, z == 0 ? nz - 1 : z - 1};
  const uint32_t zrow[3] = {static_cast<uint32_t>(zsrc[0]) * plane,
                            static_cast<uint32_t>(zsrc[1]) * plane,
                            static_cast<uint32_t>(zsrc[2]) * plane};
  int ysrc[3];
  bool mir_lo = false, mir_hi = false, lid_hi = false;
  if constexpr (SIDE == TLBM_PERIODIC_Y) {
    ysrc[0] = y == ny - 1 ? 0 : y + 1;
    ysrc[2] = y == 0 ? ny - 1 : y - 1;
  } else {
    // specular reflection off the floor (and the ceiling for SlipY): the source stays in this
    // row and the population is the y-mirrored one.  The lid replaces ceiling links instead.
    ysrc[0] = y == ny - 1 ? y : y + 1;
    ysrc[2] = y == 0 ? y : y - 1;
    mir_lo = (y == 0);
    if constexpr (SIDE == TLBM_SLIP_Y) mir_hi = (y == ny - 1);
    else lid_hi = (y == ny - 1);
  }
  ysrc[1] = y;

  // ---- pull streaming (lbm.py:410-426) ----
  float f[kQ];
  static_for<0, kQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
    if constexpr (SIDE == TLBM_LID_Y && cy == -1) {
      if (lid_hi) {
        f[q] = 0.0f;  // replaced by the lid bounce-back below
        return;
      }
    }
    const uint32_t src = zrow[1 - cz] + static_cast<uint32_t>(ysrc[1 - cy]) * snx +
                         static_cast<uint32_t>(xsrc[1 - cx]);
    int qs = q;
    if constexpr (SIDE != TLBM_PERIODIC_Y && cy == 1) qs = mir_lo ? miry(q) : q;
    if constexpr (SIDE == TLBM_SLIP_Y && cy == -1) qs = mir_hi ? miry(q) : q;
    f[q] = St::load(fin + static_cast<uint64_t>(qs) * a.ns, src, kW[q]);
  });

  // ---- moving lid: f_q = f_opp(x) + 6 w_q c_q . u_lid on ceiling links ----
  if constexpr (SIDE == TLBM_LID_Y) {
    if (lid_hi) {
      static_for<0, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (kCY[q] == -1) {
          constexpr int o = opp(q);
          const float cdu = kCX[q] * a.lid[0] + kCY[q] * a.lid[1] + kCZ[q] * a.lid[2];
          f[q] = St::load(fin + static_cast<uint64_t>(o) * a.ns, idx, kW[o]) +
                 6.0f * kW[q] * cdu;
        }
      });
    }
  }

  // ---- obstacle: halfway / moving-wall bounce-back and momentum exchange (lbm.py:416-432) ----
  float here = 0.0f;
  float mx = 0.0f, my = 0.0f, mz = 0.0f;
  bool wrap = (x == 0) || (x == nx - 1) || (z == 0) || (z == nz - 1);
  if constexpr (SIDE == TLBM_PERIODIC_Y) wrap = wrap || (y == 0) || (y == ny - 1);
  if (wrap || Obs::near(a, x, y, z)) {
    here = Obs::solid(a, x, y, z);
    const float fluid = 1.0f - here;
    float omega_rot = 0.0f;
    if constexpr (Obs::kMoving) omega_rot = a.omega_rot ? *a.omega_rot : 0.0f;
    static_for<1, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
      constexpr int o = opp(q);
      if constexpr (SIDE == TLBM_LID_Y && cy == -1) {
        if (lid_hi) return;
      }
      const float wall = Obs::solid(a, xsrc[1 - cx], ysrc[1 - cy], zsrc[1 - cz]);
      if (wall != 0.0f) {
        const float b = St::load(fin + static_cast<uint64_t>(o) * a.ns, idx, kW[o]);
        float t = 0.0f;
        if constexpr (Obs::kMoving) {
          // wall velocity at the link midpoint (lbm.py:418-425)
          float uwx, uwy, uwz;
          Obs::wall_vel(a, static_cast<float>(x) - 0.5f * cx, static_cast<float>(y) - 0.5f * cy,
                        static_cast<float>(z) - 0.5f * cz, uwx, uwy, uwz);
          t = 6.0f * kW[q] * (cx * (omega_rot * uwx) + cy * (omega_rot * uwy) +
                              cz * (omega_rot * uwz));
        }
        f[q] = b + t;
        const float hit = fluid * (2.0f * b + t);
        mx += hit * static_cast<float>(kCX[o]);
        my += hit * static_cast<float>(kCY[o]);
        mz += hit * static_cast<float>(kCZ[o]);
      }
    });
  }
  const float fluid = 1.0f - here;

  // ---- moments (lbm.py:433-444) ----
  float rho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
  static_for<0, kQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    rho += f[q];
    if constexpr (kCX[q] == 1) jx += f[q];
    if constexpr (kCX[q] == -1) jx -= f[q];
    if constexpr (kCY[q] == 1) jy += f[q];
    if constexpr (kCY[q] == -1) jy -= f[q];
    if constexpr (kCZ[q] == 1) jz += f[q];
    if constexpr (kCZ[q] == -1) jz -= f[q];
  });
  const float ux = fluid * jx / rho;
  const float uy = fluid * jy / rho;
  const float uz = fluid * jz / rho;

  // ---- collision: d[q] = post_q - feq_q (lbm.py:60-139) ----
  float e[kQ];
  equilibria(rho, ux, uy, uz, e);
  float d[kQ];
  static_for<0, kQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    d[q] = f[q] - e[q];
  });
  if constexpr (COLL == TLBM_BGK) {
    const float k = 1.0f - a.omega;
    static_for<0, kQ>([&](auto qc) { d[decltype(qc)::value] *= k; });
  } else if constexpr (COLL == TLBM_LES) {
    const float tau = smagorinsky_tau(d, rho, a.tau0, a.smag);
    const float k = 1.0f - 1.0f / tau;
    static_for<0, kQ>([&](auto qc) { d[decltype(qc)::value] *= k; });
  } else {
    const float tau = smagorinsky_tau(d, rho, a.tau0, a.smag);
    const float kp = 1.0f - 1.0f / tau;
    const float km = 1.0f - 1.0f / (a.magic / (tau - 0.5f) + 0.5f);
    d[0] *= kp;
    static_for<0, 9>([&](auto pc) {
      constexpr int q = 2 * decltype(pc)::value + 1;
      const float np = 0.5f * (d[q] + d[q + 1]);
      const float nm = 0.5f * (d[q] - d[q + 1]);
      d[q] = kp * np + km * nm;
      d[q + 1] = kp * np - km * nm;
    });
  }

  // ---- sponge and store (lbm.py:446-456) ----
  const float t = fmaxf(0.0f, (static_cast<float>(x) - a.s0) / a.s_len);
  const float sigma = fmaxf(here, a.s_k * t * t);
  const float rho_s = rho + sigma * (1.0f - rho);
  const float ux_s = ux + sigma * (a.u0 * fluid - ux);
  const float uy_s = (1.0f - sigma) * uy;
  const float uz_s = (1.0f - sigma) * uz;
  float es[kQ];
  equilibria(rho_s, ux_s, uy_s, uz_s, es);
  const float keep = 1.0f - sigma;
  if (active) {
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      St::store(fout + static_cast<uint64_t>(q) * a.ns, idx, es[q] + keep * d[q], kW[q]);
    });
  }

  // ---- report fields and FP64 sums (lbm.py:493-513, 649-661) ----
  if constexpr (REPORT) {
    const tlbm_fields& F = a.fld;
    const float p = (rho - 1.0f) / 3.0f;
    const float usq = ux * ux + uy * uy + uz * uz;
    const float ke = 0.5f * rho * usq;
    double v[TLBM_NSUMS];
    if (active) {
      const float ox = F.ux[idx], oy = F.uy[idx], oz = F.uz[idx], op = F.p[idx];
      const float ok = 0.5f * (ox * ox + oy * oy + oz * oz);
      const float r_rho = (3.0f * (p - op)) * (3.0f * (p - op));
      const float r_ux = (ux - ox) * (ux - ox);
      const float r_uy = (uy - oy) * (uy - oy);
      const float r_uz = (uz - oz) * (uz - oz);
      const float r_ke = (ke - ok) * (ke - ok);
      const float vmag = sqrtf(usq);
      F.r_rho[idx] = r_rho;
      F.r_ux[idx] = r_ux;
      F.r_uy[idx] = r_uy;
      F.r_uz[idx] = r_uz;
      F.r_ke[idx] = r_ke;
      F.ux[idx] = ux;
      F.uy[idx] = uy;
      F.uz[idx] = uz;
      F.p[idx] = p;
      F.vmag[idx] = vmag;
      F.fx[idx] = mx;
      F.fy[idx] = my;
      F.fz[idx] = mz;
      const float vals[TLBM_NSUMS] = {ux,   uy,   uz,   p,  vmag, r_rho, r_ux,
                                      r_uy, r_uz, r_ke, mx, my,   mz,    rho};
#pragma unroll
      for (int k = 0; k < TLBM_NSUMS; ++k) v[k] = static_cast<double>(vals[k]);
    } else {
#pragma unroll
      for (int k = 0; k < TLBM_NSUMS; ++k) v[k] = 0.0;
    }
    if (a.partials) {
      const uint32_t bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
      block_sums<256>(v, a.partials, a.nblocks, bid);
    }
  }
}

// ---- lbm_init (lbm.py:383-394) with the seeded velocity noise ---------------------------------
__device__ __forceinline__ float hash_uniform(uint64_t seed, uint64_t idx, uint32_t comp) {
  uint64_t s = seed * 0x9E3779B97F4A7C15ull;
  uint64_t k = idx * 3ull + comp + 1ull;
  uint64_t z = s + k * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return static_cast<float>(static_cast<uint32_t>(z >> 40)) * (1.0f / 16777216.0f);
}

template <int OBS, int STORAGE>
__global__ void __launch_bounds__(256) lbm_init_kernel(const StepArgs a, float seed_amp,
                                                       float noise_amp, uint64_t noise_seed) {
  using Obs = ObstacleT<OBS>;
  using St = StorageT<STORAGE>;
  using T = typename St::T;
  T* __restrict__ f = static_cast<T*>(a.fout);
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= a.nx) return;
  const uint32_t idx = (static_cast<uint32_t>(z) * a.ny + y) * static_cast<uint32_t>(a.nx) + x;
  const float fluid = 1.0f - Obs::solid(a, x, y, z);
  const float two_pi = 6.2831853f;
  float ux = a.u0 * fluid;
  float uy = seed_amp * a.u0 * fluid * sinf(two_pi * static_cast<float>(x) / static_cast<float>(a.nx)) *
             cosf(two_pi * static_cast<float>(z) / static_cast<float>(a.nz));
  float uz = 0.0f;
  if (noise_amp != 0.0f) {
    ux += fluid * (noise_amp * (2.0f * hash_uniform(noise_seed, idx, 0) - 1.0f));
    uy += fluid * (noise_amp * (2.0f * hash_uniform(noise_seed, idx, 1) - 1.0f));
    uz += fluid * (noise_amp * (2.0f * hash_uniform(noise_seed, idx, 2) - 1.0f));
  }
  float e[kQ];
  equilibria(1.0f, ux, uy, uz, e);
  static_for<0, kQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    St::store(f + static_cast<uint64_t>(q) * a.ns, idx, e[q], kW[q]);
  });
}

}  // namespace tlbm
