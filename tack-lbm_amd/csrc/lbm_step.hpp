// Device code of the fused D3Q19 step for gfx950.
//
// Mapping.  One thread per cell; each block steps consecutive cells in index order (see
// lbm_step_kernel), so each population plane is read and written in coalesced wave-wide runs
// of whole aligned cache lines; the +-1 x offsets of the pull scheme are plain misaligned dword
// loads, served by the same cache lines.  Measured on MI355X, one cell per thread streams as
// fast as 8/16-byte per-lane vector loads with lane shuffles or two cells per thread, and runs
// the 16-bit storage formats faster (fewer instructions, more waves per SIMD).  What matters most is that the 19 pull
// loads of a thread are issued back to back, with no branch between them.
//
// Specialisation.  The kernel is a template over collision x obstacle x side x report x storage:
// each combination the reference's Tack compiler would specialise (lbm.py:1-7) is its
// own instantiation, with no runtime branch on the configuration.
//
// Arithmetic follows lbm.py:397-513, rearranged only in ways that are exact or within FP32
// rounding of the reference's expressions:
//   * walls: neighbour solid tests run only for cells whose 3x3x3 neighbourhood can reach the
//     obstacle (an exact bound, Obstacle::near, plus the periodic wraps when the obstacle comes
//     within reach of a domain face); the bounced population (lbm.py:427) is loaded only on wall
//     links and the moving-wall term is formed only there;
//   * moments: summed in the reference's q order, skipping the c = 0 terms (adding +0 is exact);
//   * equilibria: once per pair (q, opp q): feq = A +- B, A = w rho (1 - 1.5 u.u + 4.5 (c.u)^2),
//     B = 3 w rho c.u;
//   * collision: post - feq written directly (BGK/LES: (1 - omega) neq; TRT: per pair
//     (1 - omega+) n+ +- (1 - omega-) n-), then the sponge blend of lbm.py:446-456; where
//     sigma = 0 the blend target feq(rho_s, u_s) equals feq(rho, u) bit for bit and is reused;
//   * FP16/BF16 storage keeps deviations d_q = f_q - w_q and all arithmetic runs on them
//     (rho = 1 + sum d, feq - w formed directly); bounce-back and the lid term are the same
//     expressions in deviations because w_opp(q) = w_q.
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

// D3Q7 of the temperature (lbm.py:20-25): rest plus the six faces, c_s^2 = 1/4
constexpr int kQ7 = 7;
constexpr int kCX7[kQ7] = {0, 1, -1, 0, 0, 0, 0};
constexpr int kCY7[kQ7] = {0, 0, 0, 1, -1, 0, 0};
constexpr int kCZ7[kQ7] = {0, 0, 0, 0, 0, 1, -1};
constexpr float kW7[kQ7] = {0.25f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f};
constexpr int opp7(int q) { return q == 0 ? 0 : (q % 2 == 1 ? q + 1 : q - 1); }

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
// value(): stored element -> the FP32 working value (f for FP32, f - w for shifted storage);
// encode(): the inverse, round to nearest even.
struct StoreF32 {
  using T = float;
  static constexpr bool kShift = false;
  __device__ static __forceinline__ float value(T r) { return r; }
  __device__ static __forceinline__ T encode(float v) { return v; }
};

struct StoreF16 {
  using T = _Float16;
  static constexpr bool kShift = true;
  __device__ static __forceinline__ float value(T r) { return static_cast<float>(r); }
  __device__ static __forceinline__ T encode(float v) { return static_cast<T>(v); }
};

// bfloat16 = upper half of the binary32 pattern.
struct StoreBF16 {
  using T = uint16_t;
  static constexpr bool kShift = true;
  __device__ static __forceinline__ float value(T r) {
    return __uint_as_float(static_cast<uint32_t>(r) << 16);
  }
  __device__ static __forceinline__ T encode(float v) {
    const uint32_t u = __float_as_uint(v);
    return static_cast<T>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  }
};

// ---- population buffers as buffer resources ------------------------------------------------
// A raw buffer resource over a population buffer: address = base + scalar offset (the plane)
// + vector offset (the lane's cell), 32-bit each, so 19 loads or stores of a cell share one
// lane offset.  Used when the buffer spans less than 2^32 bytes (StepArgs::off32); the
// descriptor's third word is the gfx9 raw-buffer format (32-bit data, no swizzle).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, -1, 0x00020000);
}

// NT: the non-temporal cache policy (the "nt" bit, 2 in the cache-policy operand), for lines
// read once
template <class T, bool NT = false>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  constexpr int pol = NT ? 2 : 0;
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(
                                     r, static_cast<int>(voff), static_cast<int>(soff), pol));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(
                                     r, static_cast<int>(voff), static_cast<int>(soff), pol));
  }
}

// NT: non-temporal at device scope (the "nt" and "sc1" bits, 2 | 16 in the cache-policy
// operand); for 2-byte elements only in the per-wave row build (ROWS), "nt" alone otherwise.
// Measured at 512x256x256 (19 planes, tiles of rows): the FP32 step's store pattern 0.817 ms
// with nt + sc1 against 0.831 with nt alone and 0.909 with sc1 alone (the FP32 step 0.829 ->
// 0.814 ms); the FP16 step 0.4072-0.4092 against 0.4101-0.4110 ms, the FP16 batched sweep
// (per-lane rows) 2% slower with nt + sc1.
template <class T, bool NT, bool ROWS = true>
__device__ __forceinline__ void buffer_store(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                             uint32_t soff) {
  constexpr int pol = NT ? ((sizeof(T) == 4 || ROWS) ? (2 | 16) : 2) : 0;
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r,
                                          static_cast<int>(voff), static_cast<int>(soff), pol);
  } else {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r,
                                          static_cast<int>(voff), static_cast<int>(soff), pol);
  }
}

// ---- unsigned division by a divisor fixed for the launch ----------------------------------
// Granlund and Montgomery, "Division by invariant integers using multiplication" (PLDI 1994),
// section 4: with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, sh1 = min(l, 1) and
// sh2 = max(l - 1, 0), n / d = (t + ((n - t) >> sh1)) >> sh2 where t = mulhi(m, n), for every
// 32-bit n and 1 <= d < 2^31.  A block-uniform n compiles to a few scalar instructions.
struct Divisor {
  uint32_t m, sh1, sh2;
};

inline Divisor make_divisor(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
  return Divisor{static_cast<uint32_t>(m), l < 1 ? l : 1u, l > 1 ? l - 1 : 0u};
}

__host__ __device__ __forceinline__ uint32_t divide(uint32_t n, const Divisor& d) {
  const uint32_t t = static_cast<uint32_t>((static_cast<uint64_t>(d.m) * n) >> 32);
  return (t + ((n - t) >> d.sh1)) >> d.sh2;
}

// ---- kernel arguments (host fills them from tlbm_params) ----------------------------------
struct StepArgs {
  const void* fin;
  void* fout;
  uint64_t ns;               // population plane stride (elements): qstride, or k nx in tiles
  uint64_t gns;              // temperature plane stride (elements): qstride, or k nx in tiles
  uint64_t g_tile_stride;    // temperature tiles: 7 k nx (7 planes of k nx per tile of k rows,
                             // the populations' tiles), 0 for q-major planes
  uint32_t g_off32;          // 1: the temperature buffer lies within 2^32 bytes
  // Population layout.  Row r = z ny + y starts at element row_elem(r) of plane 0:
  //   q-major planes: tile_stride = 0, tile_log2 = 31, tile_mask = all ones -> r nx;
  //   tiles of k rows: tile_stride = 19 k nx, tile_log2 = log2 k, tile_mask = k - 1 ->
  //   (r / k) 19 k nx + (r % k) nx, each tile holding its 19 planes of k nx cells together.
  uint64_t tile_stride;
  uint32_t tile_log2, tile_mask;
  uint32_t row_waves;        // 1: nx is a multiple of 64, so each wave of 64 cells is in one row
  uint32_t xcd_group;        // quiet step: runs of this many consecutive blocks per XCD (xcd_chunk)
  uint32_t nchunks;          // quiet step: blocks of cells (the grid may hold more slots)
  int32_t nx, ny, nz;
  int32_t wrap;              // bit 0/1/2: the obstacle reaches across the x/y/z periodic faces
  uint32_t stream_stores;    // 1: launch the non-temporal-store (NT) quiet kernel (host side)
  uint32_t off32;            // 1: the whole population buffer lies within 2^32 bytes of its
                             // start, so it is addressed as one buffer resource with 32-bit
                             // plane (scalar) and lane (vector) offsets
  uint32_t pull_off[19];     // off32, q-major planes: byte offset of plane q less the pull
                             // shift (c_x + c_y nx + c_z nx ny) of population q
  uint32_t plane_off[19];    // off32: byte offset of plane q
  float omega;               // BGK
  float tau0, smag, magic;   // smag = 25.456 cs^2 (lbm.py:93)
  float o[8];                // obstacle parameters
  float r2;                  // r*r (cylinder, sphere)
  float reach[3];            // half extents + 2 (cells) for the neighbourhood bound
  float u0, s0, s_inv, s_k;  // sponge: start, 1 / (nx - start), strength
  float lid[3];
  const float* omega_rot;
  const float* phi;          // TLBM_FIELD solid fractions
  tlbm_fields fld;
  double* partials;          // report: per-block FP64 partial sums [TLBM_NSUMS][nblocks]
  uint32_t nblocks;
  Divisor div_nx, div_ny;    // cell index -> (x, y, z)
  // temperature variant (lbm.py:457-492)
  const float* g_in;
  float* g_out;
  const float* t_solid;
  float* q_solid;
  float* temp;
  float alpha, nu_mol, pr_t;
};

// First element of row r (= z ny + y) in population plane 0 (see StepArgs).
__device__ __forceinline__ uint64_t row_elem(const StepArgs& a, uint32_t r) {
  return static_cast<uint64_t>(r >> a.tile_log2) * a.tile_stride +
         (r & a.tile_mask) * static_cast<uint32_t>(a.nx);
}

// ---- obstacles: lbm.py:144-227 ---------------------------------------------------------------
// solid() uses one rounding per subtract, multiply and add (no contraction), as a float32
// evaluation of the reference's expression.  near() is a conservative bound: when it is false
// for a cell, none of its 19 pull sources that do not wrap around a periodic face is solid.
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
  __device__ static __forceinline__ float sink(const StepArgs& a, int x, int y, int z) {
    return solid(a, x, y, z);
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
  __device__ static __forceinline__ float sink(const StepArgs& a, int x, int y, int z) {
    return solid(a, x, y, z);
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
  __device__ static __forceinline__ float sink(const StepArgs& a, int x, int y, int z) {
    return solid(a, x, y, z);
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

// A per-cell solid fraction phi in [0, 1]: solid() is phi itself, so every link is a partial
// bounce-back f = pulled + phi (bounced - pulled) and every cell takes the wall path.
struct Field {
  static constexpr bool kMoving = false;
  __device__ static __forceinline__ float solid(const StepArgs& a, int x, int y, int z) {
    return a.phi[(static_cast<uint32_t>(z) * a.ny + y) * static_cast<uint32_t>(a.nx) + x];
  }
  __device__ static __forceinline__ float sink(const StepArgs& a, int x, int y, int z) {
    return solid(a, x, y, z);
  }
  __device__ static __forceinline__ bool near(const StepArgs&, int, int, int) { return true; }
  __device__ static __forceinline__ void wall_vel(const StepArgs&, float, float, float, float& u,
                                                  float& v, float& w) {
    u = v = w = 0.0f;
  }
};

// lbm.py:231-268: a finned heat sink on the floor of a duct.  o = {x0, x1, z0, z1, y_base,
// y_top, fin_thickness, fin_pitch}.  The sink (base plate and fins along x) is solid, and so are
// the duct's floor and ceiling rows y = 0 and y = ny - 1 (adiabatic walls).  Resting walls.
struct HeatSink {
  static constexpr bool kMoving = false;
  __device__ static __forceinline__ float sink(const StepArgs& a, int x, int y, int z) {
    const float fx = static_cast<float>(x), fy = static_cast<float>(y), fz = static_cast<float>(z);
    const bool inside = fx >= a.o[0] && fx < a.o[1] && fz >= a.o[2] && fz < a.o[3] && fy >= 1.0f;
    const bool base = fy < a.o[4];
    // z >= z0 wherever `inside` holds, so the remainder is that of a non-negative number
    const bool fin = fy < a.o[5] && (z - static_cast<int>(a.o[2])) % static_cast<int>(a.o[7]) <
                                        static_cast<int>(a.o[6]);
    return inside && (base || fin) ? 1.0f : 0.0f;
  }
  __device__ static __forceinline__ float solid(const StepArgs& a, int x, int y, int z) {
    return (y <= 0 || y >= a.ny - 1) ? 1.0f : sink(a, x, y, z);
  }
  __device__ static __forceinline__ bool near(const StepArgs& a, int x, int y, int z) {
    const float fx = static_cast<float>(x), fy = static_cast<float>(y), fz = static_cast<float>(z);
    return y <= 1 || y >= a.ny - 2 ||
           (fx >= a.o[0] - 1.0f && fx < a.o[1] + 1.0f && fz >= a.o[2] - 1.0f &&
            fz < a.o[3] + 1.0f && fy < a.o[5] + 1.0f);
  }
  __device__ static __forceinline__ void wall_vel(const StepArgs&, float, float, float, float& u,
                                                  float& v, float& w) {
    u = v = w = 0.0f;
  }
};

template <int OBS>
using ObstacleT = std::conditional_t<
    OBS == TLBM_CYLINDER, Cylinder,
    std::conditional_t<
        OBS == TLBM_SPHERE, Sphere,
        std::conditional_t<OBS == TLBM_BOX, Box,
                           std::conditional_t<OBS == TLBM_FIELD, Field, HeatSink>>>>;

template <int STORAGE>
using StorageT = std::conditional_t<STORAGE == TLBM_F32, StoreF32,
                                    std::conditional_t<STORAGE == TLBM_F16, StoreF16, StoreBF16>>;

// ---- equilibria ------------------------------------------------------------------------------
// feq_q = w rho (1 + cu + cu^2/2 - 1.5 u.u), cu = 3 c.u  (lbm.py:52-55), for all q at once;
// SHIFT: feq_q - w_q instead.
template <bool SHIFT>
__device__ __forceinline__ void equilibria(float rho, float ux, float uy, float uz,
                                           float (&e)[kQ]) {
  const float base = 1.0f - 1.5f * (ux * ux + uy * uy + uz * uz);
  if constexpr (SHIFT) e[0] = kW[0] * (rho * base - 1.0f);
  else e[0] = kW[0] * rho * base;
  const float wr1 = kW[1] * rho, wr2 = kW[7] * rho;
  static_for<0, 9>([&](auto pc) {
    constexpr int q = 2 * decltype(pc)::value + 1;
    constexpr int o = q + 1;
    float s = 0.0f;
    if constexpr (kCX[q] == 1) s += ux;
    if constexpr (kCX[q] == -1) s -= ux;
    if constexpr (kCY[q] == 1) s += uy;
    if constexpr (kCY[q] == -1) s -= uy;
    if constexpr (kCZ[q] == 1) s += uz;
    if constexpr (kCZ[q] == -1) s -= uz;
    const float wr = (q < 7) ? wr1 : wr2;
    float A = wr * (base + 4.5f * s * s);
    if constexpr (SHIFT) A -= kW[q];
    const float B = (3.0f * wr) * s;
    e[q] = A + B;
    e[o] = A - B;
  });
}

// Reciprocal and square root with single hardware instructions (v_rcp_f32 / v_sqrt_f32, within
// 1 ulp) instead of the correctly rounded multi-instruction sequences: the step's FP32 error
// budget against the FP64 oracle is ~1e-7 relative, far above one ulp of these factors.
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// tau with the Smagorinsky eddy viscosity, lbm.py:72-93, from d = f - feq.  The stress only
// involves the pair sums d_q + d_opp(q).  inv_rho = 1 / rho.
__device__ __forceinline__ float smagorinsky_tau(const float (&d)[kQ], float inv_rho, float tau0,
                                                 float smag) {
  float pxx = 0.0f, pyy = 0.0f, pzz = 0.0f, pxy = 0.0f, pxz = 0.0f, pyz = 0.0f;
  static_for<0, 9>([&](auto pc) {
    constexpr int q = 2 * decltype(pc)::value + 1;
    const float s = d[q] + d[q + 1];
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
  const float pi = fsqrt(pxx * pxx + pyy * pyy + pzz * pzz +
                         2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
  return 0.5f * (tau0 + fsqrt(tau0 * tau0 + smag * pi * inv_rho));
}

// ---- FP64 block reduction for the report variant -------------------------------------------
__device__ __forceinline__ void block_sums(double (&v)[TLBM_NSUMS], double* partials,
                                           uint32_t nblocks, uint32_t bid) {
  __shared__ double red[4][TLBM_NSUMS];  // up to 256 threads = 4 waves
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

// ---- one cell's collision and sponge (lbm.py:433-456) ------------------------------------------
// v: the 19 streamed working values in, the values to store out.
// What the temperature step reads from the flow step (lbm.py:445, 447-451).
struct CollideExtra {
  float tau;                // the relaxation time the collision returned
  float sigma;              // sponge blend
  float uxs, uys, uzs;      // sponge target velocity
};

template <int COLL, bool SHIFT, bool EXTRA = false>
__device__ __forceinline__ void collide_cell(const StepArgs& a, float (&v)[kQ], float x,
                                             float here, float& rho_out, float& ux_out,
                                             float& uy_out, float& uz_out,
                                             CollideExtra* ex = nullptr) {
  // moments (lbm.py:433-444); shifted storage: rho = 1 + sum d
  float rho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
  static_for<0, kQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    rho += v[q];
    if constexpr (kCX[q] == 1) jx += v[q];
    if constexpr (kCX[q] == -1) jx -= v[q];
    if constexpr (kCY[q] == 1) jy += v[q];
    if constexpr (kCY[q] == -1) jy -= v[q];
    if constexpr (kCZ[q] == 1) jz += v[q];
    if constexpr (kCZ[q] == -1) jz -= v[q];
  });
  if constexpr (SHIFT) rho += 1.0f;
  const float fluid = 1.0f - here;
  const float inv_rho = rcp(rho);
  const float scale = fluid * inv_rho;
  const float ux = jx * scale;
  const float uy = jy * scale;
  const float uz = jz * scale;

  // collision: d[q] = post_q - feq_q (lbm.py:60-139)
  float e[kQ];
  equilibria<SHIFT>(rho, ux, uy, uz, e);
  float d[kQ];
  static_for<0, kQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    d[q] = v[q] - e[q];
  });
  if constexpr (COLL == TLBM_BGK) {
    const float k = 1.0f - a.omega;
    static_for<0, kQ>([&](auto qc) { d[decltype(qc)::value] *= k; });
    if constexpr (EXTRA) ex->tau = 1.0f / a.omega;
  } else if constexpr (COLL == TLBM_LES) {
    const float tau = smagorinsky_tau(d, inv_rho, a.tau0, a.smag);
    const float k = 1.0f - rcp(tau);
    static_for<0, kQ>([&](auto qc) { d[decltype(qc)::value] *= k; });
    if constexpr (EXTRA) ex->tau = tau;
  } else {
    const float tau = smagorinsky_tau(d, inv_rho, a.tau0, a.smag);
    if constexpr (EXTRA) ex->tau = tau;
    const float kp = 1.0f - rcp(tau);
    const float km = 1.0f - rcp(a.magic * rcp(tau - 0.5f) + 0.5f);
    d[0] *= kp;
    static_for<0, 9>([&](auto pc) {
      constexpr int q = 2 * decltype(pc)::value + 1;
      const float np = 0.5f * (d[q] + d[q + 1]);
      const float nm = 0.5f * (d[q] - d[q + 1]);
      d[q] = kp * np + km * nm;
      d[q + 1] = kp * np - km * nm;
    });
  }

  // sponge toward the inflow state (lbm.py:446-456)
  const float t = fmaxf(0.0f, (x - a.s0) * a.s_inv);
  const float sigma = fmaxf(here, a.s_k * t * t);
  if (sigma != 0.0f) {
    const float rho_s = rho + sigma * (1.0f - rho);
    const float ux_s = ux + sigma * (a.u0 * fluid - ux);
    const float uy_s = (1.0f - sigma) * uy;
    const float uz_s = (1.0f - sigma) * uz;
    float es[kQ];
    equilibria<SHIFT>(rho_s, ux_s, uy_s, uz_s, es);
    const float keep = 1.0f - sigma;
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      v[q] = es[q] + keep * d[q];
    });
    if constexpr (EXTRA) {
      ex->uxs = ux_s;
      ex->uys = uy_s;
      ex->uzs = uz_s;
    }
  } else {
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      v[q] = e[q] + d[q];
    });
    if constexpr (EXTRA) {
      ex->uxs = ux;
      ex->uys = uy;
      ex->uzs = uz;
    }
  }
  if constexpr (EXTRA) ex->sigma = sigma;
  rho_out = rho;
  ux_out = ux;
  uy_out = uy;
  uz_out = uz;
}

// ---- the 19 pull loads of one cell (lbm.py:410-426), raw stored values ------------------------
// Sources: x and z periodic (lbm.py:412, 415), y by side (lbm.py:342-369): periodic, or the
// y-mirrored population of the same row at a slip face; ceiling links of the lid read their
// in-bounds (clamped) row and are overwritten by the caller.
// ROWS: rows are whole waves (nx a multiple of 64; blocks of consecutive cells then put every
// wave in one row), so y and z are the same for all lanes of a wave, and so are the source
// row, plane and y mirror of every population, on the faces too.  The wave forms each source
// row's address in scalar registers and the lanes add only their source column (x, or x -+ 1
// with the row's periodic wrap): 16-bit storage through one buffer resource (plane and row as
// the scalar offset), FP32 as global loads from a scalar base.  A compile-time choice (the host
// launches the ROWS build where nx % 64 == 0): with a runtime branch the compiler merged the
// two load sequences into one with per-lane 64-bit addresses.
// NTL (with ROWS): the nine populations with c_x = 0 are loaded with the non-temporal policy.
// Their lines are read by this wave alone, once per step, while the lines of the ten x-shifted
// loads are shared with the neighbouring wave (the +-1 cell at each wave's edge) and keep the
// default policy.  Measured on MI355X at 512x256x256 (LES, sphere, lid, tiles of rows,
// non-temporal stores), in one process against plain loads: FP32 0.834 -> 0.824 ms, FP16
// 0.409-0.413 -> 0.407-0.408 ms (all 19 loads non-temporal: slower than plain).
template <bool NT, class T>
__device__ __forceinline__ T pull_load(const T* __restrict__ p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

template <int SIDE, bool ROWS = false, bool NTL = false, class T>
__device__ __forceinline__ void pull_raw(const StepArgs& a, const T* __restrict__ fin, int x, int y,
                                         int z, T (&raw)[kQ]) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t snx = static_cast<uint32_t>(nx);
  const int xsrc[3] = {x == nx - 1 ? 0 : x + 1, x, x == 0 ? nx - 1 : x - 1};
  // No branch between the 19 loads of a lane, so that all of them are in flight together.
  if constexpr (ROWS) {
    const uint32_t yu = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(y));
    const uint32_t zu = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(z));
    const uint32_t uny = static_cast<uint32_t>(ny), unz = static_cast<uint32_t>(nz);
    const uint32_t zs[3] = {zu == unz - 1 ? 0u : zu + 1, zu, zu == 0 ? unz - 1 : zu - 1};
    uint32_t ys[3];
    bool lo = false, hi = false;
    if constexpr (SIDE == TLBM_PERIODIC_Y) {
      ys[0] = yu == uny - 1 ? 0u : yu + 1;
      ys[2] = yu == 0 ? uny - 1 : yu - 1;
    } else {
      ys[0] = yu == uny - 1 ? yu : yu + 1;
      ys[2] = yu == 0 ? yu : yu - 1;
      lo = (yu == 0);
      if constexpr (SIDE == TLBM_SLIP_Y) hi = (yu == uny - 1);
    }
    ys[1] = yu;
    const uint32_t col[3] = {static_cast<uint32_t>(xsrc[0]) * static_cast<uint32_t>(sizeof(T)),
                             static_cast<uint32_t>(x) * static_cast<uint32_t>(sizeof(T)),
                             static_cast<uint32_t>(xsrc[2]) * static_cast<uint32_t>(sizeof(T))};
    // the off32 test stays outside the loads: a branch between them would hold back the later
    // ones (and the compiler does not hoist it out of the unrolled sequence)
    if (sizeof(T) == 2 && a.off32) {
      // the whole buffer within 2^32 bytes: 32-bit row offsets
      const auto rin = plane_rsrc(fin);
      const uint32_t ts = static_cast<uint32_t>(a.tile_stride);
      if (yu >= 1 && yu < uny - 1 && zu >= 1 && zu < unz - 1) {
        // a wave on no y or z face (98% of them at 512x256x256): source row r0 - (c_y + c_z ny),
        // no wrap or mirror -- the fewest scalar instructions, which the VALU-heavy 16-bit step
        // measured (1% faster than the general form below)
        const uint32_t r0 = zu * uny + yu;
        static_for<0, kQ>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          constexpr int cx = kCX[q];
          const uint32_t r = r0 - static_cast<uint32_t>(kCY[q] + kCZ[q] * ny);
          const uint32_t so = ((r >> a.tile_log2) * ts + (r & a.tile_mask) * snx) *
                                  static_cast<uint32_t>(sizeof(T)) + a.plane_off[q];
          raw[q] = buffer_load<T, NTL && cx == 0>(rin, col[1 + cx], so);
        });
        return;
      }
      static_for<0, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
        int qs = q;
        if constexpr (SIDE != TLBM_PERIODIC_Y && cy == 1) qs = lo ? miry(q) : q;
        if constexpr (SIDE == TLBM_SLIP_Y && cy == -1) qs = hi ? miry(q) : q;
        const uint32_t r = zs[1 + cz] * uny + ys[1 + cy];
        const uint32_t so = ((r >> a.tile_log2) * ts + (r & a.tile_mask) * snx) *
                                static_cast<uint32_t>(sizeof(T)) + a.plane_off[qs];
        raw[q] = buffer_load<T, NTL && cx == 0>(rin, col[1 + cx], so);
      });
    } else {
      static_for<0, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
        int qs = q;
        if constexpr (SIDE != TLBM_PERIODIC_Y && cy == 1) qs = lo ? miry(q) : q;
        if constexpr (SIDE == TLBM_SLIP_Y && cy == -1) qs = hi ? miry(q) : q;
        const uint64_t e = static_cast<uint64_t>(qs) * a.ns + row_elem(a, zs[1 + cz] * uny + ys[1 + cy]);
        // a scalar base plus a 32-bit lane offset: the global_load's saddr form
        raw[q] = pull_load<NTL && cx == 0>(
            reinterpret_cast<const T*>(reinterpret_cast<const char*>(fin + e) + col[1 + cx]));
      });
    }
  } else {
    const uint32_t idx = (static_cast<uint32_t>(z) * ny + y) * snx + x;
    const int zsrc[3] = {z == nz - 1 ? 0 : z + 1, z, z == 0 ? nz - 1 : z - 1};
    int ysrc[3];
    bool mir_lo = false, mir_hi = false;
    if constexpr (SIDE == TLBM_PERIODIC_Y) {
      ysrc[0] = y == ny - 1 ? 0 : y + 1;
      ysrc[2] = y == 0 ? ny - 1 : y - 1;
    } else {
      ysrc[0] = y == ny - 1 ? y : y + 1;
      ysrc[2] = y == 0 ? y : y - 1;
      mir_lo = (y == 0);
      if constexpr (SIDE == TLBM_SLIP_Y) mir_hi = (y == ny - 1);
    }
    ysrc[1] = y;
    // Other rows, 16-bit storage, q-major planes, interior waves (no lane on a y or z face): the
    // source of population q is idx - (cx + cy nx + cz nx ny) for every lane, except the x wrap
    // of one lane at each row end, so the loads go through one buffer resource with the plane
    // offsets less the pull shift precomputed on the host (pull_off) and three shared lane
    // offsets.  The rest (FP32, face waves, tiles, buffers beyond 2^32 bytes) forms a 64-bit
    // address per load and lane.  Measured on the FP16 LES step before the buffer paths: 472
    // VALU instructions per wave against 347, on a box where the step's VALU was ~100% busy
    // (0.46 ms at 512x256x256).
    const bool inner = y >= 1 && y < ny - 1 && z >= 1 && z < nz - 1;
    if (sizeof(T) == 2 && a.off32 && a.tile_stride == 0 &&
        __builtin_amdgcn_ballot_w64(!inner) == 0) {
      constexpr uint32_t B = sizeof(T);
      const auto rin = plane_rsrc(fin);
      const uint32_t v0 = idx * B;
      const uint32_t vp = (idx + (x == 0 ? snx : 0u)) * B;       // c_x = +1: x - 1, wrapped at 0
      const uint32_t vm = (idx - (x == nx - 1 ? snx : 0u)) * B;  // c_x = -1: x + 1, wrapped
      static_for<0, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = kCX[q];
        raw[q] = buffer_load<T>(rin, cx == 1 ? vp : (cx == -1 ? vm : v0), a.pull_off[q]);
      });
    } else {
      static_for<0, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
        int qs = q;
        if constexpr (SIDE != TLBM_PERIODIC_Y && cy == 1) qs = mir_lo ? miry(q) : q;
        if constexpr (SIDE == TLBM_SLIP_Y && cy == -1) qs = mir_hi ? miry(q) : q;
        const T* __restrict__ row =
            fin + static_cast<uint64_t>(qs) * a.ns +
            row_elem(a, static_cast<uint32_t>(zsrc[1 + cz]) * ny + static_cast<uint32_t>(ysrc[1 + cy]));
        raw[q] = row[xsrc[1 + cx]];
      });
    }
  }
}

// ---- the 19 stores of one cell -------------------------------------------------------------------
// NT: non-temporal stores, for lattices whose two population buffers are far larger than the
// Infinity Cache (measured on MI355X: 1.6% faster at FP32, 3.4% at FP16, 512x256x256); below
// that size the next step's loads hit the cache lines plain stores leave there.
// row: the first element of the cell's row in plane 0 (row_elem), x its column.  ROWS (rows of
// whole waves, see pull_raw), FP32: the row is the same for the whole wave and its address is
// formed in scalar registers (global stores from a scalar base: 0.832 -> 0.828 ms at
// 512x256x256).  16-bit storage keeps one lane offset for all 19 planes and the plane offsets
// as the scalar operand (with the row in the scalar offset too, the FP16 step measured 1.3%
// slower).
template <bool NT, bool ROWS = false, class T>
__device__ __forceinline__ void store_raw(const StepArgs& a, T* __restrict__ fout, uint64_t row,
                                          uint32_t x, const T (&raw)[kQ]) {
  constexpr uint32_t B = sizeof(T);
  if constexpr (ROWS && sizeof(T) == 4) {
    const uint64_t ru = (static_cast<uint64_t>(__builtin_amdgcn_readfirstlane(
                             static_cast<uint32_t>(row >> 32))) << 32) |
                        __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(row));
    if (NT && a.off32) {
      // non-temporal stores at device scope need the buffer form (buffer_store)
      const auto rout = plane_rsrc(fout);
      const uint32_t rb = static_cast<uint32_t>(ru) * B;
      static_for<0, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        buffer_store<T, NT>(raw[q], rout, x * B, rb + a.plane_off[q]);
      });
      return;
    }
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      T* const dst = reinterpret_cast<T*>(
          reinterpret_cast<char*>(fout + (static_cast<uint64_t>(q) * a.ns + ru)) + x * B);
      if constexpr (NT) __builtin_nontemporal_store(raw[q], dst);
      else *dst = raw[q];
    });
  } else if ((sizeof(T) == 2 || NT) && a.off32) {
    // 16-bit, and FP32 non-temporal stores (device scope, buffer_store): one lane offset
    const auto rout = plane_rsrc(fout);
    const uint32_t vo = static_cast<uint32_t>(row + x) * B;
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      buffer_store<T, NT, ROWS>(raw[q], rout, vo, a.plane_off[q]);
    });
  } else {
    const uint64_t own = row + x;
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      T* const dst = (fout + static_cast<uint64_t>(q) * a.ns) + own;
      if constexpr (NT) __builtin_nontemporal_store(raw[q], dst);
      else *dst = raw[q];
    });
  }
}

// ---- the fused step of one cell ------------------------------------------------------------------
// Pull, bounce-back, collision, sponge and store for cell (x, y, z) of the lattice described by
// `a`, reading `fin` and writing `fout` (plane stride a.ns); x, y and z are per lane.  REPORT: writes the 13 fields at `fidx` (when F.ux is not null) and returns the 14 FP64
// summands in acc (zero for inactive lanes).
// ---- the temperature step of one cell (lbm.py:457-492) --------------------------------------------
// D3Q7 pull, periodic in x, y and z; a link from a solid voxel bounces back, and from a sink voxel
// anti-bounces to its temperature: g = pulled + wall (b (1 - 2 hot) + 2 hot w t_solid - pulled).
// BGK relaxation with omega_g = 1 / (4 alpha_eff + 1/2) and the flow's sponge blend.  REPORT:
// writes the cell temperature and adds fluid * temp to the q_solid of each sink voxel it faces.
template <int OBS, bool REPORT, bool NT = false>
__device__ __forceinline__ void heat_cell(const StepArgs& a, int x, int y, int z, uint32_t idx,
                                          float here, float ux, float uy, float uz,
                                          const CollideExtra& ex, bool active) {
  using Obs = ObstacleT<OBS>;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t snx = static_cast<uint32_t>(nx);
  const int xs[3] = {x == nx - 1 ? 0 : x + 1, x, x == 0 ? nx - 1 : x - 1};
  const int ys[3] = {y == ny - 1 ? 0 : y + 1, y, y == 0 ? ny - 1 : y - 1};
  const int zs[3] = {z == nz - 1 ? 0 : z + 1, z, z == 0 ? nz - 1 : z - 1};
  const float* __restrict__ gin = a.g_in;
  // first element of row r (= z ny + y) in temperature plane 0: q-major planes (r nx) or the
  // populations' tiles of rows with 7 planes per tile
  auto g_row = [&](uint32_t r) {
    return static_cast<uint64_t>(r >> a.tile_log2) * a.g_tile_stride + (r & a.tile_mask) * snx;
  };
  const uint64_t g_own = g_row(static_cast<uint32_t>(z) * ny + y) + static_cast<uint32_t>(x);
  float g[kQ7];
  uint32_t nidx[kQ7];  // source cells (t_solid, q_solid)
  static_for<0, kQ7>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    constexpr int cx = kCX7[q], cy = kCY7[q], cz = kCZ7[q];
    const uint32_t r = static_cast<uint32_t>(zs[1 + cz]) * ny + static_cast<uint32_t>(ys[1 + cy]);
    nidx[q] = r * snx + static_cast<uint32_t>(xs[1 + cx]);
    g[q] = gin[static_cast<uint64_t>(q) * a.gns + g_row(r) + static_cast<uint32_t>(xs[1 + cx])];
  });
  const bool wrap = ((a.wrap & 1) && (x == 0 || x == nx - 1)) ||
                    ((a.wrap & 2) && (y == 0 || y == ny - 1)) ||
                    ((a.wrap & 4) && (z == 0 || z == nz - 1));
  const bool walls = wrap || Obs::near(a, x, y, z);
  if (walls) {
    // the cell's own populations and the facing sinks' temperatures are loaded together,
    // before any is used (one memory round trip, not one per wall link); t_solid only where a
    // sink faces the link
    float own[kQ7], wall[kQ7], hot[kQ7], ts[kQ7];
    static_for<0, kQ7>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = kCX7[q], cy = kCY7[q], cz = kCZ7[q];
      own[q] = gin[static_cast<uint64_t>(opp7(q)) * a.gns + g_own];
      wall[q] = q == 0 ? here : Obs::solid(a, xs[1 + cx], ys[1 + cy], zs[1 + cz]);
      hot[q] = wall[q] != 0.0f ? Obs::sink(a, xs[1 + cx], ys[1 + cy], zs[1 + cz]) : 0.0f;
    });
    static_for<0, kQ7>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      ts[q] = hot[q] != 0.0f ? a.t_solid[nidx[q]] : 0.0f;
    });
    static_for<0, kQ7>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      if (wall[q] != 0.0f) {
        float at_wall = own[q] * (1.0f - 2.0f * hot[q]);
        if (hot[q] != 0.0f) at_wall += hot[q] * 2.0f * kW7[q] * ts[q];
        g[q] = g[q] + wall[q] * (at_wall - g[q]);
      }
    });
  }
  float temp = 0.0f;
  static_for<0, kQ7>([&](auto qc) { temp += g[decltype(qc)::value]; });
  const float alpha = a.alpha + fmaxf(0.0f, (ex.tau - 0.5f) / 3.0f - a.nu_mol) / a.pr_t;
  const float omega_g = 1.0f / (4.0f * alpha + 0.5f);
  const float keep = 1.0f - ex.sigma;
  const float temp_s = keep * temp;
  if (active) {
    float gout[kQ7];
    static_for<0, kQ7>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr float cx = kCX7[q], cy = kCY7[q], cz = kCZ7[q];
      const float eq = kW7[q] * temp * (1.0f + 4.0f * (cx * ux + cy * uy + cz * uz));
      const float post = g[q] - omega_g * (g[q] - eq);
      const float eqs = kW7[q] * temp_s * (1.0f + 4.0f * (cx * ex.uxs + cy * ex.uys + cz * ex.uzs));
      gout[q] = eqs + keep * (post - eq);
    });
    if (NT && a.g_off32) {
      // non-temporal stores at device scope (buffer_store), the buffer within 2^32 bytes
      const auto rg = plane_rsrc(a.g_out);
      static_for<0, kQ7>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        buffer_store<float, true>(gout[q], rg, static_cast<uint32_t>(g_own) * 4u,
                                  static_cast<uint32_t>(q * a.gns * 4u));
      });
    } else {
      static_for<0, kQ7>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        float* dst = &a.g_out[static_cast<uint64_t>(q) * a.gns + g_own];
        if constexpr (NT) __builtin_nontemporal_store(gout[q], dst);
        else *dst = gout[q];
      });
    }
  }
  if constexpr (REPORT) {
    if (active) {
      const float fluid = 1.0f - here;
      float tv = fluid * temp;
      if (here != 0.0f) tv += here * a.t_solid[idx];
      a.temp[idx] = tv;
      if (walls && fluid != 0.0f) {
        static_for<1, kQ7>([&](auto qc) {
          constexpr int q = decltype(qc)::value;
          constexpr int cx = kCX7[q], cy = kCY7[q], cz = kCZ7[q];
          const float hot = Obs::sink(a, xs[1 + cx], ys[1 + cy], zs[1 + cz]);
          if (hot != 0.0f) atomicAdd(&a.q_solid[nidx[q]], fluid * hot * temp);
        });
      }
    }
  }
}

// NT: non-temporal population stores (and, with ROWS, non-temporal loads of the unshifted
// pulls); ROWS: rows of whole waves, addressed per wave (pull_raw).
template <int COLL, int OBS, int SIDE, bool REPORT, int STORAGE, bool HEAT = false,
          bool NT = false, bool ROWS = false>
__device__ __forceinline__ void step_cell(const StepArgs& a, const typename StorageT<STORAGE>::T* __restrict__ fin,
                                          typename StorageT<STORAGE>::T* __restrict__ fout,
                                          int x, int y, int z, bool active,
                                          const tlbm_fields& F, uint32_t fidx,
                                          double (&acc)[TLBM_NSUMS]) {
  using Obs = ObstacleT<OBS>;
  using St = StorageT<STORAGE>;
  using T = typename St::T;
  constexpr bool kShift = St::kShift;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t snx = static_cast<uint32_t>(nx);
  const uint32_t idx = (static_cast<uint32_t>(z) * ny + y) * snx + x;
  // the cell's row and own element in population plane 0 (idx for q-major planes)
  const uint64_t own_row = row_elem(a, static_cast<uint32_t>(z) * ny + y);
  const uint64_t own_e = own_row + static_cast<uint32_t>(x);

  // pull sources (x, z periodic: lbm.py:412, 415; y by side: lbm.py:342-369), indexed [1 + c]:
  // the source of a population moving with +c is one cell back.
  const int xsrc[3] = {x == nx - 1 ? 0 : x + 1, x, x == 0 ? nx - 1 : x - 1};
  const int zsrc[3] = {z == nz - 1 ? 0 : z + 1, z, z == 0 ? nz - 1 : z - 1};
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
  {
    T raw[kQ];
    pull_raw<SIDE, ROWS, NT && ROWS>(a, fin, x, y, z, raw);
    static_for<0, kQ>([&](auto qc) { f[decltype(qc)::value] = St::value(raw[decltype(qc)::value]); });
  }

  // ---- moving lid: f_q = f_opp(x) + 6 w_q c_q . u_lid on ceiling links ----
  if constexpr (SIDE == TLBM_LID_Y) {
    if (lid_hi) {
      static_for<0, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (kCY[q] == -1) {
          constexpr int o = opp(q);
          const float t =
              6.0f * kW[q] * (kCX[q] * a.lid[0] + kCY[q] * a.lid[1] + kCZ[q] * a.lid[2]);
          f[q] = St::value(fin[static_cast<uint64_t>(o) * a.ns + own_e]) + t;
        }
      });
    }
  }

  // ---- obstacle: halfway / moving-wall bounce-back and momentum exchange (lbm.py:416-432) ----
  float here = 0.0f;
  float mx = 0.0f, my = 0.0f, mz = 0.0f;
  bool wrap = ((a.wrap & 1) && (x == 0 || x == nx - 1)) ||
              ((a.wrap & 4) && (z == 0 || z == nz - 1));
  if constexpr (SIDE == TLBM_PERIODIC_Y) wrap = wrap || ((a.wrap & 2) && (y == 0 || y == ny - 1));
  if (wrap || Obs::near(a, x, y, z)) {
    here = Obs::solid(a, x, y, z);
    const float fluid = 1.0f - here;
    float omega_rot = 0.0f;
    if constexpr (Obs::kMoving) omega_rot = a.omega_rot ? *a.omega_rot : 0.0f;
    // The cell's own populations, bounced back on wall links (lbm.py:427): all 18 loads are
    // issued together, before any is needed, so a wave near the wall waits for memory once
    // rather than once per wall link.
    T own[kQ];
    if (sizeof(T) == 2 && a.off32) {
      // one lane offset, the planes' offsets in scalar registers (as the 16-bit pulls)
      const auto rin = plane_rsrc(fin);
      static_for<1, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        own[q] = buffer_load<T>(rin, static_cast<uint32_t>(own_e) * static_cast<uint32_t>(sizeof(T)),
                                a.plane_off[opp(q)]);
      });
    } else {
      static_for<1, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        own[q] = (fin + static_cast<uint64_t>(opp(q)) * a.ns)[own_e];
      });
    }
    static_for<1, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
      constexpr int o = opp(q);
      if constexpr (SIDE == TLBM_LID_Y && cy == -1) {
        if (lid_hi) return;
      }
      const float wall = Obs::solid(a, xsrc[1 + cx], ysrc[1 + cy], zsrc[1 + cz]);
      if (wall != 0.0f) {
        const float b = St::value(own[q]);
        float t = 0.0f;
        if constexpr (Obs::kMoving) {
          // wall velocity at the link midpoint (lbm.py:418-425); a surface at rest (the
          // launch-wide omega_rot = 0) gives t = 0 exactly, so its divisions are skipped
          if (omega_rot != 0.0f) {
            float uwx, uwy, uwz;
            Obs::wall_vel(a, static_cast<float>(x) - 0.5f * cx, static_cast<float>(y) - 0.5f * cy,
                          static_cast<float>(z) - 0.5f * cz, uwx, uwy, uwz);
            t = 6.0f * kW[q] *
                (cx * (omega_rot * uwx) + cy * (omega_rot * uwy) + cz * (omega_rot * uwz));
          }
        }
        // lbm.py:428-429; with a 0/1 obstacle wall = 1 here, so f = b + t exactly
        if constexpr (OBS == TLBM_FIELD) f[q] = f[q] + wall * (b - f[q] + t);
        else f[q] = b + t;
        if constexpr (REPORT) {
          const float hit = wall * fluid * (2.0f * (kShift ? b + kW[o] : b) + t);
          mx += hit * static_cast<float>(kCX[o]);
          my += hit * static_cast<float>(kCY[o]);
          mz += hit * static_cast<float>(kCZ[o]);
        }
      }
    });
  }

  // ---- collision, sponge, store ----
  float rho, ux, uy, uz;
  CollideExtra ex;
  collide_cell<COLL, kShift, HEAT>(a, f, static_cast<float>(x), here, rho, ux, uy, uz, &ex);
  if (active) {
    T raw[kQ];
    static_for<0, kQ>([&](auto qc) { raw[decltype(qc)::value] = St::encode(f[decltype(qc)::value]); });
    store_raw<NT, ROWS>(a, fout, own_row, static_cast<uint32_t>(x), raw);
  }
  if constexpr (HEAT) heat_cell<OBS, REPORT, NT>(a, x, y, z, idx, here, ux, uy, uz, ex, active);

  // ---- report fields and FP64 summands (lbm.py:493-513, 649-661) ----
  if constexpr (REPORT) {
#pragma unroll
    for (int i = 0; i < TLBM_NSUMS; ++i) acc[i] = 0.0;
    if (active) {
      const float p = (rho - 1.0f) * (1.0f / 3.0f);
      const float usq = ux * ux + uy * uy + uz * uz;
      const float ke = 0.5f * rho * usq;
      const float vmag = fsqrt(usq);
      float r_rho = 0.0f, r_ux = 0.0f, r_uy = 0.0f, r_uz = 0.0f, r_ke = 0.0f;
      if (F.ux) {
        const float ox = F.ux[fidx], oy = F.uy[fidx], oz = F.uz[fidx], op = F.p[fidx];
        const float ok = 0.5f * (ox * ox + oy * oy + oz * oz);
        r_rho = (3.0f * (p - op)) * (3.0f * (p - op));
        r_ux = (ux - ox) * (ux - ox);
        r_uy = (uy - oy) * (uy - oy);
        r_uz = (uz - oz) * (uz - oz);
        r_ke = (ke - ok) * (ke - ok);
        F.r_rho[fidx] = r_rho;
        F.r_ux[fidx] = r_ux;
        F.r_uy[fidx] = r_uy;
        F.r_uz[fidx] = r_uz;
        F.r_ke[fidx] = r_ke;
        F.ux[fidx] = ux;
        F.uy[fidx] = uy;
        F.uz[fidx] = uz;
        F.p[fidx] = p;
        F.vmag[fidx] = vmag;
        F.fx[fidx] = mx;
        F.fy[fidx] = my;
        F.fz[fidx] = mz;
      }
      const float vals[TLBM_NSUMS] = {ux,   uy,   uz,   p,  vmag, r_rho, r_ux,
                                      r_uy, r_uz, r_ke, mx, my,   mz,    rho};
#pragma unroll
      for (int i = 0; i < TLBM_NSUMS; ++i) acc[i] = static_cast<double>(vals[i]);
    }
  }
}

// The chunk (block of consecutive cells) hardware block slot s steps.  Slot s runs on XCD s mod 8
// (blocks are dispatched to the eight XCDs round-robin); with group g > 1 the slots of one XCD
// take runs of g consecutive chunks, so the cache lines two neighbouring chunks share (the
// x-shifted pulls across their common edge) are fetched into one L2 rather than two; g = 1 is
// plain order.  A grid of a multiple of 8 g slots covers every chunk once; slots whose chunk lies
// past the last one have nothing to do.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t s, uint32_t g) {
  if (g <= 1) return s;
  const uint32_t x = s & 7u, j = s >> 3;
  return g * (8u * (j / g) + x) + j % g;
}

// (x, y, z) of a cell index (x fastest): two divisions by invariant integers, exact for every
// 32-bit index.
__device__ __forceinline__ void cell_coords(uint32_t cell, uint32_t nx, uint32_t ny,
                                            const Divisor& div_nx, const Divisor& div_ny,
                                            uint32_t& x, uint32_t& y, uint32_t& z) {
  const uint32_t row = divide(cell, div_nx);
  x = cell - row * nx;
  z = divide(row, div_ny);
  y = row - z * ny;
}

// ---- the step kernel of one lattice -------------------------------------------------------------
// HEAT: with the temperature step (FP32 populations only).
// Block b takes cells [b * B, b * B + B) in index order, so every wave reads and writes whole
// aligned cache lines of each plane whatever nx is; y and z are per lane.  Measured on MI355X
// against blocks along x-rows (y and z uniform per block, row blocks kept on one XCD): equal or
// up to 4% faster on rows that are a multiple of 64 cells, and 1.5-2x faster on rows that are
// not (their waves started off a cache line, and two XCDs wrote the line a row ends in).
template <int COLL, int OBS, int SIDE, bool REPORT, int STORAGE, bool HEAT = false,
          bool NT = false, bool ROWS = false>
__global__ void __launch_bounds__(256) lbm_step_kernel(const StepArgs a) {
  static_assert(!HEAT || STORAGE == TLBM_F32, "the temperature variant stores FP32");
  using T = typename StorageT<STORAGE>::T;
  const uint32_t n = static_cast<uint32_t>(a.nx) * a.ny * static_cast<uint32_t>(a.nz);
  // report steps keep plain order: their FP64 block partials are indexed by block
  uint32_t chunk = blockIdx.x;
  if constexpr (!REPORT) {
    chunk = xcd_chunk(blockIdx.x, a.xcd_group);
    if (chunk >= a.nchunks) return;
  }
  const uint32_t cell = chunk * blockDim.x + threadIdx.x;
  const bool active = cell < n;
  if constexpr (!REPORT) {
    if (!active) return;
  }
  // in the report variant every lane reaches the block reduction: lanes past the last cell
  // mirror it (in-bounds loads) and store nothing
  const uint32_t idx = active ? cell : n - 1;
  uint32_t x, y, z;
  if constexpr (ROWS) {
    // rows of whole waves: the wave's row, y and z in scalar registers (the face, lid and row
    // addressing logic then runs on the scalar unit), only x per lane
    const uint32_t row = divide(__builtin_amdgcn_readfirstlane(idx), a.div_nx);
    x = idx - row * static_cast<uint32_t>(a.nx);
    z = divide(row, a.div_ny);
    y = row - z * static_cast<uint32_t>(a.ny);
  } else {
    cell_coords(idx, a.nx, a.ny, a.div_nx, a.div_ny, x, y, z);
  }
  double acc[TLBM_NSUMS];
  step_cell<COLL, OBS, SIDE, REPORT, STORAGE, HEAT, NT, ROWS>(
      a, static_cast<const T*>(a.fin), static_cast<T*>(a.fout), static_cast<int>(x),
      static_cast<int>(y), static_cast<int>(z), active, a.fld, idx, acc);
  if constexpr (REPORT) {
    if (a.partials) block_sums(acc, a.partials, a.nblocks, blockIdx.x);
  }
}

// ---- the step's data movement alone --------------------------------------------------------------
// f_out[q][idx] = f_in[q][source of q] for the 19 populations, the sources of a periodic lattice,
// with the step's mapping, index split, loads and stores and no arithmetic: the bandwidth this
// access pattern reaches is the ceiling of the fused step.  T: the stored element (4 or 2 bytes).
template <class T, bool NT, bool ROWS = false>
__global__ void __launch_bounds__(256) pull_copy_kernel(const StepArgs a) {
  const uint32_t n = static_cast<uint32_t>(a.nx) * a.ny * static_cast<uint32_t>(a.nz);
  const uint32_t chunk = xcd_chunk(blockIdx.x, a.xcd_group);
  if (chunk >= a.nchunks) return;
  const uint32_t idx = chunk * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  uint32_t x, y, z;
  cell_coords(idx, a.nx, a.ny, a.div_nx, a.div_ny, x, y, z);
  T raw[kQ];
  pull_raw<TLBM_PERIODIC_Y, ROWS, NT && ROWS>(a, static_cast<const T*>(a.fin), static_cast<int>(x),
                            static_cast<int>(y), static_cast<int>(z), raw);
  store_raw<NT, ROWS>(a, static_cast<T*>(a.fout), row_elem(a, z * a.ny + y), x, raw);
}

// ---- batched lattices: many independent domains in one persistent launch -------------------------
// Domain d holds its populations at elem_off[d] (plane stride = its cell count) and its report
// fields at cell_off[d] of the packed per-cell arrays.  The cells of all domains are cut into
// chunks of blockDim.x consecutive cells that never straddle two domains; a (persistent) grid
// walks the chunk slots with a block stride.
struct BatchDomain {
  StepArgs a;          // the domain's parameters (pointers in it are unused)
  uint64_t elem_off;   // first population element of the domain
  uint32_t cell_off;   // first cell of the domain in packed per-cell arrays
  uint32_t chunk_off;  // first chunk of the domain
};

struct BatchArgs {
  const void* fin;
  void* fout;
  const BatchDomain* dom;
  const uint16_t* chunk_dom;  // domain of each chunk
  int32_t ndom;
  uint32_t nchunks;
  tlbm_fields fld;   // packed over all domains (cell_off), or all null
  double* partials;  // report: [TLBM_NSUMS][nchunks]
  uint32_t group;    // chunks per XCD run (1: plain order); slots = nchunks rounded up to 8 group
  uint32_t nslots;
};

template <int COLL, int OBS, int SIDE, bool REPORT, int STORAGE, bool NT = false>
__global__ void __launch_bounds__(256) lbm_batch_kernel(const BatchArgs b) {
  using T = typename StorageT<STORAGE>::T;
  const T* __restrict__ fin = static_cast<const T*>(b.fin);
  T* __restrict__ fout = static_cast<T*>(b.fout);
  // Chunk order (xcd_chunk): plain, consecutive chunks on consecutive XCDs, so the eight
  // XCDs stream through nearby addresses of every plane together; or runs of `group`
  // consecutive chunks per XCD, so the lines neighbouring chunks share stay in one L2.
  // Handing each XCD a contiguous eighth of all chunks measured 7-9% slower at FP32 on
  // MI355X: eight streams per plane, each an eighth of the lattice apart, contend for the
  // same HBM channels.
  for (uint32_t slot = blockIdx.x; slot < b.nslots; slot += gridDim.x) {
    const uint32_t c = xcd_chunk(slot, b.group);
    if (c >= b.nchunks) continue;  // uniform over the block
    const BatchDomain& D = b.dom[b.chunk_dom[c]];
    const StepArgs& a = D.a;
    const uint32_t n = static_cast<uint32_t>(a.nx) * a.ny * a.nz;
    // lanes past the domain's last cell mirror it (in-bounds loads) and store nothing
    const uint32_t local = (c - D.chunk_off) * blockDim.x + threadIdx.x;
    const bool active = local < n;
    const uint32_t cell = active ? local : n - 1;
    uint32_t x, y, z;
    cell_coords(cell, a.nx, a.ny, a.div_nx, a.div_ny, x, y, z);
    double acc[TLBM_NSUMS];
    // per-lane addressing and plain loads (ROWS false): the domains' rows need not be whole
    // waves (non-temporal unshifted loads measured 4% slower on the FP32 sweep)
    step_cell<COLL, OBS, SIDE, REPORT, STORAGE, false, NT, false>(a, fin + D.elem_off, fout + D.elem_off,
                                                static_cast<int>(x), static_cast<int>(y),
                                                static_cast<int>(z), active, b.fld,
                                                D.cell_off + cell, acc);
    if constexpr (REPORT) {
      if (b.partials) block_sums(acc, b.partials, b.nchunks, c);
      __syncthreads();  // the block reuses its reduction scratch for the next chunk
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
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= static_cast<uint32_t>(a.nx) * a.ny * static_cast<uint32_t>(a.nz)) return;
  uint32_t xr, yr, zr;
  cell_coords(idx, a.nx, a.ny, a.div_nx, a.div_ny, xr, yr, zr);
  const int x = static_cast<int>(xr), y = static_cast<int>(yr), z = static_cast<int>(zr);
  const float fluid = 1.0f - Obs::solid(a, x, y, z);
  const float two_pi = 6.2831853f;
  float ux = a.u0 * fluid;
  float uy = seed_amp * a.u0 * fluid *
             sinf(two_pi * static_cast<float>(x) / static_cast<float>(a.nx)) *
             cosf(two_pi * static_cast<float>(z) / static_cast<float>(a.nz));
  float uz = 0.0f;
  if (noise_amp != 0.0f) {
    ux += fluid * (noise_amp * (2.0f * hash_uniform(noise_seed, idx, 0) - 1.0f));
    uy += fluid * (noise_amp * (2.0f * hash_uniform(noise_seed, idx, 1) - 1.0f));
    uz += fluid * (noise_amp * (2.0f * hash_uniform(noise_seed, idx, 2) - 1.0f));
  }
  float e[kQ];
  equilibria<St::kShift>(1.0f, ux, uy, uz, e);
  const uint64_t own_e = row_elem(a, zr * a.ny + yr) + xr;
  static_for<0, kQ>([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    f[static_cast<uint64_t>(q) * a.ns + own_e] = St::encode(e[q]);
  });
}

}  // namespace tlbm
