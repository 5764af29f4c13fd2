// Discrete adjoint of the fused step for a BGK lattice with a solid-fraction obstacle
// (TLBM_FIELD), periodic in all directions, FP32 populations.
//
// Forward step t at cell x (lbm_step_kernel with the same parameters), with h = phi(x),
// fl = 1 - h, link weights w_q = phi(x - c_q):
//   g_q = p_q + w_q (b_q - p_q),  p_q = f_t[q](x - c_q),  b_q = f_t[opp q](x)      (pull + bounce)
//   rho = sum g,  u = fl j / rho,  e = feq(rho, u)
//   sigma = max(h, k t(x)^2),  rho_s = rho + sigma (1 - rho),
//   u_s = (u_x + sigma (u0 fl - u_x), (1 - sigma) u_y, (1 - sigma) u_z)
//   f_{t+1}[q](x) = feq_q(rho_s, u_s) + K (g_q - e_q),  K = (1 - sigma)(1 - omega)
//   drag D_t(x) = sum_q w_q fl (2 b_q) c_{opp q, x} = -2 fl sum_q w_q b_q c_qx   (lbm.py:429-430)
// Objective J = sum_t lam_t sum_x D_t(x) (lam_t: the averaging weight of step t).
//
// With mu^t(x) = dJ/dg^t(x), one reverse step t computes, per cell x:
//   Lam_r = dJ/df_{t+1}[r](x) = (1 - phi(x)) mu^{t+1}_r(x + c_r)
//           + phi(x + c_r) (mu^{t+1}_{opp r}(x) + 2 lam_{t+1} fl(x) c_rx)      (reverse streaming)
//   mu^t_r = K Lam_r + A_rho + sum_a A_a (fl c_ra - u_a) / rho               (transposed Jacobian)
//   with A_rho = sum_q Lam_q ((1 - sigma) es_q / rho_s - K e_q / rho),
//        A_a = (1 - sigma) rho_s P_a(u_s) - K rho P_a(u),
//        P_a(v) = sum_q Lam_q w_q (3 c_qa + 9 (c_q.v) c_qa - 3 v_a)  (w_q: lattice weights).
// and, for cells of the parameter block, dJ/dphi(x) accumulates
//   (a) step t, through h = phi(x): u (d u_a / dh = -j_a / rho), the explicit fl of u_s_x,
//       sigma (when sigma = h) and the drag;
//   (b) step t + 1, through the links whose wall is phi(x): cell y = x + c_r has
//       w_r(y) = phi(x), so dJ/dphi(x) += mu^{t+1}_r(y) (b_r(y) - p_r(y)) - 2 lam_{t+1} fl(y)
//       b_r(y) c_rx, with p_r(y) = f_{t+1}[r](x), b_r(y) = f_{t+1}[opp r](y).
// Every term lands on the thread's own cell: no atomics, no races.
#pragma once

#include "lbm_step.hpp"

namespace tlbm {

struct AdjArgs {
  const float* f_t;   // state at step t (input of forward step t)
  const float* f_t1;  // state at step t + 1
  const float* mu_in;  // mu^{t+1}
  float* mu_out;       // mu^t
  const float* phi;
  uint64_t ns;
  int32_t nx, ny, nz;
  float omega, u0, s0, s_inv, s_k;
  float lam_t, lam_t1;   // drag weights of steps t and t + 1
  int32_t blk[3];        // parameter block origin
  int32_t bn[3];         // parameter block size
  double* gphi;          // dJ/dphi over the parameter block, x fastest
  Divisor div_nx, div_ny;  // cell index -> (x, y, z)
};

// One thread per cell, blocks of consecutive cells in index order, as lbm_step_kernel maps
// them (on the 128x64x64 adjoint workload 3% faster than blocks along x-rows).
template <bool FINAL>
__global__ void __launch_bounds__(256) lbm_adjoint_kernel(const AdjArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t snx = static_cast<uint32_t>(nx), plane = snx * static_cast<uint32_t>(ny);
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= plane * static_cast<uint32_t>(nz)) return;
  uint32_t xr, yr, zr;
  cell_coords(idx, snx, ny, a.div_nx, a.div_ny, xr, yr, zr);
  const int x = static_cast<int>(xr), y = static_cast<int>(yr), z = static_cast<int>(zr);
  // neighbour coordinates, indexed [1 + c]: x - c (upstream) ...
  const int xm[3] = {x == nx - 1 ? 0 : x + 1, x, x == 0 ? nx - 1 : x - 1};
  const int ym[3] = {y == ny - 1 ? 0 : y + 1, y, y == 0 ? ny - 1 : y - 1};
  const int zm[3] = {z == nz - 1 ? 0 : z + 1, z, z == 0 ? nz - 1 : z - 1};
  // ... and x + c (downstream) is the same table indexed [1 - c]
  auto at = [&](int xx, int yy, int zz) {
    return (static_cast<uint32_t>(zz) * ny + yy) * snx + static_cast<uint32_t>(xx);
  };

  const int bx = x - a.blk[0], by = y - a.blk[1], bz = z - a.blk[2];
  const bool in_block = bx >= 0 && bx < a.bn[0] && by >= 0 && by < a.bn[1] && bz >= 0 &&
                        bz < a.bn[2];
  const float h = a.phi[idx];
  const float fl = 1.0f - h;

  // (b): links of step t + 1 whose wall fraction is phi(x)
  double gsum = 0.0;
  if (in_block) {
    static_for<1, kQ>([&](auto qc) {
      constexpr int r = decltype(qc)::value;
      constexpr int cx = kCX[r], cy = kCY[r], cz = kCZ[r];
      const uint32_t yi = at(xm[1 - cx], ym[1 - cy], zm[1 - cz]);  // y = x + c_r
      const float mu = a.mu_in[static_cast<uint64_t>(r) * a.ns + yi];
      const float bp = a.f_t1[static_cast<uint64_t>(opp(r)) * a.ns + yi];
      const float pp = a.f_t1[static_cast<uint64_t>(r) * a.ns + idx];
      const float fly = 1.0f - a.phi[yi];
      gsum += static_cast<double>(mu * (bp - pp) - 2.0f * a.lam_t1 * fly * bp * cx);
    });
  }
  if constexpr (FINAL) {
    if (in_block) {
      const uint32_t bi = (static_cast<uint32_t>(bz) * a.bn[1] + by) * a.bn[0] + bx;
      a.gphi[bi] += gsum;
    }
    return;
  } else {
    // ---- Lam = dJ/df_{t+1}(x): reverse streaming + reverse bounce-back + drag source ----
    // wn[r] = phi(x + c_r): the wall fraction of link opp(r) of this cell, w_{opp r}(x).
    // Where any link has a wall, the cell's own populations (mu^{t+1} and f_t at x) are loaded
    // together, before they are needed, rather than one round trip per wall link.
    float L[kQ], wn[kQ];
    bool walls = false;
    L[0] = a.mu_in[idx];
    static_for<1, kQ>([&](auto qc) {
      constexpr int r = decltype(qc)::value;
      constexpr int cx = kCX[r], cy = kCY[r], cz = kCZ[r];
      const uint32_t yi = at(xm[1 - cx], ym[1 - cy], zm[1 - cz]);  // x + c_r
      L[r] = fl * a.mu_in[static_cast<uint64_t>(r) * a.ns + yi];
      wn[r] = a.phi[yi];
      walls = walls || wn[r] != 0.0f;
    });
    if (walls) {
      float own[kQ];
      static_for<1, kQ>([&](auto qc) {
        constexpr int r = decltype(qc)::value;
        own[r] = a.mu_in[static_cast<uint64_t>(opp(r)) * a.ns + idx];
      });
      static_for<1, kQ>([&](auto qc) {
        constexpr int r = decltype(qc)::value;
        if (wn[r] != 0.0f)
          L[r] += wn[r] * (own[r] + 2.0f * a.lam_t1 * fl * static_cast<float>(kCX[r]));
      });
    }

    // ---- forward recompute of step t at x ----
    float g[kQ];
    float dragw = 0.0f;  // sum_q w_q b_q c_qx
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr int cx = kCX[q], cy = kCY[q], cz = kCZ[q];
      const uint32_t si = at(xm[1 + cx], ym[1 + cy], zm[1 + cz]);  // x - c_q
      g[q] = a.f_t[static_cast<uint64_t>(q) * a.ns + si];
    });
    if (walls) {
      float own[kQ];
      static_for<1, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        own[q] = a.f_t[static_cast<uint64_t>(opp(q)) * a.ns + idx];
      });
      static_for<1, kQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float w = wn[opp(q)];  // phi(x - c_q)
        if (w != 0.0f) {
          const float p = g[q], b = own[q];
          g[q] = p + w * (b - p);
          dragw += w * b * static_cast<float>(kCX[q]);
        }
      });
    }
    float rho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      rho += g[q];
      if constexpr (kCX[q] != 0) jx += kCX[q] * g[q];
      if constexpr (kCY[q] != 0) jy += kCY[q] * g[q];
      if constexpr (kCZ[q] != 0) jz += kCZ[q] * g[q];
    });
    const float inv_rho = 1.0f / rho;
    const float ux = fl * jx * inv_rho, uy = fl * jy * inv_rho, uz = fl * jz * inv_rho;
    const float tt = fmaxf(0.0f, (static_cast<float>(x) - a.s0) * a.s_inv);
    const float ks = a.s_k * tt * tt;
    const float sigma = fmaxf(h, ks);
    const float rho_s = rho + sigma * (1.0f - rho);
    const float usx = ux + sigma * (a.u0 * fl - ux), usy = (1.0f - sigma) * uy,
                usz = (1.0f - sigma) * uz;
    const float K = (1.0f - sigma) * (1.0f - a.omega);

    // ---- transposed Jacobian: moments of Lam ----
    // e_q / rho and es_q / rho_s (the equilibrium polynomials), sums over Lam
    const float base = 1.0f - 1.5f * (ux * ux + uy * uy + uz * uz);
    const float base_s = 1.0f - 1.5f * (usx * usx + usy * usy + usz * usz);
    float Arho_e = 0.0f, Arho_s = 0.0f, dge = 0.0f;  // sum Lam e/rho, sum Lam es/rho_s, sum Lam (g - e)
    float S0 = 0.0f, S1x = 0.0f, S1y = 0.0f, S1z = 0.0f;
    float Sxx = 0.0f, Syy = 0.0f, Szz = 0.0f, Sxy = 0.0f, Sxz = 0.0f, Syz = 0.0f;
    static_for<0, kQ>([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      constexpr float cx = kCX[q], cy = kCY[q], cz = kCZ[q];
      const float cu = cx * ux + cy * uy + cz * uz;
      const float cus = cx * usx + cy * usy + cz * usz;
      const float pe = kW[q] * (base + 3.0f * cu + 4.5f * cu * cu);
      const float ps = kW[q] * (base_s + 3.0f * cus + 4.5f * cus * cus);
      const float lq = L[q];
      Arho_e += lq * pe;
      Arho_s += lq * ps;
      dge += lq * (g[q] - rho * pe);
      const float lw = lq * kW[q];
      S0 += lw;
      if constexpr (kCX[q] != 0) S1x += cx * lw;
      if constexpr (kCY[q] != 0) S1y += cy * lw;
      if constexpr (kCZ[q] != 0) S1z += cz * lw;
      if constexpr (kCX[q] != 0) Sxx += lw;
      if constexpr (kCY[q] != 0) Syy += lw;
      if constexpr (kCZ[q] != 0) Szz += lw;
      if constexpr (kCX[q] * kCY[q] != 0) Sxy += cx * cy * lw;
      if constexpr (kCX[q] * kCZ[q] != 0) Sxz += cx * cz * lw;
      if constexpr (kCY[q] * kCZ[q] != 0) Syz += cy * cz * lw;
    });
    // P_a(v) = 3 S1_a + 9 (S2 v)_a - 3 v_a S0
    auto P = [&](float vx, float vy, float vz, float& px, float& py, float& pz) {
      px = 3.0f * S1x + 9.0f * (Sxx * vx + Sxy * vy + Sxz * vz) - 3.0f * vx * S0;
      py = 3.0f * S1y + 9.0f * (Sxy * vx + Syy * vy + Syz * vz) - 3.0f * vy * S0;
      pz = 3.0f * S1z + 9.0f * (Sxz * vx + Syz * vy + Szz * vz) - 3.0f * vz * S0;
    };
    float Pux, Puy, Puz, Psx, Psy, Psz;
    P(ux, uy, uz, Pux, Puy, Puz);
    P(usx, usy, usz, Psx, Psy, Psz);
    const float Arho = (1.0f - sigma) * Arho_s - K * Arho_e;
    const float Ax = (1.0f - sigma) * rho_s * Psx - K * rho * Pux;
    const float Ay = (1.0f - sigma) * rho_s * Psy - K * rho * Puy;
    const float Az = (1.0f - sigma) * rho_s * Psz - K * rho * Puz;
    static_for<0, kQ>([&](auto qc) {
      constexpr int r = decltype(qc)::value;
      const float m = K * L[r] + Arho +
                      (Ax * (fl * kCX[r] - ux) + Ay * (fl * kCY[r] - uy) + Az * (fl * kCZ[r] - uz)) *
                          inv_rho;
      a.mu_out[static_cast<uint64_t>(r) * a.ns + idx] = m;
    });

    // ---- (a): step t through h = phi(x) ----
    if (in_block) {
      // via u = fl j / rho
      float dh = -(Ax * jx + Ay * jy + Az * jz) * inv_rho;
      // via the explicit fl of u_s,x = u_x + sigma (u0 fl - u_x): d(Lam.es)/du_s,x = rho_s Psx
      dh -= rho_s * Psx * sigma * a.u0;
      // via sigma = max(h, k t^2), when h is the larger
      if (h > ks) {
        const float Brho = Arho_s;  // sum Lam es / rho_s
        dh += Brho * (1.0f - rho) + rho_s * (Psx * (a.u0 * fl - ux) - Psy * uy - Psz * uz) -
              (1.0f - a.omega) * dge;
      }
      // via the drag: D = -2 fl sum w b c_x
      dh += 2.0f * a.lam_t * dragw;
      const uint32_t bi = (static_cast<uint32_t>(bz) * a.bn[1] + by) * a.bn[0] + bx;
      a.gphi[bi] += gsum + static_cast<double>(dh);
    }
  }
}

}  // namespace tlbm
