/*
 * CPU oracle, plain C (FP64, OpenMP over z): a second restatement of the reference step for
 * runs too long for NumPy (thousands of steps at 10^5 - 10^7 cells).
 *
 * TEST INFRASTRUCTURE.  Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline, through oracle/c_oracle.py, as the checker; the product never links it.
 *
 * It follows lbm.py line by line (cited per block) with the same conventions as
 * oracle/lbm_oracle.py, against which tests/test_oracle.py pins it (agreement to ~1e-15):
 *   - layout f[q*n + idx], idx = (z*ny + y)*nx + x (lbm.py:386-388, 426, 456);
 *   - obstacle tests in float32 with one rounding per operation (build with -ffp-contract=off);
 *   - the wall-velocity term only on wall links (the reference multiplies it by wall = 0
 *     elsewhere, which differs only when it is not finite);
 *   - the moving lid (side 2) of the build: f_q = f_opp + 6 w_q c_q.u_lid on ceiling links.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* lbm.py:15-18 */
static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static double W[19];
static int OPP[19], MIRY[19];

typedef struct {
  int32_t nx, ny, nz;
  int32_t collision; /* 0 bgk, 1 trt, 2 les */
  int32_t obstacle;  /* 0 cylinder, 1 sphere, 2 box */
  int32_t side;      /* 0 periodic, 1 slip, 2 lid */
  double omega, tau0, cs, magic;
  double obs[6];
  double u0, sponge_start, sponge_strength;
  double lid[3];
  double omega_rot;
  const float* phi; /* obstacle 3: per-cell solid fraction */
} oracle_params;

/* lbm.py:28-35 by search */
static void tables(void) {
  static int done = 0;
  if (done) return;
  for (int q = 0; q < 19; ++q) {
    W[q] = q == 0 ? 1.0 / 3 : (q < 7 ? 1.0 / 18 : 1.0 / 36);
    for (int p = 0; p < 19; ++p) {
      if (CX[p] == -CX[q] && CY[p] == -CY[q] && CZ[p] == -CZ[q]) OPP[q] = p;
      if (CX[p] == CX[q] && CY[p] == -CY[q] && CZ[p] == CZ[q]) MIRY[q] = p;
    }
  }
  done = 1;
}

/* lbm.py:52-55 */
static double feq(int q, double rho, double ux, double uy, double uz) {
  const double cu = 3.0 * (CX[q] * ux + CY[q] * uy + CZ[q] * uz);
  return W[q] * rho * (1.0 + cu + 0.5 * cu * cu - 1.5 * (ux * ux + uy * uy + uz * uz));
}

/* lbm.py:150-153, 178-182, 208-211: float32, one rounding per operation */
static double solid(const oracle_params* p, int x, int y, int z) {
  const float fx = (float)x, fy = (float)y, fz = (float)z;
  if (p->obstacle == 3) return p->phi[((int64_t)z * p->ny + y) * p->nx + x];
  if (p->obstacle == 0) {
    const float dx = fx - (float)p->obs[0], dy = fy - (float)p->obs[1];
    const float r = (float)p->obs[2];
    const float a = dx * dx, b = dy * dy, rr = r * r;
    const float s = a + b;
    return s < rr ? 1.0 : 0.0;
  }
  if (p->obstacle == 1) {
    const float dx = fx - (float)p->obs[0], dy = fy - (float)p->obs[1], dz = fz - (float)p->obs[2];
    const float r = (float)p->obs[3];
    const float a = dx * dx, b = dy * dy, c = dz * dz, rr = r * r;
    const float s = (a + b) + c;
    return s < rr ? 1.0 : 0.0;
  }
  {
    const float ax = fabsf(fx - (float)p->obs[0]) - (float)p->obs[3];
    const float ay = fabsf(fy - (float)p->obs[1]) - (float)p->obs[4];
    const float az = fabsf(fz - (float)p->obs[2]) - (float)p->obs[5];
    const float d = fmaxf(fmaxf(ax, ay), az);
    return d < 0.0f ? 1.0 : 0.0;
  }
}

/* lbm.py:159-169 (cylinder tangent); sphere and box walls rest */
static void wall_vel(const oracle_params* p, double xm, double ym, double zm, double* u) {
  (void)zm;
  if (p->obstacle == 0) {
    const double xc = (float)p->obs[0], yc = (float)p->obs[1], r = (float)p->obs[2];
    u[0] = -(ym - yc) / r;
    u[1] = (xm - xc) / r;
    u[2] = 0.0;
  } else {
    u[0] = u[1] = u[2] = 0.0;
  }
}

/* lbm.py:72-93 (e[q] = feq(q, rho, u)) */
static double smagorinsky_tau(const double* f, const double* e, double rho, double tau0,
                              double cs) {
  double pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
  for (int q = 0; q < 19; ++q) {
    const double neq = f[q] - e[q];
    const double ex = CX[q], ey = CY[q], ez = CZ[q];
    pxx += ex * ex * neq;
    pyy += ey * ey * neq;
    pzz += ez * ez * neq;
    pxy += ex * ey * neq;
    pxz += ex * ez * neq;
    pyz += ey * ez * neq;
  }
  const double pi = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + 2.0 * (pxy * pxy + pxz * pxz + pyz * pyz));
  return 0.5 * (tau0 + sqrt(tau0 * tau0 + 25.456 * cs * cs * pi / rho));
}

/* lbm.py:60-139.  e[q] = feq(q, rho, u): the reference evaluates the same function on the same
 * arguments each time it needs it, so evaluating it once gives identical values. */
static void collide(const oracle_params* p, const double* f, const double* e, double* post,
                    double rho) {
  if (p->collision == 0) {
    for (int q = 0; q < 19; ++q) post[q] = f[q] - p->omega * (f[q] - e[q]);
  } else if (p->collision == 1) {
    const double tau = smagorinsky_tau(f, e, rho, p->tau0, p->cs);
    const double om_p = 1.0 / tau;
    const double om_m = 1.0 / (p->magic / (tau - 0.5) + 0.5);
    for (int q = 0; q < 19; ++q) {
      const int o = OPP[q];
      const double eq = e[q], eo = e[o];
      post[q] = f[q] - om_p * (0.5 * (f[q] + f[o]) - 0.5 * (eq + eo)) -
                om_m * (0.5 * (f[q] - f[o]) - 0.5 * (eq - eo));
    }
  } else {
    const double tau = smagorinsky_tau(f, e, rho, p->tau0, p->cs);
    const double om = 1.0 / tau;
    for (int q = 0; q < 19; ++q) post[q] = f[q] - om * (f[q] - e[q]);
  }
}

/* lbm.py:397-456 (+ 493-513 when fields != NULL: 13 arrays of n, FIELDS order of lbm.py:545;
 * rho_out, when not NULL, receives the streamed density of every cell). */
int oracle_step(const oracle_params* p, const double* fin, double* fout, double* fields,
                double* rho_out) {
  tables();
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const int64_t n = (int64_t)nx * ny * nz;
  /* solid() of every cell, evaluated once (the reference evaluates it per use) */
  double* mask = (double*)malloc((size_t)n * sizeof(double));
  if (!mask) return 1;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) mask[((int64_t)z * ny + y) * nx + x] = solid(p, x, y, z);
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t idx = ((int64_t)z * ny + y) * nx + x;
        const double here = mask[idx];
        const double fluid = 1.0 - here;
        double f[19], post[19];
        double mx = 0, my = 0, mz = 0;
        for (int q = 0; q < 19; ++q) {
          const int o = OPP[q];
          /* (x - c + n) % n without the division: c is -1, 0 or 1 */
          int xs = x - CX[q];
          xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
          int zs = z - CZ[q];
          zs = zs < 0 ? zs + nz : (zs >= nz ? zs - nz : zs);
          int ys = y - CY[q], qs = q, lid = 0;
          if (p->side == 0) {
            ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
          } else if (ys < 0 || ys >= ny) {
            if (p->side == 2 && ys >= ny) lid = 1;
            else qs = MIRY[q];
            ys = y;
          }
          const double w6 = 6.0 * W[q];
          const double bounced = fin[(int64_t)o * n + idx];
          if (lid) {
            const double cdl = CX[q] * p->lid[0] + CY[q] * p->lid[1] + CZ[q] * p->lid[2];
            f[q] = bounced + w6 * cdl;
            continue;
          }
          const double wall = mask[((int64_t)zs * ny + ys) * nx + xs];
          double cdu = 0.0;
          if (wall > 0) {
            double uw[3];
            wall_vel(p, x - 0.5 * CX[q], y - 0.5 * CY[q], z - 0.5 * CZ[q], uw);
            cdu = CX[q] * (p->omega_rot * uw[0]) + CY[q] * (p->omega_rot * uw[1]) +
                  CZ[q] * (p->omega_rot * uw[2]);
          }
          const double pulled = fin[(int64_t)qs * n + ((int64_t)zs * ny + ys) * nx + xs];
          f[q] = pulled + wall * (bounced - pulled + w6 * cdu);
          const double hit = wall * fluid * (2.0 * bounced + w6 * cdu);
          mx += hit * CX[o];
          my += hit * CY[o];
          mz += hit * CZ[o];
        }
        double rho = 0, jx = 0, jy = 0, jz = 0;
        for (int q = 0; q < 19; ++q) {
          rho += f[q];
          jx += f[q] * CX[q];
          jy += f[q] * CY[q];
          jz += f[q] * CZ[q];
        }
        const double ux = fluid * jx / rho, uy = fluid * jy / rho, uz = fluid * jz / rho;
        double e[19];
        for (int q = 0; q < 19; ++q) e[q] = feq(q, rho, ux, uy, uz);
        collide(p, f, e, post, rho);
        const double t = fmax(0.0, ((double)x - p->sponge_start) / ((double)nx - p->sponge_start));
        const double sigma = fmax(here, p->sponge_strength * t * t);
        const double rho_s = rho + sigma * (1.0 - rho);
        const double ux_s = ux + sigma * (p->u0 * fluid - ux);
        const double uy_s = (1.0 - sigma) * uy, uz_s = (1.0 - sigma) * uz;
        /* sigma = 0 gives rho_s = rho, u_s = u exactly, so feq(rho_s, u_s) = e bit for bit */
        for (int q = 0; q < 19; ++q) {
          const double neq = post[q] - e[q];
          const double es = sigma != 0.0 ? feq(q, rho_s, ux_s, uy_s, uz_s) : e[q];
          fout[(int64_t)q * n + idx] = es + (1.0 - sigma) * neq;
        }
        if (rho_out) rho_out[idx] = rho;
        if (fields) {
          double* F[13];
          for (int k = 0; k < 13; ++k) F[k] = fields + (int64_t)k * n;
          const double pr = (rho - 1.0) / 3.0;
          const double ke = 0.5 * rho * (ux * ux + uy * uy + uz * uz);
          const double ox = F[0][idx], oy = F[1][idx], oz = F[2][idx], op = F[3][idx];
          const double ok = 0.5 * (ox * ox + oy * oy + oz * oz);
          F[5][idx] = (3.0 * (pr - op)) * (3.0 * (pr - op));
          F[6][idx] = (ux - ox) * (ux - ox);
          F[7][idx] = (uy - oy) * (uy - oy);
          F[8][idx] = (uz - oz) * (uz - oz);
          F[9][idx] = (ke - ok) * (ke - ok);
          F[0][idx] = ux;
          F[1][idx] = uy;
          F[2][idx] = uz;
          F[3][idx] = pr;
          F[4][idx] = sqrt(ux * ux + uy * uy + uz * uz);
          F[10][idx] = mx;
          F[11][idx] = my;
          F[12][idx] = mz;
        }
      }
  free(mask);
  return 0;
}

/* splitmix64 of (seed, 3 idx + comp + 1), top 24 bits -> [0, 1) (the build's seeded noise) */
static double hash_uniform(uint64_t seed, uint64_t idx, uint32_t comp) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + (idx * 3ull + comp + 1ull) * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 40) * (1.0 / 16777216.0);
}

/* lbm.py:383-394 plus the seeded velocity noise on fluid cells */
int oracle_init(const oracle_params* p, double* f, double seed_amp, double noise_amp,
                uint64_t noise_seed) {
  tables();
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const int64_t n = (int64_t)nx * ny * nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int64_t idx = ((int64_t)z * ny + y) * nx + x;
        const double fluid = 1.0 - solid(p, x, y, z);
        double ux = p->u0 * fluid;
        double uy = seed_amp * p->u0 * fluid * sin(6.2831853 * x / (double)nx) *
                    cos(6.2831853 * z / (double)nz);
        double uz = 0.0;
        if (noise_amp != 0.0) {
          ux += fluid * (noise_amp * (2.0 * hash_uniform(noise_seed, (uint64_t)idx, 0) - 1.0));
          uy += fluid * (noise_amp * (2.0 * hash_uniform(noise_seed, (uint64_t)idx, 1) - 1.0));
          uz += fluid * (noise_amp * (2.0 * hash_uniform(noise_seed, (uint64_t)idx, 2) - 1.0));
        }
        for (int q = 0; q < 19; ++q) f[(int64_t)q * n + idx] = feq(q, 1.0, ux, uy, uz);
      }
  return 0;
}
