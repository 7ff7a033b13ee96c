// Multilayer shallow-water model: fused RHS + RK-stage kernel, boundary
// conditions and the C ABI entry points (somaxhip_swm_*).
//
// Reference: somax/_src/models/swm/multilayer.py -- vector_field (:150-201),
// apply_boundary_conditions (:203-223) and the wall helpers (:383-410).
// Staggering: arrays of one shape (Ny, Nx); h[j,i] at the cell centre, u[j,i]
// on its east face, v[j,i] on its north face, X point [j,i] at its north-east
// corner.  Intermediate fields are evaluated wherever their stencil lies in
// the array; tendencies are written on the interior and are zero on the ghost
// ring, except wind (top layer) and bottom drag (bottom layer), which act on
// the full arrays.
//
// One kernel evaluates the RHS and the RK stage combination, one workgroup per
// tile and layer: it stages h, u, v of its layer (boundary condition applied on
// load, 1-cell halo) through LDS; each thread forms the cumulative pressure
// gradient sum_{l<=k} g'_l (h_l[e] - h_l[w]) of its points in registers, the
// layers above read from global memory in layer order, so the hydrostatic
// coupling costs no extra pass and the differences of h are exact in FP32 (the
// cumulative pressure itself is ~g*H, four orders of magnitude above its
// cell-to-cell change).

#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "somaxhip.h"
#include "common.h"
#include "kernels.h"

namespace somax {

constexpr int SWM_MAX_LAYERS = 8;

struct SwmArgs {
  const float* hs;  // stage state (raw; BC applied on load when apply_bc)
  const float* us;
  const float* vs;
  const float* hn;  // step start state (nullable: out = c1*s + c2*rhs)
  const float* un;
  const float* vn;
  float* ho;
  float* uo;
  float* vo;
  int Nx, Ny, nl;
  long plane;
  float gprime[SWM_MAX_LAYERS];
  float hoff[SWM_MAX_LAYERS];  // thickness = stored h + hoff[k] (rest thickness when h holds deviations)
  const float* f_row;     // Ny
  const float* wind_row;  // Ny
  float wind_coef;        // tau0 / H[0]
  float nu, kappa;
  float idx, idy, idx2, idy2;
  float c1, c2;
  int bc;        // SOMAXHIP_BC_*
  int apply_bc;  // 1: RHS of BC(stage state)
};

namespace {

constexpr int STX = 64;        // tile width
constexpr int SRW = 1;         // rows per thread
constexpr int SWV = 4;         // waves per workgroup (one tile row each; 2 measured slower at 512^2)
constexpr int SNT = 64 * SWV;  // threads per workgroup
constexpr int STY = SWV * SRW; // tile height
constexpr int SPW = STX + 2; // LDS pitch (halo 1)
constexpr int SPH = STY + 2;

enum { F_H = 0, F_U = 1, F_V = 2 };

// Value of field `f` at (gy, gx) after the boundary condition, from the raw
// array `p` (0 outside the array).  Sources are interior points, so the
// mapping reads nothing the condition itself would change.
__device__ inline float bc_value(const float* p, int f, int gy, int gx, int Ny, int Nx, int bc, int apply) {
  if (gy < 0 || gy >= Ny || gx < 0 || gx >= Nx) return 0.0f;
  if (!apply) return p[(long)gy * Nx + gx];
  int sy = gy, sx = gx;
  if (bc == SOMAXHIP_BC_PERIODIC || bc == SOMAXHIP_BC_CHANNEL) {
    sx = gx == 0 ? Nx - 2 : (gx == Nx - 1 ? 1 : gx);
  }
  if (bc == SOMAXHIP_BC_PERIODIC) {
    sy = gy == 0 ? Ny - 2 : (gy == Ny - 1 ? 1 : gy);
  } else {
    // wall rules on the rows (wall and channel)
    if (f == F_V) {
      if (gy == 0 || gy == Ny - 2 || gy == Ny - 1) return 0.0f;
    } else {
      sy = gy == 0 ? 1 : (gy == Ny - 1 ? Ny - 2 : gy);
    }
    if (bc == SOMAXHIP_BC_WALL) {
      if (f == F_U) {
        if (gx == 0 || gx == Nx - 2 || gx == Nx - 1) return 0.0f;
      } else {
        sx = gx == 0 ? 1 : (gx == Nx - 1 ? Nx - 2 : gx);
      }
    }
  }
  return p[(long)sy * Nx + sx];
}


// Source index of row gy (or column gx) for each field after the boundary
// condition; -1 where the condition sets the value to 0 or outside the array.
// The conditions are separable: a row rule and a column rule per field.
struct BcAxis {
  int h, u, v;
};
__device__ inline BcAxis bc_rows(int gy, int Ny, int bc, int apply) {
  if (gy < 0 || gy >= Ny) return {-1, -1, -1};
  if (!apply) return {gy, gy, gy};
  if (bc == SOMAXHIP_BC_PERIODIC) {
    const int s = gy == 0 ? Ny - 2 : (gy == Ny - 1 ? 1 : gy);
    return {s, s, s};
  }
  const int s = gy == 0 ? 1 : (gy == Ny - 1 ? Ny - 2 : gy);  // zero gradient / free slip
  const int sv = (gy == 0 || gy == Ny - 2 || gy == Ny - 1) ? -1 : gy;  // no normal flow
  return {s, s, sv};
}
__device__ inline BcAxis bc_cols(int gx, int Nx, int bc, int apply) {
  if (gx < 0 || gx >= Nx) return {-1, -1, -1};
  if (!apply) return {gx, gx, gx};
  if (bc == SOMAXHIP_BC_PERIODIC || bc == SOMAXHIP_BC_CHANNEL) {
    const int s = gx == 0 ? Nx - 2 : (gx == Nx - 1 ? 1 : gx);
    return {s, s, s};
  }
  const int s = gx == 0 ? 1 : (gx == Nx - 1 ? Nx - 2 : gx);
  const int su = (gx == 0 || gx == Nx - 2 || gx == Nx - 1) ? -1 : gx;
  return {s, su, s};
}

// Linear block -> tile, each XCD taking a contiguous run of tiles in
// row-major order (vertically adjacent tiles share their halo rows in L2).
__device__ inline void swm_tile(int tiles_x, int& ty, int& tx) {
  const int T = gridDim.x;
  const int b = blockIdx.x;
  const int xcd = b & 7, slot = b >> 3;
  const int per = T >> 3, extra = T & 7;
  const int t = xcd * per + (xcd < extra ? xcd : extra) + slot;
  ty = t / tiles_x;
  tx = t - ty * tiles_x;
}

}  // namespace

// PV (zeta + f_X) / h_X at X point (gy, gx) from the global arrays of one
// layer, each value read through the boundary condition when `apply` (the
// same mapping as the tile loads).  Used only for the X points past the last
// row / column that the energy-enstrophy flux needs (their periodic images).
__device__ float pv_global(const SwmArgs& a, const float* h, const float* u, const float* v, float hk, int gy,
                           int gx) {
  const int Ny = a.Ny, Nx = a.Nx, bc = a.bc, ap = a.apply_bc;
  const float zeta = (bc_value(v, F_V, gy, gx + 1, Ny, Nx, bc, ap) - bc_value(v, F_V, gy, gx, Ny, Nx, bc, ap)) * a.idx -
                     (bc_value(u, F_U, gy + 1, gx, Ny, Nx, bc, ap) - bc_value(u, F_U, gy, gx, Ny, Nx, bc, ap)) * a.idy;
  const float hx = hk + 0.25f * (bc_value(h, F_H, gy, gx, Ny, Nx, bc, ap) + bc_value(h, F_H, gy, gx + 1, Ny, Nx, bc, ap) +
                                 bc_value(h, F_H, gy + 1, gx, Ny, Nx, bc, ap) +
                                 bc_value(h, F_H, gy + 1, gx + 1, Ny, Nx, bc, ap));
  return (zeta + 0.5f * (a.f_row[gy] + a.f_row[gy + 1])) * __builtin_amdgcn_rcpf(hx);
}

// EEN: the energy-enstrophy-conserving (triad) vorticity flux instead of the
// reference's average form.  Its stencil reaches the X points one row and
// one column further north / east, so the tile carries a 2-cell halo on those
// sides and the tile's PV is formed once into LDS.  X points past the last
// row / column take the value of their periodic image (row / column 1) in a
// periodic direction and 0 beside a wall.
template <bool EEN>
__global__ __launch_bounds__(SNT) void swm_stage_kernel(SwmArgs a) {
  constexpr int PW = EEN ? STX + 3 : SPW;  // LDS pitch
  constexpr int PH = EEN ? STY + 3 : SPH;  // LDS rows
  constexpr int QW = STX + 2;              // X points of the tile and its 1-point halo
  constexpr int QH = STY + 2;
  __shared__ float sh[PH * PW];
  __shared__ float su[PH * PW];
  __shared__ float sv[PH * PW];
  __shared__ float sfx[STY + 2];  // Coriolis at the corner rows y0-1 .. y0+STY
  __shared__ float sq[EEN ? QH * QW : 1];

  const int tid = threadIdx.x;
  const int tx = tid & 63;
  const int ty = tid >> 6;
  const int Nx = a.Nx, Ny = a.Ny;
  int tyi, txi;
  swm_tile((Nx + STX - 1) / STX, tyi, txi);
  const int x0 = txi * STX;
  const int y0 = tyi * STY;

  if (tid <= STY + 1) {
    const int gy = y0 - 1 + tid;
    sfx[tid] = (gy >= 0 && gy + 1 < Ny) ? 0.5f * (a.f_row[gy] + a.f_row[gy + 1]) : 0.0f;
  }
  // column maps of this lane's (at most two) load columns
  const BcAxis cm0 = bc_cols(x0 - 1 + tx, Nx, a.bc, a.apply_bc);
  const BcAxis cm1 = bc_cols(x0 - 1 + tx + 64, Nx, a.bc, a.apply_bc);

  // One workgroup per tile and layer k (grid y).  The pressure gradient of
  // layer k sums g'_j times the thickness differences of the layers j <= k at
  // this thread's points (U and V): those of the layers above come straight
  // from global memory through the boundary condition (the values the tile
  // loads would hold), added in layer order, so the sums equal a layer loop's.
  const int k = blockIdx.y;
  float dpx[SRW], dpy[SRW];
#pragma unroll
  for (int r = 0; r < SRW; ++r) {
    dpx[r] = dpy[r] = 0.0f;
    const int gy = y0 + ty + SWV * r, gx = x0 + tx;
    if (k == 0 || !(gy >= 1 && gy <= Ny - 2 && gx >= 1 && gx <= Nx - 2)) continue;
    const BcAxis r0 = bc_rows(gy, Ny, a.bc, a.apply_bc), r1 = bc_rows(gy + 1, Ny, a.bc, a.apply_bc);
    const BcAxis c0 = bc_cols(gx, Nx, a.bc, a.apply_bc), c1 = bc_cols(gx + 1, Nx, a.bc, a.apply_bc);
    for (int j = 0; j < k; ++j) {
      const float* hj = a.hs + (long)j * a.plane;
      const float h00 = (r0.h >= 0 && c0.h >= 0) ? hj[r0.h * Nx + c0.h] : 0.0f;
      const float h01 = (r0.h >= 0 && c1.h >= 0) ? hj[r0.h * Nx + c1.h] : 0.0f;
      const float h10 = (r1.h >= 0 && c0.h >= 0) ? hj[r1.h * Nx + c0.h] : 0.0f;
      dpx[r] += a.gprime[j] * (h01 - h00);
      dpy[r] += a.gprime[j] * (h10 - h00);
    }
  }

  {
    const long off = (long)k * a.plane;
    const float* hs = a.hs + off;
    const float* us = a.us + off;
    const float* vs = a.vs + off;
    // This thread's raw stage and step-start values, fetched before the tile
    // so that their latency overlaps the tile loads.
    float hs_r[SRW], us_r[SRW], vs_r[SRW], hn_r[SRW], un_r[SRW], vn_r[SRW];
#pragma unroll
    for (int r = 0; r < SRW; ++r) {
      const int gy = y0 + ty + SWV * r, gx = x0 + tx;
      const bool ok = gy < Ny && gx < Nx;
      const long idx = off + (gy * Nx + gx);
      hs_r[r] = ok ? a.hs[idx] : 0.0f;
      us_r[r] = ok ? a.us[idx] : 0.0f;
      vs_r[r] = ok ? a.vs[idx] : 0.0f;
      hn_r[r] = (ok && a.hn) ? a.hn[idx] : 0.0f;
      un_r[r] = (ok && a.hn) ? a.un[idx] : 0.0f;
      vn_r[r] = (ok && a.hn) ? a.vn[idx] : 0.0f;
    }
    // Tile loads: a fixed count per thread, all issued before the first LDS
    // store so that one memory latency covers them (in-plane offsets fit in
    // 32 bits: a plane is below 2^31 cells).
    constexpr int NR = (PH + SWV - 1) / SWV;
    float th[NR][2], tu[NR][2], tv[NR][2];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int r = ty + SWV * j;
      const BcAxis rm = bc_rows(y0 - 1 + r, Ny, a.bc, a.apply_bc);
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const bool in = r < PH && tx + 64 * half < PW;
        const BcAxis& cm = half ? cm1 : cm0;
        th[j][half] = (in && rm.h >= 0 && cm.h >= 0) ? hs[rm.h * Nx + cm.h] : 0.0f;
        tu[j][half] = (in && rm.u >= 0 && cm.u >= 0) ? us[rm.u * Nx + cm.u] : 0.0f;
        tv[j][half] = (in && rm.v >= 0 && cm.v >= 0) ? vs[rm.v * Nx + cm.v] : 0.0f;
      }
    }
#pragma unroll
    for (int j = 0; j < NR; ++j) {
      const int r = ty + SWV * j;
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int c = tx + 64 * half;
        if (r < PH && c < PW) {
          sh[r * PW + c] = th[j][half];
          su[r * PW + c] = tu[j][half];
          sv[r * PW + c] = tv[j][half];
        }
      }
    }
    __syncthreads();
    const float gp = a.gprime[k];
    const float hk = a.hoff[k];
    const bool top = k == 0, bottom = k == a.nl - 1;
    if (EEN) {
      // PV of the tile's X points (rows y0-1 .. y0+STY, columns x0-1 .. x0+STX);
      // X point (y0-1+xr, x0-1+xc) reads tile rows xr, xr+1 and columns xc, xc+1
      for (int i = tid; i < QH * QW; i += SNT) {
        const int xr = i / QW, xc = i - xr * QW;
        const int gy = y0 - 1 + xr, gx = x0 - 1 + xc;
        float qv = 0.0f;
        if (gy >= 0 && gx >= 0 && gy <= Ny - 2 && gx <= Nx - 2) {
          const int c = xr * PW + xc;
          const float zeta = (sv[c + 1] - sv[c]) * a.idx - (su[c + PW] - su[c]) * a.idy;
          const float hx = hk + 0.25f * (sh[c] + sh[c + 1] + sh[c + PW] + sh[c + PW + 1]);
          qv = (zeta + sfx[xr]) * __builtin_amdgcn_rcpf(hx);
        } else if (gy >= 0 && gx >= 0 && gy <= Ny - 1 && gx <= Nx - 1) {
          // past the last row / column: periodic image or 0 beside a wall
          const bool xper = a.bc != SOMAXHIP_BC_WALL, yper = a.bc == SOMAXHIP_BC_PERIODIC;
          int sy = gy, sx = gx;
          bool zero = false;
          if (gx == Nx - 1) {
            zero |= !xper;
            sx = 1;
          }
          if (gy == Ny - 1) {
            zero |= !yper;
            sy = 1;
          }
          if (!zero) qv = pv_global(a, hs, us, vs, hk, sy, sx);
        }
        sq[i] = qv;
      }
      __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < SRW; ++r) {
      const int ly = ty + SWV * r;
      const int gy = y0 + ly, gx = x0 + tx;
      if (gy >= Ny || gx >= Nx) continue;
      const int c = (ly + 1) * PW + tx + 1;  // LDS index of (gy, gx)
#define H(dj, di) sh[c + (dj) * PW + (di)]
#define U(dj, di) su[c + (dj) * PW + (di)]
#define V(dj, di) sv[c + (dj) * PW + (di)]
      const bool interior = gy >= 1 && gy <= Ny - 2 && gx >= 1 && gx <= Nx - 2;
      float dh = 0.0f, du = 0.0f, dv = 0.0f;
      if (interior) {
        dpx[r] += gp * (H(0, 1) - H(0, 0));
        dpy[r] += gp * (H(1, 0) - H(0, 0));
        // mass: upwind fluxes on the four faces of the cell
        const float fe = U(0, 0) * (hk + (U(0, 0) > 0.0f ? H(0, 0) : H(0, 1)));
        const float fw = U(0, -1) * (hk + (U(0, -1) > 0.0f ? H(0, -1) : H(0, 0)));
        const float gn = V(0, 0) * (hk + (V(0, 0) > 0.0f ? H(0, 0) : H(1, 0)));
        const float gs = V(-1, 0) * (hk + (V(-1, 0) > 0.0f ? H(-1, 0) : H(0, 0)));
        dh = -((fe - fw) * a.idx + (gn - gs) * a.idy);
        auto vh = [&](int dj, int di) { return (hk + 0.5f * (H(dj, di) + H(dj + 1, di))) * V(dj, di); };
        auto uh = [&](int dj, int di) { return (hk + 0.5f * (H(dj, di) + H(dj, di + 1))) * U(dj, di); };
        auto ke = [&](int dj, int di) {
          return 0.25f * (U(dj, di) * U(dj, di) + U(dj, di - 1) * U(dj, di - 1) + V(dj, di) * V(dj, di) +
                          V(dj - 1, di) * V(dj - 1, di));
        };
        float fu, fv;
        if (EEN) {
          // triads (three corners of a cell, /12) at the corner each U / V pair shares
          auto Q = [&](int dj, int di) { return sq[(ly + 1 + dj) * QW + tx + 1 + di]; };
          const float q00 = Q(0, 0), q0m = Q(0, -1), qm0 = Q(-1, 0), qmm = Q(-1, -1);
          const float q01 = Q(0, 1), qm1 = Q(-1, 1), q10 = Q(1, 0), q1m = Q(1, -1);
          const float t_ne = q0m + q00 + qm0;  // cell (gy, gx): NE triad
          const float t_se = q00 + qm0 + qmm;  //                SE triad
          const float t_nw = qmm + q0m + q00;  //                NW triad
          const float t_nw_e = qm0 + q00 + q01;  // cell (gy, gx+1): NW triad
          const float t_sw_e = qm1 + qm0 + q00;  //                  SW triad
          const float t_sw_n = q00 + q0m + q1m;  // cell (gy+1, gx): SW triad
          const float t_se_n = q10 + q00 + q0m;  //                  SE triad
          constexpr float r12 = 1.0f / 12.0f;
          fu = r12 * (t_ne * vh(0, 0) + t_nw_e * vh(0, 1) + t_se * vh(-1, 0) + t_sw_e * vh(-1, 1));
          fv = -r12 * (t_sw_n * uh(1, -1) + t_se_n * uh(1, 0) + t_nw * uh(0, -1) + t_ne * uh(0, 0));
        } else {
          // PV at X points (dj, di) relative to (gy, gx)
          auto q = [&](int dj, int di) {
            const float zeta = (V(dj, di + 1) - V(dj, di)) * a.idx - (U(dj + 1, di) - U(dj, di)) * a.idy;
            const float hx = hk + 0.25f * (H(dj, di) + H(dj, di + 1) + H(dj + 1, di) + H(dj + 1, di + 1));
            return (zeta + sfx[ly + 1 + dj]) * __builtin_amdgcn_rcpf(hx);
          };
          const float q00 = q(0, 0);
          const float qU = 0.5f * (q00 + q(-1, 0));
          const float vhU = 0.25f * (vh(0, 0) + vh(0, 1) + vh(-1, 0) + vh(-1, 1));
          const float qV = 0.5f * (q00 + q(0, -1));
          const float uhV = 0.25f * (uh(0, 0) + uh(0, -1) + uh(1, 0) + uh(1, -1));
          fu = qU * vhU;
          fv = -qV * uhV;
        }
        const float ke0 = ke(0, 0);
        du = fu - ((ke(0, 1) - ke0) + dpx[r]) * a.idx;
        dv = fv - ((ke(1, 0) - ke0) + dpy[r]) * a.idy;
        if (a.nu != 0.0f) {
          du += a.nu * ((U(0, 1) - 2.0f * U(0, 0) + U(0, -1)) * a.idx2 + (U(1, 0) - 2.0f * U(0, 0) + U(-1, 0)) * a.idy2);
          dv += a.nu * ((V(0, 1) - 2.0f * V(0, 0) + V(0, -1)) * a.idx2 + (V(1, 0) - 2.0f * V(0, 0) + V(-1, 0)) * a.idy2);
        }
      }
      if (top) du += a.wind_coef * a.wind_row[gy];
      if (bottom) {
        du -= a.kappa * U(0, 0);
        dv -= a.kappa * V(0, 0);
      }
#undef H
#undef U
#undef V
      const long idx = off + (long)gy * Nx + gx;
      if (a.hn) {
        const float hn = hn_r[r], un = un_r[r], vn = vn_r[r];
        a.ho[idx] = hn + a.c1 * (hs_r[r] - hn) + a.c2 * dh;
        a.uo[idx] = un + a.c1 * (us_r[r] - un) + a.c2 * du;
        a.vo[idx] = vn + a.c1 * (vs_r[r] - vn) + a.c2 * dv;
      } else {
        a.ho[idx] = a.c1 * hs_r[r] + a.c2 * dh;
        a.uo[idx] = a.c1 * us_r[r] + a.c2 * du;
        a.vo[idx] = a.c1 * vs_r[r] + a.c2 * dv;
      }
    }
  }
}

// out = BC(in) for the three fields of every layer (apply_boundary_conditions).
__global__ __launch_bounds__(256) void swm_bc_kernel(const float* h, const float* u, const float* v, float* ho,
                                                     float* uo, float* vo, int Ny, int Nx, long plane, int bc) {
  const int gx = blockIdx.x * 256 + threadIdx.x;
  const int gy = blockIdx.y;
  const long off = (long)blockIdx.z * plane;
  if (gx >= Nx) return;
  const long i = off + (long)gy * Nx + gx;
  ho[i] = bc_value(h + off, F_H, gy, gx, Ny, Nx, bc, 1);
  uo[i] = bc_value(u + off, F_U, gy, gx, Ny, Nx, bc, 1);
  vo[i] = bc_value(v + off, F_V, gy, gx, Ny, Nx, bc, 1);
}

// Diagnostics (multilayer.py:225-255): PV and relative vorticity at X points,
// kinetic energy at T points (each where its stencil lies in the array), and
// per layer the FP64 sums over the interior of KE + 0.5 g' h^2 (energy) and
// 0.5 q^2 h_X (potential enstrophy); partial sums per block, reduced later
// in a fixed order.
__global__ __launch_bounds__(256) void swm_diag_kernel(const float* h, const float* u, const float* v, float* qo,
                                                       float* zo, float* keo, int Nx, int Ny, long plane,
                                                       const float* f_row, SwmArgs g, double* partial) {
  __shared__ double sa[256], sb[256];
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int y = blockIdx.y;
  const int k = blockIdx.z;
  double e = 0.0, en = 0.0;
  if (x < Nx) {
    const long s = k * plane + (long)y * Nx + x;
    const float* H = h + s;
    const float* U = u + s;
    const float* V = v + s;
    float qv = 0.0f, zv = 0.0f, hx = 0.0f;
    if (y <= Ny - 2 && x <= Nx - 2) {
      zv = (V[1] - V[0]) * g.idx - (U[Nx] - U[0]) * g.idy;
      hx = 0.25f * (H[0] + H[1] + H[Nx] + H[Nx + 1]);
      qv = (zv + 0.5f * (f_row[y] + f_row[y + 1])) / hx;
    }
    float kv = 0.0f;
    if (y >= 1 && x >= 1) kv = 0.25f * (U[0] * U[0] + U[-1] * U[-1] + V[0] * V[0] + V[-Nx] * V[-Nx]);
    if (qo) qo[s] = qv;
    if (zo) zo[s] = zv;
    if (keo) keo[s] = kv;
    if (y >= 1 && y <= Ny - 2 && x >= 1 && x <= Nx - 2) {
      const double hv = H[0];
      e = (double)kv + 0.5 * (double)g.gprime[k] * hv * hv;
      en = 0.5 * (double)qv * (double)qv * (double)hx;
    }
  }
  const int t = threadIdx.x;
  sa[t] = e;
  sb[t] = en;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      sa[t] += sa[t + w];
      sb[t] += sb[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const long blk = (long)k * gridDim.y * gridDim.x + (long)y * gridDim.x + blockIdx.x;
    partial[2 * blk] = sa[0];
    partial[2 * blk + 1] = sb[0];
  }
}

// f[k] += sign * H[k] over every point of nl planes (total thickness <->
// deviation from the rest thickness; exact in FP32 while |h - H| <= H).
__global__ __launch_bounds__(256) void swm_shift_kernel(float* f, long plane, SwmArgs a, float sign) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= plane) return;
  f[k * plane + i] += sign * a.hoff[k];
}

}  // namespace somax

using namespace somax;

struct somaxhip_swm_plan_s {
  int nx = 0, ny = 0, nl = 0, Nx = 0, Ny = 0, bc = 0;
  int vflux = SOMAXHIP_VFLUX_AVERAGE;
  double dx = 0.0, dy = 0.0;
  std::vector<float> gprime;
  std::vector<float> Hk;
  double H0 = 1.0;
  float* f_row = nullptr;
  float* wind_row = nullptr;
  long plane = 0;
  size_t ws_bytes = 0;
  ~somaxhip_swm_plan_s() {
    if (f_row) (void)hipFree(f_row);
    if (wind_row) (void)hipFree(wind_row);
  }
};

namespace {

// Stage launches bracketed by event pairs; the pairs are read after the
// whole step so that no host synchronisation sits between launches.
struct SwmProf {
  somaxhip_kernel_stat* stats;
  int max_stats;
  int n;
  hipError_t err;
  std::vector<hipEvent_t> ev;  // e0, e1 per launch
  std::vector<double> bytes;   // algorithmic bytes per launch
};
thread_local SwmProf* g_swm_prof = nullptr;

void swm_stage(somaxhip_swm_plan p, const float* hs, const float* us, const float* vs, const float* hn,
               const float* un, const float* vn, float* ho, float* uo, float* vo, const somaxhip_swm_params* prm,
               float c1, float c2, int apply_bc, bool deviation, hipStream_t st) {
  SwmArgs a{};
  a.hs = hs;
  a.us = us;
  a.vs = vs;
  a.hn = hn;
  a.un = un;
  a.vn = vn;
  a.ho = ho;
  a.uo = uo;
  a.vo = vo;
  a.Nx = p->Nx;
  a.Ny = p->Ny;
  a.nl = p->nl;
  a.plane = p->plane;
  for (int k = 0; k < p->nl; ++k) {
    a.gprime[k] = p->gprime[k];
    a.hoff[k] = deviation ? p->Hk[k] : 0.0f;
  }
  a.f_row = p->f_row;
  a.wind_row = p->wind_row;
  a.wind_coef = (float)(prm->tau0 / p->H0);
  a.nu = (float)prm->nu;
  a.kappa = (float)prm->kappa;
  a.idx = (float)(1.0 / p->dx);
  a.idy = (float)(1.0 / p->dy);
  a.idx2 = (float)(1.0 / (p->dx * p->dx));
  a.idy2 = (float)(1.0 / (p->dy * p->dy));
  a.c1 = c1;
  a.c2 = c2;
  a.bc = p->bc;
  a.apply_bc = apply_bc;
  dim3 grid((unsigned)(((p->Nx + STX - 1) / STX) * ((p->Ny + STY - 1) / STY)), (unsigned)p->nl, 1);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (g_swm_prof) {
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, st);
  }
  if (p->vflux == SOMAXHIP_VFLUX_ENERGY_ENSTROPHY)
    hipLaunchKernelGGL(swm_stage_kernel<true>, grid, dim3(SNT), 0, st, a);
  else
    hipLaunchKernelGGL(swm_stage_kernel<false>, grid, dim3(SNT), 0, st, a);
  if (g_swm_prof) {
    (void)hipEventRecord(e1, st);
    SwmProf& pr = *g_swm_prof;
    pr.ev.push_back(e0);
    pr.ev.push_back(e1);
    // algorithmic bytes: read the stage state (3 fields), the step start
    // state when combined, write 3 fields; interior cells
    const double cells = (double)p->nl * p->nx * p->ny;
    pr.bytes.push_back(cells * 4.0 * 3.0 * (hn ? 3.0 : 2.0));
  }
}

int check_plan(somaxhip_swm_plan p, const char* who) {
  if (!p) {
    set_error(std::string(who) + ": null plan");
    return SOMAXHIP_ERR_ARG;
  }
  return SOMAXHIP_OK;
}

}  // namespace

extern "C" {

int somaxhip_swm_plan_create(const somaxhip_swm_desc* d, somaxhip_swm_plan* out) {
  if (!d || !out) {
    set_error("somaxhip_swm_plan_create: null argument");
    return SOMAXHIP_ERR_ARG;
  }
  *out = nullptr;
  if (d->nx < 2 || d->ny < 2 || d->nl < 1 || d->nl > SWM_MAX_LAYERS || !(d->dx > 0) || !(d->dy > 0)) {
    set_error("somaxhip_swm_plan_create: need nx, ny >= 2, 1 <= nl <= 8, dx, dy > 0");
    return SOMAXHIP_ERR_ARG;
  }
  if ((long)(d->nx + 2) * (d->ny + 2) > 0x7fffffffL) {
    set_error("somaxhip_swm_plan_create: a plane of (nx + 2) * (ny + 2) cells must be below 2^31");
    return SOMAXHIP_ERR_ARG;
  }
  if (!d->g_prime || !d->H || !d->f_row || !d->wind_row) {
    set_error("somaxhip_swm_plan_create: a required table pointer is null");
    return SOMAXHIP_ERR_ARG;
  }
  if (d->bc != SOMAXHIP_BC_PERIODIC && d->bc != SOMAXHIP_BC_WALL && d->bc != SOMAXHIP_BC_CHANNEL) {
    set_error("somaxhip_swm_plan_create: unknown boundary condition");
    return SOMAXHIP_ERR_ARG;
  }
  if (!(d->H[0] > 0)) {
    set_error("somaxhip_swm_plan_create: H[0] must be positive");
    return SOMAXHIP_ERR_ARG;
  }
  if (d->vorticity_flux != SOMAXHIP_VFLUX_AVERAGE && d->vorticity_flux != SOMAXHIP_VFLUX_ENERGY_ENSTROPHY) {
    set_error("somaxhip_swm_plan_create: unknown vorticity flux");
    return SOMAXHIP_ERR_ARG;
  }
  somaxhip_swm_plan p = new (std::nothrow) somaxhip_swm_plan_s();
  if (!p) {
    set_error("somaxhip_swm_plan_create: out of host memory");
    return SOMAXHIP_ERR_ALLOC;
  }
  p->nx = d->nx;
  p->ny = d->ny;
  p->nl = d->nl;
  p->Nx = d->nx + 2;
  p->Ny = d->ny + 2;
  p->dx = d->dx;
  p->dy = d->dy;
  p->bc = d->bc;
  p->vflux = d->vorticity_flux;
  p->H0 = d->H[0];
  for (int k = 0; k < d->nl; ++k) {
    p->gprime.push_back((float)d->g_prime[k]);
    p->Hk.push_back((float)d->H[k]);
  }
  p->plane = (long)p->Ny * p->Nx;
  std::vector<float> f(p->Ny), w(p->Ny);
  for (int j = 0; j < p->Ny; ++j) {
    f[j] = (float)d->f_row[j];
    w[j] = (float)d->wind_row[j];
  }
  hipError_t e = hipMalloc(&p->f_row, p->Ny * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&p->wind_row, p->Ny * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(p->f_row, f.data(), p->Ny * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->wind_row, w.data(), p->Ny * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("somaxhip_swm_plan_create: ") + hipGetErrorString(e));
    delete p;
    return SOMAXHIP_ERR_ALLOC;
  }
  // workspace: thickness deviation of the step start and two stage states
  p->ws_bytes = (size_t)7 * p->nl * p->plane * sizeof(float);
  *out = p;
  return SOMAXHIP_OK;
}

int somaxhip_swm_plan_destroy(somaxhip_swm_plan p) {
  delete p;
  return SOMAXHIP_OK;
}

size_t somaxhip_swm_workspace_bytes(somaxhip_swm_plan p) { return p ? p->ws_bytes : 0; }

int somaxhip_swm_rhs(somaxhip_swm_plan p, const float* h, const float* u, const float* v, float* dh, float* du,
                     float* dv, const somaxhip_swm_params* prm, int apply_bc, somaxhip_stream_t stream) {
  int rc = check_plan(p, "somaxhip_swm_rhs");
  if (rc) return rc;
  if (!h || !u || !v || !dh || !du || !dv || !prm) {
    set_error("somaxhip_swm_rhs: null field or params");
    return SOMAXHIP_ERR_ARG;
  }
  swm_stage(p, h, u, v, nullptr, nullptr, nullptr, dh, du, dv, prm, 0.0f, 1.0f, apply_bc ? 1 : 0, false,
            reinterpret_cast<hipStream_t>(stream));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_swm_apply_bc(somaxhip_swm_plan p, const float* h, const float* u, const float* v, float* ho, float* uo,
                          float* vo, somaxhip_stream_t stream) {
  int rc = check_plan(p, "somaxhip_swm_apply_bc");
  if (rc) return rc;
  if (!h || !u || !v || !ho || !uo || !vo) {
    set_error("somaxhip_swm_apply_bc: null field");
    return SOMAXHIP_ERR_ARG;
  }
  if (ho == h || uo == u || vo == v) {
    set_error("somaxhip_swm_apply_bc: output must not alias the input");
    return SOMAXHIP_ERR_ARG;
  }
  dim3 grid((p->Nx + 255) / 256, p->Ny, p->nl);
  hipLaunchKernelGGL(swm_bc_kernel, grid, dim3(256), 0, reinterpret_cast<hipStream_t>(stream), h, u, v, ho, uo, vo,
                     p->Ny, p->Nx, p->plane, p->bc);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_swm_diagnose(somaxhip_swm_plan p, const float* h, const float* u, const float* v, float* q, float* zeta,
                          float* ke, double* energy_enstrophy, void* ws, somaxhip_stream_t stream) {
  int rc = check_plan(p, "somaxhip_swm_diagnose");
  if (rc) return rc;
  if (!h || !u || !v || !energy_enstrophy || !ws) {
    set_error("somaxhip_swm_diagnose: null field, output or workspace");
    return SOMAXHIP_ERR_ARG;
  }
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  SwmArgs g{};
  for (int k = 0; k < p->nl; ++k) g.gprime[k] = p->gprime[k];
  g.idx = (float)(1.0 / p->dx);
  g.idy = (float)(1.0 / p->dy);
  const int bx = (p->Nx + 255) / 256;
  double* partial = static_cast<double*>(ws);  // 2 * nl * Ny * bx doubles, well inside the workspace
  hipLaunchKernelGGL(swm_diag_kernel, dim3(bx, p->Ny, p->nl), dim3(256), 0, st, h, u, v, q, zeta, ke, p->Nx, p->Ny,
                     p->plane, p->f_row, g, partial);
  launch_pair_reduce(partial, bx * p->Ny, p->nl, p->dx * p->dy, energy_enstrophy, st);
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_swm_step(somaxhip_swm_plan p, float* h, float* u, float* v, int nsteps, double dt,
                      const somaxhip_swm_params* prm, void* ws, somaxhip_stream_t stream) {
  int rc = check_plan(p, "somaxhip_swm_step");
  if (rc) return rc;
  if (!h || !u || !v || !prm || !ws || nsteps < 0) {
    set_error("somaxhip_swm_step: null field, params or workspace, or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  if (nsteps == 0) return SOMAXHIP_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const long n = (long)p->nl * p->plane;
  float* w = static_cast<float*>(ws);
  float *e0 = w, *e1 = w + n, *u1 = w + 2 * n, *v1 = w + 3 * n, *e2 = w + 4 * n, *u2 = w + 5 * n, *v2 = w + 6 * n;
  // The thickness is carried as its deviation from the rest thickness during
  // the call: FP32 total thickness (~500 m) keeps only ~3e-5 m of the
  // interface signal (~1 m) whose differences drive the pressure gradient.
  SOMAX_HIP_CHECK(hipMemcpyAsync(e0, h, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SwmArgs sh{};
  for (int k = 0; k < p->nl; ++k) sh.hoff[k] = p->Hk[k];
  const dim3 sg((unsigned)((p->plane + 255) / 256), p->nl, 1);
  hipLaunchKernelGGL(swm_shift_kernel, sg, dim3(256), 0, st, e0, p->plane, sh, -1.0f);
  const float s = (float)dt;
  for (int k = 0; k < nsteps; ++k) {
    swm_stage(p, e0, u, v, nullptr, nullptr, nullptr, e1, u1, v1, prm, 1.0f, s, 1, true, st);
    swm_stage(p, e1, u1, v1, e0, u, v, e2, u2, v2, prm, 0.25f, 0.25f * s, 1, true, st);
    swm_stage(p, e2, u2, v2, e0, u, v, e0, u, v, prm, 2.0f / 3.0f, 2.0f / 3.0f * s, 1, true, st);
  }
  hipLaunchKernelGGL(swm_shift_kernel, sg, dim3(256), 0, st, e0, p->plane, sh, 1.0f);
  SOMAX_HIP_CHECK(hipMemcpyAsync(h, e0, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

size_t somaxhip_swm_tsit5_workspace_bytes(somaxhip_swm_plan p) {
  return p ? (size_t)9 * 3 * p->nl * p->plane * sizeof(float) : 0;
}

// Tsit5 (FSAL) on (h, u, v), as somaxhip_qg_step_tsit5; thickness carried as
// its deviation from the rest thickness.  tws: Y0 (step start), Y (stage
// state), K0..K6, each the three fields back to back.
int somaxhip_swm_step_tsit5(somaxhip_swm_plan p, float* h, float* u, float* v, int nsteps, double dt,
                            const somaxhip_swm_params* prm, void* tws, somaxhip_stream_t stream) {
  int rc = check_plan(p, "somaxhip_swm_step_tsit5");
  if (rc) return rc;
  if (!h || !u || !v || !prm || !tws || nsteps < 0) {
    set_error("somaxhip_swm_step_tsit5: null field, params or Tsit5 workspace, or negative nsteps");
    return SOMAXHIP_ERR_ARG;
  }
  if (nsteps == 0) return SOMAXHIP_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const long n = (long)p->nl * p->plane;
  float* Y0 = static_cast<float*>(tws);
  float* Y = Y0 + 3 * n;
  float* K[7];
  for (int i = 0; i < 7; ++i) K[i] = Y + (i + 1) * 3 * n;
  SOMAX_HIP_CHECK(hipMemcpyAsync(Y0, h, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipMemcpyAsync(Y0 + n, u, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipMemcpyAsync(Y0 + 2 * n, v, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SwmArgs sh{};
  for (int k = 0; k < p->nl; ++k) sh.hoff[k] = p->Hk[k];
  const dim3 sg((unsigned)((p->plane + 255) / 256), p->nl, 1);
  hipLaunchKernelGGL(swm_shift_kernel, sg, dim3(256), 0, st, Y0, p->plane, sh, -1.0f);
  auto f = [&](const float* y, float* k) {
    swm_stage(p, y, y + n, y + 2 * n, nullptr, nullptr, nullptr, k, k + n, k + 2 * n, prm, 0.0f, 1.0f, 1, true, st);
  };
  f(Y0, K[0]);
  for (int s = 0; s < nsteps; ++s) {
    for (int i = 1; i <= 6; ++i) {
      LinComb c{};
      c.y = Y0;
      c.nk = i;
      for (int j = 0; j < i; ++j) {
        c.k[j] = K[j];
        c.w[j] = dt * TSIT5_A[i][j];
      }
      c.n = 3 * n;
      c.out = i < 6 ? Y : Y0;
      launch_lincomb(c, st);
      f(c.out, K[i]);
    }
    std::swap(K[0], K[6]);
  }
  hipLaunchKernelGGL(swm_shift_kernel, sg, dim3(256), 0, st, Y0, p->plane, sh, 1.0f);
  SOMAX_HIP_CHECK(hipMemcpyAsync(h, Y0, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipMemcpyAsync(u, Y0 + n, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipMemcpyAsync(v, Y0 + 2 * n, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

size_t somaxhip_rqg_workspace_bytes(somaxhip_swm_plan p) {
  return p ? (size_t)11 * 3 * p->nl * p->plane * sizeof(float) : 0;
}

// ReparameterizedQG time stepping: y' = f(y) with f = RHS(P(BC(y))) (the
// projection inside apply_boundary_conditions, reparameterized.py:186-189),
// Tsit5 (FSAL) or SSP-RK3 in Butcher form, thickness carried as the
// deviation eta.  tws: Y0, Y, K0..K6, B (after the BC), Pj (projected).
int somaxhip_rqg_step(somaxhip_swm_plan p, somaxhip_qg_plan qp, int solver, double f0, const double* H,
                      const double* M, float* h, float* u, float* v, int nsteps, double dt,
                      const somaxhip_swm_params* prm, void* qg_ws, void* tws, somaxhip_stream_t stream) {
  int rc = check_plan(p, "somaxhip_rqg_step");
  if (rc) return rc;
  if (!qp || !H || !M || !h || !u || !v || !prm || !qg_ws || !tws || nsteps < 0 ||
      (solver != SOMAXHIP_SOLVER_SSPRK3 && solver != SOMAXHIP_SOLVER_TSIT5)) {
    set_error("somaxhip_rqg_step: null argument, negative nsteps or unknown solver");
    return SOMAXHIP_ERR_ARG;
  }
  if (qg_plan_f64(qp)) {
    set_error("somaxhip_rqg_step: the QG plan computes in FP64; the projection takes an FP32 / BF16 plan");
    return SOMAXHIP_ERR_ARG;
  }
  if (nsteps == 0) return SOMAXHIP_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const long n = (long)p->nl * p->plane;
  float* Y0 = static_cast<float*>(tws);
  float* Y = Y0 + 3 * n;
  float* K[7];
  for (int i = 0; i < 7; ++i) K[i] = Y + (i + 1) * 3 * n;
  float* B = Y + 8 * 3 * n;
  float* Pj = B + 3 * n;
  SOMAX_HIP_CHECK(hipMemcpyAsync(Y0, h, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipMemcpyAsync(Y0 + n, u, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipMemcpyAsync(Y0 + 2 * n, v, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SwmArgs sh{};
  for (int k = 0; k < p->nl; ++k) sh.hoff[k] = p->Hk[k];
  const dim3 sg((unsigned)((p->plane + 255) / 256), p->nl, 1);
  hipLaunchKernelGGL(swm_shift_kernel, sg, dim3(256), 0, st, Y0, p->plane, sh, -1.0f);
  int err = SOMAXHIP_OK;
  auto f = [&](const float* y, float* k) {
    hipLaunchKernelGGL(swm_bc_kernel, dim3((p->Nx + 255) / 256, p->Ny, p->nl), dim3(256), 0, st, y, y + n,
                       y + 2 * n, B, B + n, B + 2 * n, p->Ny, p->Nx, p->plane, p->bc);
    const int r = somaxhip_qg_project(qp, f0, H, M, B, B + n, B + 2 * n, 1, Pj, Pj + n, Pj + 2 * n, nullptr, qg_ws,
                                      stream);
    if (r && !err) err = r;
    swm_stage(p, Pj, Pj + n, Pj + 2 * n, nullptr, nullptr, nullptr, k, k + n, k + 2 * n, prm, 0.0f, 1.0f, 0, true,
              st);
  };
  auto comb = [&](const double* w, int nk, float* out) {
    LinComb c{};
    c.y = Y0;
    c.nk = nk;
    for (int j = 0; j < nk; ++j) {
      c.k[j] = K[j];
      c.w[j] = dt * w[j];
    }
    c.n = 3 * n;
    c.out = out;
    launch_lincomb(c, st);
  };
  if (solver == SOMAXHIP_SOLVER_TSIT5) {
    f(Y0, K[0]);
    for (int s = 0; s < nsteps && !err; ++s) {
      for (int i = 1; i <= 6; ++i) {
        comb(TSIT5_A[i], i, i < 6 ? Y : Y0);  // row 6: the new state, in place
        f(i < 6 ? Y : Y0, K[i]);
      }
      std::swap(K[0], K[6]);  // first same as last
    }
  } else {
    // SSP-RK3 in Butcher form: c = (0, 1, 1/2), b = (1/6, 1/6, 2/3)
    static const double a1[1] = {1.0}, a2[2] = {0.25, 0.25}, b3[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
    for (int s = 0; s < nsteps && !err; ++s) {
      f(Y0, K[0]);
      comb(a1, 1, Y);
      f(Y, K[1]);
      comb(a2, 2, Y);
      f(Y, K[2]);
      comb(b3, 3, Y0);
    }
  }
  if (err) return err;
  hipLaunchKernelGGL(swm_shift_kernel, sg, dim3(256), 0, st, Y0, p->plane, sh, 1.0f);
  SOMAX_HIP_CHECK(hipMemcpyAsync(h, Y0, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipMemcpyAsync(u, Y0 + n, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipMemcpyAsync(v, Y0 + 2 * n, n * sizeof(float), hipMemcpyDeviceToDevice, st));
  SOMAX_HIP_CHECK(hipGetLastError());
  return SOMAXHIP_OK;
}

int somaxhip_swm_step_profile(somaxhip_swm_plan p, float* h, float* u, float* v, int nsteps, double dt,
                              const somaxhip_swm_params* prm, void* ws, somaxhip_stream_t stream,
                              somaxhip_kernel_stat* stats, int max_stats, int* nstats) {
  if (!stats || !nstats || max_stats < 1) {
    set_error("somaxhip_swm_step_profile: null stats");
    return SOMAXHIP_ERR_ARG;
  }
  SwmProf prof{stats, max_stats, 0, hipSuccess, {}, {}};
  g_swm_prof = &prof;
  const int rc = somaxhip_swm_step(p, h, u, v, nsteps, dt, prm, ws, stream);
  g_swm_prof = nullptr;
  hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
  std::memset(&stats[0], 0, sizeof(stats[0]));
  std::strncpy(stats[0].name, "swm_stage", sizeof(stats[0].name) - 1);
  for (size_t i = 0; i < prof.bytes.size(); ++i) {
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, prof.ev[2 * i], prof.ev[2 * i + 1]);
    stats[0].launches += 1;
    stats[0].total_ms += ms;
    stats[0].bytes += prof.bytes[i];
  }
  for (hipEvent_t ev : prof.ev) (void)hipEventDestroy(ev);
  *nstats = 1;
  if (rc) return rc;
  SOMAX_HIP_CHECK(e);
  return SOMAXHIP_OK;
}

}  // extern "C"
