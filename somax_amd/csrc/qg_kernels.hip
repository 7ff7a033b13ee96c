// QG stencil, spectral and residual kernels (gfx950).
//
// qg_stage_kernel fuses one evaluation of the multilayer QG right-hand side
// (somax/_src/models/qg/baroclinic.py:161-190, with the boundary condition of
// :192-195 applied to the stage state on load) with the Runge-Kutta stage
// combination, so a stage reads q_s, psi and q_n once and writes q_{s+1} once.
// The tile (64 x 16 cells of one layer) is staged through LDS with a 2-cell
// halo of q (biharmonic) and a 1-cell halo of psi.

#include "common.h"
#include "kernels.h"

namespace somax {

namespace {
constexpr int TX = 64;
constexpr int TY_DEFAULT = 16;  // tile rows unless a batch chooses 8
constexpr int QW = TX + 4;  // q tile pitch (halo 2)
constexpr int PW = TX + 2;  // psi / lap tile pitch (halo 1)
}  // namespace

// Tile (bx, by) of plane z, of a grid of nbx x nby tiles per plane.
template <bool BIHARM, int TY>
__device__ __forceinline__ void qg_stage_arakawa(const StageArgs& a, int bx, int by, int z, int nbx, int nby) {
  __shared__ float sq[(TY + 4) * QW];
  __shared__ float sp[(TY + 2) * PW];
  __shared__ float sl[BIHARM ? (TY + 2) * PW : 1];
  __shared__ float sdb[TY + 4];

  const int tid = threadIdx.x;
  // Tiles cover the interior block (rows and columns 1 .. N-2); the edge
  // tiles also write the ghost ring (after the interior outputs).
  const int x0 = 1 + bx * TX;
  const int y0 = 1 + by * TY;
  const int Nx = a.Nx, Ny = a.Ny;
  const long off = (long)z * a.plane;
  const float* qs = a.qs + off;
  const float* psi = a.psi + off;
  const int m = a.t.member[z];
  const MemberParams pr = a.mp ? a.mp[m] : a.u;
  const float* dbeta_row = a.dbeta_rows + (long)m * Ny;
  const float* wind_row = a.wind_rows + (long)m * Ny;

  // This thread's outputs (column tx, rows tyb + 4 r): the previous stage
  // value q_n and the wind row are fetched first, so their latency overlaps
  // the tile loads instead of following the last barrier.
  const int tx = tid & 63;
  const int tyb = tid >> 6;
  float qn_r[TY / 4], wind_r[TY / 4];
#pragma unroll
  for (int r = 0; r < TY / 4; ++r) {
    const int gy = y0 + tyb + 4 * r, gx = x0 + tx;
    const bool ok = gy <= Ny - 2 && gx <= Nx - 2;
    const int idx = gy * Nx + gx;
    qn_r[r] = (ok && a.qn) ? a.qn[off + idx] : 0.0f;
    wind_r[r] = ok ? wind_row[gy] : 0.0f;
  }

  // Tile loads: a fixed number per thread, all issued before the first LDS
  // store, so one memory latency covers them.  q with the boundary condition
  // (ghost ring and outside the array read 0); psi with a 1-cell halo.
  constexpr int NQ = ((TY + 4) * QW + 255) / 256;
  constexpr int NPS = ((TY + 2) * PW + 255) / 256;
  float vq[NQ], vp[NPS];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int i = tid + 256 * k;
    const int ly = i / QW, lx = i - ly * QW;
    const int gy = y0 - 2 + ly, gx = x0 - 2 + lx;
    const bool inside = i < (TY + 4) * QW && gy >= 0 && gy < Ny && gx >= 0 && gx < Nx;
    const bool ghost = gy == 0 || gx == 0 || gy == Ny - 1 || gx == Nx - 1;
    vq[k] = (inside && !(ghost && a.apply_bc)) ? qs[gy * Nx + gx] : 0.0f;  // plane < 2^31 cells
  }
#pragma unroll
  for (int k = 0; k < NPS; ++k) {
    const int i = tid + 256 * k;
    const int ly = i / PW, lx = i - ly * PW;
    const int gy = y0 - 1 + ly, gx = x0 - 1 + lx;
    const bool inside = i < (TY + 2) * PW && gy >= 0 && gy < Ny && gx >= 0 && gx < Nx;
    vp[k] = inside ? psi[gy * Nx + gx] : 0.0f;
  }
  float vdb = 0.0f;
  if (tid < TY + 4) {
    const int gy = y0 - 2 + tid;
    vdb = (gy >= 0 && gy < Ny) ? dbeta_row[gy] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < NQ; ++k)
    if (tid + 256 * k < (TY + 4) * QW) sq[tid + 256 * k] = vq[k];
#pragma unroll
  for (int k = 0; k < NPS; ++k)
    if (tid + 256 * k < (TY + 2) * PW) sp[tid + 256 * k] = vp[k];
  if (tid < TY + 4) sdb[tid] = vdb;
  __syncthreads();

  if (BIHARM) {
    // inner Laplacian of q on the 1-halo region; its ghost ring is 0
    for (int i = tid; i < (TY + 2) * PW; i += 256) {
      const int ly = i / PW, lx = i - ly * PW;
      const int gy = y0 - 1 + ly, gx = x0 - 1 + lx;
      float v = 0.0f;
      if (gy >= 1 && gy <= Ny - 2 && gx >= 1 && gx <= Nx - 2) {
        const float* c = &sq[(ly + 1) * QW + lx + 1];
        v = (c[1] - 2.0f * c[0] + c[-1]) * a.idx2 + (c[QW] - 2.0f * c[0] + c[-QW]) * a.idy2;
      }
      sl[i] = v;
    }
    __syncthreads();
  }

  const float wc = pr.tau0 * a.t.wind_scale[z];
  const bool bottom = (a.t.layer[z] == a.t.mnl[m] - 1) && (pr.kappa != 0.0f);
  // Stage combination qn + c1 (qs - qn) + c2 rhs: the FP32 weights of the
  // Shu-Osher form (1/3 + 2/3 != 1 in FP32) would otherwise scale q by
  // 1 + 3e-8 per step.  Without qn: c1 qs + c2 rhs.
  auto combine = [&](float qn, float qraw, float rhs) {
    float o;
    if (a.qn) {
      o = qn + a.c1 * (qraw - qn) + a.c2 * rhs;
    } else {
      o = a.c2 * rhs;
      if (a.c1 != 0.0f) o += a.c1 * qraw;
    }
    return o;
  };
#pragma unroll
  for (int r = 0; r < TY / 4; ++r) {
    const int ly = tyb + 4 * r;
    const int gy = y0 + ly, gx = x0 + tx;
    if (gy > Ny - 2 || gx > Nx - 2) continue;
    float rhs = 0.0f;
    {
      const float* q = &sq[(ly + 2) * QW + tx + 2];
      const float* p = &sp[(ly + 1) * PW + tx + 1];
      // J(psi, q + b) = J(psi, q) + J(psi, b), b = beta*(y - y0) a function of
      // the row only.  Adding b to q first would cancel most of q's bits in
      // the differences; for b the Arakawa sum reduces exactly to
      //   2 (pE - pW)(bN - bS) + dN (pNE - pNW) + dS (pSE - pSW)
      // with the row increments dN = bN - bC, dS = bC - bS.
      const float dN = sdb[ly + 2], dS = sdb[ly + 1];
      const float tE = q[1], tW = q[-1], tN = q[QW], tS = q[-QW];
      const float tNE = q[QW + 1], tNW = q[QW - 1], tSE = q[-QW + 1], tSW = q[-QW - 1];
      const float pE = p[1], pW = p[-1], pN = p[PW], pS = p[-PW];
      const float pNE = p[PW + 1], pNW = p[PW - 1], pSE = p[-PW + 1], pSW = p[-PW - 1];
      const float jpp = (pE - pW) * (tN - tS) - (pN - pS) * (tE - tW);
      const float jpx = pE * (tNE - tSE) - pW * (tNW - tSW) - pN * (tNE - tNW) + pS * (tSE - tSW);
      const float jxp = tN * (pNE - pNW) - tS * (pSE - pSW) - tE * (pNE - pSE) + tW * (pNW - pSW);
      const float jb = 2.0f * (pE - pW) * (dN + dS) + dN * (pNE - pNW) + dS * (pSE - pSW);
      rhs = -((jpp + jpx + jxp) + jb) * a.iarakawa;
      if (pr.nu != 0.0f) rhs += pr.nu * ((q[1] - 2.0f * q[0] + q[-1]) * a.idx2 + (q[QW] - 2.0f * q[0] + q[-QW]) * a.idy2);
      if (BIHARM) {
        const float* l = &sl[(ly + 1) * PW + tx + 1];
        rhs -= pr.nu4 * ((l[1] - 2.0f * l[0] + l[-1]) * a.idx2 + (l[PW] - 2.0f * l[0] + l[-PW]) * a.idy2);
      }
      if (bottom) rhs -= pr.kappa * ((pE - 2.0f * p[0] + pW) * a.idx2 + (pN - 2.0f * p[0] + pS) * a.idy2);
    }
    rhs += wc * wind_r[r];
    const long idx = off + (long)gy * Nx + gx;
    a.out[idx] = combine(qn_r[r], sq[(ly + 2) * QW + tx + 2], rhs);
  }

  // Ghost ring: no Jacobian or viscous terms, only the wind row, and the raw
  // stage value (the tile holds 0 there under the boundary condition).
  // Wave 0: row 0 under the first tile row; wave 1: row Ny-1 under the last;
  // wave 2: column 0 (lanes 0..TY-1, first tile column) and column Nx-1
  // (lanes 32..32+TY-1, last tile column); wave 3: the four corners.
  const int wv = __builtin_amdgcn_readfirstlane(tyb);
  const int gx = x0 + tx;
  const bool first_x = bx == 0, last_x = bx == nbx - 1;
  const bool first_y = by == 0, last_y = by == nby - 1;
  int ry = -1, rx = -1;
  if (wv == 0 && first_y && gx <= Nx - 2) {
    ry = 0;
    rx = gx;
  } else if (wv == 1 && last_y && gx <= Nx - 2) {
    ry = Ny - 1;
    rx = gx;
  } else if (wv == 2 && tx < TY && first_x && y0 + tx <= Ny - 2) {
    ry = y0 + tx;
    rx = 0;
  } else if (wv == 2 && tx >= 32 && tx < 32 + TY && last_x && y0 + tx - 32 <= Ny - 2) {
    ry = y0 + tx - 32;
    rx = Nx - 1;
  } else if (wv == 3 && tx < 4) {
    const bool top = tx < 2, left = (tx & 1) == 0;
    if ((top ? first_y : last_y) && (left ? first_x : last_x)) {
      ry = top ? 0 : Ny - 1;
      rx = left ? 0 : Nx - 1;
    }
  }
  if (ry >= 0) {
    const long idx = off + (long)ry * Nx + rx;
    const float qn = a.qn ? a.qn[idx] : 0.0f;
    a.out[idx] = combine(qn, a.qs[idx], wc * wind_row[ry]);
  }
}

// ---------------------------------------------------------------------------
// WENO-Z flux-form advection stage (extension of the reference, which only
// has the Arakawa Jacobian)
//
// Face velocities come from psi at cell corners (4-point average inside the
// interior block, the wall value psi_wall on its boundary lines), so they are
// discretely divergence free and vanish through the walls.  Face values of
// total PV q + b(y) are upwind WENO-Z: 5th order where the five upwind-biased
// cells are interior, 3rd order where three are, else first-order upwind.
// The reconstruction works on differences: along x b is constant on a row
// (WENO commutes with adding a constant), along y the row increments of b are
// added to the q increments, and the b_row * div(u) term of the flux
// divergence is dropped because div(u) = 0 identically.
// ---------------------------------------------------------------------------
namespace {
constexpr int WX = 64;
constexpr int WY = 16;
constexpr int WQ = WX + 6;  // q tile pitch (halo 3)
constexpr float WENO_EPS = 1e-20f;

// Face value minus the upwind cell's value, from the four increments
// d1..d4 around upwind cell i (d2 = q_i - q_{i-1}, d3 = q_{i+1} - q_i).
__device__ inline float weno5_corr(float d1, float d2, float d3, float d4) {
  const float b0 = (13.0f / 12.0f) * (d2 - d1) * (d2 - d1) + 0.25f * (3.0f * d2 - d1) * (3.0f * d2 - d1);
  const float b1 = (13.0f / 12.0f) * (d3 - d2) * (d3 - d2) + 0.25f * (d2 + d3) * (d2 + d3);
  const float b2 = (13.0f / 12.0f) * (d4 - d3) * (d4 - d3) + 0.25f * (d4 - 3.0f * d3) * (d4 - 3.0f * d3);
  const float tau = fabsf(b0 - b2);
  // the nonlinear weights need no correctly rounded division: the hardware
  // reciprocal (1 ulp) moves a face value by about one FP32 ulp
  const float a0 = 0.1f * (1.0f + tau * __builtin_amdgcn_rcpf(b0 + WENO_EPS));
  const float a1 = 0.6f * (1.0f + tau * __builtin_amdgcn_rcpf(b1 + WENO_EPS));
  const float a2 = 0.3f * (1.0f + tau * __builtin_amdgcn_rcpf(b2 + WENO_EPS));
  const float c0 = (5.0f * d2 - 2.0f * d1) * (1.0f / 6.0f);
  const float c1 = (d2 + 2.0f * d3) * (1.0f / 6.0f);
  const float c2 = (4.0f * d3 - d4) * (1.0f / 6.0f);
  return (a0 * c0 + a1 * c1 + a2 * c2) * __builtin_amdgcn_rcpf(a0 + a1 + a2);
}

// 3rd-order WENO-Z correction from d2 = q_i - q_{i-1}, d3 = q_{i+1} - q_i.
__device__ inline float weno3_corr(float d2, float d3) {
  const float b0 = d2 * d2, b1 = d3 * d3;
  const float tau = fabsf(b0 - b1);
  const float a0 = (1.0f / 3.0f) * (1.0f + tau * __builtin_amdgcn_rcpf(b0 + WENO_EPS));
  const float a1 = (2.0f / 3.0f) * (1.0f + tau * __builtin_amdgcn_rcpf(b1 + WENO_EPS));
  return (a0 * (0.5f * d2) + a1 * (0.5f * d3)) * __builtin_amdgcn_rcpf(a0 + a1);
}

// A compile-time flag for the tile bodies: true where every face of the tile
// has its five upwind cells inside the interior block (no wall tests).
template <bool B>
struct Flag {
  static constexpr bool value = B;
};
// x, hidden from the optimiser: without the wall selects of the general tile
// path the compiler would fuse a helper's last operation into its caller's
// (an FMA) and round interior tiles differently from wall tiles.
__device__ inline float opaque(float x) {
  asm("" : "+v"(x));
  return x;
}

// Upwind face value minus the upwind cell value for the face between cells
// f and f+1 (1-based interior indices 1..n along the line).  dm[k] holds the
// increment q_{f+k} - q_{f+k-1} for k = -1..3 (dm index k+1), i.e. the
// increments from cell f-2 up to cell f+3.  vel > 0: upwind cell f.  INT: the
// caller guarantees 3 <= f and f + 3 <= n (the 5th-order stencil always fits).
template <bool INT>
__device__ inline float face_corr(int f, int n, float vel, const float* dm) {
  // the upwind side by selects rather than branches, so the lanes of a wave
  // with velocities of both signs share one reconstruction (mirror image
  // around cell f+1 when vel <= 0)
  const bool pos = vel > 0.0f;
  const int c = pos ? f : f + 1;  // upwind cell
  const float e1 = pos ? dm[0] : -dm[4];
  const float e2 = pos ? dm[1] : -dm[3];
  const float e3 = pos ? dm[2] : -dm[2];
  const float e4 = pos ? dm[3] : -dm[1];
  if (INT) return opaque(weno5_corr(e1, e2, e3, e4));
  if (c - 2 >= 1 && c + 2 <= n) return weno5_corr(e1, e2, e3, e4);
  if (c - 1 >= 1 && c + 1 <= n) return weno3_corr(e2, e3);
  return 0.0f;
}
}  // namespace

// Tile (bx, by) of plane z, of a grid of nbx x nby tiles per plane.
template <bool BIHARM>
__device__ __forceinline__ void qg_stage_weno(const StageArgs& a, int bx, int by, int z, int nbx, int nby) {
  constexpr int QH = WY + 6;    // q tile rows (halo 3)
  constexpr int PH = WY + 2;    // psi / Laplacian tile rows (halo 1)
  constexpr int PW2 = WX + 2;   // psi / Laplacian tile pitch
  __shared__ float sq[QH * WQ];                 // q (BC), halo 3
  __shared__ float sp[PH * PW2];                // psi, halo 1
  __shared__ float sfx[WY * (WX + 1)];          // x-face fluxes u*qface
  __shared__ float sfl[(WY + 1) * WX];          // y-face fluxes v*(qface - b of the lower cell)
  __shared__ float sfh[(WY + 1) * WX];          // the same relative to b of the upper cell
  __shared__ float sl[BIHARM ? PH * PW2 : 1];
  __shared__ float sdb[QH];                     // b[j+1] - b[j] for rows y0-3 ..

  const int tid = threadIdx.x;
  // Tiles cover the interior block (rows and columns 1 .. N-2); the edge
  // tiles also write the ghost ring (after the interior outputs).
  const int x0 = 1 + bx * WX;
  const int y0 = 1 + by * WY;
  const int Nx = a.Nx, Ny = a.Ny;
  const int nxi = Nx - 2, nyi = Ny - 2;
  const long off = (long)z * a.plane;
  const float* qs = a.qs + off;
  const float* psi = a.psi + off;
  const int m = a.t.member[z];
  const MemberParams pr = a.mp ? a.mp[m] : a.u;
  const float* dbeta_row = a.dbeta_rows + (long)m * Ny;
  const float* wind_row = a.wind_rows + (long)m * Ny;
  const float pw = a.t.psi_wall ? a.t.psi_wall[z] : 0.0f;
  // Lane tx owns column gx = x0 + tx of every tile; wave wv owns rows
  // wv, wv + 4, ...  Row indices and row tests are therefore wave-uniform
  // (scalar), and the only per-lane index is the column.
  const int tx = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int gx = x0 + tx;
  const bool gx_in = gx < Nx;
  const bool gx_out = gx <= Nx - 2;  // an interior output column
  const bool gx_ring = gx == Nx - 1;  // (gx >= 1)
  const bool bc = a.apply_bc != 0;

  // This thread's outputs (rows wv + 4 r): the step-start value and the wind
  // row are fetched first, so their latency overlaps the tile loads.
  float qn_r[WY / 4], wind_r[WY / 4];
#pragma unroll
  for (int r = 0; r < WY / 4; ++r) {
    const int gy = y0 + wv + 4 * r;
    const bool ok = gy <= Ny - 2 && gx_out;
    const int idx = gy * Nx + gx;  // plane < 2^31 cells
    qn_r[r] = (ok && a.qn) ? a.qn[off + idx] : 0.0f;
    wind_r[r] = ok ? wind_row[gy] : 0.0f;
  }
  // Tile loads, all issued before the first LDS store so that one memory
  // latency covers them.  q with the boundary condition (ghost ring and
  // outside the array read 0), halo 3: the lane's own column for the wave's
  // rows, then the six halo columns of every row by threads 0..6*QH-1.  psi
  // with halo 1 the same way (two halo columns).
  constexpr int NQM = (QH + 3) / 4;
  constexpr int NPM = (PH + 3) / 4;
  float vq[NQM], vp[NPM];
#pragma unroll
  for (int k = 0; k < NQM; ++k) {
    const int gy = y0 - 3 + wv + 4 * k;
    const bool row_ok = wv + 4 * k < QH && gy >= 0 && gy < Ny;
    const bool ghost = gy == 0 || gy == Ny - 1 || gx_ring;
    vq[k] = (row_ok && gx_in && !(ghost && bc)) ? qs[gy * Nx + gx] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < NPM; ++k) {
    const int gy = y0 - 1 + wv + 4 * k;
    const bool row_ok = wv + 4 * k < PH && gy >= 0 && gy < Ny;
    vp[k] = (row_ok && gx_in) ? psi[gy * Nx + gx] : 0.0f;
  }
  float vqe = 0.0f, vpe = 0.0f, vdb = 0.0f;
  int qe_at = -1, pe_at = -1;
  if (tid < 6 * QH) {
    const int ly = tid / 6, e = tid - 6 * ly;
    const int lx = e < 3 ? e : WX + e;  // tile columns 0..2 and WX+3..WX+5
    const int gy = y0 - 3 + ly, hx = x0 - 3 + lx;
    const bool inside = gy >= 0 && gy < Ny && hx >= 0 && hx < Nx;
    const bool ghost = gy == 0 || hx == 0 || gy == Ny - 1 || hx == Nx - 1;
    vqe = (inside && !(ghost && bc)) ? qs[gy * Nx + hx] : 0.0f;
    qe_at = ly * WQ + lx;
  }
  if (tid < 2 * PH) {
    const int ly = tid >> 1;
    const int lx = (tid & 1) ? WX + 1 : 0;
    const int gy = y0 - 1 + ly, hx = x0 - 1 + lx;
    const bool inside = gy >= 0 && gy < Ny && hx >= 0 && hx < Nx;
    vpe = inside ? psi[gy * Nx + hx] : 0.0f;
    pe_at = ly * PW2 + lx;
  }
  if (tid < QH) {
    const int gy = y0 - 3 + tid;
    vdb = (gy >= 0 && gy < Ny) ? dbeta_row[gy] : 0.0f;
  }
#pragma unroll
  for (int k = 0; k < NQM; ++k)
    if (wv + 4 * k < QH) sq[(wv + 4 * k) * WQ + tx + 3] = vq[k];
#pragma unroll
  for (int k = 0; k < NPM; ++k)
    if (wv + 4 * k < PH) sp[(wv + 4 * k) * PW2 + tx + 1] = vp[k];
  if (qe_at >= 0) sq[qe_at] = vqe;
  if (pe_at >= 0) sp[pe_at] = vpe;
  if (tid < QH) sdb[tid] = vdb;
  __syncthreads();

  // Every helper below takes a Flag: true for tiles away from the walls
  // (`interior` below), where all of their wall tests hold and are left out;
  // the arithmetic is the same either way.
  // psi at corner (cy, cx), between T rows cy, cy+1 and columns cx, cx+1
  // (cy in y0-1 .. y0+WY-1, cx in x0-1 .. x0+WX-1, always inside the tile):
  // the 4-point mean inside the interior block, the wall value on its
  // boundary lines
  auto corner = [&](auto in, int cy, int cx) {
    const float* p = &sp[(cy - y0 + 1) * PW2 + (cx - x0 + 1)];
    const float v = 0.25f * (p[0] + p[1] + p[PW2] + p[PW2 + 1]);
    if (decltype(in)::value) return opaque(v);
    return (cy >= 1 && cy <= Ny - 3 && cx >= 1 && cx <= Nx - 3) ? v : pw;
  };
  // inner Laplacian of q at tile position (ly, lx) of the 1-halo region
  // (gy = y0-1+ly, hx = x0-1+lx); its ghost ring is 0
  auto inner_lap = [&](auto in, int ly, int lx) {
    const int gy = y0 - 1 + ly, hx = x0 - 1 + lx;
    const float* c = &sq[(ly + 2) * WQ + lx + 2];
    const float v = (c[1] - 2.0f * c[0] + c[-1]) * a.idx2 + (c[WQ] - 2.0f * c[0] + c[-WQ]) * a.idy2;
    if (decltype(in)::value) return opaque(v);
    return (gy >= 1 && gy <= Ny - 2 && hx >= 1 && hx <= Nx - 2) ? v : 0.0f;
  };
  // flux through the x-face between T columns f and f+1 on row gy = y0+ly
  auto xface = [&](auto in, int ly, int f) {
    constexpr bool I = decltype(in)::value;
    const int gy = y0 + ly;
    float flux = 0.0f;
    if (I || (gy >= 1 && gy <= nyi && f >= 1 && f <= nxi - 1)) {
      const float u = -(corner(in, gy, f) - corner(in, gy - 1, f)) * a.idy;
      const float* qr = &sq[(ly + 3) * WQ + (f - x0 + 3)];  // cell f
      float dm[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) dm[k] = qr[k - 1] - qr[k - 2];  // q_{f+k-1} - q_{f+k-2}
      const float qup = u > 0.0f ? qr[0] : qr[1];
      flux = u * (qup + face_corr<I>(f, nxi, u, dm));
    }
    return I ? opaque(flux) : flux;
  };
  // y-face between T rows g and g+1 on column gx (g = y0-1 .. y0+WY-1):
  // v * (face value of q + b minus b of the lower cell g) into sfl, and the
  // same minus b of the upper cell into sfh
  auto yface = [&](auto in, int g) {
    constexpr bool I = decltype(in)::value;
    const float v = (corner(in, g, gx) - corner(in, g, gx - 1)) * a.idx;
    const int lq = g - y0 + 3;  // local q row of cell g
    float flux = 0.0f;
    if (I || (gx >= 1 && gx <= nxi && g >= 1 && g <= nyi - 1)) {
      float dm[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        // increment of q + b from cell g+k-2 to g+k-1
        const int r = lq + k - 1;
        dm[k] = (sq[r * WQ + tx + 3] - sq[(r - 1) * WQ + tx + 3]) + sdb[r - 1];
      }
      // face value of q+b minus b_g: upwind g -> q_g + corr ; upwind g+1 -> q_{g+1} + db_g + corr
      const float base = v > 0.0f ? sq[lq * WQ + tx + 3] : sq[(lq + 1) * WQ + tx + 3] + sdb[lq];
      flux = v * (base + face_corr<I>(g, nyi, v, dm));
      if (I) flux = opaque(flux);
    }
    const int lg = g - y0 + 1;
    sfl[lg * WX + tx] = flux;
    sfh[lg * WX + tx] = flux - v * sdb[lq];
  };

  auto faces = [&](auto in) {
    if (BIHARM) {
#pragma unroll
      for (int k = 0; k < NPM; ++k)
        if (wv + 4 * k < PH) sl[(wv + 4 * k) * PW2 + tx + 1] = inner_lap(in, wv + 4 * k, tx + 1);
      if (tid < 2 * PH) {
        const int ly = tid >> 1;
        const int lx = (tid & 1) ? WX + 1 : 0;
        sl[ly * PW2 + lx] = inner_lap(in, ly, lx);
      }
    }
    // faces of the lane's cells: east x-faces and north y-faces of rows
    // wv + 4 r; wave 0 adds the south y-faces of the tile's first row, wave 1
    // (16 lanes) the west x-faces of its first column
#pragma unroll
    for (int r = 0; r < WY / 4; ++r) {
      const int ly = wv + 4 * r;
      sfx[ly * (WX + 1) + tx + 1] = xface(in, ly, gx);
      yface(in, y0 + ly);
    }
    if (wv == 0) yface(in, y0 - 1);
    if (wv == 1 && tx < WY) sfx[tx * (WX + 1)] = xface(in, tx, x0 - 1);
  };
  // Away from the walls every face has its 5th-order stencil inside the
  // interior block: upwind cells reach x0-1 .. x0+WX (and y0-1 .. y0+WY), so
  // x0 - 3 >= 1 and x0 + WX + 2 <= nxi (likewise in y) cover every test.
  const bool interior = x0 >= 4 && x0 + WX + 2 <= nxi && y0 >= 4 && y0 + WY + 2 <= nyi;
  if (interior) faces(Flag<true>{});
  else faces(Flag<false>{});
  __syncthreads();

  const float wc = pr.tau0 * a.t.wind_scale[z];
  const bool bottom = (a.t.layer[z] == a.t.mnl[m] - 1) && (pr.kappa != 0.0f);
  // Stage combination qn + c1 (qs - qn) + c2 rhs (the FP32 weights of the
  // Shu-Osher form would otherwise not sum to 1); without qn: c1 qs + c2 rhs.
  auto combine = [&](float qn, float qraw, float rhs) {
    float o;
    if (a.qn) {
      o = qn + a.c1 * (qraw - qn) + a.c2 * rhs;
    } else {
      o = a.c2 * rhs;
      if (a.c1 != 0.0f) o += a.c1 * qraw;
    }
    return o;
  };
#pragma unroll
  for (int r = 0; r < WY / 4; ++r) {
    const int ly = wv + 4 * r;
    const int gy = y0 + ly;
    if (gy > Ny - 2 || !gx_out) continue;
    float rhs = 0.0f;
    {
      const float fe = sfx[ly * (WX + 1) + tx + 1], fw = sfx[ly * (WX + 1) + tx];
      // y-faces above/below cell gy, both relative to b of cell gy
      const float gn = sfl[(ly + 1) * WX + tx], gs = sfh[ly * WX + tx];
      rhs = -((fe - fw) * a.idx + (gn - gs) * a.idy);
      const float* q = &sq[(ly + 3) * WQ + tx + 3];
      const float* p = &sp[(ly + 1) * PW2 + tx + 1];
      if (pr.nu != 0.0f) rhs += pr.nu * ((q[1] - 2.0f * q[0] + q[-1]) * a.idx2 + (q[WQ] - 2.0f * q[0] + q[-WQ]) * a.idy2);
      if (BIHARM) {
        const float* l = &sl[(ly + 1) * PW2 + tx + 1];
        rhs -= pr.nu4 * ((l[1] - 2.0f * l[0] + l[-1]) * a.idx2 + (l[PW2] - 2.0f * l[0] + l[-PW2]) * a.idy2);
      }
      if (bottom) rhs -= pr.kappa * ((p[1] - 2.0f * p[0] + p[-1]) * a.idx2 + (p[PW2] - 2.0f * p[0] + p[-PW2]) * a.idy2);
    }
    rhs += wc * wind_r[r];
    const long idx = off + (long)gy * Nx + gx;
    a.out[idx] = combine(qn_r[r], sq[(ly + 3) * WQ + tx + 3], rhs);
  }

  // Ghost ring: no flux or viscous terms, only the wind row, and the raw
  // stage value (the tile holds 0 there under the boundary condition).
  // Wave 0: row 0 under the first tile row; wave 1: row Ny-1 under the last;
  // wave 2: column 0 (lanes 0..WY-1, first tile column) and column Nx-1
  // (lanes 32..32+WY-1, last tile column); wave 3: the four corners.
  const bool first_x = bx == 0, last_x = bx == nbx - 1;
  const bool first_y = by == 0, last_y = by == nby - 1;
  int ry = -1, rx = -1;
  if (wv == 0 && first_y && gx_out) {
    ry = 0;
    rx = gx;
  } else if (wv == 1 && last_y && gx_out) {
    ry = Ny - 1;
    rx = gx;
  } else if (wv == 2 && tx < WY && first_x && y0 + tx <= Ny - 2) {
    ry = y0 + tx;
    rx = 0;
  } else if (wv == 2 && tx >= 32 && tx < 32 + WY && last_x && y0 + tx - 32 <= Ny - 2) {
    ry = y0 + tx - 32;
    rx = Nx - 1;
  } else if (wv == 3 && tx < 4) {
    const bool top = tx < 2, left = (tx & 1) == 0;
    if ((top ? first_y : last_y) && (left ? first_x : last_x)) {
      ry = top ? 0 : Ny - 1;
      rx = left ? 0 : Nx - 1;
    }
  }
  if (ry >= 0) {
    const long idx = off + (long)ry * Nx + rx;
    const float qn = a.qn ? a.qn[idx] : 0.0f;
    a.out[idx] = combine(qn, a.qs[idx], wc * wind_row[ry]);
  }
}

template <bool BIHARM, int TY>
__global__ __launch_bounds__(256) void qg_stage_arakawa_kernel(StageArgs a) {
  int bx, by, z;
  xcd_tile3(bx, by, z);
  qg_stage_arakawa<BIHARM, TY>(a, bx, by, z, gridDim.x, gridDim.y);
}
template <bool BIHARM>
__global__ __launch_bounds__(256) void qg_stage_weno_kernel(StageArgs a) {
  int bx, by, z;
  xcd_tile3(bx, by, z);
  qg_stage_weno<BIHARM>(a, bx, by, z, gridDim.x, gridDim.y);
}

// Merged launch over the batches of a group: the tiles of all batches in one
// XCD-contiguous order (the order xcd_tile3 gives one batch).
struct StageGroup {
  BlockTable t;
  StageArgs a[MAX_GROUP];
};
template <bool BIHARM, int TY>
__global__ __launch_bounds__(256) void qg_stage_arakawa_group_kernel(StageGroup g) {
  int c, l, bx, by, z;
  if (!table_class(g.t, blockIdx.x, c, l)) return;
  class_block(g.t, c, xcd_order(l, g.t.count[c]), bx, by, z);
  qg_stage_arakawa<BIHARM, TY>(g.a[c], bx, by, z, g.t.gx[c], g.t.gy[c]);
}
template <bool BIHARM>
__global__ __launch_bounds__(256) void qg_stage_weno_group_kernel(StageGroup g) {
  int c, l, bx, by, z;
  if (!table_class(g.t, blockIdx.x, c, l)) return;
  class_block(g.t, c, xcd_order(l, g.t.count[c]), bx, by, z);
  qg_stage_weno<BIHARM>(g.a[c], bx, by, z, g.t.gx[c], g.t.gy[c]);
}

namespace {
dim3 stage_grid(const StageArgs& a) {
  const int tw = a.advection == 1 ? WX : TX;
  const int th = a.advection == 1 ? WY : (a.tile_rows ? a.tile_rows : TY_DEFAULT);
  return dim3((a.Nx - 2 + tw - 1) / tw, (a.Ny - 2 + th - 1) / th, a.P);  // interior tiles
}
}  // namespace

void launch_qg_stage(const StageArgs& a, hipStream_t st) {
  const dim3 g = stage_grid(a);
  if (a.advection == 1) {
    if (a.biharm) hipLaunchKernelGGL(qg_stage_weno_kernel<true>, g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(qg_stage_weno_kernel<false>, g, dim3(256), 0, st, a);
    return;
  }
  const int ty = a.tile_rows ? a.tile_rows : TY_DEFAULT;
  if (ty == 8) {
    if (a.biharm) hipLaunchKernelGGL((qg_stage_arakawa_kernel<true, 8>), g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((qg_stage_arakawa_kernel<false, 8>), g, dim3(256), 0, st, a);
  } else {
    if (a.biharm) hipLaunchKernelGGL((qg_stage_arakawa_kernel<true, 16>), g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((qg_stage_arakawa_kernel<false, 16>), g, dim3(256), 0, st, a);
  }
}

void launch_qg_stage_group(const StageArgs* a, int n, hipStream_t st) {
  StageGroup g{};
  for (int i = 0; i < n; ++i) {
    g.a[i] = a[i];
    const dim3 d = stage_grid(a[i]);
    table_add(g.t, (int)d.x, (int)d.y, (int)d.z);
  }
  const dim3 grid(g.t.start[g.t.n]);
  if (a[0].advection == 1) {
    if (a[0].biharm) hipLaunchKernelGGL(qg_stage_weno_group_kernel<true>, grid, dim3(256), 0, st, g);
    else hipLaunchKernelGGL(qg_stage_weno_group_kernel<false>, grid, dim3(256), 0, st, g);
    return;
  }
  const int ty = a[0].tile_rows ? a[0].tile_rows : TY_DEFAULT;
  if (ty == 8) {
    if (a[0].biharm) hipLaunchKernelGGL((qg_stage_arakawa_group_kernel<true, 8>), grid, dim3(256), 0, st, g);
    else hipLaunchKernelGGL((qg_stage_arakawa_group_kernel<false, 8>), grid, dim3(256), 0, st, g);
  } else {
    if (a[0].biharm) hipLaunchKernelGGL((qg_stage_arakawa_group_kernel<true, 16>), grid, dim3(256), 0, st, g);
    else hipLaunchKernelGGL((qg_stage_arakawa_group_kernel<false, 16>), grid, dim3(256), 0, st, g);
  }
}

__global__ __launch_bounds__(256) void zero_ghost_kernel(float* f, int Ny, int Nx) {
  const long plane = (long)Ny * Nx;
  float* g = f + blockIdx.y * plane;
  const int ring = 2 * Nx + 2 * (Ny - 2);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ring) return;
  long idx;
  if (i < Nx) idx = i;                                     // row 0
  else if (i < 2 * Nx) idx = (long)(Ny - 1) * Nx + (i - Nx);  // row Ny-1
  else {
    const int j = i - 2 * Nx;                              // rows 1..Ny-2, col 0 / Nx-1
    idx = (long)(1 + (j >> 1)) * Nx + ((j & 1) ? Nx - 1 : 0);
  }
  g[idx] = 0.0f;
}

void launch_zero_ghost(float* f, int nl, int Ny, int Nx, hipStream_t st) {
  const int ring = 2 * Nx + 2 * (Ny - 2);
  dim3 grid((ring + 255) / 256, nl, 1);
  hipLaunchKernelGGL(zero_ghost_kernel, grid, dim3(256), 0, st, f, Ny, Nx);
}

// ---------------------------------------------------------------------------
// Helmholtz residual for iterative refinement
// ---------------------------------------------------------------------------
namespace {
__device__ inline void store_pieces(uint16_t* o, long piece, int n, float v) {
  const uint16_t p0 = bf16_rne(v);
  o[0] = p0;
  if (n > 1) {
    const float r1 = v - bf16_bits_to_f32(p0);
    const uint16_t p1 = bf16_rne(r1);
    o[piece] = p1;
    if (n > 2) o[2 * piece] = bf16_rne(r1 - bf16_bits_to_f32(p1));
  }
}
}  // namespace

// block (bx, z) of the (ceil(Ny * op_kpad / 256), P) grid: row-major over
// the operand's Ny x op_kpad
__device__ __forceinline__ void residual_plain(const ResidualArgs& a, int bx, int z) {
  const int i = bx * blockDim.x + threadIdx.x;
  const int y = i / a.op_kpad, x = i - y * a.op_kpad;
  if (y >= a.Ny) return;
  uint16_t* o = a.r + z * a.op_batch + (long)y * a.op_ld + x;
  if (x >= a.Nx) {
    store_pieces(o, a.op_piece, 2, 0.0f);
    return;
  }
  const long s = z * a.plane + (long)y * a.Nx + x;
  const float* p = a.psi + s;
  const float c = p[0];
  const float e = (x + 1 < a.Nx) ? p[1] : 0.0f;
  const float w = (x > 0) ? p[-1] : 0.0f;
  const float n = (y + 1 < a.Ny) ? p[a.Nx] : 0.0f;
  const float so = (y > 0) ? p[-a.Nx] : 0.0f;
  const float r = a.q[s] - ((e - 2.0f * c + w) * a.idx2 + (n - 2.0f * c + so) * a.idy2) + a.lam[z] * c;
  store_pieces(o, a.op_piece, 2, r);
}

// Folded residual on the parity components (block b = 2p + q of plane m at
// b * P + m): the 5-point Laplacian of a component on the half array, with
// zero exterior and the mirror rule at the fold line (the neighbour past the
// last physical index h-1 is +-the value itself: even / odd parity).
// Each wave walks RES_ROWS folded rows of a 64-column strip, keeping the
// rows above and below in registers (a sliding window), so a row is fetched
// about once and no per-element index division is needed.  A block is four
// waves stacked along the rows: 64 columns x 4 RES_ROWS rows.
// block (bx, z) of the (ceil(op_kpad / 64) * ceil(hy / (4 RES_ROWS)), 4P) grid
constexpr int RES_ROWS = 8;
__device__ __forceinline__ void residual_fold(const ResidualArgs& a, int bx, int z) {
  const int hx = a.Nx / 2, hy = a.Ny / 2;
  const int ncx = (a.op_kpad + 63) / 64;
  const int bry = bx / ncx, bcx = bx - bry * ncx;  // uniform over the block
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int kx = bcx * 64 + ((int)threadIdx.x & 63);
  const int ky0 = (bry * 4 + wv) * RES_ROWS;
  if (kx >= a.op_kpad || ky0 >= hy) return;
  const int kend = ky0 + RES_ROWS < hy ? ky0 + RES_ROWS : hy;
  uint16_t* o = a.r + z * a.op_batch + kx;
  if (kx >= hx) {
    for (int ky = ky0; ky < kend; ++ky) store_pieces(o + (long)ky * a.op_ld, a.op_piece, 2, 0.0f);
    return;
  }
  const int blk = z / a.P, m = z - blk * a.P;
  const float sx = (blk & 1) ? -1.0f : 1.0f, sy = (blk >> 1) ? -1.0f : 1.0f;
  const float* Z = a.psi + (long)z * hx * hy;
  const float* Q = a.q + (long)z * hx * hy;
  const float lam = a.lam[m];
  // physical column in the first half; its neighbours mapped back to folded
  // columns (all in range)
  const int x = fold_phys(kx, hx);
  const int xw = x == 0 ? 0 : (x == 1 ? hx - 1 : x - 2);
  // Folded row ky is physical row y = fold_phys(ky): its north neighbour is
  // folded row ky + 1 (row 0 for the ghost pair ky = hy - 1, whose y is 0),
  // its south neighbour folded row ky - 1 (row hy - 1 for ky = 0).
  auto col = [&](int k) { return Z[(long)k * hx + kx]; };
  float below = ky0 == 0 ? col(hy - 1) : col(ky0 - 1);
  float cur = col(ky0);
#pragma unroll
  for (int r = 0; r < RES_ROWS; ++r) {
    const int ky = ky0 + r;
    if (ky >= kend) break;
    const float above = ky + 1 < hy ? col(ky + 1) : col(0);
    const int y = fold_phys(ky, hy);
    const float* Zr = Z + (long)ky * hx;
    const float c = cur, ze = Zr[x], zw = Zr[xw];
    const float qv = Q[(long)ky * hx + kx];
    const float e = x == hx - 1 ? sx * c : ze;
    const float w = x == 0 ? 0.0f : zw;
    const float n = y == hy - 1 ? sy * c : above;
    const float so = y == 0 ? 0.0f : below;
    // the x term fused into the y product's sum, written out so that the
    // rounding does not depend on the compiler's contraction choices
    const float lap = __builtin_fmaf(e - 2.0f * c + w, a.idx2, (n - 2.0f * c + so) * a.idy2);
    const float res = qv - 4.0f * (lap - lam * c);
    store_pieces(o + (long)ky * a.op_ld, a.op_piece, 2, res);
    below = cur;
    cur = above;
  }
}

__global__ __launch_bounds__(256) void residual_kernel(ResidualArgs a) { residual_plain(a, blockIdx.x, blockIdx.y); }
__global__ __launch_bounds__(256) void residual_fold_kernel(ResidualArgs a) { residual_fold(a, blockIdx.x, blockIdx.y); }

struct ResidualGroup {
  BlockTable t;
  ResidualArgs a[MAX_GROUP];
};
__global__ __launch_bounds__(256) void residual_group_kernel(ResidualGroup g) {
  int c, l, x, y, z;
  if (!table_class(g.t, blockIdx.x, c, l)) return;
  class_block(g.t, c, l, x, y, z);
  if (g.a[c].fold) residual_fold(g.a[c], x, y);
  else residual_plain(g.a[c], x, y);
}

namespace {
dim3 residual_grid(const ResidualArgs& a) {
  if (a.fold) {
    const long rows = (a.Ny / 2 + 4 * RES_ROWS - 1) / (4 * RES_ROWS);
    return dim3((unsigned)(((a.op_kpad + 63) / 64) * rows), 4 * a.P);
  }
  return dim3((unsigned)(((long)a.op_kpad * a.Ny + 255) / 256), a.P);
}
}  // namespace

void launch_residual(const ResidualArgs& a, hipStream_t st) {
  const dim3 g = residual_grid(a);
  if (a.fold) hipLaunchKernelGGL(residual_fold_kernel, g, dim3(256), 0, st, a);
  else hipLaunchKernelGGL(residual_kernel, g, dim3(256), 0, st, a);
}

void launch_residual_group(const ResidualArgs* a, int n, hipStream_t st) {
  ResidualGroup g{};
  for (int i = 0; i < n; ++i) {
    g.a[i] = a[i];
    const dim3 d = residual_grid(a[i]);
    table_add(g.t, (int)d.x, (int)d.y, (int)d.z);
  }
  hipLaunchKernelGGL(residual_group_kernel, dim3(g.t.start[g.t.n]), dim3(256), 0, st, g);
}

// ---------------------------------------------------------------------------
// layer -> mode projection in FP64, writing the first GEMM operand
// (one thread per cell of one member: its layers are read once)
// ---------------------------------------------------------------------------
// block (bx, mem) of the (ceil(Ny * width / 256), M) grid: row-major over
// Ny x width, width = max(Nx, op_kpad + off)
template <int NLM>
__device__ __forceinline__ void to_modal_plain(const ModalArgs& a, int bx, int mem) {
  const int off = a.apply_bc ? 1 : 0;
  const int width = a.Nx > a.op_kpad + off ? a.Nx : a.op_kpad + off;
  const int e = bx * blockDim.x + threadIdx.x;
  const int y = e / width, x = e - y * width;  // x: state column
  if (y >= a.Ny) return;
  const int rows = a.Ny - 2 * off, cols = a.Nx - 2 * off;
  const int ox = x - off, oy = y - off;  // operand coordinates
  const bool inside = x < a.Nx;
  const bool wop = oy >= 0 && oy < rows && ox >= 0 && ox < a.op_kpad;
  if (!inside && !wop) return;
  const int first = a.t.mfirst[mem], nl = a.t.mnl[mem];
  const double* cl2m = a.t.cl2m + mem * MAX_LAYERS * MAX_LAYERS;
  const int s = y * a.Nx + (inside ? x : 0);  // in-plane offset (a plane is below 2^31 cells)
  const bool ghost = x == 0 || y == 0 || x == a.Nx - 1 || y == a.Ny - 1;
  // Every load is issued before the first use (layers past nl re-read the
  // member's first layer and are ignored), so one memory latency covers them.
  float qf[NLM];
#pragma unroll
  for (int k = 0; k < NLM; ++k) qf[k] = a.q[(long)(first + (k < nl ? k : 0)) * a.plane + s];
  const bool zero = !inside || (ghost && a.apply_bc);
  double qk[NLM];
#pragma unroll
  for (int k = 0; k < NLM; ++k) qk[k] = zero ? 0.0 : (double)qf[k];
#pragma unroll
  for (int p = 0; p < NLM; ++p) {
    if (p >= nl) break;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < NLM; ++k)
      if (k < nl) acc += cl2m[p * MAX_LAYERS + k] * qk[k];
    const float v = (float)acc;
    const long z = first + p;
    if (inside) a.qm[z * a.plane + s] = v;
    if (wop) {
      const float w = ox < cols ? v : 0.0f;  // zero padding past the operand width
      const long o = z * a.op_batch + (long)oy * a.op_ld + ox;
      if (a.op_pieces == 0) static_cast<float*>(a.op)[o] = w;
      else store_pieces(static_cast<uint16_t*>(a.op) + o, a.op_piece, a.op_pieces, w);
    }
  }
}

// Folded variant: one thread per folded point (k'y, k'x) of one member reads
// the four mirror points of every layer, projects them to modes in FP64 and
// writes the four parity fold sums of each mode.
// block (bx, mem) of the (ceil(hy * width / 256), M) grid
template <int NLM>
__device__ __forceinline__ void to_modal_fold(const ModalArgs& a, int bx, int mem) {
  const int hx = a.Nx / 2, hy = a.Ny / 2;
  const int width = hx > a.op_kpad ? hx : a.op_kpad;
  const int e = bx * blockDim.x + threadIdx.x;  // row-major over hy x width
  if (e >= hy * width) return;
  const int ky = e / width, kx = e - ky * width;
  const int K0 = a.apply_bc ? hx - 1 : hx;  // operand columns carrying data
  const int R0 = a.apply_bc ? hy - 1 : hy;  // operand rows
  const bool inside = kx < hx;
  const bool wop = ky < R0 && kx < a.op_kpad;
  if (!inside && !wop) return;
  const int first = a.t.mfirst[mem], nl = a.t.mnl[mem];
  const double* cl2m = a.t.cl2m + mem * MAX_LAYERS * MAX_LAYERS;
  const int y0 = fold_phys(ky, hy), y1 = a.Ny - 1 - y0;
  const int x0 = inside ? fold_phys(kx, hx) : 0, x1 = a.Nx - 1 - x0;
  // in-plane offsets (a plane is below 2^31 cells)
  const int pts[4] = {y0 * a.Nx + x0, y0 * a.Nx + x1, y1 * a.Nx + x0, y1 * a.Nx + x1};
  // the ghost pair (last folded index) holds only ghost points
  const bool zero = !inside || (a.apply_bc && (ky == hy - 1 || kx == hx - 1));
  // Every load is issued before the first use (layers past nl re-read the
  // member's first layer and are ignored), so one memory latency covers them.
  float qf[NLM][4];
#pragma unroll
  for (int k = 0; k < NLM; ++k) {
    const float* src = a.q + (long)(first + (k < nl ? k : 0)) * a.plane;
#pragma unroll
    for (int c = 0; c < 4; ++c) qf[k][c] = src[pts[c]];
  }
  double qk[NLM][4];
#pragma unroll
  for (int k = 0; k < NLM; ++k)
#pragma unroll
    for (int c = 0; c < 4; ++c) qk[k][c] = zero ? 0.0 : (double)qf[k][c];
  const long fplane = (long)hx * hy;
#pragma unroll
  for (int p = 0; p < NLM; ++p) {
    if (p >= nl) break;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NLM; ++k)
      if (k < nl)
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] += cl2m[p * MAX_LAYERS + k] * qk[k][c];
    // block b = 2 py + qx: sum over mirror points (sy, sx) of (-1)^(py sy + qx sx)
    const double f[4] = {v[0] + v[1] + v[2] + v[3], v[0] - v[1] + v[2] - v[3], v[0] + v[1] - v[2] - v[3],
                         v[0] - v[1] - v[2] + v[3]};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const long z = (long)b * a.P + first + p;
      const float fv = (float)f[b];
      if (inside) a.qm[z * fplane + (long)ky * hx + kx] = fv;
      if (wop) {
        const float w = kx < K0 ? fv : 0.0f;  // zero padding past the operand width
        const long o = z * a.op_batch + (long)ky * a.op_ld + kx;
        if (a.op_pieces == 0) static_cast<float*>(a.op)[o] = w;
        else store_pieces(static_cast<uint16_t*>(a.op) + o, a.op_piece, a.op_pieces, w);
      }
    }
  }
}

// Layer-count bound of the projection kernels' unrolled loops: the smallest
// instance covering the batch's deepest member.
#define SOMAX_NL_LAUNCH(NL, KERNEL, GRID, ST, ARGS)                                      \
  do {                                                                                  \
    const int nl_ = (NL) > 0 ? (NL) : MAX_LAYERS;                                       \
    if (nl_ <= 1) hipLaunchKernelGGL(KERNEL<1>, GRID, dim3(256), 0, ST, ARGS);           \
    else if (nl_ <= 2) hipLaunchKernelGGL(KERNEL<2>, GRID, dim3(256), 0, ST, ARGS);      \
    else if (nl_ <= 3) hipLaunchKernelGGL(KERNEL<3>, GRID, dim3(256), 0, ST, ARGS);      \
    else if (nl_ <= 4) hipLaunchKernelGGL(KERNEL<4>, GRID, dim3(256), 0, ST, ARGS);      \
    else hipLaunchKernelGGL(KERNEL<MAX_LAYERS>, GRID, dim3(256), 0, ST, ARGS);           \
  } while (0)

template <int NLM>
__global__ __launch_bounds__(256) void to_modal_kernel(ModalArgs a) {
  to_modal_plain<NLM>(a, blockIdx.x, blockIdx.y);
}
template <int NLM>
__global__ __launch_bounds__(256) void to_modal_fold_kernel(ModalArgs a) {
  to_modal_fold<NLM>(a, blockIdx.x, blockIdx.y);
}
struct ModalGroup {
  BlockTable t;
  ModalArgs a[MAX_GROUP];
};
template <int NLM>
__global__ __launch_bounds__(256) void to_modal_group_kernel(ModalGroup g) {
  int c, l, x, y, z;
  if (!table_class(g.t, blockIdx.x, c, l)) return;
  class_block(g.t, c, l, x, y, z);
  if (g.a[c].fold) to_modal_fold<NLM>(g.a[c], x, y);
  else to_modal_plain<NLM>(g.a[c], x, y);
}

namespace {
dim3 to_modal_grid(const ModalArgs& a) {
  if (a.fold) {
    const int hx = a.Nx / 2;
    const int width = hx > a.op_kpad ? hx : a.op_kpad;
    return dim3((unsigned)(((long)width * (a.Ny / 2) + 255) / 256), a.M);
  }
  const int off = a.apply_bc ? 1 : 0;
  const int width = a.Nx > a.op_kpad + off ? a.Nx : a.op_kpad + off;
  return dim3((unsigned)(((long)width * a.Ny + 255) / 256), a.M);
}
int nl_of(int nlmax) { return nlmax > 0 ? nlmax : MAX_LAYERS; }
}  // namespace

void launch_to_modal(const ModalArgs& a, hipStream_t st) {
  const dim3 grid = to_modal_grid(a);
  if (a.fold) SOMAX_NL_LAUNCH(a.nlmax, to_modal_fold_kernel, grid, st, a);
  else SOMAX_NL_LAUNCH(a.nlmax, to_modal_kernel, grid, st, a);
}

void launch_to_modal_group(const ModalArgs* a, int n, hipStream_t st) {
  ModalGroup g{};
  int nl = 0;
  for (int i = 0; i < n; ++i) {
    g.a[i] = a[i];
    const dim3 d = to_modal_grid(a[i]);
    table_add(g.t, (int)d.x, (int)d.y, (int)d.z);
    nl = nl_of(a[i].nlmax) > nl ? nl_of(a[i].nlmax) : nl;
  }
  SOMAX_NL_LAUNCH(nl, to_modal_group_kernel, dim3(g.t.start[g.t.n]), st, g);
}

// ---------------------------------------------------------------------------
// mode -> layer projection of psi
// ---------------------------------------------------------------------------
// The mass-constraint partials that thread t stores for member_alpha
// (indices t and t + blockDim; nl * nblk <= 8 * 64 = 2 * blockDim), loaded by
// member_partials before the caller's other loads so that the wait for them
// does not wait for those.
struct AlphaPartials {
  double v[2];
};
__device__ inline AlphaPartials member_partials(const LayerArgs& a, int first, int nl) {
  AlphaPartials r{{0.0, 0.0}};
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int i = threadIdx.x + j * (int)blockDim.x;
    if (i < nl * a.nblk) {
      const int p = i / a.nblk, k = i - p * a.nblk;
      r.v[j] = a.partial[2L * ((long)(first + p) * a.nblk + k)];
    }
  }
  return r;
}

// Mass correction of a block's member: the alphas of its nl modes from the
// FP64 partial sums (each mode's partials summed in order by one thread, so
// every block and every run gets the same bits), into s_alpha; the member's
// first block also writes the wall value of every layer.  Every thread of the
// block calls this before any returns.
__device__ inline void member_alpha(const LayerArgs& a, int first, int nl, const float* cm2l, bool write_wall,
                                    const AlphaPartials& part, float* s_alpha) {
  __shared__ double sp[MAX_LAYERS][64];
  __shared__ double sal[MAX_LAYERS];
  const int t = threadIdx.x;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int i = t + j * (int)blockDim.x;
    if (i < nl * a.nblk) {
      const int p = i / a.nblk, k = i - p * a.nblk;
      sp[p][k] = part.v[j];
    }
  }
  __syncthreads();
  if (t < nl) {
    double sum = 0.0;
    for (int k = 0; k < a.nblk; ++k) sum += sp[t][k];
    const double hm = a.hmean[first + t];
    sal[t] = hm != 0.0 ? -sum * a.inv_count / hm : 0.0;
  }
  __syncthreads();
  if (t < nl) {
    s_alpha[t] = (float)sal[t];
    if (write_wall) {
      double w = 0.0;
      for (int p = 0; p < nl; ++p) w += (double)cm2l[t * MAX_LAYERS + p] * sal[p];
      a.psi_wall[first + t] = (float)w;
    }
  }
  __syncthreads();
}

// block (bx, mem) of the (ceil(Ny * Nx / 256), M) grid: row-major over the plane
template <int NLM>
__device__ __forceinline__ void to_layer_plain(const LayerArgs& a, int bx, int mem) {
  const int e = bx * blockDim.x + threadIdx.x;
  const int y = e / a.Nx, x = e - y * a.Nx;
  const int first = a.t.mfirst[mem], nl = a.t.mnl[mem];
  const float* cm2l = a.t.cm2l + mem * MAX_LAYERS * MAX_LAYERS;
  const bool in = y < a.Ny;
  const int s = in ? e : 0;  // in-plane offset (a plane is below 2^31 cells)
  const bool ghost = a.t.mask[mem] && (x == 0 || y == 0 || x == a.Nx - 1 || y == a.Ny - 1);
  // loads first, before the mass constraint's barriers, so their latency
  // overlaps the alphas (layers past nl re-read the member's first layer,
  // unused)
  const AlphaPartials part = a.homsol ? member_partials(a, first, nl) : AlphaPartials{{0.0, 0.0}};
  float zp[NLM], zd[NLM], zh[NLM];
#pragma unroll
  for (int p = 0; p < NLM; ++p) {
    const long i = (long)(first + (p < nl ? p : 0)) * a.plane + s;
    zp[p] = a.pm[i];
    zd[p] = a.d ? a.d[i] : 0.0f;
    zh[p] = a.homsol ? a.homsol[i] : 0.0f;
  }
  __shared__ float s_alpha[MAX_LAYERS];
  if (a.homsol) member_alpha(a, first, nl, cm2l, bx == 0, part, s_alpha);
  if (!in) return;
  float pm[NLM];
#pragma unroll
  for (int p = 0; p < NLM; ++p) {
    pm[p] = 0.0f;
    if (p < nl) {
      pm[p] = zp[p] + (a.d ? zd[p] : 0.0f);
      if (a.homsol) pm[p] += s_alpha[p] * zh[p];
    }
  }
#pragma unroll
  for (int l = 0; l < NLM; ++l) {
    if (l >= nl) break;
    float v = 0.0f;
#pragma unroll
    for (int p = 0; p < NLM; ++p)
      if (p < nl) v += cm2l[l * MAX_LAYERS + p] * pm[p];
    a.psi[(long)(first + l) * a.plane + s] = ghost ? 0.0f : v;
  }
}

// Folded variant: one thread per folded point of one member unfolds the
// parity components of every mode at the four mirror points, adds the mass
// correction and mixes modes into layers.
// block (bx, mem) of the (ceil(hy * hx / 256), M) grid
template <int NLM>
__device__ __forceinline__ void to_layer_fold(const LayerArgs& a, int bx, int mem) {
  const int hx = a.Nx / 2, hy = a.Ny / 2;
  const int e = bx * blockDim.x + threadIdx.x;  // row-major over hy x hx
  const int first = a.t.mfirst[mem], nl = a.t.mnl[mem];
  const float* cm2l = a.t.cm2l + mem * MAX_LAYERS * MAX_LAYERS;
  const bool in = e < hy * hx;
  const int ee = in ? e : 0;  // threads past the array load point 0 and return after the alphas
  const int ky = ee / hx, kx = ee - ky * hx;
  const int y0 = fold_phys(ky, hy), y1 = a.Ny - 1 - y0;
  const int x0 = fold_phys(kx, hx), x1 = a.Nx - 1 - x0;
  // in-plane offsets (a plane is below 2^31 cells)
  const int pts[4] = {y0 * a.Nx + x0, y0 * a.Nx + x1, y1 * a.Nx + x0, y1 * a.Nx + x1};
  const bool ghost = a.t.mask[mem] && (ky == hy - 1 || kx == hx - 1);
  const long fplane = (long)hx * hy;
  const int fo = ky * hx + kx;
  // Every load is issued before the first use and before the mass
  // constraint's barriers (layers past nl re-read the member's first layer and
  // are ignored), so one memory latency covers them and the alphas.
  const AlphaPartials part = a.homsol ? member_partials(a, first, nl) : AlphaPartials{{0.0, 0.0}};
  float zb[NLM][4], zd[NLM][4], zh[NLM][4];
#pragma unroll
  for (int p = 0; p < NLM; ++p) {
    const int pp = first + (p < nl ? p : 0);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const long i = ((long)b * a.P + pp) * fplane + fo;
      zb[p][b] = a.pm[i];
      zd[p][b] = a.d ? a.d[i] : 0.0f;
      zh[p][b] = a.homsol ? a.homsol[(long)pp * a.plane + pts[b]] : 0.0f;
    }
  }
  __shared__ float s_alpha[MAX_LAYERS];
  if (a.homsol) member_alpha(a, first, nl, cm2l, bx == 0, part, s_alpha);
  if (!in) return;
  float pm[NLM][4];
#pragma unroll
  for (int p = 0; p < NLM; ++p) {
#pragma unroll
    for (int c = 0; c < 4; ++c) pm[p][c] = 0.0f;
    if (p < nl) {
      float z[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) z[b] = zb[p][b] + (a.d ? zd[p][b] : 0.0f);
      pm[p][0] = (z[0] + z[1]) + (z[2] + z[3]);
      pm[p][1] = (z[0] - z[1]) + (z[2] - z[3]);
      pm[p][2] = (z[0] + z[1]) - (z[2] + z[3]);
      pm[p][3] = (z[0] - z[1]) - (z[2] - z[3]);
      if (a.homsol) {
        const float al = s_alpha[p];
#pragma unroll
        for (int c = 0; c < 4; ++c) pm[p][c] += al * zh[p][c];
      }
    }
  }
#pragma unroll
  for (int l = 0; l < NLM; ++l) {
    if (l >= nl) break;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float v = 0.0f;
#pragma unroll
      for (int p = 0; p < NLM; ++p)
        if (p < nl) v += cm2l[l * MAX_LAYERS + p] * pm[p][c];
      a.psi[(long)(first + l) * a.plane + pts[c]] = ghost ? 0.0f : v;
    }
  }
}

template <int NLM>
__global__ __launch_bounds__(256) void to_layer_kernel(LayerArgs a) {
  to_layer_plain<NLM>(a, blockIdx.x, blockIdx.y);
}
template <int NLM>
__global__ __launch_bounds__(256) void to_layer_fold_kernel(LayerArgs a) {
  to_layer_fold<NLM>(a, blockIdx.x, blockIdx.y);
}
struct LayerGroup {
  BlockTable t;
  LayerArgs a[MAX_GROUP];
};
template <int NLM>
__global__ __launch_bounds__(256) void to_layer_group_kernel(LayerGroup g) {
  int c, l, x, y, z;
  if (!table_class(g.t, blockIdx.x, c, l)) return;
  class_block(g.t, c, l, x, y, z);
  if (g.a[c].fold) to_layer_fold<NLM>(g.a[c], x, y);
  else to_layer_plain<NLM>(g.a[c], x, y);
}

namespace {
dim3 to_layer_grid(const LayerArgs& a) {
  if (a.fold) return dim3((unsigned)(((long)(a.Nx / 2) * (a.Ny / 2) + 255) / 256), a.M);
  return dim3((unsigned)(((long)a.Nx * a.Ny + 255) / 256), a.M);
}
}  // namespace

void launch_to_layer(const LayerArgs& a, hipStream_t st) {
  const dim3 grid = to_layer_grid(a);
  if (a.fold) SOMAX_NL_LAUNCH(a.nlmax, to_layer_fold_kernel, grid, st, a);
  else SOMAX_NL_LAUNCH(a.nlmax, to_layer_kernel, grid, st, a);
}

void launch_to_layer_group(const LayerArgs* a, int n, hipStream_t st) {
  LayerGroup g{};
  int nl = 0;
  for (int i = 0; i < n; ++i) {
    g.a[i] = a[i];
    const dim3 d = to_layer_grid(a[i]);
    table_add(g.t, (int)d.x, (int)d.y, (int)d.z);
    nl = nl_of(a[i].nlmax) > nl ? nl_of(a[i].nlmax) : nl;
  }
  SOMAX_NL_LAUNCH(nl, to_layer_group_kernel, dim3(g.t.start[g.t.n]), st, g);
}

}  // namespace somax

namespace somax {

// ---------------------------------------------------------------------------
// diagnostics (baroclinic.py:197-228, barotropic.py:159-183)
// ---------------------------------------------------------------------------
size_t qg_diag_partial_doubles(int Nx, int Ny, int P) { return (size_t)2 * P * Ny * ((Nx + 255) / 256); }

namespace {
// block sum of a pair of doubles (256 threads), written by thread 0
__device__ inline void block_pair_sum(double a, double b, double* dst) {
  __shared__ double sa[256], sb[256];
  const int t = threadIdx.x;
  sa[t] = a;
  sb[t] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      sa[t] += sa[t + s];
      sb[t] += sb[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    dst[0] = sa[0];
    dst[1] = sb[0];
  }
}
}  // namespace

__global__ __launch_bounds__(256) void qg_diag_kernel(QGDiagArgs a) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int y = blockIdx.y;
  const int z = blockIdx.z;
  const int Nx = a.Nx, Ny = a.Ny;
  double ke = 0.0, en = 0.0;
  if (x < Nx) {
    const long s = z * a.plane + (long)y * Nx + x;
    const float* p = a.psi + s;
    if (a.u) {
      float uv = 0.0f;
      if (y <= Ny - 2) uv = -(p[Nx] - p[0]) * a.idy;
      a.u[s] = uv;
    }
    if (a.v) {
      float vv = 0.0f;
      if (x <= Nx - 2) vv = (p[1] - p[0]) * a.idx;
      a.v[s] = vv;
    }
    const bool interior = y >= 1 && y <= Ny - 2 && x >= 1 && x <= Nx - 2;
    if (a.zeta) {
      float zv = 0.0f;
      if (interior) zv = (p[1] - 2.0f * p[0] + p[-1]) * a.idx2 + (p[Nx] - 2.0f * p[0] + p[-Nx]) * a.idy2;
      a.zeta[s] = zv;
    }
    if (interior) {
      // face averages: -(psi[y+1] - psi[y-1]) / (2 dy), (psi[x+1] - psi[x-1]) / (2 dx)
      const double uT = -0.5 * ((double)(p[Nx] - p[0]) + (double)(p[0] - p[-Nx])) * a.idy;
      const double vT = 0.5 * ((double)(p[1] - p[0]) + (double)(p[0] - p[-1])) * a.idx;
      ke = uT * uT + vT * vT;
      const double qv = a.q[s];
      en = qv * qv;
    }
  }
  const long blk = (long)z * gridDim.y * gridDim.x + (long)y * gridDim.x + blockIdx.x;
  block_pair_sum(ke, en, a.partial + 2 * blk);
}

__global__ __launch_bounds__(256) void pair_reduce_kernel(const double* partial, int n, double scale, double* out) {
  const int z = blockIdx.x;
  const double* p = partial + 2L * n * z;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    a += p[2 * i];
    b += p[2 * i + 1];
  }
  __shared__ double res[2];
  block_pair_sum(a, b, res);
  if (threadIdx.x == 0) {
    out[2 * z] = scale * res[0];
    out[2 * z + 1] = scale * res[1];
  }
}

void launch_qg_diag(const QGDiagArgs& a, hipStream_t st) {
  const int bx = (a.Nx + 255) / 256;
  hipLaunchKernelGGL(qg_diag_kernel, dim3(bx, a.Ny, a.P), dim3(256), 0, st, a);
  launch_pair_reduce(a.partial, bx * a.Ny, a.P, 0.5 * a.area, a.out, st);
}

void launch_pair_reduce(const double* partial, int n, int P, double scale, double* out, hipStream_t st) {
  hipLaunchKernelGGL(pair_reduce_kernel, dim3(P), dim3(256), 0, st, partial, n, scale, out);
}

}  // namespace somax

namespace somax {

// ---------------------------------------------------------------------------
// mass correction (homogeneous solution method)
// ---------------------------------------------------------------------------
// Per-block FP64 partial sums of a[i] + b[i] (b nullable) over planes of n
// contiguous floats at stride `plane`: block k of plane z takes the chunk
// [k * chunk, (k + 1) * chunk), 1024 threads with four loads each in flight,
// and a fixed summation order, so the result is deterministic.
constexpr int SUM_THREADS = 1024;
// block k of plane z
__device__ __forceinline__ void plane_sums_block(const PlaneSumArgs& p, int k, int z) {
  __shared__ double red[SUM_THREADS];
  const int t = threadIdx.x;
  const long lo = (long)k * p.chunk;
  const long hi = lo + p.chunk < p.n ? lo + p.chunk : p.n;
  const float* pa = p.a + z * p.plane;
  const float* pb = p.b ? p.b + z * p.plane : nullptr;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (long i = lo + t; i < hi; i += 4 * SUM_THREADS) {
    float va[4], vb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long j = i + u * SUM_THREADS;
      va[u] = j < hi ? pa[j] : 0.0f;
      vb[u] = (pb && j < hi) ? pb[j] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) s[u] += (double)va[u] + (double)vb[u];
  }
  red[t] = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  for (int w = SUM_THREADS / 2; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) {
    double* dst = p.partial + 2 * ((long)z * p.nblk + k);
    dst[0] = red[0];
    dst[1] = 0.0;
  }
}

__global__ __launch_bounds__(SUM_THREADS) void plane_sums_kernel(PlaneSumArgs p) {
  plane_sums_block(p, blockIdx.x, blockIdx.y);
}
struct PlaneSumGroup {
  BlockTable t;
  PlaneSumArgs a[MAX_GROUP];
};
__global__ __launch_bounds__(SUM_THREADS) void plane_sums_group_kernel(PlaneSumGroup g) {
  int c, l, k, z, unused;
  if (!table_class(g.t, blockIdx.x, c, l)) return;
  class_block(g.t, c, l, k, z, unused);
  plane_sums_block(g.a[c], k, z);
}

PlaneSumArgs plane_sums_args(const float* a, const float* b, int Nx, int Ny, int P, long plane, double* partial) {
  // about 4096 values per block (one round of loads for 1024 threads), at
  // most 64 blocks per plane and never more than Ny (the partial buffer holds
  // 2 * P * Ny * ceil(Nx / 256) doubles); small planes of a large batch thus
  // take one or two blocks each rather than 64 nearly idle ones
  const long n = (long)Nx * Ny;
  long want = (n + 4095) / 4096;
  if (want > 64) want = 64;
  if (want > Ny) want = Ny;
  const int nblk = want < 1 ? 1 : (int)want;
  return PlaneSumArgs{a, b, n, plane, (n + nblk - 1) / nblk, partial, P, nblk};
}

int launch_plane_sums(const float* a, const float* b, int Nx, int Ny, int P, long plane, double* partial,
                      hipStream_t st) {
  const PlaneSumArgs p = plane_sums_args(a, b, Nx, Ny, P, plane, partial);
  hipLaunchKernelGGL(plane_sums_kernel, dim3(p.nblk, P), dim3(SUM_THREADS), 0, st, p);
  return p.nblk;
}

void launch_plane_sums_group(const PlaneSumArgs* a, int n, hipStream_t st) {
  PlaneSumGroup g{};
  for (int i = 0; i < n; ++i) {
    g.a[i] = a[i];
    table_add(g.t, a[i].nblk, a[i].P, 1);
  }
  hipLaunchKernelGGL(plane_sums_group_kernel, dim3(g.t.start[g.t.n]), dim3(SUM_THREADS), 0, st, g);
}

}  // namespace somax

namespace somax {

// ---------------------------------------------------------------------------
// Runge-Kutta stage combination (Tsit5): out = y + sum_j w_j k_j in FP64
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void lincomb_kernel(LinComb a) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
    double v = (double)a.y[i];
#pragma unroll
    for (int j = 0; j < LINCOMB_MAX; ++j)
      if (j < a.nk) v += a.w[j] * (double)a.k[j][i];
    a.out[i] = (float)v;
  }
}

void launch_lincomb(const LinComb& a, hipStream_t st) {
  long blocks = (a.n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(lincomb_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a);
}

}  // namespace somax

namespace somax {

// ---------------------------------------------------------------------------
// geostrophic projection (ReparameterizedQG.project, reparameterized.py:142-178)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void project_pv_kernel(ProjArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  const int k = blockIdx.z;
  if (x >= a.Nx) return;
  const long s = (long)k * a.plane + (long)y * a.Nx + x;
  float zeta = 0.0f;
  if (y + 1 < a.Ny && x + 1 < a.Nx)
    zeta = (a.v[s + 1] - a.v[s]) * a.idx - (a.u[s + a.Nx] - a.u[s]) * a.idy;
  const float eta = a.hdev ? a.h[s] : a.h[s] - a.H[k];
  a.q[s] = zeta - a.f0 / a.H[k] * eta;
}

__global__ __launch_bounds__(256) void project_geo_kernel(ProjArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.Nx) return;
  const long c = (long)y * a.Nx + x;
  float ps[MAX_LAYERS];
#pragma unroll
  for (int k = 0; k < MAX_LAYERS; ++k) ps[k] = k < a.nl ? a.psi[k * a.plane + c] : 0.0f;
#pragma unroll
  for (int k = 0; k < MAX_LAYERS; ++k) {
    if (k >= a.nl) break;
    const float* p = a.psi + k * a.plane;
    const float s_ = y > 0 ? p[c - a.Nx] : 0.0f;
    const float w = x > 0 ? p[c - 1] : 0.0f;
    a.uo[k * a.plane + c] = -(ps[k] - s_) * a.idy;
    a.vo[k * a.plane + c] = (ps[k] - w) * a.idx;
    float m = 0.0f;
#pragma unroll
    for (int l = 0; l < MAX_LAYERS; ++l)
      if (l < a.nl) m += a.M[k * MAX_LAYERS + l] * ps[l];
    const float eta = a.f0 * a.H[k] * m;
    a.ho[k * a.plane + c] = a.hdev ? eta : a.H[k] + eta;
  }
}

void launch_project_pv(const ProjArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(project_pv_kernel, dim3((a.Nx + 255) / 256, a.Ny, a.nl), dim3(256), 0, st, a);
}

void launch_project_geo(const ProjArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(project_geo_kernel, dim3((a.Nx + 255) / 256, a.Ny), dim3(256), 0, st, a);
}

}  // namespace somax
