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
constexpr int TY = 16;
constexpr int QW = TX + 4;  // q tile pitch (halo 2)
constexpr int PW = TX + 2;  // psi / lap tile pitch (halo 1)
}  // namespace

template <bool BIHARM>
__global__ __launch_bounds__(256) void qg_stage_arakawa_kernel(StageArgs a) {
  __shared__ float sq[(TY + 4) * QW];
  __shared__ float sp[(TY + 2) * PW];
  __shared__ float sl[BIHARM ? (TY + 2) * PW : 1];
  __shared__ float sb[TY + 4];

  const int tid = threadIdx.x;
  const int layer = blockIdx.z;
  const int x0 = blockIdx.x * TX;
  const int y0 = blockIdx.y * TY;
  const int Nx = a.Nx, Ny = a.Ny;
  const long off = (long)layer * a.plane;
  const float* qs = a.qs + off;
  const float* psi = a.psi + off;

  // q with the boundary condition (ghost ring and outside the array read 0)
  for (int i = tid; i < (TY + 4) * QW; i += 256) {
    const int ly = i / QW, lx = i - ly * QW;
    const int gy = y0 - 2 + ly, gx = x0 - 2 + lx;
    float v = 0.0f;
    const bool inside = gy >= 0 && gy < Ny && gx >= 0 && gx < Nx;
    const bool ghost = gy == 0 || gx == 0 || gy == Ny - 1 || gx == Nx - 1;
    if (inside && !(ghost && a.apply_bc)) v = qs[(long)gy * Nx + gx];
    sq[i] = v;
  }
  for (int i = tid; i < (TY + 2) * PW; i += 256) {
    const int ly = i / PW, lx = i - ly * PW;
    const int gy = y0 - 1 + ly, gx = x0 - 1 + lx;
    float v = 0.0f;
    if (gy >= 0 && gy < Ny && gx >= 0 && gx < Nx) v = psi[(long)gy * Nx + gx];
    sp[i] = v;
  }
  if (tid < TY + 4) {
    const int gy = y0 - 2 + tid;
    sb[tid] = (gy >= 0 && gy < Ny) ? a.beta_row[gy] : 0.0f;
  }
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

  const int tx = tid & 63;
  const int tyb = tid >> 6;
  const float wc = a.wind_coef[layer];
  const bool bottom = (layer == a.nl - 1) && (a.kappa != 0.0f);
#pragma unroll
  for (int r = 0; r < TY / 4; ++r) {
    const int ly = tyb + 4 * r;
    const int gy = y0 + ly, gx = x0 + tx;
    if (gy >= Ny || gx >= Nx) continue;
    float rhs = 0.0f;
    if (gy >= 1 && gy <= Ny - 2 && gx >= 1 && gx <= Nx - 2) {
      const float* q = &sq[(ly + 2) * QW + tx + 2];
      const float* p = &sp[(ly + 1) * PW + tx + 1];
      const float bN = sb[ly + 3], bC = sb[ly + 2], bS = sb[ly + 1];
      // total PV q + beta*y around the point (Arakawa reads the ghost ring)
      const float tE = q[1] + bC, tW = q[-1] + bC, tN = q[QW] + bN, tS = q[-QW] + bS;
      const float tNE = q[QW + 1] + bN, tNW = q[QW - 1] + bN, tSE = q[-QW + 1] + bS, tSW = q[-QW - 1] + bS;
      const float pE = p[1], pW = p[-1], pN = p[PW], pS = p[-PW];
      const float pNE = p[PW + 1], pNW = p[PW - 1], pSE = p[-PW + 1], pSW = p[-PW - 1];
      const float jpp = (pE - pW) * (tN - tS) - (pN - pS) * (tE - tW);
      const float jpx = pE * (tNE - tSE) - pW * (tNW - tSW) - pN * (tNE - tNW) + pS * (tSE - tSW);
      const float jxp = tN * (pNE - pNW) - tS * (pSE - pSW) - tE * (pNE - pSE) + tW * (pNW - pSW);
      rhs = -(jpp + jpx + jxp) * a.iarakawa;
      if (a.nu != 0.0f) rhs += a.nu * ((q[1] - 2.0f * q[0] + q[-1]) * a.idx2 + (q[QW] - 2.0f * q[0] + q[-QW]) * a.idy2);
      if (BIHARM) {
        const float* l = &sl[(ly + 1) * PW + tx + 1];
        rhs -= a.nu4 * ((l[1] - 2.0f * l[0] + l[-1]) * a.idx2 + (l[PW] - 2.0f * l[0] + l[-PW]) * a.idy2);
      }
      if (bottom) rhs -= a.kappa * ((pE - 2.0f * p[0] + pW) * a.idx2 + (pN - 2.0f * p[0] + pS) * a.idy2);
    }
    rhs += wc * a.wind_row[gy];
    const long idx = off + (long)gy * Nx + gx;
    float o = a.c2 * rhs;
    if (a.c1 != 0.0f) o += a.c1 * a.qs[idx];
    if (a.c0 != 0.0f) o += a.c0 * a.qn[idx];
    a.out[idx] = o;
  }
}

void launch_qg_stage(const StageArgs& a, hipStream_t st) {
  dim3 grid((a.Nx + TX - 1) / TX, (a.Ny + TY - 1) / TY, a.nl);
  if (a.nu4 != 0.0f)
    hipLaunchKernelGGL(qg_stage_arakawa_kernel<true>, grid, dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(qg_stage_arakawa_kernel<false>, grid, dim3(256), 0, st, a);
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
// spectral operator (in place)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void spectral_kernel(SpectralArgs a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;  // y mode
  const int i = blockIdx.y;                              // x mode
  if (j >= a.nym) return;
  const float lxy = a.lamx[i] + a.lamy[j];
  float z[MAX_LAYERS], t[MAX_LAYERS];
  const long base = (long)i * a.ld + j;
  for (int k = 0; k < a.nl; ++k) z[k] = a.Z[k * a.bs + base];
  for (int m = 0; m < a.nl; ++m) {
    float s = 0.0f;
    for (int k = 0; k < a.nl; ++k) s += a.cl2m[m * MAX_LAYERS + k] * z[k];
    t[m] = s / (lxy - a.lam[m]);
  }
  for (int l = 0; l < a.nl; ++l) {
    float s = 0.0f;
    for (int m = 0; m < a.nl; ++m) s += a.cm2l[l * MAX_LAYERS + m] * t[m];
    a.Z[l * a.bs + base] = s;
  }
}

void launch_spectral(const SpectralArgs& a, hipStream_t st) {
  dim3 grid((a.nym + 255) / 256, a.nxm, 1);
  hipLaunchKernelGGL(spectral_kernel, grid, dim3(256), 0, st, a);
}

// ---------------------------------------------------------------------------
// Helmholtz residual for iterative refinement
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void residual_kernel(ResidualArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.Nx) return;
  float pc[MAX_LAYERS];
  for (int k = 0; k < a.nl; ++k) pc[k] = a.psi[k * a.p_bs + (long)y * a.p_ld + x];
  const bool inner = (y >= 1 && y <= a.Ny - 2 && x >= 1 && x <= a.Nx - 2);
  for (int l = 0; l < a.nl; ++l) {
    const float* p = a.psi + l * a.p_bs + (long)y * a.p_ld + x;
    const float e = (x + 1 < a.Nx) ? p[1] : 0.0f;
    const float w = (x > 0) ? p[-1] : 0.0f;
    const float n = (y + 1 < a.Ny) ? p[a.p_ld] : 0.0f;
    const float s = (y > 0) ? p[-a.p_ld] : 0.0f;
    float v = (inner || !a.apply_bc) ? a.q[l * a.q_bs + (long)y * a.q_ld + x] : 0.0f;
    v -= (e - 2.0f * pc[l] + w) * a.idx2 + (n - 2.0f * pc[l] + s) * a.idy2;
    for (int k = 0; k < a.nl; ++k) v += a.B[l * MAX_LAYERS + k] * pc[k];
    a.r[l * a.r_bs + (long)y * a.r_ld + x] = v;
  }
}

void launch_residual(const ResidualArgs& a, hipStream_t st) {
  dim3 grid((a.Nx + 255) / 256, a.Ny, 1);
  hipLaunchKernelGGL(residual_kernel, grid, dim3(256), 0, st, a);
}

}  // namespace somax
