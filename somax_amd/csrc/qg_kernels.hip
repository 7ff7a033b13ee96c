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
  __shared__ float sdb[TY + 4];

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
    sdb[tid] = (gy >= 0 && gy < Ny) ? a.dbeta_row[gy] : 0.0f;
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
      if (a.nu != 0.0f) rhs += a.nu * ((q[1] - 2.0f * q[0] + q[-1]) * a.idx2 + (q[QW] - 2.0f * q[0] + q[-QW]) * a.idy2);
      if (BIHARM) {
        const float* l = &sl[(ly + 1) * PW + tx + 1];
        rhs -= a.nu4 * ((l[1] - 2.0f * l[0] + l[-1]) * a.idx2 + (l[PW] - 2.0f * l[0] + l[-PW]) * a.idy2);
      }
      if (bottom) rhs -= a.kappa * ((pE - 2.0f * p[0] + pW) * a.idx2 + (pN - 2.0f * p[0] + pS) * a.idy2);
    }
    rhs += wc * a.wind_row[gy];
    const long idx = off + (long)gy * Nx + gx;
    // Stage combination qn + c1 (qs - qn) + c2 rhs: the FP32 weights of the
    // Shu-Osher form (1/3 + 2/3 != 1 in FP32) would otherwise scale q by
    // 1 + 3e-8 per step.  Without qn: c1 qs + c2 rhs.
    float o;
    if (a.qn) {
      const float qn = a.qn[idx];
      o = qn + a.c1 * (a.qs[idx] - qn) + a.c2 * rhs;
    } else {
      o = a.c2 * rhs;
      if (a.c1 != 0.0f) o += a.c1 * a.qs[idx];
    }
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
// Helmholtz residual for iterative refinement
// ---------------------------------------------------------------------------
namespace {
__device__ inline void store_pieces(uint16_t* o, long piece, int n, float v) {
  const uint16_t p0 = f32_to_bf16_bits(v);
  o[0] = p0;
  if (n > 1) {
    const float r1 = v - bf16_bits_to_f32(p0);
    const uint16_t p1 = f32_to_bf16_bits(r1);
    o[piece] = p1;
    if (n > 2) o[2 * piece] = f32_to_bf16_bits(r1 - bf16_bits_to_f32(p1));
  }
}
}  // namespace

__global__ __launch_bounds__(256) void residual_kernel(ResidualArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;  // up to op_kpad
  const int y = blockIdx.y;
  const int m = blockIdx.z;
  if (x >= a.op_kpad) return;
  uint16_t* o = a.r + m * a.op_batch + (long)y * a.op_ld + x;
  if (x >= a.Nx) {
    store_pieces(o, a.op_piece, 2, 0.0f);
    return;
  }
  const long s = m * a.plane + (long)y * a.Nx + x;
  const float* p = a.psi + s;
  const float c = p[0];
  const float e = (x + 1 < a.Nx) ? p[1] : 0.0f;
  const float w = (x > 0) ? p[-1] : 0.0f;
  const float n = (y + 1 < a.Ny) ? p[a.Nx] : 0.0f;
  const float so = (y > 0) ? p[-a.Nx] : 0.0f;
  const float r = a.q[s] - ((e - 2.0f * c + w) * a.idx2 + (n - 2.0f * c + so) * a.idy2) + a.lam[m] * c;
  store_pieces(o, a.op_piece, 2, r);
}

void launch_residual(const ResidualArgs& a, hipStream_t st) {
  dim3 grid((a.op_kpad + 255) / 256, a.Ny, a.nl);
  hipLaunchKernelGGL(residual_kernel, grid, dim3(256), 0, st, a);
}

// ---------------------------------------------------------------------------
// layer -> mode projection in FP64, writing the first GEMM operand
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void to_modal_kernel(ModalArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;  // state column
  const int y = blockIdx.y;
  const int off = a.apply_bc ? 1 : 0;
  const int rows = a.Ny - 2 * off, cols = a.Nx - 2 * off;
  const int ox = x - off, oy = y - off;  // operand coordinates
  const bool inside = x < a.Nx;
  const bool wop = oy >= 0 && oy < rows && ox >= 0 && ox < a.op_kpad;
  if (!inside && !wop) return;
  const long s = (long)y * a.Nx + x;
  const bool ghost = x == 0 || y == 0 || x == a.Nx - 1 || y == a.Ny - 1;
  double qk[MAX_LAYERS];
  for (int k = 0; k < a.nl; ++k)
    qk[k] = (!inside || (ghost && a.apply_bc)) ? 0.0 : (double)a.q[k * a.plane + s];
  for (int m = 0; m < a.nl; ++m) {
    double acc = 0.0;
    for (int k = 0; k < a.nl; ++k) acc += a.cl2m[m * MAX_LAYERS + k] * qk[k];
    const float v = (float)acc;
    if (inside) a.qm[m * a.plane + s] = v;
    if (wop) {
      const float w = ox < cols ? v : 0.0f;  // zero padding past the operand width
      const long o = m * a.op_batch + (long)oy * a.op_ld + ox;
      if (a.op_pieces == 0) static_cast<float*>(a.op)[o] = w;
      else store_pieces(static_cast<uint16_t*>(a.op) + o, a.op_piece, a.op_pieces, w);
    }
  }
}

void launch_to_modal(const ModalArgs& a, hipStream_t st) {
  const int off = a.apply_bc ? 1 : 0;
  const int width = a.Nx > a.op_kpad + off ? a.Nx : a.op_kpad + off;
  dim3 grid((width + 255) / 256, a.Ny, 1);
  hipLaunchKernelGGL(to_modal_kernel, grid, dim3(256), 0, st, a);
}

// ---------------------------------------------------------------------------
// mode -> layer projection of psi
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void to_layer_kernel(LayerArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= a.Nx) return;
  const long s = (long)y * a.Nx + x;
  const bool ghost = a.mask && (x == 0 || y == 0 || x == a.Nx - 1 || y == a.Ny - 1);
  float pm[MAX_LAYERS];
  for (int m = 0; m < a.nl; ++m) pm[m] = a.p[m * a.plane + s] + (a.d ? a.d[m * a.plane + s] : 0.0f);
  for (int l = 0; l < a.nl; ++l) {
    float v = 0.0f;
    for (int m = 0; m < a.nl; ++m) v += a.cm2l[l * MAX_LAYERS + m] * pm[m];
    a.psi[l * a.plane + s] = ghost ? 0.0f : v;
  }
}

void launch_to_layer(const LayerArgs& a, hipStream_t st) {
  dim3 grid((a.Nx + 255) / 256, a.Ny, 1);
  hipLaunchKernelGGL(to_layer_kernel, grid, dim3(256), 0, st, a);
}

}  // namespace somax
