// Internal kernel interfaces of the somax HIP library.
#pragma once

#include "common.h"

namespace somax {

enum GemmPrec { GP_FP32 = 0, GP_BF16X3 = 1, GP_BF16X1 = 2 };
enum GemmEpi {
  GE_STORE = 0,  // O = acc
  GE_FINAL = 1   // O = [Oadd +] acc, then the ghost ring of the physical array set to 0 if mask_ghost
};

// O[r][c] (or O[c][r]) = sum_{k<K} X[r*ldx + k] * S[c*lds + k],  r < R, c < C.
struct GemmArgs {
  const float* X;
  long ldx, bsx;  // row pitch, batch stride (floats)
  const float* Sf;       // FP32 sine matrix (GP_FP32)
  const uint16_t* Sh;    // bf16 hi part (GP_BF16X*), zero padded
  const uint16_t* Sl;    // bf16 lo part (GP_BF16X3)
  long lds;              // pitch of the sine matrix (elements)
  float* O;
  long ldo, bso;
  const float* Oadd;     // GE_FINAL: optional addend with the same layout as O
  int R, C, K;
  int mask_ghost;        // GE_FINAL: zero the ghost ring of the phys_rows x phys_cols array
  int phys_rows, phys_cols;
};

template <int EPI>
__device__ inline void gemm_store(const GemmArgs& a, float* O, long z, int ro, int co, float v) {
  const long idx = (long)ro * a.ldo + co;
  if (EPI == GE_FINAL) {
    if (a.Oadd) v += a.Oadd[z * a.bso + idx];
    if (a.mask_ghost && (ro == 0 || co == 0 || ro == a.phys_rows - 1 || co == a.phys_cols - 1)) v = 0.0f;
  }
  O[idx] = v;
}

void launch_dst_gemm(const GemmArgs& a, int prec, bool trans, int epi, int batch, hipStream_t st);

// Spectral operator of the layer-space inversion, applied in place to
// Zt[l][i][j] (i = x mode, j = y mode):
//   W_l = sum_m cm2l[l][m] / (lamx[i] + lamy[j] - lambda_m) * sum_k cl2m[m][k] Z_k
constexpr int MAX_LAYERS = 8;
struct SpectralArgs {
  float* Z;
  long ld, bs;
  int nxm, nym;  // rows (x modes), cols (y modes)
  int nl;
  const float* lamx;  // nxm
  const float* lamy;  // nym
  float cl2m[MAX_LAYERS * MAX_LAYERS];
  float cm2l[MAX_LAYERS * MAX_LAYERS];
  float lam[MAX_LAYERS];
};
void launch_spectral(const SpectralArgs& a, hipStream_t st);

// Helmholtz residual of the layer-space inversion on the full ghosted array:
//   r_l = BC(q)_l - lap_h(psi_l) + sum_k B[l][k] psi_k,  B = cm2l diag(lambda) cl2m,
// lap_h with zero values outside the array (the operator DST-I diagonalises).
struct ResidualArgs {
  const float* q;  // state layout (ghost ring read as 0)
  long q_ld, q_bs;
  const float* psi;
  long p_ld, p_bs;
  float* r;
  long r_ld, r_bs;
  int Nx, Ny, nl;
  int apply_bc;  // 1: ghost ring of q read as 0
  float idx2, idy2;
  float B[MAX_LAYERS * MAX_LAYERS];
};
void launch_residual(const ResidualArgs& a, hipStream_t st);

// One QG RHS evaluation fused with the RK stage combination:
//   out = c0*qn + c1*qs + c2*rhs(BC(qs), psi)
struct StageArgs {
  const float* qs;   // stage state (raw, ghost values kept in the combination)
  const float* psi;  // streamfunction of BC(qs)
  const float* qn;   // step start state (may be null when c0 == 0)
  float* out;
  int Nx, Ny, nl;
  long plane;        // Ny*Nx
  const float* beta_row;  // Ny
  const float* wind_row;  // Ny
  float wind_coef[MAX_LAYERS];  // tau0 * wind_scale[l]
  float nu, nu4, kappa;         // kappa acts on layer nl-1 only
  float idx2, idy2, iarakawa;   // 1/dx^2, 1/dy^2, 1/(12 dx dy)
  float idx, idy;               // 1/dx, 1/dy (WENO)
  float psi_wall[MAX_LAYERS];   // WENO corner value on the walls
  float c0, c1, c2;
  int advection;
  int apply_bc;  // 1: ghost ring of qs read as 0 in the RHS (SomaxModel.build_terms)
};
void launch_qg_stage(const StageArgs& a, hipStream_t st);

// Ghost ring of nl planes of Ny x Nx set to 0 (finitevolx.zero_boundaries).
void launch_zero_ghost(float* f, int nl, int Ny, int Nx, hipStream_t st);

}  // namespace somax
