// Internal kernel interfaces of the somax HIP library.
#pragma once

#include "common.h"

namespace somax {

// O[a][b] = sum_k A[a][k] B[b][k] with the data plane X in the A slot
// (trans == 0: O = X S) or in the B slot (trans == 1: O = (X S)^T), S the
// symmetric sine matrix.  FP32 operands when Xf != null, else bf16 pieces.
struct GemmArgs {
  // data operand X: rows xrows, k contiguous, zero padded in k to the k-tile
  const float* Xf;
  const uint16_t* Xb;
  long ldx, x_piece, x_batch;  // pitch, piece stride, batch stride (elements)
  int xrows;
  // sine operand S: rows padded to 128, k zero padded to 32
  const float* Sf;
  const uint16_t* Sb;
  long lds, s_piece;
  int K;
  int trans;
  // per batch plane precision: planes z < z_split use np_lo products and
  // write op_lo output pieces, the rest np_hi / op_hi (pieces 0 = FP32 out)
  int z_split, np_lo, np_hi, op_lo, op_hi;
  int acc64;  // add each k-tile's FP32 partial into an FP64 running sum
  // output O[ro][co], ro < RA, co < RB; columns [RB, RBpad) written as 0
  float* Of;
  uint16_t* Ob;
  long ldo, o_piece, o_batch;
  int RA, RB, RBpad;
  // optional spectral epilogue: O /= (lamA[ro] + lamB[co] - lam[z])
  const float* lamA;
  const float* lamB;
  const float* lam;  // device array, one value per batch plane
};

void launch_dst_gemm(const GemmArgs& a, int batch, hipStream_t st);

constexpr int MAX_LAYERS = 8;

// Layer -> mode projection with FP64 accumulation (the stretching term makes
// layer PV much larger than the barotropic modal PV, so an FP32 sum cancels):
//   qm[m] = sum_k cl2m[m][k] * q[k]   (ghost ring of q read as 0 when apply_bc)
// qm is written in FP32 (state layout) and as the first GEMM's operand:
// interior or full rows/cols (interior == apply_bc), pitch op_ld, zero padded
// to op_kpad columns, as op_pieces bf16 pieces or FP32 (op_pieces == 0).
struct ModalArgs {
  const float* q;
  float* qm;
  int Nx, Ny, nl;
  long plane;
  int apply_bc;
  double cl2m[MAX_LAYERS * MAX_LAYERS];
  void* op;
  int op_pieces;
  long op_ld, op_piece, op_batch;
  int op_kpad;
};
void launch_to_modal(const ModalArgs& a, hipStream_t st);

// Per-mode Helmholtz residual on the full ghosted array,
//   r_m = qm_m - lap_h(psi_m) + lambda_m psi_m,
// lap_h with zero values outside the array (the operator DST-I diagonalises),
// written as 2 bf16 pieces in the GEMM operand layout (pitch op_ld, zero
// padded to op_kpad columns).
struct ResidualArgs {
  const float* q;    // modal PV, state layout
  const float* psi;  // modal streamfunction, state layout
  uint16_t* r;
  long op_ld, op_piece, op_batch;
  int op_kpad;
  long plane;
  int Nx, Ny, nl;
  float idx2, idy2;
  float lam[MAX_LAYERS];
};
void launch_residual(const ResidualArgs& a, hipStream_t st);

// Mode -> layer projection of the streamfunction:
//   psi_l = sum_m cm2l[l][m] * (p_m + d_m)   (d nullable), ghost ring 0 if mask
struct LayerArgs {
  const float* p;
  const float* d;
  float* psi;
  int Nx, Ny, nl;
  long plane;
  int mask;
  float cm2l[MAX_LAYERS * MAX_LAYERS];
};
void launch_to_layer(const LayerArgs& a, hipStream_t st);

// One QG RHS evaluation fused with the RK stage combination:
//   out = qn + c1*(qs - qn) + c2*rhs(BC(qs), psi)   (qn given)
//   out = c1*qs + c2*rhs                           (qn null)
struct StageArgs {
  const float* qs;   // stage state (raw, ghost values kept in the combination)
  const float* psi;  // streamfunction of BC(qs)
  const float* qn;   // step start state (may be null when c0 == 0)
  float* out;
  int Nx, Ny, nl;
  long plane;        // Ny*Nx
  const float* dbeta_row;  // Ny: beta_y[j+1] - beta_y[j] (last entry unused)
  const float* wind_row;  // Ny
  float wind_coef[MAX_LAYERS];  // tau0 * wind_scale[l]
  float nu, nu4, kappa;         // kappa acts on layer nl-1 only
  float idx2, idy2, iarakawa;   // 1/dx^2, 1/dy^2, 1/(12 dx dy)
  float idx, idy;               // 1/dx, 1/dy (WENO)
  float psi_wall[MAX_LAYERS];   // WENO corner value on the walls
  float c1, c2;
  int advection;
  int apply_bc;  // 1: ghost ring of qs read as 0 in the RHS (SomaxModel.build_terms)
};
void launch_qg_stage(const StageArgs& a, hipStream_t st);

// Ghost ring of nl planes of Ny x Nx set to 0 (finitevolx.zero_boundaries).
void launch_zero_ghost(float* f, int nl, int Ny, int Nx, hipStream_t st);

}  // namespace somax
